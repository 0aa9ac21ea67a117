/*
 * gdp_oracle.c -- CPU restatement of GMAP's banded affine-gap DP core.
 *
 * TEST INFRASTRUCTURE ONLY (see gdp_oracle.h).  Scalar C, one cell at a time,
 * in the same visiting order as the reference's AVX2 build: row blocks of
 * 32 (int8) or 16 (int16) query rows, columns left to right inside a block,
 * the E/H update for every row of the block, then the vertical (F) chain over
 * the in-band rows in 32-bit ints.  Reference paths are relative to src/.
 */
#include "gdp_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- tables */

#define MM_HIGHQ (-3)  /* dynprog.h:48-50 */
#define MM_MEDQ (-2)
#define MM_LOWQ (-1)
#define MM_ENDQ (-5)   /* dynprog.c:103 */
#define SC_FULLMATCH 3 /* dynprog.h:40-42 */
#define SC_HALFMATCH 1
#define SC_AMBIGUOUS 3

static short pd[4][128][128];
static int nt2int[128];
static unsigned char consistent_tab[3][128][128];
static int use8p[4];

static int lc(int ch) { return (ch >= 'A' && ch <= 'Z') ? ch - 'A' + 'a' : ch; }

/* permute_cases (dynprog.c:908-977): every case combination, both orders. */
static void set_pair(int a, int b, short score, int mode) {
  int ca[2] = {a, lc(a)}, cb[2] = {b, lc(b)};
  int stranded = (mode == 0 || mode == 1 || mode == 3 || mode == 5);
  for (int i = 0; i < 2; i++) {
    for (int j = 0; j < 2; j++) {
      int x = ca[i], y = cb[j];
      if (stranded) {
        consistent_tab[0][x][y] = consistent_tab[0][y][x] = 1;
      } else {
        consistent_tab[1][x][y] = consistent_tab[1][y][x] = 1;
        consistent_tab[2][x][y] = consistent_tab[2][y][x] = 1;
      }
      for (int t = 0; t < 4; t++) pd[t][x][y] = pd[t][y][x] = score;
    }
  }
}

/* permute_cases_oneway (dynprog.c:980-1003): one direction, one genestrand. */
static void set_oneway(int a, int b, short score, int genestrand) {
  int ca[2] = {a, lc(a)}, cb[2] = {b, lc(b)};
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      consistent_tab[genestrand][ca[i]][cb[j]] = 1;
      for (int t = 0; t < 4; t++) pd[t][ca[i]][cb[j]] = score;
    }
}

/* Dynprog_init (dynprog.c:1006-1196). */
void orc_init(int mode) {
  static const int mm[4] = {MM_HIGHQ, MM_MEDQ, MM_LOWQ, MM_ENDQ};
  memset(pd, 0, sizeof(pd));
  memset(consistent_tab, 0, sizeof(consistent_tab));
  for (int j = 0; j < 128; j++) nt2int[j] = 4;
  nt2int['A'] = nt2int['a'] = 0;
  nt2int['C'] = nt2int['c'] = 1;
  nt2int['G'] = nt2int['g'] = 2;
  nt2int['T'] = nt2int['t'] = 3;
  for (int t = 0; t < 4; t++) use8p[t] = -128 / mm[t] - 1;

  /* Mismatch everywhere in 'A'..'z' x 'A'..'y' (note: second index < 'z'). */
  for (int c1 = 'A'; c1 <= 'z'; c1++)
    for (int c2 = 'A'; c2 < 'z'; c2++)
      for (int t = 0; t < 4; t++) pd[t][c1][c2] = (short)mm[t];

  for (int c = 'A'; c < 'Z'; c++) set_pair(c, c, SC_FULLMATCH, mode);
  set_pair('U', 'T', SC_FULLMATCH, mode);
  static const char *half[] = {"RA", "RG", "YT", "YC", "WA", "WT", "SG", "SC", "MA", "MC", "KG", "KT"};
  for (unsigned k = 0; k < sizeof(half) / sizeof(half[0]); k++) set_pair(half[k][0], half[k][1], SC_HALFMATCH, mode);
  static const char *amb[] = {"HA", "HT", "HC", "BG", "BC", "BT", "VG", "VA", "VC", "DG", "DA", "DT",
                              "NT", "NC", "NA", "NG", "XT", "XC", "XA", "XG", "NN", "XX"};
  for (unsigned k = 0; k < sizeof(amb) / sizeof(amb[0]); k++) set_pair(amb[k][0], amb[k][1], SC_AMBIGUOUS, mode);

  switch (mode) {
  case 1: set_oneway('T', 'C', SC_FULLMATCH, 0); break;                                        /* CMET_STRANDED */
  case 2: set_oneway('T', 'C', SC_FULLMATCH, 1); set_oneway('A', 'G', SC_FULLMATCH, 2); break; /* CMET_NONSTRANDED */
  case 3: set_oneway('G', 'A', SC_FULLMATCH, 0); break;                                        /* ATOI_STRANDED */
  case 4: set_oneway('G', 'A', SC_FULLMATCH, 1); set_oneway('C', 'T', SC_FULLMATCH, 2); break; /* ATOI_NONSTRANDED */
  case 5: set_oneway('C', 'T', SC_FULLMATCH, 0); break;                                        /* TTOC_STRANDED */
  case 6: set_oneway('C', 'T', SC_FULLMATCH, 1); set_oneway('G', 'A', SC_FULLMATCH, 2); break; /* TTOC_NONSTRANDED */
  default: break;
  }
}

const short *orc_pairdistance(int mismatchtype) { return &pd[mismatchtype][0][0]; }
const int *orc_nt_to_int(void) { return nt2int; }
int orc_use8p_size(int mismatchtype) { return use8p[mismatchtype]; }

/* Dynprog_consistent_p (dynprog.c:893-900). */
int orc_consistent(int genestrand, int c, int g) { return consistent_tab[genestrand][c & 127][g & 127]; }

void orc_compute_bands(int *lband, int *uband, int rlength, int glength, int extraband, int widebandp) {
  if (!widebandp) {
    *lband = extraband;
    *uband = extraband;
  } else if (glength >= rlength) {
    *uband = glength - rlength + extraband;
    *lband = extraband;
  } else {
    *lband = rlength - glength + extraband;
    *uband = extraband;
  }
}

/* ------------------------------------------------------------ fill helpers */

static const char LETTERS[5] = {'A', 'C', 'G', 'T', 'N'};

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int clampw(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
/* Tie rule: jump-late uses >=, jump-early uses >. */
static inline int wins(int a, int b, int late) { return late ? (a >= b) : (a > b); }

static inline int seq_at(const char *s, int revp, int pos1) {
  /* Reference indexing: forward s[pos-1], reversed s[1-pos]. */
  return (int)(signed char)(revp ? s[1 - pos1] : s[pos1 - 1]);
}

#define DIR_H 1
#define DIR_V 2
#define DIR_E 4
#define DIR_F 8

/* ------------------------------------------------------------- full band */

void orc_fill_full(int width, const char *rseq, const char *gseq, const char *gseq_alt,
                   int rl, int gl, int mismatchtype, int open, int extend,
                   int lband, int uband, int late, int revp,
                   short *H, unsigned char *dirs) {
  const int NEG = width == 8 ? -128 : -32768, POS = width == 8 ? 127 : 32767;
  const int S = width == 8 ? 32 : 16; /* SIMD_NCHARS / SIMD_NSHORTS, dynprog.h:122-155 */
  const int ld = rl + 1;
  int (*ps)[5] = malloc(sizeof(int[5]) * (size_t)(rl + 1));
  int *FF = malloc(sizeof(int) * (size_t)(gl + 1));
  int Hprev[32], Hold[32], E[32];

  /* Query profile over the five genome classes (dynprog_simd.c:3164-3189). */
  for (int x = 0; x < 5; x++) ps[0][x] = pd[mismatchtype]['N'][(int)LETTERS[x]];
  for (int r = 1; r <= rl; r++) {
    int q = seq_at(rseq, revp, r);
    for (int x = 0; x < 5; x++) ps[r][x] = pd[mismatchtype][q & 127][(int)LETTERS[x]];
  }

  for (int rlo = 0; rlo <= rl; rlo += S) {
    int rhigh = imin(rlo + S - 1, rl), nl = rhigh - rlo + 1;
    int c = imax(0, rlo - lband);
    /* Block start (dynprog_simd.c:3221-3254): E = -inf (+1 when jump-early),
       H = -inf - open so that H + open = -inf. */
    for (int i = 0; i < S; i++) {
      Hprev[i] = NEG - open;
      E[i] = late ? NEG : NEG + 1;
    }
    for (; c <= rhigh + uband && c <= gl; c++) {
      int x = 4, xa = 4, X;
      if (c == 0) {
        X = rlo == 0 ? 0 : NEG;
      } else {
        x = nt2int[seq_at(gseq, revp, c) & 127];
        xa = nt2int[seq_at(gseq_alt, revp, c) & 127];
        X = rlo == 0 ? NEG : H[(size_t)(c - 1) * ld + rlo - 1];
      }
      memcpy(Hold, Hprev, sizeof(Hold));
      for (int i = 0; i < nl; i++) {
        int r = rlo + i;
        /* E: horizontal gap (dynprog_simd.c:3316-3333). */
        int T1 = clampw(Hold[i] + open, NEG, POS);
        int dE = wins(E[i], T1, late);
        E[i] = clampw(imax(E[i], T1) + extend, NEG, POS);
        /* H from the diagonal plus the best of the two genome alternatives
           (3337-3388); column 0 scores are 0 at r=0 and -inf below. */
        int diag = i == 0 ? X : Hold[i - 1];
        int sc = c == 0 ? (r == 0 ? 0 : NEG) : imax(ps[r][x], ps[r][xa]);
        int Hd = clampw(diag + sc, NEG, POS);
        int dH = wins(E[i], Hd, late);
        H[(size_t)c * ld + r] = (short)imax(Hd, E[i]);
        dirs[(size_t)c * ld + r] = (unsigned char)((dH ? DIR_H : 0) | (dE ? DIR_E : 0));
      }

      /* Bottom band edge: diagonal only, E is kept (3390-3416). */
      int rhigh_calc = rhigh;
      if (rhigh >= c + lband) {
        rhigh_calc = c + lband;
        if (c > 0) {
          int r = rhigh_calc;
          int sc = imax(ps[r][x], ps[r][xa]);
          H[(size_t)c * ld + r] = (short)clampw(H[(size_t)(c - 1) * ld + r - 1] + sc, NEG, POS);
          dirs[(size_t)c * ld + r] &= (unsigned char)~(DIR_H | DIR_E);
        }
      }

      /* F: vertical gap in 32-bit ints over the in-band rows (3419-3476). */
      int rlo_calc = imax(rlo, c - uband);
      int cg, last;
      if (rlo == 0 || c >= rlo + uband) {
        cg = last = -32768;
      } else {
        cg = FF[c];
        last = H[(size_t)c * ld + rlo_calc - 1];
      }
      int r = rlo_calc;
      if (r == c - uband) { /* top band edge: F restarts, no F direction */
        cg = last + open + extend;
        last = H[(size_t)c * ld + r];
        r++;
      }
      for (; r <= rhigh_calc; r++) {
        size_t k = (size_t)c * ld + r;
        if (wins(cg, last + open, late)) {
          cg += extend;
          dirs[k] |= DIR_F;
        } else {
          cg = last + open + extend;
        }
        last = H[k];
        if (wins(cg, last, late)) {
          last = cg;
          H[k] = (short)clampw(cg, NEG, POS);
          dirs[k] = (unsigned char)((dirs[k] & ~DIR_H) | DIR_V);
        }
      }
      FF[c] = cg;
      for (int i = 0; i < nl; i++) Hprev[i] = H[(size_t)c * ld + rlo + i];
    }
  }
  free(FF);
  free(ps);
}

/* -------------------------------------------------------- one-sided fills */

/* Shared body of the upper (row-blocked, [c][r]) and lower (column-blocked,
   [r][c]) triangles.  "lanes" are the blocked index (query rows for upper,
   genome columns for lower); "steps" are the swept index.  prof[lane][x] is
   the lane's score against class x; cls[step] the class of the swept index.
   E is capped to -inf on and beyond the main diagonal and the diagonal's
   directions are forced to DIAG (dynprog_simd.c:7867-7960, 8759-8860). */
static void onesided(int width, int nlanes_total, int nsteps_total, int band,
                     const int (*prof)[5], const int *cls, const int *cls_alt,
                     int open, int extend, int late, short *H, unsigned char *dirs, int ld) {
  const int NEG = width == 8 ? -128 : -32768, POS = width == 8 ? 127 : 32767;
  const int S = width == 8 ? 32 : 16;
  int Hprev[32], Hold[32], E[32];
  for (int lo = 0; lo <= nlanes_total; lo += S) {
    int hi = imin(lo + S - 1, nlanes_total), nl = hi - lo + 1;
    for (int i = 0; i < S; i++) {
      Hprev[i] = NEG - open;
      E[i] = late ? NEG : NEG + 1;
    }
    for (int s = lo; s <= hi + band && s <= nsteps_total; s++) {
      int X = s == 0 ? 0 : (lo == 0 ? NEG : H[(size_t)(lo - 1) * ld + s - 1]);
      memcpy(Hold, Hprev, sizeof(Hold));
      for (int i = 0; i < nl; i++) {
        int lane = lo + i;
        int cap = lane >= s ? NEG : POS;
        int e0 = imin(E[i], cap);
        int T1 = clampw(Hold[i] + open, NEG, POS);
        int dE = wins(e0, T1, late);
        E[i] = imin(clampw(imax(e0, T1) + extend, NEG, POS), cap);
        int diag = i == 0 ? X : Hold[i - 1];
        int sc = imax(prof[lane][cls[s]], prof[lane][cls_alt[s]]);
        int Hd = clampw(diag + sc, NEG, POS);
        int dH = wins(E[i], Hd, late);
        H[(size_t)lane * ld + s] = (short)imax(Hd, E[i]);
        dirs[(size_t)lane * ld + s] = (unsigned char)((dH ? DIR_H : 0) | (dE ? DIR_E : 0));
      }
      if (hi >= s) dirs[(size_t)s * ld + s] = 0;
      for (int i = 0; i < nl; i++) Hprev[i] = H[(size_t)(lo + i) * ld + s];
    }
  }
}

/* The helper indexes H[lane*ld + step]; upper stores [c][r] = [step][lane],
   so upper runs on a transposed scratch and is copied back. */
void orc_fill_upper(int width, const char *rseq, const char *gseq, const char *gseq_alt,
                    int rl, int gl, int mismatchtype, int open, int extend,
                    int uband, int late, int revp, short *H, unsigned char *dirs) {
  int (*prof)[5] = malloc(sizeof(int[5]) * (size_t)(rl + 1));
  int *cls = malloc(sizeof(int) * (size_t)(gl + 1)), *clsa = malloc(sizeof(int) * (size_t)(gl + 1));
  short *Ht = malloc(sizeof(short) * (size_t)(rl + 1) * (gl + 1));
  unsigned char *Dt = malloc((size_t)(rl + 1) * (gl + 1));
  memset(Ht, 0, sizeof(short) * (size_t)(rl + 1) * (gl + 1));
  memset(Dt, 0, (size_t)(rl + 1) * (gl + 1));
  /* Query profile (dynprog_simd.c:7795-7820); column 0 uses class N (7841). */
  for (int x = 0; x < 5; x++) prof[0][x] = pd[mismatchtype]['N'][(int)LETTERS[x]];
  for (int r = 1; r <= rl; r++) {
    int q = seq_at(rseq, revp, r);
    for (int x = 0; x < 5; x++) prof[r][x] = pd[mismatchtype][q & 127][(int)LETTERS[x]];
  }
  cls[0] = clsa[0] = 4;
  for (int c = 1; c <= gl; c++) {
    cls[c] = nt2int[seq_at(gseq, revp, c) & 127];
    clsa[c] = nt2int[seq_at(gseq_alt, revp, c) & 127];
  }
  onesided(width, rl, gl, uband, (const int(*)[5])prof, cls, clsa, open, extend, late, Ht, Dt, gl + 1);
  /* Only cells the reference computes are defined; copy those back. */
  const int S = width == 8 ? 32 : 16;
  for (int rlo = 0; rlo <= rl; rlo += S) {
    int rhigh = imin(rlo + S - 1, rl);
    for (int c = rlo; c <= rhigh + uband && c <= gl; c++)
      for (int r = rlo; r <= rhigh; r++) {
        H[(size_t)c * (rl + 1) + r] = Ht[(size_t)r * (gl + 1) + c];
        dirs[(size_t)c * (rl + 1) + r] = Dt[(size_t)r * (gl + 1) + c];
      }
  }
  free(Dt);
  free(Ht);
  free(clsa);
  free(cls);
  free(prof);
}

void orc_fill_lower(int width, const char *rseq, const char *gseq, const char *gseq_alt,
                    int rl, int gl, int mismatchtype, int open, int extend,
                    int lband, int late, int revp, short *H, unsigned char *dirs) {
  int (*prof)[5] = malloc(sizeof(int[5]) * (size_t)(gl + 1));
  int *cls = malloc(sizeof(int) * (size_t)(rl + 1));
  short *Ht = malloc(sizeof(short) * (size_t)(rl + 1) * (gl + 1));
  unsigned char *Dt = malloc((size_t)(rl + 1) * (gl + 1));
  memset(Ht, 0, sizeof(short) * (size_t)(rl + 1) * (gl + 1));
  memset(Dt, 0, (size_t)(rl + 1) * (gl + 1));
  /* Genome profile over the five query classes, max over the two genome
     alternatives taken up front (dynprog_simd.c:8660-8720).  Column 0 is
     scored against 'N' in the int16 kernel (8658) but against the byte
     value 4 in the int8 kernel (5457), which the table leaves at 0. */
  for (int x = 0; x < 5; x++) prof[0][x] = pd[mismatchtype][(int)LETTERS[x]][width == 8 ? 4 : 'N'];
  for (int c = 1; c <= gl; c++) {
    int g = seq_at(gseq, revp, c) & 127, ga = seq_at(gseq_alt, revp, c) & 127;
    for (int x = 0; x < 5; x++) prof[c][x] = imax(pd[mismatchtype][(int)LETTERS[x]][g], pd[mismatchtype][(int)LETTERS[x]][ga]);
  }
  cls[0] = 4;
  for (int r = 1; r <= rl; r++) cls[r] = nt2int[seq_at(rseq, revp, r) & 127];
  /* lanes = genome columns, steps = query rows: the helper stores
     Ht[c*(rl+1) + r]. */
  onesided(width, gl, rl, lband, (const int(*)[5])prof, cls, cls, open, extend, late, Ht, Dt, rl + 1);
  const int S = width == 8 ? 32 : 16;
  for (int clo = 0; clo <= gl; clo += S) {
    int chigh = imin(clo + S - 1, gl);
    for (int r = clo; r <= chigh + lband && r <= rl; r++)
      for (int c = clo; c <= chigh; c++) {
        H[(size_t)r * (gl + 1) + c] = Ht[(size_t)c * (rl + 1) + r];
        dirs[(size_t)r * (gl + 1) + c] = Dt[(size_t)c * (rl + 1) + r];
      }
  }
  free(Dt);
  free(Ht);
  free(cls);
  free(prof);
}

/* ------------------------------------------------------------- traceback */

#define TB_MATCH 1 /* scores.h:4-9 */
#define TB_MISMATCH (-3)
#define TB_QOPEN (-3)
#define TB_QINDEL (-1)
#define TB_TOPEN (-3)
#define TB_TINDEL (-1)
#define MICROINTRON_LENGTH 9 /* pairpool.c:18 */

typedef struct {
  int *pairs;
  int n, cap;
} pairbuf_t;

/* Pairpool_push (pairpool.c:179-259) drops pairs at negative coordinates. */
static void push_pair(pairbuf_t *pb, int qpos, int gpos, int cdna, int comp, int g, int galt) {
  if (qpos < 0 || gpos < 0) return;
  if (pb->n < pb->cap) {
    int *p = pb->pairs + (size_t)pb->n * ORC_PAIR_FIELDS;
    p[0] = qpos; p[1] = gpos; p[2] = cdna; p[3] = comp; p[4] = g; p[5] = galt; p[6] = 0; p[7] = 0;
  }
  pb->n++;
}

/* Pairpool_push_gapholder (pairpool.c:374-452) with queryjump 0. */
static void push_gapholder(pairbuf_t *pb, int genomejump) {
  if (pb->n < pb->cap) {
    int *p = pb->pairs + (size_t)pb->n * ORC_PAIR_FIELDS;
    p[0] = -1; p[1] = -1; p[2] = ' '; p[3] = ' '; p[4] = ' '; p[5] = ' '; p[6] = 1; p[7] = genomejump;
  }
  pb->n++;
}

static int complement_char(int ch) {
  switch (ch) {
  case 'A': return 'T';
  case 'C': return 'G';
  case 'G': return 'C';
  case 'T': return 'A';
  case 'a': return 't';
  case 'c': return 'g';
  case 'g': return 'c';
  case 't': return 'a';
  default: return ch; /* N, n, '*' map to themselves in COMPLEMENT_LC */
  }
}

static int genome_char(const orc_genome_t *genome, unsigned int pos) {
  if (genome == NULL || pos >= genome->length) return 'N';
  const unsigned int count = genome->count ? genome->count : genome->length;
  if (pos < genome->base || pos - genome->base >= count) return '?';  /* outside the arrays given */
  pos -= genome->base;
  if (genome->nflags && genome->nflags[pos]) return 'N';
  return "ACGT"[genome->codes[pos] & 3];
}

/* get_genomic_nt (pairpool.c:1016-1063); the alternate genome is the genome. */
static int genomic_nt(int *alt, const orc_genome_t *genome, int genomicpos,
                      unsigned int chroffset, unsigned int chrhigh, int watsonp) {
  unsigned int pos;
  if (watsonp) {
    pos = chroffset + (unsigned int)genomicpos;
    if (pos < chroffset || pos >= chrhigh) { *alt = '*'; return '*'; }
    int ch = genome_char(genome, pos);
    *alt = ch;
    return ch;
  }
  pos = chrhigh - (unsigned int)genomicpos;
  if (pos < chroffset || pos >= chrhigh) { *alt = '*'; return '*'; }
  int ch = complement_char(genome_char(genome, pos));
  *alt = ch;
  return ch;
}

/* Pairpool_add_queryskip (pairpool.c:980-1010). */
static void queryskip(pairbuf_t *pb, int r, int c, int dist, const char *rseq,
                      int queryoffset, int genomeoffset, int revp) {
  int qc = r - 1, gc = c - 1, step = -1;
  if (revp) { qc = -qc; gc = -gc; step = +1; }
  for (int j = 0; j < dist; j++) {
    push_pair(pb, queryoffset + qc, genomeoffset + gc, rseq[qc], '-', ' ', ' ');
    qc += step;
  }
}

/* Pairpool_add_genomeskip (pairpool.c:1067-1180) with genomesequence NULL. */
static int genomeskip(pairbuf_t *pb, int r, int c, int dist, int queryoffset, int genomeoffset,
                      const orc_genome_t *genome, int revp, unsigned int chroffset,
                      unsigned int chrhigh, int watsonp) {
  int qc = r - 1, left = c - dist, right = c - 1, step = -1;
  if (revp) {
    int t = left;
    qc = -qc;
    left = -right;
    right = -t;
    step = +1;
  }
  if (dist < MICROINTRON_LENGTH) {
    int gc = revp ? left : right;
    for (int j = 0; j < dist; j++) {
      int alt, ch = genomic_nt(&alt, genome, genomeoffset + gc, chroffset, chrhigh, watsonp);
      push_pair(pb, queryoffset + qc, genomeoffset + gc, ' ', '-', ch, alt);
      gc += step;
    }
    return 1;
  }
  push_gapholder(pb, dist);
  return 0;
}

static void diag_pair(pairbuf_t *pb, int *counters, int r, int c, const char *rseq, const char *rsequc,
                      const char *gseq, const char *gseq_alt, int queryoffset, int genomeoffset,
                      int revp, int genestrand) {
  int qc = r - 1, gc = c - 1;
  if (revp) { qc = -qc; gc = -gc; }
  int c1 = rseq[qc], c1uc = rsequc[qc], c2 = gseq[gc], c2a = gseq_alt[gc];
  if (c2 == '*') return; /* past the chromosome end: nothing pushed */
  if (c1uc == c2 || c1uc == c2a) {
    counters[0] += TB_MATCH; counters[1] += 1;
    push_pair(pb, queryoffset + qc, genomeoffset + gc, c1, '*', c2, c2a);
  } else if (orc_consistent(genestrand, c1uc, c2) || orc_consistent(genestrand, c1uc, c2a)) {
    counters[0] += TB_MATCH; counters[1] += 1;
    push_pair(pb, queryoffset + qc, genomeoffset + gc, c1, ':', c2, c2a);
  } else {
    counters[0] += TB_MISMATCH; counters[2] += 1;
    push_pair(pb, queryoffset + qc, genomeoffset + gc, c1, ' ', c2, c2a);
  }
}

/* traceback_nogaps (dynprog_end.c:648-722): diagonal pairs from (r, c) while
   r > 0 and c > 0, scored as in the gapped walks, no closing skip. */
int orc_traceback_nogaps(int r, int c, const char *rseq, const char *rsequc, const char *gseq, const char *gseq_alt,
                         int queryoffset, int genomeoffset, int revp, int genestrand, int *counters, int *pairs,
                         int maxpairs) {
  pairbuf_t pb = {pairs, 0, maxpairs};
  while (r > 0 && c > 0) {
    diag_pair(&pb, counters, r, c, rseq, rsequc, gseq, gseq_alt, queryoffset, genomeoffset, revp, genestrand);
    r--;
    c--;
  }
  return pb.n;
}

int orc_traceback(int kind, const unsigned char *dirs, int rl, int gl, int r, int c,
                  const char *rseq, const char *rsequc, const char *gseq, const char *gseq_alt,
                  int queryoffset, int genomeoffset, const orc_genome_t *genome,
                  int revp, unsigned int chroffset, unsigned int chrhigh, int watsonp,
                  int genestrand, int dynprogindex, int *counters, int *pairs, int maxpairs) {
  pairbuf_t pb = {pairs, 0, maxpairs};
  (void)dynprogindex;
  const int ldc = rl + 1, ldr = gl + 1;
#define DCR(cc, rr) dirs[(size_t)(cc) * ldc + (rr)] /* [c][r] for full/upper */
#define DRC(rr, cc) dirs[(size_t)(rr) * ldr + (cc)] /* [r][c] for lower */
  if (kind == ORC_KIND_FULL) {
    while (r > 0 && c > 0) {
      int d = DCR(c, r);
      if ((d & DIR_H) && !(d & DIR_V)) {
        int dist = 1;
        while (c > 0 && (DCR(c--, r) & DIR_E)) dist++;
        if (genomeskip(&pb, r, c + dist, dist, queryoffset, genomeoffset, genome, revp, chroffset, chrhigh, watsonp)) {
          counters[0] += TB_TOPEN + dist * TB_TINDEL; counters[3] += 1; counters[4] += dist;
        }
      } else if (d & DIR_V) {
        int dist = 1;
        while (r > 0 && (DCR(c, r--) & DIR_F)) dist++;
        queryskip(&pb, r + dist, c, dist, rseq, queryoffset, genomeoffset, revp);
        counters[0] += TB_QOPEN + dist * TB_QINDEL; counters[3] += 1; counters[4] += dist;
      } else {
        diag_pair(&pb, counters, r, c, rseq, rsequc, gseq, gseq_alt, queryoffset, genomeoffset, revp, genestrand);
        r--; c--;
      }
    }
    if (r == 0 && c == 0) {
    } else if (c == 0) {
      queryskip(&pb, r, 1, r, rseq, queryoffset, genomeoffset, revp);
      counters[0] += TB_QOPEN + r * TB_QINDEL; counters[3] += 1; counters[4] += r;
    } else {
      int dist = c;
      if (genomeskip(&pb, 1, c, dist, queryoffset, genomeoffset, genome, revp, chroffset, chrhigh, watsonp)) {
        counters[0] += TB_TOPEN + dist * TB_TINDEL; counters[3] += 1; counters[4] += dist;
      }
    }
  } else if (kind == ORC_KIND_UPPER) {
    while (r > 0 && c > 0) {
      if (DCR(c, r) & (DIR_H | DIR_V)) {
        int dist = 1;
        while (DCR(c--, r) & DIR_E) dist++;
        if (genomeskip(&pb, r, c + dist, dist, queryoffset, genomeoffset, genome, revp, chroffset, chrhigh, watsonp)) {
          counters[0] += TB_TOPEN + dist * TB_TINDEL; counters[3] += 1; counters[4] += dist;
        }
      } else {
        diag_pair(&pb, counters, r, c, rseq, rsequc, gseq, gseq_alt, queryoffset, genomeoffset, revp, genestrand);
        r--; c--;
      }
    }
    if (c != 0) {
      int dist = c;
      if (genomeskip(&pb, 1, c, dist, queryoffset, genomeoffset, genome, revp, chroffset, chrhigh, watsonp)) {
        counters[0] += TB_TOPEN + dist * TB_TINDEL; counters[3] += 1; counters[4] += dist;
      }
    }
  } else {
    while (r > 0 && c > 0) {
      if (DRC(r, c) & (DIR_H | DIR_V)) {
        int dist = 1;
        while (DRC(r--, c) & DIR_E) dist++;
        queryskip(&pb, r + dist, c, dist, rseq, queryoffset, genomeoffset, revp);
        counters[0] += TB_QOPEN + dist * TB_QINDEL; counters[3] += 1; counters[4] += dist;
      } else {
        diag_pair(&pb, counters, r, c, rseq, rsequc, gseq, gseq_alt, queryoffset, genomeoffset, revp, genestrand);
        r--; c--;
      }
    }
    if (r != 0) {
      queryskip(&pb, r, 1, r, rseq, queryoffset, genomeoffset, revp);
      counters[0] += TB_QOPEN + r * TB_QINDEL; counters[3] += 1; counters[4] += r;
    }
  }
#undef DCR
#undef DRC
  return pb.n;
}

/* traceback_local_{8,16}_upper / _lower (dynprog_end.c:932-1135; the int8
   forms 727-931 are the same walk): the one-sided walk of the splice-junction
   end variants.  It stops at genome column endc instead of 0 -- the upper
   walk's gap runs stop there too (960) --, pushes every diagonal pair (no
   '*' test), reads its genome-skip characters from the fill's own genome
   sequence (Pairpool_add_genomeskip with genomesequence, pairpool.c:1142-1149,
   genomealt = genome), and closes with a genome skip of c - endc columns at
   row 1 (upper, 1019-1031) or a query skip of r rows at column endc + 1
   (lower, 1121-1133).  end_rc receives where the walk stopped. */
int orc_traceback_local(int kind, const unsigned char *dirs, int rl, int gl, int r, int c, int endc,
                        const char *rseq, const char *rsequc, const char *gseq, const char *gseq_alt,
                        int queryoffset, int genomeoffset, int revp, int genestrand,
                        int *counters, int *pairs, int maxpairs, int *end_rc) {
  pairbuf_t pb = {pairs, 0, maxpairs};
  const int ldc = rl + 1, ldr = gl + 1;
#define DCR(cc, rr) dirs[(size_t)(cc) * ldc + (rr)]
#define DRC(rr, cc) dirs[(size_t)(rr) * ldr + (cc)]
  while (r > 0 && c > endc) {
    const int d = kind == ORC_KIND_UPPER ? DCR(c, r) : DRC(r, c);
    if (d & (DIR_H | DIR_V)) {
      int dist = 1;
      if (kind == ORC_KIND_UPPER) {
        while (c > endc && (DCR(c--, r) & DIR_E)) dist++;
        /* the genome skip of (r, c + dist, dist), characters from gseq */
        int qc = r - 1, left = c + dist - dist, right = c + dist - 1, step = -1;
        if (revp) { int t = left; qc = -qc; left = -right; right = -t; step = +1; }
        if (dist < MICROINTRON_LENGTH) {
          int gc = revp ? left : right;
          for (int j = 0; j < dist; j++, gc += step)
            push_pair(&pb, queryoffset + qc, genomeoffset + gc, ' ', '-', gseq[gc], gseq[gc]);
          counters[0] += TB_TOPEN + dist * TB_TINDEL; counters[3] += 1; counters[4] += dist;
        } else {
          push_gapholder(&pb, dist);
        }
      } else {
        while (DRC(r--, c) & DIR_E) dist++;
        queryskip(&pb, r + dist, c, dist, rseq, queryoffset, genomeoffset, revp);
        counters[0] += TB_QOPEN + dist * TB_QINDEL; counters[3] += 1; counters[4] += dist;
      }
    } else {
      int qc = r - 1, gc = c - 1;
      if (revp) { qc = -qc; gc = -gc; }
      const int c1 = rseq[qc], c1uc = rsequc[qc], c2 = gseq[gc], c2a = gseq_alt[gc];
      if (c1uc == c2 || c1uc == c2a) {
        counters[0] += TB_MATCH; counters[1] += 1;
        push_pair(&pb, queryoffset + qc, genomeoffset + gc, c1, '*', c2, c2a);
      } else if (orc_consistent(genestrand, c1uc, c2) || orc_consistent(genestrand, c1uc, c2a)) {
        counters[0] += TB_MATCH; counters[1] += 1;
        push_pair(&pb, queryoffset + qc, genomeoffset + gc, c1, ':', c2, c2a);
      } else {
        counters[0] += TB_MISMATCH; counters[2] += 1;
        push_pair(&pb, queryoffset + qc, genomeoffset + gc, c1, ' ', c2, c2a);
      }
      r--; c--;
    }
  }
  end_rc[0] = r;
  end_rc[1] = c;
  if (kind == ORC_KIND_UPPER) {
    if (c != endc) {
      const int dist = c - endc;
      int qc = 0, left = c - dist, right = c - 1, step = -1;
      if (revp) { int t = left; qc = -qc; left = -right; right = -t; step = +1; }
      if (dist < MICROINTRON_LENGTH) {
        int gc = revp ? left : right;
        for (int j = 0; j < dist; j++, gc += step)
          push_pair(&pb, queryoffset + qc, genomeoffset + gc, ' ', '-', gseq[gc], gseq[gc]);
        counters[0] += TB_TOPEN + dist * TB_TINDEL; counters[3] += 1; counters[4] += dist;
      } else {
        push_gapholder(&pb, dist);
      }
    }
  } else if (r != 0) {
    queryskip(&pb, r, endc + 1, r, rseq, queryoffset, genomeoffset, revp);
    counters[0] += TB_QOPEN + r * TB_QINDEL; counters[3] += 1; counters[4] += r;
  }
#undef DCR
#undef DRC
  return pb.n;
}

/* ------------------------------------------------------ best end-point scan */

void orc_best_endpoint(int queryend_indels, const short *Hupper, const short *Hlower,
                       int width, int rl, int gl, int lband, int uband, int late,
                       int *finalscore, int *bestr, int *bestc) {
  int best = queryend_indels ? (width == 8 ? -128 : -32768) : 0;
  int r0 = queryend_indels ? rl : 1;
  *bestr = queryend_indels ? rl : 0;
  *bestc = 0;
  for (int r = r0; r <= rl; r++) {
    int clo = imax(1, r - lband), chigh = imin(gl, r + uband), c;
    for (c = clo; c < r; c++) {
      int v = Hlower[(size_t)r * (gl + 1) + c];
      if (wins(v, best, late)) { *bestr = r; *bestc = c; best = v; }
    }
    for (; c <= chigh; c++) {
      int v = Hupper[(size_t)c * (rl + 1) + r];
      if (wins(v, best, late)) { *bestr = r; *bestc = c; best = v; }
    }
  }
  *finalscore = best;
}

/* ------------------------------------------------------------ cDNA bridge */

/* bridge_cdna_gap_{8,16}_ud (dynprog_cdna.c:386-644; the int8 form 123-381
   differs only in its initial score).  Candidates are visited in the
   reference's order -- cL ascending, cR descending from gl-cL, rL
   ascending, rR ascending -- and a
   candidate replaces the best under >= (jump late) or > (jump early).
   Each end's rows run from max(1, c-uband) up the upper triangle to the
   diagonal (no cap at rl-1 there, 417-418), then down the lower triangle to
   min(c+lband, rl-1) (407-414); the right end's rows also stay below
   rightoffset-leftoffset-rL (422-423).  A genomic insertion of
   k = gl-cL-cR columns costs 0 when k = 0 and open otherwise: the loop body
   resets the penalty to open-extend (496-497) before the step adds extend
   back, so it never grows with k.  An upper row beyond rl, reachable
   only when rl < gl-1, reads rows the fill never wrote in the reference
   (stale workspace); here such candidates are skipped. */
static inline int bridge_score(const short *U, const short *Lo, int gl, int rl, int r, int c, int upper) {
  return upper ? U[(size_t)c * (rl + 1) + r] : Lo[(size_t)r * (gl + 1) + c];
}

void orc_bridge_cdna(int width, const short *LU, const short *LL, const short *RU, const short *RL,
                     int gl, int rlL, int rlR, int lbandL, int ubandL, int lbandR, int ubandR,
                     int open, int extend, int leftoffset, int rightoffset, int late, int *out) {
  int best = width == 8 ? -128 : -32768;
  int bcL = -1, bcR = -1, brL = -1, brR = -1;
  for (int cL = 1; cL < gl; cL++) {
    for (int cR = gl - cL; cR >= 0; cR--) {
      const int k = gl - cL - cR; /* genomic columns left between the two ends */
      const int pen = k == 0 ? 0 : open;
      const int rloL = imax(1, cL - ubandL), rhighL = imin(rlL - 1, cL + lbandL);
      const int rloR = imax(1, cR - ubandR), rhighR = imin(rlR - 1, cR + lbandR);
      for (int rL = rloL; rL < cL || rL <= rhighL; rL++) {
        const int upL = rL < cL;
        if (upL && rL > rlL) continue;
        const int sL = bridge_score(LU, LL, gl, rlL, rL, cL, upL);
        const int lim = rightoffset - leftoffset - rL;
        for (int rR = rloR; (rR < cR || rR <= rhighR) && rR < lim; rR++) {
          const int upR = rR < cR;
          if (upR && rR > rlR) continue;
          const int v = sL + bridge_score(RU, RL, gl, rlR, rR, cR, upR) + pen;
          if (wins(v, best, late)) {
            best = v;
            bcL = cL;
            bcR = cR;
            brL = rL;
            brR = rR;
          }
        }
      }
    }
  }
  out[0] = best;
  out[1] = bcL;
  out[2] = bcR;
  out[3] = brL;
  out[4] = brR;
}

/* ---------------------------------------------------------- intron bridge */

/* Dinucleotide codes of intron.h:10-21 and the intron types their AND gives
   (intron.h:26-32): a left site reads the two bases after the left exon, a
   right site the two before the right exon; either allele may supply each
   base (dynprog_genome.c:1791-1843). */
static int left_site(int l1, int l1a, int l2, int l2a) {
#define EITHER(x, xa, b) ((x) == (b) || (xa) == (b))
  if (EITHER(l1, l1a, 'G') && EITHER(l2, l2a, 'T')) return 0x21;
  if (EITHER(l1, l1a, 'G') && EITHER(l2, l2a, 'C')) return 0x10;
  if (EITHER(l1, l1a, 'A') && EITHER(l2, l2a, 'T')) return 0x08;
  if (EITHER(l1, l1a, 'C') && EITHER(l2, l2a, 'T')) return 0x06;
  return 0;
}

static int right_site(int r2, int r2a, int r1, int r1a) {
  if (EITHER(r2, r2a, 'A') && EITHER(r1, r1a, 'G')) return 0x30;
  if (EITHER(r2, r2a, 'A') && EITHER(r1, r1a, 'C')) return 0x0C;
  if (EITHER(r2, r2a, 'G') && EITHER(r1, r1a, 'C')) return 0x02;
  if (EITHER(r2, r2a, 'A') && EITHER(r1, r1a, 'T')) return 0x01;
  return 0;
#undef EITHER
}

/* intron_score_setup (dynprog_genome.c:143-186) with the constants of its
   build (76-115: USE_SCOREI, no USE_WEAK_SCOREI): per intron type the
   reward of the sense (cdna_direction > 0), antisense (< 0) or either (0)
   table, final or preliminary.  The either tables keep the reference's
   values as set, including the preliminary GT-AG forward reward of 16 and
   the final CT-AC reward of 14. */
static int intron_reward(int cdna_direction, int finalp, int type) {
  const int fwd = type == 0x20 || type == 0x10 || type == 0x08;
  const int rev = type == 0x04 || type == 0x02 || type == 0x01;
  if (cdna_direction > 0 && !fwd) return 0;
  if (cdna_direction < 0 && !rev) return 0;
  if (!fwd && !rev) return 0;
  if (cdna_direction == 0) {
    static const int either_final[0x21] = {[0x20] = 16, [0x10] = 10, [0x08] = 8, [0x04] = 14, [0x02] = 10, [0x01] = 8};
    static const int either_prelim[0x21] = {[0x20] = 16, [0x10] = 8, [0x08] = 4, [0x04] = 14, [0x02] = 8, [0x01] = 4};
    return finalp ? either_final[type] : either_prelim[type];
  }
  const int canon = type == 0x20 || type == 0x04, gcag = type == 0x10 || type == 0x02;
  if (canon) return finalp ? 16 : 14;
  if (gcag) return finalp ? 10 : 8;
  return finalp ? 8 : 4; /* AT-AC */
}

void orc_bridge_intron(int width, const short *LU, const short *LL, const short *RU, const short *RL,
                       const char *gL, const char *gLalt, const char *gR, const char *gRalt, const double *probL,
                       const double *probR, int rlength, int glengthL, int glengthR, int lbandL, int ubandL,
                       int lbandR, int ubandR, int leftoffset, int rightoffset, int cdna_direction, int finalp,
                       int halfp, int *out, double *prob_out) {
  int *ldi = (int *)calloc((size_t)glengthL + 1, sizeof(int));
  int *rdi = (int *)calloc((size_t)glengthR + 1, sizeof(int));
  for (int c = 0; c < glengthL - 1; c++) ldi[c] = left_site(gL[c], gLalt[c], gL[c + 1], gLalt[c + 1]);
  /* the right fill reads its segment backwards: column c is gR[c-1], so the
     base before the exon at column c is gR[c] and the one before it gR[c+1] */
  for (int c = 0; c < glengthR - 1; c++) rdi[c] = right_site(gR[c + 1], gRalt[c + 1], gR[c], gRalt[c]);

  const int neg = width == 8 ? -128 : -32768;
  const int D = rightoffset - leftoffset;
  int best = neg, brL = -1, brR = -1, bcL = -1, bcR = -1;
  double bestprob = 0.0;
  int dbest = neg, drL = -1, drR = -1, dcL = -1, dcR = -1;
  double dprob = 0.0;
#define LUS(c, r) LU[(size_t)(c) * (rlength + 1) + (r)]
#define LLS(r, c) LL[(size_t)(r) * (glengthL + 1) + (c)]
#define RUS(c, r) RU[(size_t)(c) * (rlength + 1) + (r)]
#define RLS(r, c) RL[(size_t)(r) * (glengthR + 1) + (c)]
  /* one candidate: replaces the best on a higher score, or on an equal score
     with a higher probability sum (1976-1996) */
#define CANDIDATE(sL, sR)                                                                   \
  do {                                                                                      \
    const int v_ = (sL) + intron_reward(cdna_direction, finalp, ldi[cL] & rdi[cR]) + (sR); \
    const double p_ = probL[cL] + probR[cR];                                                \
    if (v_ > best || (v_ == best && p_ > bestprob)) {                                       \
      best = v_;                                                                            \
      brL = rL;                                                                             \
      brR = rR;                                                                             \
      bcL = cL;                                                                             \
      bcR = cR;                                                                             \
      bestprob = p_;                                                                        \
    }                                                                                       \
  } while (0)

  for (int rL = 1, rR = rlength - 1; rL < rlength; rL++, rR--) {
    const int cloL = imax(1, rL - lbandL), chighL = imin(rL + ubandL, glengthL - 1);
    const int cloR = imax(1, rR - lbandR), chighR = imin(rR + ubandR, glengthR - 1);
    int cL, cR;
    /* both ends on their diagonals; the only candidates of the dinucleotide
       track (1958-2012) */
    cL = rL;
    cR = rR;
    CANDIDATE(LUS(cL, rL), RUS(cR, rR));
    if (intron_reward(cdna_direction, finalp, ldi[cL] & rdi[cR]) > 0 && probL[cL] + probR[cR] > dprob) {
      dbest = LUS(cL, rL) + intron_reward(cdna_direction, finalp, ldi[cL] & rdi[cR]) + RUS(cR, rR);
      dcL = cL;
      dcR = cR;
      drL = rL;
      drR = rR;
      dprob = probL[cL] + probR[cR];
    }
    /* left end on its diagonal, right end off it: lower triangle up to the
       diagonal, then the upper one past it (2014-2111) */
    cL = rL;
    for (cR = cloR; cR < rR && cR < D - cL; cR++) CANDIDATE(LUS(cL, rL), RLS(rR, cR));
    for (cR++; cR < chighR && cR < D - cL; cR++) CANDIDATE(LUS(cL, rL), RUS(cR, rR));
    /* right end on its diagonal, left end off it (2113-2206) */
    cR = rR;
    for (cL = cloL; cL < rL && cL < D - cR; cL++) CANDIDATE(LLS(rL, cL), RUS(cR, rR));
    for (cL++; cL < chighL && cL < D - cR; cL++) CANDIDATE(LUS(cL, rL), RUS(cR, rR));
  }
#undef CANDIDATE
#undef LUS
#undef LLS
#undef RUS
#undef RLS

  /* the canonical-dinucleotide candidate takes over unless the best already
     has a probability sum above 2 * PROB_CEILING, none was found, or it
     scores below 0 or more than 9 under the best (2210-2237) */
  if (!(bestprob > 2 * 0.85) && dprob != 0.0 && !(dbest < 0 || dbest < best - 9)) {
    best = dbest;
    brL = drL;
    brR = drR;
    bcL = dcL;
    bcR = dcR;
  }
  int final = best;
  if (best >= 0 && halfp) final = best - intron_reward(cdna_direction, finalp, ldi[bcL] & rdi[bcR]) / 2;
  out[0] = final;
  out[1] = brL;
  out[2] = brR;
  out[3] = bcL;
  out[4] = bcR;
  /* get_splicesite_probs at the chosen sites (2343-2348, 332-401); halfp
     takes half the chosen sites' reward off a non-negative score (2248-2255) */
  prob_out[0] = final >= 0 ? probL[bcL] : 0.0;
  prob_out[1] = final >= 0 ? probR[bcR] : 0.0;
  free(ldi);
  free(rdi);
}

/* ---------------------------------------------------------- genome segment */

static void revcomp_inplace(char *s, unsigned int n) {
  for (unsigned int i = 0, j = n ? n - 1 : 0; i < n / 2; i++, j--) {
    char t = (char)complement_char(s[i]);
    s[i] = (char)complement_char(s[j]);
    s[j] = t;
  }
  if (n % 2) s[n / 2] = (char)complement_char(s[n / 2]);
}

void orc_segment_right(char *out, const orc_genome_t *genome, unsigned int left,
                       unsigned int length, unsigned int chrhigh, int revcomp) {
  unsigned int oob = 0;
  if (length == 0) { out[0] = 0; return; }
  if (left >= chrhigh) {
    memset(out, '*', length);
    out[length] = 0;
    return;
  }
  if (left + length >= chrhigh) oob = left + length - chrhigh;
  for (unsigned int i = 0; i < length - oob; i++) out[i] = (char)genome_char(genome, left + i);
  for (unsigned int i = length - oob; i < length; i++) out[i] = '*';
  out[length] = 0;
  if (revcomp) revcomp_inplace(out, length);
}

void orc_segment_left(char *out, const orc_genome_t *genome, unsigned int right,
                      unsigned int length, unsigned int chroffset, int revcomp) {
  unsigned int oob = 0;
  if (length == 0) { out[0] = 0; return; }
  if (right < chroffset) {
    memset(out, '*', length);
    out[length] = 0;
    return;
  }
  if (right < chroffset + length) oob = chroffset + length - right;
  for (unsigned int i = 0; i < oob; i++) out[i] = '*';
  for (unsigned int i = oob; i < length; i++) out[i] = (char)genome_char(genome, right - length + i);
  out[length] = 0;
  if (revcomp) revcomp_inplace(out, length);
}

/* ------------------------------------------------------------- cell counts */

long long orc_cells_full(int rl, int gl, int lband, int uband) {
  long long n = 0;
  for (int c = 1; c <= gl; c++) {
    int lo = imax(1, c - uband), hi = imin(rl, c + lband);
    if (hi >= lo) n += hi - lo + 1;
  }
  return n;
}

long long orc_cells_upper(int rl, int gl, int uband) {
  long long n = 0;
  for (int c = 1; c <= gl; c++) {
    int lo = imax(1, c - uband), hi = imin(c, rl);
    if (hi >= lo) n += hi - lo + 1;
  }
  return n;
}

long long orc_cells_lower(int rl, int gl, int lband) {
  long long n = 0;
  for (int r = 1; r <= rl; r++) {
    int lo = imax(1, r - lband), hi = imin(r, gl);
    if (hi >= lo) n += hi - lo + 1;
  }
  return n;
}

/* ------------------------------------------------- batch timing (baseline) */

/* Problem record layout of the batched interface (include/gmapdp.h). */
typedef struct {
  long long qoff, goff, galtoff;
  int rl, gl;
  short lband, uband;
  signed char kind, width, mismatchtype, jump_late, open, extend, keep_scores, pad;
  int pad2;
} orc_problem_t;

long long orc_run_batch(const void *problems, int n, const char *qbuf, const char *gbuf, int do_traceback) {
  const orc_problem_t *pr = (const orc_problem_t *)problems;
  long long check = 0;
  for (int i = 0; i < n; i++) {
    const orc_problem_t *p = &pr[i];
    int rl = p->rl, gl = p->gl;
    size_t nc = (size_t)(rl + 1) * (gl + 1);
    short *H = malloc(nc * sizeof(short));
    unsigned char *D = calloc(nc, 1);
    const char *q = qbuf + p->qoff, *g = gbuf + p->goff, *ga = gbuf + p->galtoff;
    orc_fill_full(p->width, q, g, ga, rl, gl, p->mismatchtype, p->open, p->extend, p->lband, p->uband,
                  p->jump_late, 0, H, D);
    check += H[(size_t)gl * (rl + 1) + rl];
    if (do_traceback) {
      int counters[5] = {0, 0, 0, 0, 0};
      int *pairs = malloc(sizeof(int) * ORC_PAIR_FIELDS * (size_t)(rl + gl + 4));
      orc_traceback(ORC_KIND_FULL, D, rl, gl, rl, gl, q, q, g, ga, 0, 0, NULL, 0, 0, 0, 1, 0, 0, counters, pairs,
                    rl + gl + 4);
      check += counters[0];
      free(pairs);
    }
    free(D);
    free(H);
  }
  return check;
}

/* ------------------------------------------------------- scalar standard */

/* Dynprog_standard (dynprog.c:1267-1785), the scalar int32 fill of the
   reference's builds without SIMD, for the full band with the upperp /
   lowerp switches and a saturation floor.  Stated cell by cell:
     band      rows max(1, c - uband) .. min(rl, c + lband) of column c;
     row 0     H(0, c) = open + c*extend for c <= uband (nogap/Egap HORIZ),
     column 0  H(r, 0) = open + r*extend for r <= lband (nogap/Fgap VERT),
               H(0, 0) = 0 (1316-1341);
     E         from the clamped H and the E of (r, c-1); a left neighbour
               outside the band reads its row's initial value (1357-1369,
               1517-1523);
     F         from the unclamped H and the F of (r-1, c); the top cell of a
               column reads F = -inf and H = H(0, c) while the band touches
               row 0, -inf below, and -inf - open + 1 in column 1
               (1409-1448, 1485-1491);
     diagonal  the clamped H(r-1, c-1), except the top cell of a column,
               which reads H(0, c-1) while the band touches row 0 and the
               unclamped top cell of the previous column below (1411-1448,
               1573-1579);
     H         diagonal + score, then E and F in that order may replace it
               (1527-1570);
   scores are stored as max(H, saturation) (1574-1575); the early variant
   (1587-1785) differs only in its strict comparisons.  Out-of-band cells
   hold 0 and DIAG (the matrices are cleared per fill, dynprog.c:206-224,
   487-500).  H (int) and dirs are
   dense [c][r] with stride rl + 1. */
void orc_fill_std(const char *rseq, const char *gseq, const char *gseq_alt, int rl, int gl, int mt, int open,
                  int extend, int lband, int uband, int late, int revp, int saturation, int upperp, int lowerp,
                  int *H, unsigned char *dirs) {
  const int NEG = -32768;
  const size_t ld = (size_t)rl + 1;
  int *Hu = (int *)calloc((size_t)(gl + 1) * ld, sizeof(int)); /* unclamped H of computed cells */
  int *E = (int *)calloc((size_t)(gl + 1) * ld, sizeof(int));
  int *F = (int *)calloc((size_t)(gl + 1) * ld, sizeof(int));
  memset(H, 0, (size_t)(gl + 1) * ld * sizeof(int));
  memset(dirs, 0, (size_t)(gl + 1) * ld);
#define AT(M, cc, rr) (M)[(size_t)(cc) * ld + (rr)]
  for (int c = 1; c <= uband && c <= gl; c++) {
    AT(H, c, 0) = open + c * extend;
    AT(dirs, c, 0) = DIR_H | DIR_E;
  }
  for (int r = 1; r <= lband && r <= rl; r++) {
    AT(H, 0, r) = open + r * extend;
    AT(dirs, 0, r) = DIR_V | DIR_F;
  }
  for (int c = 1; c <= gl; c++) {
    const int rlo = imax(1, c - uband), rhi = imin(rl, c + lband);
    const int g = seq_at(gseq, revp, c), ga = seq_at(gseq_alt, revp, c);
    for (int r = rlo; r <= rhi; r++) {
      const int q = seq_at(rseq, revp, r);
      /* left neighbour (r, c-1) */
      int hl, el;
      if (c - 1 >= 1 && r <= imin(rl, c - 1 + lband)) {
        hl = AT(H, c - 1, r);
        el = AT(E, c - 1, r);
      } else {
        hl = r <= lband ? open + r * extend : NEG;
        el = NEG;
      }
      /* upper neighbour (r-1, c) and the diagonal */
      int hu, fu, dg;
      if (r > rlo) {
        hu = AT(Hu, c, r - 1);
        fu = AT(F, c, r - 1);
        dg = (c - 1 >= 1) ? AT(H, c - 1, r - 1) : ((r - 1 <= lband) ? open + (r - 1) * extend : NEG);
      } else {
        fu = NEG;
        if (c == 1) {
          hu = NEG - open + 1;
          dg = 0;
        } else if (c <= uband) {
          hu = open + c * extend;
          dg = open + (c - 1) * extend;
        } else if (rlo == 1) {
          hu = NEG;
          dg = open + (c - 1) * extend;
        } else {
          hu = NEG;
          dg = AT(Hu, c - 1, r - 1);
        }
      }
      unsigned char d = 0;
      int e = hl + open, f = hu + open;
      if (upperp && wins(el, e, late)) {
        e = el + extend;
        d |= DIR_E;
      } else {
        e = e + extend;
      }
      if (lowerp && wins(fu, f, late)) {
        f = fu + extend;
        d |= DIR_F;
      } else {
        f = f + extend;
      }
      int sc = pd[mt][q & 127][g & 127];
      if (pd[mt][q & 127][ga & 127] > sc) sc = pd[mt][q & 127][ga & 127];
      int h = dg + sc, nd = 0;
      if (upperp && wins(e, h, late)) {
        h = e;
        nd = DIR_H;
      }
      if (lowerp && wins(f, h, late)) {
        h = f;
        nd = DIR_V;
      }
      AT(Hu, c, r) = h;
      AT(E, c, r) = e;
      AT(F, c, r) = f;
      AT(H, c, r) = h < saturation ? saturation : h;
      AT(dirs, c, r) = (unsigned char)(d | nd);
    }
  }
#undef AT
  free(Hu);
  free(E);
  free(F);
}
