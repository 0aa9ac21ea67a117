/*
 * ref_shim.c -- thin C harness over the reference's own DP objects.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
 * reference sources (dynprog.c, dynprog_simd.c, pairpool.c, list.c, mem.c,
 * except.c, assert.c) read in place from the reference tree, into
 * oracle/_ref/libgmapref.so.  It exposes the same plain-array interface as
 * gdp_oracle.h so tests can compare the CPU restatement with the reference
 * cell for cell.  refh_traceback passes no genome (callers pass chrhigh = 0,
 * which makes every genome-skip character '*', pairpool.c:1037-1040);
 * refh_traceback_genome passes a Genome_T over a genomecomp image, read by
 * the reference's own Genome_get_char (genome.c, compiled in place too).
 */
#include <stdlib.h>
#include <string.h>

#include "dynprog.h"
#include "dynprog_simd.h"
#include "genome.h"
#include "list.h"
#include "pairdef.h"
#include "pairpool.h"

static Pairpool_T pool = NULL;

/* Direction tables of the most recent fill of each kind. */
static void **last_nogap[3], **last_Egap[3], **last_Fgap[3];

void refh_init(int mode) {
  Dynprog_init((Mode_T)mode);
  if (pool == NULL) pool = Pairpool_new();
}

void *refh_new(int maxrl, int maxgl, int doublep) {
  int extra = maxgl > maxrl ? maxgl - maxrl : 0;
  return (void *)Dynprog_new(/*maxlookback*/ maxrl, /*extraquerygap*/ extra, /*maxpeelback*/ 0,
                             /*extramaterial_end*/ 0, /*extramaterial_paired*/ 0, doublep ? true : false);
}

void refh_free(void *ws) {
  Dynprog_T t = (Dynprog_T)ws;
  Dynprog_free(&t);
}

int refh_max_rlength(void *ws) { return ((Dynprog_T)ws)->max_rlength; }
int refh_max_glength(void *ws) { return ((Dynprog_T)ws)->max_glength; }
int refh_use8p(int t) { return use8p_size[t]; }

void refh_pairdistance(int t, short *out) {
  for (int i = 0; i < 128; i++)
    for (int j = 0; j < 128; j++) out[i * 128 + j] = pairdistance_array[t][i][j];
}

int refh_consistent(int genestrand, int c, int g) { return consistent_array[genestrand][c][g] ? 1 : 0; }

/* Fills the part of a workspace space a fill of this shape will address with
   one byte value, so that cells the fill never writes hold a known value.
   Running a problem twice with different values shows which outputs depend
   on workspace contents left over from earlier calls. */
static void prefill(void **spaces, int nspaces, int nouter, int ninner_ceil, int elemsize, int garbage) {
  for (int k = 0; k < nspaces; k++)
    memset(spaces[k], garbage, (size_t)(nouter + 1) * ninner_ceil * elemsize);
}

static unsigned char enc(int nogap, int egap, int fgap) {
  unsigned char d = 0;
  if (nogap == HORIZ) d |= 1;
  else if (nogap != DIAG) d |= 2;
  if (egap != DIAG) d |= 4;
  if (fgap != DIAG) d |= 8;
  return d;
}

/* Copies score/direction tables into dense arrays: ncols outer x nrows inner. */
#define COPY_OUT(MAT, NO, EG, FG, NOUT, NIN)                                        \
  do {                                                                              \
    for (int o = 0; o <= (NOUT); o++)                                               \
      for (int i = 0; i <= (NIN); i++) {                                            \
        size_t k = (size_t)o * ((NIN) + 1) + i;                                     \
        H[k] = (short)(MAT)[o][i];                                                  \
        dirs[k] = enc((NO)[o][i], (EG)[o][i], (FG) ? (FG)[o][i] : DIAG);            \
      }                                                                             \
  } while (0)

void refh_fill_full(void *ws, int width, char *rseq, char *gseq, char *galt, int rl, int gl,
                    int mt, int open, int extend, int lband, int uband, int late, int revp,
                    short *H, unsigned char *dirs, int garbage) {
  Dynprog_T t = (Dynprog_T)ws;
  void *sp[4] = {t->aligned.one.matrix_space, t->aligned.one.directions_space_0,
                 t->aligned.one.directions_space_1, t->aligned.one.directions_space_2};
  int S = width == 8 ? SIMD_NCHARS : SIMD_NSHORTS;
  prefill(sp, 4, gl, ((rl + S) / S) * S, width / 8, garbage);
  if (width == 8) {
    Direction8_T **no, **eg, **fg;
    Score8_T **m = Dynprog_simd_8(&no, &eg, &fg, (Dynprog_T)ws, rseq, gseq, galt, rl, gl, (Mismatchtype_T)mt,
                                  open, extend, lband, uband, late ? true : false, revp ? true : false);
    COPY_OUT(m, no, eg, fg, gl, rl);
    last_nogap[0] = (void **)no; last_Egap[0] = (void **)eg; last_Fgap[0] = (void **)fg;
  } else {
    Direction16_T **no, **eg, **fg;
    Score16_T **m = Dynprog_simd_16(&no, &eg, &fg, (Dynprog_T)ws, rseq, gseq, galt, rl, gl, (Mismatchtype_T)mt,
                                    open, extend, lband, uband, late ? true : false, revp ? true : false);
    COPY_OUT(m, no, eg, fg, gl, rl);
    last_nogap[0] = (void **)no; last_Egap[0] = (void **)eg; last_Fgap[0] = (void **)fg;
  }
}

void refh_fill_upper(void *ws, int width, char *rseq, char *gseq, char *galt, int rl, int gl,
                     int mt, int open, int extend, int uband, int late, int revp,
                     short *H, unsigned char *dirs, int garbage) {
  Dynprog_T t = (Dynprog_T)ws;
  void *sp[3] = {t->aligned.two.upper_matrix_space, t->aligned.two.upper_directions_space_0,
                 t->aligned.two.upper_directions_space_1};
  int S = width == 8 ? SIMD_NCHARS : SIMD_NSHORTS;
  prefill(sp, 3, gl, ((rl + S) / S) * S, width / 8, garbage);
  if (width == 8) {
    Direction8_T **no, **eg, **fg = NULL;
    Score8_T **m = Dynprog_simd_8_upper(&no, &eg, (Dynprog_T)ws, rseq, gseq, galt, rl, gl, (Mismatchtype_T)mt,
                                        open, extend, uband, late ? true : false, revp ? true : false);
    COPY_OUT(m, no, eg, fg, gl, rl);
    last_nogap[1] = (void **)no; last_Egap[1] = (void **)eg;
  } else {
    Direction16_T **no, **eg, **fg = NULL;
    Score16_T **m = Dynprog_simd_16_upper(&no, &eg, (Dynprog_T)ws, rseq, gseq, galt, rl, gl, (Mismatchtype_T)mt,
                                          open, extend, uband, late ? true : false, revp ? true : false);
    COPY_OUT(m, no, eg, fg, gl, rl);
    last_nogap[1] = (void **)no; last_Egap[1] = (void **)eg;
  }
}

void refh_fill_lower(void *ws, int width, char *rseq, char *gseq, char *galt, int rl, int gl,
                     int mt, int open, int extend, int lband, int late, int revp,
                     short *H, unsigned char *dirs, int garbage) {
  Dynprog_T t = (Dynprog_T)ws;
  void *sp[3] = {t->aligned.two.lower_matrix_space, t->aligned.two.lower_directions_space_0,
                 t->aligned.two.lower_directions_space_1};
  int S = width == 8 ? SIMD_NCHARS : SIMD_NSHORTS;
  prefill(sp, 3, rl, ((gl + S) / S) * S, width / 8, garbage);
  if (width == 8) {
    Direction8_T **no, **eg, **fg = NULL;
    Score8_T **m = Dynprog_simd_8_lower(&no, &eg, (Dynprog_T)ws, rseq, gseq, galt, rl, gl, (Mismatchtype_T)mt,
                                        open, extend, lband, late ? true : false, revp ? true : false);
    COPY_OUT(m, no, eg, fg, rl, gl);
    last_nogap[2] = (void **)no; last_Egap[2] = (void **)eg;
  } else {
    Direction16_T **no, **eg, **fg = NULL;
    Score16_T **m = Dynprog_simd_16_lower(&no, &eg, (Dynprog_T)ws, rseq, gseq, galt, rl, gl, (Mismatchtype_T)mt,
                                          open, extend, lband, late ? true : false, revp ? true : false);
    COPY_OUT(m, no, eg, fg, rl, gl);
    last_nogap[2] = (void **)no; last_Egap[2] = (void **)eg;
  }
}

/* Runs the reference traceback over the tables of the last fill of `kind`
   (0 full, 1 upper, 2 lower) and flattens the pair list into push order:
   8 ints per pair {querypos, genomepos, cdna, comp, genome, genomealt, gapp,
   genomejump}, gapp 2 for a known-site gap holder (pair->knowngapp).  counters[5] receive {score, nmatches, nmismatches, nopens,
   nindels}.  Returns the number of pairs. */
static int traceback_with(Genome_T genome, int kind, int width, int r, int c, char *rseq, char *rsequc, char *gseq,
                          char *galt, int queryoffset, int genomeoffset, int revp, unsigned int chroffset,
                          unsigned int chrhigh, int watsonp, int genestrand, int dynprogindex, int *counters, int *out,
                          int maxpairs) {
  List_T pairs = NULL, p;
  int sc = 0, nm = 0, nmm = 0, no = 0, ni = 0;
  Pairpool_reset(pool);
  if (kind == 0) {
    if (width == 8)
      pairs = Dynprog_traceback_8(NULL, &sc, &nm, &nmm, &no, &ni, (Direction8_T **)last_nogap[0],
                                  (Direction8_T **)last_Egap[0], (Direction8_T **)last_Fgap[0], r, c, rseq, rsequc,
                                  gseq, galt, queryoffset, genomeoffset, genome, genome, pool, revp ? true : false,
                                  chroffset, chrhigh, watsonp ? true : false, genestrand, dynprogindex);
    else
      pairs = Dynprog_traceback_16(NULL, &sc, &nm, &nmm, &no, &ni, (Direction16_T **)last_nogap[0],
                                   (Direction16_T **)last_Egap[0], (Direction16_T **)last_Fgap[0], r, c, rseq, rsequc,
                                   gseq, galt, queryoffset, genomeoffset, genome, genome, pool, revp ? true : false,
                                   chroffset, chrhigh, watsonp ? true : false, genestrand, dynprogindex);
  } else if (kind == 1) {
    if (width == 8)
      pairs = Dynprog_traceback_8_upper(NULL, &sc, &nm, &nmm, &no, &ni, (Direction8_T **)last_nogap[1],
                                        (Direction8_T **)last_Egap[1], r, c, rseq, rsequc, gseq, galt, queryoffset,
                                        genomeoffset, genome, genome, pool, revp ? true : false, chroffset, chrhigh,
                                        watsonp ? true : false, genestrand, dynprogindex);
    else
      pairs = Dynprog_traceback_16_upper(NULL, &sc, &nm, &nmm, &no, &ni, (Direction16_T **)last_nogap[1],
                                         (Direction16_T **)last_Egap[1], r, c, rseq, rsequc, gseq, galt, queryoffset,
                                         genomeoffset, genome, genome, pool, revp ? true : false, chroffset, chrhigh,
                                         watsonp ? true : false, genestrand, dynprogindex);
  } else {
    if (width == 8)
      pairs = Dynprog_traceback_8_lower(NULL, &sc, &nm, &nmm, &no, &ni, (Direction8_T **)last_nogap[2],
                                        (Direction8_T **)last_Egap[2], r, c, rseq, rsequc, gseq, galt, queryoffset,
                                        genomeoffset, pool, genestrand, revp ? true : false, dynprogindex);
    else
      pairs = Dynprog_traceback_16_lower(NULL, &sc, &nm, &nmm, &no, &ni, (Direction16_T **)last_nogap[2],
                                         (Direction16_T **)last_Egap[2], r, c, rseq, rsequc, gseq, galt, queryoffset,
                                         genomeoffset, pool, genestrand, revp ? true : false, dynprogindex);
  }
  counters[0] += sc; counters[1] += nm; counters[2] += nmm; counters[3] += no; counters[4] += ni;
  int n = 0;
  for (p = pairs; p != NULL; p = p->rest) n++;
  int k = n - 1; /* list head is the last push */
  for (p = pairs; p != NULL; p = p->rest, k--) {
    if (k >= maxpairs) continue;
    Pair_T pr = (Pair_T)p->first;
    int *o = out + (size_t)k * 8;
    o[0] = pr->querypos; o[1] = (int)pr->genomepos; o[2] = pr->cdna; o[3] = pr->comp;
    o[4] = pr->genome; o[5] = pr->genomealt; o[6] = pr->gapp ? (pr->knowngapp ? 2 : 1) : 0; o[7] = pr->genomejump;
  }
  return n;
}

int refh_traceback(int kind, int width, int r, int c, char *rseq, char *rsequc, char *gseq, char *galt,
                   int queryoffset, int genomeoffset, int revp, unsigned int chroffset,
                   unsigned int chrhigh, int watsonp, int genestrand, int dynprogindex,
                   int *counters, int *out, int maxpairs) {
  return traceback_with(NULL, kind, width, r, c, rseq, rsequc, gseq, galt, queryoffset, genomeoffset, revp,
                        chroffset, chrhigh, watsonp, genestrand, dynprogindex, counters, out, maxpairs);
}

/* As refh_traceback, with genome skips read from a genome: a Genome_T whose
   blocks are the genomecomp image `blocks` of `length` bases (the genome and
   its alternate are the same image), through the reference's own
   Pairpool_add_genomeskip -> get_genomic_nt -> Genome_get_char
   (pairpool.c:1016-1180, genome.c:11003-11017). */
int refh_traceback_genome(const unsigned int *blocks, unsigned int length, int kind, int width, int r, int c,
                          char *rseq, char *rsequc, char *gseq, char *galt, int queryoffset, int genomeoffset,
                          int revp, unsigned int chroffset, unsigned int chrhigh, int watsonp, int genestrand,
                          int dynprogindex, int *counters, int *out, int maxpairs) {
  struct Genome_T g;
  memset(&g, 0, sizeof g);
  g.blocks = (Genomecomp_T *)blocks;
  g.genomelength = length;
  return traceback_with(&g, kind, width, r, c, rseq, rsequc, gseq, galt, queryoffset, genomeoffset, revp, chroffset,
                        chrhigh, watsonp, genestrand, dynprogindex, counters, out, maxpairs);
}

/* The direction tables of the last fill of `kind` (1 upper, 2 lower), for
   the walks of ref_end_shim.c. */
void refh_last_tables(int kind, void ***nogap, void ***egap) {
  *nogap = last_nogap[kind];
  *egap = last_Egap[kind];
}

/* ------------------------------------------------- batch timing (baseline) */

typedef struct {
  long long qoff, goff, galtoff;
  int rl, gl;
  short lband, uband;
  signed char kind, width, mismatchtype, jump_late, open, extend, keep_scores, pad;
  int pad2;
} refh_problem_t;

/* Runs full-band problems the way Dynprog_single_gap does (fill, read the
   corner score, traceback), on a private workspace and pair pool so several
   threads can time disjoint slices.  Genome skips resolve to '*' (chrhigh 0). */
long long refh_run_batch(const void *problems, int n, const char *qbuf, const char *gbuf, int do_traceback,
                         int maxrl, int maxgl) {
  const refh_problem_t *pr = (const refh_problem_t *)problems;
  Dynprog_T ws = (Dynprog_T)refh_new(maxrl, maxgl, 0);
  Pairpool_T pp = Pairpool_new();
  long long check = 0;
  for (int i = 0; i < n; i++) {
    const refh_problem_t *p = &pr[i];
    char *q = (char *)qbuf + p->qoff, *g = (char *)gbuf + p->goff, *ga = (char *)gbuf + p->galtoff;
    int sc = 0, nm = 0, nmm = 0, no = 0, ni = 0;
    Pairpool_reset(pp);
    if (p->width == 8) {
      Direction8_T **dn, **de, **df;
      Score8_T **m = Dynprog_simd_8(&dn, &de, &df, ws, q, g, ga, p->rl, p->gl, (Mismatchtype_T)p->mismatchtype,
                                    p->open, p->extend, p->lband, p->uband, p->jump_late ? true : false, false);
      check += m[p->gl][p->rl];
      if (do_traceback)
        Dynprog_traceback_8(NULL, &sc, &nm, &nmm, &no, &ni, dn, de, df, p->rl, p->gl, q, q, g, ga, 0, 0, NULL, NULL,
                            pp, false, 0, 0, true, 0, 0);
    } else {
      Direction16_T **dn, **de, **df;
      Score16_T **m = Dynprog_simd_16(&dn, &de, &df, ws, q, g, ga, p->rl, p->gl, (Mismatchtype_T)p->mismatchtype,
                                      p->open, p->extend, p->lband, p->uband, p->jump_late ? true : false, false);
      check += m[p->gl][p->rl];
      if (do_traceback)
        Dynprog_traceback_16(NULL, &sc, &nm, &nmm, &no, &ni, dn, de, df, p->rl, p->gl, q, q, g, ga, 0, 0, NULL,
                             NULL, pp, false, 0, 0, true, 0, 0);
    }
    check += sc;
  }
  Pairpool_free(&pp);
  Dynprog_free(&ws);
  return check;
}

/* Pair pools and list flattening for tests that call a Dynprog_* library
   directly (the pair list comes back as the reference's List_T). */
void *refh_pool_new(void) { return (void *)Pairpool_new(); }
void refh_pool_reset(void *pool) { Pairpool_reset((Pairpool_T)pool); }

int refh_flatten(void *list, int *out, int maxpairs) {
  List_T p;
  int n = 0;
  for (p = (List_T)list; p != NULL; p = p->rest) n++;
  int k = n - 1;
  for (p = (List_T)list; p != NULL; p = p->rest, k--) {
    if (k >= maxpairs) continue;
    Pair_T pr = (Pair_T)p->first;
    int *o = out + (size_t)k * 8;
    o[0] = pr->querypos; o[1] = (int)pr->genomepos; o[2] = pr->cdna; o[3] = pr->comp;
    o[4] = pr->genome; o[5] = pr->genomealt; o[6] = pr->gapp ? (pr->knowngapp ? 2 : 1) : 0; o[7] = pr->genomejump;
  }
  return n;
}
