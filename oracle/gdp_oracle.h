/*
 * gdp_oracle.h -- CPU restatement of GMAP's banded affine-gap DP core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * Every function restates the reference algorithm in plain scalar C,
 * following the reference file:line cited beside it (paths relative to the
 * reference's src/ directory).  Output encodings are this project's own:
 *
 *   score matrices : int16, full/upper kernels store [c][r] with stride rl+1,
 *                    the lower kernel stores [r][c] with stride gl+1
 *                    (mirrors the reference's matrix[c][r] / matrix[r][c]).
 *   direction byte : bit0 nogap==HORIZ, bit1 nogap==VERT, bit2 Egap!=DIAG,
 *                    bit3 Fgap!=DIAG (VERT excludes HORIZ in bit0/1).
 *   pair record    : 8 x int32 {querypos, genomepos, cdna, comp, genome,
 *                    genomealt, gapp, genomejump}, in push order
 *                    (the reference pushes from the end cell back to (0,0));
 *                    gapp is 1 for a gap holder and 2 for a known-site gap
 *                    holder (Pairpool_push_gapholder with knownp,
 *                    pairpool.c:374-452).
 */
#ifndef GDP_ORACLE_H
#define GDP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_HIGHQ 0
#define ORC_MEDQ 1
#define ORC_LOWQ 2
#define ORC_ENDQ 3

#define ORC_KIND_FULL 0
#define ORC_KIND_UPPER 1
#define ORC_KIND_LOWER 2

#define ORC_PAIR_FIELDS 8

/* Tables (dynprog.c:1006-1196).  mode follows mode.h (0 = STANDARD). */
void orc_init(int mode);
const short *orc_pairdistance(int mismatchtype); /* 128*128, [query][genome] */
const int *orc_nt_to_int(void);                  /* 128 */
int orc_consistent(int genestrand, int c, int g);
int orc_use8p_size(int mismatchtype);

/* Dynprog_compute_bands (dynprog.c:1245-1261). */
void orc_compute_bands(int *lband, int *uband, int rlength, int glength, int extraband, int widebandp);

/* Full-band fills: Dynprog_simd_8 (dynprog_simd.c:2986-3834) and
   Dynprog_simd_16 (dynprog_simd.c:6561-7344), AVX2 block heights 32/16.
   width is 8 or 16.  H and dirs are dense [c][r], (gl+1)*(rl+1). */
void orc_fill_full(int width, const char *rseq, const char *gseq, const char *gseq_alt,
                   int rl, int gl, int mismatchtype, int open, int extend,
                   int lband, int uband, int jump_late, int revp,
                   short *H, unsigned char *dirs);

/* One-sided fills: Dynprog_simd_{8,16}_upper (dynprog_simd.c:4303-4863,
   7713-8192) store [c][r]; Dynprog_simd_{8,16}_lower (5339-5927, 8585-9092)
   store [r][c] with stride gl+1. */
void orc_fill_upper(int width, const char *rseq, const char *gseq, const char *gseq_alt,
                    int rl, int gl, int mismatchtype, int open, int extend,
                    int uband, int jump_late, int revp, short *H, unsigned char *dirs);
void orc_fill_lower(int width, const char *rseq, const char *gseq, const char *gseq_alt,
                    int rl, int gl, int mismatchtype, int open, int extend,
                    int lband, int jump_late, int revp, short *H, unsigned char *dirs);

/* Dynprog_standard (dynprog.c:1264-1785): the scalar int32 fill of the
   reference's builds without SIMD, full band, with the upperp/lowerp
   switches and a saturation floor.  H (int) and dirs dense [c][r]. */
void orc_fill_std(const char *rseq, const char *gseq, const char *gseq_alt, int rl, int gl, int mismatchtype,
                  int open, int extend, int lband, int uband, int jump_late, int revp, int saturation, int upperp,
                  int lowerp, int *H, unsigned char *dirs);

/* Genome model used for genome-skip characters in the tracebacks
   (pairpool.c:1016-1063 get_genomic_nt over Genome_get_char genome.c:11003).
   codes: 2-bit A=0,C=1,G=2,T=3 one byte per base; nflags: nonzero => 'N'.
   length is the genome's length (positions at or past it read 'N'); the
   arrays hold positions base .. base + count - 1 (count 0: all of them from
   base 0), and any other position reads '?', which no check accepts. */
typedef struct {
  const unsigned char *codes;
  const unsigned char *nflags;
  unsigned int length;
  unsigned int base;
  unsigned int count;
} orc_genome_t;

/* Tracebacks: Dynprog_traceback_{8,16} (dynprog_simd.c:9153-9312, 9552-9711),
   _upper (9318-9435, 9715-9832), _lower (9438-9548, 9835-9945).
   dirs layout as produced by the matching fill.  Returns the number of
   pair records written to pairs (capacity maxpairs records).
   counters = {traceback_score, nmatches, nmismatches, nopens, nindels} (added to). */
int orc_traceback_nogaps(int r, int c, const char *rseq, const char *rsequc, const char *gseq, const char *gseq_alt,
                         int queryoffset, int genomeoffset, int revp, int genestrand, int *counters, int *pairs,
                         int maxpairs);
int orc_traceback(int kind, const unsigned char *dirs, int rl, int gl, int r, int c,
                  const char *rseq, const char *rsequc, const char *gseq, const char *gseq_alt,
                  int queryoffset, int genomeoffset, const orc_genome_t *genome,
                  int revp, unsigned int chroffset, unsigned int chrhigh, int watsonp,
                  int genestrand, int dynprogindex, int *counters, int *pairs, int maxpairs);

/* traceback_local_{8,16}_{upper,lower} (dynprog_end.c:727-1135): kind 1
   (upper, dirs [c][r]) or 2 (lower, dirs [r][c]) walk stopping at genome
   column endc; end_rc[2] = the cell where it stopped.  Returns the number of
   pairs (push order). */
int orc_traceback_local(int kind, const unsigned char *dirs, int rl, int gl, int r, int c, int endc,
                        const char *rseq, const char *rsequc, const char *gseq, const char *gseq_alt,
                        int queryoffset, int genomeoffset, int revp, int genestrand,
                        int *counters, int *pairs, int maxpairs, int *end_rc);

/* find_best_endpoint_{8,16} (dynprog_end.c:143-290) when queryend_indels==0,
   find_best_endpoint_to_queryend_indels_{8,16} (dynprog_end.c:358-508) otherwise.
   Hupper is [c][r] stride rl+1, Hlower is [r][c] stride gl+1. */
void orc_best_endpoint(int queryend_indels, const short *Hupper, const short *Hlower,
                       int width, int rl, int gl, int lband, int uband, int jump_late,
                       int *finalscore, int *bestr, int *bestc);

/* bridge_cdna_gap_{8,16}_ud (dynprog_cdna.c:123-381, 386-644): the best join of
   a left and a right pair of one-sided fills over one genomic gap of gl
   columns.  LU/LL are the left fills (rsequenceL forward), RU/RL the right
   ones (rev_rsequenceR backward); upper [c][r] stride rl+1, lower [r][c]
   stride gl+1, with rlL / rlR rows.  out = {bestscore, bestcL, bestcR,
   bestrL, bestrR}; the cells stay -1 when no candidate reaches the
   initial -128 / -32768 (the reference leaves them unset then). */
void orc_bridge_cdna(int width, const short *LU, const short *LL, const short *RU, const short *RL,
                     int gl, int rlL, int rlR, int lbandL, int ubandL, int lbandR, int ubandR,
                     int open, int extend, int leftoffset, int rightoffset, int jump_late, int *out);

/* bridge_intron_gap_{8,16}_ud (dynprog_genome.c:1387-1488, 2262-2363) on its
   site-level path, bridge_intron_gap_{8,16}_site_level (866-1386,
   1741-2261): the best intron between a left pair of one-sided fills
   (rsequence forward over the left segment) and a right pair (rev_rsequence
   against the right segment read backwards), the ends' rows summing to
   rlength.  LU/LL and RU/RL as orc_bridge_cdna, all with rlength rows; gL,
   gLalt and gR, gRalt are each end's genome in fill order (column c reads
   g[c-1]: the right segment reversed).  probL / probR hold each column's
   splice-site probability (the site-level loops' left_probabilities and
   right_probabilities, 1845-1933); entries up to glength-2 are read, and
   glengthL, glengthR must exceed rlength (the reference reads past its
   arrays otherwise).  out = {finalscore, bestrL, bestrR, bestcL, bestcR}
   (cells -1 when no candidate beats -128 / -32768); prob_out = {left_prob,
   right_prob}, 0.0 when finalscore < 0. */
void orc_bridge_intron(int width, const short *LU, const short *LL, const short *RU, const short *RL,
                       const char *gL, const char *gLalt, const char *gR, const char *gRalt, const double *probL,
                       const double *probR, int rlength, int glengthL, int glengthR, int lbandL, int ubandL,
                       int lbandR, int ubandR, int leftoffset, int rightoffset, int cdna_direction, int finalp,
                       int halfp, int *out, double *prob_out);

/* Genome_get_segment_right / _left (genome.c:11022-11120) over orc_genome_t,
   without SNP alternates (segmentalt == segment). out has length+1 bytes. */
void orc_segment_right(char *out, const orc_genome_t *genome, unsigned int left,
                       unsigned int length, unsigned int chrhigh, int revcomp);
void orc_segment_left(char *out, const orc_genome_t *genome, unsigned int right,
                      unsigned int length, unsigned int chroffset, int revcomp);

/* Count of in-band cells (algorithmic work unit). */
long long orc_cells_full(int rl, int gl, int lband, int uband);
long long orc_cells_upper(int rl, int gl, int uband);
long long orc_cells_lower(int rl, int gl, int lband);

#ifdef __cplusplus
}
#endif
#endif
