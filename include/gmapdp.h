/*
 * gmapdp.h -- C ABI of the MI355X (gfx950) banded affine-gap DP core.
 *
 * This is the batched interface: many independent DP problems are described
 * in one array, their sequences sit in two byte buffers in device memory, and
 * each call launches the kernels the whole array needs.  The per-call
 * interface with the reference's own signatures is in gmapdp_dynprog.h.
 *
 * Semantics follow the reference's AVX2 build of GMAP's DP core
 * (src/dynprog_simd.c, src/dynprog.c, src/dynprog_end.c, src/pairpool.c);
 * each entry point cites the reference function it stands in for.
 *
 * Conventions
 *  - All "d_" pointers are device pointers on the current HIP device; the
 *    caller allocates them (hipMalloc, or any framework allocator).
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *  - Functions return 0 on success and a negative GDP_E* code otherwise;
 *    gdp_last_error() returns a message for the last failure on the thread.
 *  - Sequences are stored in the order the fill consumes them: row r of a
 *    problem reads query byte qbuf[qoff + r - 1] and column c reads genome
 *    byte gbuf[goff + c - 1] (and gbuf[galtoff + c - 1] for the alternate
 *    allele).  A caller holding the reference's reversed pointers
 *    (revp = true, s[1 - r]) stores the reversed sequence.
 */
#ifndef GMAPDP_H
#define GMAPDP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDP_OK 0
#define GDP_EINVAL (-1)
#define GDP_ETOOBIG (-2)
#define GDP_EHIP (-3)
#define GDP_ENOINIT (-4)
#define GDP_ENOMEM (-5)

#define GDP_KIND_FULL 0  /* Dynprog_simd_8 / Dynprog_simd_16 */
#define GDP_KIND_UPPER 1 /* Dynprog_simd_8_upper / Dynprog_simd_16_upper */
#define GDP_KIND_LOWER 2 /* Dynprog_simd_8_lower / Dynprog_simd_16_lower */

#define GDP_HIGHQ 0 /* Mismatchtype_T, dynprog.h:53 */
#define GDP_MEDQ 1
#define GDP_LOWQ 2
#define GDP_ENDQ 3

/* Largest band span a problem may use (lband + uband, or the one-sided band). */
#define GDP_MAX_SPAN 960
#define GDP_MAX_LEN 8000

/* One DP problem (48 bytes, pad fields zero).  width 8 or 16 selects the reference's int8 or
   int16 kernel; open/extend are the (negative) gap penalties; jump_late is
   the reference's jump_late_p (ties prefer the gap); keep_scores asks the
   fill to keep every computed score (needed by gdp_best_endpoint). */
typedef struct gdp_problem {
  int64_t qoff;
  int64_t goff;
  int64_t galtoff;
  int32_t rl;
  int32_t gl;
  int16_t lband;
  int16_t uband;
  int8_t kind;
  int8_t width;
  int8_t mismatchtype;
  int8_t jump_late;
  int8_t open;
  int8_t extend;
  int8_t keep_scores;
  int8_t pad;
  int32_t pad2;
} gdp_problem;

/* One traceback request: walk problem `fill` from cell (r, c) (query row,
   genome column) back to row or column 0, as Dynprog_traceback_{8,16}
   (dynprog_simd.c:9153, 9552), _upper (9318, 9715) or _lower (9438, 9835)
   do for the problem's kind.  Offsets, chromosome bounds, strand and
   genestrand have the reference's meaning. */
typedef struct gdp_tb_request {
  int32_t fill;
  int32_t r;
  int32_t c;
  int32_t queryoffset;
  int32_t genomeoffset;
  uint32_t chroffset;
  uint32_t chrhigh;
  int8_t watsonp;
  int8_t revp;
  int8_t genestrand;
  int8_t flags; /* GDP_TB_OPS: emit pair-list operations instead of pairs */
  int64_t pair_off;
  int32_t max_pairs;
  int32_t pad2;
} gdp_tb_request;

/* With GDP_TB_OPS the walk records, instead of pairs, the calls the
   reference's traceback makes on its pair pool, in order: cdna = 'D' for
   Pairpool_push of a diagonal pair at (querypos, genomepos) = (r, c) with
   comp set; 'H' for Pairpool_add_genomeskip(r, c, dist) and 'V' for
   Pairpool_add_queryskip(r, c, dist), dist in genomejump.  A caller that owns
   a host pair pool and genome replays them (gmapdp_dynprog.h). */
#define GDP_TB_OPS 1

/* With GDP_TB_NOGAPS the walk is traceback_nogaps (dynprog_end.c:648-722):
   diagonal pairs only, from (r, c) while r > 0 and c > 0, no directions
   read and no closing skip.  Its start cell is the closed form of
   find_best_endpoint_to_queryend_nogaps (dynprog_end.c:576-586):
   r = c = min(rl, gl). */
#define GDP_TB_NOGAPS 2

/* One traceback output record, in the order the reference pushes them onto
   its pair list (from the end cell back to the start).  gapp = 1 marks a
   gap holder (pairpool.c:374-452) whose genomejump is the skipped length. */
typedef struct gdp_pair {
  int32_t querypos;
  int32_t genomepos;
  int8_t cdna;
  int8_t comp;
  int8_t genome;
  int8_t genomealt;
  int16_t gapp;
  int16_t genomejump;
} gdp_pair;

/* Builds the score tables for a GMAP mode (mode.h: 0 = STANDARD ...), as
   Dynprog_init does (dynprog.c:1006-1196).  Must precede every other call. */
int gdp_init(int mode);

/* As gdp_init, but with tables the host program already built (its own
   Dynprog_init): pd = pairdistance_array as 4*128*128 int16 [mt][query][genome],
   nt = nt_to_int_array (128), cons = consistent_array as 3*128*128 bytes. */
int gdp_init_tables(const int16_t *pd, const int32_t *nt, const uint8_t *cons);

/* Builds the host-side tables only (no device needed); read them back with
   gdp_pairdistance / gdp_nt_to_int / gdp_consistent / gdp_use8p_size. */
void gdp_host_tables(int mode);
const int16_t *gdp_pairdistance(int mismatchtype);
const int32_t *gdp_nt_to_int(void);
int32_t gdp_use8p_size(int mismatchtype);
int32_t gdp_consistent(int genestrand, int c, int g);

/* Dynprog_compute_bands (dynprog.c:1245-1261). */
void gdp_compute_bands(int32_t *lband, int32_t *uband, int32_t rlength, int32_t glength, int32_t extraband,
                       int32_t widebandp);

/* Message describing the last failure on this thread. */
const char *gdp_last_error(void);

/* Size in bytes of one device problem descriptor. */
int32_t gdp_descriptor_bytes(void);

/* Bytes gdp_plan writes for n problems: n descriptors followed by the wave
   lists of the paired sweep. */
int64_t gdp_plan_bytes(int32_t n);

/* Host-side planning: validates the problems and writes n device
   descriptors (gdp_descriptor_bytes() each) into descs_out, followed by the
   wave lists (gdp_plan_bytes(n) in all), assigning each problem its slice of
   the direction and score buffers.  Full-band int16 problems without kept
   scores are grouped into waves of problems that share band and tie rule
   and are swept two per lane; gdp_fill must then be given the whole array
   with the same n.  Returns the 32-bit direction words and int16 score
   slots the batch needs.  No counterpart in the reference (the batched
   form the drop-in functions are built on). */
int gdp_plan(const gdp_problem *probs, int32_t n, void *descs_out, int64_t *dir_words, int64_t *score_slots);

/* As gdp_plan; flags GDP_PLAN_SINGLE gives every problem a sweep of its own,
   so that any single descriptor can be filled by itself (gdp_fill on
   descs + i*gdp_descriptor_bytes() with n = 1): the one-launch-per-problem
   form. */
#define GDP_PLAN_SINGLE 1
int gdp_plan_ex(const gdp_problem *probs, int32_t n, int32_t flags, void *descs_out, int64_t *dir_words,
                int64_t *score_slots);

/* In-band cell count of a problem (the throughput work unit). */
int64_t gdp_cells(const gdp_problem *prob);

/* Fills every problem of a plan: the single-sweep problems in one launch,
   then the paired waves in persistent launches that take waves off a queue
   kept in the descriptor array (so two fills of the same descriptors must
   not overlap in time).  d_final[i] receives the score of cell (rl, gl) for
   full-band problems (the value Dynprog_single_gap reads,
   dynprog_single.c:603), or INT32_MIN when that cell is outside the band. */
int gdp_fill(const void *d_descs, int32_t n, const char *d_qbuf, const char *d_gbuf, uint32_t *d_dirs,
             int16_t *d_scores, int32_t *d_final, void *stream);

/* The kernels a plan needs, read from the host copy of its descriptors:
   bit 0 single sweeps, bits 1..3 the paired wave classes.  gdp_fill_ex
   launches only those (gdp_fill launches all, and empty ones return at
   once). */
#define GDP_FILL_ALL 0xFu
uint32_t gdp_plan_kernels(const void *descs, int32_t n);
int gdp_fill_ex(const void *d_descs, int32_t n, uint32_t kernels, const char *d_qbuf, const char *d_gbuf,
                uint32_t *d_dirs, int16_t *d_scores, int32_t *d_final, void *stream);

/* Best end cell of end problems: for each of m pairs (upper, lower) of
   one-sided fills with keep_scores set, find_best_endpoint_{8,16}
   (dynprog_end.c:143-290) when queryend_indels = 0, or
   find_best_endpoint_to_queryend_indels_{8,16} (358-508) otherwise.
   d_out[3*k] = {finalscore, bestr, bestc}. */
int gdp_best_endpoint(const void *d_descs, const int32_t *d_ul, int32_t m, int32_t queryend_indels,
                      const int16_t *d_scores, int32_t *d_out, void *stream);

/* Tracebacks, one GPU thread per request.  d_qbuf, d_gbuf and d_dirs are
   read in aligned 16-byte windows: they must be 16-byte aligned and readable
   for 16 bytes past their last used byte (the dir_words of gdp_plan include
   that padding).  d_genome (may be NULL) holds the
   chromosome as ASCII bases indexed by absolute position, used for genome
   skips exactly as get_genomic_nt (pairpool.c:1016-1063) uses
   Genome_get_char; positions outside [chroffset, chrhigh) give '*'.
   d_counts[6*k] = {npairs, traceback_score, nmatches, nmismatches, nopens,
   nindels}. */
int gdp_traceback(const void *d_descs, const gdp_tb_request *d_reqs, int32_t nreq, const char *d_qbuf,
                  const char *d_gbuf, const uint32_t *d_dirs, const char *d_genome, uint32_t genome_len,
                  gdp_pair *d_pairs, int32_t *d_counts, void *stream);

/* Writes one problem's directions (and scores if kept) in the reference's
   dense layout: [c][r] with stride rl+1 for full/upper, [r][c] with stride
   gl+1 for lower; direction bytes use bit0 nogap HORIZ, bit1 nogap VERT,
   bit2 Egap, bit3 Fgap.  Cells the fill does not compute are 0. */
int gdp_export_dense(const void *d_descs, int32_t idx, const uint32_t *d_dirs, const int16_t *d_scores,
                     uint8_t *d_dense_dirs, int16_t *d_dense_scores, void *stream);

/* Writes one problem in the reference's own workspace layout and encoding
   (aligned_score{8,16}_alloc / aligned_directions{8,16}_alloc,
   dynprog_simd.c:2158-2282): element (outer, inner) at outer*stride + inner,
   outer = genome column c and inner = query row r for full/upper, outer = r
   and inner = c for lower; stride is the reference's rlength_ceil /
   glength_ceil.  Elements are int8 (width 8) or int16 (width 16): scores,
   nogap in {0 DIAG, -1 HORIZ, -2 VERT}, Egap in {0, -1}, Fgap in {0, -2}
   (d_fgap may be NULL for one-sided problems).  Cells not computed are 0. */
int gdp_export_reference(const void *d_descs, int32_t idx, const uint32_t *d_dirs, const int16_t *d_scores,
                         int32_t stride, void *d_score, void *d_nogap, void *d_egap, void *d_fgap, void *stream);

#ifdef __cplusplus
}
#endif
#endif
