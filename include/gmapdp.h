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
#define GDP_KIND_STD 3   /* Dynprog_standard (dynprog.c:1267), width 32 */

/* std_flags of a GDP_KIND_STD problem: Dynprog_standard's upperp / lowerp. */
#define GDP_STD_UPPER 1
#define GDP_STD_LOWER 2

#define GDP_HIGHQ 0 /* Mismatchtype_T, dynprog.h:53 */
#define GDP_MEDQ 1
#define GDP_LOWQ 2
#define GDP_ENDQ 3

/* keep_scores values (gdp_problem) */
#define GDP_KEEP_NONE 0
#define GDP_KEEP_ALL 1
#define GDP_KEEP_BEST 2          /* best end cell, find_best_endpoint_{8,16} */
#define GDP_KEEP_BEST_QUERYEND 3 /* best end cell, ..._to_queryend_indels_{8,16} */
#define GDP_BEST_SLOTS 4         /* int16 score slots of a best-cell record: score, r, c, 0 */

/* Largest band span a problem may use (lband + uband, or the one-sided band),
   after bands wider than the matrix are narrowed to one past it (lband to
   rl + 1 and uband to gl + 1), which changes no cell.  Full-band int16 problems with spans
   up to 960 are swept in pairs; wider ones one to a 16-lane group with a
   ring of up to 8192 entries. */
#define GDP_MAX_SPAN 8176
/* GDP_KIND_STD carries 32-bit H and F between row batches, twice the ring
   space per column, so its band span is capped lower. */
#define GDP_MAX_SPAN_STD 526
#define GDP_MAX_LEN 8000

/* One DP problem (48 bytes, pad fields zero).  width 8 or 16 selects the reference's int8 or
   int16 kernel; open/extend are the (negative) gap penalties; jump_late is
   the reference's jump_late_p (ties prefer the gap); keep_scores = 1 asks
   the fill to keep every computed score (gdp_best_endpoint can scan them,
   and gdp_export_* return them).  For the one-sided fills of end problems,
   keep_scores = GDP_KEEP_BEST (or GDP_KEEP_BEST_QUERYEND) keeps only the
   fill's part of gdp_best_endpoint's scan, computed during the fill: the
   best cell of its triangle (the upper fill's cells with c >= r, the lower
   fill's with c < r) in row-major order under the problem's tie rule, from
   score 0 at (0, 0) (or, for queryend_indels, -inf at (rl, 0) over row rl
   only).  It takes GDP_BEST_SLOTS score slots instead of one per cell.
   GDP_KIND_STD is the reference's scalar fill of its non-SIMD builds
   (int32 scores, the full band, genome characters scored as they are):
   width must be 32, std_flags holds upperp / lowerp and saturation is the
   floor of the stored scores (Dynprog_standard's saturation argument,
   >= -32768; callers pass NEG_INFINITY_INT = -32768).  Its kept scores are
   those of the cells with r, c >= 1; row 0 and column 0 read 0. */
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
  int8_t std_flags;   /* GDP_KIND_STD: GDP_STD_UPPER | GDP_STD_LOWER; else 0 */
  int32_t saturation; /* GDP_KIND_STD: score floor; else 0 */
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
  int32_t endc; /* GDP_TB_LOCAL: the genome column the walk stops at */
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

/* With GDP_TB_LOCAL an upper or lower walk is traceback_local_{8,16}_upper /
   _lower (dynprog_end.c:727-1135), the walk of the splice-junction end
   variants: it stops at genome column endc instead of 0 (the upper walk's
   gap runs stop there too), its closing skip is c - endc genome columns
   (upper) or r query rows at column endc + 1 (lower), its diagonal pairs
   skip no '*' column, and its genome skips read the fill's own genome
   sequence rather than the device genome.  gdp_traceback_ends reports the
   cell where the walk stopped, from which the caller's next walk starts.
   A full-band fill with this flag, or a negative stop column, gets
   counts[0] = -1 and no pairs; a stop column at or right of the start
   column stops the walk at once, as the reference's walk does. */
#define GDP_TB_LOCAL 4

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

/* Lane-group choice of the paired sweep.  Full-band int16 problems without
   kept scores and with lband + uband <= g16_max_span (at most 240) are swept
   by 16-lane groups, eight problems to a wave; those with a wider band (up
   to GDP_MAX_SPAN) or with rl >= g64_min_rows by 64-lane groups, two to a
   wave, whose sweep takes about a quarter of the steps (lower latency for a
   long problem, more padding per cell).  Defaults: {240, GDP_MAX_LEN + 1}. */
typedef struct gdp_plan_params {
  int32_t g16_max_span;
  int32_t g64_min_rows;
  /* mix_span > 0: the problems that do not fill a wave of their own band
     share mixed waves, in which each problem's lower and upper band lies
     within mix_span of the wave's largest; the sweep is laid out for the
     largest and every problem keeps its own band edges.  0: a wave holds
     problems of one band only. */
  int32_t mix_span;
  /* With mix_span > 0, how mixed waves are formed.  0: the problems left
     over from full waves of one band, in order of band width, lower band
     and rl.  1: every problem binned by lband / mix_span and
     uband / mix_span and taken in order of rl within its bin (a wave's
     problems then agree in rows as well as in bands), full waves first,
     what is left binned again by 2 * mix_span. */
  int32_t mix_order;
} gdp_plan_params;

/* As gdp_plan_ex with explicit lane-group parameters (NULL = defaults). */
int gdp_plan_tuned(const gdp_problem *probs, int32_t n, int32_t flags, const gdp_plan_params *params,
                   void *descs_out, int64_t *dir_words, int64_t *score_slots);

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
   bit 0 single sweeps, bits 1..3 the paired wave classes, bit 4 standard
   fills (GDP_KIND_STD), bit 5 narrow one-sided sweeps (one-sided fills with
   a band under 32 columns that keep at most their best end cell, one per
   lane), bit 6 the wide single sweep (batch periods past 1024 steps: one
   wave per workgroup, rings of up to 8192 entries), bits 8..11 the log2 of
   the single sweep's LDS ring, bits 12..15 that of the standard fill's and
   bits 16..19 that of the wide sweep's (the next power of two >= the
   largest batch period among their problems, in ring words; 0 means the
   largest ring).  gdp_fill_ex launches only those (gdp_fill launches all,
   and empty ones return at once). */
#define GDP_FILL_STD 0x10u
#define GDP_FILL_NARROW 0x20u
#define GDP_FILL_WIDE 0x40u
#define GDP_FILL_ALL 0x7Fu
uint32_t gdp_plan_kernels(const void *descs, int32_t n);
int gdp_fill_ex(const void *d_descs, int32_t n, uint32_t kernels, const char *d_qbuf, const char *d_gbuf,
                uint32_t *d_dirs, int16_t *d_scores, int32_t *d_final, void *stream);

/* The one-launch-per-problem form, problem after problem: for k = 0..n-1,
   problem k is filled alone (one launch of its own kernel, as
   Dynprog_simd_{8,16} fills one problem per call), then request k is walked
   (one traceback launch), and the stream is synchronized before the next
   problem, as a caller that needs each result before it goes on would.
   descs is the host copy of a GDP_PLAN_SINGLE plan (it gives each problem's
   kernel); d_descs its device copy; d_reqs[k] the request of problem k.
   No counterpart in the reference (the per-call baseline of the batched
   form). */
int gdp_align_each(const void *descs, const void *d_descs, int32_t n, const char *d_qbuf, const char *d_gbuf,
                   uint32_t *d_dirs, int16_t *d_scores, int32_t *d_final, const gdp_tb_request *d_reqs,
                   const void *d_genome, int32_t genome_format, uint32_t genome_len, gdp_pair *d_pairs,
                   int32_t *d_counts, void *stream);

/* A fill followed by a traceback, captured once into a HIP graph and
   replayed with gdp_graph_launch: the whole batched step in one launch
   from the host.  Arguments are those of gdp_fill_ex and gdp_traceback_ex;
   the buffers they name must stay allocated while the graph lives, and
   their contents may change between replays.  The persistent fill resets
   its queues at the end of each run, so replays need no host work.  The
   capture runs on a stream of the library's own; nothing executes until
   gdp_graph_launch. */
typedef struct gdp_graph gdp_graph;

/* General form: gdp_graph_begin hands out a stream of the library's in
   capture mode; every gdp_* call enqueued on it (fills, best end cells,
   tracebacks, segment extraction) is recorded, not run; gdp_graph_end
   instantiates the graph.  Capture is per thread: enqueue from the thread
   that began it. */
int gdp_graph_begin(gdp_graph **g, void **stream);
int gdp_graph_end(gdp_graph *g);

gdp_graph *gdp_graph_create(const void *d_descs, int32_t n, uint32_t kernels, const char *d_qbuf, const char *d_gbuf,
                            uint32_t *d_dirs, int16_t *d_scores, int32_t *d_final, const gdp_tb_request *d_reqs,
                            int32_t nreq, const void *d_genome, int32_t genome_format, uint32_t genome_len,
                            gdp_pair *d_pairs, int32_t *d_counts);
int gdp_graph_launch(gdp_graph *g, void *stream);
void gdp_graph_destroy(gdp_graph *g);

/* Best end cell of end problems: for each of m pairs (upper, lower) of
   one-sided fills with keep_scores set, find_best_endpoint_{8,16}
   (dynprog_end.c:143-290) when queryend_indels = 0, or
   find_best_endpoint_to_queryend_indels_{8,16} (358-508) otherwise.
   d_out[3*k] = {finalscore, bestr, bestc}.  Fills kept with GDP_KEEP_ALL are
   scanned; fills kept with GDP_KEEP_BEST (queryend_indels = 0) or
   GDP_KEEP_BEST_QUERYEND (queryend_indels = 1) hand over their triangle's
   best cell and the pair's two are merged.  A pair whose fills kept nothing,
   or kept the other mode, gets finalscore INT32_MIN and cell (-1, -1). */
int gdp_best_endpoint(const void *d_descs, const int32_t *d_ul, int32_t m, int32_t queryend_indels,
                      const int16_t *d_scores, int32_t *d_out, void *stream);

/* As gdp_best_endpoint; with d_tb_template and d_tb_out (both or neither)
   it also writes the traceback request of each pair: d_tb_out[k] =
   d_tb_template[k] with fill, r and c set to the best cell and the triangle
   holding it (the upper fill when bestc >= bestr, dynprog_end.c:1576-1607),
   so the tracebacks can follow on the device without a host round trip.
   A best cell of (0, 0) gives a request that pushes no pairs. */
int gdp_best_endpoint_ex(const void *d_descs, const int32_t *d_ul, int32_t m, int32_t queryend_indels,
                         const int16_t *d_scores, int32_t *d_out, const gdp_tb_request *d_tb_template,
                         gdp_tb_request *d_tb_out, void *stream);

/* cDNA gap bridge: the step of Dynprog_cdna_gap (dynprog_cdna.c:786-1180)
   that joins a left and a right end across one genomic gap,
   bridge_cdna_gap_{8,16}_ud (dynprog_cdna.c:123-381, 386-644; called at
   977 and 1106), batched.  For each of m gaps d_quads[5k .. 5k+4] =
   {left upper, left lower, right upper, right lower, rightoffset -
   leftoffset}: indices of four one-sided fills of one width over the gap's
   gl genome columns, kept with GDP_KEEP_ALL -- the left pair over
   rsequenceL forward (jump_late_p), the right pair over rev_rsequenceR read
   backwards against the reversed genome (revp, !jump_late_p) -- and the
   two ends' query offsets' difference, which keeps the ends' rows apart.
   Open penalty and tie rule come from the left upper fill.
   d_out[5k .. 5k+4] = {bestscore, bestcL, bestcR, bestrL, bestrR}; when no
   join reaches the initial score (-128 / -32768) bestscore is that score and
   the cells are -1 (the reference leaves them unset); gaps whose fills do
   not fit (kinds, widths, lengths, kept scores; or a right band of more
   than 1023 rows in a gap too large for the kernel's 8192-word LDS tables:
   gl + 1 words, one per in-band right-end cell, and twice the left band)
   get INT32_MIN.  Rows of an upper triangle beyond its rl, which
   the reference reads from its workspace only when rl < gl-1, are not
   candidates.  With d_tb_template and d_tb_out (both or neither)
   d_tb_out[2k] is the right end's traceback request and d_tb_out[2k+1] the
   left end's (d_tb_template's entries with fill, r and c set: the upper
   fill when bestc >= bestr, dynprog_cdna.c:995-1060); cells of -1 give
   requests at (0, 0) that push no pairs.  Indices must name problems of
   d_descs. */
int gdp_bridge_cdna(const void *d_descs, const int32_t *d_quads, int32_t m, const int16_t *d_scores,
                    int32_t *d_out, const gdp_tb_request *d_tb_template, gdp_tb_request *d_tb_out, void *stream);

/* One intron gap of gdp_bridge_intron (32 bytes, pad zero). */
typedef struct gdp_intron_gap {
  int32_t fills[4];       /* left upper, left lower, right upper, right lower */
  int32_t span;           /* rightoffset - leftoffset (rev_goffsetR - goffsetL) */
  int8_t cdna_direction;  /* > 0 sense, < 0 antisense, 0 either */
  int8_t finalp;          /* the final intron rewards instead of the preliminary ones */
  int8_t halfp;           /* half the chosen sites' reward off the score */
  int8_t pad;
  int64_t prob_off;       /* d_probs index of the left end's first column; the right end's follow */
} gdp_intron_gap;

/* Intron gap bridge: the step of Dynprog_genome_gap (dynprog_genome.c:
   3287-3770) that picks where an intron joins a left and a right exon end,
   bridge_intron_gap_{8,16}_ud (1387-1488, 2262-2363; called at 3548, 3690) on
   its site-level path (bridge_intron_gap_{8,16}_site_level, 866-1386,
   1741-2261: novel splicing, no splicing IIT), batched.  For each of m gaps,
   d_gaps[k].fills name four one-sided fills of one width and rlength rows,
   kept with GDP_KEEP_ALL: the left pair over rsequence forward against the
   left segment (glengthL columns, jump_late_p), the right pair over
   rev_rsequence against the right segment read backwards (glengthR columns,
   stored reversed, !jump_late_p).  Candidates pair row rL of the left end
   with row rlength-rL of the right end, at least one end on its diagonal.
   A candidate scores left + intron reward + right: the reward of the
   canonical dinucleotides (GT-AG, GC-AG, AT-AC and their antisense forms,
   either allele) read from the fills' genome bytes in d_gbuf after the left
   exon's column and before the right exon's, under cdna_direction and
   finalp (intron_score_setup, 143-186).  Ties on score go to the higher sum
   of the two columns' splice-site probabilities, read from d_probs (the
   caller's per-column site model values, left_probabilities and
   right_probabilities of 1845-1933: glengthL doubles from prob_off, then
   glengthR; entries glength-1 are not read); the first candidate of the
   reference's scan keeps an exact tie.  A diagonal pair with a canonical
   intron and a higher probability sum replaces the best as the reference's
   dinucleotide rule decides (2210-2237).  d_out[5k .. 5k+4] = {finalscore,
   bestrL, bestrR, bestcL, bestcR} (halfp applied; cells -1 when no candidate
   beats the initial -128 / -32768); d_out_probs[2k, 2k+1] = the chosen
   columns' probabilities when finalscore >= 0, else 0.0 (get_splicesite_probs,
   332-401).  Gaps whose fills do not fit (kinds, widths, row counts,
   kept scores, glengthL or glengthR not above rlength -- the reference
   reads past its probability arrays there -- or above GDP_MAX_LEN) get
   INT32_MIN.  With d_tb_template and d_tb_out (both or neither), d_tb_out[2k]
   is the right end's traceback request and d_tb_out[2k+1] the left end's,
   from the template with fill, r and c set: the upper fill when bestc >= bestr
   (3575-3633, 3718-3774); a gap with finalscore < 0 gets requests at (0, 0), which push
   no pairs. */
int gdp_bridge_intron(const void *d_descs, const gdp_intron_gap *d_gaps, int32_t m, const int16_t *d_scores,
                      const char *d_gbuf, const double *d_probs, int32_t *d_out, double *d_out_probs,
                      const gdp_tb_request *d_tb_template, gdp_tb_request *d_tb_out, void *stream);

/* ---- Genome in device memory ------------------------------------------ */

/* Copies a batch's walk results to pinned host memory (hipHostMalloc'd
   buffers as large as d_counts and d_pairs): each request's six counts and
   only the pair records its walk wrote (counts[6k], at most max_pairs), at
   the request's pair_off, as a kernel on the stream that writes across PCIe.
   The rest of h_pairs is left as it was.  No counterpart in the reference. */
int gdp_download_results(const gdp_tb_request *d_reqs, int32_t nreq, const int32_t *d_counts, const gdp_pair *d_pairs,
                         int32_t *h_counts, gdp_pair *h_pairs, void *stream);

/* Formats of a device genome handed to gdp_traceback_ex and
   gdp_genome_segments: ASCII bytes indexed by position, or GMAP's genomecomp
   image.  A genomecomp image holds three 32-bit words per 32 bases,
   {high, low, flags}: bases 0..15 of the block two bits each in low (base k at
   bits 2k), bases 16..31 in high, A C G T = 0 1 2 3, and flag bit k set where
   base k reads as 'N' (uncompress_one_char, genome.c:9723-9760).  Positions
   at or past the genome length read 'N' in both formats. */
#define GDP_GENOME_ASCII 0
#define GDP_GENOME_COMP 1

/* 32-bit words of the genomecomp image of n bases: 3 per started block of 32
   plus the two trailing words the reference writes (compress-write.c:633-637). */
int64_t gdp_genomecomp_words(uint64_t nbases);

/* Packs n ASCII bases into a genomecomp image of gdp_genomecomp_words(n)
   words, as Compress_compress does (compress-write.c:535-640): A C G T in
   either case by code, N as a flag, X as T with a flag, any other character
   as a flag; the last block is padded with X and two 0xFFFFFFFF words
   follow.  Byte-identical to the .genomecomp file gmap_build writes. */
int gdp_genomecomp_pack(const char *seq, uint64_t nbases, uint32_t *out);

/* One genome segment to extract.  left = 0: Genome_get_segment_right
   (genome.c:11022-11075) of the bases [coord, coord + length), those at or
   past bound = chrhigh read '*'; left = 1: Genome_get_segment_left
   (genome.c:11078-11120) of [coord - length, coord), those below
   bound = chroffset read '*'.  revcomp reverses and complements the segment.
   reverse = 1 then writes it last byte first: the order in which a fill
   reading the segment backwards from its last base (revp) takes it.  The
   length bytes are written at d_out + out_off (no terminator). */
typedef struct gdp_segment {
  uint32_t coord;
  uint32_t bound;
  int32_t length;
  int8_t left;
  int8_t revcomp;
  int16_t reverse;
  int64_t out_off;
} gdp_segment;

/* Extracts n segments of a device genome (format GDP_GENOME_*, genome_len
   bases) on the device, one wave per segment. */
int gdp_genome_segments(const void *d_genome, int32_t format, uint32_t genome_len, const gdp_segment *d_segs,
                        int32_t n, char *d_out, void *stream);

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

/* As gdp_traceback, with the genome skips read from a device genome of either
   format (GDP_GENOME_ASCII as gdp_traceback, or GDP_GENOME_COMP). */
int gdp_traceback_ex(const void *d_descs, const gdp_tb_request *d_reqs, int32_t nreq, const char *d_qbuf,
                     const char *d_gbuf, const uint32_t *d_dirs, const void *d_genome, int32_t genome_format,
                     uint32_t genome_len, gdp_pair *d_pairs, int32_t *d_counts, void *stream);

/* As gdp_traceback_ex; d_ends (2 int32 per request, or NULL) receives the
   cell (r, c) where each walk stopped before its closing skip -- what
   traceback_local_{8,16}_{upper,lower} (dynprog_end.c:727-1135) leave in
   *r and *c, from which Dynprog_end5/3_splicejunction (1652-1922,
   2248-2745) start their second walk past the gap holder. */
int gdp_traceback_ends(const void *d_descs, const gdp_tb_request *d_reqs, int32_t nreq, const char *d_qbuf,
                       const char *d_gbuf, const uint32_t *d_dirs, const void *d_genome, int32_t genome_format,
                       uint32_t genome_len, gdp_pair *d_pairs, int32_t *d_counts, int32_t *d_ends, void *stream);

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

/* ---- Splice-junction ends: the whole caller on the device ------------- */

/* One end of Dynprog_end5_splicejunction (dynprog_end.c:1652-1922) or
   Dynprog_end3_splicejunction (2248-2745), 56 bytes, pad zero.  The read
   end and its genome window are stored in sweep order: the 5' form's
   rev_rsequence and rev_gsequence reversed (the bases from the anchor
   outwards), the 3' form's rsequence and gsequence as they are; the window's
   alternate allele follows it (gl bytes each from goff). */
typedef struct gdp_sj_end {
  int64_t qoff;        /* the read end in qbuf */
  int64_t goff;        /* its genome window in gbuf, then the alternate allele */
  int32_t rl, gl;      /* length1 (read), length2 (window) */
  int32_t roffset;     /* revoffset1 (5') / offset1 (3') */
  int32_t anchor;      /* revoffset2_anchor (5') / offset2_anchor (3') */
  int32_t far;         /* revoffset2_far (5') / offset2_far (3') */
  int32_t contlength;  /* the continuation column where the first walk stops */
  int16_t extraband;   /* extraband_end */
  int8_t genestrand;
  int8_t jump_late;    /* the caller's jump_late_p */
  int32_t pad;
  double defect_rate;
} gdp_sj_end;

/* A pair record of a splice-junction end: gdp_pair with a 32-bit genomejump,
   since the known-site gap holder (gapp 2) spans the intron. */
typedef struct gdp_sj_pair {
  int32_t querypos;
  int32_t genomepos;
  int8_t cdna;
  int8_t comp;
  int8_t genome;
  int8_t genomealt;
  int16_t gapp;
  int16_t pad;
  int32_t genomejump;
} gdp_sj_pair;

/* Per end: npairs -1 for the reference's NULL list, else the number of
   records at pair_off of the batch's pair array, in the order the returned
   list holds them from its last element to its head (the 5' form's push
   order; the 3' form's reversed); counters = {traceback_score (finalscore),
   missscore, nmatches, nmismatches, nopens, nindels}. */
typedef struct gdp_sj_result {
  int64_t pair_off;
  int32_t npairs;
  int32_t counters[6];
  int32_t pad;
} gdp_sj_result;

/* A batch object owning the device and host buffers of gdp_sj_run; they
   grow to the largest batch run through it and are reused. */
typedef struct gdp_sj_batch gdp_sj_batch;
gdp_sj_batch *gdp_sj_create(void);
void gdp_sj_destroy(gdp_sj_batch *b);

/* Runs n ends (five = 1: the 5' caller, 0: the 3' caller) from host buffers
   qbuf / gbuf of qbytes / gbytes bytes: per end the upper and lower fill of
   the read against its window (ENDQ scores, END_OPEN by defect rate with
   END_EXTEND -2, bands widened to the far corner, int8 when the read or the
   window is shorter than use8p_size[ENDQ], the 5' fills with the opposite
   tie rule), the best cell of the last row
   (find_best_endpoint_to_queryend_indels_{8,16}), and when it is not below
   0 the two local walks (to contlength, then from where that one stopped to
   column 0) joined by the known gap holder of far - anchor (5') or anchor -
   far (3') genome columns, with indel pairs at the read's end dropped
   (1797-1910, 2385-2490).  Ends of length 0 or past max_rlength /
   max_glength, and ends whose best cell scores below 0, get npairs -1.  The
   fills, scans, walks and the join all run on the device; the results and
   pairs are copied back to the batch's host arrays before it returns. */
int gdp_sj_run(gdp_sj_batch *b, int32_t five, const gdp_sj_end *ends, int32_t n, const char *qbuf, int64_t qbytes,
               const char *gbuf, int64_t gbytes, int32_t max_rlength, int32_t max_glength, void *stream);
/* The last run's pair count, per-end results and pair records (owned by the
   batch, valid until its next run). */
int64_t gdp_sj_npairs(const gdp_sj_batch *b);
const gdp_sj_result *gdp_sj_results(const gdp_sj_batch *b);
const gdp_sj_pair *gdp_sj_pairs(const gdp_sj_batch *b);

/* ---- Single gaps: the whole Dynprog_single_gap caller on the device --- */

/* One Dynprog_single_gap call (dynprog_single.c:428-675), 48 bytes, pad
   zero.  The read (rsequence) is rl bytes at qoff of the batch's query
   buffer; rsequenceuc is its upper case, as GMAP builds it.  The genome
   window comes from a device genome as the caller takes it
   (dynprog_single.c:544-550): Genome_get_segment_right(chroffset + goffset,
   gl, chrhigh) on the plus strand, the reverse complement
   Genome_get_segment_left(chrhigh - goffset + 1, gl, chroffset) on the minus
   strand; the genome is its own alternate allele. */
typedef struct gdp_sg_call {
  int64_t qoff;
  int32_t rl, gl;          /* rlength, glength */
  int32_t roffset, goffset;
  uint32_t chroffset, chrhigh;
  int32_t extraband;       /* extraband_single */
  int8_t watsonp;
  int8_t genestrand;
  int8_t jump_late;        /* jump_late_p */
  int8_t widebandp;
  double defect_rate;
} gdp_sg_call;

/* Per call: npairs -1 for the reference's NULL list, else the number of
   records at pair_off of the batch's pair array, in the order of the
   returned list from its head; counters = {traceback_score, nmatches,
   nmismatches, nopens, nindels}; simple = 1 when single_gap_simple gave the
   list.  Every call advances the caller's dynprogindex by one step, as the
   reference's returns here all do. */
typedef struct gdp_sg_result {
  int64_t pair_off;
  int32_t npairs;
  int32_t counters[5];
  int32_t simple;
  int32_t pad;
} gdp_sg_result;

/* A batch object owning the device and host buffers of gdp_sg_run; they
   grow to the largest batch run through it and are reused. */
typedef struct gdp_sg_batch gdp_sg_batch;
gdp_sg_batch *gdp_sg_create(void);
void gdp_sg_destroy(gdp_sg_batch *b);

/* Runs n calls of Dynprog_single_gap with reads from the host buffer qbuf
   (qbytes bytes) against a device genome (GDP_GENOME_* format, genome_len
   bases), the reference's steps for each:
   - lengths of 0 or past max_rlength / max_glength: NULL, traceback_score
     NEG_INFINITY_32 (dynprog_single.c:508-520);
   - mismatch type and gap penalties by defect rate (DEFECT_HIGHQ 0.003,
     DEFECT_MEDQ 0.014; SINGLE_OPEN/EXTEND of dynprog.h:61-67), or
     user_open / user_extend when user_dynprog is set (GMAP's
     --indel-open / --indel-extend);
   - the window from the device genome;
   - glength == rlength: single_gap_simple, kept when it pushes a pair with
     at most one mismatch;
   - else the bands of Dynprog_compute_bands, the int8 fill when rlength and
     glength are both below use8p_size[mismatchtype], the int16 fill
     otherwise, and the traceback from (rlength, glength), returned in push
     order (the reference reverses the pushed list).
   GMAP's homopolymer mode (homopolymerp) is not part of it.  Segments,
   fills, walks and the choice between the two lists run on the device; the
   results and pairs are copied back to the batch's host arrays before it
   returns. */
int gdp_sg_run(gdp_sg_batch *b, const gdp_sg_call *calls, int32_t n, const char *qbuf, int64_t qbytes,
               const void *d_genome, int32_t genome_format, uint32_t genome_len, int32_t max_rlength,
               int32_t max_glength, int32_t user_dynprog, int32_t user_open, int32_t user_extend, void *stream);
/* The last run's pair count, per-call results and pair records (owned by
   the batch, valid until its next run). */
int64_t gdp_sg_npairs(const gdp_sg_batch *b);
const gdp_sg_result *gdp_sg_results(const gdp_sg_batch *b);
const gdp_pair *gdp_sg_pairs(const gdp_sg_batch *b);

/* ---- End gaps: the whole Dynprog_end5_gap / Dynprog_end3_gap caller ---- */

/* One call of Dynprog_end5_gap (dynprog_end.c:1293-1650) or Dynprog_end3_gap
   (1923-2243), 48 bytes, pad zero.  The read is rl bytes at qoff of the
   batch's query buffer, first base first (the 5' caller is handed a pointer
   to its last base, rev_rsequence = qbuf + qoff + rl - 1, and reads it
   backwards).  roffset / goffset are the 5' caller's rev_roffset /
   rev_goffset or the 3' caller's roffset / goffset.  endalign is Endalign_T
   (dynprog.h:24): 0 QUERYEND_GAP, 1 QUERYEND_INDELS, 2 QUERYEND_NOGAPS,
   3 BEST_LOCAL.  The genome window comes from a device genome as the
   callers take it: 5' Genome_get_segment_left(chroffset + rev_goffset + 1,
   gl, chroffset) on the plus strand and the reverse complement
   Genome_get_segment_right(chrhigh - rev_goffset, gl, chrhigh) on the
   minus strand; 3' as Dynprog_single_gap takes its window. */
typedef struct gdp_eg_call {
  int64_t qoff;
  int32_t rl, gl;
  int32_t roffset, goffset;
  uint32_t chroffset, chrhigh;
  int32_t extraband;       /* extraband_end */
  int8_t watsonp;
  int8_t genestrand;
  int8_t jump_late;        /* jump_late_p */
  int8_t endalign;
  double defect_rate;
} gdp_eg_call;

/* Per call: npairs -1 for NULL, else the records at pair_off in the order
   of the returned list from its head; counters = {traceback_score,
   nmatches, nmismatches, nopens, nindels} as the caller leaves them;
   finalscore = the best end cell's score (0 for QUERYEND_NOGAPS), which
   the reference computes but does not return; advanced = 1 when the
   reference advances its dynprogindex (every return but the early NULLs for
   lengths and offsets). */
typedef struct gdp_eg_result {
  int64_t pair_off;
  int32_t npairs;
  int32_t counters[5];
  int32_t finalscore;
  int32_t advanced;
} gdp_eg_result;

typedef struct gdp_eg_batch gdp_eg_batch;
gdp_eg_batch *gdp_eg_create(void);
void gdp_eg_destroy(gdp_eg_batch *b);

/* Runs n calls of the 5' caller (five = 1) or the 3' caller (five = 0)
   against a device genome, the reference's steps for each:
   - rl <= 0, gl <= 0, and on the 5' side rev_goffset < 0: NULL with every
     counter 0;
   - rl and gl chopped to max_rlength / max_glength outside
     QUERYEND_NOGAPS;
   - ENDQ scores, END_OPEN by defect rate and END_EXTEND -2
     (dynprog_end.c:88-94), or user_open / user_extend when user_dynprog;
   - QUERYEND_GAP and BEST_LOCAL: the upper and lower fills over the wide
     bands of Dynprog_compute_bands, int8 when rl or gl is below
     use8p_size[ENDQ], the 5' fills with the opposite tie rule, and the walk
     from find_best_endpoint's cell in the triangle holding it;
     QUERYEND_INDELS the same over the narrow bands from
     find_best_endpoint_to_queryend_indels' cell; QUERYEND_NOGAPS the
     diagonal walk traceback_nogaps from (min(rl, gl), min(rl, gl));
   - require_pos_score_p: NULL outside QUERYEND_NOGAPS (the reference tests
     the score it has just set to 0), every counter 0;
   - QUERYEND_GAP and BEST_LOCAL: NULL with traceback_score 0 when
     nmatches + 1 < nmismatches;
   - else the pushed list less its first pairs while they are indels,
     returned reversed (5') or as pushed (3').
   Every step runs on the device; results and pairs are copied back to the
   batch's host arrays before it returns. */
int gdp_eg_run(gdp_eg_batch *b, int32_t five, const gdp_eg_call *calls, int32_t n, const char *qbuf, int64_t qbytes,
               const void *d_genome, int32_t genome_format, uint32_t genome_len, int32_t max_rlength,
               int32_t max_glength, int32_t user_dynprog, int32_t user_open, int32_t user_extend,
               int32_t require_pos_score, void *stream);
int64_t gdp_eg_npairs(const gdp_eg_batch *b);
const gdp_eg_result *gdp_eg_results(const gdp_eg_batch *b);
const gdp_pair *gdp_eg_pairs(const gdp_eg_batch *b);

/* ---- Batched Dynprog_cdna_gap ----------------------------------------- */

/* One call of Dynprog_cdna_gap (dynprog_cdna.c:786-1180), as stage 3 makes
   it: the left end reads qbuf[qoffL ...] forward (rsequenceL, roffsetL its
   query position), the right end reads qbuf[qoffR], qbuf[qoffR - 1], ...
   backwards (rev_rsequenceR, rev_roffsetR), over the glength genome columns
   from goffset (rev_goffset = goffset + glength - 1).  The left end's bytes
   past rlengthL up to the right end's first base are read too: the cDNA
   insertion of a join copies them. */
typedef struct gdp_cd_call {
  int64_t qoffL, qoffR;
  int32_t rlengthL, rlengthR, glength;
  int32_t roffsetL, rev_roffsetR, goffset;
  uint32_t chroffset, chrhigh;
  int32_t extraband;       /* extraband_paired */
  int8_t watsonp;
  int8_t genestrand;
  int8_t jump_late;        /* jump_late_p */
  int8_t pad;
  double defect_rate;
} gdp_cd_call;

/* Per call: npairs -1 for NULL, else the records at pair_off in the order of
   the returned list from its head (the right end's walk, the cDNA and genome
   insertion pairs or the gap holder, the left end's walk); traceback_score
   and incomplete (incompletep) as the caller leaves them; finalscore = the
   bridge's score; advanced = 1 when the caller advances its dynprogindex
   (every return but the NULL for glength <= 1). */
typedef struct gdp_cd_result {
  int64_t pair_off;
  int32_t npairs;
  int32_t traceback_score;
  int32_t incomplete;
  int32_t advanced;
  int32_t finalscore;
  int32_t pad;
} gdp_cd_result;

typedef struct gdp_cd_batch gdp_cd_batch;
gdp_cd_batch *gdp_cd_create(void);
void gdp_cd_destroy(gdp_cd_batch *b);

/* Runs n calls of Dynprog_cdna_gap against a device genome, the reference's
   steps for each:
   - glength <= 1: NULL, counters 0, dynprogindex kept;
   - glength > max_glength, or rlengthR or rlengthL > max_rlength: NULL,
     dynprogindex advanced;
   - the mismatch type by defect rate and the cDNA penalties (-10, -7), the
     int8 fills when glength or both ends are under use8p_size;
   - the genome windows cut from the device genome (Genome_get_segment_right
     from goffset and Genome_get_segment_left to rev_goffset on the plus
     strand, their reverse complements on the minus strand), the four
     one-sided fills with bands from Dynprog_compute_bands (widebandp), the
     right end's over the reversed read and window with the other tie rule;
   - the bridge (gdp_bridge_cdna), the right end's walk, then the cDNA and
     genome insertion pairs when the join leaves 9 query and 9 genome
     positions (INSERT_PAIRS), else a gap holder with incomplete set, then
     the left end's walk; a list of the gap holder alone is NULL.
   Every step runs on the device; results and pairs are copied back to the
   batch's host arrays before it returns. */
int gdp_cd_run(gdp_cd_batch *b, const gdp_cd_call *calls, int32_t n, const char *qbuf, int64_t qbytes,
               const void *d_genome, int32_t genome_format, uint32_t genome_len, int32_t max_rlength,
               int32_t max_glength, void *stream);
int64_t gdp_cd_npairs(const gdp_cd_batch *b);
const gdp_cd_result *gdp_cd_results(const gdp_cd_batch *b);
const gdp_pair *gdp_cd_pairs(const gdp_cd_batch *b);

/* ---- Serving: continuous batching of alignment requests -------------- */

/* A request loop over a device-resident genome.  Each request is the DP of
   Dynprog_single_gap (dynprog_single.c:590-629): bands from
   Dynprog_compute_bands, the int8 kernel when rl and gl are both below
   use8p_size[mismatchtype] and the int16 kernel otherwise, the final score
   matrix[gl][rl], and the traceback from (rl, gl).  The genome segment is
   taken on the device as the reference takes it (dynprog_single.c:544-550):
   Genome_get_segment_right(chroffset + goffset, gl, chrhigh) on the plus
   strand, the reverse complement Genome_get_segment_left(chrhigh - goffset + 1,
   gl, chroffset) on the minus strand.  (The caller-side shortcuts of
   Dynprog_single_gap, single_gap_simple for gl == rl and the homopolymer
   mode, are not part of the request.)

   Each batch slot has a thread of its own, which takes submitted requests in
   arrival order into a batch whenever its slot is free, as far as the
   direction words of the batch fit the free part of a fixed device budget,
   plans it, and runs it on the slot's stream; slots plan in parallel and
   batches overlap on the device, and results are returned as their batch
   completes.  A request's result does not depend on the batch it ran in. */
typedef struct gdp_server gdp_server;

typedef struct gdp_server_config {
  int64_t dir_budget_bytes; /* device memory for the direction words of all batches in flight */
  int64_t pair_capacity;    /* pair records one batch may produce (sum over requests of rl + gl + 2) */
  int32_t max_batch;        /* requests per batch */
  int32_t slots;            /* batches in flight at once, one HIP stream each (1..16); batches overlap
                               when each stream has a hardware queue of its own (GPU_MAX_HW_QUEUES) */
  int32_t max_rl;           /* longest read accepted (<= GDP_MAX_LEN) */
  int32_t max_gl;           /* longest genome segment accepted (<= GDP_MAX_LEN) */
  int32_t result_buffers;   /* pinned result buffers (>= slots; 0 = slots + 2) */
  int32_t long_rows;        /* reads of at least this many bases batch apart from shorter ones (0 = one class) */
  gdp_plan_params plan;     /* lane groups of the batches (all zero = gdp_plan's defaults) */
} gdp_server_config;

/* One alignment request (64 bytes).  read: rl query bytes, copied by
   gdp_server_submit.  roffset/goffset/chroffset/chrhigh/watsonp/genestrand
   have Dynprog_single_gap's meaning; extraband and widebandp feed
   Dynprog_compute_bands; open/extend/mismatchtype are the fill's scoring,
   or, with scoring = 1, chosen from defect_rate as Dynprog_single_gap
   chooses them (dynprog_single.c:458-481: HIGHQ -8/-3 below 0.003, MEDQ
   -7/-2 below 0.014, LOWQ -6/-1 above). */
typedef struct gdp_align_request {
  uint64_t tag;
  const char *read;
  int32_t rl;
  int32_t gl;
  int32_t roffset;
  int32_t goffset;
  uint32_t chroffset;
  uint32_t chrhigh;
  int32_t extraband;
  int8_t watsonp;
  int8_t widebandp;
  int8_t jump_late;
  int8_t genestrand;
  int8_t mismatchtype;
  int8_t open;
  int8_t extend;
  int8_t scoring;          /* 0: mismatchtype/open/extend as given; 1: by defect_rate */
  int32_t pad;
  double defect_rate;
} gdp_align_request;

/* One result (72 bytes).  status GDP_OK, or GDP_ETOOBIG for a request the
   reference also refuses (rl or gl <= 0 or above the server's limits:
   traceback_score = -32768, NEG_INFINITY_32, no pairs).  counts =
   {npairs, traceback_score, nmatches, nmismatches, nopens, nindels}; the
   pairs are in the traceback's push order (Dynprog_single_gap returns them
   reversed) at `pairs`, in pinned host memory the server owns, valid until
   the next gdp_server_poll on this server (NULL when npairs is 0).  Times
   are gdp_server_now() seconds. */
typedef struct gdp_align_result {
  uint64_t tag;
  int32_t status;
  int32_t finalscore;
  int32_t counts[6];
  int32_t width;
  int32_t pad;
  const gdp_pair *pairs;
  double t_submit;
  double t_done;
} gdp_align_result;

typedef struct gdp_server_stats {
  int64_t submitted;
  int64_t completed;
  int64_t batches;
  int64_t dir_bytes_peak; /* most direction bytes in flight at once */
  int32_t inflight_peak;  /* most batches in flight at once */
  int32_t pad;
  double prep_seconds;    /* slot threads, summed: forming, planning and launching batches */
  double retire_seconds;  /* slot threads, summed: turning finished batches into results */
  double batch_seconds;   /* sum over batches of launch-to-results wall time */
  double device_seconds;  /* sum over batches of stream time from upload to download */
} gdp_server_stats;

/* Starts a server over a device genome (GDP_GENOME_* format, genome_len
   bases) that the caller keeps alive until gdp_server_destroy.  gdp_init
   must have run.  Returns NULL on failure (gdp_last_error). */
gdp_server *gdp_server_create(const gdp_server_config *cfg, const void *d_genome, int32_t genome_format,
                              uint32_t genome_len);

/* Queues n requests (thread-safe; any number of submitting threads).  Each
   request gets exactly one result.  Once a HIP call of the server has failed,
   no batch starts again: the requests still waiting and every request
   submitted later get results with status GDP_EHIP, and submit returns
   GDP_EHIP. */
int gdp_server_submit(gdp_server *s, const gdp_align_request *reqs, int32_t n);

/* Moves up to max finished results into out, waiting up to wait_us
   microseconds for the first, and gives the pinned buffers of the results
   returned by the previous call back to the server (their pairs pointers
   become invalid).  One thread polls.  Returns the number of results.  When
   the consumer holds every result buffer, no new batch starts until it
   polls again. */
int32_t gdp_server_poll(gdp_server *s, gdp_align_result *out, int32_t max, int32_t wait_us);

/* Copies the statistics; returns GDP_EHIP (message in gdp_last_error) once
   the server has stopped on a failure. */
int gdp_server_get_stats(gdp_server *s, gdp_server_stats *st);

/* Finishes the batches in flight and stops the slots' threads. */
void gdp_server_destroy(gdp_server *s);

/* Seconds on the server's monotonic clock. */
double gdp_server_now(void);

#ifdef __cplusplus
}
#endif
#endif
