/*
 * gmapdp_dynprog.h -- drop-in replacement for GMAP's src/dynprog_simd.h.
 *
 * libgmapdp_dynprog.so exports the twelve functions of the reference's
 * dynprog_simd.h (src/dynprog_simd.h:16-126, non-DEBUG prototypes) with the
 * same C signatures and the same memory contract, so src/dynprog_single.c,
 * src/dynprog_end.c, src/dynprog_genome.c and src/dynprog_cdna.c link
 * against it unchanged in place of dynprog_simd.o.  The fills run on the
 * GPU (gfx950) and return host Score/Direction tables laid out exactly as
 * the reference's aligned_*_alloc (dynprog_simd.c:2158-2282) lays them out,
 * inside the caller's Dynprog_T workspace; the tracebacks walk the GPU copy
 * of the directions and build the pair list through the caller's own
 * Pairpool_push / Pairpool_add_queryskip / Pairpool_add_genomeskip.
 *
 * Link-time dependencies supplied by the host program (GMAP's dynprog.o,
 * pairpool.o): pairdistance_array, consistent_array, Dynprog_T workspaces
 * from Dynprog_new, and the three Pairpool_* functions.
 *
 * The types below are ABI-compatible declarations of the reference's
 * (dynprog.h:81-90, bool.h, list.h, pairpool.h, genome.h, univcoord.h).
 * Translation units that already include the reference's headers should
 * include dynprog_simd.h instead; the symbols are the same.
 */
#ifndef GMAPDP_DYNPROG_H
#define GMAPDP_DYNPROG_H

#ifndef DYNPROG_INCLUDED
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef char Score8_T;
typedef char Direction8_T;
typedef short Score16_T;
typedef short Direction16_T;
typedef enum { HIGHQ, MEDQ, LOWQ, ENDQ } Mismatchtype_T;
/* Opaque handles (pointers in the reference too); the tags differ from the
   reference's so the names stay usable from C++. */
typedef struct gdp_ref_dynprog *Dynprog_T;
typedef struct gdp_ref_list *List_T;
typedef struct gdp_ref_pairpool *Pairpool_T;
typedef struct gdp_ref_genome *Genome_T;
/* Univcoord_T is 64 bits in GMAP's large-genome builds (gmapl, LARGE_GENOMES,
   univcoord.h:10-12): libgmapdp_dynprog_large.so is this library built with
   GDP_LARGE_GENOMES for them. */
#if defined(LARGE_GENOMES) || defined(GDP_LARGE_GENOMES)
typedef unsigned long long Univcoord_T;
#else
typedef unsigned int Univcoord_T;
#endif

/* Dynprog_simd_16 without device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same arguments, result and tables as Dynprog_simd_16, and
   Dynprog_traceback_16 walks its tables as it walks Dynprog_simd_16's; but
   the fill's device buffers (descriptor, sequences, direction words, scores,
   the exported tables and the traceback's requests and pairs) are parts of
   one fixed array of GDP_NOALLOC_SCRATCH_BYTES in the library's device
   image, one per GPU, held until the process ends; the call makes no
   hipMalloc or hipFree.  A problem whose buffers do not fit (for example rl
   and gl of 1000 with any band fit) gets NULL, with the reason in
   gdp_last_error() (include/gmapdp.h).  Every call on one GPU shares that
   GPU's array: a fill's tables may be walked until the next call of this
   function on the same GPU, and calls that overlap on one GPU (threads
   sharing a device) race; serialize them.  Nothing in the array needs to
   start at zero (every byte read is written by the same call first). */
#define GDP_NOALLOC_SCRATCH_BYTES (16u << 20)
Score16_T **Dynprog_simd_16_noalloc(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                                    Direction16_T ***directions_Fgap, Dynprog_T dynprog, char *rsequence,
                                    char *gsequence, char *gsequence_alt, int rlength, int glength,
                                    Mismatchtype_T mismatchtype, int open, int extend, int lband, int uband,
                                    bool jump_late_p, bool revp);

#ifdef __cplusplus
}
#endif
#endif /* DYNPROG_INCLUDED */

#ifdef __cplusplus
extern "C" {
#endif

/* Full-band fills: Dynprog_simd_8 (dynprog_simd.c:2986), Dynprog_simd_16 (6561). */
Score8_T **Dynprog_simd_8(Direction8_T ***directions_nogap, Direction8_T ***directions_Egap,
                          Direction8_T ***directions_Fgap, Dynprog_T dynprog, char *rsequence, char *gsequence,
                          char *gsequence_alt, int rlength, int glength, Mismatchtype_T mismatchtype, int open,
                          int extend, int lband, int uband, bool jump_late_p, bool revp);
Score16_T **Dynprog_simd_16(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                            Direction16_T ***directions_Fgap, Dynprog_T dynprog, char *rsequence, char *gsequence,
                            char *gsequence_alt, int rlength, int glength, Mismatchtype_T mismatchtype, int open,
                            int extend, int lband, int uband, bool jump_late_p, bool revp);

/* Upper triangle (genome gaps): Dynprog_simd_8_upper (4303), _16_upper (7713). */
Score8_T **Dynprog_simd_8_upper(Direction8_T ***directions_nogap, Direction8_T ***directions_Egap, Dynprog_T dynprog,
                                char *rsequence, char *gsequence, char *gsequence_alt, int rlength, int glength,
                                Mismatchtype_T mismatchtype, int open, int extend, int uband, bool jump_late_p,
                                bool revp);
Score16_T **Dynprog_simd_16_upper(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                                  Dynprog_T dynprog, char *rsequence, char *gsequence, char *gsequence_alt,
                                  int rlength, int glength, Mismatchtype_T mismatchtype, int open, int extend,
                                  int uband, bool jump_late_p, bool revp);

/* Lower triangle (query gaps): Dynprog_simd_8_lower (5339), _16_lower (8585). */
Score8_T **Dynprog_simd_8_lower(Direction8_T ***directions_nogap, Direction8_T ***directions_Egap, Dynprog_T dynprog,
                                char *rsequence, char *gsequence, char *gsequence_alt, int rlength, int glength,
                                Mismatchtype_T mismatchtype, int open, int extend, int lband, bool jump_late_p,
                                bool revp);
Score16_T **Dynprog_simd_16_lower(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                                  Dynprog_T dynprog, char *rsequence, char *gsequence, char *gsequence_alt,
                                  int rlength, int glength, Mismatchtype_T mismatchtype, int open, int extend,
                                  int lband, bool jump_late_p, bool revp);

/* Tracebacks: Dynprog_traceback_8 (9153), _16 (9552), _8_upper (9318),
   _16_upper (9715), _8_lower (9438), _16_lower (9835).  The direction
   tables must be the ones returned by the matching fill above. */
List_T Dynprog_traceback_8(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches, int *nopens,
                           int *nindels, Direction8_T **directions_nogap, Direction8_T **directions_Egap,
                           Direction8_T **directions_Fgap, int r, int c, char *rsequence, char *rsequenceuc,
                           char *gsequence, char *gsequence_alt, int queryoffset, int genomeoffset, Genome_T genome,
                           Genome_T genomealt, Pairpool_T pairpool, bool revp, Univcoord_T chroffset,
                           Univcoord_T chrhigh, bool watsonp, int genestrand, int dynprogindex);
List_T Dynprog_traceback_16(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches, int *nopens,
                            int *nindels, Direction16_T **directions_nogap, Direction16_T **directions_Egap,
                            Direction16_T **directions_Fgap, int r, int c, char *rsequence, char *rsequenceuc,
                            char *gsequence, char *gsequence_alt, int queryoffset, int genomeoffset, Genome_T genome,
                            Genome_T genomealt, Pairpool_T pairpool, bool revp, Univcoord_T chroffset,
                            Univcoord_T chrhigh, bool watsonp, int genestrand, int dynprogindex);
List_T Dynprog_traceback_8_upper(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches, int *nopens,
                                 int *nindels, Direction8_T **directions_nogap, Direction8_T **directions_Egap, int r,
                                 int c, char *rsequence, char *rsequenceuc, char *gsequence, char *gsequence_alt,
                                 int queryoffset, int genomeoffset, Genome_T genome, Genome_T genomealt,
                                 Pairpool_T pairpool, bool revp, Univcoord_T chroffset, Univcoord_T chrhigh,
                                 bool watsonp, int genestrand, int dynprogindex);
List_T Dynprog_traceback_16_upper(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches, int *nopens,
                                  int *nindels, Direction16_T **directions_nogap, Direction16_T **directions_Egap,
                                  int r, int c, char *rsequence, char *rsequenceuc, char *gsequence,
                                  char *gsequence_alt, int queryoffset, int genomeoffset, Genome_T genome,
                                  Genome_T genomealt, Pairpool_T pairpool, bool revp, Univcoord_T chroffset,
                                  Univcoord_T chrhigh, bool watsonp, int genestrand, int dynprogindex);
List_T Dynprog_traceback_8_lower(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches, int *nopens,
                                 int *nindels, Direction8_T **directions_nogap, Direction8_T **directions_Egap, int r,
                                 int c, char *rsequence, char *rsequenceuc, char *gsequence, char *gsequence_alt,
                                 int queryoffset, int genomeoffset, Pairpool_T pairpool, int genestrand, bool revp,
                                 int dynprogindex);
List_T Dynprog_traceback_16_lower(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches, int *nopens,
                                  int *nindels, Direction16_T **directions_nogap, Direction16_T **directions_Egap,
                                  int r, int c, char *rsequence, char *rsequenceuc, char *gsequence,
                                  char *gsequence_alt, int queryoffset, int genomeoffset, Pairpool_T pairpool,
                                  int genestrand, bool revp, int dynprogindex);

/* Dynprog_simd_16 without device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Same arguments, result and tables as Dynprog_simd_16, and
   Dynprog_traceback_16 walks its tables as it walks Dynprog_simd_16's; but
   the fill's device buffers (descriptor, sequences, direction words, scores,
   the exported tables and the traceback's requests and pairs) are parts of
   one fixed array of GDP_NOALLOC_SCRATCH_BYTES in the library's device
   image, one per GPU, held until the process ends; the call makes no
   hipMalloc or hipFree.  A problem whose buffers do not fit (for example rl
   and gl of 1000 with any band fit) gets NULL, with the reason in
   gdp_last_error() (include/gmapdp.h).  Every call on one GPU shares that
   GPU's array: a fill's tables may be walked until the next call of this
   function on the same GPU, and calls that overlap on one GPU (threads
   sharing a device) race; serialize them.  Nothing in the array needs to
   start at zero (every byte read is written by the same call first). */
#define GDP_NOALLOC_SCRATCH_BYTES (16u << 20)
Score16_T **Dynprog_simd_16_noalloc(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                                    Direction16_T ***directions_Fgap, Dynprog_T dynprog, char *rsequence,
                                    char *gsequence, char *gsequence_alt, int rlength, int glength,
                                    Mismatchtype_T mismatchtype, int open, int extend, int lband, int uband,
                                    bool jump_late_p, bool revp);

#ifdef __cplusplus
}
#endif
#endif
