/*
 * ref_callers_shim.c -- the reference's own DP callers over plain arrays.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's
 * dynprog_single.c, dynprog_end.c, dynprog_cdna.c and dynprog_genome.c
 * (the last two through ref_cdna_shim.c and ref_genome_shim.c) in place
 * together with dynprog.c, pairpool.c, genome.c, maxent_hr.c, pair.c and
 * their helpers, twice:
 *   _ref/libgmapref_callers.so  with the reference's dynprog_simd.o (AVX2);
 *   _ref/libgmapref_callers_gpu.so with libgmapdp_dynprog.so in place of
 *                               dynprog_simd.o (the drop-in, on the GPU).
 * The same calls through both show that the callers, unchanged, get the
 * same pair lists and counters from the drop-in as from the reference's
 * SIMD code.  Genome skips and segments read a Genome_T over a genomecomp
 * image through the reference's genome.c.
 */
#include <stdlib.h>
#include <string.h>

#include "dynprog.h"
#include "dynprog_cdna.h"
#include "dynprog_end.h"
#include "dynprog_genome.h"
#include "dynprog_single.h"
#include "genome.h"
#include "list.h"
#include "pairdef.h"
#include "pairpool.h"

static Pairpool_T pool = NULL;
static struct Genome_T genome_s;
static Genome_T genome = NULL;

/* Tables, caller settings (no user penalties, no homopolymer mode, novel
   splicing without a splicing IIT) and one genome: the genomecomp image `blocks` of `length`
   bases, kept by the caller while calls run. */
void refc_init(int mode, const unsigned int *blocks, unsigned int length) {
  Dynprog_init((Mode_T)mode);
  Dynprog_single_setup(/*user_open*/ 0, /*user_extend*/ 0, /*user_dynprog_p*/ false, /*homopolymerp*/ false);
  Dynprog_end_setup(NULL, NULL, NULL, 0, NULL, NULL, NULL, NULL, 0, 0, false);
  Dynprog_genome_setup(/*novelsplicingp*/ true, /*splicing_iit*/ NULL, NULL, -1, -1, 0, 0, false);
  if (pool == NULL) pool = Pairpool_new();
  memset(&genome_s, 0, sizeof genome_s);
  genome_s.blocks = (Genomecomp_T *)blocks;
  genome_s.genomelength = length;
  genome = &genome_s;
}

/* A workspace as GMAP's stage 3 makes one (dynprog.c:629-877). */
void *refc_new(int maxlookback, int extraquerygap, int maxpeelback, int extramaterial_end, int doublep) {
  return (void *)Dynprog_new(maxlookback, extraquerygap, maxpeelback, extramaterial_end,
                             /*extramaterial_paired*/ 0, doublep ? true : false);
}

/* The returned list from its head: 8 ints per pair {querypos, genomepos,
   cdna, comp, genome, genomealt, gapp (2: known-site gap holder),
   genomejump}. */
static int flatten(List_T pairs, int *out, int maxpairs) {
  int n = 0;
  for (List_T p = pairs; p != NULL; p = p->rest, n++) {
    if (n >= maxpairs) continue;
    Pair_T pr = (Pair_T)p->first;
    int *o = out + (size_t)n * 8;
    o[0] = pr->querypos; o[1] = (int)pr->genomepos; o[2] = pr->cdna; o[3] = pr->comp;
    o[4] = pr->genome; o[5] = pr->genomealt; o[6] = pr->gapp ? (pr->knowngapp ? 2 : 1) : 0; o[7] = pr->genomejump;
  }
  return n;
}

/* Dynprog_single_gap (dynprog_single.c:428-675).  counters = {dynprogindex
   after the call, traceback_score, nmatches, nmismatches, nopens, nindels}.
   Returns the number of pairs of the returned list, -1 for NULL. */
int refc_single_gap(void *ws, char *rseq, char *rsequc, int rl, int gl, int roffset, int goffset,
                    unsigned int chroffset, unsigned int chrhigh, int watsonp, int genestrand, int jump_late,
                    int extraband, int widebandp, double defect_rate, int *counters, int *out, int maxpairs) {
  int idx = 1, sc = 0, nm = 0, nmm = 0, no = 0, ni = 0;
  Pairpool_reset(pool);
  List_T pairs = Dynprog_single_gap(&idx, &sc, &nm, &nmm, &no, &ni, (Dynprog_T)ws, rseq, rsequc, rl, gl, roffset,
                                    goffset, chroffset, chrhigh, watsonp ? true : false, genestrand,
                                    jump_late ? true : false, genome, genome, pool, extraband,
                                    widebandp ? true : false, defect_rate);
  counters[0] = idx; counters[1] = sc; counters[2] = nm; counters[3] = nmm; counters[4] = no; counters[5] = ni;
  return pairs == NULL ? -1 : flatten(pairs, out, maxpairs);
}

/* Dynprog_end5_gap (dynprog_end.c:1293-1650) when five, else
   Dynprog_end3_gap (1923-2243): rseq points at the read's first base; the
   5' form receives the pointer to its last base, as stage 3 passes it.
   endalign is Endalign_T (dynprog.h:24).  counters and the return value as
   refc_single_gap. */
int refc_end_gap(int five, void *ws, char *rseq, char *rsequc, int rl, int gl, int roffset, int goffset,
                 unsigned int chroffset, unsigned int chrhigh, int watsonp, int genestrand, int jump_late,
                 int extraband, double defect_rate, int endalign, int *counters, int *out, int maxpairs) {
  int idx = 1, sc = 0, nm = 0, nmm = 0, no = 0, ni = 0;
  List_T pairs;
  Pairpool_reset(pool);
  if (five)
    pairs = Dynprog_end5_gap(&idx, &sc, &nm, &nmm, &no, &ni, (Dynprog_T)ws, &rseq[rl - 1], &rsequc[rl - 1], rl, gl,
                             roffset, goffset, chroffset, chrhigh, watsonp ? true : false, genestrand,
                             jump_late ? true : false, genome, genome, pool, extraband, defect_rate,
                             (Endalign_T)endalign, /*require_pos_score_p*/ false);
  else
    pairs = Dynprog_end3_gap(&idx, &sc, &nm, &nmm, &no, &ni, (Dynprog_T)ws, rseq, rsequc, rl, gl, roffset, goffset,
                             chroffset, chrhigh, watsonp ? true : false, genestrand, jump_late ? true : false,
                             genome, genome, pool, extraband, defect_rate, (Endalign_T)endalign,
                             /*require_pos_score_p*/ false);
  counters[0] = idx; counters[1] = sc; counters[2] = nm; counters[3] = nmm; counters[4] = no; counters[5] = ni;
  return pairs == NULL ? -1 : flatten(pairs, out, maxpairs);
}

/* Dynprog_genome_gap (dynprog_genome.c:3287-3770) with two workspaces, as
   stage 3 calls it (stage3.c:9567-9576): rseq points at the gap's first read
   base.  counters = {dynprogindex after the call, new_leftgenomepos,
   new_rightgenomepos, traceback_score, nmatches, nmismatches, nopens,
   nindels, exonhead, introntype}; probs = {left_prob, right_prob}.  Returns
   the number of pairs of the returned list (its gap holder carries the
   intron), -1 for NULL. */
int refc_genome_gap(void *wsL, void *wsR, char *rseq, char *rsequc, int rl, int glL, int glR, int roffset,
                    int goffsetL, int rev_goffsetR, unsigned int chroffset, unsigned int chrhigh, int cdna_direction,
                    int watsonp, int genestrand, int jump_late, int extraband, double defect_rate, int maxpeelback,
                    int halfp, int finalp, int *counters, double *probs, int *out, int maxpairs) {
  int idx = 1, nl = 0, nr = 0, sc = 0, nm = 0, nmm = 0, no = 0, ni = 0, eh = 0, it = 0;
  double lp = 0.0, rp = 0.0;
  Pairpool_reset(pool);
  List_T pairs = Dynprog_genome_gap(&idx, &nl, &nr, &lp, &rp, &sc, &nm, &nmm, &no, &ni, &eh, &it, (Dynprog_T)wsL,
                                    (Dynprog_T)wsR, rseq, rsequc, rl, glL, glR, roffset, goffsetL, rev_goffsetR,
                                    /*chrnum*/ 1, chroffset, chrhigh, cdna_direction, watsonp ? true : false,
                                    genestrand, jump_late ? true : false, genome, genome, pool, extraband,
                                    defect_rate, maxpeelback, halfp ? true : false, finalp ? true : false);
  counters[0] = idx; counters[1] = nl; counters[2] = nr; counters[3] = sc; counters[4] = nm;
  counters[5] = nmm; counters[6] = no; counters[7] = ni; counters[8] = eh; counters[9] = it;
  probs[0] = lp;
  probs[1] = rp;
  return pairs == NULL ? -1 : flatten(pairs, out, maxpairs);
}

/* Dynprog_cdna_gap (dynprog_cdna.c:786-1180) with two workspaces, as stage 3
   calls it (stage3.c:9310-9321): the left end reads query[qdp5 ...] forward
   and the right end query[qdp3] backwards, rlengthL and rlengthR bases, over
   the glength genome columns from goffset.  counters = {dynprogindex after
   the call, traceback_score, incompletep}.  Returns the number of pairs of
   the returned list, -1 for NULL. */
int refc_cdna_gap(void *wsL, void *wsR, char *query, char *queryuc, int qdp5, int qdp3, int rlengthL, int rlengthR,
                  int glength, int goffset, unsigned int chroffset, unsigned int chrhigh, int watsonp, int genestrand,
                  int jump_late, int extraband, double defect_rate, int *counters, int *out, int maxpairs) {
  int idx = 1, sc = 0;
  bool incomplete = false;
  Pairpool_reset(pool);
  List_T pairs = Dynprog_cdna_gap(&idx, &sc, &incomplete, (Dynprog_T)wsL, (Dynprog_T)wsR, &query[qdp5],
                                  &queryuc[qdp5], &query[qdp3], &queryuc[qdp3], rlengthL, rlengthR, glength, qdp5,
                                  qdp3, goffset, chroffset, chrhigh, watsonp ? true : false, genestrand,
                                  jump_late ? true : false, genome, genome, pool, extraband, defect_rate);
  counters[0] = idx;
  counters[1] = sc;
  counters[2] = incomplete ? 1 : 0;
  return pairs == NULL ? -1 : flatten(pairs, out, maxpairs);
}
