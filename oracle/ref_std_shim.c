/*
 * ref_std_shim.c -- thin C harness over the reference's scalar DP build.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
 * reference's dynprog.c, pairpool.c, list.c, mem.c, except.c and assert.c,
 * read in place from the reference tree with no SIMD configuration (the
 * reference's *.nosimd builds), into oracle/_ref_std/libgmapref_std.so.
 * There dynprog.c provides Dynprog_standard (dynprog.c:1264-1785) and
 * Dynprog_traceback_std (1794-1947), which this file exposes over plain
 * arrays.  Genome_get_char is never reached: tracebacks run with chrhigh 0
 * (genome skips give '*'), and fills always receive the genome sequence.
 */
#include <stdlib.h>
#include <string.h>

#include "dynprog.h"
#include "list.h"
#include "pairdef.h"
#include "pairpool.h"

static Pairpool_T pool = NULL;
static Direction32_T **last_nogap, **last_Egap, **last_Fgap;

void refs_init(int mode) {
  Dynprog_init((Mode_T)mode);
  if (pool == NULL) pool = Pairpool_new();
}

void *refs_new(int maxrl, int maxgl) {
  int extra = maxgl > maxrl ? maxgl - maxrl : 0;
  return (void *)Dynprog_new(/*maxlookback*/ maxrl, /*extraquerygap*/ extra, /*maxpeelback*/ 0,
                             /*extramaterial_end*/ 0, /*extramaterial_paired*/ 0, false);
}

static unsigned char enc(int nogap, int egap, int fgap) {
  unsigned char d = 0;
  if (nogap == HORIZ) d |= 1;
  else if (nogap != DIAG) d |= 2;
  if (egap != DIAG) d |= 4;
  if (fgap != DIAG) d |= 8;
  return d;
}

/* Dense copies, [c][r] with stride rl + 1: scores as int and directions in
   the nibble encoding of the oracle (bit0 HORIZ, bit1 VERT, bit2 Egap, bit3
   Fgap). */
void refs_fill(void *ws, char *rseq, char *gseq, char *galt, int rl, int gl, int mt, int open, int extend,
               int lband, int uband, int late, int revp, int saturation, int upperp, int lowerp, int *H,
               unsigned char *dirs) {
  Direction32_T **no, **eg, **fg;
  Score32_T **m = Dynprog_standard(&no, &eg, &fg, (Dynprog_T)ws, rseq, gseq, galt, rl, gl, /*goffset*/ 0, NULL,
                                   NULL, 0, 0, true, (Mismatchtype_T)mt, open, extend, lband, uband,
                                   late ? true : false, revp ? true : false, saturation, upperp ? true : false,
                                   lowerp ? true : false);
  for (int c = 0; c <= gl; c++)
    for (int r = 0; r <= rl; r++) {
      size_t k = (size_t)c * (rl + 1) + r;
      H[k] = m[c][r];
      dirs[k] = enc(no[c][r], eg[c][r], fg[c][r]);
    }
  last_nogap = no;
  last_Egap = eg;
  last_Fgap = fg;
}

/* Dynprog_traceback_std over the last fill, flattened in push order as the
   SIMD harness does: 8 ints per pair. */
int refs_traceback(int r, int c, char *rseq, char *rsequc, char *gseq, char *galt, int queryoffset,
                   int genomeoffset, int revp, int genestrand, int *counters, int *out, int maxpairs) {
  List_T pairs, p;
  int sc = 0, nm = 0, nmm = 0, no = 0, ni = 0;
  Pairpool_reset(pool);
  pairs = Dynprog_traceback_std(NULL, &sc, &nm, &nmm, &no, &ni, last_nogap, last_Egap, last_Fgap, r, c, rseq, rsequc,
                                gseq, galt, queryoffset, genomeoffset, NULL, NULL, pool, revp ? true : false, 0, 0,
                                true, genestrand, 0);
  counters[0] += sc; counters[1] += nm; counters[2] += nmm; counters[3] += no; counters[4] += ni;
  int n = 0;
  for (p = pairs; p != NULL; p = p->rest) n++;
  int k = n - 1; /* list head is the last push */
  for (p = pairs; p != NULL; p = p->rest, k--) {
    if (k >= maxpairs) continue;
    Pair_T pr = (Pair_T)p->first;
    int *o = out + (size_t)k * 8;
    o[0] = pr->querypos; o[1] = (int)pr->genomepos; o[2] = pr->cdna; o[3] = pr->comp;
    o[4] = pr->genome; o[5] = pr->genomealt; o[6] = pr->gapp ? 1 : 0; o[7] = pr->genomejump;
  }
  return n;
}
