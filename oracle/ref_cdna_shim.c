/*
 * ref_cdna_shim.c -- reaches the reference's cDNA gap bridge.
 *
 * TEST INFRASTRUCTURE ONLY.  bridge_cdna_gap_{8,16}_ud are static functions
 * of the reference's dynprog_cdna.c, so this file compiles that source in
 * place (an #include of the file as it lies in the reference tree; the
 * Makefile puts that tree on the include path) and exposes the bridge over
 * plain arrays into oracle/_ref/libgmapref.so.  Dynprog_cdna_gap itself is
 * never called: it reads the genome through Genome_get_segment_*, which stay
 * unresolved (the library is loaded with lazy binding).
 */
#include "dynprog_cdna.c"

/* Score tables of the reference's matrix[c][r] (upper) and matrix[r][c]
   (lower) shape over dense int16 arrays of this project's oracle layout:
   upper [c][r] stride rl+1, lower [r][c] stride gl+1. */
static void *rows16(const short *a, int n, int stride) {
  Score16_T **t = (Score16_T **)malloc(sizeof(Score16_T *) * (size_t)n);
  for (int i = 0; i < n; i++) t[i] = (Score16_T *)(a + (size_t)i * stride);
  return t;
}

static void *rows8(const short *a, int n, int stride, Score8_T **keep) {
  Score8_T *b = (Score8_T *)malloc((size_t)n * stride);
  Score8_T **t = (Score8_T **)malloc(sizeof(Score8_T *) * (size_t)n);
  for (size_t k = 0; k < (size_t)n * stride; k++) b[k] = (Score8_T)a[k];
  for (int i = 0; i < n; i++) t[i] = b + (size_t)i * stride;
  *keep = b;
  return t;
}

/* out = {bestscore, bestcL, bestcR, bestrL, bestrR}; the cells start at -1
   since the reference leaves them unset when nothing reaches the initial
   score. */
void refc_bridge_cdna(int width, const short *LU, const short *LL, const short *RU, const short *RL, int gl,
                      int rlL, int rlR, int lbandL, int ubandL, int lbandR, int ubandR, int open, int extend,
                      int leftoffset, int rightoffset, int jump_late, int *out) {
  int bcL = -1, bcR = -1, brL = -1, brR = -1, best;
  void *lu, *ll, *ru, *rl;
  Score8_T *k[4] = {NULL, NULL, NULL, NULL};
  if (width == 8) {
    lu = rows8(LU, gl + 1, rlL + 1, &k[0]);
    ll = rows8(LL, rlL + 1, gl + 1, &k[1]);
    ru = rows8(RU, gl + 1, rlR + 1, &k[2]);
    rl = rows8(RL, rlR + 1, gl + 1, &k[3]);
    best = bridge_cdna_gap_8_ud(&bcL, &bcR, &brL, &brR, (Score8_T **)lu, (Score8_T **)ll, (Score8_T **)ru,
                                (Score8_T **)rl, gl, rlL, rlR, lbandL, ubandL, lbandR, ubandR, open, extend,
                                leftoffset, rightoffset, jump_late != 0);
  } else {
    lu = rows16(LU, gl + 1, rlL + 1);
    ll = rows16(LL, rlL + 1, gl + 1);
    ru = rows16(RU, gl + 1, rlR + 1);
    rl = rows16(RL, rlR + 1, gl + 1);
    best = bridge_cdna_gap_16_ud(&bcL, &bcR, &brL, &brR, (Score16_T **)lu, (Score16_T **)ll, (Score16_T **)ru,
                                 (Score16_T **)rl, gl, rlL, rlR, lbandL, ubandL, lbandR, ubandR, open, extend,
                                 leftoffset, rightoffset, jump_late != 0);
  }
  for (int i = 0; i < 4; i++) free(k[i]);
  free(lu);
  free(ll);
  free(ru);
  free(rl);
  out[0] = best;
  out[1] = bcL;
  out[2] = bcR;
  out[3] = brL;
  out[4] = brR;
}
