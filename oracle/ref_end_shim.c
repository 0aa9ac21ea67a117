/*
 * ref_end_shim.c -- reaches the reference's end-cell scans and nogaps walk.
 *
 * TEST INFRASTRUCTURE ONLY.  find_best_endpoint_{8,16},
 * find_best_endpoint_to_queryend_indels_{8,16} and traceback_nogaps are
 * static functions of the reference's dynprog_end.c, so this file compiles
 * that source in place (an #include of the file as it lies in the reference
 * tree; the Makefile puts that tree on the include path) and exposes them
 * over plain arrays in oracle/_ref/libgmapref.so.  The end-gap callers
 * themselves are never called: they read the genome through
 * Genome_get_segment_*, which stay unresolved (lazy binding).
 */
#include "dynprog_end.c"

static Pairpool_T end_pool = NULL;

/* Row tables of the reference's matrix[c][r] (upper, stride rl+1) and
   matrix[r][c] (lower, stride gl+1) shapes over dense int16 arrays; int8
   copies for width 8 (freed by the caller through *keep). */
static void *rows_of(const short *a, int n, int stride, int width, void **keep) {
  void **t = (void **)malloc(sizeof(void *) * (size_t)n);
  if (width == 8) {
    Score8_T *b = (Score8_T *)malloc((size_t)n * stride);
    for (size_t k = 0; k < (size_t)n * stride; k++) b[k] = (Score8_T)a[k];
    for (int i = 0; i < n; i++) t[i] = b + (size_t)i * stride;
    *keep = b;
  } else {
    for (int i = 0; i < n; i++) t[i] = (void *)(a + (size_t)i * stride);
    *keep = NULL;
  }
  return t;
}

/* out = {finalscore, bestr, bestc} of the scan the end-gap callers run
   (dynprog_end.c:143-290, 358-508). */
void refe_best_endpoint(int queryend_indels, int width, const short *Hupper, const short *Hlower, int rl, int gl,
                        int lband, int uband, int jump_late, int *out) {
  void *ku, *kl;
  void *mu = rows_of(Hupper, gl + 1, rl + 1, width, &ku), *ml = rows_of(Hlower, rl + 1, gl + 1, width, &kl);
  int fs = 0, br = -1, bc = -1;
  bool late = jump_late != 0;
  if (width == 8) {
    if (queryend_indels)
      find_best_endpoint_to_queryend_indels_8(&fs, &br, &bc, (Score8_T **)mu, (Score8_T **)ml, rl, gl, lband, uband,
                                              late);
    else
      find_best_endpoint_8(&fs, &br, &bc, (Score8_T **)mu, (Score8_T **)ml, rl, gl, lband, uband, late);
  } else {
    if (queryend_indels)
      find_best_endpoint_to_queryend_indels_16(&fs, &br, &bc, (Score16_T **)mu, (Score16_T **)ml, rl, gl, lband,
                                               uband, late);
    else
      find_best_endpoint_16(&fs, &br, &bc, (Score16_T **)mu, (Score16_T **)ml, rl, gl, lband, uband, late);
  }
  free(ku);
  free(kl);
  free(mu);
  free(ml);
  out[0] = fs;
  out[1] = br;
  out[2] = bc;
}

int refh_flatten(void *list, int *out, int maxpairs);

/* traceback_nogaps (dynprog_end.c:648-722) from (r, c); pairs in push order
   as refh_flatten writes them, counters += {score, nmatches, nmismatches}.
   Returns the number of pairs. */
int refe_traceback_nogaps(int r, int c, char *rseq, char *rsequc, char *gseq, char *galt, int queryoffset,
                          int genomeoffset, int revp, int genestrand, int *counters, int *out, int maxpairs) {
  int sc = 0, nm = 0, nmm = 0;
  if (end_pool == NULL) end_pool = Pairpool_new();
  Pairpool_reset(end_pool);
  List_T pairs = traceback_nogaps(NULL, &sc, &nm, &nmm, r, c, rseq, rsequc, gseq, galt, queryoffset, genomeoffset,
                                  end_pool, genestrand, revp != 0, 0);
  counters[0] += sc;
  counters[1] += nm;
  counters[2] += nmm;
  return refh_flatten(pairs, out, maxpairs);
}

void refh_last_tables(int kind, void ***nogap, void ***egap);

/* traceback_local_{8,16}_{upper,lower} (dynprog_end.c:727-1135) over the
   last upper (kind 1) or lower (kind 2) fill of ref_shim.c, from (r, c) to
   genome column endc; pairs in push order, counters += {score, nmatches,
   nmismatches, nopens, nindels}, end_rc = where the walk stopped. */
int refe_traceback_local(int kind, int width, int r, int c, int endc, char *rseq, char *rsequc, char *gseq,
                         char *galt, int queryoffset, int genomeoffset, int revp, int genestrand, int *counters,
                         int *out, int maxpairs, int *end_rc) {
  int sc = 0, nm = 0, nmm = 0, no = 0, ni = 0;
  void **nogap, **egap;
  List_T pairs;
  if (end_pool == NULL) end_pool = Pairpool_new();
  Pairpool_reset(end_pool);
  refh_last_tables(kind, &nogap, &egap);
  if (kind == 1) {
    if (width == 8)
      pairs = traceback_local_8_upper(NULL, &sc, &nm, &nmm, &no, &ni, (Direction8_T **)nogap, (Direction8_T **)egap,
                                      &r, &c, endc, rseq, rsequc, gseq, galt, queryoffset, genomeoffset, NULL, NULL,
                                      end_pool, revp != 0, 0, 0, true, genestrand, 0);
    else
      pairs = traceback_local_16_upper(NULL, &sc, &nm, &nmm, &no, &ni, (Direction16_T **)nogap,
                                       (Direction16_T **)egap, &r, &c, endc, rseq, rsequc, gseq, galt, queryoffset,
                                       genomeoffset, NULL, NULL, end_pool, revp != 0, 0, 0, true, genestrand, 0);
  } else {
    if (width == 8)
      pairs = traceback_local_8_lower(NULL, &sc, &nm, &nmm, &no, &ni, (Direction8_T **)nogap, (Direction8_T **)egap,
                                      &r, &c, endc, rseq, rsequc, gseq, galt, queryoffset, genomeoffset, end_pool,
                                      genestrand, revp != 0, 0);
    else
      pairs = traceback_local_16_lower(NULL, &sc, &nm, &nmm, &no, &ni, (Direction16_T **)nogap,
                                       (Direction16_T **)egap, &r, &c, endc, rseq, rsequc, gseq, galt, queryoffset,
                                       genomeoffset, end_pool, genestrand, revp != 0, 0);
  }
  counters[0] += sc;
  counters[1] += nm;
  counters[2] += nmm;
  counters[3] += no;
  counters[4] += ni;
  end_rc[0] = r;
  end_rc[1] = c;
  return refh_flatten(pairs, out, maxpairs);
}

/* Dynprog_end5_splicejunction (dynprog_end.c:1652-1922) on workspace ws
   (a doublep Dynprog_T of ref_shim.c), sequences as the caller passes them
   (read backwards from the pointers).  out: pairs in push order of the
   returned list; res = {traceback_score, missscore, nmatches, nmismatches,
   nopens, nindels}.  Returns the number of pairs, or -1 for a NULL list. */
int refe_end5_splicejunction(void *ws, char *rev_rseq, char *rev_rsequc, char *rev_gseq, char *rev_galt, int rlength,
                             int glength, int rev_roffset, int rev_goffset_anchor, int rev_goffset_far,
                             int genestrand, int jump_late, int extraband_end, double defect_rate, int contlength,
                             int *res, int *out, int maxpairs) {
  int dpi = 1, ts = 0, miss = 0, nm = 0, nmm = 0, no = 0, ni = 0;
  if (end_pool == NULL) end_pool = Pairpool_new();
  Pairpool_reset(end_pool);
  List_T pairs = Dynprog_end5_splicejunction(&dpi, &ts, &miss, &nm, &nmm, &no, &ni, (Dynprog_T)ws, rev_rseq,
                                             rev_rsequc, rev_gseq, rev_galt, rlength, glength, rev_roffset,
                                             rev_goffset_anchor, rev_goffset_far, 0, 0, true, genestrand,
                                             jump_late != 0, NULL, NULL, end_pool, extraband_end, defect_rate,
                                             contlength);
  res[0] = ts;
  res[1] = miss;
  res[2] = nm;
  res[3] = nmm;
  res[4] = no;
  res[5] = ni;
  if (pairs == NULL) return -1;
  return refh_flatten(pairs, out, maxpairs);
}

/* Dynprog_end3_splicejunction (dynprog_end.c:2248-2745): as
   refe_end5_splicejunction, sequences forward from the pointers. */
int refe_end3_splicejunction(void *ws, char *rseq, char *rsequc, char *gseq, char *galt, int rlength, int glength,
                             int roffset, int goffset_anchor, int goffset_far, int genestrand, int jump_late,
                             int extraband_end, double defect_rate, int contlength, int *res, int *out,
                             int maxpairs) {
  int dpi = 1, ts = 0, miss = 0, nm = 0, nmm = 0, no = 0, ni = 0;
  if (end_pool == NULL) end_pool = Pairpool_new();
  Pairpool_reset(end_pool);
  List_T pairs = Dynprog_end3_splicejunction(&dpi, &ts, &miss, &nm, &nmm, &no, &ni, (Dynprog_T)ws, rseq, rsequc,
                                             gseq, galt, rlength, glength, roffset, goffset_anchor, goffset_far, 0,
                                             0, true, genestrand, jump_late != 0, NULL, NULL, end_pool,
                                             extraband_end, defect_rate, contlength);
  res[0] = ts;
  res[1] = miss;
  res[2] = nm;
  res[3] = nmm;
  res[4] = no;
  res[5] = ni;
  if (pairs == NULL) return -1;
  return refh_flatten(pairs, out, maxpairs);
}
