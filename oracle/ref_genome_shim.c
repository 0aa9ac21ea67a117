/*
 * ref_genome_shim.c -- reaches the reference's intron bridge.
 *
 * TEST INFRASTRUCTURE ONLY.  bridge_intron_gap_{8,16}_ud are static functions
 * of the reference's dynprog_genome.c, so this file compiles that source in
 * place (an #include of the file as it lies in the reference tree; the
 * Makefile puts that tree on the include path) and exposes the bridge over
 * plain arrays.  The splice-site probabilities come from the reference's
 * maxent_hr.c (compiled in place beside it) over a Genome_T that wraps a
 * genomecomp image, as GMAP's stage 3 reads them.  No splicing IIT is set
 * (novel splicing, no known sites), so the bridge takes its site-level path.
 */
#include "dynprog_genome.c"

static struct Genome_T refg_genome_s;
static Genome_T refg_genome = NULL;

/* Intron reward tables and caller settings as GMAP sets them without a
   splicing IIT (Dynprog_genome_setup, dynprog_genome.c:191-216), and the
   genome the probabilities read: the genomecomp image `blocks` of `length`
   bases, kept by the caller while calls run. */
void refg_setup(const unsigned int *blocks, unsigned int length) {
  Dynprog_genome_setup(/*novelsplicingp*/ true, /*splicing_iit*/ NULL, /*crosstable*/ NULL,
                       /*donor_typeint*/ -1, /*acceptor_typeint*/ -1, /*user_open*/ 0, /*user_extend*/ 0,
                       /*user_dynprog_p*/ false);
  memset(&refg_genome_s, 0, sizeof refg_genome_s);
  refg_genome_s.blocks = (Genomecomp_T *)blocks;
  refg_genome_s.genomelength = length;
  refg_genome = &refg_genome_s;
}

/* Per column, the probability the site-level bridge reads for a splice
   site there (left_probabilities / right_probabilities,
   dynprog_genome.c:1845-1933): the four Maxent_hr models at the positions
   those loops use, chosen by strand and cdna_direction.  No site is known,
   so none is 1.0.  Entries glength-1 are left 0. */
void refg_site_probs(double *left, double *right, int glengthL, int glengthR, unsigned int leftoffset,
                     unsigned int rightoffset, unsigned int chroffset, unsigned int chrhigh, int cdna_direction,
                     int watsonp) {
  Genome_T g = refg_genome;
  for (int c = 0; c < glengthL; c++) left[c] = 0.0;
  for (int c = 0; c < glengthR; c++) right[c] = 0.0;
  for (int c = 0; c < glengthL - 1; c++) {
    if (watsonp) {
      Univcoord_T pos = chroffset + leftoffset + c;
      left[c] = cdna_direction > 0 ? Maxent_hr_donor_prob(g, g, pos, chroffset)
                                   : Maxent_hr_antiacceptor_prob(g, g, pos, chroffset);
    } else {
      Univcoord_T pos = chrhigh - leftoffset - c + 1;
      left[c] = cdna_direction > 0 ? Maxent_hr_antidonor_prob(g, g, pos, chroffset)
                                   : Maxent_hr_acceptor_prob(g, g, pos, chroffset);
    }
  }
  for (int c = 0; c < glengthR - 1; c++) {
    if (watsonp) {
      Univcoord_T pos = chroffset + rightoffset - c + 1;
      right[c] = cdna_direction > 0 ? Maxent_hr_acceptor_prob(g, g, pos, chroffset)
                                    : Maxent_hr_antidonor_prob(g, g, pos, chroffset);
    } else {
      Univcoord_T pos = chrhigh - rightoffset + c;
      right[c] = cdna_direction > 0 ? Maxent_hr_antiacceptor_prob(g, g, pos, chroffset)
                                    : Maxent_hr_donor_prob(g, g, pos, chroffset);
    }
  }
}

static Score16_T **rows16(const short *a, int n, int stride) {
  Score16_T **t = (Score16_T **)malloc(sizeof(Score16_T *) * (size_t)n);
  for (int i = 0; i < n; i++) t[i] = (Score16_T *)(a + (size_t)i * stride);
  return t;
}

static Score8_T **rows8(const short *a, int n, int stride, Score8_T **keep) {
  Score8_T *b = (Score8_T *)malloc((size_t)n * stride);
  Score8_T **t = (Score8_T **)malloc(sizeof(Score8_T *) * (size_t)n);
  for (size_t k = 0; k < (size_t)n * stride; k++) b[k] = (Score8_T)a[k];
  for (int i = 0; i < n; i++) t[i] = b + (size_t)i * stride;
  *keep = b;
  return t;
}

/* bridge_intron_gap_{8,16}_ud (dynprog_genome.c:1387, 2262) as
   Dynprog_genome_gap calls it (3548-3557, 3690-3699): score tables in the oracle's
   layout (upper [c][r] stride rlength+1, lower [r][c] stride glength+1),
   the left segment gsequenceL in genome order and the right segment
   rev_gsequenceR in genome order too (the bridge reads it backwards from its
   last base, as the caller passes &rev_gsequenceR[glengthR-1]).  out =
   {finalscore, bestrL, bestrR, bestcL, bestcR} (cells start at -1),
   probs = {left_prob, right_prob} (0.0 unless finalscore >= 0). */
void refg_bridge_intron(int width, const short *LU, const short *LL, const short *RU, const short *RL, char *gL,
                        char *gLalt, char *gRseg, char *gRsegalt, int rlength, int glengthL, int glengthR,
                        int cdna_direction, int watsonp, int lbandL, int ubandL, int lbandR, int ubandR,
                        unsigned int leftoffset, unsigned int rightoffset, unsigned int chroffset,
                        unsigned int chrhigh, int halfp, int finalp, int jump_late, int *out, double *probs) {
  int brL = -1, brR = -1, bcL = -1, bcR = -1, introntype = NONINTRON, best;
  double lp = 0.0, rp = 0.0;
  char *gR = &gRseg[glengthR - 1], *gRa = &gRsegalt[glengthR - 1];
  if (width == 8) {
    Score8_T *k[4];
    Score8_T **lu = rows8(LU, glengthL + 1, rlength + 1, &k[0]), **ll = rows8(LL, rlength + 1, glengthL + 1, &k[1]);
    Score8_T **ru = rows8(RU, glengthR + 1, rlength + 1, &k[2]), **rl = rows8(RL, rlength + 1, glengthR + 1, &k[3]);
    best = bridge_intron_gap_8_ud(&brL, &brR, &bcL, &bcR, &introntype, &lp, &rp, lu, ll, ru, rl, NULL, NULL, NULL,
                                  NULL, gL, gLalt, gR, gRa, (int)leftoffset, (int)rightoffset, rlength, glengthL,
                                  glengthR, cdna_direction, watsonp ? true : false, lbandL, ubandL, lbandR, ubandR,
                                  (int)leftoffset, (int)rightoffset, refg_genome, refg_genome, /*chrnum*/ 1,
                                  chroffset, chrhigh, halfp ? true : false, finalp ? true : false,
                                  jump_late ? true : false);
    for (int i = 0; i < 4; i++) free(k[i]);
    free(lu);
    free(ll);
    free(ru);
    free(rl);
  } else {
    Score16_T **lu = rows16(LU, glengthL + 1, rlength + 1), **ll = rows16(LL, rlength + 1, glengthL + 1);
    Score16_T **ru = rows16(RU, glengthR + 1, rlength + 1), **rl = rows16(RL, rlength + 1, glengthR + 1);
    best = bridge_intron_gap_16_ud(&brL, &brR, &bcL, &bcR, &introntype, &lp, &rp, lu, ll, ru, rl, NULL, NULL, NULL,
                                   NULL, gL, gLalt, gR, gRa, (int)leftoffset, (int)rightoffset, rlength, glengthL,
                                   glengthR, cdna_direction, watsonp ? true : false, lbandL, ubandL, lbandR, ubandR,
                                   (int)leftoffset, (int)rightoffset, refg_genome, refg_genome, /*chrnum*/ 1,
                                   chroffset, chrhigh, halfp ? true : false, finalp ? true : false,
                                   jump_late ? true : false);
    free(lu);
    free(ll);
    free(ru);
    free(rl);
  }
  out[0] = best;
  out[1] = brL;
  out[2] = brR;
  out[3] = bcL;
  out[4] = bcR;
  probs[0] = lp;
  probs[1] = rp;
}
