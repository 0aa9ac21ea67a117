// gdp_tables.cpp -- GMAP's DP score tables, restated from dynprog.c.
//
// pairdistance[mt][query][genome]: mismatch everywhere in 'A'..'z' x 'A'..'y'
// (dynprog.c:1108-1126), then full matches on identical letters, the
// IUPAC half matches and ambiguity codes (permute_cases, every case
// combination and both orders, 1128-1169), then the mode-specific one-way
// matches (1173-1195).  Everything outside those ranges stays 0.
#include "gdp_tables.h"

#include <cstring>

namespace {
int16_t pd[4][128][128];
int32_t nt2int[128];
uint8_t cons[3][128][128];
int32_t use8p[4];
const int kMismatch[4] = {-3, -2, -1, -5};  // HIGHQ, MEDQ, LOWQ (dynprog.h:48-50), ENDQ (dynprog.c:103)
const char kLetters[5] = {'A', 'C', 'G', 'T', 'N'};

int lower(int ch) { return (ch >= 'A' && ch <= 'Z') ? ch + 32 : ch; }

void both_ways(int a, int b, int16_t score, int mode) {
  const int ca[2] = {a, lower(a)}, cb[2] = {b, lower(b)};
  const bool stranded = mode == 0 || mode == 1 || mode == 3 || mode == 5;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      const int x = ca[i], y = cb[j];
      if (stranded) {
        cons[0][x][y] = cons[0][y][x] = 1;
      } else {
        cons[1][x][y] = cons[1][y][x] = 1;
        cons[2][x][y] = cons[2][y][x] = 1;
      }
      for (int t = 0; t < 4; t++) pd[t][x][y] = pd[t][y][x] = score;
    }
}

void one_way(int a, int b, int16_t score, int genestrand) {
  const int ca[2] = {a, lower(a)}, cb[2] = {b, lower(b)};
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      cons[genestrand][ca[i]][cb[j]] = 1;
      for (int t = 0; t < 4; t++) pd[t][ca[i]][cb[j]] = score;
    }
}
}  // namespace

void gdp_tables_build(int mode) {
  memset(pd, 0, sizeof pd);
  memset(cons, 0, sizeof cons);
  for (int j = 0; j < 128; j++) nt2int[j] = 4;
  nt2int['A'] = nt2int['a'] = 0;
  nt2int['C'] = nt2int['c'] = 1;
  nt2int['G'] = nt2int['g'] = 2;
  nt2int['T'] = nt2int['t'] = 3;
  for (int t = 0; t < 4; t++) use8p[t] = -128 / kMismatch[t] - 1;  // 41, 63, 127, 24
  for (int c1 = 'A'; c1 <= 'z'; c1++)
    for (int c2 = 'A'; c2 < 'z'; c2++)
      for (int t = 0; t < 4; t++) pd[t][c1][c2] = (int16_t)kMismatch[t];
  for (int c = 'A'; c < 'Z'; c++) both_ways(c, c, 3, mode);
  both_ways('U', 'T', 3, mode);
  static const char *half[] = {"RA", "RG", "YT", "YC", "WA", "WT", "SG", "SC", "MA", "MC", "KG", "KT"};
  for (const char *h : half) both_ways(h[0], h[1], 1, mode);
  static const char *amb[] = {"HA", "HT", "HC", "BG", "BC", "BT", "VG", "VA", "VC", "DG", "DA",
                              "DT", "NT", "NC", "NA", "NG", "XT", "XC", "XA", "XG", "NN", "XX"};
  for (const char *a : amb) both_ways(a[0], a[1], 3, mode);
  switch (mode) {
  case 1: one_way('T', 'C', 3, 0); break;
  case 2: one_way('T', 'C', 3, 1); one_way('A', 'G', 3, 2); break;
  case 3: one_way('G', 'A', 3, 0); break;
  case 4: one_way('G', 'A', 3, 1); one_way('C', 'T', 3, 2); break;
  case 5: one_way('C', 'T', 3, 0); break;
  case 6: one_way('C', 'T', 3, 1); one_way('G', 'A', 3, 2); break;
  default: break;
  }
}

void gdp_tables_pack(uint32_t *PQ, uint32_t *PG, int8_t *NT) {
  for (int t = 0; t < 4; t++)
    for (int ch = 0; ch < 128; ch++) {
      uint32_t q = 0, g = 0;
      for (int x = 0; x < 5; x++) {
        q |= (uint32_t)(pd[t][ch][(int)kLetters[x]] & 15) << (4 * x);
        g |= (uint32_t)(pd[t][(int)kLetters[x]][ch] & 15) << (4 * x);
      }
      PQ[t * 128 + ch] = q;
      PG[t * 128 + ch] = g;
    }
  for (int ch = 0; ch < 128; ch++) NT[ch] = (int8_t)nt2int[ch];
}

// Genome characters whose pairdistance columns agree for every query
// character and mismatch type score alike; Dynprog_standard looks scores up
// by the raw characters (dynprog.c:1527-1530), so its fill works on these
// classes.  SC[ch] = class of genome character ch; SP[(t*128 + q)*4 + w]
// packs the 4-bit scores of query q against classes 8w .. 8w+7.  Returns the
// class count, or -1 when the tables need more than 32 classes or a score
// does not fit 4 signed bits.
int gdp_tables_pack_std(uint32_t *SP, int8_t *SC) {
  int ncls = 0, rep[128];
  for (int g = 0; g < 128; g++) {
    int k = 0;
    for (; k < ncls; k++) {
      bool same = true;
      for (int t = 0; t < 4 && same; t++)
        for (int q = 0; q < 128 && same; q++) same = pd[t][q][g] == pd[t][q][rep[k]];
      if (same) break;
    }
    if (k == ncls) {
      if (ncls == 32) return -1;
      rep[ncls++] = g;
    }
    SC[g] = (int8_t)k;
  }
  for (int t = 0; t < 4; t++)
    for (int q = 0; q < 128; q++) {
      uint32_t w[4] = {0, 0, 0, 0};
      for (int k = 0; k < ncls; k++) {
        const int v = pd[t][q][rep[k]];
        if (v < -8 || v > 7) return -1;
        w[k >> 3] |= (uint32_t)(v & 15) << (4 * (k & 7));
      }
      for (int j = 0; j < 4; j++) SP[(t * 128 + q) * 4 + j] = w[j];
    }
  return ncls;
}

const uint8_t *gdp_tables_consistent(void) { return &cons[0][0][0]; }

void gdp_tables_set(const int16_t *pd_in, const int32_t *nt_in, const uint8_t *cons_in) {
  memcpy(pd, pd_in, sizeof pd);
  memcpy(nt2int, nt_in, sizeof nt2int);
  memcpy(cons, cons_in, sizeof cons);
  for (int t = 0; t < 4; t++) use8p[t] = -128 / kMismatch[t] - 1;
}

extern "C" const int16_t *gdp_pairdistance(int mismatchtype) { return &pd[mismatchtype & 3][0][0]; }
extern "C" const int32_t *gdp_nt_to_int(void) { return nt2int; }
extern "C" int32_t gdp_use8p_size(int mismatchtype) { return use8p[mismatchtype & 3]; }
extern "C" int32_t gdp_consistent(int genestrand, int c, int g) { return cons[genestrand % 3][c & 127][g & 127]; }

// Dynprog_compute_bands (dynprog.c:1245-1261).
extern "C" void gdp_compute_bands(int32_t *lband, int32_t *uband, int32_t rlength, int32_t glength,
                                  int32_t extraband, int32_t widebandp) {
  if (!widebandp) {
    *lband = extraband;
    *uband = extraband;
  } else if (glength >= rlength) {
    *uband = glength - rlength + extraband;
    *lband = extraband;
  } else {
    *lband = rlength - glength + extraband;
    *uband = extraband;
  }
}
