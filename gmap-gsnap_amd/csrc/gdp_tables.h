// gdp_tables.h -- GMAP's DP score tables (dynprog.c:1006-1196), built on the
// host and packed for the kernels.
#pragma once
#include <stdint.h>

void gdp_tables_build(int mode);
// Packs 4-bit score fields: PQ[mt*128+ch] field x = score(query ch, letter x),
// PG[mt*128+ch] field x = score(letter x, genome ch); NT = nucleotide class.
void gdp_tables_pack(uint32_t *PQ, uint32_t *PG, int8_t *NT);
// Score classes of genome characters for the scalar standard fill:
// SP[(mt*128 + q)*4 + w] field k = score(q, class 8w + k), SC[ch] = class.
// Returns the class count, -1 if the tables need more than 32 or wider scores.
int gdp_tables_pack_std(uint32_t *SP, int8_t *SC);
// consistent_array[genestrand][c][g] as 3*128*128 bytes.
const uint8_t *gdp_tables_consistent(void);
// Installs tables built elsewhere (the host program's Dynprog_init).
void gdp_tables_set(const int16_t *pd, const int32_t *nt, const uint8_t *cons);

extern "C" {
// Read-only views for callers and tests.
const int16_t *gdp_pairdistance(int mismatchtype);  // 128*128, [query][genome]
const int32_t *gdp_nt_to_int(void);
int32_t gdp_use8p_size(int mismatchtype);
int32_t gdp_consistent(int genestrand, int c, int g);
void gdp_compute_bands(int32_t *lband, int32_t *uband, int32_t rlength, int32_t glength, int32_t extraband,
                       int32_t widebandp);
}
