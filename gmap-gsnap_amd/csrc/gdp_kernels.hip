// gdp_kernels.hip -- gfx950 kernels of the banded affine-gap DP core.
//
// Fill: a 16-lane systolic sweep per problem.  Lane i owns rows i, i+16, ...;
// cell (r, c) of row batch b = r/16 is computed at step b*P + (r%16) + c, so
// at every step a lane's upper neighbour (r-1, c) was produced one step
// earlier by lane i-1 (DPP row shift) or, for lane 0, by lane 15 of the
// previous batch (LDS ring).  Each cell applies the reference's rules for
// the AVX2 build exactly (dynprog_simd.c): saturating H and E at the kernel
// width, a 32-bit vertical gap F, 32/16-row block starts with their
// out-of-band prefix, the forced-diagonal bottom edge, the restarted top
// edge, and the jump-late/early tie rule.
//
// Directions are packed 4 bits per cell and stored step-major
// (word = step/8 * 16 + lane), one coalesced 64-byte store per group every 8
// steps.  The traceback maps (r, c) back to its step.
#include <hip/hip_runtime.h>
#include <limits.h>

#include <atomic>

#include "gdp_internal.h"

#define GDP_BLOCK 256
#define GDP_GROUPS_PER_BLOCK (GDP_BLOCK / GDP_GROUP)

// Score tables, uploaded by gdp_init (gdp_api.cpp):
//   gdp_tabPQ[mt][ch]: five 4-bit signed scores of query char ch against
//                      A,C,G,T,N (pairdistance_array[mt][ch][letter]);
//   gdp_tabPG[mt][ch]: scores of A,C,G,T,N against genome char ch
//                      (pairdistance_array[mt][letter][ch]);
//   gdp_tabNT[ch]:     nt_to_int_array (dynprog.c:1012-1018);
//   gdp_tabCons[gs][c][g]: consistent_array (dynprog.c:893-900).
__device__ uint32_t gdp_tabPQ[4 * 128];
__device__ uint32_t gdp_tabPG[4 * 128];
__device__ int8_t gdp_tabNT[128];
__device__ uint8_t gdp_tabCons[3 * 128 * 128];
// The same as bits: bit g of word [(gs*128 + c)*4 + g/32].
__device__ uint32_t gdp_tabConsBits[3 * 128 * 4];
// Scalar standard fill (scores by raw character, dynprog.c:1527-1530):
//   gdp_tabSC[ch]:     class of genome char ch (characters whose
//                      pairdistance columns agree share a class, <= 32);
//   gdp_tabSP[mt][q]:  4-bit scores of query char q against the classes,
//                      eight per word.
__device__ int8_t gdp_tabSC[128];
__device__ uint4 gdp_tabSP[4 * 128];


// Packed five-field maximum (lower-triangle profiles take the max over the
// two genome alternatives first, dynprog_simd.c:8670-8690).
__device__ __forceinline__ uint32_t fieldmax(uint32_t a, uint32_t b) {
  uint32_t out = 0;
#pragma unroll
  for (int x = 0; x < 5; x++) out |= ((uint32_t)(gdp_imax(gdp_field4(a, x), gdp_field4(b, x)) & 15)) << (4 * x);
  return out;
}

struct gdp_tables_lds {
  uint32_t PQ[4 * 128];
  uint32_t PG[4 * 128];
  int8_t NT[128];
  int8_t SC[128];
};

__device__ __forceinline__ void load_tables(gdp_tables_lds &t) {
  for (int k = threadIdx.x; k < 4 * 128; k += blockDim.x) {
    t.PQ[k] = gdp_tabPQ[k];
    t.PG[k] = gdp_tabPG[k];
  }
  for (int k = threadIdx.x; k < 128; k += blockDim.x) {
    t.NT[k] = gdp_tabNT[k];
    t.SC[k] = gdp_tabSC[k];
  }
}

// Does candidate (v, r, c) replace the best (bv, br, bc) of an end scan?  The
// reference's scan (dynprog_end.c:143-290, 358-508) runs in row-major order
// and keeps the first maximum (the last under jump-late), so any partition
// of the cells can be scanned apart and merged by value, then by position.
__device__ __forceinline__ bool gdp_best_before(int v, int r, int c, int bv, int br, int bc, int late) {
  if (v != bv) return v > bv;
  const bool later = r != br ? r > br : c > bc;
  return late ? later : !later;
}

// The characters that score row `row`: the query character, or for the
// lower triangle (rows are genome positions) the genome character and its
// alternate (g | ga << 8).  Rows past the last read nothing.
__device__ __forceinline__ uint32_t row_chars(const gdp_dev_problem &p, const gdp_geom &G, const char *qbuf,
                                              const char *gbuf, int row) {
  if (row < 1 || row > G.R) return 0;
  if (p.kind != GDP_KIND_LOWER) return (uint32_t)(uint8_t)qbuf[p.qoff + row - 1];
  return (uint32_t)(uint8_t)gbuf[p.goff + row - 1] | ((uint32_t)(uint8_t)gbuf[p.galtoff + row - 1] << 8);
}

// Row profile: the five scores of row `row` (characters `ch` from row_chars)
// against the column classes.
__device__ __forceinline__ uint32_t row_profile(const gdp_tables_lds &t, const gdp_dev_problem &p, uint32_t ch,
                                                int row) {
  const int mt = p.mismatchtype;
  if (p.kind != GDP_KIND_LOWER) return t.PQ[mt * 128 + (row == 0 ? 'N' : (int)(ch & 127))];
  if (row == 0) return t.PG[mt * 128 + (p.width == 8 ? 4 : 'N')];
  const int g = ch & 127, ga = (ch >> 8) & 127;
  uint32_t a = t.PG[mt * 128 + g];
  return ga == g ? a : fieldmax(a, t.PG[mt * 128 + ga]);
}

// Column classes (genome class and alternate class) of column `col`.
__device__ __forceinline__ void col_class(const gdp_tables_lds &t, const gdp_dev_problem &p, const char *qbuf,
                                          const char *gbuf, int col, int &x, int &xa) {
  if (col == 0) {
    x = xa = 4;
  } else if (p.kind == GDP_KIND_LOWER) {
    x = xa = t.NT[qbuf[p.qoff + col - 1] & 127];
  } else {
    x = t.NT[gbuf[p.goff + col - 1] & 127];
    xa = t.NT[gbuf[p.galtoff + col - 1] & 127];
  }
}

// Column classes cached per group in LDS for problems with at most
// GDP_CCAP - 1 columns: byte c = genome class | alternate class << 4.
#define GDP_CCAP 512

// One problem of the single sweep, the band kind fixed at compile time.
template <int KIND>
__device__ __forceinline__ void v1_sweep(const gdp_tables_lds &tab, const gdp_dev_problem &pin, int pid,
                                         const char *__restrict__ qbuf, const char *__restrict__ gbuf,
                                         uint32_t *__restrict__ dirs, int16_t *__restrict__ scores,
                                         int32_t *__restrict__ final_out, uint32_t *ring, int rmask, uint8_t *ccls,
                                         int lane) {
  constexpr bool FULL = KIND == GDP_KIND_FULL;
  gdp_dev_problem p = pin;
  p.kind = KIND;  // known at compile time from here on
  const gdp_geom G = gdp_make_geom(p);
  const gdp_consts K = gdp_make_consts(p);
  const bool keep = p.score_off >= 0 && p.keep_mode == GDP_KEEP_ALL;
  // GDP_KEEP_BEST(_QUERYEND): this triangle's part of the end scan.  A lane
  // visits a sweep row's cells in column order, which is scan order within
  // the row (the lower fill's sweep row is one genome column, its columns
  // query rows), so within a row the scan's own rule applies (> or >=); the
  // row's best (rbv at sweep column rbc) joins the lane's best (query row br,
  // genome column bc) by value and position when the lane moves on, and the
  // group's lanes merge at the end.  A row's eligible sweep columns are its
  // computed columns from emin on: c >= r for the upper fill, c > r for the
  // lower, and for queryend_indels only query row rl (the upper fill's row
  // rl, the lower fill's last column).
  const int bmode = FULL ? 0 : p.keep_mode;
  const bool upper = p.kind == GDP_KIND_UPPER;
  const int qrl = upper ? G.R : G.C;  // query length
  int bv = bmode == GDP_KEEP_BEST_QUERYEND ? K.NEG : 0, br = bmode == GDP_KEEP_BEST_QUERYEND ? qrl : 0, bc = 0;
  int rbv = INT_MIN, rbc = 0, emin = 1 << 30;
  auto best_row = [&](int row) {  // eligible columns of sweep row `row`
    emin = upper ? row : row + 1;
    if (bmode == GDP_KEEP_BEST_QUERYEND) emin = upper ? (row == qrl ? row : 1 << 30) : gdp_imax(emin, qrl);
    if (row < 1) emin = 1 << 30;
    rbv = INT_MIN;
  };
  auto best_merge = [&](int row) {  // the row's best into the lane's
    if (rbv == INT_MIN) return;
    const int qr = upper ? row : rbc, qc = upper ? rbc : row;
    if (gdp_best_before(rbv, qr, qc, bv, br, bc, K.late)) {
      bv = rbv;
      br = qr;
      bc = qc;
    }
  };

  // Final cell (rl, gl) of a full-band problem; a single writer per problem.
  const bool final_in_band = FULL && G.R <= G.C + G.lb && G.R >= G.C - G.ub;
  if (FULL && lane == 0 && !final_in_band) final_out[pid] = INT_MIN;

  // the columns' classes, once per problem
  const bool cached = G.C < GDP_CCAP;
  if (cached) {
    for (int c = lane; c <= G.C; c += GDP_GROUP) {
      int x, xa;
      col_class(tab, p, qbuf, gbuf, c, x, xa);
      ccls[c] = (uint8_t)(x | (xa << 4));
    }
    // the group's lanes read what the others wrote: one wave, in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }

  int T = 0, r = lane;
  int cs = 0, ce = -1;
  uint32_t prof = 0;
  if (r <= G.R) {
    cs = gdp_cstart(G, r);
    ce = gdp_cend(G, r);
    prof = row_profile(tab, p, row_chars(p, G, qbuf, gbuf, r), r);
  }
  // the characters of the lane's next row, loaded a row ahead so that a row
  // switch waits on no global load
  uint32_t nxt = row_chars(p, G, qbuf, gbuf, r + GDP_GROUP);
  if (!FULL && bmode >= GDP_KEEP_BEST) best_row(r);
  int Hl = K.INIT_H, El = K.E_INIT, uprev = K.NEG, Hout = K.NEG, Fout = GDP_F_FLOOR;
  uint32_t word = 0;
  uint32_t *dout = dirs + p.dir_off;

  for (int s = 0; s < G.nsteps; s++) {
    int c = s - T - lane;
    if (r <= G.R && c > ce) {  // this lane's row is done: take row r + 16
      if (!FULL && bmode >= GDP_KEEP_BEST) {
        best_merge(r);
        best_row(r + GDP_GROUP);
      }
      T += G.P;
      r += GDP_GROUP;
      if (r <= G.R) {
        cs = gdp_cstart(G, r);
        ce = gdp_cend(G, r);
        prof = row_profile(tab, p, nxt, r);
        nxt = row_chars(p, G, qbuf, gbuf, r + GDP_GROUP);
      }
      c = s - T - lane;
    }
    const bool act = (r <= G.R) && (c >= cs) && (c <= ce);

    // Upper neighbour (r-1, c) from the previous step.
    int Hup = __builtin_amdgcn_update_dpp(Hout, Hout, 0x111, 0xF, 0xF, false);
    int Fup = FULL ? __builtin_amdgcn_update_dpp(Fout, Fout, 0x111, 0xF, 0xF, false) : GDP_F_FLOOR;
    if (lane == 0) {
      uint32_t ent = ring[c & rmask];
      Hup = (int)(int16_t)(ent & 0xffffu);
      if (FULL) Fup = (int)(int16_t)(ent >> 16) - 128;
    }
    const int Hdg = uprev;  // (r-1, c-1), received one step ago
    uprev = Hup;

    int H = K.NEG, F = GDP_F_FLOOR, E = El;
    uint32_t nib = 0;
    if (act) {
      int x, xa;
      if (cached) {
        const int b = ccls[c];
        x = b & 15;
        xa = b >> 4;
      } else {
        col_class(tab, p, qbuf, gbuf, c, x, xa);
      }
      gdp_cell_k<FULL>(K, G, r, c, cs, Hl, El, Hdg, Hup, Fup, prof, x, xa, H, F, E, nib);
      if (FULL && r == G.R && c == G.C) final_out[pid] = H;
      // the scan's rule within the row: > (>= under jump-late)
      if (!FULL && bmode >= GDP_KEEP_BEST && c >= emin && H + K.late > rbv) {
        rbv = H;
        rbc = c;
      }
      if (lane == GDP_GROUP - 1) ring[c & rmask] = (uint32_t)(H & 0xffff) | ((uint32_t)(F + 128) << 16);
    }
    if (keep) scores[p.score_off + (int64_t)s * GDP_GROUP + lane] = (int16_t)(act ? H : 0);
    Hout = H;
    if (FULL) Fout = F;
    Hl = H;
    El = E;
    word |= nib << (4 * (s & 7));
    if ((s & 7) == 7) {
      dout[(int64_t)(s >> 3) * GDP_GROUP + lane] = word;
      word = 0;
    }
  }
  if (G.nsteps & 7) dout[(int64_t)(G.nsteps >> 3) * GDP_GROUP + lane] = word;
  if (!FULL && bmode >= GDP_KEEP_BEST) {
    if (r <= G.R) best_merge(r);
    for (int d = 1; d < GDP_GROUP; d <<= 1) {
      const int ov = __shfl_xor(bv, d, 64), orr = __shfl_xor(br, d, 64), oc = __shfl_xor(bc, d, 64);
      if (gdp_best_before(ov, orr, oc, bv, br, bc, K.late)) {
        bv = ov;
        br = orr;
        bc = oc;
      }
    }
    if (lane < GDP_BEST_SLOTS) scores[p.score_off + lane] = (int16_t)(lane == 0 ? bv : lane == 1 ? br : lane == 2 ? bc : 0);
  }
  // the next problem rewrites this group's column cache
  __builtin_amdgcn_wave_barrier();
}

// ------------------------------------------------ scalar standard fill

// Score of the row's query character against genome class cls: the row's
// 32 four-bit scores as two 64-bit halves (a register pair each; selecting
// a word of a 4-word vector by index would go through scratch memory).
__device__ __forceinline__ int std_score(uint64_t lo, uint64_t hi, int cls) {
  const uint64_t v = cls < 16 ? lo : hi;
  return ((int)(uint32_t)(v >> (4 * (cls & 15))) << 28) >> 28;
}

// Dynprog_standard (dynprog.c:1267-1785) for one problem: the int32 fill of
// the reference's non-SIMD builds, full band, with its upperp / lowerp
// switches and saturation floor, swept like v1_sweep with rows dealt to the
// 16 lanes (no emulated row blocks: row r computes columns
// max(0, r - lb) .. min(gl, r + ub)).  Each lane passes the unclamped H and
// the F of its cell on; stored scores, and the H that E and the diagonal
// read, are max(H, saturation), except that row 0 and column 0 keep their
// initial values unclamped.  Rules, cell by cell:
//   row 0 / column 0 ......... open + k*extend, HORIZ / VERT (1316-1341);
//   a row's first cell ....... E and H from the left are -inf (1357-1369);
//   a column's top cell ...... F from -inf, and H above = H(0, c) while the
//                              band touches row 0, -inf below it and
//                              -inf - open + 1 in column 1; the diagonal is
//                              the unclamped top cell of the previous column
//                              (1409-1448, 1578);
//   F, E, then H ............. 1485-1491, 1517-1523, 1527-1570 (jump early:
//                              1672-1757, strict comparisons).
__device__ __forceinline__ void v1_sweep_std(const gdp_tables_lds &tab, const gdp_dev_problem &p, int pid,
                                             const char *__restrict__ qbuf, const char *__restrict__ gbuf,
                                             uint32_t *__restrict__ dirs, int16_t *__restrict__ scores,
                                             int32_t *__restrict__ final_out, uint32_t *ring, int ringsz,
                                             uint8_t *ccls, int lane) {
  const gdp_geom G = gdp_make_geom(p);
  const int NEG = -32768, o = p.open, e = p.extend, late = p.jump_late ? 1 : 0, sat = p.saturation;
  const bool up = (p.std_flags & GDP_STD_UPPER) != 0, lo = (p.std_flags & GDP_STD_LOWER) != 0;
  const bool keep = p.score_off >= 0;
  const bool final_in_band = G.R <= G.C + G.lb && G.R >= G.C - G.ub;
  if (lane == 0 && !final_in_band) final_out[pid] = INT_MIN;

  // genome classes of the columns, two bytes each, for up to GDP_CCAP/2 - 1 columns
  uint16_t *cc = (uint16_t *)ccls;
  const bool cached = G.C < GDP_CCAP / 2;
  if (cached) {
    for (int c = lane + 1; c <= G.C; c += GDP_GROUP)
      cc[c] = (uint16_t)((uint8_t)tab.SC[gbuf[p.goff + c - 1] & 127] |
                         ((uint16_t)(uint8_t)tab.SC[gbuf[p.galtoff + c - 1] & 127] << 8));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // ring entry per column: (H, F) of the last lane's row, 32 bits each
  uint2 *ring2 = (uint2 *)ring;
  const int rmask = ringsz / 2 - 1;

  int T = 0, r = lane, cs = 0, ce = -1;
  uint64_t plo = 0, phi = 0;  // the row's scores against classes 0-15 and 16-31
  if (r <= G.R) {
    cs = gdp_cstart(G, r);
    ce = gdp_cend(G, r);
    if (r > 0) {
      const uint4 pf = gdp_tabSP[p.mismatchtype * 128 + (qbuf[p.qoff + r - 1] & 127)];
      plo = pf.x | ((uint64_t)pf.y << 32);
      phi = pf.z | ((uint64_t)pf.w << 32);
    }
  }
  int Hl = NEG, El = NEG, uprev = NEG, Hout = NEG, Fout = NEG;
  uint32_t word = 0;
  uint32_t *dout = dirs + p.dir_off;

  for (int s = 0; s < G.nsteps; s++) {
    int c = s - T - lane;
    if (r <= G.R && c > ce) {  // this lane's row is done: take row r + 16
      T += G.P;
      r += GDP_GROUP;
      if (r <= G.R) {
        cs = gdp_cstart(G, r);
        ce = gdp_cend(G, r);
        const uint4 pf = gdp_tabSP[p.mismatchtype * 128 + (qbuf[p.qoff + r - 1] & 127)];
        plo = pf.x | ((uint64_t)pf.y << 32);
        phi = pf.z | ((uint64_t)pf.w << 32);
      }
      c = s - T - lane;
    }
    const bool act = (r <= G.R) && (c >= cs) && (c <= ce);

    // upper neighbour (r-1, c): the previous lane one step ago, or the ring
    int Hup = __builtin_amdgcn_update_dpp(Hout, Hout, 0x111, 0xF, 0xF, false);
    int Fup = __builtin_amdgcn_update_dpp(Fout, Fout, 0x111, 0xF, 0xF, false);
    if (lane == 0) {
      const uint2 ent = ring2[c & rmask];
      Hup = (int)ent.x;
      Fup = (int)ent.y;
    }
    const int Hdg = uprev;  // (r-1, c-1), unclamped
    uprev = Hup;

    int H = NEG, F = NEG, E = NEG;
    uint32_t nib = 0;
    if (act) {
      if (r == 0) {
        H = c == 0 ? 0 : o + c * e;
        nib = c == 0 ? 0u : (uint32_t)(GDP_DIR_H | GDP_DIR_E);
      } else if (c == 0) {
        H = o + r * e;
        nib = GDP_DIR_V | GDP_DIR_F;
      } else {
        const bool rowstart = c == cs;  // r > lb: the left neighbour lies outside the band
        const bool top = r == 1 || c == r + G.ub;
        const int hl = rowstart ? NEG : (c == 1 ? Hl : gdp_imax(Hl, sat));
        const int el = rowstart ? NEG : El;
        int hu, fu, dg;
        if (top) {
          fu = NEG;
          hu = c == 1 ? NEG - o + 1 : (r == 1 ? Hup : NEG);
          dg = c == 1 ? 0 : Hdg;
        } else {
          fu = Fup;
          hu = Hup;
          dg = c == 1 ? Hdg : gdp_imax(Hdg, sat);
        }
        const int e0 = hl + o, f0 = hu + o;
        const int dE = up && gdp_wins(el, e0, late);
        const int dF = lo && gdp_wins(fu, f0, late);
        E = (dE ? el : e0) + e;
        F = (dF ? fu : f0) + e;
        int x, xa;
        if (cached) {
          const int b = cc[c];
          x = b & 255;
          xa = b >> 8;
        } else {
          x = tab.SC[gbuf[p.goff + c - 1] & 127];
          xa = tab.SC[gbuf[p.galtoff + c - 1] & 127];
        }
        int h = dg + gdp_imax(std_score(plo, phi, x), std_score(plo, phi, xa));
        uint32_t nd = 0;
        if (up && gdp_wins(E, h, late)) {
          h = E;
          nd = GDP_DIR_H;
        }
        if (lo && gdp_wins(F, h, late)) {
          h = F;
          nd = GDP_DIR_V;
        }
        H = h;
        nib = nd | (dE ? GDP_DIR_E : 0u) | (dF ? GDP_DIR_F : 0u);
        if (r == G.R && c == G.C) final_out[pid] = gdp_imax(h, sat);
      }
      if (lane == GDP_GROUP - 1) ring2[c & rmask] = make_uint2((uint32_t)H, (uint32_t)F);
    }
    if (keep)
      scores[p.score_off + (int64_t)s * GDP_GROUP + lane] = (int16_t)((act && r > 0 && c > 0) ? gdp_imax(H, sat) : 0);
    Hout = H;
    Fout = F;
    Hl = H;
    El = E;
    word |= nib << (4 * (s & 7));
    if ((s & 7) == 7) {
      dout[(int64_t)(s >> 3) * GDP_GROUP + lane] = word;
      word = 0;
    }
  }
  if (G.nsteps & 7) dout[(int64_t)(G.nsteps >> 3) * GDP_GROUP + lane] = word;
  __builtin_amdgcn_wave_barrier();
}

// The single-sweep launches.  STD = false sweeps the SIMD-rule problems
// (gdp_fill_kernel), STD = true the standard fills (gdp_fill_std_kernel); the
// two are separate kernels so that neither carries the other's registers.
// Dynamic LDS per group: ringsz ring words (a power of two, at least the
// largest batch period of the plan's problems of that launch, two words per
// column for the standard fill), then GDP_CCAP bytes of column classes.  A
// launch sweeps the problems whose ring need (batch period, twice that for
// the standard fill) lies in (ring_lo, ringsz]: the wide launch, one wave per
// workgroup with rings of up to GDP_RING_WIDE words, takes those past
// GDP_RING.
template <bool STD>
__device__ __forceinline__ void single_sweeps(const gdp_dev_problem *__restrict__ probs, int n,
                                              const char *__restrict__ qbuf, const char *__restrict__ gbuf,
                                              uint32_t *__restrict__ dirs, int16_t *__restrict__ scores,
                                              int32_t *__restrict__ final_out, int ringsz, int ring_lo) {
  __shared__ gdp_tables_lds tab;
  extern __shared__ uint32_t ring_all[];
  // A plan lists its single-sweep problems after its header (the SIMD-rule
  // ones, then the standard fills); a bare descriptor array (one problem of a
  // GDP_PLAN_SINGLE plan) has no header and is walked whole.
  const gdp_pair_hdr *hdr = (const gdp_pair_hdr *)(probs + n);
  const bool listed = hdr->magic == (GDP_PAIR_MAGIC ^ (uint64_t)n) && hdr->n == n;
  const int n1 = listed ? (STD ? hdr->n1std : hdr->n1) : n;
  if (n1 == 0) return;
  const int32_t *list1 =
      listed ? (const int32_t *)((const uint8_t *)probs + hdr->list1_off) + (STD ? hdr->n1 : 0) : nullptr;
  load_tables(tab);
  __syncthreads();

  const int lane = threadIdx.x & (GDP_GROUP - 1);
  const int grp = threadIdx.x / GDP_GROUP;
  const int groups = blockDim.x / GDP_GROUP;
  uint32_t *ring = ring_all + grp * ringsz;
  uint8_t *ccls = (uint8_t *)(ring_all + groups * ringsz) + grp * GDP_CCAP;
  // groups walk the list grid-stride; no barrier follows
  for (int slot = blockIdx.x * groups + grp; slot < n1; slot += gridDim.x * groups) {
    const int pid = list1 ? list1[slot] : slot;
    const gdp_dev_problem p = probs[pid];
    if (p.layout != 0 || (p.kind == GDP_KIND_STD) != STD) continue;  // another launch sweeps it
    const int need = (STD ? 2 : 1) * gdp_make_geom(p).P;
    if (need <= ring_lo || need > ringsz) continue;  // the other ring size's launch sweeps it
    // one loop per band kind; the plan's list keeps each kind together, so
    // the groups of a wave run the same one
    if (STD)
      v1_sweep_std(tab, p, pid, qbuf, gbuf, dirs, scores, final_out, ring, ringsz, ccls, lane);
    else if (p.kind == GDP_KIND_FULL)
      v1_sweep<GDP_KIND_FULL>(tab, p, pid, qbuf, gbuf, dirs, scores, final_out, ring, ringsz - 1, ccls, lane);
    else if (p.kind == GDP_KIND_UPPER)
      v1_sweep<GDP_KIND_UPPER>(tab, p, pid, qbuf, gbuf, dirs, scores, final_out, ring, ringsz - 1, ccls, lane);
    else
      v1_sweep<GDP_KIND_LOWER>(tab, p, pid, qbuf, gbuf, dirs, scores, final_out, ring, ringsz - 1, ccls, lane);
  }
}

// Five waves per SIMD: the register budget (96) that holds the sweep without
// spills.
__global__ __launch_bounds__(GDP_BLOCK, 5) void gdp_fill_kernel(const gdp_dev_problem *__restrict__ probs, int n,
                                                             const char *__restrict__ qbuf,
                                                             const char *__restrict__ gbuf, uint32_t *__restrict__ dirs,
                                                             int16_t *__restrict__ scores, int32_t *__restrict__ final_out,
                                                             int ringsz, int ring_lo) {
  single_sweeps<false>(probs, n, qbuf, gbuf, dirs, scores, final_out, ringsz, ring_lo);
}

__global__ __launch_bounds__(GDP_BLOCK) void gdp_fill_std_kernel(const gdp_dev_problem *__restrict__ probs, int n,
                                                                 const char *__restrict__ qbuf,
                                                                 const char *__restrict__ gbuf,
                                                                 uint32_t *__restrict__ dirs,
                                                                 int16_t *__restrict__ scores,
                                                                 int32_t *__restrict__ final_out, int ringsz,
                                                                 int ring_lo) {
  single_sweeps<true>(probs, n, qbuf, gbuf, dirs, scores, final_out, ringsz, ring_lo);
}

// ------------------------------------------------------------ narrow sweep

// One-sided fills (layout 2) with a band under 32 columns, one problem per
// lane.  The upper fill's E is -inf on and below the diagonal and its
// diagonal cells take the diagonal (dynprog_simd.c:4478-4626, 7867-7960; the
// lower fill the same, transposed), so in-band cells never read the
// emulated out-of-band prefix of a row block: a lane computes row r's
// in-band cells c = r .. r + ub only, left to right, with the same cell
// rule as gdp_cell_k restricted to those cells (the diagonal cell's E is
// -inf and its H the diagonal's; right of it E and H as in the full rule
// without F), so the block-start seeds never enter.
// Per lane, in LDS laid out lane-minor (no bank conflicts): the row above,
// where cell r + k reads the diagonal H(r-1, r-1+k) from slot k and
// overwrites it, and a ring of column class bytes (genome class |
// alternate << 4) that gains one column per row.  Directions go out as P
// nibbles per row, one word per 8 cells.
#define GDP_NARROW_BW 32    // widest row slot (cells) of a narrow sweep
#define GDP_NARROW_RING 64  // column ring per lane
template <int BW>
__device__ __forceinline__ void narrow_sweep(const gdp_tables_lds &tab, const gdp_dev_problem &p,
                                             const char *__restrict__ qbuf, const char *__restrict__ gbuf,
                                             uint32_t *__restrict__ dirs, int16_t *__restrict__ scores,
                                             int16_t *hrow, uint8_t *cring, int lane) {
  const gdp_geom G = gdp_make_geom(p);
  const gdp_consts K = gdp_make_consts(p);
  const bool upper = p.kind == GDP_KIND_UPPER;
  const int B = G.ub, R = G.R, C = G.C, nw = p.P;  // nw = BW / 8 for every lane of the wave
  uint32_t *dout = dirs + p.dir_off;
  // the characters of column col (the upper fill's columns are genome
  // positions, the lower fill's query positions), then their class byte
  auto col_raw = [&](int col) -> uint32_t {
    if (col < 1 || col > C) return 0u;
    if (upper) return (uint32_t)(uint8_t)gbuf[p.goff + col - 1] | ((uint32_t)(uint8_t)gbuf[p.galtoff + col - 1] << 8);
    const uint32_t q = (uint8_t)qbuf[p.qoff + col - 1];
    return q | (q << 8);
  };
  auto col_cls = [&](int col, uint32_t raw) -> uint8_t {
    if (col == 0) return (uint8_t)(4u | (4u << 4));
    return (uint8_t)((uint32_t)tab.NT[raw & 127] | ((uint32_t)tab.NT[(raw >> 8) & 127] << 4));
  };
#pragma unroll 1
  for (int c = 0; c < BW; c++) cring[(c & (GDP_NARROW_RING - 1)) * 64 + lane] = col_cls(c, col_raw(c));

  // best end cell (GDP_KEEP_BEST / _QUERYEND): the upper fill is visited in
  // scan order; the lower fill's rows are genome columns, so each row's best
  // joins the lane's by value and position
  const int bmode = p.keep_mode;
  const int qrl = upper ? R : C;
  int bv = bmode == GDP_KEEP_BEST_QUERYEND ? K.NEG : 0, br = bmode == GDP_KEEP_BEST_QUERYEND ? qrl : 0, bc = 0;

  uint32_t ch = row_chars(p, G, qbuf, gbuf, 0), chn = row_chars(p, G, qbuf, gbuf, 1);
  for (int r = 0; r <= R; r++) {
    const uint32_t prof = row_profile(tab, p, ch, r);
    ch = chn;
    chn = row_chars(p, G, qbuf, gbuf, r + 2);
    const uint32_t nextraw = col_raw(r + BW);  // joins the ring after this row
    const int nact = gdp_imin(B, C - r) + 1;  // in-band cells of this row
    // the end scan's cells of this row are k in [kmin, kmax]: the upper
    // fill's from the diagonal on (queryend_indels: row rl only), the lower
    // fill's right of it (queryend_indels: query row rl only)
    int kmin = upper ? 0 : 1, kmax = nact - 1;
    if (bmode == GDP_KEEP_BEST_QUERYEND) {
      if (upper && r != R) kmax = -1;
      if (!upper) kmin = gdp_imax(kmin, C - r);
    }
    if (r < 1 || bmode < GDP_KEEP_BEST) kmax = -1;
    const unsigned kspan = kmax >= kmin ? (unsigned)(kmax - kmin) : 0u;
    if (kmax < kmin) kmin = 1 << 20;
    int Hl = K.NEG, El = K.NEG, rbv = INT_MIN, rbk = 0;
#pragma unroll 1
    for (int kk = 0; kk < BW / 8; kk++) {
      uint32_t word = 0;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int k = 8 * kk + j, c = r + k;
        const bool act = k < nact;
        const uint32_t cb = cring[(c & (GDP_NARROW_RING - 1)) * 64 + lane];
        const int hd = hrow[k * 64 + lane];
        // score of the cell: the better of the two genome alternatives
        // (4-bit fields of the row profile, one per class)
        const int sc = gdp_imax(__builtin_amdgcn_sbfe((int)prof, (int)((cb & 15u) << 2), 4),
                                __builtin_amdgcn_sbfe((int)prof, (int)((cb >> 4) << 2), 4));
        int H, E;
        uint32_t nib = 0;
        if (j == 0 && kk == 0) {
          // the diagonal cell c = r: E is -inf there and H takes the
          // diagonal (row 0: H(0, 0) = 0 + score)
          E = K.NEG;
          H = gdp_clampw((r == 0 ? 0 : hd) + sc, K.NEG, K.POS);
        } else {
          // E from the left cell, then H = max(diagonal + score, E)
          // (gdp_cell_k's one-sided rule right of the diagonal; row 0 reads
          // a diagonal of -inf)
          const int T1 = gdp_imax(Hl + K.o, K.NEG);
          const bool dE = El + K.late > T1;
          E = gdp_imax(gdp_imax(El, T1) + K.e, K.NEG);
          const int Hd = gdp_clampw((r == 0 ? K.NEG : hd) + sc, K.NEG, K.POS);
          const bool dH = E + K.late > Hd;
          H = gdp_imax(Hd, E);
          nib = (dH ? (uint32_t)GDP_DIR_H : 0u) | (dE ? (uint32_t)GDP_DIR_E : 0u);
        }
        Hl = H;
        El = E;
        hrow[k * 64 + lane] = (int16_t)(act ? H : hd);
        word |= (act ? nib : 0u) << (4 * j);
        const bool upd = (unsigned)(k - kmin) <= kspan && H + K.late > rbv;
        rbv = upd ? H : rbv;
        rbk = upd ? k : rbk;
      }
      dout[((int64_t)r * nw + kk) * 64] = word;  // lanes of the wave: consecutive words
    }
    if (rbv != INT_MIN) {
      const int qr = upper ? r : r + rbk, qc = upper ? r + rbk : r;
      if (upper ? (rbv + K.late > bv) : gdp_best_before(rbv, qr, qc, bv, br, bc, K.late)) {
        bv = rbv;
        br = qr;
        bc = qc;
      }
    }
    cring[((r + BW) & (GDP_NARROW_RING - 1)) * 64 + lane] = col_cls(r + BW, nextraw);
  }
  if (bmode >= GDP_KEEP_BEST) {
    scores[p.score_off] = (int16_t)bv;
    scores[p.score_off + 1] = (int16_t)br;
    scores[p.score_off + 2] = (int16_t)bc;
    scores[p.score_off + 3] = 0;
  }
}

// One lane per problem of the plan's narrow list; a wave takes 64
// consecutive entries (sorted by row stride and rows) and runs the loop of
// the widest stride among them.
__global__ __launch_bounds__(GDP_BLOCK, 4) void gdp_fill_narrow_kernel(const gdp_dev_problem *__restrict__ probs, int n,
                                                                    const char *__restrict__ qbuf,
                                                                    const char *__restrict__ gbuf,
                                                                    uint32_t *__restrict__ dirs,
                                                                    int16_t *__restrict__ scores) {
  __shared__ gdp_tables_lds tab;
  __shared__ int16_t hrow_all[GDP_BLOCK / 64][GDP_NARROW_BW * 64];
  __shared__ uint8_t ring_all[GDP_BLOCK / 64][GDP_NARROW_RING * 64];
  const gdp_pair_hdr *hdr = (const gdp_pair_hdr *)(probs + n);
  if (hdr->magic != (GDP_PAIR_MAGIC ^ (uint64_t)n) || hdr->n != n || hdr->n2 == 0) return;
  const int n2 = hdr->n2;
  const int32_t *list2 = (const int32_t *)((const uint8_t *)probs + hdr->list2_off);
  load_tables(tab);
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nwave = (n2 + 63) / 64;
  for (int w = blockIdx.x * (GDP_BLOCK / 64) + wv; w < nwave; w += gridDim.x * (GDP_BLOCK / 64)) {
    const int slot = w * 64 + lane;
    if (slot >= n2) continue;  // lanes past the list's end
    const gdp_dev_problem p = probs[list2[slot]];
    // the wave's row stride (the plan gives all its lanes the widest) picks the loop
    const int bw = 8 * __builtin_amdgcn_readfirstlane(p.P);
    if (bw == 32) narrow_sweep<32>(tab, p, qbuf, gbuf, dirs, scores, hrow_all[wv], ring_all[wv], lane);
    else if (bw == 16) narrow_sweep<16>(tab, p, qbuf, gbuf, dirs, scores, hrow_all[wv], ring_all[wv], lane);
    else narrow_sweep<8>(tab, p, qbuf, gbuf, dirs, scores, hrow_all[wv], ring_all[wv], lane);
  }
}

// ------------------------------------------------------------------ lookups

__device__ __forceinline__ uint32_t nib_at(const uint32_t *dirs, const gdp_dev_problem &p, const gdp_geom &G, int row,
                                           int col) {
  if (row < 0 || row > G.R || col < 0 || col > G.C) return 0;
  if (col < gdp_cstart(G, row) || col > gdp_cend(G, row)) return 0;  // not computed: DIAG
  if (p.layout == 2) {  // narrow sweep: in-band cells only, interleaved words
    if (col < row) return 0;
    const int k = col - row;
    return (dirs[p.dir_off + ((int64_t)row * p.P + (k >> 3)) * 64] >> (4 * (k & 7))) & 15;
  }
  if (p.layout == 1) {  // paired sweep: byte 0 Egap, 1 HORIZ, 2 Fgap, 3 VERT; bit = step & 7
    const int g = 1 << p.glog;
    const int t = gdp_p_step(g, p.P, p.slb, row, col);  // the wave's sweep layout
    const uint32_t w = dirs[p.dir_off + (int64_t)(t >> 3) * g + (row & (g - 1))] >> (t & 7);
    return ((w >> 24) & 1 ? GDP_DIR_V : ((w >> 8) & 1 ? GDP_DIR_H : 0)) | (w & 1 ? GDP_DIR_E : 0) |
           ((w >> 16) & 1 ? GDP_DIR_F : 0);
  }
  int s = gdp_step(G, row, col);
  uint32_t w = dirs[p.dir_off + (int64_t)(s >> 3) * GDP_GROUP + (row % GDP_GROUP)];
  return (w >> (4 * (s & 7))) & 15;
}

__device__ __forceinline__ int score_at(const int16_t *scores, const gdp_dev_problem &p, const gdp_geom &G, int row,
                                        int col) {
  int s = gdp_step(G, row, col);
  return scores[p.score_off + (int64_t)s * GDP_GROUP + (row % GDP_GROUP)];
}

// ---------------------------------------------------------------- traceback

#define TB_MATCH 1 /* scores.h:4-9 */
#define TB_MISMATCH (-3)
#define TB_QOPEN (-3)
#define TB_QINDEL (-1)
#define TB_TOPEN (-3)
#define TB_TINDEL (-1)
#define MICROINTRON_LENGTH 9 /* pairpool.c:18 */

// complCode (COMPLEMENT_LC, complement.h:30) stated as rules: nucleotide and
// IUPAC letters go to their complements in the same case (U to A), the
// bracket pairs () <> [] {} swap, control characters, DEL and the characters
// ! " ? @ and backslash become '?', and every other character stays.
__device__ __forceinline__ int compl_char(int ch) {
  if (ch < 32 || ch > 126) return '?';
  switch (ch) {
  case '!': case '"': case '?': case '@': case '\\': return '?';
  case '(': return ')';
  case ')': return '(';
  case '<': return '>';
  case '>': return '<';
  case '[': return ']';
  case ']': return '[';
  case '{': return '}';
  case '}': return '{';
  default: break;
  }
  const int lower = ch & 0x20, up = ch & ~0x20;
  if (up < 'A' || up > 'Z') return ch;
  int r;
  switch (up) {
  case 'A': r = 'T'; break;
  case 'T': case 'U': r = 'A'; break;
  case 'C': r = 'G'; break;
  case 'G': r = 'C'; break;
  case 'B': r = 'V'; break;
  case 'V': r = 'B'; break;
  case 'D': r = 'H'; break;
  case 'H': r = 'D'; break;
  case 'K': r = 'M'; break;
  case 'M': r = 'K'; break;
  case 'R': r = 'Y'; break;
  case 'Y': r = 'R'; break;
  default: r = up; break;
  }
  return r | lower;
}

// One base of a genome in device memory, as Genome_get_char reads it
// (genome.c:11003-11019): ASCII bytes indexed by position, or GMAP's
// genomecomp blocks decoded as uncompress_one_char with flag character 'N'
// (genome.c:9723-9760): per 32 bases three words {high, low, flags}, bases
// 0..15 two bits each in low, 16..31 in high, A C G T = 0 1 2 3, a set flag
// bit reads 'N'.  Positions at or past `glen` read 'N'.
__device__ __forceinline__ int genome_at(const void *genome, int fmt, uint32_t glen, uint32_t pos) {
  if (genome == nullptr || pos >= glen) return 'N';
  if (fmt == GDP_GENOME_ASCII) return ((const char *)genome)[pos];
  const uint32_t *b = (const uint32_t *)genome + (size_t)(pos >> 5) * 3;
  const int bit = (int)(pos & 31);
  if ((b[2] >> bit) & 1u) return 'N';
  const uint32_t w = bit < 16 ? b[1] : b[0];
  return (int)((0x54474341u >> (8 * ((w >> (2 * (bit & 15))) & 3u))) & 0xffu);
}

// Genome_get_segment_right / Genome_get_segment_left (genome.c:11022-11120)
// with the alternate genome equal to the genome, one wave per segment.
// Right: bases [coord, coord + length) with those at or past chrhigh = bound
// as '*'; left: bases [coord - length, coord) with those below chroffset =
// bound as '*'; revcomp reverses and complements the whole segment
// (make_complement_inplace, genome.c:331-345).
__global__ __launch_bounds__(256) void gdp_segments_kernel(const void *__restrict__ genome, int fmt, uint32_t glen,
                                                           const gdp_segment *__restrict__ segs, int n,
                                                           char *__restrict__ out) {
  const int s = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  if (s >= n) return;
  const gdp_segment sg = segs[s];
  if (sg.length <= 0) return;
  char *o = out + sg.out_off;
  const int64_t len = sg.length, coord = sg.coord, bound = sg.bound;
  // source index j is out of bounds below `lo` or at/after `hi`
  int64_t lo = 0, hi = len, start;
  if (!sg.left) {
    start = coord;
    if (coord >= bound) hi = 0;
    else if (coord + len >= bound) hi = len - (coord + len - bound);
  } else {
    start = coord - len;
    if (coord < bound) lo = len;
    else if (coord < bound + len) lo = bound + len - coord;
  }
  for (int64_t i = lane; i < len; i += 64) {
    const int64_t j = sg.revcomp ? len - 1 - i : i;
    int ch = (j < lo || j >= hi) ? '*' : genome_at(genome, fmt, glen, (uint32_t)(start + j));
    if (sg.revcomp) ch = compl_char(ch);
    o[sg.reverse ? len - 1 - i : i] = (char)ch;
  }
}

// Direction lookup for the walks: the geometry held in registers, the
// divisions of gdp_step / gdp_p_step turned into shifts.
struct tb_dirs {
  const uint32_t *base;
  int layout, glog, gmask, P, lb, ub, R, C, full, smask, slb;
  __device__ __forceinline__ void init(const uint32_t *dirs, const gdp_dev_problem &p, const gdp_geom &G) {
    base = dirs + p.dir_off;
    layout = p.layout;
    glog = p.layout == 1 ? p.glog : 4;
    gmask = (1 << glog) - 1;
    P = p.layout != 0 ? p.P : G.P;
    lb = G.lb;
    ub = G.ub;
    R = G.R;
    C = G.C;
    full = G.full;
    smask = ~(G.S - 1);
    slb = p.layout == 1 ? p.slb : G.lb;  // the paired sweep's layout follows the wave's band
  }
  // Word holding sweep cell (row, col) and the cell's step within it;
  // ok = false outside the computed cells (direction DIAG there).
  __device__ __forceinline__ const uint32_t *where(int row, int col, int &t, bool &ok) const {
    const int blk = row & smask;
    const int cs = full ? (blk - lb > 0 ? blk - lb : 0) : blk;
    const int ce = row + ub < C ? row + ub : C;
    ok = (unsigned)row <= (unsigned)R && col >= cs && col <= ce && col >= 0;
    if (layout == 2) {  // narrow sweep: in-band cells, interleaved words
      ok = ok && col >= row;
      const int k = col - row;
      t = k & 7;
      return ok ? base + ((int64_t)row * P + (k >> 3)) * 64 : base;
    }
    const int b = row >> glog, i = row & gmask;
    if (layout == 1) {
      const int c0 = (b << glog) - slb;
      t = b * P + i + col - (c0 > 0 ? c0 : 0);
      return ok ? base + ((t >> 3) << glog) + i : base;
    }
    t = b * P + i + col;
    return ok ? base + ((t >> 3) << 4) + i : base;
  }
  __device__ __forceinline__ uint32_t nib(uint32_t w, int t, bool ok) const {
    if (!ok) return 0;
    if (layout == 1) {
      w >>= t & 7;
      return ((w >> 24) & 1 ? GDP_DIR_V : ((w >> 8) & 1 ? GDP_DIR_H : 0)) | (w & 1 ? GDP_DIR_E : 0) |
             ((w >> 16) & 1 ? GDP_DIR_F : 0);
    }
    return (w >> (4 * (t & 7))) & 15;
  }
};

// ---- cooperative walk: 16 lanes per traceback ----------------------------
//
// The reference walks one cell at a time (dynprog_simd.c:9153-9945); every
// step of a single-thread walk would wait on one dependent memory read.  A
// group of TBG lanes instead evaluates TBG steps of the walk's current
// straight stretch at once, each lane reading its own cell: in a diagonal
// stretch lane j looks at (r - j, c - j) and, where the walk would pair the
// cell, forms that pair; in a gap run lane j looks at the j-th cell of the
// run.  A ballot finds where the stretch ends, all lanes advance the shared
// state by that many steps, and the pairs are written in walk order
// (prefix counts of the ballot).  The group's state (r, c, counters, pair
// count) is identical in all its lanes, so control flow is uniform per group.
#define TBG 16

// The pair list entry at index idx (stored when idx < cap).
#define TB_STORE(idx, qpos, gpos, cdna_, comp_, g_, ga_, gapp_, jump_)                        \
  do {                                                                                          \
    if ((idx) < cap) {                                                                          \
      gdp_pair pr_;                                                                             \
      pr_.querypos = (qpos);                                                                    \
      pr_.genomepos = (gpos);                                                                   \
      pr_.cdna = (int8_t)(cdna_);                                                               \
      pr_.comp = (int8_t)(comp_);                                                               \
      pr_.genome = (int8_t)(g_);                                                                \
      pr_.genomealt = (int8_t)(ga_);                                                            \
      pr_.gapp = (int16_t)(gapp_);                                                              \
      pr_.genomejump = (int16_t)(jump_);                                                        \
      obuf[(idx)] = pr_;                                                                        \
    }                                                                                           \
  } while (0)

enum { TB_DECIDE, TB_RUN_H, TB_RUN_V, TB_TAIL, TB_DONE };

__device__ __forceinline__ uint32_t tb_ballot(bool p, int g0) {
  return (uint32_t)((__ballot(p) >> g0) & 0xFFFFull);
}

// Eight waves per SIMD (64 registers): the walks are latency-bound, and
// more of them in flight hide more of each round's memory wait.
__global__ __launch_bounds__(256, 8) void gdp_traceback_kernel(
    const gdp_dev_problem *__restrict__ probs, const gdp_tb_request *__restrict__ reqs, int nreq,
    const char *__restrict__ qbuf, const char *__restrict__ gbuf, const uint32_t *__restrict__ dirs,
    const void *__restrict__ genome, int gfmt, uint32_t glen, gdp_pair *__restrict__ pairs,
    int32_t *__restrict__ counts, int32_t *__restrict__ ends) {
  __shared__ uint32_t consb[3 * 128 * 4];
  for (int j = threadIdx.x; j < 3 * 128 * 4; j += blockDim.x) consb[j] = gdp_tabConsBits[j];
  __syncthreads();
  const int lane = (int)(threadIdx.x & 63), g0 = lane & ~(TBG - 1), j = lane & (TBG - 1);
  const uint32_t below = (1u << j) - 1u;  // group lanes before this one
  const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) / TBG);
  if (k >= nreq) return;  // whole groups leave together
  const gdp_tb_request q = reqs[k];
  const int queryoffset = q.queryoffset, genomeoffset = q.genomeoffset, watsonp = q.watsonp, revp = q.revp;
  const uint32_t chroffset = q.chroffset, chrhigh = q.chrhigh;
  const int consrow = q.genestrand * 128;
  const bool ops = (q.flags & GDP_TB_OPS) != 0;
  const bool nogaps = (q.flags & GDP_TB_NOGAPS) != 0;
  // traceback_local_{8,16}_{upper,lower} (dynprog_end.c:727-1135): the walk
  // stops at genome column endc, its diagonal pairs skip no '*' column and
  // its genome skips read the fill's own genome sequence
  const bool local = (q.flags & GDP_TB_LOCAL) != 0;
  const int endc = local ? q.endc : 0;
  const int cap = q.max_pairs;
  gdp_pair *const obuf = pairs + q.pair_off;
  int32_t *cnt = counts + 6 * (int64_t)k;
  const gdp_dev_problem *pp = probs + q.fill;
  // the standard fill's directions follow the full band's traceback
  // (Dynprog_traceback_std, dynprog.c:1795-1947, walks them alike)
  const int kind = pp->kind == GDP_KIND_STD ? GDP_KIND_FULL : pp->kind;
  const int64_t p_qoff = pp->qoff, p_goff = pp->goff, p_galtoff = pp->galtoff;
  if (q.r < 0 || q.r > pp->rl || q.c < 0 || q.c > pp->gl ||  // start cell outside the matrix
      (local && kind != GDP_KIND_UPPER && kind != GDP_KIND_LOWER) ||  // local walks are one-sided
      (local && endc < 0)) {  // a stop column left of the matrix: column -1 is no cell
    // (a stop column right of the start is the reference's case too: the
    // walk stops at once and its closing skip is empty)
    if (j == 0) cnt[0] = -1;
    return;
  }
  tb_dirs TD;
  TD.init(dirs, *pp, gdp_make_geom(*pp));
  const bool lower = kind == GDP_KIND_LOWER;
  const uint32_t vbit = kind == GDP_KIND_FULL ? GDP_DIR_F : GDP_DIR_E;

  int n = 0, sc = 0, nm = 0, nmm = 0, nop = 0, nind = 0;
  int r = q.r, c = q.c;
  int mode = TB_DECIDE, dist = 0;

  while (mode != TB_DONE) {
    int sk_kind = 0, sk_rr = 0, sk_cc = 0, sk_d = 0;
    if (mode == TB_DECIDE) {
      if (r <= 0 || c <= endc) {
        mode = TB_TAIL;
        continue;
      }
      // Lane j: the cell the walk reaches after j diagonal steps.
      const int rr = r - j, cc = c - j;
      const bool inm = rr > 0 && cc > endc;
      int c1 = 0, c2 = 0, c2a = 0, m = TB_DECIDE;
      if (inm) {
        c1 = qbuf[p_qoff + rr - 1];
        c2 = gbuf[p_goff + cc - 1];
        c2a = gbuf[p_galtoff + cc - 1];
        if (!nogaps) {
          int t;
          bool ok;
          const uint32_t *a = lower ? TD.where(cc, rr, t, ok) : TD.where(rr, cc, t, ok);
          const uint32_t d = TD.nib(*a, t, ok);
          if (kind == GDP_KIND_FULL) m = (d & GDP_DIR_V) ? TB_RUN_V : ((d & GDP_DIR_H) ? TB_RUN_H : TB_DECIDE);
          else m = (d & (GDP_DIR_H | GDP_DIR_V)) ? (kind == GDP_KIND_UPPER ? TB_RUN_H : TB_RUN_V) : TB_DECIDE;
        }
      }
      const uint32_t dm = tb_ballot(inm && m == TB_DECIDE, g0);
      const int J = __builtin_ctz(~dm | 0x10000u);  // diagonal steps taken this round
      // a diagonal pair (dynprog_simd.c:9245-9275); '*' columns past the
      // chromosome end push nothing (dynprog_simd.c:9248)
      const bool paired = j < J && (local || c2 != '*');
      int comp = ' ';
      if (paired) {
        const int c1uc = (c1 >= 'a' && c1 <= 'z') ? c1 - 32 : c1;
        const int row = (consrow + (c1uc & 127)) << 2;
        if (c1uc == c2 || c1uc == c2a) comp = '*';
        else if (((consb[row + ((c2 & 127) >> 5)] >> (c2 & 31)) & 1) ||
                 ((consb[row + ((c2a & 127) >> 5)] >> (c2a & 31)) & 1))
          comp = ':';  // consistent_array (dynprog.c:893-900)
      }
      const uint32_t mm_ = tb_ballot(paired && comp == ' ', g0), ma = tb_ballot(paired && comp != ' ', g0);
      nm += __popc(ma);
      nmm += __popc(mm_);
      sc += TB_MATCH * __popc(ma) + TB_MISMATCH * __popc(mm_);
      const int qpos = queryoffset + (revp ? 1 - rr : rr - 1), gpos = genomeoffset + (revp ? 1 - cc : cc - 1);
      // Pairpool_push drops pairs at negative coordinates (pairpool.c:189-191)
      const bool emit = paired && (ops || (qpos >= 0 && gpos >= 0));
      const uint32_t em = tb_ballot(emit, g0);
      if (emit) {
        const int idx = n + __popc(em & below);
        if (ops) TB_STORE(idx, rr, cc, 'D', comp, 0, 0, 0, 0);
        else TB_STORE(idx, qpos, gpos, c1, comp, c2, c2a, 0, 0);
      }
      n += __popc(em);
      if (J < TBG) {
        // the walk stops at lane J's cell: outside the matrix, or a gap starts
        const int mJ = __shfl(m, g0 + J, 64);
        r -= J;
        c -= J;
        if (r > 0 && c > endc) {
          mode = mJ;
          dist = 1;
        }
      } else {
        r -= TBG;
        c -= TBG;
      }
      continue;
    } else if (mode == TB_RUN_H || mode == TB_RUN_V) {
      // Lane j: iteration j of the run loop (dynprog_simd.c:9230-9243):
      // full follows Egap leftwards / Fgap upwards while c > 0 / r > 0,
      // upper and lower follow their Egap unguarded.
      const bool h = mode == TB_RUN_H;
      const int rr = h ? r : r - j, cc = h ? c - j : c;
      // (the local upper walk stops its run at endc, dynprog_end.c:960)
      const bool guard = local ? (!h || cc > endc) : (kind != GDP_KIND_FULL || (h ? cc > 0 : rr > 0));
      bool more = false;
      if (guard) {
        int t;
        bool ok;
        const uint32_t *a = lower ? TD.where(cc, rr, t, ok) : TD.where(rr, cc, t, ok);
        more = (TD.nib(*a, t, ok) & (h ? GDP_DIR_E : vbit)) != 0;
      }
      const uint32_t mr = tb_ballot(more, g0);
      const int K = __builtin_ctz(~mr | 0x10000u);
      if (K == TBG) {
        dist += TBG;
        if (h) c -= TBG;
        else r -= TBG;
        continue;
      }
      const int gK = __shfl((int)guard, g0 + K, 64);
      const int dec = K + gK;
      if (h) c -= dec;
      else r -= dec;
      dist += K;
      sk_kind = h ? 1 : 2;
      sk_rr = h ? r : r + dist;
      sk_cc = h ? c + dist : c;
      sk_d = dist;
      mode = TB_DECIDE;
    } else {  // TB_TAIL: the leftover skip once r or c reaches 0
      mode = TB_DONE;
      if (nogaps) {
        // traceback_nogaps (dynprog_end.c:648-722) stops without a skip
      } else if (kind == GDP_KIND_FULL) {
        if (r == 0 && c == 0) {
        } else if (c == 0) {
          sk_kind = 2, sk_rr = r, sk_cc = 1, sk_d = r;
        } else {
          sk_kind = 1, sk_rr = 1, sk_cc = c, sk_d = c;
        }
      } else if (kind == GDP_KIND_UPPER) {
        if (c != endc) sk_kind = 1, sk_rr = 1, sk_cc = c, sk_d = c - endc;
      } else {
        if (r != 0) sk_kind = 2, sk_rr = r, sk_cc = endc + 1, sk_d = r;
      }
      if (ends != nullptr && j == 0) {
        // where the walk stopped: traceback_local leaves it in *r, *c
        ends[2 * (int64_t)k] = r;
        ends[2 * (int64_t)k + 1] = c;
      }
    }
    if (sk_kind == 1) {
      // Pairpool_add_genomeskip (pairpool.c:1067-1180) with genomesequence
      // NULL: a gap of 9 or more (MICROINTRON_LENGTH) becomes one gap holder
      // and is not scored; a shorter one pushes its pairs, lane j the j-th.
      int qc = sk_rr - 1, lo = sk_cc - sk_d, hi = sk_cc - 1, stp = -1;
      if (revp) {
        const int tmp = lo;
        qc = -qc;
        lo = -hi;
        hi = -tmp;
        stp = +1;
      }
      if (sk_d < MICROINTRON_LENGTH) {
        sc += TB_TOPEN + sk_d * TB_TINDEL;
        nop += 1;
        nind += sk_d;
      }
      if (ops || sk_d >= MICROINTRON_LENGTH) {
        if (j == 0) {
          if (ops) TB_STORE(n, sk_rr, sk_cc, 'H', 0, 0, 0, 0, sk_d);
          else TB_STORE(n, -1, -1, ' ', ' ', ' ', ' ', 1, sk_d);  // Pairpool_push_gapholder (pairpool.c:374-452)
        }
        n += 1;
      } else if (sk_d > 0) {
        const int gpos = genomeoffset + (revp ? lo : hi) + j * stp;
        bool emit = false;
        int ch = 0;
        if (j < sk_d) {
          // get_genomic_nt (pairpool.c:1016-1063) over the device genome
          const uint32_t pos = watsonp ? chroffset + (uint32_t)gpos : chrhigh - (uint32_t)gpos;
          if (local) {
            // Pairpool_add_genomeskip with the walk's genomesequence
            // (pairpool.c:1142-1149): the fill's genome in sweep order
            const int gc = (revp ? lo : hi) + j * stp;
            ch = gbuf[p_goff + (gc < 0 ? -gc : gc)];
          } else if (pos < chroffset || pos >= chrhigh) ch = '*';
          else {
            ch = genome_at(genome, gfmt, glen, pos);
            if (!watsonp) ch = compl_char(ch);
          }
          emit = queryoffset + qc >= 0 && gpos >= 0;
        }
        const uint32_t em = tb_ballot(emit, g0);
        if (emit) TB_STORE(n + __popc(em & below), queryoffset + qc, gpos, ' ', '-', ch, ch, 0, 0);
        n += __popc(em);
      }
    } else if (sk_kind == 2) {
      // Pairpool_add_queryskip (pairpool.c:980-1010), lanes j, j + TBG, ...
      int qc = sk_rr - 1, gc = sk_cc - 1, stp = -1;
      if (revp) {
        qc = -qc;
        gc = -gc;
        stp = +1;
      }
      sc += TB_QOPEN + sk_d * TB_QINDEL;
      nop += 1;
      nind += sk_d;
      if (ops) {
        if (j == 0) TB_STORE(n, sk_rr, sk_cc, 'V', 0, 0, 0, 0, sk_d);
        n += 1;
      } else {
        for (int base = 0; base < sk_d; base += TBG) {
          const int eq = qc + (base + j) * stp;
          bool emit = false;
          int ch = 0;
          if (base + j < sk_d) {
            // rsequence[eq] is q(|eq| + 1) in sweep order for both orientations
            ch = qbuf[p_qoff + (eq < 0 ? -eq : eq)];
            emit = queryoffset + eq >= 0 && genomeoffset + gc >= 0;
          }
          const uint32_t em = tb_ballot(emit, g0);
          if (emit) TB_STORE(n + __popc(em & below), queryoffset + eq, genomeoffset + gc, ch, '-', ' ', ' ', 0, 0);
          n += __popc(em);
        }
      }
    }
  }
  if (j == 0) {
    cnt[0] = n;
    cnt[1] = sc;
    cnt[2] = nm;
    cnt[3] = nmm;
    cnt[4] = nop;
    cnt[5] = nind;
  }
}

// ------------------------------------------------------------ best end cell

// find_best_endpoint_{8,16} / find_best_endpoint_to_queryend_indels_{8,16}
// (dynprog_end.c:143-290, 358-508): the best score over both triangles in
// row-major order, starting from score 0 at (0, 0) (or -inf at (rl, 0), last
// row only, for queryend_indels); a later cell replaces the best when it is
// greater (>= under jump-late).  The scan keeps the first (last under
// jump-late) occurrence of the maximum, so TBG lanes can scan disjoint rows
// (columns of the single row for queryend_indels) and merge by value, then
// by position.  One group of TBG lanes per (upper, lower) pair.
__global__ __launch_bounds__(256) void gdp_best_endpoint_kernel(
    const gdp_dev_problem *__restrict__ probs, const int32_t *__restrict__ ul, int m, int queryend_indels,
    const int16_t *__restrict__ scores, int32_t *__restrict__ out, const gdp_tb_request *__restrict__ tmpl,
    gdp_tb_request *__restrict__ tbout) {
  const int j = (int)(threadIdx.x & (TBG - 1));
  const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) / TBG);
  if (k >= m) return;  // whole groups leave together
  const gdp_dev_problem pu = probs[ul[2 * k]], pl = probs[ul[2 * k + 1]];
  int best, bestr, bestc;
  const int want = queryend_indels ? GDP_KEEP_BEST_QUERYEND : GDP_KEEP_BEST;
  if (pu.score_off < 0 || pl.score_off < 0 || pu.kind != GDP_KIND_UPPER || pl.kind != GDP_KIND_LOWER ||
      pu.keep_mode != pl.keep_mode || (pu.keep_mode != GDP_KEEP_ALL && pu.keep_mode != want)) {
    best = INT_MIN;  // scores were not kept (or kept for the other scan), or the pair is not (upper, lower)
    bestr = -1;
    bestc = -1;
  } else if (pu.keep_mode == want) {
    // the fills scanned their triangles: merge the two records
    const int16_t *a = scores + pu.score_off, *b = scores + pl.score_off;
    best = a[0];
    bestr = a[1];
    bestc = a[2];
    if (gdp_best_before(b[0], b[1], b[2], best, bestr, bestc, pu.jump_late)) {
      best = b[0];
      bestr = b[1];
      bestc = b[2];
    }
  } else {
    const gdp_geom Gu = gdp_make_geom(pu), Gl = gdp_make_geom(pl);
    const int rl = pu.rl, gl = pu.gl, lband = pl.lband, uband = pu.uband, late = pu.jump_late;
    best = queryend_indels ? (pu.width == 8 ? -128 : -32768) : 0;
    bestr = queryend_indels ? rl : 0;
    bestc = 0;
    const int r0 = queryend_indels ? rl : 1;
    for (int r = r0 + (queryend_indels ? 0 : j); r <= rl; r += queryend_indels ? rl + 1 : TBG) {
      const int clo = gdp_imax(1, r - lband), chigh = gdp_imin(gl, r + uband);
      for (int c = clo + (queryend_indels ? j : 0); c <= chigh; c += queryend_indels ? TBG : 1) {
        // lower triangle left of the diagonal (its rows are genome columns)
        const int v = c < r ? score_at(scores, pl, Gl, c, r) : score_at(scores, pu, Gu, r, c);
        if (late ? v >= best : v > best) {
          best = v;
          bestr = r;
          bestc = c;
        }
      }
    }
    // merge the lanes' results in scan order
    for (int d = 1; d < TBG; d <<= 1) {
      const int ov = __shfl_xor(best, d, 64), orr = __shfl_xor(bestr, d, 64), oc = __shfl_xor(bestc, d, 64);
      if (gdp_best_before(ov, orr, oc, best, bestr, bestc, late)) {
        best = ov;
        bestr = orr;
        bestc = oc;
      }
    }
  }
  if (j != 0) return;
  out[3 * k] = best;
  out[3 * k + 1] = bestr;
  out[3 * k + 2] = bestc;
  if (tbout != nullptr) {
    // the walk from the best cell runs in the triangle that holds it
    // (dynprog_end.c:1576-1607): upper when bestc >= bestr, lower otherwise
    gdp_tb_request q = tmpl[k];
    q.fill = bestc >= bestr ? ul[2 * k] : ul[2 * k + 1];
    q.r = bestr;
    q.c = bestc;
    tbout[k] = q;
  }
}

// ------------------------------------------------------------- cDNA bridge

#define GDP_BRIDGE_ROWS 1024  // rows of one end's column held in LDS
#define GDP_BRIDGE_LDS 8192   // words of LDS: right prefix tables and one left column
#define GDP_BRIDGE_LDS_SMALL 4096  // the launch for gaps whose tables fit in half of it
#define GDP_BRIDGE_LDS_TINY 2048   // and the one for gaps whose tables fit in a quarter

// A score (int16 range) and its row in one word: the score in the high half.
__device__ __forceinline__ int bridge_pack(int v, int r) { return (int)(((uint32_t)v << 16) | (uint32_t)(r & 0xffff)); }

// Of two candidates of the bridge, the one the reference's sequential scan
// keeps: the higher score; on a tie the later one in scan order under
// jump-late (>=), the earlier one otherwise (>).  key is the scan position.
__device__ __forceinline__ bool bridge_wins(int v, int64_t key, int bv, int64_t bkey, int late) {
  return v > bv || (v == bv && (late ? key > bkey : key < bkey));
}

// Score of cell (r, c) of one end: rows below the diagonal (r >= c) come
// from the lower fill, which is swept transposed, the others from the upper
// one (dynprog_cdna.c:417-418, 454-455).
__device__ __forceinline__ int bridge_score(const int16_t *scores, const gdp_dev_problem &U, const gdp_geom &GU,
                                            const gdp_dev_problem &Lo, const gdp_geom &GL, int r, int c) {
  return r < c ? score_at(scores, U, GU, r, c) : score_at(scores, Lo, GL, c, r);
}

// bridge_cdna_gap_{8,16}_ud (dynprog_cdna.c:123-381, 386-644), one wave per
// gap.  The reference scans cL ascending, cR descending from gl-cL, rL and rR
// ascending over each end's rows: from max(1, c-uband) up the upper triangle
// to the diagonal, then down the lower one to min(c+lband, rl-1), the right
// end's rows also below D - rL with D = rightoffset-leftoffset.  A genomic
// insertion between the ends costs 0 when empty and open otherwise (the loop
// resets its penalty each step, 496-497).  Candidates merge by value, then
// scan position, which makes the result independent of the order the lanes
// visit them.  Upper rows beyond rl (reachable when rl < gl-1) are rows the
// reference's fill never wrote; they are not candidates here.
//
// For one (cL, cR) the rows rL <= D-1-top(cR) see the whole right column, so
// their best join is the left column's prefix maximum up to that row plus
// the right column's maximum (ties to the later row under jump-late, as the
// scan keeps it): O(1) per (cL, cR) from the right columns' prefix tables
// (LDS, built once per gap) and the current left column's (LDS, one wave scan
// per cL), with lanes over cR.  Each row above that bound takes one lookup in
// the right column's table at its own bound; those rows are counted per cR
// and shared evenly over the lanes.  Gaps wider than the LDS tables take the
// direct form: per cR a prefix table of the right column, lanes over cL and
// every left row.
// Three launches share the gaps: LDSW = GDP_BRIDGE_LDS_TINY takes the gaps
// whose tables fit in that many words (four times the blocks per CU of the
// full size), GDP_BRIDGE_LDS_SMALL those that fit in its words (twice), the
// full size all others, including those that fit neither form.
template <int LDSW>
__global__ __launch_bounds__(64) void gdp_bridge_cdna_kernel(
    const gdp_dev_problem *__restrict__ probs, const int32_t *__restrict__ quads, int m,
    const int16_t *__restrict__ scores, int32_t *__restrict__ out, const gdp_tb_request *__restrict__ tmpl,
    gdp_tb_request *__restrict__ tbout) {
  __shared__ int lds[LDSW];
  const int k = blockIdx.x, lane = threadIdx.x;
  if (k >= m) return;
  const int32_t *q = quads + 5 * (int64_t)k;
  const int iLu = q[0], iLl = q[1], iRu = q[2], iRl = q[3], D = q[4];
  const gdp_dev_problem Lu = probs[iLu], Ll = probs[iLl], Ru = probs[iRu], Rl = probs[iRl];
  const int gl = Lu.gl, rlL = Lu.rl, rlR = Ru.rl;
  const int ubL = Lu.uband, lbL = Ll.lband, ubR = Ru.uband, lbR = Rl.lband;
  const int late = Lu.jump_late, open = Lu.open;
  const int init = Lu.width == 8 ? -128 : -32768;
  const bool ok0 = Lu.kind == GDP_KIND_UPPER && Ll.kind == GDP_KIND_LOWER && Ru.kind == GDP_KIND_UPPER &&
                  Rl.kind == GDP_KIND_LOWER && Lu.keep_mode == GDP_KEEP_ALL && Ll.keep_mode == GDP_KEEP_ALL &&
                  Ru.keep_mode == GDP_KEEP_ALL && Rl.keep_mode == GDP_KEEP_ALL && Lu.score_off >= 0 &&
                  Ll.score_off >= 0 && Ru.score_off >= 0 && Rl.score_off >= 0 && Ll.gl == gl && Ru.gl == gl &&
                  Rl.gl == gl && Ll.rl == rlL && Rl.rl == rlR && Ll.width == Lu.width && Ru.width == Lu.width &&
                  Rl.width == Lu.width;
  // the table form needs its offsets in LDS; the direct form one right column
  const bool fast = ok0 && gl + 1 <= LDSW && ubL + lbL + 1 <= GDP_BRIDGE_ROWS;
  int bv = INT_MIN;
  int64_t bkey = -1;
  const int64_t RR = (int64_t)gdp_imax(rlL, rlR) + 2;  // radix of the row fields of a scan position
  // the table form when the right end's prefix tables, their offsets and one
  // left column fit in LDS (a uniform decision: every input is the gap's)
  int nRsum = 0;
  for (int c0 = 0; fast && c0 < gl; c0 += 64) {
    const int c = c0 + lane;
    const int n = c < gl ? gdp_imax(0, gdp_imax(gdp_imin(c - 1, rlR), gdp_imin(rlR - 1, c + lbR)) -
                                           gdp_imax(1, c - ubR) + 1)
                         : 0;
    int x = n;
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(x, d, 64);
      if (lane >= d) x += o;
    }
    if (fast && c <= gl) lds[c] = nRsum + x - n;  // offset of column c's table (column gl: the end)
    nRsum += __shfl(x, 63, 64);
  }
  if (fast && lane == 0) lds[gl] = nRsum;
  const int nLmax = ubL + lbL + 1;
  const bool table = fast && 2 * (gl + 1) + nRsum + 2 * nLmax <= LDSW;
  // the same test at the small size, in both launches (nRsum is the same)
  // the launch a gap belongs to, from the same test in every launch (nRsum
  // and fast are the same wherever the tables fit the smaller sizes)
  const int need = 2 * (gl + 1) + nRsum + 2 * nLmax;
  const int cls = fast && need <= GDP_BRIDGE_LDS_TINY ? 0 : (fast && need <= GDP_BRIDGE_LDS_SMALL ? 1 : 2);
  if (cls != (LDSW == GDP_BRIDGE_LDS_TINY ? 0 : (LDSW == GDP_BRIDGE_LDS_SMALL ? 1 : 2))) return;  // another launch's
  const bool ok = ok0 && (table || ubR + lbR + 1 <= GDP_BRIDGE_ROWS);
  if (ok && table) {
    const gdp_geom GLu = gdp_make_geom(Lu), GLl = gdp_make_geom(Ll), GRu = gdp_make_geom(Ru), GRl = gdp_make_geom(Rl);
    const int *offR = lds;
    int *rpm = lds + gl + 1, *lp = rpm + nRsum, *lraw = lp + nLmax, *bpre = lraw + nLmax;
    __syncthreads();  // offsets written
    // prefix maxima of every right column, ties to the later row under
    // jump-late; eight rows' loads are issued before the first is used
    for (int cR = lane; cR < gl; cR += 64) {
      const int rloR = gdp_imax(1, cR - ubR), w0 = offR[cR], n = offR[cR + 1] - w0;
      int v0 = INT_MIN, r0 = 0;
      for (int i0 = 0; i0 < n; i0 += 8) {
        int vals[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
          vals[u] = i0 + u < n ? bridge_score(scores, Ru, GRu, Rl, GRl, rloR + i0 + u, cR) : 0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          if (i0 + u < n) {
            if (late ? vals[u] >= v0 : vals[u] > v0) {
              v0 = vals[u];
              r0 = rloR + i0 + u;
            }
            rpm[w0 + i0 + u] = bridge_pack(v0, r0);
          }
        }
      }
    }
    // the first 64 rows of left column c, one per lane (0 past its rows)
    auto left_rows = [&](int c) {
      const int lo = gdp_imax(1, c - ubL), top = gdp_imax(gdp_imin(c - 1, rlL), gdp_imin(rlL - 1, c + lbL));
      return lane <= top - lo ? bridge_score(scores, Lu, GLu, Ll, GLl, lo + lane, c) : 0;
    };
    int pre = gl > 1 ? left_rows(1) : 0;
    for (int cL = 1; cL < gl; cL++) {
      const int rloL = gdp_imax(1, cL - ubL), rtopL = gdp_imax(gdp_imin(cL - 1, rlL), gdp_imin(rlL - 1, cL + lbL));
      const int nL = rtopL - rloL + 1;
      const int cur = pre;
      if (cL + 1 < gl) pre = left_rows(cL + 1);  // the next column's loads run under this one's work
      if (nL <= 0) continue;
      __syncthreads();  // right tables written; the previous left column has been read
      int cv = INT_MIN, cr = -1;
      for (int base = 0; base < nL; base += 64) {
        const int j = base + lane;
        int v = INT_MIN, r = -1;
        if (j < nL) {
          r = rloL + j;
          v = base == 0 ? cur : bridge_score(scores, Lu, GLu, Ll, GLl, r, cL);
          lraw[j] = v;
        }
        for (int d = 1; d < 64; d <<= 1) {
          const int ov = __shfl_up(v, d, 64), orr = __shfl_up(r, d, 64);
          if (lane >= d && !(late ? v >= ov : v > ov)) {
            v = ov;
            r = orr;
          }
        }
        if (!(late ? v >= cv : v > cv)) {
          v = cv;
          r = cr;
        }
        if (j < nL) lp[j] = bridge_pack(v, r);
        cv = __shfl(v, 63, 64);
        cr = __shfl(r, 63, 64);
      }
      __syncthreads();
      const int64_t kc0 = (int64_t)cL * (gl + 1);
      const int ncr = gl - cL + 1;  // right columns cR = 0 .. gl-cL
      // lanes over cR: the rows up to B = D-1-top(cR) see the whole right
      // column; the rows above B (each with its own bound) are counted into
      // bpre so that the lanes can then share them evenly
      int btot = 0;
      for (int c0 = 0; c0 < ncr; c0 += 64) {
        const int cR = c0 + lane;
        int nb = 0;
        if (cR < ncr) {
          const int o = offR[cR], nR = offR[cR + 1] - o;
          if (nR > 0) {
            const int rloR = gdp_imax(1, cR - ubR), rtopR = rloR + nR - 1;
            const int B = D - 1 - rtopR;
            const int ub = gdp_imin(rtopL, B);
            if (ub >= rloL) {
              const int x = lp[ub - rloL], y = rpm[o + nR - 1];
              const int v = (x >> 16) + (y >> 16) + (cR == gl - cL ? 0 : open);
              const int64_t key = ((kc0 + (gl - cR)) * RR + (x & 0xffff)) * RR + (y & 0xffff);
              if (bridge_wins(v, key, bv, bkey, late)) {
                bv = v;
                bkey = key;
              }
            }
            // rows above B whose bound D-1-rL still reaches the column's first row
            nb = gdp_imax(0, gdp_imin(rtopL, D - 1 - rloR) - gdp_imax(rloL, B + 1) + 1);
          }
        }
        int x = nb;
        for (int d = 1; d < 64; d <<= 1) {
          const int o2 = __shfl_up(x, d, 64);
          if (lane >= d) x += o2;
        }
        if (cR < ncr) bpre[cR] = btot + x - nb;
        btot += __shfl(x, 63, 64);
      }
      __syncthreads();
      // the bounded rows, item f in column cR = the last with bpre[cR] <= f
      for (int f = lane; f < btot; f += 64) {
        int lo = 0, hi = ncr - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (bpre[mid] <= f) lo = mid;
          else hi = mid - 1;
        }
        const int cR = lo, o = offR[cR], nR = offR[cR + 1] - o;
        const int rloR = gdp_imax(1, cR - ubR), rtopR = rloR + nR - 1;
        const int rL = gdp_imax(rloL, D - rtopR) + (f - bpre[cR]);
        const int y = rpm[o + (D - rL - 1) - rloR];  // right rows up to D-1-rL
        const int v = lraw[rL - rloL] + (y >> 16) + (cR == gl - cL ? 0 : open);
        const int64_t key = ((kc0 + (gl - cR)) * RR + rL) * RR + (y & 0xffff);
        if (bridge_wins(v, key, bv, bkey, late)) {
          bv = v;
          bkey = key;
        }
      }
    }
  } else if (ok) {
    __syncthreads();  // the table form's offsets are no longer read
    const gdp_geom GLu = gdp_make_geom(Lu), GLl = gdp_make_geom(Ll), GRu = gdp_make_geom(Ru), GRl = gdp_make_geom(Rl);
    int *pmv = lds, *pmr = lds + GDP_BRIDGE_ROWS;
    for (int cR = 0; cR < gl; cR++) {
      const int rloR = gdp_imax(1, cR - ubR), rhighR = gdp_imin(rlR - 1, cR + lbR);
      const int rtopR = gdp_imax(gdp_imin(cR - 1, rlR), rhighR);
      const int nR = rtopR - rloR + 1;
      if (nR <= 0) continue;
      __syncthreads();  // the previous column's table has been read
      int cv = INT_MIN, cr = -1;
      for (int base = 0; base < nR; base += 64) {
        const int j = base + lane;
        int v = INT_MIN, r = -1;
        if (j < nR) {
          r = rloR + j;
          v = bridge_score(scores, Ru, GRu, Rl, GRl, r, cR);
        }
        // inclusive prefix maximum over the lanes, then over the earlier chunks
        for (int d = 1; d < 64; d <<= 1) {
          const int ov = __shfl_up(v, d, 64), orr = __shfl_up(r, d, 64);
          if (lane >= d && !(late ? v >= ov : v > ov)) {
            v = ov;
            r = orr;
          }
        }
        if (!(late ? v >= cv : v > cv)) {
          v = cv;
          r = cr;
        }
        if (j < nR) {
          pmv[j] = v;
          pmr[j] = r;
        }
        cv = __shfl(v, 63, 64);
        cr = __shfl(r, 63, 64);
      }
      __syncthreads();
      const int cLmax = gdp_imin(gl - 1, gl - cR);
      for (int cL = 1 + lane; cL <= cLmax; cL += 64) {
        const int pen = cR == gl - cL ? 0 : open;
        const int rloL = gdp_imax(1, cL - ubL), rhighL = gdp_imin(rlL - 1, cL + lbL);
        const int rtopL = gdp_imax(gdp_imin(cL - 1, rlL), rhighL);
        const int64_t kc = ((int64_t)cL * (gl + 1) + (gl - cR)) * RR;
        for (int rL = rloL; rL <= rtopL; rL++) {
          const int t = gdp_imin(rtopR, D - rL - 1);
          if (t < rloR) break;  // the bound only falls as rL grows
          const int v = bridge_score(scores, Lu, GLu, Ll, GLl, rL, cL) + pmv[t - rloR] + pen;
          const int64_t key = (kc + rL) * RR + pmr[t - rloR];
          if (bridge_wins(v, key, bv, bkey, late)) {
            bv = v;
            bkey = key;
          }
        }
      }
    }
  }
  for (int d = 32; d >= 1; d >>= 1) {
    const int ov = __shfl_xor(bv, d, 64);
    const int64_t okey = __shfl_xor(bkey, d, 64);
    if (bridge_wins(ov, okey, bv, bkey, late)) {
      bv = ov;
      bkey = okey;
    }
  }
  if (lane != 0) return;
  const bool has = ok && (late ? bv >= init : bv > init);
  int cL = -1, cR = -1, rL = -1, rR = -1;
  if (has) {
    rR = (int)(bkey % RR);
    rL = (int)((bkey / RR) % RR);
    const int64_t kc = bkey / RR / RR;
    cR = gl - (int)(kc % (gl + 1));
    cL = (int)(kc / (gl + 1));
  }
  int32_t *o = out + 5 * (int64_t)k;
  o[0] = ok ? (has ? bv : init) : INT_MIN;
  o[1] = cL;
  o[2] = cR;
  o[3] = rL;
  o[4] = rR;
  if (tbout != nullptr) {
    // Dynprog_cdna_gap walks the right end first, each in the triangle that
    // holds its best cell (dynprog_cdna.c:995-1013, 1045-1060)
    gdp_tb_request a = tmpl[2 * (int64_t)k], b = tmpl[2 * (int64_t)k + 1];
    a.fill = has && cR < rR ? iRl : iRu;
    a.r = has ? rR : 0;
    a.c = has ? cR : 0;
    b.fill = has && cL < rL ? iLl : iLu;
    b.r = has ? rL : 0;
    b.c = has ? cL : 0;
    tbout[2 * (int64_t)k] = a;
    tbout[2 * (int64_t)k + 1] = b;
  }
}

// ----------------------------------------------------------- intron bridge

#define GDP_INTRON_SITES 8192        // LDS bytes of one end's dinucleotide codes (glength + 1 entries)
#define GDP_INTRON_SITES_SMALL 1024  // the launch for gaps of both ends under this many columns

// Dinucleotide codes (intron.h:10-21): a left site from the two bases after
// the left exon, a right site from the two before the right exon, either
// allele supplying each base (dynprog_genome.c:1791-1843).  Their AND is the
// intron type (intron.h:26-32).
__device__ __forceinline__ bool gdp_either(int x, int xa, int b) { return x == b || xa == b; }

__device__ __forceinline__ int gdp_left_site(int l1, int l1a, int l2, int l2a) {
  if (gdp_either(l1, l1a, 'G') && gdp_either(l2, l2a, 'T')) return 0x21;  // GT
  if (gdp_either(l1, l1a, 'G') && gdp_either(l2, l2a, 'C')) return 0x10;  // GC
  if (gdp_either(l1, l1a, 'A') && gdp_either(l2, l2a, 'T')) return 0x08;  // AT
  if (gdp_either(l1, l1a, 'C') && gdp_either(l2, l2a, 'T')) return 0x06;  // CT
  return 0;
}

__device__ __forceinline__ int gdp_right_site(int r2, int r2a, int r1, int r1a) {
  if (gdp_either(r2, r2a, 'A') && gdp_either(r1, r1a, 'G')) return 0x30;  // AG
  if (gdp_either(r2, r2a, 'A') && gdp_either(r1, r1a, 'C')) return 0x0C;  // AC
  if (gdp_either(r2, r2a, 'G') && gdp_either(r1, r1a, 'C')) return 0x02;  // GC
  if (gdp_either(r2, r2a, 'A') && gdp_either(r1, r1a, 'T')) return 0x01;  // AT
  return 0;
}

// The reward of an intron type (intron_score_setup, dynprog_genome.c:143-186,
// with its build's constants, 76-115): GT-AG / CT-AC 14 preliminary, 16
// final; GC-AG and its antisense 8 / 10; AT-AC and its antisense 4 / 8; the
// sense table (cdna_direction > 0) holds the forward types, the antisense
// one (< 0) the reverse ones, the either table both -- with the reference's
// preliminary forward GT-AG at 16 and final reverse CT-AC at 14.
__device__ __forceinline__ int gdp_intron_reward(int dir, int finalp, int type) {
  const bool fwd = type == 0x20 || type == 0x10 || type == 0x08;
  const bool rev = type == 0x04 || type == 0x02 || type == 0x01;
  if ((dir > 0 && !fwd) || (dir < 0 && !rev) || (!fwd && !rev)) return 0;
  if (type == 0x20 || type == 0x04) {
    if (dir == 0) return type == 0x20 ? 16 : 14;
    return finalp ? 16 : 14;
  }
  if (type == 0x10 || type == 0x02) return finalp ? 10 : 8;
  return finalp ? 8 : 4;
}

// Of two bridge states, whether (v, p, row) comes first: the higher score,
// then the higher probability sum, then the earlier row of the scan (row -1
// is the initial state, which precedes every candidate).
__device__ __forceinline__ bool gdp_intron_before(int v, double p, int row, int bv, double bp, int brow) {
  return v > bv || (v == bv && (p > bp || (p == bp && row < brow)));
}

// bridge_intron_gap_{8,16}_site_level (dynprog_genome.c:866-1386,
// 1741-2261) as bridge_intron_gap_{8,16}_ud runs it (1387-1488, 2262-2363),
// one wave per gap.  The reference scans rL = 1 .. rlength-1 with rR =
// rlength-rL: the diagonal pair (cL, cR) = (rL, rR); the left end on its
// diagonal with cR from max(1, rR-lbandR) below the right diagonal, then
// past it up to (not including) min(rR+ubandR, glengthR-1); the right end on
// its diagonal with cL likewise; every cR below span - cL.  A candidate
// replaces the best on a higher score, or an equal score with a higher
// probability sum, so the result is the scan's first candidate of the
// highest (score, probability sum): lanes take whole rows, keep their
// rows' first best, and merge by (score, sum, row).  The dinucleotide track
// (diagonal pairs with a positive reward, the first of the highest sum above
// 0) merges by (sum, row).  The dinucleotide codes of both ends sit in LDS.
// Two launches share the gaps: SITES = GDP_INTRON_SITES_SMALL takes those
// whose ends both have fewer columns (eight times the blocks per CU), the
// full size all others, including those that do not fit.
template <int SITES>
__global__ __launch_bounds__(64) void gdp_bridge_intron_kernel(
    const gdp_dev_problem *__restrict__ probs, const gdp_intron_gap *__restrict__ gaps, int m,
    const int16_t *__restrict__ scores, const char *__restrict__ gbuf, const double *__restrict__ prob,
    int32_t *__restrict__ out, double *__restrict__ outp, const gdp_tb_request *__restrict__ tmpl,
    gdp_tb_request *__restrict__ tbout) {
  __shared__ uint8_t ldi[SITES], rdi[SITES];
  __shared__ int rew[64];
  const int k = blockIdx.x, lane = threadIdx.x;
  if (k >= m) return;
  const gdp_intron_gap q = gaps[k];
  const gdp_dev_problem Lu = probs[q.fills[0]], Ll = probs[q.fills[1]], Ru = probs[q.fills[2]],
                        Rl = probs[q.fills[3]];
  const int rl = Lu.rl, glL = Lu.gl, glR = Ru.gl;
  const bool ok = Lu.kind == GDP_KIND_UPPER && Ll.kind == GDP_KIND_LOWER && Ru.kind == GDP_KIND_UPPER &&
                  Rl.kind == GDP_KIND_LOWER && Lu.keep_mode == GDP_KEEP_ALL && Ll.keep_mode == GDP_KEEP_ALL &&
                  Ru.keep_mode == GDP_KEEP_ALL && Rl.keep_mode == GDP_KEEP_ALL && Lu.score_off >= 0 &&
                  Ll.score_off >= 0 && Ru.score_off >= 0 && Rl.score_off >= 0 && Ll.width == Lu.width &&
                  Ru.width == Lu.width && Rl.width == Lu.width && Ll.rl == rl && Ru.rl == rl && Rl.rl == rl &&
                  Ll.gl == glL && Rl.gl == glR && glL >= rl + 1 && glR >= rl + 1 && glL < GDP_INTRON_SITES &&
                  glR < GDP_INTRON_SITES;
  const bool small = ok && glL < GDP_INTRON_SITES_SMALL && glR < GDP_INTRON_SITES_SMALL;
  if (small != (SITES == GDP_INTRON_SITES_SMALL)) return;  // the other launch's gap
  if (!ok) {
    if (lane == 0) {
      int32_t *o = out + 5 * (int64_t)k;
      o[0] = INT_MIN;
      o[1] = o[2] = o[3] = o[4] = -1;
      outp[2 * (int64_t)k] = outp[2 * (int64_t)k + 1] = 0.0;
      if (tbout != nullptr) {
        gdp_tb_request a = tmpl[2 * (int64_t)k], b = tmpl[2 * (int64_t)k + 1];
        a.fill = q.fills[2];
        b.fill = q.fills[0];
        a.r = a.c = b.r = b.c = 0;
        tbout[2 * (int64_t)k] = a;
        tbout[2 * (int64_t)k + 1] = b;
      }
    }
    return;
  }
  // column c of a fill reads g[c-1]: the left site after column c is bytes
  // c, c+1 of the left fill's genome; the right fill runs backwards, so the
  // site before its column c is bytes c+1, c of its genome
  const char *gL = gbuf + Lu.goff, *gLa = gbuf + Lu.galtoff, *gR = gbuf + Ru.goff, *gRa = gbuf + Ru.galtoff;
  for (int c = lane; c <= glL; c += 64) ldi[c] = c < glL - 1 ? gdp_left_site(gL[c], gLa[c], gL[c + 1], gLa[c + 1]) : 0;
  for (int c = lane; c <= glR; c += 64)
    rdi[c] = c < glR - 1 ? gdp_right_site(gR[c + 1], gRa[c + 1], gR[c], gRa[c]) : 0;
  rew[lane] = gdp_intron_reward(q.cdna_direction, q.finalp, lane);
  __syncthreads();

  const gdp_geom GLu = gdp_make_geom(Lu), GLl = gdp_make_geom(Ll), GRu = gdp_make_geom(Ru), GRl = gdp_make_geom(Rl);
  const int lbL = Ll.lband, ubL = Lu.uband, lbR = Rl.lband, ubR = Ru.uband, D = q.span;
  const double *pL = prob + q.prob_off, *pR = pL + glL;
  const int neg = Lu.width == 8 ? -128 : -32768;
  int bv = neg, brow = -1, bcL = -1, bcR = -1;
  double bp = 0.0;
  double dp = 0.0;
  int dv = 0, drow = -1;
  for (int rL = 1 + lane; rL < rl; rL += 64) {
    const int rR = rl - rL;
    const int cloL = gdp_imax(1, rL - lbL), chighL = gdp_imin(rL + ubL, glL - 1);
    const int cloR = gdp_imax(1, rR - lbR), chighR = gdp_imin(rR + ubR, glR - 1);
    const int sLd = score_at(scores, Lu, GLu, rL, rL), sRd = score_at(scores, Ru, GRu, rR, rR);
    const int siteL = ldi[rL], siteR = rdi[rR];
    const double pLd = pL[rL], pRd = pR[rR];
    // the diagonal pair, also the dinucleotide track's candidate
    {
      const int r = rew[siteL & siteR];
      const int v = sLd + r + sRd;
      const double p = pLd + pRd;
      if (v > bv || (v == bv && p > bp)) {
        bv = v;
        bp = p;
        brow = rL;
        bcL = rL;
        bcR = rR;
      }
      if (r > 0 && p > dp) {
        dp = p;
        dv = v;
        drow = rL;
      }
    }
    // left end on its diagonal, right end off it: the reference's loop
    // runs cR from cloR while cR < rR and cR < span - rL, then, skipping the
    // column it stopped at, on while cR < chighR and cR < span - rL
    const int e2 = gdp_imin(rR, D - rL), e3 = gdp_imin(chighR, D - rL);
#pragma unroll 4
    for (int cR = cloR; cR < e2; cR++) {
      const int v = sLd + rew[siteL & rdi[cR]] + score_at(scores, Rl, GRl, cR, rR);
      const double p = pLd + pR[cR];
      if (v > bv || (v == bv && p > bp)) {
        bv = v;
        bp = p;
        brow = rL;
        bcL = rL;
        bcR = cR;
      }
    }
#pragma unroll 4
    for (int cR = gdp_imax(cloR, e2) + 1; cR < e3; cR++) {
      const int v = sLd + rew[siteL & rdi[cR]] + score_at(scores, Ru, GRu, rR, cR);
      const double p = pLd + pR[cR];
      if (v > bv || (v == bv && p > bp)) {
        bv = v;
        bp = p;
        brow = rL;
        bcL = rL;
        bcR = cR;
      }
    }
    // right end on its diagonal, left end off it, likewise
    const int f2 = gdp_imin(rL, D - rR), f3 = gdp_imin(chighL, D - rR);
#pragma unroll 4
    for (int cL = cloL; cL < f2; cL++) {
      const int v = score_at(scores, Ll, GLl, cL, rL) + rew[ldi[cL] & siteR] + sRd;
      const double p = pL[cL] + pRd;
      if (v > bv || (v == bv && p > bp)) {
        bv = v;
        bp = p;
        brow = rL;
        bcL = cL;
        bcR = rR;
      }
    }
#pragma unroll 4
    for (int cL = gdp_imax(cloL, f2) + 1; cL < f3; cL++) {
      const int v = score_at(scores, Lu, GLu, rL, cL) + rew[ldi[cL] & siteR] + sRd;
      const double p = pL[cL] + pRd;
      if (v > bv || (v == bv && p > bp)) {
        bv = v;
        bp = p;
        brow = rL;
        bcL = cL;
        bcR = rR;
      }
    }
  }
  for (int d = 32; d >= 1; d >>= 1) {
    const int ov = __shfl_xor(bv, d, 64), orow = __shfl_xor(brow, d, 64), ocL = __shfl_xor(bcL, d, 64),
              ocR = __shfl_xor(bcR, d, 64);
    const double op = __shfl_xor(bp, d, 64);
    if (gdp_intron_before(ov, op, orow, bv, bp, brow)) {
      bv = ov;
      bp = op;
      brow = orow;
      bcL = ocL;
      bcR = ocR;
    }
    const double odp = __shfl_xor(dp, d, 64);
    const int odv = __shfl_xor(dv, d, 64), odrow = __shfl_xor(drow, d, 64);
    if (odp > dp || (odp == dp && odrow < drow)) {
      dp = odp;
      dv = odv;
      drow = odrow;
    }
  }
  if (lane != 0) return;
  int brL = brow, brR = brow >= 0 ? rl - brow : -1;
  // the canonical-dinucleotide candidate takes over unless the best's
  // probability sum exceeds 2 * PROB_CEILING (0.85), none was found, or it
  // scores below 0 or more than 9 under the best (2210-2237)
  if (!(bp > 2 * 0.85) && dp != 0.0 && !(dv < 0 || dv < bv - 9)) {
    bv = dv;
    brL = bcL = drow;
    brR = bcR = rl - drow;
  }
  int fin = bv;
  if (bv >= 0 && q.halfp) fin = bv - rew[ldi[bcL] & rdi[bcR]] / 2;  // 2248-2255
  int32_t *o = out + 5 * (int64_t)k;
  o[0] = fin;
  o[1] = brL;
  o[2] = brR;
  o[3] = bcL;
  o[4] = bcR;
  outp[2 * (int64_t)k] = fin >= 0 ? pL[bcL] : 0.0;
  outp[2 * (int64_t)k + 1] = fin >= 0 ? pR[bcR] : 0.0;
  if (tbout != nullptr) {
    // Dynprog_genome_gap walks the right end first, each in the triangle that
    // holds its cell (3575-3633, 3718-3774)
    const bool has = fin >= 0;
    gdp_tb_request a = tmpl[2 * (int64_t)k], b = tmpl[2 * (int64_t)k + 1];
    a.fill = has && bcR < brR ? q.fills[3] : q.fills[2];
    a.r = has ? brR : 0;
    a.c = has ? bcR : 0;
    b.fill = has && bcL < brL ? q.fills[1] : q.fills[0];
    b.r = has ? brL : 0;
    b.c = has ? bcL : 0;
    tbout[2 * (int64_t)k] = a;
    tbout[2 * (int64_t)k + 1] = b;
  }
}

// --------------------------------------------------- splice-junction ends

// The first walk of each end of Dynprog_end5_splicejunction /
// Dynprog_end3_splicejunction (dynprog_end.c:1797-1830, 2385-2420): from its
// best last-row cell, in the triangle that holds it (fill 2k upper, 2k+1
// lower); an end whose best scores below 0 is not walked (a request at
// (0, 0), which pushes nothing).
__global__ void gdp_sj_req1_kernel(const int32_t *__restrict__ best, int m, const gdp_tb_request *__restrict__ tmpl,
                                   gdp_tb_request *__restrict__ req) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= m) return;
  gdp_tb_request q = tmpl[k];
  const bool go = best[3 * (int64_t)k] >= 0;
  const int r = go ? best[3 * (int64_t)k + 1] : 0, c = go ? best[3 * (int64_t)k + 2] : 0;
  q.fill = 2 * k + (c < r ? 1 : 0);
  q.r = r;
  q.c = c;
  req[k] = q;
}

// The second walk: from the cell where the first one stopped.
__global__ void gdp_sj_req2_kernel(const int32_t *__restrict__ best, const int32_t *__restrict__ stops, int m,
                                   const gdp_tb_request *__restrict__ tmpl, gdp_tb_request *__restrict__ req) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= m) return;
  gdp_tb_request q = tmpl[k];
  int r = stops[2 * (int64_t)k], c = stops[2 * (int64_t)k + 1];
  if (best[3 * (int64_t)k] < 0 || r < 0 || c < 0) r = c = 0;
  q.fill = 2 * k + (c < r ? 1 : 0);
  q.r = r;
  q.c = c;
  req[k] = q;
}

// Length of each end's returned list: the first walk's pairs, the gap
// holder, the second walk's pairs, less the indel pairs at the read's end
// (the first pushed, which the reference drops after reversing the list,
// 1905-1910, 2485-2490); -1 for an end that is not walked.
__global__ void gdp_sj_len_kernel(const int32_t *__restrict__ best, int m, const gdp_pair *__restrict__ p1,
                                  const gdp_tb_request *__restrict__ req1, const int32_t *__restrict__ c1,
                                  const int32_t *__restrict__ c2, int32_t *__restrict__ len,
                                  int32_t *__restrict__ trim) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= m) return;
  if (best[3 * (int64_t)k] < 0) {
    len[k] = -1;
    trim[k] = 0;
    return;
  }
  const int n1 = gdp_imax(0, c1[6 * (int64_t)k]), n2 = gdp_imax(0, c2[6 * (int64_t)k]);
  const gdp_pair *a = p1 + req1[k].pair_off;
  int t = 0;
  while (t < n1 && a[t].comp == '-') t++;
  len[k] = n1 + 1 + n2 - t;
  trim[k] = t;
}

// Writes each walked end's list at its offset, one wave per end: the 5'
// form in push order, the 3' form reversed; the gap holder between the walks
// is Pairpool_push_gapholder with knownp (pairpool.c:374-452).  aux[2k] = the
// read's length (missscore = traceback_score - rl * FULLMATCH), aux[2k+1] =
// the gap holder's genomejump.
__global__ __launch_bounds__(256) void gdp_sj_join_kernel(
    int five, int m, const gdp_pair *__restrict__ p1, const gdp_pair *__restrict__ p2,
    const gdp_tb_request *__restrict__ req1, const gdp_tb_request *__restrict__ req2, const int32_t *__restrict__ c1,
    const int32_t *__restrict__ c2, const int32_t *__restrict__ len, const int32_t *__restrict__ trim,
    const int64_t *__restrict__ off, const int32_t *__restrict__ aux, gdp_sj_pair *__restrict__ out,
    gdp_sj_result *__restrict__ res) {
  const int lane = (int)(threadIdx.x & 63);
  const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (k >= m) return;
  const int L = len[k];
  const int64_t o = off[k];
  if (lane == 0) {
    gdp_sj_result r;
    r.pair_off = o;
    r.npairs = L;
    r.pad = 0;
    for (int i = 0; i < 6; i++) r.counters[i] = 0;
    if (L >= 0) {
      const int32_t *a = c1 + 6 * (int64_t)k, *b = c2 + 6 * (int64_t)k;
      const int sc = a[1] + b[1];
      r.counters[0] = sc;
      r.counters[1] = sc - aux[2 * (int64_t)k] * 3;  // FULLMATCH (dynprog.h:42)
      for (int i = 2; i < 6; i++) r.counters[i] = a[i] + b[i];
    }
    res[k] = r;
  }
  if (L <= 0) return;
  const int n1 = gdp_imax(0, c1[6 * (int64_t)k]), n2 = gdp_imax(0, c2[6 * (int64_t)k]), N = n1 + 1 + n2;
  const int t = trim[k];
  const gdp_pair *a = p1 + req1[k].pair_off, *b = p2 + req2[k].pair_off;
  const int jump = aux[2 * (int64_t)k + 1];
  for (int j = lane; j < L; j += 64) {
    const int i = five ? t + j : N - 1 - j;
    gdp_sj_pair w;
    w.pad = 0;
    if (i == n1) {
      w.querypos = w.genomepos = -1;
      w.cdna = w.comp = w.genome = w.genomealt = ' ';
      w.gapp = 2;
      w.genomejump = jump;
    } else {
      const gdp_pair &s = i < n1 ? a[i] : b[i - n1 - 1];
      w.querypos = s.querypos;
      w.genomepos = s.genomepos;
      w.cdna = s.cdna;
      w.comp = s.comp;
      w.genome = s.genome;
      w.genomealt = s.genomealt;
      w.gapp = s.gapp;
      w.genomejump = s.genomejump;
    }
    out[o + j] = w;
  }
}

extern "C" int gdp_launch_sj(int stage, int five, int m, const int32_t *best, const int32_t *stops,
                             const gdp_tb_request *tmpl, gdp_tb_request *req, const gdp_pair *p1, const gdp_pair *p2,
                             const gdp_tb_request *req1, const gdp_tb_request *req2, const int32_t *c1,
                             const int32_t *c2, int32_t *len, int32_t *trim, const int64_t *off, const int32_t *aux,
                             gdp_sj_pair *out, gdp_sj_result *res, hipStream_t stream) {
  if (m <= 0) return 0;
  const dim3 g((m + 255) / 256), b(256);
  switch (stage) {
  case 0: hipLaunchKernelGGL(gdp_sj_req1_kernel, g, b, 0, stream, best, m, tmpl, req); break;
  case 1: hipLaunchKernelGGL(gdp_sj_req2_kernel, g, b, 0, stream, best, stops, m, tmpl, req); break;
  case 2: hipLaunchKernelGGL(gdp_sj_len_kernel, g, b, 0, stream, best, m, p1, req1, c1, c2, len, trim); break;
  default:
    hipLaunchKernelGGL(gdp_sj_join_kernel, dim3((m + 3) / 4), b, 0, stream, five, m, p1, p2, req1, req2, c1, c2, len,
                       trim, off, aux, out, res);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------ dense export

__global__ void gdp_export_kernel(const gdp_dev_problem *__restrict__ probs, int idx, const uint32_t *__restrict__ dirs,
                                  const int16_t *__restrict__ scores, uint8_t *__restrict__ ddirs,
                                  int16_t *__restrict__ dscores) {
  const gdp_dev_problem p = probs[idx];
  const gdp_geom G = gdp_make_geom(p);
  const int64_t ncell = (int64_t)(G.R + 1) * (G.C + 1);
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ncell; k += (int64_t)gridDim.x * blockDim.x) {
    const int col = (int)(k / (G.R + 1)), row = (int)(k % (G.R + 1));
    const bool comp = col >= gdp_cstart(G, row) && col <= gdp_cend(G, row);
    ddirs[k] = (uint8_t)(comp ? nib_at(dirs, p, G, row, col) : 0);
    if (dscores != nullptr)
      dscores[k] = (int16_t)((comp && p.score_off >= 0 && p.keep_mode == GDP_KEEP_ALL) ? score_at(scores, p, G, row, col) : 0);
  }
}

// Reference workspace layout and encoding (dynprog_simd.c:2158-2282).
template <typename T>
__global__ void gdp_export_ref_kernel(const gdp_dev_problem *__restrict__ probs, int idx,
                                      const uint32_t *__restrict__ dirs, const int16_t *__restrict__ scores, int stride,
                                      T *__restrict__ sc, T *__restrict__ nogap, T *__restrict__ egap,
                                      T *__restrict__ fgap) {
  const gdp_dev_problem p = probs[idx];
  const gdp_geom G = gdp_make_geom(p);
  const int64_t ncell = (int64_t)(G.C + 1) * stride;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ncell; k += (int64_t)gridDim.x * blockDim.x) {
    const int col = (int)(k / stride), row = (int)(k % stride);
    const bool comp = row <= G.R && col >= gdp_cstart(G, row) && col <= gdp_cend(G, row);
    const uint32_t nib = comp ? nib_at(dirs, p, G, row, col) : 0;
    sc[k] = (T)((comp && p.score_off >= 0 && p.keep_mode == GDP_KEEP_ALL) ? score_at(scores, p, G, row, col) : 0);
    nogap[k] = (T)((nib & GDP_DIR_V) ? -2 : ((nib & GDP_DIR_H) ? -1 : 0));
    egap[k] = (T)((nib & GDP_DIR_E) ? -1 : 0);
    if (fgap != nullptr) fgap[k] = (T)((nib & GDP_DIR_F) ? -2 : 0);
  }
}

#include "gdp_fill2.inc"

// ------------------------------------------------------------ host launchers

extern "C" int gdp_launch_export_ref(const void *d_descs, int idx, int width, const uint32_t *d_dirs,
                                     const int16_t *d_scores, int stride, void *sc, void *nogap, void *egap,
                                     void *fgap, hipStream_t stream) {
  if (width == 8)
    hipLaunchKernelGGL(gdp_export_ref_kernel<int8_t>, dim3(64), dim3(256), 0, stream,
                       (const gdp_dev_problem *)d_descs, idx, d_dirs, d_scores, stride, (int8_t *)sc, (int8_t *)nogap,
                       (int8_t *)egap, (int8_t *)fgap);
  else
    hipLaunchKernelGGL(gdp_export_ref_kernel<int16_t>, dim3(64), dim3(256), 0, stream,
                       (const gdp_dev_problem *)d_descs, idx, d_dirs, d_scores, stride, (int16_t *)sc,
                       (int16_t *)nogap, (int16_t *)egap, (int16_t *)fgap);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Persistent grid of the paired fill: as many workgroups as fit on the device
// at once, each wave taking waves off the descriptor's queue.
template <bool GENK>
static int p2_launch(const void *d_descs, int n, const char *d_qbuf, const char *d_gbuf, uint32_t *d_dirs,
                     int32_t *d_final, hipStream_t stream) {
  typedef p2lds<GENK> L;
  // resident workgroups per device, found once (launches may come from
  // several host threads: each finds the same value)
  static std::atomic<int> cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
  int res = cached[dev].load(std::memory_order_relaxed);
  if (res == 0) {
    int ncu = 0, per = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gdp_fill2_kernel<GENK>, P2_THREADS, L::LDS_BYTES) !=
        hipSuccess)
      return -1;
    res = ncu * (per > 0 ? per : 1);
    cached[dev].store(res, std::memory_order_relaxed);
  }
  // resident workgroups, but no more than the waves there can be (a wave
  // holds at least one of the n problems)
  const int blocks = gdp_imin(res, (n + P2_THREADS / 64 - 1) / (P2_THREADS / 64));
  hipLaunchKernelGGL((gdp_fill2_kernel<GENK>), dim3(blocks), dim3(P2_THREADS), L::LDS_BYTES, stream,
                     (const uint8_t *)d_descs, n, d_qbuf, d_gbuf, d_dirs, d_final);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// kernels: bit 0 single sweeps, bits 1..3 the paired classes, bit 4 standard
// fills, bit 5 narrow sweeps (gdp_plan_kernels).
extern "C" int gdp_launch_fill(const void *d_descs, int n, uint32_t kernels, const char *d_qbuf, const char *d_gbuf,
                               uint32_t *d_dirs, int16_t *d_scores, int32_t *d_final, hipStream_t stream) {
  if (n <= 0) return 0;
  // resident blocks only; they walk the plan's list of single sweeps with
  // the ring size the plan asks for (GDP_RING, or GDP_RING_WIDE for the
  // wide launch, when it does not say).  The wide launch runs one wave of
  // four groups per workgroup, so that four rings of GDP_RING_WIDE words fit
  // the LDS of a workgroup.
  auto single = [&](bool std_fill, int rlog, bool wide) -> int {
    const int cap = wide ? GDP_RING_WIDE : GDP_RING;
    const int ringsz = rlog ? 1 << rlog : cap;
    if (ringsz > cap || (wide && std_fill)) return -1;
    const int threads = wide ? 64 : GDP_BLOCK, groups = threads / GDP_GROUP;
    const size_t lds = (size_t)groups * (ringsz * sizeof(uint32_t) + GDP_CCAP);
    static std::atomic<int> cached[64][2][2][16];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
    const int slot = rlog ? rlog : (wide ? 13 : 10);
    auto kern = std_fill ? gdp_fill_std_kernel : gdp_fill_kernel;
    std::atomic<int> &cres = cached[dev][std_fill][wide][slot];
    int res = cres.load(std::memory_order_relaxed);
    if (res == 0) {
      int ncu = 0, per = 0;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, lds) != hipSuccess) return -1;
      res = ncu * (per > 0 ? per : 1);
      cres.store(res, std::memory_order_relaxed);
    }
    const int blocks = gdp_imin(res, (n + groups - 1) / groups);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), lds, stream, (const gdp_dev_problem *)d_descs, n,
                       d_qbuf, d_gbuf, d_dirs, d_scores, d_final, ringsz, wide ? GDP_RING : 0);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  };
  if ((kernels & 1u) && single(false, (kernels >> 8) & 15, false)) return -1;
  if ((kernels & GDP_FILL_WIDE) && single(false, (kernels >> 16) & 15, true)) return -1;
  if ((kernels & GDP_FILL_STD) && single(true, (kernels >> 12) & 15, false)) return -1;
  if (kernels & GDP_FILL_NARROW) {
    static std::atomic<int> cached_n[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
    int res = cached_n[dev].load(std::memory_order_relaxed);
    if (res == 0) {
      int ncu = 0, per = 0;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gdp_fill_narrow_kernel, GDP_BLOCK, 0) != hipSuccess)
        return -1;
      res = ncu * (per > 0 ? per : 1);
      cached_n[dev].store(res, std::memory_order_relaxed);
    }
    // resident blocks, but no more than the list's waves need
    const int blocks = gdp_imin(res, (n + GDP_BLOCK - 1) / GDP_BLOCK);
    hipLaunchKernelGGL(gdp_fill_narrow_kernel, dim3(blocks), dim3(GDP_BLOCK), 0, stream,
                       (const gdp_dev_problem *)d_descs, n, d_qbuf, d_gbuf, d_dirs, d_scores);
    if (hipGetLastError() != hipSuccess) return -1;
  }
  if (kernels & 0xEu) {
    // all paired classes in one persistent launch, then the general-lookup pass
    if (p2_launch<false>(d_descs, n, d_qbuf, d_gbuf, d_dirs, d_final, stream)) return -1;
    if (p2_launch<true>(d_descs, n, d_qbuf, d_gbuf, d_dirs, d_final, stream)) return -1;
  }
  return 0;
}

extern "C" int gdp_launch_traceback(const void *d_descs, const gdp_tb_request *d_reqs, int nreq, const char *d_qbuf,
                                    const char *d_gbuf, const uint32_t *d_dirs, const void *d_genome, int gfmt,
                                    uint32_t glen, gdp_pair *d_pairs, int32_t *d_counts, int32_t *d_ends,
                                    hipStream_t stream) {
  if (nreq <= 0) return 0;
  hipLaunchKernelGGL(gdp_traceback_kernel, dim3((nreq + 256 / TBG - 1) / (256 / TBG)), dim3(256), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_reqs, nreq, d_qbuf, d_gbuf, d_dirs, d_genome, gfmt, glen,
                     d_pairs, d_counts, d_ends);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------- batched Dynprog_single_gap

// single_gap_simple (dynprog_single.c:344-419) for a window as long as its
// read, one wave per call: every column pairs read base r with genome base r;
// '*' columns (past the chromosome) push nothing; a base equal to either
// allele is a match ('*'), one consistent with either (consistent_array) an
// ambiguous pair (':'), else a mismatch (' ').  The pairs are written in the
// order of the returned list from its head (the last pushed, r = rlength,
// first); counts = {npairs, traceback_score, nmatches, nmismatches, 0, 0},
// npairs -1 when the call was not tried.  More than one mismatch makes the
// reference return NULL and go on to the DP; the caller sees it in counts.
__global__ __launch_bounds__(256) void gdp_sg_simple_kernel(const gdp_sg_dev *__restrict__ calls, int m,
                                                            const char *__restrict__ qbuf,
                                                            const char *__restrict__ gbuf, gdp_pair *__restrict__ pairs,
                                                            int32_t *__restrict__ counts) {
  __shared__ uint32_t consb[3 * 128 * 4];
  for (int j = threadIdx.x; j < 3 * 128 * 4; j += blockDim.x) consb[j] = gdp_tabConsBits[j];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (k >= m) return;
  const gdp_sg_dev c = calls[k];
  int32_t *cnt = counts + 6 * (int64_t)k;
  if (c.simple_off < 0) {
    if (lane == 0) {
      cnt[0] = -1;
      for (int x = 1; x < 6; x++) cnt[x] = 0;
    }
    return;
  }
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int row = (c.genestrand * 128) << 2;
  int n = 0, nm = 0, nmm = 0;
  gdp_pair *out = pairs + c.simple_off;
  for (int base = c.rl - 1; base >= 0; base -= 64) {
    const int i = base - lane;
    const bool valid = i >= 0;
    int c1 = 'N', c2 = '*', c2a = '*';
    if (valid) {
      c1 = qbuf[c.qoff + i];
      c2 = gbuf[c.goff + i];
      c2a = c2;  // one genome: the alternate allele is the reference's
    }
    const bool paired = valid && c2 != '*';
    int comp = ' ';
    if (paired) {
      const int c1uc = (c1 >= 'a' && c1 <= 'z') ? c1 - 32 : c1;
      const int r4 = row + ((c1uc & 127) << 2);
      if (c1uc == c2 || c1uc == c2a) comp = '*';
      else if (((consb[r4 + ((c2 & 127) >> 5)] >> (c2 & 31)) & 1) || ((consb[r4 + ((c2a & 127) >> 5)] >> (c2a & 31)) & 1))
        comp = ':';
    }
    const uint64_t ma = __ballot(paired && comp != ' '), mm_ = __ballot(paired && comp == ' ');
    nm += __popcll(ma);
    nmm += __popcll(mm_);
    const int qpos = c.roffset + i, gpos = c.goffset + i;
    // Pairpool_push drops pairs at negative coordinates (pairpool.c:189-191)
    const bool emit = paired && qpos >= 0 && gpos >= 0;
    const uint64_t em = __ballot(emit);
    if (emit) {
      gdp_pair &pr = out[n + __popcll(em & below)];
      pr.querypos = qpos;
      pr.genomepos = gpos;
      pr.cdna = (int8_t)c1;
      pr.comp = (int8_t)comp;
      pr.genome = (int8_t)c2;
      pr.genomealt = (int8_t)c2a;
      pr.gapp = 0;
      pr.genomejump = 0;
    }
    n += __popcll(em);
  }
  if (lane == 0) {
    cnt[0] = n;
    cnt[1] = TB_MATCH * nm + TB_MISMATCH * nmm;
    cnt[2] = nm;
    cnt[3] = nmm;
    cnt[4] = cnt[5] = 0;
  }
}

// Per call, which list the reference returns: single_gap_simple's when it
// was tried, found at most one mismatch and pushed a pair, else the DP
// walk's (a walk that pushed nothing is NULL).  len = the chosen list's
// length (-1 NULL), choice = 1 simple / 0 DP.
__global__ void gdp_sg_len_kernel(int m, const int32_t *__restrict__ cs, const int32_t *__restrict__ cd,
                                  int32_t *__restrict__ len, int32_t *__restrict__ choice) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const bool simple = cs[6 * k] > 0 && cs[6 * k + 3] <= 1;
  choice[k] = simple ? 1 : 0;
  len[k] = simple ? cs[6 * k] : (cd[6 * k] > 0 ? cd[6 * k] : -1);
}

// Copies each call's chosen list to its place in the compact output and
// writes its result, one workgroup per call.
__global__ void gdp_sg_join_kernel(int m, const gdp_sg_dev *__restrict__ calls, const gdp_tb_request *__restrict__ reqs,
                                   const gdp_pair *__restrict__ ps, const gdp_pair *__restrict__ pd,
                                   const int32_t *__restrict__ cs, const int32_t *__restrict__ cd,
                                   const int32_t *__restrict__ len, const int32_t *__restrict__ choice,
                                   const int64_t *__restrict__ off, gdp_pair *__restrict__ out,
                                   gdp_sg_result *__restrict__ res) {
  for (int k = blockIdx.x; k < m; k += gridDim.x) {
    const bool simple = choice[k] != 0;
    const int n = len[k];
    const gdp_pair *src = simple ? ps + calls[k].simple_off : pd + reqs[k].pair_off;
    for (int j = threadIdx.x; j < n; j += blockDim.x) out[off[k] + j] = src[j];
    if (threadIdx.x == 0) {
      const int32_t *c = simple ? cs + 6 * (int64_t)k : cd + 6 * (int64_t)k;
      gdp_sg_result r;
      r.pair_off = off[k];
      r.npairs = n;
      for (int x = 0; x < 5; x++) r.counters[x] = c[1 + x];
      r.simple = simple ? 1 : 0;
      r.pad = 0;
      res[k] = r;
    }
  }
}

extern "C" int gdp_launch_sg(int stage, int m, const gdp_sg_dev *calls, const char *qbuf, const char *gbuf,
                             gdp_pair *ps, int32_t *cs, const gdp_tb_request *reqs, const gdp_pair *pd,
                             const int32_t *cd, int32_t *len, int32_t *choice, const int64_t *off, gdp_pair *out,
                             gdp_sg_result *res, hipStream_t stream) {
  if (m <= 0) return 0;
  if (stage == 0)
    hipLaunchKernelGGL(gdp_sg_simple_kernel, dim3((m + 3) / 4), dim3(256), 0, stream, calls, m, qbuf, gbuf, ps, cs);
  else if (stage == 1)
    hipLaunchKernelGGL(gdp_sg_len_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, m, (const int32_t *)cs, cd,
                       len, choice);
  else
    hipLaunchKernelGGL(gdp_sg_join_kernel, dim3(gdp_imin(m, 8192)), dim3(64), 0, stream, m, calls, reqs,
                       (const gdp_pair *)ps, pd, (const int32_t *)cs, cd, (const int32_t *)len,
                       (const int32_t *)choice, off, out, res);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------- batched Dynprog_end{5,3}_gap

// The ends' post-processing after their walks (dynprog_end.c:1622-1650,
// 2214-2243), per call: flags = endalign (bits 0-1, Endalign_T) | 4 when
// require_pos_score_p.  counts = the walk's {npairs, traceback_score,
// nmatches, nmismatches, nopens, nindels}.  The reference returns NULL
// - for require_pos_score_p outside QUERYEND_NOGAPS (it tests the score it
//   has just set to 0), with every counter 0;
// - for QUERYEND_GAP and BEST_LOCAL when nmatches + 1 < nmismatches, with
//   traceback_score 0;
// and otherwise drops the pushed list's first pairs while they are indels
// (comp '-'); an empty list is NULL.  len = the kept pairs (-1 NULL), trim =
// the dropped ones; the counters are fixed up in place.
__global__ void gdp_eg_len_kernel(int m, const int32_t *__restrict__ flags, const gdp_tb_request *__restrict__ reqs,
                                  const gdp_pair *__restrict__ pd, int32_t *__restrict__ cd, int32_t *__restrict__ len,
                                  int32_t *__restrict__ trim) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  int32_t *c = cd + 6 * (int64_t)k;
  const int f = flags[k], endalign = f & 3;
  const bool gap_or_local = endalign == 0 || endalign == 3;  // QUERYEND_GAP, BEST_LOCAL
  const int n = c[0] > 0 ? c[0] : 0;
  int t = 0, l = -1;
  if ((f & 4) && endalign != 2) {
    for (int x = 1; x < 6; x++) c[x] = 0;
  } else if (gap_or_local && c[2] + 1 < c[3]) {
    c[1] = 0;
  } else {
    const gdp_pair *p = pd + reqs[k].pair_off;
    while (t < n && p[t].comp == '-') t++;
    l = n - t > 0 ? n - t : -1;
  }
  len[k] = l;
  trim[k] = t;
}

// Copies each call's kept pairs to the compact output in the order of the
// returned list from its head: the 5' caller returns the pushed list
// reversed, the 3' caller as pushed.  One workgroup per call.
__global__ void gdp_eg_join_kernel(int m, int five, const gdp_tb_request *__restrict__ reqs,
                                   const gdp_pair *__restrict__ pd, const int32_t *__restrict__ cd,
                                   const int32_t *__restrict__ best, const int32_t *__restrict__ len,
                                   const int32_t *__restrict__ trim, const int64_t *__restrict__ off,
                                   gdp_pair *__restrict__ out, gdp_eg_result *__restrict__ res) {
  for (int k = blockIdx.x; k < m; k += gridDim.x) {
    const int n = len[k], t = trim[k];
    const gdp_pair *src = pd + reqs[k].pair_off + t;
    for (int j = threadIdx.x; j < n; j += blockDim.x) out[off[k] + j] = src[five ? n - 1 - j : j];
    if (threadIdx.x == 0) {
      gdp_eg_result r;
      r.pair_off = off[k];
      r.npairs = n;
      for (int x = 0; x < 5; x++) r.counters[x] = cd[6 * (int64_t)k + 1 + x];
      r.finalscore = best[3 * (int64_t)k];
      r.advanced = 1;
      res[k] = r;
    }
  }
}

extern "C" int gdp_launch_eg(int stage, int m, int five, const int32_t *flags, const gdp_tb_request *reqs,
                             const gdp_pair *pd, int32_t *cd, const int32_t *best, int32_t *len, int32_t *trim,
                             const int64_t *off, gdp_pair *out, gdp_eg_result *res, hipStream_t stream) {
  if (m <= 0) return 0;
  if (stage == 0)
    hipLaunchKernelGGL(gdp_eg_len_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, m, flags, reqs, pd, cd, len,
                       trim);
  else
    hipLaunchKernelGGL(gdp_eg_join_kernel, dim3(gdp_imin(m, 8192)), dim3(64), 0, stream, m, five, reqs, pd,
                       (const int32_t *)cd, best, (const int32_t *)len, (const int32_t *)trim, off, out, res);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------ Dynprog_cdna_gap's join

// get_genomic_nt (pairpool.c:1016-1063) over the device genome, the
// alternate genome equal to it: '*' outside the chromosome, the complement
// on the minus strand.
__device__ __forceinline__ int gdp_genomic_nt(const void *genome, int fmt, uint32_t glen, uint32_t chroffset,
                                              uint32_t chrhigh, int watsonp, int gpos) {
  const uint32_t pos = watsonp ? chroffset + (uint32_t)gpos : chrhigh - (uint32_t)gpos;
  if (pos < chroffset || pos >= chrhigh) return '*';
  const int ch = genome_at(genome, fmt, glen, pos);
  return watsonp ? ch : compl_char(ch);
}

// Query and genome positions the bridge's cells leave between the two
// ends' walks (dynprog_cdna.c:1067-1068).
__device__ __forceinline__ void gdp_cd_jumps(const gdp_cd_dev &c, const int32_t *o, int &qj, int &gj) {
  qj = (c.rev_roffsetR - o[4]) - (c.roffsetL + o[3]) + 1;
  gj = (c.rev_goffset - o[2]) - (c.goffset + o[1]) + 1;
}

// Per call the length of the returned list: the right walk, 18 insertion
// pairs when 9 query and 9 genome positions are left (INSERT_PAIRS,
// dynprog_cdna.c:39, 1007-1030 and 1136-1159) or else one gap holder, the left walk; -1 (NULL)
// for the gap holder alone (1170-1173).
__global__ void gdp_cd_len_kernel(int m, const gdp_cd_dev *__restrict__ calls, const int32_t *__restrict__ bout,
                                  const int32_t *__restrict__ cd, int32_t *__restrict__ len) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  int qj, gj;
  gdp_cd_jumps(calls[k], bout + 5 * (int64_t)k, qj, gj);
  const int nR = gdp_imax(cd[12 * (int64_t)k], 0), nL = gdp_imax(cd[12 * (int64_t)k + 6], 0);
  const int tot = nR + (qj == 9 && gj == 9 ? 18 : 1) + nL;
  len[k] = tot == 1 ? -1 : tot;
}

// Writes each call's returned list from its head: the right end's walk
// reversed (its pairs were reversed before the rest was pushed, and the
// whole list at the end), then the insertion pairs or the gap holder and the
// left end's walk in push order (dynprog_cdna.c:1062-1175).  One workgroup
// per call.
__global__ void gdp_cd_join_kernel(int m, const gdp_cd_dev *__restrict__ calls, const char *__restrict__ qbuf,
                                   const void *__restrict__ genome, int fmt, uint32_t glen,
                                   const int32_t *__restrict__ bout, const gdp_tb_request *__restrict__ reqs,
                                   const gdp_pair *__restrict__ pd, const int32_t *__restrict__ cd,
                                   const int32_t *__restrict__ len, const int64_t *__restrict__ off,
                                   gdp_pair *__restrict__ out, gdp_cd_result *__restrict__ res) {
  for (int k = blockIdx.x; k < m; k += gridDim.x) {
    const gdp_cd_dev c = calls[k];
    const int32_t *o = bout + 5 * (int64_t)k;
    int qj, gj;
    gdp_cd_jumps(c, o, qj, gj);
    const bool ins = qj == 9 && gj == 9;
    const int n = len[k];
    const int32_t *cR = cd + 12 * (int64_t)k, *cL = cR + 6;
    if (n > 0) {
      const int nR = gdp_imax(cR[0], 0), nL = gdp_imax(cL[0], 0), mid = ins ? 18 : 1;
      const gdp_pair *pR = pd + reqs[2 * (int64_t)k].pair_off, *pL = pd + reqs[2 * (int64_t)k + 1].pair_off;
      gdp_pair *dst = out + off[k];
      for (int j = threadIdx.x; j < nR; j += blockDim.x) dst[j] = pR[nR - 1 - j];
      for (int j = threadIdx.x; j < mid; j += blockDim.x) {
        gdp_pair p;
        if (!ins) {  // Pairpool_push_gapholder (pairpool.c:374-452), knownp false
          p.querypos = -1;
          p.genomepos = -1;
          p.cdna = p.comp = p.genome = p.genomealt = ' ';
          p.gapp = 1;
          p.genomejump = (int16_t)gj;
        } else if (j < 9) {  // the cDNA insertion, query positions descending
          const int q = c.rev_roffsetR - o[4] - j;
          p.querypos = q;
          p.genomepos = c.rev_goffset - o[2] + 1;
          p.cdna = qbuf[c.qoffL + (q - c.roffsetL)];
          p.comp = '~';  // SHORTGAP_COMP (comp.h:10)
          p.genome = p.genomealt = ' ';
          p.gapp = 0;
          p.genomejump = 0;
        } else {  // the genome insertion, genome positions descending
          const int g = c.rev_goffset - o[2] - (j - 9);
          const int ch = gdp_genomic_nt(genome, fmt, glen, c.chroffset, c.chrhigh, c.watsonp, g);
          p.querypos = c.roffsetL + o[3];
          p.genomepos = g;
          p.cdna = ' ';
          p.comp = '~';
          p.genome = p.genomealt = (int8_t)ch;
          p.gapp = 0;
          p.genomejump = 0;
        }
        dst[nR + j] = p;
      }
      for (int j = threadIdx.x; j < nL; j += blockDim.x) dst[nR + mid + j] = pL[j];
    }
    if (threadIdx.x == 0) {
      gdp_cd_result r;
      r.pair_off = off[k];
      r.npairs = n;
      r.traceback_score = cR[1] + cL[1];
      r.incomplete = ins ? 0 : 1;
      r.advanced = 1;
      r.finalscore = o[0];
      r.pad = 0;
      res[k] = r;
    }
  }
}

extern "C" int gdp_launch_cd(int stage, int m, const gdp_cd_dev *calls, const char *qbuf, const void *genome,
                             int fmt, uint32_t glen, const int32_t *bout, const gdp_tb_request *reqs,
                             const gdp_pair *pd, const int32_t *cd, int32_t *len, const int64_t *off, gdp_pair *out,
                             gdp_cd_result *res, hipStream_t stream) {
  if (m <= 0) return 0;
  if (stage == 0)
    hipLaunchKernelGGL(gdp_cd_len_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, m, calls, bout, cd, len);
  else
    hipLaunchKernelGGL(gdp_cd_join_kernel, dim3(gdp_imin(m, 8192)), dim3(64), 0, stream, m, calls, qbuf, genome, fmt,
                       glen, bout, reqs, pd, cd, (const int32_t *)len, off, out, res);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Copies between device memory and pinned host memory mapped into the
// device's address space, as kernels on the caller's stream.  A copy engine
// serves the streams of a process in submission order, so a batch's upload
// would wait for another stream's download; a kernel keeps every batch's
// transfers inside its own stream.
__global__ void gdp_copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, int64_t n16) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n16; k += (int64_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
  __threadfence_system();
}

// A served batch's results to the host: the final scores and counters
// (head_bytes from the start of the download area), then each request's
// pair records only (counts[6 * i] of them at its pair_off), one workgroup
// per request.
__global__ void gdp_serve_download_kernel(const uint8_t *__restrict__ d_down, uint8_t *__restrict__ h_dst,
                                          int64_t head_bytes, int64_t off_counts, int64_t off_pairs,
                                          const gdp_tb_request *__restrict__ reqs, int n) {
  const int64_t nh = head_bytes / 16;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nh; k += (int64_t)gridDim.x * blockDim.x)
    ((uint4 *)h_dst)[k] = ((const uint4 *)d_down)[k];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int32_t np = ((const int32_t *)(d_down + off_counts))[6 * i];
    const int64_t po = reqs[i].pair_off;
    const uint4 *src = (const uint4 *)(d_down + off_pairs) + po;
    uint4 *dst = (uint4 *)(h_dst + off_pairs) + po;
    for (int j = threadIdx.x; j < np; j += blockDim.x) dst[j] = src[j];
  }
  __threadfence_system();
}

// A batch's walk results to host memory mapped into the device's address
// space: each request's counts and only the pair records its walk wrote,
// at the same offsets as on the device.  One wave per request.
__global__ __launch_bounds__(256) void gdp_download_results_kernel(const gdp_tb_request *__restrict__ reqs, int n,
                                                                   const int32_t *__restrict__ counts,
                                                                   const gdp_pair *__restrict__ pairs,
                                                                   int32_t *__restrict__ h_counts,
                                                                   gdp_pair *__restrict__ h_pairs) {
  const int lane = threadIdx.x & 63;
  for (int k = blockIdx.x * 4 + (threadIdx.x >> 6); k < n; k += gridDim.x * 4) {
    const int32_t *c = counts + 6 * (int64_t)k;
    if (lane < 6) h_counts[6 * (int64_t)k + lane] = c[lane];
    const int np = c[0] < reqs[k].max_pairs ? c[0] : reqs[k].max_pairs;
    const uint4 *src = (const uint4 *)(pairs + reqs[k].pair_off);
    uint4 *dst = (uint4 *)(h_pairs + reqs[k].pair_off);
    for (int j = lane; j < np; j += 64) dst[j] = src[j];
  }
  __threadfence_system();
}

extern "C" int gdp_launch_download_results(const gdp_tb_request *d_reqs, int n, const int32_t *d_counts,
                                           const gdp_pair *d_pairs, int32_t *h_counts, gdp_pair *h_pairs,
                                           hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gdp_download_results_kernel, dim3(gdp_imin((n + 3) / 4, 65536)), dim3(256), 0, stream, d_reqs,
                     n, d_counts, d_pairs, h_counts, h_pairs);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_copy(void *dst, const void *src, int64_t bytes, hipStream_t stream) {
  if (bytes <= 0) return 0;
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return -1;
  const int64_t n16 = bytes / 16;
  const int blocks = (int)gdp_imin(1024, (int)((n16 + 255) / 256));
  hipLaunchKernelGGL(gdp_copy_kernel, dim3(blocks), dim3(256), 0, stream, (uint4 *)dst, (const uint4 *)src, n16);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_serve_download(const uint8_t *d_down, uint8_t *h_dst, int64_t head_bytes,
                                         int64_t off_counts, int64_t off_pairs, const gdp_tb_request *d_reqs, int n,
                                         hipStream_t stream) {
  if (n <= 0) return 0;
  if ((head_bytes & 15) || (off_pairs & 15) || ((uintptr_t)d_down & 15) || ((uintptr_t)h_dst & 15)) return -1;
  hipLaunchKernelGGL(gdp_serve_download_kernel, dim3(gdp_imin(n, 4096)), dim3(64), 0, stream, d_down, h_dst,
                     head_bytes, off_counts, off_pairs, d_reqs, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_segments(const void *d_genome, int fmt, uint32_t glen, const gdp_segment *d_segs, int n,
                                   char *d_out, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gdp_segments_kernel, dim3((n + 3) / 4), dim3(256), 0, stream, d_genome, fmt, glen, d_segs, n,
                     d_out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_best_endpoint(const void *d_descs, const int32_t *d_ul, int m, int queryend_indels,
                                        const int16_t *d_scores, int32_t *d_out, const gdp_tb_request *d_tmpl,
                                        gdp_tb_request *d_tbout, hipStream_t stream) {
  if (m <= 0) return 0;
  hipLaunchKernelGGL(gdp_best_endpoint_kernel, dim3((m + 256 / TBG - 1) / (256 / TBG)), dim3(256), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_ul, m, queryend_indels, d_scores, d_out, d_tmpl, d_tbout);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_bridge_cdna(const void *d_descs, const int32_t *d_quads, int m, const int16_t *d_scores,
                                      int32_t *d_out, const gdp_tb_request *d_tmpl, gdp_tb_request *d_tbout,
                                      hipStream_t stream) {
  if (m <= 0) return 0;
  hipLaunchKernelGGL(gdp_bridge_cdna_kernel<GDP_BRIDGE_LDS_TINY>, dim3(m), dim3(64), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_quads, m, d_scores, d_out, d_tmpl, d_tbout);
  hipLaunchKernelGGL(gdp_bridge_cdna_kernel<GDP_BRIDGE_LDS_SMALL>, dim3(m), dim3(64), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_quads, m, d_scores, d_out, d_tmpl, d_tbout);
  hipLaunchKernelGGL(gdp_bridge_cdna_kernel<GDP_BRIDGE_LDS>, dim3(m), dim3(64), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_quads, m, d_scores, d_out, d_tmpl, d_tbout);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_bridge_intron(const void *d_descs, const gdp_intron_gap *d_gaps, int m,
                                        const int16_t *d_scores, const char *d_gbuf, const double *d_probs,
                                        int32_t *d_out, double *d_outp, const gdp_tb_request *d_tmpl,
                                        gdp_tb_request *d_tbout, hipStream_t stream) {
  if (m <= 0) return 0;
  hipLaunchKernelGGL(gdp_bridge_intron_kernel<GDP_INTRON_SITES_SMALL>, dim3(m), dim3(64), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_gaps, m, d_scores, d_gbuf, d_probs, d_out, d_outp, d_tmpl,
                     d_tbout);
  hipLaunchKernelGGL(gdp_bridge_intron_kernel<GDP_INTRON_SITES>, dim3(m), dim3(64), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_gaps, m, d_scores, d_gbuf, d_probs, d_out, d_outp, d_tmpl,
                     d_tbout);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_export(const void *d_descs, int idx, const uint32_t *d_dirs, const int16_t *d_scores,
                                 uint8_t *d_ddirs, int16_t *d_dscores, hipStream_t stream) {
  hipLaunchKernelGGL(gdp_export_kernel, dim3(64), dim3(256), 0, stream, (const gdp_dev_problem *)d_descs, idx, d_dirs,
                     d_scores, d_ddirs, d_dscores);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_upload_std_tables(const uint32_t *SP, const int8_t *SC) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabSP), SP, sizeof(uint32_t) * 4 * 128 * 4) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabSC), SC, 128) != hipSuccess) return -1;
  return 0;
}

extern "C" int gdp_upload_tables(const uint32_t *PQ, const uint32_t *PG, const int8_t *NT, const uint8_t *cons) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabPQ), PQ, sizeof(uint32_t) * 4 * 128) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabPG), PG, sizeof(uint32_t) * 4 * 128) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabNT), NT, 128) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabCons), cons, 3 * 128 * 128) != hipSuccess) return -1;
  uint32_t bits[3 * 128 * 4] = {0};
  for (int k = 0; k < 3 * 128 * 128; k++)
    if (cons[k]) bits[(k >> 7) * 4 + ((k & 127) >> 5)] |= 1u << (k & 31);
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabConsBits), bits, sizeof bits) != hipSuccess) return -1;
  return 0;
}
