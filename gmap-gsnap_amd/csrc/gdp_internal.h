// gdp_internal.h -- device-side problem layout and shared constants of the
// banded affine-gap DP core (GMAP dynprog_simd.c semantics) for gfx950.
#pragma once
#include <stdint.h>

#include "../../include/gmapdp.h"

// Lanes that sweep one problem.  Rows of the DP matrix are dealt round robin
// to these lanes; a lane passes H and F of its cell to the next lane one step
// later through a DPP row shift, and the last lane hands row r+15 to lane 0
// through an LDS ring indexed by column.
#define GDP_GROUP 16
// Ring entries per group (power of two, > batch period - 15).  Problems
// whose batch period exceeds GDP_RING (bands wider than about 1000 columns)
// take the wide launch of the single sweep, with rings of up to
// GDP_RING_WIDE entries.
#define GDP_RING 1024
#define GDP_RING_WIDE 8192

// Packed direction nibble, one per cell, same bits as the test oracle:
// bit0 nogap==HORIZ, bit1 nogap==VERT, bit2 Egap!=DIAG, bit3 Fgap!=DIAG.
#define GDP_DIR_H 1
#define GDP_DIR_V 2
#define GDP_DIR_E 4
#define GDP_DIR_F 8

// F is carried in 32 bits by the reference (dynprog_simd.c:2995); values at
// or below this floor can never win a comparison against H + open, so F is
// clamped here to keep it in 16 bits (+128 bias) in the LDS ring.
#define GDP_F_FLOOR (-32896)

#if defined(__HIPCC__)
#define GDP_HD __host__ __device__ __forceinline__
#else
#define GDP_HD inline
#endif

// Device copy of one problem (96 bytes).  Row/column roles: for the full band
// and the upper triangle rows are query positions r = 0..rl and columns are
// genome positions c = 0..gl; the lower triangle is swept transposed (rows are
// genome positions, columns query positions), which is how the reference's
// column-blocked lower kernel visits it (dynprog_simd.c:8744-8860).
//
// layout 0: the problem is swept alone by a 16-lane group (gdp_fill_kernel);
// layout 1: it is one half of a lane pair swept together with problems of the
// same band and tie rule (gdp_fill2_kernel), over the padded sweep sR x sC
// with batch period P and 2^glog lanes per group;
// layout 2: a one-sided fill with a band under 32 columns, swept by one lane
// (gdp_fill_narrow_kernel) over its in-band cells only; the 64 problems of a
// wave interleave their direction words: word kk (cells 8kk .. 8kk+7 right
// of the diagonal) of row r at dir_off + (r * P + kk) * 64, P words per row.
struct gdp_dev_problem {
  int64_t qoff;     // q(r) = qbuf[qoff + r - 1], r = 1..rl (already in sweep order)
  int64_t goff;     // g(c) = gbuf[goff + c - 1]
  int64_t galtoff;  // galt(c) = gbuf[galtoff + c - 1]
  int64_t dir_off;  // first 32-bit direction word of this problem
  int64_t score_off;// first int16 score slot (one per lane-step), -1 = not kept
  int32_t rl, gl;
  int16_t lband, uband;
  int8_t kind, width, mismatchtype, jump_late;
  int8_t open, extend, layout, glog;
  int32_t sR, sC;   // layout 1: padded sweep extent shared by the wave
  int32_t P;        // layout 1: steps between row batches
  int32_t nsteps;   // layout 1: steps of the sweep (multiple of 8)
  int32_t saturation;  // GDP_KIND_STD: floor of the stored scores
  int8_t std_flags;    // GDP_KIND_STD: bit 0 upperp, bit 1 lowerp
  int8_t keep_mode;    // gdp_problem.keep_scores: 1 every score, 2/3 the best end cell only
  int16_t slb;         // layout 1: the lower band of the wave's sweep (its largest lband), which fixes
                       // the step of each cell (gdp_p_step); the problem's own band is lband/uband
  int32_t pad[3];
};
static_assert(sizeof(gdp_dev_problem) == 96, "gdp_dev_problem must stay 96 bytes");
static_assert(sizeof(gdp_problem) == 48, "gdp_problem is 48 bytes");
static_assert(sizeof(gdp_tb_request) == 48, "gdp_tb_request is 48 bytes");
// Sets the message gdp_last_error() returns on the calling thread.
void gdp_set_error(const char *msg);
// Bytes of a plan of n problems that the kernels read: the descriptors, the
// header, the wave lists and the index lists (gdp_plan_bytes reserves room
// for one wave per problem).
int64_t gdp_plan_used_bytes(const void *descs, int32_t n);

static_assert(sizeof(gdp_pair) == 16, "gdp_pair is 16 bytes");
static_assert(sizeof(gdp_intron_gap) == 32, "gdp_intron_gap is 32 bytes");
static_assert(sizeof(gdp_segment) == 24, "gdp_segment is 24 bytes");
static_assert(sizeof(gdp_align_request) == 64, "gdp_align_request is 64 bytes");
static_assert(sizeof(gdp_align_result) == 72, "gdp_align_result is 72 bytes");

// One wave of the paired fill: 64 lanes = 64/G groups of G lanes, each lane
// carrying two problems in the low and high 16 bits of its registers.  Slot
// 2*g + h is group g, half h.  Every problem of a wave has the same tie rule;
// rows and columns are padded to the largest problem.  The sweep is laid out
// for the largest lower and upper bands of the wave (lb, ub); in a mixed wave
// (mixed = 1) the problems' own bands differ, lie within [lbmin, lb] and
// [ubmin, ub], and each problem's band edges are applied in its own half.
struct gdp_wave {
  int32_t prob[8];      // problem index per slot, -1 = empty
  int32_t fin_t[8];     // steps at which final cells (rl, gl) are computed, ascending, INT32_MAX pads
  int8_t fin_slot[8];   // slot of each fin_t entry
  int32_t R, C;         // padded sweep: rows 0..R, columns 0..C
  int32_t P, nb, nsteps;
  int16_t lb, ub;
  int8_t late, glog, mixed, pad0;
  int16_t lbmin, ubmin;
  int32_t pad[6];
};
static_assert(sizeof(gdp_wave) == 128, "gdp_wave must stay 128 bytes");

// Header written after the n problem descriptors, followed by the wave
// lists of the classes (16-lane groups with P <= 128, 16-lane groups with
// P <= 256, 64-lane groups) and one int32 list per class of waves handed to
// the general-lookup pass.  The fill checks magic and n before it touches
// the lists, so a descriptor array handed over with another n never does.
#define GDP_PAIR_MAGIC 0x6764703270616972ull
#define GDP_NCLS 3
struct gdp_pair_hdr {
  uint64_t magic;           // GDP_PAIR_MAGIC ^ n
  int32_t n;
  int32_t ncls;
  int32_t nw[4];            // waves per class
  int64_t off[4];           // byte offset of each class's wave list from the descriptor base
  int64_t gen_off[4];       // byte offset of each class's general-lookup list
  int32_t queue[4][2];      // per class: next wave to hand out, waves that have left the queue
  int32_t gen_queue[4][2];  // the same for the general-lookup pass
  int32_t gen_count[4];     // waves in each general-lookup list
  int64_t sink_off;         // direction words written by the empty slots of partial waves
  int32_t n1;               // problems swept alone (layout 0) by the SIMD-rule single sweep
  int32_t n1std;            // standard fills, listed after them
  int64_t list1_off;        // byte offset of their int32 index list
  int32_t n2;               // narrow one-sided sweeps (layout 2)
  int32_t pad2;
  int64_t list2_off;        // byte offset of their int32 index list
  int32_t pad[10];
};
static_assert(sizeof(gdp_pair_hdr) == 256, "gdp_pair_hdr must stay 256 bytes");

// One call of the batched single-gap caller on the device: its read in the
// batch's query buffer, its genome window in the batch's genome buffer, and
// where single_gap_simple writes its pairs (-1: glength != rlength, not
// tried).
struct gdp_sg_dev {
  int64_t qoff;
  int64_t goff;
  int64_t simple_off;
  int32_t rl, gl, roffset, goffset;
  int32_t genestrand, pad;
};
static_assert(sizeof(gdp_sg_dev) == 48, "gdp_sg_dev is 48 bytes");

// One Dynprog_cdna_gap call as its join kernels read it: where the left
// end's read starts in the query buffer, and the query and genome
// positions of both ends (rev_goffset = goffset + glength - 1).
struct gdp_cd_dev {
  int64_t qoffL;
  int32_t roffsetL, rev_roffsetR, goffset, rev_goffset;
  uint32_t chroffset, chrhigh;
  int32_t watsonp, pad;
};
static_assert(sizeof(gdp_cd_dev) == 40, "gdp_cd_dev is 40 bytes");

// Row stride (direction nibbles) of a narrow sweep over band width w
// (uband of an upper fill, lband of a lower one), 0 when w + 1 > 32.
GDP_HD int gdp_narrow_stride(int w) { return w < 8 ? 8 : (w < 16 ? 16 : (w < 32 ? 32 : 0)); }

// Sweep geometry shared by the fill, traceback, export and scan kernels.
struct gdp_geom {
  int R, C;        // last row / column index
  int lb, ub;      // band: cells with c - ub <= r <= c + lb (one-sided: lb unused)
  int S;           // emulated SIMD block height (32 int8 / 16 int16)
  int full;        // 1 = full band, 0 = one-sided triangle
  int P;           // steps between consecutive row batches
  int nsteps;      // steps the sweep takes
};


GDP_HD int gdp_imax(int a, int b) { return a > b ? a : b; }
GDP_HD int gdp_imin(int a, int b) { return a < b ? a : b; }

GDP_HD gdp_geom gdp_make_geom(const gdp_dev_problem &p) {
  gdp_geom g;
  // the scalar standard fill is a full band without emulated row blocks
  g.full = p.kind == GDP_KIND_FULL || p.kind == GDP_KIND_STD;
  g.R = p.kind == GDP_KIND_LOWER ? p.gl : p.rl;
  g.C = p.kind == GDP_KIND_LOWER ? p.rl : p.gl;
  g.S = p.kind == GDP_KIND_STD ? 1 : (p.width == 8 ? 32 : 16);
  if (g.full) {
    g.lb = p.lband;
    g.ub = p.uband;
  } else {
    g.lb = 0;
    g.ub = p.kind == GDP_KIND_UPPER ? p.uband : p.lband;
  }
  // A lane's next row starts P+16 steps after its current row.  The largest
  // span of one row (including the emulated out-of-band prefix of its block)
  // bounds P from below; 16 keeps the handoff through the ring causal.  The
  // standard fill adds one step so that lane 0 idles at least one step
  // before each new row and reads the diagonal of its first cell from the
  // ring under that row's column.
  int span = (g.full ? g.lb + g.ub : g.ub) + g.S - GDP_GROUP + (p.kind == GDP_KIND_STD ? 1 : 0);
  g.P = gdp_imax(GDP_GROUP, span);
  int nb = (g.R + GDP_GROUP) / GDP_GROUP;
  g.nsteps = (nb - 1) * g.P + (GDP_GROUP - 1) + g.C + 1;
  return g;
}

// First and last column a row computes.  The full band starts at the column
// where the row's emulated block starts (dynprog_simd.c:3221-3222); the
// one-sided triangles start at the block's first row (7853).
GDP_HD int gdp_cstart(const gdp_geom &g, int r) {
  int blk = r & -g.S;  // S is 1, 16 or 32
  return g.full ? gdp_imax(0, blk - g.lb) : blk;
}
GDP_HD int gdp_cend(const gdp_geom &g, int r) { return gdp_imin(g.C, r + g.ub); }

// Step at which row r, column c is computed, and its direction word/nibble.
GDP_HD int gdp_step(const gdp_geom &g, int r, int c) {
  return (r / GDP_GROUP) * g.P + (r % GDP_GROUP) + c;
}
GDP_HD int64_t gdp_dir_words(const gdp_geom &g) { return (int64_t)((g.nsteps + 7) / 8) * GDP_GROUP; }

// Paired sweep (layout 1).  Row r = G*b + i is handled by lane i of its
// group during batch b; its cell in column c is computed at step
//   t = b*P + i + (c - cs0(b)),   cs0(b) = max(0, G*b - lb),
// so cell (r-1, c) is always one step older than (r, c).  Each lane spends
// P steps per row: from cs0(b) on (the emulated 16-row block start lies at
// or after it) to past the top band edge.
// P >= 2G + 8 keeps the ring distance D = P - (G - 1) - (cs0(b) - cs0(b-1))
// at 8 or more, so a block of 8 steps never reads what it writes.
GDP_HD int gdp_p_period(int G, int lb, int ub) {
  int P = gdp_imax(2 * G + 8, G + lb + ub);
  return (P + 7) & ~7;
}
GDP_HD int gdp_p_cs0(int G, int lb, int b) { return gdp_imax(0, G * b - lb); }
GDP_HD int gdp_p_nsteps(int G, int P, int R) {
  int nb = (R + G) / G;
  return (nb * P + G + 7) & ~7;
}
GDP_HD int gdp_p_step(int G, int P, int lb, int r, int c) {
  int b = r / G;
  return b * P + (r % G) + (c - gdp_p_cs0(G, lb, b));
}

// Per-problem constants of the cell update.
struct gdp_consts {
  int NEG, POS;  // saturation bounds of the kernel width
  int o, e;      // gap open / extend (negative)
  int late;      // 1: ties go to the gap (>=), 0: strict (>)
  int INIT_H;    // H seen left of a block start: -inf - open (dynprog_simd.c:3238)
  int E_INIT;    // E at a block start: -inf, or -inf + 1 for jump-early (3506-3510)
};

GDP_HD gdp_consts gdp_make_consts(const gdp_dev_problem &p) {
  gdp_consts k;
  k.NEG = p.width == 8 ? -128 : -32768;
  k.POS = p.width == 8 ? 127 : 32767;
  k.o = p.open;
  k.e = p.extend;
  k.late = p.jump_late ? 1 : 0;
  k.INIT_H = k.NEG - k.o;
  k.E_INIT = k.late ? k.NEG : k.NEG + 1;
  return k;
}

GDP_HD int gdp_clampw(int x, int lo, int hi) { return gdp_imin(gdp_imax(x, lo), hi); }
GDP_HD int gdp_wins(int a, int b, int late) { return late ? (a >= b) : (a > b); }
GDP_HD int gdp_field4(uint32_t w, int x) { return ((int)(w << (28 - 4 * x))) >> 28; }

// One cell (r, c) of a row whose first computed column is cs.  Inputs: the
// lane's left neighbour (Hl, El), the diagonal Hdg = H(r-1, c-1), the upper
// neighbour (Hup, Fup) = (H, F)(r-1, c), the row profile and the column's two
// genome classes.  Outputs H, F (F clamped at GDP_F_FLOOR), E and the
// direction nibble.  Rules of the reference's AVX2 build:
//   E/H update .................... dynprog_simd.c:3316-3388 (6862-6919)
//   column 0 and row 0 ............ 3164-3189, 3259-3291
//   block start (E, H seeds) ...... 3221-3254
//   bottom band edge .............. 3390-3416
//   F chain, top band edge ........ 3419-3476
//   one-sided E cap, diagonal ..... 4478-4626, 7867-7960
template <bool FULL>
GDP_HD void gdp_cell_k(const gdp_consts &k, const gdp_geom &G, int r, int c, int cs, int Hl, int El, int Hdg, int Hup,
                       int Fup, uint32_t prof, int x, int xa, int &H, int &F, int &E, uint32_t &nib) {
  const bool rowstart = c == cs;
  const int hl = rowstart ? k.INIT_H : Hl;
  int el = rowstart ? k.E_INIT : El;
  int diag = (rowstart && (r % G.S) != 0) ? k.INIT_H : Hdg;
  if (r == 0) diag = c == 0 ? 0 : k.NEG;
  const int sc = gdp_imax(gdp_field4(prof, x), gdp_field4(prof, xa));
  const int capE = (!FULL && r >= c) ? k.NEG : k.POS;
  el = gdp_imin(el, capE);
  const int T1 = gdp_clampw(hl + k.o, k.NEG, k.POS);
  int dE = gdp_wins(el, T1, k.late);
  E = gdp_imin(gdp_clampw(gdp_imax(el, T1) + k.e, k.NEG, k.POS), capE);
  int Hd;
  if (FULL && c == 0) Hd = r == 0 ? 0 : k.NEG;
  else Hd = gdp_clampw(diag + sc, k.NEG, k.POS);
  int dH = gdp_wins(E, Hd, k.late);
  int H0 = gdp_imax(Hd, E);
  int dF = 0, dV = 0;
  F = GDP_F_FLOOR;
  if (FULL) {
    const int kd = c - r;
    if (kd == -G.lb && c > 0) {
      H0 = Hd;
      dE = 0;
      dH = 0;
    }
    if (kd >= -G.lb) {
      if (kd == G.ub) {
        F = -32768 + k.o + k.e;
        H = H0;
      } else {
        const int cg = r == 0 ? -32768 : Fup;
        const int last = r == 0 ? -32768 : Hup;
        dF = gdp_wins(cg, last + k.o, k.late);
        F = dF ? cg + k.e : last + k.o + k.e;
        dV = gdp_wins(F, H0, k.late);
        H = dV ? F : H0;
      }
    } else {
      H = H0;
    }
    F = gdp_imax(F, GDP_F_FLOOR);
  } else {
    if (c == r) {
      dE = 0;
      dH = 0;
    }
    H = H0;
  }
  nib = (uint32_t)((dV ? GDP_DIR_V : (dH ? GDP_DIR_H : 0)) | (dE ? GDP_DIR_E : 0) | (dF ? GDP_DIR_F : 0));
}

GDP_HD void gdp_cell(const gdp_consts &k, const gdp_geom &G, int r, int c, int cs, int Hl, int El, int Hdg, int Hup,
                     int Fup, uint32_t prof, int x, int xa, int &H, int &F, int &E, uint32_t &nib) {
  if (G.full) gdp_cell_k<true>(k, G, r, c, cs, Hl, El, Hdg, Hup, Fup, prof, x, xa, H, F, E, nib);
  else gdp_cell_k<false>(k, G, r, c, cs, Hl, El, Hdg, Hup, Fup, prof, x, xa, H, F, E, nib);
}
