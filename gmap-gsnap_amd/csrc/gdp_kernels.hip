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
};

__device__ __forceinline__ void load_tables(gdp_tables_lds &t) {
  for (int k = threadIdx.x; k < 4 * 128; k += blockDim.x) {
    t.PQ[k] = gdp_tabPQ[k];
    t.PG[k] = gdp_tabPG[k];
  }
  for (int k = threadIdx.x; k < 128; k += blockDim.x) t.NT[k] = gdp_tabNT[k];
}

// Row profile: the five scores of row `row` against the column classes.
__device__ __forceinline__ uint32_t row_profile(const gdp_tables_lds &t, const gdp_dev_problem &p, const char *qbuf,
                                                const char *gbuf, int row) {
  const int mt = p.mismatchtype;
  if (p.kind != GDP_KIND_LOWER) {
    int ch = row == 0 ? 'N' : (qbuf[p.qoff + row - 1] & 127);
    return t.PQ[mt * 128 + ch];
  }
  if (row == 0) return t.PG[mt * 128 + (p.width == 8 ? 4 : 'N')];
  int g = gbuf[p.goff + row - 1] & 127, ga = gbuf[p.galtoff + row - 1] & 127;
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

__global__ __launch_bounds__(GDP_BLOCK) void gdp_fill_kernel(const gdp_dev_problem *__restrict__ probs, int n,
                                                             const char *__restrict__ qbuf,
                                                             const char *__restrict__ gbuf, uint32_t *__restrict__ dirs,
                                                             int16_t *__restrict__ scores, int32_t *__restrict__ final_out) {
  __shared__ gdp_tables_lds tab;
  __shared__ uint32_t ring_all[GDP_GROUPS_PER_BLOCK * GDP_RING];
  load_tables(tab);
  __syncthreads();

  const int lane = threadIdx.x & (GDP_GROUP - 1);
  const int grp = threadIdx.x / GDP_GROUP;
  const int pid = blockIdx.x * GDP_GROUPS_PER_BLOCK + grp;
  if (pid >= n) return;
  uint32_t *ring = ring_all + grp * GDP_RING;

  const gdp_dev_problem p = probs[pid];
  if (p.layout != 0) return;  // swept by gdp_fill2_kernel
  const gdp_geom G = gdp_make_geom(p);
  const gdp_consts K = gdp_make_consts(p);
  const bool full = G.full;
  const bool keep = p.score_off >= 0;

  // Final cell (rl, gl) of a full-band problem; a single writer per problem.
  const bool final_in_band = full && G.R <= G.C + G.lb && G.R >= G.C - G.ub;
  if (lane == 0 && !final_in_band) final_out[pid] = INT_MIN;

  int T = 0, r = lane;
  int cs = 0, ce = -1;
  uint32_t prof = 0;
  if (r <= G.R) {
    cs = gdp_cstart(G, r);
    ce = gdp_cend(G, r);
    prof = row_profile(tab, p, qbuf, gbuf, r);
  }
  int Hl = K.INIT_H, El = K.E_INIT, uprev = K.NEG, Hout = K.NEG, Fout = GDP_F_FLOOR;
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
        prof = row_profile(tab, p, qbuf, gbuf, r);
      }
      c = s - T - lane;
    }
    const bool act = (r <= G.R) && (c >= cs) && (c <= ce);

    // Upper neighbour (r-1, c) from the previous step.
    int Hup = __builtin_amdgcn_update_dpp(Hout, Hout, 0x111, 0xF, 0xF, false);
    int Fup = __builtin_amdgcn_update_dpp(Fout, Fout, 0x111, 0xF, 0xF, false);
    if (lane == 0) {
      uint32_t ent = ring[c & (GDP_RING - 1)];
      Hup = (int)(int16_t)(ent & 0xffffu);
      Fup = (int)(int16_t)(ent >> 16) - 128;
    }
    const int Hdg = uprev;  // (r-1, c-1), received one step ago
    uprev = Hup;

    int H = K.NEG, F = GDP_F_FLOOR, E = El;
    uint32_t nib = 0;
    if (act) {
      int x, xa;
      col_class(tab, p, qbuf, gbuf, c, x, xa);
      gdp_cell(K, G, r, c, cs, Hl, El, Hdg, Hup, Fup, prof, x, xa, H, F, E, nib);
      if (full && r == G.R && c == G.C) final_out[pid] = H;
      if (lane == GDP_GROUP - 1) ring[c & (GDP_RING - 1)] = (uint32_t)(H & 0xffff) | ((uint32_t)(F + 128) << 16);
    }
    if (keep) scores[p.score_off + (int64_t)s * GDP_GROUP + lane] = (int16_t)(act ? H : 0);
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
}

// ------------------------------------------------------------------ lookups

__device__ __forceinline__ uint32_t nib_at(const uint32_t *dirs, const gdp_dev_problem &p, const gdp_geom &G, int row,
                                           int col) {
  if (row < 0 || row > G.R || col < 0 || col > G.C) return 0;
  if (col < gdp_cstart(G, row) || col > gdp_cend(G, row)) return 0;  // not computed: DIAG
  if (p.layout == 1) {  // paired sweep: byte 0 Egap, 1 HORIZ, 2 Fgap, 3 VERT; bit = step & 7
    const int g = 1 << p.glog;
    const int t = gdp_p_step(g, p.P, p.lband, row, col);
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

struct tb_out {
  gdp_pair *pairs;
  int n, cap;
};

// Pairpool_push (pairpool.c:179-259): pairs at negative coordinates are dropped.
__device__ __forceinline__ void push_pair(tb_out &o, int qpos, int gpos, int cdna, int comp, int g, int ga) {
  if (qpos < 0 || gpos < 0) return;
  if (o.n < o.cap) {
    gdp_pair pr;
    pr.querypos = qpos;
    pr.genomepos = gpos;
    pr.cdna = (int8_t)cdna;
    pr.comp = (int8_t)comp;
    pr.genome = (int8_t)g;
    pr.genomealt = (int8_t)ga;
    pr.gapp = 0;
    pr.genomejump = 0;
    o.pairs[o.n] = pr;
  }
  o.n++;
}

__device__ __forceinline__ int compl_char(int ch) {
  switch (ch) {
  case 'A': return 'T';
  case 'C': return 'G';
  case 'G': return 'C';
  case 'T': return 'A';
  case 'a': return 't';
  case 'c': return 'g';
  case 'g': return 'c';
  case 't': return 'a';
  default: return ch;
  }
}

// get_genomic_nt (pairpool.c:1016-1063) over an ASCII device genome.
__device__ int genomic_nt(int *alt, const char *genome, uint32_t glen, int genomicpos, uint32_t chroffset,
                          uint32_t chrhigh, int watsonp) {
  uint32_t pos = watsonp ? chroffset + (uint32_t)genomicpos : chrhigh - (uint32_t)genomicpos;
  if (pos < chroffset || pos >= chrhigh) {
    *alt = '*';
    return '*';
  }
  int ch = (genome != nullptr && pos < glen) ? genome[pos] : 'N';
  if (!watsonp) ch = compl_char(ch);
  *alt = ch;
  return ch;
}

__global__ void gdp_traceback_kernel(const gdp_dev_problem *__restrict__ probs, const gdp_tb_request *__restrict__ reqs,
                                     int nreq, const char *__restrict__ qbuf, const char *__restrict__ gbuf,
                                     const uint32_t *__restrict__ dirs, const char *__restrict__ genome, uint32_t glen,
                                     gdp_pair *__restrict__ pairs, int32_t *__restrict__ counts) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nreq) return;
  const gdp_tb_request q = reqs[k];
  int32_t *cnt = counts + 6 * (int64_t)k;
  const gdp_dev_problem p = probs[q.fill];
  const gdp_geom G = gdp_make_geom(p);
  if (q.r < 0 || q.r > p.rl || q.c < 0 || q.c > p.gl) {  // start cell outside the matrix
    cnt[0] = -1;
    return;
  }
  const bool lower = p.kind == GDP_KIND_LOWER;
  const int revp = q.revp;
  tb_out out{pairs + q.pair_off, 0, q.max_pairs};
  int sc = 0, nm = 0, nmm = 0, nop = 0, nind = 0;
  int r = q.r, c = q.c;

  // Direction nibble at query row rr, genome column cc.
  auto dir = [&](int rr, int cc) -> uint32_t { return lower ? nib_at(dirs, p, G, cc, rr) : nib_at(dirs, p, G, rr, cc); };
  auto qch = [&](int rr) -> int { return qbuf[p.qoff + rr - 1]; };

  const bool ops = (q.flags & GDP_TB_OPS) != 0;
  auto push_op = [&](int type, int rr, int cc, int arg, int comp) {
    if (out.n < out.cap) {
      gdp_pair pr;
      pr.querypos = rr;
      pr.genomepos = cc;
      pr.cdna = (int8_t)type;
      pr.comp = (int8_t)comp;
      pr.genome = 0;
      pr.genomealt = 0;
      pr.gapp = 0;
      pr.genomejump = (int16_t)arg;
      out.pairs[out.n] = pr;
    }
    out.n++;
  };
  // Pairpool_add_queryskip (pairpool.c:980-1010).
  auto queryskip = [&](int rr, int cc, int dist) {
    if (ops) {
      push_op('V', rr, cc, dist, 0);
      return;
    }
    int qc = rr - 1, gc = cc - 1, step = -1;
    if (revp) {
      qc = -qc;
      gc = -gc;
      step = +1;
    }
    for (int j = 0; j < dist; j++) {
      // rsequence[qc] is q(|qc| + 1) in sweep order for both orientations.
      push_pair(out, q.queryoffset + qc, q.genomeoffset + gc, qch((qc < 0 ? -qc : qc) + 1), '-', ' ', ' ');
      qc += step;
    }
  };
  // Pairpool_add_genomeskip (pairpool.c:1067-1180) with genomesequence NULL.
  auto genomeskip = [&](int rr, int cc, int dist) -> bool {
    int qc = rr - 1, left = cc - dist, right = cc - 1, step = -1;
    if (revp) {
      int t = left;
      qc = -qc;
      left = -right;
      right = -t;
      step = +1;
    }
    if (ops) {
      push_op('H', rr, cc, dist, 0);
      return dist < MICROINTRON_LENGTH;
    }
    if (dist < MICROINTRON_LENGTH) {
      int gc = revp ? left : right;
      for (int j = 0; j < dist; j++) {
        int alt, ch = genomic_nt(&alt, genome, glen, q.genomeoffset + gc, q.chroffset, q.chrhigh, q.watsonp);
        push_pair(out, q.queryoffset + qc, q.genomeoffset + gc, ' ', '-', ch, alt);
        gc += step;
      }
      return true;
    }
    if (out.n < out.cap) {  // Pairpool_push_gapholder (pairpool.c:374-452)
      gdp_pair pr;
      pr.querypos = -1;
      pr.genomepos = -1;
      pr.cdna = ' ';
      pr.comp = ' ';
      pr.genome = ' ';
      pr.genomealt = ' ';
      pr.gapp = 1;
      pr.genomejump = (int16_t)dist;
      out.pairs[out.n] = pr;
    }
    out.n++;
    return false;
  };
  auto diag_pair = [&](int rr, int cc) {
    int qc = rr - 1, gc = cc - 1;
    if (revp) {
      qc = -qc;
      gc = -gc;
    }
    int c1 = qch(rr), c1uc = (c1 >= 'a' && c1 <= 'z') ? c1 - 32 : c1;
    int c2 = gbuf[p.goff + cc - 1], c2a = gbuf[p.galtoff + cc - 1];
    if (c2 == '*') return;
    int comp;
    if (c1uc == c2 || c1uc == c2a) {
      comp = '*';
    } else if (gdp_tabCons[(q.genestrand * 128 + (c1uc & 127)) * 128 + (c2 & 127)] ||
               gdp_tabCons[(q.genestrand * 128 + (c1uc & 127)) * 128 + (c2a & 127)]) {
      comp = ':';
    } else {
      comp = ' ';
    }
    if (comp == ' ') {
      sc += TB_MISMATCH;
      nmm += 1;
    } else {
      sc += TB_MATCH;
      nm += 1;
    }
    if (ops) push_op('D', rr, cc, 0, comp);
    else push_pair(out, q.queryoffset + qc, q.genomeoffset + gc, c1, comp, c2, c2a);
  };

  if (p.kind == GDP_KIND_FULL) {
    while (r > 0 && c > 0) {
      uint32_t d = dir(r, c);
      if ((d & GDP_DIR_H) && !(d & GDP_DIR_V)) {
        int dist = 1;
        while (c > 0 && (dir(r, c--) & GDP_DIR_E)) dist++;
        if (genomeskip(r, c + dist, dist)) {
          sc += TB_TOPEN + dist * TB_TINDEL;
          nop += 1;
          nind += dist;
        }
      } else if (d & GDP_DIR_V) {
        int dist = 1;
        while (r > 0 && (dir(r--, c) & GDP_DIR_F)) dist++;
        queryskip(r + dist, c, dist);
        sc += TB_QOPEN + dist * TB_QINDEL;
        nop += 1;
        nind += dist;
      } else {
        diag_pair(r, c);
        r--;
        c--;
      }
    }
    if (r == 0 && c == 0) {
    } else if (c == 0) {
      queryskip(r, 1, r);
      sc += TB_QOPEN + r * TB_QINDEL;
      nop += 1;
      nind += r;
    } else {
      int dist = c;
      if (genomeskip(1, c, dist)) {
        sc += TB_TOPEN + dist * TB_TINDEL;
        nop += 1;
        nind += dist;
      }
    }
  } else if (p.kind == GDP_KIND_UPPER) {
    while (r > 0 && c > 0) {
      if (dir(r, c) & (GDP_DIR_H | GDP_DIR_V)) {
        int dist = 1;
        while (dir(r, c--) & GDP_DIR_E) dist++;
        if (genomeskip(r, c + dist, dist)) {
          sc += TB_TOPEN + dist * TB_TINDEL;
          nop += 1;
          nind += dist;
        }
      } else {
        diag_pair(r, c);
        r--;
        c--;
      }
    }
    if (c != 0) {
      int dist = c;
      if (genomeskip(1, c, dist)) {
        sc += TB_TOPEN + dist * TB_TINDEL;
        nop += 1;
        nind += dist;
      }
    }
  } else {
    while (r > 0 && c > 0) {
      if (dir(r, c) & (GDP_DIR_H | GDP_DIR_V)) {
        int dist = 1;
        while (dir(r--, c) & GDP_DIR_E) dist++;
        queryskip(r + dist, c, dist);
        sc += TB_QOPEN + dist * TB_QINDEL;
        nop += 1;
        nind += dist;
      } else {
        diag_pair(r, c);
        r--;
        c--;
      }
    }
    if (r != 0) {
      queryskip(r, 1, r);
      sc += TB_QOPEN + r * TB_QINDEL;
      nop += 1;
      nind += r;
    }
  }
  cnt[0] = out.n;
  cnt[1] = sc;
  cnt[2] = nm;
  cnt[3] = nmm;
  cnt[4] = nop;
  cnt[5] = nind;
}

// ------------------------------------------------------------ best end cell

__global__ void gdp_best_endpoint_kernel(const gdp_dev_problem *__restrict__ probs, const int32_t *__restrict__ ul, int m,
                                         int queryend_indels, const int16_t *__restrict__ scores,
                                         int32_t *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const gdp_dev_problem pu = probs[ul[2 * k]], pl = probs[ul[2 * k + 1]];
  const gdp_geom Gu = gdp_make_geom(pu), Gl = gdp_make_geom(pl);
  if (pu.score_off < 0 || pl.score_off < 0 || pu.kind != GDP_KIND_UPPER || pl.kind != GDP_KIND_LOWER) {
    out[3 * k] = INT_MIN;  // scores were not kept, or the pair is not (upper, lower)
    out[3 * k + 1] = -1;
    out[3 * k + 2] = -1;
    return;
  }
  const int rl = pu.rl, gl = pu.gl, lband = pl.lband, uband = pu.uband, late = pu.jump_late;
  int best = queryend_indels ? (pu.width == 8 ? -128 : -32768) : 0;
  int bestr = queryend_indels ? rl : 0, bestc = 0;
  for (int r = queryend_indels ? rl : 1; r <= rl; r++) {
    int clo = gdp_imax(1, r - lband), chigh = gdp_imin(gl, r + uband), c;
    for (c = clo; c < r; c++) {
      int v = score_at(scores, pl, Gl, c, r);  // lower: rows are genome columns
      if (gdp_wins(v, best, late)) {
        bestr = r;
        bestc = c;
        best = v;
      }
    }
    for (; c <= chigh; c++) {
      int v = score_at(scores, pu, Gu, r, c);
      if (gdp_wins(v, best, late)) {
        bestr = r;
        bestc = c;
        best = v;
      }
    }
  }
  out[3 * k] = best;
  out[3 * k + 1] = bestr;
  out[3 * k + 2] = bestc;
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
    if (dscores != nullptr) dscores[k] = (int16_t)((comp && p.score_off >= 0) ? score_at(scores, p, G, row, col) : 0);
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
    sc[k] = (T)((comp && p.score_off >= 0) ? score_at(scores, p, G, row, col) : 0);
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
template <int G>
static int p2_grid() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
  if (cached[dev] == 0) {
    int ncu = 0, per = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gdp_fill2_kernel<G>, P2_THREADS, p2cfg<G>::LDS_BYTES) !=
        hipSuccess)
      return -1;
    cached[dev] = ncu * (per > 0 ? per : 1);
  }
  return cached[dev];
}

extern "C" int gdp_launch_fill(const void *d_descs, int n, const char *d_qbuf, const char *d_gbuf, uint32_t *d_dirs,
                               int16_t *d_scores, int32_t *d_final, hipStream_t stream) {
  if (n <= 0) return 0;
  int blocks = (n + GDP_GROUPS_PER_BLOCK - 1) / GDP_GROUPS_PER_BLOCK;
  hipLaunchKernelGGL(gdp_fill_kernel, dim3(blocks), dim3(GDP_BLOCK), 0, stream, (const gdp_dev_problem *)d_descs, n,
                     d_qbuf, d_gbuf, d_dirs, d_scores, d_final);
  if (hipGetLastError() != hipSuccess) return -1;
  const int g16 = p2_grid<16>(), g64 = p2_grid<64>();
  if (g16 <= 0 || g64 <= 0) return -1;
  hipLaunchKernelGGL(gdp_fill2_kernel<16>, dim3(g16), dim3(P2_THREADS), p2cfg<16>::LDS_BYTES, stream,
                     (const uint8_t *)d_descs, n, d_qbuf, d_gbuf, d_dirs, d_final);
  if (hipGetLastError() != hipSuccess) return -1;
  hipLaunchKernelGGL(gdp_fill2_kernel<64>, dim3(g64), dim3(P2_THREADS), p2cfg<64>::LDS_BYTES, stream,
                     (const uint8_t *)d_descs, n, d_qbuf, d_gbuf, d_dirs, d_final);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_traceback(const void *d_descs, const gdp_tb_request *d_reqs, int nreq, const char *d_qbuf,
                                    const char *d_gbuf, const uint32_t *d_dirs, const char *d_genome, uint32_t glen,
                                    gdp_pair *d_pairs, int32_t *d_counts, hipStream_t stream) {
  if (nreq <= 0) return 0;
  hipLaunchKernelGGL(gdp_traceback_kernel, dim3((nreq + 127) / 128), dim3(128), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_reqs, nreq, d_qbuf, d_gbuf, d_dirs, d_genome, glen, d_pairs,
                     d_counts);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_best_endpoint(const void *d_descs, const int32_t *d_ul, int m, int queryend_indels,
                                        const int16_t *d_scores, int32_t *d_out, hipStream_t stream) {
  if (m <= 0) return 0;
  hipLaunchKernelGGL(gdp_best_endpoint_kernel, dim3((m + 127) / 128), dim3(128), 0, stream,
                     (const gdp_dev_problem *)d_descs, d_ul, m, queryend_indels, d_scores, d_out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_launch_export(const void *d_descs, int idx, const uint32_t *d_dirs, const int16_t *d_scores,
                                 uint8_t *d_ddirs, int16_t *d_dscores, hipStream_t stream) {
  hipLaunchKernelGGL(gdp_export_kernel, dim3(64), dim3(256), 0, stream, (const gdp_dev_problem *)d_descs, idx, d_dirs,
                     d_scores, d_ddirs, d_dscores);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int gdp_upload_tables(const uint32_t *PQ, const uint32_t *PG, const int8_t *NT, const uint8_t *cons) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabPQ), PQ, sizeof(uint32_t) * 4 * 128) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabPG), PG, sizeof(uint32_t) * 4 * 128) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabNT), NT, 128) != hipSuccess) return -1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(gdp_tabCons), cons, 3 * 128 * 128) != hipSuccess) return -1;
  return 0;
}
