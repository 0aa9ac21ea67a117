// gdp_api.cpp -- host side of the batched C ABI (include/gmapdp.h): score
// tables, problem validation and planning, and the launch wrappers.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cstdarg>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <algorithm>
#include <utility>
#include <vector>

#include "gdp_internal.h"
#include "gdp_tables.h"

extern "C" int gdp_launch_fill(const void *, int, uint32_t, const char *, const char *, uint32_t *, int16_t *,
                               int32_t *, hipStream_t);
extern "C" int gdp_launch_traceback(const void *, const gdp_tb_request *, int, const char *, const char *,
                                    const uint32_t *, const void *, int, uint32_t, gdp_pair *, int32_t *, int32_t *,
                                    hipStream_t);
extern "C" int gdp_launch_segments(const void *, int, uint32_t, const gdp_segment *, int, char *, hipStream_t);
extern "C" int gdp_launch_best_endpoint(const void *, const int32_t *, int, int, const int16_t *, int32_t *,
                                        const gdp_tb_request *, gdp_tb_request *, hipStream_t);
extern "C" int gdp_launch_bridge_cdna(const void *, const int32_t *, int, const int16_t *, int32_t *,
                                      const gdp_tb_request *, gdp_tb_request *, hipStream_t);
extern "C" int gdp_launch_bridge_intron(const void *, const gdp_intron_gap *, int, const int16_t *, const char *,
                                        const double *, int32_t *, double *, const gdp_tb_request *, gdp_tb_request *,
                                        hipStream_t);
extern "C" int gdp_launch_export(const void *, int, const uint32_t *, const int16_t *, uint8_t *, int16_t *,
                                 hipStream_t);
extern "C" int gdp_launch_export_ref(const void *, int, int, const uint32_t *, const int16_t *, int, void *, void *,
                                     void *, void *, hipStream_t);
extern "C" int gdp_upload_tables(const uint32_t *, const uint32_t *, const int8_t *, const uint8_t *);
extern "C" int gdp_upload_std_tables(const uint32_t *, const int8_t *);

namespace {
thread_local std::string g_err;
std::mutex g_mu;
bool g_init = false;
bool g_std_ok = false;  // the installed tables fit the standard fill's score classes
uint32_t g_SP[4 * 128 * 4];
int8_t g_SC[128];

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(const char *what, hipError_t e) {
  return fail(GDP_EHIP, "%s: %s", what, hipGetErrorString(e));
}
}  // namespace

extern "C" const char *gdp_last_error(void) { return g_err.c_str(); }

void gdp_set_error(const char *msg) { g_err = msg; }

static int upload_current_tables() {
  uint32_t PQ[4 * 128], PG[4 * 128];
  int8_t NT[128];
  gdp_tables_pack(PQ, PG, NT);
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(GDP_EHIP, "gdp_init: no HIP device (%s)", hipGetErrorString(e));
  if (gdp_upload_tables(PQ, PG, NT, gdp_tables_consistent()) != 0)
    return fail(GDP_EHIP, "gdp_init: table upload failed");
  g_std_ok = gdp_tables_pack_std(g_SP, g_SC) > 0;
  if (g_std_ok && gdp_upload_std_tables(g_SP, g_SC) != 0) return fail(GDP_EHIP, "gdp_init: table upload failed");
  g_init = true;
  return GDP_OK;
}

extern "C" int gdp_init(int mode) {
  std::lock_guard<std::mutex> lk(g_mu);
  gdp_tables_build(mode);
  return upload_current_tables();
}

extern "C" int gdp_init_tables(const int16_t *pd, const int32_t *nt, const uint8_t *cons) {
  if (!pd || !nt || !cons) return fail(GDP_EINVAL, "gdp_init_tables: null table");
  std::lock_guard<std::mutex> lk(g_mu);
  gdp_tables_set(pd, nt, cons);
  return upload_current_tables();
}

extern "C" void gdp_host_tables(int mode) {
  std::lock_guard<std::mutex> lk(g_mu);
  gdp_tables_build(mode);
  g_std_ok = gdp_tables_pack_std(g_SP, g_SC) > 0;
}

extern "C" int32_t gdp_descriptor_bytes(void) { return (int32_t)sizeof(gdp_dev_problem); }

// Bands wider than the matrix narrowed to one past it.  No cell changes: a
// lower band past rl already holds every row of every column (r - lband < 0
// for every row, so each row starts at column 0 and the bottom edge
// r = c + lband lies below the matrix), and an upper band past gl never
// reaches its top edge r = c - uband; a one-sided band past its triangle's
// length covers the triangle.  Each band stays past the matrix, so every
// comparison of a cell with a band edge comes out as before.
static void narrow_bands(const gdp_problem &p, int &lb, int &ub) {
  if (p.kind == GDP_KIND_FULL || p.kind == GDP_KIND_STD) {
    lb = gdp_imin(lb, p.rl + 1);
    ub = gdp_imin(ub, p.gl + 1);
  } else if (p.kind == GDP_KIND_UPPER) {
    ub = gdp_imin(ub, p.gl + 1);
  } else if (p.kind == GDP_KIND_LOWER) {
    lb = gdp_imin(lb, p.rl + 1);
  }
}

static int validate(const gdp_problem &p, int i) {
  if (p.kind < 0 || p.kind > 3) return fail(GDP_EINVAL, "problem %d: bad kind %d", i, p.kind);
  if (p.kind == GDP_KIND_STD) {
    if (p.width != 32) return fail(GDP_EINVAL, "problem %d: the standard fill has width 32", i);
    if (p.std_flags & ~(GDP_STD_UPPER | GDP_STD_LOWER)) return fail(GDP_EINVAL, "problem %d: bad std_flags", i);
    if (p.saturation < -32768 || p.saturation > 32767)
      return fail(GDP_EINVAL, "problem %d: saturation %d outside [-32768, 32767]", i, p.saturation);
    if (!g_std_ok)
      return fail(GDP_EINVAL, "problem %d: the installed score tables need more than 32 genome classes", i);
  } else {
    if (p.width != 8 && p.width != 16) return fail(GDP_EINVAL, "problem %d: width must be 8 or 16", i);
    if (p.std_flags != 0 || p.saturation != 0)
      return fail(GDP_EINVAL, "problem %d: std_flags and saturation belong to GDP_KIND_STD", i);
  }
  if (p.mismatchtype < 0 || p.mismatchtype > 3) return fail(GDP_EINVAL, "problem %d: bad mismatchtype", i);
  if (p.rl < 1 || p.gl < 1 || p.rl > GDP_MAX_LEN || p.gl > GDP_MAX_LEN)
    return fail(GDP_ETOOBIG, "problem %d: lengths %d x %d outside [1, %d]", i, p.rl, p.gl, GDP_MAX_LEN);
  if (p.lband < 0 || p.uband < 0) return fail(GDP_EINVAL, "problem %d: negative band", i);
  if (p.keep_scores < 0 || p.keep_scores > 3) return fail(GDP_EINVAL, "problem %d: keep_scores must be 0-3", i);
  if (p.keep_scores >= 2 && p.kind != GDP_KIND_UPPER && p.kind != GDP_KIND_LOWER)
    return fail(GDP_EINVAL, "problem %d: keep_scores %d (best end cell) needs an upper or lower fill", i,
                p.keep_scores);
  if (p.open > 0 || p.open < -127 || p.extend > 0 || p.extend < -127)
    return fail(GDP_EINVAL, "problem %d: gap penalties must lie in [-127, 0]", i);
  int lb = p.lband, ub = p.uband;
  narrow_bands(p, lb, ub);
  int span = p.kind == GDP_KIND_FULL || p.kind == GDP_KIND_STD ? lb + ub : (p.kind == GDP_KIND_UPPER ? ub : lb);
  const int maxspan = p.kind == GDP_KIND_STD ? GDP_MAX_SPAN_STD : GDP_MAX_SPAN;
  if (span > maxspan) return fail(GDP_ETOOBIG, "problem %d: band span %d > %d", i, span, maxspan);
  return GDP_OK;
}

extern "C" int64_t gdp_plan_bytes(int32_t n) {
  if (n < 0) n = 0;
  return (int64_t)n * (int64_t)(sizeof(gdp_dev_problem) + sizeof(gdp_wave) + 3 * sizeof(int32_t)) +
         (int64_t)sizeof(gdp_pair_hdr);
}

// Full-band int16 fills without kept scores are swept in pairs, 2*(64/G) per
// wave; the rest (one-sided triangles, int8, kept scores, very wide bands)
// each get a 16-lane group of their own.  16-lane groups serve bands up to
// g16_max_span; wider bands, and queries of at least g64_min_rows rows (a
// sweep a quarter as long, for callers that wait on single long problems),
// take 64-lane groups.
static gdp_plan_params g_default_params = {240, GDP_MAX_LEN + 1, 0, 0};

static int pair_lanes(const gdp_problem &p, const gdp_plan_params &pp) {
  if (p.kind != GDP_KIND_FULL || p.width != 16 || p.keep_scores) return 0;
  const int W = p.lband + p.uband;
  if (W > 960) return 0;
  if (W <= pp.g16_max_span && p.rl < pp.g64_min_rows) return 16;
  return 64;
}

// Wave class: the ring of a class holds at least P step slots.
static int pair_class(int G, int P) { return G == 64 ? 2 : (P <= 128 ? 0 : 1); }

// Sorts indices by a key computed once per index, ties in their present
// order: what a stable sort of the indices by that key gives.  An LSD radix
// sort over 11-bit digits, each pass stable, skipping the digits in which
// all keys agree.
template <class KeyFn>
static void sort_by_key(std::vector<int32_t> &ix, KeyFn key) {
  const size_t n = ix.size();
  if (n < 2) return;
  std::vector<uint64_t> k(n), k2(n);
  std::vector<int32_t> v2(n);
  uint64_t all_or = 0, all_and = ~0ull;
  for (size_t j = 0; j < n; j++) {
    k[j] = key(ix[j]);
    all_or |= k[j];
    all_and &= k[j];
  }
  const uint64_t varies = all_or ^ all_and;
  constexpr int D = 11, NB = 1 << D;
  std::vector<uint32_t> cnt(NB);
  for (int sh = 0; sh < 64; sh += D) {
    if (((varies >> sh) & (NB - 1)) == 0) continue;
    std::fill(cnt.begin(), cnt.end(), 0u);
    for (size_t j = 0; j < n; j++) cnt[(k[j] >> sh) & (NB - 1)]++;
    uint32_t s = 0;
    for (int b = 0; b < NB; b++) {
      const uint32_t c = cnt[b];
      cnt[b] = s;
      s += c;
    }
    for (size_t j = 0; j < n; j++) {
      const uint32_t pos = cnt[(k[j] >> sh) & (NB - 1)]++;
      k2[pos] = k[j];
      v2[pos] = ix[j];
    }
    k.swap(k2);
    ix.swap(v2);
  }
}

extern "C" int gdp_plan_tuned(const gdp_problem *probs_in, int32_t n, int32_t flags, const gdp_plan_params *params,
                              void *descs_out, int64_t *dir_words, int64_t *score_slots) {
  if (n < 0 || (n > 0 && probs_in == nullptr) || descs_out == nullptr)
    return fail(GDP_EINVAL, "gdp_plan: bad arguments");
  // the problems with their bands narrowed to the matrix: waves group
  // problems by the bands they are swept with
  std::vector<gdp_problem> narrowed(probs_in, probs_in + n);
  for (gdp_problem &p : narrowed) {
    if (p.lband < 0 || p.uband < 0 || p.rl < 1 || p.gl < 1) continue;  // validate refuses it
    int lb = p.lband, ub = p.uband;
    narrow_bands(p, lb, ub);
    p.lband = (int16_t)lb;
    p.uband = (int16_t)ub;
  }
  const gdp_problem *probs = narrowed.data();
  const gdp_plan_params pp = params ? *params : g_default_params;
  if (pp.g16_max_span < 0 || pp.g16_max_span > 240 || pp.g64_min_rows < 1 || pp.mix_span < 0 ||
      pp.mix_span > 64 || pp.mix_order < 0 || pp.mix_order > 1)
    return fail(GDP_EINVAL,
                "gdp_plan: g16_max_span must lie in [0, 240], g64_min_rows be positive, mix_span in [0, 64] and "
                "mix_order 0 or 1");
  uint8_t *base = (uint8_t *)descs_out;
  gdp_dev_problem *d = (gdp_dev_problem *)descs_out;
  int64_t dw = 0, ss = 0;
  std::vector<int32_t> paired;
  for (int32_t i = 0; i < n; i++) {
    const gdp_problem &p = probs[i];
    int rc = validate(p, i);
    if (rc) return rc;
    gdp_dev_problem q;
    memset(&q, 0, sizeof q);
    q.qoff = p.qoff;
    q.goff = p.goff;
    q.galtoff = p.galtoff;
    q.rl = p.rl;
    q.gl = p.gl;
    q.lband = p.lband;
    q.uband = p.uband;
    q.kind = p.kind;
    q.width = p.width;
    q.mismatchtype = p.mismatchtype;
    q.jump_late = p.jump_late ? 1 : 0;
    q.open = p.open;
    q.extend = p.extend;
    q.saturation = p.saturation;
    q.std_flags = p.std_flags;
    q.keep_mode = p.keep_scores;
    q.score_off = -1;
    if (!(flags & GDP_PLAN_SINGLE) && pair_lanes(p, pp)) {
      d[i] = q;
      paired.push_back(i);
      continue;
    }
    gdp_geom G = gdp_make_geom(q);
    q.dir_off = dw;
    const int nstride = (p.kind == GDP_KIND_UPPER || p.kind == GDP_KIND_LOWER) && p.keep_scores != GDP_KEEP_ALL &&
                                !(flags & GDP_PLAN_SINGLE)
                            ? gdp_narrow_stride(G.ub)
                            : 0;
    if (nstride) {  // one lane sweeps its in-band cells; words allocated per wave below
      q.layout = 2;
      q.P = nstride;
    } else {
      dw += gdp_dir_words(G);
    }
    if (p.keep_scores == 1) {
      q.score_off = ss;
      ss += (int64_t)G.nsteps * GDP_GROUP;
    } else if (p.keep_scores >= 2) {
      q.score_off = ss;  // the best end cell's record: score, r, c
      ss += GDP_BEST_SLOTS;
    }
    d[i] = q;
  }

  // Waves: problems with the same tie rule and band, in order of size; with
  // pp.mix_span > 0 the problems left over from full waves of one band go
  // to mixed waves, whose problems' lower and upper bands each lie within
  // mix_span of the wave's largest (the sweep is laid out for those).
  // key: lane groups, tie rule, lband, uband, rl, gl (bands and lengths are
  // below 2^15 after narrowing, so each fits its field)
  auto band_key = [&](int32_t i) {
    const gdp_problem &x = probs[i];
    return ((uint64_t)(pair_lanes(x, pp) == 64) << 63) | ((uint64_t)(x.jump_late != 0) << 62) |
           ((uint64_t)(uint16_t)x.lband << 46) | ((uint64_t)(uint16_t)x.uband << 30) | ((uint64_t)x.rl << 15) |
           (uint64_t)x.gl;
  };
  sort_by_key(paired, band_key);
  std::vector<gdp_wave> waves[GDP_NCLS];
  auto emit = [&](const int32_t *ix, int cnt, int G) {
    gdp_wave w;
    memset(&w, 0, sizeof w);
    for (int s = 0; s < 8; s++) {
      w.prob[s] = -1;
      w.fin_t[s] = INT32_MAX;
    }
    int R = 0, C = 0, lb = 0, ub = 0, lbmin = INT32_MAX, ubmin = INT32_MAX;
    for (int j = 0; j < cnt; j++) {
      const gdp_problem &x = probs[ix[j]];
      R = gdp_imax(R, x.rl);
      C = gdp_imax(C, x.gl);
      lb = gdp_imax(lb, x.lband);
      ub = gdp_imax(ub, x.uband);
      lbmin = gdp_imin(lbmin, x.lband);
      ubmin = gdp_imin(ubmin, x.uband);
    }
    // A half whose lower band is below the wave's starts its rows and meets
    // its bottom edge up to lb - lbmin steps later than the sweep's own band
    // would; the period keeps those events inside their batch.
    const int P = gdp_imax(gdp_p_period(G, lb, ub), (2 * G + 8 + (lb - lbmin) + 7) & ~7);
    w.R = R;
    w.C = C;
    w.P = P;
    w.nb = (R + G) / G;
    w.nsteps = gdp_p_nsteps(G, P, R);
    w.lb = (int16_t)lb;
    w.ub = (int16_t)ub;
    w.lbmin = (int16_t)lbmin;
    w.ubmin = (int16_t)ubmin;
    w.mixed = (lbmin != lb || ubmin != ub) ? 1 : 0;
    w.late = probs[ix[0]].jump_late ? 1 : 0;
    w.glog = G == 16 ? 4 : 6;
    std::pair<int32_t, int8_t> fin[8];
    int nfin = 0;
    for (int s = 0; s < cnt; s++) {
      const int idx = ix[s];
      const gdp_problem &x = probs[idx];
      w.prob[s] = idx;
      gdp_dev_problem &q = d[idx];
      q.layout = 1;
      q.glog = (int8_t)w.glog;
      q.sR = R;
      q.sC = C;
      q.P = P;
      q.nsteps = w.nsteps;
      q.slb = (int16_t)lb;
      q.dir_off = dw;
      dw += (int64_t)(w.nsteps / 8) * G;
      if (x.rl <= x.gl + x.lband && x.rl >= x.gl - x.uband)
        fin[nfin++] = {gdp_p_step(G, P, lb, x.rl, x.gl), (int8_t)s};
    }
    std::sort(fin, fin + nfin);
    for (int j = 0; j < nfin; j++) {
      w.fin_t[j] = fin[j].first;
      w.fin_slot[j] = fin[j].second;
    }
    waves[pair_class(G, P)].push_back(w);
  };
  std::vector<int32_t> rest;  // problems left for mixed waves
  size_t k = 0;
  if (pp.mix_span > 0 && pp.mix_order == 1) {
    // Bucketed mixed waves.  A wave costs (its padded row batches) x (the
    // period of its largest bands), so its problems should agree in rows as
    // much as in bands: problems are binned by lband / B and uband / B and
    // taken in order of rl within a bin, full waves first; what is left of
    // the bins is binned again by 2B.
    auto bucketed = [&](std::vector<int32_t> &list, int B, bool full_only, std::vector<int32_t> &left) {
      sort_by_key(list, [&](int32_t i) {
        const gdp_problem &x = probs[i];
        return ((uint64_t)(pair_lanes(x, pp) == 64) << 63) | ((uint64_t)(x.jump_late != 0) << 62) |
               ((uint64_t)(x.lband / B) << 46) | ((uint64_t)(x.uband / B) << 30) | ((uint64_t)x.rl << 15) |
               (uint64_t)x.gl;
      });
      size_t a = 0;
      while (a < list.size()) {
        const gdp_problem &p0 = probs[list[a]];
        const int G = pair_lanes(p0, pp), slots = 2 * (64 / G);
        const int maxspan = G == 16 ? 240 : 960;
        int lb = p0.lband, ub = p0.uband;
        size_t e = a + 1;
        while (e < list.size() && e - a < (size_t)slots) {
          const gdp_problem &x = probs[list[e]];
          if (pair_lanes(x, pp) != G || (x.jump_late != 0) != (p0.jump_late != 0) || x.lband / B != p0.lband / B ||
              x.uband / B != p0.uband / B || gdp_imax(lb, x.lband) + gdp_imax(ub, x.uband) > maxspan)
            break;
          lb = gdp_imax(lb, x.lband);
          ub = gdp_imax(ub, x.uband);
          e++;
        }
        if (full_only && e - a < (size_t)slots) {
          for (size_t j = a; j < e; j++) left.push_back(list[j]);
        } else {
          emit(list.data() + a, (int)(e - a), G);
        }
        a = e;
      }
    };
    std::vector<int32_t> left, none;
    bucketed(paired, pp.mix_span, true, left);
    bucketed(left, 2 * pp.mix_span, false, none);
    k = paired.size();
  }
  while (k < paired.size()) {
    const gdp_problem &p0 = probs[paired[k]];
    const int G = pair_lanes(p0, pp), slots = 2 * (64 / G);
    size_t e = k;
    while (e < paired.size() && e - k < (size_t)slots) {
      const gdp_problem &x = probs[paired[e]];
      if (pair_lanes(x, pp) != G || (x.jump_late != 0) != (p0.jump_late != 0) || x.lband != p0.lband ||
          x.uband != p0.uband)
        break;
      e++;
    }
    if (pp.mix_span > 0 && e - k < (size_t)slots) {
      for (size_t j = k; j < e; j++) rest.push_back(paired[j]);
    } else {
      emit(paired.data() + k, (int)(e - k), G);
    }
    k = e;
  }
  if (!rest.empty()) {
    // neighbours in band width and lower band share a wave
    sort_by_key(rest, [&](int32_t i) {
      const gdp_problem &x = probs[i];
      return ((uint64_t)(pair_lanes(x, pp) == 64) << 63) | ((uint64_t)(x.jump_late != 0) << 62) |
             ((uint64_t)(x.lband + x.uband) << 44) | ((uint64_t)(uint16_t)x.lband << 28) | (uint64_t)x.rl;
    });
    size_t a = 0;
    while (a < rest.size()) {
      const gdp_problem &p0 = probs[rest[a]];
      const int G = pair_lanes(p0, pp), slots = 2 * (64 / G);
      // the sweep's span must fit the class's ring: 240 for 16-lane groups,
      // 960 for 64-lane groups
      const int maxspan = G == 16 ? 240 : 960;  // (and 2G + 8 + mix_span stays below the ring)
      int lb = p0.lband, ub = p0.uband, lbmin = lb, ubmin = ub;
      size_t e = a + 1;
      while (e < rest.size() && e - a < (size_t)slots) {
        const gdp_problem &x = probs[rest[e]];
        if (pair_lanes(x, pp) != G || (x.jump_late != 0) != (p0.jump_late != 0)) break;
        const int nlb = gdp_imax(lb, x.lband), nub = gdp_imax(ub, x.uband);
        const int nlbmin = gdp_imin(lbmin, x.lband), nubmin = gdp_imin(ubmin, x.uband);
        if (nlb - nlbmin > pp.mix_span || nub - nubmin > pp.mix_span || nlb + nub > maxspan ||
            2 * G + 8 + (nlb - nlbmin) > (G == 16 ? 256 : 1024))
          break;
        lb = nlb;
        ub = nub;
        lbmin = nlbmin;
        ubmin = nubmin;
        e++;
      }
      emit(rest.data() + a, (int)(e - a), G);
      a = e;
    }
  }
  // Empty slots of partial waves store into one shared sink.
  int64_t sink_words = 0;
  for (int c = 0; c < GDP_NCLS; c++)
    for (const gdp_wave &w : waves[c]) sink_words = std::max<int64_t>(sink_words, (int64_t)(w.nsteps / 8) * (1 << w.glog));
  const int64_t sink_off = dw;
  dw += sink_words;
  // Longest sweeps first, so the queue ends on short ones (ties in the
  // order the waves were formed).
  std::vector<int32_t> worder[GDP_NCLS];
  for (int c = 0; c < GDP_NCLS; c++) {
    worder[c].resize(waves[c].size());
    for (size_t j = 0; j < waves[c].size(); j++) worder[c][j] = (int32_t)j;
    sort_by_key(worder[c], [&](int32_t j) { return (uint64_t)(INT32_MAX - waves[c][j].nsteps); });
  }

  gdp_pair_hdr h;
  memset(&h, 0, sizeof h);
  h.magic = GDP_PAIR_MAGIC ^ (uint64_t)n;
  h.n = n;
  h.ncls = GDP_NCLS;
  const int64_t hoff = (int64_t)n * sizeof(gdp_dev_problem);
  int64_t off = hoff + (int64_t)sizeof(gdp_pair_hdr);
  for (int c = 0; c < GDP_NCLS; c++) {
    h.nw[c] = (int32_t)waves[c].size();
    h.off[c] = off;
    gdp_wave *wo = (gdp_wave *)(base + off);
    for (size_t j = 0; j < waves[c].size(); j++) memcpy(wo + j, &waves[c][worder[c][j]], sizeof(gdp_wave));
    off += (int64_t)waves[c].size() * (int64_t)sizeof(gdp_wave);
  }
  for (int c = 0; c < GDP_NCLS; c++) {
    h.gen_off[c] = off;
    off += (int64_t)waves[c].size() * (int64_t)sizeof(int32_t);
  }
  h.sink_off = sink_off;
  // problems swept alone, for the single-sweep launch to walk
  h.list1_off = off;
  // The four 16-lane groups of a wave take consecutive list entries and run
  // their sweeps in lockstep until the longest one ends; the list is sorted
  // by kind, width and sweep length (longest first) so that they match.
  int32_t *list1 = (int32_t *)(base + off);
  for (int32_t i = 0; i < n; i++)
    if (d[i].layout == 0 && d[i].kind != GDP_KIND_STD) list1[h.n1++] = i;
  for (int32_t i = 0; i < n; i++)
    if (d[i].kind == GDP_KIND_STD) list1[h.n1 + h.n1std++] = i;
  {
    std::vector<int32_t> steps(n, 0);
    for (int32_t i = 0; i < n; i++)
      if (d[i].layout == 0) steps[i] = gdp_make_geom(d[i]).nsteps;
    auto order = [&](int32_t a, int32_t b) {
      if (d[a].kind != d[b].kind) return d[a].kind < d[b].kind;
      if (d[a].width != d[b].width) return d[a].width < d[b].width;
      if (steps[a] != steps[b]) return steps[a] > steps[b];
      return a < b;
    };
    std::sort(list1, list1 + h.n1, order);
    std::sort(list1 + h.n1, list1 + h.n1 + h.n1std, order);
    off += (int64_t)(h.n1 + h.n1std) * (int64_t)sizeof(int32_t);
    // narrow sweeps: a wave takes 64 consecutive entries, one per lane, and
    // runs until its longest one ends; sorted by row stride, then rows
    h.list2_off = off;
    int32_t *list2 = (int32_t *)(base + off);
    auto rows = [&](int32_t a) { return d[a].kind == GDP_KIND_LOWER ? d[a].gl : d[a].rl; };
    // a counting sort: strides 32, 16, 8, then rows from GDP_MAX_LEN down,
    // indices ascending within a bucket
    {
      const int nb = 3 * (GDP_MAX_LEN + 1);
      auto bucket = [&](int32_t a) {
        const int sc = d[a].P == 32 ? 0 : (d[a].P == 16 ? 1 : 2);
        return sc * (GDP_MAX_LEN + 1) + (GDP_MAX_LEN - gdp_imin(rows(a), GDP_MAX_LEN));
      };
      std::vector<int32_t> start(nb + 1, 0);
      for (int32_t i = 0; i < n; i++)
        if (d[i].layout == 2) start[bucket(i) + 1]++;
      for (int b = 0; b < nb; b++) start[b + 1] += start[b];
      h.n2 = start[nb];
      for (int32_t i = 0; i < n; i++)
        if (d[i].layout == 2) list2[start[bucket(i)]++] = i;
    }
    // Each 64 consecutive entries are one wave's lanes.  Their direction
    // words are interleaved so that the wave's stores are coalesced: word kk
    // of row r of lane l at base + (r * nw + kk) * 64 + l, with nw the
    // wave's largest row stride in words (P becomes nw).
    for (int32_t w0 = 0; w0 < h.n2; w0 += 64) {
      const int32_t w1 = std::min(h.n2, w0 + 64);
      int nw = 1, rmax = 0;
      for (int32_t j = w0; j < w1; j++) {
        nw = std::max(nw, d[list2[j]].P / 8);
        rmax = std::max(rmax, rows(list2[j]));
      }
      for (int32_t j = w0; j < w1; j++) {
        d[list2[j]].P = nw;
        d[list2[j]].dir_off = dw + (j - w0);
      }
      dw += (int64_t)(rmax + 1) * nw * 64;
    }
  }
  memcpy(base + hoff, &h, sizeof h);
  dw += 4;  // the traceback reads direction words in aligned 16-byte windows
  if (dir_words) *dir_words = dw;
  if (score_slots) *score_slots = ss;
  return GDP_OK;
}

int64_t gdp_plan_used_bytes(const void *descs, int32_t n) {
  const gdp_pair_hdr *h = (const gdp_pair_hdr *)((const uint8_t *)descs + (int64_t)n * sizeof(gdp_dev_problem));
  return h->list2_off + (int64_t)h->n2 * (int64_t)sizeof(int32_t);
}

extern "C" int gdp_plan_ex(const gdp_problem *probs, int32_t n, int32_t flags, void *descs_out, int64_t *dir_words,
                           int64_t *score_slots) {
  return gdp_plan_tuned(probs, n, flags, nullptr, descs_out, dir_words, score_slots);
}

extern "C" int gdp_plan(const gdp_problem *probs, int32_t n, void *descs_out, int64_t *dir_words,
                        int64_t *score_slots) {
  return gdp_plan_tuned(probs, n, 0, nullptr, descs_out, dir_words, score_slots);
}

extern "C" int64_t gdp_cells(const gdp_problem *p) {
  int64_t n = 0;
  if (p->kind == GDP_KIND_FULL || p->kind == GDP_KIND_STD) {
    for (int c = 1; c <= p->gl; c++) {
      int lo = gdp_imax(1, c - p->uband), hi = gdp_imin(p->rl, c + p->lband);
      if (hi >= lo) n += hi - lo + 1;
    }
  } else if (p->kind == GDP_KIND_UPPER) {
    for (int c = 1; c <= p->gl; c++) {
      int lo = gdp_imax(1, c - p->uband), hi = gdp_imin(c, p->rl);
      if (hi >= lo) n += hi - lo + 1;
    }
  } else {
    for (int r = 1; r <= p->rl; r++) {
      int lo = gdp_imax(1, r - p->lband), hi = gdp_imin(r, p->gl);
      if (hi >= lo) n += hi - lo + 1;
    }
  }
  return n;
}

extern "C" uint32_t gdp_plan_kernels(const void *descs, int32_t n) {
  if (descs == nullptr || n < 0) return GDP_FILL_ALL;
  const gdp_dev_problem *d = (const gdp_dev_problem *)descs;
  gdp_pair_hdr h;
  memcpy(&h, (const uint8_t *)descs + (int64_t)n * sizeof(gdp_dev_problem), sizeof h);
  if (h.magic != (GDP_PAIR_MAGIC ^ (uint64_t)n)) return GDP_FILL_ALL;
  uint32_t mask = 0;
  int pmax = 0, pstd = 0, pwide = 0;
  for (int32_t i = 0; i < n; i++) {
    if (d[i].layout == 2) mask |= GDP_FILL_NARROW;
    if (d[i].layout != 0) continue;
    const int P = gdp_make_geom(d[i]).P;
    if (d[i].kind == GDP_KIND_STD) {
      mask |= GDP_FILL_STD;
      pstd = gdp_imax(pstd, 2 * P);  // ring entries of two words (H and F in 32 bits)
    } else if (P > GDP_RING) {
      mask |= GDP_FILL_WIDE;
      pwide = gdp_imax(pwide, P);
    } else {
      mask |= 1u;
      pmax = gdp_imax(pmax, P);
    }
  }
  // LDS rings of the single sweeps: the next power of two >= the largest P
  // (in words), bits 8..11 for the SIMD-rule sweep and 12..15 for the
  // standard fill
  auto rlog = [](int need) {
    int k = 6;
    while ((1 << k) < need) k++;
    return (uint32_t)k;
  };
  if (mask & 1u) mask |= rlog(pmax) << 8;
  if (mask & GDP_FILL_STD) mask |= rlog(pstd) << 12;
  if (mask & GDP_FILL_WIDE) mask |= rlog(pwide) << 16;
  for (int c = 0; c < GDP_NCLS; c++)
    if (h.nw[c] > 0) mask |= 2u << c;
  return mask;
}

extern "C" int gdp_fill_ex(const void *d_descs, int32_t n, uint32_t kernels, const char *d_qbuf, const char *d_gbuf,
                           uint32_t *d_dirs, int16_t *d_scores, int32_t *d_final, void *stream) {
  if (!g_init) return fail(GDP_ENOINIT, "gdp_fill: gdp_init has not been called");
  if (n < 0 || (n > 0 && (!d_descs || !d_qbuf || !d_gbuf || !d_dirs || !d_final)))
    return fail(GDP_EINVAL, "gdp_fill: null argument");
  if (gdp_launch_fill(d_descs, n, kernels, d_qbuf, d_gbuf, d_dirs, d_scores, d_final, (hipStream_t)stream))
    return hip_fail("gdp_fill launch", hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_fill(const void *d_descs, int32_t n, const char *d_qbuf, const char *d_gbuf, uint32_t *d_dirs,
                        int16_t *d_scores, int32_t *d_final, void *stream) {
  return gdp_fill_ex(d_descs, n, GDP_FILL_ALL, d_qbuf, d_gbuf, d_dirs, d_scores, d_final, stream);
}

// The launch mask of one descriptor filled by itself (a GDP_PLAN_SINGLE
// problem): its own kernel, with the ring its batch period needs.
static uint32_t kernels_of_one(const gdp_dev_problem &d) {
  const int P = gdp_make_geom(d).P;
  const int need = d.kind == GDP_KIND_STD ? 2 * P : P;
  int k = 6;
  while ((1 << k) < need) k++;
  if (d.kind == GDP_KIND_STD) return GDP_FILL_STD | ((uint32_t)k << 12);
  return need > GDP_RING ? (GDP_FILL_WIDE | ((uint32_t)k << 16)) : (1u | ((uint32_t)k << 8));
}

extern "C" int gdp_align_each(const void *descs, const void *d_descs, int32_t n, const char *d_qbuf,
                              const char *d_gbuf, uint32_t *d_dirs, int16_t *d_scores, int32_t *d_final,
                              const gdp_tb_request *d_reqs, const void *d_genome, int32_t genome_format,
                              uint32_t genome_len, gdp_pair *d_pairs, int32_t *d_counts, void *stream) {
  if (!g_init) return fail(GDP_ENOINIT, "gdp_align_each: gdp_init has not been called");
  if (n < 0 || (n > 0 && (!descs || !d_descs || !d_qbuf || !d_gbuf || !d_dirs || !d_final || !d_reqs || !d_pairs ||
                          !d_counts)))
    return fail(GDP_EINVAL, "gdp_align_each: null argument");
  const gdp_dev_problem *h = (const gdp_dev_problem *)descs;
  for (int32_t k = 0; k < n; k++)
    if (h[k].layout != 0)
      return fail(GDP_EINVAL, "gdp_align_each: problem %d was not planned with GDP_PLAN_SINGLE", k);
  hipStream_t st = (hipStream_t)stream;
  for (int32_t k = 0; k < n; k++) {
    const uint8_t *dk = (const uint8_t *)d_descs + (int64_t)k * sizeof(gdp_dev_problem);
    if (gdp_launch_fill(dk, 1, kernels_of_one(h[k]), d_qbuf, d_gbuf, d_dirs, d_scores, d_final + k, st))
      return hip_fail("gdp_align_each fill launch", hipGetLastError());
    // the request names its problem by index in the whole array; the walk
    // reads the array, so it sees problem k's descriptor
    if (gdp_launch_traceback(d_descs, d_reqs + k, 1, d_qbuf, d_gbuf, d_dirs, d_genome, genome_format, genome_len,
                             d_pairs, d_counts + 6 * (int64_t)k, nullptr, st))
      return hip_fail("gdp_align_each traceback launch", hipGetLastError());
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail("gdp_align_each", e);
  }
  return GDP_OK;
}

extern "C" int gdp_best_endpoint_ex(const void *d_descs, const int32_t *d_ul, int32_t m, int32_t queryend_indels,
                                    const int16_t *d_scores, int32_t *d_out, const gdp_tb_request *d_tb_template,
                                    gdp_tb_request *d_tb_out, void *stream) {
  if (!g_init) return fail(GDP_ENOINIT, "gdp_best_endpoint: gdp_init has not been called");
  if (m < 0 || (m > 0 && (!d_descs || !d_ul || !d_scores || !d_out)) || ((d_tb_template == nullptr) != (d_tb_out == nullptr)))
    return fail(GDP_EINVAL, "gdp_best_endpoint: null argument");
  if (gdp_launch_best_endpoint(d_descs, d_ul, m, queryend_indels, d_scores, d_out, d_tb_template, d_tb_out,
                               (hipStream_t)stream))
    return hip_fail("gdp_best_endpoint launch", hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_best_endpoint(const void *d_descs, const int32_t *d_ul, int32_t m, int32_t queryend_indels,
                                 const int16_t *d_scores, int32_t *d_out, void *stream) {
  return gdp_best_endpoint_ex(d_descs, d_ul, m, queryend_indels, d_scores, d_out, nullptr, nullptr, stream);
}

extern "C" int gdp_bridge_cdna(const void *d_descs, const int32_t *d_quads, int32_t m, const int16_t *d_scores,
                               int32_t *d_out, const gdp_tb_request *d_tb_template, gdp_tb_request *d_tb_out,
                               void *stream) {
  if (!g_init) return fail(GDP_ENOINIT, "gdp_bridge_cdna: gdp_init has not been called");
  if (m < 0 || (m > 0 && (!d_descs || !d_quads || !d_scores || !d_out)) ||
      ((d_tb_template == nullptr) != (d_tb_out == nullptr)))
    return fail(GDP_EINVAL, "gdp_bridge_cdna: null argument");
  if (gdp_launch_bridge_cdna(d_descs, d_quads, m, d_scores, d_out, d_tb_template, d_tb_out, (hipStream_t)stream))
    return hip_fail("gdp_bridge_cdna launch", hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_bridge_intron(const void *d_descs, const gdp_intron_gap *d_gaps, int32_t m, const int16_t *d_scores,
                                 const char *d_gbuf, const double *d_probs, int32_t *d_out, double *d_out_probs,
                                 const gdp_tb_request *d_tb_template, gdp_tb_request *d_tb_out, void *stream) {
  if (!g_init) return fail(GDP_ENOINIT, "gdp_bridge_intron: gdp_init has not been called");
  if (m < 0 || (m > 0 && (!d_descs || !d_gaps || !d_scores || !d_gbuf || !d_probs || !d_out || !d_out_probs)) ||
      ((d_tb_template == nullptr) != (d_tb_out == nullptr)))
    return fail(GDP_EINVAL, "gdp_bridge_intron: null argument");
  if (gdp_launch_bridge_intron(d_descs, d_gaps, m, d_scores, d_gbuf, d_probs, d_out, d_out_probs, d_tb_template,
                               d_tb_out, (hipStream_t)stream))
    return hip_fail("gdp_bridge_intron launch", hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_traceback_ends(const void *d_descs, const gdp_tb_request *d_reqs, int32_t nreq,
                                  const char *d_qbuf, const char *d_gbuf, const uint32_t *d_dirs, const void *d_genome,
                                  int32_t genome_format, uint32_t genome_len, gdp_pair *d_pairs, int32_t *d_counts,
                                  int32_t *d_ends, void *stream) {
  if (!g_init) return fail(GDP_ENOINIT, "gdp_traceback: gdp_init has not been called");
  if (nreq < 0 || (nreq > 0 && (!d_descs || !d_reqs || !d_qbuf || !d_gbuf || !d_dirs || !d_pairs || !d_counts)))
    return fail(GDP_EINVAL, "gdp_traceback: null argument");
  if (genome_format != GDP_GENOME_ASCII && genome_format != GDP_GENOME_COMP)
    return fail(GDP_EINVAL, "gdp_traceback: unknown genome format %d", genome_format);
  if (gdp_launch_traceback(d_descs, d_reqs, nreq, d_qbuf, d_gbuf, d_dirs, d_genome, genome_format, genome_len, d_pairs,
                           d_counts, d_ends, (hipStream_t)stream))
    return hip_fail("gdp_traceback launch", hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_traceback_ex(const void *d_descs, const gdp_tb_request *d_reqs, int32_t nreq, const char *d_qbuf,
                                const char *d_gbuf, const uint32_t *d_dirs, const void *d_genome, int32_t genome_format,
                                uint32_t genome_len, gdp_pair *d_pairs, int32_t *d_counts, void *stream) {
  return gdp_traceback_ends(d_descs, d_reqs, nreq, d_qbuf, d_gbuf, d_dirs, d_genome, genome_format, genome_len, d_pairs,
                            d_counts, nullptr, stream);
}

extern "C" int gdp_traceback(const void *d_descs, const gdp_tb_request *d_reqs, int32_t nreq, const char *d_qbuf,
                             const char *d_gbuf, const uint32_t *d_dirs, const char *d_genome, uint32_t genome_len,
                             gdp_pair *d_pairs, int32_t *d_counts, void *stream) {
  return gdp_traceback_ex(d_descs, d_reqs, nreq, d_qbuf, d_gbuf, d_dirs, d_genome, GDP_GENOME_ASCII, genome_len,
                          d_pairs, d_counts, stream);
}

extern "C" int gdp_launch_download_results(const gdp_tb_request *, int, const int32_t *, const gdp_pair *,
                                           int32_t *, gdp_pair *, hipStream_t);

extern "C" int gdp_download_results(const gdp_tb_request *d_reqs, int32_t nreq, const int32_t *d_counts,
                                    const gdp_pair *d_pairs, int32_t *h_counts, gdp_pair *h_pairs, void *stream) {
  if (nreq < 0 || (nreq > 0 && (!d_reqs || !d_counts || !d_pairs || !h_counts || !h_pairs)))
    return fail(GDP_EINVAL, "gdp_download_results: null or negative argument");
  if (nreq == 0) return GDP_OK;
  void *dc = nullptr, *dp = nullptr;
  if (hipHostGetDevicePointer(&dc, h_counts, 0) != hipSuccess ||
      hipHostGetDevicePointer(&dp, h_pairs, 0) != hipSuccess) {
    // The refused pointer leaves its error as the thread's last HIP error,
    // which every later launch check (hipGetLastError) would report as its
    // own failure: consume it here, where it is answered.
    (void)hipGetLastError();
    return fail(GDP_EINVAL, "gdp_download_results: the host buffers are not pinned (hipHostMalloc)");
  }
  if (gdp_launch_download_results(d_reqs, nreq, d_counts, d_pairs, (int32_t *)dc, (gdp_pair *)dp,
                                  (hipStream_t)stream) != 0)
    return hip_fail("gdp_download_results", hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_genome_segments(const void *d_genome, int32_t format, uint32_t genome_len, const gdp_segment *d_segs,
                                   int32_t n, char *d_out, void *stream) {
  if (n < 0 || (n > 0 && (!d_segs || !d_out))) return fail(GDP_EINVAL, "gdp_genome_segments: bad argument");
  if (format != GDP_GENOME_ASCII && format != GDP_GENOME_COMP)
    return fail(GDP_EINVAL, "gdp_genome_segments: unknown genome format %d", format);
  if (gdp_launch_segments(d_genome, format, genome_len, d_segs, n, d_out, (hipStream_t)stream))
    return hip_fail("gdp_genome_segments launch", hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_export_dense(const void *d_descs, int32_t idx, const uint32_t *d_dirs, const int16_t *d_scores,
                                uint8_t *d_dense_dirs, int16_t *d_dense_scores, void *stream) {
  if (!g_init) return fail(GDP_ENOINIT, "gdp_export_dense: gdp_init has not been called");
  if (!d_descs || !d_dirs || !d_dense_dirs) return fail(GDP_EINVAL, "gdp_export_dense: null argument");
  if (gdp_launch_export(d_descs, idx, d_dirs, d_scores, d_dense_dirs, d_dense_scores, (hipStream_t)stream))
    return hip_fail("gdp_export_dense launch", hipGetLastError());
  return GDP_OK;
}

extern "C" int gdp_export_reference(const void *d_descs, int32_t idx, const uint32_t *d_dirs, const int16_t *d_scores,
                                    int32_t stride, void *d_score, void *d_nogap, void *d_egap, void *d_fgap,
                                    void *stream) {
  if (!g_init) return fail(GDP_ENOINIT, "gdp_export_reference: gdp_init has not been called");
  if (!d_descs || !d_dirs || !d_score || !d_nogap || !d_egap || stride <= 0)
    return fail(GDP_EINVAL, "gdp_export_reference: bad argument");
  gdp_dev_problem p;
  if (hipMemcpy(&p, (const gdp_dev_problem *)d_descs + idx, sizeof p, hipMemcpyDeviceToHost) != hipSuccess)
    return hip_fail("gdp_export_reference descriptor read", hipGetLastError());
  if (p.kind == GDP_KIND_STD)
    return fail(GDP_EINVAL, "gdp_export_reference: the standard fill has no SIMD workspace layout");
  gdp_geom G = gdp_make_geom(p);
  if (stride < G.R + 1) return fail(GDP_EINVAL, "gdp_export_reference: stride %d < %d rows", stride, G.R + 1);
  if (gdp_launch_export_ref(d_descs, idx, p.width, d_dirs, d_scores, stride, d_score, d_nogap, d_egap, d_fgap,
                            (hipStream_t)stream))
    return hip_fail("gdp_export_reference launch", hipGetLastError());
  return GDP_OK;
}

struct gdp_graph {
  hipStream_t stream = nullptr;  // the capture stream
  bool capturing = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int gdp_graph_begin(gdp_graph **out, void **stream) {
  if (!out || !stream) return fail(GDP_EINVAL, "gdp_graph_begin: null argument");
  *out = nullptr;
  *stream = nullptr;
  gdp_graph *g = new gdp_graph;
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    gdp_graph_destroy(g);
    return hip_fail("gdp_graph_begin", e);
  }
  g->capturing = true;
  *out = g;
  *stream = g->stream;
  return GDP_OK;
}

extern "C" int gdp_graph_end(gdp_graph *g) {
  if (!g || !g->capturing) return fail(GDP_EINVAL, "gdp_graph_end: no capture in progress");
  g->capturing = false;
  hipError_t e = hipStreamEndCapture(g->stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  return e == hipSuccess ? GDP_OK : hip_fail("gdp_graph_end", e);
}

extern "C" gdp_graph *gdp_graph_create(const void *d_descs, int32_t n, uint32_t kernels, const char *d_qbuf,
                                       const char *d_gbuf, uint32_t *d_dirs, int16_t *d_scores, int32_t *d_final,
                                       const gdp_tb_request *d_reqs, int32_t nreq, const void *d_genome,
                                       int32_t genome_format, uint32_t genome_len, gdp_pair *d_pairs,
                                       int32_t *d_counts) {
  if (!g_init) {
    fail(GDP_ENOINIT, "gdp_graph_create: gdp_init has not been called");
    return nullptr;
  }
  gdp_graph *g = nullptr;
  void *st = nullptr;
  if (gdp_graph_begin(&g, &st) != GDP_OK) return nullptr;
  int rc = gdp_fill_ex(d_descs, n, kernels, d_qbuf, d_gbuf, d_dirs, d_scores, d_final, st);
  if (rc == GDP_OK && nreq > 0)
    rc = gdp_traceback_ex(d_descs, d_reqs, nreq, d_qbuf, d_gbuf, d_dirs, d_genome, genome_format, genome_len,
                          d_pairs, d_counts, st);
  std::string err = rc == GDP_OK ? "" : g_err;
  const int rc2 = gdp_graph_end(g);
  if (rc != GDP_OK || rc2 != GDP_OK) {
    if (rc != GDP_OK) g_err = err;
    gdp_graph_destroy(g);
    return nullptr;
  }
  return g;
}

extern "C" int gdp_graph_launch(gdp_graph *g, void *stream) {
  if (!g || !g->exec) return fail(GDP_EINVAL, "gdp_graph_launch: no graph");
  const hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
  return e == hipSuccess ? GDP_OK : hip_fail("gdp_graph_launch", e);
}

extern "C" void gdp_graph_destroy(gdp_graph *g) {
  if (!g) return;
  if (g->capturing) {
    hipGraph_t tmp = nullptr;
    if (hipStreamEndCapture(g->stream, &tmp) == hipSuccess && tmp) (void)hipGraphDestroy(tmp);
  }
  if (g->stream) (void)hipStreamDestroy(g->stream);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}
