// gdp_dynprog.hip -- the reference's dynprog_simd.h entry points over the GPU
// core (include/gmapdp_dynprog.h).  One call = one problem: the sequences go
// to the device, one fill launch runs, and the result comes back into the
// caller's Dynprog_T workspace in the reference's layout.  Tracebacks walk
// the device copy of the last fill of the same workspace and replay the
// walk's pair-pool calls on the caller's own Pairpool.
//
// Device buffers: each workspace's fills get buffers of their own, allocated
// (and grown) on first use.  Dynprog_simd_16_noalloc instead carves them out
// of `scratch`, a fixed array in this library's device image, one per GPU.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/gmapdp.h"
#include "../../include/gmapdp_dynprog.h"

// ---------------------------------------------------------------------------
// Host-program symbols (GMAP's dynprog.o and pairpool.o).
extern "C" {
extern short **pairdistance_array[4];  // dynprog.h:166
extern bool **consistent_array[3];      // dynprog.h:170
extern int *nt_to_int_array;            // dynprog.h:171
List_T Pairpool_push(List_T list, Pairpool_T pool, int querypos, int genomepos, char cdna, char comp, char genome,
                     char genomealt, int dynprogindex);
List_T Pairpool_add_queryskip(List_T pairs, int r, int c, int dist, char *querysequence, int queryoffset,
                              int genomeoffset, Pairpool_T pairpool, bool revp, int dynprogindex);
List_T Pairpool_add_genomeskip(bool *add_dashes_p, List_T pairs, int r, int c, int dist, char *genomesequence,
                               int queryoffset, int genomeoffset, Genome_T genome, Genome_T genomealt,
                               Pairpool_T pairpool, bool revp, Univcoord_T chroffset, Univcoord_T chrhigh,
                               bool watsonp, int dynprogindex);
int gdp_init_tables(const int16_t *pd, const int32_t *nt, const uint8_t *cons);
}
void gdp_set_error(const char *msg);  // libgmapdp: the message gdp_last_error() returns

// Device buffers of Dynprog_simd_16_noalloc: one array per GPU, kept until
// the process ends (GDP_NOALLOC_SCRATCH_BYTES, include/gmapdp_dynprog.h).
// Host code uses the address hipGetSymbolAddress gives on each device, never
// the array's name.
static __device__ __align__(256) uint8_t scratch[GDP_NOALLOC_SCRATCH_BYTES];

namespace {

// struct Dynprog_T of the reference's SSE2/AVX2 build (dynprog.h:175-220).
struct RefSingle {
  void **matrix_ptrs, *matrix_space;
  void **directions_ptrs_0, *directions_space_0;
  void **directions_ptrs_1, *directions_space_1;
  void **directions_ptrs_2, *directions_space_2;
};
struct RefDouble {
  void **upper_matrix_ptrs, *upper_matrix_space;
  void **upper_directions_ptrs_0, *upper_directions_space_0;
  void **upper_directions_ptrs_1, *upper_directions_space_1;
  void **lower_matrix_ptrs, *lower_matrix_space;
  void **lower_directions_ptrs_0, *lower_directions_space_0;
  void **lower_directions_ptrs_1, *lower_directions_space_1;
};
struct RefDynprog {
  int max_rlength;
  int max_glength;
  union {
    RefSingle one;
    RefDouble two;
  } aligned;
  int nspaces;
};

[[noreturn]] void die(const char *what) {
  fprintf(stderr, "gmapdp drop-in: %s: %s\n", what, gdp_last_error());
  abort();
}

void hipcheck(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "gmapdp drop-in: %s: %s\n", what, hipGetErrorString(e));
    abort();
  }
}

// A device buffer: allocated and grown on demand, or (fixed) a part of the
// scratch array that the allocation-free fill sized before use.
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  bool fixed = false;
  T *need(size_t want) {
    if (want > n) {
      if (fixed) {  // sized by scratch_layout for this fill; never reached
        fprintf(stderr, "gmapdp drop-in: scratch part of %zu elements asked for %zu\n", n, want);
        abort();
      }
      if (p) hipcheck(hipFree(p), "hipFree");
      hipcheck(hipMalloc((void **)&p, want * sizeof(T)), "hipMalloc");
      n = want;
    }
    return p;
  }
  void place(uint8_t *base, size_t &off, size_t count) {  // the next part of the scratch array
    p = (T *)(base + off);
    n = count;
    fixed = true;
    off += (count * sizeof(T) + 255) & ~(size_t)255;
  }
};

// Device state of one fill slot (full, upper or lower) of one workspace.
struct Slot {
  DevBuf<uint8_t> desc, q, g, exp[4];
  DevBuf<uint32_t> dirs;
  DevBuf<int16_t> scores;
  DevBuf<int32_t> fin, counts;
  DevBuf<gdp_pair> ops;
  DevBuf<gdp_tb_request> req;
  void **nogap_table = nullptr;  // host table handed out by the last fill
  int rl = 0, gl = 0, width = 0;
  bool revp = false;
};

struct Workspace {
  Slot slot[3];
};

std::mutex g_mu;
std::unordered_map<const void *, Workspace *> g_ws;
bool g_tables = false;

// The allocation-free fill's state per device: the scratch address on that
// device and the slot of its last fill (its buffers point into scratch).
constexpr int kMaxDevices = 64;
uint8_t *g_scratch_ptr[kMaxDevices] = {nullptr};
Slot g_static_slot[kMaxDevices];

// The kernels use the host program's own tables (built by its Dynprog_init).
void ensure_tables() {
  if (g_tables) return;
  std::vector<int16_t> pd(4 * 128 * 128);
  std::vector<int32_t> nt(128);
  std::vector<uint8_t> cons(3 * 128 * 128, 0);
  for (int t = 0; t < 4; t++)
    for (int i = 0; i < 128; i++)
      for (int j = 0; j < 128; j++) pd[(t * 128 + i) * 128 + j] = pairdistance_array[t][i][j];
  for (int i = 0; i < 128; i++) nt[i] = nt_to_int_array[i];
  for (int gs = 0; gs < 3; gs++)
    if (consistent_array[gs] != nullptr)
      for (int i = 0; i < 128; i++)
        for (int j = 0; j < 128; j++) cons[(gs * 128 + i) * 128 + j] = consistent_array[gs][i][j] ? 1 : 0;
  if (gdp_init_tables(pd.data(), nt.data(), cons.data()) != 0) die("table upload");
  g_tables = true;
}

Slot &slot_of(Dynprog_T ws, int kind) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_tables();
  Workspace *&w = g_ws[(const void *)ws];
  if (w == nullptr) w = new Workspace();
  return w->slot[kind];
}

Slot &slot_for_table(void **nogap, int kind) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_ws)
    if (kv.second->slot[kind].nogap_table == nogap) return kv.second->slot[kind];
  int dev = 0;
  if (kind == GDP_KIND_FULL && hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kMaxDevices &&
      g_static_slot[dev].nogap_table == nogap)
    return g_static_slot[dev];
  fprintf(stderr, "gmapdp drop-in: traceback on direction tables this library did not produce\n");
  abort();
}

// A workspace's tables name the last fill written into them: forget any
// other slot that handed out the same tables.
void forget_table(void **nogap, const Slot *keep) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_ws)
    for (Slot &s : kv.second->slot)
      if (&s != keep && s.nogap_table == nogap) s.nogap_table = nullptr;
  for (Slot &s : g_static_slot)
    if (&s != keep && s.nogap_table == nogap) s.nogap_table = nullptr;
}

// Sequence in sweep order: reversed reads use s[1 - pos] (dynprog_end.c:1415).
void sweep_copy(std::vector<char> &out, const char *s, int n, bool revp) {
  out.resize(n);
  for (int i = 1; i <= n; i++) out[i - 1] = revp ? s[1 - i] : s[i - 1];
}

// Runs one fill and copies it into the workspace spaces.  Returns the score
// table; fills the direction tables.
// The reference layout of a fill's tables: the inner stride rounded up to
// the SIMD width, as dynprog_simd.c:3044, 6618, 8626 round it.
size_t ref_table_bytes(int kind, int width, int rl, int gl, int *stride_out) {
  const int S = width == 8 ? 32 : 16;
  const int inner = kind == GDP_KIND_LOWER ? gl : rl, outer = kind == GDP_KIND_LOWER ? rl : gl;
  const int stride = ((inner + S) / S) * S;
  if (stride_out) *stride_out = stride;
  return (size_t)(outer + 1) * stride * (width / 8);
}

gdp_problem make_problem(int kind, int width, int rl, int gl, int mt, int open, int extend, int lband, int uband,
                         bool late) {
  gdp_problem p;
  memset(&p, 0, sizeof p);
  p.qoff = 0;
  p.goff = 0;
  p.galtoff = gl;
  p.rl = rl;
  p.gl = gl;
  p.lband = (int16_t)lband;
  p.uband = (int16_t)uband;
  p.kind = (int8_t)kind;
  p.width = (int8_t)width;
  p.mismatchtype = (int8_t)mt;
  p.jump_late = late ? 1 : 0;
  p.open = (int8_t)open;
  p.extend = (int8_t)extend;
  p.keep_scores = 1;
  return p;
}

void **run_fill(Slot &sl, Dynprog_T dynprog, int kind, int width, char *rseq, char *gseq, char *galt, int rl,
                int gl, int mt, int open, int extend, int lband, int uband, bool late, bool revp, void ***nogap,
                void ***egap, void ***fgap) {
  RefDynprog *ws = (RefDynprog *)dynprog;
  std::vector<char> q, g, ga;
  sweep_copy(q, rseq, rl, revp);
  sweep_copy(g, gseq, gl, revp);
  sweep_copy(ga, galt, gl, revp);
  const gdp_problem p = make_problem(kind, width, rl, gl, mt, open, extend, lband, uband, late);
  std::vector<uint8_t> desc((size_t)gdp_plan_bytes(1));
  int64_t dw = 0, ss = 0;
  if (gdp_plan(&p, 1, desc.data(), &dw, &ss) != 0) die("gdp_plan");
  hipcheck(hipMemcpy(sl.desc.need(desc.size()), desc.data(), desc.size(), hipMemcpyHostToDevice), "upload descriptor");
  hipcheck(hipMemcpy(sl.q.need(rl + 16), q.data(), rl, hipMemcpyHostToDevice), "upload query");
  uint8_t *dg = sl.g.need(2 * (size_t)gl + 16);
  hipcheck(hipMemcpy(dg, g.data(), gl, hipMemcpyHostToDevice), "upload genome");
  hipcheck(hipMemcpy(dg + gl, ga.data(), gl, hipMemcpyHostToDevice), "upload genome alt");
  // only the launch this problem needs (the plan's kernel mask)
  if (gdp_fill_ex(sl.desc.p, 1, gdp_plan_kernels(desc.data(), 1), (const char *)sl.q.p, (const char *)sl.g.p,
                  sl.dirs.need(dw), sl.scores.need(ss), sl.fin.need(1), nullptr) != 0)
    die("gdp_fill");

  // Reference layout: outer index (columns of the reference's tables) and the
  // SIMD-rounded inner stride (dynprog_simd.c:3044, 6618, 8626).
  const int outer = kind == GDP_KIND_LOWER ? rl : gl;
  int stride = 0;
  const size_t bytes = ref_table_bytes(kind, width, rl, gl, &stride);
  void *e[4];
  for (int k = 0; k < 4; k++) e[k] = sl.exp[k].need(bytes);
  if (gdp_export_reference(sl.desc.p, 0, sl.dirs.p, sl.scores.p, stride, e[0], e[1], e[2],
                           kind == GDP_KIND_FULL ? e[3] : nullptr, nullptr) != 0)
    die("gdp_export_reference");

  void **ptrs[4], *space[4];
  if (kind == GDP_KIND_FULL) {
    ptrs[0] = ws->aligned.one.matrix_ptrs;
    space[0] = ws->aligned.one.matrix_space;
    ptrs[1] = ws->aligned.one.directions_ptrs_0;
    space[1] = ws->aligned.one.directions_space_0;
    ptrs[2] = ws->aligned.one.directions_ptrs_1;
    space[2] = ws->aligned.one.directions_space_1;
    ptrs[3] = ws->aligned.one.directions_ptrs_2;
    space[3] = ws->aligned.one.directions_space_2;
  } else if (kind == GDP_KIND_UPPER) {
    ptrs[0] = ws->aligned.two.upper_matrix_ptrs;
    space[0] = ws->aligned.two.upper_matrix_space;
    ptrs[1] = ws->aligned.two.upper_directions_ptrs_0;
    space[1] = ws->aligned.two.upper_directions_space_0;
    ptrs[2] = ws->aligned.two.upper_directions_ptrs_1;
    space[2] = ws->aligned.two.upper_directions_space_1;
    ptrs[3] = nullptr;
    space[3] = nullptr;
  } else {
    ptrs[0] = ws->aligned.two.lower_matrix_ptrs;
    space[0] = ws->aligned.two.lower_matrix_space;
    ptrs[1] = ws->aligned.two.lower_directions_ptrs_0;
    space[1] = ws->aligned.two.lower_directions_space_0;
    ptrs[2] = ws->aligned.two.lower_directions_ptrs_1;
    space[2] = ws->aligned.two.lower_directions_space_1;
    ptrs[3] = nullptr;
    space[3] = nullptr;
  }
  for (int k = 0; k < 4; k++) {
    if (space[k] == nullptr) continue;
    hipcheck(hipMemcpy(space[k], e[k], bytes, hipMemcpyDeviceToHost), "download result");
    char *base = (char *)space[k];
    for (int o = 0; o <= outer; o++) ptrs[k][o] = base + (size_t)o * stride * (width / 8);
  }
  forget_table(ptrs[1], &sl);
  sl.nogap_table = ptrs[1];
  sl.rl = rl;
  sl.gl = gl;
  sl.width = width;
  sl.revp = revp;
  *nogap = ptrs[1];
  *egap = ptrs[2];
  if (fgap) *fgap = ptrs[3];
  return ptrs[0];
}

List_T run_traceback(int kind, List_T pairs, int *traceback_score, int *nmatches, int *nmismatches, int *nopens,
                     int *nindels, void **nogap, int r, int c, char *rsequence, char *rsequenceuc, char *gsequence,
                     char *gsequence_alt, int queryoffset, int genomeoffset, Genome_T genome, Genome_T genomealt,
                     Pairpool_T pairpool, bool revp, Univcoord_T chroffset, Univcoord_T chrhigh, bool watsonp,
                     int genestrand, int dynprogindex) {
  Slot &sl = slot_for_table(nogap, kind);
  // The walk compares rsequenceuc and gsequence as passed to this call.
  std::vector<char> q, g, ga;
  sweep_copy(q, rsequenceuc, sl.rl, revp);
  sweep_copy(g, gsequence, sl.gl, revp);
  sweep_copy(ga, gsequence_alt, sl.gl, revp);
  hipcheck(hipMemcpy(sl.q.p, q.data(), sl.rl, hipMemcpyHostToDevice), "upload query");
  hipcheck(hipMemcpy(sl.g.p, g.data(), sl.gl, hipMemcpyHostToDevice), "upload genome");
  hipcheck(hipMemcpy(sl.g.p + sl.gl, ga.data(), sl.gl, hipMemcpyHostToDevice), "upload genome alt");
  gdp_tb_request rq;
  memset(&rq, 0, sizeof rq);
  rq.fill = 0;
  rq.r = r;
  rq.c = c;
  rq.queryoffset = queryoffset;
  rq.genomeoffset = genomeoffset;
  // The walk records its pair-pool calls (GDP_TB_OPS): genome skips come
  // back as calls that the host's Pairpool_add_genomeskip replays with its
  // own genome and bounds, so the device never reads these (nor truncates a
  // 64-bit Univcoord_T of a large-genome build).
  rq.chroffset = 0;
  rq.chrhigh = 0;
  rq.watsonp = watsonp ? 1 : 0;
  rq.revp = revp ? 1 : 0;
  rq.genestrand = (int8_t)genestrand;
  rq.flags = GDP_TB_OPS;
  rq.pair_off = 0;
  rq.max_pairs = sl.rl + sl.gl + 4;  // the scratch layout reserves as many
  hipcheck(hipMemcpy(sl.req.need(1), &rq, sizeof rq, hipMemcpyHostToDevice), "upload request");
  if (gdp_traceback(sl.desc.p, sl.req.p, 1, (const char *)sl.q.p, (const char *)sl.g.p, sl.dirs.p, nullptr, 0,
                    sl.ops.need(rq.max_pairs), sl.counts.need(6), nullptr) != 0)
    die("gdp_traceback");
  int32_t cnt[6];
  hipcheck(hipMemcpy(cnt, sl.counts.p, sizeof cnt, hipMemcpyDeviceToHost), "download counts");
  if (cnt[0] < 0 || cnt[0] > rq.max_pairs) die("traceback start outside the matrix");
  std::vector<gdp_pair> ops(cnt[0]);
  hipcheck(hipMemcpy(ops.data(), sl.ops.p, sizeof(gdp_pair) * ops.size(), hipMemcpyDeviceToHost), "download ops");
  // Replay the walk's pair-pool calls (dynprog_simd.c:9552-9711).
  for (const gdp_pair &op : ops) {
    if (op.cdna == 'D') {
      int qc = op.querypos - 1, gc = op.genomepos - 1;
      if (revp) {
        qc = -qc;
        gc = -gc;
      }
      pairs = Pairpool_push(pairs, pairpool, queryoffset + qc, genomeoffset + gc, rsequence[qc], op.comp,
                            gsequence[gc], gsequence_alt[gc], dynprogindex);
    } else if (op.cdna == 'H') {
      bool add_dashes_p;
      pairs = Pairpool_add_genomeskip(&add_dashes_p, pairs, op.querypos, op.genomepos, op.genomejump, nullptr,
                                      queryoffset, genomeoffset, genome, genomealt, pairpool, revp, chroffset,
                                      chrhigh, watsonp, dynprogindex);
    } else {
      pairs = Pairpool_add_queryskip(pairs, op.querypos, op.genomepos, op.genomejump, rsequence, queryoffset,
                                     genomeoffset, pairpool, revp, dynprogindex);
    }
  }
  *traceback_score += cnt[1];
  *nmatches += cnt[2];
  *nmismatches += cnt[3];
  *nopens += cnt[4];
  *nindels += cnt[5];
  return pairs;
}

}  // namespace

extern "C" {

Score8_T **Dynprog_simd_8(Direction8_T ***directions_nogap, Direction8_T ***directions_Egap,
                          Direction8_T ***directions_Fgap, Dynprog_T dynprog, char *rsequence, char *gsequence,
                          char *gsequence_alt, int rlength, int glength, Mismatchtype_T mismatchtype, int open,
                          int extend, int lband, int uband, bool jump_late_p, bool revp) {
  return (Score8_T **)run_fill(slot_of(dynprog, GDP_KIND_FULL), dynprog, GDP_KIND_FULL, 8, rsequence, gsequence, gsequence_alt, rlength, glength,
                               mismatchtype, open, extend, lband, uband, jump_late_p, revp,
                               (void ***)directions_nogap, (void ***)directions_Egap, (void ***)directions_Fgap);
}

Score16_T **Dynprog_simd_16(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                            Direction16_T ***directions_Fgap, Dynprog_T dynprog, char *rsequence, char *gsequence,
                            char *gsequence_alt, int rlength, int glength, Mismatchtype_T mismatchtype, int open,
                            int extend, int lband, int uband, bool jump_late_p, bool revp) {
  return (Score16_T **)run_fill(slot_of(dynprog, GDP_KIND_FULL), dynprog, GDP_KIND_FULL, 16, rsequence, gsequence, gsequence_alt, rlength, glength,
                                mismatchtype, open, extend, lband, uband, jump_late_p, revp,
                                (void ***)directions_nogap, (void ***)directions_Egap, (void ***)directions_Fgap);
}

Score16_T **Dynprog_simd_16_noalloc(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                                    Direction16_T ***directions_Fgap, Dynprog_T dynprog, char *rsequence,
                                    char *gsequence, char *gsequence_alt, int rlength, int glength,
                                    Mismatchtype_T mismatchtype, int open, int extend, int lband, int uband,
                                    bool jump_late_p, bool revp) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    gdp_set_error("Dynprog_simd_16_noalloc: no current HIP device");
    return nullptr;
  }
  uint8_t *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ensure_tables();
    if (g_scratch_ptr[dev] == nullptr) {
      // once per device, with that device current: the array's address on it
      void *ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || ptr == nullptr) {
        char msg[160];
        snprintf(msg, sizeof msg, "Dynprog_simd_16_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                 hipGetErrorString(e));
        gdp_set_error(msg);
        return nullptr;
      }
      g_scratch_ptr[dev] = (uint8_t *)ptr;
    }
    base = g_scratch_ptr[dev];
  }
  // The fill's buffer sizes, from the same plan run_fill makes.
  const gdp_problem p = make_problem(GDP_KIND_FULL, 16, rlength, glength, mismatchtype, open, extend, lband, uband,
                                     jump_late_p);
  if (rlength < 1 || glength < 1 || rlength > GDP_MAX_LEN || glength > GDP_MAX_LEN) {
    gdp_set_error("Dynprog_simd_16_noalloc: lengths outside [1, GDP_MAX_LEN]");
    return nullptr;
  }
  std::vector<uint8_t> desc((size_t)gdp_plan_bytes(1));
  int64_t dw = 0, ss = 0;
  if (gdp_plan(&p, 1, desc.data(), &dw, &ss) != 0) return nullptr;  // gdp_plan set the message
  const size_t table = ref_table_bytes(GDP_KIND_FULL, 16, rlength, glength, nullptr);
  const size_t pairs = (size_t)rlength + glength + 4;
  Slot &sl = g_static_slot[dev];
  size_t off = 0;
  sl.desc.place(base, off, desc.size());
  sl.q.place(base, off, (size_t)rlength + 16);
  sl.g.place(base, off, 2 * (size_t)glength + 16);
  sl.dirs.place(base, off, (size_t)dw);
  sl.scores.place(base, off, (size_t)(ss > 0 ? ss : 1));
  sl.fin.place(base, off, 1);
  for (int k = 0; k < 4; k++) sl.exp[k].place(base, off, table);
  sl.counts.place(base, off, 6);
  sl.ops.place(base, off, pairs);
  sl.req.place(base, off, 1);
  if (off > GDP_NOALLOC_SCRATCH_BYTES) {
    char msg[200];
    snprintf(msg, sizeof msg,
             "Dynprog_simd_16_noalloc: a %d x %d fill with band %d+%d needs %zu bytes of device buffers, more than "
             "the %u of the scratch array",
             rlength, glength, lband, uband, off, GDP_NOALLOC_SCRATCH_BYTES);
    gdp_set_error(msg);
    sl.nogap_table = nullptr;
    return nullptr;
  }
  return (Score16_T **)run_fill(sl, dynprog, GDP_KIND_FULL, 16, rsequence, gsequence, gsequence_alt, rlength, glength,
                                mismatchtype, open, extend, lband, uband, jump_late_p, revp,
                                (void ***)directions_nogap, (void ***)directions_Egap, (void ***)directions_Fgap);
}

Score8_T **Dynprog_simd_8_upper(Direction8_T ***directions_nogap, Direction8_T ***directions_Egap, Dynprog_T dynprog,
                                char *rsequence, char *gsequence, char *gsequence_alt, int rlength, int glength,
                                Mismatchtype_T mismatchtype, int open, int extend, int uband, bool jump_late_p,
                                bool revp) {
  return (Score8_T **)run_fill(slot_of(dynprog, GDP_KIND_UPPER), dynprog, GDP_KIND_UPPER, 8, rsequence, gsequence, gsequence_alt, rlength, glength,
                               mismatchtype, open, extend, 0, uband, jump_late_p, revp, (void ***)directions_nogap,
                               (void ***)directions_Egap, nullptr);
}

Score16_T **Dynprog_simd_16_upper(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                                  Dynprog_T dynprog, char *rsequence, char *gsequence, char *gsequence_alt,
                                  int rlength, int glength, Mismatchtype_T mismatchtype, int open, int extend,
                                  int uband, bool jump_late_p, bool revp) {
  return (Score16_T **)run_fill(slot_of(dynprog, GDP_KIND_UPPER), dynprog, GDP_KIND_UPPER, 16, rsequence, gsequence, gsequence_alt, rlength, glength,
                                mismatchtype, open, extend, 0, uband, jump_late_p, revp, (void ***)directions_nogap,
                                (void ***)directions_Egap, nullptr);
}

Score8_T **Dynprog_simd_8_lower(Direction8_T ***directions_nogap, Direction8_T ***directions_Egap, Dynprog_T dynprog,
                                char *rsequence, char *gsequence, char *gsequence_alt, int rlength, int glength,
                                Mismatchtype_T mismatchtype, int open, int extend, int lband, bool jump_late_p,
                                bool revp) {
  return (Score8_T **)run_fill(slot_of(dynprog, GDP_KIND_LOWER), dynprog, GDP_KIND_LOWER, 8, rsequence, gsequence, gsequence_alt, rlength, glength,
                               mismatchtype, open, extend, lband, 0, jump_late_p, revp, (void ***)directions_nogap,
                               (void ***)directions_Egap, nullptr);
}

Score16_T **Dynprog_simd_16_lower(Direction16_T ***directions_nogap, Direction16_T ***directions_Egap,
                                  Dynprog_T dynprog, char *rsequence, char *gsequence, char *gsequence_alt,
                                  int rlength, int glength, Mismatchtype_T mismatchtype, int open, int extend,
                                  int lband, bool jump_late_p, bool revp) {
  return (Score16_T **)run_fill(slot_of(dynprog, GDP_KIND_LOWER), dynprog, GDP_KIND_LOWER, 16, rsequence, gsequence, gsequence_alt, rlength, glength,
                                mismatchtype, open, extend, lband, 0, jump_late_p, revp, (void ***)directions_nogap,
                                (void ***)directions_Egap, nullptr);
}

#define TB_FULL(W)                                                                                                   \
  List_T Dynprog_traceback_##W(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches, int *nopens,     \
                               int *nindels, Direction##W##_T **directions_nogap, Direction##W##_T **directions_Egap, \
                               Direction##W##_T **directions_Fgap, int r, int c, char *rsequence, char *rsequenceuc, \
                               char *gsequence, char *gsequence_alt, int queryoffset, int genomeoffset,              \
                               Genome_T genome, Genome_T genomealt, Pairpool_T pairpool, bool revp,                 \
                               Univcoord_T chroffset, Univcoord_T chrhigh, bool watsonp, int genestrand,            \
                               int dynprogindex) {                                                                   \
    (void)directions_Egap;                                                                                           \
    (void)directions_Fgap;                                                                                           \
    return run_traceback(GDP_KIND_FULL, pairs, traceback_score, nmatches, nmismatches, nopens, nindels,             \
                         (void **)directions_nogap, r, c, rsequence, rsequenceuc, gsequence, gsequence_alt,         \
                         queryoffset, genomeoffset, genome, genomealt, pairpool, revp, chroffset, chrhigh, watsonp, \
                         genestrand, dynprogindex);                                                                  \
  }
TB_FULL(8)
TB_FULL(16)

#define TB_UPPER(W)                                                                                                  \
  List_T Dynprog_traceback_##W##_upper(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches,          \
                                       int *nopens, int *nindels, Direction##W##_T **directions_nogap,               \
                                       Direction##W##_T **directions_Egap, int r, int c, char *rsequence,            \
                                       char *rsequenceuc, char *gsequence, char *gsequence_alt, int queryoffset,     \
                                       int genomeoffset, Genome_T genome, Genome_T genomealt, Pairpool_T pairpool,   \
                                       bool revp, Univcoord_T chroffset, Univcoord_T chrhigh, bool watsonp,          \
                                       int genestrand, int dynprogindex) {                                           \
    (void)directions_Egap;                                                                                           \
    return run_traceback(GDP_KIND_UPPER, pairs, traceback_score, nmatches, nmismatches, nopens, nindels,            \
                         (void **)directions_nogap, r, c, rsequence, rsequenceuc, gsequence, gsequence_alt,         \
                         queryoffset, genomeoffset, genome, genomealt, pairpool, revp, chroffset, chrhigh, watsonp, \
                         genestrand, dynprogindex);                                                                  \
  }
TB_UPPER(8)
TB_UPPER(16)

#define TB_LOWER(W)                                                                                                  \
  List_T Dynprog_traceback_##W##_lower(List_T pairs, int *traceback_score, int *nmatches, int *nmismatches,          \
                                       int *nopens, int *nindels, Direction##W##_T **directions_nogap,               \
                                       Direction##W##_T **directions_Egap, int r, int c, char *rsequence,            \
                                       char *rsequenceuc, char *gsequence, char *gsequence_alt, int queryoffset,     \
                                       int genomeoffset, Pairpool_T pairpool, int genestrand, bool revp,             \
                                       int dynprogindex) {                                                           \
    (void)directions_Egap;                                                                                           \
    return run_traceback(GDP_KIND_LOWER, pairs, traceback_score, nmatches, nmismatches, nopens, nindels,            \
                         (void **)directions_nogap, r, c, rsequence, rsequenceuc, gsequence, gsequence_alt,         \
                         queryoffset, genomeoffset, nullptr, nullptr, pairpool, revp, 0, 0, true, genestrand,       \
                         dynprogindex);                                                                              \
  }
TB_LOWER(8)
TB_LOWER(16)

}  // extern "C"
