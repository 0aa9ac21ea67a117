// gdp_ends.cpp -- GMAP's DP callers for a batch, every step on the device.
//
// Dynprog_end5_splicejunction / Dynprog_end3_splicejunction (reference
// dynprog_end.c:1652-1922, 2248-2745): the two one-sided fills per end
// keeping their best last-row cell, the merge of the two cells, both local
// walks, and the join of the walks around the known-site gap holder.
//
// Dynprog_single_gap (dynprog_single.c:428-675): the genome windows cut from
// a device genome, single_gap_simple for windows as long as their read, the
// fill and walk of the rest, and the choice of the list the reference
// returns.
//
// Dynprog_end5_gap / Dynprog_end3_gap (dynprog_end.c:1293-1650,
// 1923-2243): the windows chopped to max_glength, the two one-sided fills
// per end (one for QUERYEND_NOGAPS), the best end cell of each end mode,
// the walk, the mismatch filter and the leading-indel trim.
//
// Dynprog_cdna_gap (dynprog_cdna.c:786-1180): the two windows cut from a
// device genome, the four one-sided fills keeping every score, the bridge,
// both ends' walks, and the join of the walks around the insertion pairs
// or the gap holder.
//
// The host builds the problems, the plan and the walk templates, and sizes
// the output from the per-call lengths the device reports.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gdp_internal.h"

extern "C" int gdp_launch_sj(int stage, int five, int m, const int32_t *best, const int32_t *stops,
                             const gdp_tb_request *tmpl, gdp_tb_request *req, const gdp_pair *p1, const gdp_pair *p2,
                             const gdp_tb_request *req1, const gdp_tb_request *req2, const int32_t *c1,
                             const int32_t *c2, int32_t *len, int32_t *trim, const int64_t *off, const int32_t *aux,
                             gdp_sj_pair *out, gdp_sj_result *res, hipStream_t stream);

extern "C" int gdp_launch_sg(int stage, int m, const gdp_sg_dev *calls, const char *qbuf, const char *gbuf,
                             gdp_pair *ps, int32_t *cs, const gdp_tb_request *reqs, const gdp_pair *pd,
                             const int32_t *cd, int32_t *len, int32_t *choice, const int64_t *off, gdp_pair *out,
                             gdp_sg_result *res, hipStream_t stream);

static_assert(sizeof(gdp_sj_end) == 56, "gdp_sj_end is 56 bytes");
static_assert(sizeof(gdp_sg_call) == 48, "gdp_sg_call is 48 bytes");
static_assert(sizeof(gdp_sg_result) == 40, "gdp_sg_result is 40 bytes");
static_assert(sizeof(gdp_sj_pair) == 20, "gdp_sj_pair is 20 bytes");
static_assert(sizeof(gdp_sj_result) == 40, "gdp_sj_result is 40 bytes");

namespace {

constexpr int kEndq = GDP_ENDQ;
constexpr int kEndExtend = -2;  // END_EXTEND (dynprog_end.c:88-94)

// END_OPEN by defect rate (dynprog_end.c:88-94, 1688-1700; DEFECT_HIGHQ
// 0.003 and DEFECT_MEDQ 0.014, dynprog.h:58-59).
int end_open(double defect_rate) { return defect_rate < 0.003 ? -10 : (defect_rate < 0.014 ? -8 : -6); }

// A device buffer that grows to the largest size asked of it.
struct Grow {
  void *p = nullptr;
  size_t cap = 0;
  bool ensure(size_t bytes) {
    if (bytes <= cap) return true;
    if (p != nullptr) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return false;
    cap = bytes;
    return true;
  }
  void release() {
    if (p != nullptr) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// A pinned host buffer that grows likewise (the results are copied back
// into it at full PCIe rate).
struct PinnedGrow {
  void *p = nullptr;
  size_t cap = 0;
  bool ensure(size_t bytes) {
    if (bytes <= cap) return true;
    if (p != nullptr) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return false;
    cap = bytes;
    return true;
  }
  void release() {
    if (p != nullptr) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

int err(int code, const std::string &msg) {
  gdp_set_error(msg.c_str());
  return code;
}

}  // namespace

struct gdp_sj_batch {
  Grow descs, q, g, dirs, scores, fin, ul, best, tmpl1, tmpl2, req1, req2, p1, p2, c1, c2, stops, len, trim, off, aux,
      out, res;
  std::vector<gdp_problem> probs;
  std::vector<uint8_t> hdescs;
  std::vector<int32_t> hul, haux, hlen, live;
  std::vector<gdp_tb_request> ht1, ht2;
  std::vector<int64_t> hoff;
  std::vector<gdp_sj_result> results;
  std::vector<gdp_sj_result> lres;
  PinnedGrow pairs;
  int64_t npairs = 0;
  ~gdp_sj_batch() {
    for (Grow *b : {&descs, &q, &g, &dirs, &scores, &fin, &ul, &best, &tmpl1, &tmpl2, &req1, &req2, &p1, &p2, &c1, &c2,
                    &stops, &len, &trim, &off, &aux, &out, &res})
      b->release();
    pairs.release();
  }
};

extern "C" gdp_sj_batch *gdp_sj_create(void) { return new gdp_sj_batch(); }
extern "C" void gdp_sj_destroy(gdp_sj_batch *b) { delete b; }
extern "C" int64_t gdp_sj_npairs(const gdp_sj_batch *b) { return b ? b->npairs : 0; }
extern "C" const gdp_sj_result *gdp_sj_results(const gdp_sj_batch *b) { return b ? b->results.data() : nullptr; }
extern "C" const gdp_sj_pair *gdp_sj_pairs(const gdp_sj_batch *b) {
  return b ? static_cast<const gdp_sj_pair *>(b->pairs.p) : nullptr;
}

extern "C" int gdp_sj_run(gdp_sj_batch *b, int32_t five, const gdp_sj_end *ends, int32_t n, const char *qbuf,
                          int64_t qbytes, const char *gbuf, int64_t gbytes, int32_t max_rlength, int32_t max_glength,
                          void *stream_) {
  if (b == nullptr || n < 0 || (n > 0 && (ends == nullptr || qbuf == nullptr || gbuf == nullptr)) || qbytes < 0 ||
      gbytes < 0)
    return err(GDP_EINVAL, "gdp_sj_run: null or negative argument");
  hipStream_t stream = (hipStream_t)stream_;
  const int use8 = gdp_use8p_size(kEndq);
  b->npairs = 0;
  b->results.assign((size_t)n, gdp_sj_result{0, -1, {0, 0, 0, 0, 0, 0}, 0});
  // ends that are empty or past the workspace give NULL (dynprog_end.c:1702-1717)
  b->live.clear();
  for (int32_t k = 0; k < n; k++) {
    const gdp_sj_end &e = ends[k];
    if (e.rl <= 0 || e.rl > max_rlength || e.gl <= 0 || e.gl > max_glength) continue;
    if (e.qoff < 0 || e.goff < 0 || e.qoff + e.rl > qbytes || e.goff + 2 * (int64_t)e.gl > gbytes)
      return err(GDP_EINVAL, "gdp_sj_run: end " + std::to_string(k) + " reads past its buffers");
    if (e.contlength < 0) return err(GDP_EINVAL, "gdp_sj_run: end " + std::to_string(k) + " has contlength < 0");
    b->live.push_back(k);
  }
  const int m = (int)b->live.size();
  if (m == 0) return GDP_OK;

  // two one-sided fills per end (1719-1779, 2300-2360)
  b->probs.assign(2 * (size_t)m, gdp_problem{});
  b->hul.resize(2 * (size_t)m);
  b->haux.resize(2 * (size_t)m);
  b->ht1.assign((size_t)m, gdp_tb_request{});
  b->ht2.assign((size_t)m, gdp_tb_request{});
  int64_t pair_cap = 0;
  for (int j = 0; j < m; j++) {
    const gdp_sj_end &e = ends[b->live[j]];
    const int lb = e.gl >= e.rl ? e.extraband : e.rl - e.gl + e.extraband;  // Dynprog_compute_bands, widebandp
    const int ub = e.gl >= e.rl ? e.gl - e.rl + e.extraband : e.extraband;
    for (int t = 0; t < 2; t++) {
      gdp_problem &p = b->probs[2 * (size_t)j + t];
      p.qoff = e.qoff;
      p.goff = e.goff;
      p.galtoff = e.goff + e.gl;
      p.rl = e.rl;
      p.gl = e.gl;
      p.lband = (int16_t)lb;
      p.uband = (int16_t)ub;
      p.kind = t == 0 ? GDP_KIND_UPPER : GDP_KIND_LOWER;
      p.width = (e.rl < use8 || e.gl < use8) ? 8 : 16;
      p.mismatchtype = kEndq;
      p.jump_late = five ? !e.jump_late : (e.jump_late != 0);  // the 5' fills take the opposite tie rule
      p.open = (int8_t)end_open(e.defect_rate);
      p.extend = kEndExtend;
      p.keep_scores = GDP_KEEP_BEST_QUERYEND;
    }
    b->hul[2 * (size_t)j] = 2 * j;
    b->hul[2 * (size_t)j + 1] = 2 * j + 1;
    b->haux[2 * (size_t)j] = e.rl;
    b->haux[2 * (size_t)j + 1] = five ? e.anchor - e.far : e.far - e.anchor;
    const int maxp = e.rl + e.gl + 4;
    for (int w = 0; w < 2; w++) {
      gdp_tb_request &q = w == 0 ? b->ht1[j] : b->ht2[j];
      q.queryoffset = e.roffset;
      q.genomeoffset = w == 0 ? e.far : e.anchor;
      q.watsonp = 1;
      q.revp = five ? 1 : 0;
      q.genestrand = e.genestrand;
      q.flags = GDP_TB_LOCAL;
      q.pair_off = pair_cap;
      q.max_pairs = maxp;
      q.endc = w == 0 ? e.contlength : 0;
    }
    pair_cap += maxp;
  }
  const int np = 2 * m;
  b->hdescs.resize((size_t)gdp_plan_bytes(np));
  int64_t dw = 0, ss = 0;
  int rc = gdp_plan_ex(b->probs.data(), np, 0, b->hdescs.data(), &dw, &ss);
  if (rc != GDP_OK) return rc;
  const uint32_t kernels = gdp_plan_kernels(b->hdescs.data(), np);

  const size_t tb = sizeof(gdp_tb_request) * (size_t)m;
  const int64_t used = gdp_plan_used_bytes(b->hdescs.data(), np);
  if (!b->descs.ensure((size_t)used) || !b->q.ensure((size_t)qbytes + 16) || !b->g.ensure((size_t)gbytes + 16) ||
      !b->dirs.ensure(sizeof(uint32_t) * (size_t)(dw > 0 ? dw : 1)) ||
      !b->scores.ensure(sizeof(int16_t) * (size_t)(ss > 0 ? ss : 1)) || !b->fin.ensure(sizeof(int32_t) * np) ||
      !b->ul.ensure(sizeof(int32_t) * np) || !b->best.ensure(sizeof(int32_t) * 3 * (size_t)m) ||
      !b->tmpl1.ensure(tb) || !b->tmpl2.ensure(tb) || !b->req1.ensure(tb) || !b->req2.ensure(tb) ||
      !b->p1.ensure(sizeof(gdp_pair) * (size_t)pair_cap) || !b->p2.ensure(sizeof(gdp_pair) * (size_t)pair_cap) ||
      !b->c1.ensure(sizeof(int32_t) * 6 * (size_t)m) || !b->c2.ensure(sizeof(int32_t) * 6 * (size_t)m) ||
      !b->stops.ensure(sizeof(int32_t) * 2 * (size_t)m) || !b->len.ensure(sizeof(int32_t) * (size_t)m) ||
      !b->trim.ensure(sizeof(int32_t) * (size_t)m) || !b->off.ensure(sizeof(int64_t) * (size_t)m) ||
      !b->aux.ensure(sizeof(int32_t) * np) || !b->res.ensure(sizeof(gdp_sj_result) * (size_t)m))
    return err(GDP_ENOMEM, "gdp_sj_run: device allocation failed");
  auto up = [&](Grow &d, const void *h, size_t bytes) {
    return hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, stream) == hipSuccess;
  };
  if (!up(b->descs, b->hdescs.data(), (size_t)used) || !up(b->q, qbuf, (size_t)qbytes) ||
      !up(b->g, gbuf, (size_t)gbytes) || !up(b->ul, b->hul.data(), sizeof(int32_t) * np) ||
      !up(b->aux, b->haux.data(), sizeof(int32_t) * np) || !up(b->tmpl1, b->ht1.data(), tb) ||
      !up(b->tmpl2, b->ht2.data(), tb))
    return err(GDP_EHIP, "gdp_sj_run: upload failed");
  // the walks' stop cells start at (-1, -1), as an unwalked end leaves them
  if (hipMemsetAsync(b->stops.p, 0xff, sizeof(int32_t) * 2 * (size_t)m, stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_sj_run: clear failed");

  const char *dq = b->q.as<char>(), *dg = b->g.as<char>();
  if ((rc = gdp_fill_ex(b->descs.p, np, kernels, dq, dg, b->dirs.as<uint32_t>(), b->scores.as<int16_t>(),
                        b->fin.as<int32_t>(), stream)) != GDP_OK)
    return rc;
  // the best cell of the last row (find_best_endpoint_to_queryend_indels)
  if ((rc = gdp_best_endpoint(b->descs.p, b->ul.as<int32_t>(), m, 1, b->scores.as<int16_t>(), b->best.as<int32_t>(),
                              stream)) != GDP_OK)
    return rc;
  if (gdp_launch_sj(0, five, m, b->best.as<int32_t>(), nullptr, b->tmpl1.as<gdp_tb_request>(),
                    b->req1.as<gdp_tb_request>(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                    nullptr, nullptr, nullptr, nullptr, nullptr, stream))
    return err(GDP_EHIP, "gdp_sj_run: first requests failed");
  if ((rc = gdp_traceback_ends(b->descs.p, b->req1.as<gdp_tb_request>(), m, dq, dg, b->dirs.as<uint32_t>(), nullptr, 0,
                               0, b->p1.as<gdp_pair>(), b->c1.as<int32_t>(), b->stops.as<int32_t>(), stream)) !=
      GDP_OK)
    return rc;
  if (gdp_launch_sj(1, five, m, b->best.as<int32_t>(), b->stops.as<int32_t>(), b->tmpl2.as<gdp_tb_request>(),
                    b->req2.as<gdp_tb_request>(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                    nullptr, nullptr, nullptr, nullptr, nullptr, stream))
    return err(GDP_EHIP, "gdp_sj_run: second requests failed");
  if ((rc = gdp_traceback_ex(b->descs.p, b->req2.as<gdp_tb_request>(), m, dq, dg, b->dirs.as<uint32_t>(), nullptr, 0, 0,
                             b->p2.as<gdp_pair>(), b->c2.as<int32_t>(), stream)) != GDP_OK)
    return rc;
  if (gdp_launch_sj(2, five, m, b->best.as<int32_t>(), nullptr, nullptr, nullptr, b->p1.as<gdp_pair>(), nullptr,
                    b->req1.as<gdp_tb_request>(), nullptr, b->c1.as<int32_t>(), b->c2.as<int32_t>(),
                    b->len.as<int32_t>(), b->trim.as<int32_t>(), nullptr, nullptr, nullptr, nullptr, stream))
    return err(GDP_EHIP, "gdp_sj_run: lengths failed");
  // output offsets from the lengths
  b->hlen.resize((size_t)m);
  if (hipMemcpyAsync(b->hlen.data(), b->len.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_sj_run: lengths copy failed");
  b->hoff.resize((size_t)m);
  int64_t total = 0;
  for (int j = 0; j < m; j++) {
    b->hoff[j] = total;
    total += b->hlen[j] > 0 ? b->hlen[j] : 0;
  }
  if (!b->out.ensure(sizeof(gdp_sj_pair) * (size_t)(total > 0 ? total : 1)) || !up(b->off, b->hoff.data(), 8 * (size_t)m))
    return err(GDP_ENOMEM, "gdp_sj_run: output allocation failed");
  if (gdp_launch_sj(3, five, m, nullptr, nullptr, nullptr, nullptr, b->p1.as<gdp_pair>(), b->p2.as<gdp_pair>(),
                    b->req1.as<gdp_tb_request>(), b->req2.as<gdp_tb_request>(), b->c1.as<int32_t>(),
                    b->c2.as<int32_t>(), b->len.as<int32_t>(), b->trim.as<int32_t>(), b->off.as<int64_t>(),
                    b->aux.as<int32_t>(), b->out.as<gdp_sj_pair>(), b->res.as<gdp_sj_result>(), stream))
    return err(GDP_EHIP, "gdp_sj_run: join failed");
  b->lres.resize((size_t)m);
  if (!b->pairs.ensure(sizeof(gdp_sj_pair) * (size_t)(total > 0 ? total : 1)))
    return err(GDP_ENOMEM, "gdp_sj_run: pinned host allocation failed");
  if (hipMemcpyAsync(b->lres.data(), b->res.p, sizeof(gdp_sj_result) * m, hipMemcpyDeviceToHost, stream) !=
          hipSuccess ||
      (total > 0 && hipMemcpyAsync(b->pairs.p, b->out.p, sizeof(gdp_sj_pair) * (size_t)total, hipMemcpyDeviceToHost,
                                   stream) != hipSuccess) ||
      hipStreamSynchronize(stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_sj_run: result copy failed");
  for (int j = 0; j < m; j++) b->results[b->live[j]] = b->lres[j];
  b->npairs = total;
  return GDP_OK;
}

// ------------------------------------------------------------ single gaps

namespace {

// Dynprog_single_gap's scoring by defect rate (dynprog_single.c:458-481;
// DEFECT_HIGHQ 0.003 and DEFECT_MEDQ 0.014, dynprog.h:58-59;
// SINGLE_OPEN/EXTEND_{HIGHQ,MEDQ,LOWQ}, dynprog.h:61-67).
void single_scoring(double defect_rate, int &mt, int &open, int &extend) {
  if (defect_rate < 0.003) {
    mt = GDP_HIGHQ;
    open = -8;
    extend = -3;
  } else if (defect_rate < 0.014) {
    mt = GDP_MEDQ;
    open = -7;
    extend = -2;
  } else {
    mt = GDP_LOWQ;
    open = -6;
    extend = -1;
  }
}

}  // namespace

struct gdp_sg_batch {
  Grow q, descs, segs, g, dirs, fin, reqs, devs, ps, cs, pd, cd, len, choice, off, out, res;
  std::vector<gdp_problem> probs;
  std::vector<uint8_t> hdescs;
  std::vector<gdp_segment> hsegs;
  std::vector<gdp_tb_request> hreqs;
  std::vector<gdp_sg_dev> hdevs;
  std::vector<int32_t> live, hlen;
  std::vector<int64_t> hoff;
  std::vector<gdp_sg_result> results, lres;
  PinnedGrow pairs;
  int64_t npairs = 0;
  ~gdp_sg_batch() {
    for (Grow *x : {&q, &descs, &segs, &g, &dirs, &fin, &reqs, &devs, &ps, &cs, &pd, &cd, &len, &choice, &off, &out,
                    &res})
      x->release();
    pairs.release();
  }
};

extern "C" gdp_sg_batch *gdp_sg_create(void) { return new gdp_sg_batch(); }
extern "C" void gdp_sg_destroy(gdp_sg_batch *b) { delete b; }
extern "C" int64_t gdp_sg_npairs(const gdp_sg_batch *b) { return b ? b->npairs : 0; }
extern "C" const gdp_sg_result *gdp_sg_results(const gdp_sg_batch *b) { return b ? b->results.data() : nullptr; }
extern "C" const gdp_pair *gdp_sg_pairs(const gdp_sg_batch *b) {
  return b ? static_cast<const gdp_pair *>(b->pairs.p) : nullptr;
}

extern "C" int gdp_sg_run(gdp_sg_batch *b, const gdp_sg_call *calls, int32_t n, const char *qbuf, int64_t qbytes,
                          const void *d_genome, int32_t genome_format, uint32_t genome_len, int32_t max_rlength,
                          int32_t max_glength, int32_t user_dynprog, int32_t user_open, int32_t user_extend,
                          void *stream_) {
  if (b == nullptr || n < 0 || (n > 0 && (calls == nullptr || qbuf == nullptr || d_genome == nullptr)) || qbytes < 0)
    return err(GDP_EINVAL, "gdp_sg_run: null or negative argument");
  if (genome_format != GDP_GENOME_ASCII && genome_format != GDP_GENOME_COMP)
    return err(GDP_EINVAL, "gdp_sg_run: unknown genome format");
  if (user_dynprog && (user_open > 0 || user_open < -127 || user_extend > 0 || user_extend < -127))
    return err(GDP_EINVAL, "gdp_sg_run: gap penalties must lie in [-127, 0]");
  hipStream_t stream = (hipStream_t)stream_;
  b->npairs = 0;
  // lengths the workspace cannot hold: NULL with NEG_INFINITY_32
  b->results.assign((size_t)n, gdp_sg_result{0, -1, {-32768, 0, 0, 0, 0}, 0, 0});
  b->live.clear();
  for (int32_t k = 0; k < n; k++) {
    const gdp_sg_call &c = calls[k];
    if (c.rl <= 0 || c.gl <= 0 || c.rl > max_rlength || c.gl > max_glength) continue;
    if (c.qoff < 0 || c.qoff + c.rl > qbytes)
      return err(GDP_EINVAL, "gdp_sg_run: call " + std::to_string(k) + " reads past its buffer");
    if (c.genestrand < 0 || c.genestrand > 2 || c.extraband < 0)
      return err(GDP_EINVAL, "gdp_sg_run: call " + std::to_string(k) + " has a bad genestrand or extraband");
    b->live.push_back(k);
  }
  const int m = (int)b->live.size();
  if (m == 0) return GDP_OK;

  b->probs.assign((size_t)m, gdp_problem{});
  b->hsegs.assign((size_t)m, gdp_segment{});
  b->hreqs.assign((size_t)m, gdp_tb_request{});
  b->hdevs.assign((size_t)m, gdp_sg_dev{});
  int64_t gpos = 0, pcap = 0, scap = 0;
  for (int j = 0; j < m; j++) {
    const gdp_sg_call &c = calls[b->live[j]];
    int mt, open, extend;
    single_scoring(c.defect_rate, mt, open, extend);
    if (user_dynprog) {
      open = user_open;
      extend = user_extend;
    }
    int32_t lb = 0, ub = 0;
    gdp_compute_bands(&lb, &ub, c.rl, c.gl, c.extraband, c.widebandp);
    const int use8 = gdp_use8p_size(mt);
    gdp_problem &p = b->probs[j];
    p.qoff = c.qoff;
    p.goff = p.galtoff = gpos;
    p.rl = c.rl;
    p.gl = c.gl;
    p.lband = (int16_t)std::min<int32_t>(lb, 32767);
    p.uband = (int16_t)std::min<int32_t>(ub, 32767);
    p.kind = GDP_KIND_FULL;
    p.width = (c.rl < use8 && c.gl < use8) ? 8 : 16;
    p.mismatchtype = (int8_t)mt;
    p.jump_late = c.jump_late ? 1 : 0;
    p.open = (int8_t)open;
    p.extend = (int8_t)extend;
    gdp_segment &sg = b->hsegs[j];
    if (c.watsonp) {
      sg.coord = c.chroffset + (uint32_t)c.goffset;
      sg.bound = c.chrhigh;
    } else {
      sg.coord = c.chrhigh - (uint32_t)c.goffset + 1u;
      sg.bound = c.chroffset;
      sg.left = 1;
      sg.revcomp = 1;
    }
    sg.length = c.gl;
    sg.out_off = gpos;
    gdp_tb_request &t = b->hreqs[j];
    t.fill = j;
    t.r = c.rl;
    t.c = c.gl;
    t.queryoffset = c.roffset;
    t.genomeoffset = c.goffset;
    t.chroffset = c.chroffset;
    t.chrhigh = c.chrhigh;
    t.watsonp = c.watsonp ? 1 : 0;
    t.genestrand = c.genestrand;
    t.pair_off = pcap;
    t.max_pairs = c.rl + c.gl + 2;
    gdp_sg_dev &d = b->hdevs[j];
    d.qoff = c.qoff;
    d.goff = gpos;
    d.simple_off = c.gl == c.rl ? scap : -1;
    d.rl = c.rl;
    d.gl = c.gl;
    d.roffset = c.roffset;
    d.goffset = c.goffset;
    d.genestrand = c.genestrand;
    if (c.gl == c.rl) scap += c.rl;
    gpos += c.gl;
    pcap += t.max_pairs;
  }
  b->hdescs.resize((size_t)gdp_plan_bytes(m));
  int64_t dw = 0, ss = 0;
  // Single gaps rarely share a band: mixed paired waves, problems binned by
  // band in steps of 4 columns and by rows (what the throughput autotuner
  // picks for this length mix).
  const gdp_plan_params mixed = {240, GDP_MAX_LEN + 1, 4, 1};
  int rc = gdp_plan_tuned(b->probs.data(), m, 0, &mixed, b->hdescs.data(), &dw, &ss);
  if (rc != GDP_OK) return rc;
  const uint32_t kernels = gdp_plan_kernels(b->hdescs.data(), m);
  const int64_t used = gdp_plan_used_bytes(b->hdescs.data(), m);
  if (!b->q.ensure((size_t)qbytes + 16) || !b->descs.ensure((size_t)used) ||
      !b->segs.ensure(sizeof(gdp_segment) * (size_t)m) || !b->g.ensure((size_t)gpos + 16) ||
      !b->dirs.ensure(sizeof(uint32_t) * (size_t)(dw > 0 ? dw : 1)) || !b->fin.ensure(sizeof(int32_t) * (size_t)m) ||
      !b->reqs.ensure(sizeof(gdp_tb_request) * (size_t)m) || !b->devs.ensure(sizeof(gdp_sg_dev) * (size_t)m) ||
      !b->ps.ensure(sizeof(gdp_pair) * (size_t)(scap > 0 ? scap : 1)) ||
      !b->cs.ensure(sizeof(int32_t) * 6 * (size_t)m) || !b->pd.ensure(sizeof(gdp_pair) * (size_t)pcap) ||
      !b->cd.ensure(sizeof(int32_t) * 6 * (size_t)m) || !b->len.ensure(sizeof(int32_t) * (size_t)m) ||
      !b->choice.ensure(sizeof(int32_t) * (size_t)m) || !b->off.ensure(sizeof(int64_t) * (size_t)m) ||
      !b->res.ensure(sizeof(gdp_sg_result) * (size_t)m))
    return err(GDP_ENOMEM, "gdp_sg_run: device allocation failed");
  auto up = [&](Grow &d, const void *h, size_t bytes) {
    return hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, stream) == hipSuccess;
  };
  if (!up(b->q, qbuf, (size_t)qbytes) || !up(b->descs, b->hdescs.data(), (size_t)used) ||
      !up(b->segs, b->hsegs.data(), sizeof(gdp_segment) * (size_t)m) ||
      !up(b->reqs, b->hreqs.data(), sizeof(gdp_tb_request) * (size_t)m) ||
      !up(b->devs, b->hdevs.data(), sizeof(gdp_sg_dev) * (size_t)m))
    return err(GDP_EHIP, "gdp_sg_run: upload failed");
  const char *dq = b->q.as<char>(), *dg = b->g.as<char>();
  if ((rc = gdp_genome_segments(d_genome, genome_format, genome_len, b->segs.as<gdp_segment>(), m, b->g.as<char>(),
                                stream)) != GDP_OK)
    return rc;
  if (gdp_launch_sg(0, m, b->devs.as<gdp_sg_dev>(), dq, dg, b->ps.as<gdp_pair>(), b->cs.as<int32_t>(), nullptr,
                    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, stream))
    return err(GDP_EHIP, "gdp_sg_run: single_gap_simple failed");
  if ((rc = gdp_fill_ex(b->descs.p, m, kernels, dq, dg, b->dirs.as<uint32_t>(), nullptr, b->fin.as<int32_t>(),
                        stream)) != GDP_OK)
    return rc;
  if ((rc = gdp_traceback_ex(b->descs.p, b->reqs.as<gdp_tb_request>(), m, dq, dg, b->dirs.as<uint32_t>(), d_genome,
                             genome_format, genome_len, b->pd.as<gdp_pair>(), b->cd.as<int32_t>(), stream)) != GDP_OK)
    return rc;
  if (gdp_launch_sg(1, m, nullptr, nullptr, nullptr, nullptr, b->cs.as<int32_t>(), nullptr, nullptr,
                    b->cd.as<int32_t>(), b->len.as<int32_t>(), b->choice.as<int32_t>(), nullptr, nullptr, nullptr,
                    stream))
    return err(GDP_EHIP, "gdp_sg_run: lengths failed");
  b->hlen.resize((size_t)m);
  if (hipMemcpyAsync(b->hlen.data(), b->len.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_sg_run: lengths copy failed");
  b->hoff.resize((size_t)m);
  int64_t total = 0;
  for (int j = 0; j < m; j++) {
    b->hoff[j] = total;
    total += b->hlen[j] > 0 ? b->hlen[j] : 0;
  }
  if (!b->out.ensure(sizeof(gdp_pair) * (size_t)(total > 0 ? total : 1)) ||
      !up(b->off, b->hoff.data(), sizeof(int64_t) * (size_t)m))
    return err(GDP_ENOMEM, "gdp_sg_run: output allocation failed");
  if (gdp_launch_sg(2, m, b->devs.as<gdp_sg_dev>(), nullptr, nullptr, b->ps.as<gdp_pair>(), b->cs.as<int32_t>(),
                    b->reqs.as<gdp_tb_request>(), b->pd.as<gdp_pair>(), b->cd.as<int32_t>(), b->len.as<int32_t>(),
                    b->choice.as<int32_t>(), b->off.as<int64_t>(), b->out.as<gdp_pair>(), b->res.as<gdp_sg_result>(),
                    stream))
    return err(GDP_EHIP, "gdp_sg_run: join failed");
  b->lres.resize((size_t)m);
  if (!b->pairs.ensure(sizeof(gdp_pair) * (size_t)(total > 0 ? total : 1)))
    return err(GDP_ENOMEM, "gdp_sg_run: pinned host allocation failed");
  if (hipMemcpyAsync(b->lres.data(), b->res.p, sizeof(gdp_sg_result) * m, hipMemcpyDeviceToHost, stream) !=
          hipSuccess ||
      (total > 0 &&
       hipMemcpyAsync(b->pairs.p, b->out.p, sizeof(gdp_pair) * (size_t)total, hipMemcpyDeviceToHost, stream) !=
           hipSuccess) ||
      hipStreamSynchronize(stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_sg_run: result copy failed");
  for (int j = 0; j < m; j++) b->results[b->live[j]] = b->lres[j];
  b->npairs = total;
  return GDP_OK;
}

// --------------------------------------------------------------- end gaps

extern "C" int gdp_launch_eg(int stage, int m, int five, const int32_t *flags, const gdp_tb_request *reqs,
                             const gdp_pair *pd, int32_t *cd, const int32_t *best, int32_t *len, int32_t *trim,
                             const int64_t *off, gdp_pair *out, gdp_eg_result *res, hipStream_t stream);

static_assert(sizeof(gdp_eg_call) == 48, "gdp_eg_call is 48 bytes");
static_assert(sizeof(gdp_eg_result) == 40, "gdp_eg_result is 40 bytes");

struct gdp_eg_batch {
  Grow q, descs, segs, g, dirs, scores, fin, ul, best, tmpl, reqs, flags, pd, cd, len, trim, off, out, res;
  std::vector<gdp_problem> probs;
  std::vector<uint8_t> hdescs;
  std::vector<char> hq;
  std::vector<gdp_segment> hsegs;
  std::vector<gdp_tb_request> hreqs;
  std::vector<int32_t> live, hul, hflags, hlen;
  std::vector<int64_t> hoff;
  std::vector<gdp_eg_result> results, lres;
  PinnedGrow pairs;
  int64_t npairs = 0;
  ~gdp_eg_batch() {
    for (Grow *x : {&q, &descs, &segs, &g, &dirs, &scores, &fin, &ul, &best, &tmpl, &reqs, &flags, &pd, &cd, &len,
                    &trim, &off, &out, &res})
      x->release();
    pairs.release();
  }
};

extern "C" gdp_eg_batch *gdp_eg_create(void) { return new gdp_eg_batch(); }
extern "C" void gdp_eg_destroy(gdp_eg_batch *b) { delete b; }
extern "C" int64_t gdp_eg_npairs(const gdp_eg_batch *b) { return b ? b->npairs : 0; }
extern "C" const gdp_eg_result *gdp_eg_results(const gdp_eg_batch *b) { return b ? b->results.data() : nullptr; }
extern "C" const gdp_pair *gdp_eg_pairs(const gdp_eg_batch *b) {
  return b ? static_cast<const gdp_pair *>(b->pairs.p) : nullptr;
}

extern "C" int gdp_eg_run(gdp_eg_batch *b, int32_t five, const gdp_eg_call *calls, int32_t n, const char *qbuf,
                          int64_t qbytes, const void *d_genome, int32_t genome_format, uint32_t genome_len,
                          int32_t max_rlength, int32_t max_glength, int32_t user_dynprog, int32_t user_open,
                          int32_t user_extend, int32_t require_pos_score, void *stream_) {
  if (b == nullptr || n < 0 || (n > 0 && (calls == nullptr || qbuf == nullptr || d_genome == nullptr)) || qbytes < 0)
    return err(GDP_EINVAL, "gdp_eg_run: null or negative argument");
  if (genome_format != GDP_GENOME_ASCII && genome_format != GDP_GENOME_COMP)
    return err(GDP_EINVAL, "gdp_eg_run: unknown genome format");
  if (user_dynprog && (user_open > 0 || user_open < -127 || user_extend > 0 || user_extend < -127))
    return err(GDP_EINVAL, "gdp_eg_run: gap penalties must lie in [-127, 0]");
  hipStream_t stream = (hipStream_t)stream_;
  b->npairs = 0;
  b->results.assign((size_t)n, gdp_eg_result{0, -1, {0, 0, 0, 0, 0}, 0, 0});
  // live calls grouped by end mode: QUERYEND_GAP / BEST_LOCAL, then
  // QUERYEND_INDELS, then QUERYEND_NOGAPS; each call's chopped lengths
  std::vector<int32_t> group[3], rlx((size_t)n), glx((size_t)n);
  for (int32_t k = 0; k < n; k++) {
    const gdp_eg_call &c = calls[k];
    if (c.endalign < 0 || c.endalign > 3)
      return err(GDP_EINVAL, "gdp_eg_run: call " + std::to_string(k) + " has a bad endalign");
    if (c.genestrand < 0 || c.genestrand > 2 || c.extraband < 0)
      return err(GDP_EINVAL, "gdp_eg_run: call " + std::to_string(k) + " has a bad genestrand or extraband");
    const bool nogaps = c.endalign == 2;
    if (c.rl <= 0 || (five && c.goffset < 0) || c.gl <= 0) continue;  // NULL, counters 0, dynprogindex kept
    if (c.qoff < 0 || c.qoff + c.rl > qbytes)
      return err(GDP_EINVAL, "gdp_eg_run: call " + std::to_string(k) + " reads past its buffer");
    rlx[k] = nogaps ? c.rl : std::min(c.rl, max_rlength);
    glx[k] = nogaps ? c.gl : std::min(c.gl, max_glength);
    if (rlx[k] > GDP_MAX_LEN || glx[k] > GDP_MAX_LEN)
      return err(GDP_ETOOBIG, "gdp_eg_run: call " + std::to_string(k) + " is longer than GDP_MAX_LEN");
    group[c.endalign == 1 ? 1 : (nogaps ? 2 : 0)].push_back(k);
  }
  b->live.clear();
  for (auto &gr : group) b->live.insert(b->live.end(), gr.begin(), gr.end());
  const int m = (int)b->live.size(), m0 = (int)group[0].size(), m1 = (int)group[1].size();
  if (m == 0) return GDP_OK;

  const int use8 = gdp_use8p_size(kEndq);
  b->probs.clear();
  int64_t qtotal = 0;
  for (int j = 0; j < m; j++) qtotal += rlx[b->live[j]];
  b->hq.resize((size_t)qtotal);
  b->hsegs.assign((size_t)m, gdp_segment{});
  b->hreqs.assign((size_t)m, gdp_tb_request{});
  b->hflags.assign((size_t)m, 0);
  b->hul.clear();
  int64_t gpos = 0, pcap = 0, qpos = 0;
  for (int j = 0; j < m; j++) {
    const int32_t k = b->live[j];
    const gdp_eg_call &c = calls[k];
    const int rl = rlx[k], gl = glx[k];
    const bool nogaps = c.endalign == 2;
    // the read in sweep order: the 5' caller reads rev_rsequence[1 - r]
    char *qd = b->hq.data() + qpos;
    if (five)
      std::reverse_copy(qbuf + c.qoff + c.rl - rl, qbuf + c.qoff + c.rl, qd);
    else
      memcpy(qd, qbuf + c.qoff, (size_t)rl);
    gdp_segment &sg = b->hsegs[j];
    if (five) {
      if (c.watsonp) {
        sg.coord = c.chroffset + (uint32_t)c.goffset + 1u;
        sg.bound = c.chroffset;
        sg.left = 1;
      } else {
        sg.coord = c.chrhigh - (uint32_t)c.goffset;
        sg.bound = c.chrhigh;
        sg.revcomp = 1;
      }
      sg.reverse = 1;  // the 5' fills read the window from its last base
    } else if (c.watsonp) {
      sg.coord = c.chroffset + (uint32_t)c.goffset;
      sg.bound = c.chrhigh;
    } else {
      sg.coord = c.chrhigh - (uint32_t)c.goffset + 1u;
      sg.bound = c.chroffset;
      sg.left = 1;
      sg.revcomp = 1;
    }
    sg.length = gl;
    sg.out_off = gpos;
    const int open = user_dynprog ? user_open : end_open(c.defect_rate);
    const int extend = user_dynprog ? user_extend : kEndExtend;
    int32_t lb = 0, ub = 0;
    gdp_compute_bands(&lb, &ub, rl, gl, c.extraband, c.endalign == 0 || c.endalign == 3);
    const int first = (int)b->probs.size();
    for (int t = 0; t < (nogaps ? 1 : 2); t++) {
      gdp_problem p;
      memset(&p, 0, sizeof p);
      p.qoff = qpos;
      p.goff = p.galtoff = gpos;
      p.rl = rl;
      p.gl = gl;
      p.kind = t == 0 ? GDP_KIND_UPPER : GDP_KIND_LOWER;
      // the nogaps walk reads no directions: its sequences ride on a
      // one-column upper fill
      p.uband = (int16_t)std::min<int32_t>(nogaps ? 0 : ub, 32767);
      p.lband = (int16_t)std::min<int32_t>(lb, 32767);
      p.width = (rl < use8 || gl < use8) ? 8 : 16;
      p.mismatchtype = kEndq;
      p.jump_late = five ? !c.jump_late : (c.jump_late != 0);
      p.open = (int8_t)open;
      p.extend = (int8_t)extend;
      p.keep_scores = nogaps ? 0 : (c.endalign == 1 ? GDP_KEEP_BEST_QUERYEND : GDP_KEEP_BEST);
      b->probs.push_back(p);
    }
    if (!nogaps) {
      b->hul.push_back(first);
      b->hul.push_back(first + 1);
    }
    gdp_tb_request &q = b->hreqs[j];
    q.fill = first;
    q.queryoffset = c.roffset;
    q.genomeoffset = c.goffset;
    q.chroffset = c.chroffset;
    q.chrhigh = c.chrhigh;
    q.watsonp = c.watsonp ? 1 : 0;
    q.revp = five ? 1 : 0;
    q.genestrand = c.genestrand;
    q.pair_off = pcap;
    q.max_pairs = rl + gl + 2;
    if (nogaps) {  // find_best_endpoint_to_queryend_nogaps (dynprog_end.c:576-586)
      q.r = q.c = std::min(rl, gl);
      q.flags = GDP_TB_NOGAPS;
    }
    b->hflags[j] = c.endalign | (require_pos_score ? 4 : 0);
    gpos += gl;
    qpos += rl;
    pcap += q.max_pairs;
  }
  const int np = (int)b->probs.size();
  b->hdescs.resize((size_t)gdp_plan_bytes(np));
  int64_t dw = 0, ss = 0;
  int rc = gdp_plan_ex(b->probs.data(), np, 0, b->hdescs.data(), &dw, &ss);
  if (rc != GDP_OK) return rc;
  const uint32_t kernels = gdp_plan_kernels(b->hdescs.data(), np);
  const int64_t used = gdp_plan_used_bytes(b->hdescs.data(), np);
  const size_t tb = sizeof(gdp_tb_request) * (size_t)m;
  if (!b->q.ensure(b->hq.size() + 16) || !b->descs.ensure((size_t)used) ||
      !b->segs.ensure(sizeof(gdp_segment) * (size_t)m) || !b->g.ensure((size_t)gpos + 16) ||
      !b->dirs.ensure(sizeof(uint32_t) * (size_t)(dw > 0 ? dw : 1)) ||
      !b->scores.ensure(sizeof(int16_t) * (size_t)(ss > 0 ? ss : 1)) || !b->fin.ensure(sizeof(int32_t) * np) ||
      !b->ul.ensure(sizeof(int32_t) * (b->hul.size() + 1)) || !b->best.ensure(sizeof(int32_t) * 3 * (size_t)m) ||
      !b->tmpl.ensure(tb) || !b->reqs.ensure(tb) || !b->flags.ensure(sizeof(int32_t) * (size_t)m) ||
      !b->pd.ensure(sizeof(gdp_pair) * (size_t)pcap) || !b->cd.ensure(sizeof(int32_t) * 6 * (size_t)m) ||
      !b->len.ensure(sizeof(int32_t) * (size_t)m) || !b->trim.ensure(sizeof(int32_t) * (size_t)m) ||
      !b->off.ensure(sizeof(int64_t) * (size_t)m) || !b->res.ensure(sizeof(gdp_eg_result) * (size_t)m))
    return err(GDP_ENOMEM, "gdp_eg_run: device allocation failed");
  auto up = [&](Grow &d, const void *h, size_t bytes) {
    return bytes == 0 || hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, stream) == hipSuccess;
  };
  if (!up(b->q, b->hq.data(), b->hq.size()) || !up(b->descs, b->hdescs.data(), (size_t)used) ||
      !up(b->segs, b->hsegs.data(), sizeof(gdp_segment) * (size_t)m) ||
      !up(b->ul, b->hul.data(), sizeof(int32_t) * b->hul.size()) || !up(b->tmpl, b->hreqs.data(), tb) ||
      !up(b->reqs, b->hreqs.data(), tb) || !up(b->flags, b->hflags.data(), sizeof(int32_t) * (size_t)m))
    return err(GDP_EHIP, "gdp_eg_run: upload failed");
  // the nogaps calls have no best cell: finalscore 0
  if (hipMemsetAsync(b->best.p, 0, sizeof(int32_t) * 3 * (size_t)m, stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_eg_run: clear failed");
  const char *dq = b->q.as<char>(), *dg = b->g.as<char>();
  if ((rc = gdp_genome_segments(d_genome, genome_format, genome_len, b->segs.as<gdp_segment>(), m, b->g.as<char>(),
                                stream)) != GDP_OK)
    return rc;
  if ((rc = gdp_fill_ex(b->descs.p, np, kernels, dq, dg, b->dirs.as<uint32_t>(), b->scores.as<int16_t>(),
                        b->fin.as<int32_t>(), stream)) != GDP_OK)
    return rc;
  // best end cells of each mode's calls, writing their walks' requests
  if (m0 > 0 &&
      (rc = gdp_best_endpoint_ex(b->descs.p, b->ul.as<int32_t>(), m0, 0, b->scores.as<int16_t>(),
                                 b->best.as<int32_t>(), b->tmpl.as<gdp_tb_request>(), b->reqs.as<gdp_tb_request>(),
                                 stream)) != GDP_OK)
    return rc;
  if (m1 > 0 && (rc = gdp_best_endpoint_ex(b->descs.p, b->ul.as<int32_t>() + 2 * m0, m1, 1, b->scores.as<int16_t>(),
                                           b->best.as<int32_t>() + 3 * m0, b->tmpl.as<gdp_tb_request>() + m0,
                                           b->reqs.as<gdp_tb_request>() + m0, stream)) != GDP_OK)
    return rc;
  if ((rc = gdp_traceback_ex(b->descs.p, b->reqs.as<gdp_tb_request>(), m, dq, dg, b->dirs.as<uint32_t>(), d_genome,
                             genome_format, genome_len, b->pd.as<gdp_pair>(), b->cd.as<int32_t>(), stream)) != GDP_OK)
    return rc;
  if (gdp_launch_eg(0, m, five, b->flags.as<int32_t>(), b->reqs.as<gdp_tb_request>(), b->pd.as<gdp_pair>(),
                    b->cd.as<int32_t>(), nullptr, b->len.as<int32_t>(), b->trim.as<int32_t>(), nullptr, nullptr,
                    nullptr, stream))
    return err(GDP_EHIP, "gdp_eg_run: lengths failed");
  b->hlen.resize((size_t)m);
  if (hipMemcpyAsync(b->hlen.data(), b->len.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_eg_run: lengths copy failed");
  b->hoff.resize((size_t)m);
  int64_t total = 0;
  for (int j = 0; j < m; j++) {
    b->hoff[j] = total;
    total += b->hlen[j] > 0 ? b->hlen[j] : 0;
  }
  if (!b->out.ensure(sizeof(gdp_pair) * (size_t)(total > 0 ? total : 1)) ||
      !up(b->off, b->hoff.data(), sizeof(int64_t) * (size_t)m))
    return err(GDP_ENOMEM, "gdp_eg_run: output allocation failed");
  if (gdp_launch_eg(1, m, five, nullptr, b->reqs.as<gdp_tb_request>(), b->pd.as<gdp_pair>(), b->cd.as<int32_t>(),
                    b->best.as<int32_t>(), b->len.as<int32_t>(), b->trim.as<int32_t>(), b->off.as<int64_t>(),
                    b->out.as<gdp_pair>(), b->res.as<gdp_eg_result>(), stream))
    return err(GDP_EHIP, "gdp_eg_run: join failed");
  b->lres.resize((size_t)m);
  if (!b->pairs.ensure(sizeof(gdp_pair) * (size_t)(total > 0 ? total : 1)))
    return err(GDP_ENOMEM, "gdp_eg_run: pinned host allocation failed");
  if (hipMemcpyAsync(b->lres.data(), b->res.p, sizeof(gdp_eg_result) * m, hipMemcpyDeviceToHost, stream) !=
          hipSuccess ||
      (total > 0 &&
       hipMemcpyAsync(b->pairs.p, b->out.p, sizeof(gdp_pair) * (size_t)total, hipMemcpyDeviceToHost, stream) !=
           hipSuccess) ||
      hipStreamSynchronize(stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_eg_run: result copy failed");
  for (int j = 0; j < m; j++) b->results[b->live[j]] = b->lres[j];
  b->npairs = total;
  return GDP_OK;
}

// --------------------------------------------------------------- cDNA gaps

extern "C" int gdp_launch_cd(int stage, int m, const gdp_cd_dev *calls, const char *qbuf, const void *genome,
                             int fmt, uint32_t glen, const int32_t *bout, const gdp_tb_request *reqs,
                             const gdp_pair *pd, const int32_t *cd, int32_t *len, const int64_t *off, gdp_pair *out,
                             gdp_cd_result *res, hipStream_t stream);

static_assert(sizeof(gdp_cd_call) == 64, "gdp_cd_call is 64 bytes");
static_assert(sizeof(gdp_cd_result) == 32, "gdp_cd_result is 32 bytes");

namespace {
constexpr int kCdnaOpen = -10, kCdnaExtend = -7;  // CDNA_OPEN_*, CDNA_EXTEND_* (dynprog_cdna.c:31-37)
}

struct gdp_cd_batch {
  Grow q, descs, segs, g, dirs, scores, fin, quads, bout, tmpl, reqs, devs, pd, cd, len, off, out, res;
  std::vector<gdp_problem> probs;
  std::vector<uint8_t> hdescs;
  std::vector<char> hq;
  std::vector<gdp_segment> hsegs;
  std::vector<gdp_tb_request> hreqs;
  std::vector<gdp_cd_dev> hdevs;
  std::vector<int32_t> live, hquads, hlen;
  std::vector<int64_t> hoff;
  std::vector<gdp_cd_result> results, lres;
  PinnedGrow pairs;
  int64_t npairs = 0;
  ~gdp_cd_batch() {
    for (Grow *x : {&q, &descs, &segs, &g, &dirs, &scores, &fin, &quads, &bout, &tmpl, &reqs, &devs, &pd, &cd, &len,
                    &off, &out, &res})
      x->release();
    pairs.release();
  }
};

extern "C" gdp_cd_batch *gdp_cd_create(void) { return new gdp_cd_batch(); }
extern "C" void gdp_cd_destroy(gdp_cd_batch *b) { delete b; }
extern "C" int64_t gdp_cd_npairs(const gdp_cd_batch *b) { return b ? b->npairs : 0; }
extern "C" const gdp_cd_result *gdp_cd_results(const gdp_cd_batch *b) { return b ? b->results.data() : nullptr; }
extern "C" const gdp_pair *gdp_cd_pairs(const gdp_cd_batch *b) {
  return b ? static_cast<const gdp_pair *>(b->pairs.p) : nullptr;
}

extern "C" int gdp_cd_run(gdp_cd_batch *b, const gdp_cd_call *calls, int32_t n, const char *qbuf, int64_t qbytes,
                          const void *d_genome, int32_t genome_format, uint32_t genome_len, int32_t max_rlength,
                          int32_t max_glength, void *stream_) {
  if (b == nullptr || n < 0 || (n > 0 && (calls == nullptr || qbuf == nullptr || d_genome == nullptr)) || qbytes < 0)
    return err(GDP_EINVAL, "gdp_cd_run: null or negative argument");
  if (genome_format != GDP_GENOME_ASCII && genome_format != GDP_GENOME_COMP)
    return err(GDP_EINVAL, "gdp_cd_run: unknown genome format");
  hipStream_t stream = (hipStream_t)stream_;
  b->npairs = 0;
  b->results.assign((size_t)n, gdp_cd_result{0, -1, 0, 0, 0, 0, 0});
  b->live.clear();
  for (int32_t k = 0; k < n; k++) {
    const gdp_cd_call &c = calls[k];
    if (c.glength <= 1) continue;  // NULL before anything else (dynprog_cdna.c:828-830)
    if (c.glength > max_glength || c.rlengthR > max_rlength || c.rlengthL > max_rlength) {
      b->results[k].advanced = 1;  // (dynprog_cdna.c:864-893)
      continue;
    }
    if (c.rlengthL <= 0 || c.rlengthR <= 0 || c.extraband < 0 || c.genestrand < 0 || c.genestrand > 2)
      return err(GDP_EINVAL, "gdp_cd_run: call " + std::to_string(k) + " has a bad length, extraband or genestrand");
    // the left read up to the right end's first base (a join's cDNA
    // insertion copies those bytes), the right read backwards
    const int64_t spanL = std::max<int64_t>(c.rlengthL, (int64_t)c.rev_roffsetR - c.roffsetL + 1);
    if (c.qoffL < 0 || c.qoffL + spanL > qbytes || c.qoffR >= qbytes || c.qoffR - c.rlengthR + 1 < 0)
      return err(GDP_EINVAL, "gdp_cd_run: call " + std::to_string(k) + " reads past its buffer");
    b->live.push_back(k);
  }
  const int m = (int)b->live.size();
  if (m == 0) return GDP_OK;

  // four one-sided fills per call: left upper, left lower, right upper,
  // right lower (dynprog_cdna.c:1016-1040); the right ones over the reversed
  // read, stored after the caller's bytes, and the reversed window
  int64_t qtotal = 0;
  for (int j = 0; j < m; j++) qtotal += calls[b->live[j]].rlengthR;
  const int64_t qbaseR = (qbytes + 15) & ~(int64_t)15;
  b->hq.resize((size_t)qtotal);
  b->probs.assign(4 * (size_t)m, gdp_problem{});
  b->hsegs.assign(2 * (size_t)m, gdp_segment{});
  b->hreqs.assign(2 * (size_t)m, gdp_tb_request{});
  b->hdevs.assign((size_t)m, gdp_cd_dev{});
  b->hquads.resize(5 * (size_t)m);
  int64_t gpos = 0, pcap = 0, qr = 0;
  for (int j = 0; j < m; j++) {
    const gdp_cd_call &c = calls[b->live[j]];
    const int gl = c.glength, rev_goffset = c.goffset + gl - 1;
    // mismatch type by defect rate (dynprog_cdna.c:846-862; DEFECT_HIGHQ
    // 0.003 and DEFECT_MEDQ 0.014, dynprog.h:58-59)
    const int mt = c.defect_rate < 0.003 ? GDP_HIGHQ : (c.defect_rate < 0.014 ? GDP_MEDQ : GDP_LOWQ);
    const int u8 = gdp_use8p_size(mt);
    const int width = (gl < u8 || (c.rlengthL < u8 && c.rlengthR <= u8)) ? 8 : 16;  // (906-912)
    int32_t lbL = 0, ubL = 0, lbR = 0, ubR = 0;
    gdp_compute_bands(&lbL, &ubL, c.rlengthL, gl, c.extraband, 1);
    gdp_compute_bands(&lbR, &ubR, c.rlengthR, gl, c.extraband, 1);
    for (int i = 0; i < c.rlengthR; i++) b->hq[(size_t)(qr + i)] = qbuf[c.qoffR - i];
    const int64_t gL = gpos, gR = gpos + gl;
    for (int t = 0; t < 4; t++) {
      const bool right = t >= 2;
      gdp_problem &p = b->probs[4 * (size_t)j + t];
      p.qoff = right ? qbaseR + qr : c.qoffL;
      p.goff = p.galtoff = right ? gR : gL;
      p.rl = right ? c.rlengthR : c.rlengthL;
      p.gl = gl;
      p.lband = (int16_t)std::min<int32_t>(right ? lbR : lbL, 32767);
      p.uband = (int16_t)std::min<int32_t>(right ? ubR : ubL, 32767);
      p.kind = (t & 1) ? GDP_KIND_LOWER : GDP_KIND_UPPER;
      p.width = (int8_t)width;
      p.mismatchtype = (int8_t)mt;
      p.jump_late = right ? !c.jump_late : (c.jump_late != 0);  // for revp the other rule
      p.open = (int8_t)kCdnaOpen;
      p.extend = (int8_t)kCdnaExtend;
      p.keep_scores = GDP_KEEP_ALL;
    }
    // the windows (dynprog_cdna.c:921-931): gsequence from goffset and
    // rev_gsequence up to rev_goffset, read backwards by the right fills
    gdp_segment &sL = b->hsegs[2 * (size_t)j], &sR = b->hsegs[2 * (size_t)j + 1];
    if (c.watsonp) {
      sL.coord = c.chroffset + (uint32_t)c.goffset;
      sL.bound = c.chrhigh;
      sR.coord = c.chroffset + (uint32_t)rev_goffset + 1u;
      sR.bound = c.chroffset;
      sR.left = 1;
    } else {
      sL.coord = c.chrhigh - (uint32_t)c.goffset + 1u;
      sL.bound = c.chroffset;
      sL.left = 1;
      sL.revcomp = 1;
      sR.coord = c.chrhigh - (uint32_t)rev_goffset;
      sR.bound = c.chrhigh;
      sR.revcomp = 1;
    }
    sL.length = sR.length = gl;
    sL.out_off = gL;
    sR.out_off = gR;
    sR.reverse = 1;
    int32_t *qd = &b->hquads[5 * (size_t)j];
    qd[0] = 4 * j;
    qd[1] = 4 * j + 1;
    qd[2] = 4 * j + 2;
    qd[3] = 4 * j + 3;
    qd[4] = c.rev_roffsetR - c.roffsetL;
    // walk templates, the right end first (the bridge sets fill, r and c)
    for (int t = 0; t < 2; t++) {
      gdp_tb_request &q = b->hreqs[2 * (size_t)j + t];
      const bool right = t == 0;
      q.queryoffset = right ? c.rev_roffsetR : c.roffsetL;
      q.genomeoffset = right ? rev_goffset : c.goffset;
      q.chroffset = c.chroffset;
      q.chrhigh = c.chrhigh;
      q.watsonp = c.watsonp ? 1 : 0;
      q.genestrand = c.genestrand;
      q.revp = right ? 1 : 0;
      q.pair_off = pcap;
      q.max_pairs = (right ? c.rlengthR : c.rlengthL) + gl + 4;
      pcap += q.max_pairs;
    }
    gdp_cd_dev &d = b->hdevs[(size_t)j];
    d.qoffL = c.qoffL;
    d.roffsetL = c.roffsetL;
    d.rev_roffsetR = c.rev_roffsetR;
    d.goffset = c.goffset;
    d.rev_goffset = rev_goffset;
    d.chroffset = c.chroffset;
    d.chrhigh = c.chrhigh;
    d.watsonp = c.watsonp ? 1 : 0;
    gpos += 2 * (int64_t)gl;
    qr += c.rlengthR;
  }
  const int np = 4 * m;
  b->hdescs.resize((size_t)gdp_plan_bytes(np));
  int64_t dw = 0, ss = 0;
  int rc = gdp_plan_ex(b->probs.data(), np, 0, b->hdescs.data(), &dw, &ss);
  if (rc != GDP_OK) return rc;
  const uint32_t kernels = gdp_plan_kernels(b->hdescs.data(), np);
  const int64_t used = gdp_plan_used_bytes(b->hdescs.data(), np);
  const size_t tb = sizeof(gdp_tb_request) * 2 * (size_t)m;
  if (!b->q.ensure((size_t)(qbaseR + qtotal) + 16) || !b->descs.ensure((size_t)used) ||
      !b->segs.ensure(sizeof(gdp_segment) * 2 * (size_t)m) || !b->g.ensure((size_t)gpos + 16) ||
      !b->dirs.ensure(sizeof(uint32_t) * (size_t)(dw > 0 ? dw : 1)) ||
      !b->scores.ensure(sizeof(int16_t) * (size_t)(ss > 0 ? ss : 1)) || !b->fin.ensure(sizeof(int32_t) * np) ||
      !b->quads.ensure(sizeof(int32_t) * 5 * (size_t)m) || !b->bout.ensure(sizeof(int32_t) * 5 * (size_t)m) ||
      !b->tmpl.ensure(tb) || !b->reqs.ensure(tb) || !b->devs.ensure(sizeof(gdp_cd_dev) * (size_t)m) ||
      !b->pd.ensure(sizeof(gdp_pair) * (size_t)pcap) || !b->cd.ensure(sizeof(int32_t) * 12 * (size_t)m) ||
      !b->len.ensure(sizeof(int32_t) * (size_t)m) || !b->off.ensure(sizeof(int64_t) * (size_t)m) ||
      !b->res.ensure(sizeof(gdp_cd_result) * (size_t)m))
    return err(GDP_ENOMEM, "gdp_cd_run: device allocation failed");
  auto up = [&](void *d, const void *h, size_t bytes) {
    return bytes == 0 || hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream) == hipSuccess;
  };
  if (!up(b->q.p, qbuf, (size_t)qbytes) || !up(b->q.as<char>() + qbaseR, b->hq.data(), b->hq.size()) ||
      !up(b->descs.p, b->hdescs.data(), (size_t)used) ||
      !up(b->segs.p, b->hsegs.data(), sizeof(gdp_segment) * 2 * (size_t)m) ||
      !up(b->quads.p, b->hquads.data(), sizeof(int32_t) * 5 * (size_t)m) || !up(b->tmpl.p, b->hreqs.data(), tb) ||
      !up(b->devs.p, b->hdevs.data(), sizeof(gdp_cd_dev) * (size_t)m))
    return err(GDP_EHIP, "gdp_cd_run: upload failed");
  const char *dq = b->q.as<char>(), *dg = b->g.as<char>();
  if ((rc = gdp_genome_segments(d_genome, genome_format, genome_len, b->segs.as<gdp_segment>(), 2 * m,
                                b->g.as<char>(), stream)) != GDP_OK)
    return rc;
  if ((rc = gdp_fill_ex(b->descs.p, np, kernels, dq, dg, b->dirs.as<uint32_t>(), b->scores.as<int16_t>(),
                        b->fin.as<int32_t>(), stream)) != GDP_OK)
    return rc;
  if ((rc = gdp_bridge_cdna(b->descs.p, b->quads.as<int32_t>(), m, b->scores.as<int16_t>(), b->bout.as<int32_t>(),
                            b->tmpl.as<gdp_tb_request>(), b->reqs.as<gdp_tb_request>(), stream)) != GDP_OK)
    return rc;
  if ((rc = gdp_traceback_ex(b->descs.p, b->reqs.as<gdp_tb_request>(), 2 * m, dq, dg, b->dirs.as<uint32_t>(),
                             d_genome, genome_format, genome_len, b->pd.as<gdp_pair>(), b->cd.as<int32_t>(),
                             stream)) != GDP_OK)
    return rc;
  if (gdp_launch_cd(0, m, b->devs.as<gdp_cd_dev>(), nullptr, nullptr, 0, 0, b->bout.as<int32_t>(), nullptr, nullptr,
                    b->cd.as<int32_t>(), b->len.as<int32_t>(), nullptr, nullptr, nullptr, stream))
    return err(GDP_EHIP, "gdp_cd_run: lengths failed");
  b->hlen.resize((size_t)m);
  if (hipMemcpyAsync(b->hlen.data(), b->len.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_cd_run: lengths copy failed");
  b->hoff.resize((size_t)m);
  int64_t total = 0;
  for (int j = 0; j < m; j++) {
    b->hoff[j] = total;
    total += b->hlen[j] > 0 ? b->hlen[j] : 0;
  }
  if (!b->out.ensure(sizeof(gdp_pair) * (size_t)(total > 0 ? total : 1)) ||
      !up(b->off.p, b->hoff.data(), sizeof(int64_t) * (size_t)m))
    return err(GDP_ENOMEM, "gdp_cd_run: output allocation failed");
  if (gdp_launch_cd(1, m, b->devs.as<gdp_cd_dev>(), dq, d_genome, genome_format, genome_len, b->bout.as<int32_t>(),
                    b->reqs.as<gdp_tb_request>(), b->pd.as<gdp_pair>(), b->cd.as<int32_t>(), b->len.as<int32_t>(),
                    b->off.as<int64_t>(), b->out.as<gdp_pair>(), b->res.as<gdp_cd_result>(), stream))
    return err(GDP_EHIP, "gdp_cd_run: join failed");
  b->lres.resize((size_t)m);
  if (!b->pairs.ensure(sizeof(gdp_pair) * (size_t)(total > 0 ? total : 1)))
    return err(GDP_ENOMEM, "gdp_cd_run: pinned host allocation failed");
  if (hipMemcpyAsync(b->lres.data(), b->res.p, sizeof(gdp_cd_result) * m, hipMemcpyDeviceToHost, stream) !=
          hipSuccess ||
      (total > 0 &&
       hipMemcpyAsync(b->pairs.p, b->out.p, sizeof(gdp_pair) * (size_t)total, hipMemcpyDeviceToHost, stream) !=
           hipSuccess) ||
      hipStreamSynchronize(stream) != hipSuccess)
    return err(GDP_EHIP, "gdp_cd_run: result copy failed");
  for (int j = 0; j < m; j++) b->results[b->live[j]] = b->lres[j];
  b->npairs = total;
  return GDP_OK;
}
