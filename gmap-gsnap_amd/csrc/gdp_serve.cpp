// gdp_serve.cpp -- continuous batching of alignment requests over a device
// genome (the gdp_server_* functions of include/gmapdp.h).
//
// One server thread owns every HIP call.  It keeps up to `slots` batches in
// flight, each on a stream of its own with its own pinned staging buffers
// and device buffers; the direction words of all batches in flight share one
// device budget handed out first-fit.  Whenever a slot is free and requests
// wait, the thread takes the oldest requests that fit the slot and the free
// budget, plans them, and queues on the slot's stream: one upload, the
// genome-segment extraction, the fills, the tracebacks, one download and an
// event.  Finished batches are turned into results as their events fire, so
// new requests enter while earlier batches still run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gdp_internal.h"

extern "C" int gdp_launch_copy(void *, const void *, int64_t, hipStream_t);
extern "C" int gdp_launch_serve_download(const uint8_t *, uint8_t *, int64_t, int64_t, int64_t,
                                         const gdp_tb_request *, int, hipStream_t);

namespace {

inline int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct Req {
  gdp_align_request r;  // r.read is not used after submit
  int16_t lband, uband;
  int8_t width;
  int8_t status;        // 0, or the negated GDP_E* code the request is refused with
  int64_t roff;         // the read's bytes in the read ring (virtual offset)
  double t_submit;
};

// Read bytes of pending requests, copied in at submit and out at launch, in
// arrival order.  Offsets grow without bound; offset v lives at v % size,
// and a read never wraps (the ring skips to the next lap instead).  Bytes
// below `freed` may be overwritten.
struct ReadRing {
  std::vector<char> buf;
  int64_t next = 0, freed = 0;
  int64_t place(int32_t len) const {  // offset a read of len bytes would take
    const int64_t size = (int64_t)buf.size();
    int64_t v = next;
    if (v % size + len > size) v += size - v % size;
    return v;
  }
  bool fits(int64_t v, int32_t len) const { return v + len - freed <= (int64_t)buf.size(); }
  char *at(int64_t v) { return buf.data() + v % (int64_t)buf.size(); }
};

struct DoneBatch {
  std::vector<gdp_align_result> res;  // pairs point into result buffer `buf`
  int buf = -1;                       // -1: no buffer (refused requests)
  size_t next = 0;
};

// First-fit allocator over [0, size) in 256-byte units.
class Arena {
 public:
  void init(int64_t size) {
    free_.clear();
    if (size > 0) free_[0] = size;
  }
  int64_t largest() const {
    int64_t m = 0;
    for (const auto &kv : free_) m = std::max(m, kv.second);
    return m;
  }
  int64_t alloc(int64_t n) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const int64_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[off + n] = len - n;
      return off;
    }
    return -1;
  }
  void release(int64_t off, int64_t n) {
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<int64_t, int64_t> free_;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr, ev0 = nullptr;  // end and start of the batch on the stream
  double t_launch = 0;
  uint8_t *h_up = nullptr, *d_up = nullptr, *d_down = nullptr;
  uint8_t *h_up_dev = nullptr;  // h_up as the device addresses it
  char *d_g = nullptr;
  int buf = -1;  // result buffer of the batch in flight
  bool busy = false;
  std::vector<Req> reqs;
  std::vector<int64_t> pair_off;
  int64_t dir_off = 0, dir_bytes = 0;
  int64_t off_counts = 0, off_pairs = 0, down_bytes = 0;
};

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct gdp_server {
  gdp_server_config cfg{};
  const void *genome = nullptr;
  int fmt = 0;
  uint32_t glen = 0;
  int device = 0;
  uint8_t *d_dirs = nullptr;
  Arena arena;
  std::vector<Slot> slots;
  int64_t up_cap = 0, g_cap = 0, down_cap = 0, q_cap = 0;
  std::vector<gdp_problem> probs;

  std::mutex mu;
  std::condition_variable cv_work;
  std::deque<Req> pending[2];  // requests waiting: short queries, long queries
  ReadRing ring;
  std::condition_variable cv_space;

  std::mutex dmu;
  std::condition_variable cv_done;
  std::deque<DoneBatch> done;
  std::vector<uint8_t *> rbuf;  // pinned result buffers (down_cap bytes each)
  std::vector<uint8_t *> rbuf_dev;  // the same as the device addresses them
  std::vector<int> rfree;       // buffers neither in flight nor held by the consumer
  std::vector<int> handed;      // buffers of results the last poll returned

  std::mutex smu;
  gdp_server_stats st{};
  int64_t dir_in_flight = 0;

  std::atomic<bool> stop{false};
  // Set once a HIP call of the server thread failed: from then on no batch
  // starts, pending and newly submitted requests are refused with GDP_EHIP.
  std::atomic<bool> failed{false};
  std::string fatal;  // what failed (under smu)
  std::thread th;

  void run();
  bool launch(Slot &s);
  void retire(Slot &s);
  void fail_all(const std::string &what);
  void post(DoneBatch &&b) {
    {
      std::lock_guard<std::mutex> lk(dmu);
      done.push_back(std::move(b));
    }
    cv_done.notify_all();
  }
};

static void set_err(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
static void set_err(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  gdp_set_error(buf);
}

// Result of a request refused before it reaches the device: the values
// Dynprog_single_gap returns for lengths it will not align
// (dynprog_single.c:508-520).
static gdp_align_result refused(const Req &q, double t) {
  gdp_align_result r;
  memset(&r, 0, sizeof r);
  r.tag = q.r.tag;
  r.status = -(int)q.status;
  r.counts[1] = -32768;  // NEG_INFINITY_32
  r.t_submit = q.t_submit;
  r.t_done = t;
  return r;
}

void gdp_server::fail_all(const std::string &what) {
  {
    std::lock_guard<std::mutex> lk(smu);
    if (fatal.empty()) fatal = what;
  }
  failed = true;
  DoneBatch b;
  const double t = now_s();
  for (Slot &s : slots) {
    if (!s.busy) continue;
    for (Req &q : s.reqs) {
      q.status = (int8_t)-GDP_EHIP;
      b.res.push_back(refused(q, t));
    }
    s.busy = false;
    if (s.buf >= 0) {
      std::lock_guard<std::mutex> lk(dmu);
      rfree.push_back(s.buf);
      s.buf = -1;
    }
  }
  std::deque<Req> rest;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &pq : pending)
      for (Req &q : pq) rest.push_back(q);
    pending[0].clear();
    pending[1].clear();
    // no read waits in the ring any more: a submitter blocked on ring space
    // wakes up and is refused
    ring.freed = ring.next;
  }
  cv_space.notify_all();
  for (Req &q : rest) {
    q.status = (int8_t)-GDP_EHIP;
    b.res.push_back(refused(q, t));
  }
  {
    std::lock_guard<std::mutex> lk(smu);
    st.completed += (int64_t)b.res.size();
  }
  post(std::move(b));
}

// Takes the oldest pending requests that fit the slot and the free direction
// budget, and queues their work on the slot's stream.  Returns false when
// nothing could be started.
bool gdp_server::launch(Slot &s) {
  {
    std::lock_guard<std::mutex> lk(dmu);
    if (rfree.empty()) return false;  // the consumer holds every result buffer
  }
  std::vector<Req> batch;
  int cls = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    // a batch holds one length class: the one whose oldest request waited longest
    if (pending[0].empty() || (!pending[1].empty() && pending[1].front().t_submit < pending[0].front().t_submit))
      cls = 1;
    std::deque<Req> &pq = pending[cls];
    int64_t qb = 0, gb = 0, pb = 0;
    while (!pq.empty() && (int32_t)batch.size() < cfg.max_batch) {
      const Req &q = pq.front();
      const int64_t pairs = (int64_t)q.r.rl + q.r.gl + 2;
      if (qb + q.r.rl + 16 > q_cap || gb + q.r.gl + 16 > g_cap || pb + pairs > cfg.pair_capacity) break;
      qb += q.r.rl;
      gb += q.r.gl;
      pb += pairs;
      batch.push_back(std::move(pq.front()));
      pq.pop_front();
    }
  }
  if (batch.empty()) return false;

  // Plan; if the direction words do not fit the free budget, give back the
  // younger half and plan again.
  const int64_t free_bytes = arena.largest();
  int n = (int)batch.size();
  int64_t dw = 0, ss = 0;
  for (;;) {
    probs.resize(n);
    int64_t qpos = 0, gpos = 0;
    for (int i = 0; i < n; i++) {
      const Req &q = batch[i];
      gdp_problem &p = probs[i];
      memset(&p, 0, sizeof p);
      p.qoff = qpos;
      p.goff = p.galtoff = gpos;
      p.rl = q.r.rl;
      p.gl = q.r.gl;
      p.lband = q.lband;
      p.uband = q.uband;
      p.kind = GDP_KIND_FULL;
      p.width = q.width;
      p.mismatchtype = q.r.mismatchtype;
      p.jump_late = q.r.jump_late ? 1 : 0;
      p.open = q.r.open;
      p.extend = q.r.extend;
      qpos += q.r.rl;
      gpos += q.r.gl;
    }
    const gdp_plan_params *pp =
        (cfg.plan.g16_max_span || cfg.plan.g64_min_rows || cfg.plan.mix_span || cfg.plan.mix_order) ? &cfg.plan
                                                                                                      : nullptr;
    if (gdp_plan_tuned(probs.data(), n, 0, pp, s.h_up, &dw, &ss) != GDP_OK) {
      // Requests and the plan parameters were validated before; a planning
      // failure is a bug: refuse this batch and stop like a HIP failure.
      const std::string why = std::string("gdp_server: planning failed: ") + gdp_last_error();
      DoneBatch b;
      const double t = now_s();
      for (Req &q : batch) {
        q.status = (int8_t)-GDP_EINVAL;
        b.res.push_back(refused(q, t));
      }
      {
        std::lock_guard<std::mutex> lk(smu);
        st.completed += (int64_t)b.res.size();
      }
      post(std::move(b));
      fail_all(why);
      return true;
    }
    const int64_t need = align_up(dw * 4, 256);
    if (need <= free_bytes) break;
    if (n == 1) {
      bool idle = true;
      for (const Slot &o : slots) idle = idle && !o.busy;
      if (!idle) {  // wait for budget
        std::lock_guard<std::mutex> lk(mu);
        pending[cls].push_front(std::move(batch[0]));
        return false;
      }
      DoneBatch b;  // one request larger than the whole budget
      batch[0].status = (int8_t)-GDP_ENOMEM;
      b.res.push_back(refused(batch[0], now_s()));
      post(std::move(b));
      return true;
    }
    const int keep = n / 2;
    {
      std::lock_guard<std::mutex> lk(mu);
      for (int i = n - 1; i >= keep; i--) pending[cls].push_front(std::move(batch[i]));
    }
    batch.resize(keep);
    n = keep;
  }

  // Staging layout: descriptors | segments | traceback requests | reads.
  const int64_t plan_bytes = gdp_plan_bytes(n);
  const int64_t off_segs = align_up(plan_bytes, 16);
  const int64_t off_tb = off_segs + align_up((int64_t)n * (int64_t)sizeof(gdp_segment), 16);
  const int64_t off_q = off_tb + (int64_t)n * (int64_t)sizeof(gdp_tb_request);
  gdp_segment *segs = (gdp_segment *)(s.h_up + off_segs);
  gdp_tb_request *tbs = (gdp_tb_request *)(s.h_up + off_tb);
  char *qh = (char *)(s.h_up + off_q);
  s.pair_off.resize(n);
  int64_t qpos = 0, gpos = 0, ppos = 0;
  for (int i = 0; i < n; i++) {
    const Req &q = batch[i];
    gdp_segment &sg = segs[i];
    memset(&sg, 0, sizeof sg);
    if (q.r.watsonp) {
      sg.coord = q.r.chroffset + (uint32_t)q.r.goffset;
      sg.bound = q.r.chrhigh;
    } else {
      sg.coord = q.r.chrhigh - (uint32_t)q.r.goffset + 1u;
      sg.bound = q.r.chroffset;
      sg.left = 1;
      sg.revcomp = 1;
    }
    sg.length = q.r.gl;
    sg.out_off = gpos;
    memcpy(qh + qpos, ring.at(q.roff), (size_t)q.r.rl);
    gdp_tb_request &t = tbs[i];
    memset(&t, 0, sizeof t);
    t.fill = i;
    t.r = q.r.rl;
    t.c = q.r.gl;
    t.queryoffset = q.r.roffset;
    t.genomeoffset = q.r.goffset;
    t.chroffset = q.r.chroffset;
    t.chrhigh = q.r.chrhigh;
    t.watsonp = q.r.watsonp ? 1 : 0;
    t.genestrand = q.r.genestrand;
    t.pair_off = ppos;
    t.max_pairs = q.r.rl + q.r.gl + 2;
    s.pair_off[i] = ppos;
    qpos += q.r.rl;
    gpos += q.r.gl;
    ppos += t.max_pairs;
  }
  memset(qh + qpos, 0, 16);
  {
    // the launched reads are the oldest; their ring space is free again
    std::lock_guard<std::mutex> lk(mu);
    int64_t oldest = ring.next;
    for (const auto &pq : pending)
      if (!pq.empty()) oldest = std::min(oldest, pq.front().roff);
    ring.freed = oldest;
  }
  cv_space.notify_all();
  const int64_t up_bytes = off_q + qpos + 16;
  const uint32_t kernels = gdp_plan_kernels(s.h_up, n);

  s.dir_bytes = align_up(dw * 4, 256);
  s.dir_off = arena.alloc(s.dir_bytes);
  s.off_counts = align_up((int64_t)n * 4, 16);
  s.off_pairs = align_up(s.off_counts + (int64_t)n * 24, 16);
  s.down_bytes = s.off_pairs + ppos * (int64_t)sizeof(gdp_pair);

  {
    std::lock_guard<std::mutex> lk(dmu);
    s.buf = rfree.back();
    rfree.pop_back();
  }
  s.reqs = std::move(batch);
  s.busy = true;
  uint8_t *d_descs = s.d_up;
  uint32_t *dirs = (uint32_t *)(d_dirs + s.dir_off);
  s.t_launch = now_s();
  hipError_t e = hipEventRecord(s.ev0, s.stream);
  // transfers run as kernels on the slot's stream (gdp_launch_copy)
  int rc = e == hipSuccess ? GDP_OK : GDP_EHIP;
  if (rc == GDP_OK && gdp_launch_copy(s.d_up, s.h_up_dev, align_up(up_bytes, 16), s.stream) != 0) rc = GDP_EHIP;
  if (rc == GDP_OK)
    rc = gdp_genome_segments(genome, fmt, glen, (const gdp_segment *)(s.d_up + off_segs), n, s.d_g, s.stream);
  if (rc == GDP_OK)
    rc = gdp_fill_ex(d_descs, n, kernels, (const char *)(s.d_up + off_q), s.d_g, dirs, nullptr,
                     (int32_t *)s.d_down, s.stream);
  if (rc == GDP_OK)
    rc = gdp_traceback_ex(d_descs, (const gdp_tb_request *)(s.d_up + off_tb), n, (const char *)(s.d_up + off_q),
                          s.d_g, dirs, genome, fmt, glen, (gdp_pair *)(s.d_down + s.off_pairs),
                          (int32_t *)(s.d_down + s.off_counts), s.stream);
  if (rc == GDP_OK &&
      gdp_launch_serve_download(s.d_down, rbuf_dev[s.buf], s.off_pairs, s.off_counts, s.off_pairs,
                                (const gdp_tb_request *)(s.d_up + off_tb), n, s.stream) != 0)
    rc = GDP_EHIP;
  if (rc == GDP_OK) {
    e = hipEventRecord(s.ev, s.stream);
    rc = e == hipSuccess ? GDP_OK : GDP_EHIP;
  }
  if (rc != GDP_OK) {
    fail_all(std::string("gdp_server launch: ") + hipGetErrorString(e != hipSuccess ? e : hipGetLastError()));
    return true;
  }
  {
    std::lock_guard<std::mutex> lk(smu);
    st.batches++;
    dir_in_flight += s.dir_bytes;
    st.dir_bytes_peak = std::max(st.dir_bytes_peak, dir_in_flight);
    int inflight = 0;
    for (const Slot &o : slots) inflight += o.busy ? 1 : 0;
    st.inflight_peak = std::max(st.inflight_peak, inflight);
  }
  return true;
}

void gdp_server::retire(Slot &s) {
  const double t = now_s();
  const int n = (int)s.reqs.size();
  const uint8_t *h = rbuf[s.buf];
  const int32_t *fin = (const int32_t *)h;
  const int32_t *cnt = (const int32_t *)(h + s.off_counts);
  const gdp_pair *pr = (const gdp_pair *)(h + s.off_pairs);
  DoneBatch b;
  b.buf = s.buf;
  b.res.resize(n);
  for (int i = 0; i < n; i++) {
    const Req &q = s.reqs[i];
    gdp_align_result &r = b.res[i];
    memset(&r, 0, sizeof r);
    r.tag = q.r.tag;
    r.status = GDP_OK;
    r.finalscore = fin[i];
    for (int k = 0; k < 6; k++) r.counts[k] = cnt[6 * i + k];
    r.width = q.width;
    r.pairs = cnt[6 * i] > 0 ? pr + s.pair_off[i] : nullptr;
    r.t_submit = q.t_submit;
    r.t_done = t;
  }
  arena.release(s.dir_off, s.dir_bytes);
  float dev_ms = 0;
  if (hipEventElapsedTime(&dev_ms, s.ev0, s.ev) != hipSuccess) dev_ms = 0;
  {
    std::lock_guard<std::mutex> lk(smu);
    st.completed += n;
    dir_in_flight -= s.dir_bytes;
    st.batch_seconds += t - s.t_launch;
    st.device_seconds += dev_ms * 1e-3;
  }
  s.reqs.clear();
  s.buf = -1;
  s.busy = false;
  post(std::move(b));
}

void gdp_server::run() {
  (void)hipSetDevice(device);
  for (;;) {
    bool progressed = false;
    for (Slot &s : slots) {
      if (!s.busy) continue;
      const hipError_t e = hipEventQuery(s.ev);
      if (e == hipSuccess) {
        const double t0 = now_s();
        retire(s);
        std::lock_guard<std::mutex> lk(smu);
        st.retire_seconds += now_s() - t0;
        progressed = true;
      } else if (e != hipErrorNotReady) {
        fail_all(std::string("gdp_server batch: ") + hipGetErrorString(e));
        progressed = true;
      }
    }
    bool any_busy = false;
    for (Slot &s : slots) {
      if (!s.busy && !failed.load() && !stop.load()) {
        const double t0 = now_s();
        if (launch(s)) {
          progressed = true;
          std::lock_guard<std::mutex> lk(smu);
          st.prep_seconds += now_s() - t0;
        }
      }
      any_busy = any_busy || s.busy;
    }
    if (stop.load() && !any_busy) break;
    if (progressed) continue;
    if (any_busy) {
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    } else {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait_for(lk, std::chrono::milliseconds(5),
                       [&] { return !pending[0].empty() || !pending[1].empty() || stop.load(); });
    }
  }
}

extern "C" double gdp_server_now(void) { return now_s(); }

extern "C" gdp_server *gdp_server_create(const gdp_server_config *cfg, const void *d_genome, int32_t genome_format,
                                         uint32_t genome_len) {
  if (!cfg || cfg->max_batch <= 0 || cfg->slots < 1 || cfg->slots > 16 || cfg->max_rl <= 0 ||
      cfg->max_rl > GDP_MAX_LEN || cfg->max_gl <= 0 || cfg->max_gl > GDP_MAX_LEN || cfg->dir_budget_bytes <= 0 ||
      cfg->pair_capacity < (int64_t)cfg->max_rl + cfg->max_gl + 2) {
    set_err("gdp_server_create: bad configuration");
    return nullptr;
  }
  // lane groups: all zero for gdp_plan's defaults, else what gdp_plan_tuned accepts
  if ((cfg->plan.g16_max_span != 0 || cfg->plan.g64_min_rows != 0 || cfg->plan.mix_span != 0 ||
       cfg->plan.mix_order != 0) &&
      (cfg->plan.g16_max_span < 0 || cfg->plan.g16_max_span > 240 || cfg->plan.g64_min_rows < 1 ||
       cfg->plan.mix_span < 0 || cfg->plan.mix_span > 64 || cfg->plan.mix_order < 0 || cfg->plan.mix_order > 1)) {
    set_err("gdp_server_create: plan.g16_max_span must lie in [0, 240], plan.g64_min_rows be positive, "
            "plan.mix_span lie in [0, 64] and plan.mix_order be 0 or 1");
    return nullptr;
  }
  if (genome_format != GDP_GENOME_ASCII && genome_format != GDP_GENOME_COMP) {
    set_err("gdp_server_create: unknown genome format %d", genome_format);
    return nullptr;
  }
  gdp_server *s = new gdp_server;
  s->cfg = *cfg;
  s->genome = d_genome;
  s->fmt = genome_format;
  s->glen = genome_len;
  (void)hipGetDevice(&s->device);
  const int64_t nb = cfg->max_batch;
  s->q_cap = nb * (int64_t)cfg->max_rl + 16;
  s->g_cap = nb * (int64_t)cfg->max_gl + 16;
  s->up_cap = align_up(gdp_plan_bytes(cfg->max_batch), 16) + align_up(nb * (int64_t)sizeof(gdp_segment), 16) +
              nb * (int64_t)sizeof(gdp_tb_request) + s->q_cap;
  s->down_cap = align_up(nb * 4, 16) + align_up(nb * 24, 16) + cfg->pair_capacity * (int64_t)sizeof(gdp_pair);
  s->arena.init(align_up(cfg->dir_budget_bytes, 256) - 256);
  s->ring.buf.resize((size_t)std::max<int64_t>(64 << 20, 4 * nb * (int64_t)cfg->max_rl));
  hipError_t e = hipMalloc((void **)&s->d_dirs, (size_t)align_up(cfg->dir_budget_bytes, 256) + 64);
  s->slots.resize(cfg->slots);
  for (Slot &sl : s->slots) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&sl.ev);
    if (e == hipSuccess) e = hipEventCreate(&sl.ev0);
    if (e == hipSuccess) e = hipHostMalloc((void **)&sl.h_up, (size_t)s->up_cap + 16, hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&sl.h_up_dev, sl.h_up, 0);
    if (e == hipSuccess) e = hipMalloc((void **)&sl.d_up, (size_t)s->up_cap);
    if (e == hipSuccess) e = hipMalloc((void **)&sl.d_down, (size_t)s->down_cap);
    if (e == hipSuccess) e = hipMalloc((void **)&sl.d_g, (size_t)s->g_cap);
    if (e == hipSuccess) e = hipMemset(sl.d_g, 0, (size_t)s->g_cap);
  }
  const int nbuf = cfg->result_buffers > 0 ? std::max(cfg->result_buffers, cfg->slots) : cfg->slots + 2;
  for (int k = 0; k < nbuf && e == hipSuccess; k++) {
    uint8_t *h = nullptr, *hd = nullptr;
    e = hipHostMalloc((void **)&h, (size_t)s->down_cap, hipHostMallocMapped);
    if (e == hipSuccess) {
      s->rbuf.push_back(h);
      e = hipHostGetDevicePointer((void **)&hd, h, 0);
    }
    if (e == hipSuccess) {
      s->rbuf_dev.push_back(hd);
      s->rfree.push_back(k);
    }
  }
  if (e != hipSuccess) {
    set_err("gdp_server_create: %s", hipGetErrorString(e));
    s->stop = true;
    gdp_server_destroy(s);
    return nullptr;
  }
  s->th = std::thread([s] { s->run(); });
  return s;
}

extern "C" int gdp_server_submit(gdp_server *s, const gdp_align_request *reqs, int32_t n) {
  if (!s || n < 0 || (n > 0 && !reqs)) {
    set_err("gdp_server_submit: bad argument");
    return GDP_EINVAL;
  }
  const double t = now_s();
  DoneBatch b;
  int32_t accepted = 0;
  std::unique_lock<std::mutex> lk(s->mu);
  for (int32_t i = 0; i < n; i++) {
    const gdp_align_request &r = reqs[i];
    Req q;
    q.r = r;
    q.r.read = nullptr;
    q.t_submit = t;
    q.status = 0;
    q.roff = 0;
    int32_t lb = 0, ub = 0;
    if (s->failed.load()) {
      q.status = (int8_t)-GDP_EHIP;  // the server stopped on an error
    } else if (r.rl <= 0 || r.gl <= 0 || r.rl > s->cfg.max_rl || r.gl > s->cfg.max_gl) {
      q.status = (int8_t)-GDP_ETOOBIG;
    } else {
      gdp_compute_bands(&lb, &ub, r.rl, r.gl, r.extraband, r.widebandp);
      if (lb < 0 || ub < 0 || lb + ub > GDP_MAX_SPAN || r.mismatchtype < 0 || r.mismatchtype > 3 || r.open > 0 ||
          r.extend > 0 || r.read == nullptr)
        q.status = (int8_t)-GDP_EINVAL;
    }
    if (q.status) {
      b.res.push_back(refused(q, t));
      continue;
    }
    q.lband = (int16_t)lb;
    q.uband = (int16_t)ub;
    // Dynprog_single_gap's kernel choice (dynprog_single.c:595).
    const int u8 = gdp_use8p_size(r.mismatchtype);
    q.width = (r.rl < u8 && r.gl < u8) ? 8 : 16;
    int64_t v = s->ring.place(r.rl);
    while (!s->ring.fits(v, r.rl) && !s->failed.load()) {  // the ring is full of waiting reads
      s->cv_work.notify_one();
      s->cv_space.wait_for(lk, std::chrono::milliseconds(1));
      v = s->ring.place(r.rl);
    }
    if (s->failed.load()) {  // failed while waiting for ring space
      q.status = (int8_t)-GDP_EHIP;
      b.res.push_back(refused(q, t));
      continue;
    }
    memcpy(s->ring.at(v), r.read, (size_t)r.rl);
    s->ring.next = v + r.rl;
    q.roff = v;
    s->pending[s->cfg.long_rows > 0 && r.rl >= s->cfg.long_rows ? 1 : 0].push_back(q);
    accepted++;
  }
  lk.unlock();
  {
    std::lock_guard<std::mutex> lk2(s->smu);
    s->st.submitted += n;
    s->st.completed += (int64_t)b.res.size();
  }
  if (!b.res.empty()) s->post(std::move(b));
  if (accepted) s->cv_work.notify_one();
  return s->failed.load() ? GDP_EHIP : GDP_OK;
}

extern "C" int32_t gdp_server_poll(gdp_server *s, gdp_align_result *out, int32_t max, int32_t wait_us) {
  if (!s || max < 0 || (max > 0 && !out)) return GDP_EINVAL;
  std::vector<DoneBatch> took;
  {
    std::unique_lock<std::mutex> lk(s->dmu);
    // results of the previous call are no longer referenced
    for (int k : s->handed) s->rfree.push_back(k);
    s->handed.clear();
    if (s->done.empty() && wait_us > 0)
      s->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !s->done.empty(); });
    int64_t have = 0;
    while (have < max && !s->done.empty()) {
      have += (int64_t)(s->done.front().res.size() - s->done.front().next);
      took.push_back(std::move(s->done.front()));
      s->done.pop_front();
    }
  }
  // copy outside the lock, so the server thread can post meanwhile
  int32_t k = 0;
  for (DoneBatch &b : took) {
    const size_t m = std::min<size_t>((size_t)(max - k), b.res.size() - b.next);
    if (m) memcpy(out + k, b.res.data() + b.next, m * sizeof(gdp_align_result));
    k += (int32_t)m;
    b.next += m;
  }
  std::lock_guard<std::mutex> lk(s->dmu);
  for (size_t i = 0; i < took.size(); i++) {
    DoneBatch &b = took[i];
    if (b.next < b.res.size()) {
      // only the last batch taken can be left in part: it goes back first
      s->done.push_front(std::move(b));
    } else if (b.buf >= 0) {
      s->handed.push_back(b.buf);
    }
  }
  return k;
}

extern "C" int gdp_server_get_stats(gdp_server *s, gdp_server_stats *st) {
  if (!s || !st) return GDP_EINVAL;
  std::lock_guard<std::mutex> lk(s->smu);
  *st = s->st;
  if (!s->fatal.empty()) {
    gdp_set_error(s->fatal.c_str());
    return GDP_EHIP;
  }
  return GDP_OK;
}

extern "C" void gdp_server_destroy(gdp_server *s) {
  if (!s) return;
  s->stop = true;
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  for (Slot &sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.h_up) (void)hipHostFree(sl.h_up);
    if (sl.d_up) (void)hipFree(sl.d_up);
    if (sl.d_down) (void)hipFree(sl.d_down);
    if (sl.d_g) (void)hipFree(sl.d_g);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.ev0) (void)hipEventDestroy(sl.ev0);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  for (uint8_t *h : s->rbuf) (void)hipHostFree(h);
  if (s->d_dirs) (void)hipFree(s->d_dirs);
  delete s;
}
