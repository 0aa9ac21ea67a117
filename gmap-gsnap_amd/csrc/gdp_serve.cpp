// gdp_serve.cpp -- continuous batching of alignment requests over a device
// genome (the gdp_server_* functions of include/gmapdp.h).
//
// The server keeps up to `slots` batches in flight, each on a HIP stream of
// its own with its own pinned staging buffers and device buffers, and each
// driven by a thread of its own; the direction words of all batches in
// flight share one device budget handed out first-fit.  Whenever its slot is
// free and requests wait, a slot's thread takes the oldest requests that fit
// the slot and the free budget, plans them, and queues on the slot's stream:
// one upload, the genome-segment extraction, the fills, the tracebacks, one
// download and an event.  It turns the finished batch into results when the
// event fires.  Slots plan in parallel, so new batches are formed while
// earlier ones still run, and the host's planning does not hold back the
// device.
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
  std::thread th;  // the slot's thread: forms, launches and retires its batches
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr, ev0 = nullptr;  // end and start of the batch on the stream
  double t_launch = 0;
  uint8_t *h_up = nullptr, *d_up = nullptr, *d_down = nullptr;
  uint8_t *h_up_dev = nullptr;  // h_up as the device addresses it
  char *d_g = nullptr;
  int buf = -1;  // result buffer of the batch in flight
  bool busy = false;
  std::vector<Req> reqs;
  std::vector<gdp_problem> probs;
  std::vector<int64_t> pair_off;
  // Ring offset of the oldest read the slot still has to copy out of the
  // read ring (INT64_MAX: none); under gdp_server::mu.
  int64_t copy_lo = INT64_MAX;
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
  std::mutex amu;  // guards arena
  Arena arena;
  int64_t arena_size = 0;
  std::vector<Slot> slots;
  std::atomic<int> nbusy{0};  // slots with a batch in flight
  int64_t up_cap = 0, g_cap = 0, down_cap = 0, q_cap = 0;

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
  // Set once a HIP call of a slot's thread failed: from then on no batch
  // starts, pending and newly submitted requests are refused with GDP_EHIP.
  // Batches already in flight on other slots still complete.
  std::atomic<bool> failed{false};
  std::string fatal;  // what failed (under smu)

  enum { LAUNCHED, IDLE, BLOCKED };
  void run_slot(Slot &s);
  int launch(Slot &s);
  void retire(Slot &s);
  void fail_all(Slot *own, const std::string &what);
  void update_freed();
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

// Read-ring bytes below the oldest read that a pending request or a slot
// still copying its batch holds may be overwritten.  Called under mu.
void gdp_server::update_freed() {
  int64_t oldest = ring.next;
  for (const auto &pq : pending)
    if (!pq.empty()) oldest = std::min(oldest, pq.front().roff);
  for (const Slot &o : slots) oldest = std::min(oldest, o.copy_lo);
  ring.freed = oldest;
}

// Stops the server after a failure: refuses the calling slot's batch (own,
// may be null) and every pending request.  Other slots' batches in flight
// complete on their own threads.
void gdp_server::fail_all(Slot *own, const std::string &what) {
  {
    std::lock_guard<std::mutex> lk(smu);
    if (fatal.empty()) fatal = what;
  }
  failed = true;
  DoneBatch b;
  const double t = now_s();
  if (own && own->busy) {
    for (Req &q : own->reqs) {
      q.status = (int8_t)-GDP_EHIP;
      b.res.push_back(refused(q, t));
    }
    own->reqs.clear();
    own->busy = false;
    nbusy--;
  }
  if (own && own->buf >= 0) {
    std::lock_guard<std::mutex> lk(dmu);
    rfree.push_back(own->buf);
    own->buf = -1;
  }
  std::deque<Req> rest;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &pq : pending)
      for (Req &q : pq) rest.push_back(q);
    pending[0].clear();
    pending[1].clear();
    if (own) own->copy_lo = INT64_MAX;
    // no read waits in the ring any more: a submitter blocked on ring space
    // wakes up and is refused
    ring.freed = ring.next;
  }
  cv_space.notify_all();
  cv_work.notify_all();
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
// budget, and queues their work on the slot's stream.  Returns LAUNCHED when
// it started a batch (or refused requests), IDLE when no request waits, and
// BLOCKED when requests wait for a result buffer or for direction budget.
int gdp_server::launch(Slot &s) {
  {
    std::lock_guard<std::mutex> lk(dmu);
    if (rfree.empty()) return BLOCKED;  // the consumer holds every result buffer
    s.buf = rfree.back();
    rfree.pop_back();
  }
  auto give_back_buffer = [&] {
    std::lock_guard<std::mutex> lk(dmu);
    rfree.push_back(s.buf);
    s.buf = -1;
  };
  std::vector<Req> batch;
  int cls = 0;
  bool more = false;
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
    // the reads of the batch stay in the ring until they are copied out
    if (!batch.empty()) s.copy_lo = batch.front().roff;
    more = !pending[0].empty() || !pending[1].empty();
  }
  if (batch.empty()) {
    give_back_buffer();
    return IDLE;
  }
  if (more) cv_work.notify_one();  // another free slot can form the next batch meanwhile
  auto release_reads = [&] {
    {
      std::lock_guard<std::mutex> lk(mu);
      s.copy_lo = INT64_MAX;
      update_freed();
    }
    cv_space.notify_all();
  };

  // Plan; if the direction words do not fit the free budget, give back the
  // younger half and plan again.
  int n = (int)batch.size();
  int64_t dw = 0, ss = 0;
  for (;;) {
    s.probs.resize(n);
    int64_t qpos = 0, gpos = 0;
    for (int i = 0; i < n; i++) {
      const Req &q = batch[i];
      gdp_problem &p = s.probs[i];
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
    if (gdp_plan_tuned(s.probs.data(), n, 0, pp, s.h_up, &dw, &ss) != GDP_OK) {
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
      release_reads();
      fail_all(&s, why);
      return LAUNCHED;
    }
    const int64_t need = align_up(dw * 4, 256);
    int64_t off = -1;
    bool empty_arena = false;
    {
      std::lock_guard<std::mutex> lk(amu);
      off = arena.alloc(need);
      if (off < 0) empty_arena = arena.largest() == arena_size;
    }
    if (off >= 0) {
      s.dir_off = off;
      s.dir_bytes = need;
      break;
    }
    if (n == 1) {
      if (!empty_arena) {  // wait for budget
        {
          std::lock_guard<std::mutex> lk(mu);
          pending[cls].push_front(std::move(batch[0]));
          s.copy_lo = INT64_MAX;
          update_freed();
        }
        give_back_buffer();
        return BLOCKED;
      }
      DoneBatch b;  // one request larger than the whole budget
      batch[0].status = (int8_t)-GDP_ENOMEM;
      b.res.push_back(refused(batch[0], now_s()));
      {
        std::lock_guard<std::mutex> lk(smu);
        st.completed += 1;
      }
      post(std::move(b));
      release_reads();
      give_back_buffer();
      return LAUNCHED;
    }
    const int keep = n / 2;
    {
      std::lock_guard<std::mutex> lk(mu);
      for (int i = n - 1; i >= keep; i--) pending[cls].push_front(std::move(batch[i]));
    }
    cv_work.notify_one();
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
  release_reads();
  const int64_t up_bytes = off_q + qpos + 16;
  const uint32_t kernels = gdp_plan_kernels(s.h_up, n);

  s.off_counts = align_up((int64_t)n * 4, 16);
  s.off_pairs = align_up(s.off_counts + (int64_t)n * 24, 16);
  s.down_bytes = s.off_pairs + ppos * (int64_t)sizeof(gdp_pair);
  s.reqs = std::move(batch);
  s.busy = true;
  const int inflight = ++nbusy;
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
    fail_all(&s, std::string("gdp_server launch: ") + hipGetErrorString(e != hipSuccess ? e : hipGetLastError()));
    return LAUNCHED;
  }
  {
    std::lock_guard<std::mutex> lk(smu);
    st.batches++;
    dir_in_flight += s.dir_bytes;
    st.dir_bytes_peak = std::max(st.dir_bytes_peak, dir_in_flight);
    st.inflight_peak = std::max(st.inflight_peak, inflight);
  }
  return LAUNCHED;
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
  {
    std::lock_guard<std::mutex> lk(amu);
    arena.release(s.dir_off, s.dir_bytes);
  }
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
  nbusy--;
  post(std::move(b));
  cv_work.notify_all();  // budget came free: a slot waiting for it may go on
}

void gdp_server::run_slot(Slot &s) {
  (void)hipSetDevice(device);
  for (;;) {
    if (s.busy) {
      const hipError_t e = hipEventQuery(s.ev);
      if (e == hipSuccess) {
        const double t0 = now_s();
        retire(s);
        std::lock_guard<std::mutex> lk(smu);
        st.retire_seconds += now_s() - t0;
      } else if (e != hipErrorNotReady) {
        fail_all(&s, std::string("gdp_server batch: ") + hipGetErrorString(e));
      } else {
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
      continue;
    }
    if (stop.load()) break;
    int r = IDLE;
    if (!failed.load()) {
      const double t0 = now_s();
      r = launch(s);
      if (r == LAUNCHED) {
        std::lock_guard<std::mutex> lk(smu);
        st.prep_seconds += now_s() - t0;
        continue;
      }
    }
    std::unique_lock<std::mutex> lk(mu);
    if (r == BLOCKED)  // woken when a batch retires or the consumer returns buffers
      cv_work.wait_for(lk, std::chrono::microseconds(200));
    else
      cv_work.wait_for(lk, std::chrono::milliseconds(5), [&] {
        return stop.load() || (!failed.load() && (!pending[0].empty() || !pending[1].empty()));
      });
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
  s->arena_size = align_up(cfg->dir_budget_bytes, 256) - 256;
  s->arena.init(s->arena_size);
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
  for (Slot &sl : s->slots) {
    Slot *p = &sl;
    sl.th = std::thread([s, p] { s->run_slot(*p); });
  }
  return s;
}

extern "C" int gdp_server_submit(gdp_server *s, const gdp_align_request *reqs, int32_t n) {
  if (!s || n < 0 || (n > 0 && !reqs)) {
    set_err("gdp_server_submit: bad argument");
    return GDP_EINVAL;
  }
  const double t = now_s();
  DoneBatch b;
  // Requests are checked and banded outside the lock, then their reads are
  // placed in the ring a chunk at a time, so the slots' threads can take
  // requests while a large submission is copied in.
  constexpr int32_t CHUNK = 1024;
  std::vector<Req> qs;
  std::vector<const char *> rd;
  qs.reserve((size_t)std::min(n, CHUNK));
  rd.reserve((size_t)std::min(n, CHUNK));
  for (int32_t i0 = 0; i0 < n; i0 += CHUNK) {
    const int32_t i1 = std::min(n, i0 + CHUNK);
    qs.clear();
    rd.clear();
    for (int32_t i = i0; i < i1; i++) {
      const gdp_align_request &r = reqs[i];
      Req q;
      q.r = r;
      q.r.read = nullptr;
      q.t_submit = t;
      q.status = 0;
      q.roff = 0;
      int32_t lb = 0, ub = 0;
      if (r.scoring == 1) {
        // Dynprog_single_gap's choice by defect rate (dynprog_single.c:458-481;
        // DEFECT_HIGHQ 0.003, DEFECT_MEDQ 0.014, dynprog.h:58-59;
        // SINGLE_OPEN/EXTEND_{HIGHQ,MEDQ,LOWQ}, dynprog.h:61-67)
        const bool hq = r.defect_rate < 0.003, mq = r.defect_rate < 0.014;
        q.r.mismatchtype = (int8_t)(hq ? GDP_HIGHQ : (mq ? GDP_MEDQ : GDP_LOWQ));
        q.r.open = (int8_t)(hq ? -8 : (mq ? -7 : -6));
        q.r.extend = (int8_t)(hq ? -3 : (mq ? -2 : -1));
      }
      if (r.scoring < 0 || r.scoring > 1) {
        q.status = (int8_t)-GDP_EINVAL;
      } else if (r.rl <= 0 || r.gl <= 0 || r.rl > s->cfg.max_rl || r.gl > s->cfg.max_gl) {
        q.status = (int8_t)-GDP_ETOOBIG;
      } else {
        gdp_compute_bands(&lb, &ub, r.rl, r.gl, r.extraband, r.widebandp);
        if (lb < 0 || ub < 0 || lb + ub > GDP_MAX_SPAN || q.r.mismatchtype < 0 || q.r.mismatchtype > 3 ||
            q.r.open > 0 || q.r.extend > 0 || r.read == nullptr)
          q.status = (int8_t)-GDP_EINVAL;
      }
      if (q.status) {
        b.res.push_back(refused(q, t));
        continue;
      }
      q.lband = (int16_t)lb;
      q.uband = (int16_t)ub;
      // Dynprog_single_gap's kernel choice (dynprog_single.c:595).
      const int u8 = gdp_use8p_size(q.r.mismatchtype);
      q.width = (r.rl < u8 && r.gl < u8) ? 8 : 16;
      qs.push_back(q);
      rd.push_back(r.read);
    }
    int32_t placed = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (size_t j = 0; j < qs.size(); j++) {
        Req &q = qs[j];
        if (!s->failed.load()) {
          int64_t v = s->ring.place(q.r.rl);
          while (!s->ring.fits(v, q.r.rl) && !s->failed.load()) {  // the ring is full of waiting reads
            s->cv_work.notify_all();
            s->cv_space.wait_for(lk, std::chrono::milliseconds(1));
            v = s->ring.place(q.r.rl);
          }
          if (!s->failed.load()) {
            memcpy(s->ring.at(v), rd[j], (size_t)q.r.rl);
            s->ring.next = v + q.r.rl;
            q.roff = v;
            s->pending[s->cfg.long_rows > 0 && q.r.rl >= s->cfg.long_rows ? 1 : 0].push_back(q);
            placed++;
            continue;
          }
        }
        q.status = (int8_t)-GDP_EHIP;  // the server stopped on an error
        b.res.push_back(refused(q, t));
      }
    }
    if (placed) s->cv_work.notify_one();
  }
  {
    std::lock_guard<std::mutex> lk2(s->smu);
    s->st.submitted += n;
    s->st.completed += (int64_t)b.res.size();
  }
  if (!b.res.empty()) s->post(std::move(b));
  return s->failed.load() ? GDP_EHIP : GDP_OK;
}

extern "C" int32_t gdp_server_poll(gdp_server *s, gdp_align_result *out, int32_t max, int32_t wait_us) {
  if (!s || max < 0 || (max > 0 && !out)) return GDP_EINVAL;
  std::vector<DoneBatch> took;
  {
    std::unique_lock<std::mutex> lk(s->dmu);
    // results of the previous call are no longer referenced
    if (!s->handed.empty()) {
      for (int k : s->handed) s->rfree.push_back(k);
      s->handed.clear();
      s->cv_work.notify_all();  // a slot waiting for a result buffer may go on
    }
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
  for (Slot &sl : s->slots)
    if (sl.th.joinable()) sl.th.join();
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
