// gdp_gen.cpp -- synthetic workloads of the benchmark configurations.
//
// Every problem is generated from its own counter-based random stream
// (splitmix64 of seed and index), so any slice [first, first+count) of a
// workload is reproducible on its own and generation parallelises.
//
//   1  single-gap, 100-nt reads from a 2%-substituted, 0.2%-indel copy of a
//      genomic span, extraband 10 (wide band), int16 HIGHQ or int8 LOWQ
//   2  single-gap, rl 150, gl U{160..200}, extraband U{5..20}, int16 HIGHQ
//   3  single-gap, rl log-uniform [20, 1000], gl = rl * U(1, 1.5),
//      extraband U{3..64}, 10% of reads at 30% divergence, width by the
//      reference's length rule (int8 when rl and gl < 41)
//   4  end extensions (5' reversed / 3' forward halves), rl U{10..150},
//      gl = rl + U{0..20}, 3% subs, 1% indels of 1-5, 1% N, ENDQ,
//      extraband_end 6, int8 when rl or gl < 24; each yields an upper and a
//      lower fill
//   5  serving: a random genome packed as genomecomp (with runs of N), and
//      alignment requests with the length and divergence mix of 3 against
//      segments of it on either strand, arriving as a Poisson process
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/gmapdp.h"
#include "gdp_tables.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  int range(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }  // inclusive
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

// use8p_size (dynprog.c:1021-1024): -128 / mismatch - 1 = 41, 63, 127, 24.
int use8p(int mismatchtype) {
  static const int mm[4] = {-3, -2, -1, -5};
  return -128 / mm[mismatchtype & 3] - 1;
}

uint64_t stream_seed(uint64_t seed, int64_t index) {
  Rng r(seed ^ (0xd1b54a32d192ed03ull * (uint64_t)(index + 1)));
  return r.next();
}

// Copies src with substitutions, indels and N's; stops at maxlen output bytes.
int mutate(Rng &rng, const char *src, int srclen, char *out, int maxlen, double sub, double indel, int maxind,
           double nfrac, int *consumed) {
  int i = 0, o = 0;
  while (i < srclen && o < maxlen) {
    double u = rng.uni();
    if (u < indel * 0.5) {  // insertion
      int k = rng.range(1, maxind);
      for (int j = 0; j < k && o < maxlen; j++) out[o++] = kBases[rng.next() & 3];
      continue;
    }
    if (u < indel) {  // deletion
      i += rng.range(1, maxind);
      continue;
    }
    char ch = src[i++];
    if (rng.uni() < sub) ch = kBases[rng.next() & 3];
    if (nfrac > 0 && rng.uni() < nfrac) ch = 'N';
    out[o++] = ch;
  }
  if (consumed) *consumed = i;
  return o;
}

}  // namespace

extern "C" {

// Genome of iid uniform bases (deterministic in seed), ASCII.
void gdp_gen_genome(uint64_t seed, char *out, int64_t length) {
  const int nthreads = 16;
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([=]() {
      int64_t lo = length * t / nthreads, hi = length * (t + 1) / nthreads;
      for (int64_t b = lo & ~(int64_t)4095; b < hi; b += 4096) {
        Rng r(stream_seed(seed ^ 0x6e6f6d65ull, b >> 12));
        for (int64_t k = b; k < b + 4096 && k < hi; k++) {
          uint64_t v = r.next();
          if (k >= lo) out[k] = kBases[v & 3];
        }
      }
    });
  for (auto &t : th) t.join();
}

// Per-problem metadata the traceback needs (queryoffset etc.).
typedef struct gdp_gen_meta {
  int32_t queryoffset;
  int32_t genomeoffset;
  uint32_t chroffset;
  uint32_t chrhigh;
  int8_t watsonp;
  int8_t revp;
  int8_t pad[2];
  int32_t pair;  // config 4: index of the partner fill (lower for an upper, -1 otherwise)
} gdp_gen_meta;

// Upper bound of query and private genome bytes one problem of a config uses.
int32_t gdp_gen_max_bytes(int config) {
  switch (config) {
  case 1: return 100 + 200;
  case 2: return 150 + 8;
  case 3: return 1000 + 8;
  case 4: return 2 * (150 + 170);
  default: return 0;
  }
}

// Generates problems [first, first+count) of a config.  Single-gap problems
// read their genome window straight from `genome` (goff = window start); the
// reversed genome windows of 5' end problems are written after the genome in
// gbuf, starting at gbuf_extra (per-problem slots of gdp_gen_max_bytes).
// Query bytes go to qbuf at qoff = (index-first) * gdp_gen_max_bytes.
// Config 4 produces two fills per index (upper at 2k, lower at 2k+1).
// variant selects config 1's int8 LOWQ run (1) instead of int16 HIGHQ (0).
int gdp_gen_workload(int config, int variant, uint64_t seed, int64_t first, int64_t count, const char *genome,
                     int64_t genome_len, gdp_problem *probs, gdp_gen_meta *meta, char *qbuf, char *gbuf,
                     int64_t gbuf_extra) {
  if (config < 1 || config > 4) return -1;
  const int stride = gdp_gen_max_bytes(config);
  const int nthreads = 16;
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([=]() {
      std::vector<char> tmp(4096), src(4096);
      for (int64_t k = count * t / nthreads; k < count * (t + 1) / nthreads; k++) {
        const int64_t idx = first + k;
        Rng rng(stream_seed(seed, idx));
        char *q = qbuf + k * stride;
        gdp_problem p;
        memset(&p, 0, sizeof p);
        gdp_gen_meta m;
        memset(&m, 0, sizeof m);
        m.watsonp = 1;
        m.chroffset = 0;
        m.chrhigh = (uint32_t)genome_len;
        m.pair = -1;
        p.qoff = k * stride;
        if (config == 1) {
          const int64_t pos = (int64_t)(rng.uni() * (double)(genome_len - 400));
          int consumed = 0;
          int rl = mutate(rng, genome + pos, 300, q, 100, 0.02, 0.002, 3, 0.0, &consumed);
          p.rl = rl;
          p.gl = consumed;
          p.goff = p.galtoff = pos;
          int lb, ub;
          gdp_compute_bands(&lb, &ub, p.rl, p.gl, 10, 1);
          p.lband = (int16_t)lb;
          p.uband = (int16_t)ub;
          p.kind = GDP_KIND_FULL;
          if (variant == 0) {
            p.mismatchtype = GDP_HIGHQ; p.open = -8; p.extend = -3;
          } else {
            p.mismatchtype = GDP_LOWQ; p.open = -6; p.extend = -1;
          }
          const int u8 = use8p(p.mismatchtype);
          p.width = (p.rl < u8 && p.gl < u8) ? 8 : 16;
          p.jump_late = (int8_t)(rng.next() & 1);
          m.genomeoffset = (int32_t)pos;
        } else if (config == 2 || config == 3) {
          int rl, gl, eb;
          double div = 0.02;
          if (config == 2) {
            rl = 150;
            gl = rng.range(160, 200);
            eb = rng.range(5, 20);
          } else {
            rl = (int)std::lround(std::exp(std::log(20.0) + rng.uni() * (std::log(1000.0) - std::log(20.0))));
            gl = (int)std::lround(rl * (1.0 + 0.5 * rng.uni()));
            eb = rng.range(3, 64);
            if (rng.uni() < 0.10) div = 0.30;
          }
          const int64_t pos = (int64_t)(rng.uni() * (double)(genome_len - 2 * gl - 8));
          // The read skips gl - rl genome bases at one point (a single gap),
          // then carries substitutions and small indels.
          const int skip = gl - rl, a = rng.range(0, rl);
          memcpy(src.data(), genome + pos, a);
          memcpy(src.data() + a, genome + pos + a + skip, rl - a);
          int got = mutate(rng, src.data(), rl, q, rl, div, 0.002, 3, 0.0, nullptr);
          for (int j = got; j < rl; j++) q[j] = kBases[rng.next() & 3];
          p.rl = rl;
          p.gl = gl;
          p.goff = p.galtoff = pos;
          p.lband = (int16_t)eb;
          p.uband = (int16_t)(gl - rl + eb);
          p.kind = GDP_KIND_FULL;
          p.mismatchtype = GDP_HIGHQ;
          p.open = -8;
          p.extend = -3;
          const int u8 = use8p(GDP_HIGHQ);
          p.width = (config == 3 && rl < u8 && gl < u8) ? 8 : 16;
          p.jump_late = (int8_t)(rng.next() & 1);
          m.genomeoffset = (int32_t)pos;
        } else {
          // End extension: the read end follows the genome from the anchor
          // outward; 5' ends are swept on reversed copies (revp).
          const int rl = rng.range(10, 150), gl = rl + rng.range(0, 20);
          const bool five = (rng.next() & 1) != 0;
          const int64_t pos = 64 + (int64_t)(rng.uni() * (double)(genome_len - 2 * gl - 128));
          char *gx = gbuf + gbuf_extra + k * stride;
          for (int j = 0; j < gl; j++) gx[j] = five ? genome[pos + gl - 1 - j] : genome[pos + j];
          int got = mutate(rng, gx, gl, q, rl, 0.03, 0.01, 5, 0.01, nullptr);
          for (int j = got; j < rl; j++) q[j] = kBases[rng.next() & 3];
          int lb, ub;
          gdp_compute_bands(&lb, &ub, rl, gl, 6, 1);
          const bool u8 = rl < use8p(GDP_ENDQ) || gl < use8p(GDP_ENDQ);
          const bool late = (rng.next() & 1) != 0;
          p.rl = rl;
          p.gl = gl;
          p.goff = p.galtoff = gbuf_extra + k * stride;
          p.mismatchtype = GDP_ENDQ;
          p.open = -10;
          p.extend = -2;
          p.width = u8 ? 8 : 16;
          // 5' callers hand the reversed sequences and the flipped tie rule
          // to the kernels (dynprog_end.c:1415-1451).
          p.jump_late = (int8_t)(five ? !late : late);
          p.keep_scores = 1;
          m.revp = five ? 1 : 0;
          m.genomeoffset = five ? (int32_t)(pos + gl - 1) : (int32_t)pos;
          m.queryoffset = five ? rl - 1 : 0;
          gdp_problem up = p, lo = p;
          up.kind = GDP_KIND_UPPER;
          up.uband = (int16_t)ub;
          up.lband = (int16_t)lb;
          lo.kind = GDP_KIND_LOWER;
          lo.uband = (int16_t)ub;
          lo.lband = (int16_t)lb;
          probs[2 * k] = up;
          probs[2 * k + 1] = lo;
          m.pair = (int32_t)(2 * k + 1);
          meta[2 * k] = m;
          gdp_gen_meta ml = m;
          ml.pair = (int32_t)(2 * k);
          meta[2 * k + 1] = ml;
          continue;
        }
        probs[k] = p;
        meta[k] = m;
      }
    });
  for (auto &t : th) t.join();
  return 0;
}

// Random genome of `nbases` bases written straight into genomecomp words
// (gdp_genomecomp_words(nbases) of them).  Each 32-base block draws its two
// code words from its own stream.  N comes in runs, as assembly gaps do: a
// quarter of the 2-Mbp chunks hold one run of 1-50 kb (flags set, codes 0,
// as the packer writes N).  The tail of the last block is X and the two
// trailing words are 0xFFFFFFFF, as in a packed image.
static bool n_block(uint64_t seed, uint64_t b) {
  Rng r(stream_seed(seed ^ 0x4e72756e73ull, (int64_t)(b >> 16)));
  if (r.next() % 4 != 0) return false;
  const uint64_t len = 32 + r.next() % 1531, start = r.next() % ((1u << 16) - len);
  const uint64_t k = b & 0xFFFF;
  return k >= start && k < start + len;
}

void gdp_gen_genomecomp(uint64_t seed, uint64_t nbases, uint32_t *out) {
  const uint64_t nblocks = (nbases + 31) / 32;
  const int nthreads = 16;
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([=]() {
      for (uint64_t b = nblocks * t / nthreads; b < nblocks * (t + 1) / nthreads; b++) {
        Rng r(stream_seed(seed ^ 0x67656e6f6d65ull, (int64_t)b));
        const uint64_t v = r.next();
        uint32_t high = (uint32_t)(v >> 32), low = (uint32_t)v, flags = 0;
        if (n_block(seed, b)) high = low = 0, flags = 0xFFFFFFFFu;
        for (int k = 0; k < 32; k++) {
          if (b * 32 + (uint64_t)k < nbases) continue;
          if (k < 16) low |= 3u << (2 * k);
          else high |= 3u << (2 * (k - 16));
          flags |= 1u << k;
        }
        out[3 * b] = high;
        out[3 * b + 1] = low;
        out[3 * b + 2] = flags;
      }
    });
  for (auto &t : th) t.join();
  out[3 * nblocks] = 0xFFFFFFFFu;
  out[3 * nblocks + 1] = 0xFFFFFFFFu;
}

}  // extern "C"

namespace {

// uncompress_one_char (genome.c:9723-9760) with flag character 'N'.
char gc_base(const uint32_t *gc, uint64_t nbases, uint64_t pos) {
  if (pos >= nbases) return 'N';
  const uint32_t *b = gc + (pos / 32) * 3;
  const int k = (int)(pos % 32);
  if ((b[2] >> k) & 1u) return 'N';
  const uint32_t w = k < 16 ? b[1] : b[0];
  return kBases[(w >> (2 * (k & 15))) & 3u];
}

char compl_base(char c) {
  switch (c) {
  case 'A': return 'T';
  case 'C': return 'G';
  case 'G': return 'C';
  case 'T': return 'A';
  default: return c;
  }
}

}  // namespace

extern "C" {

// Serving requests [first, first+count) over a genomecomp genome of nbases
// bases cut into chromosomes of 2^27 bases (one chromosome below 2^29
// bases): each request's chroffset and chrhigh are its chromosome's first
// and last base and goffset its window's offset in it.  Lengths,
// band and divergence follow config 3; the strand is drawn per request and
// the read is taken from the segment the server will extract for it
// (the reverse complement on the minus strand).  Read bytes go to
// reads + (index - first) * 1008; out[k].read points there.
int gdp_gen_serve_requests(uint64_t seed, int64_t first, int64_t count, const uint32_t *gc, uint64_t nbases,
                           gdp_align_request *out, char *reads) {
  if (nbases < 4096) return -1;
  // chromosomes of 2^27 bases (134 Mbp, about a human chromosome's mean),
  // or one chromosome for a genome of fewer than four of them
  const uint64_t kChrom = nbases >= (4ull << 27) ? (1ull << 27) : nbases;
  const int stride = 1008;
  const int nthreads = 16;
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([=]() {
      std::vector<char> seg(2048), src(2048);
      for (int64_t k = count * t / nthreads; k < count * (t + 1) / nthreads; k++) {
        const int64_t idx = first + k;
        Rng rng(stream_seed(seed, idx));
        const int rl = (int)std::lround(std::exp(std::log(20.0) + rng.uni() * (std::log(1000.0) - std::log(20.0))));
        const int gl = (int)std::lround(rl * (1.0 + 0.5 * rng.uni()));
        const int eb = rng.range(3, 64);
        const double div = rng.uni() < 0.10 ? 0.30 : 0.02;
        const bool watson = (rng.next() & 1) != 0;
        // The genome is cut into chromosomes of kChrom bases (the last one
        // shorter); a request picks a position uniformly over the genome and
        // aligns within that position's chromosome: chroffset is the
        // chromosome's first base, chrhigh its last, goffset the offset of
        // the window in it (GMAP's convention: an offset within one
        // chromosome, below 2^31 however long the genome).  Reads come from
        // sequenced genome: draw again while the segment touches a run of N
        // (at most 16 draws).
        uint64_t chroffset = 0, chrhigh = 0;
        int64_t goffset = 0;
        for (int tries = 0; tries < 16; tries++) {
          uint64_t chrom = (uint64_t)(rng.uni() * (double)nbases) / kChrom;
          chroffset = chrom * kChrom;
          uint64_t chrlen = std::min<uint64_t>(kChrom, nbases - chroffset);
          if (chrlen < (uint64_t)gl + 64) {  // a stub at the genome's end: take the chromosome before it
            chroffset -= kChrom;
            chrlen = kChrom;
          }
          chrhigh = chroffset + chrlen - 1;
          goffset = 8 + (int64_t)(rng.uni() * (double)(chrlen - (uint64_t)gl - 32));
          const uint64_t lo = watson ? chroffset + (uint64_t)goffset : chrhigh - (uint64_t)goffset + 1 - (uint64_t)gl;
          bool clean = true;
          for (uint64_t blk = lo / 32; blk <= (lo + (uint64_t)gl - 1) / 32 && clean; blk++) clean = gc[3 * blk + 2] == 0;
          if (clean) break;
        }
        if (watson) {
          for (int j = 0; j < gl; j++) seg[j] = gc_base(gc, nbases, chroffset + (uint64_t)goffset + j);
        } else {
          const uint64_t right = chrhigh - (uint64_t)goffset + 1;
          for (int j = 0; j < gl; j++) seg[j] = compl_base(gc_base(gc, nbases, right - 1 - (uint64_t)j));
        }
        const int skip = gl - rl, a = rng.range(0, rl);
        memcpy(src.data(), seg.data(), a);
        memcpy(src.data() + a, seg.data() + a + skip, rl - a);
        char *q = reads + k * stride;
        int got = mutate(rng, src.data(), rl, q, rl, div, 0.002, 3, 0.0, nullptr);
        for (int j = got; j < rl; j++) q[j] = kBases[rng.next() & 3];
        gdp_align_request &r = out[k];
        memset(&r, 0, sizeof r);
        r.tag = (uint64_t)idx;
        r.read = q;
        r.rl = rl;
        r.gl = gl;
        r.roffset = 0;
        r.goffset = (int32_t)goffset;
        r.chroffset = (uint32_t)chroffset;
        r.chrhigh = (uint32_t)chrhigh;
        r.extraband = eb;
        r.watsonp = watson ? 1 : 0;
        r.widebandp = 1;
        r.jump_late = (int8_t)(rng.next() & 1);
        r.genestrand = 0;
        r.mismatchtype = GDP_HIGHQ;
        r.open = -8;
        r.extend = -3;
      }
    });
  for (auto &t : th) t.join();
  return 0;
}

// Open-loop load: request k arrives at t0 + arrival[k], where the gaps are
// exponential with mean 1/rate (rate <= 0: all arrive at t0).  A submitter
// thread hands every request to the server as its arrival time passes; the
// calling thread polls results and records latency[tag - tag0] = completion
// time - arrival time.  Returns the number of results with status GDP_OK and
// stores the time from t0 to the last completion in *elapsed.
int64_t gdp_serve_drive(gdp_server *s, const gdp_align_request *reqs, int64_t n, double rate, uint64_t seed,
                        double *latency, double *elapsed) {
  if (!s || n <= 0 || !reqs || !latency) return -1;
  std::vector<double> arrival(n);
  Rng rng(seed ^ 0x61727269ull);
  double t = 0;
  for (int64_t k = 0; k < n; k++) {
    if (rate > 0) t += -std::log(1.0 - rng.uni()) / rate;
    arrival[k] = t;
  }
  const uint64_t tag0 = reqs[0].tag;
  const double t0 = gdp_server_now() + 0.002;
  std::atomic<bool> sub_done{false};
  std::thread sub([&]() {
    int64_t k = 0;
    while (k < n) {
      const double now = gdp_server_now() - t0;
      int64_t e = k;
      while (e < n && arrival[e] <= now) e++;
      if (e > k) {
        for (int64_t j = k; j < e; j += 65536)
          gdp_server_submit(s, reqs + j, (int32_t)std::min<int64_t>(65536, e - j));
        k = e;
      } else {
        const double wait = arrival[k] - now;
        if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait * 0.5e6)));
        else std::this_thread::yield();
      }
    }
    sub_done = true;
  });
  std::vector<gdp_align_result> res(65536);
  int64_t got = 0, ok = 0;
  double last = t0;
  while (got < n) {
    const int32_t m = gdp_server_poll(s, res.data(), (int32_t)res.size(), 1000);
    if (m < 0) break;
    if (m == 0 && sub_done.load()) {
      // a server stopped on a failure refuses what is left; once nothing
      // more arrives, stop waiting
      gdp_server_stats st;
      if (gdp_server_get_stats(s, &st) != GDP_OK && gdp_server_poll(s, res.data(), 0, 0) == 0 &&
          st.completed >= st.submitted)
        break;
    }
    for (int32_t i = 0; i < m; i++) {
      const int64_t k = (int64_t)(res[i].tag - tag0);
      if (k < 0 || k >= n) continue;
      latency[k] = res[i].t_done - (t0 + arrival[k]);
      last = std::max(last, res[i].t_done);
      if (res[i].status == GDP_OK) ok++;
    }
    got += m;
  }
  sub.join();
  if (elapsed) *elapsed = last - t0;
  return ok;
}

}  // extern "C"
