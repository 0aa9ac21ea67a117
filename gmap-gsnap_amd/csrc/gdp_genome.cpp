// gdp_genome.cpp -- host side of the device genome: GMAP's genomecomp image.
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/gmapdp.h"

namespace {

// Two-bit code and flag of one base as Compress_compress classifies it
// (compress-write.c:592-608): A C G T in either case by code, N a flag with
// code 0, X code 3 (T) with a flag, every other character a flag with code 0.
inline void classify(unsigned char ch, uint32_t *code, uint32_t *flag) {
  switch (ch | 0x20) {
  case 'a': *code = 0; *flag = 0; return;
  case 'c': *code = 1; *flag = 0; return;
  case 'g': *code = 2; *flag = 0; return;
  case 't': *code = 3; *flag = 0; return;
  case 'x': *code = 3; *flag = 1; return;
  default: *code = 0; *flag = 1; return;
  }
}

}  // namespace

extern "C" int64_t gdp_genomecomp_words(uint64_t nbases) { return (int64_t)((nbases + 31) / 32) * 3 + 2; }

// Each block of 32 bases is independent: base k of the block lands at bits
// 2k of low (k < 16) or 2(k - 16) of high, flag bit k; bases past the end are
// X (code 3 and a flag), as the padding loop of compress-write.c:619-631.
extern "C" int gdp_genomecomp_pack(const char *seq, uint64_t nbases, uint32_t *out) {
  if ((nbases > 0 && seq == nullptr) || out == nullptr) return GDP_EINVAL;
  const uint64_t nblocks = (nbases + 31) / 32;
  const int nthreads = nblocks > (1u << 16) ? 16 : 1;
  std::vector<std::thread> th;
  auto work = [=](uint64_t b0, uint64_t b1) {
    for (uint64_t b = b0; b < b1; b++) {
      uint32_t high = 0, low = 0, flags = 0;
      for (int k = 0; k < 32; k++) {
        const uint64_t pos = b * 32 + (uint64_t)k;
        uint32_t code = 3, flag = 1;
        if (pos < nbases) classify((unsigned char)seq[pos], &code, &flag);
        if (k < 16) low |= code << (2 * k);
        else high |= code << (2 * (k - 16));
        flags |= flag << k;
      }
      out[3 * b] = high;
      out[3 * b + 1] = low;
      out[3 * b + 2] = flags;
    }
  };
  if (nthreads == 1) {
    work(0, nblocks);
  } else {
    for (int t = 0; t < nthreads; t++) th.emplace_back(work, nblocks * t / nthreads, nblocks * (t + 1) / nthreads);
    for (auto &t : th) t.join();
  }
  out[3 * nblocks] = 0xFFFFFFFFu;
  out[3 * nblocks + 1] = 0xFFFFFFFFu;
  return GDP_OK;
}
