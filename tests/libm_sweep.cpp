// Host program for tests/test_libm.py: p2z-tests_amd/csrc/p2z_libm.hpp (the
// sin/cos the reference-typing mode uses on the GPU) compiled for the host and
// compared with the C library's sinf, cosf and sincosf on a range of float bit
// patterns (default: all 2^32), bit for bit (NaN results compare equal to NaN).
//   libm_sweep [lo hi]   -> "mismatch sinf N cosf N sincosf_s N sincosf_c N"
//                           plus up to 8 differing inputs
//   libm_sweep table     -> the 24 words of the 2/pi table, one per line
#define __host__
#define __device__
#include <math.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "p2z_libm.hpp"

namespace {
// Called through pointers so that the compiler cannot merge a sinf and a cosf
// of one argument into sincosf: each library entry is checked on its own.
float (*volatile lib_sinf)(float) = sinf;
float (*volatile lib_cosf)(float) = cosf;
void (*volatile lib_sincosf)(float, float*, float*) = sincosf;

bool same(float a, float b) {
  if (std::isnan(a) && std::isnan(b)) return true;
  uint32_t ua, ub;
  std::memcpy(&ua, &a, 4);
  std::memcpy(&ub, &b, 4);
  return ua == ub;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc == 2 && std::strcmp(argv[1], "table") == 0) {
    for (uint32_t w : p2z::glibc_sincosf::kInvPio4) std::printf("%08x\n", w);
    return 0;
  }
  uint64_t lo = 0, hi = 1ull << 32;
  if (argc == 3) {
    lo = std::strtoull(argv[1], nullptr, 0);
    hi = std::strtoull(argv[2], nullptr, 0);
  }
  std::atomic<long long> bad[4] = {0, 0, 0, 0};
  std::atomic<int> shown{0};
  const unsigned nth = std::max(1u, std::thread::hardware_concurrency());
  const uint64_t span = hi - lo, chunk = (span + nth - 1) / nth;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nth; ++t) {
    pool.emplace_back([&, t] {
      const uint64_t a = lo + t * chunk, b = std::min(hi, a + chunk);
      long long local[4] = {0, 0, 0, 0};
      for (uint64_t i = a; i < b; ++i) {
        const uint32_t u = (uint32_t)i;
        float y;
        std::memcpy(&y, &u, 4);
        float s, c, ls, lc;
        p2z::sincosf_glibc(y, &s, &c);
        lib_sincosf(y, &ls, &lc);
        const bool ok[4] = {same(lib_sinf(y), s), same(lib_cosf(y), c), same(ls, s), same(lc, c)};
        for (int k = 0; k < 4; ++k)
          if (!ok[k]) {
            ++local[k];
            if (shown.fetch_add(1) < 8) std::printf("differs %08x (%a): sin %a cos %a\n", u, y, s, c);
          }
      }
      for (int k = 0; k < 4; ++k) bad[k] += local[k];
    });
  }
  for (auto& th : pool) th.join();
  std::printf("mismatch sinf %lld cosf %lld sincosf_s %lld sincosf_c %lld over %llu inputs\n",
              bad[0].load(), bad[1].load(), bad[2].load(), bad[3].load(), (unsigned long long)span);
  return 0;
}
