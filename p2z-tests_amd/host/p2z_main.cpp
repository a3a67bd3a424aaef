// Drop-in for the reference program (src/propagate-toz-test_OMP.cpp main,
// OMP:684-883): same inputs, same printout, the path on the GPU.
//
//   p2z_main [--preset ref|config0] [--nevts N] [--ntrks N] [--bsize W]
//            [--nlayer L] [--niter K] [--math fast|ref|fp64gain]
//            [--include-data] [--device-summary] [--per-block] [--device D]
//
// Defaults are the reference's (nevts 100, ntrks 9600, bsize 32, nlayer 20),
// with NITER 10 as its Makefile sets it, and its arithmetic (--math ref: the
// same bits as the reference build, so the same printout; --math fast and
// --math fp64gain are the production modes).  --include-data times host->device
// and device->host copies of every pass (the reference's INCLUDE_DATA mode);
// otherwise the timed region is the device-resident passes.  --per-block runs
// the reference's own loop instead (OMP:733-747: per event and block, copy the
// track block, then per layer propagateToZ and KalmanUpdate_v2 on that block)
// through the per-block functions of p2z.hpp -- the same results, one GPU call
// per block and step, so for small sizes only.  The report is computed on the
// host exactly as the reference does, or on the GPU with --device-summary.
#include <sys/time.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "p2z.hpp"

namespace {

struct Args {
  int preset = P2Z_PRESET_REF;
  int bsize = P2Z_BSIZE;
  bool include_data = false;
  bool device_summary = false;
  bool per_block = false;
};

// The reference's driver loop (OMP:733-747) over the per-block functions.
template <int W>
void per_block_loop(const MPTRK_<W>* trk, const MPHIT_<W>* hit, MPTRK_<W>* outtrk) {
  const P2zConfig& c = p2zConfig();
  for (int64_t ie = 0; ie < c.nevts; ++ie)
    for (int64_t ib = 0; ib < nbOf<W>(); ++ib) {
      const MPTRK_<W>* btracks = bTk(trk, ie, ib);
      MPTRK_<W>* obtracks = bTk(outtrk, ie, ib);
      *obtracks = *btracks;
      for (int layer = 0; layer < c.nlayer; ++layer) {
        const MPHIT_<W>* bhits = bHit(hit, ie, ib, layer);
        propagateToZ(&obtracks->cov, &obtracks->par, &obtracks->q, &bhits->pos, &obtracks->cov, &obtracks->par);
        KalmanUpdate_v2(&obtracks->cov, &obtracks->par, &bhits->cov, &bhits->pos);
      }
    }
}

long now_ms() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return (long)tv.tv_sec * 1000 + (long)tv.tv_usec / 1000;
}

[[noreturn]] void usage(const char* msg) {
  std::fprintf(stderr, "p2z_main: %s\n", msg);
  std::exit(2);
}

template <int W>
int run(const Args& a) {
  P2zConfig& c = p2zConfig();
  const int64_t nb = nbOf<W>();
  std::printf("bsize: %d, nb:%d, gpu device:%d\n", W, (int)nb, c.device);

  ATRK inputtrk;
  std::vector<AHIT> inputhits(26);
  double smear = 0;
  const int nh = p2z_preset_input(a.preset, reinterpret_cast<p2z_atrk*>(&inputtrk),
                                  reinterpret_cast<p2z_ahit*>(inputhits.data()), 26, &smear);
  if (nh < 0) usage("unknown preset");
  c.smear = smear;
  if (c.nlayer > (int)inputhits.size()) usage("nlayer larger than the number of input hits");

  std::printf("track in pos: x=%f, y=%f, z=%f, r=%f, pt=%f, phi=%f, theta=%f \n", inputtrk.par[0],
              inputtrk.par[1], inputtrk.par[2],
              sqrtf(inputtrk.par[0] * inputtrk.par[0] + inputtrk.par[1] * inputtrk.par[1]),
              1. / inputtrk.par[3], inputtrk.par[4], inputtrk.par[5]);
  std::printf("track in cov: %.2e, %.2e, %.2e \n", inputtrk.cov[0], inputtrk.cov[2], inputtrk.cov[5]);
  for (int lay = 0; lay < c.nlayer; lay++)
    std::printf("hit in layer=%d, pos: x=%f, y=%f, z=%f, r=%f \n", lay, inputhits[lay].pos[0],
                inputhits[lay].pos[1], inputhits[lay].pos[2],
                sqrtf(inputhits[lay].pos[0] * inputhits[lay].pos[0] + inputhits[lay].pos[1] * inputhits[lay].pos[1]));
  std::printf("produce nevts=%i ntrks=%i smearing by=%2.1e \n", (int)c.nevts, (int)c.ntrks, c.smear);
  std::printf("NITER=%d\n", c.niter);

  const long setup_start = now_ms();
  p2z_srand(1);
  MPTRK_<W>* trk = prepareTracks<W>(inputtrk);
  MPHIT_<W>* hit = prepareHits<W>(inputhits);
  auto* outtrk = static_cast<MPTRK_<W>*>(std::malloc(sizeof(MPTRK_<W>) * c.nevts * nb));
  const long setup_stop = now_ms();
  std::printf("done preparing!\n");

  const auto wall_start = std::chrono::high_resolution_clock::now();
  double gpu_s = 0;
  if (a.per_block) {
    for (int it = 0; it < c.niter; ++it) per_block_loop<W>(trk, hit, outtrk);
  } else {
    gpu_s = p2zRunDriver<W>(trk, hit, outtrk, c.niter, a.include_data);
  }
  const auto wall_stop = std::chrono::high_resolution_clock::now();
  const double wall_time =
      std::chrono::duration_cast<std::chrono::microseconds>(wall_stop - wall_start).count() / 1e6;
  const double setup_s = (setup_stop - setup_start) * 0.001;
  const long long ntot = (long long)c.nevts * c.ntrks * c.niter;
  std::printf("setup time time=%f (s)\n", setup_s);
  std::printf("done ntracks=%lld tot time=%f (s) time/trk=%e (s)\n", ntot, wall_time, wall_time / ntot);
  std::printf("formatted %i %i %i %i %i %f 0 %f %i\n", c.niter, (int)c.nevts, (int)c.ntrks, W, (int)nb,
              wall_time, setup_s, 1);
  if (!a.per_block)
    std::printf("gpu time=%f (s) time/trk=%e (s) track-layer updates/s=%e\n", gpu_s, gpu_s / ntot,
                (double)ntot * c.nlayer / gpu_s);

  const float last[3] = {inputhits[c.nlayer - 1].pos[0], inputhits[c.nlayer - 1].pos[1],
                         inputhits[c.nlayer - 1].pos[2]};
  p2z_summary s;
  if (a.device_summary) {
    p2z_ctx* ctx = p2zContext();
    void* d = nullptr;
    const size_t bytes = sizeof(MPTRK_<W>) * c.nevts * nb;
    p2zCheck(p2z_dev_alloc(ctx, &d, bytes), "alloc");
    p2zCheck(p2z_memcpy_h2d(ctx, d, outtrk, bytes), "h2d");
    p2zCheck(p2z_summary_device(ctx, d, c.nevts, nb, W, c.ntrks, last, &s), "summary");
    p2z_dev_free(ctx, d);
  } else {
    p2zCheck(p2z_summary_host(outtrk, c.nevts, nb, W, c.ntrks, last, &s), "summary");
  }
  char buf[2048];
  p2zCheck(p2z_format_summary(&s, buf, sizeof buf) < 0 ? P2Z_EINVAL : P2Z_OK, "format");
  std::fputs(buf, stdout);

  std::free(trk);
  std::free(hit);
  std::free(outtrk);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Args a;
  P2zConfig& c = p2zConfig();
  for (int i = 1; i < argc; ++i) {
    const std::string k = argv[i];
    auto val = [&]() -> const char* {
      if (i + 1 >= argc) usage(("missing value for " + k).c_str());
      return argv[++i];
    };
    if (k == "--preset") {
      const std::string v = val();
      a.preset = v == "ref" ? P2Z_PRESET_REF : v == "config0" ? P2Z_PRESET_CONFIG0 : -1;
      if (a.preset < 0) usage("preset must be ref or config0");
    } else if (k == "--nevts") c.nevts = std::atoll(val());
    else if (k == "--ntrks") c.ntrks = std::atoll(val());
    else if (k == "--bsize") a.bsize = std::atoi(val());
    else if (k == "--nlayer") c.nlayer = std::atoi(val());
    else if (k == "--niter") c.niter = std::atoi(val());
    else if (k == "--device") c.device = std::atoi(val());
    else if (k == "--math") {
      const std::string v = val();
      c.math = v == "fast" ? P2Z_MATH_FAST : v == "ref" ? P2Z_MATH_REF : v == "fp64gain" ? P2Z_MATH_FP64GAIN : -1;
      if (c.math < 0) usage("math must be fast, ref or fp64gain");
    } else if (k == "--include-data") a.include_data = true;
    else if (k == "--device-summary") a.device_summary = true;
    else if (k == "--per-block") a.per_block = true;
    else usage(("unknown option " + k).c_str());
  }
  if (c.nevts <= 0 || c.ntrks <= 0 || c.nlayer <= 0 || c.niter <= 0) usage("sizes must be positive");
  try {
    switch (a.bsize) {
      case 16: return run<16>(a);
      case 32: return run<32>(a);
      case 64: return run<64>(a);
      case 128: return run<128>(a);
    }
    usage("bsize must be 16, 32, 64 or 128");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "p2z_main: %s\n", e.what());
    return 1;
  }
}
