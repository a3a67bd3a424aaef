// Host side of the p2z C ABI: error text, the input generator and the printed
// report.  Compiled by the host compiler with -ffp-contract=off: the generator
// must round exactly as the reference writes its expressions.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../csrc/p2z_capi_util.hpp"
#include "p2z.h"

namespace p2z {
namespace {
thread_local std::string g_last_error;
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

}  // namespace p2z

namespace {

// Marsaglia polar normal variate over libc rand() with one cached variate
// (reference randn, OMP:127-145).  As in the reference C++ build, pow(float,
// int) promotes to double while log and sqrt of a float stay in float.
struct Randn {
  bool have = false;
  float x2 = 0.f;
  float operator()(float mu, float sigma) {
    if (have) {
      have = false;
      return mu + sigma * x2;
    }
    float u1, u2, w;
    do {
      u1 = -1 + ((float)rand() / RAND_MAX) * 2;
      u2 = -1 + ((float)rand() / RAND_MAX) * 2;
      w = (float)(std::pow((double)u1, 2) + std::pow((double)u2, 2));
    } while (w >= 1 || w == 0);
    const float mult = std::sqrt((-2 * std::log(w)) / w);
    const float x1 = u1 * mult;
    x2 = u2 * mult;
    have = true;
    return mu + sigma * x1;
  }
};
Randn g_randn;

constexpr int kTrkWords = 28, kHitWords = 9;
bool valid_width(int w) { return w == 16 || w == 32 || w == 64 || w == 128; }

}  // namespace

extern "C" {

const char* p2z_last_error(void) { return p2z::g_last_error.c_str(); }

void p2z_srand(unsigned seed) {
  srand(seed);
  g_randn.have = false;
}

float p2z_randn(float mu, float sigma) { return g_randn(mu, sigma); }

int p2z_prepare_tracks(const p2z_atrk* in, void* out, int64_t nevts, int64_t nb, int32_t width,
                       double smear) {
  if (!in || !out) return p2z::fail(P2Z_EINVAL, "null argument");
  if (!valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (nevts < 0 || nb < 0) return p2z::fail(P2Z_EINVAL, "negative size");
  float* res = static_cast<float*>(out);
  const int W = width;
  // prepareTracks (OMP:243-263): events, blocks, lanes; 6 then 21 draws per lane
  for (int64_t ie = 0; ie < nevts; ++ie)
    for (int64_t ib = 0; ib < nb; ++ib) {
      float* blk = res + (ib + nb * ie) * (int64_t)kTrkWords * W;
      for (int it = 0; it < W; ++it) {
        for (int ip = 0; ip < 6; ++ip)
          blk[ip * W + it] = (float)((1 + smear * g_randn(0, 1)) * in->par[ip]);
        for (int ip = 0; ip < 21; ++ip)
          blk[(6 + ip) * W + it] = (float)((1 + smear * g_randn(0, 1)) * in->cov[ip] * 100);
        reinterpret_cast<int32_t*>(blk)[27 * W + it] = in->q;
      }
    }
  return P2Z_OK;
}

int p2z_prepare_hits(const p2z_ahit* hits, int32_t nhits, void* out, int64_t nevts, int64_t nb,
                     int32_t width, int32_t nlayer, double smear) {
  if (!hits || !out) return p2z::fail(P2Z_EINVAL, "null argument");
  if (!valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (nevts < 0 || nb < 0 || nlayer < 0) return p2z::fail(P2Z_EINVAL, "negative size");
  if (nlayer > nhits)
    return p2z::fail(P2Z_EINVAL, "nlayer %d exceeds the %d input hits", nlayer, nhits);
  float* res = static_cast<float*>(out);
  const int W = width;
  // prepareHits (OMP:265-293): the layer loop is outermost
  for (int lay = 0; lay < nlayer; ++lay) {
    const p2z_ahit& h = hits[lay];
    for (int64_t ie = 0; ie < nevts; ++ie)
      for (int64_t ib = 0; ib < nb; ++ib) {
        float* blk = res + (lay + nlayer * (ib + nb * ie)) * (int64_t)kHitWords * W;
        for (int it = 0; it < W; ++it) {
          for (int ip = 0; ip < 3; ++ip)
            blk[ip * W + it] = (float)((1 + smear * g_randn(0, 1)) * h.pos[ip]);
          for (int ip = 0; ip < 6; ++ip)
            blk[(3 + ip) * W + it] = (float)((1 + smear * g_randn(0, 1)) * h.cov[ip]);
        }
      }
  }
  return P2Z_OK;
}

// Report statistics (OMP:758-851): per track in double, non-finite tracks
// counted as nans, tracks outside the reference's window counted as failed,
// means and RMS over the rest; RMS goes through sqrtf as in the reference.
int p2z_summary_host(const void* out, int64_t nevts, int64_t nb, int32_t width, int64_t ntrks,
                     const float lasthit[3], p2z_summary* s) {
  if (!out || !lasthit || !s) return p2z::fail(P2Z_EINVAL, "null argument");
  if (!valid_width(width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", width);
  if (nevts < 0 || ntrks < 0 || ntrks > nb * width)
    return p2z::fail(P2Z_EINVAL, "need 0 <= ntrks <= nb*width");
  const float* o = static_cast<const float*>(out);
  const int W = width;
  std::memset(s, 0, sizeof *s);
  const double hx = lasthit[0], hy = lasthit[1], hz = lasthit[2];
  auto visit = [&](auto&& body) {
    for (int64_t ie = 0; ie < nevts; ++ie)
      for (int64_t it = 0; it < ntrks; ++it) {
        const float* blk = o + ((it / W) + nb * ie) * (int64_t)kTrkWords * W;
        const int64_t l = it % W;
        const double x = blk[0 * W + l], y = blk[1 * W + l], z = blk[2 * W + l];
        const double pt = 1. / blk[3 * W + l];
        const double phi = blk[4 * W + l], theta = blk[5 * W + l];
        int state = 0;
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z) || !std::isfinite(pt) ||
            !std::isfinite(phi) || !std::isfinite(theta))
          state = 1;
        else if (std::fabs((x - hx) / hx) > 1. || std::fabs((y - hy) / hy) > 1. ||
                 std::fabs((z - hz) / hz) > 1. || std::fabs((pt - 12.) / 12.) > 1. ||
                 std::fabs((phi - 1.3) / 1.3) > 1. || std::fabs((theta - 2.8) / 2.8) > 1.)
          state = 2;
        body(state, x, y, z, pt, phi, theta);
      }
  };
  visit([&](int st, double x, double y, double z, double pt, double phi, double theta) {
    if (st == 1) { s->nnans++; return; }
    if (st == 2) { s->nfail++; return; }
    s->avgpt += pt; s->avgphi += phi; s->avgtheta += theta;
    s->avgx += x; s->avgy += y; s->avgz += z;
    s->avgdx += (x - hx) / x; s->avgdy += (y - hy) / y; s->avgdz += (z - hz) / z;
    s->ngood++;
  });
  const double n = (double)s->ngood;
  s->avgpt /= n; s->avgphi /= n; s->avgtheta /= n;
  s->avgx /= n; s->avgy /= n; s->avgz /= n;
  s->avgdx /= n; s->avgdy /= n; s->avgdz /= n;
  visit([&](int st, double x, double y, double z, double, double, double) {
    if (st != 0) return;
    s->stdx += (x - s->avgx) * (x - s->avgx);
    s->stdy += (y - s->avgy) * (y - s->avgy);
    s->stdz += (z - s->avgz) * (z - s->avgz);
    const double ddx = (x - hx) / x - s->avgdx, ddy = (y - hy) / y - s->avgdy, ddz = (z - hz) / z - s->avgdz;
    s->stddx += ddx * ddx; s->stddy += ddy * ddy; s->stddz += ddz * ddz;
  });
  s->stdx = sqrtf((float)(s->stdx / n)); s->stdy = sqrtf((float)(s->stdy / n));
  s->stdz = sqrtf((float)(s->stdz / n)); s->stddx = sqrtf((float)(s->stddx / n));
  s->stddy = sqrtf((float)(s->stddy / n)); s->stddz = sqrtf((float)(s->stddz / n));
  return P2Z_OK;
}

int p2z_format_summary(const p2z_summary* s, char* buf, int32_t cap) {
  if (!s || !buf || cap <= 0) return p2z::fail(P2Z_EINVAL, "null argument");
  const int n = snprintf(buf, (size_t)cap,
                         "track x avg=%f std/avg=%f\n"
                         "track y avg=%f std/avg=%f\n"
                         "track z avg=%f std/avg=%f\n"
                         "track dx/x avg=%f std=%f\n"
                         "track dy/y avg=%f std=%f\n"
                         "track dz/z avg=%f std=%f\n"
                         "track pt avg=%f\n"
                         "track phi avg=%f\n"
                         "track theta avg=%f\n"
                         "number of tracks with nans=%i\n"
                         "number of tracks failed=%i\n",
                         s->avgx, std::fabs(s->stdx / s->avgx), s->avgy, std::fabs(s->stdy / s->avgy),
                         s->avgz, std::fabs(s->stdz / s->avgz), s->avgdx, s->stddx, s->avgdy, s->stddy,
                         s->avgdz, s->stddz, s->avgpt, s->avgphi, s->avgtheta, s->nnans, s->nfail);
  return n < cap ? n : p2z::fail(P2Z_EINVAL, "buffer too small");
}

}  // extern "C"
