// Synthetic workloads and the ragged-event format (declared in include/p2z.h).
//
// Each event is generated from its own random stream (splitmix64-seeded
// xoshiro256**, Box-Muller normals), so the result depends on the seed only,
// not on the thread count.  Tracks follow the path's own helix model in
// double: from (x0, y0, z0, phi0) with signed radius rho = k / ipt
// (k = q * kfact) the track reaches the plane z after turning by
// alpha = (z - z0) tan(theta) / rho, at x0 + rho (sin(phi0+alpha) - sin phi0),
// y0 + rho (cos phi0 - cos(phi0+alpha)).
#include <omp.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../csrc/p2z_capi_util.hpp"
#include "p2z.h"

struct p2z_workload {
  p2z_workload_spec spec{};
  p2z_workload_info info{};
  std::vector<float> trk, hit;
  std::vector<int32_t> lane_nlay;
  std::vector<p2z_block_desc> blocks, blocks_lpt;
  std::vector<int64_t> event_blocks;
  std::vector<int32_t> event_tracks;
};

namespace {

// A file closed when it goes out of scope.
using File = std::unique_ptr<FILE, int (*)(FILE*)>;

constexpr int kTrkWords = 28, kHitWords = 9;
constexpr double kKfact64 = 100. / (-0.299792458 * 3.8112);
constexpr double kPi = 3.14159265358979323846;

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// xoshiro256** (Blackman and Vigna), seeded through splitmix64.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (auto& v : s) v = splitmix64(seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * 0x1.0p-53; }            // [0, 1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double loguniform(double a, double b) { return std::exp(uniform(std::log(a), std::log(b))); }
  int integer(int lo, int hi) { return lo + (int)(uniform() * (hi - lo + 1)); }  // [lo, hi]
  double normal() {                                                    // Box-Muller
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * uniform());
  }
};

uint64_t event_seed(uint64_t seed, int64_t e, uint64_t stream) {
  uint64_t s = seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(e + 1)) ^ (stream << 56);
  return splitmix64(s);
}

bool valid_width(int w) { return w == 16 || w == 32 || w == 64 || w == 128; }

struct TrackSpec {
  int nlay;
};

// Sizes of one event: track count and each track's layer count.
std::vector<int> event_layers(const p2z_workload_spec& sp, int64_t e) {
  Rng r(event_seed(sp.seed, e, 1));
  int n;
  if (sp.kind == P2Z_WORKLOAD_RAGGED) {
    n = (int)std::lround(2000.0 * std::exp(r.normal()));
    n = std::min(20000, std::max(50, n));
  } else {
    n = sp.ntracks;
  }
  std::vector<int> lay(n);
  for (int t = 0; t < n; ++t) lay[t] = sp.kind == P2Z_WORKLOAD_RAGGED ? r.integer(8, 26) : sp.nlayer;
  return lay;
}

// One track: writes the starting state into lane l of the track block and its
// hits into lane l of the block's hit slots.
void make_track(Rng& r, int kind, int nlay, float* tb, float* hb0, int W, int l) {
  const int q = r.uniform() < 0.5 ? -1 : 1;
  const double ipt = r.loguniform(0.01, 2.0);
  const double phi0 = r.uniform(-kPi, kPi);
  double theta;
  if (kind == P2Z_WORKLOAD_RAGGED) {
    do theta = r.uniform(0.1, 3.04); while (std::fabs(std::cos(theta)) < 0.35);
  } else {
    theta = r.uniform(0.1, 3.04);
  }
  const double x0 = 0.01 * r.normal(), y0 = 0.01 * r.normal(), z0 = 5.0 * r.normal();
  const double rho = q * kKfact64 / ipt;
  const double tanT = std::tan(theta);
  const double dir = std::cos(theta) >= 0 ? 1.0 : -1.0;
  double z = z0;
  const double vlo = kind == P2Z_WORKLOAD_RAGGED ? 1e-6 : 1e-8, vhi = 1e-3;
  const double rho_max = kind == P2Z_WORKLOAD_RAGGED ? 0.5 : 0.9;
  for (int L = 0; L < nlay; ++L) {
    z += dir * r.uniform(5.0, 25.0);
    const double a = (z - z0) * tanT / rho;
    const double xt = x0 + rho * (std::sin(phi0 + a) - std::sin(phi0));
    const double yt = y0 + rho * (std::cos(phi0) - std::cos(phi0 + a));
    const double vxx = r.loguniform(vlo, vhi), vyy = r.loguniform(vlo, vhi);
    const double corr = r.uniform(-rho_max, rho_max);
    const double vxy = corr * std::sqrt(vxx * vyy);
    // correlated noise through the Cholesky factor of [[vxx, vxy], [vxy, vyy]]
    const double g1 = r.normal(), g2 = r.normal();
    const double l11 = std::sqrt(vxx), l21 = vxy / l11, l22 = std::sqrt(std::max(vyy - l21 * l21, 0.0));
    float* hb = hb0 + (int64_t)L * kHitWords * W;
    hb[0 * W + l] = (float)(xt + l11 * g1);
    hb[1 * W + l] = (float)(yt + l21 * g1 + l22 * g2);
    hb[2 * W + l] = (float)z;
    hb[3 * W + l] = (float)vxx;
    hb[4 * W + l] = (float)vxy;
    hb[5 * W + l] = (float)vyy;
    hb[6 * W + l] = 0.f;
    hb[7 * W + l] = 0.f;
    hb[8 * W + l] = 1e-6f;
  }
  // starting state: truth smeared by its covariance (diagonal)
  const double sig[6] = {0.01, 0.01, 0.0, 0.05 * ipt, 0.005, 0.005};
  const double truth[6] = {x0, y0, z0, ipt, phi0, theta};
  for (int e = 0; e < 6; ++e) tb[e * W + l] = (float)(truth[e] + sig[e] * r.normal());
  float P[21] = {};
  const double var[6] = {1e-4, 1e-4, 1e-6, sig[3] * sig[3], sig[4] * sig[4], sig[5] * sig[5]};
  for (int i = 0; i < 6; ++i) P[i * (i + 1) / 2 + i] = (float)var[i];
  for (int e = 0; e < 21; ++e) tb[(6 + e) * W + l] = P[e];
  reinterpret_cast<int32_t*>(tb)[27 * W + l] = q;
}

int build(const p2z_workload_spec& sp, p2z_workload& w) {
  const int W = sp.width;
  const int64_t ne = sp.nevents;
  const int nthr = sp.nthreads > 0 ? sp.nthreads : omp_get_max_threads();
  // phase 1: sizes and offsets
  std::vector<std::vector<int>> lays(ne);
  // an exception may not leave a parallel region: a failed allocation is
  // flagged here and thrown after the loop
  int alloc_failed = 0;
#pragma omp parallel for num_threads(nthr) schedule(dynamic, 8)
  for (int64_t e = 0; e < ne; ++e) {
    try {
      lays[e] = event_layers(sp, e);
    } catch (const std::bad_alloc&) {
#pragma omp atomic write
      alloc_failed = 1;
      continue;
    }
    std::stable_sort(lays[e].begin(), lays[e].end(), std::greater<int>());
  }
  if (alloc_failed) throw std::bad_alloc();
  w.event_blocks.assign(ne + 1, 0);
  w.event_tracks.assign(ne, 0);
  std::vector<int64_t> hit_off(ne + 1, 0);
  int64_t ntracks = 0, tl = 0;
  for (int64_t e = 0; e < ne; ++e) {
    const int64_t n = (int64_t)lays[e].size();
    const int64_t nbk = (n + W - 1) / W;
    w.event_tracks[e] = (int32_t)n;
    w.event_blocks[e + 1] = w.event_blocks[e] + nbk;
    int64_t slots = 0;
    for (int64_t b = 0; b < nbk; ++b) slots += lays[e][b * W];  // first lane has the most layers
    hit_off[e + 1] = hit_off[e] + slots;
    ntracks += n;
    for (int v : lays[e]) tl += v;
  }
  const int64_t nblocks = w.event_blocks[ne], nhit = hit_off[ne];
  w.trk.assign((size_t)nblocks * kTrkWords * W, 0.f);
  w.hit.assign((size_t)nhit * kHitWords * W, 0.f);
  w.lane_nlay.assign((size_t)nblocks * W, 0);
  w.blocks.resize(nblocks);
  // phase 2: physics, one stream per event
#pragma omp parallel for num_threads(nthr) schedule(dynamic, 4)
  for (int64_t e = 0; e < ne; ++e) {
    Rng r(event_seed(sp.seed, e, 2));
    const int64_t n = (int64_t)lays[e].size();
    int64_t hb = hit_off[e];
    for (int64_t b = w.event_blocks[e]; b < w.event_blocks[e + 1]; ++b) {
      const int64_t t0 = (b - w.event_blocks[e]) * W;
      const int nslot = lays[e][t0];
      w.blocks[b] = p2z_block_desc{b, hb, b, nslot, (int32_t)e};
      for (int l = 0; l < W && t0 + l < n; ++l) {
        const int nl = lays[e][t0 + l];
        w.lane_nlay[(size_t)b * W + l] = nl;
        make_track(r, sp.kind, nl, w.trk.data() + (size_t)b * kTrkWords * W,
                   w.hit.data() + (size_t)hb * kHitWords * W, W, l);
      }
      hb += nslot;
    }
  }
  w.blocks_lpt = w.blocks;
  std::stable_sort(w.blocks_lpt.begin(), w.blocks_lpt.end(),
                   [](const p2z_block_desc& a, const p2z_block_desc& b) { return a.nlayer > b.nlayer; });
  w.info = p2z_workload_info{ne, ntracks, nblocks, nhit, tl, W, sp.kind};
  return P2Z_OK;
}

uint64_t fnv1a(const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  return h;
}

template <typename T>
void put(FILE* f, const std::vector<T>& v) {
  if (!v.empty()) fwrite(v.data(), sizeof(T), v.size(), f);
}

template <typename T>
bool get(FILE* f, std::vector<T>& v, size_t n) {
  v.resize(n);
  return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

bool field(const std::string& h, const char* key, long long* v) {
  const std::string k = std::string("\"") + key + "\":";
  const size_t p = h.find(k);
  if (p == std::string::npos) return false;
  return sscanf(h.c_str() + p + k.size(), "%lld", v) == 1;
}

bool field_u(const std::string& h, const char* key, unsigned long long* v) {
  const std::string k = std::string("\"") + key + "\":";
  const size_t p = h.find(k);
  if (p == std::string::npos) return false;
  return sscanf(h.c_str() + p + k.size(), "%llu", v) == 1;
}

}  // namespace

extern "C" {

static int workload_create_impl(const p2z_workload_spec* sp, p2z_workload** out) {
  if (!sp || !out) return p2z::fail(P2Z_EINVAL, "null argument");
  *out = nullptr;
  if (sp->kind != P2Z_WORKLOAD_PRECISION && sp->kind != P2Z_WORKLOAD_RAGGED)
    return p2z::fail(P2Z_EINVAL, "unknown workload kind %d", sp->kind);
  if (!valid_width(sp->width)) return p2z::fail(P2Z_EINVAL, "width must be 16/32/64/128 (got %d)", sp->width);
  if (sp->nevents < 0 || sp->nevents > (1 << 24)) return p2z::fail(P2Z_EINVAL, "bad event count");
  if (sp->kind == P2Z_WORKLOAD_PRECISION && (sp->ntracks <= 0 || sp->nlayer <= 0 || sp->nlayer > 1000))
    return p2z::fail(P2Z_EINVAL, "precision workload needs ntracks > 0 and 0 < nlayer <= 1000");
  std::unique_ptr<p2z_workload> w(new p2z_workload);
  w->spec = *sp;
  const int rc = build(*sp, *w);
  if (rc != P2Z_OK) return rc;
  *out = w.release();
  return P2Z_OK;
}

int p2z_workload_create(const p2z_workload_spec* sp, p2z_workload** out) {
  return p2z::guarded("p2z_workload_create", [&] { return workload_create_impl(sp, out); });
}

int p2z_workload_destroy(p2z_workload* w) {
  delete w;
  return P2Z_OK;
}

int p2z_workload_get_info(const p2z_workload* w, p2z_workload_info* info) {
  if (!w || !info) return p2z::fail(P2Z_EINVAL, "null argument");
  *info = w->info;
  return P2Z_OK;
}

const void* p2z_workload_tracks(const p2z_workload* w) { return w ? w->trk.data() : nullptr; }
const void* p2z_workload_hits(const p2z_workload* w) { return w ? w->hit.data() : nullptr; }
const int32_t* p2z_workload_lane_nlayer(const p2z_workload* w) { return w ? w->lane_nlay.data() : nullptr; }
const p2z_block_desc* p2z_workload_blocks(const p2z_workload* w) { return w ? w->blocks.data() : nullptr; }
const p2z_block_desc* p2z_workload_blocks_lpt(const p2z_workload* w) { return w ? w->blocks_lpt.data() : nullptr; }
const int64_t* p2z_workload_event_blocks(const p2z_workload* w) { return w ? w->event_blocks.data() : nullptr; }
const int32_t* p2z_workload_event_tracks(const p2z_workload* w) { return w ? w->event_tracks.data() : nullptr; }

static int workload_save_impl(const p2z_workload* w, const char* path) {
  if (!w || !path) return p2z::fail(P2Z_EINVAL, "null argument");
  FILE* f = fopen(path, "wb");
  if (!f) return p2z::fail(P2Z_EINVAL, "cannot open %s for writing", path);
  const p2z_workload_info& i = w->info;
  const p2z_workload_spec& s = w->spec;
  char hdr[1024];
  const int n = snprintf(
      hdr, sizeof hdr,
      "{\"format\":\"p2z-workload\",\"version\":1,\"kind\":%d,\"width\":%d,\"nevents\":%lld,\"seed\":%llu,"
      "\"spec_ntracks\":%d,\"spec_nlayer\":%d,\"ntracks\":%lld,\"nblocks\":%lld,\"nhit_blocks\":%lld,"
      "\"track_layers\":%lld,\"sum_tracks\":%llu,\"sum_hits\":%llu,\"sum_lanes\":%llu,\"sum_blocks\":%llu,"
      "\"sum_lpt\":%llu,\"sum_events\":%llu,\"sum_evtracks\":%llu}\n",
      s.kind, s.width, (long long)i.nevents, (unsigned long long)s.seed, s.ntracks, s.nlayer,
      (long long)i.ntracks, (long long)i.nblocks, (long long)i.nhit_blocks, (long long)i.track_layers,
      (unsigned long long)fnv1a(w->trk.data(), w->trk.size() * 4),
      (unsigned long long)fnv1a(w->hit.data(), w->hit.size() * 4),
      (unsigned long long)fnv1a(w->lane_nlay.data(), w->lane_nlay.size() * 4),
      (unsigned long long)fnv1a(w->blocks.data(), w->blocks.size() * sizeof(p2z_block_desc)),
      (unsigned long long)fnv1a(w->blocks_lpt.data(), w->blocks_lpt.size() * sizeof(p2z_block_desc)),
      (unsigned long long)fnv1a(w->event_blocks.data(), w->event_blocks.size() * 8),
      (unsigned long long)fnv1a(w->event_tracks.data(), w->event_tracks.size() * 4));
  fwrite(hdr, 1, (size_t)n, f);
  put(f, w->trk); put(f, w->hit); put(f, w->lane_nlay); put(f, w->blocks); put(f, w->blocks_lpt);
  put(f, w->event_blocks); put(f, w->event_tracks);
  const bool ok = ferror(f) == 0;
  fclose(f);
  return ok ? P2Z_OK : p2z::fail(P2Z_EINVAL, "write error on %s", path);
}

int p2z_workload_save(const p2z_workload* w, const char* path) {
  return p2z::guarded("p2z_workload_save", [&] { return workload_save_impl(w, path); });
}

static int workload_load_impl(const char* path, p2z_workload** out) {
  if (!path || !out) return p2z::fail(P2Z_EINVAL, "null argument");
  *out = nullptr;
  const File file(fopen(path, "rb"), &fclose);
  FILE* f = file.get();
  if (!f) return p2z::fail(P2Z_EINVAL, "cannot open %s", path);
  std::string hdr;
  for (int c; (c = fgetc(f)) != EOF && c != '\n';) hdr.push_back((char)c);
  long long kind, width, ne, snt, snl, nt, nb, nh, tl;
  unsigned long long seed, c[7];
  const char* sums[7] = {"sum_tracks", "sum_hits", "sum_lanes", "sum_blocks", "sum_lpt", "sum_events", "sum_evtracks"};
  bool ok = hdr.find("\"format\":\"p2z-workload\"") != std::string::npos && field(hdr, "kind", &kind) &&
            field(hdr, "width", &width) && field(hdr, "nevents", &ne) && field_u(hdr, "seed", &seed) &&
            field(hdr, "spec_ntracks", &snt) && field(hdr, "spec_nlayer", &snl) && field(hdr, "ntracks", &nt) &&
            field(hdr, "nblocks", &nb) && field(hdr, "nhit_blocks", &nh) && field(hdr, "track_layers", &tl);
  for (int k = 0; ok && k < 7; ++k) ok = field_u(hdr, sums[k], &c[k]);
  if (!ok || !valid_width((int)width) || ne < 0 || nb < 0 || nh < 0)
    return p2z::fail(P2Z_EINVAL, "%s: not a p2z workload file", path);
  std::unique_ptr<p2z_workload> w(new p2z_workload);
  w->spec = p2z_workload_spec{(int32_t)kind, (int32_t)width, ne, seed, (int32_t)snt, (int32_t)snl, 0, 0};
  w->info = p2z_workload_info{ne, nt, nb, nh, tl, (int32_t)width, (int32_t)kind};
  const size_t W = (size_t)width;
  ok = get(f, w->trk, (size_t)nb * kTrkWords * W) && get(f, w->hit, (size_t)nh * kHitWords * W) &&
       get(f, w->lane_nlay, (size_t)nb * W) && get(f, w->blocks, (size_t)nb) && get(f, w->blocks_lpt, (size_t)nb) &&
       get(f, w->event_blocks, (size_t)ne + 1) && get(f, w->event_tracks, (size_t)ne);
  if (ok) {
    const uint64_t got[7] = {fnv1a(w->trk.data(), w->trk.size() * 4), fnv1a(w->hit.data(), w->hit.size() * 4),
                             fnv1a(w->lane_nlay.data(), w->lane_nlay.size() * 4),
                             fnv1a(w->blocks.data(), w->blocks.size() * sizeof(p2z_block_desc)),
                             fnv1a(w->blocks_lpt.data(), w->blocks_lpt.size() * sizeof(p2z_block_desc)),
                             fnv1a(w->event_blocks.data(), w->event_blocks.size() * 8),
                             fnv1a(w->event_tracks.data(), w->event_tracks.size() * 4)};
    for (int k = 0; k < 7; ++k) ok = ok && got[k] == c[k];
  }
  if (!ok) return p2z::fail(P2Z_EINVAL, "%s: truncated or corrupt (checksum mismatch)", path);
  *out = w.release();
  return P2Z_OK;
}

int p2z_workload_load(const char* path, p2z_workload** out) {
  return p2z::guarded("p2z_workload_load", [&] { return workload_load_impl(path, out); });
}

int p2z_check_blocks(const p2z_block_desc* blocks, int64_t nblocks, int64_t pool_blocks, int64_t pool_hit_blocks,
                     int64_t out_blocks) {
  if (nblocks < 0 || (nblocks > 0 && !blocks)) return p2z::fail(P2Z_EINVAL, "null or negative descriptor table");
  for (int64_t i = 0; i < nblocks; ++i) {
    const p2z_block_desc& d = blocks[i];
    if (d.trk_block < 0 || d.trk_block >= pool_blocks || d.nlayer < 0 || d.hit_block < 0 ||
        d.hit_block > pool_hit_blocks - d.nlayer || d.out_block < 0 || d.out_block >= out_blocks)
      return p2z::fail(P2Z_EINVAL,
                       "descriptor %lld outside the buffers (trk_block %lld of %lld, hit blocks %lld+%d of %lld, "
                       "out_block %lld of %lld)",
                       (long long)i, (long long)d.trk_block, (long long)pool_blocks, (long long)d.hit_block,
                       d.nlayer, (long long)pool_hit_blocks, (long long)d.out_block, (long long)out_blocks);
  }
  return P2Z_OK;
}

}  // extern "C"
