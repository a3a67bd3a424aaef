// The host code of the p2z library (generator, report, workloads and their
// files, descriptor checks) built on its own with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py and run here:
// valid calls, argument errors, and damaged workload files.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "p2z.h"
int main(int argc, char** argv) {
  const char* path = argc > 1 ? argv[1] : "p2z_sanitize_workload.bin";
  // generator + report
  p2z_atrk t; p2z_ahit h[26]; double smear = 0;
  int nh = p2z_preset_input(P2Z_PRESET_REF, &t, h, 26, &smear);
  if (nh < 20) return 1;
  const int W = 16, nevts = 2, nb = 10, nl = 20;
  std::vector<float> trk((size_t)nevts * nb * 28 * W), hit((size_t)nevts * nb * nl * 9 * W);
  p2z_srand(1);
  if (p2z_prepare_tracks(&t, trk.data(), nevts, nb, W, smear) != P2Z_OK) return 2;
  if (p2z_prepare_hits(h, nh, hit.data(), nevts, nb, W, nl, smear) != P2Z_OK) return 3;
  if (p2z_prepare_hits(h, 5, hit.data(), nevts, nb, W, nl, smear) != P2Z_EINVAL) return 4;
  p2z_summary s;
  if (p2z_summary_host(trk.data(), nevts, nb, W, nb * W, h[nl - 1].pos, &s) != P2Z_OK) return 5;
  char buf[4096];
  if (p2z_format_summary(&s, buf, sizeof buf) < 0) return 6;
  if (p2z_format_summary(&s, buf, 8) >= 0 && buf[7] != 0) return 7;
  // workloads, save / load, corrupted load
  for (int kind : {P2Z_WORKLOAD_RAGGED, P2Z_WORKLOAD_PRECISION}) {
    for (int w : {16, 128}) {
      p2z_workload_spec sp{kind, w, 7, 11, 300, 12, 0, 0};
      p2z_workload* wl = nullptr;
      if (p2z_workload_create(&sp, &wl) != P2Z_OK) return 8;
      p2z_workload_info info;
      p2z_workload_get_info(wl, &info);
      if (p2z_check_blocks(p2z_workload_blocks(wl), info.nblocks, info.nblocks, info.nhit_blocks, info.nblocks)) return 9;
      if (p2z_workload_save(wl, path) != P2Z_OK) return 10;
      p2z_workload* w2 = nullptr;
      if (p2z_workload_load(path, &w2) != P2Z_OK) return 11;
      p2z_workload_destroy(w2);
      FILE* f = std::fopen(path, "r+b");
      std::fseek(f, 10, SEEK_SET); std::fputc('#', f); std::fclose(f);   // break the header
      if (p2z_workload_load(path, &w2) == P2Z_OK) return 12;
      f = std::fopen(path, "wb"); std::fputs("{}\n", f); std::fclose(f);   // empty header
      if (p2z_workload_load(path, &w2) == P2Z_OK) return 13;
      p2z_workload_destroy(wl);
    }
  }
  std::printf("host code clean under the sanitizers\n");
  return 0;
}
