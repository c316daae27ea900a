// Host-only check of the run-kernel launch planner (include/physis/hip/run.h):
// prints, for each case, the z chunk and window bytes that PlanRunFor picks.
// Built and run by tests/test_planner.py; no device is touched.
#include <cstdio>
#include <cstdlib>

#include "physis/hip/run.h"

namespace {

__PSGrid MakeGrid(int nx, int ny, int nz, int elm) {
  __PSGrid g = {};
  g.num_dims = 3;
  g.dim[0] = nx;
  g.dim[1] = ny;
  g.dim[2] = nz;
  g.elm_size = elm;
  g.num_elms = (int64_t)nx * ny * nz;
  return g;
}

// One launch: tiles of tx x ty, `resident` blocks per wave of the device,
// windows of `margin` planes per side.  Prints
//   nx ny nz elm margin tiles_z zchunk blocks window_bytes
void Case(int nx, int ny, int nz, int elm, int margin, int tx, int ty,
          long resident) {
  __PSGrid g = MakeGrid(nx, ny, nz, elm);
  __PSDomain dom = {};
  for (int d = 0; d < 3; ++d) {
    dom.local_min[d] = dom.local_max[d] = 0;
  }
  dom.local_max[0] = nx;
  dom.local_max[1] = ny;
  dom.local_max[2] = nz;
  const int mc = physis_hip::WindowChunkLimit({&g, &g}, margin);
  const physis_hip::RunShape s =
      physis_hip::PlanRunFor(dom, tx, ty, resident, mc);
  const long long window =
      (long long)(s.zchunk + 2 * margin) * nx * ny * (long long)elm;
  std::printf("%d %d %d %d %d %d %d %d %lld\n", nx, ny, nz, elm, margin,
              s.tiles_z, s.zchunk, s.blocks, window);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc > 1) {
    // one case from the command line (the rejection path exits here)
    if (argc != 9) return 2;
    Case(std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3]),
         std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6]),
         std::atoi(argv[7]), std::atol(argv[8]));
    return 0;
  }
  // shapes whose single-chunk windows exceed 2 GiB, and ordinary ones
  Case(8192, 8192, 16, 4, 2, 256, 14, 256);
  Case(4096, 4096, 64, 4, 2, 256, 14, 256);
  Case(4096, 4096, 64, 4, 3, 256, 12, 256);
  Case(2048, 2048, 255, 8, 2, 256, 14, 256);
  Case(512, 512, 512, 4, 2, 256, 14, 256);
  Case(512, 256, 256, 4, 2, 256, 6, 256);
  return 0;
}
