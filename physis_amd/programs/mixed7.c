/*
 * A 3-D 7-point map weighted by 1-D coefficient grids along each axis,
 * written in the Physis DSL: the kernel of
 * tests/system_tests/test_cases/test_mixed-dim2.c:12-23 (operand order
 * kept).  physisc-hip runs it on the 2.5D kernel: the 3-D grid is staged,
 * the 1-D grids are read through bounds-checked gets.
 */
#include "physis/physis.h"

static void mixed7_kernel(const int x, const int y, const int z,
                          PSGrid3DFloat g, PSGrid3DFloat g2, PSGrid1DFloat i,
                          PSGrid1DFloat j, PSGrid1DFloat k) {
  float v = PSGridGet(g, x, y, z) +
      PSGridGet(g, x-1, y, z) * PSGridGet(i, x-1) +
      PSGridGet(g, x+1, y, z) * PSGridGet(i, x+1) +
      PSGridGet(g, x, y-1, z) * PSGridGet(j, y-1) +
      PSGridGet(g, x, y+1, z) * PSGridGet(j, y+1) +
      PSGridGet(g, x, y, z-1) * PSGridGet(k, z-1) +
      PSGridGet(g, x, y, z+1) * PSGridGet(k, z+1);
  PSGridEmit(g2, v);
}

/* One sweep g2 <- kernel(g) over the interior [1, n-1) of every axis. */
void mixed7_run(PSGrid3DFloat g, PSGrid3DFloat g2, PSGrid1DFloat cx,
                PSGrid1DFloat cy, PSGrid1DFloat cz) {
  PSDomain3D d = PSDomain3DNew(1, PSGridDim(g, 0) - 1, 1, PSGridDim(g, 1) - 1,
                               1, PSGridDim(g, 2) - 1);
  PSStencilRun(PSStencilMap(mixed7_kernel, d, g, g2, cx, cy, cz));
}
