/*
 * 2-D 5- and 9-point stencils, written in the Physis DSL: the kernels of
 * tests/system_tests/test_cases/test_5-pt-2d.c:12-20 and test_9-pt-2d.c:
 * 12-24 (the points summed in the same order), scaled by w so that repeated
 * sweeps stay bounded, applied as ping-pong map pairs over the interior
 * [1, nx-1) x [1, ny-1).  physisc-hip translates it into stencil2d.hip.cpp;
 * its 2-D maps run on the 2.5D kernel over the grids' rows view.
 */
#include "physis/physis.h"

static void s2d5(const int x, const int y, PSGrid2DFloat g, PSGrid2DFloat g2,
                 float w) {
  float v = PSGridGet(g, x, y) + PSGridGet(g, x-1, y) + PSGridGet(g, x+1, y) +
            PSGridGet(g, x, y-1) + PSGridGet(g, x, y+1);
  PSGridEmit(g2, v * w);
}

static void s2d9(const int x, const int y, PSGrid2DFloat g, PSGrid2DFloat g2,
                 float w) {
  float v = PSGridGet(g, x, y) + PSGridGet(g, x-1, y) + PSGridGet(g, x+1, y) +
            PSGridGet(g, x, y-1) + PSGridGet(g, x, y+1) +
            PSGridGet(g, x-1, y-1) + PSGridGet(g, x+1, y-1) +
            PSGridGet(g, x-1, y+1) + PSGridGet(g, x+1, y+1);
  PSGridEmit(g2, v * w);
}

/* count pairs of sweeps a -> b -> a of the 5-point (points = 5) or 9-point
 * stencil. */
void stencil2d_run(PSGrid2DFloat a, PSGrid2DFloat b, int points, int count,
                   float w) {
  PSDomain2D d = PSDomain2DNew(1, PSGridDim(a, 0) - 1, 1, PSGridDim(a, 1) - 1);
  if (points == 5)
    PSStencilRun(PSStencilMap(s2d5, d, a, b, w), PSStencilMap(s2d5, d, b, a, w),
                 count);
  else
    PSStencilRun(PSStencilMap(s2d9, d, a, b, w), PSStencilMap(s2d9, d, b, a, w),
                 count);
}

/* One sweep a -> b. */
void stencil2d_sweep(PSGrid2DFloat a, PSGrid2DFloat b, int points, float w) {
  PSDomain2D d = PSDomain2DNew(1, PSGridDim(a, 0) - 1, 1, PSGridDim(a, 1) - 1);
  if (points == 5)
    PSStencilRun(PSStencilMap(s2d5, d, a, b, w));
  else
    PSStencilRun(PSStencilMap(s2d9, d, a, b, w));
}
