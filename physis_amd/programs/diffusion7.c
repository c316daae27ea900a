/*
 * 7-point diffusion with clamped (edge-value) boundaries, written in the
 * Physis DSL.  physisc-hip translates it into diffusion7.hip.cpp.
 *
 * The stencil body keeps the operand order of the reference benchmark
 * kernel (examples/diffusion-benchmark/diffusion3d_physis.c:29-58) so that
 * the HIP result is bit-identical to the REFERENCE target's.
 */
#include "physis/physis.h"

static void diffusion7_kernel(const int x, const int y, const int z,
                              PSGrid3DFloat g1, PSGrid3DFloat g2,
                              float ce, float cw, float cn, float cs,
                              float ct, float cb, float cc) {
  int nx, ny, nz;
  nx = PSGridDim(g1, 0);
  ny = PSGridDim(g1, 1);
  nz = PSGridDim(g1, 2);

  float c, w, e, n, s, b, t;
  c = PSGridGet(g1, x, y, z);
  if (x == 0) w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x-1, y, z);
  if (x == nx-1) e = PSGridGet(g1, x, y, z); else e = PSGridGet(g1, x+1, y, z);
  if (y == 0) n = PSGridGet(g1, x, y, z); else n = PSGridGet(g1, x, y-1, z);
  if (y == ny-1) s = PSGridGet(g1, x, y, z); else s = PSGridGet(g1, x, y+1, z);
  if (z == 0) b = PSGridGet(g1, x, y, z); else b = PSGridGet(g1, x, y, z-1);
  if (z == nz-1) t = PSGridGet(g1, x, y, z); else t = PSGridGet(g1, x, y, z+1);
  PSGridEmit(g2, cc*c + cw*w + ce*e + cs*s + cn*n + cb*b + ct*t);
}

/* The same diffusion with periodic wrap on every axis: every neighbour is
 * read with PSGridGetPeriodic (wrap semantics of
 * tests/system_tests/test_cases/test_7-pt-periodic.c:12-24), operand order
 * as above. */
static void diffusion7_periodic_kernel(const int x, const int y, const int z,
                                       PSGrid3DFloat g1, PSGrid3DFloat g2,
                                       float ce, float cw, float cn, float cs,
                                       float ct, float cb, float cc) {
  float c, w, e, n, s, b, t;
  c = PSGridGet(g1, x, y, z);
  w = PSGridGetPeriodic(g1, x-1, y, z);
  e = PSGridGetPeriodic(g1, x+1, y, z);
  n = PSGridGetPeriodic(g1, x, y-1, z);
  s = PSGridGetPeriodic(g1, x, y+1, z);
  b = PSGridGetPeriodic(g1, x, y, z-1);
  t = PSGridGetPeriodic(g1, x, y, z+1);
  PSGridEmit(g2, cc*c + cw*w + ce*e + cs*s + cn*n + cb*b + ct*t);
}

/* count sweeps over the whole grid, ping-ponging g1 -> g2 -> g1; count is
 * even, so the result ends in g1. */
void diffusion7_run(PSGrid3DFloat g1, PSGrid3DFloat g2, int count,
                    float ce, float cw, float cn, float cs,
                    float ct, float cb, float cc) {
  PSDomain3D dom = PSDomain3DNew(0, PSGridDim(g1, 0), 0, PSGridDim(g1, 1),
                                 0, PSGridDim(g1, 2));
  PSStencilRun(PSStencilMap(diffusion7_kernel, dom, g1, g2,
                            ce, cw, cn, cs, ct, cb, cc),
               PSStencilMap(diffusion7_kernel, dom, g2, g1,
                            ce, cw, cn, cs, ct, cb, cc),
               count / 2);
}

void diffusion7_periodic_run(PSGrid3DFloat g1, PSGrid3DFloat g2, int count,
                             float ce, float cw, float cn, float cs,
                             float ct, float cb, float cc) {
  PSDomain3D dom = PSDomain3DNew(0, PSGridDim(g1, 0), 0, PSGridDim(g1, 1),
                                 0, PSGridDim(g1, 2));
  PSStencilRun(PSStencilMap(diffusion7_periodic_kernel, dom, g1, g2,
                            ce, cw, cn, cs, ct, cb, cc),
               PSStencilMap(diffusion7_periodic_kernel, dom, g2, g1,
                            ce, cw, cn, cs, ct, cb, cc),
               count / 2);
}

/* One sweep g1 -> g2 over an arbitrary sub-domain (used to check domain
 * handling: points outside the domain keep g2's previous contents). */
void diffusion7_sweep_domain(PSGrid3DFloat g1, PSGrid3DFloat g2,
                             int x0, int x1, int y0, int y1, int z0, int z1,
                             int periodic, float ce, float cw, float cn,
                             float cs, float ct, float cb, float cc) {
  PSDomain3D dom = PSDomain3DNew(x0, x1, y0, y1, z0, z1);
  if (periodic)
    PSStencilRun(PSStencilMap(diffusion7_periodic_kernel, dom, g1, g2,
                              ce, cw, cn, cs, ct, cb, cc));
  else
    PSStencilRun(PSStencilMap(diffusion7_kernel, dom, g1, g2,
                              ce, cw, cn, cs, ct, cb, cc));
}
