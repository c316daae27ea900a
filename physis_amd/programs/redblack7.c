/*
 * Red-black Gauss-Seidel relaxation of the 7-point diffusion operator, in
 * place, written in the Physis DSL (the red-black map of
 * include/physis/physis_user.h:100-105 with the in-place 7-point kernel of
 * tests/system_tests/test_cases/test_redblack.c:13-21, weighted as the
 * diffusion benchmark weighs its neighbours).  physisc-hip translates it
 * into redblack7.hip.cpp.
 *
 * Every neighbour of a point is of the other colour, so a colour sweep
 * reads only values that sweep does not write: the result does not depend
 * on the order of the updates, and the REFERENCE target's loop order gives
 * the same bits.
 */
#include "physis/physis.h"

static void redblack7_kernel(const int x, const int y, const int z,
                             PSGrid3DFloat g, float ce, float cw, float cn,
                             float cs, float ct, float cb, float cc) {
  float c, w, e, n, s, b, t;
  c = PSGridGet(g, x, y, z);
  w = PSGridGet(g, x-1, y, z);
  e = PSGridGet(g, x+1, y, z);
  n = PSGridGet(g, x, y-1, z);
  s = PSGridGet(g, x, y+1, z);
  b = PSGridGet(g, x, y, z-1);
  t = PSGridGet(g, x, y, z+1);
  PSGridEmit(g, cc*c + cw*w + ce*e + cs*s + cn*n + cb*b + ct*t);
}

/* count iterations, each a red then a black sweep over the interior
 * [1, n-1) of every axis; the faces keep their values. */
void redblack7_run(PSGrid3DFloat g, int count, float ce, float cw, float cn,
                   float cs, float ct, float cb, float cc) {
  PSDomain3D dom = PSDomain3DNew(1, PSGridDim(g, 0) - 1, 1,
                                 PSGridDim(g, 1) - 1, 1, PSGridDim(g, 2) - 1);
  PSStencilRun(PSStencilMapRedBlack(redblack7_kernel, dom, g,
                                    ce, cw, cn, cs, ct, cb, cc), count);
}
