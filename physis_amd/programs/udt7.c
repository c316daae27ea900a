/*
 * A 7-point stencil over a grid of user-defined points, written in the
 * Physis DSL: the point type and kernel shape of
 * tests/system_tests/test_cases/test_user-defined-type-7-pt.c:13-27 (read
 * member p of the seven points, emit member q of the centre), weighted as
 * the diffusion benchmark weighs its neighbours.  physisc-hip translates it
 * into udt7.hip.cpp; the run kernel is the generic one (one thread per
 * column), which this program measures.
 */
#include "physis/physis.h"

struct Point {
  float p;
  float q;
};

DeclareGrid3D(Point, struct Point);

static void udt7_kernel(const int x, const int y, const int z,
                        PSGrid3DPoint g, float ce, float cw, float cn,
                        float cs, float ct, float cb, float cc) {
  float v = cc * PSGridGet(g, x, y, z).p + cw * PSGridGet(g, x-1, y, z).p +
            ce * PSGridGet(g, x+1, y, z).p + cs * PSGridGet(g, x, y+1, z).p +
            cn * PSGridGet(g, x, y-1, z).p + cb * PSGridGet(g, x, y, z-1).p +
            ct * PSGridGet(g, x, y, z+1).p;
  PSGridEmitUtype(g.q, v);
}

PSGrid3DPoint udt7_new(int nx, int ny, int nz) {
  return PSGrid3DPointNew(nx, ny, nz);
}

void udt7_copyin(PSGrid3DPoint g, struct Point *in) { PSGridCopyin(g, in); }

void udt7_copyout(PSGrid3DPoint g, struct Point *out) {
  PSGridCopyout(g, out);
}

void udt7_free(PSGrid3DPoint g) { PSGridFree(g); }

/* count sweeps q <- 7-point(p) over the interior [1, n-1) of every axis. */
void udt7_run(PSGrid3DPoint g, int count, float ce, float cw, float cn,
              float cs, float ct, float cb, float cc) {
  PSDomain3D dom = PSDomain3DNew(1, PSGridDim(g, 0) - 1, 1,
                                 PSGridDim(g, 1) - 1, 1, PSGridDim(g, 2) - 1);
  PSStencilRun(PSStencilMap(udt7_kernel, dom, g, ce, cw, cn, cs, ct, cb, cc),
               count);
}
