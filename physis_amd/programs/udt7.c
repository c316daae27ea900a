/*
 * A 7-point stencil over a grid of user-defined points, written in the
 * Physis DSL: the point type and kernel shape of
 * tests/system_tests/test_cases/test_user-defined-type-7-pt.c:13-27 (read
 * member p of the seven points, emit member q of the centre), weighted as
 * the diffusion benchmark weighs its neighbours.  physisc-hip translates it
 * into udt7.hip.cpp; the translator splits the grid into its member planes
 * and runs the 2.5D kernel over them, which this program measures.
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

/* The same sum over whole points (the pattern of
 * test_user-defined-type-7-pt-periodic-complex.c:13-31: struct-valued gets,
 * a struct-valued emit): the point (p, q') with q' the weighted 7-point sum
 * of p, emitted into a second grid. */
static void udt7_whole_kernel(const int x, const int y, const int z,
                              PSGrid3DPoint g, PSGrid3DPoint g2, float ce,
                              float cw, float cn, float cs, float ct, float cb,
                              float cc) {
  struct Point c = PSGridGet(g, x, y, z);
  struct Point w = PSGridGet(g, x-1, y, z);
  struct Point e = PSGridGet(g, x+1, y, z);
  struct Point s = PSGridGet(g, x, y+1, z);
  struct Point n = PSGridGet(g, x, y-1, z);
  struct Point b = PSGridGet(g, x, y, z-1);
  struct Point t = PSGridGet(g, x, y, z+1);
  struct Point v;
  v.p = c.p;
  v.q = cc * c.p + cw * w.p + ce * e.p + cs * s.p + cn * n.p + cb * b.p +
        ct * t.p;
  PSGridEmit(g2, v);
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

/* One sweep g2 <- (p, 7-point(p)) over the interior, whole points. */
void udt7_whole_run(PSGrid3DPoint g, PSGrid3DPoint g2, float ce, float cw,
                    float cn, float cs, float ct, float cb, float cc) {
  PSDomain3D dom = PSDomain3DNew(1, PSGridDim(g, 0) - 1, 1,
                                 PSGridDim(g, 1) - 1, 1, PSGridDim(g, 2) - 1);
  PSStencilRun(PSStencilMap(udt7_whole_kernel, dom, g, g2, ce, cw, cn, cs,
                            ct, cb, cc));
}
