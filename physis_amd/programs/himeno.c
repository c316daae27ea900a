/*
 * Himeno 19-point variable-coefficient Jacobi sweep for the pressure Poisson
 * equation, written in the Physis DSL; physisc-hip translates it into
 * himeno.hip.cpp.
 *
 * The stencil body keeps the operand order of the reference's Physis
 * version (examples/himeno/himenobmtxpa_physis.c:331-361), which is the
 * original benchmark's arithmetic with the axes relabelled (x = the
 * original's fastest index).  The kernel also returns ss*ss, the point's
 * contribution to the residual gosa of the original benchmark
 * (himenobmtxpa_original.c:334), which himeno_run_gosa reduces per sweep.
 */
#include "physis/physis.h"

static float jacobi_kernel(int i, int j, int k,
                           PSGrid3DFloat p0, PSGrid3DFloat p1,
                           PSGrid3DFloat a0, PSGrid3DFloat a1, PSGrid3DFloat a2,
                           PSGrid3DFloat a3, PSGrid3DFloat b0, PSGrid3DFloat b1,
                           PSGrid3DFloat b2, PSGrid3DFloat c0, PSGrid3DFloat c1,
                           PSGrid3DFloat c2, PSGrid3DFloat bnd,
                           PSGrid3DFloat wrk1, float omega) {
  float s0, ss;
  s0 = PSGridGet(a0, i, j, k) * PSGridGet(p0, i, j, k+1)
      + PSGridGet(a1, i, j, k) * PSGridGet(p0, i, j+1, k)
      + PSGridGet(a2, i, j, k) * PSGridGet(p0, i+1, j, k)
      + PSGridGet(b0, i, j, k)
      * (PSGridGet(p0, i, j+1, k+1) - PSGridGet(p0, i, j-1, k+1)
         - PSGridGet(p0, i, j+1, k-1) + PSGridGet(p0, i, j-1, k-1))
      + PSGridGet(b1, i, j, k)
      * (PSGridGet(p0, i+1, j+1, k) - PSGridGet(p0, i+1, j-1, k)
         - PSGridGet(p0, i-1, j+1, k) + PSGridGet(p0, i-1, j-1, k))
      + PSGridGet(b2, i, j, k)
      * (PSGridGet(p0, i+1, j, k+1) - PSGridGet(p0, i+1, j, k-1)
         - PSGridGet(p0, i-1, j, k+1) + PSGridGet(p0, i-1, j, k-1))
      + PSGridGet(c0, i, j, k) * PSGridGet(p0, i, j, k-1)
      + PSGridGet(c1, i, j, k) * PSGridGet(p0, i, j-1, k)
      + PSGridGet(c2, i, j, k) * PSGridGet(p0, i-1, j, k)
      + PSGridGet(wrk1, i, j, k);
  ss = (s0 * PSGridGet(a3, i, j, k) - PSGridGet(p0, i, j, k))
      * PSGridGet(bnd, i, j, k);
  float v = PSGridGet(p0, i, j, k) + omega * ss;
  PSGridEmit(p1, v);
  return ss * ss;
}

/* nn Jacobi sweeps over the inner domain [1, n-1)^3, ping-ponging p0 -> p1
 * -> p0 (nn even: result in p0). */
void himeno_run(int nn, PSGrid3DFloat p0, PSGrid3DFloat p1,
                PSGrid3DFloat a0, PSGrid3DFloat a1, PSGrid3DFloat a2,
                PSGrid3DFloat a3, PSGrid3DFloat b0, PSGrid3DFloat b1,
                PSGrid3DFloat b2, PSGrid3DFloat c0, PSGrid3DFloat c1,
                PSGrid3DFloat c2, PSGrid3DFloat bnd, PSGrid3DFloat wrk1,
                float omega) {
  PSDomain3D inner = PSDomain3DNew(1, PSGridDim(p0, 0) - 1,
                                   1, PSGridDim(p0, 1) - 1,
                                   1, PSGridDim(p0, 2) - 1);
  PSStencilRun(PSStencilMap(jacobi_kernel, inner, p0, p1, a0, a1, a2, a3,
                            b0, b1, b2, c0, c1, c2, bnd, wrk1, omega),
               PSStencilMap(jacobi_kernel, inner, p1, p0, a0, a1, a2, a3,
                            b0, b1, b2, c0, c1, c2, bnd, wrk1, omega),
               nn / 2);
}

/* As himeno_run, also storing each sweep's residual sum of ss*ss:
 * gosa_even[i] after sweep 2i, gosa_odd[i] after sweep 2i+1. */
void himeno_run_gosa(int nn, PSGrid3DFloat p0, PSGrid3DFloat p1,
                     PSGrid3DFloat a0, PSGrid3DFloat a1, PSGrid3DFloat a2,
                     PSGrid3DFloat a3, PSGrid3DFloat b0, PSGrid3DFloat b1,
                     PSGrid3DFloat b2, PSGrid3DFloat c0, PSGrid3DFloat c1,
                     PSGrid3DFloat c2, PSGrid3DFloat bnd, PSGrid3DFloat wrk1,
                     float omega, float *gosa_even, float *gosa_odd) {
  PSDomain3D inner = PSDomain3DNew(1, PSGridDim(p0, 0) - 1,
                                   1, PSGridDim(p0, 1) - 1,
                                   1, PSGridDim(p0, 2) - 1);
  PSStencilRun(PSStencilMapReduce(gosa_even, PS_SUM, jacobi_kernel, inner,
                                  p0, p1, a0, a1, a2, a3, b0, b1, b2,
                                  c0, c1, c2, bnd, wrk1, omega),
               PSStencilMapReduce(gosa_odd, PS_SUM, jacobi_kernel, inner,
                                  p1, p0, a0, a1, a2, a3, b0, b1, b2,
                                  c0, c1, c2, bnd, wrk1, omega),
               nn / 2);
}
