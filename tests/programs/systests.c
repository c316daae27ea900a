/*
 * Physis programs for the translator's system tests: each function runs one
 * of the reference's system-test scenarios (tests/system_tests/test_cases/
 * <name>.c) through the DSL and copies the result out, so that the test can
 * compare the dump with the reference's golden output.  The stencil bodies
 * keep the reference tests' operand order; the host code is restated as a
 * library function (no main, no printing).
 */
#include "physis/physis.h"

#define N 32

/* test_7-pt.c:12-20 */
static void k7(const int x, const int y, const int z, PSGrid3DFloat g,
               PSGrid3DFloat g2) {
  float v = PSGridGet(g, x, y, z) +
      PSGridGet(g, x+1, y, z) + PSGridGet(g, x-1, y, z) +
      PSGridGet(g, x, y+1, z) + PSGridGet(g, x, y-1, z) +
      PSGridGet(g, x, y, z-1) + PSGridGet(g, x, y, z+1);
  PSGridEmit(g2, v);
  return;
}

/* test_7-pt-periodic.c:12-24 */
static void k7p(const int x, const int y, const int z, PSGrid3DFloat g,
                PSGrid3DFloat g2) {
  float v =
      PSGridGetPeriodic(g, x, y, z) +
      PSGridGetPeriodic(g, x+1, y, z) +
      PSGridGetPeriodic(g, x-1, y, z) +
      PSGridGetPeriodic(g, x, y+1, z) +
      PSGridGetPeriodic(g, x, y-1, z) +
      PSGridGetPeriodic(g, x, y, z+1) +
      PSGridGetPeriodic(g, x, y, z-1);
  PSGridEmit(g2, v);
  return;
}

/* test_7-pt-neumann-cond.c:15-32 */
static void k7n(const int x, const int y, const int z, PSGrid3DFloat g1,
                PSGrid3DFloat g2) {
  int nx, ny, nz;
  nx = PSGridDim(g1, 0);
  ny = PSGridDim(g1, 1);
  nz = PSGridDim(g1, 2);
  float c, w, e, n, s, b, t;
  c = PSGridGet(g1, x, y, z);
  if (x == 0)    w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x-1, y, z);
  if (x == nx-1) e = PSGridGet(g1, x, y, z); else e = PSGridGet(g1, x+1, y, z);
  if (y == 0)    n = PSGridGet(g1, x, y, z); else n = PSGridGet(g1, x, y-1, z);
  if (y == ny-1) s = PSGridGet(g1, x, y, z); else s = PSGridGet(g1, x, y+1, z);
  if (z == 0)    b = PSGridGet(g1, x, y, z); else b = PSGridGet(g1, x, y, z-1);
  if (z == nz-1) t = PSGridGet(g1, x, y, z); else t = PSGridGet(g1, x, y, z+1);
  PSGridEmit(g2, c + w + e + s + n + b + t);
  return;
}

/* test_asymmetric.c:12-18 */
static void kasym(const int x, const int y, const int z, PSGrid3DFloat g,
                  PSGrid3DFloat g2) {
  float v = PSGridGet(g, x, y, z) +
      PSGridGet(g, x+1, y+1, z+1) + PSGridGet(g, x-2, y-2, z-2);
  PSGridEmit(g2, v);
  return;
}

/* test_asymmetric-periodic.c:12-18 */
static void kasymp(const int x, const int y, const int z, PSGrid3DFloat g,
                   PSGrid3DFloat g2) {
  float v = PSGridGet(g, x, y, z) +
      PSGridGet(g, x-1, y, z) + PSGridGetPeriodic(g, x+2, y, z);
  PSGridEmit(g2, v);
  return;
}

/* test_27-pt.c:10-32 */
static void k27(const int x, const int y, const int z, PSGrid3DFloat g,
                PSGrid3DFloat g2) {
  float v =
      PSGridGet(g, x, y, z-1) + PSGridGet(g, x+1, y, z-1) +
      PSGridGet(g, x-1, y, z-1) + PSGridGet(g, x, y+1, z-1) +
      PSGridGet(g, x+1, y+1, z-1) + PSGridGet(g, x-1, y+1, z-1) +
      PSGridGet(g, x, y-1, z-1) + PSGridGet(g, x+1, y-1, z-1) +
      PSGridGet(g, x-1, y-1, z-1) +
      PSGridGet(g, x, y, z) + PSGridGet(g, x+1, y, z) +
      PSGridGet(g, x-1, y, z) + PSGridGet(g, x, y+1, z) +
      PSGridGet(g, x+1, y+1, z) + PSGridGet(g, x-1, y+1, z) +
      PSGridGet(g, x, y-1, z) + PSGridGet(g, x+1, y-1, z) +
      PSGridGet(g, x-1, y-1, z) +
      PSGridGet(g, x, y, z+1) + PSGridGet(g, x+1, y, z+1) +
      PSGridGet(g, x-1, y, z+1) + PSGridGet(g, x, y+1, z+1) +
      PSGridGet(g, x+1, y+1, z+1) + PSGridGet(g, x-1, y+1, z+1) +
      PSGridGet(g, x, y-1, z+1) + PSGridGet(g, x+1, y-1, z+1) +
      PSGridGet(g, x-1, y-1, z+1);
  PSGridEmit(g2, v);
  return;
}

/* test_27-pt-periodic.c:13-32 */
static void k27p(const int x, const int y, const int z, PSGrid3DFloat g,
                 PSGrid3DFloat g2) {
  float v =
      PSGridGetPeriodic(g,x,y,z-1) + PSGridGetPeriodic(g,x+1,y,z-1) +
      PSGridGetPeriodic(g,x-1,y,z-1) + PSGridGetPeriodic(g,x,y+1,z-1) +
      PSGridGetPeriodic(g,x+1,y+1,z-1) + PSGridGetPeriodic(g,x-1,y+1,z-1) +
      PSGridGetPeriodic(g,x,y-1,z-1) + PSGridGetPeriodic(g,x+1,y-1,z-1) +
      PSGridGetPeriodic(g,x-1,y-1,z-1) +
      PSGridGetPeriodic(g,x,y,z) + PSGridGetPeriodic(g,x+1,y,z) +
      PSGridGetPeriodic(g,x-1,y,z) + PSGridGetPeriodic(g,x,y+1,z) +
      PSGridGetPeriodic(g,x+1,y+1,z) + PSGridGetPeriodic(g,x-1,y+1,z) +
      PSGridGetPeriodic(g,x,y-1,z) + PSGridGetPeriodic(g,x+1,y-1,z) +
      PSGridGetPeriodic(g,x-1,y-1,z) +
      PSGridGetPeriodic(g,x,y,z+1) + PSGridGetPeriodic(g,x+1,y,z+1) +
      PSGridGetPeriodic(g,x-1,y,z+1) + PSGridGetPeriodic(g,x,y+1,z+1) +
      PSGridGetPeriodic(g,x+1,y+1,z+1) + PSGridGetPeriodic(g,x-1,y+1,z+1) +
      PSGridGetPeriodic(g,x,y-1,z+1) + PSGridGetPeriodic(g,x+1,y-1,z+1) +
      PSGridGetPeriodic(g,x-1,y-1,z+1);
  PSGridEmit(g2, v);
  return;
}

/* test_27-pt-reduction.c:14-40: int grids, the sum goes through a float */
static void k27i(const int x, const int y, const int z, PSGrid3DInt g,
                 PSGrid3DInt g2) {
  float v =
      PSGridGet(g, x, y, z-1) + PSGridGet(g, x+1, y, z-1) +
      PSGridGet(g, x-1, y, z-1) + PSGridGet(g, x, y+1, z-1) +
      PSGridGet(g, x+1, y+1, z-1) + PSGridGet(g, x-1, y+1, z-1) +
      PSGridGet(g, x, y-1, z-1) + PSGridGet(g, x+1, y-1, z-1) +
      PSGridGet(g, x-1, y-1, z-1) +
      PSGridGet(g, x, y, z) + PSGridGet(g, x+1, y, z) +
      PSGridGet(g, x-1, y, z) + PSGridGet(g, x, y+1, z) +
      PSGridGet(g, x+1, y+1, z) + PSGridGet(g, x-1, y+1, z) +
      PSGridGet(g, x, y-1, z) + PSGridGet(g, x+1, y-1, z) +
      PSGridGet(g, x-1, y-1, z) +
      PSGridGet(g, x, y, z+1) + PSGridGet(g, x+1, y, z+1) +
      PSGridGet(g, x-1, y, z+1) + PSGridGet(g, x, y+1, z+1) +
      PSGridGet(g, x+1, y+1, z+1) + PSGridGet(g, x-1, y+1, z+1) +
      PSGridGet(g, x, y-1, z+1) + PSGridGet(g, x+1, y-1, z+1) +
      PSGridGet(g, x-1, y-1, z+1);
  PSGridEmit(g2, v);
  return;
}

/* Runs one map of `which` over dom on n^3 float grids initialised to i. */
static void run_one(int which, int n, int lo, int hi, float *out) {
  PSGrid3DFloat g1 = PSGrid3DFloatNew(n, n, n);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(n, n, n);
  PSDomain3D d = PSDomain3DNew(lo, hi, lo, hi, lo, hi);
  size_t nelms = n * n * n;
  float *indata = (float *)malloc(sizeof(float) * nelms);
  size_t i;
  for (i = 0; i < nelms; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSGridCopyin(g2, indata);
  if (which == 0) PSStencilRun(PSStencilMap(k7, d, g1, g2));
  if (which == 1) PSStencilRun(PSStencilMap(k7p, d, g1, g2));
  if (which == 2) PSStencilRun(PSStencilMap(kasym, d, g1, g2));
  if (which == 3) PSStencilRun(PSStencilMap(kasymp, d, g1, g2));
  if (which == 4) PSStencilRun(PSStencilMap(k27, d, g1, g2));
  if (which == 5) PSStencilRun(PSStencilMap(k27p, d, g1, g2));
  PSGridCopyout(g2, out);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

void systest_7pt(float *out) { run_one(0, N, 1, N - 1, out); }
void systest_7pt_periodic(float *out) { run_one(1, N, 0, N, out); }
void systest_asymmetric(float *out) { run_one(2, N, 2, N - 2, out); }
void systest_asymmetric_periodic(float *out) { run_one(3, N, 2, N - 2, out); }
void systest_27pt(float *out) { run_one(4, N, 1, N - 1, out); }
void systest_27pt_periodic(float *out) { run_one(5, N, 0, N, out); }

/* test_7-pt-neumann-cond.c main: N = 8, g2 starts zeroed. */
void systest_7pt_neumann(float *out) {
  int n = 8;
  PSGrid3DFloat g1 = PSGrid3DFloatNew(n, n, n);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(n, n, n);
  PSDomain3D d = PSDomain3DNew(0, n, 0, n, 0, n);
  float *indata = (float *)malloc(sizeof(float) * n * n * n);
  int i;
  for (i = 0; i < n * n * n; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSStencilRun(PSStencilMap(k7n, d, g1, g2));
  PSGridCopyout(g2, out);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

/* test_7-pt-multi-iterations.c main: ITER = 5 ping-pong iterations. */
void systest_7pt_multi_iterations(float *out) {
  PSGrid3DFloat g1 = PSGrid3DFloatNew(N, N, N);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(N, N, N);
  PSDomain3D d = PSDomain3DNew(1, N - 1, 1, N - 1, 1, N - 1);
  float *indata = (float *)malloc(sizeof(float) * N * N * N);
  int i;
  for (i = 0; i < N * N * N; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSGridCopyin(g2, indata);
  PSStencilRun(PSStencilMap(k7, d, g1, g2), PSStencilMap(k7, d, g2, g1), 5);
  PSGridCopyout(g1, out);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

/* test_27-pt-reduction.c main: two int sums after forward and back maps. */
void systest_27pt_reduction(int *v1, int *v2) {
  int n = 16;
  PSGrid3DInt g1 = PSGrid3DIntNew(n, n, n);
  PSGrid3DInt g2 = PSGrid3DIntNew(n, n, n);
  PSDomain3D d = PSDomain3DNew(1, n - 1, 1, n - 1, 1, n - 1);
  int *indata = (int *)malloc(sizeof(int) * n * n * n);
  int i;
  for (i = 0; i < n * n * n; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSGridCopyin(g2, indata);
  PSStencilRun(PSStencilMap(k27i, d, g1, g2));
  PSReduce(v1, PS_SUM, g2);
  PSStencilRun(PSStencilMap(k27i, d, g2, g1));
  PSReduce(v2, PS_SUM, g1);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

/* test_reduction-3d-sum.c main: N = 16, value i. */
void systest_reduction_3d_sum(float *v) {
  int n = 16;
  PSGrid3DFloat g1 = PSGrid3DFloatNew(n, n, n);
  float *indata = (float *)malloc(sizeof(float) * n * n * n);
  int i;
  for (i = 0; i < n * n * n; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSReduce(v, PS_SUM, g1);
  PSGridFree(g1);
  free(indata);
}
