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

/* ---- red-black, 1-D, 2-D, mixed-dimension and hole stencils */

/* test_redblack.c:13-21: the 7-point sum written back in place. */
static void krb(const int x, const int y, const int z, PSGrid3DFloat g) {
  float v = PSGridGet(g, x, y, z) +
      PSGridGet(g, x+1, y, z) + PSGridGet(g, x-1, y, z) +
      PSGridGet(g, x, y+1, z) + PSGridGet(g, x, y-1, z) +
      PSGridGet(g, x, y, z-1) + PSGridGet(g, x, y, z+1);
  PSGridEmit(g, v);
  return;
}

/* test_redblack-periodic.c:13-21 */
static void krbp(const int x, const int y, const int z, PSGrid3DFloat g) {
  float v = PSGridGetPeriodic(g, x, y, z) +
      PSGridGetPeriodic(g, x+1, y, z) + PSGridGetPeriodic(g, x-1, y, z) +
      PSGridGetPeriodic(g, x, y+1, z) + PSGridGetPeriodic(g, x, y-1, z) +
      PSGridGetPeriodic(g, x, y, z-1) + PSGridGetPeriodic(g, x, y, z+1);
  PSGridEmit(g, v);
  return;
}

/* One grid of N^3 initialised to i, swept by a red-black map (mode 0: both
 * colours in one map; 1: red and black as two maps of one run; 2: periodic
 * over the whole grid). */
static void run_redblack(int mode, float *out) {
  PSGrid3DFloat g = PSGrid3DFloatNew(N, N, N);
  PSDomain3D d = PSDomain3DNew(1, N - 1, 1, N - 1, 1, N - 1);
  PSDomain3D dall = PSDomain3DNew(0, N, 0, N, 0, N);
  float *indata = (float *)malloc(sizeof(float) * N * N * N);
  int i;
  for (i = 0; i < N * N * N; i++) indata[i] = i;
  PSGridCopyin(g, indata);
  if (mode == 0) PSStencilRun(PSStencilMapRedBlack(krb, d, g));
  if (mode == 1)
    PSStencilRun(PSStencilMapRed(krb, d, g), PSStencilMapBlack(krb, d, g));
  if (mode == 2) PSStencilRun(PSStencilMapRedBlack(krbp, dall, g));
  PSGridCopyout(g, out);
  PSGridFree(g);
  free(indata);
}

void systest_redblack(float *out) { run_redblack(0, out); }
void systest_redblack_separated(float *out) { run_redblack(1, out); }
void systest_redblack_periodic(float *out) { run_redblack(2, out); }

/* test_3-pt-1d.c:12-19 */
static void k3(const int x, PSGrid1DFloat g, PSGrid1DFloat g2) {
  float v = PSGridGet(g, x-1) + PSGridGet(g, x) + PSGridGet(g, x+1);
  PSGridEmit(g2, v);
  return;
}

/* test_5-pt-2d.c:12-21 */
static void k5(const int x, const int y, PSGrid2DFloat g, PSGrid2DFloat g2) {
  float v = PSGridGet(g, x, y) + PSGridGet(g, x-1, y) + PSGridGet(g, x+1, y) +
      PSGridGet(g, x, y-1) + PSGridGet(g, x, y+1);
  PSGridEmit(g2, v);
  return;
}

/* test_3-pt-1d.c main: N = 1024, domain [1, N-1). */
void systest_3pt_1d(float *out) {
  int n = 1024, i;
  PSGrid1DFloat g1 = PSGrid1DFloatNew(n);
  PSGrid1DFloat g2 = PSGrid1DFloatNew(n);
  PSDomain1D d = PSDomain1DNew(1, n - 1);
  float *indata = (float *)malloc(sizeof(float) * n);
  for (i = 0; i < n; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSGridCopyin(g2, indata);
  PSStencilRun(PSStencilMap(k3, d, g1, g2));
  PSGridCopyout(g2, out);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

/* test_5-pt-2d.c main: N x N, domain [1, N-1)^2. */
void systest_5pt_2d(float *out) {
  int i;
  PSGrid2DFloat g1 = PSGrid2DFloatNew(N, N);
  PSGrid2DFloat g2 = PSGrid2DFloatNew(N, N);
  PSDomain2D d = PSDomain2DNew(1, N - 1, 1, N - 1);
  float *indata = (float *)malloc(sizeof(float) * N * N);
  for (i = 0; i < N * N; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSGridCopyin(g2, indata);
  PSStencilRun(PSStencilMap(k5, d, g1, g2));
  PSGridCopyout(g2, out);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

/* test_mixed-dim3.c:10-20: a 3-D stencil weighted by a 1-D grid along x and
 * a 2-D grid over (y, z). */
static void kmix(const int x, const int y, const int z, PSGrid3DFloat g,
                 PSGrid3DFloat g2, PSGrid1DFloat i, PSGrid2DFloat j) {
  float v = PSGridGet(g, x, y, z) +
      PSGridGet(g, x-1, y, z) * PSGridGet(i, x-1) +
      PSGridGet(g, x+1, y, z) * PSGridGet(i, x+1) +
      PSGridGet(g, x, y-1, z) * PSGridGet(j, y-1, z) +
      PSGridGet(g, x, y+1, z) * PSGridGet(j, y+1, z) +
      PSGridGet(g, x, y, z-1) * PSGridGet(j, y, z-1) +
      PSGridGet(g, x, y, z+1) * PSGridGet(j, y, z+1);
  PSGridEmit(g2, v);
  return;
}

/* test_mixed-dim3.c main: grids N x (N+2) x N, weights 1 or 2 alternating. */
void systest_mixed_dim3(float *out) {
  int ny = N + 2, i;
  size_t nelms = (size_t)N * ny * N;
  PSGrid3DFloat g1 = PSGrid3DFloatNew(N, ny, N);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(N, ny, N);
  PSGrid1DFloat cx = PSGrid1DFloatNew(N);
  PSGrid2DFloat cy = PSGrid2DFloatNew(ny, N);
  PSDomain3D d = PSDomain3DNew(1, N - 1, 1, ny - 1, 1, N - 1);
  float *indata = (float *)malloc(sizeof(float) * nelms);
  for (i = 0; i < (int)nelms; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSGridCopyin(g2, indata);
  for (i = 0; i < N; i++) indata[i] = 1 + (i % 2);
  PSGridCopyin(cx, indata);
  for (i = 0; i < ny * N; i++) indata[i] = 1 + (i % 2);
  PSGridCopyin(cy, indata);
  PSStencilRun(PSStencilMap(kmix, d, g1, g2, cx, cy));
  PSGridCopyout(g2, out);
  PSGridFree(g1);
  PSGridFree(g2);
  PSGridFree(cx);
  PSGridFree(cy);
  free(indata);
}

/* test_stencil-hole.c:12-18: x offsets of +-2 with no +-1 in between. */
static void khole(const int x, const int y, const int z, PSGrid3DFloat g,
                  PSGrid3DFloat g2) {
  float v = PSGridGet(g, x, y, z) +
      PSGridGet(g, x+2, y, z) + PSGridGet(g, x-2, y, z);
  PSGridEmit(g2, v);
  return;
}

void systest_stencil_hole(float *out) {
  int i;
  PSGrid3DFloat g1 = PSGrid3DFloatNew(N, N, N);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(N, N, N);
  PSDomain3D d = PSDomain3DNew(2, N - 2, 2, N - 2, 2, N - 2);
  float *indata = (float *)malloc(sizeof(float) * N * N * N);
  for (i = 0; i < N * N * N; i++) indata[i] = i;
  PSGridCopyin(g1, indata);
  PSGridCopyin(g2, indata);
  PSStencilRun(PSStencilMap(khole, d, g1, g2));
  PSGridCopyout(g2, out);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

/* ---- user-defined point types */

/* test_user-defined-type1.c:13-36: three scalar members, whole-point emit. */
struct Vec3 {
  float x;
  float y;
  float z;
};
DeclareGrid3D(Vec3, struct Vec3);

static void kvec3(const int x, const int y, const int z, PSGrid3DVec3 g1,
                  PSGrid3DVec3 g2) {
  float p = (PSGridGet(g1, x, y, z).x +
             PSGridGet(g1, x+1, y, z).x +
             PSGridGet(g1, x-1, y, z).x) / 3.3f;
  float q = (PSGridGet(g1, x, y, z).y +
             PSGridGet(g1, x, y+1, z).y +
             PSGridGet(g1, x, y-1, z).y +
             PSGridGet(g1, x, y, z+1).y +
             PSGridGet(g1, x, y, z-1).y) / 5.5f;
  float r = (PSGridGet(g1, x, y, z).z +
             PSGridGet(g1, x, y, z+1).z +
             PSGridGet(g1, x, y, z-1).z) / 3.3f;
  struct Vec3 v = {p, q, r};
  PSGridEmit(g2, v);
  return;
}

/* test_user-defined-type1.c main: points (i, i+1, i+2), 5 ping-pong pairs. */
void systest_udt_vec3(struct Vec3 *out) {
  size_t nelms = (size_t)N * N * N, i;
  PSGrid3DVec3 g1 = PSGrid3DVec3New(N, N, N);
  PSGrid3DVec3 g2 = PSGrid3DVec3New(N, N, N);
  PSDomain3D d = PSDomain3DNew(1, N - 1, 1, N - 1, 1, N - 1);
  struct Vec3 *indata = (struct Vec3 *)malloc(sizeof(struct Vec3) * nelms);
  for (i = 0; i < nelms; i++) {
    indata[i].x = i;
    indata[i].y = i + 1;
    indata[i].z = i + 2;
  }
  PSGridCopyin(g1, indata);
  PSGridCopyin(g2, indata);
  PSStencilRun(PSStencilMap(kvec3, d, g1, g2), PSStencilMap(kvec3, d, g2, g1),
               5);
  PSGridCopyout(g1, out);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

/* test_user-defined-type-multi-dim-member.c:13-32: a 3x2 array member, one
 * component emitted in place. */
struct Tensor {
  float p[3][2];
};
DeclareGrid3D(Tensor, struct Tensor);

static void ktensor(const int x, const int y, const int z, PSGrid3DTensor g) {
  float v1 = PSGridGet(g, x, y, z).p[0][0] +
      PSGridGet(g, x-1, y, z).p[0][0] +
      PSGridGet(g, x+1, y, z).p[0][0] +
      PSGridGet(g, x, y-1, z).p[0][0] +
      PSGridGet(g, x, y+1, z).p[0][0] +
      PSGridGet(g, x, y, z-1).p[0][0] +
      PSGridGet(g, x, y, z+1).p[0][0];
  float v2 = PSGridGet(g, x, y, z).p[2][1] +
      PSGridGet(g, x-1, y, z).p[2][1] +
      PSGridGet(g, x+1, y, z).p[2][1] +
      PSGridGet(g, x, y-1, z).p[2][1] +
      PSGridGet(g, x, y+1, z).p[2][1] +
      PSGridGet(g, x, y, z-1).p[2][1] +
      PSGridGet(g, x, y, z+1).p[2][1];
  PSGridEmitUtype(g.p[1][0], v1 + v2);
  return;
}

/* test_user-defined-type-multi-dim-member.c main: p[j][k] = i + j + k. */
void systest_udt_tensor(struct Tensor *out) {
  int nelms = N * N * N, i, j, k;
  PSGrid3DTensor g = PSGrid3DTensorNew(N, N, N);
  PSDomain3D d = PSDomain3DNew(1, N - 1, 1, N - 1, 1, N - 1);
  struct Tensor *indata = (struct Tensor *)malloc(sizeof(struct Tensor) * nelms);
  for (i = 0; i < nelms; i++)
    for (j = 0; j < 3; ++j)
      for (k = 0; k < 2; ++k) indata[i].p[j][k] = i + j + k;
  PSGridCopyin(g, indata);
  PSStencilRun(PSStencilMap(ktensor, d, g));
  PSGridCopyout(g, out);
  PSGridFree(g);
  free(indata);
}

/* test_user-defined-type-7-pt-periodic-complex.c:13-31: whole-point periodic
 * gets. */
struct Cplx {
  float r;
  float i;
};
DeclareGrid3D(Cplx, struct Cplx);

static void kcplx(const int x, const int y, const int z, PSGrid3DCplx g1,
                  PSGrid3DCplx g2) {
  struct Cplx t = PSGridGetPeriodic(g1, x, y, z);
  struct Cplx t1 = PSGridGetPeriodic(g1, x+1, y, z);
  struct Cplx t2 = PSGridGetPeriodic(g1, x-1, y, z);
  struct Cplx t3 = PSGridGetPeriodic(g1, x, y+1, z);
  struct Cplx t4 = PSGridGetPeriodic(g1, x, y-1, z);
  struct Cplx t5 = PSGridGetPeriodic(g1, x, y, z+1);
  struct Cplx t6 = PSGridGetPeriodic(g1, x, y, z-1);
  float r = t.r + t1.r + t2.r + t3.r + t4.r + t5.r + t6.r;
  float i = t.i + t1.i + t2.i + t3.i + t4.i + t5.i + t6.i;
  struct Cplx v = {r, i};
  PSGridEmit(g2, v);
  return;
}

/* test_user-defined-type-7-pt-periodic-complex.c main: (i, i+1), domain
 * [0, N)^3. */
void systest_udt_complex(struct Cplx *out) {
  size_t nelms = (size_t)N * N * N, i;
  PSGrid3DCplx g1 = PSGrid3DCplxNew(N, N, N);
  PSGrid3DCplx g2 = PSGrid3DCplxNew(N, N, N);
  PSDomain3D d = PSDomain3DNew(0, N, 0, N, 0, N);
  struct Cplx *indata = (struct Cplx *)malloc(sizeof(struct Cplx) * nelms);
  for (i = 0; i < nelms; i++) {
    indata[i].r = i;
    indata[i].i = i + 1;
  }
  PSGridCopyin(g1, indata);
  PSStencilRun(PSStencilMap(kcplx, d, g1, g2));
  PSGridCopyout(g2, out);
  PSGridFree(g1);
  PSGridFree(g2);
  free(indata);
}

/* ---- host-side element access (PSGridSet / PSGridGet in host code) */

/* test_set.c-style: elements written one by one with PSGridSet, one read
 * back with PSGridGet after a map doubled them. */
static void kdouble(const int x, const int y, const int z, PSGrid3DFloat g,
                    PSGrid3DFloat g2) {
  PSGridEmit(g2, PSGridGet(g, x, y, z) * 2.0f);
}

float systest_host_get(int x, int y, int z) {
  int i, j, k;
  float v = 0;
  PSGrid3DFloat g = PSGrid3DFloatNew(4, 3, 2);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(4, 3, 2);
  PSDomain3D d = PSDomain3DNew(0, 4, 0, 3, 0, 2);
  for (k = 0; k < 2; ++k)
    for (j = 0; j < 3; ++j)
      for (i = 0; i < 4; ++i) {
        PSGridSet(g, i, j, k, v);
        v += 1;
      }
  PSStencilRun(PSStencilMap(kdouble, d, g, g2));
  v = PSGridGet(g2, x, y, z);
  PSGridFree(g);
  PSGridFree(g2);
  return v;
}

/* An index parameter shadowed by a block-scope declaration: the get reads
 * g at (x + 1, y, z) of the point (x is rebound to x + 2 inside the block).
 * The translator must not take the inner x for the coordinate. */
static void kshadow(const int x, const int y, const int z, PSGrid3DFloat g,
                    PSGrid3DFloat g2) {
  float v;
  int xx = x;
  {
    int x = xx + 2;
    v = PSGridGet(g, x - 1, y, z);
  }
  PSGridEmit(g2, v);
}

void systest_shadowed_index(const float *in, float *out, int nx, int ny,
                            int nz) {
  PSGrid3DFloat g = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(nx, ny, nz);
  PSDomain3D d = PSDomain3DNew(0, nx, 0, ny, 0, nz);
  PSGridCopyin(g, in);
  PSStencilRun(PSStencilMap(kshadow, d, g, g2));
  PSGridCopyout(g2, out);
  PSGridFree(g);
  PSGridFree(g2);
}

/* Seven reducing maps in one run (the most a run may hold): map i sums
 * (i + 1) * g over its own domain, so every result differs. */
static float kscaled(const int x, const int y, const int z, PSGrid3DFloat g,
                     PSGrid3DFloat g2, float s) {
  float v = PSGridGet(g, x, y, z) * s;
  PSGridEmit(g2, v);
  return v;
}

void systest_seven_reductions(float *sums) {
  int n = 16, i;
  PSGrid3DFloat g = PSGrid3DFloatNew(n, n, n);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(n, n, n);
  float *indata = (float *)malloc(sizeof(float) * n * n * n);
  for (i = 0; i < n * n * n; i++) indata[i] = (float)(i % 7);
  PSGridCopyin(g, indata);
  PSDomain3D d0 = PSDomain3DNew(0, n, 0, n, 0, n);
  PSDomain3D d1 = PSDomain3DNew(0, n, 0, n, 0, n / 2);
  PSStencilRun(PSStencilMapReduce(&sums[0], PS_SUM, kscaled, d0, g, g2, 1.0f),
               PSStencilMapReduce(&sums[1], PS_SUM, kscaled, d1, g, g2, 2.0f),
               PSStencilMapReduce(&sums[2], PS_SUM, kscaled, d0, g, g2, 3.0f),
               PSStencilMapReduce(&sums[3], PS_SUM, kscaled, d1, g, g2, 4.0f),
               PSStencilMapReduce(&sums[4], PS_SUM, kscaled, d0, g, g2, 5.0f),
               PSStencilMapReduce(&sums[5], PS_SUM, kscaled, d1, g, g2, 6.0f),
               PSStencilMapReduce(&sums[6], PS_SUM, kscaled, d0, g, g2, 7.0f),
               1);
  PSGridFree(g);
  PSGridFree(g2);
  free(indata);
}
