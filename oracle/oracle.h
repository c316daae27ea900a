/* Oracle: a CPU restatement of the Physis REFERENCE target for the hot path
 * (PSStencilRun over 3-D grids, PSReduce), used ONLY as test infrastructure:
 * by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's
 * cpu_baseline leg.  Nothing in the product (physis_amd/, include/) links,
 * loads or calls it.
 *
 * What it restates (reference file:line):
 *   - grids are dense, x fastest, zero-filled on creation
 *     (runtime/libphysis_rt_ref.cc:53-71, include/physis/physis_ref.h:57-74);
 *   - a stencil map runs the kernel over its domain in the order z, y, x,
 *     each index over [min, max) (translator/reference_runtime_builder.cc:
 *     605-664); PSStencilRun repeats the maps in order iter times (:837-893);
 *   - PSReduce folds all elements sequentially in the element type, starting
 *     from element 0 (runtime/libphysis_rt_ref.cc:19-30, runtime/reduce.h:
 *     14-49); integer SUM/PROD wrap in two's complement.
 *   - stencil bodies keep the operand order of the reference programs named
 *     at each function, compiled without FMA contraction.
 *
 * Parity pinning (how each function is checked, see tests/test_oracle.py):
 *   - 7-point kernels and the system-test kernels: sha256 of the %f dumps of
 *     the reference's own hand-written oracles (tests/system_tests/
 *     test_cases/<name>.manual.ref.c, hashes in tests/golden/ref_outputs.json) and
 *     oracle/_ref binaries built from those files;
 *   - clamped diffusion: RMS error against the closed-form cosine solution
 *     (examples/diffusion-benchmark/baseline.cc:51-60) and bit equality with
 *     the reference's own C++ baseline loop built in oracle/_ref;
 *   - Himeno sweep: bit equality with the reference's original Himeno
 *     benchmark (examples/himeno/himenobmtxpa_original.c:298-346) built in
 *     oracle/_ref on its own initial state, plus the fixed-point invariant
 *     (constant p with b = 0, a = c = 1, a3 = 1/6, wrk1 = 0 is unchanged);
 *   - reductions: reference-fixture values (8386560, 154463400, -733463896)
 *     and, for float sums at scale, the FP64 sum (the REF float order is
 *     "parity unpinned" there).
 */
#ifndef PHYSIS_ORACLE_H_
#define PHYSIS_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Domain [min0,max0) x [min1,max1) x [min2,max2) as six ints. */
typedef struct {
  int min[3];
  int max[3];
} oracle_domain;

/* ---- reductions (op: 0 MAX, 1 MIN, 2 SUM, 3 PROD) */
float oracle_reduce_float(int op, const float *p, int64_t n);
double oracle_reduce_double(int op, const double *p, int64_t n);
int oracle_reduce_int(int op, const int *p, int64_t n);
long oracle_reduce_long(int op, const long *p, int64_t n);
/* FP64 accumulation of a float grid (sum / max / min), for error bounds. */
double oracle_reduce_float_f64(int op, const float *p, int64_t n);
/* Sum of |x| in FP64 (condition of a float sum). */
double oracle_abs_sum_f64(const float *p, int64_t n);

/* ---- 7-point diffusion (examples/diffusion-benchmark/diffusion3d_physis.c:
 * 29-58): one sweep in -> out over the domain.  periodic != 0 reads every
 * neighbour with PSGridGetPeriodic instead of clamping at the faces. */
void oracle_diffusion7_sweep(const float *in, float *out, int nx, int ny,
                             int nz, const oracle_domain *dom, int periodic,
                             float ce, float cw, float cn, float cs, float ct,
                             float cb, float cc);
/* count sweeps ping-ponging g1 -> g2 -> g1 over the whole grid (count even:
 * result in g1), as PSStencilRun(map(g1,g2), map(g2,g1), count/2). */
void oracle_diffusion7_run(float *g1, float *g2, int nx, int ny, int nz,
                           int count, int periodic, float ce, float cw,
                           float cn, float cs, float ct, float cb, float cc);
/* The same run restated in the MPI-OpenMP target's generated shape: the
 * domain split into `div` z units run by "#pragma omp parallel for
 * schedule(static,1)" (translator/mpi_openmp_create_kernel.cc:319), with
 * the REF loop nest inside each unit.  Bit-identical to oracle_diffusion7_run.
 * Returns the number of OpenMP threads used. */
int oracle_diffusion7_run_omp(float *g1, float *g2, int nx, int ny, int nz,
                              int count, int periodic, float ce, float cw,
                              float cn, float cs, float ct, float cb,
                              float cc, int div);

/* ---- 19-point Himeno Jacobi (examples/himeno/himenobmtxpa_physis.c:
 * 331-361) over the inner domain [1, n-1)^3.  coef points at 12 grids in the
 * order a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 bnd wrk1.  If gosa is non-NULL the
 * residual sum of ss*ss is accumulated into it in FP64, z-y-x order. */
void oracle_himeno_sweep(const float *p0, float *p1, const float *const *coef,
                         int nx, int ny, int nz, float omega, double *gosa);
/* nn sweeps ping-ponging p0 -> p1 -> p0 (nn even: result in p0); gosa[i]
 * receives sweep i's FP64 residual sum if gosa is non-NULL. */
void oracle_himeno_run(float *p0, float *p1, const float *const *coef, int nx,
                       int ny, int nz, int nn, float omega, double *gosa);
/* The same sweeps with the z range split into div units run by OpenMP
 * threads (the shape of the reference's MPI-OpenMP target); gosa[i], if
 * non-NULL, is the sum of the units' FP64 partials in unit order.  Returns
 * the number of threads used. */
int oracle_himeno_run_omp(float *p0, float *p1, const float *const *coef,
                          int nx, int ny, int nz, int nn, float omega,
                          int div, double *gosa);

/* ---- red-black 7-point relaxation (physis_amd/programs/redblack7.c):
 * count iterations of a red then a black in-place sweep over [1, n-1)^3
 * in REFERENCE loop order; coef = ce cw cn cs ct cb cc. */
void oracle_redblack7_run(float *g, int nx, int ny, int nz, int count,
                          const float *coef);

/* ---- 7-point map weighted by 1-D axis coefficient grids
 * (physis_amd/programs/mixed7.c): one sweep o <- kernel(g) over [1, n-1)^3. */
void oracle_mixed7_sweep(const float *g, float *o, int nx, int ny, int nz,
                         const float *cx, const float *cy, const float *cz);

/* ---- 2-D 5- and 9-point maps (physis_amd/programs/stencil2d.c): `sweeps`
 * sweeps alternating a -> b, b -> a over [1, nx-1) x [1, ny-1); out = (sum of
 * the points in the reference tests' order) * w. */
void oracle_stencil2d_run(float *a, float *b, int nx, int ny, int points,
                          float w, int sweeps);

/* ---- 7-point stencil on user-defined points (physis_amd/programs/udt7.c):
 * one sweep q <- weighted 7-point sum of p over [1, n-1)^3; coef = ce cw cn
 * cs ct cb cc. */
void oracle_udt7_sweep(const float *p, float *q, int nx, int ny, int nz,
                       const float *coef);

/* ---- system-test kernels (tests/system_tests/test_cases), N^3 float/int
 * grids, each one map application in -> out over the test's domain. */
void oracle_test_7pt(const float *in, float *out, int n);           /* test_7-pt.c:12-20 */
void oracle_test_7pt_periodic(const float *in, float *out, int n);  /* test_7-pt-periodic.c:12-24 */
void oracle_test_7pt_neumann(const float *in, float *out, int n);   /* test_7-pt-neumann-cond.c:15-32 */
void oracle_test_asymmetric(const float *in, float *out, int n);    /* test_asymmetric.c:12-18 */
void oracle_test_asymmetric_periodic(const float *in, float *out, int n); /* test_asymmetric-periodic.c:12-18 */
void oracle_test_27pt(const float *in, float *out, int n, int periodic);  /* test_27-pt{,-periodic}.c */
void oracle_test_27pt_int(const int *in, int *out, int n);          /* test_27-pt-reduction.c:14-40 */

/* ---- Poisson solver extensions (oracle/poisson.c).  storage: 0 FP32,
 * 1 bfloat16, 2 binary16 (16-bit values as uint16 bit patterns).  Grids hold
 * nbatch problems of nz/nbatch planes stacked along z; faces of every problem
 * keep their values.  Coefficient grids: a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 bnd
 * wrk1 (the sweep's last field is its right-hand side). */
uint16_t oracle_f32_to_bf16(float f);
float oracle_bf16_to_f32(uint16_t h);
uint16_t oracle_f32_to_f16(float f);
float oracle_f16_to_f32(uint16_t h);
void oracle_to_storage(int storage, const float *in, void *out, int64_t n);
void oracle_poisson_sweep(int storage, const void *pin, void *pout,
                          const void *const *f, int nx, int ny, int nz,
                          int nbatch, float omega);
void oracle_poisson_residual(int storage, const void *p, int p64,
                             double *p_out, const void *e, double inv_sigma,
                             double sigma, void *rhs, const float *const *cf,
                             int nx, int ny, int nz, int nbatch,
                             double *sums);
void oracle_poisson_solve_mixed(int storage, const float *const *cf,
                                double *p, int nx, int ny, int nz, int nbatch,
                                float omega, double tol, int inner,
                                long max_sweeps, double *stats);
double oracle_poisson_jacobi_fp64(const float *const *cf, double *p, int nx,
                                  int ny, int nz, int nbatch, float omega,
                                  double tol, int check, long max_sweeps,
                                  long *sweeps_out);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_ORACLE_H_ */
