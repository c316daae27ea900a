/* Poisson solver extensions of the HIP target (libphysis_poisson_hip):
 * the 19-point variable-coefficient Jacobi operator of the Himeno benchmark
 * (reference examples/himeno/himenobmtxpa_physis.c:331-361) solved to a
 * residual tolerance, with
 *   - a precision tier: the inner Jacobi sweeps read and write grids stored
 *     in bfloat16 or binary16 (FP32 arithmetic), while the residual, its
 *     global reduction and the refinement correction are FP64 on the GPU
 *     (mixed-precision iterative refinement);
 *   - the plain FP32 Jacobi solve it is measured against.
 * Neither has a counterpart in the reference, which only runs a fixed number
 * of sweeps.  C ABI; grids are the runtime's __PSGrid handles.
 *
 * Operator.  With the 12 FP32 coefficient grids a0..a3, b0..b2, c0..c2, bnd,
 * wrk1 (the Himeno jacobi_kernel's arguments), one Jacobi sweep is
 *   s0  = a0 p(z+1) + a1 p(y+1) + a2 p(x+1)
 *       + b0 (p(y+1,z+1) - p(y-1,z+1) - p(y+1,z-1) + p(y-1,z-1))
 *       + b1 (p(x+1,y+1) - p(x+1,y-1) - p(x-1,y+1) + p(x-1,y-1))
 *       + b2 (p(x+1,z+1) - p(x+1,z-1) - p(x-1,z+1) + p(x-1,z-1))
 *       + c0 p(z-1) + c1 p(y-1) + c2 p(x-1) + wrk1
 *   p'  = p + omega (a3 s0 - p) bnd,
 * the fixed point of which solves A p = wrk1 with
 *   (A p) = p / a3 - (s0 - wrk1)   where bnd = 1;
 * points with bnd = 0 keep their value (Dirichlet data).  The residual is
 *   r = (s0 - p / a3) bnd        (= wrk1 - A p on bnd = 1 points),
 * and the relative residual of a problem is ||r||_2 / ||bnd wrk1||_2, both in
 * FP64.  Every grid point is swept; bnd must be 0 on the faces of each
 * problem.
 *
 * Batches.  nbatch independent problems of nx x ny x nzs may be stacked along
 * z into grids of nx x ny x (nbatch nzs); bnd = 0 on every problem's faces
 * decouples them.  Convergence means every problem's relative residual is
 * at most tol.
 *
 * Threading: the runtime's stream, one host thread; calls are synchronous
 * on return. */
#ifndef PHYSIS_PHYSIS_POISSON_H_
#define PHYSIS_PHYSIS_POISSON_H_

#include "physis/physis_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Storage of the inner-sweep grids. */
enum PSStorage {
  PS_STORAGE_FP32 = 0,
  PS_STORAGE_BF16 = 1,
  PS_STORAGE_FP16 = 2
};

/* Fields of a solver that tests and benchmarks can read and write. */
enum PSPoissonField {
  PS_POISSON_E = 0,      /* inner iterate (correction), storage type */
  PS_POISSON_RHS = 1,    /* inner right-hand side (scaled residual), storage */
  PS_POISSON_COEF0 = 2   /* + k: storage-type copy of coefficient k (0..10) */
};

typedef struct __PSPoisson __PSPoisson;

typedef struct {
  int converged;        /* 1: every problem reached tol */
  int stagnated;        /* 1: stopped because the residual stopped falling */
  int outer;            /* residual evaluations after the initial one */
  long sweeps;          /* Jacobi sweeps */
  double rel_residual;  /* max over problems of ||r|| / ||b||, FP64, at exit */
  double seconds;       /* wall time of the solve, device work included */
  /* Optional trace, set by the caller (NULL / 0: none): after every
   * residual evaluation the solver appends (sweeps, rel_residual, seconds)
   * while history_len < history_cap. */
  double *history;
  int history_cap;
  int history_len;
  /* > 0: print (sweeps, residual, seconds) to stderr every that many
   * seconds of solving, at residual checks (long solves). */
  double progress_seconds;
} PSPoissonStats;

/* A solver over the 12 FP32 coefficient grids (a0 a1 a2 a3 b0 b1 b2 c0 c1 c2
 * bnd wrk1), all 3-D of the same size with nz = nbatch * nzs.  storage selects
 * the inner-sweep grids; the FP32 grids are read, not copied, and must outlive
 * the solver.  Returns NULL (and prints why) on bad arguments or when device
 * memory runs out. */
__PSPoisson *PSPoissonNew(__PSGrid *const coef[12], int nbatch, int storage);
void PSPoissonFree(__PSPoisson *s);

/* Mixed-precision iterative refinement.  p: FP64 grid, the initial guess on
 * entry (its bnd = 0 points are the boundary values) and the solution on
 * return.  Each refinement step runs inner_sweeps Jacobi sweeps (even,
 * >= 2) on A e = r in the storage type from e = 0, then p += e and r = wrk1 -
 * A p in FP64.  Stops when converged or after max_sweeps sweeps.  Returns
 * st->converged. */
int PSPoissonSolveMixed(__PSPoisson *s, __PSGrid *p, float omega, double tol,
                        int inner_sweeps, long max_sweeps, PSPoissonStats *st);

/* Plain FP32 Jacobi: p (FP32 grid) in/out, the FP64 relative residual checked
 * every `check` sweeps (even, >= 2).  Stops when converged, after
 * max_sweeps, or when 20 consecutive checks spanning at least 20000 sweeps
 * fail to lower the best residual by 0.1% (FP32 stagnation;
 * st->stagnated = 1; the mixed solve applies the same rule).  Returns
 * st->converged. */
int PSPoissonSolvePlain(__PSPoisson *s, __PSGrid *p, float omega, double tol,
                        int check, long max_sweeps, PSPoissonStats *st);

/* ---- single steps (measurement and parity tests) */
/* n sweeps (n even) of the storage-type Jacobi on A e = rhs, E in/out. */
void PSPoissonInnerSweeps(__PSPoisson *s, int n, float omega);
/* n sweeps (n even) of the FP32 Jacobi on A p = wrk1, p (FP32 grid) in/out. */
void PSPoissonPlainSweeps(__PSPoisson *s, __PSGrid *p, int n, float omega);
/* FP64 residual.  p: FP32 or FP64 grid.  If p_out is non-NULL (FP64 grids
 * only) the pass first forms p + inv_sigma * E at every point, writes it to
 * p_out and takes the residual of that.  If sigma != 0 it writes the storage
 * type's RHS field as sigma * r (sigma a power of two).  sums receives 3
 * doubles per problem: sum r^2, max |r|, sum (bnd wrk1)^2. */
void PSPoissonResidual(__PSPoisson *s, const __PSGrid *p, __PSGrid *p_out,
                       double inv_sigma, double sigma, double *sums);
/* Host copies of a field (nx*ny*nz elements of the storage type: uint16 bit
 * patterns for BF16/FP16, float for FP32).  Return 0, or -1 on a bad field. */
int PSPoissonSetField(__PSPoisson *s, int field, const void *host);
int PSPoissonGetField(__PSPoisson *s, int field, void *host);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_PHYSIS_POISSON_H_ */
