/* Oracle of the Poisson solver extensions (precision tier and plain FP32
 * solve), test infrastructure only: used by tests/ and bench.py's
 * cpu_baseline leg, never by the product.  The extensions have no reference
 * counterpart; the sweep arithmetic is the Himeno jacobi_kernel
 * (reference examples/himeno/himenobmtxpa_physis.c:340-359, restated as
 * oracle_himeno_sweep in oracle.c), and everything else follows the
 * definitions in include/physis/physis_poisson.h.
 *
 * Parity pinning (tests/test_oracle_poisson.py):
 *   - bfloat16 / binary16 conversions: against numpy float16 and torch
 *     bfloat16 (both round to nearest even) on random values, ties and
 *     boundary cases;
 *   - the FP32-storage sweep: bit equality with oracle_himeno_sweep, which
 *     is pinned to the reference's original Himeno benchmark;
 *   - the FP64 residual: omega * a3 * r equals the Himeno update p' - p up to
 *     FP32 rounding, and r = 0 at the solution of a direct sparse solve;
 *   - the FP64 Jacobi solve: against scipy.sparse.linalg.spsolve of the same
 *     operator assembled as a matrix, on small grids.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define OFF(x, y, z) ((int64_t)(x) + (int64_t)(y) * nx + (int64_t)(z) * nx * ny)

/* ------------------------------------------------------- 16-bit formats */
static float f_of_u(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t u_of_f(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

uint16_t oracle_f32_to_bf16(float f) {
  uint32_t u = u_of_f(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

float oracle_bf16_to_f32(uint16_t h) { return f_of_u((uint32_t)h << 16); }

/* IEEE binary16, round to nearest even, with subnormals, overflow to inf. */
uint16_t oracle_f32_to_f16(float f) {
  const uint32_t u = u_of_f(f);
  const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
  const uint32_t a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return sign | 0x7e00u;          /* NaN */
  if (a >= 0x477ff000u) return sign | 0x7c00u;          /* rounds to inf */
  if (a >= 0x38800000u) {                               /* normal half */
    const uint32_t e = (a >> 23) - 127 + 15;
    uint32_t m = a & 0x7fffffu;
    uint32_t h = (e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1; /* may carry */
    return sign | (uint16_t)h;
  }
  if (a < 0x33000000u) return sign;                     /* < 2^-25: zero */
  /* subnormal half: value = m_full * 2^(e-150); half unit 2^-24 */
  {
    const int e = (int)(a >> 23);
    const uint32_t mf = (a & 0x7fffffu) | 0x800000u;
    const int shift = 126 - e; /* 14 .. 24: mf >> (shift) in units 2^-24 */
    uint32_t h = mf >> shift;
    const uint32_t rem = mf & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) h += 1;
    return sign | (uint16_t)h;
  }
}

float oracle_f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0) {
    const float v = (float)m * 5.9604644775390625e-08f; /* 2^-24, exact */
    return f_of_u(u_of_f(v) | sign);
  }
  if (e == 31) return f_of_u(sign | 0x7f800000u | (m << 13));
  return f_of_u(sign | ((e - 15 + 127) << 23) | (m << 13));
}

static float ld(int storage, const void *p, int64_t i) {
  switch (storage) {
    case 1: return oracle_bf16_to_f32(((const uint16_t *)p)[i]);
    case 2: return oracle_f16_to_f32(((const uint16_t *)p)[i]);
    default: return ((const float *)p)[i];
  }
}
static void st(int storage, void *p, int64_t i, float v) {
  switch (storage) {
    case 1: ((uint16_t *)p)[i] = oracle_f32_to_bf16(v); break;
    case 2: ((uint16_t *)p)[i] = oracle_f32_to_f16(v); break;
    default: ((float *)p)[i] = v;
  }
}

void oracle_to_storage(int storage, const float *in, void *out, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) st(storage, out, i, in[i]);
}

/* A point is swept when it is not on a face of its problem. */
static int interior(int x, int y, int z, int nx, int ny, int nzs) {
  const int zl = z % nzs;
  return x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && zl > 0 &&
         zl < nzs - 1;
}

/* ---------------------------------------------------- storage-type sweep */
void oracle_poisson_sweep(int storage, const void *pin, void *pout,
                          const void *const *f, int nx, int ny, int nz,
                          int nbatch, float omega) {
  const int nzs = nz / nbatch;
#define P(a, b, c) ld(storage, pin, OFF(a, b, c))
  /* planes are independent (pin is read only): splitting z over threads
   * leaves every point's arithmetic unchanged */
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t o = OFF(x, y, z);
        const float pc = ld(storage, pin, o);
        const float bnd = ld(storage, f[10], o);
        if (!interior(x, y, z, nx, ny, nzs) || bnd == 0.0f) {
          st(storage, pout, o, pc);
          continue;
        }
        float c[12];
        for (int g = 0; g < 12; ++g) c[g] = ld(storage, f[g], o);
        const float s0 =
            c[0] * P(x, y, z + 1) + c[1] * P(x, y + 1, z) + c[2] * P(x + 1, y, z) +
            c[4] * (P(x, y + 1, z + 1) - P(x, y - 1, z + 1) - P(x, y + 1, z - 1) +
                    P(x, y - 1, z - 1)) +
            c[5] * (P(x + 1, y + 1, z) - P(x + 1, y - 1, z) - P(x - 1, y + 1, z) +
                    P(x - 1, y - 1, z)) +
            c[6] * (P(x + 1, y, z + 1) - P(x + 1, y, z - 1) - P(x - 1, y, z + 1) +
                    P(x - 1, y, z - 1)) +
            c[7] * P(x, y, z - 1) + c[8] * P(x, y - 1, z) + c[9] * P(x - 1, y, z) +
            c[11];
        const float ss = (s0 * c[3] - pc) * c[10];
        st(storage, pout, o, pc + omega * ss);
      }
#undef P
}

/* --------------------------------------------------------- FP64 residual */
void oracle_poisson_residual(int storage, const void *p, int p64,
                             double *p_out, const void *e, double inv_sigma,
                             double sigma, void *rhs, const float *const *cf,
                             int nx, int ny, int nz, int nbatch,
                             double *sums) {
  const int nzs = nz / nbatch;
  const int64_t n = (int64_t)nx * ny * nz;
  double *pv = (double *)malloc(sizeof(double) * (size_t)n);
  /* per-plane partial sums, folded in plane order below */
  double *ps = (double *)calloc(3 * (size_t)nz, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double v = p64 ? ((const double *)p)[i] : (double)((const float *)p)[i];
    if (p_out) v = v + (double)ld(storage, e, i) * inv_sigma;
    pv[i] = v;
  }
  if (p_out) memcpy(p_out, pv, sizeof(double) * (size_t)n);
  for (int q = 0; q < nbatch; ++q) sums[3 * q] = sums[3 * q + 1] = sums[3 * q + 2] = 0.0;
#define P(a, b, c) pv[OFF(a, b, c)]
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t o = OFF(x, y, z);
        double c[12];
        for (int g = 0; g < 12; ++g) c[g] = (double)cf[g][o];
        double r = 0.0;
        if (interior(x, y, z, nx, ny, nzs) && c[10] != 0.0) {
          const double s0 =
              c[0] * P(x, y, z + 1) + c[1] * P(x, y + 1, z) + c[2] * P(x + 1, y, z) +
              c[4] * (P(x, y + 1, z + 1) - P(x, y - 1, z + 1) - P(x, y + 1, z - 1) +
                      P(x, y - 1, z - 1)) +
              c[5] * (P(x + 1, y + 1, z) - P(x + 1, y - 1, z) - P(x - 1, y + 1, z) +
                      P(x - 1, y - 1, z)) +
              c[6] * (P(x + 1, y, z + 1) - P(x + 1, y, z - 1) - P(x - 1, y, z + 1) +
                      P(x - 1, y, z - 1)) +
              c[7] * P(x, y, z - 1) + c[8] * P(x, y - 1, z) + c[9] * P(x - 1, y, z) +
              c[11];
          r = (s0 - P(x, y, z) / c[3]) * c[10];
        }
        if (rhs && sigma != 0.0) st(storage, rhs, o, (float)(r * sigma));
        double *s = ps + 3 * (size_t)z;
        const double b = c[11] * c[10];
        s[0] += r * r;
        if (fabs(r) > s[1]) s[1] = fabs(r);
        s[2] += b * b;
      }
#undef P
  for (int z = 0; z < nz; ++z) {
    double *s = sums + 3 * (z / nzs);
    const double *t = ps + 3 * (size_t)z;
    s[0] += t[0];
    if (t[1] > s[1]) s[1] = t[1];
    s[2] += t[2];
  }
  free(ps);
  free(pv);
}

static void summarise(const double *sums, int nbatch, double *rel,
                      double *rmax) {
  *rel = 0.0;
  *rmax = 0.0;
  for (int q = 0; q < nbatch; ++q) {
    const double r2 = sums[3 * q], b2 = sums[3 * q + 2];
    const double v = b2 > 0.0 ? sqrt(r2 / b2) : (r2 > 0.0 ? INFINITY : 0.0);
    if (v > *rel) *rel = v;
    if (sums[3 * q + 1] > *rmax) *rmax = sums[3 * q + 1];
  }
}

static double scale_for(double rmax) {
  int e;
  if (!(rmax > 0.0) || !isfinite(rmax)) return 1.0;
  frexp(rmax, &e);
  return ldexp(1.0, -e);
}

/* ------------------------------------------- mixed-precision refinement */
/* The refinement schedule of PSPoissonSolveMixed, step for step.  stats:
 * [outer, sweeps, rel_residual, stagnated]. */
void oracle_poisson_solve_mixed(int storage, const float *const *cf,
                                double *p, int nx, int ny, int nz, int nbatch,
                                float omega, double tol, int inner,
                                long max_sweeps, double *stats) {
  const int64_t n = (int64_t)nx * ny * nz;
  const size_t es = storage == 0 ? 4 : 2;
  void *lo[12];
  for (int g = 0; g < 11; ++g) {
    lo[g] = malloc(es * (size_t)n);
    oracle_to_storage(storage, cf[g], lo[g], n);
  }
  void *rhs = calloc((size_t)n, es);
  lo[11] = rhs;
  void *e0 = calloc((size_t)n, es), *e1 = calloc((size_t)n, es);
  double *pb = (double *)malloc(sizeof(double) * (size_t)n);
  double *sums = (double *)malloc(sizeof(double) * 3 * (size_t)nbatch);
  double rel, rmax, best;
  long sweeps = 0, best_at = 0;
  int outer = 0, no_gain = 0, rhs_ready = 0, stagnated = 0;
  oracle_poisson_residual(storage, p, 1, NULL, NULL, 0, 0, NULL, cf, nx, ny,
                          nz, nbatch, sums);
  summarise(sums, nbatch, &rel, &rmax);
  double sigma = scale_for(rmax);
  best = rel;
  double *pa = p;
  while (rel > tol && sweeps < max_sweeps) {
    if (!rhs_ready) {
      oracle_poisson_residual(storage, pa, 1, NULL, NULL, 0, sigma, rhs, cf,
                              nx, ny, nz, nbatch, sums);
      rhs_ready = 1;
    }
    memset(e0, 0, es * (size_t)n);
    for (int i = 0; i < inner; i += 2) {
      oracle_poisson_sweep(storage, e0, e1, (const void *const *)lo, nx, ny,
                           nz, nbatch, omega);
      oracle_poisson_sweep(storage, e1, e0, (const void *const *)lo, nx, ny,
                           nz, nbatch, omega);
    }
    sweeps += inner;
    const double next_sigma = scale_for(rmax);
    oracle_poisson_residual(storage, pa, 1, pb, e0, 1.0 / sigma, next_sigma,
                            rhs, cf, nx, ny, nz, nbatch, sums);
    sigma = next_sigma;
    memcpy(pa, pb, sizeof(double) * (size_t)n);
    outer += 1;
    summarise(sums, nbatch, &rel, &rmax);
    if (rel < best * (1.0 - 1e-3)) {
      best = rel;
      no_gain = 0;
      best_at = sweeps;
    } else if (++no_gain >= 20 && sweeps - best_at >= 20000) {
      stagnated = 1;
      break;
    }
  }
  stats[0] = outer;
  stats[1] = (double)sweeps;
  stats[2] = rel;
  stats[3] = stagnated;
  for (int g = 0; g < 12; ++g) free(lo[g]);
  free(e0);
  free(e1);
  free(pb);
  free(sums);
}

/* ------------------------------------------------- FP64 plain Jacobi */
/* p' = p + omega (a3 s0 - p) bnd in FP64 over the interior of every problem,
 * from FP32 coefficients; stops at relative residual tol (checked every
 * `check` sweeps) or max_sweeps.  Returns the final relative residual; the
 * sweep count goes to *sweeps_out. */
double oracle_poisson_jacobi_fp64(const float *const *cf, double *p, int nx,
                                  int ny, int nz, int nbatch, float omega,
                                  double tol, int check, long max_sweeps,
                                  long *sweeps_out) {
  const int nzs = nz / nbatch;
  const int64_t n = (int64_t)nx * ny * nz;
  double *t = (double *)malloc(sizeof(double) * (size_t)n);
  double *sums = (double *)malloc(sizeof(double) * 3 * (size_t)nbatch);
  double rel, rmax;
  long sweeps = 0;
  const double w = (double)omega;
  oracle_poisson_residual(0, p, 1, NULL, NULL, 0, 0, NULL, cf, nx, ny, nz,
                          nbatch, sums);
  summarise(sums, nbatch, &rel, &rmax);
  while (rel > tol && sweeps < max_sweeps) {
    for (int it = 0; it < check; ++it) {
      const double *in = (it % 2 == 0) ? p : t;
      double *out = (it % 2 == 0) ? t : p;
#define P(a, b, c) in[OFF(a, b, c)]
#pragma omp parallel for schedule(static)
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            const int64_t o = OFF(x, y, z);
            double c[12];
            for (int g = 0; g < 12; ++g) c[g] = (double)cf[g][o];
            if (!interior(x, y, z, nx, ny, nzs) || c[10] == 0.0) {
              out[o] = in[o];
              continue;
            }
            const double s0 =
                c[0] * P(x, y, z + 1) + c[1] * P(x, y + 1, z) + c[2] * P(x + 1, y, z) +
                c[4] * (P(x, y + 1, z + 1) - P(x, y - 1, z + 1) - P(x, y + 1, z - 1) +
                        P(x, y - 1, z - 1)) +
                c[5] * (P(x + 1, y + 1, z) - P(x + 1, y - 1, z) - P(x - 1, y + 1, z) +
                        P(x - 1, y - 1, z)) +
                c[6] * (P(x + 1, y, z + 1) - P(x + 1, y, z - 1) - P(x - 1, y, z + 1) +
                        P(x - 1, y, z - 1)) +
                c[7] * P(x, y, z - 1) + c[8] * P(x, y - 1, z) + c[9] * P(x - 1, y, z) +
                c[11];
            out[o] = in[o] + w * ((s0 * c[3] - in[o]) * c[10]);
          }
#undef P
    }
    if (check % 2) memcpy(p, t, sizeof(double) * (size_t)n);
    sweeps += check;
    oracle_poisson_residual(0, p, 1, NULL, NULL, 0, 0, NULL, cf, nx, ny, nz,
                            nbatch, sums);
    summarise(sums, nbatch, &rel, &rmax);
  }
  if (sweeps_out) *sweeps_out = sweeps;
  free(t);
  free(sums);
  return rel;
}
