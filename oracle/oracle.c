/* Oracle implementation: CPU restatement of the Physis REFERENCE target for
 * the hot path.  Test infrastructure only (see oracle.h); build with
 * oracle/Makefile (gcc, -ffp-contract=off, no -ffast-math). */
#include "oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OFF(x, y, z, nx, ny) \
  ((int64_t)(x) + (int64_t)(y) * (nx) + (int64_t)(z) * (nx) * (ny))

/* ---------------------------------------------------------------- reduce */
/* Sequential fold from element 0 (runtime/libphysis_rt_ref.cc:19-30) with
 * the operators of runtime/reduce.h:14-49. */
#define DEFINE_REDUCE(NAME, T, UT)                                  \
  T NAME(int op, const T *p, int64_t n) {                           \
    T v = p[0];                                                     \
    for (int64_t i = 1; i < n; ++i) {                               \
      T d = p[i];                                                   \
      switch (op) {                                                 \
        case 0: v = (v > d) ? v : d; break;                         \
        case 1: v = (v < d) ? v : d; break;                         \
        case 2: v = (T)((UT)v + (UT)d); break;                      \
        default: v = (T)((UT)v * (UT)d); break;                     \
      }                                                             \
    }                                                               \
    return v;                                                       \
  }
DEFINE_REDUCE(oracle_reduce_float, float, float)
DEFINE_REDUCE(oracle_reduce_double, double, double)
DEFINE_REDUCE(oracle_reduce_int, int, unsigned int)
DEFINE_REDUCE(oracle_reduce_long, long, unsigned long)

double oracle_reduce_float_f64(int op, const float *p, int64_t n) {
  double v = p[0];
  for (int64_t i = 1; i < n; ++i) {
    double d = p[i];
    switch (op) {
      case 0: v = (v > d) ? v : d; break;
      case 1: v = (v < d) ? v : d; break;
      case 2: v += d; break;
      default: v *= d; break;
    }
  }
  return v;
}

double oracle_abs_sum_f64(const float *p, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += fabs((double)p[i]);
  return s;
}

/* -------------------------------------------------------- 7-pt diffusion */
static inline int wrap(int i, int n) { return (i + n) % n; }

/* Body of kernel_physis (examples/diffusion-benchmark/diffusion3d_physis.c:
 * 29-58, the if/else form), and its periodic twin. */
static inline float diffusion7_point(const float *g, int x, int y, int z,
                                     int nx, int ny, int nz, int periodic,
                                     float ce, float cw, float cn, float cs,
                                     float ct, float cb, float cc) {
  float c, w, e, n, s, b, t;
  if (!periodic) {
    c = g[OFF(x, y, z, nx, ny)];
    if (x == 0) w = g[OFF(x, y, z, nx, ny)]; else w = g[OFF(x - 1, y, z, nx, ny)];
    if (x == nx - 1) e = g[OFF(x, y, z, nx, ny)]; else e = g[OFF(x + 1, y, z, nx, ny)];
    if (y == 0) n = g[OFF(x, y, z, nx, ny)]; else n = g[OFF(x, y - 1, z, nx, ny)];
    if (y == ny - 1) s = g[OFF(x, y, z, nx, ny)]; else s = g[OFF(x, y + 1, z, nx, ny)];
    if (z == 0) b = g[OFF(x, y, z, nx, ny)]; else b = g[OFF(x, y, z - 1, nx, ny)];
    if (z == nz - 1) t = g[OFF(x, y, z, nx, ny)]; else t = g[OFF(x, y, z + 1, nx, ny)];
  } else {
    c = g[OFF(x, y, z, nx, ny)];
    w = g[OFF(wrap(x - 1, nx), y, z, nx, ny)];
    e = g[OFF(wrap(x + 1, nx), y, z, nx, ny)];
    n = g[OFF(x, wrap(y - 1, ny), z, nx, ny)];
    s = g[OFF(x, wrap(y + 1, ny), z, nx, ny)];
    b = g[OFF(x, y, wrap(z - 1, nz), nx, ny)];
    t = g[OFF(x, y, wrap(z + 1, nz), nx, ny)];
  }
  return cc * c + cw * w + ce * e + cs * s + cn * n + cb * b + ct * t;
}

static void diffusion7_zrange(const float *in, float *out, int nx, int ny,
                              int nz, const oracle_domain *d, int z0, int z1,
                              int periodic, float ce, float cw, float cn,
                              float cs, float ct, float cb, float cc) {
  for (int z = z0; z < z1; ++z)
    for (int y = d->min[1]; y < d->max[1]; ++y)
      for (int x = d->min[0]; x < d->max[0]; ++x)
        out[OFF(x, y, z, nx, ny)] = diffusion7_point(
            in, x, y, z, nx, ny, nz, periodic, ce, cw, cn, cs, ct, cb, cc);
}

void oracle_diffusion7_sweep(const float *in, float *out, int nx, int ny,
                             int nz, const oracle_domain *d, int periodic,
                             float ce, float cw, float cn, float cs, float ct,
                             float cb, float cc) {
  diffusion7_zrange(in, out, nx, ny, nz, d, d->min[2], d->max[2], periodic,
                    ce, cw, cn, cs, ct, cb, cc);
}

void oracle_diffusion7_run(float *g1, float *g2, int nx, int ny, int nz,
                           int count, int periodic, float ce, float cw,
                           float cn, float cs, float ct, float cb, float cc) {
  oracle_domain d = {{0, 0, 0}, {nx, ny, nz}};
  for (int i = 0; i < count / 2; ++i) {
    oracle_diffusion7_sweep(g1, g2, nx, ny, nz, &d, periodic, ce, cw, cn, cs,
                            ct, cb, cc);
    oracle_diffusion7_sweep(g2, g1, nx, ny, nz, &d, periodic, ce, cw, cn, cs,
                            ct, cb, cc);
  }
}

/* Division units along z (mpi_openmp_translator.h:9-15 splits x, y, z;
 * only z is divided here, as in the default {1, 1, div}). */
static inline void unit_range(int lo, int hi, int div, int u, int *a,
                              int *b) {
  int len = hi - lo, chunk = len / div, rem = len % div;
  *a = lo + chunk * u + (u < rem ? u : rem);
  *b = *a + chunk + (u < rem ? 1 : 0);
}

int oracle_diffusion7_run_omp(float *g1, float *g2, int nx, int ny, int nz,
                              int count, int periodic, float ce, float cw,
                              float cn, float cs, float ct, float cb,
                              float cc, int div) {
  oracle_domain d = {{0, 0, 0}, {nx, ny, nz}};
  int threads = 1;
  if (div < 1) div = 1;
  for (int i = 0; i < count; ++i) {
    const float *in = (i % 2 == 0) ? g1 : g2;
    float *out = (i % 2 == 0) ? g2 : g1;
#pragma omp parallel for schedule(static, 1)
    for (int u = 0; u < div; ++u) {
      int z0, z1;
      unit_range(0, nz, div, u, &z0, &z1);
      diffusion7_zrange(in, out, nx, ny, nz, &d, z0, z1, periodic, ce, cw, cn,
                        cs, ct, cb, cc);
#ifdef _OPENMP
      if (u == 0) threads = omp_get_num_threads();
#endif
    }
  }
  return threads;
}

/* ----------------------------------------------------------- 19-pt Himeno */
/* jacobi_kernel of examples/himeno/himenobmtxpa_physis.c:331-361; i is x,
 * j is y, k is z.  Returns ss. */
static inline float himeno_point(const float *p, float *out,
                                 const float *const *cf, int i, int j, int k,
                                 int nx, int ny, float omega) {
#define P(a, b, c) p[OFF(a, b, c, nx, ny)]
  const int64_t o = OFF(i, j, k, nx, ny);
  const float *a0 = cf[0], *a1 = cf[1], *a2 = cf[2], *a3 = cf[3];
  const float *b0 = cf[4], *b1 = cf[5], *b2 = cf[6];
  const float *c0 = cf[7], *c1 = cf[8], *c2 = cf[9];
  const float *bnd = cf[10], *wrk1 = cf[11];
  float s0, ss;
  s0 = a0[o] * P(i, j, k + 1) + a1[o] * P(i, j + 1, k) + a2[o] * P(i + 1, j, k) +
       b0[o] * (P(i, j + 1, k + 1) - P(i, j - 1, k + 1) - P(i, j + 1, k - 1) +
                P(i, j - 1, k - 1)) +
       b1[o] * (P(i + 1, j + 1, k) - P(i + 1, j - 1, k) - P(i - 1, j + 1, k) +
                P(i - 1, j - 1, k)) +
       b2[o] * (P(i + 1, j, k + 1) - P(i + 1, j, k - 1) - P(i - 1, j, k + 1) +
                P(i - 1, j, k - 1)) +
       c0[o] * P(i, j, k - 1) + c1[o] * P(i, j - 1, k) + c2[o] * P(i - 1, j, k) +
       wrk1[o];
  ss = (s0 * a3[o] - P(i, j, k)) * bnd[o];
  out[o] = P(i, j, k) + omega * ss;
#undef P
  return ss;
}

static double himeno_zrange(const float *p0, float *p1,
                            const float *const *coef, int nx, int ny, int z0,
                            int z1, float omega) {
  double g = 0.0;
  for (int k = z0; k < z1; ++k)
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        float ss = himeno_point(p0, p1, coef, i, j, k, nx, ny, omega);
        g += (double)(ss * ss);
      }
  return g;
}

void oracle_himeno_sweep(const float *p0, float *p1, const float *const *coef,
                         int nx, int ny, int nz, float omega, double *gosa) {
  double g = himeno_zrange(p0, p1, coef, nx, ny, 1, nz - 1, omega);
  if (gosa) *gosa = g;
}

void oracle_himeno_run(float *p0, float *p1, const float *const *coef, int nx,
                       int ny, int nz, int nn, float omega, double *gosa) {
  for (int i = 0; i < nn; ++i) {
    const float *in = (i % 2 == 0) ? p0 : p1;
    float *out = (i % 2 == 0) ? p1 : p0;
    oracle_himeno_sweep(in, out, coef, nx, ny, nz, omega,
                        gosa ? &gosa[i] : NULL);
  }
}

int oracle_himeno_run_omp(float *p0, float *p1, const float *const *coef,
                          int nx, int ny, int nz, int nn, float omega,
                          int div, double *gosa) {
  int threads = 1;
  if (div < 1) div = 1;
  double *part = gosa ? (double *)calloc((size_t)div, sizeof(double)) : NULL;
  for (int i = 0; i < nn; ++i) {
    const float *in = (i % 2 == 0) ? p0 : p1;
    float *out = (i % 2 == 0) ? p1 : p0;
#pragma omp parallel for schedule(static, 1)
    for (int u = 0; u < div; ++u) {
      int z0, z1;
      unit_range(1, nz - 1, div, u, &z0, &z1);
      double g = himeno_zrange(in, out, coef, nx, ny, z0, z1, omega);
      if (part) part[u] = g;
#ifdef _OPENMP
      if (u == 0) threads = omp_get_num_threads();
#endif
    }
    if (part) {
      /* unit partials (each in z-y-x order) folded in unit order */
      double g = 0.0;
      for (int u = 0; u < div; ++u) g += part[u];
      gosa[i] = g;
    }
  }
  free(part);
  return threads;
}

/* ------------------------------------------ red-black 7-point relaxation */
/* count iterations of physis_amd/programs/redblack7.c, each a red then a
 * black colour sweep in place over [1, n-1) of every axis, in the
 * REFERENCE target's loop order (reference_runtime_builder.cc:585-599,
 * 617-646: z, y, then x from min + ((min & 1) ^ ((c + y + z) % 2)) in
 * steps of 2, i.e. the points with x + y + z + c even).  coef: ce cw cn cs
 * ct cb cc, summed in the kernel's operand order. */
void oracle_redblack7_run(float *g, int nx, int ny, int nz, int count,
                          const float *coef) {
  const float ce = coef[0], cw = coef[1], cn = coef[2], cs = coef[3];
  const float ct = coef[4], cb = coef[5], cc = coef[6];
#define G(a, b, c) g[OFF(a, b, c, nx, ny)]
  for (int it = 0; it < count; ++it)
    for (int c = 0; c < 2; ++c)
      for (int z = 1; z < nz - 1; ++z)
        for (int y = 1; y < ny - 1; ++y)
          for (int x = 1 + ((1 & 1) ^ ((c + y + z) % 2)); x < nx - 1;
               x += 2) {
            float v = cc * G(x, y, z) + cw * G(x - 1, y, z) +
                      ce * G(x + 1, y, z) + cs * G(x, y + 1, z) +
                      cn * G(x, y - 1, z) + cb * G(x, y, z - 1) +
                      ct * G(x, y, z + 1);
            G(x, y, z) = v;
          }
#undef G
}

/* ------------------------------- 7-point map with 1-D axis coefficients */
/* One sweep of physis_amd/programs/mixed7.c over [1, n-1)^3, the kernel of
 * test_mixed-dim2.c:12-23 in its operand order: v = g + g[x-1] cx[x-1] +
 * g[x+1] cx[x+1] + g[y-1] cy[y-1] + g[y+1] cy[y+1] + g[z-1] cz[z-1] +
 * g[z+1] cz[z+1] (points outside keep o). */
void oracle_mixed7_sweep(const float *g, float *o, int nx, int ny, int nz,
                         const float *cx, const float *cy, const float *cz) {
#define G(a, b, c) g[OFF(a, b, c, nx, ny)]
  for (int z = 1; z < nz - 1; ++z)
    for (int y = 1; y < ny - 1; ++y)
      for (int x = 1; x < nx - 1; ++x)
        o[OFF(x, y, z, nx, ny)] =
            G(x, y, z) + G(x - 1, y, z) * cx[x - 1] +
            G(x + 1, y, z) * cx[x + 1] + G(x, y - 1, z) * cy[y - 1] +
            G(x, y + 1, z) * cy[y + 1] + G(x, y, z - 1) * cz[z - 1] +
            G(x, y, z + 1) * cz[z + 1];
#undef G
}

/* ------------------------------------------------- 2-D 5- and 9-point maps */
/* sweeps of physis_amd/programs/stencil2d.c, alternating a -> b and b -> a
 * (the first from a), each over [1, nx-1) x [1, ny-1) in the REFERENCE
 * target's loop order (y, then x): v = the points summed in the order of
 * test_5-pt-2d.c:13-18 / test_9-pt-2d.c:13-22, out = v * w.  Points outside
 * the domain keep their value. */
void oracle_stencil2d_run(float *a, float *b, int nx, int ny, int points,
                          float w, int sweeps) {
  for (int s = 0; s < sweeps; ++s) {
    const float *g = (s & 1) ? b : a;
    float *o = (s & 1) ? a : b;
#define G(i, j) g[(long)(j) * nx + (i)]
    for (int y = 1; y < ny - 1; ++y)
      for (int x = 1; x < nx - 1; ++x) {
        float v = G(x, y) + G(x - 1, y) + G(x + 1, y) + G(x, y - 1) +
                  G(x, y + 1);
        if (points == 9)
          v = v + G(x - 1, y - 1) + G(x + 1, y - 1) + G(x - 1, y + 1) +
              G(x + 1, y + 1);
        o[(long)y * nx + x] = v * w;
      }
#undef G
  }
}

/* ------------------------------------ 7-point stencil on user-defined points */
/* One sweep of physis_amd/programs/udt7.c over [1, n-1) of every axis:
 * q[c] = cc p[c] + cw p[x-1] + ce p[x+1] + cs p[y+1] + cn p[y-1] + cb p[z-1]
 * + ct p[z+1], left to right (points outside keep q).  p, q: the two
 * members of the points, each a dense array (the runtime's layout). */
void oracle_udt7_sweep(const float *p, float *q, int nx, int ny, int nz,
                       const float *coef) {
  const float ce = coef[0], cw = coef[1], cn = coef[2], cs = coef[3];
  const float ct = coef[4], cb = coef[5], cc = coef[6];
#define P(a, b, c) p[OFF(a, b, c, nx, ny)]
  for (int z = 1; z < nz - 1; ++z)
    for (int y = 1; y < ny - 1; ++y)
      for (int x = 1; x < nx - 1; ++x)
        q[OFF(x, y, z, nx, ny)] =
            cc * P(x, y, z) + cw * P(x - 1, y, z) + ce * P(x + 1, y, z) +
            cs * P(x, y + 1, z) + cn * P(x, y - 1, z) + cb * P(x, y, z - 1) +
            ct * P(x, y, z + 1);
#undef P
}

/* ---------------------------------------------------- system-test kernels */
#define FOR_DOMAIN(lo, hi)                  \
  for (int z = (lo); z < (hi); ++z)         \
    for (int y = (lo); y < (hi); ++y)       \
      for (int x = (lo); x < (hi); ++x)

void oracle_test_7pt(const float *g, float *o, int n) {
#define G(a, b, c) g[OFF(a, b, c, n, n)]
  FOR_DOMAIN(1, n - 1) {
    float v = G(x, y, z) + G(x + 1, y, z) + G(x - 1, y, z) + G(x, y + 1, z) +
              G(x, y - 1, z) + G(x, y, z - 1) + G(x, y, z + 1);
    o[OFF(x, y, z, n, n)] = v;
  }
#undef G
}

void oracle_test_7pt_periodic(const float *g, float *o, int n) {
#define G(a, b, c) g[OFF(wrap(a, n), wrap(b, n), wrap(c, n), n, n)]
  FOR_DOMAIN(0, n) {
    float v = G(x, y, z) + G(x + 1, y, z) + G(x - 1, y, z) + G(x, y + 1, z) +
              G(x, y - 1, z) + G(x, y, z + 1) + G(x, y, z - 1);
    o[OFF(x, y, z, n, n)] = v;
  }
#undef G
}

void oracle_test_7pt_neumann(const float *g, float *o, int n) {
#define G(a, b, c) g[OFF(a, b, c, n, n)]
  FOR_DOMAIN(0, n) {
    float c, w, e, nn, s, b, t;
    c = G(x, y, z);
    if (x == 0) w = G(x, y, z); else w = G(x - 1, y, z);
    if (x == n - 1) e = G(x, y, z); else e = G(x + 1, y, z);
    if (y == 0) nn = G(x, y, z); else nn = G(x, y - 1, z);
    if (y == n - 1) s = G(x, y, z); else s = G(x, y + 1, z);
    if (z == 0) b = G(x, y, z); else b = G(x, y, z - 1);
    if (z == n - 1) t = G(x, y, z); else t = G(x, y, z + 1);
    o[OFF(x, y, z, n, n)] = c + w + e + s + nn + b + t;
  }
#undef G
}

void oracle_test_asymmetric(const float *g, float *o, int n) {
#define G(a, b, c) g[OFF(a, b, c, n, n)]
  FOR_DOMAIN(2, n - 2) {
    float v = G(x, y, z) + G(x + 1, y + 1, z + 1) + G(x - 2, y - 2, z - 2);
    o[OFF(x, y, z, n, n)] = v;
  }
#undef G
}

void oracle_test_asymmetric_periodic(const float *g, float *o, int n) {
  FOR_DOMAIN(2, n - 2) {
    float v = g[OFF(x, y, z, n, n)] + g[OFF(x - 1, y, z, n, n)] +
              g[OFF(wrap(x + 2, n), y, z, n, n)];
    o[OFF(x, y, z, n, n)] = v;
  }
}

/* The 27 neighbours in the reference's listing order: z-1, z, z+1 planes,
 * each (0,0) (+1,0) (-1,0) (0,+1) (+1,+1) (-1,+1) (0,-1) (+1,-1) (-1,-1). */
static const int k27[9][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {1, 1},
                              {-1, 1}, {0, -1}, {1, -1}, {-1, -1}};

void oracle_test_27pt(const float *g, float *o, int n, int periodic) {
  int lo = periodic ? 0 : 1, hi = periodic ? n : n - 1;
  FOR_DOMAIN(lo, hi) {
    float v = 0.0f;
    int first = 1;
    for (int dz = -1; dz <= 1; ++dz)
      for (int k = 0; k < 9; ++k) {
        int a = x + k27[k][0], b = y + k27[k][1], c = z + dz;
        if (periodic) {
          a = wrap(a, n);
          b = wrap(b, n);
          c = wrap(c, n);
        }
        float e = g[OFF(a, b, c, n, n)];
        v = first ? e : v + e;
        first = 0;
      }
    o[OFF(x, y, z, n, n)] = v;
  }
}

void oracle_test_27pt_int(const int *g, int *o, int n) {
  FOR_DOMAIN(1, n - 1) {
    int v = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int k = 0; k < 9; ++k)
        v += g[OFF(x + k27[k][0], y + k27[k][1], z + dz, n, n)];
    float f = (float)v;
    o[OFF(x, y, z, n, n)] = (int)f;
  }
}
