/* Driver for the reference's original Himeno benchmark, built by
 * oracle/Makefile from examples/himeno/himenobmtxpa_original.c where it
 * lies (its main renamed away).  Test infrastructure only.
 *
 * Usage: himeno_ref NX NY NZ NN SEED > out
 * Builds the 14 Himeno arrays from a seeded generator in the reference's
 * Mat layout (slowest index i = Physis z, fastest k = Physis x), runs the
 * reference's jacobi() for NN sweeps and prints "gosa %.9e" and the final p
 * in Physis order (x fastest) as raw float bits in hex, one per line. */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>

/* Layout of the benchmark's matrix struct (himenobmtxpa_original.c:44-50). */
struct Mat {
  float *m;
  int mnums;
  int mrows;
  int mcols;
  int mdeps;
};
typedef struct Mat Matrix;
extern Matrix a, b, c, p, bnd, wrk1, wrk2;
extern float omega;
int newMat(Matrix *Mat, int mnums, int mrows, int mcols, int mdeps);
float jacobi(int nn, Matrix *a, Matrix *b, Matrix *c, Matrix *p, Matrix *bnd,
             Matrix *wrk1, Matrix *wrk2);

/* splitmix64 -> uniform float in [lo, hi) (same generator as
 * oracle/seeded.py). */
static uint64_t sm_state;
static uint64_t splitmix64(void) {
  uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static float uniform01(void) { return (float)(splitmix64() >> 40) * (1.0f / 16777216.0f); }

/* Element (x, y, z) of the Physis-order array lives at MR(mat, l, z, y, x). */
#define MIDX(M, l, r, c_, d) \
  ((size_t)(l) * (M).mrows * (M).mcols * (M).mdeps + (size_t)(r) * (M).mcols * (M).mdeps + (size_t)(c_) * (M).mdeps + (d))

int main(int argc, char **argv) {
  if (argc != 6) { fprintf(stderr, "usage: himeno_ref NX NY NZ NN SEED\n"); return 2; }
  int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]), nn = atoi(argv[4]);
  sm_state = strtoull(argv[5], NULL, 10);
  newMat(&p, 1, nz, ny, nx);
  newMat(&bnd, 1, nz, ny, nx);
  newMat(&wrk1, 1, nz, ny, nx);
  newMat(&wrk2, 1, nz, ny, nx);
  newMat(&a, 4, nz, ny, nx);
  newMat(&b, 3, nz, ny, nx);
  newMat(&c, 3, nz, ny, nx);
  /* Fill order = the Physis arrays a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 bnd wrk1 p,
   * each x fastest, values drawn in that order. */
  Matrix *mats[13] = {&a, &a, &a, &a, &b, &b, &b, &c, &c, &c, &bnd, &wrk1, &p};
  int sub[13] = {0, 1, 2, 3, 0, 1, 2, 0, 1, 2, 0, 0, 0};
  for (int g = 0; g < 13; ++g)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
          mats[g]->m[MIDX(*mats[g], sub[g], z, y, x)] = uniform01();
  memcpy(wrk2.m, p.m, sizeof(float) * (size_t)nx * ny * nz);
  float gosa = jacobi(nn, &a, &b, &c, &p, &bnd, &wrk1, &wrk2);
  printf("gosa %.9e\n", gosa);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        float v = p.m[MIDX(p, 0, z, y, x)];
        uint32_t u;
        memcpy(&u, &v, 4);
        printf("%08x\n", u);
      }
  return 0;
}
