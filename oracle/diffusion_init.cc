// Closed-form state of the 7-point diffusion benchmark, restated for the
// oracle (test infrastructure only).  The cosine solution of
// examples/diffusion-benchmark/diffusion3d.cc:14-39 and the coefficients of
// diffusion3d.h:47-61 (kappa = 0.1, unit cube, dt = 0.1*dx^2/kappa), written
// in C++ so that float arguments pick the same math overloads.
#include <math.h>

namespace {
struct Params {
  float kappa, dx, dy, dz, kx, ky, kz, dt;
  float ce, cw, cn, cs, ct, cb, cc;
};

Params Make(int nx, int ny, int nz) {
  Params p;
  p.kappa = 0.1f;
  float l = 1.0f;
  p.dx = l / nx;
  p.dy = l / ny;
  p.dz = l / nz;
  p.kx = p.ky = p.kz = 2.0 * 3.1415926535897932384626;
  p.dt = 0.1 * p.dx * p.dx / p.kappa;
  p.ce = p.cw = p.kappa * p.dt / (p.dx * p.dx);
  p.cn = p.cs = p.kappa * p.dt / (p.dy * p.dy);
  p.ct = p.cb = p.kappa * p.dt / (p.dz * p.dz);
  p.cc = 1.0 - (p.ce + p.cw + p.cn + p.cs + p.ct + p.cb);
  return p;
}
}  // namespace

extern "C" {

// coef = {ce, cw, cn, cs, ct, cb, cc, dt}
void oracle_diffusion_coefs(int nx, int ny, int nz, float *coef) {
  Params p = Make(nx, ny, nz);
  coef[0] = p.ce; coef[1] = p.cw; coef[2] = p.cn; coef[3] = p.cs;
  coef[4] = p.ct; coef[5] = p.cb; coef[6] = p.cc; coef[7] = p.dt;
}

// Analytic solution after `count` steps of dt (count = 0: initial state).
void oracle_diffusion_init(float *buf, int nx, int ny, int nz, int count) {
  Params p = Make(nx, ny, nz);
  float time = count * p.dt;
  float ax = exp(-p.kappa * time * (p.kx * p.kx));
  float ay = exp(-p.kappa * time * (p.ky * p.ky));
  float az = exp(-p.kappa * time * (p.kz * p.kz));
  for (int jz = 0; jz < nz; jz++)
    for (int jy = 0; jy < ny; jy++)
      for (int jx = 0; jx < nx; jx++) {
        int j = jz * nx * ny + jy * nx + jx;
        float x = p.dx * ((float)(jx + 0.5));
        float y = p.dy * ((float)(jy + 0.5));
        float z = p.dz * ((float)(jz + 0.5));
        float f0 = (float)0.125 * (1.0 - ax * cos(p.kx * x)) *
                   (1.0 - ay * cos(p.ky * y)) * (1.0 - az * cos(p.kz * z));
        buf[j] = f0;
      }
}

}  // extern "C"
