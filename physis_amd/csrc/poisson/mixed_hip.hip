// libphysis_poisson_hip: the precision tier (mixed-precision iterative
// refinement) and the plain FP32 Jacobi solve of the 19-point Himeno
// operator.  The operator, residual and batch layout are defined in
// include/physis/physis_poisson.h.
//
// Kernels (all 2.5D z-marching, one LDS ring of planes per staged grid,
// built from the same Stage / BufWin blocks as the Physis run kernels):
//   SweepKernel<M>     one Jacobi sweep; p and the 12 coefficient/rhs grids
//                      stored as M (float, bfloat16 or binary16), FP32
//                      arithmetic in the operand order of the Himeno
//                      jacobi_kernel (reference examples/himeno/
//                      himenobmtxpa_physis.c:340-359).  Bytes per point:
//                      14 * sizeof(M).  The 12 centre-only fields are read
//                      by the compute waves only, non-temporally when they
//                      exceed the last-level cache.
//   ResidualKernel     FP64 residual r = (s0 - p / a3) bnd from FP32
//                      coefficients, optionally fused with the refinement
//                      update p' = p + E / sigma (written to a second FP64
//                      grid) and with the write of the next inner
//                      right-hand side sigma' r in storage type M; per-block
//                      partial sums of r^2, max |r| and (bnd wrk1)^2.
// Each block owns a 64*VX x (BY-2) column tile and a z chunk that never
// straddles two problems of a batch, so partials fold per problem.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "physis/hip/run.h"
#include "physis/hip/stencil.h"
#include "physis/physis_poisson.h"
#include "../runtime/runtime_hip.h"

#pragma clang fp contract(off)

namespace physis_poisson {

using physis_hip::BufWin;
using physis_hip::Convert;
using physis_hip::RingSlot;
using physis_hip::Stage;
using physis_hip::TileCoord;
using physis_hip::TileOf;
using physis_hip::bf16_t;
using physis_hip::fp16_t;
using physis_hip::kWave;

constexpr int kBY = 8;    // waves per block: 6 compute rows + 2 halo rows
constexpr int kPD = 2;    // staged planes in flight ahead of the ring
constexpr int kNumF = 12; // a0 a1 a2 a3 b0 b1 b2 c0 c1 c2 bnd rhs/wrk1
// A solve stops as stagnated when 20 consecutive residual checks spanning
// at least this many sweeps fail to lower the best residual by 0.1%: Jacobi
// on large grids gains that little per check long before it stops gaining.
constexpr long kStallSweeps = 20000;

// Block -> (x tile, y tile, problem q, chunk j of problem q).
struct Plan {
  int tiles_x, tiles_y;
  int tz;      // z chunks per problem
  int zchunk;  // planes per chunk
  int nzs;     // planes per problem
  int blocks;
};

template <typename M>
struct Fields {
  const M *g[kNumF];
};

// The 9 LDS row segments a 19-point point reads (plane z-1/z/z+1, row
// y-1/y/y+1); segments with x-1 .. x+1 are VX + 2 long, seg[i + 1] = x + i.
template <typename T, int VX>
struct Rows19 {
  T zp_y[VX + 2], zp_yp[VX], zp_ym[VX];
  T z_y[VX + 2], z_yp[VX + 2], z_ym[VX + 2];
  T zm_y[VX + 2], zm_yp[VX], zm_ym[VX];

  template <typename S>
  __device__ __forceinline__ void Read(const T *lds, int base, int wave,
                                       int lane) {
    const T *zm = lds + RingSlot<4>(base, 0) * S::kPlane;
    const T *zc = lds + RingSlot<4>(base, 1) * S::kPlane;
    const T *zp = lds + RingSlot<4>(base, 2) * S::kPlane;
    S::template ReadRow<-1, 1>(zp, wave, lane, zp_y);
    S::template ReadRow<0, 0>(zp, wave + 1, lane, zp_yp);
    S::template ReadRow<0, 0>(zp, wave - 1, lane, zp_ym);
    S::template ReadRow<-1, 1>(zc, wave, lane, z_y);
    S::template ReadRow<-1, 1>(zc, wave + 1, lane, z_yp);
    S::template ReadRow<-1, 1>(zc, wave - 1, lane, z_ym);
    S::template ReadRow<-1, 1>(zm, wave, lane, zm_y);
    S::template ReadRow<0, 0>(zm, wave + 1, lane, zm_yp);
    S::template ReadRow<0, 0>(zm, wave - 1, lane, zm_ym);
  }

  // s0 of point i with coefficients c[0..11] (a0 a1 a2 a3 b0 b1 b2 c0 c1 c2
  // bnd wrk1), summed left to right as the Himeno kernel writes it.
  __device__ __forceinline__ T S0(const T (&c)[kNumF], int i) const {
    return c[0] * zp_y[i + 1] + c[1] * z_yp[i + 1] + c[2] * z_y[i + 2] +
           c[4] * (zp_yp[i] - zp_ym[i] - zm_yp[i] + zm_ym[i]) +
           c[5] * (z_yp[i + 2] - z_ym[i + 2] - z_yp[i] + z_ym[i]) +
           c[6] * (zp_y[i + 2] - zm_y[i + 2] - zp_y[i] + zm_y[i]) +
           c[7] * zm_y[i + 1] + c[8] * z_ym[i + 1] + c[9] * z_y[i] + c[11];
  }
  __device__ __forceinline__ T Centre(int i) const { return z_y[i + 1]; }
};

// z range of a block: chunk j of problem q.
struct ZRange {
  int zb, ze;
};
__device__ __forceinline__ ZRange BlockZ(const Plan &pl, int bz) {
  const int q = bz / pl.tz, j = bz - q * pl.tz;
  ZRange r;
  r.zb = q * pl.nzs + j * pl.zchunk;
  r.ze = min(r.zb + pl.zchunk, (q + 1) * pl.nzs);
  return r;
}

// ------------------------------------------------------------- Jacobi sweep
// LAUX: cache policy of the coefficient loads (physis_hip::kStreamLoad when
// the fields exceed the last-level cache, else 0: re-read every sweep, they
// then stay cached).
template <typename M, int VX, int LAUX>
__global__ __launch_bounds__(kWave *kBY) void SweepKernel(
    const M *__restrict__ pin, M *__restrict__ pout, Fields<M> f, int nx,
    int ny, int nz, float omega, Plan pl) {
  constexpr int TX = kWave * VX, TY = kBY - 2;
  typedef Stage<float, VX, kBY, 1, 1, 4, false, M> S;
  __shared__ __attribute__((aligned(16))) float lds[S::kLdsElems];
  const int lane = threadIdx.x % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const TileCoord tc = TileOf(pl.tiles_x, pl.tiles_y);
  const ZRange zr = BlockZ(pl, tc.bz);
  const int zb = zr.zb, ze = zr.ze;
  if (zb >= ze) return;  // uniform over the block
  const int x0 = tc.bx * TX, y0 = tc.by * TY;
  const int yrow = y0 - 1 + wave;
  const bool computes = wave >= 1 && wave < kBY - 1 && yrow < ny;
  const int xs = x0 + lane * VX;
  // centre-only grids: halo waves issue the same loads at an out-of-range
  // offset, which the buffer range check drops (no memory access), so
  // every wave keeps the same count of loads in flight
  const bool dl = wave >= 1 && wave < kBY - 1;
  const int drow = dl ? yrow : y0;
  const int nxy = nx * ny, n = nxy * nz;
  const int dim[3] = {nx, ny, nz};
  const BufWin<M> pw =
      BufWin<M>::Make(pin, n, (long)(zb - 2) * nxy, (long)(ze + 2) * nxy);
  S st[kPD];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    st[0].Load(pw, pin, dim, x0, yrow, zb - 1 + k, lane);
    st[0].Store(lds + k * S::kPlane, wave, lane);
  }
#pragma unroll
  for (int k = 0; k < kPD; ++k)
    st[k].Load(pw, pin, dim, x0, yrow, zb + 2 + k, lane);
  int base = 0;
  const BufWin<M> ow =
      BufWin<M>::Make(pout, n, (long)zb * nxy, (long)ze * nxy);
  BufWin<M> fw[kNumF];
  M fv[kNumF][2][VX];
  const int doff = xs + drow * nx;
#pragma unroll
  for (int g = 0; g < kNumF; ++g) {
    fw[g] = BufWin<M>::Make(f.g[g], n, (long)zb * nxy, (long)ze * nxy);
    fw[g].template Load<VX, LAUX>(
        dl ? doff + zb * nxy : physis_hip::kDrop, fv[g][0]);
  }
  __syncthreads();
  for (int zc = zb; zc < ze; zc += kPD) {
#pragma unroll
    for (int k = 0; k < kPD; ++k) {
      const int z = zc + k, cur = k & 1, nxt = cur ^ 1;
#pragma unroll
      for (int g = 0; g < kNumF; ++g)
        fw[g].template Load<VX, LAUX>(
            dl ? doff + (z + 1) * nxy : physis_hip::kDrop, fv[g][nxt]);
      if (computes && z < ze) {
        Rows19<float, VX> r;
        r.template Read<S>(lds, base, wave, lane);
        M out[VX];
        unsigned mask = 0;
#pragma unroll
        for (int i = 0; i < VX; ++i) {
          float c[kNumF];
#pragma unroll
          for (int g = 0; g < kNumF; ++g) c[g] = Convert<float, M>(fv[g][cur][i]);
          const float pc = r.Centre(i);
          const float s0 = r.S0(c, i);
          const float ss = (s0 * c[3] - pc) * c[10];
          // bnd = 0 points keep their value whatever their neighbours hold
          out[i] = Convert<M, float>(c[10] != 0.0f ? pc + omega * ss : pc);
          if (xs + i < nx) mask |= 1u << i;
        }
        physis_hip::StoreMasked<M, VX>(ow, xs + yrow * nx + z * nxy, out,
                                       mask);
      }
      st[k].Store(lds + RingSlot<4>(base, 3) * S::kPlane, wave, lane);
      __syncthreads();
      st[k].Load(pw, pin, dim, x0, yrow, z + 2 + kPD, lane);
      base = RingSlot<4>(base, 1);
    }
  }
}

// s0 of the Himeno operator at one point whose neighbours come from p(dx,
// dy, dz), in the operand order of Rows19::S0.
template <typename P>
__device__ __forceinline__ float S0At(const float (&c)[kNumF], P p) {
  return c[0] * p(0, 0, 1) + c[1] * p(0, 1, 0) + c[2] * p(1, 0, 0) +
         c[4] * (p(0, 1, 1) - p(0, -1, 1) - p(0, 1, -1) + p(0, -1, -1)) +
         c[5] * (p(1, 1, 0) - p(1, -1, 0) - p(-1, 1, 0) + p(-1, -1, 0)) +
         c[6] * (p(1, 0, 1) - p(1, 0, -1) - p(-1, 0, 1) + p(-1, 0, -1)) +
         c[7] * p(0, 0, -1) + c[8] * p(0, -1, 0) + c[9] * p(-1, 0, 0) + c[11];
}

// One Jacobi update in storage precision: the value a sweep stores (M) at
// a point with centre pc, returned as the float the next sweep loads.
template <typename M>
__device__ __forceinline__ M UpdateM(const float (&c)[kNumF], float pc,
                                     float s0, float omega) {
  const float ss = (s0 * c[3] - pc) * c[10];
  // bnd = 0 points keep their value whatever their neighbours hold
  return Convert<M, float>(c[10] != 0.0f ? pc + omega * ss : pc);
}

// ------------------------------------------------------ two-sweep passes
// Two Jacobi sweeps per launch (temporal blocking; the technique of the
// reference's examples/diffusion-benchmark/
// diffusion3d_cuda_temporal_blocking.cu:78-161): every field is read once
// per two sweeps.  Ring A holds the source planes s-1..s+1 with a
// two-point halo, ring B the intermediate planes s-3..s-1 with a one-point
// halo, as the floats of the storage-type values the first sweep would
// store (so the pass is bit-identical to two SweepKernel launches).  Step s
// computes E'(s) on rows y0-1 .. y0+TY (the columns x0-1 and x0+TX by lanes
// of wave 0) and E''(s-2) on the tile; one barrier per step.  Points with
// bnd = 0 (every problem's faces, and points outside the grid, whose
// fields load as 0) keep their value, so the E' values the pass computes
// beyond the grid feed only points that ignore them.  With 16-bit storage
// and two points per lane the fields of plane s are read once, one step
// ahead of E'(s), and stay in a register ring of 4 planes until E''(s) two
// steps later (the rows of E'' are rows of E'); otherwise (the ring would
// not fit the register budget) they are read again for E''.
template <typename M, int VX, int BY>
__global__ __launch_bounds__(kWave *BY) void TwoSweepKernel(
    const M *__restrict__ pin, M *__restrict__ pout, Fields<M> f, int nx,
    int ny, int nz, float omega, Plan pl) {
  constexpr int TX = kWave * VX, TY = BY - 4;
  typedef Stage<float, VX, BY, 2, 2, 4, false, M> SA;
  typedef Stage<float, VX, BY, 1, 1, 4, false> SB;
  __shared__ __attribute__((aligned(16))) float a_lds[SA::kLdsElems];
  __shared__ __attribute__((aligned(16))) float b_lds[SB::kLdsElems];
  const int lane = threadIdx.x % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const TileCoord tc = TileOf(pl.tiles_x, pl.tiles_y);
  const ZRange zr = BlockZ(pl, tc.bz);
  const int zb = zr.zb, ze = zr.ze;
  if (zb >= ze) return;  // uniform over the block
  const int x0 = tc.bx * TX, y0 = tc.by * TY;
  const int yrow = y0 - 2 + wave;
  const bool mid_row = wave >= 1 && wave < BY - 1;
  const bool out_row = wave >= 2 && wave < BY - 2 && yrow < ny;
  const int xs = x0 + lane * VX;
  const bool hlane = wave == 0 && lane < 2 * (BY - 2);
  const int hr = 1 + (lane >> 1);
  const int px = (lane & 1) ? TX : -1;
  const int hx = x0 + px, hy = y0 - 2 + hr;
  const int nxy = nx * ny, n = nxy * nz;
  const int dim[3] = {nx, ny, nz};
  const BufWin<M> aw =
      BufWin<M>::Make(pin, n, (long)(zb - 3) * nxy, (long)(ze + 3) * nxy);
  const BufWin<M> ow =
      BufWin<M>::Make(pout, n, (long)zb * nxy, (long)ze * nxy);
  BufWin<M> fw[kNumF];
#pragma unroll
  for (int g = 0; g < kNumF; ++g)
    fw[g] = BufWin<M>::Make(f.g[g], n, (long)(zb - 1) * nxy,
                            (long)(ze + 1) * nxy);
  SA st[kPD];
  const typename SA::Addr ra = SA::Prepare(dim, x0, yrow, lane);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    st[0].Load(aw, pin, dim, ra, zb - 2 + k);
    st[0].Store(a_lds + k * SA::kPlane, wave, lane);
  }
#pragma unroll
  for (int k = 0; k < kPD; ++k) st[k].Load(aw, pin, dim, ra, zb + 1 + k);
  constexpr bool kRing = sizeof(M) == 2 && VX == 2;
  // kRing: plane s in slot (s - zb + 1) & 3; else slots 0/1 alternate
  M f1[kNumF][kRing ? 4 : 2][VX], f2[kNumF][kRing ? 1 : VX];
  // fields of plane z for E': the lane's VX points on rows y0-1 .. y0+TY,
  // its halo point in wave 0
  auto load1 = [&](int z, int slot) {
    if (wave == 0) {
      const int o = hlane ? hx + hy * nx + z * nxy : physis_hip::kDrop;
#pragma unroll
      for (int g = 0; g < kNumF; ++g)
        fw[g].template Load<1>(o, *reinterpret_cast<M(*)[1]>(&f1[g][slot][0]));
    } else {
      const int o = mid_row ? xs + yrow * nx + z * nxy : physis_hip::kDrop;
#pragma unroll
      for (int g = 0; g < kNumF; ++g)
        fw[g].template Load<VX>(o, f1[g][slot]);
    }
  };
  load1(zb - 1, 0);
  __syncthreads();
  for (int sc = zb - 1; sc <= ze + 1; sc += 4) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int s = sc + k;
      const int cur = kRing ? k : k & 1;
      const int nxt = kRing ? (k + 1) & 3 : cur ^ 1;
      if (s > ze + 1) break;  // uniform over the block
      load1(s + 1, nxt);
      if constexpr (!kRing) {
        const int o2 = out_row && s - 2 >= zb
                           ? xs + yrow * nx + (s - 2) * nxy
                           : physis_hip::kDrop;
#pragma unroll
        for (int g = 0; g < kNumF; ++g)
          fw[g].template Load<VX, physis_hip::kStreamLoad>(o2, f2[g]);
      }
      if (mid_row && s <= ze) {
        Rows19<float, VX> r;
        r.template Read<SA>(a_lds, k, wave, lane);
        float *brow = b_lds + RingSlot<4>(k, 3) * SB::kPlane + wave * SB::kRow +
                      SB::kPad + lane * VX;
#pragma unroll
        for (int i = 0; i < VX; ++i) {
          float c[kNumF];
#pragma unroll
          for (int g = 0; g < kNumF; ++g)
            c[g] = Convert<float, M>(f1[g][cur][i]);
          brow[i] = Convert<float, M>(
              UpdateM<M>(c, r.Centre(i), r.S0(c, i), omega));
        }
      }
      if (hlane && s <= ze) {
        const auto p = [&](int dx, int dy, int dz) {
          return a_lds[RingSlot<4>(k, dz + 1) * SA::kPlane +
                       (hr + dy) * SA::kRow + SA::kPad + px + dx];
        };
        float c[kNumF];
#pragma unroll
        for (int g = 0; g < kNumF; ++g) c[g] = Convert<float, M>(f1[g][cur][0]);
        b_lds[RingSlot<4>(k, 3) * SB::kPlane + hr * SB::kRow + SB::kPad + px] =
            Convert<float, M>(UpdateM<M>(c, p(0, 0, 0), S0At(c, p), omega));
      }
      if (out_row && s - 2 >= zb) {
        Rows19<float, VX> r;
        r.template Read<SB>(b_lds, k, wave, lane);
        M out[VX];
        unsigned mask = 0;
#pragma unroll
        for (int i = 0; i < VX; ++i) {
          float c[kNumF];
#pragma unroll
          for (int g = 0; g < kNumF; ++g)
            c[g] = Convert<float, M>(kRing ? f1[g][(k + 2) & 3][i]
                                           : f2[g][kRing ? 0 : i]);
          out[i] = UpdateM<M>(c, r.Centre(i), r.S0(c, i), omega);
          if (xs + i < nx) mask |= 1u << i;
        }
        physis_hip::StoreMasked<M, VX>(ow, xs + yrow * nx + (s - 2) * nxy,
                                       out, mask);
      }
      st[k % kPD].Store(a_lds + RingSlot<4>(k, 3) * SA::kPlane, wave, lane);
      __syncthreads();
      st[k % kPD].Load(aw, pin, dim, ra, s + 2 + kPD);
    }
  }
}

// ----------------------------------------------------------- FP64 residual
// PT: element type of p (float or double).  ME: storage type of E and RHS.
// UPDATE: p' = p + inv_sigma * E formed on load, written to pout, residual
// of p'.  WRHS: RHS = M(float(sigma * r)).
template <typename PT, typename ME, int VX, bool UPDATE, bool WRHS>
__global__ __launch_bounds__(kWave *kBY) void ResidualKernel(
    const PT *__restrict__ pin, double *__restrict__ pout,
    const ME *__restrict__ e, ME *__restrict__ rhs, Fields<float> f, int nx,
    int ny, int nz, double inv_sigma, double sigma, Plan pl,
    double *__restrict__ partials) {
  static_assert(!UPDATE || __is_same(PT, double), "update needs FP64 p");
  constexpr int TX = kWave * VX, TY = kBY - 2;
  typedef Stage<double, VX, kBY, 1, 1, 4, false, PT> S;
  typedef Stage<double, VX, kBY, 1, 1, 4, false, ME> SE;
  __shared__ __attribute__((aligned(16))) double lds[S::kLdsElems];
  __shared__ double red[3][kBY];
  const int lane = threadIdx.x % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const TileCoord tc = TileOf(pl.tiles_x, pl.tiles_y);
  const ZRange zr = BlockZ(pl, tc.bz);
  const int zb = zr.zb, ze = zr.ze;
  double acc_r2 = 0.0, acc_max = 0.0, acc_b2 = 0.0;
  if (zb < ze) {
    const int x0 = tc.bx * TX, y0 = tc.by * TY;
    const int yrow = y0 - 1 + wave;
    const bool computes = wave >= 1 && wave < kBY - 1 && yrow < ny;
    const int xs = x0 + lane * VX;
    const int drow = wave < 1 ? y0 : (wave >= kBY - 1 ? y0 + TY - 1 : yrow);
    const int nxy = nx * ny, n = nxy * nz;
    const int dim[3] = {nx, ny, nz};
    const BufWin<PT> pw =
        BufWin<PT>::Make(pin, n, (long)(zb - 2) * nxy, (long)(ze + 2) * nxy);
    BufWin<ME> ew;
    if constexpr (UPDATE)
      ew = BufWin<ME>::Make(e, n, (long)(zb - 2) * nxy, (long)(ze + 2) * nxy);
    S st[kPD];
    SE se[kPD];
    auto load = [&](int slot, int z) {
      st[slot].Load(pw, pin, dim, x0, yrow, z, lane);
      if constexpr (UPDATE) se[slot].Load(ew, e, dim, x0, yrow, z, lane);
    };
    auto store = [&](int slot, double *dst) {
      if constexpr (UPDATE) {
#pragma unroll
        for (int i = 0; i < VX; ++i)
          st[slot].own[i] =
              st[slot].own[i] + Convert<double, ME>(se[slot].own[i]) * inv_sigma;
        st[slot].halo =
            st[slot].halo + Convert<double, ME>(se[slot].halo) * inv_sigma;
      }
      st[slot].Store(dst, wave, lane);
    };
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      load(0, zb - 1 + k);
      store(0, lds + k * S::kPlane);
    }
#pragma unroll
    for (int k = 0; k < kPD; ++k) load(k, zb + 2 + k);
    int base = 0;
    BufWin<double> ow;
    if constexpr (UPDATE)
      ow = BufWin<double>::Make(pout, n, (long)zb * nxy, (long)ze * nxy);
    BufWin<ME> rw;
    if constexpr (WRHS)
      rw = BufWin<ME>::Make(rhs, n, (long)zb * nxy, (long)ze * nxy);
    BufWin<float> fw[kNumF];
    float fv[kNumF][2][VX];
    const int doff = xs + drow * nx;
#pragma unroll
    for (int g = 0; g < kNumF; ++g) {
      fw[g] = BufWin<float>::Make(f.g[g], n, (long)zb * nxy, (long)ze * nxy);
      fw[g].template Load<VX>(doff + zb * nxy, fv[g][0]);
    }
    __syncthreads();
    for (int zc = zb; zc < ze; zc += kPD) {
#pragma unroll
      for (int k = 0; k < kPD; ++k) {
        const int z = zc + k, cur = k & 1, nxt = cur ^ 1;
#pragma unroll
        for (int g = 0; g < kNumF; ++g)
          fw[g].template Load<VX>(doff + (z + 1) * nxy, fv[g][nxt]);
        if (computes && z < ze) {
          Rows19<double, VX> r;
          r.template Read<S>(lds, base, wave, lane);
          double pn[VX];
          ME rv[VX];
          unsigned mask = 0;
#pragma unroll
          for (int i = 0; i < VX; ++i) {
            double c[kNumF];
#pragma unroll
            for (int g = 0; g < kNumF; ++g) c[g] = (double)fv[g][cur][i];
            const double pc = r.Centre(i);
            const double s0 = r.S0(c, i);
            const double rr = c[10] != 0.0 ? (s0 - pc / c[3]) * c[10] : 0.0;
            pn[i] = pc;
            rv[i] = Convert<ME, float>((float)(rr * sigma));
            if (xs + i < nx) {
              mask |= 1u << i;
              const double bb = c[11] * c[10];
              acc_r2 += rr * rr;
              acc_max = fmax(acc_max, fabs(rr));
              acc_b2 += bb * bb;
            }
          }
          const int off = xs + yrow * nx + z * nxy;
          if constexpr (UPDATE)
            physis_hip::StoreMasked<double, VX>(ow, off, pn, mask);
          if constexpr (WRHS)
            physis_hip::StoreMasked<ME, VX>(rw, off, rv, mask);
        }
        store(k, lds + RingSlot<4>(base, 3) * S::kPlane);
        __syncthreads();
        load(k, z + 2 + kPD);
        base = RingSlot<4>(base, 1);
      }
    }
  }
  const double r2 =
      physis_hip::BlockReduce<double, PS_SUM, kWave * kBY>(acc_r2, red[0]);
  const double rm =
      physis_hip::BlockReduce<double, PS_MAX, kWave * kBY>(acc_max, red[1]);
  const double b2 =
      physis_hip::BlockReduce<double, PS_SUM, kWave * kBY>(acc_b2, red[2]);
  if (threadIdx.x == 0) {
    const int logical = tc.bx + pl.tiles_x * (tc.by + pl.tiles_y * tc.bz);
    partials[3 * logical + 0] = r2;
    partials[3 * logical + 1] = rm;
    partials[3 * logical + 2] = b2;
  }
}

// --------------------------------------------------- storage conversion
template <typename M>
__global__ void ToStorageKernel(const float *__restrict__ in,
                                M *__restrict__ out, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = Convert<M, float>(in[i]);
}

// ------------------------------------------------------------------- host
// Tiles nx x ny into 64*VX x (BY-2) tiles and picks z chunks per problem
// minimising waves * (chunk + 2), as the Physis run planner does.  Chunks
// are capped so that every buffer window (the chunk plus two planes per side
// of the widest, FP64, field) stays within the 2 GiB reach of a buffer range.
Plan MakePlan(int nx, int ny, int nzs, int nbatch, int vx, int per_cu,
              int ty = kBY - 2) {
  Plan p{};
  const long max_chunk =
      (long)INT32_MAX / ((long)nx * ny * (long)sizeof(double)) - 4;
  if (max_chunk < 1)
    PS_FATAL("Poisson planes of %d x %d points exceed the reach of a buffer "
             "window", nx, ny);
  const int tx = kWave * vx;
  p.tiles_x = (nx + tx - 1) / tx;
  p.tiles_y = (ny + ty - 1) / ty;
  p.nzs = nzs;
  const long xy = (long)p.tiles_x * p.tiles_y * nbatch;
  const long resident = (long)__PSHipNumCUs() * std::max(per_cu, 1);
  long best = -1;
  for (int tz = 1; tz <= nzs; ++tz) {
    const int chunk = (nzs + tz - 1) / tz;
    if (chunk > max_chunk) continue;
    if ((long)(tz - 1) * chunk >= nzs) continue;
    const long blocks = xy * tz;
    if (blocks > (1L << 28)) break;
    const long cost = ((blocks + resident - 1) / resident) * (chunk + 2);
    if (best < 0 || cost < best) {
      best = cost;
      p.tz = tz;
      p.zchunk = chunk;
    }
  }
  p.blocks = (int)(xy * p.tz);
  return p;
}

}  // namespace physis_poisson

using namespace physis_poisson;

struct __PSPoisson {
  int nx, ny, nz, nzs, nbatch, storage;
  size_t elems, esize;
  __PSGrid *coef[kNumF];
  void *lo[kNumF - 1];  // storage-type a0..c2, bnd (FP32: the masters)
  bool own_lo;
  void *rhs, *e[2];
  double *p2 = nullptr;  // FP64 partner grid of the refinement
  float *t32 = nullptr;  // FP32 partner grid of the plain solve
  double *partials = nullptr;
  double *host_partials = nullptr;
  int sweep_vx;
  Plan sweep_plan, res_plan;
  int sweep2_vx;     // two-sweep passes
  Plan sweep2_plan;
};

namespace {

hipStream_t Stream() { return (hipStream_t)__PSHipStream(); }

void CheckLaunch() { PS_HIP_CHECK(hipGetLastError()); }

template <typename M>
Fields<M> InnerFields(const __PSPoisson *s) {
  Fields<M> f;
  for (int g = 0; g < kNumF - 1; ++g) f.g[g] = (const M *)s->lo[g];
  f.g[kNumF - 1] = (const M *)s->rhs;
  return f;
}

Fields<float> MasterFields(const __PSPoisson *s) {
  Fields<float> f;
  for (int g = 0; g < kNumF; ++g) f.g[g] = (const float *)s->coef[g]->p;
  return f;
}

// MI355X's last-level cache (AMD Infinity Cache): 256 MiB.
constexpr size_t kLastLevelCacheBytes = size_t(256) << 20;

template <typename M, int VX>
void LaunchSweepVX(const __PSPoisson *s, const M *in, M *out, Fields<M> f,
                   float omega) {
  const dim3 grid(s->sweep_plan.blocks), block(kWave * kBY);
  const size_t field_bytes =
      (size_t)kNumF * s->nx * s->ny * s->nz * sizeof(M);
  if (field_bytes > kLastLevelCacheBytes)
    hipLaunchKernelGGL((SweepKernel<M, VX, physis_hip::kStreamLoad>), grid,
                       block, 0, Stream(), in, out, f, s->nx, s->ny, s->nz,
                       omega, s->sweep_plan);
  else
    hipLaunchKernelGGL((SweepKernel<M, VX, 0>), grid, block, 0, Stream(), in,
                       out, f, s->nx, s->ny, s->nz, omega, s->sweep_plan);
}

template <typename M>
void LaunchSweep(const __PSPoisson *s, const M *in, M *out, Fields<M> f,
                 float omega) {
  if (s->sweep_vx == 4)
    LaunchSweepVX<M, 4>(s, in, out, f, omega);
  else
    LaunchSweepVX<M, 2>(s, in, out, f, omega);
  CheckLaunch();
}

// Waves per block of the two-sweep kernel (TY = BY - 4 output rows) and its
// points per lane.  The fields take 2 * VX values of E' in flight and VX of
// E'' per lane.  Measured at 512^3 (ms per sweep; pairs of single sweeps:
// 0.807 BF16, 1.448 FP32): 16-bit fields run best at VX = 2 in 12-wave
// blocks (0.655; 130 VGPRs, three waves per SIMD; VX = 4 in 8-wave blocks
// 0.827, in 12-wave blocks 0.732 with spills), FP32 fields at VX = 4 in
// 8-wave blocks (0.984; VX = 2 in 12-wave blocks 1.104).
template <typename M> struct TwoSweepBY { static constexpr int value = 12; };
template <> struct TwoSweepBY<float> { static constexpr int value = 8; };
template <typename M>
int TwoSweepVX(int nx) { return 2; }
template <>
int TwoSweepVX<float>(int nx) { return (nx % 4 == 0 && nx > 128) ? 4 : 2; }

template <typename M>
const void *TwoSweepFn(int vx) {
  constexpr int BY = TwoSweepBY<M>::value;
  return vx == 4 ? (const void *)TwoSweepKernel<M, 4, BY>
                 : (const void *)TwoSweepKernel<M, 2, BY>;
}

// Two sweeps in -> out in one launch.
template <typename M>
void LaunchTwoSweep(const __PSPoisson *s, const M *in, M *out, Fields<M> f,
                    float omega) {
  constexpr int BY = TwoSweepBY<M>::value;
  const dim3 grid(s->sweep2_plan.blocks), block(kWave * BY);
  if (s->sweep2_vx == 4)
    hipLaunchKernelGGL((TwoSweepKernel<M, 4, BY>), grid, block, 0, Stream(),
                       in, out, f, s->nx, s->ny, s->nz, omega, s->sweep2_plan);
  else
    hipLaunchKernelGGL((TwoSweepKernel<M, 2, BY>), grid, block, 0, Stream(),
                       in, out, f, s->nx, s->ny, s->nz, omega, s->sweep2_plan);
  CheckLaunch();
}

// n (even) sweeps a -> t -> a ... ending in a: pairs of sweeps run as
// two-sweep passes a -> t, t -> a; an odd pair count starts with one pair
// of single sweeps.
template <typename M>
void SweepPairs(__PSPoisson *s, M *a, M *t, Fields<M> f, int n,
                float omega) {
  int pairs = n / 2;
  if (pairs % 2) {
    LaunchSweep<M>(s, a, t, f, omega);
    LaunchSweep<M>(s, t, a, f, omega);
    --pairs;
  }
  for (int i = 0; i < pairs; i += 2) {
    LaunchTwoSweep<M>(s, a, t, f, omega);
    LaunchTwoSweep<M>(s, t, a, f, omega);
  }
}

template <typename M>
void InnerSweeps(__PSPoisson *s, int n, float omega) {
  SweepPairs<M>(s, (M *)s->e[0], (M *)s->e[1], InnerFields<M>(s), n, omega);
}

void PlainSweeps(__PSPoisson *s, float *p, int n, float omega) {
  SweepPairs<float>(s, p, s->t32, MasterFields(s), n, omega);
}

template <typename PT, typename ME, bool UPDATE, bool WRHS>
void LaunchResidual(const __PSPoisson *s, const PT *p, double *pout,
                    double inv_sigma, double sigma) {
  const dim3 grid(s->res_plan.blocks), block(kWave * kBY);
  hipLaunchKernelGGL((ResidualKernel<PT, ME, 2, UPDATE, WRHS>), grid, block,
                     0, Stream(), p, pout, (const ME *)s->e[0], (ME *)s->rhs,
                     MasterFields(s), s->nx, s->ny, s->nz, inv_sigma, sigma,
                     s->res_plan, s->partials);
  CheckLaunch();
}

template <typename ME>
void ResidualT(const __PSPoisson *s, const void *p, bool p64, double *pout,
               double inv_sigma, double sigma) {
  const bool w = sigma != 0.0;
  if (!p64) {
    if (w)
      LaunchResidual<float, ME, false, true>(s, (const float *)p, nullptr, 0, sigma);
    else
      LaunchResidual<float, ME, false, false>(s, (const float *)p, nullptr, 0, 0);
  } else if (pout) {
    if (w)
      LaunchResidual<double, ME, true, true>(s, (const double *)p, pout, inv_sigma, sigma);
    else
      LaunchResidual<double, ME, true, false>(s, (const double *)p, pout, inv_sigma, 0);
  } else {
    if (w)
      LaunchResidual<double, ME, false, true>(s, (const double *)p, nullptr, 0, sigma);
    else
      LaunchResidual<double, ME, false, false>(s, (const double *)p, nullptr, 0, 0);
  }
}

// Runs the residual pass and folds the partials per problem into sums
// (3 doubles per problem).
void Residual(const __PSPoisson *s, const void *p, bool p64, double *pout,
              double inv_sigma, double sigma, double *sums) {
  switch (s->storage) {
    case PS_STORAGE_BF16: ResidualT<bf16_t>(s, p, p64, pout, inv_sigma, sigma); break;
    case PS_STORAGE_FP16: ResidualT<fp16_t>(s, p, p64, pout, inv_sigma, sigma); break;
    default: ResidualT<float>(s, p, p64, pout, inv_sigma, sigma); break;
  }
  const Plan &pl = s->res_plan;
  PS_HIP_CHECK(hipMemcpyAsync(s->host_partials, s->partials,
                              sizeof(double) * 3 * pl.blocks,
                              hipMemcpyDeviceToHost, Stream()));
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  for (int q = 0; q < s->nbatch; ++q) sums[3 * q] = sums[3 * q + 1] = sums[3 * q + 2] = 0.0;
  const int xy = pl.tiles_x * pl.tiles_y;
  for (int b = 0; b < pl.blocks; ++b) {
    const int q = (b / xy) / pl.tz;
    double *o = sums + 3 * q;
    const double *h = s->host_partials + 3 * b;
    o[0] += h[0];
    o[1] = std::max(o[1], h[1]);
    o[2] += h[2];
  }
}

// max over problems of ||r|| / ||b|| and of max |r|.
void Summarise(const __PSPoisson *s, const std::vector<double> &sums,
               double *rel, double *rmax) {
  *rel = 0.0;
  *rmax = 0.0;
  for (int q = 0; q < s->nbatch; ++q) {
    const double r2 = sums[3 * q], b2 = sums[3 * q + 2];
    const double v = b2 > 0.0 ? std::sqrt(r2 / b2) : (r2 > 0.0 ? INFINITY : 0.0);
    *rel = std::max(*rel, v);
    *rmax = std::max(*rmax, sums[3 * q + 1]);
  }
}

// Power of two 2^-e with 2^(e-1) < rmax <= 2^e: scales r into (1/2, 1].
double ScaleFor(double rmax) {
  if (!(rmax > 0.0) || !std::isfinite(rmax)) return 1.0;
  int e;
  std::frexp(rmax, &e);  // rmax = f * 2^e, f in [0.5, 1)
  return std::ldexp(1.0, -e);
}

bool SameShape(const __PSPoisson *s, const __PSGrid *g) {
  return g && g->p && g->num_dims == 3 && g->dim[0] == s->nx &&
         g->dim[1] == s->ny && g->dim[2] == s->nz;
}

void *DevAlloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}

size_t StorageSize(int storage) { return storage == PS_STORAGE_FP32 ? 4 : 2; }

void *FieldPtr(__PSPoisson *s, int field) {
  if (field == PS_POISSON_E) return s->e[0];
  if (field == PS_POISSON_RHS) return s->rhs;
  const int k = field - PS_POISSON_COEF0;
  if (k >= 0 && k < kNumF - 1) return s->lo[k];
  return nullptr;
}

// Appends (sweeps, rel, seconds since t0) to the caller's trace.
void Trace(PSPoissonStats *st, double rel, double t0);

double Now() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

void Trace(PSPoissonStats *st, double rel, double t0) {
  if (st->progress_seconds > 0.0) {
    static thread_local double last = 0.0;
    const double now = Now();
    if (st->sweeps == 0) last = now;
    if (now - last >= st->progress_seconds) {
      std::fprintf(stderr, "[physis-hip] solve: %ld sweeps, residual %.3e, "
                   "%.1f s\n", st->sweeps, rel, now - t0);
      last = now;
    }
  }
  if (!st->history || st->history_len >= st->history_cap) return;
  double *h = st->history + 3 * st->history_len;
  h[0] = (double)st->sweeps;
  h[1] = rel;
  h[2] = Now() - t0;
  st->history_len += 1;
}

}  // namespace

extern "C" {

__PSPoisson *PSPoissonNew(__PSGrid *const coef[12], int nbatch, int storage) {
  if (storage < PS_STORAGE_FP32 || storage > PS_STORAGE_FP16 || nbatch < 1) {
    std::fprintf(stderr, "[physis-hip] PSPoissonNew: bad storage or batch\n");
    return nullptr;
  }
  for (int g = 0; g < kNumF; ++g) {
    const __PSGrid *c = coef[g];
    if (!c || !c->p || c->num_dims != 3 || c->type != PS_FLOAT ||
        c->dim[0] != coef[0]->dim[0] || c->dim[1] != coef[0]->dim[1] ||
        c->dim[2] != coef[0]->dim[2]) {
      std::fprintf(stderr,
                   "[physis-hip] PSPoissonNew: coefficient %d is not a 3-D "
                   "float grid of the common size\n", g);
      return nullptr;
    }
  }
  const int nx = coef[0]->dim[0], ny = coef[0]->dim[1], nz = coef[0]->dim[2];
  if (nx % 2 || nz % nbatch || nz / nbatch < 3 || nx < 3 || ny < 3 ||
      (long)nx * ny * nz >= (long)PSINDEX_MAX / 2) {
    std::fprintf(stderr,
                 "[physis-hip] PSPoissonNew: need even nx, nz divisible by "
                 "the batch, >= 3 points per axis and < 2^30 points\n");
    return nullptr;
  }
  __PSPoisson *s = new __PSPoisson();
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->nbatch = nbatch;
  s->nzs = nz / nbatch;
  s->storage = storage;
  s->elems = (size_t)nx * ny * nz;
  s->esize = StorageSize(storage);
  for (int g = 0; g < kNumF; ++g) s->coef[g] = coef[g];
  s->own_lo = storage != PS_STORAGE_FP32;
  bool ok = true;
  for (int g = 0; g < kNumF - 1; ++g) {
    s->lo[g] = s->own_lo ? DevAlloc(s->elems * s->esize) : coef[g]->p;
    ok = ok && s->lo[g];
  }
  s->rhs = DevAlloc(s->elems * s->esize);
  s->e[0] = DevAlloc(s->elems * s->esize);
  s->e[1] = DevAlloc(s->elems * s->esize);
  ok = ok && s->rhs && s->e[0] && s->e[1];
  s->sweep_vx = (nx % 4 == 0 && nx > 128) ? 4 : 2;
  int per_cu_sweep = 1, per_cu_res = 1;
  switch (storage) {
    case PS_STORAGE_BF16:
      per_cu_sweep = physis_hip::BlocksPerCU(
          s->sweep_vx == 4 ? (const void *)SweepKernel<bf16_t, 4, 0>
                           : (const void *)SweepKernel<bf16_t, 2, 0>, kWave * kBY);
      break;
    case PS_STORAGE_FP16:
      per_cu_sweep = physis_hip::BlocksPerCU(
          s->sweep_vx == 4 ? (const void *)SweepKernel<fp16_t, 4, 0>
                           : (const void *)SweepKernel<fp16_t, 2, 0>, kWave * kBY);
      break;
    default:
      per_cu_sweep = physis_hip::BlocksPerCU(
          s->sweep_vx == 4 ? (const void *)SweepKernel<float, 4, 0>
                           : (const void *)SweepKernel<float, 2, 0>, kWave * kBY);
  }
  per_cu_res = physis_hip::BlocksPerCU(
      (const void *)ResidualKernel<double, bf16_t, 2, true, true>, kWave * kBY);
  s->sweep_plan = MakePlan(nx, ny, s->nzs, nbatch, s->sweep_vx, per_cu_sweep);
  {
    s->sweep2_vx = storage == PS_STORAGE_FP32 ? TwoSweepVX<float>(nx)
                                              : TwoSweepVX<bf16_t>(nx);
    const void *fn = storage == PS_STORAGE_BF16   ? TwoSweepFn<bf16_t>(s->sweep2_vx)
                     : storage == PS_STORAGE_FP16 ? TwoSweepFn<fp16_t>(s->sweep2_vx)
                                                  : TwoSweepFn<float>(s->sweep2_vx);
    const int by = storage == PS_STORAGE_FP32 ? TwoSweepBY<float>::value
                                              : TwoSweepBY<bf16_t>::value;
    s->sweep2_plan = MakePlan(nx, ny, s->nzs, nbatch, s->sweep2_vx,
                              physis_hip::BlocksPerCU(fn, kWave * by), by - 4);
  }
  s->res_plan = MakePlan(nx, ny, s->nzs, nbatch, 2, per_cu_res);
  s->partials = (double *)DevAlloc(sizeof(double) * 3 * s->res_plan.blocks);
  ok = ok && s->partials;
  if (ok && hipHostMalloc((void **)&s->host_partials,
                          sizeof(double) * 3 * s->res_plan.blocks) != hipSuccess) {
    (void)hipGetLastError();
    s->host_partials = nullptr;
    ok = false;
  }
  if (!ok) {
    std::fprintf(stderr, "[physis-hip] PSPoissonNew: out of device memory\n");
    PSPoissonFree(s);
    return nullptr;
  }
  if (s->own_lo) {
    const int threads = 256;
    const int blocks = (int)std::min<size_t>((s->elems + threads - 1) / threads, 8192);
    for (int g = 0; g < kNumF - 1; ++g) {
      if (storage == PS_STORAGE_BF16)
        hipLaunchKernelGGL(ToStorageKernel<bf16_t>, dim3(blocks), dim3(threads), 0,
                           Stream(), (const float *)coef[g]->p, (bf16_t *)s->lo[g],
                           (long)s->elems);
      else
        hipLaunchKernelGGL(ToStorageKernel<fp16_t>, dim3(blocks), dim3(threads), 0,
                           Stream(), (const float *)coef[g]->p, (fp16_t *)s->lo[g],
                           (long)s->elems);
      CheckLaunch();
    }
  }
  PS_HIP_CHECK(hipMemsetAsync(s->rhs, 0, s->elems * s->esize, Stream()));
  PS_HIP_CHECK(hipMemsetAsync(s->e[0], 0, s->elems * s->esize, Stream()));
  PS_HIP_CHECK(hipMemsetAsync(s->e[1], 0, s->elems * s->esize, Stream()));
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  return s;
}

void PSPoissonFree(__PSPoisson *s) {
  if (!s) return;
  (void)hipStreamSynchronize(Stream());
  if (s->own_lo)
    for (int g = 0; g < kNumF - 1; ++g)
      if (s->lo[g]) PS_HIP_CHECK(hipFree(s->lo[g]));
  void *dev[] = {s->rhs, s->e[0], s->e[1], s->p2, s->t32, s->partials};
  for (void *p : dev)
    if (p) PS_HIP_CHECK(hipFree(p));
  if (s->host_partials) PS_HIP_CHECK(hipHostFree(s->host_partials));
  delete s;
}

void PSPoissonInnerSweeps(__PSPoisson *s, int n, float omega) {
  if (n < 0 || n % 2) PS_FATAL("PSPoissonInnerSweeps: n must be even, got %d", n);
  switch (s->storage) {
    case PS_STORAGE_BF16: InnerSweeps<bf16_t>(s, n, omega); break;
    case PS_STORAGE_FP16: InnerSweeps<fp16_t>(s, n, omega); break;
    default: InnerSweeps<float>(s, n, omega);
  }
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
}

void PSPoissonPlainSweeps(__PSPoisson *s, __PSGrid *p, int n, float omega) {
  if (n < 0 || n % 2) PS_FATAL("PSPoissonPlainSweeps: n must be even, got %d", n);
  if (!SameShape(s, p) || p->type != PS_FLOAT)
    PS_FATAL("PSPoissonPlainSweeps: p must be a float grid of the solver's size");
  if (!s->t32) {
    s->t32 = (float *)DevAlloc(s->elems * sizeof(float));
    if (!s->t32) PS_FATAL("PSPoissonPlainSweeps: out of device memory");
  }
  PlainSweeps(s, (float *)p->p, n, omega);
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
}

void PSPoissonResidual(__PSPoisson *s, const __PSGrid *p, __PSGrid *p_out,
                       double inv_sigma, double sigma, double *sums) {
  if (!SameShape(s, p) || (p->type != PS_FLOAT && p->type != PS_DOUBLE))
    PS_FATAL("PSPoissonResidual: p must be a float or double grid of the "
             "solver's size");
  const bool p64 = p->type == PS_DOUBLE;
  if (p_out && (!p64 || !SameShape(s, p_out) || p_out->type != PS_DOUBLE ||
                p_out->p == p->p))
    PS_FATAL("PSPoissonResidual: p_out must be a distinct double grid and p "
             "a double grid");
  Residual(s, p->p, p64, p_out ? (double *)p_out->p : nullptr, inv_sigma, sigma,
           sums);
}

int PSPoissonSetField(__PSPoisson *s, int field, const void *host) {
  void *d = FieldPtr(s, field);
  if (!d || (field >= PS_POISSON_COEF0 && !s->own_lo)) return -1;
  PS_HIP_CHECK(hipMemcpy(d, host, s->elems * s->esize, hipMemcpyHostToDevice));
  return 0;
}

int PSPoissonGetField(__PSPoisson *s, int field, void *host) {
  void *d = FieldPtr(s, field);
  if (!d) return -1;
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  PS_HIP_CHECK(hipMemcpy(host, d, s->elems * s->esize, hipMemcpyDeviceToHost));
  return 0;
}

int PSPoissonSolveMixed(__PSPoisson *s, __PSGrid *p, float omega, double tol,
                        int inner_sweeps, long max_sweeps, PSPoissonStats *st) {
  if (!SameShape(s, p) || p->type != PS_DOUBLE)
    PS_FATAL("PSPoissonSolveMixed: p must be a double grid of the solver's size");
  if (inner_sweeps < 2 || inner_sweeps % 2)
    PS_FATAL("PSPoissonSolveMixed: inner_sweeps must be even and >= 2");
  PSPoissonStats local{};
  if (!st) st = &local;
  {
    double *h = st->history;
    const int cap = st->history_cap;
    const double prog = st->progress_seconds;
    *st = PSPoissonStats{};
    st->history = h;
    st->history_cap = cap;
    st->progress_seconds = prog;
  }
  if (!s->p2) {
    s->p2 = (double *)DevAlloc(s->elems * sizeof(double));
    if (!s->p2) PS_FATAL("PSPoissonSolveMixed: out of device memory");
  }
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  const double t0 = Now();
  // The update pass writes every point of pb (E is 0 where bnd = 0, so
  // boundary values carry over), so pb needs no initialisation.
  double *pa = (double *)p->p, *pb = s->p2;
  std::vector<double> sums(3 * s->nbatch);
  double rel, rmax;
  Residual(s, pa, true, nullptr, 0.0, 0.0, sums.data());
  Summarise(s, sums, &rel, &rmax);
  Trace(st, rel, t0);
  double sigma = ScaleFor(rmax);  // scale of the RHS field
  bool rhs_ready = false;
  double best = rel;
  int no_gain = 0;
  long best_at = 0;
  while (rel > tol && st->sweeps < max_sweeps) {
    if (!rhs_ready) {
      Residual(s, pa, true, nullptr, 0.0, sigma, sums.data());
      rhs_ready = true;
    }
    PS_HIP_CHECK(hipMemsetAsync(s->e[0], 0, s->elems * s->esize, Stream()));
    switch (s->storage) {
      case PS_STORAGE_BF16: InnerSweeps<bf16_t>(s, inner_sweeps, omega); break;
      case PS_STORAGE_FP16: InnerSweeps<fp16_t>(s, inner_sweeps, omega); break;
      default: InnerSweeps<float>(s, inner_sweeps, omega);
    }
    st->sweeps += inner_sweeps;
    // p_b = p_a + E / sigma; r of p_b; RHS = sigma' r with sigma' from the
    // last known max |r| (one step stale, so |sigma' r| stays near 1).
    const double next_sigma = ScaleFor(rmax);
    Residual(s, pa, true, pb, 1.0 / sigma, next_sigma, sums.data());
    sigma = next_sigma;
    std::swap(pa, pb);
    st->outer += 1;
    Summarise(s, sums, &rel, &rmax);
    Trace(st, rel, t0);
    if (rel < best * (1.0 - 1e-3)) {
      best = rel;
      no_gain = 0;
      best_at = st->sweeps;
    } else if (++no_gain >= 20 && st->sweeps - best_at >= kStallSweeps) {
      st->stagnated = 1;
      break;
    }
  }
  if (pa != (double *)p->p)
    PS_HIP_CHECK(hipMemcpyAsync(p->p, pa, s->elems * sizeof(double),
                                hipMemcpyDeviceToDevice, Stream()));
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  st->seconds = Now() - t0;
  st->rel_residual = rel;
  st->converged = rel <= tol;
  return st->converged;
}

int PSPoissonSolvePlain(__PSPoisson *s, __PSGrid *p, float omega, double tol,
                        int check, long max_sweeps, PSPoissonStats *st) {
  if (!SameShape(s, p) || p->type != PS_FLOAT)
    PS_FATAL("PSPoissonSolvePlain: p must be a float grid of the solver's size");
  if (check < 2 || check % 2)
    PS_FATAL("PSPoissonSolvePlain: check must be even and >= 2");
  PSPoissonStats local{};
  if (!st) st = &local;
  {
    double *h = st->history;
    const int cap = st->history_cap;
    const double prog = st->progress_seconds;
    *st = PSPoissonStats{};
    st->history = h;
    st->history_cap = cap;
    st->progress_seconds = prog;
  }
  if (!s->t32) {
    s->t32 = (float *)DevAlloc(s->elems * sizeof(float));
    if (!s->t32) PS_FATAL("PSPoissonSolvePlain: out of device memory");
  }
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  const double t0 = Now();
  std::vector<double> sums(3 * s->nbatch);
  double rel, rmax;
  Residual(s, p->p, false, nullptr, 0.0, 0.0, sums.data());
  Summarise(s, sums, &rel, &rmax);
  Trace(st, rel, t0);
  double best = rel;
  int no_gain = 0;
  long best_at = 0;
  while (rel > tol && st->sweeps < max_sweeps) {
    PlainSweeps(s, (float *)p->p, check, omega);
    st->sweeps += check;
    Residual(s, p->p, false, nullptr, 0.0, 0.0, sums.data());
    st->outer += 1;
    Summarise(s, sums, &rel, &rmax);
    Trace(st, rel, t0);
    if (rel < best * (1.0 - 1e-3)) {
      best = rel;
      no_gain = 0;
      best_at = st->sweeps;
    } else if (++no_gain >= 20 && st->sweeps - best_at >= kStallSweeps) {
      st->stagnated = 1;
      break;
    }
  }
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  st->seconds = Now() - t0;
  st->rel_residual = rel;
  st->converged = rel <= tol;
  return st->converged;
}

}  // extern "C"
