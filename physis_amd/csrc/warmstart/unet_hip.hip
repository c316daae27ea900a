// libphysis_warmstart_hip: the U-Net warm start of the Poisson solve
// (include/physis/physis_warmstart.h), forward, backward and Adam on the GPU.
//
// Tensors are channels-last BF16 ([position][channel], position
// ((n*D + z)*H + y)*W + x, channels a multiple of 8), so the 8 K-values an
// MFMA lane feeds are one 16-byte load.
//
// 3x3x3 convolution as an implicit GEMM on v_mfma_f32_16x16x32_bf16:
//   forward   D[16 x-positions][16 cout] += A[16 pos][32 k] B[32 k][16 cout],
//             k = tap * CIN + ci, A read from an LDS-staged input tile with
//             zero padding, B the weights stored [cout][k] (K padded to 32);
//   backward  dX = the forward kernel on dZ with the flipped, transposed
//             weights Wf[ci][tap' * COUT + co] = W[co][26 - tap'][ci];
//   filter    dW[co][k] = sum over positions of dZ[p][co] X[p + off(k)][ci]:
//             GEMM with the positions as K, split over blocks of up to
//             16384 positions into FP32 partials summed in block order
//             (deterministic); both operands staged transposed in LDS for
//             rows of a multiple of 32 positions.
// MFMA lane layout (checked on the device against a host product): lane l
// feeds A row l%16 and B column l%16 with k = 8*(l/16) .. +7, and receives
// D rows 4*(l/16) .. +3 of column l%16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "physis/hip/device.h"
#include "physis/hip/stencil.h"
#include "physis/physis_warmstart.h"
#include "../runtime/runtime_hip.h"

#pragma clang fp contract(off)

namespace physis_unet {

typedef unsigned short u16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWgradChunk = 16384;  // max positions per filter-gradient block
constexpr int kPool = 4;           // pooling / upsampling factor per axis
// Weight fragments (K steps x 16-channel tiles) that the z-marching
// convolution keeps in registers; more are staged in LDS.  21 covers the
// 24 -> 8 layer (6% faster than staging its weights).
constexpr int kRegWeightSteps = 21;

__device__ __forceinline__ float Bf2F(u16 h) {
  return __uint_as_float((unsigned)h << 16);
}
__device__ __forceinline__ u16 F2Bf(float f) {
  return physis_hip::Convert<physis_hip::bf16_t, float>(f).bits;
}
__device__ __forceinline__ bf16x8 Load8(const u16 *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(p);
  return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 Zero8() {
  return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
}

constexpr int KPad(int cin) { return ((27 * cin + 31) / 32) * 32; }

// ------------------------------------------------------------ convolution
// Block (x tile, 4-row group, n*D + z) copies planes z-1..z+1, rows
// y0-1..y0+4, x0-1 .. x0+XT of the input (channels-last, zero outside the
// grid) into LDS with 16-byte loads; wave w then computes row y0+w over the
// tile in segments of 16 x: D[16 x][16 cout] accumulated over K steps of
// 32, every A operand (8 channels of one tap at one position) one aligned
// 16-byte LDS read.  The weight fragments of all K steps stay in registers
// when there are at most 14 of them (more cost occupancy: 21 for 24 -> 8
// held 200 VGPRs), else they are read per step (L1-resident).
template <int CIN, int COUT, bool RELU, int XT>
__global__ __launch_bounds__(256) void ConvKernel(
    const u16 *__restrict__ x, const u16 *__restrict__ wt,
    const float *__restrict__ bias, u16 *__restrict__ y, int D, int H,
    int W) {
  constexpr int K = 27 * CIN, KB = (K + 31) / 32, KP = KB * 32;
  constexpr int NT = (COUT + 15) / 16;
  constexpr bool kPre = KB * NT <= 14;
  constexpr int XS = XT + 2, NV = CIN / 8;
  __shared__ __attribute__((aligned(16))) u16 tile[18 * XS * CIN];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int tiles_x = (W + XT - 1) / XT;
  const int x0 = (blockIdx.x % tiles_x) * XT, y0 = (blockIdx.x / tiles_x) * 4;
  const int nz = blockIdx.y;
  const int n = nz / D, z = nz - n * D;
  for (int v = threadIdx.x; v < 18 * XS * NV; v += 256) {
    const int c8 = v % NV, t = v / NV;
    const int xp = t % XS, pr = t / XS;
    const int zz = z + pr / 6 - 1, yy = y0 + pr % 6 - 1, xx = x0 - 1 + xp;
    uint4 u = make_uint4(0, 0, 0, 0);
    if (zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W)
      u = *reinterpret_cast<const uint4 *>(
          x + ((((long)n * D + zz) * H + yy) * W + xx) * CIN + c8 * 8);
    *reinterpret_cast<uint4 *>(tile + t * CIN + c8 * 8) = u;
  }
  bf16x8 bw[kPre ? KB : 1][NT];
  if constexpr (kPre) {
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + r;
        bw[kb][t] = co < COUT ? Load8(wt + (long)co * KP + kb * 32 + 8 * g)
                              : Zero8();
      }
  }
  float bb[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = t * 16 + r;
    bb[t] = (bias && co < COUT) ? bias[co] : 0.0f;
  }
  __syncthreads();
  const int yy = y0 + wv;
  if (yy >= H) return;  // whole wave; no barrier follows
  const long row = (((long)n * D + z) * H + yy) * W;
  for (int xs0 = 0; xs0 < XT && x0 + xs0 < W; xs0 += 16) {
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int xa = xs0 + r;
#pragma unroll 4
    for (int kb = 0; kb < KB; ++kb) {
      const int k0 = kb * 32 + 8 * g;
      bf16x8 a = Zero8();
      if (k0 < K) {
        const int tap = k0 / CIN, ci0 = k0 - tap * CIN;
        // staged plane dz+1 = tap/9, row wv + dy+1, x xa + dx+1
        a = *reinterpret_cast<const bf16x8 *>(
            tile + (((tap / 9) * 6 + wv + (tap / 3) % 3) * XS + xa + tap % 3) *
                       CIN + ci0);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        bf16x8 b;
        if constexpr (kPre) {
          b = bw[kb][t];
        } else {
          const int co = t * 16 + r;
          b = co < COUT ? Load8(wt + (long)co * KP + kb * 32 + 8 * g)
                        : Zero8();
        }
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int co = t * 16 + r;
      if (co >= COUT) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int xo = x0 + xs0 + 4 * g + v;
        if (xo >= W) continue;
        float val = acc[t][v] + bb[t];
        if (RELU) val = val > 0.0f ? val : 0.0f;
        y[(row + xo) * COUT + co] = F2Bf(val);
      }
    }
  }
}

// 3x3x3 convolution marching z (the 2.5D scheme of the stencil run
// kernels).  Block (x tile, y tile, problem n, z chunk) owns XT x TY output
// positions of every plane of its chunk.  An LDS ring holds the input planes
// z-1 .. z+1 over rows y0-1 .. y0+TY and x0-1 .. x0+XT (zero outside the
// grid) and the plane z+2 being staged; the plane after that is loaded into
// registers a step ahead.  Each input plane is thus staged once per tile
// (the per-plane kernel above stages three), with one barrier per step.
// The GEMM is D[cout][16 x] = W[cout][32 k] X[32 k][16 x]: the weights are
// the A operand (registers, or LDS when they are many), the staged input the
// B operand (one aligned 16-byte LDS read per lane), and a lane's result is
// four consecutive channels of one position, stored as one 8-byte word.
template <int CIN, int COUT, bool RELU, int XT, int TY>
__global__ __launch_bounds__(64 * TY) void ConvZKernel(
    const u16 *__restrict__ x, const u16 *__restrict__ wt,
    const float *__restrict__ bias, u16 *__restrict__ y, int D, int H, int W,
    int tz, int zchunk) {
  constexpr int K = 27 * CIN, KB = (K + 31) / 32, KP = KB * 32;
  constexpr int MT = (COUT + 15) / 16;
  constexpr bool kRegW = KB * MT <= kRegWeightSteps;
  constexpr int XS = XT + 2, RS = TY + 2, NV = CIN / 8;
  constexpr int PLANE = RS * XS * CIN;  // u16 elements of a staged plane
  constexpr int SV = RS * XS * NV;      // 16-byte vectors of a staged plane
  constexpr int NTH = 64 * TY;
  constexpr int VPT = (SV + NTH - 1) / NTH;
  constexpr int WL = kRegW ? 1 : MT * 16 * KP;  // weights staged in LDS
  __shared__ __attribute__((aligned(16))) u16 ring[4 * PLANE];
  __shared__ __attribute__((aligned(16))) u16 wlds[WL];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int tiles_x = (W + XT - 1) / XT;
  const int x0 = (blockIdx.x % tiles_x) * XT, y0 = (blockIdx.x / tiles_x) * TY;
  const int n = blockIdx.y / tz;
  const int zb = (blockIdx.y - n * tz) * zchunk;
  const int ze = min(zb + zchunk, D);
  if (zb >= ze) return;  // uniform over the block
  auto fetch = [&](int zz, uint4 (&buf)[VPT]) {
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int v = threadIdx.x + i * NTH;
      const int c8 = v % NV, t = v / NV;
      const int xp = t % XS, pr = t / XS;
      const int yy = y0 - 1 + pr, xx = x0 - 1 + xp;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (v < SV && zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 &&
          xx < W)
        u = *reinterpret_cast<const uint4 *>(
            x + ((((long)n * D + zz) * H + yy) * W + xx) * CIN + c8 * 8);
      buf[i] = u;
    }
  };
  auto put = [&](int slot, const uint4 (&buf)[VPT]) {
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (v < SV)
        *reinterpret_cast<uint4 *>(ring + slot * PLANE + (v / NV) * CIN +
                                   (v % NV) * 8) = buf[i];
    }
  };
  uint4 buf[VPT];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    fetch(zb - 1 + k, buf);
    put(k, buf);
  }
  fetch(zb + 2, buf);
  bf16x8 aw[kRegW ? KB : 1][MT];
  if constexpr (kRegW) {
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const int co = t * 16 + r;
        aw[kb][t] = co < COUT ? Load8(wt + (long)co * KP + kb * 32 + 8 * g)
                              : Zero8();
      }
  } else {
    for (int v = threadIdx.x; v < MT * 16 * KP / 8; v += NTH) {
      const int co = (v * 8) / KP;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (co < COUT) u = *reinterpret_cast<const uint4 *>(wt + (long)v * 8);
      *reinterpret_cast<uint4 *>(wlds + v * 8) = u;
    }
  }
  // With the weights in registers and at most 14 K steps the K loop is
  // unrolled and the B operand's LDS offset of every K step (u16 elements
  // within a staged plane, relative to the lane's row and x) is decoded
  // once, with the tap's plane (0..2) in bits 24 and up; -1 past K.
  // (Measured at level 1: backward data 8 -> 24 392 -> 279 us, forward
  // 8 -> 8 198 -> 187 us; the 21-step 24 -> 8 forward went 617 -> 729 us
  // and keeps the per-step decode.)
  constexpr bool kDecode = kRegW && KB <= 14;
  int boff[kDecode ? KB : 1];
  if constexpr (kDecode) {
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int k0 = kb * 32 + 8 * g;
      if (k0 < K) {
        const int tap = k0 / CIN, ci0 = k0 - tap * CIN;
        boff[kb] = ((((tap / 3) % 3) * XS + tap % 3) * CIN + ci0) |
                   ((tap / 9) << 24);
      } else {
        boff[kb] = -1;
      }
    }
  }
  // bias of the four channels co = 16 t + 4 g + v this lane receives
  float bb[MT][4];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int co = t * 16 + 4 * g + v;
      bb[t][v] = (bias && co < COUT) ? bias[co] : 0.0f;
    }
  __syncthreads();
  const int yy = y0 + wv;
  for (int z = zb; z < ze; ++z) {
    const int base = z - zb;  // ring slot of plane z - 1 (mod 4)
    if (yy < H) {
      const long row = (((long)n * D + z) * H + yy) * W;
      for (int xs0 = 0; xs0 < XT && x0 + xs0 < W; xs0 += 16) {
        f32x4 acc[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int xa = xs0 + r;
        if constexpr (kDecode) {
          const u16 *rb = ring + (wv * XS + xa) * CIN;
          const int so[3] = {(base & 3) * PLANE, ((base + 1) & 3) * PLANE,
                             ((base + 2) & 3) * PLANE};
#pragma unroll
          for (int kb = 0; kb < KB; ++kb) {
            bf16x8 b = Zero8();
            if (boff[kb] >= 0) {
              const int pz = boff[kb] >> 24;
              const int sp = pz == 0 ? so[0] : (pz == 1 ? so[1] : so[2]);
              b = *reinterpret_cast<const bf16x8 *>(rb + sp +
                                                    (boff[kb] & 0xffffff));
            }
#pragma unroll
            for (int t = 0; t < MT; ++t)
              acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                  aw[kb][t], b, acc[t], 0, 0, 0);
          }
        } else {
#pragma unroll 4
          for (int kb = 0; kb < KB; ++kb) {
            const int k0 = kb * 32 + 8 * g;
            bf16x8 b = Zero8();
            if (k0 < K) {
              const int tap = k0 / CIN, ci0 = k0 - tap * CIN;
              const int slot = (base + tap / 9) & 3;
              b = *reinterpret_cast<const bf16x8 *>(
                  ring + slot * PLANE +
                  ((wv + (tap / 3) % 3) * XS + xa + tap % 3) * CIN + ci0);
            }
#pragma unroll
            for (int t = 0; t < MT; ++t) {
              bf16x8 a;
              if constexpr (kRegW)
                a = aw[kb][t];
              else
                a = *reinterpret_cast<const bf16x8 *>(
                    wlds + (t * 16 + r) * KP + kb * 32 + 8 * g);
              acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t],
                                                               0, 0, 0);
            }
          }
        }
        const int xo = x0 + xs0 + r;
        if (xo < W) {
#pragma unroll
          for (int t = 0; t < MT; ++t) {
            const int co = t * 16 + 4 * g;
            if (co >= COUT) continue;
            unsigned w2[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              float v0 = acc[t][2 * h] + bb[t][2 * h];
              float v1 = acc[t][2 * h + 1] + bb[t][2 * h + 1];
              if (RELU) {
                v0 = v0 > 0.0f ? v0 : 0.0f;
                v1 = v1 > 0.0f ? v1 : 0.0f;
              }
              w2[h] = (unsigned)F2Bf(v0) | ((unsigned)F2Bf(v1) << 16);
            }
            *reinterpret_cast<uint2 *>(y + (row + xo) * COUT + co) =
                make_uint2(w2[0], w2[1]);
          }
        }
      }
    }
    put((base + 3) & 3, buf);
    __syncthreads();
    fetch(z + 3, buf);
  }
}

// The z-marching convolution for 8 output channels, two output rows per
// wave.  With COUT = 8 half of the 16 A rows of the tap-major kernel are
// zero; here row q = 8 h + co of the A operand is output channel co of
// output row y + h (h = 0, 1), and K runs over the input rows y - 1 .. y + 2
// of both: k = ((dz 4 + ry) 3 + dx) CIN + ci, A[q][k] = W[co][9 dz + 3 (ry -
// h) + dx][ci] when 0 <= ry - h <= 2, else 0.  One B operand (8 channels of
// one staged input position) then serves both rows: 36 CIN / 32 K steps per
// 16 x positions of two rows instead of 2 x 27 CIN / 32 (CIN = 24: 27
// instead of 42 MFMAs and B reads).  Block (x tile, 2 TY rows, problem, z
// chunk); the ring and the staging are those of ConvZKernel.
template <int CIN, bool RELU, int XT, int TY>
__global__ __launch_bounds__(64 * TY) void ConvZPairKernel(
    const u16 *__restrict__ x, const u16 *__restrict__ wt,
    const float *__restrict__ bias, u16 *__restrict__ y, int D, int H, int W,
    int tz, int zchunk) {
  constexpr int COUT = 8;
  constexpr int KW = 27 * CIN, KPW = ((KW + 31) / 32) * 32;  // wt layout
  constexpr int K = 36 * CIN, KB = (K + 31) / 32;
  constexpr int ROWS = 2 * TY;
  constexpr int XS = XT + 2, RS = ROWS + 2, NV = CIN / 8;
  constexpr int PLANE = RS * XS * CIN;  // u16 elements of a staged plane
  constexpr int SV = RS * XS * NV;      // 16-byte vectors of a staged plane
  constexpr int NTH = 64 * TY;
  constexpr int VPT = (SV + NTH - 1) / NTH;
  __shared__ __attribute__((aligned(16))) u16 ring[4 * PLANE];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int tiles_x = (W + XT - 1) / XT;
  const int x0 = (blockIdx.x % tiles_x) * XT;
  const int y0 = (blockIdx.x / tiles_x) * ROWS;
  const int n = blockIdx.y / tz;
  const int zb = (blockIdx.y - n * tz) * zchunk;
  const int ze = min(zb + zchunk, D);
  if (zb >= ze) return;  // uniform over the block
  auto fetch = [&](int zz, uint4 (&buf)[VPT]) {
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int v = threadIdx.x + i * NTH;
      const int c8 = v % NV, t = v / NV;
      const int xp = t % XS, pr = t / XS;
      const int yy = y0 - 1 + pr, xx = x0 - 1 + xp;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (v < SV && zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 &&
          xx < W)
        u = *reinterpret_cast<const uint4 *>(
            x + ((((long)n * D + zz) * H + yy) * W + xx) * CIN + c8 * 8);
      buf[i] = u;
    }
  };
  auto put = [&](int slot, const uint4 (&buf)[VPT]) {
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (v < SV)
        *reinterpret_cast<uint4 *>(ring + slot * PLANE + v * 8) = buf[i];
    }
  };
  uint4 buf[VPT];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    fetch(zb - 1 + k, buf);
    put(k, buf);
  }
  fetch(zb + 2, buf);
  // A fragments (registers) and the B operand's offsets of every K step:
  // u16 elements within a staged plane relative to the wave's first row and
  // the lane's x, with the input plane (0..2) in bits 24 and up; -1 past K
  const int co = r & 7, h = r >> 3;
  bf16x8 aw[KB];
  int boff[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    const int k0 = kb * 32 + 8 * g;
    aw[kb] = Zero8();
    boff[kb] = -1;
    if (k0 < K) {
      const int cell = k0 / CIN, ci0 = k0 - cell * CIN;
      const int dz = cell / 12, ry = (cell / 3) % 4, dx = cell % 3;
      const int dy = ry - h;
      if (dy >= 0 && dy <= 2)
        aw[kb] = Load8(wt + (long)co * KPW + (9 * dz + 3 * dy + dx) * CIN + ci0);
      boff[kb] = ((ry * XS + dx) * CIN + ci0) | (dz << 24);
    }
  }
  // bias of the four outputs (row y + (g >> 1), channels 4 (g & 1) + v)
  float bb[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) bb[v] = bias ? bias[4 * (g & 1) + v] : 0.0f;
  __syncthreads();
  const int yy0 = y0 + 2 * wv;  // the wave's first output row
  const int yo = yy0 + (g >> 1);  // the row of this lane's outputs
  for (int z = zb; z < ze; ++z) {
    const int base = z - zb;  // ring slot of plane z - 1 (mod 4)
    if (yy0 < H) {
      const long row = (((long)n * D + z) * H + yo) * W;
      const int so[3] = {(base & 3) * PLANE, ((base + 1) & 3) * PLANE,
                         ((base + 2) & 3) * PLANE};
      for (int xs0 = 0; xs0 < XT && x0 + xs0 < W; xs0 += 16) {
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        const u16 *rb = ring + (2 * wv * XS + xs0 + r) * CIN;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          bf16x8 b = Zero8();
          if (boff[kb] >= 0) {
            const int pz = boff[kb] >> 24;
            const int sp = pz == 0 ? so[0] : (pz == 1 ? so[1] : so[2]);
            b = *reinterpret_cast<const bf16x8 *>(rb + sp +
                                                  (boff[kb] & 0xffffff));
          }
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[kb], b, acc, 0, 0,
                                                        0);
        }
        const int xo = x0 + xs0 + r;
        if (xo < W && yo < H) {
          unsigned w2[2];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            float v0 = acc[2 * q] + bb[2 * q];
            float v1 = acc[2 * q + 1] + bb[2 * q + 1];
            if (RELU) {
              v0 = v0 > 0.0f ? v0 : 0.0f;
              v1 = v1 > 0.0f ? v1 : 0.0f;
            }
            w2[q] = (unsigned)F2Bf(v0) | ((unsigned)F2Bf(v1) << 16);
          }
          *reinterpret_cast<uint2 *>(y + (row + xo) * COUT + 4 * (g & 1)) =
              make_uint2(w2[0], w2[1]);
        }
      }
    }
    put((base + 3) & 3, buf);
    __syncthreads();
    fetch(z + 3, buf);
  }
}

// Filter gradient partials marching the input planes: block (x tile, y
// tile, problem, z chunk [zb, ze) of output-gradient planes) sums over its
// XT x TY x zchunk positions
//   part[block][co][k] = sum_p dz[p][co] x[p + off(k)][ci],  k = tap CIN + ci
// as a GEMM over positions (K) on v_mfma_f32_16x16x32_bf16.  Step z' stages
// input plane z' (transposed, xT[row][ci][x], 3 slots) and uses it for the
// three taps whose z offset t - 1 maps it to output-gradient plane
// z' + 1 - t (t = 0, 1, 2; planes outside [zb, ze) stage as zeros, dzT
// [co][y][x] in a ring of 4).  B = the shifted input (8 consecutive x of
// one channel at one (dy, dx) offset: column (3 dy + dx) CIN + ci, read as
// 32-bit words with a funnel shift for odd offsets) is read once per step
// and feeds every t: the A operand's rows pack (t, co) pairs, row
// q = t COUT + co, so COUT = 8 needs 2 MFMAs per B operand (the tap-major
// order needed 3, each half empty, and read B three times).  Waves split
// into WT groups over column tiles and TY / WT groups over the 32-position
// steps of a plane; the step groups' sums are added in a fixed order at
// the end (deterministic).  One barrier per step.
template <int CIN, int COUT, int XT, int TY, int WT>
__global__ __launch_bounds__(64 * TY) void WgradZKernel(
    const u16 *__restrict__ x, const u16 *__restrict__ dz,
    float *__restrict__ part, float *__restrict__ bpart, int D, int H, int W,
    int tz, int zchunk) {
  static_assert(XT % 32 == 0, "steps of 32 positions along x");
  static_assert(TY % WT == 0, "wave groups");
  constexpr int K = 27 * CIN, NC = 9 * CIN;
  constexpr int NTN = (NC + 15) / 16;       // column tiles
  constexpr int MM = (3 * COUT + 15) / 16;  // MFMAs per B operand
  constexpr int WS = TY / WT;               // step groups
  constexpr int TPW = (NTN + WT - 1) / WT;  // column tiles per wave
  constexpr int NSTEP = TY * XT / 32;       // 32-position steps per plane
  static_assert(NSTEP % WS == 0, "steps split evenly");
  constexpr int XR = XT + 2, RS = TY + 2;
  constexpr int RW = CIN * XR / 2;          // 32-bit words per staged row
  constexpr int PW = RS * RW;               // words per staged plane
  constexpr int NTH = 64 * TY;
  constexpr int XV = RS * XR * (CIN / 8);   // 16-byte input vectors per plane
  constexpr int XVT = (XV + NTH - 1) / NTH;
  constexpr int DROW = TY * XT + 8;         // dzT row (one channel), padded
  constexpr int DPL = COUT * DROW;          // u16 per staged dz plane
  constexpr int DV = TY * XT * (COUT / 8);  // 16-byte dz vectors per plane
  constexpr int DVT = (DV + NTH - 1) / NTH;
  constexpr int LOOP_WORDS = 3 * PW + 4 + 4 * DPL / 2;
  constexpr int RED_WORDS = (WS - 1) * WT * (TPW * MM * 4 + MM) * 64;
  constexpr int LDS_WORDS = LOOP_WORDS > RED_WORDS ? LOOP_WORDS : RED_WORDS;
  static_assert(XR % 2 == 0 && DROW % 8 == 0, "alignment");
  __shared__ __attribute__((aligned(16))) unsigned lds[LDS_WORDS];
  unsigned *const xT = lds;
  u16 *const dzT = reinterpret_cast<u16 *>(lds + 3 * PW + 4);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wt = wv % WT, ws = wv / WT;
  const int g = lane >> 4, r = lane & 15;
  const int tiles_x = (W + XT - 1) / XT;
  const int x0 = (blockIdx.x % tiles_x) * XT, y0 = (blockIdx.x / tiles_x) * TY;
  const int n = blockIdx.y / tz;
  const int zb = (blockIdx.y - n * tz) * zchunk;
  const int ze = min(zb + zchunk, D);
  const long bid = (long)blockIdx.y * gridDim.x + blockIdx.x;
  // element offset (within a staged plane) of each tile column's (dy, ci,
  // dx) origin; -1 past the last column
  int lofs[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    const int col = (wt + WT * i) * 16 + r;
    if (col < NC) {
      const int o = col / CIN, ci = col - o * CIN;
      lofs[i] = ((o / 3) * CIN + ci) * XR + o % 3;
    } else {
      lofs[i] = -1;
    }
  }
  // A rows of this lane: q = 16 m + r -> (t, co); t = 3 means none
  int at[MM], aco[MM];
#pragma unroll
  for (int m = 0; m < MM; ++m) {
    const int q = 16 * m + r;
    at[m] = q < 3 * COUT ? q / COUT : 3;
    aco[m] = q < 3 * COUT ? q % COUT : 0;
  }
  f32x4 acc[TPW][MM];
#pragma unroll
  for (int i = 0; i < TPW; ++i)
#pragma unroll
    for (int m = 0; m < MM; ++m) acc[i][m] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum[MM];  // dz sums of the t = 1 rows (bias gradient)
#pragma unroll
  for (int m = 0; m < MM; ++m) bsum[m] = 0.0f;
  // input plane zz, rows y0-1 .. y0+TY, x0-1 .. x0+XT
  auto fetchx = [&](int zz, uint4 (&buf)[XVT]) {
#pragma unroll
    for (int i = 0; i < XVT; ++i) {
      const int v = threadIdx.x + i * NTH;
      const int c8 = v % (CIN / 8), t = v / (CIN / 8);
      const int xp = t % XR, row = t / XR;
      const int yy = y0 - 1 + row, xx = x0 - 1 + xp;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (v < XV && zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 &&
          xx < W)
        u = *reinterpret_cast<const uint4 *>(
            x + ((((long)n * D + zz) * H + yy) * W + xx) * CIN + c8 * 8);
      buf[i] = u;
    }
  };
  // transposed into slot `slot` of xT: [row][ci][x] (16-bit elements)
  auto putx = [&](int slot, const uint4 (&buf)[XVT]) {
    u16 *base = reinterpret_cast<u16 *>(xT + slot * PW);
#pragma unroll
    for (int i = 0; i < XVT; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (v >= XV) continue;
      const int c8 = v % (CIN / 8), t = v / (CIN / 8);
      const int xp = t % XR, row = t / XR;
      const unsigned w4[4] = {buf[i].x, buf[i].y, buf[i].z, buf[i].w};
      u16 *dst = base + (row * CIN + c8 * 8) * XR + xp;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        dst[(2 * j) * XR] = (u16)(w4[j] & 0xffffu);
        dst[(2 * j + 1) * XR] = (u16)(w4[j] >> 16);
      }
    }
  };
  // output-gradient plane zz (zeros outside [zb, ze))
  auto fetchd = [&](int zz, uint4 (&buf)[DVT]) {
#pragma unroll
    for (int i = 0; i < DVT; ++i) {
      const int v = threadIdx.x + i * NTH;
      const int c8 = v % (COUT / 8), t = v / (COUT / 8);
      const int xp = t % XT, row = t / XT;
      const int yy = y0 + row, xx = x0 + xp;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (v < DV && zz >= zb && zz < ze && yy < H && xx < W)
        u = *reinterpret_cast<const uint4 *>(
            dz + ((((long)n * D + zz) * H + yy) * W + xx) * COUT + c8 * 8);
      buf[i] = u;
    }
  };
  auto putd = [&](int slot, const uint4 (&buf)[DVT]) {
    u16 *base = dzT + slot * DPL;
#pragma unroll
    for (int i = 0; i < DVT; ++i) {
      const int v = threadIdx.x + i * NTH;
      if (v >= DV) continue;
      const int c8 = v % (COUT / 8), t = v / (COUT / 8);
      const unsigned w4[4] = {buf[i].x, buf[i].y, buf[i].z, buf[i].w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        base[(c8 * 8 + 2 * j) * DROW + t] = (u16)(w4[j] & 0xffffu);
        base[(c8 * 8 + 2 * j + 1) * DROW + t] = (u16)(w4[j] >> 16);
      }
    }
  };
  if (zb < ze) {
    // x plane zz in slot (zz - zb + 1) % 3, dz plane zz in (zz - zb + 2) & 3
    uint4 bx[XVT], bd[DVT];
    fetchx(zb - 1, bx);
    putx(0, bx);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      fetchd(zb - 2 + k, bd);
      putd(k, bd);
    }
    fetchx(zb, bx);
    fetchd(zb + 1, bd);
    __syncthreads();
    for (int zp = zb - 1; zp <= ze; ++zp) {
      const int xs = (zp - zb + 1) % 3;
      const int dbase = zp - zb + 3;  // dz slot of t is (dbase - t) & 3
      const unsigned *xplane = xT + xs * PW;
#pragma unroll
      for (int j = 0; j < NSTEP / WS; ++j) {
        const int st = (ws * (NSTEP / WS) + j) * 32;  // first position
        const int ry = st / XT, rx = st - ry * XT;
        bf16x8 a[MM];
#pragma unroll
        for (int m = 0; m < MM; ++m) {
          a[m] = at[m] < 3
                     ? *reinterpret_cast<const bf16x8 *>(
                           dzT + ((dbase - at[m]) & 3) * DPL + aco[m] * DROW +
                           st + 8 * g)
                     : Zero8();
          if (wt == 0 && at[m] == 1) {
#pragma unroll
            for (int e = 0; e < 8; ++e) bsum[m] += (float)a[m][e];
          }
        }
#pragma unroll
        for (int i = 0; i < TPW; ++i) {
          if (wt + WT * i >= NTN) break;  // wave-uniform
          bf16x8 b = Zero8();
          if (lofs[i] >= 0) {
            const int e = lofs[i] + ry * CIN * XR + rx + 8 * g;
            const unsigned *wsrc = xplane + (e >> 1);
            unsigned w[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) w[k] = wsrc[k];
            uint4 o;
            if (e & 1)
              o = make_uint4(__builtin_amdgcn_alignbit(w[1], w[0], 16),
                             __builtin_amdgcn_alignbit(w[2], w[1], 16),
                             __builtin_amdgcn_alignbit(w[3], w[2], 16),
                             __builtin_amdgcn_alignbit(w[4], w[3], 16));
            else
              o = make_uint4(w[0], w[1], w[2], w[3]);
            b = __builtin_bit_cast(bf16x8, o);
          }
#pragma unroll
          for (int m = 0; m < MM; ++m)
            acc[i][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b,
                                                                acc[i][m], 0,
                                                                0, 0);
        }
      }
      // x plane zp + 1 and dz plane zp + 2 into the slots no wave reads in
      // this step (x: plane zp - 2, dz: plane zp - 2)
      putx((zp - zb + 2) % 3, bx);
      putd((zp - zb + 4) & 3, bd);
      __syncthreads();
      fetchx(zp + 2, bx);
      fetchd(zp + 3, bd);
    }
  }
  // step groups 1 .. WS-1 hand their sums to group 0 through LDS
  if constexpr (WS > 1) {
    __syncthreads();  // the staged planes are no longer read
    float *red = reinterpret_cast<float *>(lds);
    constexpr int PER = TPW * MM * 4 + MM;  // floats per lane and wave
    if (ws > 0) {
      float *o = red + ((ws - 1) * WT + wt) * PER * 64;
#pragma unroll
      for (int i = 0; i < TPW; ++i)
#pragma unroll
        for (int m = 0; m < MM; ++m)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            o[((i * MM + m) * 4 + v) * 64 + lane] = acc[i][m][v];
#pragma unroll
      for (int m = 0; m < MM; ++m) o[(TPW * MM * 4 + m) * 64 + lane] = bsum[m];
    }
    __syncthreads();
    if (ws == 0) {
      for (int u = 1; u < WS; ++u) {
        const float *o = red + ((u - 1) * WT + wt) * PER * 64;
#pragma unroll
        for (int i = 0; i < TPW; ++i)
#pragma unroll
          for (int m = 0; m < MM; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              acc[i][m][v] += o[((i * MM + m) * 4 + v) * 64 + lane];
#pragma unroll
        for (int m = 0; m < MM; ++m)
          bsum[m] += o[(TPW * MM * 4 + m) * 64 + lane];
      }
    }
  }
  if (ws == 0) {
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int col = (wt + WT * i) * 16 + r;
      if (wt + WT * i >= NTN || col >= NC) continue;
#pragma unroll
      for (int m = 0; m < MM; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int q = 16 * m + 4 * g + v;
          if (q >= 3 * COUT) continue;
          const int t = q / COUT, co = q - t * COUT;
          part[(bid * COUT + co) * K + 9 * t * CIN + col] = acc[i][m][v];
        }
    }
    if (wt == 0) {
      // lanes r of the four groups g hold disjoint positions of row q = 16 m
      // + r: fold the groups in order
#pragma unroll
      for (int m = 0; m < MM; ++m) {
        float sm = bsum[m];
        sm += __shfl_xor(sm, 16, 64);
        sm += __shfl_xor(sm, 32, 64);
        if (g == 0 && at[m] == 1) bpart[bid * COUT + aco[m]] = sm;
      }
    }
  }
}

// Filter gradient partials: block (b, t) covers positions [b*chunk,
// (b+1)*chunk) and column tiles 4t .. 4t+3 of k (one per wave) and writes
// part[b][co][k]; blocks (b, 0) also write bpart[b][co].  A lane decodes its
// first position of a 32-position step once and walks the next 7
// incrementally.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void WgradKernel(
    const u16 *__restrict__ x, const u16 *__restrict__ dz,
    float *__restrict__ part, float *__restrict__ bpart, int P, int D, int H,
    int W, int chunk) {
  constexpr int K = 27 * CIN, NTN = (K + 15) / 16, MT = (COUT + 15) / 16;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int p0 = blockIdx.x * chunk;
  const int p1 = p0 + chunk < P ? p0 + chunk : P;
  const int nt = blockIdx.y * 4 + wv;
  if (nt < NTN) {
    f32x4 acc[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int col = nt * 16 + r;
    const bool colv = col < K;
    const int tap = colv ? col / CIN : 0, ci = colv ? col - tap * CIN : 0;
    const int dzo = tap / 9 - 1, dyo = (tap / 3) % 3 - 1, dxo = tap % 3 - 1;
    for (int kp = p0; kp < p1; kp += 32) {
      bf16x8 b = Zero8();
      bf16x8 a[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m) a[m] = Zero8();
      const int pb = kp + 8 * g;
      int q = pb;
      int xx = q % W;
      q /= W;
      int yy = q % H;
      q /= H;
      int zz = q % D;
      int nn = q / D;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int pos = pb + j;
        if (pos < p1) {
          const int xs = xx + dxo, ys = yy + dyo, zs = zz + dzo;
          if (colv && xs >= 0 && xs < W && ys >= 0 && ys < H && zs >= 0 &&
              zs < D)
            b[j] = __builtin_bit_cast(
                __bf16, x[((long)((nn * D + zs) * H + ys) * W + xs) * CIN + ci]);
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            const int co = m * 16 + r;
            if (co < COUT)
              a[m][j] = __builtin_bit_cast(__bf16, dz[(long)pos * COUT + co]);
          }
        }
        if (++xx == W) {
          xx = 0;
          if (++yy == H) {
            yy = 0;
            if (++zz == D) {
              zz = 0;
              ++nn;
            }
          }
        }
      }
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[m], 0, 0, 0);
    }
    if (colv) {
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int co = m * 16 + 4 * g + v;
          if (co < COUT)
            part[((long)blockIdx.x * COUT + co) * K + col] = acc[m][v];
        }
    }
  }
  // bias (blocks of the first tile group): thread t sums channel t % COUT
  // over every (256 / COUT)-th position, then the partial sums of each
  // channel are added in order
  if (blockIdx.y != 0) return;
  constexpr int kLanes = 256 / COUT;
  __shared__ float red[256];
  float s = 0.0f;
  const int c = threadIdx.x % COUT, lid = threadIdx.x / COUT;
  if (lid < kLanes)
    for (int p = p0 + lid; p < p1; p += kLanes)
      s += Bf2F(dz[(long)p * COUT + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < COUT) {
    float t = 0.0f;
    for (int i = 0; i < kLanes; ++i) t += red[threadIdx.x + COUT * i];
    bpart[(long)blockIdx.x * COUT + threadIdx.x] = t;
  }
}

// Filter gradient partials through LDS, for rows whose width is a multiple
// of 32.  Block b covers positions [b*chunk, (b+1)*chunk) (a multiple of 32)
// and every column of k; each step takes 32 positions of one row:
//   - dz of the 32 positions is staged transposed, dzT[co][p], so the A
//     operand of a lane (channel co, positions 8g .. 8g+7) is one 16-byte
//     LDS read;
//   - x of the 3 x 3 neighbouring rows (z-1..z+1, y-1..y+1) at x0-1 ..
//     x0+32 is staged transposed, xT[row][ci][x], zero outside the grid, so
//     the B operand of column (tap, ci) is 8 consecutive LDS elements;
//   - wave w accumulates column tiles w, w+4, w+8, ... against the shared
//     A operands.
// Both operand tiles come from 16-byte global loads of whole channel
// vectors, in place of the 2-byte gathers of WgradKernel.  Partials and
// bias partials have WgradKernel's layout; the bias is summed in position
// order by threads 0 .. COUT-1.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void WgradLdsKernel(
    const u16 *__restrict__ x, const u16 *__restrict__ dz,
    float *__restrict__ part, float *__restrict__ bpart, int P, int D, int H,
    int W, int chunk) {
  constexpr int K = 27 * CIN, NTN = (K + 15) / 16, MT = (COUT + 15) / 16;
  constexpr int TPW = (NTN + 3) / 4;  // column tiles per wave
  constexpr int XR = 34;              // staged x: x0-1 .. x0+32
  // xT[row][ci][x] as 16-bit elements, written and read as 32-bit words
  // (pairs of neighbouring x); 4 words of padding for the last row's
  // 5-word reads
  __shared__ unsigned xT[9 * CIN * XR / 2 + 4];
  __shared__ __attribute__((aligned(16))) u16 dzT[COUT * 32];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int p0 = blockIdx.x * chunk;
  const int p1 = p0 + chunk < P ? p0 + chunk : P;
  // LDS offset of each tile's column for this lane (-1: padding column)
  int lofs[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    const int col = (wv + 4 * i) * 16 + r;
    if (col < K) {
      const int tap = col / CIN, ci = col - tap * CIN;
      lofs[i] = ((tap / 3) * CIN + ci) * XR + tap % 3;
    } else {
      lofs[i] = -1;
    }
  }
  f32x4 acc[TPW][MT];
#pragma unroll
  for (int i = 0; i < TPW; ++i)
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[i][m] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.0f;
  for (int s = p0; s < p1; s += 32) {
    int q = s;
    const int x0 = q % W;
    q /= W;
    const int y = q % H;
    q /= H;
    const int z = q % D;
    const int n = q / D;
    // x pairs (xp, xp+1), xp even: one 32-bit word per channel
    for (int v = threadIdx.x; v < 9 * (XR / 2) * (CIN / 8); v += 256) {
      const int c8 = v % (CIN / 8), t = v / (CIN / 8);
      const int xp = 2 * (t % (XR / 2)), row = t / (XR / 2);
      const int zz = z + row / 3 - 1, yy = y + row % 3 - 1, xx = x0 - 1 + xp;
      uint4 u0 = make_uint4(0, 0, 0, 0), u1 = u0;
      if (zz >= 0 && zz < D && yy >= 0 && yy < H) {
        const u16 *src =
            x + ((((long)n * D + zz) * H + yy) * W + xx) * CIN + c8 * 8;
        if (xx >= 0 && xx < W) u0 = *reinterpret_cast<const uint4 *>(src);
        if (xx + 1 >= 0 && xx + 1 < W)
          u1 = *reinterpret_cast<const uint4 *>(src + CIN);
      }
      const unsigned a4[4] = {u0.x, u0.y, u0.z, u0.w};
      const unsigned b4[4] = {u1.x, u1.y, u1.z, u1.w};
      unsigned *dst = xT + ((row * CIN + c8 * 8) * XR + xp) / 2;
#pragma unroll
      for (int j = 0; j < 4; ++j) {  // channels 2j, 2j+1 of the vector
        dst[(2 * j) * (XR / 2)] = (a4[j] & 0xffffu) | (b4[j] << 16);
        dst[(2 * j + 1) * (XR / 2)] = (a4[j] >> 16) | (b4[j] & 0xffff0000u);
      }
    }
    for (int v = threadIdx.x; v < 32 * (COUT / 8); v += 256) {
      const int c8 = v % (COUT / 8), pp = v / (COUT / 8);
      const uint4 u = *reinterpret_cast<const uint4 *>(
          dz + (long)(s + pp) * COUT + c8 * 8);
      const unsigned w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        dzT[(c8 * 8 + 2 * j) * 32 + pp] = (u16)(w4[j] & 0xffffu);
        dzT[(c8 * 8 + 2 * j + 1) * 32 + pp] = (u16)(w4[j] >> 16);
      }
    }
    __syncthreads();
    if (threadIdx.x < COUT)
      for (int pp = 0; pp < 32; ++pp) bsum += Bf2F(dzT[threadIdx.x * 32 + pp]);
    bf16x8 a[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int co = m * 16 + r;
      a[m] = co < COUT
                 ? *reinterpret_cast<const bf16x8 *>(dzT + co * 32 + 8 * g)
                 : Zero8();
    }
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      if (wv + 4 * i >= NTN) break;  // wave-uniform
      bf16x8 b = Zero8();
      if (lofs[i] >= 0) {
        // elements e .. e+7 from the 5 words that cover them; an odd e
        // shifts each output word by one element
        const int e = lofs[i] + 8 * g;
        const unsigned *wsrc = xT + (e >> 1);
        unsigned w[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) w[k] = wsrc[k];
        uint4 o;
        if (e & 1) {
          o = make_uint4(__builtin_amdgcn_alignbit(w[1], w[0], 16),
                         __builtin_amdgcn_alignbit(w[2], w[1], 16),
                         __builtin_amdgcn_alignbit(w[3], w[2], 16),
                         __builtin_amdgcn_alignbit(w[4], w[3], 16));
        } else {
          o = make_uint4(w[0], w[1], w[2], w[3]);
        }
        b = __builtin_bit_cast(bf16x8, o);
      }
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[i][m] =
            __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[i][m], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    const int col = (wv + 4 * i) * 16 + r;
    if (wv + 4 * i >= NTN || col >= K) continue;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int co = m * 16 + 4 * g + v;
        if (co < COUT)
          part[((long)blockIdx.x * COUT + co) * K + col] = acc[i][m][v];
      }
  }
  if (threadIdx.x < COUT) bpart[(long)blockIdx.x * COUT + threadIdx.x] = bsum;
}

// Deterministic sum over blocks of part[b][i]: stage 1 sums segment s of
// the blocks in order into tmp[s][i], stage 2 sums the segments in order.
__global__ void SumSegmentsKernel(const float *__restrict__ part, int nparts,
                                  long stride, int seglen,
                                  float *__restrict__ tmp, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b0 = blockIdx.y * seglen;
  const int b1 = b0 + seglen < nparts ? b0 + seglen : nparts;
  float s = 0.0f;
  for (int b = b0; b < b1; ++b) s += part[b * stride + i];
  tmp[blockIdx.y * n + i] = s;
}

__global__ void SumPartsKernel(const float *__restrict__ part, int nparts,
                               long stride, float *__restrict__ out, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int b = 0; b < nparts; ++b) s += part[b * stride + i];
  out[i] = s;
}

// ------------------------------------------------------ elementwise layers
// Threads own 8 consecutive channels (one 16-byte vector) of one position;
// positions are decoded with 32-bit arithmetic (< 2^31 positions).
struct V8 {
  float v[8];
};
__device__ __forceinline__ V8 Load8F(const u16 *p) {
  const uint4 u = *reinterpret_cast<const uint4 *>(p);
  const unsigned w[4] = {u.x, u.y, u.z, u.w};
  V8 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    r.v[2 * i] = __uint_as_float(w[i] << 16);
    r.v[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
  }
  return r;
}
__device__ __forceinline__ void Store8F(u16 *p, const V8 &r) {
  unsigned w[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    w[i] = (unsigned)F2Bf(r.v[2 * i]) | ((unsigned)F2Bf(r.v[2 * i + 1]) << 16);
  *reinterpret_cast<uint4 *>(p) = make_uint4(w[0], w[1], w[2], w[3]);
}

struct Pos {
  int n, z, y, x;
};
__device__ __forceinline__ Pos Decode(int p, int D, int H, int W) {
  Pos r;
  r.x = p % W;
  p /= W;
  r.y = p % H;
  p /= H;
  r.z = p % D;
  r.n = p / D;
  return r;
}

// 4x4x4 average pool: out[coarse][c] = (FP32 sum of the 64 inputs) / 64.
__global__ void PoolKernel(const u16 *__restrict__ in, u16 *__restrict__ out,
                           int C, int Dc, int Hc, int Wc, int total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int G = C / 8;
  if (i >= total) return;
  const int gi = i % G;
  const Pos q = Decode(i / G, Dc, Hc, Wc);
  const int D = Dc * kPool, H = Hc * kPool, W = Wc * kPool;
  V8 s;
#pragma unroll
  for (int k = 0; k < 8; ++k) s.v[k] = 0.0f;
  for (int dz = 0; dz < kPool; ++dz)
    for (int dy = 0; dy < kPool; ++dy)
      for (int dx = 0; dx < kPool; ++dx) {
        const long f = (((long)q.n * D + q.z * kPool + dz) * H + q.y * kPool +
                        dy) * W + q.x * kPool + dx;
        const V8 a = Load8F(in + f * C + gi * 8);
#pragma unroll
        for (int k = 0; k < 8; ++k) s.v[k] += a.v[k];
      }
#pragma unroll
  for (int k = 0; k < 8; ++k) s.v[k] *= (1.0f / 64.0f);
  Store8F(out + (long)i * 8, s);
}

// out[fine p][0:CA] = coarse[p / 4][0:CA]; out[p][CA:CA+CB] = skip[p].
__global__ void UpConcatKernel(const u16 *__restrict__ coarse,
                               const u16 *__restrict__ skip,
                               u16 *__restrict__ out, int CA, int CB, int D,
                               int H, int W, int total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int G = (CA + CB) / 8, GA = CA / 8;
  if (i >= total) return;
  const int gi = i % G, p = i / G;
  uint4 v;
  if (gi >= GA) {
    v = *reinterpret_cast<const uint4 *>(skip + (long)p * CB + (gi - GA) * 8);
  } else {
    const Pos q = Decode(p, D, H, W);
    const int Dc = D / kPool, Hc = H / kPool, Wc = W / kPool;
    const long c = (((long)q.n * Dc + q.z / kPool) * Hc + q.y / kPool) * Wc +
                   q.x / kPool;
    v = *reinterpret_cast<const uint4 *>(coarse + c * CA + gi * 8);
  }
  *reinterpret_cast<uint4 *>(out + (long)i * 8) = v;
}

// Gradient of the upsampled part of a concatenation: gc[coarse][c] = FP32
// sum over the 64 fine positions of g[fine][c], c < CA of C channels.
__global__ void UpBackKernel(const u16 *__restrict__ g, u16 *__restrict__ gc,
                             int CA, int C, int Dc, int Hc, int Wc,
                             int total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int GA = CA / 8;
  if (i >= total) return;
  const int gi = i % GA;
  const Pos q = Decode(i / GA, Dc, Hc, Wc);
  const int D = Dc * kPool, H = Hc * kPool, W = Wc * kPool;
  V8 s;
#pragma unroll
  for (int k = 0; k < 8; ++k) s.v[k] = 0.0f;
  for (int dz = 0; dz < kPool; ++dz)
    for (int dy = 0; dy < kPool; ++dy)
      for (int dx = 0; dx < kPool; ++dx) {
        const long f = (((long)q.n * D + q.z * kPool + dz) * H + q.y * kPool +
                        dy) * W + q.x * kPool + dx;
        const V8 a = Load8F(g + f * C + gi * 8);
#pragma unroll
        for (int k = 0; k < 8; ++k) s.v[k] += a.v[k];
      }
  Store8F(gc + (long)i * 8, s);
}

// dz[p][c] = (gpool[p/4][c] / 64 + gskip[p][CA + c]) if act[p][c] > 0 else 0:
// the gradient reaching a pooled encoder output through its pool and its
// skip connection (gskip: the concatenation gradient, CA + CB channels),
// gated by its ReLU.  FP32 sum, one rounding.
__global__ void PoolBackSkipGateKernel(const u16 *__restrict__ gpool,
                                       const u16 *__restrict__ gskip,
                                       const u16 *__restrict__ act,
                                       u16 *__restrict__ dz, int CB, int CA,
                                       int D, int H, int W, int total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int G = CB / 8;
  if (i >= total) return;
  const int gi = i % G, p = i / G;
  const Pos q = Decode(p, D, H, W);
  const int Dc = D / kPool, Hc = H / kPool, Wc = W / kPool;
  const long c = (((long)q.n * Dc + q.z / kPool) * Hc + q.y / kPool) * Wc +
                 q.x / kPool;
  const V8 gp = Load8F(gpool + c * CB + gi * 8);
  const V8 gs = Load8F(gskip + (long)p * (CA + CB) + CA + gi * 8);
  const V8 a = Load8F(act + (long)i * 8);
  V8 o;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    o.v[k] = a.v[k] > 0.0f ? gp.v[k] * (1.0f / 64.0f) + gs.v[k] : 0.0f;
  Store8F(dz + (long)i * 8, o);
}

// g *= (act > 0): ReLU gradient, in place, 8 values per thread.
__global__ void GateKernel(u16 *__restrict__ g, const u16 *__restrict__ act,
                           long n8) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n8) return;
  uint4 gv = *reinterpret_cast<const uint4 *>(g + i * 8);
  const uint4 av = *reinterpret_cast<const uint4 *>(act + i * 8);
  unsigned *gw = reinterpret_cast<unsigned *>(&gv);
  const unsigned aw[4] = {av.x, av.y, av.z, av.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool lo = __uint_as_float(aw[k] << 16) > 0.0f;
    const bool hi = __uint_as_float(aw[k] & 0xffff0000u) > 0.0f;
    gw[k] = (lo ? gw[k] & 0xffffu : 0u) | (hi ? gw[k] & 0xffff0000u : 0u);
  }
  *reinterpret_cast<uint4 *>(g + i * 8) = gv;
}

// Network input: ch0 = rhs, ch1 = bnd, ch2..7 = 0 (problems [first, ...)).
__global__ void InputKernel(const float *__restrict__ rhs,
                            const float *__restrict__ bnd,
                            u16 *__restrict__ x, long off, long P) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint4 v = make_uint4(0, 0, 0, 0);
  v.x = (unsigned)F2Bf(rhs[off + p]) | ((unsigned)F2Bf(bnd[off + p]) << 16);
  *reinterpret_cast<uint4 *>(x + p * 8) = v;
}

// 1x1 convolution 8 -> 1: out[p] = sum_c w[c] in[p][c] + b (FP32).
__global__ void HeadKernel(const u16 *__restrict__ in,
                           const float *__restrict__ w,
                           float *__restrict__ out, long P) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= P) return;
  float s = 0.0f;
  for (int c = 0; c < 8; ++c) s += w[c] * Bf2F(in[p * 8 + c]);
  out[p] = s + w[8];
}

// Smooth-error terms of the loss.  The error d = out - target * inv_scale
// averaged over the 4x4x4 cells of the pooling grid: one wave per cell
// (lane l at offset (l & 3, (l >> 2) & 3, l >> 4)), summed across the wave
// in a fixed order.  e4 has the level-2 layout ((n D2 + Z) H2 + Y) W2 + X.
__global__ __launch_bounds__(256) void SmoothCellKernel(
    const float *__restrict__ out, const float *__restrict__ target,
    long toff, float inv_scale, float *__restrict__ e4, int D1, int H1,
    int W1, long P2) {
  const long c = (blockIdx.x * (long)blockDim.x + threadIdx.x) >> 6;
  const int l = threadIdx.x & 63;
  if (c >= P2) return;  // uniform over the wave
  const int W2 = W1 / kPool, H2 = H1 / kPool, D2 = D1 / kPool;
  const int X = (int)(c % W2), Y = (int)((c / W2) % H2);
  const long nz = c / ((long)W2 * H2);  // n D2 + Z
  const long n = nz / D2, Z = nz - n * D2;
  const long p = ((n * D1 + Z * kPool + (l >> 4)) * H1 + Y * kPool +
                  ((l >> 2) & 3)) * W1 + X * kPool + (l & 3);
  float v = out[p] - target[toff + p] * inv_scale;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  if (l == 0) e4[c] = v * (1.0f / 64.0f);
}

// e16: the means of e4 over 4x4x4 cells of the level-2 grid (the error
// averaged over 16x16x16 cells), one thread per cell, summed z-y-x.
__global__ void SmoothCoarseKernel(const float *__restrict__ e4,
                                   float *__restrict__ e16, int D2, int H2,
                                   int W2, long P3) {
  const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= P3) return;
  const int W3 = W2 / kPool, H3 = H2 / kPool, D3 = D2 / kPool;
  const int X = (int)(c % W3), Y = (int)((c / W3) % H3);
  const long nz = c / ((long)W3 * H3);
  const long n = nz / D3, Z = nz - n * D3;
  float s = 0.0f;
  for (int dz = 0; dz < kPool; ++dz)
    for (int dy = 0; dy < kPool; ++dy)
      for (int dx = 0; dx < kPool; ++dx)
        s += e4[((n * D2 + Z * kPool + dz) * H2 + Y * kPool + dy) * W2 +
                X * kPool + dx];
  e16[c] = s * (1.0f / 64.0f);
}

// Loss partials and its gradient dout.  With d = out - target * inv_scale,
// e4 / e16 the means of d over the point's 4^3 / 16^3 cell:
//   loss = (1/P) sum_p (d^2 + w4 e4^2 + w16 e16^2)
//        = mean d^2 + w4 mean over 4^3 cells e4^2 + w16 mean e16^2,
//   dout = 2 (d + w4 e4 + w16 e16) / P.
// SMOOTH false: w4 = w16 = 0 (e4, e16 unused).
template <bool SMOOTH>
__global__ __launch_bounds__(256) void LossKernel(
    const float *__restrict__ out, const float *__restrict__ target,
    long toff, float inv_scale, const float *__restrict__ e4,
    const float *__restrict__ e16, float w4, float w16, int D1, int H1,
    int W1, float *__restrict__ dout, double *__restrict__ part, long P) {
  __shared__ double red[4];
  double s = 0.0;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P;
       p += (long)gridDim.x * blockDim.x) {
    const float d = out[p] - target[toff + p] * inv_scale;
    if constexpr (SMOOTH) {
      const int x = (int)(p % W1), y = (int)((p / W1) % H1);
      const long nz = p / ((long)W1 * H1);
      const long n = nz / D1, z = nz - n * D1;
      const int W2 = W1 / kPool, H2 = H1 / kPool, D2 = D1 / kPool;
      const long c2 = ((n * D2 + z / kPool) * H2 + y / kPool) * W2 +
                      x / kPool;
      const int W3 = W2 / kPool, H3 = H2 / kPool, D3 = D2 / kPool;
      const long c3 = ((n * D3 + z / (kPool * kPool)) * H3 +
                       y / (kPool * kPool)) * W3 + x / (kPool * kPool);
      const float a = e4[c2], b = e16[c3];
      dout[p] = 2.0f * (d + w4 * a + w16 * b) / (float)P;
      s += (double)d * d + (double)w4 * a * a + (double)w16 * b * b;
    } else {
      dout[p] = 2.0f * d / (float)P;
      s += (double)d * d;
    }
  }
  s = physis_hip::BlockReduce<double, PS_SUM, 256>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Head backward: din[p][c] = w[c] dout[p] (BF16, gated by act), and
// per-block partials of dw[c] = sum dout[p] in[p][c], db = sum dout[p].
__global__ __launch_bounds__(256) void HeadBackKernel(
    const u16 *__restrict__ in, const float *__restrict__ w,
    const float *__restrict__ dout, u16 *__restrict__ din,
    float *__restrict__ part, long P) {
  __shared__ float red[9][4];
  float acc[9];
  for (int c = 0; c < 9; ++c) acc[c] = 0.0f;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < P;
       p += (long)gridDim.x * blockDim.x) {
    const float d = dout[p];
    for (int c = 0; c < 8; ++c) {
      const float a = Bf2F(in[p * 8 + c]);
      acc[c] += d * a;
      din[p * 8 + c] = a > 0.0f ? F2Bf(w[c] * d) : (u16)0;
    }
    acc[8] += d;
  }
  for (int c = 0; c < 9; ++c) {
    const float s = physis_hip::BlockReduce<float, PS_SUM, 256>(acc[c], red[c]);
    if (threadIdx.x == 0) part[blockIdx.x * 9 + c] = s;
  }
}

// Prediction: p[off + i] = scale * out[i] * bnd[off + i] (FP32 or FP64 p).
template <typename T>
__global__ void PredictKernel(const float *__restrict__ out,
                              const float *__restrict__ bnd, T *__restrict__ p,
                              long off, float scale, long P) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < P) p[off + i] = (T)(scale * out[i]) * (T)bnd[off + i];
}

// Weights: wt[co][k] = bf16(W[co][k]) (zero for k >= 27 cin) and the
// flipped, transposed wf[ci][tap' cout + co] = bf16(W[co][26 - tap'][ci]).
__global__ void PackKernel(const float *__restrict__ w, u16 *__restrict__ wt,
                           u16 *__restrict__ wf, int cin, int cout) {
  const int K = 27 * cin, kp = KPad(cin), kf = KPad(cout);
  const int n1 = cout * kp, n2 = cin * kf;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n1 + n2;
       i += gridDim.x * blockDim.x) {
    if (i < n1) {
      const int co = i / kp, k = i - co * kp;
      wt[i] = k < K ? F2Bf(w[co * K + k]) : (u16)0;
    } else {
      const int j = i - n1, ci = j / kf, k = j - ci * kf;
      u16 v = 0;
      if (k < 27 * cout) {
        const int tp = k / cout, co = k - tp * cout;
        v = F2Bf(w[co * K + (26 - tp) * cin + ci]);
      }
      wf[j] = v;
    }
  }
}

// Reflected copy of whole problems: dst problem q' = dst_first + q holds
// src problem first + q with the axes whose bit is set in `flips` (1 x,
// 2 y, 4 z) reversed inside the problem.
__global__ void ReflectKernel(const float *__restrict__ src,
                              float *__restrict__ dst, int nx, int ny,
                              int nzs, long soff, long doff, int flips,
                              long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long m = (long)nx * ny * nzs;
  const long q = i / m;
  long r = i - q * m;
  const int z = (int)(r / ((long)nx * ny));
  r -= (long)z * nx * ny;
  const int y = (int)(r / nx), x = (int)(r - (long)y * nx);
  const int sx = flips & 1 ? nx - 1 - x : x;
  const int sy = flips & 2 ? ny - 1 - y : y;
  const int sz = flips & 4 ? nzs - 1 - z : z;
  dst[doff + i] = src[soff + q * m + ((long)sz * ny + sy) * nx + sx];
}

__global__ void AdamKernel(float *__restrict__ p, const float *__restrict__ g,
                           float *__restrict__ m, float *__restrict__ v,
                           long n, float lr, float c1, float c2) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// ------------------------------------------------------------ dispatch
hipStream_t Stream() { return (hipStream_t)__PSHipStream(); }
void Check() { PS_HIP_CHECK(hipGetLastError()); }
unsigned Blocks(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template <int CIN, int COUT, int XT>
void ConvLaunch(const u16 *x, const u16 *wt, const float *b, bool relu,
                u16 *y, int N, int D, int H, int W) {
  const int tiles_x = (W + XT - 1) / XT;
  const dim3 grid(tiles_x * ((H + 3) / 4), N * D), block(256);
  if (relu)
    hipLaunchKernelGGL((ConvKernel<CIN, COUT, true, XT>), grid, block, 0,
                       Stream(), x, wt, b, y, D, H, W);
  else
    hipLaunchKernelGGL((ConvKernel<CIN, COUT, false, XT>), grid, block, 0,
                       Stream(), x, wt, b, y, D, H, W);
  Check();
}

// Blocks per compute unit that the z-marching launches aim for (measured
// at level 1 of batch 8 x 128^3: the plane-marching filter gradient is
// fastest with 2, e.g. 8 -> 8 213 us against 223 with 4 and 237 with 8,
// and so are the convolutions).
constexpr int kConvZBlocksPerCU = 2;
constexpr int kWgradZBlocksPerCU = 2;

// z chunks of a z-marching launch: enough blocks for per_cu per compute
// unit over `tiles` xy tiles of N problems of D planes.
void ZChunks(long tiles, int N, int D, int per_cu, int *tz, int *zchunk) {
  const long want = (long)per_cu * __PSHipNumCUs();
  long t = (want + tiles * N - 1) / (tiles * N);
  if (t < 1) t = 1;
  if (t > D) t = D;
  *zchunk = (int)((D + t - 1) / t);
  *tz = (D + *zchunk - 1) / *zchunk;
}

template <int CIN, int COUT, int XT, int TY>
void ConvZLaunch(const u16 *x, const u16 *wt, const float *b, bool relu,
                 u16 *y, int N, int D, int H, int W) {
  const int tiles = ((W + XT - 1) / XT) * ((H + TY - 1) / TY);
  int tz, zchunk;
  ZChunks(tiles, N, D, kConvZBlocksPerCU, &tz, &zchunk);
  const dim3 grid(tiles, N * tz), block(64 * TY);
  if (relu)
    hipLaunchKernelGGL((ConvZKernel<CIN, COUT, true, XT, TY>), grid, block, 0,
                       Stream(), x, wt, b, y, D, H, W, tz, zchunk);
  else
    hipLaunchKernelGGL((ConvZKernel<CIN, COUT, false, XT, TY>), grid, block,
                       0, Stream(), x, wt, b, y, D, H, W, tz, zchunk);
  Check();
}

template <int CIN, int XT, int TY>
void ConvZPairLaunch(const u16 *x, const u16 *wt, const float *b, bool relu,
                     u16 *y, int N, int D, int H, int W) {
  const int tiles = ((W + XT - 1) / XT) * ((H + 2 * TY - 1) / (2 * TY));
  int tz, zchunk;
  ZChunks(tiles, N, D, kConvZBlocksPerCU, &tz, &zchunk);
  const dim3 grid(tiles, N * tz), block(64 * TY);
  if (relu)
    hipLaunchKernelGGL((ConvZPairKernel<CIN, true, XT, TY>), grid, block, 0,
                       Stream(), x, wt, b, y, D, H, W, tz, zchunk);
  else
    hipLaunchKernelGGL((ConvZPairKernel<CIN, false, XT, TY>), grid, block, 0,
                       Stream(), x, wt, b, y, D, H, W, tz, zchunk);
  Check();
}

// z-marching convolution: x tile of 64 positions below 24 input channels,
// else 32, 16 for rows of at most 16; 8 rows per block below 24 input
// channels, else 4 (the ring of 4 staged planes then stays within LDS).
template <int CIN, int COUT>
void ConvT(const u16 *x, const u16 *wt, const float *b, bool relu, u16 *y,
           int N, int D, int H, int W) {
  if constexpr (COUT == 8) {
    // two output rows per wave: 8 rows per block of 4 waves
    if constexpr (CIN < 24) {
      if (W > 32)
        return ConvZPairLaunch<CIN, 64, 4>(x, wt, b, relu, y, N, D, H, W);
    }
    if (W > 16)
      return ConvZPairLaunch<CIN, 32, 4>(x, wt, b, relu, y, N, D, H, W);
    return ConvZPairLaunch<CIN, 16, 4>(x, wt, b, relu, y, N, D, H, W);
  }
  constexpr int TY = CIN < 24 ? 8 : 4;
  if constexpr (CIN < 24) {
    if (W > 32)
      return ConvZLaunch<CIN, COUT, 64, TY>(x, wt, b, relu, y, N, D, H, W);
  }
  if (W > 16)
    ConvZLaunch<CIN, COUT, 32, TY>(x, wt, b, relu, y, N, D, H, W);
  else
    ConvZLaunch<CIN, COUT, 16, TY>(x, wt, b, relu, y, N, D, H, W);
}

// Positions per filter-gradient block: about P / 512 (a multiple of 32),
// between 256 and kWgradChunk, so that small levels still spread over many
// blocks.
int WgradChunk(long P) {
  long c = ((P / 512 + 31) / 32) * 32;
  if (c < 256) c = 256;
  if (c > kWgradChunk) c = kWgradChunk;
  return (int)c;
}

// Launch shape of the z-marching filter gradient (rows of a multiple of 32
// positions): x tile 64 when the row allows it and the input has at most
// 24 channels (24 -> 8 at level 1: 479 us against 526 with 32), else 32;
// 4 rows per block; z chunks as the convolution.
// Returns the block count (= partials).
constexpr int kWgradZRows = 4;
// Wave groups over the column tiles (9 cin / 16 of them): one up to 5
// tiles, two up to 14, else four (the rest of the 4 waves split the steps).
constexpr int WgradZGroups(int cin) {
  return (9 * cin + 15) / 16 <= 5 ? 1 : ((9 * cin + 15) / 16 <= 14 ? 2 : 4);
}
int WgradZXT(int cin, int W) { return (W % 64 == 0 && cin <= 24) ? 64 : 32; }
long WgradZBlocks(int cin, int N, int D, int H, int W, int *tz, int *zchunk) {
  const int xt = WgradZXT(cin, W);
  const long tiles = (long)((W + xt - 1) / xt) *
                     ((H + kWgradZRows - 1) / kWgradZRows);
  ZChunks(tiles, N, D, kWgradZBlocksPerCU, tz, zchunk);
  return tiles * N * *tz;
}

// Partials (blocks) of a filter-gradient launch over N problems of
// D x H x W positions with cin input channels: every buffer of partials is
// sized by this.
long WgradBlocks(int cin, int N, int D, int H, int W) {
  if (W % 32 == 0) {
    int tz, zchunk;
    return WgradZBlocks(cin, N, D, H, W, &tz, &zchunk);
  }
  const long P = (long)N * D * H * W;
  const long chunk = WgradChunk(P);
  return (P + chunk - 1) / chunk;
}

// Launches the filter gradient into part / bpart, which hold `capacity`
// blocks of partials; *nparts receives the blocks written.
template <int CIN, int COUT>
void WgradT(const u16 *x, const u16 *dz, float *part, float *bpart, int N,
            int D, int H, int W, long capacity, int *nparts) {
  constexpr int NTN = (27 * CIN + 15) / 16;
  const long P = (long)N * D * H * W;
  if (P >= (1L << 31) - kWgradChunk)
    PS_FATAL("filter gradient: %ld positions exceed 32-bit indexing", P);
  if (WgradBlocks(CIN, N, D, H, W) > capacity)
    PS_FATAL("filter gradient: %ld partial blocks exceed the %ld allocated",
             WgradBlocks(CIN, N, D, H, W), capacity);
  // rows of a multiple of 32 positions: the z-marching kernel
  if (W % 32 == 0) {
    int tz, zchunk;
    const long nb = WgradZBlocks(CIN, N, D, H, W, &tz, &zchunk);
    const int tiles = (int)(nb / ((long)N * tz));
    const dim3 grid(tiles, N * tz), block(64 * kWgradZRows);
    if (WgradZXT(CIN, W) == 64) {
      if constexpr (CIN <= 24)
        hipLaunchKernelGGL(
            (WgradZKernel<CIN, COUT, 64, kWgradZRows, WgradZGroups(CIN)>),
            grid, block, 0, Stream(), x, dz, part, bpart, D, H, W, tz,
            zchunk);
    } else {
      hipLaunchKernelGGL(
          (WgradZKernel<CIN, COUT, 32, kWgradZRows, WgradZGroups(CIN)>), grid,
          block, 0, Stream(), x, dz, part, bpart, D, H, W, tz, zchunk);
    }
    Check();
    *nparts = (int)nb;
    return;
  }
  const int chunk = WgradChunk(P);
  const int nb = (int)((P + chunk - 1) / chunk);
  hipLaunchKernelGGL((WgradKernel<CIN, COUT>), dim3(nb, (NTN + 3) / 4),
                     dim3(256), 0, Stream(), x, dz, part, bpart, (int)P, D,
                     H, W, chunk);
  Check();
  *nparts = nb;
}

// out[0..n) = sum over nparts blocks of part[b * n + i], via tmp
// (kSumSegments * n floats).
constexpr int kSumSegments = 64;
void SumParts(const float *part, int nparts, long n, float *tmp, float *out) {
  const int seg = nparts < kSumSegments ? nparts : kSumSegments;
  const int seglen = (nparts + seg - 1) / seg;
  const int nseg = (nparts + seglen - 1) / seglen;
  hipLaunchKernelGGL(SumSegmentsKernel, dim3(Blocks(n), nseg), dim3(256), 0,
                     Stream(), part, nparts, n, seglen, tmp, n);
  Check();
  hipLaunchKernelGGL(SumPartsKernel, dim3(Blocks(n)), dim3(256), 0, Stream(),
                     tmp, nseg, n, out, n);
  Check();
}

#define PS_UNET_PAIRS(X) \
  X(8, 8) X(8, 16) X(16, 16) X(16, 32) X(32, 32) X(48, 16) X(24, 8) \
  X(16, 8) X(32, 16) X(16, 48) X(8, 24)

void Conv(int cin, int cout, const u16 *x, const u16 *wt, const float *b,
          bool relu, u16 *y, int N, int D, int H, int W) {
#define PS_CASE(I, O) \
  if (cin == I && cout == O) return ConvT<I, O>(x, wt, b, relu, y, N, D, H, W);
  PS_UNET_PAIRS(PS_CASE)
#undef PS_CASE
  PS_FATAL("no convolution kernel for %d -> %d channels", cin, cout);
}

void Wgrad(int cin, int cout, const u16 *x, const u16 *dz, float *part,
           float *bpart, int N, int D, int H, int W, long capacity,
           int *nparts) {
#define PS_CASE(I, O) \
  if (cin == I && cout == O) \
    return WgradT<I, O>(x, dz, part, bpart, N, D, H, W, capacity, nparts);
  PS_UNET_PAIRS(PS_CASE)
#undef PS_CASE
  PS_FATAL("no filter-gradient kernel for %d -> %d channels", cin, cout);
}

struct Layer {
  int cin, cout;
  size_t w_off, b_off;  // into the flat parameter vector
  u16 *wt = nullptr, *wf = nullptr;
};

void *Alloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 16;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}

uint64_t SplitMix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace physis_unet

using namespace physis_unet;

struct __PSUNet {
  int nx, ny, nzs, max_batch;
  float out_scale;
  std::vector<Layer> layers;  // 10 convolutions
  size_t head_off, nparams;
  float *params = nullptr, *grad = nullptr, *adam_m = nullptr,
        *adam_v = nullptr;
  long step = 0;
  // activations (BF16) at batch max_batch
  u16 *x0 = nullptr, *a[10] = {}, *q1 = nullptr, *q2 = nullptr,
      *c2 = nullptr, *c1 = nullptr;
  float *out = nullptr, *dout = nullptr;
  // gradient scratch
  u16 *g1a = nullptr, *g1b = nullptr, *gc1 = nullptr;  // level 1
  u16 *g2a = nullptr, *g2b = nullptr, *gc2 = nullptr;  // level 2
  u16 *g3a = nullptr, *g3b = nullptr;                  // level 3
  float *part = nullptr, *bpart = nullptr, *tmp = nullptr;
  double *lpart = nullptr, *lhost = nullptr;
  size_t part_elems = 0, bpart_elems = 0;
  // smooth-error loss weights and the cell means of the error
  float w4 = 0.0f, w16 = 0.0f;
  float *e4 = nullptr, *e16 = nullptr;
};

namespace {

const int kLayerCh[10][2] = {{8, 8},  {8, 8},   {8, 16},  {16, 16}, {16, 32},
                             {32, 32}, {48, 16}, {16, 16}, {24, 8},  {8, 8}};
constexpr int kLossBlocks = 1024;
constexpr int kHeadBlocks = 1024;

void Upload(__PSUNet *net) {  // master weights -> BF16 copies
  for (auto &L : net->layers) {
    hipLaunchKernelGGL(PackKernel, dim3(64), dim3(256), 0, Stream(),
                       net->params + L.w_off, L.wt, L.wf, L.cin, L.cout);
    Check();
  }
}

struct Dims {
  int N, D1, H1, W1, D2, H2, W2, D3, H3, W3;
  long P1, P2, P3;
};

Dims DimsOf(const __PSUNet *net, int count) {
  Dims d;
  d.N = count;
  d.D1 = net->nzs; d.H1 = net->ny; d.W1 = net->nx;
  d.D2 = d.D1 / kPool; d.H2 = d.H1 / kPool; d.W2 = d.W1 / kPool;
  d.D3 = d.D2 / kPool; d.H3 = d.H2 / kPool; d.W3 = d.W2 / kPool;
  d.P1 = (long)count * d.D1 * d.H1 * d.W1;
  d.P2 = (long)count * d.D2 * d.H2 * d.W2;
  d.P3 = (long)count * d.D3 * d.H3 * d.W3;
  return d;
}

void ConvL(const __PSUNet *net, int l, const u16 *in, u16 *out, int N, int D,
           int H, int W) {
  const Layer &L = net->layers[l];
  Conv(L.cin, L.cout, in, L.wt, net->params + L.b_off, true, out, N, D, H, W);
}

void CheckGrid(const __PSUNet *net, const __PSGrid *g, int first, int count,
               const char *what) {
  if (!g || !g->p || g->num_dims != 3 || g->dim[0] != net->nx ||
      g->dim[1] != net->ny || g->dim[2] % net->nzs != 0)
    PS_FATAL("U-Net: %s is not a 3-D grid of %d x %d x (k * %d)", what,
             net->nx, net->ny, net->nzs);
  const int nb = g->dim[2] / net->nzs;
  if (first < 0 || count < 1 || count > net->max_batch || first + count > nb)
    PS_FATAL("U-Net: problems [%d, %d) outside 0..%d or above the batch %d",
             first, first + count, nb, net->max_batch);
}

// Forward pass of problems [first, first+count): fills the activations and
// net->out.
void Forward(__PSUNet *net, const __PSGrid *rhs, const __PSGrid *bnd,
             int first, int count) {
  CheckGrid(net, rhs, first, count, "rhs");
  CheckGrid(net, bnd, first, count, "bnd");
  if (rhs->type != PS_FLOAT || bnd->type != PS_FLOAT)
    PS_FATAL("U-Net: rhs and bnd must be float grids");
  const Dims d = DimsOf(net, count);
  const long off = (long)first * net->nx * net->ny * net->nzs;
  hipLaunchKernelGGL(InputKernel, dim3(Blocks(d.P1)), dim3(256), 0, Stream(),
                     (const float *)rhs->p, (const float *)bnd->p, net->x0,
                     off, d.P1);
  Check();
  ConvL(net, 0, net->x0, net->a[0], d.N, d.D1, d.H1, d.W1);
  ConvL(net, 1, net->a[0], net->a[1], d.N, d.D1, d.H1, d.W1);
  hipLaunchKernelGGL(PoolKernel, dim3(Blocks(d.P2)), dim3(256), 0, Stream(),
                     net->a[1], net->q1, 8, d.D2, d.H2, d.W2, (int)d.P2);
  Check();
  ConvL(net, 2, net->q1, net->a[2], d.N, d.D2, d.H2, d.W2);
  ConvL(net, 3, net->a[2], net->a[3], d.N, d.D2, d.H2, d.W2);
  hipLaunchKernelGGL(PoolKernel, dim3(Blocks(d.P3 * 2)), dim3(256), 0,
                     Stream(), net->a[3], net->q2, 16, d.D3, d.H3, d.W3,
                     (int)(d.P3 * 2));
  Check();
  ConvL(net, 4, net->q2, net->a[4], d.N, d.D3, d.H3, d.W3);
  ConvL(net, 5, net->a[4], net->a[5], d.N, d.D3, d.H3, d.W3);
  hipLaunchKernelGGL(UpConcatKernel, dim3(Blocks(d.P2 * 6)), dim3(256), 0,
                     Stream(), net->a[5], net->a[3], net->c2, 32, 16, d.D2,
                     d.H2, d.W2, (int)(d.P2 * 6));
  Check();
  ConvL(net, 6, net->c2, net->a[6], d.N, d.D2, d.H2, d.W2);
  ConvL(net, 7, net->a[6], net->a[7], d.N, d.D2, d.H2, d.W2);
  hipLaunchKernelGGL(UpConcatKernel, dim3(Blocks(d.P1 * 3)), dim3(256), 0,
                     Stream(), net->a[7], net->a[1], net->c1, 16, 8, d.D1,
                     d.H1, d.W1, (int)(d.P1 * 3));
  Check();
  ConvL(net, 8, net->c1, net->a[8], d.N, d.D1, d.H1, d.W1);
  ConvL(net, 9, net->a[8], net->a[9], d.N, d.D1, d.H1, d.W1);
  hipLaunchKernelGGL(HeadKernel, dim3(Blocks(d.P1)), dim3(256), 0, Stream(),
                     net->a[9], net->params + net->head_off, net->out, d.P1);
  Check();
}

// Filter gradient of layer l from input `in` and gated output gradient dz.
void LayerWgrad(__PSUNet *net, int l, const u16 *in, const u16 *dz, int N,
                int D, int H, int W) {
  const Layer &L = net->layers[l];
  int np = 0;
  const long nw = (long)L.cout * 27 * L.cin;
  const long cap = std::min<long>((long)net->part_elems / nw,
                                  (long)net->bpart_elems / L.cout);
  Wgrad(L.cin, L.cout, in, dz, net->part, net->bpart, N, D, H, W, cap, &np);
  SumParts(net->part, np, nw, net->tmp, net->grad + L.w_off);
  SumParts(net->bpart, np, (long)L.cout, net->tmp, net->grad + L.b_off);
}

// Data gradient of layer l: dx = conv(dz, flipped W), no bias, no ReLU.
void LayerDgrad(__PSUNet *net, int l, const u16 *dz, u16 *dx, int N, int D,
                int H, int W) {
  const Layer &L = net->layers[l];
  Conv(L.cout, L.cin, dz, L.wf, nullptr, false, dx, N, D, H, W);
}

void Gate(u16 *g, const u16 *act, long n) {  // n: a multiple of 8
  hipLaunchKernelGGL(GateKernel, dim3(Blocks(n / 8)), dim3(256), 0, Stream(),
                     g, act, n / 8);
  Check();
}

// Loss of the current forward pass against target, and the full gradient
// into net->grad.
double Backward(__PSUNet *net, const __PSGrid *target, int first, int count) {
  CheckGrid(net, target, first, count, "target");
  if (target->type != PS_FLOAT) PS_FATAL("U-Net: target must be a float grid");
  const Dims d = DimsOf(net, count);
  const long off = (long)first * net->nx * net->ny * net->nzs;
  const float inv_scale = 1.0f / net->out_scale;
  if (net->w4 != 0.0f || net->w16 != 0.0f) {
    hipLaunchKernelGGL(SmoothCellKernel, dim3(Blocks(d.P2 * 64)), dim3(256),
                       0, Stream(), net->out, (const float *)target->p, off,
                       inv_scale, net->e4, d.D1, d.H1, d.W1, d.P2);
    Check();
    hipLaunchKernelGGL(SmoothCoarseKernel, dim3(Blocks(d.P3)), dim3(256), 0,
                       Stream(), net->e4, net->e16, d.D2, d.H2, d.W2, d.P3);
    Check();
    hipLaunchKernelGGL(LossKernel<true>, dim3(kLossBlocks), dim3(256), 0,
                       Stream(), net->out, (const float *)target->p, off,
                       inv_scale, net->e4, net->e16, net->w4, net->w16, d.D1,
                       d.H1, d.W1, net->dout, net->lpart, d.P1);
  } else {
    hipLaunchKernelGGL(LossKernel<false>, dim3(kLossBlocks), dim3(256), 0,
                       Stream(), net->out, (const float *)target->p, off,
                       inv_scale, nullptr, nullptr, 0.0f, 0.0f, d.D1, d.H1,
                       d.W1, net->dout, net->lpart, d.P1);
  }
  Check();
  // head: g1a = gated dD1 (layer 9's dz), head partials in part
  hipLaunchKernelGGL(HeadBackKernel, dim3(kHeadBlocks), dim3(256), 0,
                     Stream(), net->a[9], net->params + net->head_off,
                     net->dout, net->g1a, net->part, d.P1);
  Check();
  SumParts(net->part, kHeadBlocks, 9L, net->tmp, net->grad + net->head_off);
  // level 1 decoder
  LayerWgrad(net, 9, net->a[8], net->g1a, d.N, d.D1, d.H1, d.W1);
  LayerDgrad(net, 9, net->g1a, net->g1b, d.N, d.D1, d.H1, d.W1);
  Gate(net->g1b, net->a[8], d.P1 * 8);
  LayerWgrad(net, 8, net->c1, net->g1b, d.N, d.D1, d.H1, d.W1);
  LayerDgrad(net, 8, net->g1b, net->gc1, d.N, d.D1, d.H1, d.W1);  // 24 ch
  hipLaunchKernelGGL(UpBackKernel, dim3(Blocks(d.P2 * 2)), dim3(256), 0,
                     Stream(), net->gc1, net->g2a, 16, 24, d.D2, d.H2, d.W2,
                     (int)(d.P2 * 2));
  Check();
  // level 2 decoder
  Gate(net->g2a, net->a[7], d.P2 * 16);
  LayerWgrad(net, 7, net->a[6], net->g2a, d.N, d.D2, d.H2, d.W2);
  LayerDgrad(net, 7, net->g2a, net->g2b, d.N, d.D2, d.H2, d.W2);
  Gate(net->g2b, net->a[6], d.P2 * 16);
  LayerWgrad(net, 6, net->c2, net->g2b, d.N, d.D2, d.H2, d.W2);
  LayerDgrad(net, 6, net->g2b, net->gc2, d.N, d.D2, d.H2, d.W2);  // 48 ch
  hipLaunchKernelGGL(UpBackKernel, dim3(Blocks(d.P3 * 4)), dim3(256), 0,
                     Stream(), net->gc2, net->g3a, 32, 48, d.D3, d.H3, d.W3,
                     (int)(d.P3 * 4));
  Check();
  // level 3
  Gate(net->g3a, net->a[5], d.P3 * 32);
  LayerWgrad(net, 5, net->a[4], net->g3a, d.N, d.D3, d.H3, d.W3);
  LayerDgrad(net, 5, net->g3a, net->g3b, d.N, d.D3, d.H3, d.W3);
  Gate(net->g3b, net->a[4], d.P3 * 32);
  LayerWgrad(net, 4, net->q2, net->g3b, d.N, d.D3, d.H3, d.W3);
  LayerDgrad(net, 4, net->g3b, net->g3a, d.N, d.D3, d.H3, d.W3);  // 16 ch
  // level 2 encoder: E2 receives the pool gradient and the skip gradient
  hipLaunchKernelGGL(PoolBackSkipGateKernel, dim3(Blocks(d.P2 * 2)),
                     dim3(256), 0, Stream(), net->g3a, net->gc2, net->a[3],
                     net->g2a, 16, 32, d.D2, d.H2, d.W2, (int)(d.P2 * 2));
  Check();
  LayerWgrad(net, 3, net->a[2], net->g2a, d.N, d.D2, d.H2, d.W2);
  LayerDgrad(net, 3, net->g2a, net->g2b, d.N, d.D2, d.H2, d.W2);
  Gate(net->g2b, net->a[2], d.P2 * 16);
  LayerWgrad(net, 2, net->q1, net->g2b, d.N, d.D2, d.H2, d.W2);
  LayerDgrad(net, 2, net->g2b, net->g2a, d.N, d.D2, d.H2, d.W2);  // 8 ch
  // level 1 encoder
  hipLaunchKernelGGL(PoolBackSkipGateKernel, dim3(Blocks(d.P1)), dim3(256),
                     0, Stream(), net->g2a, net->gc1, net->a[1], net->g1a, 8,
                     16, d.D1, d.H1, d.W1, (int)d.P1);
  Check();
  LayerWgrad(net, 1, net->a[0], net->g1a, d.N, d.D1, d.H1, d.W1);
  LayerDgrad(net, 1, net->g1a, net->g1b, d.N, d.D1, d.H1, d.W1);
  Gate(net->g1b, net->a[0], d.P1 * 8);
  LayerWgrad(net, 0, net->x0, net->g1b, d.N, d.D1, d.H1, d.W1);
  PS_HIP_CHECK(hipMemcpyAsync(net->lhost, net->lpart,
                              sizeof(double) * kLossBlocks,
                              hipMemcpyDeviceToHost, Stream()));
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  double loss = 0.0;
  for (int b = 0; b < kLossBlocks; ++b) loss += net->lhost[b];
  return loss / (double)d.P1;
}

void FreeNet(__PSUNet *net) {
  (void)hipStreamSynchronize(Stream());
  std::vector<void *> ptrs = {net->params, net->grad, net->adam_m,
                              net->adam_v, net->x0, net->q1, net->q2,
                              net->c2, net->c1, net->out, net->dout, net->g1a,
                              net->g1b, net->gc1, net->g2a, net->g2b,
                              net->gc2, net->g3a, net->g3b, net->part,
                              net->bpart, net->tmp, net->lpart, net->e4,
                              net->e16};
  for (int i = 0; i < 10; ++i) ptrs.push_back(net->a[i]);
  for (auto &L : net->layers) {
    ptrs.push_back(L.wt);
    ptrs.push_back(L.wf);
  }
  for (void *p : ptrs)
    if (p) PS_HIP_CHECK(hipFree(p));
  if (net->lhost) PS_HIP_CHECK(hipHostFree(net->lhost));
  delete net;
}

// Host-side device buffer for the single-layer entry points.
struct DevBuf {
  void *p = nullptr;
  explicit DevBuf(size_t bytes) {
    p = Alloc(bytes);
    if (!p) PS_FATAL("out of device memory");
  }
  ~DevBuf() { (void)hipFree(p); }
};

void CheckLayerArgs(int n, int d, int h, int w, int cin, int cout) {
  if (n < 1 || d < 1 || h < 1 || w < 1 || cin % 8 || cout < 1)
    PS_FATAL("convolution: bad sizes n=%d d=%d h=%d w=%d cin=%d cout=%d", n,
             d, h, w, cin, cout);
}

}  // namespace

extern "C" {

__PSUNet *PSUNetNew(int nx, int ny, int nzs, int max_batch, float out_scale,
                    uint64_t seed) {
  if (nx % 16 || ny % 16 || nzs % 16 || nx < 16 || ny < 16 || nzs < 16 ||
      max_batch < 1 || !(out_scale > 0.0f)) {
    std::fprintf(stderr,
                 "[physis-hip] PSUNetNew: sizes must be positive multiples "
                 "of 16, batch >= 1, scale > 0\n");
    return nullptr;
  }
  __PSUNet *net = new __PSUNet();
  net->nx = nx;
  net->ny = ny;
  net->nzs = nzs;
  net->max_batch = max_batch;
  net->out_scale = out_scale;
  size_t off = 0;
  for (int l = 0; l < 10; ++l) {
    Layer L;
    L.cin = kLayerCh[l][0];
    L.cout = kLayerCh[l][1];
    L.w_off = off;
    off += (size_t)L.cout * 27 * L.cin;
    L.b_off = off;
    off += L.cout;
    net->layers.push_back(L);
  }
  net->head_off = off;
  off += 9;
  net->nparams = off;
  // He-normal weights (std sqrt(2 / fan_in)), zero biases; head std 1/sqrt(8)
  std::vector<float> h(off, 0.0f);
  uint64_t s = seed;
  auto normal = [&]() {
    const double u1 = ((SplitMix(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = (SplitMix(s) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (auto &L : net->layers) {
    const double sd = std::sqrt(2.0 / (27.0 * L.cin));
    for (size_t i = 0; i < (size_t)L.cout * 27 * L.cin; ++i)
      h[L.w_off + i] = (float)(sd * normal());
  }
  for (int c = 0; c < 8; ++c) h[net->head_off + c] = (float)(normal() / std::sqrt(8.0));
  const Dims d = DimsOf(net, max_batch);
  bool ok = true;
  auto A16 = [&](long elems) {
    u16 *p = (u16 *)Alloc(sizeof(u16) * (size_t)elems);
    ok = ok && p;
    return p;
  };
  auto A32 = [&](long elems) {
    float *p = (float *)Alloc(sizeof(float) * (size_t)elems);
    ok = ok && p;
    return p;
  };
  net->params = A32(off);
  net->grad = A32(off);
  net->adam_m = A32(off);
  net->adam_v = A32(off);
  for (auto &L : net->layers) {
    L.wt = A16((long)L.cout * KPad(L.cin));
    L.wf = A16((long)L.cin * KPad(L.cout));
  }
  net->x0 = A16(d.P1 * 8);
  const long act[10] = {d.P1 * 8,  d.P1 * 8,  d.P2 * 16, d.P2 * 16, d.P3 * 32,
                        d.P3 * 32, d.P2 * 16, d.P2 * 16, d.P1 * 8,  d.P1 * 8};
  for (int l = 0; l < 10; ++l) net->a[l] = A16(act[l]);
  net->q1 = A16(d.P2 * 8);
  net->q2 = A16(d.P3 * 16);
  net->c2 = A16(d.P2 * 48);
  net->c1 = A16(d.P1 * 24);
  net->out = A32(d.P1);
  net->dout = A32(d.P1);
  net->g1a = A16(d.P1 * 8);
  net->g1b = A16(d.P1 * 8);
  net->gc1 = A16(d.P1 * 24);
  net->g2a = A16(d.P2 * 16);
  net->g2b = A16(d.P2 * 16);
  net->gc2 = A16(d.P2 * 48);
  net->g3a = A16(d.P3 * 32);
  net->g3b = A16(d.P3 * 32);
  size_t pe = (size_t)kHeadBlocks * 9, be = 0, te = 9;
  for (int l = 0; l < 10; ++l) {
    const int lv = l < 2 || l > 7 ? 1 : (l < 4 || l > 5 ? 2 : 3);
    const long P = lv == 1 ? d.P1 : (lv == 2 ? d.P2 : d.P3);
    const int W = lv == 1 ? d.W1 : (lv == 2 ? d.W2 : d.W3);
    const int H = lv == 1 ? d.H1 : (lv == 2 ? d.H2 : d.H3);
    const int D = lv == 1 ? d.D1 : (lv == 2 ? d.D2 : d.D3);
    (void)P;
    // the block count depends on the batch: the largest over every batch a
    // call may use
    size_t nb = 0;
    for (int nbat = 1; nbat <= max_batch; ++nbat)
      nb = std::max<size_t>(
          nb, (size_t)WgradBlocks(net->layers[l].cin, nbat, D, H, W));
    const size_t nw = (size_t)net->layers[l].cout * 27 * net->layers[l].cin;
    pe = std::max(pe, nb * nw);
    be = std::max(be, nb * net->layers[l].cout);
    te = std::max(te, nw);
  }
  net->part_elems = pe;
  net->bpart_elems = be;
  net->part = A32((long)pe);
  net->bpart = A32((long)be);
  net->tmp = A32((long)(te * kSumSegments));
  net->e4 = A32(d.P2);
  net->e16 = A32(d.P3);
  net->lpart = (double *)Alloc(sizeof(double) * kLossBlocks);
  ok = ok && net->lpart;
  if (ok && hipHostMalloc((void **)&net->lhost, sizeof(double) * kLossBlocks) !=
                hipSuccess) {
    (void)hipGetLastError();
    ok = false;
  }
  if (!ok) {
    std::fprintf(stderr, "[physis-hip] PSUNetNew: out of device memory\n");
    FreeNet(net);
    return nullptr;
  }
  PS_HIP_CHECK(hipMemcpy(net->params, h.data(), sizeof(float) * off,
                         hipMemcpyHostToDevice));
  PS_HIP_CHECK(hipMemset(net->adam_m, 0, sizeof(float) * off));
  PS_HIP_CHECK(hipMemset(net->adam_v, 0, sizeof(float) * off));
  PS_HIP_CHECK(hipMemset(net->grad, 0, sizeof(float) * off));
  Upload(net);
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  return net;
}

void PSUNetFree(__PSUNet *net) {
  if (net) FreeNet(net);
}

size_t PSUNetNumParams(const __PSUNet *net) { return net->nparams; }

void PSUNetGetParams(const __PSUNet *net, float *host) {
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  PS_HIP_CHECK(hipMemcpy(host, net->params, sizeof(float) * net->nparams,
                         hipMemcpyDeviceToHost));
}

void PSUNetSetParams(__PSUNet *net, const float *host) {
  PS_HIP_CHECK(hipMemcpy(net->params, host, sizeof(float) * net->nparams,
                         hipMemcpyHostToDevice));
  Upload(net);
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
}

void PSUNetPredict(__PSUNet *net, const __PSGrid *rhs, const __PSGrid *bnd,
                   int first, int count, __PSGrid *p) {
  CheckGrid(net, p, first, count, "p");
  if (p->type != PS_FLOAT && p->type != PS_DOUBLE)
    PS_FATAL("U-Net: p must be a float or double grid");
  Forward(net, rhs, bnd, first, count);
  const Dims d = DimsOf(net, count);
  const long off = (long)first * net->nx * net->ny * net->nzs;
  if (p->type == PS_DOUBLE)
    hipLaunchKernelGGL(PredictKernel<double>, dim3(Blocks(d.P1)), dim3(256), 0,
                       Stream(), net->out, (const float *)bnd->p,
                       (double *)p->p, off, net->out_scale, d.P1);
  else
    hipLaunchKernelGGL(PredictKernel<float>, dim3(Blocks(d.P1)), dim3(256), 0,
                       Stream(), net->out, (const float *)bnd->p,
                       (float *)p->p, off, net->out_scale, d.P1);
  Check();
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
}

void PSUNetReflect(const __PSGrid *src, int first, int count, __PSGrid *dst,
                   int dst_first, int nx, int ny, int nzs, int flips) {
  if (!src || !dst || src->type != PS_FLOAT || dst->type != PS_FLOAT ||
      src->num_dims != 3 || dst->num_dims != 3 || src->dim[0] != nx ||
      dst->dim[0] != nx || src->dim[1] != ny || dst->dim[1] != ny ||
      nzs < 1 || src->dim[2] % nzs || dst->dim[2] % nzs || first < 0 ||
      dst_first < 0 || count < 0 || flips < 0 || flips > 7 ||
      (long)(first + count) * nzs > src->dim[2] ||
      (long)(dst_first + count) * nzs > dst->dim[2] ||
      (src->p == dst->p && count > 0))
    PS_FATAL("PSUNetReflect: bad grids or problem range");
  const long m = (long)nx * ny * nzs, n = m * count;
  if (n == 0) return;
  hipLaunchKernelGGL(ReflectKernel, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, Stream(), (const float *)src->p,
                     (float *)dst->p, nx, ny, nzs, m * first, m * dst_first,
                     flips, n);
  Check();
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
}

double PSUNetLossGrad(__PSUNet *net, const __PSGrid *rhs, const __PSGrid *bnd,
                      const __PSGrid *target, int first, int count,
                      float *grad_host) {
  Forward(net, rhs, bnd, first, count);
  const double loss = Backward(net, target, first, count);
  if (grad_host)
    PS_HIP_CHECK(hipMemcpy(grad_host, net->grad, sizeof(float) * net->nparams,
                           hipMemcpyDeviceToHost));
  return loss;
}

void PSUNetSetLossWeights(__PSUNet *net, float w4, float w16) {
  if (!(w4 >= 0.0f) || !(w16 >= 0.0f))
    PS_FATAL("U-Net: loss weights must be >= 0 (got %g, %g)", w4, w16);
  net->w4 = w4;
  net->w16 = w16;
}

double PSUNetTrainStep(__PSUNet *net, const __PSGrid *rhs,
                       const __PSGrid *bnd, const __PSGrid *target, int first,
                       int count, float lr) {
  Forward(net, rhs, bnd, first, count);
  const double loss = Backward(net, target, first, count);
  net->step += 1;
  const float c1 = 1.0f - std::pow(0.9f, (float)net->step);
  const float c2 = 1.0f - std::pow(0.999f, (float)net->step);
  hipLaunchKernelGGL(AdamKernel, dim3(Blocks((long)net->nparams)), dim3(256),
                     0, Stream(), net->params, net->grad, net->adam_m,
                     net->adam_v, (long)net->nparams, lr, c1, c2);
  Check();
  Upload(net);
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  return loss;
}

void PSConv3dForwardHost(const uint16_t *x, const uint16_t *w,
                         const float *bias, int relu, int n, int d, int h,
                         int wd, int cin, int cout, uint16_t *y) {
  CheckLayerArgs(n, d, h, wd, cin, cout);
  const long P = (long)n * d * h * wd;
  const size_t nw = (size_t)cout * 27 * cin;
  DevBuf dx(P * cin * 2), dw32(nw * 4), dwt((size_t)cout * KPad(cin) * 2),
      dwf((size_t)cin * KPad(cout) * 2), db(cout * 4), dy(P * cout * 2);
  std::vector<float> wf(nw);
  for (size_t i = 0; i < nw; ++i) {
    const uint32_t u = (uint32_t)w[i] << 16;
    std::memcpy(&wf[i], &u, 4);
  }
  PS_HIP_CHECK(hipMemcpy(dx.p, x, P * cin * 2, hipMemcpyHostToDevice));
  PS_HIP_CHECK(hipMemcpy(dw32.p, wf.data(), nw * 4, hipMemcpyHostToDevice));
  PS_HIP_CHECK(hipMemcpy(db.p, bias, cout * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(PackKernel, dim3(64), dim3(256), 0, Stream(),
                     (const float *)dw32.p, (u16 *)dwt.p, (u16 *)dwf.p, cin,
                     cout);
  Check();
  Conv(cin, cout, (const u16 *)dx.p, (const u16 *)dwt.p, (const float *)db.p,
       relu != 0, (u16 *)dy.p, n, d, h, wd);
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  PS_HIP_CHECK(hipMemcpy(y, dy.p, P * cout * 2, hipMemcpyDeviceToHost));
}

void PSConv3dBackwardDataHost(const uint16_t *dy, const uint16_t *w, int n,
                              int d, int h, int wd, int cin, int cout,
                              uint16_t *dx) {
  CheckLayerArgs(n, d, h, wd, cin, cout);
  if (cout % 8) PS_FATAL("convolution: cout must be a multiple of 8");
  const long P = (long)n * d * h * wd;
  const size_t nw = (size_t)cout * 27 * cin;
  DevBuf ddy(P * cout * 2), dw32(nw * 4), dwt((size_t)cout * KPad(cin) * 2),
      dwf((size_t)cin * KPad(cout) * 2), ddx(P * cin * 2);
  std::vector<float> wf(nw);
  for (size_t i = 0; i < nw; ++i) {
    const uint32_t u = (uint32_t)w[i] << 16;
    std::memcpy(&wf[i], &u, 4);
  }
  PS_HIP_CHECK(hipMemcpy(ddy.p, dy, P * cout * 2, hipMemcpyHostToDevice));
  PS_HIP_CHECK(hipMemcpy(dw32.p, wf.data(), nw * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(PackKernel, dim3(64), dim3(256), 0, Stream(),
                     (const float *)dw32.p, (u16 *)dwt.p, (u16 *)dwf.p, cin,
                     cout);
  Check();
  Conv(cout, cin, (const u16 *)ddy.p, (const u16 *)dwf.p, nullptr, false,
       (u16 *)ddx.p, n, d, h, wd);
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  PS_HIP_CHECK(hipMemcpy(dx, ddx.p, P * cin * 2, hipMemcpyDeviceToHost));
}

void PSConv3dBackwardFilterHost(const uint16_t *x, const uint16_t *dy, int n,
                                int d, int h, int wd, int cin, int cout,
                                float *dw, float *db) {
  CheckLayerArgs(n, d, h, wd, cin, cout);
  const long P = (long)n * d * h * wd;
  const long nb = WgradBlocks(cin, n, d, h, wd);
  const size_t nw = (size_t)cout * 27 * cin;
  DevBuf ddx(P * cin * 2), ddy(P * cout * 2), part(nb * nw * 4),
      bpart(nb * cout * 4), dw_d(nw * 4), db_d(cout * 4),
      tmp(nw * kSumSegments * 4);
  PS_HIP_CHECK(hipMemcpy(ddx.p, x, P * cin * 2, hipMemcpyHostToDevice));
  PS_HIP_CHECK(hipMemcpy(ddy.p, dy, P * cout * 2, hipMemcpyHostToDevice));
  int np = 0;
  Wgrad(cin, cout, (const u16 *)ddx.p, (const u16 *)ddy.p, (float *)part.p,
        (float *)bpart.p, n, d, h, wd, nb, &np);
  SumParts((const float *)part.p, np, (long)nw, (float *)tmp.p,
           (float *)dw_d.p);
  SumParts((const float *)bpart.p, np, (long)cout, (float *)tmp.p,
           (float *)db_d.p);
  PS_HIP_CHECK(hipStreamSynchronize(Stream()));
  PS_HIP_CHECK(hipMemcpy(dw, dw_d.p, nw * 4, hipMemcpyDeviceToHost));
  PS_HIP_CHECK(hipMemcpy(db, db_d.p, cout * 4, hipMemcpyDeviceToHost));
}

}  // extern "C"
