// Device-side building blocks shared by libphysis_rt_hip and the run kernels
// that physisc-hip generates: reduction operators, wave64/block reductions,
// vectorised grid loads with the REF target's out-of-range semantics, and the
// XCD-aware block numbering used by every 2.5D z-marching kernel.
//
// Written for gfx950 (CDNA4): 64-lane wavefronts, 8 XCDs that each own an L2,
// workgroups dispatched to XCDs round-robin by linear block id.
#ifndef PHYSIS_HIP_DEVICE_H_
#define PHYSIS_HIP_DEVICE_H_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "physis/reduce.h"

namespace physis_hip {

constexpr int kWave = 64;
constexpr int kNumXCD = 8;

// ---------------------------------------------------------------- operators
// Identity values are the mathematical ones (-inf for MAX on floats).  The
// reference CUDA runtime seeds MAX with FLT_MIN (runtime/reduce.h:61-63),
// which is wrong for all-negative data; the REF target seeds with element 0
// (runtime/libphysis_rt_ref.cc:24-28), which these identities reproduce.
template <typename T> struct Lim;
template <> struct Lim<float> {
  __device__ static float lowest() { return -__builtin_huge_valf(); }
  __device__ static float highest() { return __builtin_huge_valf(); }
};
template <> struct Lim<double> {
  __device__ static double lowest() { return -__builtin_huge_val(); }
  __device__ static double highest() { return __builtin_huge_val(); }
};
template <> struct Lim<int> {
  __device__ static int lowest() { return INT32_MIN; }
  __device__ static int highest() { return INT32_MAX; }
};
template <> struct Lim<long> {
  __device__ static long lowest() { return INT64_MIN; }
  __device__ static long highest() { return INT64_MAX; }
};

// Integer SUM/PROD wrap in two's complement like the sequential REF loop
// (an oracle output of the 27-point test overflows int32); done in unsigned
// arithmetic so the device code has no signed-overflow UB.
template <typename T> struct Wrap { typedef T U; };
template <> struct Wrap<int> { typedef unsigned int U; };
template <> struct Wrap<long> { typedef unsigned long U; };

template <typename T, int OP> struct ReduceOp;
template <typename T> struct ReduceOp<T, PS_SUM> {
  __device__ static T identity() { return T(0); }
  __device__ static T apply(T a, T b) {
    typedef typename Wrap<T>::U U;
    return (T)((U)a + (U)b);
  }
};
template <typename T> struct ReduceOp<T, PS_PROD> {
  __device__ static T identity() { return T(1); }
  __device__ static T apply(T a, T b) {
    typedef typename Wrap<T>::U U;
    return (T)((U)a * (U)b);
  }
};
template <typename T> struct ReduceOp<T, PS_MAX> {
  __device__ static T identity() { return Lim<T>::lowest(); }
  __device__ static T apply(T a, T b) { return (a > b) ? a : b; }
};
template <typename T> struct ReduceOp<T, PS_MIN> {
  __device__ static T identity() { return Lim<T>::highest(); }
  __device__ static T apply(T a, T b) { return (a < b) ? a : b; }
};

// ---------------------------------------------------------------- reductions
// Butterfly over the 64 lanes of a wavefront; every lane ends with the
// result.  The combination order is fixed, so results are deterministic.
template <typename T, int OP>
__device__ __forceinline__ T WaveReduce(T v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1)
    v = ReduceOp<T, OP>::apply(v, __shfl_xor(v, m, kWave));
  return v;
}

// Block-wide reduction through LDS; `lds` holds one T per wavefront.  The
// value is valid in thread 0.
template <typename T, int OP, int NTHREADS>
__device__ __forceinline__ T BlockReduce(T v, T *lds) {
  constexpr int kWaves = NTHREADS / kWave;
  v = WaveReduce<T, OP>(v);
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  if (kWaves == 1) return v;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T r = ReduceOp<T, OP>::identity();
  if (wid == 0) {
    r = lane < kWaves ? lds[lane] : r;
    r = WaveReduce<T, OP>(r);
  }
  return r;
}

// ------------------------------------------------------- XCD-aware numbering
// Hardware dispatch places block b on XCD (b % 8).  Renumbering so that each
// XCD receives a contiguous range of logical tiles keeps a tile's y- and
// z-neighbours (whose planes it re-reads as halo) in the same L2.
// With nblocks = 8q + r, XCD j receives q + (j < r) blocks; its range of
// logical ids starts at j*q + min(j, r).  A bijection for every nblocks.
__device__ __forceinline__ int XcdLogicalBlock(int b, int nblocks) {
  const int q = nblocks / kNumXCD, r = nblocks % kNumXCD;
  const int xcd = b % kNumXCD;
  return xcd * q + (xcd < r ? xcd : r) + b / kNumXCD;
}

// ----------------------------------------------------------- grid sampling
// Wrap used by PSGridGetPeriodic; equals the reference's (i + n) % n
// (include/physis/physis_ref.h:63-74) for every i >= -n and stays in range
// for any i.
__device__ __forceinline__ int WrapIndex(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// WrapIndex for indices near the grid ([-n, 2n) without a division; others
// fall back to it): plane and row indices of periodic stencils.
__device__ __forceinline__ int WrapNear(int i, int n) {
  if (i >= 0 && i < n) return i;
  if (i < 0 && i >= -n) return i + n;
  if (i >= n && i - n < n) return i - n;
  return WrapIndex(i, n);
}

}  // namespace physis_hip

#endif  // PHYSIS_HIP_DEVICE_H_
