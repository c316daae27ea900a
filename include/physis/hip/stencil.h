// Building blocks of the 2.5D z-marching run kernels that physisc-hip emits
// for 3-D stencil maps on gfx950.
//
// Shape of a run kernel (replaces the CUDA target's one-thread-per-column
// kernel, reference translator/cuda_runtime_builder.cc:1099-1228):
//   * a block of BX x BY threads owns a TX x TY tile of the xy plane,
//     TX = BX * VX: each thread owns VX consecutive x points of one row, so
//     grid rows move as 16-byte vector loads when VX = 4;
//   * the block marches z through a chunk [zb, ze) of the domain;
//   * grids read at xy offsets ("staged" grids) live in an LDS ring of
//     planes with a halo of HX columns and HY rows; the ring holds the planes
//     z+ZL .. z+ZR the stencil reads plus the plane being filled, so there is
//     one barrier per z step;
//   * the next plane of every staged grid, and the next values of grids read
//     only at the centre point ("direct" grids), are loaded into registers at
//     the top of a step and consumed one step later, so HBM latency overlaps
//     the arithmetic;
//   * PSGridEmit stores straight to global memory.
//
// Out-of-range reads keep the REF target's meaning (reference
// include/physis/physis_ref.h:57-74): a non-periodic get reads the element
// at the linear offset x + y*nx + z*nx*ny whenever that offset lies inside
// the grid, and yields 0 otherwise (the REF target would read outside the
// allocation there); a periodic get wraps each axis.
#ifndef PHYSIS_HIP_STENCIL_H_
#define PHYSIS_HIP_STENCIL_H_

#include "physis/hip/device.h"

namespace physis_hip {

// ------------------------------------------------------ vector load / store
// A compile-time flag passed to generic lambdas (the interior and general
// forms of a wave step).
template <bool B> struct Flag { static constexpr bool value = B; };

template <typename T, int N> struct NativeVec { typedef T V; };
template <> struct NativeVec<float, 2> {
  typedef float V __attribute__((ext_vector_type(2)));
};
template <> struct NativeVec<float, 4> {
  typedef float V __attribute__((ext_vector_type(4)));
};
template <> struct NativeVec<int, 2> {
  typedef int V __attribute__((ext_vector_type(2)));
};
template <> struct NativeVec<int, 4> {
  typedef int V __attribute__((ext_vector_type(4)));
};
template <> struct NativeVec<double, 2> {
  typedef double V __attribute__((ext_vector_type(2)));
};
template <> struct NativeVec<long, 2> {
  typedef long V __attribute__((ext_vector_type(2)));
};

// ------------------------------------------------ 16-bit storage formats
// Grids of the precision tier hold bfloat16 or IEEE binary16 values as raw
// bit patterns; arithmetic converts them to float.  float -> 16-bit rounds to
// nearest even.
struct bf16_t { unsigned short bits; };
struct fp16_t { unsigned short bits; };

template <typename To, typename From>
__device__ __forceinline__ To Convert(From v) { return (To)v; }
template <> __device__ __forceinline__ float Convert<float, bf16_t>(bf16_t v) {
  return __uint_as_float((unsigned)v.bits << 16);
}
template <> __device__ __forceinline__ double Convert<double, bf16_t>(bf16_t v) {
  return (double)__uint_as_float((unsigned)v.bits << 16);
}
template <> __device__ __forceinline__ float Convert<float, fp16_t>(fp16_t v) {
  return (float)__builtin_bit_cast(_Float16, v.bits);
}
template <> __device__ __forceinline__ double Convert<double, fp16_t>(fp16_t v) {
  return (double)(float)__builtin_bit_cast(_Float16, v.bits);
}
template <> __device__ __forceinline__ bf16_t Convert<bf16_t, float>(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return bf16_t{(unsigned short)0x7fc0};
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf16_t{(unsigned short)(u >> 16)};
}
template <> __device__ __forceinline__ fp16_t Convert<fp16_t, float>(float f) {
  return fp16_t{__builtin_bit_cast(unsigned short, (_Float16)f)};
}

// True when N consecutive T can move as one native vector.
template <typename T, int N> struct CanVec {
  static constexpr bool value =
      N > 1 && (sizeof(T) * N == 8 || sizeof(T) * N == 16);
};

// ------------------------------------------------------- buffer windows
// Grid memory is accessed through raw buffer instructions whose resource
// descriptor covers a window of the grid.  The hardware range check turns
// every access outside the window into "load 0 / drop the store" without a
// branch, which (a) implements the REF rule that out-of-grid non-periodic
// reads yield 0 here, and (b) keeps every load unconditional, so the
// compiler can wait for exactly the plane it needs (s_waitcnt vmcnt(N))
// instead of draining all in-flight loads.
// Offsets passed to the accessors are global linear element offsets; the
// window start is subtracted here.  Windows are < 2 GiB (the host planners
// cap z chunks so that every window fits: run.h WindowChunkLimit), so a
// negative offset and the kDrop sentinel are both out of range.
constexpr int kBufDword3 = 0x00020000;  // raw 32-bit dwords, gfx9 family
constexpr int kDrop = INT32_MIN;
// Non-temporal (streaming) cache policy bit of buffer stores on gfx950.
constexpr int kStreamStore = 2;
// The same bit on buffer loads: data read once (centre-only coefficient
// grids) is marked for early eviction, which leaves the L2 to the staged
// planes whose halo rows neighbouring blocks read again.
constexpr int kStreamLoad = 2;

template <typename T>
struct BufWin {
  __amdgpu_buffer_rsrc_t rsrc;
  int start;  // element offset of the window start

  // Window over elements [lo, hi) of a grid of n elements (clamped).
  __device__ __forceinline__ static BufWin Make(const T *p, int n, long lo,
                                                long hi) {
    lo = lo < 0 ? 0 : (lo > n ? n : lo);
    hi = hi < lo ? lo : (hi > n ? n : hi);
    BufWin w;
    w.start = (int)lo;
    w.rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T *>(p) + lo, 0, (int)((hi - lo) * (long)sizeof(T)),
        kBufDword3);
    return w;
  }

  __device__ __forceinline__ int Bytes(int off) const {
    return off == kDrop ? kDrop : (off - start) * (int)sizeof(T);
  }

  // AUX is the cache-policy operand of the buffer load (kStreamLoad or 0).
  template <int N, int AUX = 0>
  __device__ __forceinline__ void Load(int off, T (&out)[N]) const {
    constexpr int kBytes = (int)sizeof(T) * N;
    const int b = Bytes(off);
    if constexpr (kBytes == 32) {
      auto v0 = __builtin_amdgcn_raw_buffer_load_b128(rsrc, b, 0, AUX);
      auto v1 = __builtin_amdgcn_raw_buffer_load_b128(
          rsrc, b == kDrop ? kDrop : b + 16, 0, AUX);
      __builtin_memcpy(out, &v0, 16);
      __builtin_memcpy((char *)out + 16, &v1, 16);
    } else if constexpr (kBytes == 16) {
      auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, b, 0, AUX);
      __builtin_memcpy(out, &v, 16);
    } else if constexpr (kBytes == 8) {
      auto v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, b, 0, AUX);
      __builtin_memcpy(out, &v, 8);
    } else if constexpr (kBytes == 4) {
      auto v = __builtin_amdgcn_raw_buffer_load_b32(rsrc, b, 0, AUX);
      __builtin_memcpy(out, &v, 4);
    } else if constexpr (kBytes == 2) {
      auto v = __builtin_amdgcn_raw_buffer_load_b16(rsrc, b, 0, AUX);
      __builtin_memcpy(out, &v, 2);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T one[1];
        Load<1, AUX>(off == kDrop ? kDrop : off + i, one);
        out[i] = one[0];
      }
    }
  }

  // AUX is the cache-policy operand of the buffer store: kStreamStore marks
  // output that is not read again before it leaves the caches.
  template <int N, int AUX = 0>
  __device__ __forceinline__ void Store(int off, const T (&v)[N]) const {
    constexpr int kBytes = (int)sizeof(T) * N;
    const int b = Bytes(off);
    if constexpr (kBytes == 32) {
      __attribute__((ext_vector_type(4))) unsigned w0, w1;
      __builtin_memcpy(&w0, v, 16);
      __builtin_memcpy(&w1, (const char *)v + 16, 16);
      __builtin_amdgcn_raw_buffer_store_b128(w0, rsrc, b, 0, AUX);
      __builtin_amdgcn_raw_buffer_store_b128(
          w1, rsrc, b == kDrop ? kDrop : b + 16, 0, AUX);
    } else if constexpr (kBytes == 16) {
      __attribute__((ext_vector_type(4))) unsigned w;
      __builtin_memcpy(&w, v, 16);
      __builtin_amdgcn_raw_buffer_store_b128(w, rsrc, b, 0, AUX);
    } else if constexpr (kBytes == 8) {
      __attribute__((ext_vector_type(2))) unsigned w;
      __builtin_memcpy(&w, v, 8);
      __builtin_amdgcn_raw_buffer_store_b64(w, rsrc, b, 0, AUX);
    } else if constexpr (kBytes == 4) {
      unsigned w;
      __builtin_memcpy(&w, v, 4);
      __builtin_amdgcn_raw_buffer_store_b32(w, rsrc, b, 0, AUX);
    } else if constexpr (kBytes == 2) {
      unsigned short w;
      __builtin_memcpy(&w, v, 2);
      __builtin_amdgcn_raw_buffer_store_b16(w, rsrc, b, 0, AUX);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const T one[1] = {v[i]};
        Store<1, AUX>(off == kDrop ? kDrop : off + i, one);
      }
    }
  }
};

// Stores the VX values of one thread; elements whose mask bit is clear are
// dropped by the range check.  A full mask is one vector store.
template <typename T, int VX, int AUX = kStreamStore>
__device__ __forceinline__ void StoreMasked(const BufWin<T> &w, int off,
                                            const T (&v)[VX], unsigned mask) {
  constexpr unsigned kAll = (1u << VX) - 1u;
  if (mask == kAll) {
    w.template Store<VX, AUX>(off, v);
  } else {
#pragma unroll
    for (int i = 0; i < VX; ++i) {
      const T one[1] = {v[i]};
      w.template Store<1, AUX>((mask >> i) & 1u ? off + i : kDrop, one);
    }
  }
}

// ---------------------------------------------------------- 3-D tile index
// Block -> (tile x, tile y, z chunk) after XCD-aware renumbering; x fastest
// so that y-neighbour tiles, which share halo rows, sit on one XCD and run
// together.
struct TileCoord {
  int bx, by, bz;
};
__device__ __forceinline__ TileCoord TileOf(int tiles_x, int tiles_y) {
  int b = XcdLogicalBlock(blockIdx.x, gridDim.x);
  TileCoord t;
  t.bx = b % tiles_x;
  b /= tiles_x;
  t.by = b % tiles_y;
  t.bz = b / tiles_y;
  return t;
}

// ------------------------------------------------------------- staged grid
// One staged grid of element type T.  A block is BX = 64 lanes (one wave)
// per row and BY waves; wave r owns tile row y0 - HY + r, so the first and
// last HY waves load halo rows only and the TY = BY - 2*HY waves between
// them compute.  Each wave also fetches the HX columns left and right of
// its row segment with one dword load in which lanes 0 .. 2*HX-1 carry the
// halo and the other lanes repeat lane 0's address.
//   HX, HY  halo width in x and y (maximum |offset| the stencil uses)
//   NSLOT   planes in the LDS ring (ZR - ZL + 2)
//   M       element type in global memory; values are converted to T as
//           they enter LDS (T = M for Physis grids)
template <typename T, int VX, int BY, int HX, int HY, int NSLOT,
          bool PERIODIC, typename M = T>
struct Stage {
  static constexpr int BX = kWave;
  static constexpr int kTX = BX * VX;
  static constexpr int kTY = BY - 2 * HY;
  static_assert(kTY > 0, "block too short for its y halo");
  static_assert(2 * HX <= BX, "x halo wider than a wavefront");
  // Left pad rounded up to VX keeps interior rows vector-aligned in LDS.
  static constexpr int kPad = HX == 0 ? 0 : ((HX + VX - 1) / VX) * VX;
  static constexpr int kRow = kPad + kTX + kPad;
  static constexpr int kPlane = kRow * BY;
  static constexpr int kSlots = NSLOT;
  static constexpr int kLdsElems = kPlane * NSLOT;

  M own[VX];
  M halo;

  // Linear offset of (x, y, z): wrapped per axis when periodic.
  __device__ __forceinline__ static int Offset(const int *dim, int x, int y,
                                               int z) {
    if (PERIODIC) {
      x = WrapIndex(x, dim[0]);
      y = WrapIndex(y, dim[1]);
      z = WrapIndex(z, dim[2]);
    }
    return x + y * dim[0] + z * dim[0] * dim[1];
  }

  // The part of a row's addresses that does not change from plane to
  // plane: the x offsets of the lane's VX points and of its halo point
  // (lanes 0 .. 2*HX-1 carry the halo, the others repeat lane 0's) and the
  // row's y offset, wrapped when periodic.  Computed once per row, so that
  // a plane step only adds the plane offset.
  struct Addr {
    int xo, ho, yo;
  };
  __device__ __forceinline__ static Addr Prepare(const int *dim, int x0, int y,
                                                 int lane) {
    const int hl = lane < 2 * HX ? lane : 0;
    const int hx = hl < HX ? x0 - HX + hl : x0 + kTX + (hl - HX);
    Addr a;
    if (PERIODIC) {
      a.xo = WrapIndex(x0 + lane * VX, dim[0]);
      a.ho = WrapIndex(hx, dim[0]);
      a.yo = WrapIndex(y, dim[1]) * dim[0];
    } else {
      a.xo = x0 + lane * VX;
      a.ho = hx;
      a.yo = y * dim[0];
    }
    return a;
  }

  // Issues the two loads of plane z for the wave's row y = y0 - HY + wave.
  // Non-periodic reads go through the block's window w (the REF linear rule
  // may spill into the neighbouring plane); periodic reads go through a
  // window of exactly the wrapped plane, which may lie anywhere in the grid.
  __device__ __forceinline__ void Load(const BufWin<M> &w, const M *p,
                                       const int *dim, const Addr &a, int z) {
    M v[1];
    if (PERIODIC) {
      const int nxy = dim[0] * dim[1];
      const int zw = WrapNear(z, dim[2]);
      const BufWin<M> pw =
          BufWin<M>::Make(p, nxy * dim[2], (long)zw * nxy, (long)(zw + 1) * nxy);
      const int zo = zw * nxy;
      pw.template Load<VX>(a.xo + a.yo + zo, own);
      pw.template Load<1>(a.ho + a.yo + zo, v);
    } else {
      const int zo = z * dim[0] * dim[1];
      w.template Load<VX>(a.xo + a.yo + zo, own);
      w.template Load<1>(a.ho + a.yo + zo, v);
    }
    halo = v[0];
  }
  __device__ __forceinline__ void Load(const BufWin<M> &w, const M *p,
                                       const int *dim, int x0, int y, int z,
                                       int lane) {
    Load(w, p, dim, Prepare(dim, x0, y, lane), z);
  }

  // Writes the wave's row of the plane into one LDS slot.
  __device__ __forceinline__ void Store(T *slot, int wave, int lane) const {
    T *row = slot + wave * kRow;
    if constexpr (__is_same(T, M)) {
      StoreVec(row + kPad + lane * VX, own);
    } else {
      T cv[VX];
#pragma unroll
      for (int i = 0; i < VX; ++i) cv[i] = Convert<T, M>(own[i]);
      StoreVec(row + kPad + lane * VX, cv);
    }
    if (lane < 2 * HX)
      row[lane < HX ? kPad - HX + lane : kPad + kTX + (lane - HX)] =
          Convert<T, M>(halo);
  }

  // Row segment of `slot` at row `row` covering x offsets XL .. VX-1+XR
  // around the thread's VX points: seg[i - XL] is the value at offset i.
  template <int XL, int XR>
  __device__ __forceinline__ static void ReadRow(const T *slot, int row,
                                                 int lane,
                                                 T (&seg)[VX + XR - XL]) {
    const T *p = slot + row * kRow + kPad + lane * VX;
    T mid[VX];
    LoadVec(p, mid);
#pragma unroll
    for (int i = XL; i < VX + XR; ++i)
      seg[i - XL] = (i >= 0 && i < VX) ? mid[i] : p[i];
  }
  // Writes the lane's VX values of row `row` of a staged plane.
  __device__ __forceinline__ static void WriteRow(T *slot, int row, int lane,
                                                  const T (&v)[VX]) {
    StoreVec(slot + row * kRow + kPad + lane * VX, v);
  }

 private:
  __device__ __forceinline__ static void StoreVec(T *dst, const T (&v)[VX]) {
    if constexpr (CanVec<T, VX>::value) {
      typedef typename NativeVec<T, VX>::V V;
      V w;
#pragma unroll
      for (int i = 0; i < VX; ++i) w[i] = v[i];
      *reinterpret_cast<V *>(dst) = w;
    } else {
#pragma unroll
      for (int i = 0; i < VX; ++i) dst[i] = v[i];
    }
  }
  __device__ __forceinline__ static void LoadVec(const T *src, T (&v)[VX]) {
    if constexpr (CanVec<T, VX>::value) {
      typedef typename NativeVec<T, VX>::V V;
      V w = *reinterpret_cast<const V *>(src);
#pragma unroll
      for (int i = 0; i < VX; ++i) v[i] = w[i];
    } else {
#pragma unroll
      for (int i = 0; i < VX; ++i) v[i] = src[i];
    }
  }
};

// ----------------------------------------------------- irregular accesses
// PSGridGet / PSGridGetPeriodic at indices the stencil analysis could not
// reduce to constant offsets: a direct global read with the same rules as
// the staged path (linear offset inside the grid, else 0; periodic wraps).
template <typename T>
__device__ __forceinline__ T GetLinear(const T *p, const int *dim, int nd,
                                       int i0, int i1 = 0, int i2 = 0) {
  long off = i0;
  long n = dim[0];
  if (nd > 1) { off += (long)i1 * dim[0]; n *= dim[1]; }
  if (nd > 2) { off += (long)i2 * dim[0] * dim[1]; n *= dim[2]; }
  return (off >= 0 && off < n) ? p[off] : T(0);
}
// Offsets of gets on grids of user-defined points (one plane per scalar
// component): the linear offset of the index, or -1 where GetLinear would
// read 0; the wrapped offset of a periodic get.
__device__ __forceinline__ long OffsetLinear(const int *dim, int nd, int i0,
                                             int i1 = 0, int i2 = 0) {
  long off = i0;
  long n = dim[0];
  if (nd > 1) { off += (long)i1 * dim[0]; n *= dim[1]; }
  if (nd > 2) { off += (long)i2 * dim[0] * dim[1]; n *= dim[2]; }
  return (off >= 0 && off < n) ? off : -1;
}
__device__ __forceinline__ long OffsetPeriodic(const int *dim, int nd, int i0,
                                               int i1 = 0, int i2 = 0) {
  long off = WrapIndex(i0, dim[0]);
  if (nd > 1) off += (long)WrapIndex(i1, dim[1]) * dim[0];
  if (nd > 2) off += (long)WrapIndex(i2, dim[2]) * dim[0] * dim[1];
  return off;
}
// Component `comp` (of ncomp) of a member at point offset `off` (n points
// per plane): 0 outside; stores outside are dropped.
template <typename T>
__device__ __forceinline__ T UdtLoad(const T *p, long n, int ncomp, long comp,
                                     long off) {
  return (off >= 0 && comp >= 0 && comp < ncomp) ? p[comp * n + off] : T(0);
}
template <typename T>
__device__ __forceinline__ void UdtStore(T *p, long n, int ncomp, long comp,
                                         long off, T v) {
  if (off >= 0 && off < n && comp >= 0 && comp < ncomp) p[comp * n + off] = v;
}

template <typename T>
__device__ __forceinline__ T GetPeriodic(const T *p, const int *dim, int nd,
                                         int i0, int i1 = 0, int i2 = 0) {
  long off = WrapIndex(i0, dim[0]);
  if (nd > 1) off += (long)WrapIndex(i1, dim[1]) * dim[0];
  if (nd > 2) off += (long)WrapIndex(i2, dim[2]) * dim[0] * dim[1];
  return p[off];
}

// Slot of the plane `k` planes past the oldest one in a ring whose oldest
// plane sits in slot `base`.
template <int NSLOT>
__device__ __forceinline__ int RingSlot(int base, int k) {
  int s = base + k;
  return s >= NSLOT ? s - NSLOT : s;
}

// ------------------------------------------------------ launch arithmetic
// Host-side tile counts of a run kernel over a domain.
struct RunShape {
  int x_first_tile;  // tile index of the first x tile (tiles anchor at 0)
  int tiles_x, tiles_y, tiles_z;
  int zchunk;
  int blocks;
};

}  // namespace physis_hip

#endif  // PHYSIS_HIP_STENCIL_H_
