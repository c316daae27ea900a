// reduce_hip: PSReduce over a whole grid on the device.
//
// Replaces the CUDA target's Thrust calls (reference runtime/reduce_cuda.h:
// 19-40, reduce_cuda.cu:14-40).  Two deterministic stages:
//   1. a fixed number of 256-thread blocks stream the grid with 16-byte
//      loads, four in flight per thread, fold them (float SUM/PROD in
//      double, everything else in the element type), and reduce wave64 ->
//      block, writing one partial per block;
//   2. one block folds the partials in a fixed order.
// Float sums are therefore the FP64 sum rounded once to float; the REF
// target's sequential float loop (libphysis_rt_ref.cc:19-30) drifts from
// that value as the running sum grows.
// The result is copied into pinned host memory and returned synchronously,
// as the reference does.  Traffic: one read of num_elms * elm_size bytes.

#include <cstring>

#include "physis/hip/device.h"
#include "runtime_hip.h"

namespace physis_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;

// 16-byte native vectors (clang ext_vector_type) so that the loads lower to
// global_load_dwordx4 and accept the nontemporal hint.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef long i64x2 __attribute__((ext_vector_type(2)));
template <typename T> struct Vec16;
template <> struct Vec16<float> { typedef f32x4 V; static constexpr int N = 4; };
template <> struct Vec16<int> { typedef i32x4 V; static constexpr int N = 4; };
template <> struct Vec16<double> { typedef f64x2 V; static constexpr int N = 2; };
template <> struct Vec16<long> { typedef i64x2 V; static constexpr int N = 2; };

// Accumulation type: float SUM and PROD fold in double, so the result is the
// FP64 value rounded once to float, independent of the fold order.  Other
// cases fold in the element type (exact for MAX/MIN, wrapping for integers).
template <typename T, int OP> struct Acc { typedef T A; };
template <> struct Acc<float, PS_SUM> { typedef double A; };
template <> struct Acc<float, PS_PROD> { typedef double A; };

template <typename T, int OP>
__device__ __forceinline__ typename Acc<T, OP>::A FoldVec(
    typename Acc<T, OP>::A acc, const typename Vec16<T>::V &v) {
  typedef typename Acc<T, OP>::A A;
  typedef ReduceOp<A, OP> Op;
#pragma unroll
  for (int i = 0; i < Vec16<T>::N; ++i) acc = Op::apply(acc, (A)v[i]);
  return acc;
}

template <typename T, int OP>
__global__ __launch_bounds__(kThreads) void ReduceStage1(
    const T *__restrict__ p, int64_t n, typename Acc<T, OP>::A *__restrict__ partials) {
  typedef typename Acc<T, OP>::A A;
  typedef ReduceOp<A, OP> Op;
  typedef typename Vec16<T>::V V;
  constexpr int N = Vec16<T>::N;
  __shared__ A lds[kThreads / kWave];

  const int64_t nvec = n / N;
  const V *pv = reinterpret_cast<const V *>(p);
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  A acc = Op::identity();
  // Four independent 16-byte loads per thread per trip keep enough bytes in
  // flight to stream at HBM rate.
  for (; i + (kUnroll - 1) * stride < nvec; i += kUnroll * stride) {
    V v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      v[u] = __builtin_nontemporal_load(pv + i + u * stride);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc = FoldVec<T, OP>(acc, v[u]);
  }
  for (; i < nvec; i += stride) acc = FoldVec<T, OP>(acc, pv[i]);
  // Scalar tail, taken by the first threads of block 0.
  if (blockIdx.x == 0) {
    int64_t t = nvec * N + threadIdx.x;
    if (t < n) acc = Op::apply(acc, (A)p[t]);
  }
  A r = BlockReduce<A, OP, kThreads>(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <typename T, int OP>
__global__ __launch_bounds__(1024) void ReduceStage2(
    const typename Acc<T, OP>::A *__restrict__ partials, int nparts,
    T *__restrict__ out) {
  typedef typename Acc<T, OP>::A A;
  typedef ReduceOp<A, OP> Op;
  __shared__ A lds[1024 / kWave];
  A acc = Op::identity();
  for (int i = threadIdx.x; i < nparts; i += 1024)
    acc = Op::apply(acc, partials[i]);
  A r = BlockReduce<A, OP, 1024>(acc, lds);
  if (threadIdx.x == 0) out[0] = (T)r;
}

// Number of stage-1 blocks for n elements: enough to stream the grid, at
// most eight per compute unit.
template <typename T>
int StageBlocks(int64_t n) {
  int64_t per_block = (int64_t)kThreads * Vec16<T>::N * kUnroll;
  int64_t want = (n + per_block - 1) / per_block;
  int64_t cap = (int64_t)rt().num_cus * 8;
  return (int)(want < 1 ? 1 : (want > cap ? cap : want));
}

// Both stages on the runtime stream: nb partials into `parts`, the result
// (one T) into `out`; then the result is copied to host_out and the stream
// drained.
template <typename T, int OP>
void LaunchStages(const T *p, int64_t n, int nb,
                  typename Acc<T, OP>::A *parts, T *out, void *host_out) {
  Runtime &r = rt();
  hipLaunchKernelGGL((ReduceStage1<T, OP>), dim3(nb), dim3(kThreads), 0,
                     r.stream, p, n, parts);
  PS_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL((ReduceStage2<T, OP>), dim3(1), dim3(1024), 0, r.stream,
                     parts, nb, out);
  PS_HIP_CHECK(hipGetLastError());
  PS_HIP_CHECK(hipMemcpyAsync(host_out, out, sizeof(T), hipMemcpyDeviceToHost,
                              r.stream));
  PS_HIP_CHECK(hipStreamSynchronize(r.stream));
}

template <typename T, int OP>
void Launch(const T *p, int64_t n, void *host_out) {
  const int nb = StageBlocks<T>(n);
  typedef typename Acc<T, OP>::A A;
  A *parts = (A *)Scratch(0, sizeof(A) * (size_t)(nb + 1));
  LaunchStages<T, OP>(p, n, nb, parts, (T *)(parts + nb), host_out);
}

template <typename T>
void Dispatch(void *buf, PSReduceOp op, const __PSGrid *g) {
  if (g->num_elms <= 0) PS_FATAL("PSReduce on an empty grid");
  void *slot = HostResultSlot();
  const T *p = (const T *)g->p;
  switch (op) {
    case PS_MAX: Launch<T, PS_MAX>(p, g->num_elms, slot); break;
    case PS_MIN: Launch<T, PS_MIN>(p, g->num_elms, slot); break;
    case PS_SUM: Launch<T, PS_SUM>(p, g->num_elms, slot); break;
    case PS_PROD: Launch<T, PS_PROD>(p, g->num_elms, slot); break;
    default: PS_FATAL("unknown reduction operator %d", (int)op);
  }
  std::memcpy(buf, slot, sizeof(T));
}

// ------------------------------------------- reduction without allocation
// Partials and result of __PSReduceGridNoAlloc live in this file-scope
// device array, one copy per GPU, held until the process ends.  Host code
// never passes `scratch` itself to a kernel or copy (in host code the name
// is a stand-in, not a device address): it uses the address that
// hipGetSymbolAddress returns for the current device, kept in g_scratch_addr.
// Stage 1 writes every partial it launches and stage 2 writes the result,
// so nothing in the array has to start at zero.
constexpr int kNoAllocMaxParts = 4096;
constexpr size_t kNoAllocScratchBytes = 8 * (kNoAllocMaxParts + 1);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
void *g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on device `dev`, looked up once per device.
void *NoAllocScratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices)
    PS_FATAL("device %d beyond the %d devices the no-allocation reduction "
             "supports", dev, kMaxDevices);
  if (!g_scratch_addr[dev]) {
    void *ptr = nullptr;
    PS_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) PS_FATAL("no device address for the reduction scratch array");
    g_scratch_addr[dev] = ptr;
  }
  return g_scratch_addr[dev];
}

template <typename T, int OP>
void LaunchNoAlloc(const T *p, int64_t n, void *buf, unsigned char *area) {
  typedef typename Acc<T, OP>::A A;
  int nb = StageBlocks<T>(n);
  if (nb > kNoAllocMaxParts) nb = kNoAllocMaxParts;
  const size_t need = sizeof(A) * (size_t)nb + sizeof(T);
  if (need > kNoAllocScratchBytes)
    PS_FATAL("reduction needs %zu scratch bytes, more than the %zu of the "
             "no-allocation area", need, kNoAllocScratchBytes);
  A *parts = (A *)area;
  LaunchStages<T, OP>(p, n, nb, parts, (T *)(parts + nb), buf);
}

template <typename T>
void DispatchNoAlloc(void *buf, PSReduceOp op, const __PSGrid *g,
                     unsigned char *area) {
  const T *p = (const T *)g->p;
  switch (op) {
    case PS_MAX: LaunchNoAlloc<T, PS_MAX>(p, g->num_elms, buf, area); break;
    case PS_MIN: LaunchNoAlloc<T, PS_MIN>(p, g->num_elms, buf, area); break;
    case PS_SUM: LaunchNoAlloc<T, PS_SUM>(p, g->num_elms, buf, area); break;
    case PS_PROD: LaunchNoAlloc<T, PS_PROD>(p, g->num_elms, buf, area); break;
    default: PS_FATAL("unknown reduction operator %d", (int)op);
  }
}

// Checks that grid g can be reduced as `type`.
void CheckReducible(const __PSGrid *g, PSType type) {
  int want_size = 0;
  switch (type) {
    case PS_FLOAT: want_size = sizeof(float); break;
    case PS_DOUBLE: want_size = sizeof(double); break;
    case PS_INT: want_size = sizeof(int); break;
    case PS_LONG: want_size = sizeof(long); break;
    default: PS_FATAL("PSReduce on unsupported element type %d", (int)type);
  }
  if (!g) PS_FATAL("PSReduce on a null grid");
  if (g->user) PS_FATAL("PSReduce on a grid of user-defined points");
  if (g->elm_size != want_size)
    PS_FATAL("PSReduce: grid element size %d does not match the %d-byte "
             "reduction type", g->elm_size, want_size);
  if (g->num_elms <= 0) PS_FATAL("PSReduce on an empty grid");
}

}  // namespace

void InitNoAllocScratch(int dev) { NoAllocScratch(dev); }

void ReduceGrid(void *buf, PSReduceOp op, const __PSGrid *g, PSType type) {
  EnsureInit();
  CheckReducible(g, type);
  switch (type) {
    case PS_FLOAT: Dispatch<float>(buf, op, g); break;
    case PS_DOUBLE: Dispatch<double>(buf, op, g); break;
    case PS_INT: Dispatch<int>(buf, op, g); break;
    case PS_LONG: Dispatch<long>(buf, op, g); break;
    default: break;
  }
}

void ReduceGridNoAlloc(void *buf, PSReduceOp op, const __PSGrid *g,
                       PSType type) {
  if (!rt().initialized)
    PS_FATAL("__PSReduceGridNoAlloc before PSInit (the runtime must be set "
             "up first)");
  CheckReducible(g, type);
  int dev = 0;
  PS_HIP_CHECK(hipGetDevice(&dev));
  unsigned char *area = (unsigned char *)NoAllocScratch(dev);
  switch (type) {
    case PS_FLOAT: DispatchNoAlloc<float>(buf, op, g, area); break;
    case PS_DOUBLE: DispatchNoAlloc<double>(buf, op, g, area); break;
    case PS_INT: DispatchNoAlloc<int>(buf, op, g, area); break;
    case PS_LONG: DispatchNoAlloc<long>(buf, op, g, area); break;
    default: break;
  }
}

// -------------------------------------------- partials of fused reductions
// Run kernels of PSStencilMapReduce maps leave one partial per block for each
// execution; this folds row r of `partials` (nparts values) into out[r], in
// a fixed order, accumulating float data in double.
template <typename T> struct FoldAcc { typedef T A; };
template <> struct FoldAcc<float> { typedef double A; };

template <typename T, int OP>
__global__ __launch_bounds__(256) void FoldRows(const T *__restrict__ partials,
                                                int nparts,
                                                T *__restrict__ out) {
  typedef typename FoldAcc<T>::A A;
  typedef ReduceOp<A, OP> Op;
  __shared__ A lds[256 / kWave];
  const T *row = partials + (size_t)blockIdx.x * nparts;
  A acc = Op::identity();
  for (int i = threadIdx.x; i < nparts; i += 256)
    acc = Op::apply(acc, (A)row[i]);
  A r = BlockReduce<A, OP, 256>(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = (T)r;
}

template <typename T>
void FoldTyped(int op, const void *partials, int nparts, int count,
               void *host_out) {
  Runtime &r = rt();
  // Results land after the partials in the scratch area of the caller, so
  // they use a separate small device buffer.
  static T *dev_out = nullptr;
  static int dev_cap = 0;
  if (count > dev_cap) {
    if (dev_out) PS_HIP_CHECK(hipFree(dev_out));
    PS_HIP_CHECK(hipMalloc(&dev_out, sizeof(T) * (size_t)count));
    dev_cap = count;
  }
  const T *p = (const T *)partials;
  switch (op) {
    case PS_MAX: hipLaunchKernelGGL((FoldRows<T, PS_MAX>), dim3(count), dim3(256), 0, r.stream, p, nparts, dev_out); break;
    case PS_MIN: hipLaunchKernelGGL((FoldRows<T, PS_MIN>), dim3(count), dim3(256), 0, r.stream, p, nparts, dev_out); break;
    case PS_SUM: hipLaunchKernelGGL((FoldRows<T, PS_SUM>), dim3(count), dim3(256), 0, r.stream, p, nparts, dev_out); break;
    case PS_PROD: hipLaunchKernelGGL((FoldRows<T, PS_PROD>), dim3(count), dim3(256), 0, r.stream, p, nparts, dev_out); break;
    default: PS_FATAL("unknown reduction operator %d", op);
  }
  PS_HIP_CHECK(hipGetLastError());
  PS_HIP_CHECK(hipMemcpyAsync(host_out, dev_out, sizeof(T) * (size_t)count,
                              hipMemcpyDeviceToHost, r.stream));
  PS_HIP_CHECK(hipStreamSynchronize(r.stream));
}

}  // namespace physis_hip

extern "C" void __PSHipReduceFold(int type, int op, const void *partials,
                                  int nparts, int count, void *host_out) {
  using namespace physis_hip;
  EnsureInit();
  if (count <= 0) return;
  if (nparts <= 0) PS_FATAL("reduction over an empty domain");
  switch (type) {
    case PS_FLOAT: FoldTyped<float>(op, partials, nparts, count, host_out); break;
    case PS_DOUBLE: FoldTyped<double>(op, partials, nparts, count, host_out); break;
    case PS_INT: FoldTyped<int>(op, partials, nparts, count, host_out); break;
    case PS_LONG: FoldTyped<long>(op, partials, nparts, count, host_out); break;
    default: PS_FATAL("unsupported reduction type %d", type);
  }
}
