// Streaming-bandwidth ceilings of one MI355X, the reference points that the
// run kernels' HBM rates are compared against: a copy (one read and one
// write stream) and a sum of 13 arrays into a 14th (the 19-point Himeno
// sweep's 13 read streams and one write stream).  Both are plain
// grid-stride loops over 16-byte vectors, four loads in flight per thread,
// non-temporal stores; nothing is reused, so every byte moves through HBM.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kThreads = 256;
constexpr int kUnroll = 4;

__global__ __launch_bounds__(kThreads) void CopyKernel(
    const f32x4 *__restrict__ src, f32x4 *__restrict__ dst, long n4) {
  const long stride = (long)gridDim.x * kThreads;
  long i = (long)blockIdx.x * kThreads + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < n4; i += kUnroll * stride) {
    f32x4 v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n4; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

struct Srcs13 {
  const f32x4 *p[13];
};

__global__ __launch_bounds__(kThreads) void Sum13Kernel(
    Srcs13 s, f32x4 *__restrict__ dst, long n4) {
  const long stride = (long)gridDim.x * kThreads;
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n4;
       i += stride) {
    f32x4 acc = __builtin_nontemporal_load(s.p[0] + i);
#pragma unroll
    for (int g = 1; g < 13; ++g) acc += __builtin_nontemporal_load(s.p[g] + i);
    __builtin_nontemporal_store(acc, dst + i);
  }
}

void Check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "[physis-hip] %s: %s\n", what, hipGetErrorString(e));
    std::exit(EXIT_FAILURE);
  }
}

}  // namespace

extern "C" {

// dst[i] = src[i] for n floats (n a multiple of 4, 16-byte aligned arrays)
// on `stream` with `blocks` blocks of 256 threads.
void PSHipStreamCopy(const float *src, float *dst, long n, int blocks,
                     void *stream) {
  if (n % 4 || blocks < 1) {
    std::fprintf(stderr, "[physis-hip] PSHipStreamCopy: n %% 4 != 0\n");
    std::exit(EXIT_FAILURE);
  }
  hipLaunchKernelGGL(CopyKernel, dim3(blocks), dim3(kThreads), 0,
                     (hipStream_t)stream, (const f32x4 *)src, (f32x4 *)dst,
                     n / 4);
  Check(hipGetLastError(), "PSHipStreamCopy");
}

// dst[i] = src[0][i] + ... + src[12][i] for n floats.
void PSHipStreamSum13(const float *const *src, float *dst, long n, int blocks,
                      void *stream) {
  if (n % 4 || blocks < 1) {
    std::fprintf(stderr, "[physis-hip] PSHipStreamSum13: n %% 4 != 0\n");
    std::exit(EXIT_FAILURE);
  }
  Srcs13 s;
  for (int g = 0; g < 13; ++g) s.p[g] = (const f32x4 *)src[g];
  hipLaunchKernelGGL(Sum13Kernel, dim3(blocks), dim3(kThreads), 0,
                     (hipStream_t)stream, s, (f32x4 *)dst, n / 4);
  Check(hipGetLastError(), "PSHipStreamSum13");
}

}  // extern "C"
