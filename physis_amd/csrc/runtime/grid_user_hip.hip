// Copy-in / copy-out of grids of user-defined point types: the host holds an
// array of C structs (array of structures), the device one plane per scalar
// component (structure of arrays, so that a stencil reading one member of
// its neighbours reads consecutive addresses).  The host array is staged in
// device memory and transposed there, one launch per member; each thread
// moves all components of one point of the member.
#include "runtime_hip.h"

namespace physis_hip {

namespace {

template <typename U>
__global__ __launch_bounds__(256) void AosToSoaKernel(
    const unsigned char *__restrict__ aos, size_t stride, size_t off,
    int comps, U *__restrict__ plane, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const U *src = reinterpret_cast<const U *>(aos + (size_t)i * stride + off);
  for (int c = 0; c < comps; ++c) plane[(int64_t)c * n + i] = src[c];
}

template <typename U>
__global__ __launch_bounds__(256) void SoaToAosKernel(
    const U *__restrict__ plane, int comps, unsigned char *__restrict__ aos,
    size_t stride, size_t off, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  U *dst = reinterpret_cast<U *>(aos + (size_t)i * stride + off);
  for (int c = 0; c < comps; ++c) dst[c] = plane[(int64_t)c * n + i];
}

template <typename U>
void LaunchIn(const UserLayout &l, int m, const void *aos, size_t stride,
              void *soa, int64_t n, hipStream_t s) {
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL((AosToSoaKernel<U>), dim3(blocks), dim3(256), 0, s,
                     (const unsigned char *)aos, stride, l.host_off[m],
                     l.comps[m], (U *)((char *)soa + l.dev_off[m]), n);
}

template <typename U>
void LaunchOut(const UserLayout &l, int m, const void *soa, void *aos,
               size_t stride, int64_t n, hipStream_t s) {
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL((SoaToAosKernel<U>), dim3(blocks), dim3(256), 0, s,
                     (const U *)((const char *)soa + l.dev_off[m]), l.comps[m],
                     (unsigned char *)aos, stride, l.host_off[m], n);
}

}  // namespace

void UserAosToSoa(const UserLayout &l, const void *aos_dev, size_t stride,
                  void *soa, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  for (int m = 0; m < l.num_members; ++m) {
    switch (l.size[m]) {
      case 1: LaunchIn<uint8_t>(l, m, aos_dev, stride, soa, n, s); break;
      case 2: LaunchIn<uint16_t>(l, m, aos_dev, stride, soa, n, s); break;
      case 4: LaunchIn<uint32_t>(l, m, aos_dev, stride, soa, n, s); break;
      case 8: LaunchIn<uint64_t>(l, m, aos_dev, stride, soa, n, s); break;
      default: PS_FATAL("member scalar of %d bytes", l.size[m]);
    }
  }
  PS_HIP_CHECK(hipGetLastError());
}

void UserSoaToAos(const UserLayout &l, const void *soa, void *aos_dev,
                  size_t stride, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  for (int m = 0; m < l.num_members; ++m) {
    switch (l.size[m]) {
      case 1: LaunchOut<uint8_t>(l, m, soa, aos_dev, stride, n, s); break;
      case 2: LaunchOut<uint16_t>(l, m, soa, aos_dev, stride, n, s); break;
      case 4: LaunchOut<uint32_t>(l, m, soa, aos_dev, stride, n, s); break;
      case 8: LaunchOut<uint64_t>(l, m, soa, aos_dev, stride, n, s); break;
      default: PS_FATAL("member scalar of %d bytes", l.size[m]);
    }
  }
  PS_HIP_CHECK(hipGetLastError());
}

}  // namespace physis_hip
