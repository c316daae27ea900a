// buffer_hip: capacity-managed host/device buffers and sub-grid transfers.
//
// Counterpart of the reference's Buffer / BufferHost / BufferCUDADev /
// BufferCUDAHost (runtime/buffer.h:11-108, buffer_cuda.{h,cu}).  A sub-grid
// of a device grid that is not one contiguous run of memory is gathered
// into (or scattered from) a linear staging buffer by a pack (unpack)
// kernel and moved with a single copy; a contiguous one is copied directly.
//
// Pack / unpack move whole sub-grid rows: one wavefront per row strip, each
// lane moving 4-byte words, so both sides of every access are coalesced.

#include <cstring>

#include "buffer_hip.h"
#include "physis/hip/device.h"
#include "runtime_hip.h"

namespace physis_hip {

// ------------------------------------------------------------------ Buffer
Buffer::~Buffer() {}

void Buffer::EnsureCapacity(size_t size) {
  if (size <= capacity_) {
    size_ = size;
    return;
  }
  Free();
  if (size) ptr_ = Allocate(size);
  capacity_ = size;
  size_ = size;
}

void Buffer::Free() {
  if (ptr_) Release(ptr_);
  ptr_ = nullptr;
  capacity_ = size_ = 0;
}

BufferHost::~BufferHost() { Free(); }
void *BufferHost::Allocate(size_t size) {
  void *p = std::calloc(1, size);
  if (!p) PS_FATAL("host allocation of %zu bytes failed", size);
  return p;
}
void BufferHost::Release(void *p) { std::free(p); }

BufferHipHost::~BufferHipHost() { Free(); }
void *BufferHipHost::Allocate(size_t size) {
  void *p = nullptr;
  PS_HIP_CHECK(hipHostMalloc(&p, size));
  std::memset(p, 0, size);
  return p;
}
void BufferHipHost::Release(void *p) { PS_HIP_CHECK(hipHostFree(p)); }

BufferHipDev::~BufferHipDev() { Free(); }
void *BufferHipDev::Allocate(size_t size) {
  EnsureInit();
  void *p = nullptr;
  PS_HIP_CHECK(hipMalloc(&p, size));
  PS_HIP_CHECK(hipMemsetAsync(p, 0, size, rt().stream));
  return p;
}
void BufferHipDev::Release(void *p) {
  PS_HIP_CHECK(hipStreamSynchronize(rt().stream));
  PS_HIP_CHECK(hipFree(p));
}

// ------------------------------------------------------- pack / unpack
namespace {

// Sub-grid geometry in 4-byte words along x.
struct SubGrid {
  int gx, gy;       // grid extent in words (x) and rows (y)
  int ox, oy, oz;   // sub-grid offset (x in words)
  int sx, sy, sz;   // sub-grid extent (x in words)
};

constexpr int kPackThreads = 256;

// Block b handles rows [b*4, b*4+4) of the sy*sz rows; each wave one row.
__global__ __launch_bounds__(kPackThreads) void PackRows(
    const unsigned *__restrict__ grid, unsigned *__restrict__ lin, SubGrid s) {
  const int row = blockIdx.x * (kPackThreads / kWave) + threadIdx.x / kWave;
  if (row >= s.sy * s.sz) return;
  const int y = row % s.sy, z = row / s.sy;
  const long src = (long)s.ox + (long)(s.oy + y) * s.gx +
                   (long)(s.oz + z) * s.gx * s.gy;
  const long dst = (long)row * s.sx;
  for (int x = threadIdx.x % kWave; x < s.sx; x += kWave)
    lin[dst + x] = grid[src + x];
}

__global__ __launch_bounds__(kPackThreads) void UnpackRows(
    unsigned *__restrict__ grid, const unsigned *__restrict__ lin, SubGrid s) {
  const int row = blockIdx.x * (kPackThreads / kWave) + threadIdx.x / kWave;
  if (row >= s.sy * s.sz) return;
  const int y = row % s.sy, z = row / s.sy;
  const long dst = (long)s.ox + (long)(s.oy + y) * s.gx +
                   (long)(s.oz + z) * s.gx * s.gy;
  const long src = (long)row * s.sx;
  for (int x = threadIdx.x % kWave; x < s.sx; x += kWave)
    grid[dst + x] = lin[src + x];
}

SubGrid MakeSubGrid(const __PSGrid *g, const PSIndex *off,
                    const PSIndex *size) {
  PSIndex o[3] = {0, 0, 0}, n[3] = {1, 1, 1};
  for (int d = 0; d < g->num_dims; ++d) {
    o[d] = off[d];
    n[d] = size[d];
    if (o[d] < 0 || n[d] < 0 || o[d] + n[d] > g->dim[d])
      PS_FATAL("sub-grid [%d, %d) out of grid extent %d on axis %d", o[d],
               o[d] + n[d], g->dim[d], d);
  }
  if (g->elm_size % 4 != 0)
    PS_FATAL("sub-grid copy needs a multiple-of-4 element size, got %d",
             g->elm_size);
  const int w = g->elm_size / 4;
  SubGrid s;
  s.gx = g->dim[0] * w;
  s.gy = g->num_dims > 1 ? g->dim[1] : 1;
  s.ox = o[0] * w;
  s.oy = o[1];
  s.oz = o[2];
  s.sx = n[0] * w;
  s.sy = n[1];
  s.sz = n[2];
  return s;
}

// True when the sub-grid is one contiguous run of the grid (the reference's
// "packing unnecessary" test, buffer_cuda.cu:117-124).
bool Contiguous(const __PSGrid *g, const SubGrid &s) {
  if (g->num_dims >= 2 && s.sx != s.gx && s.sy * s.sz > 1) return false;
  if (g->num_dims >= 3 && s.sy != s.gy && s.sz > 1) return false;
  return true;
}

// Never destroyed: a static destructor would run after the HIP runtime has
// been torn down at process exit.
BufferHipDev &Staging() {
  static BufferHipDev *b = new BufferHipDev;
  return *b;
}

}  // namespace

void CopySubgrid(__PSGrid *g, void *host, const PSIndex *off,
                 const PSIndex *size, bool to_host) {
  EnsureInit();
  const SubGrid s = MakeSubGrid(g, off, size);
  const size_t words = (size_t)s.sx * s.sy * s.sz;
  const size_t bytes = words * 4;
  if (bytes == 0) return;
  hipStream_t st = rt().stream;
  char *base = (char *)g->p;
  if (Contiguous(g, s)) {
    const size_t start =
        ((size_t)s.ox + (size_t)s.oy * s.gx + (size_t)s.oz * s.gx * s.gy) * 4;
    if (to_host)
      PS_HIP_CHECK(hipMemcpyAsync(host, base + start, bytes,
                                  hipMemcpyDeviceToHost, st));
    else
      PS_HIP_CHECK(hipMemcpyAsync(base + start, host, bytes,
                                  hipMemcpyHostToDevice, st));
    PS_HIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  BufferHipDev &tmp = Staging();
  tmp.EnsureCapacity(bytes);
  const int rows = s.sy * s.sz;
  const int blocks = (rows + kPackThreads / kWave - 1) / (kPackThreads / kWave);
  if (to_host) {
    hipLaunchKernelGGL(PackRows, dim3(blocks), dim3(kPackThreads), 0, st,
                       (const unsigned *)g->p, (unsigned *)tmp.Get(), s);
    PS_HIP_CHECK(hipGetLastError());
    PS_HIP_CHECK(hipMemcpyAsync(host, tmp.Get(), bytes, hipMemcpyDeviceToHost,
                                st));
  } else {
    PS_HIP_CHECK(hipMemcpyAsync(tmp.Get(), host, bytes, hipMemcpyHostToDevice,
                                st));
    hipLaunchKernelGGL(UnpackRows, dim3(blocks), dim3(kPackThreads), 0, st,
                       (unsigned *)g->p, (const unsigned *)tmp.Get(), s);
    PS_HIP_CHECK(hipGetLastError());
  }
  PS_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace physis_hip

extern "C" {

void __PSGridCopyoutSubgrid(__PSGrid *g, void *dst, const PSIndex *offset,
                            const PSIndex *size) {
  physis_hip::CopySubgrid(g, dst, offset, size, true);
}

void __PSGridCopyinSubgrid(__PSGrid *g, const void *src, const PSIndex *offset,
                           const PSIndex *size) {
  physis_hip::CopySubgrid(g, const_cast<void *>(src), offset, size, false);
}

}  // extern "C"
