// Capacity-managed buffers of libphysis_rt_hip (reference runtime/buffer.h:
// 11-108 and buffer_cuda.h): a buffer grows on EnsureCapacity and keeps its
// allocation when asked for less.
#ifndef PHYSIS_AMD_BUFFER_HIP_H_
#define PHYSIS_AMD_BUFFER_HIP_H_

#include <cstddef>

#include "physis/physis_hip.h"

namespace physis_hip {

class Buffer {
 public:
  virtual ~Buffer();
  void EnsureCapacity(size_t size);
  void Free();
  void *Get() { return ptr_; }
  const void *Get() const { return ptr_; }
  size_t size() const { return size_; }
  size_t capacity() const { return capacity_; }

 protected:
  virtual void *Allocate(size_t size) = 0;
  virtual void Release(void *p) = 0;

 private:
  void *ptr_ = nullptr;
  size_t size_ = 0;
  size_t capacity_ = 0;
};

// Pageable host memory.
class BufferHost : public Buffer {
 public:
  ~BufferHost() override;

 protected:
  void *Allocate(size_t size) override;
  void Release(void *p) override;
};

// Pinned host memory (hipHostMalloc) for staging transfers.
class BufferHipHost : public Buffer {
 public:
  ~BufferHipHost() override;

 protected:
  void *Allocate(size_t size) override;
  void Release(void *p) override;
};

// Zero-filled device memory.
class BufferHipDev : public Buffer {
 public:
  ~BufferHipDev() override;

 protected:
  void *Allocate(size_t size) override;
  void Release(void *p) override;
};

// Moves the sub-grid [offset, offset+size) of a device grid to (to_host) or
// from a dense host array, packing on the device when it is not contiguous.
void CopySubgrid(__PSGrid *g, void *host, const PSIndex *offset,
                 const PSIndex *size, bool to_host);

}  // namespace physis_hip

#endif  // PHYSIS_AMD_BUFFER_HIP_H_
