// Internal declarations of libphysis_rt_hip: error checking, the runtime
// singleton and helpers shared by grid, buffer and reduction code.
#ifndef PHYSIS_AMD_RUNTIME_HIP_H_
#define PHYSIS_AMD_RUNTIME_HIP_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "physis/physis_hip.h"

// Abort-on-error contract of the reference (runtime_common_cuda.h:16-27).
#define PS_HIP_CHECK(call)                                                   \
  do {                                                                       \
    hipError_t ps_err_ = (call);                                             \
    if (ps_err_ != hipSuccess) {                                             \
      std::fprintf(stderr, "[physis-hip] %s:%d: %s failed: %s\n", __FILE__,  \
                   __LINE__, #call, hipGetErrorString(ps_err_));             \
      std::exit(EXIT_FAILURE);                                               \
    }                                                                        \
  } while (0)

#define PS_FATAL(...)                                                        \
  do {                                                                       \
    std::fprintf(stderr, "[physis-hip] %s:%d: ", __FILE__, __LINE__);        \
    std::fprintf(stderr, __VA_ARGS__);                                       \
    std::fprintf(stderr, "\n");                                              \
    std::exit(EXIT_FAILURE);                                                 \
  } while (0)

namespace physis_hip {

struct Runtime {
  bool initialized = false;
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;  // null stream: the reference's default
  static constexpr int kScratchSlots = 10;
  void *scratch[kScratchSlots] = {};  // partial results, grown on demand
  size_t scratch_bytes[kScratchSlots] = {};
  void *host_result = nullptr;   // pinned landing slot for scalar results
};

Runtime &rt();
// Initialises the device lazily so that grids created before PSInit (as some
// test programs do) still work.
void EnsureInit();
void *Scratch(int slot, size_t bytes);
void *HostResultSlot();

// reduce_hip.hip
void ReduceGrid(void *buf, PSReduceOp op, const __PSGrid *g, PSType type);
// The same reduction with its partials in a fixed device array (no
// allocation); InitNoAllocScratch looks up that array's address on device
// dev and is called once per device at set-up.
void ReduceGridNoAlloc(void *buf, PSReduceOp op, const __PSGrid *g,
                       PSType type);
void InitNoAllocScratch(int dev);

// Storage of a grid of user-defined points (type PS_USER with members):
// member by member, each scalar component of a member (array members have
// several, row-major) a dense plane of num_elms values in the grid's own
// index order; member m's first plane starts dev_off[m] bytes into the
// buffer (256-byte aligned), component c of it c * num_elms * size[m] bytes
// later.  host_off[m] is the member's byte offset in the C struct the host
// arrays hold.
struct UserLayout {
  static constexpr int kMaxMembers = 64;
  int num_members = 0;
  int size[kMaxMembers];        // bytes of one scalar of the member
  int comps[kMaxMembers];       // scalars per point (product of extents)
  size_t host_off[kMaxMembers];
  size_t dev_off[kMaxMembers];
  size_t bytes = 0;             // device buffer size
};

// grid_user_hip.hip: host AoS array (staged in device memory) <-> planes
void UserAosToSoa(const UserLayout &l, const void *aos_dev, size_t stride,
                  void *soa, int64_t n, hipStream_t s);
void UserSoaToAos(const UserLayout &l, const void *soa, void *aos_dev,
                  size_t stride, int64_t n, hipStream_t s);

}  // namespace physis_hip

#endif  // PHYSIS_AMD_RUNTIME_HIP_H_
