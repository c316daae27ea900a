// libphysis_rt_hip: the Physis runtime for one MI355X.
//
// Replaces the CUDA target runtime (reference runtime/libphysis_rt_cuda.cc)
// entry point for entry point; see include/physis/physis_hip.h for the
// contract of each function.

#include <cstdarg>
#include <cstdint>
#include <cstring>
#include <string>

#include "runtime_hip.h"

FILE *__ps_trace = nullptr;

namespace physis_hip {

Runtime &rt() {
  static Runtime r;
  return r;
}

void EnsureInit() {
  Runtime &r = rt();
  if (r.initialized) return;
  int ndev = 0;
  PS_HIP_CHECK(hipGetDeviceCount(&ndev));
  if (ndev < 1) PS_FATAL("no HIP device found");
  int dev = 0;
  if (const char *s = std::getenv("PHYSIS_HIP_DEVICE")) dev = std::atoi(s);
  if (dev < 0 || dev >= ndev) PS_FATAL("PHYSIS_HIP_DEVICE=%d out of range", dev);
  PS_HIP_CHECK(hipSetDevice(dev));
  hipDeviceProp_t prop;
  PS_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  r.device = dev;
  r.num_cus = prop.multiProcessorCount;
  r.stream = nullptr;
  InitNoAllocScratch(dev);
  r.initialized = true;
}

void *Scratch(int slot, size_t bytes) {
  Runtime &r = rt();
  EnsureInit();
  if (slot < 0 || slot >= Runtime::kScratchSlots)
    PS_FATAL("scratch slot %d out of range", slot);
  if (bytes > r.scratch_bytes[slot]) {
    if (r.scratch[slot]) {
      PS_HIP_CHECK(hipStreamSynchronize(r.stream));
      PS_HIP_CHECK(hipFree(r.scratch[slot]));
    }
    size_t cap = bytes < (1u << 20) ? (1u << 20) : bytes;
    PS_HIP_CHECK(hipMalloc(&r.scratch[slot], cap));
    r.scratch_bytes[slot] = cap;
  }
  return r.scratch[slot];
}

void *HostResultSlot() {
  Runtime &r = rt();
  EnsureInit();
  if (!r.host_result) PS_HIP_CHECK(hipHostMalloc(&r.host_result, 256));
  return r.host_result;
}

namespace {

// Removes "--physis-trace" from argv and turns tracing on, as the reference
// runtime does (runtime/runtime.h:20-30, runtime_common.cc:14-35).
void ParseTraceOption(int *argc, char ***argv) {
  if (!argc || !argv || !*argv) return;
  int out = 0;
  for (int i = 0; i < *argc; ++i) {
    const char *a = (*argv)[i];
    if (a && (std::strcmp(a, "--physis-trace") == 0 ||
              std::strcmp(a, "-physis-trace") == 0)) {
      __ps_trace = stderr;
      continue;
    }
    (*argv)[out++] = (*argv)[i];
  }
  *argc = out;
}

PSType TypeFromSize(const __PSGridTypeInfo *ti) {
  if (ti) return ti->type;
  return PS_FLOAT;
}

// The member layout of a grid of user-defined points.  Host offsets follow
// the C rules for a struct of scalars and arrays of scalars (each member at
// the next multiple of its scalar size, the struct padded to its largest
// one); a type info that does not add up to the struct's size is rejected.
UserLayout *NewUserLayout(const __PSGridTypeInfo *ti, int64_t num_elms) {
  if (ti->num_members > UserLayout::kMaxMembers)
    PS_FATAL("point type with %d members (at most %d)", ti->num_members,
             UserLayout::kMaxMembers);
  UserLayout *l = new UserLayout;
  l->num_members = ti->num_members;
  size_t host = 0, dev = 0;
  int align = 1;
  for (int m = 0; m < ti->num_members; ++m) {
    const __PSGridTypeMemberInfo &mi = ti->members[m];
    const int sz = mi.size;
    if (sz != 1 && sz != 2 && sz != 4 && sz != 8)
      PS_FATAL("member %d: scalars of %d bytes are not supported", m, sz);
    if (mi.rank < 0 || mi.rank > PS_GRID_USER_TYPE_MAX_ARRAY_RANK)
      PS_FATAL("member %d: array rank %d", m, mi.rank);
    int comps = 1;
    for (int r = 0; r < mi.rank; ++r) comps *= mi.dim[r];
    l->size[m] = sz;
    l->comps[m] = comps;
    host = (host + sz - 1) / sz * sz;
    l->host_off[m] = host;
    host += (size_t)sz * comps;
    align = sz > align ? sz : align;
    dev = (dev + 255) / 256 * 256;
    l->dev_off[m] = dev;
    dev += (size_t)sz * comps * (size_t)num_elms;
  }
  host = (host + align - 1) / align * align;
  if ((int64_t)host != ti->size)
    PS_FATAL("point type of %d bytes whose members add up to %zu", ti->size,
             host);
  l->bytes = dev;
  return l;
}

void CopyIndices(va_list vl, int nd, PSIndex *idx) {
  for (int i = 0; i < nd; ++i) idx[i] = va_arg(vl, PSIndex);
}

int64_t LinearOffset(const __PSGrid *g, const PSIndex *idx) {
  int64_t off = 0, base = 1;
  for (int i = 0; i < g->num_dims; ++i) {
    if (idx[i] < 0 || idx[i] >= g->dim[i])
      PS_FATAL("grid index %d out of range [0,%d) on axis %d", idx[i],
               g->dim[i], i);
    off += (int64_t)idx[i] * base;
    base *= g->dim[i];
  }
  return off;
}

}  // namespace
}  // namespace physis_hip

using namespace physis_hip;

extern "C" {

void PSInit(int *argc, char ***argv, int grid_num_dims, ...) {
  (void)grid_num_dims;
  ParseTraceOption(argc, argv);
  EnsureInit();
}

void PSFinalize(void) {
  Runtime &r = rt();
  if (!r.initialized) return;
  PS_HIP_CHECK(hipDeviceSynchronize());
  for (int i = 0; i < Runtime::kScratchSlots; ++i) {
    if (r.scratch[i]) PS_HIP_CHECK(hipFree(r.scratch[i]));
    r.scratch[i] = nullptr;
    r.scratch_bytes[i] = 0;
  }
  if (r.host_result) PS_HIP_CHECK(hipHostFree(r.host_result));
  r.host_result = nullptr;
}

int __PSGridGetID(__PSGrid *g) {
  (void)g;
  return 0;
}

__PSGrid *__PSGridNew(__PSGridTypeInfo *type_info, int num_dims,
                      PSVectorInt dim, __PSGrid_devNewFunc func) {
  EnsureInit();
  if (num_dims < 1 || num_dims > PS_MAX_DIM)
    PS_FATAL("unsupported grid dimensionality %d", num_dims);
  __PSGrid *g = (__PSGrid *)std::calloc(1, sizeof(__PSGrid));
  if (!g) return INVALID_GRID;
  g->elm_size = type_info ? type_info->size : (int)sizeof(float);
  g->num_dims = num_dims;
  g->type = TypeFromSize(type_info);
  g->num_elms = 1;
  for (int i = 0; i < PS_MAX_DIM; ++i) g->dim[i] = 1;
  for (int i = 0; i < num_dims; ++i) {
    if (dim[i] < 0) PS_FATAL("negative grid extent %d", dim[i]);
    g->dim[i] = dim[i];
    g->num_elms *= dim[i];
  }
  if (func) {
    g->dev = (__PSGrid_dev *)func(num_dims, dim);
    g->p = nullptr;
    return g;
  }
  size_t bytes = (size_t)g->num_elms * (size_t)g->elm_size;
  if (type_info && type_info->num_members > 0) {
    UserLayout *l = NewUserLayout(type_info, g->num_elms);
    g->user = l;
    g->type = PS_USER;
    bytes = l->bytes;
  }
  void *buf = nullptr;
  if (bytes > 0) {
    hipError_t e = hipMalloc(&buf, bytes);
    if (e != hipSuccess || !buf) {
      (void)hipGetLastError();  // clear the sticky allocation error
      std::fprintf(stderr,
                   "[physis-hip] __PSGridNew: cannot allocate %zu bytes: %s\n",
                   bytes, hipGetErrorString(e));
      delete (UserLayout *)g->user;
      std::free(g);
      return INVALID_GRID;
    }
    // Grids start zero-filled like the reference (cudaMemset / calloc).
    PS_HIP_CHECK(hipMemsetAsync(buf, 0, bytes, rt().stream));
  }
  g->p = buf;
  g->dev = (__PSGrid_dev *)std::calloc(1, sizeof(__PSGrid_dev));
  for (int i = 0; i < 3; ++i) g->dev->dim[i] = g->dim[i];
  g->dev->p = buf;
  return g;
}

void __PSGridFree(void *gv, __PSGrid_devFreeFunc func) {
  __PSGrid *g = (__PSGrid *)gv;
  if (!g) return;
  if (g->dev) {
    if (func) {
      func(g->dev);
    } else {
      if (g->p) {
        PS_HIP_CHECK(hipStreamSynchronize(rt().stream));
        PS_HIP_CHECK(hipFree(g->p));
      }
      std::free(g->dev);
    }
    g->dev = nullptr;
  }
  delete (UserLayout *)g->user;
  g->user = nullptr;
  g->p = nullptr;
  std::free(g);
}

void __PSGridCopyin(void *gv, const void *src_array,
                    __PSGrid_devCopyinFunc func) {
  __PSGrid *g = (__PSGrid *)gv;
  if (func) {
    func(g->dev, src_array, (size_t)g->num_elms);
    return;
  }
  size_t bytes = (size_t)g->num_elms * (size_t)g->elm_size;
  if (bytes == 0) return;
  if (g->user) {
    // stage the host structs on the device, then split them into planes
    void *stage = Scratch(0, bytes);
    PS_HIP_CHECK(hipMemcpyAsync(stage, src_array, bytes, hipMemcpyHostToDevice,
                                rt().stream));
    UserAosToSoa(*(const UserLayout *)g->user, stage, (size_t)g->elm_size,
                 g->p, g->num_elms, rt().stream);
    PS_HIP_CHECK(hipStreamSynchronize(rt().stream));
    return;
  }
  PS_HIP_CHECK(hipMemcpyAsync(g->p, src_array, bytes, hipMemcpyHostToDevice,
                              rt().stream));
  PS_HIP_CHECK(hipStreamSynchronize(rt().stream));
}

void __PSGridCopyout(void *gv, void *dst_array,
                     __PSGrid_devCopyoutFunc func) {
  __PSGrid *g = (__PSGrid *)gv;
  if (func) {
    func(g->dev, dst_array, (size_t)g->num_elms);
    return;
  }
  size_t bytes = (size_t)g->num_elms * (size_t)g->elm_size;
  if (bytes == 0) return;
  if (g->user) {
    // gather the planes into structs on the device (padding bytes zero),
    // then copy them out
    void *stage = Scratch(0, bytes);
    PS_HIP_CHECK(hipMemsetAsync(stage, 0, bytes, rt().stream));
    UserSoaToAos(*(const UserLayout *)g->user, g->p, stage,
                 (size_t)g->elm_size, g->num_elms, rt().stream);
    PS_HIP_CHECK(hipMemcpyAsync(dst_array, stage, bytes, hipMemcpyDeviceToHost,
                                rt().stream));
    PS_HIP_CHECK(hipStreamSynchronize(rt().stream));
    return;
  }
  PS_HIP_CHECK(hipMemcpyAsync(dst_array, g->p, bytes, hipMemcpyDeviceToHost,
                              rt().stream));
  PS_HIP_CHECK(hipStreamSynchronize(rt().stream));
}

void *__PSGridUserMember(__PSGrid *g, int member) {
  if (!g || !g->user)
    PS_FATAL("__PSGridUserMember on a grid of a primitive type");
  const UserLayout *l = (const UserLayout *)g->user;
  if (member < 0 || member >= l->num_members)
    PS_FATAL("member %d of a point type with %d members", member,
             l->num_members);
  return (char *)g->p + l->dev_off[member];
}

void __PSGridSwap(__PSGrid *g) { (void)g; }

void __PSGridSet(__PSGrid *g, void *buf, ...) {
  PSIndex idx[PS_MAX_DIM] = {0, 0, 0};
  va_list vl;
  va_start(vl, buf);
  CopyIndices(vl, g->num_dims, idx);
  va_end(vl);
  if (g->user) PS_FATAL("PSGridSet on a grid of user-defined points");
  int64_t off = LinearOffset(g, idx) * g->elm_size;
  PS_HIP_CHECK(hipMemcpyAsync((char *)g->p + off, buf, g->elm_size,
                              hipMemcpyHostToDevice, rt().stream));
  PS_HIP_CHECK(hipStreamSynchronize(rt().stream));
}

void __PSGridGet(__PSGrid *g, void *buf, ...) {
  PSIndex idx[PS_MAX_DIM] = {0, 0, 0};
  va_list vl;
  va_start(vl, buf);
  CopyIndices(vl, g->num_dims, idx);
  va_end(vl);
  if (g->user) PS_FATAL("PSGridGet on a grid of user-defined points");
  int64_t off = LinearOffset(g, idx) * g->elm_size;
  PS_HIP_CHECK(hipMemcpyAsync(buf, (const char *)g->p + off, g->elm_size,
                              hipMemcpyDeviceToHost, rt().stream));
  PS_HIP_CHECK(hipStreamSynchronize(rt().stream));
}

PSDomain1D PSDomain1DNew(PSIndex minx, PSIndex maxx) {
  PSDomain1D d = {{minx, 0, 0}, {maxx, 0, 0}, {minx, 0, 0}, {maxx, 0, 0}};
  return d;
}

PSDomain2D PSDomain2DNew(PSIndex minx, PSIndex maxx, PSIndex miny,
                         PSIndex maxy) {
  PSDomain2D d = {{minx, miny, 0}, {maxx, maxy, 0},
                  {minx, miny, 0}, {maxx, maxy, 0}};
  return d;
}

PSDomain3D PSDomain3DNew(PSIndex minx, PSIndex maxx, PSIndex miny,
                         PSIndex maxy, PSIndex minz, PSIndex maxz) {
  PSDomain3D d = {{minx, miny, minz}, {maxx, maxy, maxz},
                  {minx, miny, minz}, {maxx, maxy, maxz}};
  return d;
}

void __PSCheckHipError(const char *message) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(rt().stream);
  if (e != hipSuccess) {
    std::fprintf(stderr, "ERROR: %s: %s\n", message, hipGetErrorString(e));
    PSAbort(1);
  }
}

void __PSReduceGridFloat(void *buf, enum PSReduceOp op, __PSGrid *g) {
  ReduceGrid(buf, op, g, PS_FLOAT);
}
void __PSReduceGridDouble(void *buf, enum PSReduceOp op, __PSGrid *g) {
  ReduceGrid(buf, op, g, PS_DOUBLE);
}
void __PSReduceGridInt(void *buf, enum PSReduceOp op, __PSGrid *g) {
  ReduceGrid(buf, op, g, PS_INT);
}
void __PSReduceGridLong(void *buf, enum PSReduceOp op, __PSGrid *g) {
  ReduceGrid(buf, op, g, PS_LONG);
}
void __PSReduceGridNoAlloc(void *buf, enum PSReduceOp op, __PSGrid *g,
                           PSType type) {
  ReduceGridNoAlloc(buf, op, g, type);
}

void *__PSHipStream(void) {
  EnsureInit();
  return (void *)rt().stream;
}

int __PSHipNumCUs(void) {
  EnsureInit();
  return rt().num_cus;
}

void *__PSHipScratch(int slot, size_t bytes) { return Scratch(slot, bytes); }

}  // extern "C"
