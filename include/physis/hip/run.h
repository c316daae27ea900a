// Host-side helpers of the run functions that physisc-hip generates: launch
// planning for 2.5D run kernels, argument checks and the run timer that backs
// --physis-trace (reference: translator/reference_runtime_builder.cc:896-940
// emits the equivalent stopwatch code inline).
#ifndef PHYSIS_HIP_RUN_H_
#define PHYSIS_HIP_RUN_H_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>

#include "physis/physis_hip.h"
#include "physis/hip/stencil.h"

namespace physis_hip {

// ------------------------------------------------ host intrinsic helpers
// PSGrid<N>D<T>New(d0, d1, d2): the type-info block the reference emits
// inline (reference_runtime_builder.cc:959-1065) folded into one call.
inline __PSGrid *NewGrid(int type, int size, int num_dims, int d0, int d1 = 1,
                         int d2 = 1) {
  __PSGridTypeInfo ti = {type, size, 0, nullptr};
  PSVectorInt dim = {d0, d1, d2};
  return __PSGridNew(&ti, num_dims, dim, nullptr);
}

// PSGrid<n>D<Name>New for a user-defined point type: ti describes its members.
inline __PSGrid *NewUserGrid(__PSGridTypeInfo *ti, int num_dims, int d0,
                             int d1 = 1, int d2 = 1) {
  PSVectorInt dim = {d0, d1, d2};
  return __PSGridNew(ti, num_dims, dim, nullptr);
}

// PSGridGet(g, i...) in host code: one element copied back (synchronous).
template <typename T>
inline T GetElem(__PSGrid *g, PSIndex i0, PSIndex i1 = 0, PSIndex i2 = 0) {
  T v;
  __PSGridGet(g, &v, i0, i1, i2);
  return v;
}

// PSGridSet(g, i..., v) with v converted to the element type.
template <typename T>
inline void SetElem(__PSGrid *g, T v, PSIndex i0, PSIndex i1 = 0,
                    PSIndex i2 = 0) {
  __PSGridSet(g, &v, i0, i1, i2);
}

// Resident blocks of `kernel` per compute unit at `threads` per block.
inline int BlocksPerCU(const void *kernel, int threads) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) !=
          hipSuccess ||
      n < 1) {
    (void)hipGetLastError();
    n = 1;
  }
  return n;
}

// Buffer windows (stencil.h: BufWin) address at most 2^31 - 1 bytes, the
// reach of a 32-bit buffer range.  The largest z chunk for which a window of
// grid g over the chunk plus `margin` planes on each side stays within that
// reach; a grid of which not even one such window fits is rejected (the
// reference's 32-bit PSIndex, include/physis/physis_common.h:33-55, bounds
// elements, not bytes).
inline int WindowChunkLimit(const __PSGrid *g, int margin) {
  if (!g) return INT32_MAX;
  int64_t plane = (int64_t)g->elm_size;
  for (int d = 0; d < g->num_dims && d < 2; ++d) plane *= g->dim[d];
  if (plane <= 0) return INT32_MAX;
  const int64_t planes = (int64_t)INT32_MAX / plane - 2 * (int64_t)margin;
  if (planes < 1) {
    std::fprintf(stderr,
                 "[physis-hip] a %lld-byte plane (with %d halo planes per "
                 "side) exceeds the 2 GiB reach of a buffer window\n",
                 (long long)plane, margin);
    std::exit(EXIT_FAILURE);
  }
  return planes > INT32_MAX ? INT32_MAX : (int)planes;
}
inline int WindowChunkLimit(std::initializer_list<const __PSGrid *> grids,
                            int margin) {
  int m = INT32_MAX;
  for (const __PSGrid *g : grids) {
    const int c = WindowChunkLimit(g, margin);
    m = c < m ? c : m;
  }
  return m;
}

// The same reach for a 2-D grid in the rows view (nx x 1 x ny): z chunks
// are chunks of rows.
inline int WindowChunkLimitRows(std::initializer_list<const __PSGrid *> grids,
                                int margin) {
  int m = INT32_MAX;
  for (const __PSGrid *g : grids) {
    if (!g) continue;
    __PSGrid row = *g;
    row.num_dims = 1;  // one row of dim[0] elements per "plane"
    const int c = WindowChunkLimit(&row, margin);
    m = c < m ? c : m;
  }
  return m;
}

// Tiles a domain into TX x TY column tiles and picks the number of z chunks.
// Each block costs (chunk + 2) plane steps (two prologue planes), blocks run
// in waves of `resident` blocks; the chunk count minimising
// waves * (chunk + 2) is chosen so that the last wave is not left mostly
// idle.  Chunks never exceed max_chunk planes (the window reach).
inline RunShape PlanRunFor(const __PSDomain &dom, int tx, int ty,
                           long resident, int max_chunk = INT32_MAX) {
  RunShape s = {0, 0, 0, 0, 0, 0};
  const int nx = dom.local_max[0] - dom.local_min[0];
  const int ny = dom.local_max[1] - dom.local_min[1];
  const int nz = dom.local_max[2] - dom.local_min[2];
  if (nx <= 0 || ny <= 0 || nz <= 0) return s;
  if (max_chunk < 1) max_chunk = 1;
  if (resident < 1) resident = 1;
  // x tiles are anchored at x = 0 so that vector loads stay aligned.
  const int xf = dom.local_min[0] >= 0 ? dom.local_min[0] / tx
                                       : -((-dom.local_min[0] + tx - 1) / tx);
  const int xl = (dom.local_max[0] - 1) >= 0
                     ? (dom.local_max[0] - 1) / tx
                     : -((-(dom.local_max[0] - 1) + tx - 1) / tx);
  s.x_first_tile = xf;
  s.tiles_x = xl - xf + 1;
  s.tiles_y = (ny + ty - 1) / ty;
  const long xy = (long)s.tiles_x * s.tiles_y;
  long best_cost = -1;
  int best_tz = 0;
  for (int tz = (nz + max_chunk - 1) / max_chunk; tz <= nz; ++tz) {
    const int chunk = (nz + tz - 1) / tz;
    if (chunk > max_chunk) continue;
    if ((long)(tz - 1) * chunk >= nz) continue;  // same chunk, fewer blocks
    const long blocks = xy * tz;
    if (blocks > (1L << 30)) break;
    const long waves = (blocks + resident - 1) / resident;
    const long cost = waves * (chunk + 2);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best_tz = tz;
    }
  }
  if (best_tz == 0) {
    std::fprintf(stderr, "[physis-hip] no launch shape fits a domain of "
                 "%d x %d x %d points\n", nx, ny, nz);
    std::exit(EXIT_FAILURE);
  }
  s.tiles_z = best_tz;
  s.zchunk = (nz + best_tz - 1) / best_tz;
  s.blocks = (int)(xy * best_tz);
  return s;
}

inline RunShape PlanRun(const __PSDomain &dom, int tx, int ty, int per_cu,
                        int max_chunk = INT32_MAX) {
  return PlanRunFor(dom, tx, ty, (long)__PSHipNumCUs() * per_cu, max_chunk);
}

// A 3-D grid argument must be live, of the stencil's element type, and small
// enough for 32-bit PSIndex offsets (the reference's default index type).
inline void CheckGrid3D(const __PSGrid *g, int type) {
  if (!g || !g->p) {
    std::fprintf(stderr, "[physis-hip] stencil map on a freed or NULL grid\n");
    std::exit(EXIT_FAILURE);
  }
  if (g->type != type) {
    std::fprintf(stderr,
                 "[physis-hip] grid element type %d does not match the "
                 "stencil's type %d\n", g->type, type);
    std::exit(EXIT_FAILURE);
  }
  if (g->num_elms >= (int64_t)PSINDEX_MAX) {
    std::fprintf(stderr,
                 "[physis-hip] grid of %lld elements exceeds 32-bit PSIndex\n",
                 (long long)g->num_elms);
    std::exit(EXIT_FAILURE);
  }
}

// Device-to-device copy of `bytes` on `stream` (scratch grids of two-sweep
// runs start as copies of the grid they stand in for); aborts on an error.
inline void CopyDevice(void *dst, const void *src, size_t bytes,
                       hipStream_t stream) {
  const hipError_t e =
      hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[physis-hip] device copy failed: %s\n",
                 hipGetErrorString(e));
    std::exit(EXIT_FAILURE);
  }
}

// The domain of a map must lie inside every grid the map emits to; the run
// kernel stores at every domain point.
// The device view of a runtime grid as a run kernel takes it.  The runtime
// keeps one form per grid, {int dim[3]; void *p}; the 1-D and 2-D views are
// {int dim[N]; T *p}, whose pointer sits at another offset, so views are
// built field by field rather than by reinterpreting that form.
template <typename View>
inline View DevView(const __PSGrid *g) {
  View v;
  constexpr int nd = (int)(sizeof(v.dim) / sizeof(v.dim[0]));
  if (!g || !g->dev || g->num_dims != nd) {
    std::fprintf(stderr, "[physis-hip] stencil map on a %s grid where a "
                 "%d-D grid is expected\n", g ? "mismatched" : "NULL", nd);
    std::exit(EXIT_FAILURE);
  }
  for (int i = 0; i < nd; ++i) v.dim[i] = g->dev->dim[i];
  v.p = static_cast<decltype(v.p)>(g->dev->p);
  return v;
}

// The domain of a whole 3-D grid (the tiling of a red-black pass, which
// writes every point of its output).
inline __PSDomain WholeDomain(const __PSGrid *g) {
  __PSDomain d;
  for (int i = 0; i < 3; ++i) {
    d.min[i] = d.local_min[i] = 0;
    d.max[i] = d.local_max[i] = g->dim[i];
  }
  return d;
}

// The rows view of a 2-D grid: the 3-D view nx x 1 x ny of the same
// memory (equal linear offsets), in which a 2-D stencil runs on the 2.5D
// path, marching y as z.
template <typename View>
inline View RowView(const __PSGrid *g) {
  View v;
  static_assert(sizeof(v.dim) / sizeof(v.dim[0]) == 3, "a 3-D view");
  if (!g || !g->dev || g->num_dims != 2) {
    std::fprintf(stderr, "[physis-hip] stencil map on a %s grid where a "
                 "2-D grid is expected\n", g ? "mismatched" : "NULL");
    std::exit(EXIT_FAILURE);
  }
  v.dim[0] = g->dev->dim[0];
  v.dim[1] = 1;
  v.dim[2] = g->dev->dim[1];
  v.p = static_cast<decltype(v.p)>(g->dev->p);
  return v;
}
// The domain of a 2-D map in the rows view: [x0, x1) x [0, 1) x [y0, y1).
inline __PSDomain RowDomain(const __PSDomain &d) {
  __PSDomain r = d;
  r.min[2] = d.min[1];
  r.max[2] = d.max[1];
  r.local_min[2] = d.local_min[1];
  r.local_max[2] = d.local_max[1];
  r.min[1] = r.local_min[1] = 0;
  r.max[1] = r.local_max[1] = 1;
  return r;
}

// The device view of one plane of a grid of user-defined points as a grid of
// its own: component `comp` (row-major; 0 for a scalar member) of member
// `member` (the runtime stores the points member by member, the components
// of an array member one after another), with the grid's extents.
template <typename View>
inline View MemberView(__PSGrid *g, int member, int comp) {
  View v;
  constexpr int nd = (int)(sizeof(v.dim) / sizeof(v.dim[0]));
  if (!g || !g->user || g->num_dims != nd) {
    std::fprintf(stderr, "[physis-hip] stencil map on a %s grid where a "
                 "%d-D grid of user-defined points is expected\n",
                 g ? "mismatched" : "NULL", nd);
    std::exit(EXIT_FAILURE);
  }
  for (int i = 0; i < nd; ++i) v.dim[i] = g->dim[i];
  v.p = static_cast<decltype(v.p)>(__PSGridUserMember(g, member)) +
        (long)comp * (long)g->num_elms;
  return v;
}

// Checks a grid of user-defined points whose member planes a 2.5D run
// kernel reads or writes as grids: 3-D, allocated, and within 32-bit
// indexing.
inline void CheckUserGrid3D(const __PSGrid *g) {
  if (!g || !g->p || !g->user || g->num_dims != 3) {
    std::fprintf(stderr, "[physis-hip] stencil map on a freed, NULL or "
                 "mismatched grid where a 3-D grid of user-defined points is "
                 "expected\n");
    std::exit(EXIT_FAILURE);
  }
  if (g->num_elms >= (int64_t)PSINDEX_MAX) {
    std::fprintf(stderr,
                 "[physis-hip] grid of %lld elements exceeds 32-bit PSIndex\n",
                 (long long)g->num_elms);
    std::exit(EXIT_FAILURE);
  }
}

inline void CheckEmitDomain(const __PSDomain &dom, const __PSGrid *g) {
  for (int d = 0; d < g->num_dims; ++d) {
    if (dom.local_max[d] <= dom.local_min[d]) return;  // empty domain
  }
  for (int d = 0; d < g->num_dims; ++d) {
    if (dom.local_min[d] < 0 || dom.local_max[d] > g->dim[d]) {
      std::fprintf(stderr,
                   "[physis-hip] domain [%d,%d) on axis %d exceeds the "
                   "emitted grid's extent %d\n",
                   dom.local_min[d], dom.local_max[d], d, g->dim[d]);
      std::exit(EXIT_FAILURE);
    }
  }
}

// Brackets a run function.  Launches stay asynchronous, as in the
// reference's run functions (cuda_runtime_builder.cc:1465-1507: they wait
// for the device only under CUDA_KERNEL_ERROR_CHECK, which the translator's
// kernel_error_check option mirrors with __PSCheckHipError); launch errors
// are checked here, faults inside a kernel surface at the next synchronising
// call (copy-out, reduction, free).  With tracing on (--physis-trace) the
// run is timed with HIP events, which waits for it.
class RunTimer {
 public:
  explicit RunTimer(hipStream_t stream) : stream_(stream) {
    if (__ps_trace) {
      (void)hipEventCreate(&start_);
      (void)hipEventCreate(&stop_);
      (void)hipEventRecord(start_, stream_);
    }
  }
  float Stop() {
    float ms = 0.0f;
    hipError_t e = hipGetLastError();
    if (__ps_trace) {
      (void)hipEventRecord(stop_, stream_);
      if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    }
    if (e != hipSuccess) {
      std::fprintf(stderr, "[physis-hip] stencil run failed: %s\n",
                   hipGetErrorString(e));
      std::exit(EXIT_FAILURE);
    }
    if (__ps_trace) {
      (void)hipEventElapsedTime(&ms, start_, stop_);
      (void)hipEventDestroy(start_);
      (void)hipEventDestroy(stop_);
    }
    return ms;
  }

 private:
  hipStream_t stream_;
  hipEvent_t start_ = nullptr, stop_ = nullptr;
};

// Scratch slots of the reducing maps of one run function: runtime scratch
// areas 1..7 (__PSHipScratch), one per map that reduces, in map order, so
// no two maps of a run share partials.
class ReduceSlots {
 public:
  int Next() {
    if (next_ > kLast) {
      std::fprintf(stderr,
                   "[physis-hip] a PSStencilRun holds more than %d "
                   "PSStencilMapReduce maps\n", kLast);
      std::exit(EXIT_FAILURE);
    }
    return next_++;
  }

 private:
  static constexpr int kLast = 7;
  int next_ = 1;
};

// Partial results of a PSStencilMapReduce map inside one run function.  The
// run kernel of execution i writes one partial per block into Slot(i);
// After(i) folds every completed batch of executions into the user's array
// (out[i] = reduction of execution i).  A NULL out makes it a no-op.
class MapReducer {
 public:
  MapReducer(void *out, int op, int type, int nblocks, int iter,
             ReduceSlots &slots)
      : out_((char *)out), op_(op), type_(type), nblocks_(nblocks),
        iter_(iter) {
    if (!out_ || iter <= 0 || nblocks <= 0) {
      out_ = nullptr;
      return;
    }
    elm_ = (type == PS_DOUBLE || type == PS_LONG) ? 8 : 4;
    batch_ = iter < kBatch ? iter : kBatch;
    base_ = (char *)__PSHipScratch(slots.Next(),
                                   (size_t)batch_ * nblocks_ * elm_);
  }
  void *Slot(int i) const {
    if (!out_) return nullptr;
    return base_ + (size_t)(i % batch_) * nblocks_ * elm_;
  }
  void After(int i) {
    if (!out_) return;
    if ((i + 1) % batch_ != 0 && i + 1 != iter_) return;
    const int first = i - (i % batch_);
    const int count = i + 1 - first;
    __PSHipReduceFold(type_, op_, base_, nblocks_, count,
                      out_ + (size_t)first * elm_);
  }

 private:
  static constexpr int kBatch = 256;
  char *out_;
  int op_, type_, nblocks_, iter_;
  int elm_ = 4, batch_ = 1;
  char *base_ = nullptr;
};

}  // namespace physis_hip

#endif  // PHYSIS_HIP_RUN_H_
