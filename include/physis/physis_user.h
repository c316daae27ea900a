/* The Physis DSL as seen by user programs (compiled with -DPHYSIS_USER).
 *
 * Grid handles are pointers to structs whose function-pointer members give
 * each grid element type its own get/emit signatures, so that a plain C
 * compiler can type-check a stencil program before physisc-hip rewrites it.
 * Same intrinsic names, macros and argument order as the reference
 * (include/physis/physis_user.h:18-108). */
#ifndef PHYSIS_PHYSIS_USER_H_
#define PHYSIS_PHYSIS_USER_H_

#include "physis/physis_common.h"

#ifdef __cplusplus
extern "C" {
#endif

#define __PS_USER_GRID_MEMBERS(type, ...)                    \
  type _type_indicator;                                      \
  void (*set)(__VA_ARGS__, type);                            \
  type (*get)(__VA_ARGS__);                                  \
  type (*get_periodic)(__VA_ARGS__);                         \
  type (*emit)(type);                                        \
  type (*emitDirichlet)(type);                               \
  type (*emitNeumann)(type, int);                            \
  type (*reduce)(void *dom, type (*kernel)(type, type));

#define DeclareGrid1D(name, type)                                        \
  struct __PSGrid1D##name { __PS_USER_GRID_MEMBERS(type, PSIndex) };     \
  typedef struct __PSGrid1D##name *PSGrid1D##name;                       \
  extern PSGrid1D##name PSGrid1D##name##New(PSIndex, ...);

#define DeclareGrid2D(name, type)                                        \
  struct __PSGrid2D##name {                                              \
    __PS_USER_GRID_MEMBERS(type, PSIndex, PSIndex)                       \
  };                                                                     \
  typedef struct __PSGrid2D##name *PSGrid2D##name;                       \
  extern PSGrid2D##name PSGrid2D##name##New(PSIndex, PSIndex, ...);

#define DeclareGrid3D(name, type)                                        \
  struct __PSGrid3D##name {                                              \
    __PS_USER_GRID_MEMBERS(type, PSIndex, PSIndex, PSIndex)              \
  };                                                                     \
  typedef struct __PSGrid3D##name *PSGrid3D##name;                       \
  extern PSGrid3D##name PSGrid3D##name##New(PSIndex, PSIndex, PSIndex, ...);

DeclareGrid1D(Float, float)
DeclareGrid1D(Double, double)
DeclareGrid1D(Int, int)
DeclareGrid1D(Long, long)
DeclareGrid2D(Float, float)
DeclareGrid2D(Double, double)
DeclareGrid2D(Int, int)
DeclareGrid2D(Long, long)
DeclareGrid3D(Float, float)
DeclareGrid3D(Double, double)
DeclareGrid3D(Int, int)
DeclareGrid3D(Long, long)

extern void PSGridCopyin(void *g, const void *src_array);
extern void PSGridCopyout(void *g, void *dst_array);
extern void PSGridFree(void *g);
extern PSIndex PSGridDim(void *g, int d);

#define PSGridGet(g, ...) g->get(__VA_ARGS__)
#define PSGridGetPeriodic(g, ...) g->get_periodic(__VA_ARGS__)
#define PSGridSet(g, ...) g->set(__VA_ARGS__)
#define PSGridEmit(g, v) g->emit(v)
#define PSGridEmitDirichlet(g, v) g->emitDirichlet(v)
#define PSGridEmitNeumann(g, v, grad) g->emitNeumann(v, grad)

typedef int PSStencil;
#define PSStencilMap(k, ...) __PSStencilMap((void *)k, __VA_ARGS__)
extern PSStencil __PSStencilMap(void *, ...);
#define PSStencilMapRedBlack(k, ...) __PSStencilMapRedBlack((void *)k, __VA_ARGS__)
extern PSStencil __PSStencilMapRedBlack(void *, ...);
#define PSStencilMapRed(k, ...) __PSStencilMapRed((void *)k, __VA_ARGS__)
extern PSStencil __PSStencilMapRed(void *, ...);
#define PSStencilMapBlack(k, ...) __PSStencilMapBlack((void *)k, __VA_ARGS__)
extern PSStencil __PSStencilMapBlack(void *, ...);
extern void PSStencilRun(PSStencil, ...);

/* HIP target extension: a stencil map whose kernel returns a value per
 * point.  Each execution i of the map inside a PSStencilRun reduces those
 * values with op and stores the result in v[i] (v holds at least iter
 * elements of the kernel's return type).  It is the reduction-kernel form
 * documented, but never implemented, for PSReduce in the reference
 * (docs/programming.md:283-315), fused into the sweep. */
#define PSStencilMapReduce(v, op, k, ...) \
  __PSStencilMapReduce((void *)(v), op, (void *)k, __VA_ARGS__)
extern PSStencil __PSStencilMapReduce(void *v, enum PSReduceOp op, void *k,
                                      ...);

/* PSReduce(&v, op, grid), or the kernel form
 * PSReduce(&v, op, kernel, domain, args...) documented in the Physis
 * programming guide (docs/programming.md:283-315). */
extern void PSReduce(void *v, ...);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_PHYSIS_USER_H_ */
