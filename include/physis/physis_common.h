/* Declarations shared by Physis user programs and the code physisc-hip
 * generates from them.
 *
 * The names, layouts and values here are the Physis DSL surface; they have to
 * stay identical to the reference so that unchanged programs compile
 * (reference: include/physis/physis_common.h:40-168). */
#ifndef PHYSIS_PHYSIS_COMMON_H_
#define PHYSIS_PHYSIS_COMMON_H_

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <assert.h>

#include "physis/types.h"
#include "physis/reduce.h"
#include "physis/stopwatch.h"
#include "physis/runtime.h"

#define PS_MAX_DIM (3)

#define PSAssert(e) assert(e)

#ifdef __cplusplus
extern "C" {
#endif

/* Grid index type: 32-bit by default, as in the reference
 * (physis_common.h:33-55). */
typedef int32_t PSIndex;
#define PSINDEX_MAX INT32_MAX
#define PSINDEX_MIN INT32_MIN

typedef int PSVectorInt[PS_MAX_DIM];
typedef PSVectorInt PSPoint;

static inline void PSVectorIntInit(PSVectorInt vec, int x) {
  for (int i = 0; i < PS_MAX_DIM; ++i) vec[i] = x;
}
static inline void PSVectorIntCopy(PSVectorInt dst, const PSVectorInt src) {
  memcpy(dst, src, sizeof(PSVectorInt));
}

/* Iteration domain of a stencil map: [min, max) per axis; local_* equal the
 * global bounds on one GPU (physis_common.h:88-96). */
typedef struct {
  PSIndex min[PS_MAX_DIM];
  PSIndex max[PS_MAX_DIM];
  PSIndex local_min[PS_MAX_DIM];
  PSIndex local_max[PS_MAX_DIM];
} __PSDomain;
typedef __PSDomain PSDomain1D;
typedef __PSDomain PSDomain2D;
typedef __PSDomain PSDomain3D;

extern void PSInit(int *argc, char ***argv, int grid_num_dims, ...);
extern void PSFinalize(void);

extern PSDomain1D PSDomain1DNew(PSIndex minx, PSIndex maxx);
extern PSDomain2D PSDomain2DNew(PSIndex minx, PSIndex maxx,
                                PSIndex miny, PSIndex maxy);
extern PSDomain3D PSDomain3DNew(PSIndex minx, PSIndex maxx,
                                PSIndex miny, PSIndex maxy,
                                PSIndex minz, PSIndex maxz);

static inline void PSAbort(int code) { exit(code); }

enum PS_GRID_ATTRIBUTE { PS_GRID_ATTRIBUTE_DUMMY = 1 << 0 };

#define PS_GRID_USER_TYPE_MAX_ARRAY_RANK (5)
typedef struct {
  PSType type;
  int size;
  int rank;
  int dim[PS_GRID_USER_TYPE_MAX_ARRAY_RANK];
} __PSGridTypeMemberInfo;

/* Element type descriptor passed to __PSGridNew (physis_common.h:155-166). */
typedef struct {
  PSType type;
  int size;
  int num_members;
  __PSGridTypeMemberInfo *members;
} __PSGridTypeInfo;

#define INVALID_GRID (NULL)

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_PHYSIS_COMMON_H_ */
