/* Trace hooks that generated run functions call around a PSStencilRun
 * (reference: include/physis/runtime.h:16-30). __ps_trace is set by PSInit
 * when the program is started with --physis-trace. */
#ifndef PHYSIS_RUNTIME_H_
#define PHYSIS_RUNTIME_H_

#include <stdio.h>
#include "physis/stopwatch.h"

#define __PS_PERIODIC(x, y) (((x) + (y)) % (y))

#ifdef __cplusplus
extern "C" {
#endif

extern FILE *__ps_trace;

static inline void __PSTraceStencilPre(const char *msg) {
  if (__ps_trace) fprintf(__ps_trace, "Physis: Stencil started (%s)\n", msg);
}

static inline void __PSTraceStencilPost(float time_ms) {
  if (__ps_trace)
    fprintf(__ps_trace, "Physis: Stencil finished (time: %f)\n", time_ms);
}

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_RUNTIME_H_ */
