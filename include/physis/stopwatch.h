/* Wall-clock stopwatch used by generated run functions when kernel tracing
 * is on (reference: include/physis/stopwatch.h:25-46). Milliseconds. */
#ifndef PHYSIS_STOPWATCH_H_
#define PHYSIS_STOPWATCH_H_

#include <time.h>

typedef struct {
  struct timespec ts;
} __PSStopwatch;

static inline void __PSStopwatchStart(__PSStopwatch *w) {
  clock_gettime(CLOCK_MONOTONIC, &w->ts);
}

static inline float __PSStopwatchStop(__PSStopwatch *w) {
  struct timespec now;
  clock_gettime(CLOCK_MONOTONIC, &now);
  return (float)((now.tv_sec - w->ts.tv_sec) * 1.0e3 +
                 (now.tv_nsec - w->ts.tv_nsec) * 1.0e-6);
}

#endif /* PHYSIS_STOPWATCH_H_ */
