/* Reduction operators accepted by PSReduce.
 * Same names and ordinal values as the reference enum PSReduceOp
 * (include/physis/reduce.h:16-21): generated code and runtime pass the value
 * through unchanged. */
#ifndef PHYSIS_REDUCE_H_
#define PHYSIS_REDUCE_H_

#ifdef __cplusplus
extern "C" {
#endif

enum PSReduceOp {
  PS_MAX = 0,
  PS_MIN = 1,
  PS_SUM = 2,
  PS_PROD = 3
};

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_REDUCE_H_ */
