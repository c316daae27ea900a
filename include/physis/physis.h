/* Single include for Physis programs. physisc-hip parses programs with
 * PHYSIS_USER defined and compiles the generated code with PHYSIS_HIP
 * (reference dispatch: include/physis/physis.h). */
#ifndef PHYSIS_PHYSIS_H_
#define PHYSIS_PHYSIS_H_

#if defined(PHYSIS_USER)
#include "physis/physis_user.h"
#elif defined(PHYSIS_HIP)
#include "physis/physis_hip.h"
#else
#error "define PHYSIS_USER (DSL source) or PHYSIS_HIP (generated code)"
#endif

#endif /* PHYSIS_PHYSIS_H_ */
