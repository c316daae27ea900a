/* Learned warm start of the Poisson solve (libphysis_warmstart_hip): a
 * 3-level 3-D convolutional U-Net that predicts the solution p of the
 * 19-point operator of physis_poisson.h from its right-hand side, trained on
 * the GPU by hand-written forward and backward kernels (3x3x3 convolutions
 * on BF16 MFMA with FP32 accumulation).  Jacobi then runs from the
 * prediction to the requested residual, so it both falls back for and
 * measures the network's error.  No counterpart in the reference.
 *
 * Network (channels-last BF16 activations, FP32 master weights):
 *   in  (8 channels: rhs, bnd, 6 zero)
 *   E1  = relu(conv 8->8), relu(conv 8->8)            full resolution
 *   E2  = relu(conv 8->16), relu(conv 16->16)         after 4x4x4 average pool
 *   E3  = relu(conv 16->32), relu(conv 32->32)        after 4x4x4 average pool
 *   D2  = relu(conv [up4(E3), E2] 48->16), relu(conv 16->16)
 *   D1  = relu(conv [up4(D2), E1] 24->8), relu(conv 8->8)
 *   out = 1x1 conv 8->1 (FP32), times the output scale
 * Every 3x3x3 convolution has zero padding 1 and a bias; up4 repeats each
 * value over a 4x4x4 block.  Problem sizes must be multiples of 16 per axis.
 *
 * Parameters are one flat FP32 vector, per convolution W[cout][tap][cin]
 * (tap = 9 dz' + 3 dy' + dx', offsets d' - 1) followed by b[cout], then the
 * 1x1 weights w[8] and bias.  Inputs: FP32 grids of nbatch problems stacked
 * along z (the layout of physis_poisson.h).
 *
 * Threading: the runtime's stream, one host thread; calls are synchronous. */
#ifndef PHYSIS_PHYSIS_WARMSTART_H_
#define PHYSIS_PHYSIS_WARMSTART_H_

#include <stddef.h>
#include <stdint.h>

#include "physis/physis_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct __PSUNet __PSUNet;

/* A network with seeded He-normal weights and zero biases, able to run
 * batches of up to max_batch problems of nx x ny x nzs.  out_scale
 * multiplies the network output (the prediction is out_scale * out).
 * Returns NULL (and prints why) on bad sizes or when memory runs out. */
__PSUNet *PSUNetNew(int nx, int ny, int nzs, int max_batch, float out_scale,
                    uint64_t seed);
void PSUNetFree(__PSUNet *net);
size_t PSUNetNumParams(const __PSUNet *net);
/* FP32 master parameters to / from the host. */
void PSUNetGetParams(const __PSUNet *net, float *host);
void PSUNetSetParams(__PSUNet *net, const float *host);

/* Prediction for problems [first, first + count) of the FP32 grids rhs and
 * bnd (nbatch problems stacked along z, count <= max_batch), written to the
 * same problems of p (FP32 or FP64 grid) as out_scale * out * bnd, so that
 * bnd = 0 points (the Dirichlet data) are 0. */
void PSUNetPredict(__PSUNet *net, const __PSGrid *rhs, const __PSGrid *bnd,
                   int first, int count, __PSGrid *p);

/* Loss and gradient without an update, for problems [first, first + count)
 * with d = out - target / out_scale:
 *   mean over points of d^2
 *   + w4  * mean over 4x4x4 cells of (the cell's mean of d)^2
 *   + w16 * mean over 16x16x16 cells of (the cell's mean of d)^2
 * (w4 = w16 = 0 unless set by PSUNetSetLossWeights: the mean square error).
 * The two cell terms weight the smooth part of the error, which Jacobi
 * sweeps remove slowest.  grad_host (NumParams floats) may be NULL. */
double PSUNetLossGrad(__PSUNet *net, const __PSGrid *rhs, const __PSGrid *bnd,
                      const __PSGrid *target, int first, int count,
                      float *grad_host);

/* Weights w4, w16 >= 0 of the smooth-error terms of the loss. */
void PSUNetSetLossWeights(__PSUNet *net, float w4, float w16);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss; returns the
 * loss before the step. */
double PSUNetTrainStep(__PSUNet *net, const __PSGrid *rhs,
                       const __PSGrid *bnd, const __PSGrid *target, int first,
                       int count, float lr);

/* Training-set augmentation: problems [first, first + count) of the FP32
 * grid src (problems of nx x ny x nzs stacked along z) copied to problems
 * [dst_first, ...) of dst with the axes whose bit is set in flips (1 x,
 * 2 y, 4 z) reversed.  The reflection of a solved problem (right-hand side
 * and solution) is the solution of the reflected problem, whose
 * coefficients (+-noise about the same values) are as likely as the
 * original's.  src and dst must be different grids. */
void PSUNetReflect(const __PSGrid *src, int first, int count, __PSGrid *dst,
                   int dst_first, int nx, int ny, int nzs, int flips);

/* ---- single layers on host arrays (parity tests).  Tensors are
 * channels-last BF16 bit patterns [n][d][h][w][c]; w is W[cout][27][cin]
 * as BF16 bit patterns, bias FP32. */
void PSConv3dForwardHost(const uint16_t *x, const uint16_t *w,
                         const float *bias, int relu, int n, int d, int h,
                         int wd, int cin, int cout, uint16_t *y);
void PSConv3dBackwardDataHost(const uint16_t *dy, const uint16_t *w, int n,
                              int d, int h, int wd, int cin, int cout,
                              uint16_t *dx);
void PSConv3dBackwardFilterHost(const uint16_t *x, const uint16_t *dy, int n,
                                int d, int h, int wd, int cin, int cout,
                                float *dw, float *db);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIS_PHYSIS_WARMSTART_H_ */
