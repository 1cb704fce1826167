/*
 * sg_adjoint.h -- a differentiable two-layer leaky network on one MI355X:
 * forward pass and hand-written adjoint (backpropagation through time)
 * with checkpoints.  No counterpart in the reference.
 *
 * Batch B, T steps, layers n_in -> n_hid -> n_out, fully connected, BF16
 * weights W1 [n_hid][n_in] and W2 [n_out][n_hid], FP32 state:
 *
 *     v1[t] = beta v1[t-1] - reset theta s1[t-1] + W1 x[t]
 *     s1[t] = S(v1[t] - theta)
 *     v2[t] = beta v2[t-1] - reset theta s2[t-1] + W2 s1[t]
 *     s2[t] = S(v2[t] - theta)
 *     count = sum over t of s2[t];   loss = mean over the batch of
 *             cross-entropy(softmax(count), label)
 *
 * Mode SG_SNN_SPIKE: S(u) = (u >= 0), with the fast-sigmoid surrogate
 * dS/du = 1 / (1 + slope |u|)^2 in the adjoint.  Mode SG_SNN_RATE: the
 * leaky rate unit S(u) = (1 + slope u / (1 + slope |u|)) / 2, whose exact
 * derivative slope / (2 (1 + slope |u|)^2) the adjoint uses, so its
 * gradient is the true gradient (checkable by finite differences).
 * `reset` scales the reset by subtraction: 1 for spiking units, 0 for
 * rate units (with reset, a rate unit's map v[t-1] -> v[t] has slope
 * beta - theta S' down to about -11 at slope 25, and the dynamics turn
 * chaotic).
 *
 * Inputs x[t] are 0/1 (rate-coded spikes) in BF16, [T][B][n_in].  The
 * input currents W1 x of all steps are one MFMA GEMM (BF16 operands, FP32
 * accumulation), kept in the workspace; the recurrences run per batch row;
 * dW1 is one MFMA GEMM over every step's adjoint states, which enter as
 * two BF16 terms (hi + mid: 16 significant bits; where n_hid is not a
 * multiple of 4, three); against the FP64 model its error is that of the
 * FP32 recurrences (dW1 relative L2 2.5e-7 in spike mode, 3.1e-6 in rate
 * mode on a 784 -> 1024 -> 10 network, batch 128, 100 steps).
 * The forward keeps the recurrent state only at every `checkpoint`-th
 * step; the backward recomputes layer 1's recurrence of each block of
 * steps from its checkpoint (the input currents, and layer 2's potentials
 * of every step, T * B * n_out floats, are kept by the forward).  The
 * workspace holds 2 * T * B * n_hid floats for the currents and the
 * adjoint states besides the checkpoints; T * B is at most 4,194,240.
 *
 * All pointers are device pointers; `stream` is a hipStream_t (NULL: the
 * default stream).  Calls are asynchronous and return 0, or an error code
 * with sg_snn_error() describing it.  There is no CPU fallback.
 */
#ifndef SG_ADJOINT_H
#define SG_ADJOINT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SG_SNN_SPIKE = 0, SG_SNN_RATE = 1 };

typedef struct sg_snn_cfg {
    int batch, steps, n_in, n_hid, n_out;
    int checkpoint;          /* steps per checkpoint; divides `steps`     */
    int mode;                /* SG_SNN_SPIKE or SG_SNN_RATE               */
    float beta, theta, slope;
    float reset;             /* 1: reset by subtraction; 0: none          */
} sg_snn_cfg;

/* Device workspace the forward fills (checkpoints, output counts, loss
 * gradient) and the backward reuses; `bytes` receives its size. */
int sg_snn_workspace_bytes(const sg_snn_cfg *cfg, size_t *bytes);

/* Forward: x [T][B][n_in] BF16 (0/1), w1 [n_hid][n_in] BF16,
 * w2 [n_out][n_hid] BF16, labels [B] int32 in [0, n_out).  Writes the
 * mean loss (1 float) and counts [B][n_out]. */
int sg_snn_forward(const sg_snn_cfg *cfg, const uint16_t *x,
                   const uint16_t *w1, const uint16_t *w2,
                   const int32_t *labels, void *workspace, float *loss,
                   float *counts, void *stream);

/* Backward after sg_snn_forward with the same arguments and workspace:
 * writes dloss/dW1 [n_hid][n_in] and dloss/dW2 [n_out][n_hid] in FP32,
 * each scaled by `grad_scale` (the incoming gradient of the loss). */
int sg_snn_backward(const sg_snn_cfg *cfg, const uint16_t *x,
                    const uint16_t *w1, const uint16_t *w2,
                    const int32_t *labels, void *workspace, float grad_scale,
                    float *dw1, float *dw2, void *stream);

const char *sg_snn_error(void);

#ifdef __cplusplus
}
#endif

#endif
