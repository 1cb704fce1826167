/*
 * sg_ensemble.h -- batched rate-encoded ensembles on one MI355X.
 *
 * Many independent rate networks of uneven width, each an input layer
 * (width w0) feeding two fully connected rate layers (w0 -> w1 -> w2), are
 * stepped together: every step is ONE kernel launch over every row of every
 * instance, and a run of many steps is captured once into a HIP graph and
 * replayed as a single graph launch.
 *
 * Per step, with the parallel cluster's one-step lag (every layer reads the
 * outputs of the previous step, engine/cluster/parallel_cluster.cpp:107-131):
 *
 *     h1 <- (1 - leak) * h1 + leak * f(W1 . x0)
 *     h2 <- (1 - leak) * h2 + leak * f(W2 . h1_previous)
 *
 * f = tanh, or the reference's rectified tanh (input > 0 ? tanh : 0) when
 * `rectify` is set.  With leak 1 and rectify this is the reference's
 * rate_encoding model (state/impl/rate_encoding_attributes.cpp:14-23) fed by
 * FULLY_CONNECTED base activators (engine/kernel/synapse_kernel.h:113-149,
 * 887-912); the leak is this library's extension and has no counterpart in
 * the reference.
 *
 * Precision SG_ENS_BF16: BF16 weights, FP32 state and accumulation.  The
 * sums W.x run on the matrix cores (MFMA 16x16x32 BF16); the FP32 input
 * vector enters as three BF16 terms (hi + mid + lo, together exact to
 * FP32), so the only rounding beyond FP32 is the weights' own BF16 storage.
 * Precision SG_ENS_FP32: FP32 weights, the plain vector-ALU path.
 *
 * Synthetic instance data come from counter-based Philox4x32-10 with key
 * (seed, workload) and counter (0, index, stream, 0), stream 4n + L for
 * instance n: x0 (L = 0) uniform [0,1); W_L (L = 1, 2) normal with mean 0
 * and standard deviation 1/sqrt(fan_in), element index row * fan_in + col,
 * rounded to BF16 (nearest even) for SG_ENS_BF16.  h1 and h2 start at 0.
 *
 * All functions return 0 on success; on failure a non-zero code and
 * sg_ens_error() describes it.  There is no CPU fallback.
 */
#ifndef SG_ENSEMBLE_H
#define SG_ENSEMBLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SG_ENS_BF16 = 0, SG_ENS_FP32 = 1 };

typedef struct sg_ens sg_ens;

/* Allocate and initialise an ensemble of n instances on the current HIP
 * device; widths holds w0, w1, w2 per instance (3n ints, each >= 1).
 * Returns NULL on failure. */
sg_ens *sg_ens_create(int n, const int *widths, int precision, float leak,
                      int rectify, uint64_t seed, uint32_t workload);
void sg_ens_destroy(sg_ens *e);

/* Launch shape: row_tiles 16-row tiles per block, each split over
 * k_split wavefronts along the fan-in (BF16: row_tiles * k_split in
 * {1,2,4,8,16} wavefronts; FP32: k_split must be 1, row_tiles 4-row
 * groups).  Changing it drops any captured graph.  It first waits for
 * every run still pending (on the ensemble's stream and on the stream of
 * the last sg_ens_run), whose kernels read the work list it rewrites. */
int sg_ens_set_tile(sg_ens *e, int row_tiles, int k_split);
int sg_ens_get_tile(const sg_ens *e, int *row_tiles, int *k_split);

/* Time every supported launch shape over `steps` steps (HIP events on the
 * ensemble's stream) and keep the fastest; the state is restored after
 * timing.  best_ms receives the fastest average step time. */
int sg_ens_autotune(sg_ens *e, int steps, int *row_tiles, int *k_split,
                    float *best_ms);

/* Run `steps` steps on `stream` (a hipStream_t, NULL = the ensemble's own
 * stream).  use_graph: capture the steps into a HIP graph (cached per step
 * count and launch shape) and launch it once.  Asynchronous. */
int sg_ens_run(sg_ens *e, int steps, int use_graph, void *stream);

/* Run `steps` steps in ONE persistent launch (BF16 ensembles): a grid of
 * as many blocks as stay resident takes each step's work items from an
 * atomic queue and meets at a grid barrier between steps.  The launch is
 * cooperative (the runtime keeps every block resident or refuses it).
 * use_graph: the launch is the one kernel node, with the cooperative
 * attribute, of a HIP graph (cached per step count and parity) that is
 * then launched; where the runtime has no such node it is launched
 * directly.  Same results as sg_ens_run (each row's sum is computed by the
 * same code, in the same order).  Asynchronous. */
int sg_ens_run_persistent(sg_ens *e, int steps, int use_graph, void *stream);

/* Grid of the last persistent launch, whether it ran from a graph, and the
 * barrier waits that gave up so far (0 unless a block was never resident;
 * the results of such a run are wrong).  Reading the count waits for the
 * runs in flight. */
int sg_ens_persistent_info(sg_ens *e, int *blocks, int *captured,
                           int *barrier_timeouts);

/* Wait for everything queued on the ensemble's own stream and on the
 * stream of the last sg_ens_run. */
int sg_ens_synchronize(sg_ens *e);

/* Average duration of one step launch over `steps` launches, measured with
 * HIP events around the launches on the ensemble's stream. */
int sg_ens_time_steps(sg_ens *e, int steps, float *avg_ms);

/* Host copies (synchronous).  layer 0 = x0, 1 = h1, 2 = h2 (current
 * values, `width` floats).  Weights: layer 1 or 2, row-major
 * [fan_out][fan_in] FP32 values as the kernels use them. */
int sg_ens_read_state(sg_ens *e, int inst, int layer, float *host);
int sg_ens_write_state(sg_ens *e, int inst, int layer, const float *host);
int sg_ens_read_weights(sg_ens *e, int inst, int layer, float *host);
/* Values are rounded to BF16 (nearest even) for SG_ENS_BF16. */
int sg_ens_write_weights(sg_ens *e, int inst, int layer, const float *host);

/* Per-step accounting: synapses = sum of w0*w1 + w1*w2; neurons = sum of
 * w1 + w2; algorithmic HBM bytes = weight bytes (2 or 4 per synapse,
 * unpadded, each read once) + 16 B per neuron of w0, w1 and w2 (state
 * read and written, input read, output written once each). */
int sg_ens_counts(const sg_ens *e, long long *synapses, long long *neurons,
                  long long *algorithmic_bytes, long long *stored_bytes);

const char *sg_ens_error(void);

#ifdef __cplusplus
}
#endif

#endif
