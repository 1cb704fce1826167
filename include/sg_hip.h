/*
 * sg_hip.h -- the thin kernel C-ABI of the MI355X simulation core.
 *
 * Everything here is plain C: POD descriptors, raw device pointers, an
 * opaque stream (a hipStream_t, or NULL for the default stream) and int
 * error codes (0 == hipSuccess).  libsyngen.so (the C++ host layer behind
 * the reference-compatible C-ABI in syngen.h) drives the simulation through
 * these entry points, and Python loads the same library with ctypes for
 * kernel-level parity tests.
 *
 * Each entry point replaces one family of the reference's per-timestep
 * kernels; the reference file:line it restates is cited beside it.  Arrays
 * called "device layout" hold per-synapse data in 64-destination tiles:
 * element (slot s, visit k) of a connection with K visits per slot lives at
 * ((s / 64) * K + k) * 64 + s % 64, so the 64 lanes of a wavefront (one
 * destination each) read 256 consecutive bytes per visit (see
 * sg_hip_tile32).  Convolutional kernels stay a plain [k] array; one-to-one
 * connections keep one value per destination.
 */
#ifndef SG_HIP_H
#define SG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Connection types (util/constants.h:71-78).  SG_SPARSE is what a
 * connection with "sparse": true becomes once its zero weights are dropped
 * (state/weight_matrix.cpp:510-663, network/connection.cpp:219-222). */
enum { SG_FULLY_CONNECTED = 0, SG_SUBSET = 1, SG_ONE_TO_ONE = 2,
       SG_CONVERGENT = 3, SG_DIVERGENT = 4, SG_SPARSE = 5 };

/* Source output types (util/constants.h:15-19). */
enum { SG_OUT_FLOAT = 0, SG_OUT_BIT = 1, SG_OUT_INT = 2 };

/* Opcodes (util/constants.h:109-121). */
enum { SG_ADD = 0, SG_SUB = 1, SG_MULT = 2, SG_DIV = 3, SG_POOL = 4,
       SG_REWARD = 5, SG_MODULATE = 6, SG_GAP = 7, SG_ADD_HEAVISIDE = 8,
       SG_SUB_HEAVISIDE = 9, SG_MULT_HEAVISIDE = 10 };

/* One connection as the kernels see it (engine/kernel/synapse_data.h:16-57
 * plus the ArborizedConfig / SubsetConfig it carries). */
typedef struct sg_conn {
    int type;            /* SG_FULLY_CONNECTED .. SG_SPARSE               */
    int opcode;          /* aggregator for the base activator             */
    int out_type;        /* SG_OUT_* of the source layer                  */
    int delay;           /* connection delay                              */
    int convolutional;   /* shared kernel of row/column field size        */
    int from_size, from_rows, from_cols;
    int to_size, to_rows, to_cols;
    /* arborized config (network/connection_config.cpp:8-45) */
    int row_field, col_field, row_stride, col_stride;
    int row_spacing, col_spacing, row_offset, col_offset, wrap;
    /* subset config (network/connection_config.cpp:112-134) */
    int from_row_start, from_row_end, from_col_start, from_col_end;
    int to_row_start, to_row_end, to_col_start, to_col_end;
    /* weights per destination slot in device layout, number of slots */
    int k_per_dest;
    int num_dest;
    int num_weights;
    /* Base activator with a BIT source and only finite weights: a visit
     * whose bit is 0 adds +0 or -0, which leaves the sum unchanged (it
     * starts at +0 and can never become -0), so its weight is not loaded. */
    int skip_silent;
    /* source output at word delay/32 (BIT) or register delay (FLOAT/INT) */
    const uint32_t *src;
    /* destination input register */
    float *dst;
    /* weights, device layout */
    const float *weights;
    /* Scratch for a BIT source's spike plane (ceil(from_size / 64) words),
     * or NULL.  With skip_silent set and a tiled connection type,
     * sg_hip_activate_base packs the bit each source shows this connection
     * into it (64 sources per word) and loads the weights of the spiking
     * visits only, summing in visit order. */
    uint64_t *plane;
    /* The plane already holds this step's bits (written by the source
     * layer's neuron kernel): skip packing it. */
    int plane_ready;
    /* Store each destination's sum itself into dst instead of aggregating
     * it into the input; the neuron kernel then aggregates it into the
     * layer's init-config value (sg_iz_layer.noise == 2). */
    int store_sum;
    /* SG_SPARSE: source index of visit k of slot s in device layout
     * (k_per_dest = the largest nonzero count), -1 past the slot's count
     * (synapse_kernel.h:305-366 loops over nonzero_counts[to]). */
    const int32_t *sparse_from;
    /* SG_SPARSE, optional: per 64-destination tile, the fewest and the most
     * nonzero visits of its slots ({lo, hi} per tile); a tile's walk ends
     * at its hi, and visits below its lo need no check. */
    const int32_t *tile_counts;
    /* The launch runs beside a long kernel on another stream: a base
     * activator on the plane path then uses blocks of four wavefronts
     * (they start as soon as a CU has four free wave slots) instead of
     * sixteen. */
    int beside;
} sg_conn;

/* Fields of the interleaved per-synapse record of a tiled Izhikevich
 * connection.  Visits are grouped by eight and the groups are the outer
 * index: element (destination slot s, visit k, field f) lives at
 * ((((k / 8) * T + s / 64) * 8 + k % 8) * SG_IZ_FIELDS + f) * 64 + s % 64,
 * T = ceil(slots / 64), so one wavefront's eight visits are 14 KB of
 * contiguous memory and the tiles' current groups sit side by side (see
 * sg_hip_record32).  The record holds ceil(K / 8) * 8 * T * 64 *
 * SG_IZ_FIELDS floats. */
enum { SG_IZ_W = 0, SG_IZ_R = 1, SG_IZ_SW = 2, SG_IZ_PRE = 3, SG_IZ_DW = 4,
       SG_IZ_ST = 5, SG_IZ_LT = 6, SG_IZ_FIELDS = 7 };

/* The compact record: the same layout with four fields.  The presynaptic
 * trace and the short-term plasticity pair (R, w) of a synapse change only
 * in the updater, and there only from their own values and the source's
 * spike (izhikevich_attributes.cpp:444-508); the updater skips synapses of
 * weight below 1e-4 and clamps every weight it writes to at least 1e-4.  So
 * every synapse of weight >= 1e-4 from one source carries the same three
 * values if they were equal at the start, and they are kept once per source
 * (sg_iz_syn.src_stp).  A synapse of weight below 1e-4 never changes its
 * weight, derivative, trace or (R, w): its DW field holds its constant
 * product R * w instead, and the host keeps its other values. */
enum { SG_IZC_W = 0, SG_IZC_DW = 1, SG_IZC_ST = 2, SG_IZC_LT = 3,
       SG_IZC_FIELDS = 4 };

/* Per-synapse and per-source state of an Izhikevich conductance connection
 * (state/impl/izhikevich_attributes.h:45-66).  Tiled connection types use
 * `record` (interleaved, above); one-to-one connections use the separate
 * arrays. */
typedef struct sg_iz_syn {
    float *record;
    float *weights;
    float *short_traces, *long_traces, *presyn_traces;
    float *stp_r, *stp_w, *weight_derivatives;
    /* per-synapse delay & 31 (the bit of the source word the synapse
     * reads); tiled types: the record's visit-group-major order with one
     * field, (((k / 8) * T + s / 64) * 8 + k % 8) * 64 + s % 64 */
    const uint8_t *delays;
    /* time since spike per source after the last activation (read by the
     * updater) */
    const int32_t *tss_cur;
    /* NULL: `record` has SG_IZ_FIELDS fields.  Otherwise `record` is the
     * compact record (SG_IZC_FIELDS) and this holds, per source, {R, w,
     * presynaptic trace, time since spike (int32 bits)} after the source's
     * last update: the update reads the time since spike from here instead
     * of tss_cur. */
    const float *src_stp;
    float baseline_conductance;
    float learning_rate;
    float max_weight;
    int stp_flag;
} sg_iz_syn;

/* Destination-side data an Izhikevich updater reads, plus the conductance
 * arrays an activator accumulates into. */
typedef struct sg_iz_dest {
    const uint32_t *out_word0;      /* destination spike history word 0 */
    const float *post_trace;
    const int32_t *time_since_spike;
    float *short_cond;             /* ampa / gabaa / mult / dopamine    */
    float *long_cond;              /* nmda / gabab, or NULL             */
    float short_tau, long_tau;
    const float *voltage;          /* gap junctions                     */
    float *input;                  /* gap junctions                     */
} sg_iz_dest;

/* Izhikevich conductance activator kinds (izhikevich_attributes.cpp:288-366). */
enum { SG_IZ_ADD = 0, SG_IZ_SUB = 1, SG_IZ_MULT = 2, SG_IZ_REWARD = 3,
       SG_IZ_GAP = 4 };

/* Per-source time since spike of one conductance connection reading this
 * layer, stepped by the layer's neuron kernel from its history words before
 * the shift (the words that step's activation read); same rule and arrays
 * as sg_hip_tss_step. */
typedef struct sg_tss_out {
    int32_t *next;
    const int32_t *cur;
    const uint8_t *visits;
    const uint8_t *rev_delays;
    int word;                     /* history word read: delay / 32 */
    /* A compact record's per-source {R, w, presynaptic trace} to step with
     * the new time since spike this step (the connection learns), or
     * NULL; stp_sub selects update_iz_sub's constants, stp_flag the
     * connection's short-term plasticity (as sg_hip_iz_source_update). */
    float *stp;
    int stp_sub, stp_flag;
} sg_tss_out;

enum { SG_MAX_TSS_OUT = 2 };

/* One Izhikevich layer (izhikevich_attributes.h:6-43). */
typedef struct sg_iz_layer {
    int size, history;
    float *input;
    uint32_t *output;             /* [history][size] */
    float *ampa, *nmda, *gabaa, *gabab, *mult, *dopamine;
    float *voltage, *recovery, *post_trace;
    int32_t *time_since_spike;
    const float *a, *b, *c, *d;
    /* Optional: plane of the new spike bits (ceil(size / 64) words), for
     * base activators that read this layer at delay 0 next step. */
    uint64_t *plane;
    /* Optional: time-since-spike steps of outgoing conductance
     * connections. */
    int ntss;
    sg_tss_out tss[SG_MAX_TSS_OUT];
    /* Init-config value drawn here instead of by sg_hip_init_random
     * (instruction.cpp:3-30 order: the value is written first, then the
     * activators aggregate into it): noise 0 = none; 1 = the input is the
     * drawn value (no base activator writes it); 2 = the input holds the
     * sum of the one base activator that covers every neuron, stored raw
     * (sg_conn.store_sum), and the input becomes
     * aggregate(noise_op, value, sum) -- the reference's own operation.
     * The result is written back to input.  dist: 0 uniform[a, b),
     * 1 normal(a, b), 2 poisson (val with probability a), 3 flat a. */
    int noise, noise_dist, noise_op;
    float noise_a, noise_b, noise_val;
    uint64_t noise_seed, noise_step;
    /* Optional: the draw's step is *noise_base + noise_step (a device
     * counter, so that a captured HIP graph draws new values on every
     * replay); NULL: noise_step itself. */
    const uint64_t *noise_base;
} sg_iz_layer;

/* Base activator: every connection type, extractor and aggregator
 * (engine/kernel/synapse_kernel.h:47-912, kernel.cpp:217,
 * extractor.cpp:8-15, aggregator.cpp:15-48). */
int sg_hip_activate_base(const sg_conn *c, void *stream);

/* The base activator for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  No counterpart in the reference.  Same kernels and results
 * as sg_hip_activate_base; the plane path's spike plane (sg_conn.plane is
 * ignored) lives in a fixed device array of the library,
 * SG_NOALLOC_PLANE_WORDS 64-bit words (512 KiB, sources up to 4,194,304),
 * one copy per GPU, held until the process ends.  A connection with more
 * sources on the plane path returns SG_HIP_ERROR_NOALLOC_LIMIT before any
 * launch.  Calls share their GPU's array: calls that overlap on one GPU
 * (other streams, other threads) race on it.  Allocates nothing. */
enum { SG_NOALLOC_PLANE_WORDS = 65536, SG_HIP_ERROR_NOALLOC_LIMIT = 20001 };
int sg_hip_activate_base_noalloc(const sg_conn *c, void *stream);

/* activate_base_second_order (synapse_kernel.h:918-935, kernel.cpp:220): a
 * second-order slave connection folds each synapse's activity into the
 * host's weight copy, sow[w] = aggregate(opcode, sow[w], extract(source) *
 * weights[w]); sow has the slave's (= the host's) device layout.  A
 * convolutional slave (its field the whole source layer, stride 0) folds
 * one-to-one over its source layer, weight i with source i
 * (kernel.cpp:223-229, synapse_kernel.h:943-970). */
int sg_hip_activate_second_order(const sg_conn *slave, float *sow,
                                 void *stream);

/* Time since spike of every source after one activation of an Izhikevich
 * conductance connection (izhikevich_attributes.cpp:258-262, applied once
 * per visit in the reference's serial order).  A source visited m times
 * ends at min(32, m - j) if its j-th (1-based) visit is the last one that
 * sees a spike, else at min(32, tss + m); only the last 32 visits can
 * matter.  rev_delays[f * 32 + i] is the delay (& 31) of the i-th visit of
 * source f counted back from its last one, visits[f] = min(32, m). */
int sg_hip_tss_step(int32_t *tss_next, const int32_t *tss_cur,
                    const uint8_t *visits, const uint8_t *rev_delays,
                    const uint32_t *src_words, int n, void *stream);

/* Izhikevich conductance / gap activators
 * (izhikevich_attributes.cpp:219-366). */
int sg_hip_iz_activate(const sg_conn *c, const sg_iz_syn *s,
                       const sg_iz_dest *d, int kind, void *stream);

/* Izhikevich STDP/iSTDP + short-term-plasticity updaters
 * (izhikevich_attributes.cpp:404-520).  sub != 0 selects update_iz_sub.
 * src_table / dst_table hold the iSTDP factor for time-since-spike 0..31. */
int sg_hip_iz_update(const sg_conn *c, const sg_iz_syn *s,
                     const sg_iz_dest *d, int sub, const float *src_table,
                     const float *dst_table, void *stream);

/* The updater of step t-1 fused with the conductance activator of step t
 * over one pass of the per-synapse arrays (same results as running
 * sg_hip_iz_update then sg_hip_iz_activate). */
int sg_hip_iz_update_activate(const sg_conn *c, const sg_iz_syn *s,
                              const sg_iz_dest *d, int kind, int sub,
                              const float *src_table, const float *dst_table,
                              void *stream);

/* One updater step of a compact record's per-source state: for every
 * source f with src_spike = (tss[f] == 0), the presynaptic trace (update_iz_add
 * only) and, with stp_flag, R and w, exactly as update_iz_add / update_iz_sub
 * step them for each synapse (izhikevich_attributes.cpp:444-508), and
 * stores tss[f] beside them.  stp holds 4 values per source
 * (sg_iz_syn.src_stp); tss is the time since spike after this step's
 * activation. */
int sg_hip_iz_source_update(float *stp, const int32_t *tss, int n, int sub,
                            int stp_flag, void *stream);

/* Neuron models (izhikevich_attributes.cpp:102-213,
 * game_of_life_attributes.cpp:14-33, rate_encoding_attributes.cpp:14-23,
 * relay_attributes.cpp:10-18) including the history shift
 * (state/attributes.h:316-333). */
int sg_hip_iz_neuron(const sg_iz_layer *l, void *stream);
/* Every layer of a list in one launch (up to SG_MAX_IZ_BATCH layers per
 * launch; longer lists are split). */
enum { SG_MAX_IZ_BATCH = 8 };
int sg_hip_iz_neurons(const sg_iz_layer *layers, int n, void *stream);

/* Spike plane of a BIT layer for one delay: bit i of plane[i / 64] is bit
 * 31 - delay % 32 of words[i] (words = history word delay / 32). */
int sg_hip_pack_plane(const uint32_t *words, int n, int delay,
                      uint64_t *plane, void *stream);
int sg_hip_gol_neuron(float *input, uint32_t *output, int size, int history,
                      int survival_min, int survival_max, int birth_min,
                      int birth_max, void *stream);
int sg_hip_rate_neuron(float *input, float *output, int size, int history,
                       void *stream);
int sg_hip_relay_neuron(float *input, float *output, int size, int history,
                        int ramp, void *stream);

/* Game of life fused step for the one-connection 3x3 convolutional case:
 * activation + rule + shift in one pass over double-buffered words. */
int sg_hip_gol_step(const uint32_t *words_in, uint32_t *words_out,
                    float *input, int rows, int cols, int wrap,
                    const float *kernel9, int survival_min, int survival_max,
                    int birth_min, int birth_max, void *stream);
/* The same for `steps` consecutive steps in one pass (1 to
 * sg_hip_gol_max_steps(kernel9, rows, cols): 8 for kernels of 0s and 1s,
 * else 4), for a layer
 * whose words nothing else reads between them: words_out holds the words
 * after the last step, `input` (if not NULL) that step's neighbour sums. */
int sg_hip_gol_max_steps(const float *kernel9, int rows, int cols);
int sg_hip_gol_steps(const uint32_t *words_in, uint32_t *words_out,
                     float *input, int rows, int cols, int wrap,
                     const float *kernel9, int survival_min, int survival_max,
                     int birth_min, int birth_max, int steps, void *stream);

/* Random init configs (engine/kernel/kernel.cpp:43-158 via
 * engine/instruction.h:208-282): dist 0 = uniform[a,b), 1 = normal(a, b),
 * 2 = poisson (fires `val` with probability a).  Counter-based Philox keyed
 * by (seed, step), so runs are reproducible.  step_base: NULL, or a device
 * counter added to `step` (as sg_iz_layer.noise_base). */
int sg_hip_init_random(float *p, int n, int dist, float a, float b, float val,
                       int overwrite, uint64_t seed, uint64_t step,
                       const uint64_t *step_base, void *stream);

/* *counter += add on the device (advances the step counter of a captured
 * run of steps at its end). */
int sg_hip_add_u64(uint64_t *counter, uint64_t add, void *stream);

/* set_data (engine/kernel/kernel.cpp:14-40). */
int sg_hip_set_f32(float *p, float val, long n, int overwrite, void *stream);

/* Layout changes: 32-bit element matrix transpose [rows][cols] ->
 * [cols][rows]. */
int sg_hip_transpose32(const void *src, void *dst, int rows, int cols,
                       void *stream);

/* Layout change between the reference's host layout [slots][K] and the
 * 64-destination tiled device layout, optionally interleaved with other
 * per-synapse arrays: element (s, k) of field f of nfields lives at
 * (((s / 64) * K + k) * nfields + f) * 64 + s % 64 (to_tiled != 0: host ->
 * tiled; else tiled -> host).  The tiled buffer holds
 * ceil(slots / 64) * 64 * K * nfields elements. */
int sg_hip_tile32(const void *src, void *dst, int slots, int K, int nfields,
                  int field, int to_tiled, void *stream);

/* Layout change between one [slots][K] host-layout array and field
 * `field` of an Izhikevich record of `nfields` fields (SG_IZ_FIELDS, or
 * SG_IZC_FIELDS for the compact record; layout above with nfields in place
 * of SG_IZ_FIELDS); to_record != 0: host -> record, else record -> host.
 * Lanes past `slots` are written as 0. */
int sg_hip_record32(const void *src, void *dst, int slots, int K, int field,
                    int nfields, int to_record, void *stream);

/* Plain copy of n 32-bit words (a known-size reference for the hardware
 * byte counters). */
int sg_hip_copy32(const void *src, void *dst, long n, void *stream);

/* Copy of n 32-bit words by a kernel on `stream`, any n, 4-byte aligned
 * pointers; either side may be pinned host memory mapped into the device's
 * address space (hipHostGetDevicePointer).  The engine's per-step
 * environment I/O: a kernel in the consuming stream's order instead of a
 * DMA copy on a stream of its own and the events between them. */
int sg_hip_copy_u32(const void *src, void *dst, long n, void *stream);

/* Device properties: number of visible devices, -1 on error. */
int sg_hip_device_count(void);

/* Last error string for a code. */
const char *sg_hip_error_string(int code);

/* calc_internal (engine/kernel/kernel.cpp:161-187): one dendritic node's
 * register folded into its parent's, dst[i] = aggregate(opcode, dst[i],
 * src[i]), and reset to the node's init value, src[i] = trail. */
int sg_hip_dendrite(float *dst, float *src, int n, int opcode, float trail,
                    void *stream);

#ifdef __cplusplus
}
#endif

#endif
