// Synapse kernels: the base activator for every connection type, the
// Izhikevich conductance activators, the STDP / iSTDP updaters and their
// fused update+activate pass.
//
// One lane per destination slot, walking that destination's synapses in the
// reference's serial order so every FP32 sum is accumulated in the same
// order as on the reference's CPU path.  Per-synapse arrays are k-major
// ("device layout"), so a wavefront's 64 lanes read 64 consecutive words for
// each k.  Compiled with -ffp-contract=off: no multiply-add is fused, every
// rounding matches the reference's separate float / double operations.
#include <hip/hip_runtime.h>

#include "sg_device.h"

namespace sg {

constexpr int kBlock = 256;

static inline int blocks_for(long n) {
    return (int)((n + kBlock - 1) / kBlock);
}

// Synapses are visited in batches of kBatch: the batch's loads are issued
// together (one coalesced wavefront load per array and k), then the
// synapses are consumed strictly in k order, so the FP32 sums keep the
// reference's sequence while each wavefront keeps dozens of loads in flight.
constexpr int kBatch = 8;


// ---------------------------------------------------------------------------
// Base activator (synapse_kernel.h:887-912): sum = 0; sum += extract * w in
// serial order; input = aggregate(input, sum).
template <int TYPE, int OUT>
__global__ __launch_bounds__(kBlock) void activate_base_kernel(sg_conn c) {
    constexpr int B = 16;
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.num_dest) return;
    SynapseIter<TYPE> it;
    it.init(c, slot);
    const int count = it.count;
    const int delay = c.delay;
    float sum = 0.0f;
    float wa[B], wb[B];
    uint32_t xa[B], xb[B];
    int fa[B], fb[B];
    auto load = [&](int k0, float *w, uint32_t *x, int *from) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            from[b] = it.k < count ? it.source() : -1;
            if (from[b] >= 0) {
                w[b] = c.weights[it.wi];
                x[b] = c.src[from[b]];
            }
            if (it.k < count) it.advance();
        }
    };
    auto consume = [&](const float *w, const uint32_t *x, const int *from) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (from[b] < 0) continue;
            const float val = extract<OUT>(x[b], delay) * w[b];
            sum += val;
        }
    };
    // Two batches in flight: load the next while consuming the current.
    load(0, wa, xa, fa);
    for (int k0 = 0; k0 < count; k0 += 2 * B) {
        if (k0 + B < count) load(k0 + B, wb, xb, fb);
        consume(wa, xa, fa);
        if (k0 + B >= count) break;
        if (k0 + 2 * B < count) load(k0 + 2 * B, wa, xa, fa);
        consume(wb, xb, fb);
    }
    const int to = slot_to_index(c, slot);
    put_sum(c, to, sum);
}

template <int TYPE>
static hipError_t launch_base_t(const sg_conn &c, hipStream_t s) {
    dim3 g(blocks_for(c.num_dest)), b(kBlock);
    switch (c.out_type) {
        case SG_OUT_BIT:
            hipLaunchKernelGGL((activate_base_kernel<TYPE, SG_OUT_BIT>), g, b,
                               0, s, c);
            break;
        case SG_OUT_INT:
            hipLaunchKernelGGL((activate_base_kernel<TYPE, SG_OUT_INT>), g, b,
                               0, s, c);
            break;
        default:
            hipLaunchKernelGGL((activate_base_kernel<TYPE, SG_OUT_FLOAT>), g,
                               b, 0, s, c);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Time since spike of each source (izhikevich_attributes.cpp:258-262).  The
// reference updates it at every visit, in serial order: spike ? 0 :
// min(32, tss + 1).  A source visited m times therefore ends at
// min(32, m - j) when its j-th visit (1-based) is the last one that sees a
// spike, and at min(32, tss + m) when none does; visits more than 32 from
// the end cannot change the result.  One lane per source scans its last
// visits backwards (their delays were laid out for it on the host) for the
// first one whose delay bit is set in the source's history word.
__global__ __launch_bounds__(kBlock) void tss_step_kernel(
        int32_t *next, const int32_t *cur, const uint8_t *visits,
        const uint8_t *rev_delays, const uint32_t *words, int n) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    next[f] = tss_next(cur[f], visits[f],
                       (const uint32_t *)(rev_delays + (long)f * 32),
                       words[f]);
}

// ---------------------------------------------------------------------------
// Izhikevich per-synapse state of one batch element.
struct IzSyn {
    float w, r, sw, pre, dw, st, lt;
    uint8_t dr;
    uint32_t word;
    int tss;
};

struct IzAct {
    float short_sum = 0.0f, long_sum = 0.0f;
};

struct IzDest {
    float dest_spike;       // MSB of the destination's history word 0
    float dest_exc_trace;
    int dest_tss;
};

// Conductance activator for one synapse (izhikevich_attributes.cpp:257-277):
//   weight = w * (r * stp_w)
//   trace' = trace + spike * weight * g ; sum += trace' ; trace = trace' * tau
template <int KIND>
__device__ __forceinline__ void iz_activate_one(const sg_iz_syn &s,
                                                const sg_iz_dest &d,
                                                IzAct &acc, IzSyn &y,
                                                int from) {
    const float spike = extract<SG_OUT_BIT>(y.word, y.dr);
    const float weight = y.w * (y.r * y.sw);
    const float inc = spike * weight * s.baseline_conductance;
    y.st = y.st + inc;
    acc.short_sum += y.st;
    y.st = y.st * d.short_tau;
    if (KIND == SG_IZ_ADD || KIND == SG_IZ_SUB) {
        y.lt = y.lt + inc;
        acc.long_sum += y.lt;
        y.lt = y.lt * d.long_tau;
    }
}

// Updater for one synapse (izhikevich_attributes.cpp:444-508); only called
// when w >= MIN_WEIGHT.
template <int SUB>
__device__ __forceinline__ void iz_update_one(const sg_iz_syn &s,
                                              const IzDest &dd,
                                              const float *src_table,
                                              const float *dst_table,
                                              IzSyn &y) {
    const int src_tss = y.tss;
    const bool src_spike = src_tss == 0;
    float src_trace, dest_trace;
    if (!SUB) {
        // STDP_ADD: presyn = spike ? 0.004 : presyn * 0.933 (double math)
        y.pre = src_spike ? (float)0.004 : (float)((double)y.pre * 0.933);
        src_trace = y.pre;
        dest_trace = dd.dest_exc_trace;
    } else {
        // STDP_SUB: the iSTDP factors depend only on the integer times since
        // spike; they are tabulated on the host with the reference's own
        // powf / expf expressions.
        src_trace = (dd.dest_spike > 0.0f && src_tss > 0 && src_tss < 32)
            ? src_table[src_tss] : 0.0f;
        dest_trace = (src_spike && dd.dest_tss > 0 && dd.dest_tss < 32)
            ? dst_table[dd.dest_tss] : 0.0f;
    }
    // weight_derivatives -= (wd - C_A) * C_TAU + dest_spike * src_trace
    //                        - src_spike * dest_trace        (double math)
    const double rhs = (((double)y.dw - 0.000001) * 0.0001 +
                        (double)(dd.dest_spike * src_trace)) -
                       (double)((float)src_spike * dest_trace);
    y.dw = (float)((double)y.dw - rhs);
    const float nw = y.w + s.learning_rate * y.dw;
    y.w = fmaxf(0.0001f, fminf(s.max_weight, nw));
    if (s.stp_flag) {
        // U, D, F of the opcode (izhikevich_attributes.cpp:432-438)
        const double U = SUB ? 0.2 : 0.5;
        const float D = SUB ? 700.0f : 800.0f;
        const double F = SUB ? 20.0 : 1000.0;
        const float fs = (float)src_spike;
        const float r = y.r, sw = y.sw;
        y.r = r + (((1.0f - r) / D) - (r * sw * fs));
        y.sw = (float)((double)sw + (((U - (double)sw) / F) +
                                     (U * (double)(1.0f - sw) * (double)fs)));
    }
}

// One kernel for the three Izhikevich synapse passes:
//   UPD = 1, ACT = 0 -- updater of the current step
//   UPD = 0, ACT = 1 -- conductance activator
//   UPD = 1, ACT = 1 -- updater of the previous step, then the activator of
//                       this step, per synapse, with the updated weight and
//                       short-term plasticity kept in registers between them
template <int TYPE, int KIND, int SUB, int UPD, int ACT>
__global__ __launch_bounds__(kBlock) void iz_synapse_kernel(
        sg_conn c, sg_iz_syn s, sg_iz_dest d, const float *src_table,
        const float *dst_table) {
    constexpr int B = kBatch;
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.num_dest) return;
    const int to = slot_to_index(c, slot);
    IzDest dd;
    if (UPD) {
        dd.dest_spike = extract<SG_OUT_BIT>(d.out_word0[to], 0);
        dd.dest_exc_trace = d.post_trace[to];
        dd.dest_tss = d.time_since_spike[to];
    }
    SynapseIter<TYPE> it;
    it.init(c, slot);
    const int count = it.count;
    IzAct acc;
    IzSyn ya[B], yb[B];
    int fa[B], fb[B];
    long ia[B], ib[B];
    auto load = [&](int k0, IzSyn *y, int *from, long *wi) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            from[b] = it.k < count ? it.source() : -1;
            wi[b] = it.wi;
            if (it.k < count) it.advance();
            if (from[b] < 0) continue;
            y[b].w = s.weights[wi[b]];
            y[b].r = s.stp_r[wi[b]];
            y[b].sw = s.stp_w[wi[b]];
            if (UPD) {
                y[b].dw = s.weight_derivatives[wi[b]];
                if (!SUB) y[b].pre = s.presyn_traces[wi[b]];
                y[b].tss = s.tss_cur[from[b]];
            }
            if (ACT) {
                y[b].dr = s.delays[wi[b]];
                y[b].word = c.src[from[b]];
                y[b].st = s.short_traces[wi[b]];
                if (KIND == SG_IZ_ADD || KIND == SG_IZ_SUB)
                    y[b].lt = s.long_traces[wi[b]];
            }
        }
    };
    auto consume = [&](IzSyn *y, const int *from, const long *wi) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (from[b] < 0) continue;
            const long i = wi[b];
            if (UPD && y[b].w >= 0.0001f) {
                iz_update_one<SUB>(s, dd, src_table, dst_table, y[b]);
                s.weights[i] = y[b].w;
                s.weight_derivatives[i] = y[b].dw;
                if (!SUB) s.presyn_traces[i] = y[b].pre;
                if (s.stp_flag) {
                    s.stp_r[i] = y[b].r;
                    s.stp_w[i] = y[b].sw;
                }
            }
            if (ACT) {
                iz_activate_one<KIND>(s, d, acc, y[b], from[b]);
                s.short_traces[i] = y[b].st;
                if (KIND == SG_IZ_ADD || KIND == SG_IZ_SUB)
                    s.long_traces[i] = y[b].lt;
            }
        }
    };
    // Two batches in flight: load the next while consuming the current.
    load(0, ya, fa, ia);
    for (int k0 = 0; k0 < count; k0 += 2 * B) {
        if (k0 + B < count) load(k0 + B, yb, fb, ib);
        consume(ya, fa, ia);
        if (k0 + B >= count) break;
        if (k0 + 2 * B < count) load(k0 + 2 * B, ya, fa, ia);
        consume(yb, fb, ib);
    }
    if (ACT) {
        d.short_cond[to] += acc.short_sum;
        if (KIND == SG_IZ_ADD || KIND == SG_IZ_SUB)
            d.long_cond[to] += acc.long_sum;
    }
}

// Gap junctions (izhikevich_attributes.cpp:354-366): sum += (v[from] - v[to])
// * w over the destination layer's voltage array.
template <int TYPE>
__global__ __launch_bounds__(kBlock) void iz_gap_kernel(sg_conn c,
                                                         sg_iz_syn s,
                                                         sg_iz_dest d) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.num_dest) return;
    const int to = slot_to_index(c, slot);
    const float v = d.voltage[to];
    float sum = 0.0f;
    for_each_synapse<TYPE>(c, slot, [&](int, int from, int wi) {
        sum += (d.voltage[from] - v) * s.weights[wi];
    });
    d.input[to] += sum;
}

// ---------------------------------------------------------------------------
// Dispatch helpers.
#define SG_TYPE_SWITCH(TYPEVAR, ...)                                       \
    switch (TYPEVAR) {                                                     \
        case SG_FULLY_CONNECTED: { constexpr int T = SG_FULLY_CONNECTED;   \
                                   __VA_ARGS__; } break;                   \
        case SG_SUBSET: { constexpr int T = SG_SUBSET; __VA_ARGS__; }      \
            break;                                                         \
        case SG_ONE_TO_ONE: { constexpr int T = SG_ONE_TO_ONE;             \
                              __VA_ARGS__; } break;                        \
        case SG_CONVERGENT: { constexpr int T = SG_CONVERGENT;             \
                              __VA_ARGS__; } break;                        \
        case SG_DIVERGENT: { constexpr int T = SG_DIVERGENT;               \
                             __VA_ARGS__; } break;                         \
        case SG_SPARSE: { constexpr int T = SG_SPARSE;                     \
                          __VA_ARGS__; } break;                            \
        default: return hipErrorInvalidValue;                              \
    }

#define SG_KIND_SWITCH(KINDVAR, ...)                                       \
    switch (KINDVAR) {                                                     \
        case SG_IZ_ADD: { constexpr int K = SG_IZ_ADD; __VA_ARGS__; }      \
            break;                                                         \
        case SG_IZ_SUB: { constexpr int K = SG_IZ_SUB; __VA_ARGS__; }      \
            break;                                                         \
        case SG_IZ_MULT: { constexpr int K = SG_IZ_MULT; __VA_ARGS__; }    \
            break;                                                         \
        case SG_IZ_REWARD: { constexpr int K = SG_IZ_REWARD;               \
                             __VA_ARGS__; } break;                         \
        default: return hipErrorInvalidValue;                              \
    }

hipError_t launch_tile_base(const sg_conn &c, hipStream_t st);
bool plane_path(const sg_conn &c);
hipError_t launch_plane_base(const sg_conn &c, hipStream_t st);
hipError_t launch_iz_tile(const sg_conn &c, const sg_iz_syn &s,
                          const sg_iz_dest &d, int kind, int sub, int upd,
                          int act, const float *st_, const float *dt_,
                          hipStream_t st);

static inline bool tiled_type(int t) {
    return t == SG_FULLY_CONNECTED || t == SG_SUBSET || t == SG_CONVERGENT ||
           t == SG_DIVERGENT || t == SG_SPARSE;
}

}  // namespace sg

using namespace sg;

extern "C" int sg_hip_activate_base(const sg_conn *c, void *stream) {
    if (c->num_dest <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (plane_path(*c)) return launch_plane_base(*c, s);
    if (tiled_type(c->type)) return launch_tile_base(*c, s);
    switch (c->type) {
        case SG_FULLY_CONNECTED: return launch_base_t<SG_FULLY_CONNECTED>(*c, s);
        case SG_SUBSET: return launch_base_t<SG_SUBSET>(*c, s);
        case SG_ONE_TO_ONE: return launch_base_t<SG_ONE_TO_ONE>(*c, s);
        case SG_CONVERGENT: return launch_base_t<SG_CONVERGENT>(*c, s);
        case SG_DIVERGENT: return launch_base_t<SG_DIVERGENT>(*c, s);
        case SG_SPARSE: return launch_base_t<SG_SPARSE>(*c, s);
    }
    return hipErrorInvalidValue;
}

// activate_base_second_order (synapse_kernel.h:918-935): one lane per
// destination slot walks its visits; every synapse of the slave updates the
// same index of the host's weight copy.
template <int TYPE, int OUT>
__global__ void second_order_kernel(sg_conn c, float *sow) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.num_dest) return;
    SynapseIter<TYPE> it;
    it.init(c, slot);
    for (int k = 0; k < it.count; ++k) {
        const int from = it.source();
        if (from >= 0) {
            const float val = extract<OUT>(c.src[from], c.delay) * c.weights[it.wi];
            sow[it.wi] = aggregate(c.opcode, sow[it.wi], val);
        }
        it.advance();
    }
}

template <int TYPE>
static hipError_t launch_second_order_t(const sg_conn &c, float *sow,
                                        hipStream_t s) {
    dim3 g((c.num_dest + 255) / 256), b(256);
    switch (c.out_type) {
        case SG_OUT_BIT:
            hipLaunchKernelGGL((second_order_kernel<TYPE, SG_OUT_BIT>), g, b,
                               0, s, c, sow);
            break;
        case SG_OUT_INT:
            hipLaunchKernelGGL((second_order_kernel<TYPE, SG_OUT_INT>), g, b,
                               0, s, c, sow);
            break;
        default:
            hipLaunchKernelGGL((second_order_kernel<TYPE, SG_OUT_FLOAT>), g,
                               b, 0, s, c, sow);
    }
    return hipGetLastError();
}

// A convolutional slave (kernel.cpp:223-229, synapse_kernel.h:943-970):
// one-to-one over its source layer, sow[i] = aggregate(sow[i],
// extract(source i) * weights[i]).
template <int OUT>
__global__ void second_order_conv_kernel(sg_conn c, float *sow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.from_size) return;
    const float val = extract<OUT>(c.src[i], c.delay) * c.weights[i];
    sow[i] = aggregate(c.opcode, sow[i], val);
}

extern "C" int sg_hip_activate_second_order(const sg_conn *c, float *sow,
                                            void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (c->convolutional) {
        if (c->from_size <= 0) return 0;
        dim3 g((c->from_size + 255) / 256), b(256);
        switch (c->out_type) {
            case SG_OUT_BIT:
                hipLaunchKernelGGL(second_order_conv_kernel<SG_OUT_BIT>, g, b,
                                   0, s, *c, sow);
                break;
            case SG_OUT_INT:
                hipLaunchKernelGGL(second_order_conv_kernel<SG_OUT_INT>, g, b,
                                   0, s, *c, sow);
                break;
            default:
                hipLaunchKernelGGL(second_order_conv_kernel<SG_OUT_FLOAT>, g,
                                   b, 0, s, *c, sow);
        }
        return hipGetLastError();
    }
    if (c->num_dest <= 0) return 0;
    switch (c->type) {
        case SG_FULLY_CONNECTED:
            return launch_second_order_t<SG_FULLY_CONNECTED>(*c, sow, s);
        case SG_SUBSET: return launch_second_order_t<SG_SUBSET>(*c, sow, s);
        case SG_ONE_TO_ONE:
            return launch_second_order_t<SG_ONE_TO_ONE>(*c, sow, s);
        case SG_CONVERGENT:
            return launch_second_order_t<SG_CONVERGENT>(*c, sow, s);
        case SG_DIVERGENT:
            return launch_second_order_t<SG_DIVERGENT>(*c, sow, s);
    }
    return hipErrorInvalidValue;
}

// calc_internal (engine/kernel/kernel.cpp:161-187).
__global__ void dendrite_kernel(float *dst, float *src, int n, int op,
                                float trail) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = aggregate(op, dst[i], src[i]);
    src[i] = trail;
}

extern "C" int sg_hip_dendrite(float *dst, float *src, int n, int opcode,
                               float trail, void *stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(dendrite_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, dst, src, n, opcode, trail);
    return hipGetLastError();
}

extern "C" int sg_hip_tss_step(int32_t *next, const int32_t *cur,
                               const uint8_t *visits,
                               const uint8_t *rev_delays,
                               const uint32_t *words, int n, void *stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(tss_step_kernel, dim3(blocks_for(n)), dim3(kBlock), 0,
                       (hipStream_t)stream, next, cur, visits, rev_delays,
                       words, n);
    return hipGetLastError();
}

extern "C" int sg_hip_iz_activate(const sg_conn *c, const sg_iz_syn *s,
                                  const sg_iz_dest *d, int kind,
                                  void *stream) {
    if (c->num_dest <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    dim3 g(blocks_for(c->num_dest)), b(kBlock);
    if (kind == SG_IZ_GAP) {
        SG_TYPE_SWITCH(c->type,
            hipLaunchKernelGGL((iz_gap_kernel<T>), g, b, 0, st, *c, *s, *d));
        return hipGetLastError();
    }
    if (tiled_type(c->type) && s->record)
        return launch_iz_tile(*c, *s, *d, kind, 0, 0, 1, nullptr, nullptr, st);
    SG_TYPE_SWITCH(c->type, SG_KIND_SWITCH(kind,
        hipLaunchKernelGGL((iz_synapse_kernel<T, K, 0, 0, 1>), g, b, 0, st,
                           *c, *s, *d, nullptr, nullptr)));
    return hipGetLastError();
}

extern "C" int sg_hip_iz_update(const sg_conn *c, const sg_iz_syn *s,
                                const sg_iz_dest *d, int sub,
                                const float *src_table, const float *dst_table,
                                void *stream) {
    if (c->num_dest <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    dim3 g(blocks_for(c->num_dest)), b(kBlock);
    if (tiled_type(c->type) && s->record)
        return launch_iz_tile(*c, *s, *d, sub ? SG_IZ_SUB : SG_IZ_ADD, sub, 1,
                              0, src_table, dst_table, st);
    if (sub) {
        SG_TYPE_SWITCH(c->type,
            hipLaunchKernelGGL((iz_synapse_kernel<T, SG_IZ_SUB, 1, 1, 0>), g,
                               b, 0, st, *c, *s, *d, src_table, dst_table));
    } else {
        SG_TYPE_SWITCH(c->type,
            hipLaunchKernelGGL((iz_synapse_kernel<T, SG_IZ_ADD, 0, 1, 0>), g,
                               b, 0, st, *c, *s, *d, src_table, dst_table));
    }
    return hipGetLastError();
}

extern "C" int sg_hip_iz_update_activate(const sg_conn *c, const sg_iz_syn *s,
                                         const sg_iz_dest *d, int kind,
                                         int sub, const float *src_table,
                                         const float *dst_table,
                                         void *stream) {
    if (c->num_dest <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    dim3 g(blocks_for(c->num_dest)), b(kBlock);
    if (tiled_type(c->type) && s->record)
        return launch_iz_tile(*c, *s, *d, kind, sub, 1, 1, src_table,
                              dst_table, st);
    if (sub) {
        SG_TYPE_SWITCH(c->type, SG_KIND_SWITCH(kind,
            hipLaunchKernelGGL((iz_synapse_kernel<T, K, 1, 1, 1>), g, b, 0,
                               st, *c, *s, *d, src_table, dst_table)));
    } else {
        SG_TYPE_SWITCH(c->type, SG_KIND_SWITCH(kind,
            hipLaunchKernelGGL((iz_synapse_kernel<T, K, 0, 1, 1>), g, b, 0,
                               st, *c, *s, *d, src_table, dst_table)));
    }
    return hipGetLastError();
}
