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

// ---------------------------------------------------------------------------
// Base activator (synapse_kernel.h:887-912): sum = 0; sum += extract * w in
// serial order; input = aggregate(input, sum).
template <int TYPE, int OUT>
__global__ __launch_bounds__(kBlock) void activate_base_kernel(sg_conn c) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.num_dest) return;
    const uint32_t *src = c.src;
    const float *w = c.weights;
    const int delay = c.delay;
    float sum = 0.0f;
    for_each_synapse<TYPE>(c, slot, [&](int, int from, int wi) {
        float val = extract<OUT>(src[from], delay) * w[wi];
        sum += val;
    });
    const int to = slot_to_index(c, slot);
    c.dst[to] = aggregate(c.opcode, c.dst[to], sum);
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
// time-since-spike of each source (izhikevich_attributes.cpp:258-262).  The
// reference updates it per visit: spike ? 0 : min(32, tss + 1).  Over one
// activation a source visited m times ends at min(32, m - j) if its last
// spiking visit is the j-th (1-based), else min(32, tss + m).  The pre-pass
// writes the second value; each spiking visit then lowers it to its own
// min(32, m - j) with atomicMin, which leaves exactly the last one's value.
__global__ __launch_bounds__(kBlock) void tss_prepass_kernel(
        int32_t *next, const int32_t *cur, const uint8_t *visits, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) next[i] = min(32, cur[i] + (int)visits[i]);
}

// ---------------------------------------------------------------------------
// Izhikevich conductance activator step for one synapse
// (izhikevich_attributes.cpp:257-277):
//   weight = w * (r * stp_w)
//   trace' = trace + spike * weight * g ; sum += trace' ; trace = trace' * tau
struct IzAct {
    float short_sum = 0.0f, long_sum = 0.0f;
};

template <int KIND>
__device__ __forceinline__ void iz_activate_synapse(
        const sg_iz_syn &s, const sg_iz_dest &d, IzAct &acc, float w,
        float r, float sw, uint32_t word, int from, int wi) {
    const uint16_t dr = s.delay_rank[wi];
    const float spike = extract<SG_OUT_BIT>(word, dr & 0xFF);
    if (spike > 0.0f) atomicMin(&s.tss_next[from], (int)(dr >> 8));
    const float weight = w * (r * sw);
    const float inc = spike * weight * s.baseline_conductance;
    const float st = s.short_traces[wi] + inc;
    acc.short_sum += st;
    s.short_traces[wi] = st * d.short_tau;
    if (KIND == SG_IZ_ADD || KIND == SG_IZ_SUB) {
        const float lt = s.long_traces[wi] + inc;
        acc.long_sum += lt;
        s.long_traces[wi] = lt * d.long_tau;
    }
}

template <int TYPE, int KIND>
__global__ __launch_bounds__(kBlock) void iz_activate_kernel(sg_conn c,
                                                              sg_iz_syn s,
                                                              sg_iz_dest d) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.num_dest) return;
    const int to = slot_to_index(c, slot);
    if (KIND == SG_IZ_GAP) {
        // izhikevich_attributes.cpp:354-366: sum += (v[from] - v[to]) * w
        // over the destination layer's voltage array.
        const float v = d.voltage[to];
        float sum = 0.0f;
        for_each_synapse<TYPE>(c, slot, [&](int, int from, int wi) {
            sum += (d.voltage[from] - v) * s.weights[wi];
        });
        d.input[to] += sum;
        return;
    }
    IzAct acc;
    for_each_synapse<TYPE>(c, slot, [&](int, int from, int wi) {
        iz_activate_synapse<KIND>(s, d, acc, s.weights[wi], s.stp_r[wi],
                                  s.stp_w[wi], c.src[from], from, wi);
    });
    d.short_cond[to] += acc.short_sum;
    if (KIND == SG_IZ_ADD || KIND == SG_IZ_SUB) d.long_cond[to] += acc.long_sum;
}

// ---------------------------------------------------------------------------
// Izhikevich updater for one synapse (izhikevich_attributes.cpp:444-508).
// Returns the (possibly) updated weight, r and stp_w in place.
struct IzDest {
    float dest_spike;       // MSB of the destination's history word 0
    float dest_exc_trace;
    int dest_tss;
};

// U, D, F per opcode (izhikevich_attributes.cpp:432-438)
template <int SUB>
__device__ __forceinline__ void iz_update_synapse(
        const sg_iz_syn &s, const IzDest &dd, const float *src_table,
        const float *dst_table, float &w, float &r, float &sw, int from,
        int wi, float max_weight) {
    if (!(w >= 0.0001f)) return;
    const int src_tss = s.tss_cur[from];
    const bool src_spike = src_tss == 0;
    float src_trace, dest_trace;
    if (!SUB) {
        // STDP_ADD: presyn = spike ? 0.004 : presyn * 0.933 (double math)
        const float p = src_spike
            ? (float)0.004
            : (float)((double)s.presyn_traces[wi] * 0.933);
        s.presyn_traces[wi] = p;
        src_trace = p;
        dest_trace = dd.dest_exc_trace;
    } else {
        // STDP_SUB: iSTDP factors depend only on the integer times since
        // spike, tabulated on the host with the reference's own
        // powf / expf expressions.
        src_trace = (dd.dest_spike > 0.0f && src_tss > 0 && src_tss < 32)
            ? src_table[src_tss] : 0.0f;
        dest_trace = (src_spike && dd.dest_tss > 0 && dd.dest_tss < 32)
            ? dst_table[dd.dest_tss] : 0.0f;
    }
    // weight_derivatives -= (wd - C_A) * C_TAU + dest_spike * src_trace
    //                        - src_spike * dest_trace        (double math)
    const float wd = s.weight_derivatives[wi];
    const double rhs = (((double)wd - 0.000001) * 0.0001 +
                        (double)(dd.dest_spike * src_trace)) -
                       (double)((float)src_spike * dest_trace);
    const float wd_new = (float)((double)wd - rhs);
    s.weight_derivatives[wi] = wd_new;
    float nw = w + s.learning_rate * wd_new;
    w = fmaxf(0.0001f, fminf(max_weight, nw));
    if (s.stp_flag) {
        const double U = SUB ? 0.2 : 0.5;
        const float D = SUB ? 700.0f : 800.0f;
        const double F = SUB ? 20.0 : 1000.0;
        const float fs = (float)src_spike;
        const float r_new = r + (((1.0f - r) / D) - (r * sw * fs));
        const float w_new = (float)((double)sw +
                                    (((U - (double)sw) / F) +
                                     (U * (double)(1.0f - sw) * (double)fs)));
        r = r_new;
        sw = w_new;
    }
}

template <int TYPE, int SUB>
__global__ __launch_bounds__(kBlock) void iz_update_kernel(
        sg_conn c, sg_iz_syn s, sg_iz_dest d, const float *src_table,
        const float *dst_table) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.num_dest) return;
    const int to = slot_to_index(c, slot);
    IzDest dd;
    dd.dest_spike = extract<SG_OUT_BIT>(d.out_word0[to], 0);
    dd.dest_exc_trace = d.post_trace[to];
    dd.dest_tss = d.time_since_spike[to];
    const float mw = s.max_weight;
    for_each_synapse<TYPE>(c, slot, [&](int, int from, int wi) {
        float w = s.weights[wi];
        if (!(w >= 0.0001f)) return;
        float r = s.stp_r[wi], sw = s.stp_w[wi];
        iz_update_synapse<SUB>(s, dd, src_table, dst_table, w, r, sw, from,
                               wi, mw);
        s.weights[wi] = w;
        if (s.stp_flag) {
            s.stp_r[wi] = r;
            s.stp_w[wi] = sw;
        }
    });
}

// Fused: updater of the previous step, then this step's activator, per
// synapse, with the weights held in registers between the two.
template <int TYPE, int KIND, int SUB>
__global__ __launch_bounds__(kBlock) void iz_update_activate_kernel(
        sg_conn c, sg_iz_syn s, sg_iz_dest d, const float *src_table,
        const float *dst_table) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.num_dest) return;
    const int to = slot_to_index(c, slot);
    IzDest dd;
    dd.dest_spike = extract<SG_OUT_BIT>(d.out_word0[to], 0);
    dd.dest_exc_trace = d.post_trace[to];
    dd.dest_tss = d.time_since_spike[to];
    const float mw = s.max_weight;
    IzAct acc;
    for_each_synapse<TYPE>(c, slot, [&](int, int from, int wi) {
        float w = s.weights[wi];
        float r = s.stp_r[wi], sw = s.stp_w[wi];
        if (w >= 0.0001f) {
            iz_update_synapse<SUB>(s, dd, src_table, dst_table, w, r, sw,
                                   from, wi, mw);
            s.weights[wi] = w;
            if (s.stp_flag) {
                s.stp_r[wi] = r;
                s.stp_w[wi] = sw;
            }
        }
        iz_activate_synapse<KIND>(s, d, acc, w, r, sw, c.src[from], from, wi);
    });
    d.short_cond[to] += acc.short_sum;
    if (KIND == SG_IZ_ADD || KIND == SG_IZ_SUB) d.long_cond[to] += acc.long_sum;
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

}  // namespace sg

using namespace sg;

extern "C" int sg_hip_activate_base(const sg_conn *c, void *stream) {
    if (c->num_dest <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    switch (c->type) {
        case SG_FULLY_CONNECTED: return launch_base_t<SG_FULLY_CONNECTED>(*c, s);
        case SG_SUBSET: return launch_base_t<SG_SUBSET>(*c, s);
        case SG_ONE_TO_ONE: return launch_base_t<SG_ONE_TO_ONE>(*c, s);
        case SG_CONVERGENT: return launch_base_t<SG_CONVERGENT>(*c, s);
        case SG_DIVERGENT: return launch_base_t<SG_DIVERGENT>(*c, s);
    }
    return hipErrorInvalidValue;
}

extern "C" int sg_hip_tss_prepass(int32_t *next, const int32_t *cur,
                                  const uint8_t *visits, int n,
                                  void *stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(tss_prepass_kernel, dim3(blocks_for(n)), dim3(kBlock),
                       0, (hipStream_t)stream, next, cur, visits, n);
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
            hipLaunchKernelGGL((iz_activate_kernel<T, SG_IZ_GAP>), g, b, 0,
                               st, *c, *s, *d));
        return hipGetLastError();
    }
    SG_TYPE_SWITCH(c->type, SG_KIND_SWITCH(kind,
        hipLaunchKernelGGL((iz_activate_kernel<T, K>), g, b, 0, st, *c, *s,
                           *d)));
    return hipGetLastError();
}

extern "C" int sg_hip_iz_update(const sg_conn *c, const sg_iz_syn *s,
                                const sg_iz_dest *d, int sub,
                                const float *src_table, const float *dst_table,
                                void *stream) {
    if (c->num_dest <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    dim3 g(blocks_for(c->num_dest)), b(kBlock);
    if (sub) {
        SG_TYPE_SWITCH(c->type,
            hipLaunchKernelGGL((iz_update_kernel<T, 1>), g, b, 0, st, *c, *s,
                               *d, src_table, dst_table));
    } else {
        SG_TYPE_SWITCH(c->type,
            hipLaunchKernelGGL((iz_update_kernel<T, 0>), g, b, 0, st, *c, *s,
                               *d, src_table, dst_table));
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
    if (sub) {
        SG_TYPE_SWITCH(c->type, SG_KIND_SWITCH(kind,
            hipLaunchKernelGGL((iz_update_activate_kernel<T, K, 1>), g, b, 0,
                               st, *c, *s, *d, src_table, dst_table)));
    } else {
        SG_TYPE_SWITCH(c->type, SG_KIND_SWITCH(kind,
            hipLaunchKernelGGL((iz_update_activate_kernel<T, K, 0>), g, b, 0,
                               st, *c, *s, *d, src_table, dst_table)));
    }
    return hipGetLastError();
}
