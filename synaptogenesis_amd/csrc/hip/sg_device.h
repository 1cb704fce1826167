// Device-side helpers shared by the synapse kernels: the source extractor,
// the opcode aggregator and the per-destination synapse enumeration of every
// connection type, in the reference's serial visiting order.
#ifndef SG_DEVICE_H
#define SG_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sg_hip.h"

namespace sg {

// Philox4x32-10 (Salmon et al., SC'11): the counter-based generator of the
// random init configs and of synthetic ensemble data.
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0,
                                       uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// ((x % n) + n) % n with C truncating %, as synapse_kernel.h:401-406.
__device__ __forceinline__ int wrap_index(int x, int n) {
    return ((x % n) + n) % n;
}

// extract (engine/kernel/extractor.cpp:8-15): FLOAT reads the float,
// INT converts the word, BIT returns bit (delay & 31) counted from the MSB.
template <int OUT>
__device__ __forceinline__ float extract(uint32_t word, int delay) {
    if (OUT == SG_OUT_FLOAT) return __uint_as_float(word);
    if (OUT == SG_OUT_INT) return (float)word;
    return (float)((word << (delay & 31)) >> 31);
}

// aggregate (engine/kernel/aggregator.cpp:15-48); POOL is fmax as on the
// serial path (util/parallel.h:67).
__device__ __forceinline__ float aggregate(int op, float prior, float in) {
    switch (op) {
        case SG_ADD: return prior + in;
        case SG_SUB: return prior - in;
        case SG_MULT: return prior * in;
        case SG_DIV: return prior / in;
        case SG_POOL: return fmaxf(prior, in);
        case SG_ADD_HEAVISIDE: return prior + (float)(in > 0.0f);
        case SG_SUB_HEAVISIDE: return prior - (float)(in > 0.0f);
        case SG_MULT_HEAVISIDE: return prior * (float)(in > 0.0f);
        default: return prior + in;
    }
}

// Writes a destination's sum: aggregated into its input, or stored raw for
// the neuron kernel to aggregate into an init-config value (store_sum).
__device__ __forceinline__ void put_sum(const sg_conn &c, int to, float sum) {
    c.dst[to] = c.store_sum ? sum : aggregate(c.opcode, c.dst[to], sum);
}

// Time since spike of one source after one activation of an Izhikevich
// conductance connection (izhikevich_attributes.cpp:258-262 applied once per
// visit in serial order): with m = min(32, visits) and rev[i] the delay of
// the i-th visit counted back from the last, the result is the first i
// whose delayed bit of `word` is set, else min(32, cur + m).
__device__ __forceinline__ int tss_next(int cur, int m, const uint32_t *rd,
                                        uint32_t word) {
    int t = min(32, cur + m);
    if (word != 0u) {
        for (int q = 0; q < 8 && q * 4 < m; ++q) {
            const uint32_t four = rd[q];
            int hit = -1;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int i = q * 4 + b;
                const int d = (four >> (8 * b)) & 31;
                if (hit < 0 && i < m && ((word << d) >> 31)) hit = i;
            }
            if (hit >= 0) { t = hit; break; }
        }
    }
    return t;
}

// The presynaptic trace (update_iz_add only) and the short-term plasticity
// pair (R, w) of a synapse after one updater step: they depend only on
// their own values and the source's spike (izhikevich_attributes.cpp:
// 444-508).  The tile kernels step them per synapse; a compact record
// steps them once per source.
template <int SUB>
__device__ __forceinline__ void iz_source_step(bool src_spike, int stp_flag,
                                               float &pre, float &r,
                                               float &sw) {
    if (!SUB) pre = src_spike ? (float)0.004 : (float)((double)pre * 0.933);
    if (stp_flag) {
        const double U = SUB ? 0.2 : 0.5;
        const float D = SUB ? 700.0f : 800.0f;
        const double F = SUB ? 20.0 : 1000.0;
        const float fs = (float)src_spike;
        const float r0 = r, sw0 = sw;
        r = r0 + (((1.0f - r0) / D) - (r0 * sw0 * fs));
        sw = (float)((double)sw0 + (((U - (double)sw0) / F) +
                                    (U * (double)(1.0f - sw0) * (double)fs)));
    }
}

// Destination index of slot `slot` (SUBSET slots cover the to-rectangle).
__device__ __forceinline__ int slot_to_index(const sg_conn &c, int slot) {
    if (c.type != SG_SUBSET) return slot;
    int sub_cols = c.to_col_end - c.to_col_start;
    return (slot / sub_cols + c.to_row_start) * c.to_cols +
           (slot % sub_cols) + c.to_col_start;
}

// Calls f(k, from_index, weight_index) for every synapse of destination
// slot `slot`, in the order of the reference's serial loops:
//   FC          synapse_kernel.h:115-126 (from ascending)
//   SUBSET      synapse_kernel.h:182-195 (source rectangle row-major)
//   ONE_TO_ONE  synapse_kernel.h:257-286
//   CONVERGENT  synapse_kernel.h:384-421 (k_index also counts skipped
//               out-of-bounds kernel cells)
//   DIVERGENT   synapse_kernel.h:506-550 (C truncating division)
//   SPARSE      synapse_kernel.h:311-322 (the compacted list; k runs to
//               the largest count, entries past a slot's own count are -1)
// weight_index is the device layout index (see tiled_index), or k for a
// shared convolutional kernel.
// Per-destination synapse arrays are tiled by 64 destinations: the k-th
// synapse of destination slot s lives at ((s / 64) * K + k) * 64 + s % 64,
// so each wavefront-wide tile of destinations streams one contiguous block
// of K * 64 elements per array.
__host__ __device__ __forceinline__ long tiled_index(int slot, int k, int K) {
    return ((long)(slot >> 6) * K + k) * 64 + (slot & 63);
}

template <int TYPE, class F>
__device__ __forceinline__ void for_each_synapse(const sg_conn &c, int slot,
                                                 F &&f) {
    const int K = c.k_per_dest;
    if (TYPE == SG_FULLY_CONNECTED) {
        for (int k = 0; k < c.from_size; ++k) f(k, k, tiled_index(slot, k, K));
    } else if (TYPE == SG_ONE_TO_ONE) {
        f(0, slot, slot);
    } else if (TYPE == SG_SUBSET) {
        int k = 0;
        for (int fr = c.from_row_start; fr < c.from_row_end; ++fr)
            for (int fc = c.from_col_start; fc < c.from_col_end; ++fc, ++k)
                f(k, fr * c.from_cols + fc, tiled_index(slot, k, K));
    } else if (TYPE == SG_CONVERGENT) {
        const int to_row = slot / c.to_cols, to_col = slot % c.to_cols;
        const int s_row = to_row * c.row_stride + c.row_spacing * c.row_offset;
        const int s_col = to_col * c.col_stride + c.col_spacing * c.col_offset;
        int k = 0;
        for (int kr = 0; kr < c.row_field; ++kr) {
            int fr = s_row + kr * c.row_spacing;
            bool row_in = true;
            if (c.wrap) fr = wrap_index(fr, c.from_rows);
            else row_in = fr >= 0 && fr < c.from_rows;
            int fc = s_col;
            if (c.wrap) fc = wrap_index(fc, c.from_cols);
            for (int kc = 0; kc < c.col_field; ++kc, ++k) {
                int col = fc;
                if (c.wrap) {
                    fc += c.col_spacing;
                    while (fc >= c.from_cols) fc -= c.from_cols;
                } else {
                    fc += c.col_spacing;
                    if (!row_in || col < 0 || col >= c.from_cols) continue;
                }
                f(k, fr * c.from_cols + col,
                  c.convolutional ? k : tiled_index(slot, k, K));
            }
        }
    } else if (TYPE == SG_SPARSE) {
        for (int k = 0; k < K; ++k) {
            const long wi = tiled_index(slot, k, K);
            const int from = c.sparse_from[wi];
            if (from >= 0) f(k, from, wi);
        }
    } else if (TYPE == SG_DIVERGENT) {
        const int to_row = slot / c.to_cols, to_col = slot % c.to_cols;
        const int sr0 = (to_row + c.row_spacing * (-c.row_offset - c.row_field +
                                                   c.row_stride)) / c.row_stride;
        const int sc0 = (to_col + c.col_spacing * (-c.col_offset - c.col_field +
                                                   c.col_stride)) / c.col_stride;
        const int sr1 = sr0 + (c.row_spacing * (c.row_field - c.row_stride) /
                               c.row_stride);
        const int sc1 = sc0 + (c.col_spacing * (c.col_field - c.col_stride) /
                               c.col_stride);
        int k = 0;
        for (int sr = sr0; sr <= sr1; sr += c.row_spacing) {
            int fr = sr;
            bool row_in = true;
            if (c.wrap) fr = wrap_index(fr, c.from_rows);
            else row_in = fr >= 0 && fr < c.from_rows;
            for (int sc = sc0; sc <= sc1; sc += c.col_spacing, ++k) {
                int fc = sc;
                if (c.wrap) fc = wrap_index(fc, c.from_cols);
                else if (!row_in || fc < 0 || fc >= c.from_cols) continue;
                f(k, fr * c.from_cols + fc,
                  c.convolutional ? k : tiled_index(slot, k, K));
            }
        }
    }
}

}  // namespace sg

#endif

#ifndef SG_DEVICE_RANDOM_ACCESS
#define SG_DEVICE_RANDOM_ACCESS
namespace sg {

// Random access into one destination's synapse list: the k-th visit of the
// reference's serial loop, so loads for a batch of k can be issued before
// the ordered accumulation consumes them.  `count` is the loop trip count
// (k runs over it even where a visit is skipped, as k_index does).
struct SynapseCursor {
    int count;        // visits in the serial loop, skipped ones included
    int row0, col0;   // first source row/column (before wrapping)
    int ncols;        // inner-loop trip count
    int rstep, cstep; // source row/column step per inner/outer iteration
};

template <int TYPE>
__host__ __device__ inline SynapseCursor cursor(const sg_conn &c, int slot) {
    SynapseCursor q;
    if (TYPE == SG_FULLY_CONNECTED) {
        q.count = c.from_size;
        q.row0 = 0; q.col0 = 0; q.ncols = c.from_cols; q.rstep = 1; q.cstep = 1;
    } else if (TYPE == SG_ONE_TO_ONE) {
        q.count = 1;
        q.row0 = slot / c.from_cols; q.col0 = slot % c.from_cols;
        q.ncols = 1; q.rstep = 1; q.cstep = 1;
    } else if (TYPE == SG_SUBSET) {
        q.ncols = c.from_col_end - c.from_col_start;
        q.count = (c.from_row_end - c.from_row_start) * q.ncols;
        q.row0 = c.from_row_start; q.col0 = c.from_col_start;
        q.rstep = 1; q.cstep = 1;
    } else if (TYPE == SG_CONVERGENT) {
        const int to_row = slot / c.to_cols, to_col = slot % c.to_cols;
        q.row0 = to_row * c.row_stride + c.row_spacing * c.row_offset;
        q.col0 = to_col * c.col_stride + c.col_spacing * c.col_offset;
        q.ncols = c.col_field;
        q.count = c.row_field * c.col_field;
        q.rstep = c.row_spacing; q.cstep = c.col_spacing;
    } else if (TYPE == SG_SPARSE) {
        q.count = c.k_per_dest;
        q.row0 = slot; q.col0 = 0; q.ncols = c.k_per_dest;
        q.rstep = 1; q.cstep = 1;
    } else {  // SG_DIVERGENT
        const int to_row = slot / c.to_cols, to_col = slot % c.to_cols;
        q.row0 = (to_row + c.row_spacing * (-c.row_offset - c.row_field +
                                            c.row_stride)) / c.row_stride;
        q.col0 = (to_col + c.col_spacing * (-c.col_offset - c.col_field +
                                            c.col_stride)) / c.col_stride;
        const int rspan = c.row_spacing * (c.row_field - c.row_stride) /
                          c.row_stride;
        const int cspan = c.col_spacing * (c.col_field - c.col_stride) /
                          c.col_stride;
        const int nrows = rspan < 0 ? 0 : rspan / c.row_spacing + 1;
        q.ncols = cspan < 0 ? 0 : cspan / c.col_spacing + 1;
        q.count = nrows * q.ncols;
        q.rstep = c.row_spacing; q.cstep = c.col_spacing;
    }
    return q;
}

// Source index of visit k, or -1 where the serial loop skips it.
template <int TYPE>
__device__ __forceinline__ int source_at(const sg_conn &c,
                                         const SynapseCursor &q, int k) {
    if (TYPE == SG_FULLY_CONNECTED) return k;
    if (TYPE == SG_ONE_TO_ONE) return q.row0 * c.from_cols + q.col0;
    if (TYPE == SG_SPARSE)
        return c.sparse_from[tiled_index(q.row0, k, c.k_per_dest)];
    const int kr = k / q.ncols, kc = k - kr * q.ncols;
    int r = q.row0 + kr * q.rstep, col = q.col0 + kc * q.cstep;
    if (TYPE == SG_SUBSET) return r * c.from_cols + col;
    if (c.wrap) {
        r = wrap_index(r, c.from_rows);
        col = wrap_index(col, c.from_cols);
    } else if (r < 0 || r >= c.from_rows || col < 0 || col >= c.from_cols) {
        return -1;
    }
    return r * c.from_cols + col;
}

// Division-free walk over one destination's visits in serial order: the
// source row/column advance by their steps and wrap by subtraction, the
// device-layout weight index advances by one k-stride per visit.
template <int TYPE>
struct SynapseIter {
    int k, count, kc, ncols;
    int r, c, c0, rstep, cstep, rows, cols;
    bool wrap;
    long wi, wstep;
    const int32_t *sf;   // SG_SPARSE: the tiled source indices

    __device__ __forceinline__ void init(const sg_conn &cn, int slot) {
        const SynapseCursor q = cursor<TYPE>(cn, slot);
        sf = cn.sparse_from;
        k = 0; kc = 0;
        count = q.count; ncols = q.ncols;
        rows = cn.from_rows; cols = cn.from_cols;
        rstep = q.rstep; cstep = q.cstep;
        wrap = (TYPE == SG_CONVERGENT || TYPE == SG_DIVERGENT) && cn.wrap;
        r = q.row0; c = q.col0;
        if (wrap) {
            r = wrap_index(r, rows); c = wrap_index(c, cols);
            // steps taken modulo the layer: one subtraction wraps a step
            rstep = wrap_index(rstep, rows); cstep = wrap_index(cstep, cols);
        }
        c0 = c;
        if (TYPE == SG_ONE_TO_ONE) { wi = slot; wstep = 0; }
        else if ((TYPE == SG_CONVERGENT || TYPE == SG_DIVERGENT) &&
                 cn.convolutional) { wi = 0; wstep = 1; }
        else { wi = tiled_index(slot, 0, cn.k_per_dest); wstep = 64; }
    }
    // Jump to visit k (one division; then advance() walks on).
    __device__ __forceinline__ void seek(const sg_conn &cn, int slot,
                                         int kk) {
        init(cn, slot);
        if (kk <= 0 || TYPE == SG_ONE_TO_ONE) return;
        const SynapseCursor q = cursor<TYPE>(cn, slot);
        k = kk;
        kc = kk % ncols;
        const int kr = kk / ncols;
        r = q.row0 + kr * q.rstep;
        c = q.col0 + kc * q.cstep;
        if (wrap) { r = wrap_index(r, rows); c = wrap_index(c, cols); }
        wi += (long)kk * wstep;
    }
    // Source index of the current visit, or -1 where the loop skips it.
    __device__ __forceinline__ int source() const {
        if (TYPE == SG_SPARSE) return sf[wi];
        if ((TYPE == SG_CONVERGENT || TYPE == SG_DIVERGENT) && !wrap &&
            (r < 0 || r >= rows || c < 0 || c >= cols))
            return -1;
        return r * cols + c;
    }
    __device__ __forceinline__ void advance() {
        ++k;
        wi += wstep;
        if (++kc == ncols) {
            kc = 0;
            r += rstep;
            if (wrap && r >= rows) r -= rows;
            c = c0;
        } else {
            c += cstep;
            if (wrap && c >= cols) c -= cols;
        }
    }
};



}  // namespace sg
#endif
