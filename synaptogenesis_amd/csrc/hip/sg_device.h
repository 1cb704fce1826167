// Device-side helpers shared by the synapse kernels: the source extractor,
// the opcode aggregator and the per-destination synapse enumeration of every
// connection type, in the reference's serial visiting order.
#ifndef SG_DEVICE_H
#define SG_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sg_hip.h"

namespace sg {

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
// weight_index is the device layout index: k * num_dest + slot, or k for a
// shared convolutional kernel.
template <int TYPE, class F>
__device__ __forceinline__ void for_each_synapse(const sg_conn &c, int slot,
                                                 F &&f) {
    const int nd = c.num_dest;
    if (TYPE == SG_FULLY_CONNECTED) {
        for (int k = 0; k < c.from_size; ++k) f(k, k, k * nd + slot);
    } else if (TYPE == SG_ONE_TO_ONE) {
        f(0, slot, slot);
    } else if (TYPE == SG_SUBSET) {
        int k = 0;
        for (int fr = c.from_row_start; fr < c.from_row_end; ++fr)
            for (int fc = c.from_col_start; fc < c.from_col_end; ++fc, ++k)
                f(k, fr * c.from_cols + fc, k * nd + slot);
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
                  c.convolutional ? k : k * nd + slot);
            }
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
                  c.convolutional ? k : k * nd + slot);
            }
        }
    }
}

}  // namespace sg

#endif
