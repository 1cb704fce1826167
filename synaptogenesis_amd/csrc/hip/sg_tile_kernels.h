// Kernel templates of the tile kernels (base activator and Izhikevich
// synapse passes), shared by sg_tile.hip (full records) and
// sg_tile_compact.hip (compact records), which instantiate them.
#ifndef SG_TILE_KERNELS_H
#define SG_TILE_KERNELS_H

// Tile kernels for every per-destination connection type (FULLY_CONNECTED,
// SUBSET, CONVERGENT, DIVERGENT): the base activator and the three
// Izhikevich synapse passes (update, activate, fused update+activate).
//
// A block owns a tile of 64 destinations.  Its W wavefronts (W = 3 or 4
// for the Izhikevich passes, chosen per launch so that the whole grid is
// resident at once where possible) split the tile's synapse list k-wise in
// chunks of 8 * W (8 visits per wavefront per chunk): every synapse's
// independent work -- loads, STDP / short-term plasticity, trace decay,
// stores -- runs with W wavefronts in flight per tile instead of one.  The
// only order-dependent part, the FP32 sum over a destination's synapses, is
// taken out of the per-synapse pass: each synapse writes the value it adds
// to a double-buffered LDS chunk, and after a barrier one wavefront per sum
// adds the chunk's values in k order.  The result is the reference's
// sequential sum, bit for bit.
//
// Per-synapse arrays are in the 64-destination tiled layout (tiled_index),
// so each wavefront load is one contiguous 256-byte line and each tile
// streams one contiguous K * 256-byte block per array.  The Izhikevich
// record is read and written once per step and is far larger than the
// caches, so its loads and stores are non-temporal: the stream does not
// evict the small, reused state (neuron arrays, spike words, other
// connections' weights) from L2 and the Infinity Cache.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sg_device.h"

namespace sg {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kPerWave = 8;      // visits per wavefront per chunk
constexpr int kSub = 8;

template <int TYPE>
__device__ __forceinline__ int visit_count(const sg_conn &c) {
    return cursor<TYPE>(c, 0).count;
}

// ---------------------------------------------------------------------------
// Base activator (synapse_kernel.h:887-912).  WAVES wavefronts share a tile
// and CHUNK visits per barrier; small destination layers use more
// wavefronts per tile to keep the GPU busy.
template <int TYPE>
__device__ __forceinline__ void ordered_sum_n(const float (*vals)[64],
                                              SynapseIter<TYPE> &walk, int n,
                                              int lane, float &sum) {
    for (int j = 0; j < n; ++j) {
        const bool valid = walk.source() >= 0;
        walk.advance();
        if (valid) sum += vals[j][lane];
    }
}

// The same sum where every skipped visit wrote +0.0f (SPARSE: a slot's
// visits past its own nonzero count).  Adding +0.0f leaves every value but
// -0.0f unchanged (quiet NaNs keep their payload), and a sum that starts at
// +0.0f is never -0.0f (x + y rounds to -0.0f only when both are -0.0f), so
// the result is bit for bit the sum over the valid visits alone -- without
// reading each visit's source index again.
__device__ __forceinline__ void dense_sum_n(const float (*vals)[64], int n,
                                            int lane, float &sum) {
    for (int j = 0; j < n; ++j) sum += vals[j][lane];
}

template <int TYPE, int OUT, int WAVES, int CHUNK>
__global__ __launch_bounds__(64 * WAVES) void tile_base_kernel(sg_conn c) {
    constexpr int PER = CHUNK / WAVES;
    __shared__ float vals[2][CHUNK][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = blockIdx.x * 64 + lane;
    const bool live = slot < c.num_dest;
    const int count = visit_count<TYPE>(c);
    const int K = c.k_per_dest;
    const bool skip = OUT == SG_OUT_BIT && c.skip_silent;
    SynapseIter<TYPE> walk;
    walk.init(c, live ? slot : 0);
    float sum = 0.0f;
    for (int k0 = 0, chunk = 0; k0 < count; k0 += CHUNK, ++chunk) {
        const int buf = chunk & 1;
        const int kw = k0 + wave * PER;
        if (live && kw < count) {
            SynapseIter<TYPE> it;
            it.seek(c, slot, kw);
            const float *wrow = c.convolutional ? c.weights + kw
                                                : c.weights +
                                                      tiled_index(slot, kw, K);
            const int wstep = c.convolutional ? 1 : 64;
            uint32_t x[PER];
            int from[PER];
#pragma unroll
            for (int b = 0; b < PER; ++b) {
                from[b] = it.k < count ? it.source() : -1;
                if (it.k < count) it.advance();
                if (from[b] >= 0) x[b] = c.src[from[b]];
                else if (TYPE == SG_SPARSE)
                    vals[buf][wave * PER + b][lane] = 0.0f;
            }
            if (skip) {
#pragma unroll
                for (int b = 0; b < PER; ++b)
                    if (from[b] >= 0) {
                        const bool on = (x[b] << (c.delay & 31)) >> 31;
                        vals[buf][wave * PER + b][lane] =
                            on ? wrow[b * wstep] : 0.0f;
                    }
            } else {
                float w[PER];
#pragma unroll
                for (int b = 0; b < PER; ++b)
                    if (from[b] >= 0) w[b] = wrow[b * wstep];
#pragma unroll
                for (int b = 0; b < PER; ++b)
                    if (from[b] >= 0)
                        vals[buf][wave * PER + b][lane] =
                            extract<OUT>(x[b], c.delay) * w[b];
            }
        }
        __syncthreads();
        if (live && wave == 0) {
            if (TYPE == SG_SPARSE)
                dense_sum_n(vals[buf], min(CHUNK, count - k0), lane, sum);
            else
                ordered_sum_n(vals[buf], walk, min(CHUNK, count - k0), lane,
                              sum);
        }
    }
    if (live && wave == 0) {
        const int to = slot_to_index(c, slot);
        put_sum(c, to, sum);
    }
}

// ---------------------------------------------------------------------------
// Izhikevich synapse passes.  UPD: STDP / iSTDP + short-term plasticity
// (izhikevich_attributes.cpp:444-508) of the previous step (fused) or of
// this step (UPD only).  ACT: conductance activator
// (izhikevich_attributes.cpp:257-284).
// Non-temporal access to field f of the record lane q points at.
__device__ __forceinline__ float rec_ld(const float *q, int f) {
    return __builtin_nontemporal_load(&q[f * 64]);
}
__device__ __forceinline__ void rec_st(float v, float *q, int f) {
    __builtin_nontemporal_store(v, &q[f * 64]);
}

// COMPACT: pre, r and sw are the source's, already stepped.
template <int SUB, int COMPACT>
__device__ __forceinline__ void tile_update(const sg_iz_syn &s, float dest_spike,
                                            float dest_exc_trace, int dest_tss,
                                            const float *src_table,
                                            const float *dst_table, int src_tss,
                                            float &w, float &r, float &sw,
                                            float &pre, float &dw) {
    const bool src_spike = src_tss == 0;
    if (!COMPACT) iz_source_step<SUB>(src_spike, s.stp_flag, pre, r, sw);
    float src_trace, dest_trace;
    if (!SUB) {
        src_trace = pre;
        dest_trace = dest_exc_trace;
    } else {
        src_trace = (dest_spike > 0.0f && src_tss > 0 && src_tss < 32)
            ? src_table[src_tss] : 0.0f;
        dest_trace = (src_spike && dest_tss > 0 && dest_tss < 32)
            ? dst_table[dest_tss] : 0.0f;
    }
    const double rhs = (((double)dw - 0.000001) * 0.0001 +
                        (double)(dest_spike * src_trace)) -
                       (double)((float)src_spike * dest_trace);
    dw = (float)((double)dw - rhs);
    const float nw = w + s.learning_rate * dw;
    w = fmaxf(0.0001f, fminf(s.max_weight, nw));
}

// COMPACT: the record has the SG_IZC_* fields; pre, R, w and the time
// since spike come from the per-source array s.src_stp.
template <int TYPE, int KIND, int SUB, int UPD, int ACT, int WAVES,
          int COMPACT>
__global__ __launch_bounds__(64 * WAVES) void iz_tile_kernel(
        sg_conn c, sg_iz_syn s, sg_iz_dest d, const float *src_table,
        const float *dst_table) {
    constexpr bool kLong = KIND == SG_IZ_ADD || KIND == SG_IZ_SUB;
    constexpr int F = COMPACT ? SG_IZC_FIELDS : SG_IZ_FIELDS;
    constexpr int FW = COMPACT ? SG_IZC_W : SG_IZ_W;
    constexpr int FDW = COMPACT ? SG_IZC_DW : SG_IZ_DW;
    constexpr int FST = COMPACT ? SG_IZC_ST : SG_IZ_ST;
    constexpr int FLT = COMPACT ? SG_IZC_LT : SG_IZ_LT;
    constexpr int CHUNK = WAVES * kPerWave;
    __shared__ float vals[2][2][CHUNK][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = blockIdx.x * 64 + lane;
    const bool live = slot < c.num_dest;
    const int to = live ? slot_to_index(c, slot) : 0;
    const int K = c.k_per_dest;
    float dest_spike = 0.0f, dest_trace = 0.0f;
    int dest_tss = 0;
    if (UPD && live) {
        dest_spike = extract<SG_OUT_BIT>(d.out_word0[to], 0);
        dest_trace = d.post_trace[to];
        dest_tss = d.time_since_spike[to];
    }
    int count = visit_count<TYPE>(c);
    // No visit is ever skipped: every k < count has a source.
    const bool noskip = TYPE == SG_FULLY_CONNECTED || TYPE == SG_SUBSET ||
        ((TYPE == SG_CONVERGENT || TYPE == SG_DIVERGENT) && c.wrap);
    // Visits below `sure` have a source in every lane of the tile.
    int sure = noskip ? count : 0;
    if (TYPE == SG_SPARSE && c.tile_counts) {
        sure = c.tile_counts[2 * blockIdx.x];
        count = c.tile_counts[2 * blockIdx.x + 1];
    }
    SynapseIter<TYPE> walk;
    walk.init(c, live ? slot : 0);
    float sum = 0.0f;
    for (int k0 = 0, chunk = 0; k0 < count; k0 += CHUNK, ++chunk) {
        const int buf = chunk & 1;
        const int kw = k0 + wave * kPerWave;
        if (live && kw < count) {
            SynapseIter<TYPE> it;
            it.seek(c, slot, kw);
            // Record of visit kw (a multiple of 8) of this lane's
            // destination: group kw / 8 of tile blockIdx.x; visit kw + b is
            // b * F * 64 further, its delay byte b * 64 further.
            float *rec = s.record +
                ((((long)(kw >> 3) * gridDim.x + blockIdx.x) * 8) * F) * 64 +
                lane;
            const uint8_t *dly = s.delays +
                (((long)(kw >> 3) * gridDim.x + blockIdx.x) * 8) * 64 + lane;
            // One batch of kSub visits.  CHECK: some visit may be skipped
            // (past the count, or a source outside a non-wrapping layer);
            // otherwise every visit is valid and no lane branches.
            auto batch = [&](int h, auto check) {
            constexpr bool CHECK = decltype(check)::value;
            float w[kSub], r[kSub], sw[kSub], pre[kSub], dw[kSub], st[kSub],
                lt[kSub];
            uint32_t x[kSub];
            int from[kSub], tss[kSub], dl[kSub];
#pragma unroll
            for (int b = 0; b < kSub; ++b) {
                if (CHECK) {
                    from[b] = it.k < count ? it.source() : -1;
                    if (it.k < count) it.advance();
                    if (from[b] < 0) continue;
                } else {
                    from[b] = it.source();
                    it.advance();
                }
                const float *q = rec + (h + b) * F * 64;
                w[b] = rec_ld(q, FW);
                if (COMPACT) {
                    const f32x4 v =
                        reinterpret_cast<const f32x4 *>(s.src_stp)[from[b]];
                    r[b] = v[0];
                    sw[b] = v[1];
                    pre[b] = v[2];
                    if (UPD) tss[b] = __float_as_int(v[3]);
                } else {
                    r[b] = rec_ld(q, SG_IZ_R);
                    sw[b] = rec_ld(q, SG_IZ_SW);
                }
                if (UPD) {
                    dw[b] = rec_ld(q, FDW);
                    if (!COMPACT) {
                        if (!SUB) pre[b] = rec_ld(q, SG_IZ_PRE);
                        tss[b] = s.tss_cur[from[b]];
                    }
                }
                if (ACT) {
                    dl[b] = dly[(h + b) * 64];
                    x[b] = c.src[from[b]];
                    st[b] = rec_ld(q, FST);
                    if (kLong) lt[b] = rec_ld(q, FLT);
                }
            }
#pragma unroll
            for (int b = 0; b < kSub; ++b) {
                if (CHECK && from[b] < 0) {
                    // SPARSE: the ordered sum adds every visit of the chunk
                    if (TYPE == SG_SPARSE && ACT) {
                        vals[buf][0][wave * kPerWave + h + b][lane] = 0.0f;
                        if (kLong)
                            vals[buf][1][wave * kPerWave + h + b][lane] = 0.0f;
                    }
                    continue;
                }
                float *q = rec + (h + b) * F * 64;
                // The updater leaves a synapse of weight below 1e-4 alone
                // and never writes a weight below it: the test is the same
                // before and after the update.  The update is computed for
                // every lane and kept where it applies; a frozen synapse
                // writes back the values it read.
                const bool active = w[b] >= 0.0001f;
                if (UPD) {
                    float w1 = w[b], r1 = r[b], sw1 = sw[b], pre1 = pre[b],
                          dw1 = dw[b];
                    tile_update<SUB, COMPACT>(s, dest_spike, dest_trace,
                                              dest_tss, src_table, dst_table,
                                              tss[b], w1, r1, sw1, pre1, dw1);
                    if (active) {
                        w[b] = w1; dw[b] = dw1; pre[b] = pre1;
                        r[b] = r1; sw[b] = sw1;
                    }
                    rec_st(w[b], q, FW);
                    rec_st(dw[b], q, FDW);
                    if (!COMPACT) {
                        if (!SUB) rec_st(pre[b], q, SG_IZ_PRE);
                        if (s.stp_flag) {
                            rec_st(r[b], q, SG_IZ_R);
                            rec_st(sw[b], q, SG_IZ_SW);
                        }
                    }
                }
                if (ACT) {
                    const float spike = extract<SG_OUT_BIT>(x[b], dl[b]);
                    // compact record, frozen synapse: R * w is its DW field
                    const float rs = (COMPACT && !active)
                        ? (UPD ? dw[b] : rec_ld(q, FDW))
                        : r[b] * sw[b];
                    const float weight = w[b] * rs;
                    const float inc = spike * weight * s.baseline_conductance;
                    const float st1 = st[b] + inc;
                    vals[buf][0][wave * kPerWave + h + b][lane] = st1;
                    rec_st(st1 * d.short_tau, q, FST);
                    if (kLong) {
                        const float lt1 = lt[b] + inc;
                        vals[buf][1][wave * kPerWave + h + b][lane] = lt1;
                        rec_st(lt1 * d.long_tau, q, FLT);
                    }
                }
            }
            };
            for (int h = 0; h < kPerWave; h += kSub) {
                if (kw + h + kSub <= sure)
                    batch(h, std::false_type{});
                else
                    batch(h, std::true_type{});
            }
        }
        if (ACT) {
            __syncthreads();
            if (live && (wave == 0 || (kLong && wave == 1))) {
                const int n = min(CHUNK, count - k0);
                // SPARSE: skipped visits wrote +0.0f (see dense_sum_n)
                if (k0 + n <= sure || TYPE == SG_SPARSE)
                    dense_sum_n(vals[buf][wave], n, lane, sum);
                else
                    ordered_sum_n(vals[buf][wave], walk, n, lane, sum);
            }
        }
    }
    if (ACT && live) {
        if (wave == 0) d.short_cond[to] += sum;
        else if (kLong && wave == 1) d.long_cond[to] += sum;
    }
}

}  // namespace sg

// Launch helpers.
namespace sg {

// Number of compute units of the current device (cached per device).
static inline int cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount,
                                  dev) != hipSuccess || n <= 0)
            n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

// Wavefronts per tile for a fused Izhikevich launch: the count that lets
// the whole grid be resident in the fewest rounds of blocks (a grid that
// overflows the GPU by a fraction runs its last round mostly idle), ties
// going to more wavefronts per tile.
template <class KA, class KB>
static inline int pick_waves(KA *four, KB *three, int tiles) {
    int nb4 = 0, nb3 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb4, reinterpret_cast<const void *>(four), 256, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb3, reinterpret_cast<const void *>(three), 192, 0) != hipSuccess ||
        nb4 <= 0 || nb3 <= 0)
        return 4;
    const long cus = cu_count();
    const long r4 = (tiles + nb4 * cus - 1) / (nb4 * cus);
    const long r3 = (tiles + nb3 * cus - 1) / (nb3 * cus);
    return r3 < r4 ? 3 : 4;
}

template <int UPD, int ACT, int COMPACT>
hipError_t launch_iz_tile_t(const sg_conn &c, const sg_iz_syn &s,
                                   const sg_iz_dest &d, int kind, int sub,
                                   const float *st_, const float *dt_,
                                   hipStream_t st) {
    const int tiles = (c.num_dest + 63) / 64;
    dim3 g(tiles);
#define SG_IZ(T, K, S)                                                       \
    do {                                                                     \
        auto *k4 = iz_tile_kernel<T, K, S, UPD, ACT, 4, COMPACT>;            \
        auto *k3 = iz_tile_kernel<T, K, S, UPD, ACT, 3, COMPACT>;            \
        if (pick_waves(k4, k3, tiles) == 3)                                  \
            hipLaunchKernelGGL(k3, g, dim3(192), 0, st, c, s, d, st_, dt_);  \
        else                                                                 \
            hipLaunchKernelGGL(k4, g, dim3(256), 0, st, c, s, d, st_, dt_);  \
    } while (0)
#define SG_IZK(T, S)                                                         \
    switch (kind) {                                                          \
        case SG_IZ_ADD: SG_IZ(T, SG_IZ_ADD, S); break;                       \
        case SG_IZ_SUB: SG_IZ(T, SG_IZ_SUB, S); break;                       \
        case SG_IZ_MULT: SG_IZ(T, SG_IZ_MULT, S); break;                     \
        case SG_IZ_REWARD: SG_IZ(T, SG_IZ_REWARD, S); break;                 \
        default: return hipErrorInvalidValue;                                \
    }
#define SG_IZT(T)                                                            \
    if (sub) { SG_IZK(T, 1) } else { SG_IZK(T, 0) }
    switch (c.type) {
        case SG_FULLY_CONNECTED: SG_IZT(SG_FULLY_CONNECTED); break;
        case SG_SUBSET: SG_IZT(SG_SUBSET); break;
        case SG_CONVERGENT: SG_IZT(SG_CONVERGENT); break;
        case SG_DIVERGENT: SG_IZT(SG_DIVERGENT); break;
        case SG_SPARSE: SG_IZT(SG_SPARSE); break;
        default: return hipErrorInvalidValue;
    }
#undef SG_IZ
#undef SG_IZK
#undef SG_IZT
    return hipGetLastError();
}

}  // namespace sg

#endif
