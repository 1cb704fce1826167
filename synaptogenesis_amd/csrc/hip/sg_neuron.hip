// Neuron-model kernels with the spike / value history shift, plus the small
// utility kernels (set, transpose).
//
// One lane per neuron; neuron state is structure-of-arrays so every access
// is a coalesced 4-byte stream.  The Izhikevich update keeps the reference's
// mixed precision exactly: the expressions that its double literals promote
// to FP64 are evaluated in FP64 here too, in the same order, and rounded to
// FP32 where the reference stores to a float (compiled with
// -ffp-contract=off, correctly rounded division, denormals preserved).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "sg_device.h"
#include "sg_hip.h"

namespace sg {

constexpr int kNeuronBlock = 256;

static inline int nblocks(long n) {
    return (int)((n + kNeuronBlock - 1) / kNeuronBlock);
}

// SHIFT_BIT_OUTPUTS (state/attributes.h:321-333): the history words of one
// neuron form one long shift register, newest bit at the MSB of word 0.
// Returns the bit that left word 0 (prev_bit).
__device__ __forceinline__ bool shift_bits(uint32_t *out, int size, int nid,
                                           int history, uint32_t new_bit) {
    uint32_t curr = out[(long)size * (history - 1) + nid];
    for (int i = history - 1; i > 0; --i) {
        const uint32_t prev = out[(long)size * (i - 1) + nid];
        out[(long)size * i + nid] = (curr >> 1) | (prev << 31);
        curr = prev;
    }
    out[nid] = (curr >> 1) | (new_bit << 31);
    return (curr >> 31) != 0;
}

// SHIFT_FLOAT_OUTPUTS (state/attributes.h:316-319).
__device__ __forceinline__ void shift_floats(float *out, int size, int nid,
                                             int history, float v) {
    for (int i = history - 1; i > 0; --i)
        out[(long)size * i + nid] = out[(long)size * (i - 1) + nid];
    out[nid] = v;
}

// Value of one init config for neuron i at `step` (engine/kernel/
// kernel.cpp:43-158 distributions, drawn from a counter-based Philox keyed
// by (seed, step, i) so runs are reproducible): dist 0 uniform[a, b),
// 1 normal(a, b) by Box-Muller, 2 poisson (val with probability a, else 0),
// 3 flat a.
__device__ __forceinline__ float init_value(int dist, float a, float b,
                                            float val, uint64_t seed,
                                            uint64_t step, int i,
                                            bool &fired) {
    fired = true;
    if (dist == 3) return a;
    uint32_t c[4] = {(uint32_t)step, (uint32_t)i, (uint32_t)(step >> 32), 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float u = (float)((double)(c[0] >> 8) * (1.0 / 16777216.0));
    if (dist == 0) return u * (b - a) + a;
    if (dist == 1) {
        const double u1 = ((double)(c[0] >> 8) + 1.0) * (1.0 / 16777216.0);
        const double u2 = (double)(c[1] >> 8) * (1.0 / 16777216.0);
        const float z =
            (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
        return z * b + a;
    }
    fired = u < a;
    return fired ? val : 0.0f;
}

// One updater step of a compact record's per-source {R, w, pre, tss}:
// tss is the time since spike the step read (the updater's next pass
// reads it from here).
__device__ __forceinline__ void source_step(float *stp, long f, int tss,
                                            int sub, int stp_flag) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    f32x4 v = reinterpret_cast<const f32x4 *>(stp)[f];
    float r = v[0], sw = v[1], pre = v[2];
    if (sub) iz_source_step<1>(tss == 0, stp_flag, pre, r, sw);
    else iz_source_step<0>(tss == 0, stp_flag, pre, r, sw);
    v[0] = r;
    v[1] = sw;
    v[2] = pre;
    v[3] = __int_as_float(tss);
    reinterpret_cast<f32x4 *>(stp)[f] = v;
}

// iz_attribute_kernel (state/impl/izhikevich_attributes.cpp:102-213) for
// neuron nid of layer l, plus the optional side outputs of sg_iz_layer:
// the init-config value, the outgoing connections' time since spike (from
// the words before the shift) and the plane of the new spike bits.  Every
// lane of the wavefront calls it (the plane is a wavefront ballot);
// nid >= l.size only takes part in the ballot.
__device__ __forceinline__ void iz_step(const sg_iz_layer &l, int nid) {
    const bool valid = nid < l.size;
    uint32_t spike = 0u;
    if (valid) {
        const float ampa = l.ampa[nid];
        const float nmda = l.nmda[nid];
        const float gabaa = l.gabaa[nid];
        const float gabab = l.gabab[nid];
        const float mult = l.mult[nid];
        float voltage = l.voltage[nid];
        float recovery = l.recovery[nid];
        float base_current = l.input[nid];
        if (l.noise) {
            bool fired;
            const uint64_t step = l.noise_base
                ? *l.noise_base + l.noise_step : l.noise_step;
            const float v = init_value(l.noise_dist, l.noise_a, l.noise_b,
                                       l.noise_val, l.noise_seed, step, nid,
                                       fired);
            base_current = l.noise == 1
                ? v : aggregate(l.noise_op, v, base_current);
            l.input[nid] = base_current;
        }
        const float a = l.a[nid];
        const float b = l.b[nid];
        // every load of the rest of the step, issued before the loop: the
        // stores in between could alias them, so the compiler would not
        // move them up by itself
        const float cc = l.c[nid], dd = l.d[nid];
        const float dopamine = l.dopamine[nid];
        const float post_trace = l.post_trace[nid];
        const int tss_self = l.time_since_spike[nid];
        int t_cur[SG_MAX_TSS_OUT], t_vis[SG_MAX_TSS_OUT];
        uint32_t t_word[SG_MAX_TSS_OUT], t_rev[SG_MAX_TSS_OUT][8];
#pragma unroll
        for (int t = 0; t < SG_MAX_TSS_OUT; ++t) {
            if (t >= l.ntss) break;
            const sg_tss_out &o = l.tss[t];
            t_cur[t] = o.cur[nid];
            t_vis[t] = o.visits[nid];
            t_word[t] = l.output[(long)o.word * l.size + nid];
            const uint4 *rv =
                reinterpret_cast<const uint4 *>(o.rev_delays + (long)nid * 32);
            const uint4 r0 = rv[0], r1 = rv[1];
            t_rev[t][0] = r0.x; t_rev[t][1] = r0.y;
            t_rev[t][2] = r0.z; t_rev[t][3] = r0.w;
            t_rev[t][4] = r1.x; t_rev[t][5] = r1.y;
            t_rev[t][6] = r1.z; t_rev[t][7] = r1.w;
        }

        for (int i = 0; i < 10 && (double)voltage < 30.0; ++i) {
            float current = -ampa * voltage;
            // powf(x, 2) is folded to x * x by the reference's compiler.
            const float x = (voltage + 80.0f) / 60.0f;
            const float temp = x * x;
            current = current - nmda * (temp / (1.0f + temp)) * voltage;
            current = current - gabaa * (voltage + 70.0f);
            current = current - gabab * (voltage + 90.0f);
            current = current * (1.0f + mult);
            current = current + base_current;

            const double dv = (0.04 * (double)voltage) * (double)voltage +
                              (double)(5.0f * voltage) + 140.0 -
                              (double)recovery + (double)current;
            const float delta_v = (float)dv;

            const float sum_g = ampa + nmda + gabaa + gabab;
            const float drive = delta_v + (gabaa * -70.0f) + (gabab * -90.0f);
            voltage = (float)(((double)voltage + 0.1 * (double)drive) /
                              (1.0 + 0.1 * (double)sum_g));
            voltage = (voltage != voltage) ? 30.0f : voltage;

            const float adjusted_tau = ((double)voltage > 30.0)
                ? (float)((double)delta_v /
                          (30.0 - (double)voltage + (double)delta_v) * 0.1)
                : (float)0.1;
            recovery = recovery + a * adjusted_tau * ((b * voltage) - recovery);
        }

        l.ampa[nid] = 0.0f;
        l.nmda[nid] = 0.0f;
        l.gabaa[nid] = 0.0f;
        l.gabab[nid] = 0.0f;
        l.mult[nid] = 0.0f;
        l.dopamine[nid] = (float)((double)dopamine * 0.95);

#pragma unroll
        for (int t = 0; t < SG_MAX_TSS_OUT; ++t) {
            if (t >= l.ntss) break;
            const sg_tss_out &o = l.tss[t];
            const int tn = tss_next(t_cur[t], t_vis[t], t_rev[t], t_word[t]);
            o.next[nid] = tn;
            if (o.stp) source_step(o.stp, nid, tn, o.stp_sub, o.stp_flag);
        }

        spike = (double)voltage >= 30.0 ? 1u : 0u;
        const bool prev_bit =
            shift_bits(l.output, l.size, nid, l.history, spike);

        l.post_trace[nid] = prev_bit
            ? (float)0.004
            : (float)((double)post_trace * 0.95);
        l.time_since_spike[nid] = prev_bit ? 0 : min(32, tss_self + 1);
        l.voltage[nid] = spike ? cc : voltage;
        l.recovery[nid] = recovery + ((float)spike * dd);
    }
    if (l.plane) {
        const uint64_t m = __ballot(spike != 0u);
        // lane 0 of a wavefront: nid is a multiple of 64
        if ((threadIdx.x & 63) == 0 && valid) l.plane[nid >> 6] = m;
    }
}

// Up to SG_MAX_IZ_BATCH layers in one grid: layer j owns blocks
// first[j] .. first[j + 1] - 1.
struct IzBatch {
    int n;
    int first[SG_MAX_IZ_BATCH + 1];
    sg_iz_layer l[SG_MAX_IZ_BATCH];
};

__global__ __launch_bounds__(256) void source_update_kernel(
        float *stp, const int32_t *tss, int n, int sub, int stp_flag) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f < n) source_step(stp, f, tss[f], sub, stp_flag);
}

__global__ __launch_bounds__(kNeuronBlock) void iz_neuron_kernel(IzBatch b) {
    int j = 0;
    while (j + 1 < b.n && (int)blockIdx.x >= b.first[j + 1]) ++j;
    iz_step(b.l[j], (blockIdx.x - b.first[j]) * kNeuronBlock + threadIdx.x);
}

// gol_attribute_kernel (state/impl/game_of_life_attributes.cpp:14-33):
// survival reads bit 0 of word 0, the state 32 steps back.
__global__ __launch_bounds__(kNeuronBlock) void gol_neuron_kernel(
        const float *input, uint32_t *output, int size, int history,
        int smin, int smax, int bmin, int bmax) {
    const int nid = blockIdx.x * blockDim.x + threadIdx.x;
    if (nid >= size) return;
    const int in = (int)input[nid];
    const bool prev_live = output[nid] & 1u;
    const bool live = prev_live ? (in >= smin && in <= smax)
                                : (in >= bmin && in <= bmax);
    shift_bits(output, size, nid, history, live ? 1u : 0u);
}

// re_attribute_kernel (state/impl/rate_encoding_attributes.cpp:14-23):
// out = input > 0 ? tanh(input) : 0.  tanh is evaluated in FP64 and
// rounded once to FP32.
__global__ __launch_bounds__(kNeuronBlock) void rate_neuron_kernel(
        const float *input, float *output, int size, int history) {
    const int nid = blockIdx.x * blockDim.x + threadIdx.x;
    if (nid >= size) return;
    const float in = input[nid];
    const float v = ((double)in > 0.0) ? (float)tanh((double)in) : 0.0f;
    shift_floats(output, size, nid, history, v);
}

// relay_attribute_kernel (state/impl/relay_attributes.cpp:10-18).
__global__ __launch_bounds__(kNeuronBlock) void relay_neuron_kernel(
        const float *input, float *output, int size, int history, int ramp) {
    const int nid = blockIdx.x * blockDim.x + threadIdx.x;
    if (nid >= size) return;
    const float in = input[nid];
    const float v = ramp ? (float)fmax(0.0, (double)in) : in;
    shift_floats(output, size, nid, history, v);
}

__global__ __launch_bounds__(kNeuronBlock) void set_f32_kernel(
        float *p, float val, long n, int overwrite) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = overwrite ? val : p[i] + val;
}

__global__ __launch_bounds__(kNeuronBlock) void init_random_kernel(
        float *p, int n, int dist, float a, float b, float val, int overwrite,
        uint64_t seed, uint64_t step, const uint64_t *step_base) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (step_base) step += *step_base;
    bool fired;
    const float v = init_value(dist, a, b, val, seed, step, i, fired);
    if (dist == 2 && !overwrite) {
        if (fired) p[i] += v;               // poisson adds only on a spike
    } else {
        p[i] = overwrite ? v : p[i] + v;
    }
}

// 32-bit transpose through a padded LDS tile.
constexpr int kTile = 64;
constexpr int kRowsPerPass = 4;

__global__ __launch_bounds__(kTile * kRowsPerPass) void transpose32_kernel(
        const uint32_t *src, uint32_t *dst, int rows, int cols) {
    __shared__ uint32_t tile[kTile][kTile + 1];
    const int bx = blockIdx.x * kTile, by = blockIdx.y * kTile;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int j = ty; j < kTile; j += kRowsPerPass) {
        const int r = by + j, c = bx + tx;
        if (r < rows && c < cols) tile[j][tx] = src[(long)r * cols + c];
    }
    __syncthreads();
    for (int j = ty; j < kTile; j += kRowsPerPass) {
        const int r = bx + j, c = by + tx;   // output row = input column
        if (r < cols && c < rows) dst[(long)r * rows + c] = tile[tx][j];
    }
}

__global__ __launch_bounds__(kNeuronBlock) void tile32_kernel(
        const uint32_t *src, uint32_t *dst, int slots, int K, int nfields,
        int field, int to_tiled) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tiles = (slots + 63) / 64;
    if (i >= tiles * 64 * K) return;
    // i enumerates one field of the tiled buffer: tile, k, lane
    const int lane = (int)(i & 63);
    const long tk = i >> 6;
    const int k = (int)(tk % K);
    const long tile = tk / K;
    const long slot = tile * 64 + lane;
    const long t = ((tile * K + k) * nfields + field) * 64 + lane;
    if (slot >= slots) {
        if (to_tiled) dst[t] = 0u;
        return;
    }
    const long h = slot * K + k;
    if (to_tiled) dst[t] = src[h];
    else dst[h] = src[t];
}

__global__ __launch_bounds__(kNeuronBlock) void record32_kernel(
        const uint32_t *src, uint32_t *dst, int slots, int K, int field,
        int nfields, int to_record) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tiles = (slots + 63) / 64;
    if (i >= tiles * 64 * K) return;
    // i enumerates (tile, k, lane) of one field
    const int lane = (int)(i & 63);
    const long tk = i >> 6;
    const int k = (int)(tk % K);
    const long tile = tk / K;
    const long slot = tile * 64 + lane;
    const long r = ((((long)(k >> 3) * tiles + tile) * 8 + (k & 7)) *
                        nfields + field) * 64 + lane;
    if (slot >= slots) {
        if (to_record) dst[r] = 0u;
        return;
    }
    const long h = slot * K + k;
    if (to_record) dst[r] = src[h];
    else dst[h] = src[r];
}

__global__ __launch_bounds__(kNeuronBlock) void copy32_kernel(
        const uint4 *src, uint4 *dst, long n4) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Any n: 16-byte words where both sides allow it, then the tail.
__global__ __launch_bounds__(kNeuronBlock) void copy_u32_kernel(
        const uint32_t *src, uint32_t *dst, long n, int vec) {
    const long stride = (long)gridDim.x * blockDim.x;
    const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long done = 0;
    if (vec) {
        const long n4 = n / 4;
        for (long i = i0; i < n4; i += stride)
            ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
        done = n4 * 4;
    }
    for (long i = done + i0; i < n; i += stride) dst[i] = src[i];
}

}  // namespace sg

using namespace sg;

extern "C" int sg_hip_iz_neurons(const sg_iz_layer *layers, int n,
                                 void *stream) {
    for (int s0 = 0; s0 < n; s0 += SG_MAX_IZ_BATCH) {
        IzBatch b;
        b.n = 0;
        int blocks = 0;
        for (int j = s0; j < n && j < s0 + SG_MAX_IZ_BATCH; ++j) {
            if (layers[j].size <= 0) continue;
            if (layers[j].ntss < 0 || layers[j].ntss > SG_MAX_TSS_OUT)
                return (int)hipErrorInvalidValue;
            b.l[b.n] = layers[j];
            b.first[b.n] = blocks;
            blocks += nblocks(layers[j].size);
            ++b.n;
        }
        if (b.n == 0) continue;
        b.first[b.n] = blocks;
        hipLaunchKernelGGL(iz_neuron_kernel, dim3(blocks), dim3(kNeuronBlock),
                           0, (hipStream_t)stream, b);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    return 0;
}

extern "C" int sg_hip_iz_neuron(const sg_iz_layer *l, void *stream) {
    return sg_hip_iz_neurons(l, 1, stream);
}

extern "C" int sg_hip_gol_neuron(float *input, uint32_t *output, int size,
                                 int history, int smin, int smax, int bmin,
                                 int bmax, void *stream) {
    if (size <= 0) return 0;
    hipLaunchKernelGGL(gol_neuron_kernel, dim3(nblocks(size)),
                       dim3(kNeuronBlock), 0, (hipStream_t)stream, input,
                       output, size, history, smin, smax, bmin, bmax);
    return hipGetLastError();
}

extern "C" int sg_hip_rate_neuron(float *input, float *output, int size,
                                  int history, void *stream) {
    if (size <= 0) return 0;
    hipLaunchKernelGGL(rate_neuron_kernel, dim3(nblocks(size)),
                       dim3(kNeuronBlock), 0, (hipStream_t)stream, input,
                       output, size, history);
    return hipGetLastError();
}

extern "C" int sg_hip_relay_neuron(float *input, float *output, int size,
                                   int history, int ramp, void *stream) {
    if (size <= 0) return 0;
    hipLaunchKernelGGL(relay_neuron_kernel, dim3(nblocks(size)),
                       dim3(kNeuronBlock), 0, (hipStream_t)stream, input,
                       output, size, history, ramp);
    return hipGetLastError();
}

extern "C" int sg_hip_init_random(float *p, int n, int dist, float a,
                                  float b, float val, int overwrite,
                                  uint64_t seed, uint64_t step,
                                  const uint64_t *step_base, void *stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(init_random_kernel, dim3(nblocks(n)),
                       dim3(kNeuronBlock), 0, (hipStream_t)stream, p, n, dist,
                       a, b, val, overwrite, seed, step, step_base);
    return hipGetLastError();
}

__global__ void add_u64_kernel(uint64_t *counter, uint64_t add) {
    if (threadIdx.x == 0) *counter += add;
}

extern "C" int sg_hip_add_u64(uint64_t *counter, uint64_t add, void *stream) {
    hipLaunchKernelGGL(add_u64_kernel, dim3(1), dim3(64), 0,
                       (hipStream_t)stream, counter, add);
    return hipGetLastError();
}

extern "C" int sg_hip_set_f32(float *p, float val, long n, int overwrite,
                              void *stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(set_f32_kernel, dim3(nblocks(n)), dim3(kNeuronBlock),
                       0, (hipStream_t)stream, p, val, n, overwrite);
    return hipGetLastError();
}

extern "C" int sg_hip_transpose32(const void *src, void *dst, int rows,
                                  int cols, void *stream) {
    if (rows <= 0 || cols <= 0) return 0;
    dim3 g((cols + kTile - 1) / kTile, (rows + kTile - 1) / kTile);
    hipLaunchKernelGGL(transpose32_kernel, g, dim3(kTile, kRowsPerPass), 0,
                       (hipStream_t)stream, (const uint32_t *)src,
                       (uint32_t *)dst, rows, cols);
    return hipGetLastError();
}

extern "C" int sg_hip_tile32(const void *src, void *dst, int slots, int K,
                             int nfields, int field, int to_tiled,
                             void *stream) {
    if (slots <= 0 || K <= 0 || nfields <= 0 || field < 0 || field >= nfields)
        return slots <= 0 || K <= 0 ? 0 : (int)hipErrorInvalidValue;
    const long n = (long)((slots + 63) / 64) * 64 * K;
    hipLaunchKernelGGL(tile32_kernel, dim3(nblocks(n)), dim3(kNeuronBlock), 0,
                       (hipStream_t)stream, (const uint32_t *)src,
                       (uint32_t *)dst, slots, K, nfields, field, to_tiled);
    return hipGetLastError();
}

extern "C" int sg_hip_record32(const void *src, void *dst, int slots,
                               int K, int field, int nfields, int to_record,
                               void *stream) {
    if (slots <= 0 || K <= 0) return 0;
    if ((nfields != SG_IZ_FIELDS && nfields != SG_IZC_FIELDS) || field < 0 ||
        field >= nfields)
        return (int)hipErrorInvalidValue;
    const long n = (long)((slots + 63) / 64) * 64 * K;
    hipLaunchKernelGGL(record32_kernel, dim3(nblocks(n)), dim3(kNeuronBlock),
                       0, (hipStream_t)stream, (const uint32_t *)src,
                       (uint32_t *)dst, slots, K, field, nfields, to_record);
    return hipGetLastError();
}

extern "C" int sg_hip_iz_source_update(float *stp, const int32_t *tss, int n,
                                       int sub, int stp_flag, void *stream) {
    if (n <= 0) return 0;
    if (!stp || !tss || ((uintptr_t)stp & 15) != 0)
        return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(source_update_kernel, dim3((n + 255) / 256), dim3(256),
                       0, (hipStream_t)stream, stp, tss, n, sub, stp_flag);
    return hipGetLastError();
}

extern "C" int sg_hip_copy32(const void *src, void *dst, long n,
                             void *stream) {
    if (n <= 0) return 0;
    if (n % 4 != 0 || ((uintptr_t)src | (uintptr_t)dst) % 16 != 0)
        return (int)hipErrorInvalidValue;
    const long n4 = n / 4;
    const int blocks = (int)std::min<long>((n4 + kNeuronBlock - 1) /
                                           kNeuronBlock, 256L * 64);
    hipLaunchKernelGGL(copy32_kernel, dim3(blocks), dim3(kNeuronBlock), 0,
                       (hipStream_t)stream, (const uint4 *)src, (uint4 *)dst,
                       n4);
    return hipGetLastError();
}

extern "C" int sg_hip_copy_u32(const void *src, void *dst, long n,
                               void *stream) {
    if (n <= 0) return 0;
    if (((uintptr_t)src | (uintptr_t)dst) % 4 != 0)
        return (int)hipErrorInvalidValue;
    const int vec = ((uintptr_t)src | (uintptr_t)dst) % 16 == 0;
    const long units = vec ? n / 4 + 3 : n;
    const int blocks = (int)std::min<long>((units + kNeuronBlock - 1) /
                                           kNeuronBlock, 1024L);
    hipLaunchKernelGGL(copy_u32_kernel, dim3(blocks), dim3(kNeuronBlock), 0,
                       (hipStream_t)stream, (const uint32_t *)src,
                       (uint32_t *)dst, n, vec);
    return hipGetLastError();
}

extern "C" int sg_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return -1;
    return n;
}

extern "C" const char *sg_hip_error_string(int code) {
    if (code == SG_HIP_ERROR_NOALLOC_LIMIT)
        return "connection has more sources than the no-allocation "
               "activator's plane holds (SG_NOALLOC_PLANE_WORDS * 64)";
    return hipGetErrorString((hipError_t)code);
}
