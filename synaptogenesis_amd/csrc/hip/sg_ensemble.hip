// Batched rate-encoded ensembles (include/sg_ensemble.h): every row of every
// instance's two fully connected layers in one launch per step.
//
// BF16 weights live in HBM already in the matrix cores' operand order: a
// layer of R rows and fan-in K is cut into tiles of 16 rows x 32 columns,
// and tile (row tile rt, column tile kt) is 1 KB at ((rt * Kt + kt) * 64 +
// lane) * 16 B, where lane l holds row rt*16 + l%16, columns kt*32 +
// (l/16)*8 + 0..7 -- the A operand of v_mfma_f32_16x16x32_bf16.  A
// wavefront streams one row tile's Kt tiles as contiguous 1 KB loads.
//
// The input vector is staged in LDS as three BF16 terms per element (hi =
// bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)); they are the B
// operand's columns 0, 1 and 2, so one MFMA yields W.hi, W.mid and W.lo and
// their FP32 sum carries the input at FP32 precision.  The other thirteen
// columns are zero: the sums are HBM-bound, the matrix cores are not.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "sg_device.h"
#include "sg_ensemble.h"

namespace sg {
namespace ens {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kTileRows = 16;   // BF16 row tile
constexpr int kTileCols = 32;   // BF16 column tile
constexpr int kGroupRows = 4;   // FP32 rows per wavefront

// One (instance, layer) product W . x with its state update.
struct Layer {
    const void *w;       // BF16 tiles, or FP32 [rows padded to 4][Kp]
    const float *x[2];   // input, by parity
    float *h[2];         // state: h[p] is read, h[p ^ 1] written
    int R, K, Kp;        // rows, fan-in, fan-in padded to 32
    int tiles;           // BF16: row tiles of 16; FP32: row groups of 4
};

struct Item {
    int layer;
    int tile0;
};

__host__ __device__ inline uint32_t bf16_rne(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

__host__ __device__ inline float bf16_value(uint32_t b) {
    const uint32_t u = b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ float rate(float y, int rectify) {
    if (rectify) return y > 0.0f ? tanhf(y) : 0.0f;
    return tanhf(y);
}

// One work item: RT row tiles of one layer, W . x and the state update,
// every thread of the block taking part (it stages the input in `lds`).
template <int RT, int KS>
__device__ __forceinline__ void bf16_item(const Layer *__restrict__ layers,
                                          const Item it, int parity,
                                          float leak, int rectify,
                                          unsigned char *lds) {
    // Fields are read through the pointer: a local copy of the descriptor
    // indexed by the run-time parity would live in scratch memory.
    const Layer &L = layers[it.layer];
    const int Kp = L.Kp;
    uint16_t *xs = (uint16_t *)lds;   // [3][Kp]
    const float *x = L.x[parity];
    for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
        const float v = k < L.K ? x[k] : 0.0f;
        const uint32_t hi = bf16_rne(v);
        const float r = v - bf16_value(hi);
        const uint32_t mid = bf16_rne(r);
        const uint32_t lo = bf16_rne(r - bf16_value(mid));
        xs[k] = (uint16_t)hi;
        xs[Kp + k] = (uint16_t)mid;
        xs[2 * Kp + k] = (uint16_t)lo;
    }
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rt = it.tile0 + wave / KS;
    const int part = wave % KS;
    const int j = lane & 15, g = lane >> 4;
    const int Kt = Kp / kTileCols;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (rt < L.tiles) {
        const int kt0 = (int)((long)part * Kt / KS);
        const int kt1 = (int)((long)(part + 1) * Kt / KS);
        const u32x4 *wp = (const u32x4 *)L.w + (long)rt * Kt * 64 + lane;
        const bool live = j < 3;
        const uint16_t *xb = xs + (live ? j : 0) * Kp + g * 8;
        const bf16x8 zero = {};
        int kt = kt0;
        for (; kt + 8 <= kt1; kt += 8) {
            u32x4 a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                a[u] = __builtin_nontemporal_load(wp + (long)(kt + u) * 64);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bf16x8 b =
                    live ? *(const bf16x8 *)(xb + (kt + u) * kTileCols) : zero;
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(bf16x8, a[u]), b, acc, 0, 0, 0);
            }
        }
        for (; kt < kt1; ++kt) {
            const u32x4 a = __builtin_nontemporal_load(wp + (long)kt * 64);
            const bf16x8 b =
                live ? *(const bf16x8 *)(xb + kt * kTileCols) : zero;
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                __builtin_bit_cast(bf16x8, a), b, acc, 0, 0, 0);
        }
    }
    if (KS > 1) {
        // partial sums of the fan-in slices, added in slice order
        f32x4 *parts = (f32x4 *)(lds + ((6 * Kp + 15) & ~15));
        parts[threadIdx.x] = acc;
        __syncthreads();
        if (part == 0)
            for (int p = 1; p < KS; ++p) acc += parts[threadIdx.x + p * 64];
    }
    if (part != 0 || rt >= L.tiles) return;
    // lane (g, j) holds column j of rows 4g..4g+3: y = (hi + mid) + lo
    f32x4 y;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float m = __shfl_down(acc[i], 1);
        const float l = __shfl_down(acc[i], 2);
        y[i] = (acc[i] + m) + l;
    }
    if (j != 0) return;
    const float keep = 1.0f - leak;
    const float *hs = L.h[parity];
    float *hd = L.h[parity ^ 1];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = rt * kTileRows + 4 * g + i;
        if (row < L.R) hd[row] = keep * hs[row] + leak * rate(y[i], rectify);
    }
}

template <int RT, int KS>
__global__ __launch_bounds__(64 * RT * KS) void ens_bf16_kernel(
    const Layer *__restrict__ layers, const Item *__restrict__ items,
    int parity, float leak, int rectify) {
    extern __shared__ __align__(16) unsigned char lds[];
    bf16_item<RT, KS>(layers, items[blockIdx.x], parity, leak, rectify, lds);
}

// Grid barrier of the persistent kernel, on two words of device memory
// (bar[0] arrivals, bar[1] generation): the last block to arrive clears
// the arrivals and advances the generation; the others wait for the
// generation to change.  The launch keeps every block resident (a
// cooperative launch, or a graph node with the cooperative attribute), so
// every block arrives; still, a wait gives up after about a second of
// polling and counts a timeout in bar[2] (the run's results are then
// wrong, and the host reports it) rather than spinning for ever.
constexpr long kBarrierPolls = 1L << 24;

__device__ __forceinline__ void grid_barrier(unsigned int *bar,
                                             unsigned int blocks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int gen = __hip_atomic_load(
            bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned int arrived = __hip_atomic_fetch_add(
            bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == blocks - 1) {
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long polls = 0;
            while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE,
                                     __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (++polls > kBarrierPolls) {
                    __hip_atomic_fetch_add(bar + 2, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// Every step of a run in one launch: the blocks, all resident, take the
// step's items from an atomic queue (longest fan-in first, as the list is
// ordered), then meet at the grid barrier, so that the next step reads
// complete states.  Step t uses queue[t & 1]; block 0 clears it after the
// barrier that ends step t, and the barrier that ends step t + 1 orders
// the clear before its next use.  Every block runs the same number of
// barriers.  sync = {queue[2], barrier[3]}.
template <int RT, int KS>
__global__ __launch_bounds__(64 * RT * KS) void ens_bf16_persistent(
    const Layer *__restrict__ layers, const Item *__restrict__ items,
    int n_items, int parity, int steps, float leak, int rectify,
    unsigned int *sync) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int next;
    unsigned int *queue = sync, *bar = sync + 2;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        queue[0] = 0u;
        queue[1] = 0u;
    }
    grid_barrier(bar, gridDim.x);
    for (int t = 0; t < steps; ++t) {
        unsigned int *q = queue + (t & 1);
        if (threadIdx.x == 0) next = (int)atomicAdd(q, 1u);
        __syncthreads();
        int i = next;
        __syncthreads();
        while (i < n_items) {
            // the next item's index is fetched while this one is computed
            // (thread 0 keeps it in a register until the item is done)
            unsigned int ahead = 0u;
            if (threadIdx.x == 0) ahead = atomicAdd(q, 1u);
            bf16_item<RT, KS>(layers, items[i], parity ^ (t & 1), leak,
                              rectify, lds);
            __syncthreads();
            if (threadIdx.x == 0) next = (int)ahead;
            __syncthreads();
            i = next;
        }
        grid_barrier(bar, gridDim.x);
        if (blockIdx.x == 0 && threadIdx.x == 0) queue[t & 1] = 0u;
    }
}

// FP32 path: each wavefront owns four rows; lanes stride the fan-in in
// float4 steps, each row's partial sums reduce across the wavefront.
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void ens_fp32_kernel(
    const Layer *__restrict__ layers, const Item *__restrict__ items,
    int parity, float leak, int rectify) {
    extern __shared__ __align__(16) unsigned char lds[];
    const Item it = items[blockIdx.x];
    const Layer &L = layers[it.layer];
    const int Kp = L.Kp;
    float *xs = (float *)lds;
    const float *x = L.x[parity];
    for (int k = threadIdx.x; k < Kp; k += blockDim.x)
        xs[k] = k < L.K ? x[k] : 0.0f;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int grp = it.tile0 + wave;
    if (grp >= L.tiles) return;
    const int K4 = Kp / 4;
    const f32x4 *w = (const f32x4 *)L.w + (long)grp * kGroupRows * K4;
    const f32x4 *x4 = (const f32x4 *)xs;
    float acc[kGroupRows] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = lane; k < K4; k += 64) {
        const f32x4 xv = x4[k];
        f32x4 wv[kGroupRows];
#pragma unroll
        for (int r = 0; r < kGroupRows; ++r)
            wv[r] = __builtin_nontemporal_load(w + (long)r * K4 + k);
#pragma unroll
        for (int r = 0; r < kGroupRows; ++r) {
            acc[r] += wv[r].x * xv.x;
            acc[r] += wv[r].y * xv.y;
            acc[r] += wv[r].z * xv.z;
            acc[r] += wv[r].w * xv.w;
        }
    }
#pragma unroll
    for (int r = 0; r < kGroupRows; ++r)
        for (int o = 32; o > 0; o >>= 1) acc[r] += __shfl_xor(acc[r], o);
    if (lane >= kGroupRows) return;
    const int row = grp * kGroupRows + lane;
    if (row >= L.R) return;
    float y = acc[0];
    for (int r = 1; r < kGroupRows; ++r)
        if (lane == r) y = acc[r];
    const float *hs = L.h[parity];
    float *hd = L.h[parity ^ 1];
    hd[row] = (1.0f - leak) * hs[row] + leak * rate(y, rectify);
}

// Synthetic data, Philox key (seed, workload), counter (0, index, stream).
__device__ __forceinline__ float normal_value(uint32_t index, uint32_t stream,
                                              uint32_t k0, uint32_t k1,
                                              double sd) {
    uint32_t c[4] = {0u, index, stream, 0u};
    philox(c, k0, k1);
    const double u1 = ((double)(c[0] >> 8) + 1.0) * (1.0 / 16777216.0);
    const double u2 = (double)(c[1] >> 8) * (1.0 / 16777216.0);
    return (float)(sd * (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2)));
}

__global__ void ens_init_bf16(uint16_t *w, int R, int K, int Kt, long total,
                              uint32_t k0, uint32_t k1, uint32_t stream,
                              double sd) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const long tile = i >> 9;
        const int within = (int)(i & 511), lane = within >> 3, e = within & 7;
        const int row = (int)(tile / Kt) * kTileRows + (lane & 15);
        const int col = (int)(tile % Kt) * kTileCols + (lane >> 4) * 8 + e;
        const float v = (row < R && col < K)
            ? normal_value((uint32_t)row * (uint32_t)K + (uint32_t)col,
                           stream, k0, k1, sd)
            : 0.0f;
        w[i] = (uint16_t)bf16_rne(v);
    }
}

__global__ void ens_init_fp32(float *w, int R, int K, int Kp, long total,
                              uint32_t k0, uint32_t k1, uint32_t stream,
                              double sd) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const int row = (int)(i / Kp), col = (int)(i % Kp);
        w[i] = (row < R && col < K)
            ? normal_value((uint32_t)row * (uint32_t)K + (uint32_t)col,
                           stream, k0, k1, sd)
            : 0.0f;
    }
}

__global__ void ens_init_uniform(float *x, int n, uint32_t k0, uint32_t k1,
                                 uint32_t stream) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c[4] = {0u, (uint32_t)i, stream, 0u};
    philox(c, k0, k1);
    x[i] = (float)((double)(c[0] >> 8) * (1.0 / 16777216.0));
}

}  // namespace ens
}  // namespace sg

using namespace sg::ens;

namespace {

thread_local std::string g_error;

int fail(const std::string &msg) {
    g_error = msg;
    return 1;
}

int hip_fail(hipError_t e, const char *what) {
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return (int)e;
}

#define ENS_CHECK(call, what)                           \
    do {                                                \
        hipError_t err_ = (call);                       \
        if (err_ != hipSuccess) return hip_fail(err_, what); \
    } while (0)

long ceil_to(long x, long m) { return (x + m - 1) / m * m; }

}  // namespace

struct sg_ens {
    int n = 0, precision = SG_ENS_BF16, rectify = 0;
    float leak = 1.0f;
    uint64_t seed = 0;
    uint32_t workload = 0;
    std::vector<int> widths;
    hipStream_t stream = nullptr;
    // stream of the last sg_ens_run (the caller's, or `stream`): work there
    // may still read the items, the layers and the state
    hipStream_t last = nullptr;
    // per instance: state offsets (floats) of x0, h1[2], h2[2]
    std::vector<long> x0_off, h1_off[2], h2_off[2];
    std::vector<long> w_off;          // per (instance, layer - 1), bytes
    std::vector<Layer> layers;        // 2 per instance
    Layer *d_layers = nullptr;
    std::vector<Item> items;
    Item *d_items = nullptr;
    size_t d_items_cap = 0;
    unsigned char *d_w = nullptr;
    size_t w_bytes = 0;
    float *d_state = nullptr, *d_backup = nullptr;
    size_t state_floats = 0;
    int kp_max = 32;
    int rt = 4, ks = 1;
    int parity = 0;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (steps, parity)
    std::map<std::pair<int, int>, hipGraph_t> graph_defs;
    // persistent runs: the work queue, the grid of the last launch, and the
    // graphs holding one persistent launch (by steps and parity)
    unsigned int *d_queue = nullptr;
    int persist_blocks = 0;
    int persist_captured = 0;
    std::map<std::pair<int, int>, hipGraphExec_t> pgraphs;
    std::map<std::pair<int, int>, hipGraph_t> pgraph_defs;
};

namespace {

bool bf16_shape_ok(int rt, int ks) {
    const int waves = rt * ks;
    return rt >= 1 && ks >= 1 && waves <= 16 && (waves & (waves - 1)) == 0;
}

size_t lds_bytes(const sg_ens *e) {
    if (e->precision == SG_ENS_FP32) return (size_t)e->kp_max * 4;
    size_t b = ((size_t)6 * e->kp_max + 15) & ~(size_t)15;
    if (e->ks > 1) b += (size_t)64 * e->rt * e->ks * 16;
    return b;
}

template <int RT, int KS>
void launch_bf16_t(const sg_ens *e, int parity, hipStream_t s) {
    hipLaunchKernelGGL((ens_bf16_kernel<RT, KS>), dim3((unsigned)e->items.size()),
                       dim3(64 * RT * KS), lds_bytes(e), s, e->d_layers,
                       e->d_items, parity, e->leak, e->rectify);
}

template <int W>
void launch_fp32_t(const sg_ens *e, int parity, hipStream_t s) {
    hipLaunchKernelGGL((ens_fp32_kernel<W>), dim3((unsigned)e->items.size()),
                       dim3(64 * W), lds_bytes(e), s, e->d_layers, e->d_items,
                       parity, e->leak, e->rectify);
}

// One step: every row of every instance, reading state `parity`.
int launch_step(const sg_ens *e, int parity, hipStream_t s) {
    if (e->items.empty()) return 0;
    if (e->precision == SG_ENS_FP32) {
        switch (e->rt) {
            case 1: launch_fp32_t<1>(e, parity, s); break;
            case 2: launch_fp32_t<2>(e, parity, s); break;
            case 4: launch_fp32_t<4>(e, parity, s); break;
            case 8: launch_fp32_t<8>(e, parity, s); break;
            case 16: launch_fp32_t<16>(e, parity, s); break;
            default: return fail("unsupported FP32 launch shape");
        }
    } else {
#define ENS_CASE(R, K) \
    if (e->rt == R && e->ks == K) { launch_bf16_t<R, K>(e, parity, s); return 0; }
        ENS_CASE(1, 1) ENS_CASE(1, 2) ENS_CASE(1, 4) ENS_CASE(1, 8)
        ENS_CASE(1, 16) ENS_CASE(2, 1) ENS_CASE(2, 2) ENS_CASE(2, 4)
        ENS_CASE(2, 8) ENS_CASE(4, 1) ENS_CASE(4, 2) ENS_CASE(4, 4)
        ENS_CASE(8, 1) ENS_CASE(8, 2) ENS_CASE(16, 1)
#undef ENS_CASE
        return fail("unsupported BF16 launch shape");
    }
    return 0;
}

void drop_graphs(sg_ens *e) {
    for (auto &kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : e->graph_defs) (void)hipGraphDestroy(kv.second);
    for (auto &kv : e->pgraphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : e->pgraph_defs) (void)hipGraphDestroy(kv.second);
    e->graphs.clear();
    e->graph_defs.clear();
    e->pgraphs.clear();
    e->pgraph_defs.clear();
}

// A persistent launch of `steps` steps from state `parity`: the grid is as
// many blocks as fit resident at once (the occupancy the runtime reports
// for this block shape and LDS size, times the compute units), at most one
// per item.  With `graph` null it is launched at once as a cooperative
// kernel (the runtime refuses one it cannot keep resident); otherwise it
// becomes the one kernel node of *graph, with the cooperative attribute.
template <int RT, int KS>
hipError_t launch_persistent_t(sg_ens *e, int steps, int parity,
                               hipStream_t s, hipGraph_t *graph) {
    auto *k = ens_bf16_persistent<RT, KS>;
    const size_t lds = lds_bytes(e);
    int per_cu = 0, dev = 0, cus = 0;
    hipError_t err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(k), 64 * RT * KS, lds);
    if (err != hipSuccess) return err;
    if ((err = hipGetDevice(&dev)) != hipSuccess) return err;
    if ((err = hipDeviceGetAttribute(
             &cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return err;
    if (per_cu <= 0 || cus <= 0) return hipErrorLaunchOutOfResources;
    const int n_items = (int)e->items.size();
    const int blocks = std::max(1, std::min(n_items, per_cu * cus));
    e->persist_blocks = blocks;
    const Layer *layers = e->d_layers;
    const Item *items = e->d_items;
    float leak = e->leak;
    int rectify = e->rectify;
    unsigned int *sync = e->d_queue;
    int n = n_items, p = parity, st = steps;
    void *args[] = {(void *)&layers, (void *)&items, (void *)&n, (void *)&p,
                    (void *)&st, (void *)&leak, (void *)&rectify,
                    (void *)&sync};
    if (!graph)
        return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k),
                                          dim3(blocks), dim3(64 * RT * KS),
                                          args, (unsigned int)lds, s);
    if ((err = hipGraphCreate(graph, 0)) != hipSuccess) return err;
    hipKernelNodeParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.func = reinterpret_cast<void *>(k);
    kp.gridDim = dim3(blocks);
    kp.blockDim = dim3(64 * RT * KS);
    kp.sharedMemBytes = (unsigned int)lds;
    kp.kernelParams = args;
    hipGraphNode_t node = nullptr;
    if ((err = hipGraphAddKernelNode(&node, *graph, nullptr, 0, &kp)) !=
        hipSuccess)
        return err;
    hipKernelNodeAttrValue v;
    std::memset(&v, 0, sizeof(v));
    v.cooperative = 1;
    return hipGraphKernelNodeSetAttribute(node, hipKernelNodeAttributeCooperative,
                                          &v);
}

hipError_t launch_persistent(sg_ens *e, int steps, int parity,
                             hipStream_t s, hipGraph_t *graph) {
#define ENS_PCASE(R, K) \
    if (e->rt == R && e->ks == K) \
        return launch_persistent_t<R, K>(e, steps, parity, s, graph);
    ENS_PCASE(1, 1) ENS_PCASE(1, 2) ENS_PCASE(1, 4) ENS_PCASE(1, 8)
    ENS_PCASE(1, 16) ENS_PCASE(2, 1) ENS_PCASE(2, 2) ENS_PCASE(2, 4)
    ENS_PCASE(2, 8) ENS_PCASE(4, 1) ENS_PCASE(4, 2) ENS_PCASE(4, 4)
    ENS_PCASE(8, 1) ENS_PCASE(8, 2) ENS_PCASE(16, 1)
#undef ENS_PCASE
    return hipErrorInvalidValue;
}

// Waits for the work of the last run, wherever it was launched.
int wait_runs(sg_ens *e) {
    if (e->last) ENS_CHECK(hipStreamSynchronize(e->last), "sync last run");
    if (e->stream) ENS_CHECK(hipStreamSynchronize(e->stream), "sync");
    return 0;
}

// Work list for the current launch shape: one item per block of rows,
// layers with the longest fan-in first so the short ones fill the tail.
// Waits for pending runs first: their kernels read the list it rewrites.
int build_items(sg_ens *e) {
    if (wait_runs(e)) return 1;
    std::vector<int> order(e->layers.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return e->layers[a].Kp > e->layers[b].Kp;
    });
    const int per_block = e->rt;
    e->items.clear();
    for (int li : order)
        for (int t = 0; t < e->layers[li].tiles; t += per_block)
            e->items.push_back(Item{li, t});
    if (e->items.size() > e->d_items_cap) {
        if (e->d_items) (void)hipFree(e->d_items);
        e->d_items = nullptr;
        ENS_CHECK(hipMalloc(&e->d_items, e->items.size() * sizeof(Item)),
                  "hipMalloc items");
        e->d_items_cap = e->items.size();
    }
    ENS_CHECK(hipMemcpy(e->d_items, e->items.data(),
                        e->items.size() * sizeof(Item), hipMemcpyHostToDevice),
              "hipMemcpy items");
    drop_graphs(e);
    return 0;
}

int layer_tiles(int precision, int R) {
    return precision == SG_ENS_FP32 ? (R + kGroupRows - 1) / kGroupRows
                                    : (R + kTileRows - 1) / kTileRows;
}

size_t layer_weight_bytes(int precision, int R, int Kp) {
    if (precision == SG_ENS_FP32)
        return (size_t)ceil_to(R, kGroupRows) * Kp * 4;
    return (size_t)ceil_to(R, kTileRows) * Kp * 2;
}

int check_inst(const sg_ens *e, int inst, int layer, int lo) {
    if (!e) return fail("null ensemble");
    if (inst < 0 || inst >= e->n) return fail("instance out of range");
    if (layer < lo || layer > 2) return fail("layer out of range");
    return 0;
}

}  // namespace

extern "C" {

const char *sg_ens_error(void) { return g_error.c_str(); }

sg_ens *sg_ens_create(int n, const int *widths, int precision, float leak,
                      int rectify, uint64_t seed, uint32_t workload) {
    if (n <= 0 || !widths) { fail("no instances"); return nullptr; }
    if (precision != SG_ENS_BF16 && precision != SG_ENS_FP32) {
        fail("unknown precision");
        return nullptr;
    }
    for (int i = 0; i < 3 * n; ++i)
        if (widths[i] < 1 || widths[i] > 8192) {
            fail("width out of range [1, 8192]");
            return nullptr;
        }
    sg_ens *e = new sg_ens();
    e->n = n;
    e->precision = precision;
    e->leak = leak;
    e->rectify = rectify ? 1 : 0;
    e->seed = seed;
    e->workload = workload;
    e->widths.assign(widths, widths + 3 * n);
    e->rt = precision == SG_ENS_FP32 ? 4 : 4;
    e->ks = 1;
    auto bail = [&](int) { sg_ens_destroy(e); return (sg_ens *)nullptr; };
    // state: x0, h1[2], h2[2] per instance, each padded to 32 floats
    long off = 0;
    e->x0_off.resize(n);
    for (int p = 0; p < 2; ++p) {
        e->h1_off[p].resize(n);
        e->h2_off[p].resize(n);
    }
    for (int i = 0; i < n; ++i) {
        const int *w = &e->widths[3 * i];
        e->x0_off[i] = off; off += ceil_to(w[0], 32);
        for (int p = 0; p < 2; ++p) { e->h1_off[p][i] = off; off += ceil_to(w[1], 32); }
        for (int p = 0; p < 2; ++p) { e->h2_off[p][i] = off; off += ceil_to(w[2], 32); }
    }
    e->state_floats = (size_t)off;
    size_t wb = 0;
    e->w_off.resize(2 * n);
    for (int i = 0; i < n; ++i)
        for (int l = 1; l <= 2; ++l) {
            const int R = e->widths[3 * i + l], K = e->widths[3 * i + l - 1];
            e->w_off[2 * i + l - 1] = (long)wb;
            wb += ceil_to(layer_weight_bytes(precision, R, ceil_to(K, 32)), 256);
            e->kp_max = std::max(e->kp_max, (int)ceil_to(K, 32));
        }
    e->w_bytes = wb;
    if (hipStreamCreate(&e->stream) != hipSuccess) { fail("hipStreamCreate"); return bail(0); }
    if (hipMalloc(&e->d_w, wb) != hipSuccess ||
        hipMalloc(&e->d_state, e->state_floats * 4) != hipSuccess ||
        hipMalloc(&e->d_layers, 2 * n * sizeof(Layer)) != hipSuccess) {
        fail("hipMalloc: ensemble does not fit");
        return bail(0);
    }
    if (hipMemset(e->d_state, 0, e->state_floats * 4) != hipSuccess) {
        fail("hipMemset");
        return bail(0);
    }
    const uint32_t k0 = (uint32_t)seed, k1 = workload;
    for (int i = 0; i < n; ++i) {
        const int *w = &e->widths[3 * i];
        hipLaunchKernelGGL(ens_init_uniform, dim3((w[0] + 255) / 256), dim3(256),
                           0, e->stream, e->d_state + e->x0_off[i], w[0], k0, k1,
                           (uint32_t)(4 * i));
        for (int l = 1; l <= 2; ++l) {
            const int R = w[l], K = w[l - 1], Kp = (int)ceil_to(K, 32);
            unsigned char *dst = e->d_w + e->w_off[2 * i + l - 1];
            const double sd = 1.0 / std::sqrt((double)K);
            const uint32_t st = (uint32_t)(4 * i + l);
            if (precision == SG_ENS_FP32) {
                const long total = ceil_to(R, kGroupRows) * (long)Kp;
                hipLaunchKernelGGL(ens_init_fp32, dim3(1024), dim3(256), 0,
                                   e->stream, (float *)dst, R, K, Kp, total,
                                   k0, k1, st, sd);
            } else {
                const long total = ceil_to(R, kTileRows) * (long)Kp;
                hipLaunchKernelGGL(ens_init_bf16, dim3(1024), dim3(256), 0,
                                   e->stream, (uint16_t *)dst, R, K, Kp / 32,
                                   total, k0, k1, st, sd);
            }
            Layer L;
            L.w = dst;
            L.R = R;
            L.K = K;
            L.Kp = Kp;
            L.tiles = layer_tiles(precision, R);
            for (int p = 0; p < 2; ++p) {
                L.x[p] = l == 1 ? e->d_state + e->x0_off[i]
                                : e->d_state + e->h1_off[p][i];
                L.h[p] = e->d_state + (l == 1 ? e->h1_off[p][i] : e->h2_off[p][i]);
            }
            e->layers.push_back(L);
        }
    }
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(e->d_layers, e->layers.data(), e->layers.size() * sizeof(Layer),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess) {
        fail("ensemble initialisation failed");
        return bail(0);
    }
    if (build_items(e)) return bail(0);
    return e;
}

void sg_ens_destroy(sg_ens *e) {
    if (!e) return;
    if (e->last) (void)hipStreamSynchronize(e->last);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    if (e->d_w) (void)hipFree(e->d_w);
    if (e->d_state) (void)hipFree(e->d_state);
    if (e->d_backup) (void)hipFree(e->d_backup);
    if (e->d_layers) (void)hipFree(e->d_layers);
    if (e->d_items) (void)hipFree(e->d_items);
    if (e->d_queue) (void)hipFree(e->d_queue);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int sg_ens_set_tile(sg_ens *e, int row_tiles, int k_split) {
    if (!e) return fail("null ensemble");
    if (e->precision == SG_ENS_FP32) {
        if (k_split != 1 || !bf16_shape_ok(row_tiles, 1))
            return fail("FP32 launch shape: k_split 1, row_tiles 1..16 (power of 2)");
    } else if (!bf16_shape_ok(row_tiles, k_split)) {
        return fail("BF16 launch shape: row_tiles * k_split in {1,2,4,8,16}");
    }
    e->rt = row_tiles;
    e->ks = k_split;
    return build_items(e);
}

int sg_ens_get_tile(const sg_ens *e, int *row_tiles, int *k_split) {
    if (!e) return fail("null ensemble");
    if (row_tiles) *row_tiles = e->rt;
    if (k_split) *k_split = e->ks;
    return 0;
}

int sg_ens_run(sg_ens *e, int steps, int use_graph, void *stream) {
    if (!e) return fail("null ensemble");
    if (steps < 0) return fail("negative step count");
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    if (steps == 0) return 0;
    e->last = s;
    if (!use_graph) {
        for (int t = 0; t < steps; ++t) {
            if (launch_step(e, e->parity, s)) return 1;
            e->parity ^= 1;
        }
        ENS_CHECK(hipGetLastError(), "step launch");
        return 0;
    }
    const std::pair<int, int> key(steps, e->parity);
    auto it = e->graphs.find(key);
    if (it == e->graphs.end()) {
        hipStream_t cap;
        ENS_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking),
                  "hipStreamCreate");
        hipGraph_t g = nullptr;
        hipError_t err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int p = e->parity, rc = 0;
        if (err == hipSuccess) {
            for (int t = 0; t < steps && rc == 0; ++t) {
                rc = launch_step(e, p, cap);
                p ^= 1;
            }
            err = hipStreamEndCapture(cap, &g);
        }
        (void)hipStreamDestroy(cap);
        if (rc) return rc;
        if (err != hipSuccess) return hip_fail(err, "graph capture");
        hipGraphExec_t ex = nullptr;
        err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (err != hipSuccess) {
            (void)hipGraphDestroy(g);
            return hip_fail(err, "hipGraphInstantiate");
        }
        e->graph_defs[key] = g;
        it = e->graphs.emplace(key, ex).first;
    }
    ENS_CHECK(hipGraphLaunch(it->second, s), "hipGraphLaunch");
    if (steps & 1) e->parity ^= 1;
    return 0;
}

int sg_ens_run_persistent(sg_ens *e, int steps, int use_graph,
                          void *stream) {
    if (!e) return fail("null ensemble");
    if (steps < 0) return fail("negative step count");
    if (e->precision != SG_ENS_BF16)
        return fail("persistent runs: BF16 ensembles only");
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    if (steps == 0 || e->items.empty()) return 0;
    if (!e->d_queue) {
        ENS_CHECK(hipMalloc(&e->d_queue, 8 * sizeof(unsigned int)),
                  "hipMalloc queue");
        ENS_CHECK(hipMemset(e->d_queue, 0, 8 * sizeof(unsigned int)),
                  "hipMemset queue");
    }
    e->last = s;
    if (!use_graph) {
        ENS_CHECK(launch_persistent(e, steps, e->parity, s, nullptr),
                  "persistent launch");
        e->persist_captured = 0;
        if (steps & 1) e->parity ^= 1;
        return 0;
    }
    const std::pair<int, int> key(steps, e->parity);
    auto it = e->pgraphs.find(key);
    if (it == e->pgraphs.end()) {
        hipGraph_t g = nullptr;
        hipError_t err = launch_persistent(e, steps, e->parity, s, &g);
        hipGraphExec_t ex = nullptr;
        if (err == hipSuccess)
            err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (err != hipSuccess) {
            // no cooperative graph node here: launch it directly (still
            // one launch for every step)
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            ENS_CHECK(launch_persistent(e, steps, e->parity, s, nullptr),
                      "persistent launch");
            e->persist_captured = 0;
            if (steps & 1) e->parity ^= 1;
            return 0;
        }
        e->pgraph_defs[key] = g;
        it = e->pgraphs.emplace(key, ex).first;
    }
    ENS_CHECK(hipGraphLaunch(it->second, s), "hipGraphLaunch");
    e->persist_captured = 1;
    if (steps & 1) e->parity ^= 1;
    return 0;
}

int sg_ens_persistent_info(sg_ens *e, int *blocks, int *captured,
                           int *barrier_timeouts) {
    if (!e) return fail("null ensemble");
    if (blocks) *blocks = e->persist_blocks;
    if (captured) *captured = e->persist_captured;
    if (barrier_timeouts) {
        unsigned int words[8] = {0};
        if (e->d_queue) {
            if (wait_runs(e)) return 1;
            ENS_CHECK(hipMemcpy(words, e->d_queue, sizeof(words),
                                hipMemcpyDeviceToHost), "read barrier");
        }
        *barrier_timeouts = (int)words[4];
    }
    return 0;
}

int sg_ens_synchronize(sg_ens *e) {
    if (!e) return fail("null ensemble");
    return wait_runs(e);
}

int sg_ens_time_steps(sg_ens *e, int steps, float *avg_ms) {
    if (!e || steps <= 0) return fail("null ensemble or no steps");
    hipEvent_t a, b;
    ENS_CHECK(hipEventCreate(&a), "hipEventCreate");
    ENS_CHECK(hipEventCreate(&b), "hipEventCreate");
    int rc = 0;
    if (hipEventRecord(a, e->stream) != hipSuccess) rc = fail("hipEventRecord");
    for (int t = 0; t < steps && rc == 0; ++t) {
        rc = launch_step(e, e->parity, e->stream);
        e->parity ^= 1;
    }
    float ms = 0.0f;
    if (rc == 0 && (hipEventRecord(b, e->stream) != hipSuccess ||
                    hipEventSynchronize(b) != hipSuccess ||
                    hipEventElapsedTime(&ms, a, b) != hipSuccess))
        rc = fail("event timing failed");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc == 0 && avg_ms) *avg_ms = ms / steps;
    return rc;
}

int sg_ens_autotune(sg_ens *e, int steps, int *row_tiles, int *k_split,
                    float *best_ms) {
    if (!e || steps <= 0) return fail("null ensemble or no steps");
    static const int bf16_shapes[][2] = {{1, 4}, {2, 2}, {4, 1}, {1, 8},
                                         {2, 4}, {4, 2}, {8, 1}, {2, 8},
                                         {4, 4}, {8, 2}, {16, 1}};
    static const int fp32_shapes[][2] = {{2, 1}, {4, 1}, {8, 1}, {16, 1}};
    const bool fp32 = e->precision == SG_ENS_FP32;
    const int count = fp32 ? 4 : 11;
    if (!e->d_backup)
        ENS_CHECK(hipMalloc(&e->d_backup, e->state_floats * 4), "hipMalloc backup");
    ENS_CHECK(hipStreamSynchronize(e->stream), "sync");
    ENS_CHECK(hipMemcpy(e->d_backup, e->d_state, e->state_floats * 4,
                        hipMemcpyDeviceToDevice), "backup state");
    const int saved_parity = e->parity;
    float best = 1e30f;
    int brt = e->rt, bks = e->ks;
    for (int c = 0; c < count; ++c) {
        const int *sh = fp32 ? fp32_shapes[c] : bf16_shapes[c];
        if (sg_ens_set_tile(e, sh[0], sh[1])) return 1;
        float ms = 0.0f;
        if (sg_ens_time_steps(e, 1, &ms)) return 1;   // first touch
        if (sg_ens_time_steps(e, steps, &ms)) return 1;
        if (ms < best) { best = ms; brt = sh[0]; bks = sh[1]; }
    }
    ENS_CHECK(hipMemcpy(e->d_state, e->d_backup, e->state_floats * 4,
                        hipMemcpyDeviceToDevice), "restore state");
    e->parity = saved_parity;
    if (sg_ens_set_tile(e, brt, bks)) return 1;
    if (row_tiles) *row_tiles = brt;
    if (k_split) *k_split = bks;
    if (best_ms) *best_ms = best;
    return 0;
}

int sg_ens_read_state(sg_ens *e, int inst, int layer, float *host) {
    if (check_inst(e, inst, layer, 0)) return 1;
    ENS_CHECK(hipDeviceSynchronize(), "sync");
    const long off = layer == 0 ? e->x0_off[inst]
                   : layer == 1 ? e->h1_off[e->parity][inst]
                                : e->h2_off[e->parity][inst];
    ENS_CHECK(hipMemcpy(host, e->d_state + off, e->widths[3 * inst + layer] * 4,
                        hipMemcpyDeviceToHost), "read state");
    return 0;
}

int sg_ens_write_state(sg_ens *e, int inst, int layer, const float *host) {
    if (check_inst(e, inst, layer, 0)) return 1;
    ENS_CHECK(hipDeviceSynchronize(), "sync");
    const long off = layer == 0 ? e->x0_off[inst]
                   : layer == 1 ? e->h1_off[e->parity][inst]
                                : e->h2_off[e->parity][inst];
    ENS_CHECK(hipMemcpy(e->d_state + off, host, e->widths[3 * inst + layer] * 4,
                        hipMemcpyHostToDevice), "write state");
    return 0;
}

// Host image of one layer's weights <-> row-major [R][K] values.
int sg_ens_read_weights(sg_ens *e, int inst, int layer, float *host) {
    if (check_inst(e, inst, layer, 1)) return 1;
    ENS_CHECK(hipDeviceSynchronize(), "sync");
    const int R = e->widths[3 * inst + layer], K = e->widths[3 * inst + layer - 1];
    const int Kp = (int)ceil_to(K, 32);
    const size_t bytes = layer_weight_bytes(e->precision, R, Kp);
    std::vector<unsigned char> img(bytes);
    ENS_CHECK(hipMemcpy(img.data(), e->d_w + e->w_off[2 * inst + layer - 1], bytes,
                        hipMemcpyDeviceToHost), "read weights");
    if (e->precision == SG_ENS_FP32) {
        const float *f = (const float *)img.data();
        for (int r = 0; r < R; ++r)
            memcpy(host + (size_t)r * K, f + (size_t)r * Kp, (size_t)K * 4);
        return 0;
    }
    const uint16_t *b = (const uint16_t *)img.data();
    const int Kt = Kp / 32;
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < K; ++c) {
            const long tile = (long)(r / 16) * Kt + c / 32;
            const int lane = (r % 16) + 16 * ((c % 32) / 8);
            host[(size_t)r * K + c] = bf16_value(b[(tile * 64 + lane) * 8 + c % 8]);
        }
    return 0;
}

int sg_ens_write_weights(sg_ens *e, int inst, int layer, const float *host) {
    if (check_inst(e, inst, layer, 1)) return 1;
    ENS_CHECK(hipDeviceSynchronize(), "sync");
    const int R = e->widths[3 * inst + layer], K = e->widths[3 * inst + layer - 1];
    const int Kp = (int)ceil_to(K, 32);
    const size_t bytes = layer_weight_bytes(e->precision, R, Kp);
    std::vector<unsigned char> img(bytes, 0);
    if (e->precision == SG_ENS_FP32) {
        float *f = (float *)img.data();
        for (int r = 0; r < R; ++r)
            memcpy(f + (size_t)r * Kp, host + (size_t)r * K, (size_t)K * 4);
    } else {
        uint16_t *b = (uint16_t *)img.data();
        const int Kt = Kp / 32;
        for (int r = 0; r < R; ++r)
            for (int c = 0; c < K; ++c) {
                const long tile = (long)(r / 16) * Kt + c / 32;
                const int lane = (r % 16) + 16 * ((c % 32) / 8);
                b[(tile * 64 + lane) * 8 + c % 8] =
                    (uint16_t)bf16_rne(host[(size_t)r * K + c]);
            }
    }
    ENS_CHECK(hipMemcpy(e->d_w + e->w_off[2 * inst + layer - 1], img.data(), bytes,
                        hipMemcpyHostToDevice), "write weights");
    return 0;
}

int sg_ens_counts(const sg_ens *e, long long *synapses, long long *neurons,
                  long long *algorithmic_bytes, long long *stored_bytes) {
    if (!e) return fail("null ensemble");
    long long syn = 0, neu = 0, widths = 0;
    for (int i = 0; i < e->n; ++i) {
        const int *w = &e->widths[3 * i];
        syn += (long long)w[0] * w[1] + (long long)w[1] * w[2];
        neu += w[1] + w[2];
        widths += (long long)w[0] + w[1] + w[2];
    }
    const int wb = e->precision == SG_ENS_FP32 ? 4 : 2;
    if (synapses) *synapses = syn;
    if (neurons) *neurons = neu;
    if (algorithmic_bytes) *algorithmic_bytes = wb * syn + 16 * widths;
    if (stored_bytes) *stored_bytes = (long long)e->w_bytes;
    return 0;
}

}  // extern "C"
