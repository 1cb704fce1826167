// Base activator for spiking (BIT) sources with finite weights, driven by a
// spike plane.
//
// For a BIT source the reference's per-synapse value is
// extract(word, delay) * w = (0 or 1) * w (synapse_kernel.h:887-912,
// extractor.cpp:8-14).  With finite weights a silent synapse adds +0 or -0,
// which leaves the destination's running sum unchanged (it starts at +0 and
// a round-to-nearest sum of nonzero terms is never -0), so the sum over the
// spiking visits alone, in visit order, is the reference's sum bit for bit.
//
// 1. pack_plane_kernel turns the source's history words into a plane: the
//    bit each source neuron shows this connection (word delay/32, bit
//    31 - delay%32), 64 neurons per uint64, one wavefront ballot each.
// 2. plane_base_kernel: a block owns a tile of 64 destinations (one lane
//    each) and stages the plane in LDS.  Its W wavefronts split the visit
//    list into windows of VPW * W visits: each lane tests its VPW visits'
//    bits in LDS and issues the weight loads of the spiking ones together
//    (coalesced across the tile), writing the weight -- or +0 for a silent
//    or skipped visit -- to an LDS window.  After a barrier wavefront 0 adds
//    the window in visit order.  Adding +0 leaves the sum unchanged, so the
//    result is the reference's sum bit for bit.
//
// Work per synapse is a bit test in LDS; weights are read only for the
// spiking visits.  Planes larger than kPlaneMax words are read from global
// memory (through the caches) instead of LDS.
#include <hip/hip_runtime.h>

#include "sg_device.h"

namespace sg {

constexpr int kPlaneMax = 4096;     // uint64 plane words staged in LDS

__global__ __launch_bounds__(256) void pack_plane_kernel(
        const uint32_t *words, int n, int delay, uint64_t *plane) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool bit = i < n && ((words[i] << (delay & 31)) >> 31) != 0;
    const uint64_t m = __ballot(bit);
    // lane 0 of each wavefront: i is a multiple of 64
    if ((threadIdx.x & 63) == 0 && i < n) plane[i >> 6] = m;
}

template <int TYPE>
__device__ __forceinline__ int plane_visits(const sg_conn &c) {
    return cursor<TYPE>(c, 0).count;
}

template <int TYPE, int W, int VPW, int NBUF, bool STAGED>
__global__ __launch_bounds__(64 * W) void plane_base_kernel(
        sg_conn c, const uint64_t *plane) {
    constexpr int WIN = VPW * W;
    extern __shared__ uint64_t splane[];          // STAGED: the plane
    __shared__ float vals[NBUF][WIN][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = blockIdx.x * 64 + lane;
    const bool live = slot < c.num_dest;
    const int count = plane_visits<TYPE>(c);
    if (STAGED) {
        const int np = (c.from_size + 63) >> 6;
        for (int i = threadIdx.x; i < np; i += 64 * W) splane[i] = plane[i];
    }
    __syncthreads();
    const uint64_t *pl = STAGED ? splane : plane;
    const float *wbase = c.convolutional
        ? c.weights
        : c.weights + tiled_index(live ? slot : 0, 0, c.k_per_dest);
    const long wstep = c.convolutional ? 1 : 64;
    float sum = 0.0f;
    for (int k0 = 0, win = 0; k0 < count; k0 += WIN, ++win) {
        const int buf = NBUF == 2 ? (win & 1) : 0;
        const int kw = k0 + wave * VPW;
        if (live && kw < count) {
            SynapseIter<TYPE> it;
            it.seek(c, slot, kw);
            const int n = min(VPW, count - kw);
            float v[VPW];
#pragma unroll
            for (int b = 0; b < VPW; ++b) {
                v[b] = 0.0f;
                if (b < n) {
                    const int from = it.source();
                    it.advance();
                    if (from >= 0 && ((pl[from >> 6] >> (from & 63)) & 1u))
                        v[b] = wbase[(kw + b) * wstep];
                }
            }
#pragma unroll
            for (int b = 0; b < VPW; ++b)
                vals[buf][wave * VPW + b][lane] = v[b];
        }
        __syncthreads();
        if (live && wave == 0) {
            const int n = min(WIN, count - k0);
            for (int j = 0; j < n; ++j) sum += vals[buf][j][lane];
        }
        // One buffer: the next window's writes wait for the sum.
        if (NBUF == 1) __syncthreads();
    }
    if (live && wave == 0) {
        const int to = slot_to_index(c, slot);
        put_sum(c, to, sum);
    }
}

// Fully connected, at most kListMax sources: every destination visits the
// sources in index order, so the spiking ones form one list shared by the
// whole layer.  A block (one wavefront, one tile of 64 destinations)
// builds it in LDS from the plane -- a popcount per plane word, a
// wavefront prefix sum over the words, each word's set bits written at its
// offset -- and then each lane adds its weights of the listed sources in
// list order (the same 256-byte row of the tiled weights for all lanes per
// entry).  Silent visits add +0 in the reference's loop, so the sum is its
// sum bit for bit.
constexpr int kListMax = 4096;

__global__ __launch_bounds__(64) void fc_list_kernel(sg_conn c,
                                                     const uint64_t *plane) {
    __shared__ int list[kListMax];
    __shared__ int total;
    const int lane = threadIdx.x;
    const int np = (c.from_size + 63) >> 6;          // <= 64 words
    const uint64_t word = lane < np ? plane[lane] : 0ull;
    const int n = __popcll(word);
    int pre = n;                                      // inclusive scan
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(pre, d, 64);
        if (lane >= d) pre += v;
    }
    int at = pre - n;
    for (uint64_t m = word; m; m &= m - 1)
        list[at++] = lane * 64 + (int)__builtin_ctzll(m);
    if (lane == 63) total = pre;
    __syncthreads();
    const int m = total;
    const int slot = blockIdx.x * 64 + lane;
    if (slot >= c.num_dest) return;
    const float *wrow = c.weights + tiled_index(slot, 0, c.k_per_dest);
    float sum = 0.0f;
    int j = 0;
    for (; j + 8 <= m; j += 8) {
        float v[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) v[b] = wrow[(long)list[j + b] * 64];
#pragma unroll
        for (int b = 0; b < 8; ++b) sum += v[b];
    }
    for (; j < m; ++j) sum += wrow[(long)list[j] * 64];
    put_sum(c, slot_to_index(c, slot), sum);
}

template <int TYPE, int W, int VPW, int NBUF>
static hipError_t launch_plane_w(const sg_conn &c, hipStream_t st) {
    dim3 g((c.num_dest + 63) / 64), b(64 * W);
    const int np = (c.from_size + 63) / 64;
    if (np <= kPlaneMax)
        hipLaunchKernelGGL((plane_base_kernel<TYPE, W, VPW, NBUF, true>), g,
                           b, (size_t)np * 8, st, c, c.plane);
    else
        hipLaunchKernelGGL((plane_base_kernel<TYPE, W, VPW, NBUF, false>), g,
                           b, 0, st, c, c.plane);
    return hipGetLastError();
}

// At most 16 visits per destination: one wavefront per tile, each lane
// walks its destination's visits in order, reading the plane words through
// the caches (no LDS staging, no barrier: a block is one short wavefront
// that fits the single wave slot a CU has left beside a long kernel).
// Every visit's plane word is loaded first, then the weights of the
// spiking visits, then the sum in visit order (silent visits add +0).
template <int TYPE>
__global__ __launch_bounds__(64) void plane_small_kernel(sg_conn c,
                                                         const uint64_t *plane) {
    constexpr int V = 16;
    const int slot = blockIdx.x * 64 + threadIdx.x;
    if (slot >= c.num_dest) return;
    const int count = plane_visits<TYPE>(c);
    const float *wbase = c.convolutional
        ? c.weights : c.weights + tiled_index(slot, 0, c.k_per_dest);
    const long wstep = c.convolutional ? 1 : 64;
    SynapseIter<TYPE> it;
    it.seek(c, slot, 0);
    int from[V];
    uint64_t word[V];
#pragma unroll
    for (int b = 0; b < V; ++b) {
        from[b] = -1;
        if (b < count) {
            from[b] = it.source();
            it.advance();
        }
        word[b] = from[b] >= 0 ? plane[from[b] >> 6] : 0ull;
    }
    float v[V];
#pragma unroll
    for (int b = 0; b < V; ++b) {
        v[b] = 0.0f;
        if (from[b] >= 0 && ((word[b] >> (from[b] & 63)) & 1u))
            v[b] = wbase[b * wstep];
    }
    float sum = 0.0f;
#pragma unroll
    for (int b = 0; b < V; ++b) sum += v[b];
    put_sum(c, slot_to_index(c, slot), sum);
}

// A destination with at most 16 visits: plane_small_kernel.  Beside a
// long kernel (sg_conn.beside): four wavefronts of 8 visits per tile, two
// LDS buffers.  Otherwise sixteen wavefronts of 8 visits share a tile
// (windows of
// 128 visits, one LDS buffer): the most wavefronts in flight per tile,
// which is what hides latency for the few-tile layers this path serves
// (measured on the E/I sheet's E->I connection: 12.9 us against 19.9 us
// with four wavefronts of 16 visits).
template <int TYPE>
static hipError_t launch_plane_t(const sg_conn &c, hipStream_t st) {
    if (cursor<TYPE>(c, 0).count <= 16) {
        hipLaunchKernelGGL(plane_small_kernel<TYPE>,
                           dim3((c.num_dest + 63) / 64), dim3(64), 0, st, c,
                           c.plane);
        return hipGetLastError();
    }
    if (c.beside) return launch_plane_w<TYPE, 4, 8, 2>(c, st);
    return launch_plane_w<TYPE, 16, 8, 1>(c, st);
}

// True when the plane path applies: a BIT source, finite weights and a
// tiled connection type.
bool plane_path(const sg_conn &c) {
    return c.plane != nullptr && c.skip_silent && c.out_type == SG_OUT_BIT &&
           c.type != SG_ONE_TO_ONE;
}

hipError_t launch_plane_base(const sg_conn &c, hipStream_t st) {
    if (!c.plane_ready) {
        const int n = c.from_size;
        hipLaunchKernelGGL(pack_plane_kernel, dim3((n + 255) / 256),
                           dim3(256), 0, st, c.src, n, c.delay, c.plane);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    switch (c.type) {
        case SG_FULLY_CONNECTED:
            if (c.from_size <= kListMax && !c.convolutional) {
                hipLaunchKernelGGL(fc_list_kernel, dim3((c.num_dest + 63) / 64),
                                   dim3(64), 0, st, c, c.plane);
                return hipGetLastError();
            }
            return launch_plane_t<SG_FULLY_CONNECTED>(c, st);
        case SG_SUBSET: return launch_plane_t<SG_SUBSET>(c, st);
        case SG_CONVERGENT: return launch_plane_t<SG_CONVERGENT>(c, st);
        case SG_DIVERGENT: return launch_plane_t<SG_DIVERGENT>(c, st);
        case SG_SPARSE: return launch_plane_t<SG_SPARSE>(c, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace sg

using namespace sg;

extern "C" int sg_hip_pack_plane(const uint32_t *words, int n, int delay,
                                 uint64_t *plane, void *stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(pack_plane_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, words, n, delay, plane);
    return hipGetLastError();
}
