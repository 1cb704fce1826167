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
// 2. plane_base_kernel: a 256-lane block owns a tile of 64 destinations.
//    It stages the plane in LDS; its four wavefronts split the tile's visit
//    list into mask words and record, per destination, which visits spike
//    (order-free work); then wavefront 0 walks each destination's spiking
//    visits in k order, loading their weights eight at a time, and adds
//    them in order.
//
// Work per synapse is a bit test in LDS; weights are read only for the
// spiking visits.  Planes larger than kPlaneMax words stay in global memory
// (read through the caches); destinations with more than kMaxVisits visits
// use the tile kernels.
#include <hip/hip_runtime.h>

#include "sg_device.h"

namespace sg {

constexpr int kPlaneThreads = 256;
constexpr int kPlaneMax = 4096;           // uint64 words staged in LDS
constexpr int kMaxVisits = 1024;
constexpr int kMaskWords = kMaxVisits / 32;
constexpr int kBatch = 8;

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

template <int TYPE, bool STAGED>
__global__ __launch_bounds__(kPlaneThreads) void plane_base_kernel(
        sg_conn c, const uint64_t *plane) {
    __shared__ uint64_t splane[STAGED ? kPlaneMax : 1];
    __shared__ uint32_t masks[kMaskWords][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = blockIdx.x * 64 + lane;
    const bool live = slot < c.num_dest;
    // count <= k_per_dest <= kMaxVisits (checked by plane_path); the clamp
    // only keeps LDS indexing in range by construction.
    const int count = min(plane_visits<TYPE>(c), kMaxVisits);
    const int nwords = (count + 31) >> 5;
    if (STAGED) {
        const int np = (c.from_size + 63) >> 6;
        for (int i = threadIdx.x; i < np; i += kPlaneThreads)
            splane[i] = plane[i];
    }
    __syncthreads();

    // Pass 1: spiking visits of mask words [j0, j1) of every destination.
    const int per = (nwords + 3) >> 2;
    const int j0 = wave * per, j1 = min(nwords, j0 + per);
    if (live && j0 < j1) {
        SynapseIter<TYPE> it;
        it.seek(c, slot, j0 * 32);
        for (int j = j0; j < j1; ++j) {
            uint32_t m = 0;
            const int n = min(32, count - j * 32);
            for (int b = 0; b < n; ++b) {
                const int from = it.source();
                it.advance();
                if (from >= 0) {
                    const uint64_t p = STAGED ? splane[from >> 6]
                                              : plane[from >> 6];
                    m |= (uint32_t)((p >> (from & 63)) & 1u) << b;
                }
            }
            masks[j][lane] = m;
        }
    }
    __syncthreads();
    if (wave != 0) return;

    // Pass 2: in-order sum over the spiking visits.
    const float *wbase = c.convolutional
        ? c.weights : c.weights + tiled_index(live ? slot : 0, 0, c.k_per_dest);
    const int wstep = c.convolutional ? 1 : 64;
    float sum = 0.0f;
    int j = 0;
    uint32_t m = (live && nwords > 0) ? masks[0][lane] : 0u;
    bool more = live && nwords > 0;
    while (__any(more)) {
        int ks[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            while (m == 0u && j + 1 < nwords && more) m = masks[++j][lane];
            if (m != 0u) {
                ks[b] = (j << 5) + __builtin_ctz(m);
                m &= m - 1u;
            } else {
                ks[b] = -1;
            }
        }
        float w[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b)
            w[b] = ks[b] >= 0 ? wbase[(long)ks[b] * wstep] : 0.0f;
#pragma unroll
        for (int b = 0; b < kBatch; ++b)
            if (ks[b] >= 0) sum += w[b];
        more = more && ks[kBatch - 1] >= 0;
    }
    if (live) {
        const int to = slot_to_index(c, slot);
        c.dst[to] = aggregate(c.opcode, c.dst[to], sum);
    }
}

template <int TYPE>
static hipError_t launch_plane_t(const sg_conn &c, const uint64_t *plane,
                                 hipStream_t st) {
    dim3 g((c.num_dest + 63) / 64), b(kPlaneThreads);
    if ((c.from_size + 63) / 64 <= kPlaneMax)
        hipLaunchKernelGGL((plane_base_kernel<TYPE, true>), g, b, 0, st, c,
                           plane);
    else
        hipLaunchKernelGGL((plane_base_kernel<TYPE, false>), g, b, 0, st, c,
                           plane);
    return hipGetLastError();
}

// True when the plane path applies: BIT source, finite weights, a tiled
// connection type and at most kMaxVisits visits per destination.
bool plane_path(const sg_conn &c) {
    return c.plane != nullptr && c.skip_silent && c.out_type == SG_OUT_BIT &&
           c.type != SG_ONE_TO_ONE && c.k_per_dest <= kMaxVisits;
}

hipError_t launch_plane_base(const sg_conn &c, hipStream_t st) {
    const int n = c.from_size;
    hipLaunchKernelGGL(pack_plane_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       st, c.src, n, c.delay, c.plane);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (c.type) {
        case SG_FULLY_CONNECTED:
            return launch_plane_t<SG_FULLY_CONNECTED>(c, c.plane, st);
        case SG_SUBSET: return launch_plane_t<SG_SUBSET>(c, c.plane, st);
        case SG_CONVERGENT:
            return launch_plane_t<SG_CONVERGENT>(c, c.plane, st);
        case SG_DIVERGENT: return launch_plane_t<SG_DIVERGENT>(c, c.plane, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace sg
