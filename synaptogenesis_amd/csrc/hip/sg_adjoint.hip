// Forward pass and adjoint of the two-layer leaky network of
// include/sg_adjoint.h.
//
// Layer 1's recurrence does not read layer 2, so the input currents of all
// steps are one GEMM and each block of `checkpoint` steps (M = checkpoint *
// batch rows, row = step * batch + b) runs as parallel passes with only
// elementwise recurrences in time:
//   forward   gemm_fwd_dma A1[T*batch][n_hid] = X . W1^T for every step
//                          (MFMA BF16; gemm_fwd where n_in is not a
//                          multiple of 8), kept for the backward
//             T <= 1024: scan1_fwd_all (W2 put in fragment order by
//             extra blocks of its launch), out_current_mfma (S1 . W2^T as
//             one MFMA product) and scan2_fwd_all, the passes below over
//             all steps; else per block:
//             scan1_fwd    one thread per (b, hidden unit): v1, s1 over
//                          the block's steps
//             out_current  A2[M][n_out] = S1 . W2^T, one wavefront per row
//             scan2_fwd    one thread per (b, output): v2, s2, counts
//                          (checkpoints of at most 16 steps: both in one
//                          launch, out_scan2_fwd, a block per batch row);
//                          v2 of every step is kept for the backward
//   backward  n_hid a multiple of 256 and checkpoints of at most 16
//             steps: scan_bwd_fused, the loss gradient and all of the
//             passes below for every block of steps in one launch (dW2
//             summed per batch row), with pack_x as extra blocks of it;
//             else per block, last first: layer 1's recurrence again from
//             the block's checkpoint (the currents are kept), recording
//             v1 and s1, then
//             scan2_bwd    one thread per (b, output), steps in reverse:
//                          d2 = dL/dv2
//             scan1_bwd    one thread per (b, hidden unit), steps in
//                          reverse: d1 = dL/dv1 from W2^T d2, kept for
//                          every step
//             dw2_partial  D2^T . S1 over 16-row chunks, accumulated over
//                          the blocks
//             and after the last block
//             pack_x       x regrouped by eight rows for the product
//             gemm_dw1_dma dW1 = D1^T . X over every step's rows (MFMA
//                          BF16, D1 as hi + mid BF16 terms; gemm_dw1,
//                          with a third term lo, where n_hid is not a
//                          multiple of 4), split into row ranges
// and add_parts / dw2_reduce add the ranges and chunks in a fixed order into dW1
// and dW2 (dw2_reduce's work as one more z slice of gemm_dw1_dma's launch).
// Every sum runs in a fixed order: results are deterministic.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <string>

#include "sg_adjoint.h"

namespace sg {
namespace snn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxOut = 16;
constexpr int kTnSplit = 8;     // row ranges of the dW1 product
constexpr int kAllSteps = 1024; // longest forward run as two passes (LDS)

__device__ __forceinline__ uint32_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ float bf16_value(uint32_t b) {
    return __uint_as_float(b << 16);
}

// S(u) and the derivative the adjoint uses (see sg_adjoint.h).
__device__ __forceinline__ float act(float u, int mode, float slope) {
    if (mode == SG_SNN_SPIKE) return u >= 0.0f ? 1.0f : 0.0f;
    return 0.5f * (1.0f + slope * u / (1.0f + slope * fabsf(u)));
}

__device__ __forceinline__ float act_grad(float u, int mode, float slope) {
    const float d = 1.0f + slope * fabsf(u);
    const float g = 1.0f / (d * d);
    return mode == SG_SNN_SPIKE ? g : 0.5f * slope * g;
}

// The same derivative with the hardware reciprocal (about 1 ulp) in place
// of the IEEE division: the backward passes' hot loop.
__device__ __forceinline__ float act_grad_fast(float u, int mode, float slope) {
    const float r = __builtin_amdgcn_rcpf(1.0f + slope * fabsf(u));
    const float g = r * r;
    return mode == SG_SNN_SPIKE ? g : 0.5f * slope * g;
}

// act and act_grad_fast for a mode fixed at compile time (MODE < 0: the
// run's mode): the same operations, without the other mode's.
template <int MODE>
__device__ __forceinline__ float act_m(float u, int mode, float slope) {
    return act(u, MODE < 0 ? mode : MODE, slope);
}
template <int MODE>
__device__ __forceinline__ float act_grad_fast_m(float u, int mode,
                                                 float slope) {
    return act_grad_fast(u, MODE < 0 ? mode : MODE, slope);
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// 8 consecutive BF16 of row `r`, columns c..c+7, zero outside [rows, cols).
__device__ __forceinline__ u32x4 load8(const uint16_t *p, int rows, int cols,
                                       int r, int c) {
    u32x4 v = {0u, 0u, 0u, 0u};
    if (r >= rows) return v;
    const uint16_t *q = p + (long)r * cols + c;
    if ((cols & 7) == 0 && c + 8 <= cols) return *(const u32x4 *)q;
    uint16_t e[8];
    for (int i = 0; i < 8; ++i) e[i] = c + i < cols ? q[i] : (uint16_t)0;
    memcpy(&v, e, 16);
    return v;
}

// C[M][N] = X[M][K] . W[N][K]^T, BF16 operands, FP32 result (the input
// currents of every step).  Block tile 128 (m) x 256 (n): four wavefronts
// of 64 x 128, each 4 x 8 MFMA 16x16x32 tiles; K in slices of 32 through
// double-buffered LDS (row pitch 40 BF16 = 80 B: the 16 rows a fragment
// read touches start on 16 different 4-bank groups), the next slice's
// global loads in flight during the current slice's MFMAs.  A wavefront
// reads 12 fragments per 32 MFMAs.
constexpr int kFm = 128, kFn = 256, kFk = 32, kFp = 40;

__global__ __launch_bounds__(256) void gemm_fwd(const uint16_t *__restrict__ X,
                                                const uint16_t *__restrict__ W,
                                                float *__restrict__ C, int M,
                                                int N, int K) {
    __shared__ __align__(16) uint16_t As[2][kFm][kFp];
    __shared__ __align__(16) uint16_t Bs[2][kFn][kFp];
    const int n0 = blockIdx.x * kFn, m0 = blockIdx.y * kFm;
    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const int wm = (w & 1) * 64, wn = (w >> 1) * 128;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    // staging: 16-byte chunk q = t + 256 i is row q / 4, columns (q % 4) * 8
    u32x4 ra[2], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int q = t + 256 * i;
            ra[i] = load8(X, M, K, m0 + (q >> 2), k0 + (q & 3) * 8);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = t + 256 * i;
            rb[i] = load8(W, N, K, n0 + (q >> 2), k0 + (q & 3) * 8);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int q = t + 256 * i;
            *(u32x4 *)&As[buf][q >> 2][(q & 3) * 8] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = t + 256 * i;
            *(u32x4 *)&Bs[buf][q >> 2][(q & 3) * 8] = rb[i];
        }
    };
    f32x4 acc[4][8] = {};
    const int ns = (K + kFk - 1) / kFk;
    load(0);
    store(0);
    __syncthreads();
    for (int sl = 0; sl < ns; ++sl) {
        const int buf = sl & 1;
        if (sl + 1 < ns) load((sl + 1) * kFk);
        bf16x8 a[4], b[8];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a[i] = *(const bf16x8 *)&As[buf][wm + 16 * i + li][lk];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            b[j] = *(const bf16x8 *)&Bs[buf][wn + 16 * j + li][lk];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[i], b[j], acc[i][j], 0, 0, 0);
        if (sl + 1 < ns) store(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int col = n0 + wn + 16 * j + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + 16 * i + (lane >> 4) * 4 + r;
                if (row < M && col < N) C[(long)row * N + col] = acc[i][j][r];
            }
        }
}

// X [M][K] regrouped for the dW1 product, whose reduction runs over the
// rows: xp[g][k][0..7] = X[8 g .. 8 g + 7][k], for g < gp (rows past M are
// 0; gp = the groups rounded up to whole 32-row slices), so one 16-byte
// read is the eight reduction values an MFMA operand lane holds.  Thread =
// (group, 8 columns), whose output is the 128 contiguous bytes of xp at
// 128 i when K is a multiple of 8: the block's 32 KB then go out through
// LDS as 16-byte stores of consecutive lanes to consecutive addresses
// (9.7 us for 784 columns x 12800 rows, 4.1 TB/s, against 12.3 us with
// each lane storing its own 128 bytes).
// One 256-thread block `block` of it (pack_x, or extra blocks of the
// fused backward scan's launch).
__device__ __forceinline__ void pack_x_block(const uint16_t *__restrict__ X,
                                             uint16_t *__restrict__ xp,
                                             int M, int K, long gp,
                                             long block) {
    __shared__ u32x4 staged[256 * 8];
    const int nc = (K + 7) / 8;
    const long n = gp * nc;
    const long i0 = block * 256;
    const long i = i0 + threadIdx.x;
    const bool whole = K % 8 == 0;
    u32x4 o[8];
    if (i < n) {
        const int g = (int)(i / nc), c = (int)(i % nc);
        u32x4 r[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] = load8(X, M, K, 8 * g + j, 8 * c);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const uint32_t lo =
                    (r[2 * p][e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
                const uint32_t hi =
                    (r[2 * p + 1][e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
                o[e][p] = lo | (hi << 16);
            }
        }
        if (!whole) {
            for (int e = 0; e < 8 && 8 * c + e < K; ++e)
                *(u32x4 *)&xp[((long)g * K + 8 * c + e) * 8] = o[e];
            return;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) staged[threadIdx.x * 8 + e] = o[e];
    }
    if (!whole) return;   // uniform over the block
    __syncthreads();
    const long left = (n - i0) * 8;   // 16-byte chunks of this block
    u32x4 *dst = (u32x4 *)xp + i0 * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int q = threadIdx.x + 256 * j;
        if (q < left) dst[q] = staged[q];
    }
}

__global__ __launch_bounds__(256) void pack_x(const uint16_t *__restrict__ X,
                                              uint16_t *__restrict__ xp,
                                              int M, int K, long gp) {
    pack_x_block(X, xp, M, K, gp, blockIdx.x);
}

// part[z][H][K] = L[rows of range z][H]^T . X[rows of range z][K] (dW1):
// L FP32 enters as three BF16 terms hi + mid + lo, so the product keeps
// FP32 precision; X through xp (pack_x).  Block tile 128 (h) x 112 (k),
// four wavefronts of 32 x 112 (2 x 7 MFMA 16x16x32 tiles, three MFMAs per
// tile and slice), 32 rows per slice through double-buffered LDS.  An
// operand lane needs eight consecutive rows of one column of L: it reads
// them from the LDS tile with eight 4-byte reads, and the tile's row groups
// of eight are stored with their columns XOR-swizzled by 16 per group, so
// that the four groups a read touches fall in four different bank ranges.
// Blocks are numbered so that each XCD (hardware block id modulo 8) gets
// whole row ranges, all k tiles of an h tile in a row: a range's rows of L
// and X are fetched into that XCD's L2 once and reused there.
constexpr int kDh = 128, kDk = 112, kDr = 32;

__device__ __forceinline__ void split3(const float (&v)[8], bf16x8 &hi,
                                       bf16x8 &mid, bf16x8 &lo) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)v[j];
        const float r1 = v[j] - (float)h;
        const __bf16 m = (__bf16)r1;
        hi[j] = h;
        mid[j] = m;
        lo[j] = (__bf16)(r1 - (float)m);
    }
}

// The first two terms of split3: 16 significant bits of each value.
__device__ __forceinline__ void split2(const float (&v)[8], bf16x8 &hi,
                                       bf16x8 &mid) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)v[j];
        hi[j] = h;
        mid[j] = (__bf16)(v[j] - (float)h);
    }
}

__global__ __launch_bounds__(256) void gemm_dw1(const float *__restrict__ L,
                                                const uint16_t *__restrict__ xp,
                                                float *__restrict__ part,
                                                int M, int H, int K, int per) {
    __shared__ __align__(16) float Ls[2][kDr][kDh];
    __shared__ __align__(16) uint16_t Xs[2][4][kDk][8];
    // block numbering: logical = xcd * (blocks / 8) + local, k tile fastest
    const int nx = gridDim.x, ny = gridDim.y;
    const int nb = nx * ny * gridDim.z;
    const int hw = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const int lg = (nb % 8 == 0) ? (hw % 8) * (nb / 8) + hw / 8 : hw;
    const int bx = lg % nx, by = (lg / nx) % ny, bz = lg / (nx * ny);
    const int h0 = by * kDh, kb = bx * kDk;
    const int rb = bz * per, re = min(M, rb + per);
    const long G = (M + 7) / 8;
    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const int li = lane & 15, gq = lane >> 4;
    const bool vec = (H & 3) == 0;
    float4 rl[4];
    u32x4 rx[2];
    // staging: L chunk q = t + 256 i is row q / 32, columns 4 (q % 32);
    // X chunk q = t + 256 i (q < 448) is group q / 112, column q % 112
    auto load = [&](int r0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = t + 256 * i;
            const int row = r0 + (q >> 5), h = h0 + 4 * (q & 31);
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (row < re) {
                const float *src = L + (long)row * H + h;
                if (vec && h + 3 < H) {
                    v = *(const float4 *)src;
                } else {
                    if (h < H) v.x = src[0];
                    if (h + 1 < H) v.y = src[1];
                    if (h + 2 < H) v.z = src[2];
                    if (h + 3 < H) v.w = src[3];
                }
            }
            rl[i] = v;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int q = t + 256 * i;
            u32x4 v = {0u, 0u, 0u, 0u};
            if (q < 4 * kDk) {
                const long g = r0 / 8 + q / kDk;
                const int k = kb + q % kDk;
                if (g < G && k < K) v = *(const u32x4 *)&xp[(g * K + k) * 8];
            }
            rx[i] = v;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = t + 256 * i;
            const int row = q >> 5;
            const int col = (4 * (q & 31)) ^ (16 * ((row >> 3) & 3));
            *(float4 *)&Ls[buf][row][col] = rl[i];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int q = t + 256 * i;
            if (q < 4 * kDk)
                *(u32x4 *)&Xs[buf][q / kDk][q % kDk][0] = rx[i];
        }
    };
    f32x4 acc[2][7] = {};
    load(rb);
    store(0);
    __syncthreads();
    for (int r0 = rb; r0 < re; r0 += kDr) {
        const int buf = ((r0 - rb) / kDr) & 1;
        const bool more = r0 + kDr < re;
        if (more) load(r0 + kDr);
        bf16x8 b[7];
#pragma unroll
        for (int n = 0; n < 7; ++n)
            b[n] = *(const bf16x8 *)&Xs[buf][gq][16 * n + li][0];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int hl = (32 * w + 16 * i + li) ^ (16 * gq);
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = Ls[buf][8 * gq + j][hl];
            bf16x8 hi, mid, lo;
            split3(v, hi, mid, lo);
#pragma unroll
            for (int n = 0; n < 7; ++n) {
                acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    hi, b[n], acc[i][n], 0, 0, 0);
                acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    mid, b[n], acc[i][n], 0, 0, 0);
                acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    lo, b[n], acc[i][n], 0, 0, 0);
            }
        }
        if (more) store(buf ^ 1);
        __syncthreads();
    }
    float *P = part + (long)bz * H * K;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int n = 0; n < 7; ++n) {
            const int k = kb + 16 * n + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int h = h0 + 32 * w + 16 * i + 4 * gq + r;
                if (h < H && k < K) P[(long)h * K + k] = acc[i][n][r];
            }
        }
}

// ---- LDS-DMA ring forms ------------------------------------------------------
// The same two products with their operand slices brought into LDS by the
// loads themselves (global_load_lds_dwordx4: each lane's 16 bytes land at
// the wavefront's LDS base + 16 * lane), three slices in a ring, two in
// flight while the third is multiplied: the slices' memory latency is
// hidden behind two slices of MFMAs instead of one.  Each wavefront waits
// for its own loads of a slice (vmcnt), then a barrier makes every
// wavefront's part visible; a second barrier keeps a slot from being
// refilled before every wavefront has read it.  The barriers are bare
// s_barrier: the waits are explicit (a workgroup fence would wait for the
// slices in flight too).  Out-of-range rows and columns are read from a
// clamped in-range address; what they feed is either never stored or
// multiplied by the zeros the other operand holds there.

constexpr int kRing = 3;

// s_waitcnt immediate (gfx9 encoding): vmcnt <= n, no LDS/scalar or export
// wait.
#define SG_WAIT_VM(n) \
    __builtin_amdgcn_s_waitcnt(((n) & 15) | (7 << 4) | (15 << 8) | (((n) >> 4) << 14))
// every LDS read of this wavefront has returned (before a slot it read may
// be refilled)
#define SG_WAIT_LDS() __builtin_amdgcn_s_waitcnt(15 | (7 << 4) | (3 << 14))
// no memory access moves across it in the compiler (no code)
#define SG_COMPILER_FENCE() asm volatile("" ::: "memory")
#define SG_BARRIER()                   \
    do {                               \
        SG_COMPILER_FENCE();           \
        __builtin_amdgcn_s_barrier();  \
        SG_COMPILER_FENCE();           \
    } while (0)

typedef __attribute__((address_space(3))) void lds_void;

// 16 bytes per lane from `src` into LDS at the wavefront's `lds_base` + 16
// * lane.  Written as the instruction itself rather than the builtin: the
// compiler tracks only a few such loads by destination and otherwise makes
// every LDS read wait for all of them, which would drain the ring at each
// slice; here the waits are the explicit vmcnt ones above.  The
// instruction returns nothing to registers.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void dma16(const void *src, void *lds_base) {
    const uint32_t off = (uint32_t)(uintptr_t)(lds_void *)lds_base;
    const uint32_t m0v = __builtin_amdgcn_readfirstlane(off);
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
                 :: "v"(src), "s"(m0v) : "memory", "m0");
}
#pragma clang diagnostic pop

// Forward currents, K a multiple of 8 (16-byte rows).  Slices of 32
// columns: a slice's rows are stored in groups of 16, each group as its
// four 16-byte column chunks one after the other, each chunk as the 16
// rows' values in row order (byte (r / 16) * 1024 + c * 256 + (r % 16) *
// 16): the 16 lanes of a fragment read take 256 contiguous bytes, and a
// load instruction fills one group, 16 rows of 64 contiguous bytes.
// Blocks that share an m tile run back to back on one XCD.
__device__ __forceinline__ int tile_off(int r, int c) {
    return (r >> 4) * 512 + c * 128 + (r & 15) * 8;   // in BF16 elements
}

__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(2, 2))) void gemm_fwd_dma(
        const uint16_t *__restrict__ X, const uint16_t *__restrict__ W,
        float *__restrict__ C, int M, int N, int K) {
    // one variable per ring slot, chosen by a constant at every use
    __shared__ __align__(16) uint16_t A0[kFm * kFk], A1[kFm * kFk],
        A2[kFm * kFk];
    __shared__ __align__(16) uint16_t B0[kFn * kFk], B1[kFn * kFk],
        B2[kFn * kFk];
    auto As = [&](int slot) { return slot == 0 ? A0 : slot == 1 ? A1 : A2; };
    auto Bs = [&](int slot) { return slot == 0 ? B0 : slot == 1 ? B1 : B2; };
    const int nx = gridDim.x, nb = gridDim.x * gridDim.y;
    const int hw = blockIdx.x + nx * blockIdx.y;
    const int lg = (nb % 8 == 0) ? (hw % 8) * (nb / 8) + hw / 8 : hw;
    const int n0 = (lg % nx) * kFn, m0 = (lg / nx) * kFm;
    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const int wm = (w & 1) * 64, wn = (w >> 1) * 128;
    const int li = lane & 15, lc = lane >> 4;
    // this lane's DMA row within a 16-row group and its column chunk
    const int dr = lane & 15, dc = 8 * (lane >> 4);
    const uint16_t *asrc[2], *bsrc[4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
        asrc[i] = X + (long)min(m0 + 16 * (2 * w + i) + dr, M - 1) * K + dc;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        bsrc[i] = W + (long)min(n0 + 16 * (4 * w + i) + dr, N - 1) * K + dc;
    auto issue = [&](int sl, int slot) {
        const int k0 = sl * kFk;
#pragma unroll
        for (int i = 0; i < 2; ++i)
            dma16(asrc[i] + k0, As(slot) + (2 * w + i) * 512);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            dma16(bsrc[i] + k0, Bs(slot) + (4 * w + i) * 512);
    };
    f32x4 acc[4][8] = {};
    auto compute = [&](int slot) {
        const uint16_t *as = As(slot), *bs = Bs(slot);
        bf16x8 a[4], b[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = *(const bf16x8 *)&as[tile_off(wm + 16 * i + li, lc)];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            b[j] = *(const bf16x8 *)&bs[tile_off(wn + 16 * j + li, lc)];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[i], b[j], acc[i][j], 0, 0, 0);
    };
    const int ns = K / kFk;   // whole slices through the ring
    auto stage = [&](int sl, int slot) {
        if (sl + 2 < ns) {
            if (sl > 0) {
                SG_COMPILER_FENCE();
                SG_WAIT_LDS();
                SG_BARRIER();
            }
            issue(sl + 2, (slot + 2) % kRing);
            // six loads per slice per wavefront: slices sl + 1 and sl + 2
            // may stay in flight
            SG_WAIT_VM(12);
        } else {
            SG_WAIT_VM(0);
        }
        SG_BARRIER();
        compute(slot);
    };
    if (ns > 0) issue(0, 0);
    if (ns > 1) issue(1, 1);
    for (int sl = 0; sl < ns; sl += kRing) {
        stage(sl, 0);
        if (sl + 1 < ns) stage(sl + 1, 1);
        if (sl + 2 < ns) stage(sl + 2, 2);
    }
    if (K % kFk) {
        // the last partial slice through registers (columns past K are 0)
        __syncthreads();
        const int k0 = ns * kFk;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int q = t + 256 * i, r = q >> 2;
            *(u32x4 *)&A0[tile_off(r, q & 3)] =
                load8(X, M, K, m0 + r, k0 + (q & 3) * 8);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = t + 256 * i, r = q >> 2;
            *(u32x4 *)&B0[tile_off(r, q & 3)] =
                load8(W, N, K, n0 + r, k0 + (q & 3) * 8);
        }
        __syncthreads();
        compute(0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int col = n0 + wn + 16 * j + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + 16 * i + lc * 4 + r;
                if (row < M && col < N) C[(long)row * N + col] = acc[i][j][r];
            }
        }
}

// dW2[o][u] = the partial sums (chunks or batch rows): block `blk` of 256
// threads covers 64 outputs, its wavefront w adds the w-th quarter of the
// slots in slot order, and the quarters are added in order (dw2_reduce, or
// extra blocks of the dW1 product's launch).
__device__ __forceinline__ void dw2_reduce_block(
        const float *__restrict__ part, float *__restrict__ dw2, int chunks,
        int H, int O, int blk) {
    __shared__ float q[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blk * 64 + lane;
    const int per = (chunks + 3) / 4;
    const int c0 = w * per, c1 = min(chunks, c0 + per);
    float acc = 0.0f;
    if (i < O * H) {
        // eight loads in flight before their in-order adds: the sum is the
        // same as one slot at a time, without a memory latency per slot
        const long st = (long)O * H;
        const float *p = part + i;
        int c = c0;
        for (; c + 8 <= c1; c += 8) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = p[(c + j) * st];
#pragma unroll
            for (int j = 0; j < 8; ++j) acc += v[j];
        }
        for (; c < c1; ++c) acc += p[c * st];
    }
    q[w][lane] = acc;
    __syncthreads();
    if (w == 0 && i < O * H)
        dw2[i] = ((q[0][lane] + q[1][lane]) + q[2][lane]) + q[3][lane];
    __syncthreads();   // q is reused by the block's next output group
}

// dW1, H a multiple of 4.  A slice is 32 rows of L (row = 128 FP32 in
// 4-float units stored at unit u ^ (4 * ((row >> 3) & 3))) and the xp
// entries of its four groups ([4][112][8], 7 KB, padded to 8 KB so that
// every wavefront loads two chunks).  Row ranges are whole slices (per is a
// multiple of 32); rows past M read row M - 1 of L, against xp's zeros.
// L enters as two BF16 terms hi + mid (split2: 16 significant bits, two
// MFMAs per tile and slice): on the 784 -> 1024 -> 10 network (batch 128,
// 100 steps) dW1's relative L2 error against the FP64 model was 2.51e-7
// (spike mode) and 3.12e-6 (rate mode) against 2.47e-7 and 3.09e-6 with
// the third term, whose products the FP32 recurrences' own rounding
// outweighs.
__global__ __launch_bounds__(256) void gemm_dw1_dma(
        const float *__restrict__ L, const uint16_t *__restrict__ xp,
        float *__restrict__ part, int M, int H, int K, int per, long gp,
        const float *__restrict__ part2, float *__restrict__ dw2, int chunks2,
        int O) {
    __shared__ __align__(16) float L0[kDr * kDh], L1[kDr * kDh],
        L2[kDr * kDh];
    __shared__ __align__(16) uint16_t X0[8 * 512], X1[8 * 512], X2[8 * 512];
    auto Ls = [&](int slot) { return slot == 0 ? L0 : slot == 1 ? L1 : L2; };
    auto Xs = [&](int slot) { return slot == 0 ? X0 : slot == 1 ? X1 : X2; };
    const int nx = gridDim.x, ny = gridDim.y;
    const int nb = nx * ny * gridDim.z;
    const int hw = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const int lg = (nb % 8 == 0) ? (hw % 8) * (nb / 8) + hw / 8 : hw;
    const int bx = lg % nx, by = (lg / nx) % ny, bz = lg / (nx * ny);
    if (dw2 && bz == (int)gridDim.z - 1) {
        // the last z slice sums dW2 (dw2_reduce's work) beside the product
        const int ng = (O * H + 63) / 64;
        for (int g = bx + nx * by; g < ng; g += nx * ny)
            dw2_reduce_block(part2, dw2, chunks2, H, O, g);
        return;
    }
    const int h0 = by * kDh, kb = bx * kDk;
    const int rb = bz * per, re = min(M, rb + per);
    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const int li = lane & 15, gq = lane >> 4;
    // DMA sources: L chunk j = 4 w + i holds rows 2 j, 2 j + 1 of the slice
    int lrow[4], lcol[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = 2 * (4 * w + i) + (lane >> 5);
        const int u = (lane & 31) ^ (4 * ((row >> 3) & 3));
        lrow[i] = row;
        lcol[i] = min(h0 + 4 * u, H - 4);
    }
    // xp chunk j = 2 w + i: entry q = 64 j + lane = group q / 112, column
    // q % 112 (q >= 448: a spare slot, loaded from a clamped address)
    int xg[2], xk[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int q = 64 * (2 * w + i) + lane;
        xg[i] = q < 4 * kDk ? q / kDk : 3;
        xk[i] = min(kb + (q < 4 * kDk ? q % kDk : 0), K - 1);
    }
    auto issue = [&](int r0, int slot) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long row = min(r0 + lrow[i], M - 1);
            dma16(L + row * H + lcol[i], Ls(slot) + (4 * w + i) * 256);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const long g = min((long)(r0 / 8 + xg[i]), gp - 1);
            dma16(xp + (g * K + xk[i]) * 8, Xs(slot) + (2 * w + i) * 512);
        }
    };
    f32x4 acc[2][7] = {};
    auto compute = [&](int slot) {
        const float *ls = Ls(slot);
        const uint16_t *xs = Xs(slot);
        bf16x8 b[7];
#pragma unroll
        for (int n = 0; n < 7; ++n)
            b[n] = *(const bf16x8 *)&xs[(gq * kDk + 16 * n + li) * 8];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int hl = (32 * w + 16 * i + li) ^ (16 * gq);
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = ls[(8 * gq + j) * kDh + hl];
            bf16x8 hi, mid;
            split2(v, hi, mid);
#pragma unroll
            for (int n = 0; n < 7; ++n) {
                acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    hi, b[n], acc[i][n], 0, 0, 0);
                acc[i][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    mid, b[n], acc[i][n], 0, 0, 0);
            }
        }
    };
    const int ns = re > rb ? (re - rb + kDr - 1) / kDr : 0;
    auto stage = [&](int sl, int slot) {
        if (sl + 2 < ns) {
            if (sl > 0) {
                SG_COMPILER_FENCE();
                SG_WAIT_LDS();
                SG_BARRIER();
            }
            issue(rb + (sl + 2) * kDr, (slot + 2) % kRing);
            // six loads per slice per wavefront: slices sl + 1 and sl + 2
            // may stay in flight
            SG_WAIT_VM(12);
        } else {
            SG_WAIT_VM(0);
        }
        SG_BARRIER();
        compute(slot);
    };
    if (ns > 0) issue(rb, 0);
    if (ns > 1) issue(rb + kDr, 1);
    for (int sl = 0; sl < ns; sl += kRing) {
        stage(sl, 0);
        if (sl + 1 < ns) stage(sl + 1, 1);
        if (sl + 2 < ns) stage(sl + 2, 2);
    }
    float *P = part + (long)bz * H * K;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int n = 0; n < 7; ++n) {
            const int k = kb + 16 * n + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int h = h0 + 32 * w + 16 * i + 4 * gq + r;
                if (h < H && k < K) P[(long)h * K + k] = acc[i][n][r];
            }
        }
}

// D[i] = part[0][i] + part[1][i] + ... in range order, four at a time
// where n and D allow it.
__global__ void add_parts(const float *__restrict__ part, float *__restrict__ D,
                          long n, int parts) {
    const long n4 = (n % 4 == 0 && ((uintptr_t)D & 15) == 0) ? n / 4 : 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4;
         i += (long)gridDim.x * blockDim.x) {
        float4 acc = ((const float4 *)part)[i];
        for (int z = 1; z < parts; ++z) {
            const float4 v = ((const float4 *)(part + (long)z * n))[i];
            acc.x += v.x;
            acc.y += v.y;
            acc.z += v.z;
            acc.w += v.w;
        }
        ((float4 *)D)[i] = acc;
    }
    for (long i = 4 * n4 + blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        float acc = part[i];
        for (int z = 1; z < parts; ++z) acc += part[(long)z * n + i];
        D[i] = acc;
    }
}

constexpr int kWin = 8;   // steps whose loads a scan issues together

struct ScanArgs {
    int B, H, O, steps;          // batch, hidden, outputs, steps in block
    int mode;
    float beta, theta, slope;
    float reset_theta;           // reset * theta
};

// Checkpoint slot: v1 [B][H], s1 [B][H], v2 [B][O], s2 [B][O].
__host__ __device__ inline long slot_floats(int B, int H, int O) {
    return 2L * B * H + 2L * B * O;
}

// Layer 1 over one block of steps; thread = (b, u).  s1 of every step goes
// to s1_out; rec_v1 (or null) records v1.
__global__ void scan1_fwd(ScanArgs a, const float *__restrict__ a1,
                          const float *__restrict__ state_in,
                          float *__restrict__ state_out,
                          float *__restrict__ s1_out,
                          float *__restrict__ rec_v1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = a.H;
    if (i >= a.B * H) return;
    const long BH = (long)a.B * H;
    float v = state_in[i], s = state_in[BH + i];
    // the block's currents in windows of kWin steps, each window's loads
    // issued together
    for (int t0 = 0; t0 < a.steps; t0 += kWin) {
        float in[kWin];
#pragma unroll
        for (int j = 0; j < kWin; ++j)
            if (t0 + j < a.steps) in[j] = a1[(long)(t0 + j) * BH + i];
#pragma unroll
        for (int j = 0; j < kWin; ++j) {
            if (t0 + j >= a.steps) break;
            const long idx = (long)(t0 + j) * BH + i;
            v = a.beta * v - a.reset_theta * s + in[j];
            s = act(v - a.theta, a.mode, a.slope);
            s1_out[idx] = s;
            if (rec_v1) rec_v1[idx] = v;
        }
    }
    if (state_out) {
        state_out[i] = v;
        state_out[BH + i] = s;
    }
}

// W2 transposed to [u][16] BF16 (outputs padded to 16 with zeros): a lane
// reads all outputs of its unit with two 16 B loads.
__global__ void transpose_w2(const uint16_t *__restrict__ w2,
                             uint16_t *__restrict__ w2t, int H, int O) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * kMaxOut) return;
    const int u = i / kMaxOut, o = i % kMaxOut;
    w2t[i] = o < O ? w2[o * H + u] : (uint16_t)0;
}

// A2[row][o] = sum over u of W2[o][u] s1[row][u], one wavefront per row:
// lane partial sums over u = lane + 64 j (j ascending), then a fixed
// butterfly across the wavefront.
// p[o] (lane-sum across the wavefront still to do) for one row.
__device__ __forceinline__ void row_partials(const float *__restrict__ s1,
                                             const uint16_t *__restrict__ w2t,
                                             long row, int H, int lane,
                                             float (&p)[kMaxOut]) {
    for (int o = 0; o < kMaxOut; ++o) p[o] = 0.0f;
    // eight units per lane in flight: the loads of a group are issued
    // before its products (the same u order as one at a time)
    constexpr int G = 8;
    for (int u0 = lane; u0 < H; u0 += 64 * G) {
        float sv[G];
        u32x4 lo[G], hi[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int u = u0 + 64 * j;
            if (u < H) {
                sv[j] = s1[row * H + u];
                lo[j] = *(const u32x4 *)&w2t[u * kMaxOut];
                hi[j] = *(const u32x4 *)&w2t[u * kMaxOut + 8];
            }
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            if (u0 + 64 * j >= H) break;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                p[2 * q] += bf16_value(lo[j][q] & 0xFFFFu) * sv[j];
                p[2 * q + 1] += bf16_value(lo[j][q] >> 16) * sv[j];
                p[8 + 2 * q] += bf16_value(hi[j][q] & 0xFFFFu) * sv[j];
                p[9 + 2 * q] += bf16_value(hi[j][q] >> 16) * sv[j];
            }
        }
    }
}

__global__ __launch_bounds__(256) void out_current(
    const float *__restrict__ s1, const uint16_t *__restrict__ w2t,
    float *__restrict__ a2, int M, int H, int O) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= M) return;
    float p[kMaxOut];
    row_partials(s1, w2t, row, H, lane, p);
#pragma unroll
    for (int o = 0; o < kMaxOut; ++o)
        if (o < O) {
            const float v = wave_sum(p[o]);
            if (lane == 0) a2[(long)row * O + o] = v;
        }
}

// Layer 2 over one block of steps; thread = (b, o).  count_mode 0: leave
// counts, 1: start them at 0, 2: add to them.
__global__ void scan2_fwd(ScanArgs a, const float *__restrict__ a2,
                          const float *__restrict__ state_in,
                          float *__restrict__ state_out,
                          float *__restrict__ counts, int count_mode,
                          float *__restrict__ rec_v2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int O = a.O;
    if (i >= a.B * O) return;
    const long BH = (long)a.B * a.H, BO = (long)a.B * O;
    float v = state_in[2 * BH + i], s = state_in[2 * BH + BO + i];
    float cnt = count_mode == 2 ? counts[i] : 0.0f;
    for (int t0 = 0; t0 < a.steps; t0 += kWin) {
        float in[kWin];
#pragma unroll
        for (int j = 0; j < kWin; ++j)
            if (t0 + j < a.steps) in[j] = a2[(long)(t0 + j) * BO + i];
#pragma unroll
        for (int j = 0; j < kWin; ++j) {
            if (t0 + j >= a.steps) break;
            v = a.beta * v - a.reset_theta * s + in[j];
            s = act(v - a.theta, a.mode, a.slope);
            cnt += s;
            if (rec_v2) rec_v2[(long)(t0 + j) * BO + i] = v;
        }
    }
    if (state_out) {
        state_out[2 * BH + i] = v;
        state_out[2 * BH + BO + i] = s;
    }
    if (count_mode != 0) counts[i] = cnt;
}

// out_current and scan2_fwd in one launch for checkpoints of at most 16
// steps: block = batch row b, wavefront t = step t of the block (row
// t * B + b) reduces W2 . s1 exactly as out_current into LDS, then threads
// o < O run layer 2's recurrence as scan2_fwd does.  Same values, one
// launch and no A2 round trip through memory.
__global__ __launch_bounds__(1024) void out_scan2_fwd(
        ScanArgs a, const float *__restrict__ s1,
        const uint16_t *__restrict__ w2t, const float *__restrict__ state_in,
        float *__restrict__ state_out, float *__restrict__ counts,
        int count_mode, float *__restrict__ rec_v2) {
    __shared__ float a2s[16][kMaxOut];
    const int b = blockIdx.x, t = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int O = a.O;
    {
        float p[kMaxOut];
        row_partials(s1, w2t, (long)t * a.B + b, a.H, lane, p);
#pragma unroll
        for (int o = 0; o < kMaxOut; ++o)
            if (o < O) {
                const float v = wave_sum(p[o]);
                if (lane == 0) a2s[t][o] = v;
            }
    }
    __syncthreads();
    if (threadIdx.x >= O) return;
    const int i = b * O + threadIdx.x;
    const long BH = (long)a.B * a.H, BO = (long)a.B * O;
    float v = state_in[2 * BH + i], s = state_in[2 * BH + BO + i];
    float cnt = count_mode == 2 ? counts[i] : 0.0f;
    for (int tt = 0; tt < a.steps; ++tt) {
        v = a.beta * v - a.reset_theta * s + a2s[tt][threadIdx.x];
        s = act(v - a.theta, a.mode, a.slope);
        cnt += s;
        if (rec_v2) rec_v2[(long)tt * BO + i] = v;
    }
    if (state_out) {
        state_out[2 * BH + i] = v;
        state_out[2 * BH + BO + i] = s;
    }
    if (count_mode != 0) counts[i] = cnt;
}

// w2f[(kk * 64 + lane) * 8 + e] = W2[lane % 16][32 kk + 8 (lane / 16) + e]
// (0 past O or H): the B fragment of units 32 kk .. 32 kk + 31; entry i,
// written by extra blocks of scan1_fwd_all's launch.
__device__ __forceinline__ void frag_w2_at(const uint16_t *__restrict__ w2,
                                           uint16_t *__restrict__ w2f, int H,
                                           int O, int nk, int i) {
    if (i >= nk * 64 * 8) return;
    const int e = i & 7, lane = (i >> 3) & 63, kk = i >> 9;
    const int o = lane & 15, u = 32 * kk + 8 * (lane >> 4) + e;
    w2f[i] = (o < O && u < H) ? w2[o * H + u] : (uint16_t)0;
}

// The forward over all T steps in two passes (T <= 1024; layer 1 does not
// read layer 2).  scan1_fwd_all: layer 1, thread = (b, u), the recurrence
// of scan1_fwd over every step; s1 of every step to s1_all (to s1_bf as
// BF16 in spike mode), the state after each block of ck steps to that
// block's checkpoint slot.
__global__ void scan1_fwd_all(ScanArgs a, int T, int ck, long slot,
                              const float *__restrict__ a1,
                              float *__restrict__ state,
                              float *__restrict__ s1_all,
                              uint16_t *__restrict__ s1_bf,
                              const uint16_t *__restrict__ w2,
                              uint16_t *__restrict__ w2f, int nk,
                              int scan_blocks) {
    // the blocks past the scan's put W2 in fragment order for the output
    // currents in the same launch
    if ((int)blockIdx.x >= scan_blocks) {
        frag_w2_at(w2, w2f, a.H, a.O, nk,
                   ((int)blockIdx.x - scan_blocks) * blockDim.x +
                       threadIdx.x);
        return;
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = a.H;
    if (i >= a.B * H) return;
    const long BH = (long)a.B * H;
    // the run starts from v1 = s1 = 0, written as checkpoint 0
    float v = 0.0f, s = 0.0f;
    state[i] = 0.0f;
    state[BH + i] = 0.0f;
    int left = ck;
    float *next = state + slot;
    for (int t0 = 0; t0 < T; t0 += kWin) {
        float in[kWin];
#pragma unroll
        for (int j = 0; j < kWin; ++j)
            if (t0 + j < T) in[j] = a1[(long)(t0 + j) * BH + i];
#pragma unroll
        for (int j = 0; j < kWin; ++j) {
            if (t0 + j >= T) break;
            v = a.beta * v - a.reset_theta * s + in[j];
            s = act(v - a.theta, a.mode, a.slope);
            // spikes (0 / 1, exact in BF16) as BF16, else FP32
            if (s1_bf)
                s1_bf[(long)(t0 + j) * BH + i] =
                    s != 0.0f ? (uint16_t)0x3F80u : (uint16_t)0u;
            else
                s1_all[(long)(t0 + j) * BH + i] = s;
            if (--left == 0) {
                left = ck;
                next[i] = v;
                next[BH + i] = s;
                next += slot;
            }
        }
    }
}

// The forward's output currents for every step as one MFMA product over
// the whole grid: A2[row][o] = sum over u of S1[row][u] W2[o][u], rows =
// t * B + b.  W2 enters as the B operand in fragment order (w2f, BF16,
// exact); S1 as the A operand, exact in BF16 in spike mode (0 / 1) and as
// three BF16 terms hi + mid + lo in rate mode, which keeps FP32 precision.
// A block of four wavefronts takes 16 rows; wavefront w multiplies units
// [w H / 4, (w + 1) H / 4) with all of its loads issued first, and the four
// partial tiles are added in wavefront order.  Every sum runs in a fixed
// order.
constexpr int kOutWaves = 4;

template <bool SPLIT>
__global__ __launch_bounds__(256) void out_current_mfma(
        const void *__restrict__ s1v, const uint16_t *__restrict__ w2f,
        float *__restrict__ a2, long rows, int H, int O, int nk) {
    __shared__ f32x4 part[kOutWaves - 1][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const long r0 = (long)blockIdx.x * 16;
    const long row = min(r0 + li, rows - 1);   // past the end: never stored
    // S1 in FP32 (SPLIT) or as BF16 spikes
    const float *srow = (const float *)s1v + row * H;
    const uint16_t *brow = (const uint16_t *)s1v + row * H;
    const int per = (nk + kOutWaves - 1) / kOutWaves;
    const int k0 = w * per, k1 = min(nk, k0 + per);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const bool vec = (H & 7) == 0;
    constexpr int F = 8;   // fragments whose loads are in flight together
    for (int kb = k0; kb < k1; kb += F) {
        float v[F][8];
        bf16x8 b[F], sb[F];
#pragma unroll
        for (int q = 0; q < F; ++q) {
            const int kk = kb + q;
            if (kk >= k1) break;
            const int u0 = 32 * kk + 8 * lk;
            if (!SPLIT) {
                if (vec && u0 + 8 <= H) {
                    sb[q] = *(const bf16x8 *)&brow[u0];
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        sb[q][e] = u0 + e < H ? __builtin_bit_cast(
                                                    __bf16, brow[u0 + e])
                                              : (__bf16)0.0f;
                }
            } else if (vec && u0 + 8 <= H) {
                const float4 x0 = *(const float4 *)&srow[u0];
                const float4 x1 = *(const float4 *)&srow[u0 + 4];
                v[q][0] = x0.x; v[q][1] = x0.y; v[q][2] = x0.z; v[q][3] = x0.w;
                v[q][4] = x1.x; v[q][5] = x1.y; v[q][6] = x1.z; v[q][7] = x1.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    v[q][e] = u0 + e < H ? srow[u0 + e] : 0.0f;
            }
            b[q] = *(const bf16x8 *)&w2f[((long)kk * 64 + lane) * 8];
        }
#pragma unroll
        for (int q = 0; q < F; ++q) {
            if (kb + q >= k1) break;
            if (SPLIT) {
                bf16x8 hi, mid, lo;
                split3(v[q], hi, mid, lo);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi, b[q], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(mid, b[q], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo, b[q], acc, 0, 0, 0);
            } else {
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sb[q], b[q], acc, 0, 0, 0);
            }
        }
    }
    if (w > 0) part[w - 1][lane] = acc;
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int p = 0; p < kOutWaves - 1; ++p) {
        const f32x4 o4 = part[p][lane];
        acc[0] += o4[0];
        acc[1] += o4[1];
        acc[2] += o4[2];
        acc[3] += o4[3];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const long rr = r0 + 4 * lk + r;
        if (rr < rows && li < O) a2[rr * O + li] = acc[r];
    }
}

// Layer 2 over every step from the all-steps currents; thread = (b, o):
// v2 of every step to rec_v2, the state after each block of ck steps to
// that block's checkpoint slot, the spike counts to counts.  The currents
// of 16 steps are loaded together, the next 16 while these are stepped
// (the grid is only batch x outputs threads: each thread's chain of loads
// and steps is the launch).  MODE as in scan_bwd_fused.
template <int MODE>
__global__ void scan2_fwd_all(ScanArgs a, int T, int ck, long slot,
                              const float *__restrict__ a2,
                              float *__restrict__ state,
                              float *__restrict__ counts,
                              float *__restrict__ rec_v2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int O = a.O;
    if (i >= a.B * O) return;
    const long BH = (long)a.B * a.H, BO = (long)a.B * O;
    // the run starts from v2 = s2 = 0, written as checkpoint 0
    float v = 0.0f, s = 0.0f;
    state[2 * BH + i] = 0.0f;
    state[2 * BH + BO + i] = 0.0f;
    float cnt = 0.0f;
    int left = ck;
    float *next = state + slot;
    constexpr int W = 16;
    float in_n[W];
#pragma unroll
    for (int j = 0; j < W; ++j)
        in_n[j] = j < T ? a2[(long)j * BO + i] : 0.0f;
    for (int t0 = 0; t0 < T; t0 += W) {
        float in[W];
#pragma unroll
        for (int j = 0; j < W; ++j) in[j] = in_n[j];
#pragma unroll
        for (int j = 0; j < W; ++j)
            if (t0 + W + j < T) in_n[j] = a2[(long)(t0 + W + j) * BO + i];
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (t0 + j >= T) break;
            v = a.beta * v - a.reset_theta * s + in[j];
            s = act_m<MODE>(v - a.theta, a.mode, a.slope);
            cnt += s;
            rec_v2[(long)(t0 + j) * BO + i] = v;
            if (--left == 0) {
                left = ck;
                next[2 * BH + i] = v;
                next[2 * BH + BO + i] = s;
                next += slot;
            }
        }
    }
    counts[i] = cnt;
}

// Adjoint of layer 2 over one block of steps, in reverse; thread = (b, o):
// d2 = (g - r theta d2') S'(v2 - theta) + beta d2' (d2': the next step's,
// carried between blocks; r = reset).
__global__ void scan2_bwd(ScanArgs a, const float *__restrict__ rec_v2,
                          const float *__restrict__ g,
                          float *__restrict__ carry2, float *__restrict__ g2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int O = a.O;
    if (i >= a.B * O) return;
    const long BO = (long)a.B * O;
    const float gb = g[i];
    float dn = carry2[i];
    for (int t1 = a.steps - 1; t1 >= 0; t1 -= kWin) {
        float vv[kWin];
#pragma unroll
        for (int j = 0; j < kWin; ++j)
            if (t1 - j >= 0) vv[j] = rec_v2[(long)(t1 - j) * BO + i];
#pragma unroll
        for (int j = 0; j < kWin; ++j) {
            const int t = t1 - j;
            if (t < 0) break;
            const float d = (gb - a.reset_theta * dn) *
                            act_grad(vv[j] - a.theta, a.mode, a.slope) +
                            a.beta * dn;
            g2[(long)t * BO + i] = d;
            dn = d;
        }
    }
    carry2[i] = dn;
}

// dL/dcount[b][o] = scale * (softmax(count[b])[o] - onehot(label[b])[o]) / B,
// the same operations wherever it is computed (loss_grad_kernel, the fused
// backward scan), so both give the same bits.
__device__ inline float loss_grad_at(const float *__restrict__ counts,
                                     const int32_t *__restrict__ labels,
                                     int B, int O, float scale, int b, int o) {
    const float *c = counts + b * O;
    float m = c[0];
    for (int k = 1; k < O; ++k) m = fmaxf(m, c[k]);
    float s = 0.0f;
    for (int k = 0; k < O; ++k) s += expf(c[k] - m);
    const float p = expf(c[o] - m) / s;
    return scale * (p - (o == labels[b] ? 1.0f : 0.0f)) / (float)B;
}

// The whole backward in one launch (n_hid a multiple of 256, checkpoints
// of at most 16 steps): a block of 256 threads = 256 hidden units of one
// batch row b; it walks the blocks of steps from the last, and per block
//  - threads o < O run layer 2's adjoint for (b, o) over the block's steps
//    exactly as scan2_bwd does, into LDS (every block of the row computes
//    the same values; the carry stays in a register);
//  - every thread runs layer 1's recurrence of its (b, u) again from the
//    block's checkpoint (the currents a1; v1 in registers, s1 = S(v1 -
//    theta) as scan1_fwd computes them), then its adjoint in reverse as
//    scan1_bwd does (d1 of every step to lam, the carry in a register),
//    and the block's share of dW2, the sum over its steps (latest first)
//    of d2[o] * s1[u], added to a running sum in block order.
// The dW2 sums go to part2[b][o][u] (dw2_reduce adds the batch rows in a
// fixed order).  No scan2_bwd, dw2_partial or recomputing launch.  MODE
// (>= 0) and OUT (> 0) fix the mode and the number of outputs at compile
// time: the per-step loops then carry no guard or select for the other
// mode or for outputs past O.
template <int MODE, int OUT>
__global__ __launch_bounds__(256) void scan_bwd_fused(
        ScanArgs a, int nblk, long slot, const uint16_t *__restrict__ w2,
        const float *__restrict__ a1, const float *__restrict__ state,
        const float *__restrict__ rec_v2, const float *__restrict__ counts,
        const int32_t *__restrict__ labels, float scale,
        float *__restrict__ lam, float *__restrict__ part2,
        const uint16_t *__restrict__ X, uint16_t *__restrict__ xp, int MT,
        int n_in, long gp, int scan_blocks) {
    // the blocks past the scan's regroup x for the dW1 product (pack_x):
    // memory work beside the scan's arithmetic, in the same launch
    if ((int)blockIdx.x >= scan_blocks) {
        pack_x_block(X, xp, MT, n_in, gp, (long)blockIdx.x - scan_blocks);
        return;
    }
    constexpr int NO = OUT > 0 ? OUT : kMaxOut;   // output slots walked
    __shared__ float d2s[16][kMaxOut];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = a.H, O = OUT > 0 ? OUT : a.O;
    const long BH = (long)a.B * H, BO = (long)a.B * O;
    const long M = (long)a.steps * a.B;   // rows of a block of steps
    const int b = i / H, u = i % H;
    const bool out_thread = threadIdx.x < O;
    const int i2 = b * O + (out_thread ? threadIdx.x : 0);
    const float gb = out_thread ? loss_grad_at(counts, labels, a.B, O, scale, b,
                                               threadIdx.x)
                                : 0.0f;
    float wc[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o)
        wc[o] = o < O ? bf16_value(w2[o * H + u]) : 0.0f;
    float dn = 0.0f, dn2 = 0.0f;
    float tot[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) tot[o] = 0.0f;
    // the loads of block blk - 1 (currents, layer 2's potentials, the
    // checkpoint) are issued at the start of block blk and land while it
    // is computed
    float in_n[16] = {}, v2_n[16] = {}, ckv_n = 0.0f, cks_n = 0.0f;
    auto fetch = [&](int blk) {
        const float *ab = a1 + blk * M * H;
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (t < a.steps) in_n[t] = ab[(long)t * BH + i];
        const float *ck = state + blk * slot;
        ckv_n = ck[i];
        cks_n = ck[BH + i];
        if (out_thread) {
            const float *rv = rec_v2 + blk * M * O;
#pragma unroll
            for (int t = 0; t < 16; ++t)
                if (t < a.steps) v2_n[t] = rv[(long)t * BO + i2];
        }
    };
    fetch(nblk - 1);
    for (int blk = nblk - 1; blk >= 0; --blk) {
        float in[16], v2[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            in[t] = in_n[t];
            v2[t] = v2_n[t];
        }
        float v = ckv_n, s = cks_n;
        if (blk > 0) fetch(blk - 1);
        if (out_thread) {
#pragma unroll
            for (int t = 15; t >= 0; --t) {
                if (t >= a.steps) continue;
                const float d = (gb - a.reset_theta * dn2) *
                                act_grad_fast_m<MODE>(v2[t] - a.theta, a.mode,
                                                      a.slope) +
                                a.beta * dn2;
                d2s[t][threadIdx.x] = d;
                dn2 = d;
            }
        }
        // layer 1 forward from the checkpoint
        float vr[16];
        __syncthreads();   // d2s of this block complete
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            if (t >= a.steps) break;
            v = a.beta * v - a.reset_theta * s + in[t];
            s = act_m<MODE>(v - a.theta, a.mode, a.slope);
            vr[t] = v;
        }
        float acc[NO];
#pragma unroll
        for (int o = 0; o < NO; ++o) acc[o] = 0.0f;
        float *lb = lam + blk * M * H;
#pragma unroll
        for (int t = 15; t >= 0; --t) {
            if (t >= a.steps) continue;
            const float sv = act_m<MODE>(vr[t] - a.theta, a.mode, a.slope);
            float gs = 0.0f;
#pragma unroll
            for (int o = 0; o < NO; ++o)
                if (OUT > 0 || o < O) {
                    gs = fmaf(wc[o], d2s[t][o], gs);
                    acc[o] = fmaf(d2s[t][o], sv, acc[o]);
                }
            const float d = (gs - a.reset_theta * dn) *
                            act_grad_fast_m<MODE>(vr[t] - a.theta, a.mode,
                                                  a.slope) +
                            a.beta * dn;
            lb[(long)t * BH + i] = d;
            dn = d;
        }
#pragma unroll
        for (int o = 0; o < NO; ++o)
            tot[o] = blk == nblk - 1 ? acc[o] : tot[o] + acc[o];
        __syncthreads();   // d2s read by every thread before it changes
    }
#pragma unroll
    for (int o = 0; o < NO; ++o)
        if (o < O) part2[((long)b * O + o) * H + u] = tot[o];
}

// Adjoint of layer 1 over one block of steps, in reverse; thread = (b, u):
// d1 = (sum over o of W2[o][u] d2[o] - r theta d1') S'(v1 - theta) +
// beta d1'.
__global__ void scan1_bwd(ScanArgs a, const uint16_t *__restrict__ w2,
                          const float *__restrict__ rec_v1,
                          const float *__restrict__ g2,
                          float *__restrict__ carry1,
                          float *__restrict__ lam) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = a.H, O = a.O;
    if (i >= a.B * H) return;
    const int b = i / H, u = i % H;
    const long BH = (long)a.B * H, BO = (long)a.B * O;
    float wc[kMaxOut];
#pragma unroll
    for (int o = 0; o < kMaxOut; ++o)
        wc[o] = o < O ? bf16_value(w2[o * H + u]) : 0.0f;
    float dn = carry1[i];
    // windows of kWin steps from the last: each window's loads (v1 and the
    // output adjoints) issued together, then the recurrence
    for (int t1 = a.steps - 1; t1 >= 0; t1 -= kWin) {
        float vv[kWin], gsum[kWin];
#pragma unroll
        for (int j = 0; j < kWin; ++j) {
            const int t = t1 - j;
            if (t < 0) break;
            vv[j] = rec_v1[(long)t * BH + i];
            const float *d2 = g2 + (long)t * BO + (long)b * O;
            float gs = 0.0f;
#pragma unroll
            for (int o = 0; o < kMaxOut; ++o)
                if (o < O) gs += wc[o] * d2[o];
            gsum[j] = gs;
        }
#pragma unroll
        for (int j = 0; j < kWin; ++j) {
            const int t = t1 - j;
            if (t < 0) break;
            const float d = (gsum[j] - a.reset_theta * dn) *
                            act_grad(vv[j] - a.theta, a.mode, a.slope) +
                            a.beta * dn;
            lam[(long)t * BH + i] = d;
            dn = d;
        }
    }
    carry1[i] = dn;
}

constexpr int kDw2Rows = 16;   // rows of one partial sum of dW2

// dW2 partial sums: part[c][o][u] += sum over the rows r of chunk c
// (ascending) of G2[r][o] * S1[r][u]; every block of steps adds to the
// same chunk slots, latest block first.
__global__ void dw2_partial(const float *__restrict__ g2,
                            const float *__restrict__ s1,
                            float *__restrict__ part, int M, int H, int O) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (u >= H) return;
    float acc[kMaxOut];
    for (int o = 0; o < kMaxOut; ++o) acc[o] = 0.0f;
    const int r0 = c * kDw2Rows, r1 = min(M, r0 + kDw2Rows);
    // the chunk's s1 values first, then the products in row order
    float sv[kDw2Rows];
#pragma unroll
    for (int j = 0; j < kDw2Rows; ++j)
        sv[j] = r0 + j < r1 ? s1[(long)(r0 + j) * H + u] : 0.0f;
#pragma unroll
    for (int j = 0; j < kDw2Rows; ++j) {
        if (r0 + j >= r1) break;
#pragma unroll
        for (int o = 0; o < kMaxOut; ++o)
            if (o < O) acc[o] += g2[(long)(r0 + j) * O + o] * sv[j];
    }
#pragma unroll
    for (int o = 0; o < kMaxOut; ++o)
        if (o < O) part[((long)c * O + o) * H + u] += acc[o];
}

__global__ __launch_bounds__(256) void dw2_reduce(
        const float *__restrict__ part, float *__restrict__ dw2, int chunks,
        int H, int O) {
    dw2_reduce_block(part, dw2, chunks, H, O, blockIdx.x);
}

// Mean softmax cross-entropy over the batch (one block; sums in batch
// order), and the counts copied out.
__global__ void loss_kernel(const float *__restrict__ counts,
                            const int32_t *__restrict__ labels, int B, int O,
                            float *__restrict__ loss,
                            float *__restrict__ counts_out) {
    extern __shared__ float lb[];
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        const float *c = counts + b * O;
        float m = c[0];
        for (int o = 1; o < O; ++o) m = fmaxf(m, c[o]);
        float s = 0.0f;
        for (int o = 0; o < O; ++o) s += expf(c[o] - m);
        const int lab = labels[b];
        lb[b] = (lab >= 0 && lab < O) ? m + logf(s) - c[lab]
                                      : __int_as_float(0x7fc00000);
        if (counts_out)
            for (int o = 0; o < O; ++o) counts_out[b * O + o] = c[o];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // in batch order; four LDS words a read
        float s = 0.0f;
        int b = 0;
        for (; b + 4 <= B; b += 4) {
            const float4 v = *(const float4 *)(lb + b);
            s += v.x;
            s += v.y;
            s += v.z;
            s += v.w;
        }
        for (; b < B; ++b) s += lb[b];
        *loss = s / (float)B;
    }
}

// dL/dcount = scale * (softmax(count) - onehot(label)) / B.
__global__ void loss_grad_kernel(const float *__restrict__ counts,
                                 const int32_t *__restrict__ labels, int B,
                                 int O, float scale, float *__restrict__ g) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    for (int o = 0; o < O; ++o)
        g[b * O + o] = loss_grad_at(counts, labels, B, O, scale, b, o);
}

__global__ void fill_kernel(float *p, long n, float v) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace snn
}  // namespace sg

using namespace sg::snn;

namespace {

thread_local std::string g_error;

int fail(const std::string &m) {
    g_error = m;
    return 1;
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace carve-up, in bytes from the base.
struct Layout {
    size_t state, counts, grad, a1, s1, a2, rv1, rv2, carry1, carry2, lam,
        g2, part2, part1, w2t, xp, a2all, w2f, total;
};

// The whole backward runs in one launch (scan_bwd_fused) where each
// 256-thread block lies in one batch row and a block of steps fits its
// LDS table of output adjoints.
bool fused_backward(const sg_snn_cfg &c) {
    return c.n_hid % 256 == 0 && c.checkpoint <= 16;
}

// dW2 partial-sum slots: one per batch row (fused backward) or per
// kDw2Rows rows of a block of steps.
size_t dw2_slots(const sg_snn_cfg &c) {
    return fused_backward(c)
        ? (size_t)c.batch
        : ((size_t)c.checkpoint * c.batch + kDw2Rows - 1) / kDw2Rows;
}

Layout layout(const sg_snn_cfg &c) {
    const size_t B = c.batch, H = c.n_hid, O = c.n_out, ck = c.checkpoint;
    const size_t slots = (size_t)c.steps / ck + 1;
    const size_t M = ck * B, MT = (size_t)c.steps * B;
    Layout l;
    size_t off = 0;
    auto take = [&](size_t floats) { size_t o = off; off += al(floats * 4); return o; };
    l.state = take(slots * slot_floats(c.batch, c.n_hid, c.n_out));
    l.counts = take(B * O);
    l.grad = take(B * O);
    l.a1 = take(MT * H);            // input currents of every step
    l.s1 = take(M * H);
    l.a2 = take(M * O);
    l.rv1 = take(M * H);
    l.rv2 = take(MT * O);           // v2 of every step, from the forward
    l.carry1 = take(B * H);
    l.carry2 = take(B * O);
    l.lam = take(MT * H);           // dL/dv1 of every step
    l.g2 = take(M * O);
    l.part2 = take(dw2_slots(c) * O * H);
    l.part1 = take((size_t)kTnSplit * H * c.n_in);
    l.w2t = take((H * kMaxOut + 1) / 2);
    l.a2all = take(MT * O);         // output currents of every step
    l.w2f = take((H + 31) / 32 * 64 * 8 / 2);   // W2 in B-fragment order
    // x regrouped by 8 rows (BF16), groups padded to whole 32-row slices
    l.xp = take(((MT + 7) / 8 + 3) / 4 * 4 * c.n_in * 4);
    l.total = off;
    return l;
}

int check(const sg_snn_cfg *c) {
    if (!c) return fail("null configuration");
    if (c->batch < 1 || c->steps < 1 || c->n_in < 1 || c->n_hid < 1 ||
        c->n_out < 1)
        return fail("sizes must be positive");
    if (c->n_hid > 8192) return fail("n_hid above 8192");
    if (c->n_out > kMaxOut) return fail("n_out above 16");
    if (c->batch > 16384) return fail("batch above 16384");
    if (c->checkpoint < 1 || c->steps % c->checkpoint != 0)
        return fail("checkpoint must divide steps");
    if ((long)c->steps * c->batch > 65535L * 64)
        return fail("steps * batch above 4194240");
    if (c->mode != SG_SNN_SPIKE && c->mode != SG_SNN_RATE)
        return fail("unknown mode");
    return 0;
}

void fill(float *p, long n, float v, hipStream_t s) {
    hipLaunchKernelGGL(fill_kernel, dim3(256), dim3(256), 0, s, p, n, v);
}

ScanArgs scan_args(const sg_snn_cfg &c) {
    ScanArgs a;
    a.B = c.batch;
    a.H = c.n_hid;
    a.O = c.n_out;
    a.steps = c.checkpoint;
    a.mode = c.mode;
    a.beta = c.beta;
    a.theta = c.theta;
    a.slope = c.slope;
    a.reset_theta = c.reset * c.theta;
    return a;
}

unsigned blocks(long threads) { return (unsigned)((threads + 255) / 256); }

// The forward passes of block `blk` of steps from checkpoint `in`: state
// to `out` (or null), counts per count_mode, v records when rec_v1 is set.
// The block's input currents are already in a1.
// One block of steps from the checkpoint `in`.  Forward (recompute
// false): both layers, the next checkpoint to `out`, v2 of every step
// kept.  Backward (recompute true): layer 1 only, recording v1 and s1.
void forward_block(const sg_snn_cfg &c, const ScanArgs &a, int blk,
                   const Layout &l, unsigned char *ws, const float *in,
                   float *out, int count_mode, bool recompute,
                   hipStream_t s) {
    const int M = c.checkpoint * c.batch;
    float *a1 = (float *)(ws + l.a1) + (long)blk * M * c.n_hid;
    float *s1 = (float *)(ws + l.s1);
    float *a2 = (float *)(ws + l.a2), *cnt = (float *)(ws + l.counts);
    float *rv2 = (float *)(ws + l.rv2) + (long)blk * M * c.n_out;
    hipLaunchKernelGGL(scan1_fwd, dim3(blocks((long)c.batch * c.n_hid)),
                       dim3(256), 0, s, a, a1, in, out, s1,
                       recompute ? (float *)(ws + l.rv1) : nullptr);
    if (recompute) return;
    if (c.checkpoint <= 16) {
        hipLaunchKernelGGL(out_scan2_fwd, dim3(c.batch),
                           dim3(64 * c.checkpoint), 0, s, a, s1,
                           (const uint16_t *)(ws + l.w2t), in, out, cnt,
                           count_mode, rv2);
        return;
    }
    hipLaunchKernelGGL(out_current, dim3((M + 3) / 4), dim3(256), 0, s, s1,
                       (const uint16_t *)(ws + l.w2t), a2, M, c.n_hid,
                       c.n_out);
    hipLaunchKernelGGL(scan2_fwd, dim3(blocks((long)c.batch * c.n_out)),
                       dim3(256), 0, s, a, a2, in, out, cnt, count_mode, rv2);
}

}  // namespace

extern "C" {

const char *sg_snn_error(void) { return g_error.c_str(); }

int sg_snn_workspace_bytes(const sg_snn_cfg *cfg, size_t *bytes) {
    if (check(cfg)) return 1;
    if (bytes) *bytes = layout(*cfg).total;
    return 0;
}

int sg_snn_forward(const sg_snn_cfg *cfg, const uint16_t *x,
                   const uint16_t *w1, const uint16_t *w2,
                   const int32_t *labels, void *workspace, float *loss,
                   float *counts, void *stream) {
    if (check(cfg)) return 1;
    if (!x || !w1 || !w2 || !labels || !workspace || !loss)
        return fail("null pointer");
    const sg_snn_cfg &c = *cfg;
    hipStream_t s = (hipStream_t)stream;
    const Layout l = layout(c);
    unsigned char *ws = (unsigned char *)workspace;
    float *state = (float *)(ws + l.state);
    float *cnt = (float *)(ws + l.counts);
    const long slot = slot_floats(c.batch, c.n_hid, c.n_out);
    const ScanArgs a = scan_args(c);
    // the initial state (checkpoint 0) is zero: written by the all-steps
    // scans themselves, cleared here for the per-block path
    if (c.steps > kAllSteps) fill(state, slot, 0.0f, s);
    // input currents of every step in one GEMM (rows t * batch + b); the
    // backward reuses them and recomputes only the recurrences
    const int MT = c.steps * c.batch;
    const dim3 fgrid((c.n_hid + kFn - 1) / kFn, (MT + kFm - 1) / kFm);
    if (c.n_in % 8 == 0)
        hipLaunchKernelGGL(gemm_fwd_dma, fgrid, dim3(256), 0, s, x, w1,
                           (float *)(ws + l.a1), MT, c.n_hid, c.n_in);
    else
        hipLaunchKernelGGL(gemm_fwd, fgrid, dim3(256), 0, s, x, w1,
                           (float *)(ws + l.a1), MT, c.n_hid, c.n_in);
    // W2 transposed for the per-block passes (the per-block forward, and
    // the backward's recomputation when it is not the fused one)
    if (c.steps > kAllSteps || !fused_backward(c))
        hipLaunchKernelGGL(transpose_w2,
                           dim3(blocks((long)c.n_hid * kMaxOut)), dim3(256), 0,
                           s, w2, (uint16_t *)(ws + l.w2t), c.n_hid, c.n_out);
    if (c.steps <= kAllSteps) {
        // two passes over all steps; s1 of every step goes to the adjoint
        // states' buffer, which only the backward writes
        float *s1_all = (float *)(ws + l.lam);
        uint16_t *w2f = (uint16_t *)(ws + l.w2f);
        const int nk = (c.n_hid + 31) / 32;
        // W2 in fragment order by extra blocks of layer 1's scan
        const unsigned scan_blocks = blocks((long)c.batch * c.n_hid);
        hipLaunchKernelGGL(scan1_fwd_all,
                           dim3(scan_blocks + blocks((long)nk * 512)),
                           dim3(256), 0, s, a, c.steps, c.checkpoint, slot,
                           (const float *)(ws + l.a1), state, s1_all,
                           c.mode == SG_SNN_SPIKE ? (uint16_t *)s1_all
                                                  : nullptr,
                           w2, w2f, nk, (int)scan_blocks);
        {
            // S1 . W2^T over the whole grid (MFMA), then layer 2's scan
            float *a2all = (float *)(ws + l.a2all);
            const dim3 og((unsigned)((MT + 15) / 16));
            if (c.mode == SG_SNN_SPIKE)
                hipLaunchKernelGGL(out_current_mfma<false>, og, dim3(256), 0,
                                   s, s1_all, w2f, a2all, (long)MT, c.n_hid,
                                   c.n_out, nk);
            else
                hipLaunchKernelGGL(out_current_mfma<true>, og, dim3(256), 0,
                                   s, s1_all, w2f, a2all, (long)MT, c.n_hid,
                                   c.n_out, nk);
            hipLaunchKernelGGL(c.mode == SG_SNN_SPIKE
                                   ? scan2_fwd_all<SG_SNN_SPIKE>
                                   : scan2_fwd_all<SG_SNN_RATE>,
                               dim3(blocks((long)c.batch * c.n_out)), dim3(256),
                               0, s, a, c.steps, c.checkpoint, slot, a2all,
                               state, cnt, (float *)(ws + l.rv2));
        }
    } else {
        for (int blk = 0; blk < c.steps / c.checkpoint; ++blk)
            forward_block(c, a, blk, l, ws, state + blk * slot,
                          state + (blk + 1) * slot, blk == 0 ? 1 : 2, false,
                          s);
    }
    hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), c.batch * 4, s, cnt,
                       labels, c.batch, c.n_out, loss, counts);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(std::string("launch: ") + hipGetErrorString(e));
    return 0;
}

int sg_snn_backward(const sg_snn_cfg *cfg, const uint16_t *x,
                    const uint16_t *w1, const uint16_t *w2,
                    const int32_t *labels, void *workspace, float grad_scale,
                    float *dw1, float *dw2, void *stream) {
    if (check(cfg)) return 1;
    if (!x || !w1 || !w2 || !labels || !workspace || !dw1 || !dw2)
        return fail("null pointer");
    const sg_snn_cfg &c = *cfg;
    hipStream_t s = (hipStream_t)stream;
    const Layout l = layout(c);
    unsigned char *ws = (unsigned char *)workspace;
    float *state = (float *)(ws + l.state);
    float *cnt = (float *)(ws + l.counts);
    float *g = (float *)(ws + l.grad);
    float *s1 = (float *)(ws + l.s1);
    float *rv1 = (float *)(ws + l.rv1), *rv2 = (float *)(ws + l.rv2);
    float *carry1 = (float *)(ws + l.carry1), *carry2 = (float *)(ws + l.carry2);
    float *lam = (float *)(ws + l.lam), *g2 = (float *)(ws + l.g2);
    float *part2 = (float *)(ws + l.part2), *part1 = (float *)(ws + l.part1);
    const int H = c.n_hid, O = c.n_out, B = c.batch;
    const int M = c.checkpoint * B;
    const long slot = slot_floats(B, H, O);
    const ScanArgs a = scan_args(c);
    const int chunks = (int)dw2_slots(c);
    const int nblk = c.steps / c.checkpoint;
    const int MT = c.steps * B;
    const long gp = ((MT + 7) / 8 + 3) / 4 * 4;
    uint16_t *xp = (uint16_t *)(ws + l.xp);
    if (fused_backward(c)) {
        // the mode and, for 10 outputs, their number fixed at compile time
        typedef void (*BK)(ScanArgs, int, long, const uint16_t *,
                           const float *, const float *, const float *,
                           const float *, const int32_t *, float, float *,
                           float *, const uint16_t *, uint16_t *, int, int,
                           long, int);
        const bool spike = c.mode == SG_SNN_SPIKE;
        const BK k = spike ? (O == 10 ? (BK)scan_bwd_fused<SG_SNN_SPIKE, 10>
                                      : (BK)scan_bwd_fused<SG_SNN_SPIKE, 0>)
                           : (O == 10 ? (BK)scan_bwd_fused<SG_SNN_RATE, 10>
                                      : (BK)scan_bwd_fused<SG_SNN_RATE, 0>);
        const int scan_blocks = (int)((long)B * H / 256);
        const long pack_blocks = blocks(gp * ((c.n_in + 7) / 8));
        hipLaunchKernelGGL(k, dim3((unsigned)(scan_blocks + pack_blocks)),
                           dim3(256), 0, s, a, nblk, slot, w2,
                           (const float *)(ws + l.a1), state, rv2, cnt, labels,
                           grad_scale, lam, part2, x, xp, MT, c.n_in, gp,
                           scan_blocks);
    } else {
        hipLaunchKernelGGL(loss_grad_kernel, dim3((B + 255) / 256), dim3(256),
                           0, s, cnt, labels, B, O, grad_scale, g);
        fill(part2, (long)chunks * O * H, 0.0f, s);
        fill(carry1, (long)B * H, 0.0f, s);
        fill(carry2, (long)B * O, 0.0f, s);
        for (int blk = nblk - 1; blk >= 0; --blk) {
            forward_block(c, a, blk, l, ws, state + blk * slot, nullptr, 0,
                          true, s);
            hipLaunchKernelGGL(scan2_bwd, dim3(blocks((long)B * O)),
                               dim3(256), 0, s, a, rv2 + (long)blk * M * O,
                               g, carry2, g2);
            hipLaunchKernelGGL(scan1_bwd, dim3(blocks((long)B * H)),
                               dim3(256), 0, s, a, w2, rv1, g2, carry1,
                               lam + (long)blk * M * H);
            hipLaunchKernelGGL(dw2_partial, dim3((H + 255) / 256, chunks),
                               dim3(256), 0, s, g2, s1, part2, M, H, O);
        }
    }
    // dW1 over every step's rows in one GEMM, split into kTnSplit row
    // ranges of whole 32-row slices (x regrouped by the fused scan's
    // launch, else here)
    if (!fused_backward(c))
        hipLaunchKernelGGL(pack_x, dim3(blocks(gp * ((c.n_in + 7) / 8))),
                           dim3(256), 0, s, x, xp, MT, c.n_in, gp);
    const int per = ((MT + kTnSplit - 1) / kTnSplit + kDr - 1) / kDr * kDr;
    const dim3 dgrid((c.n_in + kDk - 1) / kDk, (H + kDh - 1) / kDh, kTnSplit);
    if (H % 4 == 0) {
        // one more z slice of blocks sums dW2
        const dim3 g2(dgrid.x, dgrid.y, kTnSplit + 1);
        hipLaunchKernelGGL(gemm_dw1_dma, g2, dim3(256), 0, s, lam, xp, part1,
                           MT, H, c.n_in, per, gp, part2, dw2, chunks, O);
    } else {
        hipLaunchKernelGGL(gemm_dw1, dgrid, dim3(256), 0, s, lam, xp, part1,
                           MT, H, c.n_in, per);
        hipLaunchKernelGGL(dw2_reduce, dim3((O * H + 63) / 64), dim3(256), 0,
                           s, part2, dw2, chunks, H, O);
    }
    hipLaunchKernelGGL(add_parts, dim3(1024), dim3(256), 0, s, part1, dw1,
                       (long)H * c.n_in, kTnSplit);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(std::string("launch: ") + hipGetErrorString(e));
    return 0;
}

}  // extern "C"
