// Fused game-of-life step for a layer whose only input is one 3x3
// CONVERGENT convolutional connection from itself (the configuration of
// networks/game_of_life.json with field size 3): the base activator
// (synapse_kernel.h:370-461 + 887-912), the rule and the shift
// (game_of_life_attributes.cpp:14-33, attributes.h:321-333) in one pass.
//
// Words are double-buffered: the step reads words_in (state after t-1) and
// writes words_out.  Each block stages a 64 x 32 (small boards: 64 x 16)
// tile of words plus a one-cell halo in LDS (1.09 / 1.16 words read per
// cell), so each word is fetched from HBM/L2 about once per step instead of
// nine times; the neighbour sum runs in the kernel's row-major k order with the
// weights given, so any kernel values (not just 0/1) match the reference.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "sg_hip.h"

namespace sg {

constexpr int kGolTx = 64;   // tile columns (one wavefront per tile row)
constexpr int kGolWy = 4;    // wavefronts per block (tile rows per pass)

// Row or column index x >= -1 of a wrapping board (a tile and its halo
// reach at most a tile past the edge) wrapped into [0, n): one add or
// subtract, a division only for boards smaller than a tile.
__device__ __forceinline__ int gol_wrap(int x, int n) {
    if (x < 0) return x + n;
    if (x >= n) return x - n < n ? x - n : x % n;
    return x;
}

// INTW: every weight is an integer of magnitude at most 2^20, so the float
// sum of the 0/1-weighted terms is exact in any order and equals the
// integer sum; out-of-range neighbours of a non-wrapping board are staged
// as 0 and add 0.  Otherwise the sum is taken in float in the kernel's
// row-major order, skipping out-of-range neighbours as the reference does.
// Each thread owns kGolRun consecutive rows of one column and reads the
// newest bits of its (kGolRun + 2) x 3 neighbourhood once.
template <bool INTW, int kGolTy>
__global__ __launch_bounds__(kGolTx * kGolWy) void gol_step_kernel(
        const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
        float *__restrict__ input, int rows, int cols, int wrap, float k0,
        float k1, float k2, float k3, float k4, float k5, float k6, float k7,
        float k8, int smin, int smax, int bmin, int bmax) {
    constexpr int kGolRun = kGolTy / kGolWy;
    __shared__ uint32_t tile[kGolTy + 2][kGolTx + 2];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int col0 = blockIdx.x * kGolTx, row0 = blockIdx.y * kGolTy;

    // Stage (kGolTy + 2) x (kGolTx + 2) words; out-of-range cells of a
    // non-wrapping board hold 0.
    const int tid = ty * kGolTx + tx;
    const bool vec = (cols & 3) == 0 && col0 + kGolTx <= cols &&
                     ((uintptr_t)in & 15) == 0;
    auto stage = [&](int lr, int lc) {
        int r = row0 + lr - 1, c = col0 + lc - 1;
        uint32_t v = 0;
        if (wrap) {
            r = gol_wrap(r, rows);
            c = gol_wrap(c, cols);
            v = in[(long)r * cols + c];
        } else if (r >= 0 && r < rows && c >= 0 && c < cols) {
            v = in[(long)r * cols + c];
        }
        tile[lr][lc] = v;
    };
    if (vec) {
        // interior columns as 16-byte loads (the tile's 64 columns are in
        // range and 16-byte aligned), the two halo columns one word each
        constexpr int Q = kGolTx / 4;
        for (int i = tid; i < (kGolTy + 2) * Q; i += kGolTx * kGolWy) {
            const int lr = i / Q, q = i % Q;
            int r = row0 + lr - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (wrap) {
                r = gol_wrap(r, rows);
                v = *reinterpret_cast<const uint4 *>(in + (long)r * cols + col0 +
                                                     4 * q);
            } else if (r >= 0 && r < rows) {
                v = *reinterpret_cast<const uint4 *>(in + (long)r * cols + col0 +
                                                     4 * q);
            }
            tile[lr][1 + 4 * q] = v.x;
            tile[lr][2 + 4 * q] = v.y;
            tile[lr][3 + 4 * q] = v.z;
            tile[lr][4 + 4 * q] = v.w;
        }
        for (int i = tid; i < 2 * (kGolTy + 2); i += kGolTx * kGolWy)
            stage(i >> 1, (i & 1) ? kGolTx + 1 : 0);
    } else {
        for (int i = tid; i < (kGolTy + 2) * (kGolTx + 2);
             i += kGolTx * kGolWy)
            stage(i / (kGolTx + 2), i % (kGolTx + 2));
    }
    __syncthreads();

    const int c = col0 + tx;
    if (c >= cols) return;
    const int ly0 = ty * kGolRun;
    uint32_t bit[kGolRun + 2][3];
#pragma unroll
    for (int j = 0; j < kGolRun + 2; ++j)
#pragma unroll
        for (int kc = 0; kc < 3; ++kc) bit[j][kc] = tile[ly0 + j][tx + kc] >> 31;
    const float kw[9] = {k0, k1, k2, k3, k4, k5, k6, k7, k8};
    int ki[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) ki[q] = (int)kw[q];
#pragma unroll
    for (int j = 0; j < kGolRun; ++j) {
        const int ly = ly0 + j, r = row0 + ly;
        if (r >= rows) break;
        int n;
        float sum;
        if (INTW) {
            n = 0;
#pragma unroll
            for (int q = 0; q < 9; ++q)
                n += bit[j + q / 3][q % 3] ? ki[q] : 0;
            sum = (float)n;
        } else {
            sum = 0.0f;
#pragma unroll
            for (int kr = 0; kr < 3; ++kr) {
#pragma unroll
                for (int kc = 0; kc < 3; ++kc) {
                    if (!wrap) {
                        const int rr = r + kr - 1, cc = c + kc - 1;
                        if (rr < 0 || rr >= rows || cc < 0 || cc >= cols)
                            continue;
                    }
                    sum += (float)bit[j + kr][kc] * kw[kr * 3 + kc];
                }
            }
            n = (int)sum;
        }
        const uint32_t self = tile[ly + 1][tx + 1];
        const bool prev_live = self & 1u;
        const bool live = prev_live ? (n >= smin && n <= smax)
                                    : (n >= bmin && n <= bmax);
        const long idx = (long)r * cols + c;
        // streamed: the next step reads it from HBM anyway
        __builtin_nontemporal_store((self >> 1) | ((live ? 1u : 0u) << 31),
                                    &out[idx]);
        if (input) input[idx] = sum;
    }
}

// x within K <= 8 of [0, n) wrapped into it: one add or subtract, a
// division only for boards narrower than the halo.
__device__ __forceinline__ int gol_wrap_k(int x, int n) {
    if (x < 0) x += n;
    else if (x >= n) x -= n;
    if ((unsigned)x >= (unsigned)n) x = ((x % n) + n) % n;
    return x;
}

// K steps (2 <= K <= 4) in one pass, for boards whose words nothing else
// reads between those steps.  The block stages its tile plus a K-cell halo
// of words; step s = 1 .. K then computes the newest bit of every cell of
// the tile and of a (K - s)-cell ring around it from step s - 1's bits
// (ring cells are computed again by each block that needs them), with the
// same sum, order and rule as the one-step kernel: the survival test of
// step s reads bit s - 1 of the staged word (bit 0 after s - 1 shifts),
// the neighbours' newest bits are step s - 1's.  The tile's words are
// written once, shifted by K with the K new bits on top, and the input
// (neighbour sum) of step K only.  Cells outside a non-wrapping board are
// no cells: their bits stay 0 and the float sum skips them, as in the
// one-step kernel.
//
// Step S of K: each thread takes one column of the step's region and a
// run of R consecutive rows (NR runs per column, as many as one pass of
// the block's threads covers for every column), reading the (R + 2) x 3
// neighbourhood bits once into registers.
template <bool INTW, int TY, int K, int S>
__device__ __forceinline__ void gol_multi_step(
        const uint32_t (&w0)[TY + 2 * K][kGolTx + 2 * K],
        uint8_t (&bits)[K + 1][TY + 2 * K][kGolTx + 2 * K],
        uint32_t *__restrict__ out, float *__restrict__ input, int rows,
        int cols, int wrap, int row0, int col0, const float (&kw)[9],
        const int (&ki)[9], int smin, int smax, int bmin, int bmax) {
    constexpr int RY = TY + 2 * K;
    constexpr int m = K - S;                   // ring of this step
    constexpr int ny = TY + 2 * m, nx = kGolTx + 2 * m;
    constexpr int NR = kGolTx * kGolWy / nx;   // row runs per column
    constexpr int R = (ny + NR - 1) / NR;
    const int t = threadIdx.y * kGolTx + threadIdx.x;
    if (t < nx * NR) {
        const int cx = t % nx;
        const int lr0 = S + (t / nx) * R;      // this thread's first row
        const int lc = S + cx;
        const int c = col0 + lc;
        uint32_t win[R + 2][3];
#pragma unroll
        for (int j = 0; j < R + 2; ++j) {
            const int lr = lr0 - 1 + j;
#pragma unroll
            for (int kc = 0; kc < 3; ++kc)
                win[j][kc] = lr < RY ? bits[S - 1][lr][lc - 1 + kc] : 0u;
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int lr = lr0 + j;
            if (lr >= RY - S) break;           // past the region
            const int r = row0 + lr;
            if (!wrap && (r < 0 || r >= rows || c < 0 || c >= cols)) {
                bits[S][lr][lc] = 0;
                continue;
            }
            int n;
            float sum;
            if (INTW) {
                n = 0;
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    n += win[j + q / 3][q % 3] ? ki[q] : 0;
                sum = (float)n;
            } else {
                sum = 0.0f;
#pragma unroll
                for (int kr = 0; kr < 3; ++kr) {
#pragma unroll
                    for (int kc = 0; kc < 3; ++kc) {
                        if (!wrap) {
                            const int rr = r + kr - 1, cc = c + kc - 1;
                            if (rr < 0 || rr >= rows || cc < 0 || cc >= cols)
                                continue;
                        }
                        sum += (float)win[j + kr][kc] * kw[kr * 3 + kc];
                    }
                }
                n = (int)sum;
            }
            const uint32_t w = w0[lr][lc];
            const bool prev_live = (w >> (S - 1)) & 1u;
            const bool live = prev_live ? (n >= smin && n <= smax)
                                        : (n >= bmin && n <= bmax);
            bits[S][lr][lc] = live ? 1 : 0;
            if (S == K && r < rows && c < cols) {
                // m == 0: a tile cell (r, c >= 0 there)
                uint32_t word = w >> K;
#pragma unroll
                for (int q = 1; q < K; ++q)
                    word |= (uint32_t)bits[q][lr][lc] << (31 - K + q);
                word |= (live ? 1u : 0u) << 31;
                const long idx = (long)r * cols + c;
                __builtin_nontemporal_store(word, &out[idx]);
                if (input) input[idx] = sum;
            }
        }
    }
}

template <bool INTW, int TY, int K>
__global__ __launch_bounds__(kGolTx * kGolWy) void gol_multi_kernel(
        const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
        float *__restrict__ input, int rows, int cols, int wrap, float k0,
        float k1, float k2, float k3, float k4, float k5, float k6, float k7,
        float k8, int smin, int smax, int bmin, int bmax) {
    constexpr int RY = TY + 2 * K, RX = kGolTx + 2 * K;
    __shared__ uint32_t w0[RY][RX];
    __shared__ uint8_t bits[K + 1][RY][RX];   // newest bit after step s
    const int tid = threadIdx.y * kGolTx + threadIdx.x;
    // board coordinates of the region's first row and column
    const int row0 = blockIdx.y * TY - K, col0 = blockIdx.x * kGolTx - K;
    for (int i = tid; i < RY * RX; i += kGolTx * kGolWy) {
        const int lr = i / RX, lc = i - lr * RX;
        int r = row0 + lr, c = col0 + lc;
        uint32_t v = 0;
        if (wrap) {
            r = gol_wrap_k(r, rows);
            c = gol_wrap_k(c, cols);
            v = in[(long)r * cols + c];
        } else if (r >= 0 && r < rows && c >= 0 && c < cols) {
            v = in[(long)r * cols + c];
        }
        w0[lr][lc] = v;
        bits[0][lr][lc] = (uint8_t)(v >> 31);
    }
    __syncthreads();
    const float kw[9] = {k0, k1, k2, k3, k4, k5, k6, k7, k8};
    int ki[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) ki[q] = (int)kw[q];
    gol_multi_step<INTW, TY, K, 1>(w0, bits, out, input, rows, cols, wrap,
                                   row0, col0, kw, ki, smin, smax, bmin, bmax);
    __syncthreads();
    gol_multi_step<INTW, TY, K, 2>(w0, bits, out, input, rows, cols, wrap,
                                   row0, col0, kw, ki, smin, smax, bmin, bmax);
    if constexpr (K >= 3) {
        __syncthreads();
        gol_multi_step<INTW, TY, K, 3>(w0, bits, out, input, rows, cols,
                                       wrap, row0, col0, kw, ki, smin, smax,
                                       bmin, bmax);
    }
    if constexpr (K >= 4) {
        __syncthreads();
        gol_multi_step<INTW, TY, K, 4>(w0, bits, out, input, rows, cols,
                                       wrap, row0, col0, kw, ki, smin, smax,
                                       bmin, bmax);
    }
}

// v_writelane_b32: old with lane `lane` set to the wavefront-uniform value
// (the compiler's intrinsic, which this HIP has no builtin name for).
__device__ int gol_writelane(int value, int lane, int old)
    __asm("llvm.amdgcn.writelane.i32");

// Lane i gets lane i - 1's value (wave_shr:1) / lane i + 1's (wave_shl:1);
// the lane with no source gets 0.
__device__ __forceinline__ uint64_t gol_from_prev_lane(uint64_t v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x138,
                                               0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32),
                                               0x138, 0xf, 0xf, true);
    return (uint64_t)(uint32_t)hi << 32 | (uint32_t)lo;
}
__device__ __forceinline__ uint64_t gol_from_next_lane(uint64_t v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x130,
                                               0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32),
                                               0x130, 0xf, 0xf, true);
    return (uint64_t)(uint32_t)hi << 32 | (uint32_t)lo;
}

// Kernels whose nine weights are each 0 or 1 (the game of life's own:
// ones around a cleared centre): K steps (1 <= K <= 8) on rows of 64 cells
// packed into 64-bit masks.  A block owns a 64 x 64 region of cells (rows
// row0 .., columns col0 ..) and writes its inner (64 - 2K) x (64 - 2K)
// tile: each step's new bits are exact one cell further in from the
// region's edge than the previous step's.  Its W wavefronts stage 64 / W
// region rows each, lane l = column col0 + l (a region inside the board
// loads from a row base with no per-lane wrap or bounds arithmetic), and
// pack each row's newest bits, the K survival bits (bit s - 1 of the word
// for step s) and, on a bounded board, the cells inside it, by ballot; lane
// j of the wavefront collects row j's masks.  One wavefront then takes all
// K steps, lane i = region row i, with the rows above and below from its
// neighbour lanes.  LIFE (neighbours = the eight around the cell, survival
// on 2-3, birth on 3): each row's horizontal sums by full adders, their sum
// as a ones mask and four twos masks; the count is 2 or 3 where exactly one
// twos mask is set, and the cell lives there if the ones mask or its
// survival bit is set.  Otherwise the count of 64 cells at once as four bit
// planes (a ripple add of each weighted neighbour mask) and the rule as the
// union of the count's equality masks for the values the survival and
// birth ranges admit.  The count equals the integer sum of the one-weighted
// bits, which is the float sum of the one-step kernel exactly; cells
// outside a bounded board hold 0 at every step.  Finally each wavefront
// writes its rows' inner words: the staged word shifted by K with the K new
// bits on top (each lane reads the half of each step's row mask that holds
// its column), and the last step's count as the input.
template <int K, bool LIFE, int W>
__global__ __launch_bounds__(64 * W) void gol_packed_kernel(
        const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
        float *__restrict__ input, int rows, int cols, int wrap,
        uint32_t wmask, uint32_t surv, uint32_t birth) {
    constexpr int TI = 64 - 2 * K;           // inner tile
    constexpr int NP = LIFE ? 5 : 4;         // count masks of the last step
    constexpr int NM = K + 2;                // masks staged per row
    constexpr int RW = 64 / W;               // region rows per wavefront
    __shared__ uint64_t staged[NM][64];      // newest, inside, survival
    __shared__ uint32_t halves[64][2][K];    // step masks, [row][half][step]
    __shared__ uint64_t planes[NP][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row0 = blockIdx.y * TI - K, col0 = blockIdx.x * TI - K;
    const int c = col0 + lane;
    const bool interior = row0 >= 0 && row0 + 64 <= rows && col0 >= 0 &&
                          col0 + 64 <= cols;
    uint32_t words[RW];
    uint32_t mlo[NM], mhi[NM];               // lane j: row j's masks
#pragma unroll
    for (int m = 0; m < NM; ++m) mlo[m] = mhi[m] = 0u;
    auto collect = [&](int j, int m, uint64_t v) {
        mlo[m] = (uint32_t)gol_writelane((int)(uint32_t)v, j, (int)mlo[m]);
        mhi[m] = (uint32_t)gol_writelane((int)(uint32_t)(v >> 32), j,
                                         (int)mhi[m]);
    };
    if (interior) {
        mlo[1] = mhi[1] = ~0u;               // every cell inside the board
        const uint32_t *base = in + (long)(row0 + wave * RW) * cols + col0;
#pragma unroll
        for (int j = 0; j < RW; ++j) words[j] = base[(long)j * cols + lane];
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            const uint32_t v = words[j];
            collect(j, 0, __ballot((int)v < 0));
#pragma unroll
            for (int s = 0; s < K; ++s)
                collect(j, 2 + s, __ballot((v >> s) & 1u));
        }
    } else {
        // addresses first (cells outside a bounded board read cell 0 and
        // are cleared), then all 16 loads at once
        const int cw = wrap ? gol_wrap_k(c, cols) : c;
        const bool col_ok = wrap || (c >= 0 && c < cols);
        bool okv[RW];
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            const int r = row0 + wave * RW + j;
            const int rw = wrap ? gol_wrap_k(r, rows) : r;
            okv[j] = col_ok && (wrap || (r >= 0 && r < rows));
            words[j] = in[okv[j] ? (long)rw * cols + cw : 0];
        }
#pragma unroll
        for (int j = 0; j < RW; ++j) {
            const uint32_t v = okv[j] ? words[j] : 0u;
            words[j] = v;
            collect(j, 0, __ballot((int)v < 0));
            collect(j, 1, __ballot(okv[j]));
#pragma unroll
            for (int s = 0; s < K; ++s)
                collect(j, 2 + s, __ballot((v >> s) & 1u));
        }
    }
    if (lane < RW) {
#pragma unroll
        for (int m = 0; m < NM; ++m)
            staged[m][wave * RW + lane] = (uint64_t)mhi[m] << 32 | mlo[m];
    }
    __syncthreads();
    if (wave == 0) {
        uint64_t b = staged[0][lane];
        const uint64_t ib = staged[1][lane];
#pragma unroll
        for (int s = 0; s < K; ++s) {
            // rows above and below by whole-wavefront DPP shifts (every
            // lane active); the edge rows get 0, which only touches the
            // rows no output depends on
            const uint64_t a = gol_from_prev_lane(b);
            const uint64_t d = gol_from_next_lane(b);
            const uint64_t p = staged[2 + s][lane];
            if (LIFE) {
                const uint64_t al = a << 1, ar = a >> 1, bl = b << 1,
                               br = b >> 1, dl = d << 1, dr = d >> 1;
                uint64_t t = al ^ a;
                const uint64_t sa = t ^ ar, ca = (al & a) | (ar & t);
                const uint64_t sb = bl ^ br, cb = bl & br;
                t = dl ^ d;
                const uint64_t sd = t ^ dr, cd = (dl & d) | (dr & t);
                t = sa ^ sb;
                const uint64_t one = t ^ sd, c1 = (sa & sb) | (sd & t);
                // count = one + 2 (ca + cb + cd + c1): 2 or 3 where exactly
                // one of the four twos masks is set
                const uint64_t p1 = ca ^ cb, q1 = cd ^ c1;
                const uint64_t t1 =
                    (p1 ^ q1) & ~((ca & cb) | (cd & c1) | (p1 & q1));
                b = t1 & (one | p) & ib;
                if (s == K - 1 && input) {
                    planes[0][lane] = one;
                    planes[1][lane] = ca;
                    planes[2][lane] = cb;
                    planes[3][lane] = cd;
                    planes[NP - 1][lane] = c1;
                }
            } else {
                const uint64_t nbr[9] = {a << 1, a, a >> 1, b << 1, b,
                                         b >> 1, d << 1, d, d >> 1};
                uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    if (!((wmask >> q) & 1u)) continue;
                    const uint64_t x = nbr[q];
                    const uint64_t k0 = c0 & x;
                    c0 ^= x;
                    const uint64_t k1 = c1 & k0;
                    c1 ^= k0;
                    const uint64_t k2 = c2 & k1;
                    c2 ^= k1;
                    c3 |= k2;
                }
                uint64_t sv = 0, bv = 0;
#pragma unroll
                for (int v = 0; v <= 9; ++v) {
                    if (!(((surv | birth) >> v) & 1u)) continue;
                    const uint64_t eq = ((v & 1) ? c0 : ~c0) &
                                        ((v & 2) ? c1 : ~c1) &
                                        ((v & 4) ? c2 : ~c2) &
                                        ((v & 8) ? c3 : ~c3);
                    if ((surv >> v) & 1u) sv |= eq;
                    if ((birth >> v) & 1u) bv |= eq;
                }
                b = ((p & sv) | (~p & bv)) & ib;
                if (s == K - 1 && input) {
                    planes[0][lane] = c0;
                    planes[1][lane] = c1;
                    planes[2][lane] = c2;
                    planes[3][lane] = c3;
                }
            }
            halves[lane][0][s] = (uint32_t)b;
            halves[lane][1][s] = (uint32_t)(b >> 32);
        }
    }
    __syncthreads();
    if (lane < K || lane >= 64 - K) return;
    if (c >= cols) return;
    const int half = lane >> 5, bit = lane & 31;
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int i = wave * RW + j;
        const int r = row0 + i;
        if (i < K || i >= 64 - K || r >= rows) continue;
        uint32_t word = words[j] >> K;
        const uint32_t *h = halves[i][half];
#pragma unroll
        for (int s = 0; s < K; ++s)
            word |= __builtin_amdgcn_ubfe(h[s], bit, 1) << (32 - K + s);
        const long idx = (long)r * cols + c;
        __builtin_nontemporal_store(word, &out[idx]);
        if (input) {
            int n;
            if (LIFE)
                n = (int)((planes[0][i] >> lane) & 1ull) +
                    2 * ((int)((planes[1][i] >> lane) & 1ull) +
                         (int)((planes[2][i] >> lane) & 1ull) +
                         (int)((planes[3][i] >> lane) & 1ull) +
                         (int)((planes[NP - 1][i] >> lane) & 1ull));
            else
                n = (int)((planes[0][i] >> lane) & 1ull) |
                    (int)((planes[1][i] >> lane) & 1ull) << 1 |
                    (int)((planes[2][i] >> lane) & 1ull) << 2 |
                    (int)((planes[3][i] >> lane) & 1ull) << 3;
            input[idx] = (float)n;
        }
    }
}


// An 8 x 8 bit matrix (byte i = row i, bit j = column j) transposed by
// three swap stages: byte j of the result holds column j (bit i = row i).
__device__ __forceinline__ uint64_t gol_transpose8(uint64_t x) {
    uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
    x = x ^ t ^ (t << 7);
    t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
    x = x ^ t ^ (t << 14);
    t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
    return x ^ t ^ (t << 28);
}

// The game of life's own rule (0/1 kernel of the eight neighbours,
// survival on 2-3, birth on 3), K steps (1 <= K <= 8), one wavefront per
// 64 x 64 region with lane = column col0 + lane: each lane loads its
// column's 64 words and builds its own 64-row masks (bit i = region row i)
// of the newest bits, the K survival bits and the cells inside the board;
// the rows above and below are bit shifts of a mask, the columns beside
// it the neighbour lanes' masks by DPP; the rule is the full-adder form of
// gol_packed_kernel with the columns in the rows' place (the neighbourhood
// is symmetric).  Masks and words stay in the lane that loaded them: no
// LDS, no barrier.  The inner (64 - 2K) x (64 - 2K) tile is written as in
// gol_packed_kernel.
template <int K>
__global__ __launch_bounds__(64) void gol_column_kernel(
        const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
        float *__restrict__ input, int rows, int cols, int wrap) {
    constexpr int TI = 64 - 2 * K;
    const int lane = threadIdx.x;
    const int row0 = blockIdx.y * TI - K, col0 = blockIdx.x * TI - K;
    const int c = col0 + lane;
    const bool interior = row0 >= 0 && row0 + 64 <= rows && col0 >= 0 &&
                          col0 + 64 <= cols;
    uint32_t words[64];
    uint64_t ib = ~0ull;
    if (interior) {
        const uint32_t *base = in + (long)row0 * cols + c;
#pragma unroll
        for (int r = 0; r < 64; ++r) words[r] = base[(long)r * cols];
    } else {
        const int cw = wrap ? gol_wrap_k(c, cols) : c;
        const bool col_ok = wrap || (c >= 0 && c < cols);
        ib = 0ull;
#pragma unroll
        for (int r = 0; r < 64; ++r) {
            const int rr = row0 + r;
            const int rw = wrap ? gol_wrap_k(rr, rows) : rr;
            const bool ok = col_ok && (wrap || (rr >= 0 && rr < rows));
            words[r] = ok ? in[(long)rw * cols + cw] : 0u;
            ib |= (uint64_t)ok << r;
        }
    }
    uint32_t nlo = 0u, nhi = 0u, klo[K], khi[K];
#pragma unroll
    for (int s = 0; s < K; ++s) klo[s] = khi[s] = 0u;
#pragma unroll
    for (int r = 0; r < 32; ++r) {
        nlo |= (words[r] >> 31) << r;
        nhi |= (words[r + 32] >> 31) << r;
    }
    // the survival masks eight rows at a time: the rows' low bytes as the
    // rows of an 8 x 8 bit matrix, transposed (three swap stages), so that
    // byte s holds bit s of the eight rows
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t *w = words + 8 * j;
        const uint32_t xl = (w[0] & 0xFFu) | (w[1] & 0xFFu) << 8 |
                            (w[2] & 0xFFu) << 16 | w[3] << 24;
        const uint32_t xh = (w[4] & 0xFFu) | (w[5] & 0xFFu) << 8 |
                            (w[6] & 0xFFu) << 16 | w[7] << 24;
        const uint64_t x = gol_transpose8((uint64_t)xh << 32 | xl);
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const uint32_t byte = (uint32_t)(x >> (8 * s)) & 0xFFu;
            if (j < 4) klo[s] |= byte << (8 * j);
            else khi[s] |= byte << (8 * (j - 4));
        }
    }
    uint64_t b = (uint64_t)nhi << 32 | nlo;
    uint64_t step[K];
    uint64_t one = 0, ca = 0, cb = 0, cd = 0, c1 = 0;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const uint64_t L = gol_from_prev_lane(b);   // the column left
        const uint64_t R = gol_from_next_lane(b);   // the column right
        const uint64_t p = (uint64_t)khi[s] << 32 | klo[s];
        const uint64_t lu = L << 1, ld = L >> 1, bu = b << 1, bd = b >> 1,
                       ru = R << 1, rd = R >> 1;
        uint64_t t = lu ^ L;
        const uint64_t sa = t ^ ld;
        ca = (lu & L) | (ld & t);
        const uint64_t sb = bu ^ bd;
        cb = bu & bd;
        t = ru ^ R;
        const uint64_t sd = t ^ rd;
        cd = (ru & R) | (rd & t);
        t = sa ^ sb;
        one = t ^ sd;
        c1 = (sa & sb) | (sd & t);
        const uint64_t p1 = ca ^ cb, q1 = cd ^ c1;
        const uint64_t t1 = (p1 ^ q1) & ~((ca & cb) | (cd & c1) | (p1 & q1));
        b = t1 & (one | p) & ib;
        step[s] = b;
    }
    if (lane < K || lane >= 64 - K || c >= cols) return;
    // each row's K new bits as one byte: per eight rows, byte s = step s's
    // bits of the rows, transposed so that byte i = row i's bits
    uint64_t rb[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        rb[j] = 0ull;
        if (8 * j + 7 < K || 8 * j >= 64 - K) continue;   // no output row
        uint64_t x = 0ull;
#pragma unroll
        for (int s = 0; s < K; ++s) x |= ((step[s] >> (8 * j)) & 0xFFull) << (8 * s);
        rb[j] = gol_transpose8(x);
    }
#pragma unroll
    for (int r = K; r < 64 - K; ++r) {
        if (row0 + r >= rows) break;
        const uint32_t bits =
            (uint32_t)(rb[r >> 3] >> (8 * (r & 7))) & ((1u << K) - 1u);
        const uint32_t word = (words[r] >> K) | bits << (32 - K);
        const long idx = (long)(row0 + r) * cols + c;
        __builtin_nontemporal_store(word, &out[idx]);
        if (input)
            input[idx] = (float)((int)((one >> r) & 1ull) +
                                 2 * ((int)((ca >> r) & 1ull) +
                                      (int)((cb >> r) & 1ull) +
                                      (int)((cd >> r) & 1ull) +
                                      (int)((c1 >> r) & 1ull)));
    }
}

}  // namespace sg

extern "C" int sg_hip_gol_step(const uint32_t *words_in, uint32_t *words_out,
                               float *input, int rows, int cols, int wrap,
                               const float *k, int smin, int smax, int bmin,
                               int bmax, void *stream) {
    using namespace sg;
    if (rows <= 0 || cols <= 0) return 0;
    // 32-row tiles where the board gives at least 2048 of them (less halo
    // per cell), else 16-row tiles (more blocks for small boards)
    const long tiles32 = (long)((cols + kGolTx - 1) / kGolTx) *
                         ((rows + 31) / 32);
    const int ty = tiles32 >= 2048 ? 32 : 16;
    dim3 g((cols + kGolTx - 1) / kGolTx, (rows + ty - 1) / ty);
    bool intw = true;
    for (int q = 0; q < 9; ++q)
        intw = intw && std::fabs(k[q]) <= 1048576.0f &&
               k[q] == (float)(int)k[q];
    auto *kern = ty == 32
        ? (intw ? gol_step_kernel<true, 32> : gol_step_kernel<false, 32>)
        : (intw ? gol_step_kernel<true, 16> : gol_step_kernel<false, 16>);
    hipLaunchKernelGGL(kern, g, dim3(kGolTx, kGolWy), 0, (hipStream_t)stream,
                       words_in, words_out, input, rows, cols, wrap, k[0],
                       k[1], k[2], k[3], k[4], k[5], k[6], k[7], k[8], smin,
                       smax, bmin, bmax);
    return hipGetLastError();
}

extern "C" int sg_hip_gol_max_steps(const float *k, int rows, int cols) {
    (void)rows;
    (void)cols;
    for (int q = 0; q < 9; ++q)
        if (k[q] != 0.0f && k[q] != 1.0f) return 4;
    // the packed form: its instructions, not the board's traffic, bound a
    // launch, and 8 steps a launch (48 x 48 tiles) measured fastest on
    // boards of 1024 and 16384 (4 steps: 1.5x the time a step on both; 10
    // to 16 steps: no faster on 1024, slower on 16384)
    return 8;
}

extern "C" int sg_hip_gol_steps(const uint32_t *words_in, uint32_t *words_out,
                                float *input, int rows, int cols, int wrap,
                                const float *k, int smin, int smax, int bmin,
                                int bmax, int steps, void *stream) {
    using namespace sg;
    if (steps < 1 || steps > sg_hip_gol_max_steps(k, rows, cols))
        return (int)hipErrorInvalidValue;
    if (rows <= 0 || cols <= 0) return 0;
    // weights of 0 and 1 only: the packed form
    uint32_t wmask = 0;
    bool bitw = true;
    for (int q = 0; q < 9; ++q) {
        if (k[q] == 1.0f) wmask |= 1u << q;
        else if (k[q] != 0.0f) bitw = false;   // -0.0f counts as 0
    }
    if (bitw) {
        // rule ranges as flags over the counts 0 .. 9
        uint32_t surv = 0, birth = 0;
        for (int v = 0; v <= 9; ++v) {
            if (v >= smin && v <= smax) surv |= 1u << v;
            if (v >= bmin && v <= bmax) birth |= 1u << v;
        }
        const int ti = 64 - 2 * steps;
        const dim3 g((cols + ti - 1) / ti, (rows + ti - 1) / ti);
        typedef void (*PK)(const uint32_t *, uint32_t *, float *, int, int,
                           int, uint32_t, uint32_t, uint32_t);
        // the game of life's own rule: eight neighbours, survival on 2-3,
        // birth on 3
        const bool life = wmask == 0x1efu && surv == 0xcu && birth == 0x8u;
        // eight wavefronts a block: 8 region rows each to stage and write
        // (7.6 against 9.5 us a launch of 8 steps on a 1024 board with four,
        // the same on 16384)
        constexpr int W = 8;
        static const PK life_k[8] = {
            gol_packed_kernel<1, true, W>, gol_packed_kernel<2, true, W>,
            gol_packed_kernel<3, true, W>, gol_packed_kernel<4, true, W>,
            gol_packed_kernel<5, true, W>, gol_packed_kernel<6, true, W>,
            gol_packed_kernel<7, true, W>, gol_packed_kernel<8, true, W>};
        static const PK any_k[8] = {
            gol_packed_kernel<1, false, W>, gol_packed_kernel<2, false, W>,
            gol_packed_kernel<3, false, W>, gol_packed_kernel<4, false, W>,
            gol_packed_kernel<5, false, W>, gol_packed_kernel<6, false, W>,
            gol_packed_kernel<7, false, W>, gol_packed_kernel<8, false, W>};
        if (life && (long)g.x * g.y >= 1024) {
            // one wavefront per region, lane = column (gol_column_kernel):
            // fewer instructions per region but one wavefront's chain per
            // region, which bounds small boards.  8 steps a launch, row
            // form against column form: 484 regions 7.8 / 10.7 us, 1024
            // regions 10.9 / 11.0, 1849 16.7 / 16.0, 7396 49.0 / 41.7, the
            // 16384 board 761 / 449
            typedef void (*CK)(const uint32_t *, uint32_t *, float *, int,
                               int, int);
            static const CK col_k[8] = {
                gol_column_kernel<1>, gol_column_kernel<2>,
                gol_column_kernel<3>, gol_column_kernel<4>,
                gol_column_kernel<5>, gol_column_kernel<6>,
                gol_column_kernel<7>, gol_column_kernel<8>};
            hipLaunchKernelGGL(col_k[steps - 1], g, dim3(64), 0,
                               (hipStream_t)stream, words_in, words_out,
                               input, rows, cols, wrap);
            return hipGetLastError();
        }
        const PK kern = (life ? life_k : any_k)[steps - 1];
        hipLaunchKernelGGL(kern, g, dim3(64 * W), 0, (hipStream_t)stream,
                           words_in, words_out, input, rows, cols, wrap,
                           wmask, surv, birth);
        return hipGetLastError();
    }
    if (steps == 1)
        return sg_hip_gol_step(words_in, words_out, input, rows, cols, wrap,
                               k, smin, smax, bmin, bmax, stream);
    const long tiles32 = (long)((cols + kGolTx - 1) / kGolTx) *
                         ((rows + 31) / 32);
    const int ty = tiles32 >= 2048 ? 32 : 16;
    dim3 g((cols + kGolTx - 1) / kGolTx, (rows + ty - 1) / ty);
    bool intw = true;
    for (int q = 0; q < 9; ++q)
        intw = intw && std::fabs(k[q]) <= 1048576.0f &&
               k[q] == (float)(int)k[q];
    typedef void (*Kern)(const uint32_t *, uint32_t *, float *, int, int, int,
                         float, float, float, float, float, float, float,
                         float, float, int, int, int, int);
#define SG_GOL_K(KS)                                                         \
    (ty == 32 ? (intw ? (Kern)gol_multi_kernel<true, 32, KS>                 \
                      : (Kern)gol_multi_kernel<false, 32, KS>)               \
              : (intw ? (Kern)gol_multi_kernel<true, 16, KS>                 \
                      : (Kern)gol_multi_kernel<false, 16, KS>))
    const Kern kern = steps == 2 ? SG_GOL_K(2)
                    : steps == 3 ? SG_GOL_K(3) : SG_GOL_K(4);
#undef SG_GOL_K
    hipLaunchKernelGGL(kern, g, dim3(kGolTx, kGolWy), 0, (hipStream_t)stream,
                       words_in, words_out, input, rows, cols, wrap, k[0],
                       k[1], k[2], k[3], k[4], k[5], k[6], k[7], k[8], smin,
                       smax, bmin, bmax);
    return hipGetLastError();
}
