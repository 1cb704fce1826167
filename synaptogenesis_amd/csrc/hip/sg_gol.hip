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
