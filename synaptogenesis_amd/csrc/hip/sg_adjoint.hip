// Forward pass and adjoint of the two-layer leaky network of
// include/sg_adjoint.h.
//
// Per block of `checkpoint` steps (M = checkpoint * batch rows, row = step
// * batch + b):
//   forward   gemm_nt      A1[M][n_hid] = X[M][n_in] . W1^T   (MFMA BF16)
//             scan_fwd     per batch row: both layers' recurrences over
//                          the block's steps, W2 . s1 reduced in LDS
//   backward  gemm_nt + scan_fwd (recomputing the block from its
//                          checkpoint, recording v and s of every step)
//             scan_bwd     per batch row, steps in reverse: the adjoint
//                          states dL/dv1, dL/dv2
//             gemm_tn      dW1[n_hid][n_in] += Lambda^T . X   (MFMA BF16,
//                          Lambda as hi + mid + lo BF16 terms)
//             dw2          dW2[n_out][n_hid] += G2^T . S1
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

constexpr int kMaxHid = 1024;   // hidden units per batch row (4 per thread)
constexpr int kMaxOut = 16;
constexpr int kScanThreads = 256;
constexpr int kUnits = kMaxHid / kScanThreads;

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

// C[M][N] = A[M][K] . B[N][K]^T, BF16 operands, FP32 result.  64 x 64
// tile per block, four wavefronts of 32 x 32 (2 x 2 MFMA 16x16x32 tiles).
__global__ __launch_bounds__(256) void gemm_nt(const uint16_t *__restrict__ A,
                                               const uint16_t *__restrict__ B,
                                               float *__restrict__ C, int M,
                                               int N, int K) {
    __shared__ __align__(16) uint16_t As[64][40];
    __shared__ __align__(16) uint16_t Bs[64][40];
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int t = threadIdx.x, lr = t >> 2, lc = (t & 3) * 8;
    const int w = t >> 6, lane = t & 63;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    f32x4 acc[2][2] = {};
    for (int k0 = 0; k0 < K; k0 += 32) {
        *(u32x4 *)&As[lr][lc] = load8(A, M, K, m0 + lr, k0 + lc);
        *(u32x4 *)&Bs[lr][lc] = load8(B, N, K, n0 + lr, k0 + lc);
        __syncthreads();
        bf16x8 a[2], b[2];
        for (int i = 0; i < 2; ++i)
            a[i] = *(const bf16x8 *)&As[wm + i * 16 + li][lk];
        for (int j = 0; j < 2; ++j)
            b[j] = *(const bf16x8 *)&Bs[wn + j * 16 + li][lk];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + j * 16 + li;
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
                if (row < M && col < N) C[(long)row * N + col] = acc[i][j][r];
            }
        }
}

// D[N][K] += L[M][N]^T . X[M][K]: FP32 L as three BF16 terms, BF16 X.
// 64 (n) x 64 (k) tile per block; the M rows are the reduction.
__global__ __launch_bounds__(256) void gemm_tn(const float *__restrict__ L,
                                               const uint16_t *__restrict__ X,
                                               float *__restrict__ D, int M,
                                               int N, int K) {
    __shared__ __align__(16) uint16_t Ah[3][64][40];
    __shared__ __align__(16) uint16_t Bs[64][40];
    const int n0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
    const int t = threadIdx.x, lr = t >> 3, lc = (t & 7) * 8;
    const int w = t >> 6, lane = t & 63;
    const int wn = (w >> 1) * 32, wk = (w & 1) * 32;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    f32x4 acc[2][2] = {};
    for (int r0 = 0; r0 < M; r0 += 32) {
        const int r = r0 + lr;
        for (int e = 0; e < 8; ++e) {
            const int n = n0 + lc + e;
            const float v = (r < M && n < N) ? L[(long)r * N + n] : 0.0f;
            const uint32_t hi = bf16_rne(v);
            const float rem = v - bf16_value(hi);
            const uint32_t mid = bf16_rne(rem);
            const uint32_t lo = bf16_rne(rem - bf16_value(mid));
            Ah[0][lc + e][lr] = (uint16_t)hi;
            Ah[1][lc + e][lr] = (uint16_t)mid;
            Ah[2][lc + e][lr] = (uint16_t)lo;
        }
        const u32x4 xv = load8(X, M, K, r, k0 + lc);
        uint16_t xe[8];
        memcpy(xe, &xv, 16);
        for (int e = 0; e < 8; ++e) Bs[lc + e][lr] = xe[e];
        __syncthreads();
        bf16x8 b[2];
        for (int j = 0; j < 2; ++j)
            b[j] = *(const bf16x8 *)&Bs[wk + j * 16 + li][lk];
        for (int s = 0; s < 3; ++s)
            for (int i = 0; i < 2; ++i) {
                const bf16x8 a = *(const bf16x8 *)&Ah[s][wn + i * 16 + li][lk];
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        a, b[j], acc[i][j], 0, 0, 0);
            }
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int col = k0 + wk + j * 16 + li;
            for (int q = 0; q < 4; ++q) {
                const int row = n0 + wn + i * 16 + (lane >> 4) * 4 + q;
                if (row < N && col < K) D[(long)row * K + col] += acc[i][j][q];
            }
        }
}

struct ScanArgs {
    int B, H, O, steps;          // batch, hidden, outputs, steps in block
    int mode;
    float beta, theta, slope;
    float reset_theta;           // reset * theta
};

// Recurrences of one block of steps for batch row blockIdx.x.  state_in /
// state_out: [B][2H + 2O] (v1, s1, v2, s2); count_mode 0: leave counts,
// 1: start them at 0, 2: add to them.  rec_*: per-step records or null.
__global__ __launch_bounds__(kScanThreads) void scan_fwd(
    ScanArgs a, const float *__restrict__ a1, const uint16_t *__restrict__ w2,
    const float *__restrict__ state_in, float *__restrict__ state_out,
    float *__restrict__ counts, int count_mode, float *__restrict__ rec_v1,
    float *__restrict__ rec_s1, float *__restrict__ rec_v2,
    float *__restrict__ rec_s2) {
    extern __shared__ float w2s[];           // [O][H]
    __shared__ float red[kScanThreads / 64][kMaxOut];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int H = a.H, O = a.O;
    for (int i = tid; i < O * H; i += kScanThreads) w2s[i] = bf16_value(w2[i]);
    const float *st = state_in + (long)b * (2 * H + 2 * O);
    float v1[kUnits], s1[kUnits];
    for (int q = 0; q < kUnits; ++q) {
        const int u = tid + q * kScanThreads;
        v1[q] = u < H ? st[u] : 0.0f;
        s1[q] = u < H ? st[H + u] : 0.0f;
    }
    float v2 = 0.0f, s2 = 0.0f, cnt = 0.0f;
    if (tid < O) {
        v2 = st[2 * H + tid];
        s2 = st[2 * H + O + tid];
        if (count_mode == 2) cnt = counts[b * O + tid];
    }
    __syncthreads();
    for (int t = 0; t < a.steps; ++t) {
        const long row = (long)t * a.B + b;
        for (int q = 0; q < kUnits; ++q) {
            const int u = tid + q * kScanThreads;
            if (u < H) {
                v1[q] = a.beta * v1[q] - a.reset_theta * s1[q] +
                        a1[row * H + u];
                s1[q] = act(v1[q] - a.theta, a.mode, a.slope);
                if (rec_v1) {
                    rec_v1[row * H + u] = v1[q];
                    rec_s1[row * H + u] = s1[q];
                }
            }
        }
        for (int o = 0; o < O; ++o) {
            float p = 0.0f;
            for (int q = 0; q < kUnits; ++q) {
                const int u = tid + q * kScanThreads;
                if (u < H) p += w2s[o * H + u] * s1[q];
            }
            p = wave_sum(p);
            if (lane == 0) red[wave][o] = p;
        }
        __syncthreads();
        if (tid < O) {
            float a2 = red[0][tid];
            for (int w = 1; w < kScanThreads / 64; ++w) a2 += red[w][tid];
            v2 = a.beta * v2 - a.reset_theta * s2 + a2;
            s2 = act(v2 - a.theta, a.mode, a.slope);
            cnt += s2;
            if (rec_v2) {
                rec_v2[row * O + tid] = v2;
                rec_s2[row * O + tid] = s2;
            }
        }
        __syncthreads();
    }
    if (state_out) {
        float *so = state_out + (long)b * (2 * H + 2 * O);
        for (int q = 0; q < kUnits; ++q) {
            const int u = tid + q * kScanThreads;
            if (u < H) {
                so[u] = v1[q];
                so[H + u] = s1[q];
            }
        }
        if (tid < O) {
            so[2 * H + tid] = v2;
            so[2 * H + O + tid] = s2;
        }
    }
    if (count_mode != 0 && tid < O) counts[b * O + tid] = cnt;
}

// Adjoint of one block of steps, in reverse, for batch row blockIdx.x
// (r = reset): dL/ds2 = g - r theta d2', d2 = dL/ds2 S'(v2 - theta) +
// beta d2'; dL/ds1 = W2^T d2 - r theta d1', d1 = dL/ds1 S'(v1 - theta) +
// beta d1' (primes: the next step's adjoint states, carried between
// blocks).
__global__ __launch_bounds__(kScanThreads) void scan_bwd(
    ScanArgs a, const uint16_t *__restrict__ w2,
    const float *__restrict__ rec_v1, const float *__restrict__ rec_v2,
    const float *__restrict__ g, float *__restrict__ carry1,
    float *__restrict__ carry2, float *__restrict__ lam,
    float *__restrict__ g2) {
    extern __shared__ float w2s[];
    __shared__ float d2s[kMaxOut];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int H = a.H, O = a.O;
    for (int i = tid; i < O * H; i += kScanThreads) w2s[i] = bf16_value(w2[i]);
    float d1n[kUnits];
    for (int q = 0; q < kUnits; ++q) {
        const int u = tid + q * kScanThreads;
        d1n[q] = u < H ? carry1[(long)b * H + u] : 0.0f;
    }
    float d2n = tid < O ? carry2[b * O + tid] : 0.0f;
    const float gb = tid < O ? g[b * O + tid] : 0.0f;
    __syncthreads();
    for (int t = a.steps - 1; t >= 0; --t) {
        const long row = (long)t * a.B + b;
        if (tid < O) {
            const float ds2 = gb - a.reset_theta * d2n;
            const float d2 = ds2 * act_grad(rec_v2[row * O + tid] - a.theta,
                                            a.mode, a.slope) + a.beta * d2n;
            g2[row * O + tid] = d2;
            d2s[tid] = d2;
            d2n = d2;
        }
        __syncthreads();
        for (int q = 0; q < kUnits; ++q) {
            const int u = tid + q * kScanThreads;
            if (u >= H) continue;
            float gs = 0.0f;
            for (int o = 0; o < O; ++o) gs += w2s[o * H + u] * d2s[o];
            const float ds1 = gs - a.reset_theta * d1n[q];
            const float d1 = ds1 * act_grad(rec_v1[row * H + u] - a.theta,
                                            a.mode, a.slope) + a.beta * d1n[q];
            lam[row * H + u] = d1;
            d1n[q] = d1;
        }
        __syncthreads();
    }
    for (int q = 0; q < kUnits; ++q) {
        const int u = tid + q * kScanThreads;
        if (u < H) carry1[(long)b * H + u] = d1n[q];
    }
    if (tid < O) carry2[b * O + tid] = d2n;
}

// dW2[o][u] += sum over rows r (ascending) of G2[r][o] * S1[r][u].
__global__ void dw2_kernel(const float *__restrict__ g2,
                           const float *__restrict__ s1, float *__restrict__ dw2,
                           int M, int H, int O) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= O * H) return;
    const int o = i / H, u = i % H;
    float acc = 0.0f;
    for (int r = 0; r < M; ++r) acc += g2[(long)r * O + o] * s1[(long)r * H + u];
    dw2[i] += acc;
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
        lb[b] = (lab >= 0 && lab < O) ? m + logf(s) - c[lab] : __int_as_float(0x7fc00000);
        if (counts_out)
            for (int o = 0; o < O; ++o) counts_out[b * O + o] = c[o];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s += lb[b];
        *loss = s / (float)B;
    }
}

// dL/dcount = scale * (softmax(count) - onehot(label)) / B.
__global__ void loss_grad_kernel(const float *__restrict__ counts,
                                 const int32_t *__restrict__ labels, int B,
                                 int O, float scale, float *__restrict__ g) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float *c = counts + b * O;
    float m = c[0];
    for (int o = 1; o < O; ++o) m = fmaxf(m, c[o]);
    float s = 0.0f;
    for (int o = 0; o < O; ++o) s += expf(c[o] - m);
    for (int o = 0; o < O; ++o) {
        const float p = expf(c[o] - m) / s;
        g[b * O + o] = scale * (p - (o == labels[b] ? 1.0f : 0.0f)) / (float)B;
    }
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
    size_t state, counts, grad, a1, rv1, rs1, rv2, rs2, carry1, carry2, lam,
        g2, total;
};

Layout layout(const sg_snn_cfg &c) {
    const size_t B = c.batch, H = c.n_hid, O = c.n_out, ck = c.checkpoint;
    const size_t slots = (size_t)c.steps / ck + 1;
    const size_t M = ck * B;
    Layout l;
    size_t off = 0;
    auto take = [&](size_t floats) { size_t o = off; off += al(floats * 4); return o; };
    l.state = take(slots * B * (2 * H + 2 * O));
    l.counts = take(B * O);
    l.grad = take(B * O);
    l.a1 = take(M * H);
    l.rv1 = take(M * H);
    l.rs1 = take(M * H);
    l.rv2 = take(M * O);
    l.rs2 = take(M * O);
    l.carry1 = take(B * H);
    l.carry2 = take(B * O);
    l.lam = take(M * H);
    l.g2 = take(M * O);
    l.total = off;
    return l;
}

int check(const sg_snn_cfg *c) {
    if (!c) return fail("null configuration");
    if (c->batch < 1 || c->steps < 1 || c->n_in < 1 || c->n_hid < 1 ||
        c->n_out < 1)
        return fail("sizes must be positive");
    if (c->n_hid > kMaxHid) return fail("n_hid above 1024");
    if (c->n_out > kMaxOut) return fail("n_out above 16");
    if ((long)c->n_out * c->n_hid > 16000)
        return fail("n_out * n_hid above 16000 (W2 is staged in LDS)");
    if (c->batch > 16384) return fail("batch above 16384");
    if (c->checkpoint < 1 || c->steps % c->checkpoint != 0)
        return fail("checkpoint must divide steps");
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

void input_current(const sg_snn_cfg &c, const uint16_t *x_block,
                   const uint16_t *w1, float *a1, hipStream_t s) {
    const int M = c.checkpoint * c.batch;
    hipLaunchKernelGGL(gemm_nt, dim3((c.n_hid + 63) / 64, (M + 63) / 64),
                       dim3(256), 0, s, x_block, w1, a1, M, c.n_hid, c.n_in);
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
    float *a1 = (float *)(ws + l.a1);
    const long slot = (long)c.batch * (2 * c.n_hid + 2 * c.n_out);
    const long xblock = (long)c.checkpoint * c.batch * c.n_in;
    const size_t lds = (size_t)c.n_out * c.n_hid * 4;
    const ScanArgs a = scan_args(c);
    fill(state, slot, 0.0f, s);
    for (int blk = 0; blk < c.steps / c.checkpoint; ++blk) {
        input_current(c, x + blk * xblock, w1, a1, s);
        hipLaunchKernelGGL(scan_fwd, dim3(c.batch), dim3(kScanThreads), lds, s,
                           a, a1, w2, state + blk * slot,
                           state + (blk + 1) * slot, cnt, blk == 0 ? 1 : 2,
                           nullptr, nullptr, nullptr, nullptr);
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
    float *a1 = (float *)(ws + l.a1);
    float *rv1 = (float *)(ws + l.rv1), *rs1 = (float *)(ws + l.rs1);
    float *rv2 = (float *)(ws + l.rv2), *rs2 = (float *)(ws + l.rs2);
    float *carry1 = (float *)(ws + l.carry1), *carry2 = (float *)(ws + l.carry2);
    float *lam = (float *)(ws + l.lam), *g2 = (float *)(ws + l.g2);
    const int H = c.n_hid, O = c.n_out, B = c.batch;
    const int M = c.checkpoint * B;
    const long slot = (long)B * (2 * H + 2 * O);
    const long xblock = (long)M * c.n_in;
    const size_t lds = (size_t)O * H * 4;
    const ScanArgs a = scan_args(c);
    hipLaunchKernelGGL(loss_grad_kernel, dim3((B + 255) / 256), dim3(256), 0, s,
                       cnt, labels, B, O, grad_scale, g);
    fill(dw1, (long)H * c.n_in, 0.0f, s);
    fill(dw2, (long)O * H, 0.0f, s);
    fill(carry1, (long)B * H, 0.0f, s);
    fill(carry2, (long)B * O, 0.0f, s);
    for (int blk = c.steps / c.checkpoint - 1; blk >= 0; --blk) {
        const uint16_t *xb = x + blk * xblock;
        input_current(c, xb, w1, a1, s);
        hipLaunchKernelGGL(scan_fwd, dim3(B), dim3(kScanThreads), lds, s, a, a1,
                           w2, state + blk * slot, (float *)nullptr, cnt, 0,
                           rv1, rs1, rv2, rs2);
        hipLaunchKernelGGL(scan_bwd, dim3(B), dim3(kScanThreads), lds, s, a, w2,
                           rv1, rv2, g, carry1, carry2, lam, g2);
        hipLaunchKernelGGL(gemm_tn, dim3((c.n_in + 63) / 64, (H + 63) / 64),
                           dim3(256), 0, s, lam, xb, dw1, M, H, c.n_in);
        hipLaunchKernelGGL(dw2_kernel, dim3((O * H + 255) / 256), dim3(256), 0,
                           s, g2, rs1, dw2, M, H, O);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(std::string("launch: ") + hipGetErrorString(e));
    return 0;
}

}  // extern "C"
