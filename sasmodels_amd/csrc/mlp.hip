// mlp.hip -- the learned surrogate: an MLP 6 -> 4x256 SiLU -> 1 that maps
// (parameters, log q) to log I, trained and served with BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulate) on gfx950.
//
// Numerics (identical in training and inference):
//   layer 1 (6 -> 256)      FP32 VALU
//   layers 2-4 (256 -> 256) inputs and weights rounded to BF16, MFMA, FP32 sums
//   layer 5 (256 -> 1)      FP32 VALU on the FP32 SiLU of layer 4
// Parameters live in one flat FP32 vector (the optimizer's), see the offsets
// below; BF16 copies of W2..W4 (both orientations) are repacked before use.
// No reference counterpart: the reference has no surrogate.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "common.h"

using namespace saship_host;

namespace mlp {

constexpr int D = 6, H = 256;
// offsets (floats) in the flat parameter vector
constexpr int OFF_W1 = 0;                 // [D][H]
constexpr int OFF_B1 = OFF_W1 + D * H;    // [H]
constexpr int OFF_W2 = OFF_B1 + H;        // [H][H]  (k-major: W[k][n])
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H * H;
constexpr int OFF_W4 = OFF_B3 + H;
constexpr int OFF_B4 = OFF_W4 + H * H;
constexpr int OFF_W5 = OFF_B4 + H;        // [H]
constexpr int OFF_B5 = OFF_W5 + H;        // [1]
constexpr int NPARAM = OFF_B5 + 1;
// BF16 pack: for l = 2..4, W_l (k-major) then W_l^T (n-major)
constexpr int PACK_PER_LAYER = 2 * H * H;
constexpr int NPACK = 3 * PACK_PER_LAYER;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

// sigmoid through the hardware reciprocal (1 ulp; exp(-z) = inf gives 0)
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// ------------------------------------------------------------------ pack
__global__ void k_pack(const float *__restrict__ theta, __bf16 *__restrict__ pack)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * H * H) return;
    const int l = i / (H * H), r = i % (H * H);
    const int k = r / H, n = r % H;
    const float w = theta[OFF_W2 + l * (H * H + H) + r];
    pack[l * PACK_PER_LAYER + r] = (__bf16)w;                 // W[k][n]
    pack[l * PACK_PER_LAYER + H * H + n * H + k] = (__bf16)w; // W^T[n][k]
}

// ------------------------------------------------------------------ layer 1 and 5 (VALU)
// Z1 = X W1 + b1; A1 = bf16(silu(Z1))
__global__ void k_layer1(const float *__restrict__ X, int B, const float *__restrict__ theta,
                         float *__restrict__ Z1, __bf16 *__restrict__ A1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * H) return;
    const int m = i / H, n = i % H;
    float z = theta[OFF_B1 + n];
#pragma unroll
    for (int k = 0; k < D; k++) z = fmaf(X[m * D + k], theta[OFF_W1 + k * H + n], z);
    Z1[i] = z;
    A1[i] = (__bf16)silu(z);
}

// y = silu(Z4) w5 + b5: one wave64 per row
__global__ void k_layer5(const float *__restrict__ Z4, int B, const float *__restrict__ theta,
                         float *__restrict__ y)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= B) return;
    float s = 0.0f;
    for (int n = lane; n < H; n += 64) s = fmaf(silu(Z4[(size_t)row * H + n]), theta[OFF_W5 + n], s);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) y[row] = s + theta[OFF_B5];
}

// ------------------------------------------------------------------ generic BF16 GEMM
// C[M][N] = op(A)[M][K] * op(B)[K][N], FP32 accumulation, 64x64 tile per
// workgroup of four wave64s (each 32x32 = 2x2 MFMA 16x16x32 tiles).
//   TA: A stored [K][M] (else [M][K]); TB: B stored [N][K] (else [K][N]).
// Epilogues:
//   EPI_FWD   z = acc + bias[n]; Zout[m][n] = z; Aout[m][n] = bf16(silu(z))
//   EPI_BWD   d = acc * dsilu(Zin[m][n]); Dout32[m][n] = d; Dout16[m][n] = bf16(d)
//   EPI_PART  P[split][m][n] = acc (split-K partial)
enum { EPI_FWD = 0, EPI_BWD = 1, EPI_PART = 2 };

struct GemmArgs {
    int M, N, K;
    const __bf16 *A;
    int lda;
    const __bf16 *B;
    int ldb;
    int k_per_split;
    const float *bias;
    const float *Zin;
    float *out32;
    __bf16 *out16;
    int ldc;
};

constexpr int BM = 64, BN = 64, BK = 32, LDSK = BK + 8;   // +8 bf16 pad against bank conflicts

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    __shared__ __align__(16) __bf16 As[BM][LDSK];
    __shared__ __align__(16) __bf16 Bs[BN][LDSK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kbeg = blockIdx.z * g.k_per_split;
    const int kend = min(g.K, kbeg + g.k_per_split);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        // ---- stage A tile (BM x BK) as As[m][k]
        if (!TA) {
            const int m = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int gm = m0 + m, gk = k0 + kk + j;
                As[m][kk + j] = (gm < g.M && gk < kend) ? g.A[(size_t)gm * g.lda + gk] : (__bf16)0.0f;
            }
        } else {
            const int kk = tid >> 3, m = (tid & 7) * 8;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int gm = m0 + m + j, gk = k0 + kk;
                As[m + j][kk] = (gm < g.M && gk < kend) ? g.A[(size_t)gk * g.lda + gm] : (__bf16)0.0f;
            }
        }
        // ---- stage B tile (BK x BN) as Bs[n][k]
        if (!TB) {
            const int kk = tid >> 3, n = (tid & 7) * 8;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int gn = n0 + n + j, gk = k0 + kk;
                Bs[n + j][kk] = (gn < g.N && gk < kend) ? g.B[(size_t)gk * g.ldb + gn] : (__bf16)0.0f;
            }
        } else {
            const int n = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int gn = n0 + n, gk = k0 + kk + j;
                Bs[n][kk + j] = (gn < g.N && gk < kend) ? g.B[(size_t)gn * g.ldb + gk] : (__bf16)0.0f;
            }
        }
        __syncthreads();
        // ---- MFMA: lane holds row (lane%16), k block (lane/16)*8
        const int r = lane & 15, kb = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = *(const bf16x8 *)&As[wm + 16 * i + r][kb];
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = *(const bf16x8 *)&Bs[wn + 16 * j + r][kb];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    // ---- epilogue: acc[i][j][v] is C[row][col], col = lane%16, row = (lane/16)*4 + v
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int gm = m0 + wm + 16 * i + (lane >> 4) * 4 + v;
                const int gn = n0 + wn + 16 * j + (lane & 15);
                if (gm >= g.M || gn >= g.N) continue;
                const float x = acc[i][j][v];
                const size_t idx = (size_t)gm * g.ldc + gn;
                if (EPI == EPI_FWD) {
                    const float z = x + g.bias[gn];
                    g.out32[idx] = z;
                    g.out16[idx] = (__bf16)silu(z);
                } else if (EPI == EPI_BWD) {
                    const float d = x * dsilu(g.Zin[idx]);
                    g.out32[idx] = d;
                    g.out16[idx] = (__bf16)d;
                } else {
                    g.out32[(size_t)blockIdx.z * g.M * g.ldc + idx] = x;
                }
            }
}

// ------------------------------------------------------------------ backward helpers
// dy = 2 (y - t) / B; loss partial sums per block
__global__ void k_loss(const float *__restrict__ y, const float *__restrict__ t, int B,
                       float *__restrict__ dy, float *__restrict__ loss_part)
{
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float e = 0.0f;
    if (i < B) {
        const float r = y[i] - t[i];
        dy[i] = 2.0f * r / (float)B;
        e = r * r;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// dZ4 = (dy w5) * dsilu(Z4), FP32 and BF16
__global__ void k_dz4(const float *__restrict__ dy, const float *__restrict__ Z4, int B,
                      const float *__restrict__ theta, float *__restrict__ d32, __bf16 *__restrict__ d16)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * H) return;
    const int m = i / H, n = i % H;
    const float d = dy[m] * theta[OFF_W5 + n] * dsilu(Z4[i]);
    d32[i] = d;
    d16[i] = (__bf16)d;
}

// Column sums over rows in fixed order: out[n] = sum_m X[m][n] (one block per 256 columns,
// threads split rows, then a tree) -- used for bias gradients and dW5, dW1.
__global__ void k_colsum(const float *__restrict__ X, int B, int ld, int N, float *__restrict__ out)
{
    // block: 256 threads = 8 row groups x 32 columns
    __shared__ float red[8][32];
    const int c = blockIdx.x * 32 + (threadIdx.x & 31);
    const int g = threadIdx.x >> 5;
    float s = 0.0f;
    if (c < N)
        for (int m = g; m < B; m += 8) s += X[(size_t)m * ld + c];
    red[g][threadIdx.x & 31] = s;
    __syncthreads();
    if (g == 0 && c < N) {
        float t = 0.0f;
        for (int k = 0; k < 8; k++) t += red[k][threadIdx.x & 31];
        out[c] = t;
    }
}

// dW5[n] = sum_m silu(Z4[m][n]) dy[m]; db5 = sum_m dy[m]
__global__ void k_dw5(const float *__restrict__ Z4, const float *__restrict__ dy, int B,
                      float *__restrict__ grad)
{
    __shared__ float red[8][32];
    const int n = blockIdx.x * 32 + (threadIdx.x & 31);
    const int g = threadIdx.x >> 5;
    float s = 0.0f;
    if (n < H)
        for (int m = g; m < B; m += 8) s += silu(Z4[(size_t)m * H + n]) * dy[m];
    red[g][threadIdx.x & 31] = s;
    __syncthreads();
    if (g == 0 && n < H) {
        float t = 0.0f;
        for (int k = 0; k < 8; k++) t += red[k][threadIdx.x & 31];
        grad[OFF_W5 + n] = t;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float t = 0.0f;
        for (int m = 0; m < B; m++) t += dy[m];
        grad[OFF_B5] = t;
    }
}

// dW1[k][n] = sum_m X[m][k] dZ1[m][n]
__global__ void k_dw1(const float *__restrict__ X, const float *__restrict__ dZ1, int B,
                      float *__restrict__ grad)
{
    __shared__ float red[8][32];
    const int idx = blockIdx.x * 32 + (threadIdx.x & 31);    // over D*H
    const int g = threadIdx.x >> 5;
    const int k = idx / H, n = idx % H;
    float s = 0.0f;
    if (idx < D * H)
        for (int m = g; m < B; m += 8) s += X[(size_t)m * D + k] * dZ1[(size_t)m * H + n];
    red[g][threadIdx.x & 31] = s;
    __syncthreads();
    if (g == 0 && idx < D * H) {
        float t = 0.0f;
        for (int q = 0; q < 8; q++) t += red[q][threadIdx.x & 31];
        grad[OFF_W1 + idx] = t;
    }
}

// out[i] = sum over splits of P[s][i]
__global__ void k_sum_splits(const float *__restrict__ P, int splits, int n, float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < splits; k++) s += P[(size_t)k * n + i];
    out[i] = s;
}

// ------------------------------------------------------------------ fused inference
// One workgroup (4 waves) takes TM rows at a time: features -> layer 1 (VALU)
// -> BF16 activations in LDS -> layers 2-4 (each wave: all rows x 64 columns,
// W^T fragments straight from global/L2) -> layer 5 (FP32) -> output.
// Measured on the 4096 x 512 workload: TM = 64 (two waves per SIMD) runs at
// 420 TFLOP/s, TM = 128 (fewer weight reads per row, but 385 registers and
// one wave per SIMD) at 365: the SiLU epilogue (two quarter-rate
// transcendentals per activation), not the weight traffic, sets the pace.
// Rows are (curve, q) pairs; features are built in the kernel:
//   x = ((log R, log L, sR, sL, log drho, log10 q) - mean) / std
constexpr int BMI = 64;

// Features of one (curve, q) row and the analytic baseline of log I.  The
// polydisperse, orientation-averaged cylinder curve scaled by
// 1e-4 contrast^2 pi R^2 L depends on q, R, L only through qR and qL, so the
// features are (log qR, log qL, radius_pd, length_pd, bL, bR) with the
// smooth envelope terms bL = log(1 + (qL/pi)^2)/2 (thin rod, pi/(qL) at large
// qL) and bR = 3/2 log(1 + (qR)^2/4) (disc cross-section, (qR)^-3); the
// network predicts the residual r = log(I/(1e-4 contrast^2 pi R^2 L)) + bL + bR,
// a quantity of order one.
__device__ __forceinline__ float features(const float *cp, float lq, float *x)
{
    const float ln10 = 2.302585093f;
    const float lqr = fmaf(lq, ln10, logf(cp[0]));
    const float lql = fmaf(lq, ln10, logf(cp[1]));
    const float qr = expf(lqr), ql = expf(lql);
    const float u = ql * 0.318309886f;
    x[0] = lqr;
    x[1] = lql;
    x[2] = cp[2];
    x[3] = cp[3];
    x[4] = 0.5f * log1pf(u * u);
    x[5] = 1.5f * log1pf(0.25f * qr * qr);
    return -(x[4] + x[5]);
}

struct InferArgs {
    const float *theta;
    const __bf16 *pack;
    const float *curves;      // [ncurves][5]: radius, length, radius_pd, length_pd, contrast
    int ncurves;
    const float *logq;        // [nq] log10 q
    int nq;
    const float *norm;        // [12]: mean[6], std[6]
    int out_iq;               // 0: network output r; 1: I = 1e-4 drho^2 pi R^2 L exp(r + baseline)
    float *y;                 // [ncurves*nq]
};

template <int TM>
__global__ __launch_bounds__(256) void k_infer(InferArgs a)
{
    __shared__ __align__(16) __bf16 act[TM][H + 8];
    __shared__ float partial[4][TM];
    __shared__ float feat[TM][D + 1];            // standardised features, then the baseline
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t rows = (int64_t)a.ncurves * a.nq;
    const int wn = wave * 64;                     // this wave's 64 output columns
    for (int64_t r0 = (int64_t)blockIdx.x * TM; r0 < rows; r0 += (int64_t)gridDim.x * TM) {
        // ---- features once per row (first wave)
        if (tid < TM) {
            const int64_t row = r0 + tid;
            float x[D] = {0, 0, 0, 0, 0, 0};
            float base = 0.0f;
            if (row < rows) {
                const int c = (int)(row / a.nq), iq = (int)(row % a.nq);
                base = features(a.curves + (size_t)c * 5, a.logq[iq], x);
            }
#pragma unroll
            for (int k = 0; k < D; k++) feat[tid][k] = (x[k] - a.norm[k]) / a.norm[6 + k];
            feat[tid][D] = base;
        }
        __syncthreads();
        // ---- layer 1: thread handles column n = tid for all TM rows
        {
            const int n = tid;
            float w1[D];
#pragma unroll
            for (int k = 0; k < D; k++) w1[k] = a.theta[OFF_W1 + k * H + n];
            const float b1 = a.theta[OFF_B1 + n];
            for (int m = 0; m < TM; m++) {
                float z = b1;
#pragma unroll
                for (int k = 0; k < D; k++) z = fmaf(feat[m][k], w1[k], z);
                act[m][n] = (__bf16)silu(z);
            }
        }
        __syncthreads();
        // ---- layers 2-4 (MFMA); layer 4 keeps FP32 and feeds layer 5
        float out5[TM / 16][4];
        for (int layer = 0; layer < 3; layer++) {
            const __bf16 *Wt = a.pack + layer * PACK_PER_LAYER + H * H;   // [n][k]
            const float *bias = a.theta + OFF_B2 + layer * (H * H + H);
            floatx4 acc[TM / 16][4];
#pragma unroll
            for (int i = 0; i < TM / 16; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            const int r = lane & 15, kb = (lane >> 4) * 8;
            for (int k0 = 0; k0 < H; k0 += 32) {
                bf16x8 b[4];
#pragma unroll
                for (int j = 0; j < 4; j++) b[j] = *(const bf16x8 *)(Wt + (size_t)(wn + 16 * j + r) * H + k0 + kb);
#pragma unroll
                for (int i = 0; i < TM / 16; i++) {
                    const bf16x8 av = *(const bf16x8 *)&act[16 * i + r][k0 + kb];
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b[j], acc[i][j], 0, 0, 0);
                }
            }
            __syncthreads();     // every wave has read the layer input
            if (layer < 2) {
#pragma unroll
                for (int i = 0; i < TM / 16; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int m = 16 * i + (lane >> 4) * 4 + v, n = wn + 16 * j + (lane & 15);
                            act[m][n] = (__bf16)silu(acc[i][j][v] + bias[n]);
                        }
                __syncthreads();
            } else {
                // layer 5 partial dot over this wave's 64 columns
#pragma unroll
                for (int i = 0; i < TM / 16; i++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        float s = 0.0f;
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const int n = wn + 16 * j + (lane & 15);
                            s = fmaf(silu(acc[i][j][v] + bias[n]), a.theta[OFF_W5 + n], s);
                        }
                        // reduce over the 16 lanes sharing these rows (lane%16 varies)
#pragma unroll
                        for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off);
                        out5[i][v] = s;
                    }
            }
        }
        if ((lane & 15) == 0) {
#pragma unroll
            for (int i = 0; i < TM / 16; i++)
#pragma unroll
                for (int v = 0; v < 4; v++) partial[wave][16 * i + (lane >> 4) * 4 + v] = out5[i][v];
        }
        __syncthreads();
        if (tid < TM) {
            const int64_t row = r0 + tid;
            if (row < rows) {
                const float yv = partial[0][tid] + partial[1][tid] + partial[2][tid] + partial[3][tid] + a.theta[OFF_B5];
                if (a.out_iq) {
                    const float *cp = a.curves + (size_t)(row / a.nq) * 5;
                    const float v0 = 3.14159265f * cp[0] * cp[0] * cp[1];
                    a.y[row] = 1e-4f * cp[4] * cp[4] * v0 * expf(yv + feat[tid][D]);
                } else {
                    a.y[row] = yv;
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace mlp

namespace {

int launch_gemm(int epi, bool ta, bool tb, const mlp::GemmArgs &g, int splits, hipStream_t st)
{
    using namespace mlp;
    const dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, splits);
#define GEMM_CASE(TA, TB, E) \
    if (ta == TA && tb == TB && epi == E) { k_gemm<TA, TB, E><<<grid, 256, 0, st>>>(g); return 0; }
    GEMM_CASE(false, true, EPI_FWD)
    GEMM_CASE(false, true, EPI_BWD)
    GEMM_CASE(true, false, EPI_PART)
#undef GEMM_CASE
    return fail(SASHIP_EINVAL, "unsupported gemm variant");
}

// Workspace layout (bytes) for a training batch of B rows.
struct Work {
    float *Z[4];
    __bf16 *A[3];
    float *d32[2];
    __bf16 *d16[2];
    float *y, *dy, *loss_part, *split;
    size_t bytes;
};

constexpr int DW_SPLITS = 16;

Work carve(char *base, int B)
{
    Work w{};
    size_t off = 0;
    auto take = [&](size_t n) {
        char *p = base ? base + off : nullptr;
        off += (n + 255) & ~(size_t)255;
        return p;
    };
    for (int i = 0; i < 4; i++) w.Z[i] = (float *)take(sizeof(float) * (size_t)B * mlp::H);
    for (int i = 0; i < 3; i++) w.A[i] = (__bf16 *)take(sizeof(__bf16) * (size_t)B * mlp::H);
    for (int i = 0; i < 2; i++) w.d32[i] = (float *)take(sizeof(float) * (size_t)B * mlp::H);
    for (int i = 0; i < 2; i++) w.d16[i] = (__bf16 *)take(sizeof(__bf16) * (size_t)B * mlp::H);
    w.y = (float *)take(sizeof(float) * B);
    w.dy = (float *)take(sizeof(float) * B);
    w.loss_part = (float *)take(sizeof(float) * ((B + 255) / 256));
    w.split = (float *)take(sizeof(float) * DW_SPLITS * mlp::H * mlp::H);
    w.bytes = off;
    return w;
}

int forward(const float *theta, const __bf16 *pack, const float *X, int B, const Work &w, hipStream_t st)
{
    using namespace mlp;
    const int nb = (B * H + 255) / 256;
    k_layer1<<<nb, 256, 0, st>>>(X, B, theta, w.Z[0], w.A[0]);
    HIP_TRY(hipGetLastError());
    for (int l = 0; l < 3; l++) {
        GemmArgs g{B, H, H, w.A[l], H, pack + l * PACK_PER_LAYER + H * H, H, H,
                   theta + OFF_B2 + l * (H * H + H), nullptr, w.Z[l + 1],
                   l < 2 ? w.A[l + 1] : w.A[2], H};
        if (l == 2) {
            // layer 4 keeps only Z4 (layer 5 takes silu(Z4) in FP32); the
            // BF16 output goes to a scratch slot that nothing reads
            g.out16 = w.d16[1];
        }
        if (int e = launch_gemm(EPI_FWD, false, true, g, 1, st)) return e;
        HIP_TRY(hipGetLastError());
    }
    k_layer5<<<(B + 3) / 4, 256, 0, st>>>(w.Z[3], B, theta, w.y);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int saship_mlp_sizes(int32_t *n_params, int32_t *n_pack, int64_t *work_bytes, int32_t batch)
{
    if (!n_params || !n_pack || !work_bytes) return fail(SASHIP_EINVAL, "NULL argument");
    if (batch < 1) return fail(SASHIP_EINVAL, "batch must be >= 1");
    *n_params = mlp::NPARAM;
    *n_pack = mlp::NPACK;
    *work_bytes = (int64_t)carve(nullptr, batch).bytes;
    return 0;
}

int saship_mlp_pack(const float *theta, void *pack, void *stream)
{
    if (!theta || !pack) return fail(SASHIP_EINVAL, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    mlp::k_pack<<<(3 * mlp::H * mlp::H + 255) / 256, 256, 0, st>>>(theta, (__bf16 *)pack);
    HIP_TRY(hipGetLastError());
    return 0;
}

int saship_mlp_forward(const float *theta, const void *pack, const float *X, int32_t batch,
                       void *work, float *y, void *stream)
{
    if (!theta || !pack || !X || !work || !y) return fail(SASHIP_EINVAL, "NULL argument");
    if (batch < 1) return fail(SASHIP_EINVAL, "batch must be >= 1");
    hipStream_t st = (hipStream_t)stream;
    Work w = carve((char *)work, batch);
    if (int e = forward(theta, (const __bf16 *)pack, X, batch, w, st)) return e;
    HIP_TRY(hipMemcpyAsync(y, w.y, sizeof(float) * batch, hipMemcpyDeviceToDevice, st));
    return 0;
}

int saship_mlp_train_step(const float *theta, const void *pack, const float *X, const float *target,
                          int32_t batch, void *work, float *grad, float *loss_parts, void *stream)
{
    using namespace mlp;
    if (!theta || !pack || !X || !target || !work || !grad) return fail(SASHIP_EINVAL, "NULL argument");
    if (batch < 1) return fail(SASHIP_EINVAL, "batch must be >= 1");
    hipStream_t st = (hipStream_t)stream;
    const int B = batch;
    const __bf16 *pk = (const __bf16 *)pack;
    Work w = carve((char *)work, B);
    if (int e = forward(theta, pk, X, B, w, st)) return e;
    const int nlb = (B + 255) / 256;
    k_loss<<<nlb, 256, 0, st>>>(w.y, target, B, w.dy, loss_parts ? loss_parts : w.loss_part);
    k_dw5<<<H / 32, 256, 0, st>>>(w.Z[3], w.dy, B, grad);
    const int nb = (B * H + 255) / 256;
    k_dz4<<<nb, 256, 0, st>>>(w.dy, w.Z[3], B, theta, w.d32[0], w.d16[0]);
    HIP_TRY(hipGetLastError());
    // layers 4, 3, 2: dW_l = A_{l-1}^T dZ_l, db_l = colsum dZ_l, dZ_{l-1} = (dZ_l W_l^T) * dsilu(Z_{l-1})
    int cur = 0;
    const int kps = ((B + DW_SPLITS - 1) / DW_SPLITS + 31) / 32 * 32;
    for (int l = 2; l >= 0; l--) {
        const __bf16 *Ain = w.A[l];                  // input of layer l+2 (A1..A3)
        GemmArgs gw{H, H, B, Ain, H, w.d16[cur], H, kps, nullptr, nullptr, w.split, nullptr, H};
        if (int e = launch_gemm(EPI_PART, true, false, gw, DW_SPLITS, st)) return e;
        k_sum_splits<<<(H * H + 255) / 256, 256, 0, st>>>(w.split, DW_SPLITS, H * H, grad + OFF_W2 + l * (H * H + H));
        k_colsum<<<H / 32, 256, 0, st>>>(w.d32[cur], B, H, H, grad + OFF_B2 + l * (H * H + H));
        // dA_{l} = dZ W^T with W stored [k][n] (row-major [H_in][H_out]): B operand
        // of the product is W^T, i.e. stored [N=H_in][K=H_out] -> the k-major W itself
        GemmArgs gd{B, H, H, w.d16[cur], H, pk + l * PACK_PER_LAYER, H, H, nullptr, w.Z[l],
                    w.d32[cur ^ 1], w.d16[cur ^ 1], H};
        if (int e = launch_gemm(EPI_BWD, false, true, gd, 1, st)) return e;
        HIP_TRY(hipGetLastError());
        cur ^= 1;
    }
    // layer 1
    k_dw1<<<(D * H + 31) / 32, 256, 0, st>>>(X, w.d32[cur], B, grad);
    k_colsum<<<H / 32, 256, 0, st>>>(w.d32[cur], B, H, H, grad + OFF_B1);
    HIP_TRY(hipGetLastError());
    return 0;
}

int saship_mlp_infer(const float *theta, const void *pack, const float *curves, int32_t ncurves,
                     const float *logq, int32_t nq, const float *norm, int32_t out_iq, float *y,
                     int32_t grid, void *stream)
{
    if (!theta || !pack || !curves || !logq || !norm || !y) return fail(SASHIP_EINVAL, "NULL argument");
    if (ncurves < 0 || nq < 0) return fail(SASHIP_EINVAL, "negative size");
    if ((int64_t)ncurves * nq == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    const int64_t rows = (int64_t)ncurves * nq;
    const int64_t tiles = (rows + mlp::BMI - 1) / mlp::BMI;
    int g = grid > 0 ? grid : cu_count(dev) * 4;
    g = (int)std::min<int64_t>(g, tiles);
    mlp::InferArgs a{theta, (const __bf16 *)pack, curves, ncurves, logq, nq, norm, out_iq ? 1 : 0, y};
    mlp::k_infer<mlp::BMI><<<g, 256, 0, st>>>(a);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
