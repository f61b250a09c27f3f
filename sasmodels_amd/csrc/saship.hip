// saship.hip -- polydispersity-mesh I(q) kernels for gfx950 and the C ABI.
//
// One evaluation (saship_kernel_call) runs three kernels on the handle's
// stream:
//   1. k_expand  (one workgroup) walks mesh steps [pd_start, pd_stop) in the
//      reference order (kernel_iq.c:642-664, :682-823: level 0 innermost,
//      index_k = (step / stride_k) % length_k), applies VALID, the projection
//      weight |cos dtheta| for oriented 2D kernels (kernel_iq.c:621-626) and
//      the strict "weight > cutoff" test (:726), and compacts the accepted
//      points into FP32 records in mesh order with every per-point constant
//      folded in.  Sum w, sum w*V and sum w*R_eff are reduced in FP64.
//   2. k_fq1d / k_iq2d: q tiled 256 per workgroup (4 wave64s), mesh split
//      over blockIdx.y; each workgroup stages its records in LDS and every
//      lane accumulates w*F^2 (and w*F) in FP32 with compensated (TwoSum)
//      addition.  Form factors and orientation quadrature run in FP32.
//   3. k_finalize: one wave64 per output reduces the per-split partial sums
//      in FP64 with a fixed shuffle tree and writes the reference result
//      layout (adding to it when pd_start > 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/saship.h"
#include "models.h"

namespace sas {

constexpr int EXPAND_THREADS = 512;
constexpr int EVAL_THREADS = 256;
constexpr int TILE_PTS = 64;

// Record widths in float4 units.
template <class M> constexpr int rv1() { return (M::REC1D + 3) / 4; }
template <class M> constexpr int rv2() { return (M::REC2D + (M::ORIENTED ? 8 : 0) + 3) / 4; }
template <class M, int KIND> constexpr int rvk() { return KIND == 0 ? rv1<M>() : rv2<M>(); }

template <class M> constexpr int nvalues()
{
    return 2 + M::NPARS + (M::NMAG ? 4 + 3 * M::NMAG : 0);
}

// Compensated sum (Knuth TwoSum): s + c carries the exact running total.
struct KSum {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x)
    {
        const float t = s + x;
        const float z = t - s;
        c += (s - (t - z)) + (x - z);
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};

// ------------------------------------------------------------------ expand
template <class M, int KIND>
__global__ __launch_bounds__(EXPAND_THREADS) void k_expand(
    const int32_t *__restrict__ det, const double *__restrict__ values,
    int pd_start, int pd_stop, double cutoff, int reff_mode,
    float4 *__restrict__ recs, int *__restrict__ meta_n, double *__restrict__ meta_sums)
{
    constexpr int NP = M::NPARS, MP = M::MAX_PD;
    constexpr bool OR2D = (KIND != 0) && M::ORIENTED;
    constexpr int RV = rvk<M, KIND>();
    const int32_t *pd_par = det, *pd_len = det + MP, *pd_off = det + 2 * MP, *pd_stride = det + 3 * MP;
    const int num_weights = det[4 * MP + 1];
    const int theta_par = det[4 * MP + 3];
    const double *pd_value = values + nvalues<M>();
    const double *pd_weight = pd_value + num_weights;

    double base[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) base[i] = values[2 + i];
    double view_theta = 0.0, view_phi = 0.0;
    if (OR2D) {
        view_theta = values[theta_par + 2];
        view_phi = values[theta_par + 3];
#pragma unroll
        for (int i = 0; i < NP; i++)
            if (i == theta_par || i == theta_par + 1) base[i] = 0.0;
    }
    double sv, cv, sp, cp;
    sincos(view_theta * 0.017453292519943295, &sv, &cv);
    sincos(view_phi * 0.017453292519943295, &sp, &cp);

    __shared__ int wave_cnt[EXPAND_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    double acc_w = 0.0, acc_v = 0.0, acc_r = 0.0;
    int running = 0;
    for (int b0 = pd_start; b0 < pd_stop; b0 += EXPAND_THREADS) {
        const int step = b0 + tid;
        bool keep = false;
        float rec[4 * RV];
        if (step < pd_stop) {
            double p[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) p[i] = base[i];
            double w0 = 1.0;
            for (int k = MP - 1; k >= 0; k--) {
                const int idx = (step / pd_stride[k]) % pd_len[k];
                const double v = pd_value[pd_off[k] + idx];
                const int par = pd_par[k];
#pragma unroll
                for (int i = 0; i < NP; i++)
                    if (i == par) p[i] = v;
                w0 = pd_weight[pd_off[k] + idx] * w0;
            }
            double dth = 0.0, dph = 0.0;
            if (OR2D) {
#pragma unroll
                for (int i = 0; i < NP; i++) {
                    if (i == theta_par) dth = p[i];
                    if (i == theta_par + 1) dph = p[i];
                }
            }
            const double weight = OR2D ? fabs(cos(dth * 0.017453292519943295)) * w0 : w0;
            if (M::valid(p) && weight > cutoff) {
                keep = true;
                const double vol = M::volume(p);
                acc_w += weight;
                acc_v += weight * vol;
                if (reff_mode != 0) acc_r += weight * M::reff(reff_mode, p);
#pragma unroll
                for (int i = 0; i < 4 * RV; i++) rec[i] = 0.0f;
                if (KIND == 0) {
                    M::rec1d(p, weight, rec);
                } else {
                    M::rec2d(p, weight, rec);
                    if (OR2D) {
                        // rows of (jitter^-1 * view^-1) acting on (qx, qy, 0);
                        // row 3 is the reference's qac_rotation (kernel_iq.c:197-230),
                        // rows 1-2 give qab directly instead of sqrt(q^2 - qc^2).
                        double sdt, cdt, sdp, cdp;
                        sincos(dth * 0.017453292519943295, &sdt, &cdt);
                        sincos(dph * 0.017453292519943295, &sdp, &cdp);
                        const double v11 = cp * cv, v12 = sp * cv, v21 = -sp, v22 = cp;
                        const double v31 = sv * cp, v32 = sp * sv;
                        const double j11 = cdt, j12 = sdp * sdt, j13 = -sdt * cdp;
                        const double j22 = cdp, j23 = sdp;
                        const double j31 = sdt, j32 = -sdp * cdt, j33 = cdp * cdt;
                        float *rot = rec + M::REC2D;
                        rot[0] = (float)(j11 * v11 + j12 * v21 + j13 * v31);
                        rot[1] = (float)(j11 * v12 + j12 * v22 + j13 * v32);
                        rot[2] = (float)(j22 * v21 + j23 * v31);
                        rot[3] = (float)(j22 * v22 + j23 * v32);
                        rot[4] = (float)(j31 * v11 + j32 * v21 + j33 * v31);
                        rot[5] = (float)(j31 * v12 + j32 * v22 + j33 * v32);
                    }
                }
            }
        }
        const uint64_t mask = __ballot(keep);
        if (lane == 0) wave_cnt[wid] = __popcll(mask);
        __syncthreads();
        int woff = 0, total = 0;
#pragma unroll
        for (int w = 0; w < EXPAND_THREADS / 64; w++) {
            const int c = wave_cnt[w];
            woff += (w < wid) ? c : 0;
            total += c;
        }
        if (keep) {
            const int pos = running + woff + __popcll(mask & lt);
            float4 *dst = recs + (size_t)pos * RV;
#pragma unroll
            for (int j = 0; j < RV; j++)
                dst[j] = make_float4(rec[4 * j], rec[4 * j + 1], rec[4 * j + 2], rec[4 * j + 3]);
        }
        running += total;
        __syncthreads();
    }

    // fixed-order FP64 tree reduction of the normalisation sums
    __shared__ double red[3][EXPAND_THREADS];
    red[0][tid] = acc_w;
    red[1][tid] = acc_v;
    red[2][tid] = acc_r;
    __syncthreads();
    for (int s = EXPAND_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] += red[0][tid + s];
            red[1][tid] += red[1][tid + s];
            red[2][tid] += red[2][tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        meta_n[0] = running;
        meta_sums[0] = red[0][0];
        meta_sums[1] = red[1][0];
        meta_sums[2] = red[1][0];   // shell volume == form volume for these solids
        meta_sums[3] = red[2][0];
    }
}

// Range of compacted points handled by split `split` of `nsplit`.
__device__ __forceinline__ void split_range(int npts, int nsplit, int split, int &s0, int &s1)
{
    const int per = (npts + nsplit - 1) / nsplit;
    s0 = min(npts, split * per);
    s1 = min(npts, s0 + per);
}

// ------------------------------------------------------------------ 1D <F>, <F^2>
template <class M, bool FIXED>
__global__ __launch_bounds__(EVAL_THREADS) void k_fq1d(
    int nq, const float *__restrict__ q32, const float4 *__restrict__ recs,
    const int *__restrict__ meta_n, int nsplit, Quad quad, double *__restrict__ partial)
{
    constexpr int RV = rv1<M>();
    __shared__ float4 tile[TILE_PTS * RV];
    const int iq = blockIdx.x * EVAL_THREADS + threadIdx.x;
    const int split = blockIdx.y;
    int s0, s1;
    split_range(*meta_n, nsplit, split, s0, s1);
    const float q = iq < nq ? q32[iq] : 0.0f;
    KSum a1, a2;
    for (int t0 = s0; t0 < s1; t0 += TILE_PTS) {
        const int cnt = min(TILE_PTS, s1 - t0);
        for (int i = threadIdx.x; i < cnt * RV; i += EVAL_THREADS) tile[i] = recs[(size_t)t0 * RV + i];
        __syncthreads();
        if (iq < nq) {
            for (int j = 0; j < cnt; j++) {
                float r[4 * RV];
#pragma unroll
                for (int v = 0; v < RV; v++) {
                    const float4 x = tile[j * RV + v];
                    r[4 * v] = x.x; r[4 * v + 1] = x.y; r[4 * v + 2] = x.z; r[4 * v + 3] = x.w;
                }
                float f1, f2;
                M::template eval1d<FIXED>(q, r, quad, f1, f2);
                a1.add(f1);
                a2.add(f2);
            }
        }
        __syncthreads();
    }
    if (iq < nq) {
        partial[((size_t)0 * nq + iq) * nsplit + split] = a2.value();
        partial[((size_t)1 * nq + iq) * nsplit + split] = a1.value();
    }
}

// ------------------------------------------------------------------ 2D, plain and magnetic
// Effective SLD for spin cross section xs (kernel_iq.c:127-175): the
// magnetisation component perpendicular to q, projected on the polarisation
// (non-spin-flip) or on the two axes normal to it (spin-flip real/imag).
__device__ __forceinline__ double mag_sld(int xs, double qx, double qy, double cmt, double smt,
                                          double cmp, double smp, double sld,
                                          double mx, double my, double mz)
{
    const double qn = sqrt(qx * qx + qy * qy);
    const double ux = qx / qn, uy = qy / qn;
    const double k = (mx * ux + my * uy) / (ux * ux + uy * uy);
    const double px = mx - k * ux, py = my - k * uy, pz = mz;
    switch (xs) {
    case 0: return sld - (smt * cmp * px + smt * smp * py + cmt * pz);
    case 1:
    case 2: return -smp * px + cmp * py;
    case 3: return sld + (smt * cmp * px + smt * smp * py + cmt * pz);
    case 4: return -(-cmt * cmp * px - cmt * smp * py + smt * pz);
    default: return -cmt * cmp * px - cmt * smp * py + smt * pz;
    }
}

// Spin cross-section weights dd, du.re, ud.re, uu, du.im, ud.im (kernel_iq.c:84-124)
__device__ __forceinline__ void spin_weights(double in_spin, double out_spin, double *w)
{
    in_spin = fmin(fmax(in_spin, 0.0), 1.0);
    out_spin = fmin(fmax(out_spin, 0.0), 1.0);
    const double norm = out_spin < 0.5 ? 1.0 - out_spin : out_spin;
    w[0] = (1.0 - in_spin) * (1.0 - out_spin) / norm;
    w[1] = (1.0 - in_spin) * out_spin / norm;
    w[2] = in_spin * (1.0 - out_spin) / norm;
    w[3] = in_spin * out_spin / norm;
    w[4] = w[1];
    w[5] = w[2];
}

template <class M, bool MAG>
__global__ __launch_bounds__(EVAL_THREADS) void k_iq2d(
    int nq, const float2 *__restrict__ q32, const double *__restrict__ q64,
    const double *__restrict__ values, const float4 *__restrict__ recs,
    const int *__restrict__ meta_n, int nsplit, double *__restrict__ partial)
{
    constexpr int RV = rv2<M>();
    constexpr int NX = MAG ? 6 : 1, NC = M::NCON, NP = M::NPARS;
    __shared__ float4 tile[TILE_PTS * RV];
    const int iq = blockIdx.x * EVAL_THREADS + threadIdx.x;
    const int split = blockIdx.y;
    int s0, s1;
    split_range(*meta_n, nsplit, split, s0, s1);
    const float2 qv = iq < nq ? q32[iq] : make_float2(0.0f, 0.0f);
    const float qx = qv.x, qy = qv.y;
    const float qmag = sqrt_f(fmaf(qx, qx, qy * qy));

    // contrast terms per spin cross section (uniform active set)
    float con[NX][NC];
    float xsw[NX];
    bool active[NX];
    {
        double sld[M::NMAG];
#pragma unroll
        for (int k = 0; k < M::NMAG; k++) sld[k] = values[2 + M::sld_param(k)];
        if constexpr (!MAG) {
            M::contrasts(sld, con[0]);
            xsw[0] = 1.0f;
            active[0] = true;
        } else {
            double w[6];
            spin_weights(values[2 + NP], values[3 + NP], w);
            double smt, cmt, smp, cmp;
            sincos(values[4 + NP] * 0.017453292519943295, &smt, &cmt);
            sincos(values[5 + NP] * 0.017453292519943295, &smp, &cmp);
            const double qxd = iq < nq ? q64[2 * iq] : 0.0, qyd = iq < nq ? q64[2 * iq + 1] : 0.0;
            const bool q_ok = qxd * qxd + qyd * qyd > 1.e-16;
#pragma unroll
            for (int x = 0; x < NX; x++) {
                active[x] = w[x] > 1.e-8;
                xsw[x] = (active[x] && q_ok) ? (float)w[x] : 0.0f;
                double eff[M::NMAG];
#pragma unroll
                for (int k = 0; k < M::NMAG; k++) {
                    const double *mv = values + 2 + NP + 4 + 3 * k;
                    eff[k] = q_ok ? mag_sld(x, qxd, qyd, cmt, smt, cmp, smp, sld[k], mv[0], mv[1], mv[2]) : 0.0;
                }
                M::contrasts(eff, con[x]);
            }
        }
    }

    KSum acc;
    for (int t0 = s0; t0 < s1; t0 += TILE_PTS) {
        const int cnt = min(TILE_PTS, s1 - t0);
        for (int i = threadIdx.x; i < cnt * RV; i += EVAL_THREADS) tile[i] = recs[(size_t)t0 * RV + i];
        __syncthreads();
        if (iq < nq) {
            for (int j = 0; j < cnt; j++) {
                float r[4 * RV];
#pragma unroll
                for (int v = 0; v < RV; v++) {
                    const float4 x = tile[j * RV + v];
                    r[4 * v] = x.x; r[4 * v + 1] = x.y; r[4 * v + 2] = x.z; r[4 * v + 3] = x.w;
                }
                float qab = qmag, qc = 0.0f;
                if constexpr (M::ORIENTED) {
                    const float *rot = r + M::REC2D;
                    const float qa = fmaf(rot[0], qx, rot[1] * qy);
                    const float qb = fmaf(rot[2], qx, rot[3] * qy);
                    qc = fmaf(rot[4], qx, rot[5] * qy);
                    qab = sqrt_f(fmaf(qa, qa, qb * qb));
                }
                float g[NC];
                M::geom2d(qab, qc, qmag, r, g);
                float f2 = 0.0f;
#pragma unroll
                for (int x = 0; x < NX; x++) {
                    if (active[x]) {
                        float F = 0.0f;
#pragma unroll
                        for (int k = 0; k < NC; k++) F = fmaf(con[x][k], g[k], F);
                        f2 = fmaf(xsw[x] * F, F, f2);
                    }
                }
                acc.add(r[M::A2D] * f2);
            }
        }
        __syncthreads();
    }
    if (iq < nq) partial[(size_t)iq * nsplit + split] = acc.value();
}

// ------------------------------------------------------------------ finalize
__global__ __launch_bounds__(256) void k_finalize(
    int nq, int nout, int nsplit, const double *__restrict__ partial,
    const double *__restrict__ meta_sums, int accumulate, double *__restrict__ result)
{
    const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave < nout * nq) {
        const int o = wave / nq, iq = wave - o * nq;
        const double *src = partial + ((size_t)o * nq + iq) * nsplit;
        double s = 0.0;
        for (int k = lane; k < nsplit; k += 64) s += src[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) {
            const size_t idx = (size_t)iq * nout + o;
            result[idx] = accumulate ? result[idx] + s : s;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        const size_t idx = (size_t)nout * nq + threadIdx.x;
        const double v = meta_sums[threadIdx.x];
        result[idx] = accumulate ? result[idx] + v : v;
    }
}

}  // namespace sas

// ====================================================================== host
using sas::Quad;

namespace {

thread_local std::string g_err;
thread_local int g_device = 0;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return fail(SASHIP_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Gauss-Legendre nodes and weights on [-1, 1], ascending, by Newton's method
// on the three-term Legendre recurrence (long double).
void gauss_legendre(int n, std::vector<double> &z, std::vector<double> &w)
{
    z.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; i++) {
        long double x = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
        long double dp = 1.0L;
        for (int it = 0; it < 100; it++) {
            long double p0 = 1.0L, p1 = x;
            for (int k = 2; k <= n; k++) {
                const long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0L);
            const long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-19L) break;
        }
        long double p0 = 1.0L, p1 = x;
        for (int k = 2; k <= n; k++) {
            const long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1.0L);
        const long double wt = 2.0L / ((1.0L - x * x) * dp * dp);
        z[n - 1 - i] = (double)x;
        z[i] = (double)-x;
        w[n - 1 - i] = w[i] = (double)wt;
    }
    if (n % 2) z[n / 2] = 0.0;
}

int cu_count(int dev)
{
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    return n;
}

}  // namespace

struct saship_model {
    int id = 0, quad_n = 0, dev = 0;
    float *tab = nullptr;     // 3 * ntab floats: a | b | w
    int ntab = 0;
    Quad quad{};
};

struct saship_kernel {
    saship_model *m = nullptr;
    int nq = 0, is2d = 0, dev = 0, cus = 256;
    hipStream_t st = nullptr;
    float *q32 = nullptr;
    double *q64 = nullptr;
    char *d_args = nullptr;
    size_t args_cap = 0;
    float4 *d_recs = nullptr;
    size_t recs_cap = 0;
    char *d_meta = nullptr;
    double *d_partial = nullptr;
    size_t partial_cap = 0;
    double *d_result = nullptr;
    char *h_stage = nullptr;
    size_t stage_cap = 0;
    double *h_result = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing = false;
    double t_ms = 0.0;
    int ncalls = 0;
    std::mutex mu;
};

namespace {

int model_npars(int id)
{
    static const int np[5] = {3, 6, 6, 8, 5};
    return np[id];
}
int model_maxpd(int id)
{
    static const int mp[5] = {1, 4, 4, 5, 2};
    return mp[id];
}
int model_nvalues(int id)
{
    static const int nv[5] = {15, 18, 18, 23, 20};
    return nv[id];
}

// Grow a device buffer to at least `bytes`.
template <class T>
int ensure(T *&ptr, size_t &cap, size_t bytes)
{
    if (bytes <= cap && ptr) return 0;
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, 256);
    HIP_TRY(hipMalloc((void **)&ptr, want));
    cap = want;
    return 0;
}

int ensure_host(char *&ptr, size_t &cap, size_t bytes)
{
    if (bytes <= cap && ptr) return 0;
    if (ptr) hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, 4096);
    HIP_TRY(hipHostMalloc((void **)&ptr, want, hipHostMallocDefault));
    cap = want;
    return 0;
}

// Launch the three kernels of one evaluation for model M.
template <class M>
int launch(saship_kernel *k, int kind, const int32_t *d_det, const double *d_val,
           int pd_start, int pd_stop, double cutoff, int reff_mode, int accumulate)
{
    using namespace sas;
    const int nq = k->nq;
    const int nout = kind == 0 ? 2 : 1;
    const int npts_max = std::max(0, pd_stop - pd_start);
    const int rv = kind == 0 ? rv1<M>() : rv2<M>();
    if (int e = ensure(k->d_recs, k->recs_cap, sizeof(float4) * (size_t)std::max(1, npts_max) * rv)) return e;
    int *meta_n = (int *)k->d_meta;
    double *meta_sums = (double *)(k->d_meta + 64);

    switch (kind) {
    case 0: k_expand<M, 0><<<1, EXPAND_THREADS, 0, k->st>>>(d_det, d_val, pd_start, pd_stop, cutoff, reff_mode, k->d_recs, meta_n, meta_sums); break;
    case 1: k_expand<M, 1><<<1, EXPAND_THREADS, 0, k->st>>>(d_det, d_val, pd_start, pd_stop, cutoff, reff_mode, k->d_recs, meta_n, meta_sums); break;
    default: k_expand<M, 2><<<1, EXPAND_THREADS, 0, k->st>>>(d_det, d_val, pd_start, pd_stop, cutoff, reff_mode, k->d_recs, meta_n, meta_sums); break;
    }
    HIP_TRY(hipGetLastError());

    // Split the mesh so that about 8 workgroups land on every CU.
    const int ntiles = std::max(1, (nq + EVAL_THREADS - 1) / EVAL_THREADS);
    const int target = std::max(1, (k->cus * 8) / std::max(1, ntiles));
    const int ppb = std::max(1, (npts_max + target - 1) / target);
    const int nsplit = std::max(1, std::min(65535, (npts_max + ppb - 1) / ppb));
    if (int e = ensure(k->d_partial, k->partial_cap, sizeof(double) * (size_t)nout * nq * nsplit)) return e;

    const dim3 grid(ntiles, nsplit);
    if (k->timing) HIP_TRY(hipEventRecord(k->ev0, k->st));
    if (kind == 0) {
        if (k->m->quad_n > 0)
            k_fq1d<M, true><<<grid, EVAL_THREADS, 0, k->st>>>(nq, k->q32, k->d_recs, meta_n, nsplit, k->m->quad, k->d_partial);
        else
            k_fq1d<M, false><<<grid, EVAL_THREADS, 0, k->st>>>(nq, k->q32, k->d_recs, meta_n, nsplit, k->m->quad, k->d_partial);
    } else if (kind == 1) {
        k_iq2d<M, false><<<grid, EVAL_THREADS, 0, k->st>>>(nq, (const float2 *)k->q32, k->q64, d_val, k->d_recs, meta_n, nsplit, k->d_partial);
    } else {
        k_iq2d<M, true><<<grid, EVAL_THREADS, 0, k->st>>>(nq, (const float2 *)k->q32, k->q64, d_val, k->d_recs, meta_n, nsplit, k->d_partial);
    }
    HIP_TRY(hipGetLastError());
    if (k->timing) HIP_TRY(hipEventRecord(k->ev1, k->st));

    const int nwaves = nout * nq;
    const int fblocks = std::max(1, (nwaves + 3) / 4);
    k_finalize<<<fblocks, 256, 0, k->st>>>(nq, nout, nsplit, k->d_partial, meta_sums, accumulate, k->d_result);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

const char *saship_last_error(void) { return g_err.c_str(); }

int saship_device_count(int32_t *count)
{
    if (!count) return fail(SASHIP_EINVAL, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return 0;
}

int saship_init(int32_t device)
{
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(SASHIP_EINVAL, "device " + std::to_string(device) + " out of range (" + std::to_string(n) + " devices)");
    HIP_TRY(hipSetDevice(device));
    g_device = device;
    return 0;
}

int saship_model_create(int32_t model_id, int32_t quad_n, saship_model **out)
{
    if (!out) return fail(SASHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (model_id < 0 || model_id > 4) return fail(SASHIP_EINVAL, "unknown model id " + std::to_string(model_id));
    if (quad_n < 0 || quad_n > 100000) return fail(SASHIP_EINVAL, "quadrature size must be 0 (adaptive) or 1..100000");
    HIP_TRY(hipSetDevice(g_device));
    saship_model *m = new (std::nothrow) saship_model();
    if (!m) return fail(SASHIP_ENOMEM, "out of host memory");
    m->id = model_id;
    m->quad_n = quad_n;
    m->dev = g_device;
    // orientation-quadrature tables (models with an orientation average in 1D)
    const bool cyl_family = model_id == 1 || model_id == 3;
    const bool ell = model_id == 2;
    std::vector<int> sizes;
    if (quad_n > 0) sizes = {quad_n};
    else sizes = {20, 76, 500, 5000};
    std::vector<float> a, b, w;
    std::vector<int> offs;
    if (cyl_family || ell) {
        for (int n : sizes) {
            std::vector<double> z, wt;
            gauss_legendre(n, z, wt);
            offs.push_back((int)a.size());
            for (int i = 0; i < n; i++) {
                if (cyl_family) {
                    const double th = z[i] * sas::PI_4_D + sas::PI_4_D;
                    a.push_back((float)sin(th));
                    b.push_back((float)cos(th));
                    w.push_back((float)(wt[i] * sin(th)));
                } else {
                    const double u = z[i] * 0.5 + 0.5;
                    a.push_back((float)(u * u));
                    b.push_back(0.0f);
                    w.push_back((float)wt[i]);
                }
            }
        }
    } else {
        offs.push_back(0);
        a.push_back(0.0f);
        b.push_back(0.0f);
        w.push_back(0.0f);
    }
    m->ntab = (int)a.size();
    hipError_t e = hipMalloc((void **)&m->tab, sizeof(float) * 3 * m->ntab);
    if (e == hipSuccess) e = hipMemcpy(m->tab, a.data(), sizeof(float) * m->ntab, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->tab + m->ntab, b.data(), sizeof(float) * m->ntab, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->tab + 2 * m->ntab, w.data(), sizeof(float) * m->ntab, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (m->tab) hipFree(m->tab);
        delete m;
        return fail(SASHIP_EHIP, std::string("quadrature upload: ") + hipGetErrorString(e));
    }
    m->quad.a = m->tab;
    m->quad.b = m->tab + m->ntab;
    m->quad.w = m->tab + 2 * m->ntab;
    m->quad.n = quad_n;
    for (int i = 0; i < 4; i++) m->quad.off[i] = i < (int)offs.size() ? offs[i] : 0;
    *out = m;
    return 0;
}

int saship_model_release(saship_model *m)
{
    if (!m) return 0;
    hipSetDevice(m->dev);
    if (m->tab) hipFree(m->tab);
    delete m;
    return 0;
}

int saship_kernel_create(saship_model *m, int32_t nq, const double *q, int32_t is2d, saship_kernel **out)
{
    if (!out) return fail(SASHIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!m) return fail(SASHIP_EINVAL, "model is NULL");
    if (nq < 0) return fail(SASHIP_EINVAL, "nq must be >= 0");
    if (nq > 0 && !q) return fail(SASHIP_EINVAL, "q is NULL");
    HIP_TRY(hipSetDevice(m->dev));
    saship_kernel *k = new (std::nothrow) saship_kernel();
    if (!k) return fail(SASHIP_ENOMEM, "out of host memory");
    k->m = m;
    k->nq = nq;
    k->is2d = is2d ? 1 : 0;
    k->dev = m->dev;
    k->cus = cu_count(m->dev);
    const size_t nqv = (size_t)nq * (k->is2d ? 2 : 1);
    std::vector<float> q32(std::max<size_t>(nqv, 1));
    for (size_t i = 0; i < nqv; i++) q32[i] = (float)q[i];
    hipError_t e = hipStreamCreateWithFlags(&k->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&k->q32, sizeof(float) * std::max<size_t>(nqv, 1));
    if (e == hipSuccess) e = hipMalloc((void **)&k->q64, sizeof(double) * std::max<size_t>(nqv, 1));
    if (e == hipSuccess) e = hipMalloc((void **)&k->d_meta, 256);
    if (e == hipSuccess) e = hipMalloc((void **)&k->d_result, sizeof(double) * (2 * (size_t)nq + 4));
    if (e == hipSuccess) e = hipHostMalloc((void **)&k->h_result, sizeof(double) * (2 * (size_t)nq + 4), hipHostMallocDefault);
    if (e == hipSuccess && nqv) e = hipMemcpy(k->q32, q32.data(), sizeof(float) * nqv, hipMemcpyHostToDevice);
    if (e == hipSuccess && nqv) e = hipMemcpy(k->q64, q, sizeof(double) * nqv, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&k->ev0);
    if (e == hipSuccess) e = hipEventCreate(&k->ev1);
    if (e != hipSuccess) {
        std::string msg = std::string("kernel create: ") + hipGetErrorString(e);
        saship_kernel_release(k);
        return fail(SASHIP_EHIP, msg);
    }
    *out = k;
    return 0;
}

int saship_kernel_call(saship_kernel *k, int32_t kind, const int32_t *details, int32_t n_details,
                       const double *values, int32_t n_values, double cutoff, int32_t reff_mode,
                       int32_t pd_start, int32_t pd_stop, double *result)
{
    if (!k) return fail(SASHIP_EINVAL, "kernel is NULL");
    if (!details || !values || !result) return fail(SASHIP_EINVAL, "details, values and result must not be NULL");
    if (kind < 0 || kind > 2) return fail(SASHIP_EINVAL, "kind must be 0 (Iq), 1 (Iqxy) or 2 (Imagnetic)");
    if ((kind == 0) == (k->is2d != 0))
        return fail(SASHIP_EINVAL, kind == 0 ? "Iq kernel needs 1D q" : "2D kernels need (qx, qy) input");
    const int id = k->m->id, mp = model_maxpd(id);
    if (n_details != 4 * mp + 4)
        return fail(SASHIP_EINVAL, "details must have 4*max_pd+4 = " + std::to_string(4 * mp + 4) + " entries");
    if (pd_start < 0 || pd_stop < pd_start) return fail(SASHIP_EINVAL, "need 0 <= pd_start <= pd_stop");
    const int32_t num_eval = details[4 * mp], num_weights = details[4 * mp + 1];
    if (pd_stop > num_eval) return fail(SASHIP_EINVAL, "pd_stop exceeds num_eval");
    if (num_weights < 0 || n_values < model_nvalues(id) + 2 * num_weights)
        return fail(SASHIP_EINVAL, "values shorter than nvalues + 2*num_weights");
    for (int i = 0; i < mp; i++) {
        const int par = details[i], len = details[mp + i], off = details[2 * mp + i], stride = details[3 * mp + i];
        if (par < 0 || par >= model_npars(id) || len < 1 || off < 0 || off + len > num_weights || stride < 1)
            return fail(SASHIP_EINVAL, "inconsistent details buffer at pd level " + std::to_string(i));
    }
    if (kind != 0 && (id == 1 || id == 2 || id == 3)) {
        const int tp = details[4 * mp + 3];
        if (tp < 0 || tp + 1 >= model_npars(id)) return fail(SASHIP_EINVAL, "theta_par out of range");
    }
    std::lock_guard<std::mutex> lock(k->mu);
    HIP_TRY(hipSetDevice(k->dev));
    const int nq = k->nq, nout = kind == 0 ? 2 : 1;
    const size_t vbytes = sizeof(double) * (size_t)n_values;
    const size_t dbytes = sizeof(int32_t) * (size_t)n_details;
    const size_t doff = (vbytes + 255) & ~(size_t)255;
    if (int e = ensure_host(k->h_stage, k->stage_cap, doff + dbytes)) return e;
    if (int e = ensure(k->d_args, k->args_cap, doff + dbytes)) return e;
    memcpy(k->h_stage, values, vbytes);
    memcpy(k->h_stage + doff, details, dbytes);
    HIP_TRY(hipMemcpyAsync(k->d_args, k->h_stage, doff + dbytes, hipMemcpyHostToDevice, k->st));
    const size_t rbytes = sizeof(double) * ((size_t)nout * nq + 4);
    const int accumulate = pd_start > 0 ? 1 : 0;
    if (accumulate) {
        memcpy(k->h_result, result, rbytes);
        HIP_TRY(hipMemcpyAsync(k->d_result, k->h_result, rbytes, hipMemcpyHostToDevice, k->st));
    }
    const int32_t *d_det = (const int32_t *)(k->d_args + doff);
    const double *d_val = (const double *)k->d_args;
    int rc;
    switch (id) {
    case 0: rc = launch<sas::Sphere>(k, kind, d_det, d_val, pd_start, pd_stop, cutoff, reff_mode, accumulate); break;
    case 1: rc = launch<sas::Cylinder>(k, kind, d_det, d_val, pd_start, pd_stop, cutoff, reff_mode, accumulate); break;
    case 2: rc = launch<sas::Ellipsoid>(k, kind, d_det, d_val, pd_start, pd_stop, cutoff, reff_mode, accumulate); break;
    case 3: rc = launch<sas::CoreShellCylinder>(k, kind, d_det, d_val, pd_start, pd_stop, cutoff, reff_mode, accumulate); break;
    default: rc = launch<sas::CoreShellSphere>(k, kind, d_det, d_val, pd_start, pd_stop, cutoff, reff_mode, accumulate); break;
    }
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(k->h_result, k->d_result, rbytes, hipMemcpyDeviceToHost, k->st));
    HIP_TRY(hipStreamSynchronize(k->st));
    memcpy(result, k->h_result, rbytes);
    if (k->timing) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, k->ev0, k->ev1));
        k->t_ms += ms;
        k->ncalls++;
    }
    return 0;
}

int saship_kernel_set_timing(saship_kernel *k, int32_t enable)
{
    if (!k) return fail(SASHIP_EINVAL, "kernel is NULL");
    std::lock_guard<std::mutex> lock(k->mu);
    k->timing = enable != 0;
    k->t_ms = 0.0;
    k->ncalls = 0;
    return 0;
}

int saship_kernel_get_timing(saship_kernel *k, double *eval_ms, int32_t *ncalls)
{
    if (!k || !eval_ms || !ncalls) return fail(SASHIP_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lock(k->mu);
    *eval_ms = k->t_ms;
    *ncalls = k->ncalls;
    return 0;
}

int saship_kernel_release(saship_kernel *k)
{
    if (!k) return 0;
    hipSetDevice(k->dev);
    if (k->st) hipStreamSynchronize(k->st);
    if (k->q32) hipFree(k->q32);
    if (k->q64) hipFree(k->q64);
    if (k->d_args) hipFree(k->d_args);
    if (k->d_recs) hipFree(k->d_recs);
    if (k->d_meta) hipFree(k->d_meta);
    if (k->d_partial) hipFree(k->d_partial);
    if (k->d_result) hipFree(k->d_result);
    if (k->h_stage) hipHostFree(k->h_stage);
    if (k->h_result) hipHostFree(k->h_result);
    if (k->ev0) hipEventDestroy(k->ev0);
    if (k->ev1) hipEventDestroy(k->ev1);
    if (k->st) hipStreamDestroy(k->st);
    delete k;
    return 0;
}

}  // extern "C"
