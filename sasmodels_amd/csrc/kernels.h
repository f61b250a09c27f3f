// kernels.h -- device kernels shared by the single-call and batched paths.
//
//   k_expand   one workgroup per parameter set walks mesh steps
//              [pd_start, pd_stop) in the reference order (kernel_iq.c:642-664,
//              :682-823: level 0 innermost, index_k = (step/stride_k) % length_k),
//              applies VALID, the projection weight |cos dtheta| of oriented 2D
//              kernels (kernel_iq.c:621-626) and the strict "weight > cutoff" test
//              (:726), and compacts accepted points into FP32 records, in mesh
//              order, with every per-point constant folded in.  Sum w, sum w*V and
//              sum w*R_eff are reduced in FP64 with a fixed tree.
//   k_fq1d     1D <F>, <F^2>: 256 q per workgroup (four wave64s), mesh split over
//              blockIdx.y; records staged in LDS, every lane sums w*F^2 and w*F in
//              FP32 with TwoSum compensation.
//   k_iq2d     2D |F|^2, plain or summed over spin cross sections, same tiling.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "models.h"

namespace sas {

constexpr int EXPAND_THREADS = 512;
// Work queue of the persistent batched kernels: WORK_QUEUES counters
// WORK_QUEUE_STRIDE ints apart (separate memory channels); counter c hands
// out the items [c n / WORK_QUEUES, (c+1) n / WORK_QUEUES).  One counter for
// every item serialises thousands of atomics per launch.
constexpr int WORK_QUEUES = 16, WORK_QUEUE_STRIDE = 1024;
constexpr int EVAL_THREADS = 256;
constexpr int TILE_PTS = 64;
// Records wider than this many float4 are not copied to registers by the
// 1D kernels (the evaluation reads them in place).
constexpr int BIG_RECORD = 8;
constexpr double DEG = 0.017453292519943295;   // pi/180

// Does model M have the FP32 2D path (REC2F, rec2f, geom2f)?
// direct-intensity models (Direct<D> in models.h): I itself, not an amplitude
template <class M, class = void> struct is_direct : std::false_type {};
template <class M> struct is_direct<M, std::void_t<decltype(M::DIRECT)>> : std::bool_constant<M::DIRECT> {};
template <class M, class = void> struct has_f32_2d : std::false_type {};
template <class M> struct has_f32_2d<M, std::void_t<decltype(M::REC2F)>> : std::true_type {};
template <class M> constexpr int rec2f_floats()
{
    if constexpr (has_f32_2d<M>::value) return M::REC2F + (M::ORIENTED ? 6 : 0);
    else return 0;
}

// Record widths in float4 units.  A 2D record is the FP64 part (model
// fields, then six FP64 rotation entries for oriented models) followed, for
// models with the FP32 path, by the FP32 part (six FP32 rotation entries
// for oriented models, then the model's FP32 fields).
template <class M> constexpr int rv1() { return (M::REC1D + 3) / 4; }
template <class M> constexpr int rv2d() { return (M::REC2D + (M::ORIENTED ? 12 : 0) + 3) / 4; }
template <class M> constexpr int rv2f() { return (rec2f_floats<M>() + 3) / 4; }
template <class M> constexpr int rv2() { return rv2d<M>() + rv2f<M>(); }
template <class M, int KIND> constexpr int rvk() { return KIND == 0 ? rv1<M>() : rv2<M>(); }

// Largest phase (radians) the FP32 2D path takes: a wave whose largest
// |q| times a record's length scale exceeds it evaluates that record in
// FP64 (rotation, argument and its reduction by pi/2).  At 400 rad the FP32
// phase is good to about 4e-5 rad.
constexpr float XF32_2D = 400.0f;

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

// Where one workgroup of k_expand finds its set and leaves its results.
struct ExpandArgs {
    const int32_t *det;         // details of set 0; set s at det + s*ndet
    const double *values;       // values of set 0; set s at values + voff[s]
    const int64_t *voff;        // per-set value offsets (NULL: single set)
    const int64_t *roff;        // per-set record offsets in records (NULL: 0)
    int ndet;
    int pd_start, pd_stop;      // pd_stop < 0: the set's whole mesh (num_eval)
    double cutoff;
    int reff_mode;
    float4 *recs;
    int *npts;                  // [nsets] accepted points
    double *sums;               // [4*nsets] sum w, sum w*V_form, sum w*V_shell, sum w*R_eff
    int *queue;                 // counters reset to 0 by workgroup 0 when not NULL
    float *spin;                // magnetic kernel: per-set spin constants (spin_floats<M>() each)
};

// Magnetic set-up that depends only on the set (written once per set by
// k_expand, read by every lane of the magnetic kernels), in floats:
//   [0, 6)   cross-section weights (0 when inactive)   [6, 12)  active flags
//   [12, 15) polarisation axis   [15, 18) spin-flip axis   [18, 21) the third axis
//   then per sld slot k, at 24 + 8k: the nuclear sld (a double in two words)
//   and the magnetisation vector (mx, my, mz).
template <class M> constexpr int spin_floats() { return 24 + 8 * (M::NMAG > 0 ? M::NMAG : 1); }
template <class M> __device__ void write_spin(const double *values, float *out);

template <class M, int KIND>
__global__ __launch_bounds__(EXPAND_THREADS) void k_expand(ExpandArgs a)
{
    constexpr int NP = M::NPARS, MP = M::MAX_PD;
    constexpr bool OR2D = (KIND != 0) && M::ORIENTED;
    constexpr int RV = rvk<M, KIND>();
    const int set = blockIdx.x;
    const int32_t *det = a.det + (size_t)set * a.ndet;
    const double *values = a.values + (a.voff ? a.voff[set] : 0);
    float4 *recs = a.recs + (a.roff ? a.roff[set] : 0) * RV;
    if (a.queue && set == 0 && threadIdx.x < WORK_QUEUES) a.queue[threadIdx.x * WORK_QUEUE_STRIDE] = 0;
    if constexpr (KIND == 2) {
        if (a.spin && threadIdx.x == 0) write_spin<M>(values, a.spin + (size_t)set * spin_floats<M>());
    }

    const int32_t *pd_par = det, *pd_len = det + MP, *pd_off = det + 2 * MP, *pd_stride = det + 3 * MP;
    const int num_eval = det[4 * MP];
    const int num_weights = det[4 * MP + 1];
    const int theta_par = det[4 * MP + 3];
    const int pd_start = a.pd_start;
    const int pd_stop = a.pd_stop < 0 ? num_eval : a.pd_stop;
    const double *pd_value = values + nvalues<M>();
    const double *pd_weight = pd_value + num_weights;

    double base[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) base[i] = values[2 + i];
    double view_theta = 0.0, view_phi = 0.0, view_psi = 0.0;
    if (OR2D) {
        view_theta = values[theta_par + 2];
        view_phi = values[theta_par + 3];
        if (M::ABC) view_psi = values[theta_par + 4];
#pragma unroll
        for (int i = 0; i < NP; i++)
            if (i == theta_par || i == theta_par + 1 || (M::ABC && i == theta_par + 2)) base[i] = 0.0;
    }
    double sv, cv, sp, cp, ss, cs;
    sincos(view_theta * DEG, &sv, &cv);
    sincos(view_phi * DEG, &sp, &cp);
    sincos(view_psi * DEG, &ss, &cs);
    // first two columns of view^-1 (kernel_iq.c:197-230, :251-296)
    double v11, v12, v21, v22;
    if (M::ABC) {
        v11 = -sp * ss + cp * cs * cv;
        v12 = sp * cs * cv + ss * cp;
        v21 = -sp * cs - ss * cp * cv;
        v22 = -sp * ss * cv + cp * cs;
    } else {
        v11 = cp * cv;
        v12 = sp * cv;
        v21 = -sp;
        v22 = cp;
    }
    const double v31 = sv * cp, v32 = sp * sv;

    __shared__ int wave_cnt[EXPAND_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    double acc_w = 0.0, acc_v = 0.0, acc_s = 0.0, acc_r = 0.0;
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
            double dth = 0.0, dph = 0.0, dps = 0.0;
            if (OR2D) {
#pragma unroll
                for (int i = 0; i < NP; i++) {
                    if (i == theta_par) dth = p[i];
                    if (i == theta_par + 1) dph = p[i];
                    if (M::ABC && i == theta_par + 2) dps = p[i];
                }
            }
            const double weight = OR2D ? fabs(cos(dth * DEG)) * w0 : w0;
            if (M::valid(p) && weight > a.cutoff) {
                keep = true;
                const double vol = M::volume(p);
                acc_w += weight;
                acc_v += weight * vol;
                if constexpr (M::HOLLOW) acc_s += weight * M::shell_volume(p);
                if (a.reff_mode != 0) acc_r += weight * M::reff(a.reff_mode, p);
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
                        // with psi (qabc models) the jitter also turns about
                        // the particle's c axis (kernel_iq.c:269-286)
                        double sdt, cdt, sdp, cdp, sds, cds;
                        sincos(dth * DEG, &sdt, &cdt);
                        sincos(dph * DEG, &sdp, &cdp);
                        sincos(dps * DEG, &sds, &cds);
                        double j11, j12, j13, j21, j22, j23;
                        if (M::ABC) {
                            j11 = cds * cdt;
                            j12 = sdp * sdt * cds + sds * cdp;
                            j13 = sdp * sds - sdt * cdp * cds;
                            j21 = -sds * cdt;
                            j22 = -sdp * sds * sdt + cdp * cds;
                            j23 = sdp * cds + sds * sdt * cdp;
                        } else {
                            j11 = cdt;
                            j12 = sdp * sdt;
                            j13 = -sdt * cdp;
                            j21 = 0.0;
                            j22 = cdp;
                            j23 = sdp;
                        }
                        const double j31 = sdt, j32 = -sdp * cdt, j33 = cdp * cdt;
                        float *rot = rec + M::REC2D;      // six FP64 entries
                        put_d(rot, 0, j11 * v11 + j12 * v21 + j13 * v31);
                        put_d(rot, 2, j11 * v12 + j12 * v22 + j13 * v32);
                        put_d(rot, 4, j21 * v11 + j22 * v21 + j23 * v31);
                        put_d(rot, 6, j21 * v12 + j22 * v22 + j23 * v32);
                        put_d(rot, 8, j31 * v11 + j32 * v21 + j33 * v31);
                        put_d(rot, 10, j31 * v12 + j32 * v22 + j33 * v32);
                    }
                    if constexpr (KIND != 0 && has_f32_2d<M>::value) {
                        float *f = rec + 4 * rv2d<M>();
                        if (OR2D) {
                            const float *rot = rec + M::REC2D;
                            for (int k = 0; k < 6; k++) f[k] = (float)get_d(rot, 2 * k);
                        }
                        M::rec2f(p, weight, f);
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
    __shared__ double red[4][EXPAND_THREADS];
    red[0][tid] = acc_w;
    red[1][tid] = acc_v;
    red[2][tid] = acc_s;
    red[3][tid] = acc_r;
    __syncthreads();
    for (int s = EXPAND_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) {
#pragma unroll
            for (int k = 0; k < 4; k++) red[k][tid] += red[k][tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        a.npts[set] = running;
        a.sums[4 * set + 0] = red[0][0];
        a.sums[4 * set + 1] = red[1][0];
        a.sums[4 * set + 2] = M::HOLLOW ? red[2][0] : red[1][0];   // solids: shell == form
        a.sums[4 * set + 3] = red[3][0];
    }
}

// Range of compacted points handled by split `split` of `nsplit`.
__device__ __forceinline__ void split_range(int npts, int nsplit, int split, int &s0, int &s1)
{
    const int per = (npts + nsplit - 1) / nsplit;
    s0 = min(npts, split * per);
    s1 = min(npts, s0 + per);
}

// Read-only view of device memory that no kernel running with it writes:
// loads through the constant address space go through the scalar cache when
// the address is wave-uniform, whatever else the kernel stores.
#define SAS_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const SAS_CONST T *as_const(const T *p)
{
    return (const SAS_CONST T *)p;
}

// Element loads through such a view (HIP's vector types cannot bind to it).
__device__ __forceinline__ float4 ld4(const SAS_CONST float4 *p)
{
    const SAS_CONST float *f = (const SAS_CONST float *)p;
    return make_float4(f[0], f[1], f[2], f[3]);
}
__device__ __forceinline__ int2 ld2(const SAS_CONST int2 *p)
{
    const SAS_CONST int *v = (const SAS_CONST int *)p;
    return make_int2(v[0], v[1]);
}

__device__ __forceinline__ void load_rec(const float4 *src, float *r, int rv)
{
    for (int v = 0; v < rv; v++) {
        const float4 x = src[v];
        r[4 * v] = x.x; r[4 * v + 1] = x.y; r[4 * v + 2] = x.z; r[4 * v + 3] = x.w;
    }
}

// ------------------------------------------------------------------ 1D <F>, <F^2>
template <class M, bool FIXED>
__global__ __launch_bounds__(EVAL_THREADS) void k_fq1d(
    int nq, const float *__restrict__ q32, const double *__restrict__ q64, const float4 *__restrict__ recs,
    const int *__restrict__ npts_p, int nsplit, Quad quad, double *__restrict__ partial)
{
    constexpr int RV = rv1<M>();
    __shared__ float4 tile[TILE_PTS * RV];
    const int iq = blockIdx.x * EVAL_THREADS + threadIdx.x;
    const int split = blockIdx.y;
    int s0, s1;
    split_range(*npts_p, nsplit, split, s0, s1);
    const float q = iq < nq ? q32[iq] : 0.0f;
    const double qd = M::FP64 && iq < nq ? q64[iq] : 0.0;
    KSum a1, a2;
    double d1 = 0.0, d2 = 0.0;     // FP64 models
    for (int t0 = s0; t0 < s1; t0 += TILE_PTS) {
        const int cnt = min(TILE_PTS, s1 - t0);
        for (int i = threadIdx.x; i < cnt * RV; i += EVAL_THREADS) tile[i] = recs[(size_t)t0 * RV + i];
        __syncthreads();
        if (iq < nq) {
            for (int j = 0; j < cnt; j++) {
                if constexpr (M::FP64 && RV > BIG_RECORD) {
                    // a record too large for registers is read where it lies in LDS
                    double f1, f2;
                    M::eval1d_d(qd, (const float *)(tile + j * RV), quad, f1, f2);
                    d1 += f1;
                    d2 += f2;
                    continue;
                }
                float r[4 * RV];
#pragma unroll
                for (int v = 0; v < RV; v++) {
                    const float4 x = tile[j * RV + v];
                    r[4 * v] = x.x; r[4 * v + 1] = x.y; r[4 * v + 2] = x.z; r[4 * v + 3] = x.w;
                }
                if constexpr (M::FP64) {
                    double f1, f2;
                    M::eval1d_d(qd, r, quad, f1, f2);
                    d1 += f1;
                    d2 += f2;
                } else {
                    float f1, f2;
                    M::template eval1d<FIXED>(q, r, quad, f1, f2);
                    a1.add(f1);
                    a2.add(f2);
                }
            }
        }
        __syncthreads();
    }
    if (iq < nq) {
        partial[((size_t)0 * nq + iq) * nsplit + split] = M::FP64 ? d2 : a2.value();
        partial[((size_t)1 * nq + iq) * nsplit + split] = M::FP64 ? d1 : a1.value();
    }
}

// ------------------------------------------------------------------ 2D, plain and magnetic
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

template <class M>
__device__ void write_spin(const double *values, float *out)
{
    constexpr int NP = M::NPARS;
    double w[6];
    spin_weights(values[2 + NP], values[3 + NP], w);
    double smt, cmt, smp, cmp;
    sincos(values[4 + NP] * DEG, &smt, &cmt);
    sincos(values[5 + NP] * DEG, &smp, &cmp);
    for (int x = 0; x < 6; x++) {
        const bool act = w[x] > 1.e-8;                    // kernel_iq.c:761
        out[x] = act ? (float)w[x] : 0.0f;
        out[6 + x] = act ? 1.0f : 0.0f;
    }
    const double axes[9] = {smt * cmp, smt * smp, cmt, -smp, cmp, 0.0, -cmt * cmp, -cmt * smp, smt};
    for (int i = 0; i < 9; i++) out[12 + i] = (float)axes[i];
    out[21] = out[22] = out[23] = 0.0f;
    for (int k = 0; k < M::NMAG; k++) {
        float *o = out + 24 + 8 * k;
        put_d(o, 0, values[2 + M::sld_param(k)]);
        const double *mv = values + 2 + NP + 4 + 3 * k;
        o[2] = (float)mv[0];
        o[3] = (float)mv[1];
        o[4] = (float)mv[2];
        o[5] = o[6] = o[7] = 0.0f;
    }
}

// Per-lane contrast terms for each spin cross section, from the set's values
// (nuclear) or its spin constants (magnetic).
template <class M, bool MAG>
struct Contrast {
    static constexpr int NX = MAG ? 6 : 1;
    static constexpr bool DIRECT = is_direct<M>::value;
    static constexpr int NM1 = M::NMAG > 0 ? M::NMAG : 1;
    float con[NX][M::NCON];
    float xsw[NX];
    bool active[NX];
    bool on = true;    // magnetic: q != 0 (the reference adds nothing at q = 0)
    // direct-intensity models: the effective slds of each cross section
    double effd[DIRECT && MAG ? NX : 1][DIRECT && MAG ? NM1 : 1];

    __device__ __forceinline__ void setup(const double *values, const float *spin, double qx, double qy)
    {
        if constexpr (!MAG) {
            double sld[NM1];
#pragma unroll
            for (int k = 0; k < M::NMAG; k++) sld[k] = values[2 + M::sld_param(k)];
            M::contrasts(values + 2, sld, con[0]);
            xsw[0] = 1.0f;
            active[0] = true;
        } else {
            // Effective slds of the six spin cross sections (kernel_iq.c:127-175):
            // the magnetisation component perpendicular to q, projected on the
            // polarisation (non-spin-flip: sld -+ P.Mperp) or on the two axes
            // normal to it (spin-flip, real and imaginary).  The axes, weights
            // and magnetisations are the set's constants (scalar cache); per
            // lane only the projection perpendicular to q, in FP32, with the
            // reference's division by the unit vector's squared norm.
            const SAS_CONST float *sc = as_const(spin);
            const bool q_ok = qx * qx + qy * qy > 1.e-16;   // kernel_iq.c:754
            on = q_ok;
            const float fx = (float)qx, fy = (float)qy;
            const float rn = q_ok ? rsqrt_f(fmaf(fx, fx, fy * fy)) : 1.0f;
            const float ux = fx * rn, uy = fy * rn;
            const float uu = q_ok ? fmaf(ux, ux, uy * uy) : 1.0f;
            const float ax[9] = {sc[12], sc[13], sc[14], sc[15], sc[16], sc[17], sc[18], sc[19], sc[20]};
            double eff[NX][NM1];
#pragma unroll
            for (int k = 0; k < M::NMAG; k++) {
                const SAS_CONST float *m = sc + 24 + 8 * k;
                double sld;
                __builtin_memcpy(&sld, (const float *)m, sizeof sld);
                const float mx = m[2], my = m[3], mz = m[4];
                const float kk = fmaf(mx, ux, my * uy) / uu;
                const float px = fmaf(-kk, ux, mx), py = fmaf(-kk, uy, my), pz = mz;
                const float pol = fmaf(ax[0], px, fmaf(ax[1], py, ax[2] * pz));
                const float flip = fmaf(ax[3], px, ax[4] * py);
                const float perp = fmaf(ax[6], px, fmaf(ax[7], py, ax[8] * pz));
                const double e[6] = {sld - pol, flip, flip, sld + pol, -perp, perp};
#pragma unroll
                for (int x = 0; x < NX; x++) eff[x][k] = q_ok ? e[x] : 0.0;
            }
#pragma unroll
            for (int x = 0; x < NX; x++) {
                active[x] = sc[6 + x] != 0.0f;
                xsw[x] = q_ok ? sc[x] : 0.0f;
                if constexpr (DIRECT) {
#pragma unroll
                    for (int k = 0; k < NM1; k++) effd[x][k] = eff[x][k];
                } else {
                    M::contrasts(values + 2, eff[x], con[x]);
                }
            }
        }
    }

    // sum over cross sections of w_xs * F_xs^2 for amplitude basis g
    __device__ __forceinline__ float f2(const float *g) const
    {
        if (!on) return 0.0f;    // also when the geometry is not finite at q = 0
        float s = 0.0f;
#pragma unroll
        for (int x = 0; x < NX; x++) {
            if (active[x]) {
                float F = 0.0f;
#pragma unroll
                for (int k = 0; k < M::NCON; k++) F = fmaf(con[x][k], g[k], F);
                s = fmaf(xsw[x] * F, F, s);
            }
        }
        return s;
    }
};

// Amplitude basis of one 2D record at (qx, qy), with q rotated into the
// particle frame in FP64.
template <class M>
__device__ __forceinline__ void geom_at(const float *r, double qx, double qy, double qmag, float *g)
{
    QRot q{qmag, 0.0, 0.0, qmag};
    if constexpr (M::ORIENTED) {
        const float *rot = r + M::REC2D;
        q.qa = fma(get_d(rot, 0), qx, get_d(rot, 2) * qy);
        q.qb = fma(get_d(rot, 4), qx, get_d(rot, 6) * qy);
        q.qc = fma(get_d(rot, 8), qx, get_d(rot, 10) * qy);
    }
    M::geom2d(q, r, g);
}

// One lane's q of a 2D kernel, in FP64 and FP32, and the wave's largest |q|.
struct Lane2D {
    double qx, qy, qmag;
    float fx, fy, fmag, wave_qmax;
    __device__ __forceinline__ void set(bool ok, double x, double y)
    {
        qx = ok ? x : 0.0;
        qy = ok ? y : 0.0;
        qmag = sqrt(qx * qx + qy * qy);
        fx = (float)qx;
        fy = (float)qy;
        fmag = (float)qmag;
        float m = fmag;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        wave_qmax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(m)));
    }
};

// w * sum over cross sections of w_xs F_xs^2 for one 2D record (rec points
// at its first float4; the wave reads it with uniform loads): the FP32 path
// when the model has one and every phase of the wave is within XF32_2D,
// else FP64 rotation and arguments.
template <class M, bool MAG, class P>
__device__ __forceinline__ float point2d(P rec, const Lane2D &L, const Contrast<M, MAG> &con)
{
    if constexpr (is_direct<M>::value) {
        // I(|q|) at the record's parameters; magnetic: once per active cross
        // section with its effective slds in place
        constexpr int RD = rv2d<M>();
        float r[4 * RD];
#pragma unroll
        for (int v = 0; v < RD; v++) {
            const float4 x = ld4(rec + v);
            r[4 * v] = x.x; r[4 * v + 1] = x.y; r[4 * v + 2] = x.z; r[4 * v + 3] = x.w;
        }
        double p[M::NPARS];
        const double w = M::load(r, p);
        if constexpr (!MAG) {
            return (float)(w * M::iq(L.qmag, p));
        } else {
            if (!con.on) return 0.0f;
            double s = 0.0;
#pragma unroll
            for (int x = 0; x < Contrast<M, MAG>::NX; x++) {
                if (con.active[x]) {
#pragma unroll
                    for (int k = 0; k < M::NMAG; k++) p[M::sld_param(k)] = con.effd[x][k];
                    s += (double)con.xsw[x] * M::iq(L.qmag, p);
                }
            }
            return (float)(w * s);
        }
    }
    float g[M::NCON];
    if constexpr (has_f32_2d<M>::value) {
        constexpr int RF = rv2f<M>();
        float f[4 * RF];
#pragma unroll
        for (int v = 0; v < RF; v++) {
            const float4 x = ld4(rec + rv2d<M>() + v);
            f[4 * v] = x.x; f[4 * v + 1] = x.y; f[4 * v + 2] = x.z; f[4 * v + 3] = x.w;
        }
        if (L.wave_qmax * f[M::XS2F] <= XF32_2D) {
            QRotF q{L.fmag, 0.0f, 0.0f, L.fmag};
            if constexpr (M::ORIENTED) {
                q.qa = fmaf(f[0], L.fx, f[1] * L.fy);
                q.qb = fmaf(f[2], L.fx, f[3] * L.fy);
                q.qc = fmaf(f[4], L.fx, f[5] * L.fy);
            }
            M::geom2f(q, f, g);
            return f[M::A2F] * con.f2(g);
        }
    }
    constexpr int RD = rv2d<M>();
    float r[4 * RD];
#pragma unroll
    for (int v = 0; v < RD; v++) {
        const float4 x = ld4(rec + v);
        r[4 * v] = x.x; r[4 * v + 1] = x.y; r[4 * v + 2] = x.z; r[4 * v + 3] = x.w;
    }
    geom_at<M>(r, L.qx, L.qy, L.qmag, g);
    return r[M::A2D] * con.f2(g);
}

// Points summed in FP32 between compensated adds (TwoSum) of the block sums:
// the block's FP32 rounding stays below 16 ulp of its terms.
constexpr int SUM_BLOCK = 16;

template <class M, bool MAG>
__global__ __launch_bounds__(EVAL_THREADS) void k_iq2d(
    int nq, const double *__restrict__ q64,
    const double *__restrict__ values, const float *__restrict__ spin, const float4 *__restrict__ recs,
    const int *__restrict__ npts_p, int nsplit, double *__restrict__ partial)
{
    constexpr int RV = rv2<M>();
    const int iq = blockIdx.x * EVAL_THREADS + threadIdx.x;
    const int split = blockIdx.y;
    int s0, s1;
    split_range(*npts_p, nsplit, split, s0, s1);
    const bool ok = iq < nq;
    Lane2D L;
    L.set(ok, ok ? q64[2 * iq] : 0.0, ok ? q64[2 * iq + 1] : 0.0);
    Contrast<M, MAG> con;
    con.setup(values, spin, L.qx, L.qy);
    const SAS_CONST float4 *rc = as_const(recs);
    KSum acc;
    for (int b0 = s0; b0 < s1; b0 += SUM_BLOCK) {
        const int b1 = min(s1, b0 + SUM_BLOCK);
        float blk = 0.0f;
        for (int j = b0; j < b1; j++) blk += point2d<M, MAG>(rc + (size_t)j * RV, L, con);
        acc.add(blk);
    }
    if (ok) partial[(size_t)iq * nsplit + split] = acc.value();
}

}  // namespace sas

// ===================================================================== separable cylinder
// For the cylinder 1D kernel the amplitude factorises,
//   f(q, a; R, L) = J(q R sin a) S(q L/2 cos a),  J = 2J1(x)/x, S = sin(x)/x,
// and a mesh that varies only radius and length is a product grid
// (R_i, L_j) with weights w_ij = wr_i wl_j wo and s_ij = drho pi R_i^2 L_j,
// so every accepted term splits into a radius factor and a length factor:
//   F  term = (wr_i R_i^2 J_i) (wl_j L_j S_j) * c1,
//   F2 term = (wr_i R_i^4 J_i^2) (wl_j L_j^2 S_j^2) * c2.
// Acceptance (VALID and w_ij > cutoff) is monotone in wl_j for a fixed row,
// so with the columns sorted by weight (descending) every row accepts a
// prefix of them, k_i columns.  With the rows sorted by k_i the per-node
// sums are one pass over the sorted columns with running prefix sums
//   P1(k) = sum_{m<k} v1_m S_m,   P2(k) = sum_{m<k} v2_m S_m^2,
// adding u1_i J_i P1(k_i) and u2_i J_i^2 P2(k_i) as each row's prefix is
// complete: NL sinc and NR J1 evaluations per (q, node) instead of NR NL of
// each ("prefix" mode; a mesh with no pruning is the case k_i = NL).
// When the accepted points are not row prefixes (a pd_start/pd_stop piece
// of the mesh, negative weights) the sums are bilinear forms J^T A S over
// tables with zeros for rejected points ("bilinear" mode), evaluated eight
// rows at a time.  Same terms as the generic kernel, summed in another
// order.  Tables are read with wave-uniform loads.
namespace sas {
namespace sep {

constexpr int NRMAX = 64;    // radius and length points per set (either axis)
constexpr int RCHUNK = 8;    // bilinear mode: rows per pass

// Moment series (rotation mode).  With J(x) = 2 J1(x)/x = sum_k a_k x^2k and
// the radius grid R_i <= Rmax, r_i = R_i / Rmax,
//   U1(qs) = sum_i u1_i J(qs R_i)   = sum_k [a_k sum_i u1_i r_i^2k] z^k,
//   U2(qs) = sum_i u2_i J(qs R_i)^2 = sum_k [b_k sum_i u2_i r_i^2k] z^k,
// z = (qs Rmax)^2, b = a * a (Cauchy product); likewise the length factors
// with sin(y)/y and its square.  Each factor becomes one FP64 polynomial in z
// whose coefficients are computed once per set; it is used while every
// argument of the wave is <= SER_X, where the truncation and the
// cancellation among the terms stay below 2e-9 of the factor.
constexpr int SER_NU1 = 16, SER_NU2 = 30, SER_NT1 = 16, SER_NT2 = 26;
constexpr int SER_U1 = 2, SER_U2 = SER_U1 + SER_NU1, SER_T1 = SER_U2 + SER_NU2, SER_T2 = SER_T1 + SER_NT1;
constexpr int SER_STRIDE = 96;   // doubles per set: [Rmax, Ymax, U1 | U2 | T1 | T2 coefficients]
constexpr int SER_MIN_POINTS = 4;   // fewer grid points: the per-point sums are cheaper
constexpr double SER_X = 6.0;
static_assert(SER_T2 + SER_NT2 <= SER_STRIDE, "series table overflows its stride");

// Per-set tables, set s at ROW/COL/GRP + s*NRMAX, AT/BT + s*nlp*NRMAX,
// dims + 4*s, SER + s*SER_STRIDE.  Rotation mode: ROW/COL in grid order
// with the grid step in .w, no groups.  Prefix mode: ROW[r] = (R, u1, u2, 0) for the rows that accept
// at least one column, by ascending accepted count; COL[m] = (L/2, v1, v2, 0)
// by descending weight; GRP[g] = (columns, rows) at the end of the g-th run
// of rows with equal accepted count.  Bilinear mode: ROW[i].x = R_i,
// COL[j].x = L_j/2 in mesh order.
struct Tables {
    double *SER;     // rotation mode: moment series (Rmax and Ymax +inf where unused)
    float4 *ROW;
    float4 *COL;
    float4 *ROWB;    // rotation mode: (1/R, u1 R^-3/2, u2 R^-3, 1 in entry 0 when usable)
    float4 *COLB;    // rotation mode: (2/L, v1 2/L, v2 (2/L)^2, 1 in entry 0 when usable)
    int2 *GRP;
    float *AT;       // bilinear mode [nlp][NRMAX]: w * 1e-2 * drho * V * pi/4 at (row, column)
    float *BT;       // bilinear mode [nlp][NRMAX]: w * 1e-4 * (drho * V)^2 * pi/4
    int *dims;       // NR, NL, mode (3 interval, 2 rotation, 1 prefix, 0 bilinear), number of row groups
    double *G64;     // [2][NRMAX] per set: the radius and the length grid in FP64, mesh order
    int nlp;         // columns per set in AT/BT (>= every set's NL)
};

inline size_t tables_bytes(int nsets, int nlp)
{
    return (size_t)nsets * (SER_STRIDE * sizeof(double) +
                            NRMAX * (4 * sizeof(float4) + sizeof(int2) + 2 * sizeof(float) * (size_t)nlp +
                                     2 * sizeof(double)) +
                            4 * sizeof(int)) + 512;
}

// Carve the tables of nsets sets out of one buffer of tables_bytes(nsets, nlp).
inline Tables carve(void *base, int nsets, int nlp)
{
    Tables t;
    const size_t n = (size_t)nsets * NRMAX;
    t.SER = (double *)base;
    t.ROW = (float4 *)(t.SER + (size_t)nsets * SER_STRIDE);
    t.COL = t.ROW + n;
    t.ROWB = t.COL + n;
    t.COLB = t.ROWB + n;
    t.GRP = (int2 *)(t.COLB + n);
    t.AT = (float *)(t.GRP + n);
    t.BT = t.AT + n * nlp;
    t.dims = (int *)(t.BT + n * nlp);
    // FP64 grids after the dims, on an 8-byte boundary
    t.G64 = (double *)(((uintptr_t)(t.dims + 4 * (size_t)nsets) + 7) & ~(uintptr_t)7);
    t.nlp = nlp;
    return t;
}

// Power series coefficients, evaluated at compile time: 2 J1(x)/x =
// sum_k a_k x^2k with a_k = (-1)^k / (4^k k! (k+1)!); its square (Cauchy
// product); sin(y)/y = sum_k (-1)^k y^2k / (2k+1)!; (sin(y)/y)^2 =
// (1 - cos 2y)/(2 y^2) = sum_k (-1)^k 2^(2k+1) y^2k / (2k+2)!.
struct SerCoef {
    double j1c[SER_NU2], j1c_sq[SER_NU2], sinc[SER_NT2], sinc_sq[SER_NT2];
};
constexpr SerCoef make_ser_coef()
{
    SerCoef c{};
    double a = 1.0;
    for (int k = 0; k < SER_NU2; k++) {
        if (k > 0) a *= -1.0 / (4.0 * k * (k + 1));
        c.j1c[k] = a;
    }
    for (int k = 0; k < SER_NU2; k++) {
        double b = 0.0;
        for (int j = 0; j <= k; j++) b += c.j1c[j] * c.j1c[k - j];
        c.j1c_sq[k] = b;
    }
    double s = 1.0, t = 1.0;
    for (int k = 0; k < SER_NT2; k++) {
        if (k > 0) {
            s *= -1.0 / ((2.0 * k) * (2.0 * k + 1.0));
            t *= -4.0 / ((2.0 * k + 1.0) * (2.0 * k + 2.0));
        }
        c.sinc[k] = s;
        c.sinc_sq[k] = t;
    }
    return c;
}
__constant__ constexpr SerCoef SER_COEF = make_ser_coef();

// One workgroup per set: find the radius and length levels (staged in LDS),
// walk the set's mesh steps (FP64 normalisation sums as k_expand,
// acceptance mask and the bilinear tables), then pick the mode: rotation
// directly when every point is accepted on equally spaced grids, else from
// the prefix, interval and group structure of the acceptance mask; write the
// tables of that mode, and for the rotation mode the scaled tables and the
// moment series.  (Templated so that both translation units can
// instantiate it.)
template <int VARIANT = 0>
__device__ __forceinline__ void sep_expand(const ExpandArgs &a, const Tables &t)
{
    constexpr int MP = Cylinder::MAX_PD, NP = Cylinder::NPARS;
    const int set = blockIdx.x, nlp = t.nlp;
    const int32_t *det = a.det + (size_t)set * a.ndet;
    const double *values = a.values + (a.voff ? a.voff[set] : 0);
    float4 *ROW = t.ROW + (size_t)set * NRMAX, *COL = t.COL + (size_t)set * NRMAX;
    int2 *GRP = t.GRP + (size_t)set * NRMAX;
    float *AT = t.AT + (size_t)set * NRMAX * nlp, *BT = t.BT + (size_t)set * NRMAX * nlp;
    const int32_t *pd_par = det, *pd_len = det + MP, *pd_off = det + 2 * MP, *pd_stride = det + 3 * MP;
    const int num_eval = det[4 * MP];
    const int num_weights = det[4 * MP + 1];
    const int pd_start = a.pd_start;
    const int pd_stop = a.pd_stop < 0 ? num_eval : a.pd_stop;
    const double *pd_value = values + nvalues<Cylinder>();
    const double *pd_weight = pd_value + num_weights;
    int kr = 0, kl = 0;
    for (int k = 0; k < MP; k++) {
        if (pd_par[k] == 2) kr = k;
        if (pd_par[k] == 3) kl = k;
    }
    const int NR = pd_len[kr], NL = pd_len[kl];
    const int tid = threadIdx.x;
    // the radius and length levels in LDS: the loops below walk them
    // serially, and each step from device memory would wait out its latency
    __shared__ double grid[4][NRMAX];
    if (tid < NR) {
        grid[0][tid] = pd_value[pd_off[kr] + tid];
        grid[1][tid] = pd_weight[pd_off[kr] + tid];
    } else if (tid >= NRMAX && tid - NRMAX < NL) {
        grid[2][tid - NRMAX] = pd_value[pd_off[kl] + tid - NRMAX];
        grid[3][tid - NRMAX] = pd_weight[pd_off[kl] + tid - NRMAX];
    }
    const double *rv = grid[0], *rw = grid[1], *lv = grid[2], *lw = grid[3];
    if (tid < NR) t.G64[(size_t)set * 2 * NRMAX + tid] = pd_value[pd_off[kr] + tid];
    else if (tid >= NRMAX && tid - NRMAX < NL) t.G64[(size_t)set * 2 * NRMAX + tid] = pd_value[pd_off[kl] + tid - NRMAX];
    __shared__ unsigned char acc[NRMAX * NRMAX];
    __shared__ int rowk[NRMAX], roworder[NRMAX], colorder[NRMAX], bad, nzero, ngroups, mode;
    __shared__ float steps[2];
    if (a.queue && set == 0 && tid < WORK_QUEUES) a.queue[tid * WORK_QUEUE_STRIDE] = 0;
    __shared__ int nacc, notap, nowhole;
    if (tid == 0) {
        bad = 0;
        nacc = 0;
        notap = 0;
        // the whole mesh, no pruning, no effective radius (checked per point below)
        nowhole = !(pd_start == 0 && pd_stop == num_eval && a.cutoff <= 0.0 && a.reff_mode == 0);
    }
    __syncthreads();
    // equally spaced grids (every distribution of the reference is a linspace)
    const double dr = NR > 1 ? (rv[NR - 1] - rv[0]) / (NR - 1) : 0.0;
    const double dl = NL > 1 ? (lv[NL - 1] - lv[0]) / (NL - 1) : 0.0;
    constexpr double WMIN = 1e-150;    // products of two such weights stay positive
    if (tid < NR) {
        if (!(fabs(rv[tid] - (rv[0] + tid * dr)) <= 1e-9 * fabs(rv[NR - 1]))) notap = 1;
        if (!(rv[tid] >= 0.0 && rw[tid] >= WMIN)) nowhole = 1;
    } else if (tid >= NRMAX && tid - NRMAX < NL) {
        const int j = tid - NRMAX;
        if (!(fabs(lv[j] - (lv[0] + j * dl)) <= 1e-9 * fabs(lv[NL - 1]))) notap = 1;
        if (!(lv[j] >= 0.0 && lw[j] >= WMIN)) nowhole = 1;
    } else if (tid >= 2 * NRMAX && tid - 2 * NRMAX < MP) {
        const int k = tid - 2 * NRMAX;
        if (k != kr && k != kl && !(pd_weight[pd_off[k]] >= WMIN)) nowhole = 1;
    }
    __syncthreads();
    // Every point of an equally spaced product mesh valid and accepted: the
    // rotation mode needs neither the acceptance mask nor the bilinear
    // tables, and the weight sums factorise over the two grids (summed in
    // another order than the walk below, which the other modes take).
    const bool whole = nowhole == 0 && notap == 0;
    constexpr int NW = EXPAND_THREADS / 64;
    __shared__ double red[3][NW];
    double base[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) base[i] = values[2 + i];
    if (!whole) {
    for (int i = tid; i < NRMAX * nlp; i += EXPAND_THREADS) {
        AT[i] = 0.0f;
        BT[i] = 0.0f;
    }
    for (int i = tid; i < NRMAX * NRMAX; i += EXPAND_THREADS) acc[i] = 0;
    __syncthreads();
    double acc_w = 0.0, acc_v = 0.0, acc_r = 0.0;
    int nmine = 0;
    // the levels other than radius and length have one point each (a
    // separable mesh): their values and weights once, the grids from LDS
    double pfix[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) pfix[i] = base[i];
    double wfix[MP];
#pragma unroll
    for (int k = 0; k < MP; k++) {
        wfix[k] = pd_weight[pd_off[k]];
        if (k == kr || k == kl) continue;
        const int par = pd_par[k];
#pragma unroll
        for (int i = 0; i < NP; i++)
            if (i == par) pfix[i] = pd_value[pd_off[k]];
    }
    const int sr = pd_stride[kr], sl = pd_stride[kl];
    for (int step = pd_start + tid; step < pd_stop; step += EXPAND_THREADS) {
        const int ir = (step / sr) % NR, il = (step / sl) % NL;
        double p[NP];
#pragma unroll
        for (int i = 0; i < NP; i++) p[i] = pfix[i];
        p[2] = rv[ir];
        p[3] = lv[il];
        // the product in mesh order (outermost level first), as k_expand
        double w0 = 1.0;
#pragma unroll
        for (int k = MP - 1; k >= 0; k--) w0 = (k == kr ? rw[ir] : k == kl ? lw[il] : wfix[k]) * w0;
        if (Cylinder::valid(p) && w0 > a.cutoff) {
            const double vol = Cylinder::volume(p);
            acc_w += w0;
            acc_v += w0 * vol;
            if (a.reff_mode != 0) acc_r += w0 * Cylinder::reff(a.reff_mode, p);
            const double s = (p[0] - p[1]) * vol;
            AT[il * NRMAX + ir] = (float)(w0 * 1e-2 * s * PI_4_D);
            BT[il * NRMAX + ir] = (float)(w0 * 1e-4 * s * s * PI_4_D);
            acc[ir * NRMAX + il] = 1;
            nmine++;
        }
    }
    if (nmine) atomicAdd(&nacc, nmine);
    // weight sums: within each wave by shuffles, then over the waves
    for (int off = 32; off > 0; off >>= 1) {
        acc_w += __shfl_down(acc_w, off);
        acc_v += __shfl_down(acc_v, off);
        acc_r += __shfl_down(acc_r, off);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = acc_w;
        red[1][tid >> 6] = acc_v;
        red[2][tid >> 6] = acc_r;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < NW; w++) {
            red[0][0] += red[0][w];
            red[1][0] += red[1][w];
            red[2][0] += red[2][w];
        }
    }
    } else if (tid == 0) {
        double sr_w = 0.0, sr_v = 0.0, sl_w = 0.0, sl_v = 0.0;
        for (int i = 0; i < NR; i++) {
            sr_w += rw[i];
            sr_v += rw[i] * (rv[i] * rv[i]);
        }
        for (int j = 0; j < NL; j++) {
            sl_w += lw[j];
            sl_v += lw[j] * lv[j];
        }
        double wo = 1.0;
        for (int k = MP - 1; k >= 0; k--)
            if (k != kr && k != kl) wo = pd_weight[pd_off[k]] * wo;
        red[0][0] = sr_w * sl_w * wo;
        red[1][0] = PI_D * sr_v * sl_v * wo;
        red[2][0] = 0.0;
        nacc = NR * NL;
    }
    __syncthreads();
    // every point accepted on equally spaced grids: rotation mode, without
    // the orders and groups below (a serial walk of the grids each)
    const bool fast = nacc == NR * NL && notap == 0;
    if (!fast) {
        // columns by descending weight (invalid lengths last), ties by index
        if (tid < NL) {
            const double key = lv[tid] >= 0.0 ? lw[tid] : -__builtin_inf();
            int rank = 0;
            for (int j = 0; j < NL; j++) {
                const double kj = lv[j] >= 0.0 ? lw[j] : -__builtin_inf();
                rank += (kj > key || (kj == key && j < tid)) ? 1 : 0;
            }
            colorder[rank] = tid;
        }
        __syncthreads();
        // accepted columns of each row, and whether they are a prefix of colorder
        if (tid < NR) {
            int k = 0;
            for (int j = 0; j < NL; j++) k += acc[tid * NRMAX + j];
            rowk[tid] = k;
            for (int m = 0; m < NL; m++)
                if ((acc[tid * NRMAX + colorder[m]] != 0) != (m < k)) bad = 1;
        }
        __syncthreads();
        if (tid < NR) {
            const int k = rowk[tid];
            int rank = 0;
            for (int i = 0; i < NR; i++) rank += (rowk[i] < k || (rowk[i] == k && i < tid)) ? 1 : 0;
            roworder[rank] = tid;
        }
        __syncthreads();
    }
    const bool prefix = bad == 0;
    if (tid == 0 && fast) {
        nzero = 0;
        ngroups = 1;
        GRP[0] = make_int2(NL, NR);
        steps[0] = (float)dr;
        steps[1] = (float)(0.5 * dl);
        mode = 2;
    } else if (tid == 0) {
        // rows accepting nothing drop out; runs of equal accepted count form groups
        int z = 0;
        while (z < NR && rowk[roworder[z]] == 0) z++;
        int g = 0;
        if (prefix) {
            for (int m = z; m < NR; m++) {
                const int k = rowk[roworder[m]];
                if (m + 1 == NR || rowk[roworder[m + 1]] != k) GRP[g++] = make_int2(k, m + 1 - z);
            }
        }
        nzero = z;
        ngroups = g;
        int md = prefix ? 1 : 0;
        const bool ap = prefix && notap == 0;
        if (ap) {
            steps[0] = (float)dr;
            steps[1] = (float)(0.5 * dl);
        }
        if (ap && z == 0 && g == 1 && rowk[roworder[0]] == NL) {
            md = 2;     // every point accepted: rotation mode
        } else if (ap && g >= 1 && g < NRMAX) {
            // interval mode: the first k columns of the weight order are a
            // grid interval growing outward from colorder[0], and the rows of
            // each group (and the rows accepting nothing, first) are taken
            // inward from the two ends of the row grid
            const int kmax = GRP[g - 1].x, c0 = colorder[0];
            bool iv = true;
            int lo = c0, hi = c0;
            for (int m = 1; m < kmax && iv; m++) {
                const int j = colorder[m];
                if (j == lo - 1) lo = j;
                else if (j == hi + 1) hi = j;
                else iv = false;
            }
            int rl = 0, rr = NR - 1, rl0 = 0, rr0 = NR - 1, kprev = 0, t = 0;
            for (int gi = -1; gi < g && iv; gi++) {
                const int end = gi < 0 ? z : z + GRP[gi].y;      // rows of this group: roworder[t..end)
                int nrl = 0;
                while (t < end && roworder[t] == rl) {
                    rl++;
                    nrl++;
                    t++;
                }
                const int nrr = end - t;
                for (int u = 0; u < nrr && iv; u++) iv = roworder[t + u] == rr - nrr + 1 + u;
                rr -= nrr;
                t = end;
                if (gi < 0) {
                    rl0 = rl;
                    rr0 = rr;
                    continue;
                }
                const int k = GRP[gi].x;
                int ncl = 0, ncr = 0;
                for (int m = kprev; m < k; m++) (colorder[m] < c0 ? ncl : ncr)++;
                kprev = k;
                GRP[gi] = make_int2(ncl | (ncr << 16), nrl | (nrr << 16));
            }
            if (iv) {
                GRP[g] = make_int2(c0, rl0 | (rr0 << 16));
                md = 3;
            } else {
                // restore the prefix-mode groups overwritten above
                for (int m = z, gi = 0; m < NR; m++) {
                    const int k = rowk[roworder[m]];
                    if (m + 1 == NR || rowk[roworder[m + 1]] != k) GRP[gi++] = make_int2(k, m + 1 - z);
                }
            }
        }
        mode = md;
    }
    __syncthreads();
    if (tid == 0) {
        a.sums[4 * set + 0] = red[0][0];
        a.sums[4 * set + 1] = red[1][0];
        a.sums[4 * set + 2] = red[1][0];
        a.sums[4 * set + 3] = red[2][0];
        int *d = t.dims + 4 * set;
        d[0] = NR;
        d[1] = NL;
        d[2] = mode;
        d[3] = ngroups;
    }
    // single-point levels multiply in, in mesh order like w0
    double wo = 1.0;
    for (int lev = MP - 1; lev >= 0; lev--)
        if (lev != kr && lev != kl) wo = pd_weight[pd_off[lev]] * wo;
    const double drho = base[0] - base[1];
    const double c1 = wo * 1e-2 * drho * PI_D * PI_4_D;
    const double c2 = wo * 1e-4 * drho * drho * PI_D * PI_D * PI_4_D;
    if (tid < NRMAX) {
        const int m = tid;
        float4 row = make_float4(0.0f, 0.0f, 0.0f, 0.0f), col = row;
        if (prefix) {
            // rotation and interval modes: grid order, grid steps in .w
            const bool rot = mode >= 2;
            if (m < (rot ? NR : NR - nzero)) {
                const int i = rot ? m : roworder[m + nzero];
                const double x = rv[i];
                row = make_float4((float)x, (float)(rw[i] * x * x), (float)(rw[i] * x * x * x * x),
                                  rot ? steps[0] : 0.0f);
            }
            if (m < NL) {
                const int j = rot ? m : colorder[m];
                const double x = lv[j];
                col = make_float4((float)(0.5 * x), (float)(lw[j] * x * c1), (float)(lw[j] * x * x * c2),
                                  rot ? steps[1] : 0.0f);
            }
        } else {
            if (m < NR) row.x = (float)rv[m];
            if (m < NL) col.x = (float)(0.5 * lv[m]);
        }
        ROW[m] = row;
        COL[m] = col;
    }
    // rotation mode: the moment series of the four factors from the same
    // FP32-rounded grid values and weights as ROW and COL (u1 = w R^2,
    // u2 = w R^4; v1 = w L c1, v2 = w L^2 c2): the weights and the powers
    // r^2k of the scaled grids in LDS, then one coefficient per thread.
    // Rotation-mode grids are equally spaced: their extremes are the ends.
    __shared__ double pw[2][SER_NU2][NRMAX];
    __shared__ double wt[4][NRMAX];
    static_assert(SER_NT2 <= SER_NU2, "power table too short for the length series");
    const double r0 = (double)(float)rv[0], r1 = (double)(float)rv[NR - 1];
    const double y0 = (double)(float)(0.5 * lv[0]), y1 = (double)(float)(0.5 * lv[NL - 1]);
    const double rmax = fmax(r0, r1), ymax = fmax(y0, y1);
    const bool rpos = fmin(r0, r1) > 0.0, lpos = fmin(y0, y1) > 0.0;
    const bool ur = mode == 2 && NR >= SER_MIN_POINTS && rmax > 0.0;
    const bool ul = mode == 2 && NL >= SER_MIN_POINTS && ymax > 0.0;
    if (tid < 2 * NRMAX) {
        const int g = tid / NRMAX, i = tid % NRMAX;
        if (i < (g ? NL : NR)) {
            const double x = g ? lv[i] : rv[i], w = g ? lw[i] : rw[i];
            const double a = (double)(float)(g ? w * x * c1 : w * x * x);
            const double b = (double)(float)(g ? w * x * x * c2 : w * x * x * x * x);
            wt[2 * g][i] = a;
            wt[2 * g + 1][i] = b;
            // rotation mode, grids without zero: the scaled tables of the
            // rotation loops (J = sqrt(2) (qs R)^-3/2 j1_mc, sin(y)/y =
            // sin(y) / (qc L/2))
            float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (mode == 2 && (g ? lpos : rpos)) {
                const double r = 1.0 / (double)(float)(g ? 0.5 * x : x);
                e = g ? make_float4((float)r, (float)(a * r), (float)(b * r * r), i == 0 ? 1.0f : 0.0f)
                      : make_float4((float)r, (float)(a * r * sqrt(r)), (float)(b * r * r * r), i == 0 ? 1.0f : 0.0f);
            }
            (g ? t.COLB : t.ROWB)[(size_t)set * NRMAX + i] = e;
            if (g ? ul : ur) {
                const double r = g ? (double)(float)(0.5 * x) / ymax : (double)(float)x / rmax;
                double p = 1.0;
                for (int k = 0; k < SER_NU2; k++) {
                    pw[g][k][i] = p;
                    p *= r * r;
                }
            }
        } else {
            (g ? t.COLB : t.ROWB)[(size_t)set * NRMAX + i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    __syncthreads();
    if (tid < SER_STRIDE) {
        const int c = tid;
        double v = 0.0;
        if (c == 0) {
            v = ur ? rmax : __builtin_inf();
        } else if (c == 1) {
            v = ul ? ymax : __builtin_inf();
        } else if (c < SER_T1 ? ur : (c < SER_T2 + SER_NT2 && ul)) {
            // term k of series U1 (J), U2 (J^2), T1 (sinc) or T2 (sinc^2)
            const int sr = c < SER_U2 ? 0 : c < SER_T1 ? 1 : c < SER_T2 ? 2 : 3;
            const int k = c - (sr == 0 ? SER_U1 : sr == 1 ? SER_U2 : sr == 2 ? SER_T1 : SER_T2);
            const int g = sr >> 1, n = g ? NL : NR;
            const double *u = wt[sr], *pk = pw[g][k];
            double v0 = 0.0, v1 = 0.0;
            int i = 0;
            for (; i + 1 < n; i += 2) {
                v0 = fma(u[i], pk[i], v0);
                v1 = fma(u[i + 1], pk[i + 1], v1);
            }
            if (i < n) v0 = fma(u[i], pk[i], v0);
            v = (v0 + v1) * (sr == 0 ? SER_COEF.j1c[k] : sr == 1 ? SER_COEF.j1c_sq[k] :
                             sr == 2 ? SER_COEF.sinc[k] : SER_COEF.sinc_sq[k]);
        }
        t.SER[(size_t)set * SER_STRIDE + c] = v;
    }
}

template <int VARIANT = 0>
__global__ __launch_bounds__(EXPAND_THREADS) void k_sep_expand(ExpandArgs a, Tables t)
{
    sep_expand<VARIANT>(a, t);
}

// Loads of the rotation-mode tables from host or device memory.
SAS_HD float4 ldv(const float4 *p) { return *p; }
__device__ __forceinline__ float4 ldv(const SAS_CONST float4 *p) { return ld4(p); }

constexpr int RESTART = 8;   // rotation mode: direct sin/cos every RESTART points

// Rotation mode (every mesh point accepted, both grids equally spaced, in
// grid order, grid step in .w): F = (sum_i u1_i J_i)(sum_j v1_j S_j) and
// F^2 = (sum_i u2_i J_i^2)(sum_j v2_j S_j^2).  Per factor and wave: the
// moment series when every argument is small, else sin and cos along the
// grid advanced by rotating through the step angle (the exact product of
// q and the step, reduced like any argument) and recomputed directly every
// RESTART points, S = sin(y)/y by its Taylor series below |y| = 2 and J by
// its polynomial below 2 where some argument of the wave is that small.

// Wave-uniform vote (every lane of the wave) on the device, the value itself on the host.
SAS_HD bool wave_all(bool p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __all(p);
#else
    return p;
#endif
}

// Smallest value over the wave, in a scalar register (every lane of the
// wave active).
__device__ __forceinline__ int wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

// One length point m: (s, c) = sin and cos of its argument, advanced by the
// step rotation (sd, cd) for the next.  MIXED: some argument of the wave
// is below 2 (Taylor series there); the two instantiations keep the series
// out of the loops of waves that never take it (a masked-off block still
// issues).  A run: n <= RESTART points from m0, unrolled when whole.
template <bool MIXED, class P4>
SAS_HD void length_point(P4 COL, int m, float qc, float &s, float &c, float sd, float cd, float &T1, float &T2)
{
    const float4 e = ldv(COL + m);
    const float y = qc * e.x;
    float S = s * rcp_f(y);
    if (MIXED && y < 2.0f) S = sinc_taylor(y);
    T1 = fmaf(e.y, S, T1);
    T2 = fmaf(e.z, S * S, T2);
    const float s2 = fmaf(s, cd, c * sd);
    c = fmaf(c, cd, -s * sd);
    s = s2;
}

template <bool MIXED, class P4>
SAS_HD void length_run(P4 COL, int m0, int n, float qc, float s, float c, float sd, float cd, float &T1,
                       float &T2)
{
    if (n == RESTART) {
#pragma unroll
        for (int t = 0; t < RESTART; t++) length_point<MIXED>(COL, m0 + t, qc, s, c, sd, cd, T1, T2);
    } else {
        for (int t = 0; t < n; t++) length_point<MIXED>(COL, m0 + t, qc, s, c, sd, cd, T1, T2);
    }
}

// Length factors T1 = sum v1 S, T2 = sum v2 S^2 (grid order, step COL[0].w).
template <class P4>
SAS_HD void length_sums(P4 COL, int NL, float qc, float &T1, float &T2)
{
    const float ya = qc * ldv(COL).x, yb = qc * ldv(COL + NL - 1).x;   // grid ends (either order)
    const float ymin = fminf(ya, yb), ymax = fmaxf(ya, yb);
    T1 = 0.0f;
    T2 = 0.0f;
    if (wave_all(ymax < 2.0f)) {
        for (int m = 0; m < NL; m++) {
            const float4 e = ldv(COL + m);
            const float S = sinc_taylor(qc * e.x);
            T1 = fmaf(e.y, S, T1);
            T2 = fmaf(e.z, S * S, T2);
        }
        return;
    }
    const bool mixed = !wave_all(ymin >= 2.0f);
    float sd, cd;
    sincos_prod(qc, ldv(COL).w, sd, cd);
    for (int m0 = 0; m0 < NL; m0 += RESTART) {
        float s, c;
        sincos_prod(qc, ldv(COL + m0).x, s, c);
        const int n = NL - m0 < RESTART ? NL - m0 : RESTART;
        if (mixed) length_run<true>(COL, m0, n, qc, s, c, sd, cd, T1, T2);
        else length_run<false>(COL, m0, n, qc, s, c, sd, cd, T1, T2);
    }
}

// One radius point and a run of them (as for the length; MIXED: some
// argument of the wave is at most 2, polynomial branch there).
template <bool MIXED, class P4>
SAS_HD void radius_point(P4 ROW, int i, float qs, float &s, float &c, float sd, float cd, float &U1, float &U2)
{
    const float4 e = ldv(ROW + i);
    const float x = qs * e.x;
    float J;
    if (MIXED && x <= 2.0f) J = j1c_small(x);
    else J = j1c_big(x, s, c);
    U1 = fmaf(e.y, J, U1);
    U2 = fmaf(e.z, J * J, U2);
    const float s2 = fmaf(s, cd, c * sd);
    c = fmaf(c, cd, -s * sd);
    s = s2;
}

template <bool MIXED, class P4>
SAS_HD void radius_run(P4 ROW, int i0, int n, float qs, float s, float c, float sd, float cd, float &U1,
                       float &U2)
{
    if (n == RESTART) {
#pragma unroll
        for (int t = 0; t < RESTART; t++) radius_point<MIXED>(ROW, i0 + t, qs, s, c, sd, cd, U1, U2);
    } else {
        for (int t = 0; t < n; t++) radius_point<MIXED>(ROW, i0 + t, qs, s, c, sd, cd, U1, U2);
    }
}

// Radius factors U1 = sum u1 J, U2 = sum u2 J^2 (grid order, step ROW[0].w).
template <class P4>
SAS_HD void radius_sums(P4 ROW, int NR, float qs, float &U1, float &U2)
{
    const float xa = qs * ldv(ROW).x, xb = qs * ldv(ROW + NR - 1).x;   // grid ends (either order)
    const float xmin = fminf(xa, xb), xmax = fmaxf(xa, xb);
    U1 = 0.0f;
    U2 = 0.0f;
    if (wave_all(xmax <= 2.0f)) {
        for (int i = 0; i < NR; i++) {
            const float4 e = ldv(ROW + i);
            const float J = j1c_small(qs * e.x);
            U1 = fmaf(e.y, J, U1);
            U2 = fmaf(e.z, J * J, U2);
        }
        return;
    }
    const bool mixed = !wave_all(xmin > 2.0f);
    float sd, cd;
    sincos_prod(qs, ldv(ROW).w, sd, cd);
    for (int i0 = 0; i0 < NR; i0 += RESTART) {
        float s, c;
        sincos_prod(qs, ldv(ROW + i0).x, s, c);
        const int n = NR - i0 < RESTART ? NR - i0 : RESTART;
        if (mixed) radius_run<true>(ROW, i0, n, qs, s, c, sd, cd, U1, U2);
        else radius_run<false>(ROW, i0, n, qs, s, c, sd, cd, U1, U2);
    }
}

// Two series in one z, FP64: P = sum_k<N1 a_k z^k and Q = sum_k<N2 b_k z^k.
// Blocks of 8 coefficients are power sums over the shared powers z..z^7
// (the coefficients, read through the scalar cache, enter as
// multiplicands: one FMA each), and the blocks combine by Horner in z^8.
// Fewer vector registers than one power sum over every power, and no copy
// of each coefficient into vector registers as a plain Horner chain needs.
template <int N, class PD>
__device__ __forceinline__ double ser_blocks(PD c, const double *zp, double z8)
{
    double u = 0.0;
#pragma unroll
    for (int b = (N - 1) / 8; b >= 0; b--) {
        double blk = c[8 * b];
#pragma unroll
        for (int k = 1; k < 8; k++)
            if (8 * b + k < N) blk = fma(c[8 * b + k], zp[k], blk);
        u = b == (N - 1) / 8 ? blk : fma(u, z8, blk);
    }
    return u;
}

template <int N1, int N2, class PD>
__device__ __forceinline__ void ser_pair(PD a, PD b, double z, double &P, double &Q)
{
    double zp[8];
    zp[0] = 1.0;
    zp[1] = z;
#pragma unroll
    for (int k = 2; k < 8; k++) zp[k] = zp[k - 1] * z;
    const double z8 = zp[7] * z;
    P = ser_blocks<N1>(a, zp, z8);
    Q = ser_blocks<N2>(b, zp, z8);
}

// Pure rotation waves (every argument of the wave above 2) on the scaled
// tables: T1 = (1/qc) sum v1' sin y, T2 = (1/qc^2) sum v2' sin^2 y with
// v' = v (2/L)^k, and U1 = sqrt(2) qs^-3/2 sum u1' j1_mc(q, sin x, cos x),
// U2 = 2 qs^-3 sum u2' j1_mc^2 with u' = u R^-3k/2 and q = (1/qs)(1/R):
// no reciprocal or square root per point.  Each run of RESTART points
// starts from an anchor advanced by the rotation through RESTART grid steps
// (sin and cos computed directly for the first point and the two steps).
__device__ __forceinline__ void rotate(float &s, float &c, float sd, float cd)
{
    const float s2 = fmaf(s, cd, c * sd);
    c = fmaf(c, cd, -s * sd);
    s = s2;
}

template <class P4>
__device__ __forceinline__ void length_sums_b(P4 COLB, float y0, float step, int NL, float qc, float &T1, float &T2)
{
    float sd, cd, sa, ca, s0, c0;
    sincos_prod(qc, step, sd, cd);
    sincos_prod(qc, RESTART * step, sa, ca);
    sincos_prod(qc, y0, s0, c0);
    float A = 0.0f, B = 0.0f;
    for (int m0 = 0; m0 < NL; m0 += RESTART) {
        float s = s0, c = c0;
        const int n = NL - m0 < RESTART ? NL - m0 : RESTART;
        if (n == RESTART) {
#pragma unroll
            for (int t = 0; t < RESTART; t++) {
                const float4 e = ldv(COLB + m0 + t);
                A = fmaf(e.y, s, A);
                B = fmaf(e.z, s * s, B);
                rotate(s, c, sd, cd);
            }
            rotate(s0, c0, sa, ca);
        } else {
            for (int t = 0; t < n; t++) {
                const float4 e = ldv(COLB + m0 + t);
                A = fmaf(e.y, s, A);
                B = fmaf(e.z, s * s, B);
                rotate(s, c, sd, cd);
            }
        }
    }
    const float r = rcp_f(qc);
    T1 = A * r;
    T2 = B * (r * r);
}

template <class P4>
__device__ __forceinline__ void radius_sums_b(P4 ROWB, float x0, float step, int NR, float qs, float &U1, float &U2)
{
    float sd, cd, sa, ca, s0, c0;
    sincos_prod(qs, step, sd, cd);
    sincos_prod(qs, RESTART * step, sa, ca);
    sincos_prod(qs, x0, s0, c0);
    const float iq = rcp_f(qs);
    float A = 0.0f, B = 0.0f;
    for (int i0 = 0; i0 < NR; i0 += RESTART) {
        float s = s0, c = c0;
        const int n = NR - i0 < RESTART ? NR - i0 : RESTART;
        if (n == RESTART) {
#pragma unroll
            for (int t = 0; t < RESTART; t++) {
                const float4 e = ldv(ROWB + i0 + t);
                const float j = j1_mc(iq * e.x, s, c);
                A = fmaf(e.y, j, A);
                B = fmaf(e.z, j * j, B);
                rotate(s, c, sd, cd);
            }
            rotate(s0, c0, sa, ca);
        } else {
            for (int t = 0; t < n; t++) {
                const float4 e = ldv(ROWB + i0 + t);
                const float j = j1_mc(iq * e.x, s, c);
                A = fmaf(e.y, j, A);
                B = fmaf(e.z, j * j, B);
                rotate(s, c, sd, cd);
            }
        }
    }
    const float f = iq * sqrt_f(iq);   // qs^-3/2
    U1 = A * (1.41421356f * f);
    U2 = B * (2.0f * f * f);
}

// The rotation-mode sums with the moment series: a factor whose arguments
// are all <= SER_X across the wave (largest argument q Rmax) is one
// polynomial in (q Rmax)^2; a factor whose arguments all exceed 2 runs on
// the scaled tables; otherwise the per-point sums above (ROW, COL).  Rmax
// is +inf where a set has no series, which fails the vote (also at q = 0).
// The rotation tables come as RB, CB (ROWB, COLB), the series table `ser`
// and the grid ends and steps `ends` = (R first, R last, R step, L/2 first,
// L/2 last, L/2 step).
template <class PF4, class PD, class PF>
__device__ __forceinline__ void length_full(const SAS_CONST float4 *COL, PF4 CB, PD ser, PF ends, int NL, float qc,
                                            float &T1, float &T2)
{
    const double yl = (double)qc * ser[1];
    if (wave_all(yl <= SER_X)) {
        double P, Q;
        ser_pair<SER_NT1, SER_NT2>(ser + SER_T1, ser + SER_T2, yl * yl, P, Q);
        T1 = (float)P;
        T2 = (float)Q;
    } else if (ldv(CB).w != 0.0f && wave_all(qc * fminf(ends[3], ends[4]) >= 2.0f)) {
        length_sums_b(CB, ends[3], ends[5], NL, qc, T1, T2);
    } else {
        length_sums(COL, NL, qc, T1, T2);
    }
}

template <class PF4, class PD, class PF>
__device__ __forceinline__ void radius_full(const SAS_CONST float4 *ROW, PF4 RB, PD ser, PF ends, int NR, float qs,
                                            float &U1, float &U2)
{
    const double xl = (double)qs * ser[0];
    if (wave_all(xl <= SER_X)) {
        double P, Q;
        ser_pair<SER_NU1, SER_NU2>(ser + SER_U1, ser + SER_U2, xl * xl, P, Q);
        U1 = (float)P;
        U2 = (float)Q;
    } else if (ldv(RB).w != 0.0f && wave_all(qs * fminf(ends[0], ends[1]) > 2.0f)) {
        radius_sums_b(RB, ends[0], ends[2], NR, qs, U1, U2);
    } else {
        radius_sums(ROW, NR, qs, U1, U2);
    }
}

template <class PF4, class PD, class PF>
__device__ __forceinline__ void rank_one_sums(const SAS_CONST float4 *ROW, const SAS_CONST float4 *COL, PF4 RB,
                                              PF4 CB, PD ser, PF ends, int NR, int NL, float q, float sn, float cs,
                                              float &f1, float &f2)
{
    float T1, T2, U1, U2;
    length_full(COL, CB, ser, ends, NL, fabsf(q * cs), T1, T2);
    radius_full(ROW, RB, ser, ends, NR, fabsf(q * sn), U1, U2);
    f1 = U1 * T1;
    f2 = U2 * T2;
}

SAS_HD int2 ldv2(const int2 *p) { return *p; }
__device__ __forceinline__ int2 ldv2(const SAS_CONST int2 *p) { return ld2(p); }

// One element of a rotation stream along a grid: sin/cos at the element
// (recomputed directly every RESTART elements of the stream, else carried
// by rotation), then advanced by the step in direction dir (+1 or -1).
struct Stream {
    float s = 0.0f, c = 1.0f;
    int n = 0;
    SAS_HD void at(float qa, float x, float sd, float cd, int dir, float &so, float &co)
    {
        if ((n & (RESTART - 1)) == 0) sincos_prod(qa, x, s, c);
        so = s;
        co = c;
        const float sdd = dir > 0 ? sd : -sd;
        const float s2 = fmaf(s, cd, c * sdd);
        c = fmaf(c, cd, -s * sdd);
        s = s2;
        n++;
    }
};

// Interval mode (pruned mesh, equally spaced grids, prefixes of the weight
// order are grid intervals): groups in ascending accepted count; group g
// extends the column interval by ncl columns to the left and ncr to the
// right of the current interval (prefix sums P1, P2 grow), then adds its
// rows, nrl taken inward from the left end and nrr from the right end,
// with J_i P(k_g).  GRP[g] = (ncl | ncr << 16, nrl | nrr << 16), GRP[G] =
// (first column, first left row | first right row << 16).  Four rotation
// streams; series below argument 2 as in rotation mode.
// Argument regimes of a grid at one (q, node), voted over the wave:
// every argument small (series only), every argument large (rotation
// only), or mixed (both, per lane).
enum { ARG_SMALL = 0, ARG_LARGE = 1, ARG_MIXED = 2 };

SAS_HD int regime(float a, float b, float bound)
{
    if (wave_all(fmaxf(a, b) <= bound)) return ARG_SMALL;
    if (wave_all(fminf(a, b) > bound)) return ARG_LARGE;
    return ARG_MIXED;
}

template <int CM, int RM, class P4, class PG>
SAS_HD void interval_sums_r(P4 ROW, P4 COL, PG GRP, int NR, int NL, int G, float qc, float qs, float &f1,
                            float &f2)
{
    float sdc = 0.0f, cdc = 1.0f, sdr = 0.0f, cdr = 1.0f;
    if (CM != ARG_SMALL) sincos_prod(qc, ldv(COL).w, sdc, cdc);
    if (RM != ARG_SMALL) sincos_prod(qs, ldv(ROW).w, sdr, cdr);
    const int2 st = ldv2(GRP + G);
    int cl = st.x - 1, cr = st.x, rl = st.y & 0xffff, rr = st.y >> 16;
    Stream CL, CR, RL, RR;
    float P1 = 0.0f, P2 = 0.0f, F1 = 0.0f, F2 = 0.0f;
    auto column = [&](Stream &str, int m, int dir) {
        const float4 e = ldv(COL + m);
        const float y = qc * e.x;
        float S;
        if (CM == ARG_SMALL) {
            S = sinc_taylor(y);
        } else {
            float sy, cy;
            str.at(qc, e.x, sdc, cdc, dir, sy, cy);
            S = sy * rcp_f(y);
            if (CM == ARG_MIXED && y <= 2.0f) S = sinc_taylor(y);
        }
        P1 = fmaf(e.y, S, P1);
        P2 = fmaf(e.z, S * S, P2);
    };
    auto row = [&](Stream &str, int i, int dir) {
        const float4 e = ldv(ROW + i);
        const float x = qs * e.x;
        float J;
        if (RM == ARG_SMALL) {
            J = j1c_small(x);
        } else {
            float sx, cx;
            str.at(qs, e.x, sdr, cdr, dir, sx, cx);
            if (RM == ARG_MIXED && x <= 2.0f) J = j1c_small(x);
            else J = j1c_big(x, sx, cx);
        }
        F1 = fmaf(e.y * J, P1, F1);
        F2 = fmaf(e.z * (J * J), P2, F2);
    };
    for (int g = 0; g < G; g++) {
        const int2 e = ldv2(GRP + g);
        const int ncl = min(e.x & 0xffff, cl + 1), ncr = min(e.x >> 16, NL - cr);
        const int nrl = min(e.y & 0xffff, NR - rl), nrr = min(e.y >> 16, rr + 1);
        for (int t = 0; t < ncl; t++) column(CL, cl - t, -1);
        for (int t = 0; t < ncr; t++) column(CR, cr + t, +1);
        cl -= ncl;
        cr += ncr;
        for (int t = 0; t < nrl; t++) row(RL, rl + t, +1);
        for (int t = 0; t < nrr; t++) row(RR, rr - t, -1);
        rl += nrl;
        rr -= nrr;
    }
    f1 = F1;
    f2 = F2;
}

template <class P4, class PG>
SAS_HD void interval_sums(P4 ROW, P4 COL, PG GRP, int NR, int NL, int G, float q, float sn, float cs, float &f1,
                          float &f2)
{
    const float qc = fabsf(q * cs), qs = fabsf(q * sn);
    const int cm = regime(qc * ldv(COL).x, qc * ldv(COL + NL - 1).x, 2.0f);
    const int rm = regime(qs * ldv(ROW).x, qs * ldv(ROW + NR - 1).x, 2.0f);
#define SAS_IV(C, R) \
    if (cm == C && rm == R) return interval_sums_r<C, R>(ROW, COL, GRP, NR, NL, G, qc, qs, f1, f2);
    SAS_IV(0, 0) SAS_IV(0, 1) SAS_IV(0, 2) SAS_IV(1, 0) SAS_IV(1, 1) SAS_IV(1, 2) SAS_IV(2, 0) SAS_IV(2, 1)
#undef SAS_IV
    interval_sums_r<2, 2>(ROW, COL, GRP, NR, NL, G, qc, qs, f1, f2);
}

// F and F^2 sums of one set at (q, Gauss node) -- sn, cs: sine and cosine of
// the node; the node weight is applied by the caller.  Every lane of the
// wave reads the same table entries (scalar loads: the tables are written
// by k_sep_expand, an earlier launch, and only read here).
// node_sums for a set whose tables are in rotation mode (dims[2] == 2):
// the rank-one sums alone, without the code of the other modes.
__device__ __forceinline__ void node_sums_rot(const Tables &t, int set, float q, float sn, float cs, float &f1,
                                              float &f2)
{
    const SAS_CONST int *dims = as_const(t.dims) + 4 * set;
    const int NR = dims[0], NL = dims[1];
    const SAS_CONST float4 *ROW = as_const(t.ROW) + (size_t)set * NRMAX;
    const SAS_CONST float4 *COL = as_const(t.COL) + (size_t)set * NRMAX;
    const float ends[6] = {ldv(ROW).x, ldv(ROW + NR - 1).x, ldv(ROW).w, ldv(COL).x, ldv(COL + NL - 1).x,
                           ldv(COL).w};
    rank_one_sums(ROW, COL, as_const(t.ROWB) + (size_t)set * NRMAX, as_const(t.COLB) + (size_t)set * NRMAX,
                  as_const(t.SER) + (size_t)set * SER_STRIDE, ends, NR, NL, q, sn, cs, f1, f2);
}

__device__ __forceinline__ void node_sums(const Tables &t, int set, float q, float sn, float cs, float &f1,
                                          float &f2)
{
    const SAS_CONST int *dims = as_const(t.dims) + 4 * set;
    const int NR = dims[0], NL = dims[1];
    const SAS_CONST float4 *ROW = as_const(t.ROW) + (size_t)set * NRMAX;
    const SAS_CONST float4 *COL = as_const(t.COL) + (size_t)set * NRMAX;
    if (dims[2] == 2) {
        const float ends[6] = {ldv(ROW).x, ldv(ROW + NR - 1).x, ldv(ROW).w, ldv(COL).x, ldv(COL + NL - 1).x,
                               ldv(COL).w};
        rank_one_sums(ROW, COL, as_const(t.ROWB) + (size_t)set * NRMAX, as_const(t.COLB) + (size_t)set * NRMAX,
                      as_const(t.SER) + (size_t)set * SER_STRIDE, ends, NR, NL, q, sn, cs, f1, f2);
        return;
    }
    if (dims[2] == 3) {
        interval_sums(ROW, COL, as_const(t.GRP) + (size_t)set * NRMAX, NR, NL, dims[3], q, sn, cs, f1, f2);
        return;
    }
    const float qc = q * cs, qs = q * sn;
    float F1 = 0.0f, F2 = 0.0f;
    if (dims[2]) {
        const SAS_CONST int2 *GRP = as_const(t.GRP) + (size_t)set * NRMAX;
        const int G = dims[3];
        float P1 = 0.0f, P2 = 0.0f;
        int m = 0, r = 0;
        // next table entries are loaded one iteration ahead (entry NRMAX of
        // a set is the next set's first entry or the following table: in bounds)
        float4 cn = ld4(COL), wn = ld4(ROW);
        for (int g = 0; g < G; g++) {
            const int2 e = ld2(GRP + g);
            for (; m < e.x; m++) {
                const float4 c = cn;
                cn = ld4(COL + m + 1);
                const float S = sinc_prod(qc, c.x);
                P1 = fmaf(c.y, S, P1);
                P2 = fmaf(c.z, S * S, P2);
            }
            for (; r < e.y; r++) {
                const float4 w = wn;
                wn = ld4(ROW + r + 1);
                const float J = j1c_f(qs * w.x);
                F1 = fmaf(w.y * J, P1, F1);
                F2 = fmaf(w.z * (J * J), P2, F2);
            }
        }
    } else {
        const SAS_CONST float *AT = as_const(t.AT) + (size_t)set * NRMAX * t.nlp;
        const SAS_CONST float *BT = as_const(t.BT) + (size_t)set * NRMAX * t.nlp;
        for (int i0 = 0; i0 < NR; i0 += RCHUNK) {
            float J[RCHUNK], t1[RCHUNK], t2[RCHUNK];
#pragma unroll
            for (int c = 0; c < RCHUNK; c++) {
                J[c] = i0 + c < NR ? j1c_f(qs * ld4(ROW + i0 + c).x) : 0.0f;
                t1[c] = 0.0f;
                t2[c] = 0.0f;
            }
            for (int j = 0; j < NL; j++) {
                const float S = sinc_prod(qc, ld4(COL + j).x), S2 = S * S;
                const SAS_CONST float *a = AT + j * NRMAX + i0, *b = BT + j * NRMAX + i0;
#pragma unroll
                for (int c = 0; c < RCHUNK; c++) {
                    t1[c] = fmaf(a[c], S, t1[c]);
                    t2[c] = fmaf(b[c], S2, t2[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < RCHUNK; c++) {
                F1 = fmaf(J[c], t1[c], F1);
                F2 = fmaf(J[c] * J[c], t2[c], F2);
            }
        }
    }
    f1 = F1;
    f2 = F2;
}

// Prefix sums along one grid for the adaptive rule: the point terms of
// the radius (R) or length (L) factors in grid order, added while the point
// index is below the lane's prefix n and captured at np (np <= n) before the
// point np is added; the walk ends where every lane's prefix is complete.  Small arguments on every lane take the polynomial,
// else sin and cos are carried along the grid by the step rotation and
// recomputed every RESTART points (as the rotation-mode sums).
template <bool RADIUS, class P4>
__device__ __forceinline__ void grid_prefix(P4 G, int N, float qa, int n, int np, float &A1, float &A2, float &P1,
                                            float &P2)
{
    A1 = A2 = P1 = P2 = 0.0f;
    const float xa = qa * ldv(G).x, xb = qa * ldv(G + N - 1).x;
    const float xmin = fminf(xa, xb), xmax = fmaxf(xa, xb);
    const bool all_small = wave_all(xmax <= 2.0f), mixed = !wave_all(xmin > 2.0f);
    float sd = 0.0f, cd = 1.0f, s = 0.0f, c = 1.0f;
    if (!all_small) sincos_prod(qa, ldv(G).w, sd, cd);
    int i = 0;
    for (; i < N; i++) {
        if (wave_all(i >= n)) break;   // every lane's prefix complete
        const float4 e = ldv(G + i);
        const float x = qa * e.x;
        if (!all_small && (i & (RESTART - 1)) == 0) sincos_prod(qa, e.x, s, c);
        float f;
        if (RADIUS) f = all_small || (mixed && x <= 2.0f) ? j1c_small(x) : j1c_big(x, s, c);
        else f = all_small || (mixed && x < 2.0f) ? sinc_taylor(x) : s * rcp_f(x);
        if (i == np) {
            P1 = A1;
            P2 = A2;
        }
        if (i < n) {
            A1 = fmaf(e.y, f, A1);
            A2 = fmaf(e.z, f * f, A2);
        }
        if (!all_small) {
            const float s2 = fmaf(s, cd, c * sd);
            c = fmaf(c, cd, -s * sd);
            s = s2;
        }
    }
    if (np >= i) {
        P1 = A1;
        P2 = A2;
    }
}

// Adaptive rule (rotation-mode tables, grids ascending): the mesh points
// evaluated with rule k at q form the rectangle of the first a radius and
// b length points (q R < T_k and q L/2 < T_k) less the rectangle of the
// first ap, bp (the smaller rules), so at a node of rule k
//   F = U1(a) T1(b) - U1(ap) T1(bp),  F^2 = U2(a) T2(b) - U2(ap) T2(bp)
// with the prefix sums U(a) = sum_{i<a} u_i J_i and T(b) = sum_{j<b} v_j S_j
// (ap <= a, bp <= b per lane).
__device__ __forceinline__ void prefix_sums(const Tables &t, int set, float q, float sn, float cs, int a, int ap,
                                            int b, int bp, float &f1, float &f2)
{
    const SAS_CONST int *dims = as_const(t.dims) + 4 * set;
    const int NR = dims[0], NL = dims[1];
    const SAS_CONST float4 *ROW = as_const(t.ROW) + (size_t)set * NRMAX;
    const SAS_CONST float4 *COL = as_const(t.COL) + (size_t)set * NRMAX;
    float U1, U2, U1p, U2p, T1, T2, T1p, T2p;
    grid_prefix<true>(ROW, NR, fabsf(q * sn), a, ap, U1, U2, U1p, U2p);
    grid_prefix<false>(COL, NL, fabsf(q * cs), b, bp, T1, T2, T1p, T2p);
    f1 = fmaf(U1, T1, -U1p * T1p);
    f2 = fmaf(U2, T2, -U2p * T2p);
}

// Band sums along one grid: the point terms of the lane's points lo <= i < hi
// (a lane with lo >= hi adds nothing); the walk covers the wave's smallest
// lo to its largest hi, with sin and cos carried as in grid_prefix.
template <bool RADIUS, class P4>
__device__ __forceinline__ void grid_band(P4 G, int N, float qa, int lo, int hi, float &A1, float &A2)
{
    A1 = A2 = 0.0f;
    const int i0 = wave_min(lo);
    const float xa = qa * ldv(G).x, xb = qa * ldv(G + N - 1).x;
    const float xmin = fminf(xa, xb), xmax = fmaxf(xa, xb);
    const bool all_small = wave_all(xmax <= 2.0f), mixed = !wave_all(xmin > 2.0f);
    float sd = 0.0f, cd = 1.0f, s = 0.0f, c = 1.0f;
    if (!all_small) sincos_prod(qa, ldv(G).w, sd, cd);
    for (int i = i0; i < N; i++) {
        if (wave_all(i >= hi)) break;
        const float4 e = ldv(G + i);
        const float x = qa * e.x;
        if (!all_small && ((i - i0) & (RESTART - 1)) == 0) sincos_prod(qa, e.x, s, c);
        float f;
        if (RADIUS) f = all_small || (mixed && x <= 2.0f) ? j1c_small(x) : j1c_big(x, s, c);
        else f = all_small || (mixed && x < 2.0f) ? sinc_taylor(x) : s * rcp_f(x);
        if (i >= lo && i < hi) {
            A1 = fmaf(e.y, f, A1);
            A2 = fmaf(e.z, f * f, A2);
        }
        if (!all_small) {
            const float s2 = fmaf(s, cd, c * sd);
            c = fmaf(c, cd, -s * sd);
            s = s2;
        }
    }
}

// Band sums on the scaled tables (every argument of the wave above 2, as
// length_sums_b and radius_sums_b): sin and cos start directly at the
// wave's first band point i0 (grid value g0) and are carried by the step
// rotation, the run anchors by the rotation through RESTART steps.
template <bool RADIUS, class P4>
__device__ __forceinline__ void band_b(P4 GB, float g0, float step, int i0, int N, float qa, int lo, int hi,
                                       float &A1, float &A2)
{
    float sd, cd, sa, ca, s0, c0;
    sincos_prod(qa, step, sd, cd);
    sincos_prod(qa, RESTART * step, sa, ca);
    sincos_prod(qa, g0, s0, c0);
    const float iq = rcp_f(qa);
    float A = 0.0f, B = 0.0f;
    for (int m0 = i0; m0 < N; m0 += RESTART) {
        if (wave_all(m0 >= hi)) break;
        float s = s0, c = c0;
        const int n = N - m0 < RESTART ? N - m0 : RESTART;
        for (int t = 0; t < n; t++) {
            const int i = m0 + t;
            const float4 e = ldv(GB + i);
            const float f = RADIUS ? j1_mc(iq * e.x, s, c) : s;
            if (i >= lo && i < hi) {
                A = fmaf(e.y, f, A);
                B = fmaf(e.z, f * f, B);
            }
            rotate(s, c, sd, cd);
        }
        rotate(s0, c0, sa, ca);
    }
    if (RADIUS) {
        const float f = iq * sqrt_f(iq);   // qa^-3/2
        A1 = A * (1.41421356f * f);
        A2 = B * (2.0f * f * f);
    } else {
        A1 = A * iq;
        A2 = B * (iq * iq);
    }
}

// The adaptive rule where one grid enters whole on every lane: with a = NR
// and ap in {0, NR} the node's sums are U(NR) times the length band
// [ap ? bp : 0, b); with b = NL and bp in {0, NL} (RADIUS_BAND) U over the
// radius band times T(NL).  The whole factor runs as in rank_one_sums
// (series, scaled tables or the rotation sums).
template <bool RADIUS_BAND>
__device__ __forceinline__ void band_sums(const Tables &t, int set, float q, float sn, float cs, int lo, int hi,
                                          float &f1, float &f2)
{
    const SAS_CONST int *dims = as_const(t.dims) + 4 * set;
    const int NR = dims[0], NL = dims[1];
    const SAS_CONST float4 *ROW = as_const(t.ROW) + (size_t)set * NRMAX;
    const SAS_CONST float4 *COL = as_const(t.COL) + (size_t)set * NRMAX;
    const SAS_CONST float4 *RB = as_const(t.ROWB) + (size_t)set * NRMAX;
    const SAS_CONST float4 *CB = as_const(t.COLB) + (size_t)set * NRMAX;
    const SAS_CONST double *ser = as_const(t.SER) + (size_t)set * SER_STRIDE;
    const float ends[6] = {ldv(ROW).x, ldv(ROW + NR - 1).x, ldv(ROW).w, ldv(COL).x, ldv(COL + NL - 1).x,
                           ldv(COL).w};
    const float qc = fabsf(q * cs), qs = fabsf(q * sn);
    float U1, U2, T1, T2;
    const int i0 = wave_min(lo);
    if (RADIUS_BAND) {
        length_full(COL, CB, ser, ends, NL, qc, T1, T2);
        if (i0 < NR && ldv(RB).w != 0.0f && wave_all(qs * fminf(ends[0], ends[1]) > 2.0f))
            band_b<true>(RB, ldv(ROW + i0).x, ends[2], i0, NR, qs, lo, hi, U1, U2);
        else
            grid_band<true>(ROW, NR, qs, lo, hi, U1, U2);
    } else {
        radius_full(ROW, RB, ser, ends, NR, qs, U1, U2);
        if (i0 < NL && ldv(CB).w != 0.0f && wave_all(qc * fminf(ends[3], ends[4]) >= 2.0f))
            band_b<false>(CB, ldv(COL + i0).x, ends[5], i0, NL, qc, lo, hi, T1, T2);
        else
            grid_band<false>(COL, NL, qc, lo, hi, T1, T2);
    }
    f1 = U1 * T1;
    f2 = U2 * T2;
}

// Number of grid values g[0..n) (ascending) with q * s * g < T, in FP64 as
// the reference's rule choice (q * fmax(radius, 0.5 * length) < T).
__device__ __forceinline__ int rule_count(const double *g, int n, double q, double scale, double T)
{
    int c = 0;
    for (int i = 0; i < n; i++) c += q * (scale * g[i]) < T ? 1 : 0;
    return c;
}

// Single set: one wave per (64-q tile, Gauss node), four nodes per
// workgroup, their four sums combined in wave order through LDS:
// partial[(blk*2 + o)*nq + iq] (blk = the workgroup's node block, o = 0
// for F^2, 1 for F) for k_finalize_sep.
template <int VARIANT = 0>
__global__ __launch_bounds__(256) void k_fq1d_sep(int nq, const float *__restrict__ q32, Tables t, Quad quad,
                                                   double *__restrict__ partial)
{
    __shared__ double red[2][4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int iq = blockIdx.x * 64 + lane;
    const int node = blockIdx.y * 4 + wave;
    double d1 = 0.0, d2 = 0.0;
    if (node < quad.n) {
        const float q = iq < nq ? q32[iq] : 0.0f;
        float f1, f2;
        node_sums(t, 0, q, quad.a[node], quad.b[node], f1, f2);
        const float w = quad.w[node];
        d2 = (double)(w * f2);
        d1 = (double)(w * f1);
    }
    red[0][wave][lane] = d2;
    red[1][wave][lane] = d1;
    __syncthreads();
    if (wave < 2 && iq < nq) {
        const double v = ((red[wave][0][lane] + red[wave][1][lane]) + red[wave][2][lane]) + red[wave][3][lane];
        partial[((size_t)blockIdx.y * 2 + wave) * nq + iq] = v;
    }
}

}  // namespace sep
}  // namespace sas
