// models.h -- the form factors as device functors.
//
// Each model supplies, from its FP64 kernel-parameter vector p[]
// (modelinfo order, no scale/background):
//   valid(p), volume(p), reff(mode, p)       FP64, as the pd loop needs them
//   rec1d(p, w, r)    FP32 record for the 1D <F>,<F^2> kernel, with the mesh
//                     weight and every per-point constant folded in
//   eval1d(q, r, quad, f1, f2)   weighted F and F^2 contributions at q
//   rec2d(p, w, r)    2D record: lengths that enter a phase in FP64 (two
//                     words each), per-layer volumes and w*1e-4 in FP32
//   geom2d(qr, r, g)  amplitude basis g[k] at the rotated q (FP64); F = sum_k con[k]*g[k]
//   contrasts(p, sld, con)       con[k] from the SLD slot values (FP64); p, the
//                                kernel parameters, gives the layer count of
//                                models with vector parameters
// The amplitude of every model is linear in its SLDs, which lets the
// magnetic kernel evaluate the geometry once and reuse it for each spin
// cross section.
//
// Reference formulas (sasmodels/models): sphere.c:1-20, cylinder.c:1-102,
// ellipsoid.c:1-105, core_shell_cylinder.c:1-123, core_shell_sphere.c:1-30
// with lib/core_shell.c:8-30, parallelepiped.c:1-150,
// triaxial_ellipsoid.c:1-160, bcc/fcc/sc_paracrystal.c with
// lib/sphere_form.c, hollow_cylinder.c, vesicle.c, fuzzy_sphere.c,
// core_multi_shell.c, onion.c, core_shell_ellipsoid.c, elliptical_cylinder.c,
// core_shell_bicelle.c.  Parameter order: sphere.py:66-72,
// cylinder.py:135-142, ellipsoid.py:145-152, core_shell_cylinder.py:118-134,
// core_shell_sphere.py:78-82.
#pragma once
#include "sasmath.h"

namespace sas {

// pi and 4*pi/3 with the values the reference's C uses (kernel_header.c:144-169)
constexpr double PI_D = 3.14159265358979323846;
constexpr double PI_4_D = 0.78539816339744830962;
constexpr double FOUR_PI_3_D = 4.18879020478639;

// Quadrature tables on the device.  For the cylinder family the nodes are
// theta = z*pi/4 + pi/4 and the tables hold sin, cos and w*sin; for the
// ellipsoid u = z/2 + 1/2 and the tables hold u^2 and w.  Adaptive
// selection (adaptive.c:11290-11297) picks one of four tables per lane.
// q in the particle frame (FP64): qa, qb across the axis, qc along it; q = |q|.
struct QRot {
    double qa, qb, qc, q;
};

// q in the particle frame in FP32 (the 2D FP32 path).  Models that have it
// define REC2F (FP32 record floats after the rotation rows of oriented
// models), A2F (slot of w*1e-4) and XS2F (slot of the largest length by
// which the geometry multiplies |q|: the phase bound of the record), and
// rec2f/geom2f, the FP32 counterparts of rec2d/geom2d with every length
// divided by 2 pi so that the phases come out in revolutions.
struct QRotF {
    float qa, qb, qc, q;
};
constexpr double INV_TWO_PI_D = 0.15915494309189533577;

struct Quad {
    const float *a;     // sin(theta) | u^2
    const float *b;     // cos(theta) | unused
    const float *w;     // w*sin(theta) | w
    const float *c;     // nested rules: outer-node functions (see the model)
    const float *d;
    const double *g64;  // FP64 models: n nodes each of sin, cos of the outer and inner angles, and w
    int n;              // fixed size, or 0 for adaptive
    int off[4];         // adaptive: offsets of the 20/76/500/5000 tables
};

SAS_HD int adaptive_pick(float qr)
{
    return qr < 10.0f ? 0 : qr < 100.0f ? 1 : qr < 750.0f ? 2 : 3;
}

template <bool FIXED>
__device__ __forceinline__ void quad_range(const Quad &t, float qr, int &off, int &n)
{
    if constexpr (FIXED) {
        off = 0;
        n = t.n;
    } else {
        const int k = adaptive_pick(qr);
        off = k == 0 ? t.off[0] : k == 1 ? t.off[1] : k == 2 ? t.off[2] : t.off[3];
        n = k == 0 ? 20 : k == 1 ? 76 : k == 2 ? 500 : 5000;
    }
}

// Nested rules (adaptive.c:11276-11298 called with outer_n = ADAPTIVE_MAX_OUTER
// for the outer integral and outer_n = n_outer for the inner one): the outer
// rule stops at 500 points and the inner rule may use one size less for every
// size the outer rule went above 20 (20 -> up to 5000, 76 -> 500, 500 -> 76).
// `level` is the outer rule's size index; a fixed rule uses t.n for both.
template <bool FIXED>
__device__ __forceinline__ void quad_nested(const Quad &t, float qr, int cap, int &level, int &off, int &n)
{
    if constexpr (FIXED) {
        level = 0;
        off = 0;
        n = t.n;
    } else {
        const int k = min(adaptive_pick(qr), cap);
        level = k;
        off = k == 0 ? t.off[0] : k == 1 ? t.off[1] : k == 2 ? t.off[2] : t.off[3];
        n = k == 0 ? 20 : k == 1 ? 76 : k == 2 ? 500 : 5000;
    }
}

// ---------------------------------------------------------------- sphere
struct Sphere {
    static constexpr int ID = 0, NPARS = 3, NMAG = 2, MAX_PD = 1, NCON = 1;
    static constexpr bool SHARED_1D = false;   // eval1d_shared: lengths shared by a chunk
    static constexpr bool ORIENTED = false;
    static constexpr bool ABC = false;       // three view angles (theta, phi, psi)
    static constexpr bool FQ = true;         // 1D kernel returns <F> as well as <F^2>
    static constexpr bool FP64 = false;      // 1D evaluation in FP64 (eval1d_d)
    static constexpr bool HOLLOW = false;    // shell volume differs from the form volume
    static constexpr int REC1D = 4, REC2D = 4;
    static constexpr int A2D = 3;   // slot of w*1e-4 in the 2D record
    SAS_HD static int sld_param(int k) { return k; }            // sld, sld_solvent
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(p[2]); }
    __device__ static double reff(int, const double *p) { return p[2]; }
    __device__ static double qr_scale(const double *p) { return p[2]; }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double s = (p[0] - p[1]) * volume(p);
        r[0] = (float)p[2];
        r[1] = (float)(w * 1e-2 * s);
        r[2] = (float)(w * 1e-4 * s * s);
        r[3] = 0.0f;
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &, float &f1, float &f2)
    {
        const float f = sph_j1c_f(q * r[0]);
        f1 = r[1] * f;
        f2 = r[2] * f * f;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[2]);
        r[2] = (float)volume(p);
        r[3] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        g[0] = r[2] * sph_j1c_d(q.q * get_d(r, 0));
    }
    // FP32 record: R/2pi, V, w*1e-4, R
    static constexpr int REC2F = 4, A2F = 2, XS2F = 3;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        f[0] = (float)(p[2] * INV_TWO_PI_D);
        f[1] = (float)volume(p);
        f[2] = (float)(w * 1e-4);
        f[3] = (float)p[2];
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
        const float t = q.q * f[0];
        g[0] = f[1] * sph_j1c_rev2(t * t);
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- cylinder
struct Cylinder {
    static constexpr int ID = 1, NPARS = 6, NMAG = 2, MAX_PD = 4, NCON = 1;
    static constexpr bool SHARED_1D = true;   // eval1d_shared: lengths shared by a chunk
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = false;       // three view angles (theta, phi, psi)
    static constexpr bool FQ = true;         // 1D kernel returns <F> as well as <F^2>
    static constexpr bool FP64 = false;      // 1D evaluation in FP64 (eval1d_d)
    static constexpr bool HOLLOW = false;    // shell volume differs from the form volume
    static constexpr int REC1D = 4, REC2D = 8;
    static constexpr int A2D = 5;   // slot of w*1e-4 in the 2D record
    SAS_HD static int sld_param(int k) { return k; }
    __device__ static bool valid(const double *p) { return p[2] >= 0.0 && p[3] >= 0.0; }
    __device__ static double volume(const double *p) { return PI_D * p[2] * p[2] * p[3]; }
    __device__ static double reff(int mode, const double *p)
    {
        const double r = p[2], l = p[3];
        switch (mode) {
        default:
        case 1: return 0.5 * cbrt(0.75 * r * (2.0 * r * l + (r + l) * (PI_D * r + l)));
        case 2: return cbrt(PI_D * r * r * l / FOUR_PI_3_D);
        case 3: return r;
        case 4: return 0.5 * l;
        case 5: return r < 0.5 * l ? r : 0.5 * l;
        case 6: return r > 0.5 * l ? r : 0.5 * l;
        case 7: return sqrt(r * r + 0.25 * l * l);
        }
    }
    __device__ static double qr_scale(const double *p) { return fmax(p[2], 0.5 * p[3]); }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double s = (p[0] - p[1]) * volume(p);
        r[0] = (float)p[2];
        r[1] = (float)(0.5 * p[3]);
        r[2] = (float)(w * 1e-2 * s * PI_4_D);
        r[3] = (float)(w * 1e-4 * s * s * PI_4_D);
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        int off, n;
        quad_range<FIXED>(t, q * fmaxf(r[0], r[1]), off, n);
        const float qr = q * r[0], qh = q * r[1];
        float a1 = 0.0f, a2 = 0.0f;
        for (int i = 0; i < n; i++) {
            const float f = j1c_f(qr * t.a[off + i]) * sinc_prod(qh, t.b[off + i]);
            const float wf = t.w[off + i] * f;
            a1 += wf;
            a2 = fmaf(wf, f, a2);
        }
        f1 = r[2] * a1;
        f2 = r[3] * a2;
    }
    // Fixed rule, records r0..r1 of a chunk with the same half length r[1]:
    // nodes outer, the length factor once per node.  Adds sum_pts r[2] w f
    // and sum_pts r[3] w f^2 of every node to a1, a2.
    // Nodes in blocks of NB: a record is read once per block instead of once
    // per node; each node's point sums and their order are unchanged.
    template <class ACC>
    __device__ static void eval1d_shared(float q, const float *rec, int rv, int n0, int n1, const Quad &t, ACC &a1,
                                         ACC &a2)
    {
        constexpr int NB = 8;
        const float qh = q * rec[n0 * rv + 1];
        int i0 = 0;
        for (; i0 + NB <= t.n; i0 += NB) {
            float s[NB], w[NB], S[NB], t1[NB], t2[NB];
#pragma unroll
            for (int k = 0; k < NB; k++) {
                s[k] = t.a[i0 + k];
                w[k] = t.w[i0 + k];
                S[k] = sinc_prod(qh, t.b[i0 + k]);
                t1[k] = 0.0f;
                t2[k] = 0.0f;
            }
            for (int j = n0; j < n1; j++) {
                const float *r = rec + j * rv;
                const float qr = q * r[0];
#pragma unroll
                for (int k = 0; k < NB; k++) {
                    const float f = j1c_f(qr * s[k]) * S[k];
                    const float wf = w[k] * f;
                    t1[k] = fmaf(r[2], wf, t1[k]);
                    t2[k] = fmaf(r[3], wf * f, t2[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < NB; k++) {
                a1.add(t1[k]);
                a2.add(t2[k]);
            }
        }
        for (int i = i0; i < t.n; i++) {
            const float s = t.a[i], w = t.w[i];
            const float S = sinc_prod(qh, t.b[i]);
            float t1 = 0.0f, t2 = 0.0f;
            for (int j = n0; j < n1; j++) {
                const float *r = rec + j * rv;
                const float f = j1c_f((q * r[0]) * s) * S;
                const float wf = w * f;
                t1 = fmaf(r[2], wf, t1);
                t2 = fmaf(r[3], wf * f, t2);
            }
            a1.add(t1);
            a2.add(t2);
        }
    }
    __device__ static bool shares_lengths(const float *rec, int rv, int n0, int n1)
    {
        bool same = true;
        for (int j = n0 + 1; j < n1; j++) same = same && rec[j * rv + 1] == rec[n0 * rv + 1];
        return same;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[2]);
        put_d(r, 2, 0.5 * p[3]);
        r[4] = (float)volume(p);
        r[5] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double qab = sqrt(q.qa * q.qa + q.qb * q.qb);
        g[0] = r[4] * j1c_d(qab * get_d(r, 0)) * sinc_d(q.qc * get_d(r, 2));
    }
    // FP32 record after the rotation rows: R/2pi, (L/2)/2pi, V, w*1e-4, max(R, L/2)
    static constexpr int REC2F = 5, A2F = 9, XS2F = 10;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        f[6] = (float)(p[2] * INV_TWO_PI_D);
        f[7] = (float)(0.5 * p[3] * INV_TWO_PI_D);
        f[8] = (float)volume(p);
        f[9] = (float)(w * 1e-4);
        f[10] = (float)fmax(p[2], 0.5 * p[3]);
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
        const float qab = sqrt_f(fmaf(q.qa, q.qa, q.qb * q.qb));
        g[0] = f[8] * j1c_rev(qab * f[6]) * sinc_rev(q.qc * f[7]);
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- ellipsoid
struct Ellipsoid {
    static constexpr int ID = 2, NPARS = 6, NMAG = 2, MAX_PD = 4, NCON = 1;
    static constexpr bool SHARED_1D = false;   // eval1d_shared: lengths shared by a chunk
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = false;       // three view angles (theta, phi, psi)
    static constexpr bool FQ = true;         // 1D kernel returns <F> as well as <F^2>
    static constexpr bool FP64 = false;      // 1D evaluation in FP64 (eval1d_d)
    static constexpr bool HOLLOW = false;    // shell volume differs from the form volume
    static constexpr int REC1D = 4, REC2D = 8;
    static constexpr int A2D = 5;   // slot of w*1e-4 in the 2D record
    SAS_HD static int sld_param(int k) { return k; }
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * p[2] * p[3] * p[3]; }
    __device__ static double reff(int mode, const double *p)
    {
        const double rp = p[2], re = p[3];
        switch (mode) {
        default:
        case 1: {
            if (rp == re) return rp;
            if (rp * re == 0.) return 0.;
            // mean curvature radius, Isihara, J. Chem. Phys. 18 (1950) 1446, eq. 26
            const double ratio = rp < re ? rp / re : re / rp;
            const double e1 = sqrt(1.0 - ratio * ratio);
            const double b1 = 1.0 + asin(e1) / (e1 * ratio);
            const double bL = (1.0 + e1) / (1.0 - e1);
            const double b2 = 1.0 + 0.5 * ratio * ratio / e1 * log(bL);
            const double delta = 0.75 * b1 * b2;
            return 0.5 * cbrt(2.0 * (delta + 1.0) * rp * re * re);
        }
        case 2: return cbrt(rp * re * re);
        case 3: return rp < re ? rp : re;
        case 4: return rp > re ? rp : re;
        }
    }
    __device__ static double qr_scale(const double *p) { return fmax(p[2], p[3]); }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double s = (p[0] - p[1]) * volume(p);
        const double ratio = p[2] / p[3];
        r[0] = (float)p[3];
        r[1] = (float)(ratio * ratio - 1.0);
        r[2] = (float)(w * 1e-2 * s * 0.5);
        r[3] = (float)(w * 1e-4 * s * s * 0.5);
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        int off, n;
        // r_max = max(rp, re) = re * max(1, sqrt(1 + v))
        const float rmax = r[0] * sqrt_f(fmaxf(1.0f, 1.0f + r[1]));
        quad_range<FIXED>(t, q * rmax, off, n);
        const float qre = q * r[0];
        float a1 = 0.0f, a2 = 0.0f;
        for (int i = 0; i < n; i++) {
            const float x = qre * sqrt_f(fmaf(t.a[off + i], r[1], 1.0f));
            const float f = sph_j1c_f(x);
            const float wf = t.w[off + i] * f;
            a1 += wf;
            a2 = fmaf(wf, f, a2);
        }
        f1 = r[2] * a1;
        f2 = r[3] * a2;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[3] * p[3]);
        put_d(r, 2, p[2] * p[2]);
        r[4] = (float)volume(p);
        r[5] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double qr = sqrt(get_d(r, 0) * (q.qa * q.qa + q.qb * q.qb) + get_d(r, 2) * (q.qc * q.qc));
        g[0] = r[4] * sph_j1c_d(qr);
    }
    // FP32 record after the rotation rows: (re/2pi)^2, (rp/2pi)^2, V, w*1e-4, max(re, rp)
    static constexpr int REC2F = 5, A2F = 9, XS2F = 10;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        const double re = p[3] * INV_TWO_PI_D, rp = p[2] * INV_TWO_PI_D;
        f[6] = (float)(re * re);
        f[7] = (float)(rp * rp);
        f[8] = (float)volume(p);
        f[9] = (float)(w * 1e-4);
        f[10] = (float)fmax(p[2], p[3]);
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
        const float t2 = fmaf(f[6], fmaf(q.qa, q.qa, q.qb * q.qb), f[7] * (q.qc * q.qc));
        g[0] = f[8] * sph_j1c_rev2(t2);
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- core_shell_cylinder
struct CoreShellCylinder {
    static constexpr int ID = 3, NPARS = 8, NMAG = 3, MAX_PD = 5, NCON = 2;
    static constexpr bool SHARED_1D = true;   // eval1d_shared: lengths shared by a chunk
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = false;       // three view angles (theta, phi, psi)
    static constexpr bool FQ = true;         // 1D kernel returns <F> as well as <F^2>
    static constexpr bool FP64 = false;      // 1D evaluation in FP64 (eval1d_d)
    static constexpr bool HOLLOW = false;    // shell volume differs from the form volume
    static constexpr int REC1D = 8, REC2D = 12;
    static constexpr int A2D = 10;  // slot of w*1e-4 in the 2D record
    SAS_HD static int sld_param(int k) { return k; }            // core, shell, solvent
    __device__ static bool valid(const double *) { return true; }
    __device__ static double vol3(double r, double t, double l) { return PI_D * (r + t) * (r + t) * (l + 2.0 * t); }
    __device__ static double volume(const double *p) { return vol3(p[3], p[4], p[5]); }
    __device__ static double reff(int mode, const double *p)
    {
        const double r = p[3], t = p[4], l = p[5];
        switch (mode) {
        default:
        case 1: {
            const double rt = r + t, lt = l + 2.0 * t;
            return 0.5 * cbrt(0.75 * rt * (2.0 * rt * lt + (rt + lt) * (PI_D * rt + lt)));
        }
        case 2: return cbrt(vol3(r, t, l) / FOUR_PI_3_D);
        case 3: return r + t;
        case 4: return 0.5 * l + t;
        case 5: return r < 0.5 * l ? r + t : 0.5 * l + t;
        case 6: return r > 0.5 * l ? r + t : 0.5 * l + t;
        case 7: {
            const double ro = r + t, lo = l + 2.0 * t;
            return sqrt(ro * ro + 0.25 * lo * lo);
        }
        }
    }
    __device__ static double qr_scale(const double *p) { return fmax(p[3] + p[4], 0.5 * p[5] + p[4]); }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double rr = p[3], t = p[4], l = p[5];
        r[0] = (float)rr;
        r[1] = (float)(0.5 * l);
        r[2] = (float)(rr + t);
        r[3] = (float)(0.5 * l + t);
        r[4] = (float)(vol3(rr, 0.0, l) * (p[0] - p[1]));
        r[5] = (float)(vol3(rr, t, l) * (p[1] - p[2]));
        r[6] = (float)(w * 1e-2 * PI_4_D);
        r[7] = (float)(w * 1e-4 * PI_4_D);
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        int off, n;
        quad_range<FIXED>(t, q * fmaxf(r[2], r[3]), off, n);
        const float qrc = q * r[0], qhc = q * r[1], qrs = q * r[2], qhs = q * r[3];
        float a1 = 0.0f, a2 = 0.0f;
        for (int i = 0; i < n; i++) {
            const float s = t.a[off + i], c = t.b[off + i];
            const float f = r[4] * sinc_prod(qhc, c) * j1c_f(qrc * s)
                          + r[5] * sinc_prod(qhs, c) * j1c_f(qrs * s);
            const float wf = t.w[off + i] * f;
            a1 += wf;
            a2 = fmaf(wf, f, a2);
        }
        f1 = r[6] * a1;
        f2 = r[7] * a2;
    }
    // Fixed rule, records n0..n1 of a chunk with the same half lengths r[1]
    // and r[3] (only the radius varies): nodes outer, the two length factors
    // once per node.  Adds sum_pts r[6] w f and sum_pts r[7] w f^2 per node.
    // Nodes in blocks of NB: a record is read once per block instead of once
    // per node; each node's point sums and their order are unchanged.
    template <class ACC>
    __device__ static void eval1d_shared(float q, const float *rec, int rv, int n0, int n1, const Quad &t, ACC &a1,
                                         ACC &a2)
    {
        constexpr int NB = 8;
        const float qhc = q * rec[n0 * rv + 1], qhs = q * rec[n0 * rv + 3];
        int i0 = 0;
        for (; i0 + NB <= t.n; i0 += NB) {
            float s[NB], w[NB], S1[NB], S2[NB], t1[NB], t2[NB];
#pragma unroll
            for (int k = 0; k < NB; k++) {
                const float c = t.b[i0 + k];
                s[k] = t.a[i0 + k];
                w[k] = t.w[i0 + k];
                S1[k] = sinc_prod(qhc, c);
                S2[k] = sinc_prod(qhs, c);
                t1[k] = 0.0f;
                t2[k] = 0.0f;
            }
            for (int j = n0; j < n1; j++) {
                const float *r = rec + j * rv;
                const float qr0 = q * r[0], qr2 = q * r[2];
#pragma unroll
                for (int k = 0; k < NB; k++) {
                    const float f = r[4] * S1[k] * j1c_f(qr0 * s[k]) + r[5] * S2[k] * j1c_f(qr2 * s[k]);
                    const float wf = w[k] * f;
                    t1[k] = fmaf(r[6], wf, t1[k]);
                    t2[k] = fmaf(r[7], wf * f, t2[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < NB; k++) {
                a1.add(t1[k]);
                a2.add(t2[k]);
            }
        }
        for (int i = i0; i < t.n; i++) {
            const float s = t.a[i], c = t.b[i], w = t.w[i];
            const float S1 = sinc_prod(qhc, c), S2 = sinc_prod(qhs, c);
            float t1 = 0.0f, t2 = 0.0f;
            for (int j = n0; j < n1; j++) {
                const float *r = rec + j * rv;
                const float f = r[4] * S1 * j1c_f((q * r[0]) * s) + r[5] * S2 * j1c_f((q * r[2]) * s);
                const float wf = w * f;
                t1 = fmaf(r[6], wf, t1);
                t2 = fmaf(r[7], wf * f, t2);
            }
            a1.add(t1);
            a2.add(t2);
        }
    }
    __device__ static bool shares_lengths(const float *rec, int rv, int n0, int n1)
    {
        bool same = true;
        for (int j = n0 + 1; j < n1; j++)
            same = same && rec[j * rv + 1] == rec[n0 * rv + 1] && rec[j * rv + 3] == rec[n0 * rv + 3];
        return same;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        const double rr = p[3], t = p[4], l = p[5];
        put_d(r, 0, rr);
        put_d(r, 2, 0.5 * l);
        put_d(r, 4, rr + t);
        put_d(r, 6, 0.5 * l + t);
        r[8] = (float)vol3(rr, 0.0, l);
        r[9] = (float)vol3(rr, t, l);
        r[10] = (float)(w * 1e-4);
        r[11] = 0.0f;
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double qab = sqrt(q.qa * q.qa + q.qb * q.qb);
        g[0] = r[8] * sinc_d(q.qc * get_d(r, 2)) * j1c_d(qab * get_d(r, 0));
        g[1] = r[9] * sinc_d(q.qc * get_d(r, 6)) * j1c_d(qab * get_d(r, 4));
    }
    // FP32 record after the rotation rows: R, L/2, R+t, L/2+t (each / 2pi),
    // V_core, V_shell, w*1e-4, max(R+t, L/2+t)
    static constexpr int REC2F = 8, A2F = 12, XS2F = 13;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        const double rr = p[3], t = p[4], l = p[5];
        f[6] = (float)(rr * INV_TWO_PI_D);
        f[7] = (float)(0.5 * l * INV_TWO_PI_D);
        f[8] = (float)((rr + t) * INV_TWO_PI_D);
        f[9] = (float)((0.5 * l + t) * INV_TWO_PI_D);
        f[10] = (float)vol3(rr, 0.0, l);
        f[11] = (float)vol3(rr, t, l);
        f[12] = (float)(w * 1e-4);
        f[13] = (float)fmax(rr + t, 0.5 * l + t);
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
        const float qab = sqrt_f(fmaf(q.qa, q.qa, q.qb * q.qb));
        g[0] = f[10] * sinc_rev(q.qc * f[7]) * j1c_rev(qab * f[6]);
        g[1] = f[11] * sinc_rev(q.qc * f[9]) * j1c_rev(qab * f[8]);
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con)
    {
        con[0] = (float)(sld[0] - sld[1]);
        con[1] = (float)(sld[1] - sld[2]);
    }
};

// ---------------------------------------------------------------- core_shell_sphere
struct CoreShellSphere {
    static constexpr int ID = 4, NPARS = 5, NMAG = 3, MAX_PD = 2, NCON = 2;
    static constexpr bool SHARED_1D = false;   // eval1d_shared: lengths shared by a chunk
    static constexpr bool ORIENTED = false;
    static constexpr bool ABC = false;       // three view angles (theta, phi, psi)
    static constexpr bool FQ = true;         // 1D kernel returns <F> as well as <F^2>
    static constexpr bool FP64 = false;      // 1D evaluation in FP64 (eval1d_d)
    static constexpr bool HOLLOW = false;    // shell volume differs from the form volume
    static constexpr int REC1D = 8, REC2D = 8;
    static constexpr int A2D = 6;   // slot of w*1e-4 in the 2D record
    SAS_HD static int sld_param(int k) { return 2 + k; }        // core, shell, solvent
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(p[0] + p[1]); }
    __device__ static double reff(int mode, const double *p) { return mode == 2 ? p[0] : p[0] + p[1]; }
    __device__ static double qr_scale(const double *p) { return p[0] + p[1]; }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double rc = p[0], rs = p[0] + p[1];
        r[0] = (float)rc;
        r[1] = (float)rs;
        r[2] = (float)(FOUR_PI_3_D * cube_d(rc) * (p[2] - p[3]));
        r[3] = (float)(FOUR_PI_3_D * cube_d(rs) * (p[3] - p[4]));
        r[4] = (float)(w * 1e-2);
        r[5] = (float)(w * 1e-4);
        r[6] = r[7] = 0.0f;
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &, float &f1, float &f2)
    {
        const float f = r[2] * sph_j1c_f(q * r[0]) + r[3] * sph_j1c_f(q * r[1]);
        f1 = r[4] * f;
        f2 = r[5] * f * f;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        const double rc = p[0], rs = p[0] + p[1];
        put_d(r, 0, rc);
        put_d(r, 2, rs);
        r[4] = (float)(FOUR_PI_3_D * cube_d(rc));
        r[5] = (float)(FOUR_PI_3_D * cube_d(rs));
        r[6] = (float)(w * 1e-4);
        r[7] = 0.0f;
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        g[0] = r[4] * sph_j1c_d(q.q * get_d(r, 0));
        g[1] = r[5] * sph_j1c_d(q.q * get_d(r, 2));
    }
    // FP32 record: r_core/2pi, r_shell/2pi, V_core, V_shell, w*1e-4, r_shell
    static constexpr int REC2F = 6, A2F = 4, XS2F = 5;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        const double rc = p[0], rs = p[0] + p[1];
        f[0] = (float)(rc * INV_TWO_PI_D);
        f[1] = (float)(rs * INV_TWO_PI_D);
        f[2] = (float)(FOUR_PI_3_D * cube_d(rc));
        f[3] = (float)(FOUR_PI_3_D * cube_d(rs));
        f[4] = (float)(w * 1e-4);
        f[5] = (float)rs;
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
        const float tc = q.q * f[0], ts = q.q * f[1];
        g[0] = f[2] * sph_j1c_rev2(tc * tc);
        g[1] = f[3] * sph_j1c_rev2(ts * ts);
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con)
    {
        con[0] = (float)(sld[0] - sld[1]);
        con[1] = (float)(sld[1] - sld[2]);
    }
};

// ---------------------------------------------------------------- parallelepiped
// p = sld, sld_solvent, length_a, length_b, length_c, theta, phi, psi
// (parallelepiped.py:216-231; parallelepiped.c:1-150).  1D: nested rules,
// outer sigma = cos(alpha) in [0, 1] (tables c = sigma, d = sqrt(1 - sigma^2)),
// inner angle pi/2 u (tables a = sin, b = cos), plain weights w.
struct Parallelepiped {
    static constexpr int ID = 6, NPARS = 8, NMAG = 2, MAX_PD = 5, NCON = 1;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = true;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = false;
    static constexpr int REC1D = 8, REC2D = 8;
    static constexpr int A2D = 7;
    SAS_HD static int sld_param(int k) { return k; }
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return p[2] * p[3] * p[4]; }
    __device__ static double excluded_volume_radius(double a, double b, double c)
    {
        // equivalent cylinder: the two closest sides make the cross-section
        const double hi = fmax(a, fmax(b, c));
        double s1 = hi, s2 = hi;
        const double len[3] = {a, b, c};
        for (int i = 0; i < 3; i++) {
            if (len[i] < s1) {
                s2 = s1;
                s1 = len[i];
            } else if (len[i] < s2) {
                s2 = len[i];
            }
        }
        double r, l;
        if (s2 - s1 > hi - s2) {
            r = sqrt(s2 * hi / PI_D);
            l = s1;
        } else {
            r = sqrt(s1 * s2 / PI_D);
            l = hi;
        }
        return 0.5 * cbrt(0.75 * r * (2.0 * r * l + (r + l) * (PI_D * r + l)));
    }
    __device__ static double reff(int mode, const double *p)
    {
        const double a = p[2], b = p[3], c = p[4];
        switch (mode) {
        default:
        case 1: return excluded_volume_radius(a, b, c);
        case 2: return cbrt(a * b * c / FOUR_PI_3_D);
        case 3: return 0.5 * a;
        case 4: return 0.5 * b;
        case 5: return 0.5 * c;
        case 6: return sqrt(a * b / PI_D);
        case 7: return 0.5 * sqrt(a * a + b * b);
        case 8: return 0.5 * sqrt(a * a + b * b + c * c);
        }
    }
    __device__ static double qr_scale(const double *p)
    {
        return 0.5 * fmax(sqrt(p[2] * p[2] + p[3] * p[3]), p[4]);
    }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double s = (p[0] - p[1]) * volume(p);
        r[0] = (float)(0.5 * p[3]);                 // b/2
        r[1] = (float)(p[2] / p[3]);                // a/b
        r[2] = (float)(p[4] / p[3]);                // c/b
        r[3] = (float)qr_scale(p);
        r[4] = (float)(w * 1e-2 * s * 0.25);        // two half-interval factors
        r[5] = (float)(w * 1e-4 * s * s * 0.25);
        r[6] = r[7] = 0.0f;
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        const float mu = q * r[0], amax = fmaxf(r[1], 1.0f), mc = mu * r[2];
        int lo, oo, no;
        quad_nested<FIXED>(t, q * r[3], 2, lo, oo, no);
        float o1 = 0.0f, o2 = 0.0f;
        for (int i = 0; i < no; i++) {
            const float mp = mu * t.d[oo + i];
            int li, io, ni;
            quad_nested<FIXED>(t, mp * amax, 3 - lo, li, io, ni);
            float i1 = 0.0f, i2 = 0.0f;
            const float ma = mp * r[1];
            for (int j = 0; j < ni; j++) {
                const float f = sinc_prod(ma, t.a[io + j]) * sinc_prod(mp, t.b[io + j]);
                const float wf = t.w[io + j] * f;
                i1 += wf;
                i2 = fmaf(wf, f, i2);
            }
            const float si = sinc_prod(mc, t.c[oo + i]);
            const float ws = t.w[oo + i] * si;
            o1 = fmaf(ws, i1, o1);
            o2 = fmaf(ws * si, i2, o2);
        }
        f1 = r[4] * o1;
        f2 = r[5] * o2;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, 0.5 * p[2]);
        put_d(r, 2, 0.5 * p[3]);
        put_d(r, 4, 0.5 * p[4]);
        r[6] = (float)volume(p);
        r[7] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        g[0] = r[6] * sinc_d(q.qa * get_d(r, 0)) * sinc_d(q.qb * get_d(r, 2)) * sinc_d(q.qc * get_d(r, 4));
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- triaxial ellipsoid
// p = sld, sld_solvent, radius_equat_minor, radius_equat_major, radius_polar,
// theta, phi, psi (triaxial_ellipsoid.py:143-158; triaxial_ellipsoid.c:1-160).
// 1D: the radii sorted (minor <= major <= polar), nested rules, outer angle
// phi = pi/4 (z + 1) (table c = sin^2 phi), inner u = (z + 1)/2 (table a = u^2),
// plain weights; the inner rule is the same for every outer node.  2D uses
// the radii in the given order.
struct TriaxialEllipsoid {
    static constexpr int ID = 7, NPARS = 8, NMAG = 2, MAX_PD = 5, NCON = 1;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = true;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = false;
    static constexpr int REC1D = 8, REC2D = 8;
    static constexpr int A2D = 7;
    SAS_HD static int sld_param(int k) { return k; }
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * p[2] * p[3] * p[4]; }
    __device__ static double curvature_radius(double ra, double rb, double rc)
    {
        // the reference's first test compares (ra == rb), a 0/1 value, with rc
        if ((double)(ra == rb) == rc) return rc;
        if (ra * rb * rc == 0.) return 0.;
        const double hi = fmax(ra, fmax(rb, rc));
        double r1 = hi, r2 = hi, r3 = hi;
        const double rad[3] = {ra, rb, rc};
        for (int i = 0; i < 3; i++) {
            if (rad[i] < r1) {
                r3 = r2;
                r2 = r1;
                r1 = rad[i];
            } else if (rad[i] < r2) {
                r2 = rad[i];
            }
        }
        double re, rp;
        if (r2 - r1 > r3 - r2) {
            re = sqrt(r2 * r3);
            rp = r1;
        } else {
            re = sqrt(r1 * r2);
            rp = r3;
        }
        // mean curvature radius of the equivalent spheroid (Isihara)
        const double ratio = rp < re ? rp / re : re / rp;
        const double e1 = sqrt(1.0 - ratio * ratio);
        const double b1 = 1.0 + asin(e1) / (e1 * ratio);
        const double bl = (1.0 + e1) / (1.0 - e1);
        const double b2 = 1.0 + 0.5 * ratio * ratio / e1 * log(bl);
        const double delta = 0.75 * b1 * b2;
        return 0.5 * cbrt(2.0 * (delta + 1.0) * rp * re * re);
    }
    __device__ static double reff(int mode, const double *p)
    {
        const double ra = p[2], rb = p[3], rc = p[4];
        switch (mode) {
        default:
        case 1: return curvature_radius(ra, rb, rc);
        case 2: return cbrt(ra * rb * rc);
        case 3: {
            const double m = ra < rb ? ra : rb;
            return m < rc ? m : rc;
        }
        case 4: {
            const double m = ra < rb ? rb : ra;
            return m > rc ? m : rc;
        }
        }
    }
    __device__ static double qr_scale(const double *p) { return fmax(p[2], fmax(p[3], p[4])); }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double lo = fmin(p[2], p[3]), hi = fmax(p[2], p[3]);
        const double mid = fmax(lo, p[4]);
        const double minor = fmin(lo, p[4]), major = fmin(hi, mid), polar = fmax(hi, mid);
        const double s = (p[0] - p[1]) * FOUR_PI_3_D * minor * major * polar;
        r[0] = (float)major;
        r[1] = (float)((minor / major) * (minor / major) - 1.0);
        r[2] = (float)((polar / major) * (polar / major) - 1.0);
        r[3] = (float)polar;
        r[4] = (float)(w * 1e-2 * s * 0.25);
        r[5] = (float)(w * 1e-4 * s * s * 0.25);
        r[6] = r[7] = 0.0f;
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        int lo, oo, no, li, io, ni;
        quad_nested<FIXED>(t, q * r[3], 2, lo, oo, no);
        quad_nested<FIXED>(t, q * r[0], 3 - lo, li, io, ni);
        const float qb = q * r[0];
        float o1 = 0.0f, o2 = 0.0f;
        for (int i = 0; i < no; i++) {
            const float pa = r[1] * t.c[oo + i];
            float i1 = 0.0f, i2 = 0.0f;
            for (int j = 0; j < ni; j++) {
                const float u2 = t.a[io + j];
                const float x = qb * sqrt_f(fmaf(pa, 1.0f - u2, fmaf(r[2], u2, 1.0f)));
                const float f = sph_j1c_f(x);
                const float wf = t.w[io + j] * f;
                i1 += wf;
                i2 = fmaf(wf, f, i2);
            }
            o1 = fmaf(t.w[oo + i], i1, o1);
            o2 = fmaf(t.w[oo + i], i2, o2);
        }
        f1 = r[4] * o1;
        f2 = r[5] * o2;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[2] * p[2]);
        put_d(r, 2, p[3] * p[3]);
        put_d(r, 4, p[4] * p[4]);
        r[6] = (float)volume(p);
        r[7] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double qr = sqrt(get_d(r, 0) * q.qa * q.qa + get_d(r, 2) * q.qb * q.qb + get_d(r, 4) * q.qc * q.qc);
        g[0] = r[6] * sph_j1c_d(qr);
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- paracrystals
// Spheres on a distorted bcc, fcc or sc lattice (bcc_paracrystal.c,
// fcc_paracrystal.c, sc_paracrystal.c with lib/sphere_form.c):
// p = dnn, d_factor, radius, sld, sld_solvent, theta, phi, psi.
// I = vf(R, dnn) P(q) Z(q), Z the paracrystal lattice factor
//   Z = -expm1(2 x)^3 / prod_k ((e^x - 2 cos(d_a a_k)) e^x + 1),
//   x = -(dnn d_factor)^2 |a|^2 / 2,
// with the lattice's projections a_k of (qa, qb, qc).  The reference marks
// these models double precision only: Z cancels near the Bragg peaks, so
// the 1D orientation average (a fixed product rule over theta and phi,
// Gauss-150 by default) and Z in 2D are evaluated in FP64.
enum Lattice { BCC = 0, FCC = 1, SC = 2 };

template <int L>
__device__ __forceinline__ double paracrystal_z(double qa, double qb, double qc, double dnn, double dfac)
{
    double a1, a2, a3, da;
    if (L == BCC) {
        a1 = (-qa + qb + qc) * 0.5;
        a2 = (qa - qb + qc) * 0.5;
        a3 = (qa + qb - qc) * 0.5;
        da = dnn / sqrt(0.75);
    } else if (L == FCC) {
        a1 = (qa + qb) * 0.5;
        a2 = (qa + qc) * 0.5;
        a3 = (qb + qc) * 0.5;
        da = dnn * sqrt(2.0);
    } else {
        a1 = qa;
        a2 = qb;
        a3 = qc;
        da = dnn;
    }
    const double sd = dnn * dfac;
    const double x = -0.5 * sd * sd * (a1 * a1 + a2 * a2 + a3 * a3);
    const double e = exp(x);
    const double m = expm1(2.0 * x);
    const double den = ((e - 2.0 * cos(da * a1)) * e + 1.0) * ((e - 2.0 * cos(da * a2)) * e + 1.0) *
                       ((e - 2.0 * cos(da * a3)) * e + 1.0);
    return -(m * m * m) / den;
}

template <int L>
struct Paracrystal {
    static constexpr int ID = 8 + L, NPARS = 8, NMAG = 2, MAX_PD = 4, NCON = 1;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = true;
    static constexpr bool FQ = false;
    static constexpr bool FP64 = true;
    static constexpr bool HOLLOW = false;
    static constexpr int REC1D = 8, REC2D = 8;
    static constexpr int A2D = 7;
    SAS_HD static int sld_param(int k) { return 3 + k; }
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(p[2]); }
    __device__ static double reff(int, const double *) { return 0.0; }
    // lattice volume fraction of the spheres
    __device__ static double vfrac(const double *p)
    {
        // 1/sqrt(2) with the reference's value (kernel_header.c:156-157)
        const double x = L == BCC ? sqrt(0.75) * p[2] / p[0] : L == FCC ? 0.70710678118654746 * p[2] / p[0]
                                                                        : p[2] / p[0];
        return (L == BCC ? 2.0 : L == FCC ? 4.0 : 1.0) * FOUR_PI_3_D * cube_d(x);
    }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double s = (p[3] - p[4]) * volume(p);
        put_d(r, 0, p[2]);
        put_d(r, 2, p[0]);
        put_d(r, 4, p[1]);
        put_d(r, 6, w * vfrac(p) * 1e-4 * s * s);
    }
    // orientation average of Z by the fixed product rule (outer theta with
    // sin(theta) weight, inner phi): bcc/fcc theta in [0, pi], phi in
    // [0, 2 pi], divided by 4 pi; sc theta, phi in [0, pi/2], divided by pi/2
    __device__ static void eval1d_d(double q, const float *r, const Quad &t, double &f1, double &f2)
    {
        const int n = t.n;
        const double *st = t.g64, *ct = t.g64 + n, *sp = t.g64 + 2 * n, *cp = t.g64 + 3 * n, *w = t.g64 + 4 * n;
        const double dnn = get_d(r, 2), dfac = get_d(r, 4);
        double outer = 0.0;
        for (int i = 0; i < n; i++) {
            const double qc = q * ct[i], qab = q * st[i];
            double inner = 0.0;
            for (int j = 0; j < n; j++) inner += w[j] * paracrystal_z<L>(qab * cp[j], qab * sp[j], qc, dnn, dfac);
            outer += w[i] * inner * st[i];
        }
        // theta_m * phi_m / (4 pi) = pi/8 (bcc, fcc); (pi/4)^2 / (pi/2) = pi/8 (sc)
        const double z = outer * (PI_D / 8.0);
        const double j = sph_j1c_dd(q * get_d(r, 0));
        f1 = 0.0;
        f2 = get_d(r, 6) * j * j * z;
    }
    template <bool FIXED>
    __device__ static void eval1d(float, const float *, const Quad &, float &f1, float &f2)
    {
        f1 = f2 = 0.0f;    // FP64 models use eval1d_d
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[2]);
        put_d(r, 2, p[0]);
        put_d(r, 4, p[1]);
        r[6] = (float)volume(p);
        r[7] = (float)(w * 1e-4 * vfrac(p));
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double z = paracrystal_z<L>(q.qa, q.qb, q.qc, get_d(r, 2), get_d(r, 4));
        const double qm = sqrt(q.qa * q.qa + q.qb * q.qb + q.qc * q.qc);
        g[0] = (float)(r[6] * sph_j1c_dd(qm * get_d(r, 0)) * sqrt(z));
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};
using BccParacrystal = Paracrystal<BCC>;
using FccParacrystal = Paracrystal<FCC>;
using ScParacrystal = Paracrystal<SC>;

// ---------------------------------------------------------------- hollow cylinder
// p = radius, thickness, length, sld, sld_solvent, theta, phi
// (hollow_cylinder.py:90-98; hollow_cylinder.c:1-140).  The amplitude is
// normalised by the shell (wall) volume; the 1D orientation rule runs over
// cos(theta) = (z + 1)/2 (tables a = sin, b = cos, plain w).
struct HollowCylinder {
    static constexpr int ID = 11, NPARS = 7, NMAG = 2, MAX_PD = 5, NCON = 1;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = false;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = true;
    static constexpr int REC1D = 8, REC2D = 8;
    static constexpr int A2D = 7;
    SAS_HD static int sld_param(int k) { return 3 + k; }
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return PI_D * p[2] * (p[0] + p[1]) * (p[0] + p[1]); }
    __device__ static double shell_volume(const double *p)
    {
        return PI_D * p[2] * ((p[0] + p[1]) * (p[0] + p[1]) - p[0] * p[0]);
    }
    __device__ static double reff(int mode, const double *p)
    {
        const double ro = p[0] + p[1], l = p[2];
        switch (mode) {
        default:
        case 1: return 0.5 * cbrt(0.75 * ro * (2.0 * ro * l + (ro + l) * (PI_D * ro + l)));
        case 2: return cbrt(PI_D * ro * ro * l / FOUR_PI_3_D);
        case 3: return ro;
        case 4: return 0.5 * l;
        case 5: return ro < 0.5 * l ? ro : 0.5 * l;
        case 6: return ro > 0.5 * l ? ro : 0.5 * l;
        case 7: return sqrt(ro * ro + 0.25 * l * l);
        }
    }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double ro = p[0] + p[1], g2 = (p[0] / ro) * (p[0] / ro);
        const double s = (p[3] - p[4]) * shell_volume(p);
        r[0] = (float)ro;
        r[1] = (float)p[0];
        r[2] = (float)(0.5 * p[2]);
        r[3] = (float)g2;
        r[4] = (float)(1.0 / (1.0 - g2));
        r[5] = (float)(w * 1e-2 * s * 0.5);
        r[6] = (float)(w * 1e-4 * s * s * 0.5);
        r[7] = 0.0f;
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        int off, n;
        quad_range<FIXED>(t, q * fmaxf(r[0], r[2]), off, n);
        const float qo = q * r[0], qi = q * r[1], qh = q * r[2];
        float a1 = 0.0f, a2 = 0.0f;
        for (int i = 0; i < n; i++) {
            const float st = t.a[off + i];
            const float psi = (j1c_f(qo * st) - r[3] * j1c_f(qi * st)) * r[4];
            const float f = psi * sinc_prod(qh, t.b[off + i]);
            const float wf = t.w[off + i] * f;
            a1 += wf;
            a2 = fmaf(wf, f, a2);
        }
        f1 = r[5] * a1;
        f2 = r[6] * a2;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[0] + p[1]);
        put_d(r, 2, p[0]);
        put_d(r, 4, 0.5 * p[2]);
        r[6] = (float)shell_volume(p);
        r[7] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double qab = sqrt(q.qa * q.qa + q.qb * q.qb);
        const double ro = get_d(r, 0), ri = get_d(r, 2);
        const double g2 = (ri / ro) * (ri / ro);
        const float psi = (float)((j1c_d(qab * ro) - g2 * j1c_d(qab * ri)) / (1.0 - g2));
        g[0] = r[6] * psi * sinc_d(q.qc * get_d(r, 4));
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- vesicle
// p = sld, sld_solvent, volfraction, radius, thickness (vesicle.py:87-96;
// vesicle.c:1-50): a shell of sld between radius and radius + thickness,
// solvent inside and out; normalised by the shell volume.
struct Vesicle {
    static constexpr int ID = 12, NPARS = 5, NMAG = 2, MAX_PD = 2, NCON = 1;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = false;
    static constexpr bool ABC = false;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = true;
    static constexpr int REC1D = 8, REC2D = 8;
    static constexpr int A2D = 6;
    SAS_HD static int sld_param(int k) { return k; }
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(p[3] + p[4]); }
    __device__ static double shell_volume(const double *p)
    {
        return FOUR_PI_3_D * (cube_d(p[3] + p[4]) - cube_d(p[3]));
    }
    __device__ static double reff(int, const double *p) { return p[3] + p[4]; }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double drho = p[0] - p[1];
        r[0] = (float)p[3];
        r[1] = (float)(p[3] + p[4]);
        r[2] = (float)(FOUR_PI_3_D * cube_d(p[3]));
        r[3] = (float)volume(p);
        r[4] = (float)(w * 1e-2 * sqrt(p[2]) * drho);
        r[5] = (float)(w * 1e-4 * p[2] * drho * drho);
        r[6] = r[7] = 0.0f;
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &, float &f1, float &f2)
    {
        const float g = r[3] * sph_j1c_f(q * r[1]) - r[2] * sph_j1c_f(q * r[0]);
        f1 = r[4] * g;
        f2 = r[5] * g * g;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[3]);
        put_d(r, 2, p[3] + p[4]);
        r[4] = (float)(FOUR_PI_3_D * cube_d(p[3]));
        r[5] = (float)volume(p);
        r[6] = (float)(w * 1e-4 * p[2]);
        r[7] = 0.0f;
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        g[0] = r[5] * sph_j1c_d(q.q * get_d(r, 2)) - r[4] * sph_j1c_d(q.q * get_d(r, 0));
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- fuzzy sphere
// p = sld, sld_solvent, radius, fuzziness (fuzzy_sphere.py:113-117;
// fuzzy_sphere.c:1-35): the sphere amplitude damped by exp(-(q fuzziness)^2/2).
struct FuzzySphere {
    static constexpr int ID = 13, NPARS = 4, NMAG = 2, MAX_PD = 2, NCON = 1;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = false;
    static constexpr bool ABC = false;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = false;
    static constexpr int REC1D = 4, REC2D = 8;
    static constexpr int A2D = 5;
    SAS_HD static int sld_param(int k) { return k; }
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(p[2]); }
    __device__ static double reff(int mode, const double *p) { return mode == 2 ? p[2] + p[3] : p[2]; }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double s = (p[0] - p[1]) * volume(p);
        r[0] = (float)p[2];
        r[1] = (float)p[3];
        r[2] = (float)(w * 1e-2 * s);
        r[3] = (float)(w * 1e-4 * s * s);
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &, float &f1, float &f2)
    {
        const float qf = q * r[1];
        const float f = sph_j1c_f(q * r[0]) * expf(-0.5f * qf * qf);
        f1 = r[2] * f;
        f2 = r[3] * f * f;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[2]);
        put_d(r, 2, p[3]);
        r[4] = (float)volume(p);
        r[5] = (float)(w * 1e-4);
        r[6] = r[7] = 0.0f;
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double qf = q.q * get_d(r, 2);
        g[0] = r[4] * sph_j1c_d(q.q * get_d(r, 0)) * (float)exp(-0.5 * qf * qf);
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- core_shell_ellipsoid
// p = radius_equat_core, x_core, thick_shell, x_polar_shell, sld_core,
// sld_shell, sld_solvent, theta, phi (core_shell_ellipsoid.py:139-149;
// core_shell_ellipsoid.c:1-170).  Core radii (re, re x_core), shell radii
// (re + t, re x_core + t x_polar_shell).  1D: u = cos(theta) = (z + 1)/2
// (table a = u^2, plain weights); the argument of each spheroid is
// q sqrt(re^2 + (rp^2 - re^2) u^2).
struct CoreShellEllipsoid {
    static constexpr int ID = 19, NPARS = 9, NMAG = 3, MAX_PD = 5, NCON = 2;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = false;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = false;
    static constexpr int REC1D = 9, REC2D = 11;
    static constexpr int A2D = 10;
    SAS_HD static int sld_param(int k) { return 4 + k; }   // core, shell, solvent
    __device__ static bool valid(const double *) { return true; }
    __device__ static void radii(const double *p, double &re_c, double &rp_c, double &re_s, double &rp_s)
    {
        re_c = p[0];
        rp_c = p[0] * p[1];
        re_s = p[0] + p[2];
        rp_s = p[0] * p[1] + p[2] * p[3];
    }
    __device__ static double volume(const double *p)
    {
        const double re_s = p[0] + p[2], rp_s = p[0] * p[1] + p[2] * p[3];
        return FOUR_PI_3_D * re_s * re_s * rp_s;
    }
    __device__ static double reff(int mode, const double *p)
    {
        const double re = p[0] + p[2], rp = p[0] * p[1] + p[2] * p[3];
        switch (mode) {
        default:
        case 1: {
            if (1.0 == p[1] && 1.0 == p[3]) return re;
            if (re * rp == 0.) return 0.;
            // mean curvature radius of the outer spheroid (Isihara 1950, eq. 26)
            const double ratio = rp < re ? rp / re : re / rp;
            const double e1 = sqrt(1.0 - ratio * ratio);
            const double b1 = 1.0 + asin(e1) / (e1 * ratio);
            const double bL = (1.0 + e1) / (1.0 - e1);
            const double b2 = 1.0 + 0.5 * ratio * ratio / e1 * log(bL);
            const double delta = 0.75 * b1 * b2;
            return 0.5 * cbrt(2.0 * (delta + 1.0) * rp * re * re);
        }
        case 2: return cbrt(volume(p) / FOUR_PI_3_D);
        case 3: return rp < re ? rp : re;
        case 4: return rp > re ? rp : re;
        }
    }
    // 1D record: re_c^2, rp_c^2 - re_c^2, re_s^2, rp_s^2 - re_s^2,
    // V_core (core - shell), V_shell (shell - solvent), w*1e-2/2, w*1e-4/2, max radius
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        double re_c, rp_c, re_s, rp_s;
        radii(p, re_c, rp_c, re_s, rp_s);
        r[0] = (float)(re_c * re_c);
        r[1] = (float)(rp_c * rp_c - re_c * re_c);
        r[2] = (float)(re_s * re_s);
        r[3] = (float)(rp_s * rp_s - re_s * re_s);
        r[4] = (float)(FOUR_PI_3_D * re_c * re_c * rp_c * (p[4] - p[5]));
        r[5] = (float)(FOUR_PI_3_D * re_s * re_s * rp_s * (p[5] - p[6]));
        r[6] = (float)(w * 1e-2 * 0.5);
        r[7] = (float)(w * 1e-4 * 0.5);
        r[8] = (float)fmax(re_s, rp_s);
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        int off, n;
        quad_range<FIXED>(t, q * r[8], off, n);
        float a1 = 0.0f, a2 = 0.0f;
        for (int i = 0; i < n; i++) {
            const float u2 = t.a[off + i];
            const float xc = q * sqrt_f(fmaf(r[1], u2, r[0]));
            const float xs = q * sqrt_f(fmaf(r[3], u2, r[2]));
            const float f = fmaf(r[4], sph_j1c_f(xc), r[5] * sph_j1c_f(xs));
            const float wf = t.w[off + i] * f;
            a1 += wf;
            a2 = fmaf(wf, f, a2);
        }
        f1 = r[6] * a1;
        f2 = r[7] * a2;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        double re_c, rp_c, re_s, rp_s;
        radii(p, re_c, rp_c, re_s, rp_s);
        put_d(r, 0, re_c * re_c);
        put_d(r, 2, rp_c * rp_c);
        put_d(r, 4, re_s * re_s);
        put_d(r, 6, rp_s * rp_s);
        r[8] = (float)(FOUR_PI_3_D * re_c * re_c * rp_c);
        r[9] = (float)(FOUR_PI_3_D * re_s * re_s * rp_s);
        r[10] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double ab = q.qa * q.qa + q.qb * q.qb, c = q.qc * q.qc;
        g[0] = r[8] * sph_j1c_d(sqrt(get_d(r, 0) * ab + get_d(r, 2) * c));
        g[1] = r[9] * sph_j1c_d(sqrt(get_d(r, 4) * ab + get_d(r, 6) * c));
    }
    // FP32 record after the rotation rows: (re_c, rp_c, re_s, rp_s)/2pi squared,
    // V_core, V_shell, w*1e-4, largest radius
    static constexpr int REC2F = 8, A2F = 12, XS2F = 13;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        double re_c, rp_c, re_s, rp_s;
        radii(p, re_c, rp_c, re_s, rp_s);
        const double k = INV_TWO_PI_D * INV_TWO_PI_D;
        f[6] = (float)(re_c * re_c * k);
        f[7] = (float)(rp_c * rp_c * k);
        f[8] = (float)(re_s * re_s * k);
        f[9] = (float)(rp_s * rp_s * k);
        f[10] = (float)(FOUR_PI_3_D * re_c * re_c * rp_c);
        f[11] = (float)(FOUR_PI_3_D * re_s * re_s * rp_s);
        f[12] = (float)(w * 1e-4);
        f[13] = (float)fmax(fmax(re_c, rp_c), fmax(re_s, rp_s));
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
        const float ab = fmaf(q.qa, q.qa, q.qb * q.qb), c = q.qc * q.qc;
        g[0] = f[10] * sph_j1c_rev2(fmaf(f[6], ab, f[7] * c));
        g[1] = f[11] * sph_j1c_rev2(fmaf(f[8], ab, f[9] * c));
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con)
    {
        con[0] = (float)(sld[0] - sld[1]);
        con[1] = (float)(sld[1] - sld[2]);
    }
};

// ---------------------------------------------------------------- elliptical_cylinder
// p = radius_minor, axis_ratio, length, sld, sld_solvent, theta, phi, psi
// (elliptical_cylinder.py:115-122; elliptical_cylinder.c:1-150).  1D: nested
// rules, outer u = cos(alpha) = (z + 1)/2 (tables c = u, d = sqrt(1 - u^2)),
// inner angle theta = (z + 1) pi/2 around the cross-section (tables
// a = cos^2(theta/2), b = sin^2(theta/2)), plain weights; the cross-section
// radius at theta is sqrt(r_minor^2 a + r_major^2 b), which is the
// reference's sqrt((r_major^2 + r_minor^2)/2 - (r_major^2 - r_minor^2)/2 cos theta)
// without its cancellation.  One inner rule serves every outer node.
struct EllipticalCylinder {
    static constexpr int ID = 20, NPARS = 8, NMAG = 2, MAX_PD = 5, NCON = 1;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = true;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = false;
    static constexpr int REC1D = 8, REC2D = 8;
    static constexpr int A2D = 7;
    SAS_HD static int sld_param(int k) { return 3 + k; }
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p) { return PI_D * p[0] * p[0] * p[1] * p[2]; }
    __device__ static double reff(int mode, const double *p)
    {
        const double rm = p[0], ratio = p[1], l = p[2];
        const double rmin = ratio > 1.0 ? rm : ratio * rm, rmax = ratio < 1.0 ? rm : ratio * rm;
        switch (mode) {
        default:
        case 1: {
            const double r = sqrt(rm * rm * ratio);
            return 0.5 * cbrt(0.75 * r * (2.0 * r * l + (r + l) * (PI_D * r + l)));
        }
        case 2: return cbrt(volume(p) / FOUR_PI_3_D);
        case 3: return 0.5 * rm * (1.0 + ratio);
        case 4: return rmin;
        case 5: return rmax;
        case 6: return sqrt(rm * rm * ratio);
        case 7: return 0.5 * l;
        case 8: return rmin < 0.5 * l ? rmin : 0.5 * l;
        case 9: return rmax > 0.5 * l ? rmax : 0.5 * l;
        case 10: {
            const double r = ratio > 1.0 ? rm * ratio : rm;
            return sqrt(r * r + 0.25 * l * l);
        }
        }
    }
    // 1D record: r_minor, axis_ratio^2, L/2, outer qr scale, inner qr scale,
    // w*1e-2 s/4, w*1e-4 s^2/4 with s = (sld - solvent) V
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        const double s = (p[3] - p[4]) * volume(p);
        const double rmaj = p[1] * p[0];
        r[0] = (float)p[0];
        r[1] = (float)(p[1] * p[1]);
        r[2] = (float)(0.5 * p[2]);
        r[3] = (float)fmax(0.5 * p[2], fmax(p[0], rmaj));
        r[4] = (float)fmax(p[0], rmaj);
        r[5] = (float)(w * 1e-2 * s * 0.25);
        r[6] = (float)(w * 1e-4 * s * s * 0.25);
        r[7] = 0.0f;
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        int lo, oo, no, li, io, ni;
        quad_nested<FIXED>(t, q * r[3], 2, lo, oo, no);
        quad_nested<FIXED>(t, q * r[4], 3 - lo, li, io, ni);
        const float qr = q * r[0], qh = q * r[2];
        float o1 = 0.0f, o2 = 0.0f;
        for (int i = 0; i < no; i++) {
            const float qs = qr * t.d[oo + i];
            float i1 = 0.0f, i2 = 0.0f;
            for (int j = 0; j < ni; j++) {
                const float be = j1c_f(qs * sqrt_f(fmaf(r[1], t.b[io + j], t.a[io + j])));
                const float wb = t.w[io + j] * be;
                i1 += wb;
                i2 = fmaf(wb, be, i2);
            }
            const float si = sinc_prod(qh, t.c[oo + i]);
            const float ws = t.w[oo + i] * si;
            o1 = fmaf(ws, i1, o1);
            o2 = fmaf(ws * si, i2, o2);
        }
        f1 = r[5] * o1;
        f2 = r[6] * o2;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        put_d(r, 0, p[0]);
        put_d(r, 2, p[1]);
        put_d(r, 4, 0.5 * p[2]);
        r[6] = (float)volume(p);
        r[7] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double k = get_d(r, 2) * q.qb;
        g[0] = r[6] * j1c_d(get_d(r, 0) * sqrt(k * k + q.qa * q.qa)) * sinc_d(q.qc * get_d(r, 4));
    }
    // FP32 record after the rotation rows: r_minor/2pi, axis_ratio, L/(4 pi), V,
    // w*1e-4, largest length
    static constexpr int REC2F = 6, A2F = 10, XS2F = 11;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        f[6] = (float)(p[0] * INV_TWO_PI_D);
        f[7] = (float)p[1];
        f[8] = (float)(0.5 * p[2] * INV_TWO_PI_D);
        f[9] = (float)volume(p);
        f[10] = (float)(w * 1e-4);
        f[11] = (float)fmax(0.5 * p[2], p[0] * fmax(p[1], 1.0));
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
        const float k = f[7] * q.qb;
        g[0] = f[9] * j1c_rev(f[6] * sqrt_f(fmaf(k, k, q.qa * q.qa))) * sinc_rev(q.qc * f[8]);
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con) { con[0] = (float)(sld[0] - sld[1]); }
};

// ---------------------------------------------------------------- core_shell_bicelle
// p = radius, thick_rim, thick_face, length, sld_core, sld_face, sld_rim,
// sld_solvent, theta, phi (core_shell_bicelle.py:140-151; core_shell_bicelle.c:1-160).
// F = V1 (core - face) S(h) B(R) + V2 (rim - solvent) S(h + tf) B(R + tr)
//   + V3 (face - rim) S(h + tf) B(R), S = sinc(qc .), B = 2 J1(qab .)/(qab .),
// h = L/2, V1 = 2 pi R^2 h, V2 = 2 pi (R + tr)^2 (h + tf), V3 = 2 pi R^2 (h + tf).
// 1D: theta = (z + 1) pi/4 (tables a = sin, b = cos, w sin: the cylinder family's).
struct CoreShellBicelle {
    static constexpr int ID = 21, NPARS = 10, NMAG = 4, MAX_PD = 5, NCON = 3;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = true;
    static constexpr bool ABC = false;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = false;
    static constexpr int REC1D = 10, REC2D = 12;
    static constexpr int A2D = 11;
    SAS_HD static int sld_param(int k) { return 4 + k; }   // core, face, rim, solvent
    __device__ static bool valid(const double *) { return true; }
    __device__ static double volume(const double *p)
    {
        const double ro = p[0] + p[1];
        return PI_D * ro * ro * (p[3] + 2.0 * p[2]);
    }
    __device__ static double reff(int mode, const double *p)
    {
        const double r = p[0] + p[1], l = p[3] + 2.0 * p[2];
        switch (mode) {
        default:
        case 1: return 0.5 * cbrt(0.75 * r * (2.0 * r * l + (r + l) * (PI_D * r + l)));
        case 2: return cbrt(volume(p) / FOUR_PI_3_D);
        case 3: return p[0] + p[1];
        case 4: return 0.5 * p[3] + p[2];
        case 5: return sqrt(r * r + 0.25 * l * l);
        }
    }
    __device__ static void vols(const double *p, double &v1, double &v2, double &v3)
    {
        const double r = p[0], h = 0.5 * p[3];
        v1 = PI_D * r * r * 2.0 * h;
        v2 = PI_D * (r + p[1]) * (r + p[1]) * 2.0 * (h + p[2]);
        v3 = PI_D * r * r * 2.0 * (h + p[2]);
    }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        double v1, v2, v3;
        vols(p, v1, v2, v3);
        r[0] = (float)p[0];
        r[1] = (float)(0.5 * p[3]);
        r[2] = (float)(p[0] + p[1]);
        r[3] = (float)(0.5 * p[3] + p[2]);
        r[4] = (float)(v1 * (p[4] - p[5]));
        r[5] = (float)(v2 * (p[6] - p[7]));
        r[6] = (float)(v3 * (p[5] - p[6]));
        r[7] = (float)(w * 1e-2 * PI_4_D);
        r[8] = (float)(w * 1e-4 * PI_4_D);
        r[9] = (float)fmax(p[0] + p[1], p[3] + p[2]);    // the reference's q r_max
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &t, float &f1, float &f2)
    {
        int off, n;
        quad_range<FIXED>(t, q * r[9], off, n);
        const float qr1 = q * r[0], qh1 = q * r[1], qr2 = q * r[2], qh2 = q * r[3];
        float a1 = 0.0f, a2 = 0.0f;
        for (int i = 0; i < n; i++) {
            const float s = t.a[off + i], c = t.b[off + i];
            const float b1 = j1c_f(qr1 * s), s2 = sinc_prod(qh2, c);
            const float f = r[4] * sinc_prod(qh1, c) * b1 + s2 * fmaf(r[5], j1c_f(qr2 * s), r[6] * b1);
            const float wf = t.w[off + i] * f;
            a1 += wf;
            a2 = fmaf(wf, f, a2);
        }
        f1 = r[7] * a1;
        f2 = r[8] * a2;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        double v1, v2, v3;
        vols(p, v1, v2, v3);
        put_d(r, 0, p[0]);
        put_d(r, 2, 0.5 * p[3]);
        put_d(r, 4, p[0] + p[1]);
        put_d(r, 6, 0.5 * p[3] + p[2]);
        r[8] = (float)v1;
        r[9] = (float)v2;
        r[10] = (float)v3;
        r[11] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double qab = sqrt(q.qa * q.qa + q.qb * q.qb);
        const float b1 = j1c_d(qab * get_d(r, 0)), s2 = sinc_d(q.qc * get_d(r, 6));
        g[0] = r[8] * sinc_d(q.qc * get_d(r, 2)) * b1;
        g[1] = r[9] * s2 * j1c_d(qab * get_d(r, 4));
        g[2] = r[10] * s2 * b1;
    }
    // FP32 record after the rotation rows: R, h, R + tr, h + tf (each / 2pi),
    // V1, V2, V3, w*1e-4, largest length
    static constexpr int REC2F = 9, A2F = 13, XS2F = 14;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        double v1, v2, v3;
        vols(p, v1, v2, v3);
        f[6] = (float)(p[0] * INV_TWO_PI_D);
        f[7] = (float)(0.5 * p[3] * INV_TWO_PI_D);
        f[8] = (float)((p[0] + p[1]) * INV_TWO_PI_D);
        f[9] = (float)((0.5 * p[3] + p[2]) * INV_TWO_PI_D);
        f[10] = (float)v1;
        f[11] = (float)v2;
        f[12] = (float)v3;
        f[13] = (float)(w * 1e-4);
        f[14] = (float)fmax(p[0] + p[1], 0.5 * p[3] + p[2]);
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
        const float qab = sqrt_f(fmaf(q.qa, q.qa, q.qb * q.qb));
        const float b1 = j1c_rev(qab * f[6]), s2 = sinc_rev(q.qc * f[9]);
        g[0] = f[10] * sinc_rev(q.qc * f[7]) * b1;
        g[1] = f[11] * s2 * j1c_rev(qab * f[8]);
        g[2] = f[12] * s2 * b1;
    }
    SAS_HD static void contrasts(const double *, const double *sld, float *con)
    {
        con[0] = (float)(sld[0] - sld[1]);
        con[1] = (float)(sld[2] - sld[3]);
        con[2] = (float)(sld[1] - sld[2]);
    }
};

// ---------------------------------------------------------------- concentric shells
// Vector parameters: a control parameter p[3] (limits [0, 10]) gives
// n = (int)(p[3] + 0.5) shells, read from the first n of ten slots per
// vector (core_multi_shell.c:12-18, onion.c:31-38).  Boundary k = 0..n
// sits at the core radius plus the first k thicknesses.
constexpr int NSHELL = 10, NBOUND = NSHELL + 1;
SAS_HD int shell_count(double fp_n)
{
    const int n = (int)(fp_n + 0.5);
    return n < 0 ? 0 : n > NSHELL ? NSHELL : n;
}

// p = sld_core, radius, sld_solvent, n, sld[10], thickness[10]
// (core_multi_shell.py:91-103; core_multi_shell.c:1-60): uniform shells;
// the amplitude is sum_k 4 pi/3 r_k^3 (sld outside - sld inside) 3 j1(q r_k)/(q r_k).
// Records carry the eleven boundaries with zero volume beyond n, so the
// evaluation skips them.
struct CoreMultiShell {
    static constexpr int ID = 17, NPARS = 4 + 2 * NSHELL, NMAG = 2 + NSHELL, MAX_PD = 5, NCON = NBOUND;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = false;
    static constexpr bool ABC = false;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = false;
    static constexpr bool HOLLOW = false;
    // 1D: (r_k, V_k * jump_k) pairs, w*1e-2, w*1e-4
    // 2D: r_k in FP64, V_k, w*1e-4
    static constexpr int REC1D = 2 * NBOUND + 2, REC2D = 3 * NBOUND + 1;
    static constexpr int A2D = 3 * NBOUND;
    SAS_HD static int sld_param(int k) { return k == 0 ? 0 : k == 1 ? 2 : 2 + k; }   // core, solvent, sld1..
    __device__ static bool valid(const double *) { return true; }
    __device__ static double outer(const double *p)
    {
        const int n = shell_count(p[3]);
        double r = p[1];
        for (int i = 0; i < n; i++) r += p[4 + NSHELL + i];
        return r;
    }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(outer(p)); }
    __device__ static double reff(int mode, const double *p) { return mode == 2 ? p[1] : outer(p); }
    // radius of boundary k, 0 beyond the last one
    __device__ static void radii(const double *p, double *r)
    {
        const int n = shell_count(p[3]);
        double rk = p[1];
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            r[k] = k <= n ? rk : 0.0;
            if (k < NSHELL && k < n) rk += p[4 + NSHELL + k];
        }
    }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        double rad[NBOUND], sld[NMAG], con[NCON];
        radii(p, rad);
#pragma unroll
        for (int k = 0; k < NMAG; k++) sld[k] = p[sld_param(k)];
        jumps(p, sld, con);
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            r[2 * k] = (float)rad[k];
            r[2 * k + 1] = (float)(FOUR_PI_3_D * cube_d(rad[k]) * con[k]);
        }
        r[2 * NBOUND] = (float)(w * 1e-2);
        r[2 * NBOUND + 1] = (float)(w * 1e-4);
    }
    template <bool FIXED>
    __device__ static void eval1d(float q, const float *r, const Quad &, float &f1, float &f2)
    {
        float f = 0.0f;
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            const float a = r[2 * k + 1];   // the same for every lane: a uniform branch
            if (a != 0.0f) f = fmaf(a, sph_j1c_f(q * r[2 * k]), f);
        }
        f1 = r[2 * NBOUND] * f;
        f2 = r[2 * NBOUND + 1] * f * f;
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        double rad[NBOUND];
        radii(p, rad);
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            put_d(r, 2 * k, rad[k]);
            r[2 * NBOUND + k] = (float)(FOUR_PI_3_D * cube_d(rad[k]));
        }
        r[A2D] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            const float v = r[2 * NBOUND + k];
            g[k] = v != 0.0f ? v * sph_j1c_d(q.q * get_d(r, 2 * k)) : 0.0f;
        }
    }
    // FP32 record: r_k/2pi, V_k, w*1e-4, outer radius
    static constexpr int REC2F = 2 * NBOUND + 2, A2F = 2 * NBOUND, XS2F = 2 * NBOUND + 1;
    __device__ static void rec2f(const double *p, double w, float *f)
    {
        double rad[NBOUND];
        radii(p, rad);
        double rmax = 0.0;
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            f[k] = (float)(rad[k] * INV_TWO_PI_D);
            f[NBOUND + k] = (float)(FOUR_PI_3_D * cube_d(rad[k]));
            rmax = fmax(rmax, rad[k]);
        }
        f[A2F] = (float)(w * 1e-4);
        f[XS2F] = (float)rmax;
    }
    __device__ static void geom2f(const QRotF &q, const float *f, float *g)
    {
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            const float v = f[NBOUND + k];
            const float t = q.q * f[k];
            g[k] = v != 0.0f ? v * sph_j1c_rev2(t * t) : 0.0f;
        }
    }
    // jump of the sld across boundary k, outside minus inside (the
    // reference's sign); zero beyond the last boundary
    SAS_HD static void jumps(const double *p, const double *sld, double *con)
    {
        const int n = shell_count(p[3]);
        double inside = sld[0];
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            const double outside = k < n ? sld[2 + (k < NSHELL ? k : 0)] : sld[1];
            con[k] = k <= n ? outside - inside : 0.0;
            inside = outside;
        }
    }
    SAS_HD static void contrasts(const double *p, const double *sld, float *con)
    {
        double c[NCON];
        jumps(p, sld, c);
#pragma unroll
        for (int k = 0; k < NCON; k++) con[k] = (float)c[k];
    }
};

// 3 (sin x - x cos x)/x^3 weighted by the exponential profile
// exp(alpha r/R) of onion.c:12-21 (qr = x): the amplitude, per unit
// volume and unit slope, of a sphere whose sld rises as exp(alpha r/R).
SAS_HD double onion_profile_amp(double alpha, double x)
{
    const double x2 = x * x, a2 = alpha * alpha;
    const double sum2 = a2 + x2;
    double s, c;
    sincos(x, &s, &c);
    const double t1 = (a2 - x2) * s / x - 2.0 * alpha * c;
    const double t2 = alpha * s / x - c;
    return -3.0 * (t1 / sum2 - t2) / sum2;
}

// p = sld_core, radius_core, sld_solvent, n_shells, sld_in[10], sld_out[10],
// thickness[10], A[10] (onion.py:323-333; onion.c:1-73), FP64 like the
// reference (single = False).  Shell i's sld runs from sld_in to sld_out as
// sld_in + slope_i (exp(A_i x) - 1), slope_i = (sld_out - sld_in)/expm1(A_i),
// over its fractional depth x; flat at sld_in for A_i = 0.  Taking the
// reference's difference of sphere amplitudes at each shell's inner and
// outer radius apart gives, with b_k = 3 j1(q r_k)/(q r_k), V_k = 4 pi/3 r_k^3,
//   F = sum_k V_k b_k (sld inside - sld outside at r_k)
//     + sum_i slope_i [V_i (b_i - u(a_i r_i)) - e^A_i V_i+1 (b_i+1 - u(a_i r_i+1))]
// with a_i = A_i/thickness_i and u = onion_profile_amp: one term per
// boundary and one per exponential shell, each linear in the slds, so the
// magnetic kernel reuses the geometry as for the other models.  At q = 0
// every sphere amplitude of the reference is its bare volume and they
// cancel: F = 0.
struct Onion {
    static constexpr int ID = 18, NPARS = 4 + 4 * NSHELL, NMAG = 2 + 2 * NSHELL, MAX_PD = 5;
    static constexpr int NCON = NBOUND + NSHELL;
    static constexpr bool SHARED_1D = false;
    static constexpr bool ORIENTED = false;
    static constexpr bool ABC = false;
    static constexpr bool FQ = true;
    static constexpr bool FP64 = true;
    static constexpr bool HOLLOW = false;
    static constexpr int P_SIN = 4, P_SOUT = 4 + NSHELL, P_T = 4 + 2 * NSHELL, P_A = 4 + 3 * NSHELL;
    // 1D record, in doubles: r_k, V_k jump_k, slope_i, e^A_i, A_i/t_i, w*1e-2, w*1e-4
    static constexpr int D_R = 0, D_C = NBOUND, D_S = 2 * NBOUND, D_E = 2 * NBOUND + NSHELL,
                         D_A = 2 * NBOUND + 2 * NSHELL, D_W = 2 * NBOUND + 3 * NSHELL;
    static constexpr int REC1D = 2 * (D_W + 2);
    // 2D record, in doubles: r_k, e^A_i (0: flat or unused shell), A_i/t_i; then w*1e-4
    static constexpr int E_R = 0, E_E = NBOUND, E_A = NBOUND + NSHELL;
    static constexpr int A2D = 2 * (NBOUND + 2 * NSHELL), REC2D = A2D + 1;
    SAS_HD static int sld_param(int k) { return k == 0 ? 0 : k == 1 ? 2 : 2 + k; }   // core, solvent, in.., out..
    __device__ static bool valid(const double *) { return true; }
    __device__ static double outer(const double *p)
    {
        const int n = shell_count(p[3]);
        double r = p[1];
        for (int i = 0; i < n; i++) r += p[P_T + i];
        return r;
    }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(outer(p)); }
    __device__ static double reff(int, const double *p) { return outer(p); }
    __device__ static bool profiled(const double *p, int i) { return i < shell_count(p[3]) && fabs(p[P_A + i]) > 0.0; }
    __device__ static void radii(const double *p, double *r)
    {
        const int n = shell_count(p[3]);
        double rk = p[1];
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            r[k] = k <= n ? rk : 0.0;
            if (k < NSHELL && k < n) rk += p[P_T + k];
        }
    }
    // jumps (inside - outside) at the boundaries, then the shells' slopes
    SAS_HD static void contrasts_d(const double *p, const double *sld, double *con)
    {
        const int n = shell_count(p[3]);
        double inside = sld[0];
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            double jump = 0.0, slope = 0.0;
            if (k < n) {
                const int i = k < NSHELL ? k : 0;
                const double in = sld[2 + i], A = p[P_A + i];
                jump = inside - in;
                inside = in;
                if (fabs(A) > 0.0) {
                    const double out = sld[2 + NSHELL + i];
                    slope = (out - in) / expm1(A);
                    inside = out;
                }
            } else if (k == n) {
                jump = inside - sld[1];
            }
            con[k] = jump;
            if (k < NSHELL) con[NBOUND + k] = slope;
        }
    }
    SAS_HD static void contrasts(const double *p, const double *sld, float *con)
    {
        double c[NCON];
        contrasts_d(p, sld, c);
#pragma unroll
        for (int k = 0; k < NCON; k++) con[k] = (float)c[k];
    }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
        double rad[NBOUND], sld[NMAG], con[NCON];
        radii(p, rad);
#pragma unroll
        for (int k = 0; k < NMAG; k++) sld[k] = p[sld_param(k)];
        contrasts_d(p, sld, con);
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            put_d(r, 2 * (D_R + k), rad[k]);
            put_d(r, 2 * (D_C + k), FOUR_PI_3_D * cube_d(rad[k]) * con[k]);
        }
#pragma unroll
        for (int i = 0; i < NSHELL; i++) {
            const bool on = profiled(p, i);
            put_d(r, 2 * (D_S + i), on ? con[NBOUND + i] : 0.0);
            put_d(r, 2 * (D_E + i), on ? exp(p[P_A + i]) : 0.0);
            put_d(r, 2 * (D_A + i), on ? p[P_A + i] / p[P_T + i] : 0.0);
        }
        put_d(r, 2 * D_W, w * 1e-2);
        put_d(r, 2 * (D_W + 1), w * 1e-4);
    }
    // the profile term of shell i between radii ri and ro
    __device__ static double shell_term(double q, double ri, double ro, double bi, double bo, double e, double a)
    {
        const double vi = ri > 0.0 ? FOUR_PI_3_D * cube_d(ri) * (bi - onion_profile_amp(a * ri, q * ri)) : 0.0;
        const double vo = ro > 0.0 ? FOUR_PI_3_D * cube_d(ro) * (bo - onion_profile_amp(a * ro, q * ro)) : 0.0;
        return vi - e * vo;
    }
    __device__ static void eval1d_d(double q, const float *r, const Quad &, double &f1, double &f2)
    {
        double f = 0.0;
        if (q != 0.0) {
            double b[NBOUND];
#pragma unroll
            for (int k = 0; k < NBOUND; k++) {
                const double c = get_d(r, 2 * (D_C + k));
                b[k] = sph_j1c_dd(q * get_d(r, 2 * (D_R + k)));
                f += c * b[k];
            }
#pragma unroll
            for (int i = 0; i < NSHELL; i++) {
                const double s = get_d(r, 2 * (D_S + i));
                if (s != 0.0)
                    f += s * shell_term(q, get_d(r, 2 * (D_R + i)), get_d(r, 2 * (D_R + i + 1)), b[i], b[i + 1],
                                        get_d(r, 2 * (D_E + i)), get_d(r, 2 * (D_A + i)));
            }
        }
        f1 = get_d(r, 2 * D_W) * f;
        f2 = get_d(r, 2 * (D_W + 1)) * f * f;
    }
    template <bool FIXED>
    __device__ static void eval1d(float, const float *, const Quad &, float &f1, float &f2)
    {
        f1 = f2 = 0.0f;    // FP64 models use eval1d_d
    }
    __device__ static void rec2d(const double *p, double w, float *r)
    {
        double rad[NBOUND];
        radii(p, rad);
#pragma unroll
        for (int k = 0; k < NBOUND; k++) put_d(r, 2 * (E_R + k), rad[k]);
#pragma unroll
        for (int i = 0; i < NSHELL; i++) {
            const bool on = profiled(p, i);
            put_d(r, 2 * (E_E + i), on ? exp(p[P_A + i]) : 0.0);
            put_d(r, 2 * (E_A + i), on ? p[P_A + i] / p[P_T + i] : 0.0);
        }
        r[A2D] = (float)(w * 1e-4);
    }
    __device__ static void geom2d(const QRot &q, const float *r, float *g)
    {
        const double qm = q.q;
        double b[NBOUND];
#pragma unroll
        for (int k = 0; k < NBOUND; k++) {
            const double rk = get_d(r, 2 * (E_R + k));
            b[k] = sph_j1c_dd(qm * rk);
            g[k] = qm != 0.0 ? (float)(FOUR_PI_3_D * cube_d(rk) * b[k]) : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < NSHELL; i++) {
            const double e = get_d(r, 2 * (E_E + i));
            g[NBOUND + i] = (qm != 0.0 && e != 0.0)
                                ? (float)shell_term(qm, get_d(r, 2 * (E_R + i)), get_d(r, 2 * (E_R + i + 1)), b[i],
                                                    b[i + 1], e, get_d(r, 2 * (E_A + i)))
                                : 0.0f;
        }
    }
};

// ---------------------------------------------------------------- direct-intensity models
// Models the reference defines by Iq alone (no <F>): I(q) is evaluated in
// FP64 at every (q, mesh point) from the point's parameters.  2D data use
// I(|q|), the reference's Iqxy of an isotropic Iq; the magnetic kernel
// evaluates I once per active spin cross section with that cross section's
// effective slds in place, as the reference does.  Records hold the kernel
// parameters and the weight in FP64.  A model D derives from DirectBase and
// supplies NPARS, iq(q, p) and, where it has them, NMAG and sld_param(k),
// MAX_PD, volume(p) and reff(mode, p).

// Gamma with the reference's extension below 1 (lib/sas_gamma.c:140-153):
// tgamma(x + 1)/x for x >= 0, the reflection formula below 0.
SAS_HD double sas_gamma_d(double x)
{
    return x < 0.0 ? PI_D / tgamma(1.0 - x) / sin(PI_D * x) : tgamma(x + 1.0) / x;
}

struct DirectBase {
    static constexpr int NMAG = 0, MAX_PD = 0;
    SAS_HD static int sld_param(int) { return 0; }
    __device__ static double volume(const double *) { return 1.0; }
    __device__ static double reff(int, const double *) { return 0.0; }
};

template <class D>
struct Direct : D {
    static constexpr int NPARS = D::NPARS, NCON = 1;
    static constexpr bool SHARED_1D = false, ORIENTED = false, ABC = false, FQ = false, FP64 = true;
    static constexpr bool HOLLOW = false, DIRECT = true;
    static constexpr int REC1D = 2 * (NPARS + 1), REC2D = REC1D, A2D = 0;
    __device__ static bool valid(const double *) { return true; }
    __device__ static void rec1d(const double *p, double w, float *r)
    {
#pragma unroll
        for (int i = 0; i < NPARS; i++) put_d(r, 2 * i, p[i]);
        put_d(r, 2 * NPARS, w);
    }
    // the record's parameters and weight
    __device__ static double load(const float *r, double *p)
    {
#pragma unroll
        for (int i = 0; i < NPARS; i++) p[i] = get_d(r, 2 * i);
        return get_d(r, 2 * NPARS);
    }
    __device__ static void eval1d_d(double q, const float *r, const Quad &, double &f1, double &f2)
    {
        double p[NPARS];
        const double w = load(r, p);
        f1 = 0.0;
        f2 = w * D::iq(q, p);
    }
    template <bool FIXED>
    __device__ static void eval1d(float, const float *, const Quad &, float &f1, float &f2)
    {
        f1 = f2 = 0.0f;    // FP64 models use eval1d_d
    }
    __device__ static void rec2d(const double *p, double w, float *r) { rec1d(p, w, r); }
    __device__ static void geom2d(const QRot &, const float *, float *g) { g[0] = 0.0f; }   // not used
    SAS_HD static void contrasts(const double *, const double *, float *con) { con[0] = 0.0f; }   // not used
};

// guinier.c: exp(-|rg| rg q^2/3)  (guinier.py; p = rg)
struct GuinierD : DirectBase {
    static constexpr int ID = 22, NPARS = 1;
    __device__ static double iq(double q, const double *p) { return exp(-(fabs(p[0]) * p[0] * q * q / 3.0)); }
};
// gaussian_peak.c: exp(-((q - q0)/sigma)^2/2)  (p = peak_pos, sigma)
struct GaussianPeakD : DirectBase {
    static constexpr int ID = 23, NPARS = 2;
    __device__ static double iq(double q, const double *p)
    {
        const double x = (q - p[0]) / p[1];
        return exp(-0.5 * x * x);
    }
};
// lorentz.c: 1/(1 + (q xi)^2)  (p = cor_length)
struct LorentzD : DirectBase {
    static constexpr int ID = 24, NPARS = 1;
    __device__ static double iq(double q, const double *p)
    {
        const double x = q * p[0];
        return 1.0 / (1.0 + x * x);
    }
};
// dab.c (Debye-Anderson-Brumberger): xi^3/(1 + (q xi)^2)^2  (p = cor_length)
struct DabD : DirectBase {
    static constexpr int ID = 25, NPARS = 1;
    __device__ static double iq(double q, const double *p)
    {
        const double x = q * p[0], d = 1.0 + x * x;
        return cube_d(p[0]) / (d * d);
    }
};
// gel_fit.c: Lorentz-like term over (1 + (D + 1)/3 (q xi)^2)^(D/2) plus a
// Guinier term  (p = guinier_scale, lorentz_scale, rg, fractal_dim, cor_length)
struct GelFitD : DirectBase {
    static constexpr int ID = 26, NPARS = 5;
    __device__ static double iq(double q, const double *p)
    {
        const double x = q * p[4], g = q * p[2];
        const double lorentz = pow(1.0 + ((p[3] + 1.0) / 3.0) * (x * x), 0.5 * p[3]);
        return p[1] / lorentz + p[0] * exp(-(g * g) / 3.0);
    }
};
// mass_fractal.c: sphere form factor times the mass-fractal S(q) with its
// limits at q = 0 and D = 1  (p = radius, fractal_dim_mass, cutoff_length)
struct MassFractalD : DirectBase {
    static constexpr int ID = 27, NPARS = 3;
    __device__ static double iq(double q, const double *p)
    {
        const double f = sph_j1c_dd(q * p[0]), c = p[2], D = p[1];
        double s;
        if (q > 0.0 && D > 1.0) {
            const double dm = D - 1.0, x = q * c;
            s = sas_gamma_d(dm) * sin(dm * atan(x)) * pow(c, dm) * pow(1.0 + x * x, -0.5 * dm) / q;
        } else if (q > 0.0) {
            s = atan(q * c) / q;
        } else if (D > 1.0) {
            s = pow(c, D) * sas_gamma_d(D);
        } else {
            s = c;
        }
        return f * f * s;
    }
};
// surface_fractal.c: sphere form factor times the surface-fractal S(q),
// negative results clipped to 0  (p = radius, fractal_dim_surf, cutoff_length)
struct SurfaceFractalD : DirectBase {
    static constexpr int ID = 28, NPARS = 3;
    __device__ static double iq(double q, const double *p)
    {
        const double f = sph_j1c_dd(q * p[0]), m = 5.0 - p[1], x = q * p[2];
        const double s = sas_gamma_d(m) * pow(p[2], m) * pow(1.0 + x * x, -0.5 * m) * sin(-m * atan(x)) / q;
        const double i = f * f * s;
        return i > 0.0 ? i : 0.0;
    }
};
// mass_surface_fractal.c: product of two generalised Guinier-Porod terms
// (p = fractal_dim_mass, fractal_dim_surf, rg_cluster, rg_primary)
struct MassSurfaceFractalD : DirectBase {
    static constexpr int ID = 29, NPARS = 4;
    __device__ static double iq(double q, const double *p)
    {
        const double dm = 0.5 * p[0], dt = 0.5 * (6.0 - (p[0] + p[1]));
        const double a = q * p[2], b = q * p[3];
        const double t1 = dm == 0.0 ? 1.0 : pow(1.0 + a * a / (3.0 * dm), -dm);
        const double t2 = dt == 0.0 ? 1.0 : pow(1.0 + b * b / (3.0 * dt), -dt);
        return t1 * t2;
    }
};
// star_polymer.c: Benoit's star of f Gaussian arms (f rounded to the
// nearest integer), the series of (v + e^-v - 1) 2/v^2 below (q Rg)^2 = 0.03
// (p = rg_squared, arms)
struct StarPolymerD : DirectBase {
    static constexpr int ID = 30, NPARS = 2;
    __device__ static double iq(double q, const double *p)
    {
        if (q == 0.0) return 1.0;
        const int n = (int)(p[1] + 0.5);
        const double u2 = p[0] * q * q;
        const double v = (u2 * n) / (3 * n - 2);
        const double t1 = u2 < 0.03
            ? 1.0 + v * (-1.0 / 3.0 + v * (1.0 / 12.0 + v * (-1.0 / 60.0 + v * (1.0 / 360.0 + v * (-1.0 / 2520.0)))))
            : 2.0 * (v + expm1(-v)) / (v * v);
        const double e = expm1(-v) / v;
        return (t1 - e * e) / n + e * e;
    }
};
// mono_gauss_coil.c: Debye's Gaussian coil, i_zero 2(e^-x + x - 1)/x^2 with
// x = (q Rg)^2 and a Pade approximant below x = 0.5  (p = i_zero, rg)
struct MonoGaussCoilD : DirectBase {
    static constexpr int ID = 31, NPARS = 2, MAX_PD = 1;
    __device__ static double reff(int mode, const double *p)
    {
        switch (mode) {
        default:
        case 1: return p[1];
        case 2: return 2.0 * p[1];
        case 3: return 3.0 * p[1];
        case 4: return sqrt(5.0 / 3.0) * p[1];
        }
    }
    __device__ static double iq(double q, const double *p)
    {
        const double qr = q * p[1], x = qr * qr;
        double g;
        if (x < 0.5) {
            // [5/5] Pade approximant of 2 (e^-x + x - 1)/x^2 about 0
            const double n = ((((-1.0 / 1995840.0 * x + 1.0 / 23760.0) * x + 1.0 / 2640.0) * x + 2.0 / 99.0) * x
                              + 1.0 / 12.0) * x + 1.0;
            const double d = ((((1.0 / 95040.0 * x + 1.0 / 2376.0) * x + 1.0 / 132.0) * x + 5.0 / 66.0) * x
                              + 5.0 / 12.0) * x + 1.0;
            g = n / d;
        } else {
            g = 2.0 * (expm1(-x) + x) / (x * x);
        }
        return p[0] * g;
    }
};
// fractal.c with lib/fractal_sq.c (Teixeira's S(q), eq. 15): spheres of
// radius r in a fractal aggregate  (p = volfraction, radius, fractal_dim,
// cor_length, sld_block, sld_solvent)
struct FractalD : DirectBase {
    static constexpr int ID = 32, NPARS = 6, NMAG = 2, MAX_PD = 1;
    SAS_HD static int sld_param(int k) { return 4 + k; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(p[1]); }
    __device__ static double iq(double q, const double *p)
    {
        const double r = p[1], D = p[2], xi = p[3];
        double term;
        if (q == 0.0) {
            term = pow(xi / r, D) * sas_gamma_d(D + 1.0);
        } else if (D == 0.0) {
            term = 1.0;
        } else if (D == 1.0) {
            term = atan(q * xi) / (q * r);
        } else {
            const double dm = D - 1.0, x = q * xi;
            term = sas_gamma_d(D + 1.0) / dm * sin(dm * atan(x)) * pow(q * r, -D) * pow(1.0 + 1.0 / (x * x), -0.5 * dm);
        }
        const double f = volume(p) * (p[4] - p[5]) * sph_j1c_dd(q * r);
        return 1e-4 * p[0] * (1.0 + term) * f * f;
    }
};

// Teixeira's fractal structure factor (lib/fractal_sq.c), shared by fractal
// and fractal_core_shell
__device__ __forceinline__ double fractal_sq_d(double q, double r, double D, double xi)
{
    double term;
    if (q == 0.0) {
        term = pow(xi / r, D) * sas_gamma_d(D + 1.0);
    } else if (D == 0.0) {
        term = 1.0;
    } else if (D == 1.0) {
        term = atan(q * xi) / (q * r);
    } else {
        const double dm = D - 1.0, x = q * xi;
        term = sas_gamma_d(D + 1.0) / dm * sin(dm * atan(x)) * pow(q * r, -D) * pow(1.0 + 1.0 / (x * x), -0.5 * dm);
    }
    return 1.0 + term;
}
// sin(x)/x in FP64 (1 at 0, as kernel_header.c:177)
SAS_HD double sinc_dd(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

// fractal_core_shell.c: core-shell spheres (lib/core_shell.c) in a fractal
// aggregate  (p = radius, thickness, sld_core, sld_shell, sld_solvent,
// volfraction, fractal_dim, cor_length)
struct FractalCoreShellD : DirectBase {
    static constexpr int ID = 33, NPARS = 8, NMAG = 3, MAX_PD = 2;
    SAS_HD static int sld_param(int k) { return 2 + k; }
    __device__ static double volume(const double *p) { return FOUR_PI_3_D * cube_d(p[0] + p[1]); }
    __device__ static double iq(double q, const double *p)
    {
        const double rc = p[0], rs = p[0] + p[1];
        const double f = FOUR_PI_3_D * cube_d(rc) * sph_j1c_dd(q * rc) * (p[2] - p[3])
                       + FOUR_PI_3_D * cube_d(rs) * sph_j1c_dd(q * rs) * (p[3] - p[4]);
        return 1.0e-4 * p[5] * fractal_sq_d(q, rs, p[6], p[7]) * f * f;
    }
};
// lamellar_hg.c: a bilayer of tails and heads, normalised by its thickness
// (p = length_tail, length_head, sld, sld_head, sld_solvent)
struct LamellarHgD : DirectBase {
    static constexpr int ID = 34, NPARS = 5, NMAG = 3, MAX_PD = 2;
    SAS_HD static int sld_param(int k) { return 2 + k; }
    __device__ static double iq(double q, const double *p)
    {
        const double lt = p[0], lh = p[1], q2 = q * q;
        const double st = sin(q * lt);
        const double a = (p[3] - p[4]) * (sin(q * (lh + lt)) - st) + (p[2] - p[4]) * st;
        const double pq = a * a * (4.0 / q2);
        return 2.0e-4 * PI_D * pq / q2 / (2.0 * (lh + lt));
    }
};
// Caille structure factor of an N-layer stack (lamellar_stack_caille.c,
// lamellar_hg_stack_caille.c): 1 + 2 sum_k (1 - k/N) cos(q d k)
// exp(-q^2 d^2 alpha_k), alpha_k = eta/(4 pi^2) (ln(pi k) + Euler's constant)
__device__ __forceinline__ double caille_sq(double q, int n, double d, double eta)
{
    const double euler = 0.577215664901533;
    double s = 0.0;
    for (int k = 1; k < n; k++) {
        const double alpha = eta / 4.0 / PI_D / PI_D * (log(PI_D * k) + euler);
        s += (1.0 - (double)k / (double)n) * cos(d * q * k) * exp(-(q * q * d * d * alpha));
    }
    return 2.0 * s + 1.0;
}
// lamellar_stack_caille.c  (p = thickness, Nlayers, d_spacing,
// Caille_parameter, sld, sld_solvent)
struct LamellarStackCailleD : DirectBase {
    static constexpr int ID = 35, NPARS = 6, NMAG = 2, MAX_PD = 2;
    SAS_HD static int sld_param(int k) { return 4 + k; }
    __device__ static double iq(double q, const double *p)
    {
        const double c = p[4] - p[5];
        const double pq = 2.0 * c * c / q / q * (1.0 - cos(q * p[0]));
        const double sq = caille_sq(q, (int)(p[1] + 0.5), p[2], p[3]);
        return 2.0 * PI_D * pq * sq / (p[2] * q * q) * 1.0e-4;
    }
};
// lamellar_hg_stack_caille.c  (p = length_tail, length_head, Nlayers,
// d_spacing, Caille_parameter, sld, sld_head, sld_solvent)
struct LamellarHgStackCailleD : DirectBase {
    static constexpr int ID = 36, NPARS = 8, NMAG = 3, MAX_PD = 3;
    SAS_HD static int sld_param(int k) { return 5 + k; }
    __device__ static double iq(double q, const double *p)
    {
        const double lt = p[0], lh = p[1];
        const double st = sin(q * lt);
        const double a = (p[6] - p[7]) * (sin(q * (lh + lt)) - st) + (p[5] - p[7]) * st;
        const double pq = a * a * (4.0 / (q * q));
        const double sq = caille_sq(q, (int)(p[2] + 0.5), p[3], p[4]);
        return 2.0 * PI_D * pq * sq / (p[3] * q * q) * 1.0e-4;
    }
};
// lamellar_stack_paracrystal.c (Pedersen): paracrystal S(q) mixed between
// the two integer layer counts around Nlayers  (p = thickness, Nlayers,
// d_spacing, sigma_d, sld, sld_solvent)
struct LamellarStackParacrystalD : DirectBase {
    static constexpr int ID = 37, NPARS = 6, NMAG = 2, MAX_PD = 1;
    SAS_HD static int sld_param(int k) { return 4 + k; }
    __device__ static double sn(double w, double q, double d, int n)
    {
        double an = 4.0 * w * w - 2.0 * (w * w * w + w) * cos(q * d);
        an -= 4.0 * pow(w, (double)(n + 2)) * cos((double)n * q * d);
        an += 2.0 * pow(w, (double)(n + 3)) * cos((double)(n - 1) * q * d);
        an += 2.0 * pow(w, (double)(n + 1)) * cos((double)(n + 1) * q * d);
        const double den = 1.0 + w * w - 2.0 * w * cos(q * d);
        return an / ((double)n * den * den);
    }
    __device__ static double iq(double q, const double *p)
    {
        const int n1 = (int)p[1], n2 = n1 + 1;
        const double xn = (double)n2 - p[1];
        const double d = p[2], x = q * p[3] * d;
        const double w = exp(-0.5 * x * x);
        double z = xn * sn(w, q, d, n1) + (1.0 - xn) * sn(w, q, d, n2);
        z += (1.0 - w * w) / (1.0 + w * w - 2.0 * w * cos(q * d));
        const double b = sinc_dd(q * p[0] / 2.0), c = p[4] - p[5];
        return 1.0e-4 * (2.0 * PI_D * c * c * (b * b) * z / (q * q));
    }
};
// linear_pearls.c: N spheres on a line, centre spacing edge_sep + 2r
// (p = radius, edge_sep, num_pearls, sld, sld_solvent)
struct LinearPearlsD : DirectBase {
    static constexpr int ID = 38, NPARS = 5, NMAG = 2;
    SAS_HD static int sld_param(int k) { return 3 + k; }
    __device__ static double iq(double q, const double *p)
    {
        const int n = (int)(p[2] + 0.5);
        const double v = FOUR_PI_3_D * cube_d(p[0]);
        const double m = (p[3] - p[4]) * v, sep = p[1] + 2.0 * p[0];
        double s = (double)n;
        for (int k = 1; k < n; k++) s += 2.0 * (n - k) * sinc_dd(q * sep * k);
        const double a = m * sph_j1c_dd(q * p[0]);
        return 1.0e-4 * s * (a * a) / (n * v);
    }
};
// polymer_micelle.c (Pedersen & Gerstenberg): core, Gaussian corona chains
// and their cross terms  (p = ndensity, v_core, v_corona, sld_solvent,
// sld_core, sld_corona, radius_core, rg, d_penetration, n_aggreg)
struct PolymerMicelleD : DirectBase {
    static constexpr int ID = 39, NPARS = 10, NMAG = 3;
    SAS_HD static int sld_param(int k) { return 3 + k; }
    __device__ static double iq(double q, const double *p)
    {
        const double bc = p[1] * (p[4] - p[3]), bk = p[2] * (p[5] - p[3]), n = p[9];
        const double fc = sph_j1c_dd(q * p[6]);
        const double t1 = (n * bc * fc) * (n * bc * fc);
        const double x = (q * p[7]) * (q * p[7]);
        const double debye = x == 0.0 ? 1.0 : 2.0 * (expm1(-x) + x) / (x * x);
        const double t2 = n * bk * bk * debye;
        const double amp = x == 0.0 ? 1.0 : -expm1(-x) / x;
        const double sk = sinc_dd(q * (p[6] + p[8] * p[7]));
        const double t3 = 2.0 * n * n * bc * bk * fc * amp * sk;
        const double u = bk * amp * sk;
        const double t4 = n * (n - 1.0) * u * u;
        return (t1 + t2 + t3 + t4) * 1.0e-13 * p[0];
    }
};
// raspberry.c (Larson-Smith et al.): large spheres decorated by small ones,
// the intensity normalised inside (form volume 1)  (p = sld_lg, sld_sm,
// sld_solvent, volfraction_lg, volfraction_sm, surface_fraction, radius_lg,
// radius_sm, penetration)
struct RaspberryD : DirectBase {
    static constexpr int ID = 40, NPARS = 9, NMAG = 3, MAX_PD = 3;
    SAS_HD static int sld_param(int k) { return k; }
    __device__ static double reff(int mode, const double *p)
    {
        return mode == 2 ? p[6] + 2.0 * p[7] - p[8] : p[6];
    }
    __device__ static double iq(double q, const double *p)
    {
        const double rl = p[6], rs = p[7], dl = fabs(p[0] - p[2]), ds = fabs(p[1] - p[2]);
        const double vl = FOUR_PI_3_D * rl * rl * rl, vs = FOUR_PI_3_D * rs * rs * rs;
        const double np = p[4] * p[5] * vl / p[3] / vs;
        const double slt = dl * vl + np * ds * vs;
        const double pl = sph_j1c_dd(q * rl), ps = sph_j1c_dd(q * rs);
        const double sc = sinc_dd(q * (rl + p[8] * rs));
        double f2 = dl * dl * vl * vl * pl * pl + np * ds * ds * vs * vs * ps * ps
                  + np * (np - 1.0) * ds * ds * vs * vs * (ps * ps * sc * sc) + 2.0 * np * dl * ds * vl * vs * (pl * ps * sc);
        if (f2 != 0.0) f2 = f2 / slt / slt;
        f2 = f2 * (p[3] * dl * dl * vl + p[4] * p[5] * np * ds * ds * vs);
        f2 += p[4] * (1.0 - p[5]) * ds * ds * vs * ps * ps;
        return f2 * 1.0e8 * 1.0e-12;
    }
};

using Guinier = Direct<GuinierD>;
using GaussianPeak = Direct<GaussianPeakD>;
using Lorentz = Direct<LorentzD>;
using Dab = Direct<DabD>;
using GelFit = Direct<GelFitD>;
using MassFractal = Direct<MassFractalD>;
using SurfaceFractal = Direct<SurfaceFractalD>;
using MassSurfaceFractal = Direct<MassSurfaceFractalD>;
using StarPolymer = Direct<StarPolymerD>;
using MonoGaussCoil = Direct<MonoGaussCoilD>;
using Fractal = Direct<FractalD>;
using FractalCoreShell = Direct<FractalCoreShellD>;
using LamellarHg = Direct<LamellarHgD>;
using LamellarStackCaille = Direct<LamellarStackCailleD>;
using LamellarHgStackCaille = Direct<LamellarHgStackCailleD>;
using LamellarStackParacrystal = Direct<LamellarStackParacrystalD>;
using LinearPearls = Direct<LinearPearlsD>;
using PolymerMicelle = Direct<PolymerMicelleD>;
using Raspberry = Direct<RaspberryD>;

}  // namespace sas
