/*
 * sasoracle.c -- FP64 CPU oracle for the sasmodels I(q) evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, the smoke check and the CPU-baseline
 * leg of bench.py may load it.  The shipped library (libsaship.so) never
 * links, calls or falls back to it.
 *
 * What it restates (reference file:line, sasmodels v1.1.0):
 *   - the polydispersity-mesh integration kernel  sasmodels/kernel_iq.c:317-858
 *     (mesh walk with restart :642-664/:682-823, weight product :654,
 *      strict "weight > cutoff" :726, VALID exclusion :721, projection
 *      weight |cos dtheta| :621-626, view angles at theta_par+2 :612-613,
 *      accumulator slots result[nout*nq + 0..3] :392-423/:826-847)
 *   - view/jitter rotation qac_rotation/qac_apply  kernel_iq.c:197-242
 *   - magnetism set_spin_weights/mag_sld/xs loop  kernel_iq.c:84-175, :750-779
 *   - Cephes J1 (double)  models/lib/sas_J1.c:107-153, 2J1(x)/x :241-245
 *   - Horner evaluation  models/lib/polevl.c:52-85
 *   - 3 j1(x)/x  models/lib/sas_3j1x_x.c:40-57 (double cutoff 0.1)
 *   - sin(x)/x  kernel_header.c:177
 *   - quadrature selection  models/lib/adaptive.c:11276-11298 and
 *     models/lib/nonadaptive.c:9-13; Gauss-Legendre nodes as made by
 *     gengauss.py:10-46 (numpy leggauss), recomputed here by Newton's method
 *   - models: sphere.c:1-20, cylinder.c:1-102, ellipsoid.c:1-105,
 *     core_shell_cylinder.c:1-123, core_shell_sphere.c:1-30 with
 *     lib/core_shell.c:8-30, parallelepiped.c:1-150,
 *     triaxial_ellipsoid.c:1-160 (nested rules adaptive.c:11243-11298),
 *     bcc_paracrystal.c, fcc_paracrystal.c, sc_paracrystal.c with
 *     lib/sphere_form.c (Gauss-150 product rule, lib/gauss150.c),
 *     hollow_cylinder.c, vesicle.c, fuzzy_sphere.c (shell volumes of the
 *     hollow shapes in the volume sums, kernel_iq.c:728-731),
 *     core_multi_shell.c and onion.c (vector parameters: n shells of
 *     sld[n], thickness[n], ... packed as ten slots each),
 *     core_shell_ellipsoid.c, elliptical_cylinder.c (nested rules),
 *     core_shell_bicelle.c; the Iq-only models guinier.c, gaussian_peak.c,
 *     lorentz.c, dab.c, gel_fit.c, mass_fractal.c, surface_fractal.c,
 *     mass_surface_fractal.c, star_polymer.c, mono_gauss_coil.c, fractal.c
 *     with lib/fractal_sq.c and lib/sas_gamma.c:140-153, fractal_core_shell.c,
 *     lamellar_hg.c, lamellar_stack_caille.c, lamellar_hg_stack_caille.c,
 *     lamellar_stack_paracrystal.c, linear_pearls.c, polymer_micelle.c,
 *     raspberry.c;
 *     view/jitter rotation with psi qabc_rotation/qabc_apply
 *     kernel_iq.c:251-310
 *   - hard-sphere structure factor hardsphere.c:1-90
 *
 * Floating-point expressions keep the reference's operand order so that,
 * compiled like the reference DLL (gcc -std=c99 -O2, no FMA contraction),
 * results agree to the last bits; the summation over q runs q-outer with
 * the per-q accumulation still in mesh order, so OpenMP threads do not
 * change any result.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Arithmetic type of the restatement: double, except in oracle/flopcount.cpp,
 * which compiles this file as C++ with a counting scalar in its place to
 * count the floating-point operations of the same formulas. */
#ifndef SASORACLE_COUNTING
typedef double oreal;
#endif

/* constants with the values the reference's C uses (kernel_header.c:144-169) */
#define O_PI 3.14159265358979323846
#define O_PI_4 0.78539816339744830962
#define O_PI_180 0.017453292519943295
#define O_4PI_3 4.18879020478639
#define O_SQRT1_PI 0.56418958354775628
#define O_SQRT1_2 0.70710678118654746

/* ------------------------------------------------------------------ */
/* Gauss-Legendre nodes: Newton iteration on P_n in long double.       */
/* ------------------------------------------------------------------ */
#define GL_MAX_TABLES 16
static int gl_n[GL_MAX_TABLES];
static oreal *gl_z[GL_MAX_TABLES], *gl_w[GL_MAX_TABLES];
static int gl_count = 0;

static void gl_compute(int n, oreal *z, oreal *w)
{
    for (int i = 0; i < (n + 1) / 2; i++) {
        /* root i counted from the top (x near +1) */
        long double x = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
        long double dp = 0.0L;
        for (int it = 0; it < 100; it++) {
            long double p0 = 1.0L, p1 = x;
            for (int k = 2; k <= n; k++) {
                long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0L);
            long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-19L) {
                /* one more pass for the derivative at the converged root */
                p0 = 1.0L; p1 = x;
                for (int k = 2; k <= n; k++) {
                    long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
                    p0 = p1; p1 = p2;
                }
                dp = n * (x * p1 - p0) / (x * x - 1.0L);
                break;
            }
        }
        long double wt = 2.0L / ((1.0L - x * x) * dp * dp);
        /* ascending order, as numpy.polynomial.legendre.leggauss returns */
        z[n - 1 - i] = (oreal)x;
        z[i] = (oreal)(-x);
        w[n - 1 - i] = (oreal)wt;
        w[i] = (oreal)wt;
    }
    if (n % 2) z[n / 2] = 0.0;
}

/* Returns the table for n points, computing it on first use.
 * Not thread safe on first use: callers prime it before parallel regions. */
static int gl_table(int n, const oreal **z, const oreal **w)
{
    for (int t = 0; t < gl_count; t++) {
        if (gl_n[t] == n) { *z = gl_z[t]; *w = gl_w[t]; return n; }
    }
    if (gl_count >= GL_MAX_TABLES || n < 1) return -1;
    oreal *zz = (oreal *)malloc(sizeof(oreal) * n);
    oreal *ww = (oreal *)malloc(sizeof(oreal) * n);
    gl_compute(n, zz, ww);
    gl_n[gl_count] = n; gl_z[gl_count] = zz; gl_w[gl_count] = ww; gl_count++;
    *z = zz; *w = ww;
    return n;
}

/* quad_n == 0: adaptive choice by q*r_max (adaptive.c:11290-11297);
 * quad_n > 0: fixed table (nonadaptive.c:9-13 after set_integration_size). */
static int quad_select(int quad_n, oreal qr, const oreal **z, const oreal **w)
{
    int n;
    if (quad_n > 0) n = quad_n;
    else if (qr < 10) n = 20;
    else if (qr < 100) n = 76;
    else if (qr < 750) n = 500;
    else n = 5000;
    return gl_table(n, z, w);
}

/* Nested rules (adaptive.c:11243-11298): the outer call passes
 * outer_n = ADAPTIVE_MAX_OUTER, the inner call the outer rule's size. */
#define O_ADAPTIVE_MAX_N 100000
#define O_ADAPTIVE_MAX_OUTER (O_ADAPTIVE_MAX_N / 500)
static int quad_select_nested(int quad_n, oreal qr, int outer_n, const oreal **z, const oreal **w)
{
    int n;
    if (quad_n > 0) n = quad_n;
    else if (qr < 10 || outer_n > O_ADAPTIVE_MAX_N / 76) n = 20;
    else if (qr < 100 || outer_n > O_ADAPTIVE_MAX_N / 500) n = 76;
    else if (qr < 750 || outer_n > O_ADAPTIVE_MAX_N / 5000) n = 500;
    else n = 5000;
    return gl_table(n, z, w);
}

static void quad_prime(int quad_n)
{
    const oreal *z, *w;
    if (quad_n > 0) { gl_table(quad_n, &z, &w); return; }
    gl_table(20, &z, &w); gl_table(76, &z, &w);
    gl_table(500, &z, &w); gl_table(5000, &z, &w);
    gl_table(150, &z, &w);   /* the paracrystals' default product rule */
}

/* ------------------------------------------------------------------ */
/* Special functions                                                   */
/* ------------------------------------------------------------------ */
/* c[0]*x^n + c[1]*x^(n-1) + ... + c[n]   (polevl.c:52-66) */
static oreal horner(oreal x, const oreal *c, int n)
{
    oreal acc = c[0];
    for (int i = 1; i <= n; i++) acc = acc * x + c[i];
    return acc;
}

/* x^n + c[0]*x^(n-1) + ... + c[n-1]   (polevl.c:73-85) */
static oreal horner_monic(oreal x, const oreal *c, int n)
{
    oreal acc = x + c[0];
    for (int i = 1; i < n; i++) acc = acc * x + c[i];
    return acc;
}

/* Cephes j1.c rational approximations (S. L. Moshier, Cephes 2.8), the
 * published coefficient sets RP/RQ on |x|<=5 and PP/PQ/QP/QQ beyond. */
static const oreal J1_RP[4] = {
    -8.99971225705559398224E8, 4.52228297998194034323E11,
    -7.27494245221818276015E13, 3.68295732863852883286E15};
static const oreal J1_RQ[8] = {
    6.20836478118054335476E2, 2.56987256757748830383E5,
    8.35146791431949253037E7, 2.21511595479792499675E10,
    4.74914122079991414898E12, 7.84369607876235854894E14,
    8.95222336184627338078E16, 5.32278620332680085395E18};
static const oreal J1_PP[7] = {
    7.62125616208173112003E-4, 7.31397056940917570436E-2,
    1.12719608129684925192E0, 5.11207951146807644818E0,
    8.42404590141772420927E0, 5.21451598682361504063E0,
    1.00000000000000000254E0};
static const oreal J1_PQ[7] = {
    5.71323128072548699714E-4, 6.88455908754495404082E-2,
    1.10514232634061696926E0, 5.07386386128601488557E0,
    8.39985554327604159757E0, 5.20982848682361821619E0,
    9.99999999999999997461E-1};
static const oreal J1_QP[8] = {
    5.10862594750176621635E-2, 4.98213872951233449420E0,
    7.58238284132545283818E1, 3.66779609360150777800E2,
    7.10856304998926107277E2, 5.97489612400613639965E2,
    2.11688757100572135698E2, 2.52070205858023719784E1};
static const oreal J1_QQ[7] = {
    7.42373277035675149943E1, 1.05644886038262816351E3,
    4.98641058337653607651E3, 9.56231892404756170795E3,
    7.99704160447350683650E3, 2.82619278517639096600E3,
    3.36093607810698293419E2};

/* J1(x), odd extension to x<0 (sas_J1.c:107-153) */
static oreal o_j1(oreal x)
{
    const oreal Z1 = 1.46819706421238932572E1;  /* first zero squared */
    const oreal Z2 = 4.92184563216946036703E1;  /* second zero squared */
    const oreal ax = x < 0 ? -x : x;
    const oreal sgn = x < 0 ? -1.0 : 1.0;
    if (ax <= 5.0) {
        const oreal z = ax * ax;
        oreal r = horner(z, J1_RP, 3) / horner_monic(z, J1_RQ, 8);
        r = r * ax * (z - Z1) * (z - Z2);
        return sgn * r;
    }
    const oreal u = 5.0 / ax;
    const oreal z = u * u;
    oreal p = horner(z, J1_PP, 6) / horner(z, J1_PQ, 6);
    const oreal qq = horner(z, J1_QP, 7) / horner_monic(z, J1_QQ, 7);
    const oreal s = sin(ax), c = cos(ax);
    /* p*cos(x-3pi/4) - u*q*sin(x-3pi/4) expanded in sin(x), cos(x) */
    p = p * (s - c) + u * qq * (s + c);
    return sgn * p * O_SQRT1_PI / sqrt(ax);
}

static oreal o_2j1x_x(oreal x) { return (x != 0.0) ? 2.0 * o_j1(x) / x : 1.0; }

static oreal o_sinx_x(oreal x) { return x == 0 ? 1.0 : sin(x) / x; }

/* 3 j1(x)/x: 4-term Taylor series below 0.1, closed form above */
static oreal o_sph_j1c(oreal x)
{
    if (fabs(x) < 0.1) {
        const oreal x2 = x * x;
        return (1.0 + x2 * (-3. / 30. + x2 * (3. / 840. + x2 * (-3. / 45360.))));
    }
    const oreal s = sin(x), c = cos(x);
    return 3.0 * (s / x - c) / (x * x);
}

static oreal sq(oreal x) { return x * x; }
static oreal cb(oreal x) { return x * x * x; }

/* ------------------------------------------------------------------ */
/* Models.  p[] is the kernel parameter vector (no scale/background).  */
/* ------------------------------------------------------------------ */

/* ---- sphere (sphere.c:1-20): p = sld, sld_solvent, radius ---- */
static oreal sph_vol(const oreal *p) { return O_4PI_3 * cb(p[2]); }
static oreal sph_reff(int mode, const oreal *p) { (void)mode; return p[2]; }
static void sph_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal bes = o_sph_j1c(q * p[2]);
    const oreal contrast = (p[0] - p[1]);
    const oreal form = contrast * sph_vol(p) * bes;
    *f1 = 1.0e-2 * form;
    *f2 = 1.0e-4 * form * form;
}

/* ---- cylinder (cylinder.c:1-102): p = sld, sld_solvent, radius, length, theta, phi ---- */
static oreal cyl_volume(oreal r, oreal l) { return O_PI * r * r * l; }
static oreal cyl_vol(const oreal *p) { return cyl_volume(p[2], p[3]); }
static int cyl_valid(const oreal *p) { return p[2] >= 0.0 && p[3] >= 0.0; }
static oreal cyl_reff(int mode, const oreal *p)
{
    const oreal r = p[2], l = p[3];
    switch (mode) {
    default:
    case 1: return 0.5 * cbrt(0.75 * r * (2.0 * r * l + (r + l) * (O_PI * r + l)));
    case 2: return cbrt(O_PI * r * r * l / O_4PI_3);
    case 3: return r;
    case 4: return 0.5 * l;
    case 5: return (r < 0.5 * l ? r : 0.5 * l);
    case 6: return (r > 0.5 * l ? r : 0.5 * l);
    case 7: return sqrt(r * r + 0.25 * l * l);
    }
}
static oreal cyl_amp(oreal qab, oreal qc, oreal r, oreal l)
{
    return o_2j1x_x(qab * r) * o_sinx_x(qc * 0.5 * l);
}
static void cyl_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal r = p[2], l = p[3];
    const oreal *z, *w;
    const int n = quad_select(quad_n, q * fmax(r, 0.5 * l), &z, &w);
    oreal t1 = 0.0, t2 = 0.0;
    for (int i = 0; i < n; i++) {
        const oreal th = z[i] * O_PI_4 + O_PI_4;
        const oreal s = sin(th), c = cos(th);
        const oreal form = cyl_amp(q * s, q * c, r, l);
        t1 += w[i] * form * s;
        t2 += w[i] * form * form * s;
    }
    t1 *= O_PI_4;
    t2 *= O_PI_4;
    const oreal sv = (p[0] - p[1]) * cyl_volume(r, l);
    *f1 = 1e-2 * sv * t1;
    *f2 = 1e-4 * sv * sv * t2;
}
static oreal cyl_iqac(oreal qab, oreal qc, const oreal *p)
{
    const oreal form = cyl_amp(qab, qc, p[2], p[3]);
    const oreal sv = (p[0] - p[1]) * cyl_volume(p[2], p[3]);
    return 1.0e-4 * sq(sv * form);
}

/* ---- ellipsoid (ellipsoid.c:1-105): p = sld, sld_solvent, rp, re, theta, phi ---- */
static oreal ell_volume(oreal rp, oreal re) { return O_4PI_3 * rp * re * re; }
static oreal ell_vol(const oreal *p) { return ell_volume(p[2], p[3]); }
static oreal ell_reff(int mode, const oreal *p)
{
    const oreal rp = p[2], re = p[3];
    switch (mode) {
    default:
    case 1: {
        if (rp == re) return rp;
        if (rp * re == 0.) return 0.;
        /* Isihara (1950) eq. 26 mean curvature radius */
        const oreal ratio = (rp < re ? rp / re : re / rp);
        const oreal e1 = sqrt(1.0 - ratio * ratio);
        const oreal b1 = 1.0 + asin(e1) / (e1 * ratio);
        const oreal bL = (1.0 + e1) / (1.0 - e1);
        const oreal b2 = 1.0 + 0.5 * ratio * ratio / e1 * log(bL);
        const oreal delta = 0.75 * b1 * b2;
        const oreal ddd = 2.0 * (delta + 1.0) * rp * re * re;
        return 0.5 * cbrt(ddd);
    }
    case 2: return cbrt(rp * re * re);
    case 3: return (rp < re ? rp : re);
    case 4: return (rp > re ? rp : re);
    }
}
static void ell_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal rp = p[2], re = p[3];
    const oreal vm1 = sq(rp / re) - 1.0;
    const oreal *z, *w;
    const int n = quad_select(quad_n, q * fmax(rp, re), &z, &w);
    oreal t2 = 0.0, t1 = 0.0;
    for (int i = 0; i < n; i++) {
        const oreal u = z[i] * 0.5 + 0.5;
        const oreal r = re * sqrt(1.0 + u * u * vm1);
        const oreal f = o_sph_j1c(q * r);
        t2 += w[i] * f * f;
        t1 += w[i] * f;
    }
    t1 *= 0.5;
    t2 *= 0.5;
    const oreal sv = (p[0] - p[1]) * ell_volume(rp, re);
    *f1 = 1e-2 * sv * t1;
    *f2 = 1e-4 * sv * sv * t2;
}
static oreal ell_iqac(oreal qab, oreal qc, const oreal *p)
{
    const oreal rp = p[2], re = p[3];
    const oreal qr = sqrt(sq(re * qab) + sq(rp * qc));
    const oreal f = o_sph_j1c(qr);
    const oreal sv = (p[0] - p[1]) * ell_volume(rp, re);
    return 1.0e-4 * sq(f * sv);
}

/* ---- core_shell_cylinder (core_shell_cylinder.c:1-123):
 *      p = sld_core, sld_shell, sld_solvent, radius, thickness, length, theta, phi ---- */
static oreal csc_volume(oreal r, oreal t, oreal l) { return O_PI * sq(r + t) * (l + 2.0 * t); }
static oreal csc_vol(const oreal *p) { return csc_volume(p[3], p[4], p[5]); }
static oreal csc_reff(int mode, const oreal *p)
{
    const oreal r = p[3], t = p[4], l = p[5];
    switch (mode) {
    default:
    case 1: {
        const oreal rt = r + t, lt = l + 2.0 * t;
        return 0.5 * cbrt(0.75 * rt * (2.0 * rt * lt + (rt + lt) * (O_PI * rt + lt)));
    }
    case 2: return cbrt(csc_volume(r, t, l) / O_4PI_3);
    case 3: return r + t;
    case 4: return 0.5 * l + t;
    case 5: return (r < 0.5 * l ? r + t : 0.5 * l + t);
    case 6: return (r > 0.5 * l ? r + t : 0.5 * l + t);
    case 7: {
        const oreal ro = r + t, lo = l + 2.0 * t;
        return sqrt(ro * ro + 0.25 * lo * lo);
    }
    }
}
static oreal csc_part(oreal vd, oreal besarg, oreal siarg)
{
    return vd * o_sinx_x(siarg) * o_2j1x_x(besarg);
}
static oreal csc_amp(oreal qab, oreal qc, const oreal *p)
{
    const oreal r = p[3], t = p[4], l = p[5];
    const oreal core_r = r;
    const oreal core_h = 0.5 * l;
    const oreal core_vd = csc_volume(r, 0, l) * (p[0] - p[1]);
    const oreal shell_r = (r + t);
    const oreal shell_h = (0.5 * l + t);
    const oreal shell_vd = csc_volume(r, t, l) * (p[1] - p[2]);
    return csc_part(core_vd, core_r * qab, core_h * qc)
         + csc_part(shell_vd, shell_r * qab, shell_h * qc);
}
static void csc_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal r = p[3], t = p[4], l = p[5];
    const oreal *z, *w;
    const int n = quad_select(quad_n, q * fmax(r + t, 0.5 * l + t), &z, &w);
    oreal t1 = 0.0, t2 = 0.0;
    for (int i = 0; i < n; i++) {
        const oreal th = z[i] * O_PI_4 + O_PI_4;
        const oreal s = sin(th), c = cos(th);
        const oreal fq = csc_amp(q * s, q * c, p);
        t1 += w[i] * fq * s;
        t2 += w[i] * fq * fq * s;
    }
    *f1 = 1.0e-2 * t1 * O_PI_4;
    *f2 = 1.0e-4 * t2 * O_PI_4;
}
static oreal csc_iqac(oreal qab, oreal qc, const oreal *p)
{
    const oreal fq = csc_amp(qab, qc, p);
    return 1.0e-4 * fq * fq;
}

/* ---- core_shell_sphere (core_shell_sphere.c:1-30, lib/core_shell.c:8-30):
 *      p = radius, thickness, sld_core, sld_shell, sld_solvent ---- */
static oreal css_vol(const oreal *p) { return O_4PI_3 * cb(p[0] + p[1]); }
static oreal css_reff(int mode, const oreal *p)
{
    switch (mode) {
    default:
    case 1: return p[0] + p[1];
    case 2: return p[0];
    }
}
static void css_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal r = p[0], t = p[1];
    const oreal core_qr = q * r;
    const oreal core_contrast = p[2] - p[3];
    const oreal core_bes = o_sph_j1c(core_qr);
    const oreal core_volume = O_4PI_3 * cb(r);
    oreal f = core_volume * core_bes * core_contrast;
    const oreal shell_qr = q * (r + t);
    const oreal shell_contrast = p[3] - p[4];
    const oreal shell_bes = o_sph_j1c(shell_qr);
    const oreal shell_volume = O_4PI_3 * cb(r + t);
    f += shell_volume * shell_bes * shell_contrast;
    *f1 = 1.0e-2 * f;
    *f2 = 1.0e-4 * f * f;
}

/* ---- parallelepiped (parallelepiped.c:1-150):
 *      p = sld, sld_solvent, length_a, length_b, length_c, theta, phi, psi ---- */
static oreal ppd_vol(const oreal *p) { return p[2] * p[3] * p[4]; }
static oreal ppd_excluded(oreal a, oreal b, oreal c)
{
    const oreal len[3] = {a, b, c};
    const oreal lmax = fmax(len[0], fmax(len[1], len[2]));
    oreal l1 = lmax, l2 = lmax;
    const oreal l3 = lmax;
    for (int i = 0; i < 3; i++) {
        if (len[i] < l1) { l2 = l1; l1 = len[i]; }
        else if (len[i] < l2) l2 = len[i];
    }
    oreal r, l;
    if (l2 - l1 > l3 - l2) { r = sqrt(l2 * l3 / O_PI); l = l1; }
    else { r = sqrt(l1 * l2 / O_PI); l = l3; }
    return 0.5 * cbrt(0.75 * r * (2.0 * r * l + (r + l) * (O_PI * r + l)));
}
static oreal ppd_reff(int mode, const oreal *p)
{
    const oreal a = p[2], b = p[3], c = p[4];
    switch (mode) {
    default:
    case 1: return ppd_excluded(a, b, c);
    case 2: return cbrt(a * b * c / O_4PI_3);
    case 3: return 0.5 * a;
    case 4: return 0.5 * b;
    case 5: return 0.5 * c;
    case 6: return sqrt(a * b / O_PI);
    case 7: return 0.5 * sqrt(a * a + b * b);
    case 8: return 0.5 * sqrt(a * a + b * b + c * c);
    }
}
static void ppd_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal a = p[2], b = p[3], c = p[4];
    const oreal mu = 0.5 * q * b;
    const oreal as = a / b, cs = c / b;
    const oreal *zo, *wo;
    const int no = quad_select_nested(quad_n, q * fmax(sqrt(a * a + b * b), c) / 2, O_ADAPTIVE_MAX_OUTER, &zo, &wo);
    oreal o1 = 0.0, o2 = 0.0;
    for (int i = 0; i < no; i++) {
        const oreal sigma = 0.5 * (zo[i] + 1.0);
        const oreal mp = mu * sqrt(1.0 - sigma * sigma);
        const oreal *zi, *wi;
        const int ni = quad_select_nested(quad_n, mp * fmax(as, 1.0), no, &zi, &wi);
        oreal i1 = 0.0, i2 = 0.0;
        for (int j = 0; j < ni; j++) {
            const oreal uu = 0.5 * (zi[j] + 1.0);
            const oreal su = sin(O_PI / 2 * uu), cu = cos(O_PI / 2 * uu);
            const oreal fq = o_sinx_x(mp * su * as) * o_sinx_x(mp * cu);
            i1 += wi[j] * fq;
            i2 += wi[j] * fq * fq;
        }
        i1 *= 0.5;
        i2 *= 0.5;
        const oreal si = o_sinx_x(mu * cs * sigma);
        o1 += wo[i] * i1 * si;
        o2 += wo[i] * i2 * si * si;
    }
    o1 *= 0.5;
    o2 *= 0.5;
    const oreal s = (p[0] - p[1]) * ppd_vol(p);
    *f1 = 1.0e-2 * s * o1;
    *f2 = 1.0e-4 * s * s * o2;
}
static oreal ppd_iqabc(oreal qa, oreal qb, oreal qc, const oreal *p)
{
    const oreal f = ppd_vol(p) * (p[0] - p[1]) * o_sinx_x(0.5 * p[2] * qa) * o_sinx_x(0.5 * p[3] * qb)
                     * o_sinx_x(0.5 * p[4] * qc);
    return 1.0e-4 * f * f;
}

/* ---- triaxial ellipsoid (triaxial_ellipsoid.c:1-160): p = sld, sld_solvent,
 *      radius_equat_minor, radius_equat_major, radius_polar, theta, phi, psi ---- */
static oreal tri_vol(const oreal *p) { return O_4PI_3 * p[2] * p[3] * p[4]; }
static oreal tri_curvature(oreal ra, oreal rb, oreal rc)
{
    /* the reference compares the 0/1 result of ra == rb with rc */
    if ((ra == rb) == rc) return rc;
    if (ra * rb * rc == 0.) return 0.;
    const oreal rad[3] = {ra, rb, rc};
    const oreal rmax = fmax(rad[0], fmax(rad[1], rad[2]));
    oreal r1 = rmax, r2 = rmax, r3 = rmax;
    for (int i = 0; i < 3; i++) {
        if (rad[i] < r1) { r3 = r2; r2 = r1; r1 = rad[i]; }
        else if (rad[i] < r2) r2 = rad[i];
    }
    oreal re, rp;
    if (r2 - r1 > r3 - r2) { re = sqrt(r2 * r3); rp = r1; }
    else { re = sqrt(r1 * r2); rp = r3; }
    const oreal ratio = (rp < re ? rp / re : re / rp);
    const oreal e1 = sqrt(1.0 - ratio * ratio);
    const oreal b1 = 1.0 + asin(e1) / (e1 * ratio);
    const oreal bL = (1.0 + e1) / (1.0 - e1);
    const oreal b2 = 1.0 + 0.5 * ratio * ratio / e1 * log(bL);
    const oreal delta = 0.75 * b1 * b2;
    const oreal ddd = 2.0 * (delta + 1.0) * rp * re * re;
    return 0.5 * cbrt(ddd);
}
static oreal tri_reff(int mode, const oreal *p)
{
    const oreal ra = p[2], rb = p[3], rc = p[4];
    switch (mode) {
    default:
    case 1: return tri_curvature(ra, rb, rc);
    case 2: return cbrt(ra * rb * rc);
    case 3: { const oreal m = (ra < rb ? ra : rb); return (m < rc ? m : rc); }
    case 4: { const oreal m = (ra < rb ? rb : ra); return (m > rc ? m : rc); }
    }
}
static void tri_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal lo = fmin(p[2], p[3]), hi = fmax(p[2], p[3]);
    const oreal mid = fmax(lo, p[4]);
    const oreal minor = fmin(lo, p[4]), major = fmin(hi, mid), polar = fmax(hi, mid);
    const oreal pa = sq(minor / major) - 1.0;
    const oreal pc = sq(polar / major) - 1.0;
    const oreal qr_inner = fmax(q * minor, q * major);
    const oreal qr_max = fmax(qr_inner, q * polar);
    const oreal *zo, *wo, *zi, *wi;
    const int no = quad_select_nested(quad_n, qr_max, O_ADAPTIVE_MAX_OUTER, &zo, &wo);
    const int ni = quad_select_nested(quad_n, qr_inner, no, &zi, &wi);
    oreal o1 = 0.0, o2 = 0.0;
    for (int i = 0; i < no; i++) {
        const oreal phi = zo[i] * O_PI_4 + O_PI_4;
        const oreal pas = pa * sq(sin(phi));
        oreal i1 = 0.0, i2 = 0.0;
        for (int j = 0; j < ni; j++) {
            const oreal usq = sq(zi[j] * 0.5 + 0.5);
            const oreal r = major * sqrt(pas * (1.0 - usq) + 1.0 + pc * usq);
            const oreal fq = o_sph_j1c(q * r);
            i1 += wi[j] * fq;
            i2 += wi[j] * fq * fq;
        }
        o1 += wo[i] * i1;
        o2 += wo[i] * i2;
    }
    o1 *= 0.25;
    o2 *= 0.25;
    const oreal s = (p[0] - p[1]) * O_4PI_3 * minor * major * polar;
    *f1 = 1.0e-2 * s * o1;
    *f2 = 1.0e-4 * s * s * o2;
}
static oreal tri_iqabc(oreal qa, oreal qb, oreal qc, const oreal *p)
{
    const oreal qr = sqrt(sq(p[2] * qa) + sq(p[3] * qb) + sq(p[4] * qc));
    const oreal f = o_sph_j1c(qr);
    return 1.0e-4 * sq(tri_vol(p) * (p[0] - p[1]) * f);
}

/* ---- paracrystals (bcc/fcc/sc_paracrystal.c with lib/sphere_form.c):
 *      p = dnn, d_factor, radius, sld, sld_solvent, theta, phi, psi ---- */
static oreal pc_zq(int lat, oreal qa, oreal qb, oreal qc, oreal dnn, oreal d_factor)
{
    oreal a1, a2, a3, d_a;
    if (lat == 0) {        /* bcc */
        a1 = (-qa + qb + qc) / 2.0; a2 = (+qa - qb + qc) / 2.0; a3 = (+qa + qb - qc) / 2.0;
        d_a = dnn / sqrt(0.75);
    } else if (lat == 1) { /* fcc */
        a1 = (qa + qb) / 2.0; a2 = (qa + qc) / 2.0; a3 = (qb + qc) / 2.0;
        d_a = dnn * sqrt(2);
    } else {               /* sc */
        a1 = qa; a2 = qb; a3 = qc;
        d_a = dnn;
    }
    const oreal arg = -0.5 * sq(dnn * d_factor) * (a1 * a1 + a2 * a2 + a3 * a3);
    const oreal exp_arg = exp(arg);
    return -cb(expm1(2.0 * arg))
        / (((exp_arg - 2.0 * cos(d_a * a1)) * exp_arg + 1.0)
           * ((exp_arg - 2.0 * cos(d_a * a2)) * exp_arg + 1.0)
           * ((exp_arg - 2.0 * cos(d_a * a3)) * exp_arg + 1.0));
}
static oreal pc_vfrac(int lat, oreal radius, oreal dnn)
{
    if (lat == 0) return 2.0 * (O_4PI_3 * cb(sqrt(0.75) * radius / dnn));
    if (lat == 1) return 4.0 * (O_4PI_3 * cb(O_SQRT1_2 * radius / dnn));
    return O_4PI_3 * cb(radius / dnn);
}
static oreal pc_sphere_form(oreal q, const oreal *p)
{
    const oreal fq = O_4PI_3 * cb(p[2]) * o_sph_j1c(q * p[2]);
    return 1.0e-4 * sq((p[3] - p[4]) * fq);
}
static oreal pc_vol(const oreal *p) { return O_4PI_3 * cb(p[2]); }
static oreal pc_reff(int mode, const oreal *p) { (void)mode; (void)p; return 0.0; }
static oreal pc_iq(int lat, oreal q, const oreal *p, int quad_n)
{
    const oreal *z, *w;
    const int n = gl_table(quad_n > 0 ? quad_n : 150, &z, &w);
    const oreal tm = lat == 2 ? O_PI_4 : O_PI / 2, tb = tm;
    const oreal pm = lat == 2 ? O_PI_4 : O_PI, pb = pm;
    oreal outer = 0.0;
    for (int i = 0; i < n; i++) {
        oreal inner = 0.0;
        const oreal theta = z[i] * tm + tb;
        const oreal st = sin(theta), ct = cos(theta);
        const oreal qc = q * ct, qab = q * st;
        for (int j = 0; j < n; j++) {
            const oreal phi = z[j] * pm + pb;
            const oreal sp = sin(phi), cp = cos(phi);
            inner += w[j] * pc_zq(lat, qab * cp, qab * sp, qc, p[0], p[1]);
        }
        inner *= pm;
        outer += w[i] * inner * st;
    }
    outer *= tm;
    const oreal zq = lat == 2 ? outer / (O_PI / 2) : outer / (4.0 * O_PI);
    return pc_vfrac(lat, p[2], p[0]) * pc_sphere_form(q, p) * zq;
}
static oreal pc_iqabc(int lat, oreal qa, oreal qb, oreal qc, const oreal *p)
{
    const oreal q = sqrt(qa * qa + qb * qb + qc * qc);
    return pc_vfrac(lat, p[2], p[0]) * pc_sphere_form(q, p) * pc_zq(lat, qa, qb, qc, p[0], p[1]);
}
#define PC_FNS(L, NAME) \
    static void NAME##_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2) \
    { *f1 = 0.0; *f2 = pc_iq(L, q, p, quad_n); } \
    static oreal NAME##_iqabc(oreal qa, oreal qb, oreal qc, const oreal *p) \
    { return pc_iqabc(L, qa, qb, qc, p); }
PC_FNS(0, bcc)
PC_FNS(1, fcc)
PC_FNS(2, scp)

/* ---- hollow cylinder (hollow_cylinder.c:1-140): p = radius, thickness,
 *      length, sld, sld_solvent, theta, phi ---- */
static oreal hcy_vol(const oreal *p) { return O_PI * p[2] * sq(p[0] + p[1]); }
static oreal hcy_svol(const oreal *p) { return O_PI * p[2] * (sq(p[0] + p[1]) - p[0] * p[0]); }
static oreal hcy_reff(int mode, const oreal *p)
{
    const oreal ro = p[0] + p[1], l = p[2];
    switch (mode) {
    default:
    case 1: return 0.5 * cbrt(0.75 * ro * (2.0 * ro * l + (ro + l) * (O_PI * ro + l)));
    case 2: return cbrt(O_PI * sq(p[0] + p[1]) * l / O_4PI_3);
    case 3: return ro;
    case 4: return 0.5 * l;
    case 5: return (ro < 0.5 * l ? ro : 0.5 * l);
    case 6: return (ro > 0.5 * l ? ro : 0.5 * l);
    case 7: return sqrt(sq(p[0] + p[1]) + 0.25 * sq(l));
    }
}
static oreal hcy_amp(oreal qab, oreal qc, const oreal *p)
{
    const oreal r = p[0], t = p[1], l = p[2];
    const oreal lam1 = o_2j1x_x((r + t) * qab);
    const oreal lam2 = o_2j1x_x(r * qab);
    const oreal gamma_sq = sq(r / (r + t));
    const oreal psi = (lam1 - gamma_sq * lam2) / (1.0 - gamma_sq);
    return psi * o_sinx_x(0.5 * l * qc);
}
static void hcy_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal *z, *w;
    const int n = quad_select(quad_n, q * fmax(p[0] + p[1], 0.5 * p[2]), &z, &w);
    oreal t1 = 0.0, t2 = 0.0;
    for (int i = 0; i < n; i++) {
        const oreal ct = 0.5 * (z[i] * (1.0 - 0.0) + 0.0 + 1.0);
        const oreal st = sqrt(1.0 - ct * ct);
        const oreal form = hcy_amp(q * st, q * ct, p);
        t1 += w[i] * form;
        t2 += w[i] * form * form;
    }
    t1 *= 0.5 * (1.0 - 0.0);
    t2 *= 0.5 * (1.0 - 0.0);
    const oreal s = (p[3] - p[4]) * hcy_svol(p);
    *f1 = 1e-2 * s * t1;
    *f2 = 1e-4 * s * s * t2;
}
static oreal hcy_iqac(oreal qab, oreal qc, const oreal *p)
{
    const oreal form = hcy_amp(qab, qc, p);
    const oreal s = (p[3] - p[4]) * hcy_svol(p);
    return 1.0e-4 * sq(s * form);
}

/* ---- vesicle (vesicle.c:1-50): p = sld, sld_solvent, volfraction, radius, thickness ---- */
static oreal ves_vol(const oreal *p) { return O_4PI_3 * cb(p[3] + p[4]); }
static oreal ves_svol(const oreal *p) { return O_4PI_3 * (cb(p[3] + p[4]) - cb(p[3])); }
static oreal ves_reff(int mode, const oreal *p) { (void)mode; return p[3] + p[4]; }
static void ves_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    oreal vol, contrast, f;
    contrast = p[1] - p[0];
    vol = O_4PI_3 * cb(p[3]);
    f = vol * o_sph_j1c(q * p[3]) * contrast;
    contrast = p[0] - p[1];
    vol = O_4PI_3 * cb(p[3] + p[4]);
    f += vol * o_sph_j1c(q * (p[3] + p[4])) * contrast;
    *f1 = 1e-2 * sqrt(p[2]) * f;
    *f2 = 1.0e-4 * p[2] * f * f;
}

/* ---- fuzzy sphere (fuzzy_sphere.c:1-35): p = sld, sld_solvent, radius, fuzziness ---- */
static oreal fuz_vol(const oreal *p) { return O_4PI_3 * cb(p[2]); }
static oreal fuz_reff(int mode, const oreal *p) { return mode == 2 ? p[2] + p[3] : p[2]; }
static void fuz_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal bes = o_sph_j1c(q * p[2]);
    const oreal qf = exp(-0.5 * sq(q * p[3]));
    const oreal form = (p[0] - p[1]) * fuz_vol(p) * bes * qf;
    *f1 = 1.0e-2 * form;
    *f2 = 1.0e-4 * form * form;
}

/* ---- core_shell_ellipsoid (core_shell_ellipsoid.c:1-170): p = radius_equat_core,
 * x_core, thick_shell, x_polar_shell, sld_core, sld_shell, sld_solvent, theta, phi.
 * Core radii (re, re x_core); shell radii (re + t, re x_core + t x_polar_shell). ---- */
static oreal cse_amp(oreal qab, oreal qc, oreal re_c, oreal rp_c, oreal re_s, oreal rp_s,
                     oreal d_core_shell, oreal d_shell_solvent)
{
    const oreal x_c = sqrt(sq(re_c * qab) + sq(rp_c * qc));
    const oreal x_s = sqrt(sq(re_s * qab) + sq(rp_s * qc));
    return o_sph_j1c(x_c) * (O_4PI_3 * re_c * re_c * rp_c) * d_core_shell
         + o_sph_j1c(x_s) * (O_4PI_3 * re_s * re_s * rp_s) * d_shell_solvent;
}
static oreal cse_vol(const oreal *p)
{
    const oreal re_s = p[0] + p[2];
    const oreal rp_s = p[0] * p[1] + p[2] * p[3];
    return O_4PI_3 * re_s * re_s * rp_s;
}
static oreal cse_reff(int mode, const oreal *p)
{
    const oreal re = p[0] + p[2], rp = p[0] * p[1] + p[2] * p[3];
    switch (mode) {
    default:
    case 1: {
        if (1.0 == p[1] && 1.0 == p[3]) return re;
        if (re * rp == 0.) return 0.;
        /* Isihara (1950) eq. 26 for the outer spheroid */
        const oreal ratio = (rp < re ? rp / re : re / rp);
        const oreal e1 = sqrt(1.0 - ratio * ratio);
        const oreal b1 = 1.0 + asin(e1) / (e1 * ratio);
        const oreal bL = (1.0 + e1) / (1.0 - e1);
        const oreal b2 = 1.0 + 0.5 * ratio * ratio / e1 * log(bL);
        const oreal delta = 0.75 * b1 * b2;
        return 0.5 * cbrt(2.0 * (delta + 1.0) * rp * re * re);
    }
    case 2: return cbrt(cse_vol(p) / O_4PI_3);
    case 3: return (rp < re ? rp : re);
    case 4: return (rp > re ? rp : re);
    }
}
static void cse_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal d1 = p[4] - p[5], d2 = p[5] - p[6];
    const oreal rp_c = p[0] * p[1], re_s = p[0] + p[2], rp_s = p[0] * p[1] + p[2] * p[3];
    const oreal *z, *w;
    const int n = quad_select(quad_n, q * fmax(re_s, rp_s), &z, &w);
    oreal t1 = 0.0, t2 = 0.0;
    for (int i = 0; i < n; i++) {
        const oreal u = z[i] * 0.5 + 0.5;
        const oreal s = sqrt(1.0 - u * u);
        const oreal f = cse_amp(q * s, q * u, p[0], rp_c, re_s, rp_s, d1, d2);
        t1 += w[i] * f;
        t2 += w[i] * f * f;
    }
    *f1 = 1.0e-2 * (t1 * 0.5);
    *f2 = 1.0e-4 * (t2 * 0.5);
}
static oreal cse_iqac(oreal qab, oreal qc, const oreal *p)
{
    const oreal f = cse_amp(qab, qc, p[0], p[0] * p[1], p[0] + p[2], p[0] * p[1] + p[2] * p[3],
                            p[4] - p[5], p[5] - p[6]);
    return 1.0e-4 * f * f;
}

/* ---- elliptical_cylinder (elliptical_cylinder.c:1-150): p = radius_minor,
 * axis_ratio, length, sld, sld_solvent, theta, phi, psi.  1D: nested rules,
 * outer cos(alpha) = (z + 1)/2 on [0, 1], inner angle (z + 1) pi/2 on [0, pi]
 * around the elliptical cross-section. ---- */
static oreal ecy_vol(const oreal *p) { return O_PI * p[0] * p[0] * p[1] * p[2]; }
static oreal ecy_reff(int mode, const oreal *p)
{
    const oreal rm = p[0], ratio = p[1], l = p[2];
    const oreal rmin = ratio > 1.0 ? rm : ratio * rm, rmax = ratio < 1.0 ? rm : ratio * rm;
    switch (mode) {
    default:
    case 1: {
        const oreal r = sqrt(rm * rm * ratio);
        return 0.5 * cbrt(0.75 * r * (2.0 * r * l + (r + l) * (O_PI * r + l)));
    }
    case 2: return cbrt(ecy_vol(p) / O_4PI_3);
    case 3: return 0.5 * rm * (1.0 + ratio);
    case 4: return rmin;
    case 5: return rmax;
    case 6: return sqrt(rm * rm * ratio);
    case 7: return 0.5 * l;
    case 8: return (rmin < 0.5 * l ? rmin : 0.5 * l);
    case 9: return (rmax > 0.5 * l ? rmax : 0.5 * l);
    case 10: {
        const oreal r = ratio > 1.0 ? rm * ratio : rm;
        return sqrt(r * r + 0.25 * l * l);
    }
    }
}
static void ecy_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal rm = p[0], rmaj = p[1] * p[0], l = p[2];
    const oreal ra = 0.5 * (sq(rmaj) + sq(rm)), rb = 0.5 * (sq(rmaj) - sq(rm));
    const oreal *zo, *wo, *zi, *wi;
    const int no = quad_select_nested(quad_n, q * fmax(0.5 * l, fmax(rm, rm * p[1])), O_ADAPTIVE_MAX_OUTER,
                                      &zo, &wo);
    const int ni = quad_select_nested(quad_n, q * fmax(rm, rmaj), no, &zi, &wi);
    oreal o1 = 0.0, o2 = 0.0;
    for (int i = 0; i < no; i++) {
        const oreal u = (zo[i] + 1.0) / 2.0;
        const oreal s = sqrt(1.0 - u * u);
        oreal i1 = 0.0, i2 = 0.0;
        for (int j = 0; j < ni; j++) {
            const oreal th = (zi[j] * O_PI + O_PI) / 2.0;
            const oreal be = o_2j1x_x(q * (s * sqrt(ra - rb * cos(th))));
            i1 += wi[j] * be;
            i2 += wi[j] * be * be;
        }
        i1 *= 0.5 * O_PI;
        i2 *= 0.5 * O_PI;
        const oreal si = o_sinx_x(q * 0.5 * l * u);
        o1 += wo[i] * i1 * si;
        o2 += wo[i] * i2 * si * si;
    }
    o1 *= 0.5 / O_PI;
    o2 *= 0.5 / O_PI;
    const oreal sv = (p[3] - p[4]) * ecy_vol(p);
    *f1 = 1.0e-2 * sv * o1;
    *f2 = 1.0e-4 * sv * sv * o2;
}
static oreal ecy_iqabc(oreal qa, oreal qb, oreal qc, const oreal *p)
{
    const oreal be = o_2j1x_x(p[0] * sqrt(sq(p[1] * qb) + sq(qa)));
    const oreal si = o_sinx_x(qc * 0.5 * p[2]);
    return 1.0e-4 * sq((p[3] - p[4]) * ecy_vol(p) * (be * si));
}

/* ---- core_shell_bicelle (core_shell_bicelle.c:1-160): p = radius, thick_rim,
 * thick_face, length, sld_core, sld_face, sld_rim, sld_solvent, theta, phi.
 * A core cylinder capped by face layers and wrapped by a rim. ---- */
static oreal csb_vol(const oreal *p) { return O_PI * sq(p[0] + p[1]) * (p[3] + 2.0 * p[2]); }
static oreal csb_reff(int mode, const oreal *p)
{
    const oreal r = p[0] + p[1], l = p[3] + 2.0 * p[2];
    switch (mode) {
    default:
    case 1: return 0.5 * cbrt(0.75 * r * (2.0 * r * l + (r + l) * (O_PI * r + l)));
    case 2: return cbrt(csb_vol(p) / O_4PI_3);
    case 3: return p[0] + p[1];
    case 4: return 0.5 * p[3] + p[2];
    case 5: return sqrt(r * r + 0.25 * l * l);
    }
}
static oreal csb_amp(oreal qab, oreal qc, const oreal *p)
{
    const oreal r = p[0], tr = p[1], tf = p[2], h = 0.5 * p[3];
    const oreal be1 = o_2j1x_x(r * qab), be2 = o_2j1x_x((r + tr) * qab);
    const oreal si1 = o_sinx_x(h * qc), si2 = o_sinx_x((h + tf) * qc);
    const oreal v1 = O_PI * sq(r) * 2.0 * h;
    const oreal v2 = O_PI * sq(r + tr) * 2.0 * (h + tf);
    const oreal v3 = O_PI * sq(r) * 2.0 * (h + tf);
    return v1 * (p[4] - p[5]) * si1 * be1 + v2 * (p[6] - p[7]) * si2 * be2 + v3 * (p[5] - p[6]) * si2 * be1;
}
static void csb_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    const oreal *z, *w;
    const int n = quad_select(quad_n, q * fmax(p[0] + p[1], p[3] + p[2]), &z, &w);
    oreal t1 = 0.0, t2 = 0.0;
    for (int i = 0; i < n; i++) {
        const oreal th = (z[i] + 1.0) * O_PI_4;
        const oreal s = sin(th), c = cos(th);
        const oreal f = csb_amp(q * s, q * c, p);
        t1 += w[i] * f * s;
        t2 += w[i] * f * f * s;
    }
    *f1 = 1.0e-2 * (t1 * O_PI_4);
    *f2 = 1.0e-4 * (t2 * O_PI_4);
}
static oreal csb_iqac(oreal qab, oreal qc, const oreal *p)
{
    const oreal f = csb_amp(qab, qc, p);
    return 1.0e-4 * f * f;
}

/* ---- models defined by Iq alone (no <F>); volume 1 unless stated ---- */
/* gamma extended below 1 (lib/sas_gamma.c:140-153) */
static oreal o_gamma(oreal x) { return x < 0. ? O_PI / tgamma(1. - x) / sin(O_PI * x) : tgamma(x + 1) / x; }
static oreal one_vol(const oreal *p) { (void)p; return 1.0; }
static oreal zero_reff(int mode, const oreal *p) { (void)mode; (void)p; return 0.0; }

/* guinier.c: p = rg */
static void gui_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal e = fabs(p[0]) * p[0] * q * q / 3.0;
    *f1 = 0.0;
    *f2 = exp(-e);
}
/* gaussian_peak.c: p = peak_pos, sigma */
static void gpk_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal x = (q - p[0]) / p[1];
    *f1 = 0.0;
    *f2 = exp(-0.5 * x * x);
}
/* lorentz.c: p = cor_length */
static void lor_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal d = 1 + (q * p[0]) * (q * p[0]);
    *f1 = 0.0;
    *f2 = 1 / d;
}
/* dab.c: p = cor_length */
static void dab_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    *f1 = 0.0;
    *f2 = cb(p[0]) / sq(1 + sq(q * p[0]));
}
/* gel_fit.c: p = guinier_scale, lorentz_scale, rg, fractal_dim, cor_length */
static void gel_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    oreal lor = sq(q * p[4]);
    lor = 1.0 + ((p[3] + 1.0) / 3.0) * lor;
    lor = pow(lor, p[3] / 2.0);
    oreal ex = sq(q * p[2]);
    ex = exp(-ex / 3.0);
    *f1 = 0.0;
    *f2 = p[1] / lor + p[0] * ex;
}
/* mass_fractal.c: p = radius, fractal_dim_mass, cutoff_length */
static void mfr_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal pq = sq(o_sph_j1c(q * p[0]));
    oreal s;
    if (q > 0. && p[1] > 1.) {
        const oreal dm = p[1] - 1.0;
        const oreal t1 = o_gamma(dm) * sin(dm * atan(q * p[2]));
        const oreal t2 = pow(p[2], dm);
        const oreal t3 = pow(1.0 + sq(q * p[2]), -0.5 * dm);
        s = t1 * t2 * t3 / q;
    } else if (q > 0.) {
        s = atan(q * p[2]) / q;
    } else if (p[1] > 1.) {
        s = pow(p[2], p[1]) * o_gamma(p[1]);
    } else {
        s = p[2];
    }
    *f1 = 0.0;
    *f2 = pq * s;
}
/* surface_fractal.c: p = radius, fractal_dim_surf, cutoff_length */
static void sfr_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal pq = sq(o_sph_j1c(q * p[0]));
    const oreal m = 5.0 - p[1];
    const oreal s = o_gamma(m) * pow(p[2], m) * pow(1.0 + sq(q * p[2]), -0.5 * m) * sin(-m * atan(q * p[2])) / q;
    const oreal r = pq * s;
    *f1 = 0.0;
    *f2 = r > 0. ? r : 0.;
}
/* mass_surface_fractal.c: p = fractal_dim_mass, fractal_dim_surf, rg_cluster, rg_primary */
static void msf_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal dm = 0.5 * p[0];
    const oreal dt = 0.5 * (6.0 - (p[0] + p[1]));
    const oreal t1 = dm == 0. ? 1.0 : pow(1.0 + sq(q * p[2]) / (3.0 * dm), -dm);
    const oreal t2 = dt == 0. ? 1.0 : pow(1.0 + sq(q * p[3]) / (3.0 * dt), -dt);
    *f1 = 0.0;
    *f2 = t1 * t2;
}
/* star_polymer.c (double-precision series cutoff 0.03): p = rg_squared, arms */
static void star_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    *f1 = 0.0;
    if (q == 0.) { *f2 = 1.0; return; }
    const int n = (int)(p[1] + 0.5);
    const oreal u2 = p[0] * q * q;
    const oreal v = (u2 * n) / (3 * n - 2);
    oreal t1;
    if (u2 < 0.03) t1 = 1. + v * (-1. / 3. + v * (1. / 12. + v * (-1. / 60. + v * (1. / 360. + v * (-1. / 2520)))));
    else t1 = 2.0 * (v + expm1(-v)) / (v * v);
    const oreal t2 = sq(expm1(-v) / v);
    *f2 = (t1 - t2) / n + t2;
}
/* mono_gauss_coil.c (double precision: Pade below x = 0.5): p = i_zero, rg */
static oreal mgc_reff(int mode, const oreal *p)
{
    switch (mode) {
    default:
    case 1: return p[1];
    case 2: return 2.0 * p[1];
    case 3: return 3.0 * p[1];
    case 4: return sqrt(5.0 / 3.0) * p[1];
    }
}
static void mgc_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal x = sq(q * p[1]);
    oreal g;
    if (x < 0.5) {
        const oreal A1 = 1. / 12., A2 = 2. / 99., A3 = 1. / 2640., A4 = 1. / 23760., A5 = -1. / 1995840.;
        const oreal B1 = 5. / 12., B2 = 5. / 66., B3 = 1. / 132., B4 = 1. / 2376., B5 = 1. / 95040.;
        g = (((((A5 * x + A4) * x + A3) * x + A2) * x + A1) * x + 1.)
          / (((((B5 * x + B4) * x + B3) * x + B2) * x + B1) * x + 1.);
    } else {
        g = 2.0 * (expm1(-x) + x) / (x * x);
    }
    *f1 = 0.0;
    *f2 = p[0] * g;
}
/* fractal.c with lib/fractal_sq.c: p = volfraction, radius, fractal_dim,
 * cor_length, sld_block, sld_solvent; volume 4/3 pi r^3 */
static oreal frc_vol(const oreal *p) { return O_4PI_3 * cb(p[1]); }
static void frc_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal r = p[1], D = p[2], xi = p[3];
    oreal term;
    if (q == 0.) {
        term = pow(xi / r, D) * o_gamma(D + 1.);
    } else if (D == 0.) {
        term = 1.0;
    } else if (D == 1.) {
        term = atan(q * xi) / (q * r);
    } else {
        const oreal dm = D - 1.0;
        const oreal t1 = o_gamma(D + 1.) / dm * sin(dm * atan(q * xi));
        const oreal t2 = pow(q * r, -D);
        const oreal t3 = pow(1.0 + 1.0 / sq(q * xi), -0.5 * dm);
        term = t1 * t2 * t3;
    }
    const oreal s = 1.0 + term;
    const oreal fq = frc_vol(p) * (p[4] - p[5]) * o_sph_j1c(q * r);
    *f1 = 0.0;
    *f2 = 1.e-4 * p[0] * s * fq * fq;
}

/* ---- fractal_core_shell.c with lib/core_shell.c: p = radius, thickness,
 * sld_core, sld_shell, sld_solvent, volfraction, fractal_dim, cor_length ---- */
static oreal o_fractal_sq(oreal q, oreal radius, oreal D, oreal xi)
{
    oreal term;
    if (q == 0.) term = pow(xi / radius, D) * o_gamma(D + 1.);
    else if (D == 0.) term = 1.0;
    else if (D == 1.) term = atan(q * xi) / (q * radius);
    else {
        const oreal dm = D - 1.0;
        const oreal t1 = o_gamma(D + 1.) / dm * sin(dm * atan(q * xi));
        const oreal t2 = pow(q * radius, -D);
        const oreal t3 = pow(1.0 + 1.0 / sq(q * xi), -0.5 * dm);
        term = t1 * t2 * t3;
    }
    return 1.0 + term;
}
static oreal fcs_vol(const oreal *p) { return O_4PI_3 * cb(p[0] + p[1]); }
static void fcs_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal s = o_fractal_sq(q, p[0] + p[1], p[6], p[7]);
    oreal f = O_4PI_3 * cb(p[0]) * o_sph_j1c(q * p[0]) * (p[2] - p[3]);
    f += O_4PI_3 * cb(p[0] + p[1]) * o_sph_j1c(q * (p[0] + p[1])) * (p[3] - p[4]);
    *f1 = 0.0;
    *f2 = 1.0e-4 * p[5] * s * f * f;
}
/* ---- lamellar_hg.c: p = length_tail, length_head, sld, sld_head, sld_solvent ---- */
static void lhg_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal qsq = q * q, drh = p[3] - p[4], drt = p[2] - p[4], qT = q * p[0];
    oreal Pq = drh * (sin(q * (p[1] + p[0])) - sin(qT)) + drt * sin(qT);
    Pq *= Pq;
    Pq *= 4.0 / (qsq);
    oreal inten = 2.0e-4 * O_PI * Pq / qsq;
    inten /= 2.0 * (p[1] + p[0]);
    *f1 = 0.0;
    *f2 = inten;
}
/* ---- Caille S(q) of lamellar_stack_caille.c / lamellar_hg_stack_caille.c ---- */
static oreal o_caille_sq(oreal q, int n, oreal dd, oreal Cp)
{
    const oreal euler = 0.577215664901533;
    oreal s = 0.0;
    for (int ii = 1; ii < n; ii++) {
        const oreal alpha = Cp / 4.0 / O_PI / O_PI * (log(O_PI * ii) + euler);
        const oreal t2 = 2.0 * q * q * dd * dd * alpha;
        oreal t = 1.0 - (oreal)ii / (oreal)n;
        t *= cos(dd * q * ii);
        t *= exp(-t2 / 2.0);
        s += t;
    }
    s *= 2.0;
    s += 1.0;
    return s;
}
/* lamellar_stack_caille.c: p = thickness, Nlayers, d_spacing, Caille_parameter, sld, sld_solvent */
static void lsc_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal contr = p[4] - p[5];
    const oreal Pq = 2.0 * contr * contr / q / q * (1.0 - cos(q * p[0]));
    const oreal Sq = o_caille_sq(q, (int)(p[1] + 0.5), p[2], p[3]);
    oreal inten = 2.0 * O_PI * Pq * Sq / (p[2] * q * q);
    inten *= 1.0e-04;
    *f1 = 0.0;
    *f2 = inten;
}
/* lamellar_hg_stack_caille.c: p = length_tail, length_head, Nlayers,
 * d_spacing, Caille_parameter, sld, sld_head, sld_solvent */
static void lhc_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    oreal Pq = (p[6] - p[7]) * (sin(q * (p[1] + p[0])) - sin(q * p[0])) + (p[5] - p[7]) * sin(q * p[0]);
    Pq *= Pq;
    Pq *= 4.0 / (q * q);
    const oreal Sq = o_caille_sq(q, (int)(p[2] + 0.5), p[3], p[4]);
    oreal inten = 2.0 * O_PI * Pq * Sq / (p[3] * q * q);
    inten *= 1.0e-04;
    *f1 = 0.0;
    *f2 = inten;
}
/* ---- lamellar_stack_paracrystal.c: p = thickness, Nlayers, d_spacing, sigma_d, sld, sld_solvent ---- */
static oreal lsp_an(oreal ww, oreal q, oreal davg, int n)
{
    oreal an = 4.0 * ww * ww - 2.0 * (ww * ww * ww + ww) * cos(q * davg);
    an -= 4.0 * pow(ww, (oreal)(n + 2)) * cos((oreal)n * q * davg);
    an += 2.0 * pow(ww, (oreal)(n + 3)) * cos((oreal)(n - 1) * q * davg);
    an += 2.0 * pow(ww, (oreal)(n + 1)) * cos((oreal)(n + 1) * q * davg);
    return an;
}
static oreal lsp_sn(oreal ww, oreal q, oreal davg, int n, oreal an)
{
    return an / ((oreal)n * sq(1.0 + ww * ww - 2.0 * ww * cos(q * davg)));
}
static void lsp_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const int n1 = (int)(p[1]), n2 = n1 + 1;
    const oreal xn = (oreal)n2 - p[1];
    const oreal ww = exp(-0.5 * sq(q * p[3] * p[2]));
    oreal Z = xn * lsp_sn(ww, q, p[2], n1, lsp_an(ww, q, p[2], n1));
    Z += (1.0 - xn) * lsp_sn(ww, q, p[2], n2, lsp_an(ww, q, p[2], n2));
    Z += (1.0 - ww * ww) / (1.0 + ww * ww - 2.0 * ww * cos(q * p[2]));
    const oreal Pbil = sq(o_sinx_x(q * (p[0] / 2.0)));
    const oreal contr = p[4] - p[5];
    const oreal inten = 2.0 * O_PI * contr * contr * Pbil * Z / (q * q);
    *f1 = 0.0;
    *f2 = 1.0e-4 * inten;
}
/* ---- linear_pearls.c: p = radius, edge_sep, num_pearls, sld, sld_solvent ---- */
static void lpe_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const int n = (int)(p[2] + 0.5);
    const oreal vol = O_4PI_3 * cb(p[0]);
    const oreal tot = n * vol;
    const oreal m = (p[3] - p[4]) * vol;
    const oreal sep = p[1] + 2.0 * p[0];
    const oreal psi = o_sph_j1c(q * p[0]);
    oreal s = (oreal)n;
    for (int k = 1; k < n; k++) s += 2.0 * (n - k) * o_sinx_x(q * sep * k);
    *f1 = 0.0;
    *f2 = 1.0e-4 * s * sq(m * psi) / tot;
}
/* ---- polymer_micelle.c: p = ndensity, v_core, v_corona, sld_solvent,
 * sld_core, sld_corona, radius_core, rg, d_penetration, n_aggreg ---- */
static void pmi_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal bc = p[1] * (p[4] - p[3]), bk = p[2] * (p[5] - p[3]), n = p[9];
    const oreal fc = o_sph_j1c(q * p[6]);
    const oreal t1 = sq(n * bc * fc);
    const oreal x = sq(q * p[7]);
    const oreal debye = (x == 0.0) ? 1.0 : 2.0 * (expm1(-x) + x) / (x * x);
    const oreal t2 = n * bk * bk * debye;
    const oreal amp = (x == 0.0) ? 1.0 : -expm1(-x) / x;
    const oreal sk = o_sinx_x(q * (p[6] + p[8] * p[7]));
    const oreal t3 = 2.0 * n * n * bc * bk * fc * amp * sk;
    const oreal t4 = n * (n - 1.0) * sq(bk * amp * sk);
    oreal i = t1 + t2 + t3 + t4;
    i *= 1.0e-13;
    i *= p[0];
    *f1 = 0.0;
    *f2 = i;
}
/* ---- raspberry.c: p = sld_lg, sld_sm, sld_solvent, volfraction_lg,
 * volfraction_sm, surface_fraction, radius_lg, radius_sm, penetration ---- */
static oreal ras_reff(int mode, const oreal *p) { return mode == 2 ? p[6] + 2.0 * p[7] - p[8] : p[6]; }
static void ras_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const oreal vfL = p[3], rL = p[6], vfS = p[4], fSs = p[5], rS = p[7], dS = p[8];
    const oreal drL = fabs(p[0] - p[2]), drS = fabs(p[1] - p[2]);
    const oreal VL = O_4PI_3 * rL * rL * rL, VS = O_4PI_3 * rS * rS * rS;
    const oreal Np = vfS * fSs * VL / vfL / VS;
    const oreal slT = drL * VL + Np * drS * VS;
    const oreal psiL = o_sph_j1c(q * rL), psiS = o_sph_j1c(q * rS);
    const oreal sfLS = psiL * psiS * o_sinx_x(q * (rL + dS * rS));
    const oreal sfSS = psiS * psiS * o_sinx_x(q * (rL + dS * rS)) * o_sinx_x(q * (rL + dS * rS));
    oreal f = drL * drL * VL * VL * psiL * psiL;
    f += Np * drS * drS * VS * VS * psiS * psiS;
    f += Np * (Np - 1) * drS * drS * VS * VS * sfSS;
    f += 2 * Np * drL * drS * VL * VS * sfLS;
    if (f != 0.0) f = f / slT / slT;
    f = f * (vfL * drL * drL * VL + vfS * fSs * Np * drS * drS * VS);
    f += vfS * (1.0 - fSs) * drS * drS * VS * psiS * psiS;
    f *= 1.0e8;
    f *= 1.0e-12;
    *f1 = 0.0;
    *f2 = f;
}

/* ---- shells of vector parameters: n = (int)(fp_n + 0.5) of a control
 * parameter with limits [0, 10] (core_multi_shell.c:12-18, onion.c:31-38) ---- */
#define O_NSHELL 10
static int shell_count(oreal fp_n) { return (int)(fp_n + 0.5); }

/* ---- core_multi_shell (core_multi_shell.c:1-60): p = sld_core, radius,
 * sld_solvent, n, sld[10], thickness[10].  Boundary k sits at radius plus
 * the first k thicknesses and carries (sld outside - sld inside). ---- */
static oreal cms_outer(const oreal *p)
{
    const int n = shell_count(p[3]);
    oreal r = p[1];
    for (int i = 0; i < n; i++) r += p[14 + i];
    return r;
}
static oreal cms_vol(const oreal *p) { return O_4PI_3 * cb(cms_outer(p)); }
static oreal cms_reff(int mode, const oreal *p) { return mode == 2 ? p[1] : cms_outer(p); }
static void cms_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const int n = shell_count(p[3]);
    oreal r = p[1], inside = p[0], f = 0.0;
    for (int i = 0; i < n; i++) {
        f += O_4PI_3 * cb(r) * (p[4 + i] - inside) * o_sph_j1c(q * r);
        inside = p[4 + i];
        r += p[14 + i];
    }
    f += O_4PI_3 * cb(r) * (p[2] - inside) * o_sph_j1c(q * r);
    *f1 = 1e-2 * f;
    *f2 = 1e-4 * f * f;
}

/* ---- onion (onion.c:1-73): p = sld_core, radius_core, sld_solvent,
 * n_shells, sld_in[10], sld_out[10], thickness[10], A[10].  Shell i has
 * sld(x) = sld_in + (sld_out - sld_in)(exp(A x) - 1)/(exp(A) - 1) over its
 * fractional depth x; flat (sld_in) when A = 0.  The amplitude of a sphere
 * of radius r is evaluated at the shell's inner (side 0) and outer (side 1)
 * radius and differenced. ---- */
static oreal oni_sphere_amp(oreal q, oreal r, oreal sld_in, oreal sld_out, oreal thickness, oreal A,
                            oreal side)
{
    const oreal vol = O_4PI_3 * cb(r);
    const oreal qr = q * r;
    const oreal bes = o_sph_j1c(qr);
    const oreal alpha = A * r / thickness;
    oreal amp;
    if (qr == 0.0) {
        amp = 1.0;
    } else if (fabs(A) > 0.0) {
        const oreal qr2 = qr * qr, alpha2 = alpha * alpha;
        const oreal sum2 = alpha2 + qr2;
        const oreal s = sin(qr), c = cos(qr);
        const oreal t1 = (alpha2 - qr2) * s / qr - 2.0 * alpha * c;
        const oreal t2 = alpha * s / qr - c;
        const oreal fun = -3.0 * (t1 / sum2 - t2) / sum2;
        const oreal slope = (sld_out - sld_in) / expm1(A);
        const oreal edge = slope * exp(A * side);
        amp = edge * fun + (sld_in - slope) * bes;
    } else {
        amp = sld_in * bes;
    }
    return vol * amp;
}
static oreal oni_outer(const oreal *p)
{
    const int n = shell_count(p[3]);
    oreal r = p[1];
    for (int i = 0; i < n; i++) r += p[24 + i];
    return r;
}
static oreal oni_vol(const oreal *p) { return O_4PI_3 * cb(oni_outer(p)); }
static oreal oni_reff(int mode, const oreal *p) { (void)mode; return oni_outer(p); }
static void oni_fq(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2)
{
    (void)quad_n;
    const int n = shell_count(p[3]);
    oreal r_out = p[1];
    oreal f = oni_sphere_amp(q, r_out, p[0], 0.0, 0.0, 0.0, 0.0);
    for (int i = 0; i < n; i++) {
        const oreal r_in = r_out;
        r_out += p[24 + i];
        f -= oni_sphere_amp(q, r_in, p[4 + i], p[14 + i], p[24 + i], p[34 + i], 0.0);
        f += oni_sphere_amp(q, r_out, p[4 + i], p[14 + i], p[24 + i], p[34 + i], 1.0);
    }
    f -= oni_sphere_amp(q, r_out, p[2], 0.0, 0.0, 0.0, 0.0);
    *f1 = 1e-2 * f;
    *f2 = 1e-4 * f * f;
}

/* ------------------------------------------------------------------ */
/* Model table                                                         */
/* ------------------------------------------------------------------ */
#define O_MAX_PARS 48   /* kernel parameters of the largest model (onion: 44) */
#define O_MAX_SLD 24    /* sld parameters of the largest model (onion: 22) */
typedef struct {
    int npars, nmag, max_pd, oriented;
    int sld_idx[O_MAX_SLD];
    int (*valid)(const oreal *p);
    oreal (*vol)(const oreal *p);
    oreal (*reff)(int mode, const oreal *p);
    void (*fq)(oreal q, const oreal *p, int quad_n, oreal *f1, oreal *f2);
    oreal (*iqac)(oreal qab, oreal qc, const oreal *p);
    oreal (*iqabc)(oreal qa, oreal qb, oreal qc, const oreal *p);
    int no_fq;   /* Iq-only model: one output per q in 1D too */
    oreal (*svol)(const oreal *p);   /* shell volume of hollow models (else the form volume) */
} OModel;

static int always_valid(const oreal *p) { (void)p; return 1; }

/* model ids: 0 sphere, 1 cylinder, 2 ellipsoid, 3 core_shell_cylinder,
 * 4 core_shell_sphere, 6 parallelepiped, 7 triaxial_ellipsoid, 8-10 bcc,
 * fcc, sc paracrystal, 11 hollow_cylinder, 12 vesicle, 13 fuzzy_sphere,
 * 17 core_multi_shell, 18 onion, 19 core_shell_ellipsoid, 20
 * elliptical_cylinder, 21 core_shell_bicelle, 22-32 the Iq-only models guinier,
 * gaussian_peak, lorentz, dab, gel_fit, mass_fractal, surface_fractal,
 * mass_surface_fractal, star_polymer, mono_gauss_coil, fractal, 33-40
 * fractal_core_shell, lamellar_hg, lamellar_stack_caille,
 * lamellar_hg_stack_caille, lamellar_stack_paracrystal, linear_pearls,
 * polymer_micelle, raspberry (the ids libsaship uses; 5 and 14-16, the
 * structure factors, have entries of their own) */
#define N_MODELS 41
static const OModel MODELS[N_MODELS] = {
    {3, 2, 1, 0, {0, 1, -1}, always_valid, sph_vol, sph_reff, sph_fq, 0, 0, 0, 0},
    {6, 2, 4, 1, {0, 1, -1}, cyl_valid, cyl_vol, cyl_reff, cyl_fq, cyl_iqac, 0, 0, 0},
    {6, 2, 4, 1, {0, 1, -1}, always_valid, ell_vol, ell_reff, ell_fq, ell_iqac, 0, 0, 0},
    {8, 3, 5, 1, {0, 1, 2}, always_valid, csc_vol, csc_reff, csc_fq, csc_iqac, 0, 0, 0},
    {5, 3, 2, 0, {2, 3, 4}, always_valid, css_vol, css_reff, css_fq, 0, 0, 0, 0},
    {0, 0, 0, 0, {-1, -1, -1}, 0, 0, 0, 0, 0, 0, 0, 0},
    {8, 2, 5, 1, {0, 1, -1}, always_valid, ppd_vol, ppd_reff, ppd_fq, 0, ppd_iqabc, 0, 0},
    {8, 2, 5, 1, {0, 1, -1}, always_valid, tri_vol, tri_reff, tri_fq, 0, tri_iqabc, 0, 0},
    {8, 2, 4, 1, {3, 4, -1}, always_valid, pc_vol, pc_reff, bcc_fq, 0, bcc_iqabc, 1, 0},
    {8, 2, 4, 1, {3, 4, -1}, always_valid, pc_vol, pc_reff, fcc_fq, 0, fcc_iqabc, 1, 0},
    {8, 2, 4, 1, {3, 4, -1}, always_valid, pc_vol, pc_reff, scp_fq, 0, scp_iqabc, 1, 0},
    {7, 2, 5, 1, {3, 4, -1}, always_valid, hcy_vol, hcy_reff, hcy_fq, hcy_iqac, 0, 0, hcy_svol},
    {5, 2, 2, 0, {0, 1, -1}, always_valid, ves_vol, ves_reff, ves_fq, 0, 0, 0, ves_svol},
    {4, 2, 2, 0, {0, 1, -1}, always_valid, fuz_vol, fuz_reff, fuz_fq, 0, 0, 0, 0},
    {0, 0, 0, 0, {-1}, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, {-1}, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, {-1}, 0, 0, 0, 0, 0, 0, 0, 0},
    {24, 12, 5, 0, {0, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13}, always_valid, cms_vol, cms_reff, cms_fq,
     0, 0, 0, 0},
    {44, 22, 5, 0, {0, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23},
     always_valid, oni_vol, oni_reff, oni_fq, 0, 0, 0, 0},
    {9, 3, 5, 1, {4, 5, 6}, always_valid, cse_vol, cse_reff, cse_fq, cse_iqac, 0, 0, 0},
    {8, 2, 5, 1, {3, 4}, always_valid, ecy_vol, ecy_reff, ecy_fq, 0, ecy_iqabc, 0, 0},
    {10, 4, 5, 1, {4, 5, 6, 7}, always_valid, csb_vol, csb_reff, csb_fq, csb_iqac, 0, 0, 0},
    {1, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, gui_fq, 0, 0, 1, 0},
    {2, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, gpk_fq, 0, 0, 1, 0},
    {1, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, lor_fq, 0, 0, 1, 0},
    {1, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, dab_fq, 0, 0, 1, 0},
    {5, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, gel_fq, 0, 0, 1, 0},
    {3, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, mfr_fq, 0, 0, 1, 0},
    {3, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, sfr_fq, 0, 0, 1, 0},
    {4, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, msf_fq, 0, 0, 1, 0},
    {2, 0, 0, 0, {-1}, always_valid, one_vol, zero_reff, star_fq, 0, 0, 1, 0},
    {2, 0, 1, 0, {-1}, always_valid, one_vol, mgc_reff, mgc_fq, 0, 0, 1, 0},
    {6, 2, 1, 0, {4, 5}, always_valid, frc_vol, zero_reff, frc_fq, 0, 0, 1, 0},
    {8, 3, 2, 0, {2, 3, 4}, always_valid, fcs_vol, zero_reff, fcs_fq, 0, 0, 1, 0},
    {5, 3, 2, 0, {2, 3, 4}, always_valid, one_vol, zero_reff, lhg_fq, 0, 0, 1, 0},
    {6, 2, 2, 0, {4, 5}, always_valid, one_vol, zero_reff, lsc_fq, 0, 0, 1, 0},
    {8, 3, 3, 0, {5, 6, 7}, always_valid, one_vol, zero_reff, lhc_fq, 0, 0, 1, 0},
    {6, 2, 1, 0, {4, 5}, always_valid, one_vol, zero_reff, lsp_fq, 0, 0, 1, 0},
    {5, 2, 0, 0, {3, 4}, always_valid, one_vol, zero_reff, lpe_fq, 0, 0, 1, 0},
    {10, 3, 0, 0, {3, 4, 5}, always_valid, one_vol, zero_reff, pmi_fq, 0, 0, 1, 0},
    {9, 3, 3, 0, {0, 1, 2}, always_valid, one_vol, ras_reff, ras_fq, 0, 0, 1, 0},
};
static int model_ok(int id) { return id >= 0 && id < N_MODELS && MODELS[id].npars > 0; }

/* ------------------------------------------------------------------ */
/* Rotation and magnetism helpers                                      */
/* ------------------------------------------------------------------ */
/* Third row of (jitter^-1 * view^-1); maps (qx, qy) to qc (kernel_iq.c:197-230) */
static void view_jitter_row(oreal theta, oreal phi, oreal dtheta, oreal dphi,
                            oreal *r31, oreal *r32)
{
    const oreal st = sin(theta * O_PI_180), ct = cos(theta * O_PI_180);
    const oreal sp = sin(phi * O_PI_180), cp = cos(phi * O_PI_180);
    const oreal v11 = cp * ct, v12 = sp * ct;
    const oreal v21 = -sp, v22 = cp;
    const oreal v31 = st * cp, v32 = sp * st;
    const oreal sdt = sin(dtheta * O_PI_180), cdt = cos(dtheta * O_PI_180);
    const oreal sdp = sin(dphi * O_PI_180), cdp = cos(dphi * O_PI_180);
    const oreal j31 = sdt, j32 = -sdp * cdt, j33 = cdp * cdt;
    *r31 = j31 * v11 + j32 * v21 + j33 * v31;
    *r32 = j31 * v12 + j32 * v22 + j33 * v32;
}

/* First two columns of (jitter^-1 * view^-1) with psi: (qx, qy) -> (qa, qb, qc)
 * (kernel_iq.c:251-296), r[0..5] = R11, R12, R21, R22, R31, R32 */
static void view_jitter_abc(oreal theta, oreal phi, oreal psi,
                            oreal dtheta, oreal dphi, oreal dpsi, oreal r[6])
{
    oreal st = sin(theta * O_PI_180), ct = cos(theta * O_PI_180);
    oreal sp = sin(phi * O_PI_180), cp = cos(phi * O_PI_180);
    oreal ss = sin(psi * O_PI_180), cs = cos(psi * O_PI_180);
    const oreal v11 = -sp * ss + cp * cs * ct;
    const oreal v12 = sp * cs * ct + ss * cp;
    const oreal v21 = -sp * cs - ss * cp * ct;
    const oreal v22 = -sp * ss * ct + cp * cs;
    const oreal v31 = st * cp;
    const oreal v32 = sp * st;
    st = sin(dtheta * O_PI_180); ct = cos(dtheta * O_PI_180);
    sp = sin(dphi * O_PI_180); cp = cos(dphi * O_PI_180);
    ss = sin(dpsi * O_PI_180); cs = cos(dpsi * O_PI_180);
    const oreal j11 = cs * ct;
    const oreal j12 = sp * st * cs + ss * cp;
    const oreal j13 = sp * ss - st * cp * cs;
    const oreal j21 = -ss * ct;
    const oreal j22 = -sp * ss * st + cp * cs;
    const oreal j23 = sp * cs + ss * st * cp;
    const oreal j31 = st;
    const oreal j32 = -sp * ct;
    const oreal j33 = cp * ct;
    r[0] = j11 * v11 + j12 * v21 + j13 * v31;
    r[1] = j11 * v12 + j12 * v22 + j13 * v32;
    r[2] = j21 * v11 + j22 * v21 + j23 * v31;
    r[3] = j21 * v12 + j22 * v22 + j23 * v32;
    r[4] = j31 * v11 + j32 * v21 + j33 * v31;
    r[5] = j31 * v12 + j32 * v22 + j33 * v32;
}

/* spin cross-section weights dd, du.re, ud.re, uu, du.im, ud.im (kernel_iq.c:84-124) */
static void spin_weights(oreal in_spin, oreal out_spin, oreal wt[6])
{
    in_spin = in_spin < 0.0 ? 0.0 : in_spin > 1.0 ? 1.0 : in_spin;
    out_spin = out_spin < 0.0 ? 0.0 : out_spin > 1.0 ? 1.0 : out_spin;
    const oreal norm = out_spin < 0.5 ? 1 - out_spin : out_spin;
    wt[0] = (1.0 - in_spin) * (1.0 - out_spin) / norm;
    wt[1] = (1.0 - in_spin) * out_spin / norm;
    wt[2] = in_spin * (1.0 - out_spin) / norm;
    wt[3] = in_spin * out_spin / norm;
    wt[4] = wt[1];
    wt[5] = wt[2];
}

static oreal dot3(const oreal *a, const oreal *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* effective sld for cross section xs (kernel_iq.c:127-175) */
static oreal magnetic_sld(int xs, oreal qx, oreal qy, oreal cmt, oreal smt,
                           oreal cmp, oreal smp, oreal sld, oreal mx, oreal my, oreal mz)
{
    const oreal qn = sqrt(qx * qx + qy * qy);
    const oreal qhat[3] = {qx / qn, qy / qn, 0};
    const oreal m[3] = {mx, my, mz};
    const oreal pol[3] = {smt * cmp, smt * smp, cmt};
    const oreal py[3] = {-smp, cmp, 0};
    const oreal pz[3] = {-cmt * cmp, -cmt * smp, smt};
    const oreal k = dot3(m, qhat) / dot3(qhat, qhat);
    const oreal mp[3] = {m[0] - k * qhat[0], m[1] - k * qhat[1], m[2] - k * qhat[2]};
    switch (xs) {
    case 0: return sld - dot3(pol, mp);
    case 1: case 2: return dot3(py, mp);
    case 3: return sld + dot3(pol, mp);
    case 4: return -dot3(pz, mp);
    default: return dot3(pz, mp);
    }
}

/* ------------------------------------------------------------------ */
/* Kernel                                                              */
/* ------------------------------------------------------------------ */
static int n_threads = 1;

void sasoracle_set_threads(int n) { n_threads = n > 0 ? n : 1; }

/*
 * Reference-shaped kernel call (kernel_iq.c:317-328, argtypes kerneldll.py:369)
 * with the model, the kernel kind (0: <model>_Iq, 1: _Iqxy, 2: _Imagnetic) and
 * the quadrature (0 adaptive, else fixed Gauss-n) named explicitly.
 * details: int32[4*max_pd+4]; values: f64; q: nq (1D) or 2*nq interleaved (2D);
 * result: nout*nq+4 with nout = 2 for the 1D kernel of models with Fq, else 1.
 * Returns 0, or -1 for a bad model/kind, -2 for a bad quadrature size.
 */
int sasoracle_kernel(int model_id, int kind, int quad_n,
                     int32_t nq, int32_t pd_start, int32_t pd_stop,
                     const int32_t *details, const oreal *values,
                     const oreal *q, oreal *result,
                     oreal cutoff, int32_t reff_mode)
{
    if (!model_ok(model_id) || kind < 0 || kind > 2) return -1;
    if (quad_n < 0) return -2;
    const OModel *m = &MODELS[model_id];
    const int np = m->npars, maxpd = m->max_pd;
    const int32_t *pd_par = details, *pd_len = details + maxpd;
    const int32_t *pd_off = details + 2 * maxpd, *pd_stride = details + 3 * maxpd;
    const int32_t num_weights = details[4 * maxpd + 1];
    const int32_t theta_par = details[4 * maxpd + 3];
    const oreal *pd_value = values + (2 + np + (m->nmag ? 4 + 3 * m->nmag : 0));
    const oreal *pd_weight = pd_value + num_weights;
    const int nout = (kind == 0 && !m->no_fq) ? 2 : 1;
    const int oriented2d = (kind != 0) && m->oriented;
    const int abc = m->iqabc != 0;
    quad_prime(quad_n);

    oreal acc[4];
    for (int k = 0; k < 4; k++) acc[k] = pd_start == 0 ? 0.0 : result[nout * nq + k];
    if (pd_start == 0) memset(result, 0, sizeof(oreal) * nout * nq);

    oreal base[O_MAX_PARS];
    for (int i = 0; i < np; i++) base[i] = values[2 + i];
    oreal view_theta = 0, view_phi = 0, view_psi = 0;
    if (oriented2d) {
        view_theta = values[theta_par + 2];
        view_phi = values[theta_par + 3];
        base[theta_par] = 0.;
        base[theta_par + 1] = 0.;
        if (abc) {
            /* psi: kernel_iq.c:541-543 */
            view_psi = values[theta_par + 4];
            base[theta_par + 2] = 0.;
        }
    }

    oreal xsw[6] = {0, 0, 0, 0, 0, 0}, smt = 0, cmt = 1, smp = 0, cmp = 1;
    if (kind == 2) {
        spin_weights(values[np + 2], values[np + 3], xsw);
        smt = sin(values[np + 4] * O_PI_180); cmt = cos(values[np + 4] * O_PI_180);
        smp = sin(values[np + 5] * O_PI_180); cmp = cos(values[np + 5] * O_PI_180);
    }

    /* Pass 1: walk the mesh in order, keep the accepted points. */
    const int npts = pd_stop > pd_start ? pd_stop - pd_start : 0;
    const int rw = np + 7;   /* parameters, weight, rotation */
    oreal *pts = (oreal *)malloc(sizeof(oreal) * (size_t)(npts ? npts : 1) * rw);
    int nacc = 0;
    for (int step = pd_start; step < pd_stop; step++) {
        oreal p[O_MAX_PARS];
        memcpy(p, base, sizeof(oreal) * np);
        oreal w0 = 1.0;
        for (int k = maxpd - 1; k >= 0; k--) {
            const int idx = (step / pd_stride[k]) % pd_len[k];
            p[pd_par[k]] = pd_value[pd_off[k] + idx];
            w0 = pd_weight[pd_off[k] + idx] * w0;
        }
        if (!m->valid(p)) continue;
        oreal weight = w0;
        oreal r31 = 0, r32 = 0;
        if (oriented2d) weight = fabs(cos(p[theta_par] * O_PI_180)) * w0;
        if (!(weight > cutoff)) continue;
        const oreal form = m->vol(p);
        acc[0] += weight;
        acc[1] += weight * form;
        acc[2] += weight * (m->svol ? m->svol(p) : form);
        if (reff_mode != 0) acc[3] += weight * m->reff(reff_mode, p);
        oreal *rec = pts + (size_t)nacc * rw;
        memcpy(rec, p, sizeof(oreal) * np);
        rec[np] = weight;
        if (oriented2d && abc) {
            view_jitter_abc(view_theta, view_phi, view_psi, p[theta_par], p[theta_par + 1],
                            p[theta_par + 2], rec + np + 1);
        } else if (oriented2d) {
            view_jitter_row(view_theta, view_phi, p[theta_par], p[theta_par + 1], &r31, &r32);
            rec[np + 1] = r31;
            rec[np + 2] = r32;
        }
        nacc++;
    }

    /* Pass 2: every q accumulates the accepted points in mesh order. */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads)
#endif
    for (int iq = 0; iq < nq; iq++) {
        for (int a = 0; a < nacc; a++) {
            const oreal *rec = pts + (size_t)a * rw;
            oreal p[O_MAX_PARS];
            memcpy(p, rec, sizeof(oreal) * np);
            const oreal weight = rec[np];
            if (kind == 0) {
                oreal f1, f2;
                m->fq(q[iq], p, quad_n, &f1, &f2);
                result[nout * iq + 0] += weight * f2;
                if (nout == 2) result[2 * iq + 1] += weight * f1;
                continue;
            }
            const oreal qx = q[2 * iq], qy = q[2 * iq + 1];
            oreal qab = 0, qc = 0, qa = 0, qb = 0;
            if (oriented2d && abc) {
                /* qabc_apply (kernel_iq.c:298-310) */
                qa = rec[np + 1] * qx + rec[np + 2] * qy;
                qb = rec[np + 3] * qx + rec[np + 4] * qy;
                qc = rec[np + 5] * qx + rec[np + 6] * qy;
            } else if (oriented2d) {
                const oreal dqc = rec[np + 1] * qx + rec[np + 2] * qy;
                const oreal dqab_sq = -dqc * dqc + qx * qx + qy * qy;
                qab = dqab_sq > 0.0 ? sqrt(dqab_sq) : 0.0;
                qc = dqc;
            }
            oreal f2 = 0.0;
            if (kind == 1) {
                if (oriented2d && abc) f2 = m->iqabc(qa, qb, qc, p);
                else if (oriented2d) f2 = m->iqac(qab, qc, p);
                else { oreal f1; m->fq(sqrt(qx * qx + qy * qy), p, quad_n, &f1, &f2); }
            } else {
                const oreal qsq = qx * qx + qy * qy;
                if (qsq > 1.e-16) {
                    for (int xs = 0; xs < 6; xs++) {
                        if (!(xsw[xs] > 1.e-8)) continue;
                        for (int sk = 0; sk < m->nmag; sk++) {
                            const int mi = np + 6 + 3 * sk;
                            const int si = m->sld_idx[sk];
                            p[si] = magnetic_sld(xs, qx, qy, cmt, smt, cmp, smp, values[si + 2],
                                                 values[mi], values[mi + 1], values[mi + 2]);
                        }
                        oreal fx;
                        if (oriented2d && abc) fx = m->iqabc(qa, qb, qc, p);
                        else if (oriented2d) fx = m->iqac(qab, qc, p);
                        else { oreal f1; m->fq(sqrt(qx * qx + qy * qy), p, quad_n, &f1, &fx); }
                        f2 += xsw[xs] * fx;
                    }
                }
            }
            result[iq] += weight * f2;
        }
    }
    free(pts);
    for (int k = 0; k < 4; k++) result[nout * nq + k] = acc[k];
    return 0;
}

/* ------------------------------------------------------------------ */
/* Per-function entry points used by the tests                         */
/* ------------------------------------------------------------------ */
oreal sasoracle_j1(oreal x) { return o_j1(x); }
oreal sasoracle_2j1x_x(oreal x) { return o_2j1x_x(x); }
oreal sasoracle_sph_j1c(oreal x) { return o_sph_j1c(x); }
oreal sasoracle_sinx_x(oreal x) { return o_sinx_x(x); }

int sasoracle_gauss(int n, oreal *z, oreal *w)
{
    const oreal *zz, *ww;
    if (gl_table(n, &zz, &ww) != n) return -1;
    memcpy(z, zz, sizeof(oreal) * n);
    memcpy(w, ww, sizeof(oreal) * n);
    return n;
}

int sasoracle_fq(int model_id, int quad_n, oreal q, const oreal *p, oreal *f1, oreal *f2)
{
    if (!model_ok(model_id)) return -1;
    quad_prime(quad_n);
    MODELS[model_id].fq(q, p, quad_n, f1, f2);
    return 0;
}

oreal sasoracle_iqac(int model_id, oreal qab, oreal qc, const oreal *p)
{
    if (!model_ok(model_id) || !MODELS[model_id].iqac) return NAN;
    return MODELS[model_id].iqac(qab, qc, p);
}

oreal sasoracle_iqabc(int model_id, oreal qa, oreal qb, oreal qc, const oreal *p)
{
    if (!model_ok(model_id) || !MODELS[model_id].iqabc) return NAN;
    return MODELS[model_id].iqabc(qa, qb, qc, p);
}

void sasoracle_view_jitter_abc(oreal theta, oreal phi, oreal psi, oreal dtheta, oreal dphi,
                               oreal dpsi, oreal *r)
{
    view_jitter_abc(theta, phi, psi, dtheta, dphi, dpsi, r);
}

oreal sasoracle_shell_volume(int model_id, const oreal *p)
{
    if (!model_ok(model_id)) return NAN;
    return MODELS[model_id].svol ? MODELS[model_id].svol(p) : MODELS[model_id].vol(p);
}

oreal sasoracle_form_volume(int model_id, const oreal *p)
{
    if (!model_ok(model_id)) return NAN;
    return MODELS[model_id].vol(p);
}

oreal sasoracle_radius_effective(int model_id, int mode, const oreal *p)
{
    if (!model_ok(model_id)) return NAN;
    return MODELS[model_id].reff(mode, p);
}

void sasoracle_view_jitter_row(oreal theta, oreal phi, oreal dtheta, oreal dphi,
                               oreal *r31, oreal *r32)
{
    view_jitter_row(theta, phi, dtheta, dphi, r31, r32);
}

oreal sasoracle_magnetic_sld(int xs, oreal qx, oreal qy, oreal mtheta_deg, oreal mphi_deg,
                              oreal sld, oreal mx, oreal my, oreal mz)
{
    return magnetic_sld(xs, qx, qy, cos(mtheta_deg * O_PI_180), sin(mtheta_deg * O_PI_180),
                        cos(mphi_deg * O_PI_180), sin(mphi_deg * O_PI_180), sld, mx, my, mz);
}

void sasoracle_spin_weights(oreal in_spin, oreal out_spin, oreal *wt)
{
    spin_weights(in_spin, out_spin, wt);
}

/*
 * Hard-sphere structure factor, Percus-Yevick closure, as
 * sasmodels/models/hardsphere.c:1-90 computes it in double precision:
 *   |R| < 1e-12            -> 1
 *   x = |2 q R| < 5e-6     -> 1/A, the q = 0 limit
 *   x < 0.05               -> 1/(1 + phi FF), FF its Taylor series to x^4
 *   otherwise              -> 1/(1 + 24 phi FF/x^2), FF from sin and cos
 * with A = ((1 + 2 phi)/(1 - phi)^2)^2, B = -6 phi ((1 + phi/2)/(1 - phi)^2)^2,
 * G = phi A / 2 (the Percus-Yevick direct-correlation coefficients).
 */
/*
 * Sticky hard spheres (sasmodels/models/stickyhardsphere.c:1-70) and the
 * square-well fluid (sasmodels/models/squarewell.c:1-60), restated step by
 * step in FP64: the perturbative Baxter solution (smaller root lambda of the
 * contact quadratic; -1 without a physical root) and the MSA square well.
 */
oreal sasoracle_stickyhardsphere(oreal q, oreal radius, oreal phi, oreal perturb, oreal stickiness)
{
    const oreal onemineps = 1.0 - perturb;
    const oreal eta = phi / onemineps / onemineps / onemineps;
    const oreal sig = 2.0 * radius;
    const oreal aa = sig / onemineps;
    const oreal etam1 = 1.0 - eta;
    const oreal etam1sq = etam1 * etam1;
    const oreal qa = eta / 6.0;
    const oreal qb = stickiness + eta / etam1;
    const oreal qc = (1.0 + eta / 2.0) / etam1sq;
    oreal radic = qb * qb - 2.0 * qa * qc;
    if (radic < 0) return -1.0;
    radic = sqrt(radic);
    oreal lam = (qb - radic) / qa;
    const oreal lam2 = (qb + radic) / qa;
    if (lam2 < lam) lam = lam2;
    const oreal test = 1.0 + 2.0 * eta;
    const oreal mu = lam * eta * etam1;
    if (mu > test) return -1.0;
    const oreal alpha = (1.0 + 2.0 * eta - mu) / etam1sq;
    const oreal beta = (mu - 3.0 * eta) / (2.0 * etam1sq);
    const oreal kk = q * aa, k2 = kk * kk, k3 = kk * k2;
    const oreal ds = sin(kk), dc = cos(kk);
    const oreal aq1 = ((ds - kk * dc) * alpha) / k3;
    const oreal aq2 = (beta * (1.0 - dc)) / k2;
    const oreal aq3 = (lam * ds) / (12.0 * kk);
    const oreal aq = 1.0 + 12.0 * eta * (aq1 + aq2 - aq3);
    const oreal bq1 = alpha * (0.5 / kk - ds / k2 + (1.0 - dc) / k3);
    const oreal bq2 = beta * (1.0 / kk - ds / k2);
    const oreal bq3 = (lam / 12.0) * ((1.0 - dc) / kk);
    const oreal bq = 12.0 * eta * (bq1 + bq2 - bq3);
    return 1.0 / (aq * aq + bq * bq);
}

oreal sasoracle_squarewell(oreal q, oreal radius, oreal phi, oreal welldepth, oreal wellwidth)
{
    const oreal sigma = radius * 2.;
    const oreal eta = phi, eta2 = eta * eta, eta3 = eta * eta2, eta4 = eta * eta3;
    const oreal etam1 = 1. - eta;
    const oreal etam14 = etam1 * etam1 * etam1 * etam1;
    oreal sk = 1.0 + 2.0 * eta;
    sk *= sk;
    const oreal alpha = (sk + eta3 * (eta - 4.0)) / etam14;
    const oreal beta = -(eta / 3.0) * (18. + 20. * eta - 12. * eta2 + eta4) / etam14;
    const oreal gamm = 0.5 * eta * (sk + eta3 * (eta - 4.)) / etam14;
    sk = q * sigma;
    const oreal sk2 = sk * sk, sk3 = sk * sk2, sk4 = sk3 * sk;
    const oreal S = sin(sk), C = cos(sk), SL = sin(wellwidth * sk), CL = cos(wellwidth * sk);
    const oreal t1 = alpha * sk3 * (S - sk * C);
    const oreal t2 = beta * sk2 * 2.0 * (sk * S - (0.5 * sk2 - 1.) * C - 1.0);
    const oreal t3 = gamm * ((4.0 * sk3 - 24. * sk) * S - (sk4 - 12.0 * sk2 + 24.0) * C + 24.0);
    const oreal t4 = -welldepth * sk3 * (SL + sk * (C - wellwidth * CL) - S);
    const oreal ck = -24.0 * eta * (t1 + t2 + t3 + t4) / sk3 / sk3;
    return 1. / (1. - ck);
}

/*
 * Hayter-Penfold rescaled MSA structure factor (sasmodels/models/hayter_msa.c:
 * Iq at :15-76, sqcoef at :78-176, sqfun at :179-467, sqhcal at :469-541;
 * J. B. Hayter and J. Penfold, Mol. Phys. 42 (1981) 109; J.-P. Hansen and
 * J. B. Hayter, Mol. Phys. 46 (1982) 651), restated with the reference's
 * steps, thresholds, iteration limits and SI constants (its answers depend on
 * which branch and root are taken): the coefficients at a trial volume
 * fraction in four modes (full MSA from the large-screening start or near
 * the Gillan condition, the Gillan-condition coefficients with the residual
 * of the other equation or with g(1+)), the secant rescaling of the volume
 * fraction, and S(q) in closed form with its small-q expansion.
 */

typedef struct {
    oreal eta0, gek0, ak0;          // volume fraction, contact potential / kT, kappa * sigma
    oreal eta, gek, ak, s;          // in use (rescaled): s = (eta0 / eta)^(1/3)
    oreal a, b, c, f, u, v, g1;     // MSA coefficients, g1 = g(1+)
} OMsa;

// Coefficients at the trial volume fraction `et` (hayter_msa.c:179-467):
// mode 1 the full MSA with the large-screening start (and the asymptotic
// form beyond k = 15), 2 the Gillan-condition coefficients and the
// residual of the other equation in `fval`, 3 the full MSA started near
// the Gillan condition, 4 the Gillan-condition coefficients with g(1+) in
// `fval`.  Returns the Newton iterations, or -2 when Newton fails.
static int omsa_coefficients(OMsa *st, oreal et, int mode, oreal *fval)
{
    oreal a2 = 0.0, a3 = 0.0, b2 = 0.0, b3 = 0.0, v2 = 0.0, v3 = 0.0, p2 = 0.0, p3 = 0.0;
    int ir = 0;
    const oreal eta2 = et * et, eta3 = eta2 * et;
    const oreal e12 = 12.0 * et, e24 = e12 + e12;
    st->s = pow(st->eta0 / et, 1.0 / 3.0);
    st->ak = st->ak0 / st->s;
    const int big = st->ak > 15.0 && mode == 1;
    st->gek = st->gek0 * st->s * exp(st->ak0 - st->ak);
    const oreal gek = st->gek, ak = st->ak;
    const oreal ak2 = ak * ak, ak1 = 1.0 + ak, dak2 = 1.0 / ak2, dak4 = dak2 * dak2;
    const oreal d = 1.0 - et, d2 = d * d, dak = d / ak, dd2 = 1.0 / d2, dd4 = dd2 * dd2;
    const oreal dd45 = dd4 * 2.0e-1;
    const oreal eta3d = 3.0 * et, eta6d = eta3d + eta3d, eta32 = eta3 + eta3;
    const oreal eta2d = et + 2.0, eta2d2 = eta2d * eta2d, eta21 = 2.0 * et + 1.0, eta22 = eta21 * eta21;
    // alpha, beta, nu, phi, tau of the paper
    const oreal al1 = -eta21 * dak, al2 = (14.0 * eta2 - 4.0 * et - 1.0) * dak2, al3 = 36.0 * eta2 * dak4;
    const oreal be1 = -(eta2 + 7.0 * et + 1.0) * dak, be2 = 9.0 * et * (eta2 + 4.0 * et - 2.0) * dak2;
    const oreal be3 = 12.0 * et * (2.0 * eta2 + 8.0 * et - 1.0) * dak4;
    const oreal vu1 = -(eta3 + 3.0 * eta2 + 45.0 * et + 5.0) * dak;
    const oreal vu2 = (eta32 + 3.0 * eta2 + 42.0 * et - 2.0e1) * dak2;
    const oreal vu3 = (eta32 + 3.0e1 * et - 5.0) * dak4;
    const oreal vu4 = vu1 + e24 * ak * vu3, vu5 = eta6d * (vu2 + 4.0 * vu3);
    const oreal ph1 = eta6d / ak, ph2 = d - e12 * dak2;
    const oreal ta1 = (et + 5.0) / (5.0 * ak), ta2 = eta2d * dak2;
    const oreal ta3 = -e12 * gek * (ta1 + ta2), ta4 = eta3d * ak2 * (ta1 * ta1 - ta2 * ta2);
    const oreal ta5 = eta3d * (et + 8.0) * 1.0e-1 - 2.0 * eta22 * dak2;
    // sinh k, cosh k (e^-k dropped beyond k = 20)
    const oreal ex1 = exp(ak), ex2 = ak < 20.0 ? exp(-ak) : 0.0;
    const oreal sk = 0.5 * (ex1 - ex2), ck = 0.5 * (ex1 + ex2);
    const oreal ckma = ck - 1.0 - ak * sk, skma = sk - ak * ck;
    const oreal a1 = (e24 * gek * (al1 + al2 + ak1 * al3) - eta22) * dd4;
    const oreal b1 = (1.5 * et * eta2d2 - e12 * gek * (be1 + be2 + ak1 * be3)) * dd4;
    const oreal v1 = (eta21 * (eta2 - 2.0 * et + 1.0e1) * 2.5e-1 - gek * (vu4 + vu5)) * dd45;
    if (!big) {
        a2 = e24 * (al3 * skma + al2 * sk - al1 * ck) * dd4;
        a3 = e24 * (eta22 * dak2 - 0.5 * d2 + al3 * ckma - al1 * sk + al2 * ck) * dd4;
        b2 = e12 * (-be3 * skma - be2 * sk + be1 * ck) * dd4;
        b3 = e12 * (0.5 * d2 * eta2d - eta3d * eta2d2 * dak2 - be3 * ckma + be1 * sk - be2 * ck) * dd4;
        v2 = (vu4 * ck - vu5 * sk) * dd45;
        v3 = ((eta3 - 6.0 * eta2 + 5.0) * d - eta6d * (2.0 * eta3 - 3.0 * eta2 + 18.0 * et + 1.0e1) * dak2 +
              e24 * vu3 + vu4 * sk - vu5 * ck) * dd45;
    }
    const oreal pp1 = ph1 * ph1, pp2 = ph2 * ph2, pp = pp1 + pp2, p1p2 = ph1 * ph2 * 2.0;
    const oreal p1 = (gek * (pp1 + pp2 - p1p2) - 0.5 * eta2d) * dd2;
    if (!big) {
        p2 = (pp * sk + p1p2 * ck) * dd2;
        p3 = (pp * ck + p1p2 * sk + pp1 - pp2) * dd2;
    }
    const oreal t1 = ta3 + ta4 * a1 + ta5 * b1;
    oreal fa, ca;
    if (big) {
        // very large screening: the asymptotic solution
        v3 = ((eta3 - 6.0 * eta2 + 5.0) * d - eta6d * (2.0 * eta3 - 3.0 * eta2 + 18.0 * et + 1.0e1) * dak2 +
              e24 * vu3) * dd45;
        const oreal t3 = ta4 * a3 + ta5 * b3 + e12 * ta2 - 4.0e-1 * et * (et + 1.0e1) - 1.0;
        p3 = (pp1 - pp2) * dd2;
        b3 = e12 * (0.5 * d2 * eta2d - eta3d * eta2d2 * dak2 + be3) * dd4;
        a3 = e24 * (eta22 * dak2 - 0.5 * d2 - al3) * dd4;
        const oreal um6 = t3 * a3 - e12 * v3 * v3, um5 = t1 * a3 + a1 * t3 - e24 * v1 * v3;
        const oreal um4 = t1 * a1 - e12 * v1 * v1;
        const oreal l6 = e12 * p3 * p3, l5 = e24 * p1 * p3 - b3 - b3 - ak2, l4 = e12 * p1 * p1 - b1 - b1;
        fa = -(um4 * l6 - l4 * um6) / (um5 * l6 - l5 * um6);
        ca = -fa;
        st->f = fa;
        st->c = ca;
        st->b = b1 + b3 * fa;
        st->a = a1 + a3 * fa;
        st->v = v1 + v3 * fa;
        st->g1 = -(p1 + p3 * fa);
        *fval = fabs(st->g1) < 1.0e-3 ? 0.0 : st->g1;
        st->eta = et;
    } else {
        const oreal t2 = ta4 * a2 + ta5 * b2 + e12 * (ta1 * ck - ta2 * sk);
        const oreal t3 = ta4 * a3 + ta5 * b3 + e12 * (ta1 * sk - ta2 * (ck - 1.0)) - 4.0e-1 * et * (et + 1.0e1) - 1.0;
        // mu
        const oreal um1 = t2 * a2 - e12 * v2 * v2, um2 = t1 * a2 + t2 * a1 - e24 * v1 * v2;
        const oreal um3 = t2 * a3 + t3 * a2 - e24 * v2 * v3, um4 = t1 * a1 - e12 * v1 * v1;
        const oreal um5 = t1 * a3 + t3 * a1 - e24 * v1 * v3, um6 = t3 * a3 - e12 * v3 * v3;
        if (mode == 1 || mode == 3) {
            // the full MSA: lambda, omega and the quartic in f
            const oreal l1 = e12 * p2 * p2, l2 = e24 * p1 * p2 - b2 - b2, l3 = e24 * p2 * p3;
            const oreal l4 = e12 * p1 * p1 - b1 - b1, l5 = e24 * p1 * p3 - b3 - b3 - ak2, l6 = e12 * p3 * p3;
            const oreal w16 = um1 * l6 - l1 * um6, w15 = um1 * l5 - l1 * um5, w14 = um1 * l4 - l1 * um4;
            const oreal w13 = um1 * l3 - l1 * um3, w12 = um1 * l2 - l1 * um2;
            const oreal w26 = um2 * l6 - l2 * um6, w25 = um2 * l5 - l2 * um5, w24 = um2 * l4 - l2 * um4;
            const oreal w36 = um3 * l6 - l3 * um6, w35 = um3 * l5 - l3 * um5, w34 = um3 * l4 - l3 * um4;
            const oreal w32 = um3 * l2 - l3 * um2;
            const oreal w46 = um4 * l6 - l4 * um6, w56 = um5 * l6 - l5 * um6;
            const oreal w3526 = w35 + w26, w3425 = w34 + w25;
            const oreal q4 = w16 * w16 - w13 * w36;
            const oreal q3 = 2.0 * w16 * w15 - w13 * w3526 - w12 * w36;
            const oreal q2 = w15 * w15 + 2.0 * w16 * w14 - w13 * w3425 - w12 * w3526;
            const oreal q1 = 2.0 * w15 * w14 - w13 * w24 - w12 * w3425;
            const oreal q0 = w14 * w14 - w12 * w24;
            oreal fap;
            if (mode == 1) {
                fap = (w14 - w34 - w46) / (w12 - w15 + w35 - w26 + w56 - w32);   // large k
            } else {
                // near the Gillan condition; the Percus-Yevick-like estimate
                // when both the contact potential and k are small
                oreal g = 0.5 * eta2d * dd2 * exp(-gek);
                if (gek <= 2.0 && gek >= 0.0 && ak <= 1.0) {
                    const oreal e24g = e24 * gek * exp(ak), pwk = sqrt(e24g);
                    const oreal qpw = (1.0 - sqrt(1.0 + 2.0 * d2 * d * pwk / eta22)) * eta21 / d;
                    g = -qpw * qpw / e24 + 0.5 * eta2d * dd2;
                }
                const oreal pg = p1 + g;
                oreal cst = ak2 * pg + 2.0 * (b3 * pg - b1 * p3) + e12 * g * g * p3;
                cst = -cst / (ak2 * p2 + 2.0 * (b3 * p2 - b2 * p3));
                fap = -(pg + p2 * cst) / p3;
            }
            // Newton on the quartic
            oreal del;
            int it = 0;
            do {
                if (++it > 40) return -2;
                fa = fap;
                const oreal fun = q0 + (q1 + (q2 + (q3 + q4 * fa) * fa) * fa) * fa;
                const oreal fund = q1 + (2.0 * q2 + (3.0 * q3 + 4.0 * q4 * fa) * fa) * fa;
                fap = fa - fun / fund;
                del = fabs((fap - fa) / fa);
            } while (del > 1.0e-6);
            ir = it;
            fa = fap;
            ca = -(w16 * fa * fa + w15 * fa + w14) / (w13 * fa + w12);
            st->g1 = -(p1 + p2 * ca + p3 * fa);
            *fval = fabs(st->g1) < 1.0e-3 ? 0.0 : st->g1;
            st->eta = et;
        } else {
            // the Gillan condition g(1+) = 0 fixes c and f
            ca = ak2 * p1 + 2.0 * (b3 * p1 - b1 * p3);
            ca = -ca / (ak2 * p2 + 2.0 * (b3 * p2 - b2 * p3));
            fa = -(p1 + p2 * ca) / p3;
            if (mode == 2) *fval = um1 * ca * ca + (um2 + um3 * fa) * ca + um4 + um5 * fa + um6 * fa * fa;
            else *fval = -(p1 + p2 * ca + p3 * fa);
        }
        st->f = fa;
        st->c = ca;
        st->b = b1 + b2 * ca + b3 * fa;
        st->a = a1 + a2 * ca + a3 * fa;
        st->v = (v1 + v2 * ca + v3 * fa) / st->a;
    }
    const oreal g24 = e24 * gek * ex1;
    st->u = (ak * ak2 * ca - g24) / (ak2 * g24);
    return ir;
}

// The coefficients for (eta0, gek0, ak0) (hayter_msa.c:88-176): the plain
// MSA when it gives g(1+) >= 0, else the volume fraction rescaled (secant
// iteration) to meet the Gillan condition.  Negative on failure.
static int omsa_solve(OMsa *st)
{
    oreal fval = 0.0;
    int ir = 0;
    int ig = 1;
    if (st->ak0 >= 1.0 + 8.0 * st->eta0) {
        ig = 0;
        ir = omsa_coefficients(st, st->eta0, 1, &fval);
        if (ir < 0 || fval >= 0.0) return ir;
        st->g1 = fval;
    }
    oreal e = fmin(st->eta0, 0.20);
    const oreal gamk = 2.0 * pow(st->eta0, 1.0 / 3.0) * st->gek0 * exp(st->ak0 - st->ak0 / pow(st->eta0, 1.0 / 3.0));
    if (!ig || gamk >= 0.15) {
        int ii = 0;
        oreal del, e2;
        do {
            if (++ii > 40) return -1;
            if (e <= 0.0) e = st->eta0 / ii;
            if (e > 0.6) e = 0.35 / ii;
            const oreal e1 = e;
            oreal f1, f2;
            omsa_coefficients(st, e1, 2, &f1);
            omsa_coefficients(st, e * 1.01, 2, &f2);
            e2 = e1 - (e * 1.01 - e1) * f1 / (f2 - f1);
            e = e2;
            del = fabs((e2 - e1) / e1);
        } while (del > 5.0e-6);
        omsa_coefficients(st, e2, 4, &fval);
        st->g1 = fval;
        st->eta = e2;
        ir = ii;
        if (!ig || e >= st->eta0) return ir;
    }
    ir = omsa_coefficients(st, st->eta0, 3, &fval);
    st->g1 = fval;
    if (ir >= 0 && fval < 0.0) ir = -3;
    return ir;
}

// S(q) from solved coefficients at q sigma (hayter_msa.c:469-541): the
// closed form of the Fourier transform of c(x), its Taylor expansion in
// (q sigma / s)^2 below 0.01 and -1/a at zero.
static oreal omsa_sq(const OMsa *st, oreal qsig)
{
    const oreal eta = st->eta, ak = st->ak, gek = st->gek;
    const oreal e24 = 24.0 * eta;
    const oreal x1 = exp(ak), x2 = ak < 20.0 ? exp(-ak) : 0.0;
    const oreal ck = 0.5 * (x1 + x2), sk = 0.5 * (x1 - x2), ak2 = ak * ak;
    const oreal qk = qsig / st->s, q2k = qk * qk;
    if (qk <= 1.0e-08) return -1.0 / st->a;
    if (qk <= 0.01) {
        const oreal aqk = st->a * (8.0 + 2.0 * eta) + 6 * st->b - 12.0 * st->f -
                           24 * (gek * (1.0 + ak) - ck * ak * st->c + st->f * (ck - 1.0) + (st->c - st->f * ak) * sk) / ak2 +
                           q2k * (-(st->a * (48.0 + 15.0 * eta) + 40.0 * st->b) / 60.0 + st->f +
                                  (4.0 / ak2) * (gek * (9.0 + 7.0 * ak) + ck * (9.0 * st->f - 7.0 * st->c * ak) +
                                                 sk * (9.0 * st->c - 7.0 * st->f * ak)));
        return 1.0 / (1.0 - eta * aqk);
    }
    const oreal qk2 = 1.0 / q2k, qk3 = qk2 / qk, qqk = 1.0 / (qk * (q2k + ak2));
    oreal sink, cosk;
    sink = sin(qk);
    cosk = cos(qk);
    const oreal asink = ak * sink, qcosk = qk * cosk;
    oreal aqk = st->a * (sink - qcosk);
    aqk = aqk + st->b * ((2.0 * qk2 - 1.0) * qcosk + 2.0 * sink - 2.0 / qk);
    const oreal inter = 24.0 * qk3 + 4.0 * (1.0 - 6.0 * qk2) * sink;
    aqk = (aqk + 0.5 * eta * st->a * (inter - (1.0 - 12.0 * qk2 + 24.0 * qk2 * qk2) * qcosk)) * qk3;
    aqk = aqk + st->c * (ck * asink - sk * qcosk) * qqk;
    aqk = aqk + st->f * (sk * asink - qk * (ck * cosk - 1.0)) * qqk;
    aqk = aqk + st->f * (cosk - 1.0) * qk2;
    aqk = aqk - gek * (asink + qcosk) * qqk;
    return 1.0 / (1.0 - e24 * aqk);
}


// Hayter-Penfold rescaled MSA S(q): sphere radius (A), volume fraction,
// charge (e), temperature (K), 1:1 salt concentration (mol/l) and the
// solvent's relative permittivity.  The ionic strength counts the
// monovalent counterions of the spheres and the salt; the Debye screening
// constant and the contact potential follow in SI units with the constants
// of the reference.  NaN when the coefficients do not converge.
oreal sasoracle_hayter_msa(oreal q, oreal radius, oreal phi, oreal charge, oreal temperature,
                             oreal csalt, oreal dielec)
{
    const oreal e_charge = 1.602189e-19, k_b = 1.380662e-23, eps0 = 8.85418782E-12;
    const oreal diam = 2.0 * radius;
    const oreal beta = 1.0 / (k_b * temperature);
    const oreal perm = dielec * eps0;
    const oreal zc = charge * e_charge;
    const oreal sidiam = diam * 1.0E-10;
    const oreal half = sidiam / 2.0;
    const oreal vp = 4.18879020478639 * (half * half * half);   // 4 pi / 3 as the reference's header rounds it
    const oreal cs = csalt * 6.022E23 * 1.0E3;
    const oreal ionst = 0.5 * e_charge * e_charge * (charge * phi / vp + 2.0 * cs);
    const oreal kappa = sqrt(2 * beta * ionst / perm);
    OMsa st_, *st = &st_;
    st->gek0 = beta * zc * zc / (3.141592653589793 * perm * sidiam * ((2.0 + kappa * sidiam) * (2.0 + kappa * sidiam)));
    st->ak0 = kappa * sidiam;
    st->eta0 = phi;
    st->eta = phi;
    if (omsa_solve(st) < 0) return NAN;
    return omsa_sq(st, q * diam);
}


oreal sasoracle_hardsphere(oreal q, oreal radius, oreal phi)
{
    if (fabs(radius) < 1e-12) return 1.0;
    const oreal inv = 1.0 / (1.0 - phi);
    const oreal d = inv * inv;
    const oreal a = sq((1.0 + 2.0 * phi) * d);
    const oreal x = fabs(2.0 * q * radius);
    if (x < 5e-6) return 1.0 / a;
    const oreal b = -6.0 * phi * sq((1.0 + 0.5 * phi) * d);
    const oreal g = 0.5 * phi * a;
    const oreal x2 = x * x;
    if (x < 0.05) {
        const oreal c0 = 8.0 * a + 6.0 * b + 4.0 * g;
        const oreal c2 = -0.8 * a - b / 1.5 - 0.5 * g;
        const oreal c4 = a / 35.0 + 0.0125 * b + 0.02 * g;
        return 1.0 / (1.0 + phi * (c0 + (c2 + c4 * x2) * x2));
    }
    const oreal s = sin(x), c = cos(x), x4 = x2 * x2;
    const oreal ta = a * (s - x * c);
    const oreal tb = b * (2.0 * x * s - (x2 - 2.0) * c - 2.0);
    const oreal tg = g * ((4.0 * x2 - 24.0) * x * s - (x4 - 12.0 * x2 + 24.0) * c + 24.0);
    const oreal ff = ((tg / x2 + tb) / x + ta) / x;
    return 1.0 / (1.0 + 24.0 * phi * ff / x2);
}
