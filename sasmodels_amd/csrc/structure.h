// structure.h -- structure factors S(q) for P@S products, in FP64.
//
// Hard spheres with the Percus-Yevick closure (sasmodels/models/hardsphere.c,
// hardsphere.py): spheres of radius R at volume fraction phi.  The Fourier
// transform of the PY direct correlation function c(r) = -(a + b t + g t^3),
// t = r / 2R < 1 (Wertheim's solution), gives, with x = 2 q R,
//
//   S(q) = 1 / (1 + 24 phi J(x)),
//   J(x) = int_0^1 (a + b t + g t^3) t^2 sin(x t)/(x t) dt,
//   a = (1 + 2 phi)^2 / (1 - phi)^4,
//   b = -6 phi (1 + phi/2)^2 / (1 - phi)^4,
//   g = phi a / 2.
//
// J is evaluated from the moments  M_n(x) = int_0^1 t^n sinc(x t) dt:
//   x >= 1: closed forms
//     M_2 = (sin x - x cos x) / x^3
//     M_3 = (2 x sin x + (2 - x^2) cos x - 2) / x^4
//     M_5 = ((4 x^3 - 24 x) sin x - (x^4 - 12 x^2 + 24) cos x + 24) / x^6
//     (cancellation there costs at most ~150 ulp: ~3e-14 relative);
//   x < 1: the power series of sinc term by term,
//     J = sum_k (-x^2)^k / (2k+1)! [a/(2k+3) + b/(2k+4) + g/(2k+6)],
//     nine terms (the tenth is below 3e-15 of the first).
// The reference switches at x = 0.05 with a series to x^4; both agree to
// the series remainder (< 2e-8 relative at the switch).
//
// Like the reference, R == 0 (|R| < 1e-12) gives S = 1.
//
// Also here: Baxter's adhesive ("sticky") hard spheres, perturbative
// solution (stickyhardsphere.c), the square-well fluid in the mean
// spherical approximation (squarewell.c) and charged spheres in the
// rescaled MSA (hayter_msa.c); all FP64 like the reference.
#pragma once
#include <hip/hip_runtime.h>

namespace sas {

// Structure-factor model ids of the C ABI.
constexpr int S_HARDSPHERE = 5, S_STICKY = 14, S_SQUAREWELL = 15, S_HAYTER = 16;

__host__ __device__ inline double hardsphere_sq(double q, double radius, double phi)
{
    if (fabs(radius) < 1e-12) return 1.0;
    const double u = 1.0 / (1.0 - phi);
    const double u4 = (u * u) * (u * u);
    const double a = (1.0 + 2.0 * phi) * (1.0 + 2.0 * phi) * u4;
    const double b = -6.0 * phi * (1.0 + 0.5 * phi) * (1.0 + 0.5 * phi) * u4;
    const double g = 0.5 * phi * a;
    const double x = fabs(2.0 * q * radius);
    double J;
    if (x < 1.0) {
        const double mx2 = -x * x;
        double term = 1.0;   // (-x^2)^k / (2k+1)!
        J = 0.0;
        for (int k = 0; k < 9; k++) {
            const double m = 2.0 * k;
            J += term * (a / (m + 3.0) + b / (m + 4.0) + g / (m + 6.0));
            term *= mx2 / ((m + 2.0) * (m + 3.0));
        }
    } else {
        double s, c;
        sincos(x, &s, &c);
        const double x2 = x * x;
        const double m2 = (s - x * c) / (x2 * x);
        const double m3 = (2.0 * x * s + (2.0 - x2) * c - 2.0) / (x2 * x2);
        const double m5 = ((4.0 * x2 - 24.0) * x * s - ((x2 - 12.0) * x2 + 24.0) * c + 24.0) / (x2 * x2 * x2);
        J = a * m2 + b * m3 + g * m5;
    }
    return 1.0 / (1.0 + 24.0 * phi * J);
}

// Sticky hard spheres: radius R, volume fraction phi, perturbation
// parameter eps (well width relative to the diameter) and stickiness tau.
// With eta = phi/(1 - eps)^3, d = 2R/(1 - eps), k = q d, the contact
// parameter lambda is the smaller root of (eta/12) l^2 - (tau + eta/(1-eta)) l
// + (1 + eta/2)/(1-eta)^2 = 0, mu = lambda eta (1 - eta), and
//   S = 1/(A^2 + B^2),
//   A = 1 + 12 eta [a (sin k - k cos k)/k^3 + b (1 - cos k)/k^2 - lambda sin k/(12 k)],
//   B = 12 eta [a (1/(2k) - sin k/k^2 + (1 - cos k)/k^3) + b (1/k - sin k/k^2)
//               - lambda (1 - cos k)/(12 k)],
//   a = (1 + 2 eta - mu)/(1 - eta)^2,  b = (mu - 3 eta)/(2 (1 - eta)^2).
// No physical root (negative discriminant, or mu > 1 + 2 eta): -1, as the
// reference returns.
__host__ __device__ inline double stickyhardsphere_sq(double q, double radius, double phi, double eps, double tau)
{
    const double one_e = 1.0 - eps;
    const double eta = phi / (one_e * one_e * one_e);
    const double d = 2.0 * radius / one_e;
    const double m = 1.0 - eta, m2 = m * m;
    const double c2 = eta / 6.0, c1 = tau + eta / m, c0 = (1.0 + 0.5 * eta) / m2;
    double disc = c1 * c1 - 2.0 * c2 * c0;
    if (disc < 0.0) return -1.0;
    disc = sqrt(disc);
    const double l1 = (c1 - disc) / c2, l2 = (c1 + disc) / c2;
    const double lam = l2 < l1 ? l2 : l1;
    const double mu = lam * eta * m;
    if (mu > 1.0 + 2.0 * eta) return -1.0;
    const double a = (1.0 + 2.0 * eta - mu) / m2;
    const double b = (mu - 3.0 * eta) / (2.0 * m2);
    const double k = q * d, k2 = k * k, k3 = k * k2;
    double sk, ck;
    sincos(k, &sk, &ck);
    const double A = 1.0 + 12.0 * eta * ((sk - k * ck) * a / k3 + b * (1.0 - ck) / k2 - lam * sk / (12.0 * k));
    const double B = 12.0 * eta * (a * (0.5 / k - sk / k2 + (1.0 - ck) / k3) + b * (1.0 / k - sk / k2) -
                                   (lam / 12.0) * ((1.0 - ck) / k));
    return 1.0 / (A * A + B * B);
}

// Square-well fluid (mean spherical approximation, Sharma and Sharma):
// spheres of diameter s = 2R at volume fraction eta with a well of depth
// e (in kT) out to l diameters.  The hard-core part is the Percus-Yevick
// direct correlation function; with k = q s,
//   c(k) = -24 eta [a k^3 (sin k - k cos k) + 2 b k^2 (k sin k - (k^2/2 - 1) cos k - 1)
//                   + g ((4k^3 - 24k) sin k - (k^4 - 12k^2 + 24) cos k + 24)
//                   - e k^3 (sin(l k) + k (cos k - l cos(l k)) - sin k)] / k^6,
//   a = ((1 + 2 eta)^2 + eta^3 (eta - 4))/(1 - eta)^4,  g = eta a / 2,
//   b = -(eta/3)(18 + 20 eta - 12 eta^2 + eta^4)/(1 - eta)^4,
// S = 1/(1 - c).
__host__ __device__ inline double squarewell_sq(double q, double radius, double eta, double depth, double width)
{
    const double m = 1.0 - eta, m4 = (m * m) * (m * m);
    const double e2 = eta * eta, e3 = e2 * eta, e4 = e2 * e2;
    const double p = (1.0 + 2.0 * eta) * (1.0 + 2.0 * eta);
    const double a = (p + e3 * (eta - 4.0)) / m4;
    const double b = -(eta / 3.0) * (18.0 + 20.0 * eta - 12.0 * e2 + e4) / m4;
    const double g = 0.5 * eta * (p + e3 * (eta - 4.0)) / m4;
    const double k = q * 2.0 * radius, k2 = k * k, k3 = k2 * k, k4 = k2 * k2;
    double sk, ck, sl, cl;
    sincos(k, &sk, &ck);
    sincos(width * k, &sl, &cl);
    const double t1 = a * k3 * (sk - k * ck);
    const double t2 = 2.0 * b * k2 * (k * sk - (0.5 * k2 - 1.0) * ck - 1.0);
    const double t3 = g * ((4.0 * k3 - 24.0 * k) * sk - (k4 - 12.0 * k2 + 24.0) * ck + 24.0);
    const double t4 = -depth * k3 * (sl + k * (ck - width * cl) - sk);
    const double c = -24.0 * eta * (t1 + t2 + t3 + t4) / (k3 * k3);
    return 1.0 / (1.0 - c);
}

// Charged spheres with a screened Coulomb (Yukawa) tail in the rescaled
// mean spherical approximation (hayter_msa.c): J. B. Hayter and J. Penfold,
// Mol. Phys. 42 (1981) 109 give the MSA direct correlation function inside
// the core in closed form with coefficients a, b, c, f (and u, v), fixed
// by a quartic in f; J.-P. Hansen and J. B. Hayter, Mol. Phys. 46
// (1982) 651 rescale the volume fraction (and with it the contact
// potential and the screening constant) until the contact value of g is
// zero (the Gillan condition) whenever the plain MSA gives g(1+) < 0.
// Steps, thresholds and iteration limits follow the reference for parity
// (its answers depend on which branch and root are taken), and so does the
// rounding: no FMA contraction here, as in the reference's C build.  The
// iterations stop at relative steps of 5e-6 (volume fraction) and 1e-6
// (the quartic's root), and S is sensitive to where they stop: a result
// moves by up to ~1e-3 when the last step does.
namespace msa {

struct State {
    double eta0, gek0, ak0;          // volume fraction, contact potential / kT, kappa * sigma
    double eta, gek, ak, s;          // in use (rescaled): s = (eta0 / eta)^(1/3)
    double a, b, c, f, u, v, g1;     // MSA coefficients, g1 = g(1+)
};

// Coefficients at the trial volume fraction `et` (hayter_msa.c:179-467):
// mode 1 the full MSA with the large-screening start (and the asymptotic
// form beyond k = 15), 2 the Gillan-condition coefficients and the
// residual of the other equation in `fval`, 3 the full MSA started near
// the Gillan condition, 4 the Gillan-condition coefficients with g(1+) in
// `fval`.  Returns the Newton iterations, or -2 when Newton fails.
__host__ __device__ inline int coefficients(State &st, double et, int mode, double &fval)
{
    #pragma clang fp contract(off)
    double a2 = 0.0, a3 = 0.0, b2 = 0.0, b3 = 0.0, v2 = 0.0, v3 = 0.0, p2 = 0.0, p3 = 0.0;
    int ir = 0;
    const double eta2 = et * et, eta3 = eta2 * et;
    const double e12 = 12.0 * et, e24 = e12 + e12;
    st.s = pow(st.eta0 / et, 1.0 / 3.0);
    st.ak = st.ak0 / st.s;
    const bool big = st.ak > 15.0 && mode == 1;
    st.gek = st.gek0 * st.s * exp(st.ak0 - st.ak);
    const double gek = st.gek, ak = st.ak;
    const double ak2 = ak * ak, ak1 = 1.0 + ak, dak2 = 1.0 / ak2, dak4 = dak2 * dak2;
    const double d = 1.0 - et, d2 = d * d, dak = d / ak, dd2 = 1.0 / d2, dd4 = dd2 * dd2;
    const double dd45 = dd4 * 2.0e-1;
    const double eta3d = 3.0 * et, eta6d = eta3d + eta3d, eta32 = eta3 + eta3;
    const double eta2d = et + 2.0, eta2d2 = eta2d * eta2d, eta21 = 2.0 * et + 1.0, eta22 = eta21 * eta21;
    // alpha, beta, nu, phi, tau of the paper
    const double al1 = -eta21 * dak, al2 = (14.0 * eta2 - 4.0 * et - 1.0) * dak2, al3 = 36.0 * eta2 * dak4;
    const double be1 = -(eta2 + 7.0 * et + 1.0) * dak, be2 = 9.0 * et * (eta2 + 4.0 * et - 2.0) * dak2;
    const double be3 = 12.0 * et * (2.0 * eta2 + 8.0 * et - 1.0) * dak4;
    const double vu1 = -(eta3 + 3.0 * eta2 + 45.0 * et + 5.0) * dak;
    const double vu2 = (eta32 + 3.0 * eta2 + 42.0 * et - 2.0e1) * dak2;
    const double vu3 = (eta32 + 3.0e1 * et - 5.0) * dak4;
    const double vu4 = vu1 + e24 * ak * vu3, vu5 = eta6d * (vu2 + 4.0 * vu3);
    const double ph1 = eta6d / ak, ph2 = d - e12 * dak2;
    const double ta1 = (et + 5.0) / (5.0 * ak), ta2 = eta2d * dak2;
    const double ta3 = -e12 * gek * (ta1 + ta2), ta4 = eta3d * ak2 * (ta1 * ta1 - ta2 * ta2);
    const double ta5 = eta3d * (et + 8.0) * 1.0e-1 - 2.0 * eta22 * dak2;
    // sinh k, cosh k (e^-k dropped beyond k = 20)
    const double ex1 = exp(ak), ex2 = ak < 20.0 ? exp(-ak) : 0.0;
    const double sk = 0.5 * (ex1 - ex2), ck = 0.5 * (ex1 + ex2);
    const double ckma = ck - 1.0 - ak * sk, skma = sk - ak * ck;
    const double a1 = (e24 * gek * (al1 + al2 + ak1 * al3) - eta22) * dd4;
    const double b1 = (1.5 * et * eta2d2 - e12 * gek * (be1 + be2 + ak1 * be3)) * dd4;
    const double v1 = (eta21 * (eta2 - 2.0 * et + 1.0e1) * 2.5e-1 - gek * (vu4 + vu5)) * dd45;
    if (!big) {
        a2 = e24 * (al3 * skma + al2 * sk - al1 * ck) * dd4;
        a3 = e24 * (eta22 * dak2 - 0.5 * d2 + al3 * ckma - al1 * sk + al2 * ck) * dd4;
        b2 = e12 * (-be3 * skma - be2 * sk + be1 * ck) * dd4;
        b3 = e12 * (0.5 * d2 * eta2d - eta3d * eta2d2 * dak2 - be3 * ckma + be1 * sk - be2 * ck) * dd4;
        v2 = (vu4 * ck - vu5 * sk) * dd45;
        v3 = ((eta3 - 6.0 * eta2 + 5.0) * d - eta6d * (2.0 * eta3 - 3.0 * eta2 + 18.0 * et + 1.0e1) * dak2 +
              e24 * vu3 + vu4 * sk - vu5 * ck) * dd45;
    }
    const double pp1 = ph1 * ph1, pp2 = ph2 * ph2, pp = pp1 + pp2, p1p2 = ph1 * ph2 * 2.0;
    const double p1 = (gek * (pp1 + pp2 - p1p2) - 0.5 * eta2d) * dd2;
    if (!big) {
        p2 = (pp * sk + p1p2 * ck) * dd2;
        p3 = (pp * ck + p1p2 * sk + pp1 - pp2) * dd2;
    }
    const double t1 = ta3 + ta4 * a1 + ta5 * b1;
    double fa, ca;
    if (big) {
        // very large screening: the asymptotic solution
        v3 = ((eta3 - 6.0 * eta2 + 5.0) * d - eta6d * (2.0 * eta3 - 3.0 * eta2 + 18.0 * et + 1.0e1) * dak2 +
              e24 * vu3) * dd45;
        const double t3 = ta4 * a3 + ta5 * b3 + e12 * ta2 - 4.0e-1 * et * (et + 1.0e1) - 1.0;
        p3 = (pp1 - pp2) * dd2;
        b3 = e12 * (0.5 * d2 * eta2d - eta3d * eta2d2 * dak2 + be3) * dd4;
        a3 = e24 * (eta22 * dak2 - 0.5 * d2 - al3) * dd4;
        const double um6 = t3 * a3 - e12 * v3 * v3, um5 = t1 * a3 + a1 * t3 - e24 * v1 * v3;
        const double um4 = t1 * a1 - e12 * v1 * v1;
        const double l6 = e12 * p3 * p3, l5 = e24 * p1 * p3 - b3 - b3 - ak2, l4 = e12 * p1 * p1 - b1 - b1;
        fa = -(um4 * l6 - l4 * um6) / (um5 * l6 - l5 * um6);
        ca = -fa;
        st.f = fa;
        st.c = ca;
        st.b = b1 + b3 * fa;
        st.a = a1 + a3 * fa;
        st.v = v1 + v3 * fa;
        st.g1 = -(p1 + p3 * fa);
        fval = fabs(st.g1) < 1.0e-3 ? 0.0 : st.g1;
        st.eta = et;
    } else {
        const double t2 = ta4 * a2 + ta5 * b2 + e12 * (ta1 * ck - ta2 * sk);
        const double t3 = ta4 * a3 + ta5 * b3 + e12 * (ta1 * sk - ta2 * (ck - 1.0)) - 4.0e-1 * et * (et + 1.0e1) - 1.0;
        // mu
        const double um1 = t2 * a2 - e12 * v2 * v2, um2 = t1 * a2 + t2 * a1 - e24 * v1 * v2;
        const double um3 = t2 * a3 + t3 * a2 - e24 * v2 * v3, um4 = t1 * a1 - e12 * v1 * v1;
        const double um5 = t1 * a3 + t3 * a1 - e24 * v1 * v3, um6 = t3 * a3 - e12 * v3 * v3;
        if (mode == 1 || mode == 3) {
            // the full MSA: lambda, omega and the quartic in f
            const double l1 = e12 * p2 * p2, l2 = e24 * p1 * p2 - b2 - b2, l3 = e24 * p2 * p3;
            const double l4 = e12 * p1 * p1 - b1 - b1, l5 = e24 * p1 * p3 - b3 - b3 - ak2, l6 = e12 * p3 * p3;
            const double w16 = um1 * l6 - l1 * um6, w15 = um1 * l5 - l1 * um5, w14 = um1 * l4 - l1 * um4;
            const double w13 = um1 * l3 - l1 * um3, w12 = um1 * l2 - l1 * um2;
            const double w26 = um2 * l6 - l2 * um6, w25 = um2 * l5 - l2 * um5, w24 = um2 * l4 - l2 * um4;
            const double w36 = um3 * l6 - l3 * um6, w35 = um3 * l5 - l3 * um5, w34 = um3 * l4 - l3 * um4;
            const double w32 = um3 * l2 - l3 * um2;
            const double w46 = um4 * l6 - l4 * um6, w56 = um5 * l6 - l5 * um6;
            const double w3526 = w35 + w26, w3425 = w34 + w25;
            const double q4 = w16 * w16 - w13 * w36;
            const double q3 = 2.0 * w16 * w15 - w13 * w3526 - w12 * w36;
            const double q2 = w15 * w15 + 2.0 * w16 * w14 - w13 * w3425 - w12 * w3526;
            const double q1 = 2.0 * w15 * w14 - w13 * w24 - w12 * w3425;
            const double q0 = w14 * w14 - w12 * w24;
            double fap;
            if (mode == 1) {
                fap = (w14 - w34 - w46) / (w12 - w15 + w35 - w26 + w56 - w32);   // large k
            } else {
                // near the Gillan condition; the Percus-Yevick-like estimate
                // when both the contact potential and k are small
                double g = 0.5 * eta2d * dd2 * exp(-gek);
                if (gek <= 2.0 && gek >= 0.0 && ak <= 1.0) {
                    const double e24g = e24 * gek * exp(ak), pwk = sqrt(e24g);
                    const double qpw = (1.0 - sqrt(1.0 + 2.0 * d2 * d * pwk / eta22)) * eta21 / d;
                    g = -qpw * qpw / e24 + 0.5 * eta2d * dd2;
                }
                const double pg = p1 + g;
                double cst = ak2 * pg + 2.0 * (b3 * pg - b1 * p3) + e12 * g * g * p3;
                cst = -cst / (ak2 * p2 + 2.0 * (b3 * p2 - b2 * p3));
                fap = -(pg + p2 * cst) / p3;
            }
            // Newton on the quartic
            double del;
            int it = 0;
            do {
                if (++it > 40) return -2;
                fa = fap;
                const double fun = q0 + (q1 + (q2 + (q3 + q4 * fa) * fa) * fa) * fa;
                const double fund = q1 + (2.0 * q2 + (3.0 * q3 + 4.0 * q4 * fa) * fa) * fa;
                fap = fa - fun / fund;
                del = fabs((fap - fa) / fa);
            } while (del > 1.0e-6);
            ir = it;
            fa = fap;
            ca = -(w16 * fa * fa + w15 * fa + w14) / (w13 * fa + w12);
            st.g1 = -(p1 + p2 * ca + p3 * fa);
            fval = fabs(st.g1) < 1.0e-3 ? 0.0 : st.g1;
            st.eta = et;
        } else {
            // the Gillan condition g(1+) = 0 fixes c and f
            ca = ak2 * p1 + 2.0 * (b3 * p1 - b1 * p3);
            ca = -ca / (ak2 * p2 + 2.0 * (b3 * p2 - b2 * p3));
            fa = -(p1 + p2 * ca) / p3;
            if (mode == 2) fval = um1 * ca * ca + (um2 + um3 * fa) * ca + um4 + um5 * fa + um6 * fa * fa;
            else fval = -(p1 + p2 * ca + p3 * fa);
        }
        st.f = fa;
        st.c = ca;
        st.b = b1 + b2 * ca + b3 * fa;
        st.a = a1 + a2 * ca + a3 * fa;
        st.v = (v1 + v2 * ca + v3 * fa) / st.a;
    }
    const double g24 = e24 * gek * ex1;
    st.u = (ak * ak2 * ca - g24) / (ak2 * g24);
    return ir;
}

// The coefficients for (eta0, gek0, ak0) (hayter_msa.c:88-176): the plain
// MSA when it gives g(1+) >= 0, else the volume fraction rescaled (secant
// iteration) to meet the Gillan condition.  Negative on failure.
__host__ __device__ inline int solve(State &st)
{
    #pragma clang fp contract(off)
    double fval = 0.0;
    int ir = 0;
    bool ig = true;
    if (st.ak0 >= 1.0 + 8.0 * st.eta0) {
        ig = false;
        ir = coefficients(st, st.eta0, 1, fval);
        if (ir < 0 || fval >= 0.0) return ir;
        st.g1 = fval;
    }
    double e = fmin(st.eta0, 0.20);
    const double gamk = 2.0 * pow(st.eta0, 1.0 / 3.0) * st.gek0 * exp(st.ak0 - st.ak0 / pow(st.eta0, 1.0 / 3.0));
    if (!ig || gamk >= 0.15) {
        int ii = 0;
        double del, e2;
        do {
            if (++ii > 40) return -1;
            if (e <= 0.0) e = st.eta0 / ii;
            if (e > 0.6) e = 0.35 / ii;
            const double e1 = e;
            double f1, f2;
            coefficients(st, e1, 2, f1);
            coefficients(st, e * 1.01, 2, f2);
            e2 = e1 - (e * 1.01 - e1) * f1 / (f2 - f1);
            e = e2;
            del = fabs((e2 - e1) / e1);
        } while (del > 5.0e-6);
        coefficients(st, e2, 4, fval);
        st.g1 = fval;
        st.eta = e2;
        ir = ii;
        if (!ig || e >= st.eta0) return ir;
    }
    ir = coefficients(st, st.eta0, 3, fval);
    st.g1 = fval;
    if (ir >= 0 && fval < 0.0) ir = -3;
    return ir;
}

// S(q) from solved coefficients at q sigma (hayter_msa.c:469-541): the
// closed form of the Fourier transform of c(x), its Taylor expansion in
// (q sigma / s)^2 below 0.01 and -1/a at zero.
__host__ __device__ inline double sq(const State &st, double qsig)
{
    #pragma clang fp contract(off)
    const double eta = st.eta, ak = st.ak, gek = st.gek;
    const double e24 = 24.0 * eta;
    const double x1 = exp(ak), x2 = ak < 20.0 ? exp(-ak) : 0.0;
    const double ck = 0.5 * (x1 + x2), sk = 0.5 * (x1 - x2), ak2 = ak * ak;
    const double qk = qsig / st.s, q2k = qk * qk;
    if (qk <= 1.0e-08) return -1.0 / st.a;
    if (qk <= 0.01) {
        const double aqk = st.a * (8.0 + 2.0 * eta) + 6 * st.b - 12.0 * st.f -
                           24 * (gek * (1.0 + ak) - ck * ak * st.c + st.f * (ck - 1.0) + (st.c - st.f * ak) * sk) / ak2 +
                           q2k * (-(st.a * (48.0 + 15.0 * eta) + 40.0 * st.b) / 60.0 + st.f +
                                  (4.0 / ak2) * (gek * (9.0 + 7.0 * ak) + ck * (9.0 * st.f - 7.0 * st.c * ak) +
                                                 sk * (9.0 * st.c - 7.0 * st.f * ak)));
        return 1.0 / (1.0 - eta * aqk);
    }
    const double qk2 = 1.0 / q2k, qk3 = qk2 / qk, qqk = 1.0 / (qk * (q2k + ak2));
    double sink, cosk;
    sincos(qk, &sink, &cosk);
    const double asink = ak * sink, qcosk = qk * cosk;
    double aqk = st.a * (sink - qcosk);
    aqk = aqk + st.b * ((2.0 * qk2 - 1.0) * qcosk + 2.0 * sink - 2.0 / qk);
    const double inter = 24.0 * qk3 + 4.0 * (1.0 - 6.0 * qk2) * sink;
    aqk = (aqk + 0.5 * eta * st.a * (inter - (1.0 - 12.0 * qk2 + 24.0 * qk2 * qk2) * qcosk)) * qk3;
    aqk = aqk + st.c * (ck * asink - sk * qcosk) * qqk;
    aqk = aqk + st.f * (sk * asink - qk * (ck * cosk - 1.0)) * qqk;
    aqk = aqk + st.f * (cosk - 1.0) * qk2;
    aqk = aqk - gek * (asink + qcosk) * qqk;
    return 1.0 / (1.0 - e24 * aqk);
}

}  // namespace msa

// Hayter-Penfold rescaled MSA for sphere radius (A), volume fraction,
// charge (e), temperature (K), 1:1 salt concentration (mol/l) and the
// solvent's relative permittivity: the unscaled volume fraction, contact
// potential and screening constant.  The ionic strength counts the
// monovalent counterions of the spheres and the salt; the Debye screening
// constant and the contact potential follow in SI units with the constants
// of the reference.
__host__ __device__ inline void hayter_msa_setup(double radius, double phi, double charge, double temperature,
                                                 double csalt, double dielec, msa::State &st)
{
    #pragma clang fp contract(off)
    const double e_charge = 1.602189e-19, k_b = 1.380662e-23, eps0 = 8.85418782E-12;
    const double diam = 2.0 * radius;
    const double beta = 1.0 / (k_b * temperature);
    const double perm = dielec * eps0;
    const double zc = charge * e_charge;
    const double sidiam = diam * 1.0E-10;
    const double half = sidiam / 2.0;
    const double vp = 4.18879020478639 * (half * half * half);   // 4 pi / 3 as the reference's header rounds it
    const double cs = csalt * 6.022E23 * 1.0E3;
    const double ionst = 0.5 * e_charge * e_charge * (charge * phi / vp + 2.0 * cs);
    const double kappa = sqrt(2 * beta * ionst / perm);
    st.gek0 = beta * zc * zc / (3.141592653589793 * perm * sidiam * ((2.0 + kappa * sidiam) * (2.0 + kappa * sidiam)));
    st.ak0 = kappa * sidiam;
    st.eta0 = phi;
    st.eta = phi;
}

// S(q) of the Hayter MSA; NaN when the coefficients do not converge.
__host__ __device__ inline double hayter_msa_sq(double q, double radius, double phi, double charge, double temperature,
                                                double csalt, double dielec)
{
    msa::State st;
    hayter_msa_setup(radius, phi, charge, temperature, csalt, dielec, st);
    if (msa::solve(st) < 0) return __builtin_nan("");
    return msa::sq(st, q * (2.0 * radius));
}

// S(q) of structure-factor model `id` with p = (radius_effective,
// volfraction, its further parameters: up to four).
__host__ __device__ inline double structure_sq(int id, double q, const double *p)
{
    if (id == S_STICKY) return stickyhardsphere_sq(q, p[0], p[1], p[2], p[3]);
    if (id == S_SQUAREWELL) return squarewell_sq(q, p[0], p[1], p[2], p[3]);
    if (id == S_HAYTER) return hayter_msa_sq(q, p[0], p[1], p[2], p[3], p[4], p[5]);
    return hardsphere_sq(q, p[0], p[1]);
}

// The q-independent part of a structure factor for one parameter set, in
// S_PREP doubles: the solved Hayter MSA coefficients (a, b, c, f, eta, gek,
// k, s; a = NaN when they do not converge), nothing for the closed forms.
// Solving once per set keeps the iteration (and its registers) out of the
// kernel that evaluates S at every q.
constexpr int S_PREP = 8;
__host__ __device__ inline void structure_prepare(int id, const double *p, double *prep)
{
    if (id != S_HAYTER) return;
    msa::State st;
    hayter_msa_setup(p[0], p[1], p[2], p[3], p[4], p[5], st);
    const bool ok = msa::solve(st) >= 0;
    prep[0] = ok ? st.a : __builtin_nan("");
    prep[1] = st.b;
    prep[2] = st.c;
    prep[3] = st.f;
    prep[4] = st.eta;
    prep[5] = st.gek;
    prep[6] = st.ak;
    prep[7] = st.s;
}

// S(q) from structure_prepare's output (same value as structure_sq).
__host__ __device__ inline double structure_sq_prepared(int id, double q, const double *p, const double *prep)
{
    if (id == S_STICKY) return stickyhardsphere_sq(q, p[0], p[1], p[2], p[3]);
    if (id == S_SQUAREWELL) return squarewell_sq(q, p[0], p[1], p[2], p[3]);
    if (id == S_HAYTER) {
        if (prep[0] != prep[0]) return prep[0];   // NaN: the iteration did not converge
        msa::State st;
        st.a = prep[0];
        st.b = prep[1];
        st.c = prep[2];
        st.f = prep[3];
        st.eta = prep[4];
        st.gek = prep[5];
        st.ak = prep[6];
        st.s = prep[7];
        return msa::sq(st, q * (2.0 * p[0]));
    }
    return hardsphere_sq(q, p[0], p[1]);
}

}  // namespace sas
