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
// solution (stickyhardsphere.c), and the square-well fluid in the mean
// spherical approximation (squarewell.c); both FP64 like the reference.
#pragma once
#include <hip/hip_runtime.h>

namespace sas {

// Structure-factor model ids of the C ABI.
constexpr int S_HARDSPHERE = 5, S_STICKY = 14, S_SQUAREWELL = 15;

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

// S(q) of structure-factor model `id` with p = (radius_effective,
// volfraction, its third and fourth parameter).
__host__ __device__ inline double structure_sq(int id, double q, const double *p)
{
    if (id == S_STICKY) return stickyhardsphere_sq(q, p[0], p[1], p[2], p[3]);
    if (id == S_SQUAREWELL) return squarewell_sq(q, p[0], p[1], p[2], p[3]);
    return hardsphere_sq(q, p[0], p[1]);
}

}  // namespace sas
