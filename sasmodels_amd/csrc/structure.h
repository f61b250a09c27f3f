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
#pragma once
#include <hip/hip_runtime.h>

namespace sas {

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

}  // namespace sas
