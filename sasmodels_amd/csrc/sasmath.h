// sasmath.h -- FP32 special functions for the I(q) kernels (gfx950 VALU).
//
// Everything here is __host__ __device__ so the same code can be checked on
// the CPU against the FP64 oracle.  Division and square root go through the
// hardware reciprocal/sqrt on the device (1 ulp), which is well inside the
// 5e-5 relative budget of the kernels.
//
// Functions and the published methods they use:
//   sincos_f      Cody-Waite reduction by pi/2 with a three-part split of
//                 pi/2 (Cody & Waite, "Software Manual for the Elementary
//                 Functions", 1980); FMA makes the first step exact for
//                 k < 2^24.  Taylor polynomials on |r| <= pi/4.  Arguments
//                 above 1e3 are reduced in FP64 from the exact product.
//   sinc_f        sin(x)/x, 1 at x == 0 (sas_sinx_x, kernel_header.c:177).
//   sph_j1c_f     3 j1(x)/x: Taylor series below 1, direct form above
//                 (the same two-branch scheme as sas_3j1x_x.c:46-57).
//   *_d           the same functions for arguments known in FP64: the
//                 reduction by pi/2 runs in FP64, the polynomials in FP32, so
//                 the phase error does not grow with x (2D kernels, where a
//                 single orientation puts pixels next to form-factor zeros).
//   j1c_f         2 J1(x)/x from the Cephes single-precision j1f
//                 approximations (S. L. Moshier, Cephes Math Library 2.8):
//                 a polynomial on |x| <= 2 and the modulus/phase asymptotic
//                 form beyond, as sas_J1.c:190-245 uses for float kernels;
//                 j1c_small / j1c_big are the two branches (j1c_big takes
//                 sin and cos of x from the caller).
//   sinc_taylor   sin(x)/x on |x| <= 2 by its Taylor series.
//   *_rev         the same functions of a phase given in revolutions,
//                 with the hardware sine and cosine (device only; the FP32
//                 path of the 2D kernels).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define SAS_HD __host__ __device__ __forceinline__

namespace sas {

SAS_HD float rcp_f(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

SAS_HD float sqrt_f(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

SAS_HD float rsqrt_f(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// sin and cos of r for |r| <= pi/4 + tiny (Taylor to r^9 / r^10:
// truncation below 2e-9 absolute, under half an FP32 ulp).
SAS_HD void sincos_poly(float r, float &s, float &c)
{
    const float r2 = r * r;
    float ps = fmaf(r2, 2.75573192e-6f, -1.98412698e-4f);   // 1/9!, -1/7!
    ps = fmaf(r2, ps, 8.33333333e-3f);                        // 1/5!
    ps = fmaf(r2, ps, -1.66666667e-1f);                       // -1/3!
    s = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.75573192e-7f, 2.48015873e-5f);   // -1/10!, 1/8!
    pc = fmaf(r2, pc, -1.38888889e-3f);                       // -1/6!
    pc = fmaf(r2, pc, 4.16666667e-2f);                        // 1/4!
    pc = fmaf(r2, pc, -0.5f);                                 // -1/2!
    c = fmaf(r2, pc, 1.0f);
}

// Rotate (sin r, cos r) by quadrant k: returns sin/cos of r + k*pi/2.
SAS_HD void quadrant(int k, float s, float c, float &so, float &co)
{
    const bool swap = (k & 1) != 0;
    const float a = swap ? c : s;
    const float b = swap ? s : c;
    so = (k & 2) ? -a : a;
    co = ((k + 1) & 2) ? -b : b;
}

// sin and cos of x = a*b (the product of two FP32 operands).
SAS_HD void sincos_prod(float a, float b, float &s, float &c)
{
    const float x = a * b;
    float r;
    int k;
    if (fabsf(x) <= 1.0e3f) {
        const float kf = rintf(x * 0x1.45f306p-1f);   // x * 2/pi
        r = fmaf(-kf, 0x1.921fb6p+0f, x);             // exact (Sterbenz, FMA)
        r = fmaf(-kf, -0x1.777a5cp-25f, r);
        r = fmaf(-kf, -0x1.ee59dap-50f, r);
        k = (int)kf;
    } else {
        // FP64 reduction of the exact product (24+24 bits fit in 53).
        const double xd = (double)a * (double)b;
        const double kd = rint(xd * 0x1.45f306dc9c883p-1);
        double rd = fma(-kd, 0x1.921fb54442d18p+0, xd);
        rd = fma(-kd, 0x1.1a62633145c07p-54, rd);
        r = (float)rd;
        k = (int)(long long)kd;
    }
    float sr, cr;
    sincos_poly(r, sr, cr);
    quadrant(k & 3, sr, cr, s, c);
}

SAS_HD void sincos_f(float x, float &s, float &c) { sincos_prod(x, 1.0f, s, c); }

SAS_HD float sin_f(float x)
{
    float s, c;
    sincos_prod(x, 1.0f, s, c);
    return s;
}

// sin(x)/x
SAS_HD float sinc_f(float x)
{
    if (x == 0.0f) return 1.0f;
    return sin_f(x) * rcp_f(x);
}

// sin(a*b)/(a*b)
SAS_HD float sinc_prod(float a, float b)
{
    const float x = a * b;
    if (x == 0.0f) return 1.0f;
    float s, c;
    sincos_prod(a, b, s, c);
    return s * rcp_f(x);
}

// 3 j1(x)/x = 3 (sin x - x cos x) / x^3
SAS_HD float sph_j1c_f(float x)
{
    const float ax = fabsf(x);
    if (ax < 1.0f) {
        // sum_k (-1)^k 3 (2k+2) x^(2k) / (2k+3)!, k = 0..5
        const float z = x * x;
        float p = fmaf(z, -5.78125e-9f /* -1/172972800 */, 7.51563251e-7f /* 1/1330560 */);
        p = fmaf(z, p, -6.61375661e-5f);   // -1/15120
        p = fmaf(z, p, 3.57142857e-3f);    // 1/280
        p = fmaf(z, p, -0.1f);             // -1/10
        return fmaf(z, p, 1.0f);
    }
    float s, c;
    sincos_f(x, s, c);
    const float inv = rcp_f(x);
    return 3.0f * (s * inv - c) * inv * inv;
}

// sin(x)/x for |x| <= 2 by its Taylor series to x^14 (truncation below
// 2e-10 at |x| = 2)
SAS_HD float sinc_taylor(float x)
{
    const float z = x * x;
    float p = fmaf(z, -7.64716373e-13f, 1.60590438e-10f);   // -1/15!, 1/13!
    p = fmaf(z, p, -2.50521084e-8f);                          // -1/11!
    p = fmaf(z, p, 2.75573192e-6f);                           // 1/9!
    p = fmaf(z, p, -1.98412698e-4f);                          // -1/7!
    p = fmaf(z, p, 8.33333333e-3f);                           // 1/5!
    p = fmaf(z, p, -1.66666667e-1f);                          // -1/3!
    return fmaf(z, p, 1.0f);
}

// 2 J1(x)/x on |x| <= 2 (Cephes j1f polynomial)
SAS_HD float j1c_small(float x)
{
    const float z = x * x;
    float p = fmaf(z, -4.878788132172128e-9f, 6.009061827883699e-7f);
    p = fmaf(z, p, -4.541343896997497e-5f);
    p = fmaf(z, p, 1.937383947804541e-3f);
    p = fmaf(z, p, -3.405537384615824e-2f);
    return 2.0f * (z - 14.681970596f) * p;
}

// Modulus and phase part of J1 for x > 2 (Cephes j1f modulus/phase form):
// with q = 1/x, sx = sin x, cx = cos x,
//   J1(x) = sqrt(q) / sqrt(2) * j1_mc(q, sx, cx),
//   j1_mc = M(q) [sin xn (sx + cx) + cos xn (sx - cx)],  xn = q PH(q^2),
// i.e. sqrt(q) M(q) cos(x + xn - 3pi/4).
SAS_HD float j1_mc(float q, float sx, float cx)
{
    float m = fmaf(q, 6.913942741265801e-2f, -2.284801500053359e-1f);
    m = fmaf(q, m, 3.138238455499697e-1f);
    m = fmaf(q, m, -2.102302420403875e-1f);
    m = fmaf(q, m, 5.435364690523026e-3f);
    m = fmaf(q, m, 1.493389585089498e-1f);
    m = fmaf(q, m, 4.976029650847191e-6f);
    m = fmaf(q, m, 7.978845453073848e-1f);
    const float q2 = q * q;
    float ph = fmaf(q2, -4.497014141919556e+1f, 5.073465654089319e+1f);
    ph = fmaf(q2, ph, -2.485774108720340e+1f);
    ph = fmaf(q2, ph, 7.222973196770240e+0f);
    ph = fmaf(q2, ph, -1.544842782180211e+0f);
    ph = fmaf(q2, ph, 3.503787691653334e-1f);
    ph = fmaf(q2, ph, -1.637986776941202e-1f);
    ph = fmaf(q2, ph, 3.749989509080821e-1f);
    const float xn = q * ph;
    // |xn| <= 0.19 here: short Taylor for its sine and cosine
    const float xn2 = xn * xn;
    const float sxn = xn * fmaf(xn2, fmaf(xn2, 8.33333333e-3f, -1.66666667e-1f), 1.0f);
    const float cxn = fmaf(xn2, fmaf(xn2, fmaf(xn2, -1.38888889e-3f, 4.16666667e-2f), -0.5f), 1.0f);
    // cos(x + xn - 3pi/4) sqrt(2) = sin xn (sin x + cos x) + cos xn (sin x - cos x)
    return m * (sxn * (sx + cx) + cxn * (sx - cx));
}

// 2 J1(x)/x for x = ax > 2 given sx = sin(ax), cx = cos(ax)
SAS_HD float j1c_big(float ax, float sx, float cx)
{
    const float q = rcp_f(ax);
    const float w = sqrt_f(q);
    const float j1 = w * 0.70710678118654752f * j1_mc(q, sx, cx);
    return 2.0f * j1 * q;   // even function: sign of x cancels
}

// 2 J1(x)/x
SAS_HD float j1c_f(float x)
{
    const float ax = fabsf(x);
    if (ax <= 2.0f) return j1c_small(x);
    float sx, cx;
    sincos_f(ax, sx, cx);
    return j1c_big(ax, sx, cx);
}

// ---- arguments known in FP64 (the 2D kernels): FP64 reduction by pi/2,
// FP32 polynomials and FP32 for every factor that does not carry the phase.
SAS_HD void sincos_d(double x, float &s, float &c)
{
    const double kd = rint(x * 0x1.45f306dc9c883p-1);
    double rd = fma(-kd, 0x1.921fb54442d18p+0, x);
    rd = fma(-kd, 0x1.1a62633145c07p-54, rd);
    float sr, cr;
    sincos_poly((float)rd, sr, cr);
    quadrant((int)(long long)kd & 3, sr, cr, s, c);
}

SAS_HD float sinc_d(double x)
{
    if (x == 0.0) return 1.0f;
    float s, c;
    sincos_d(x, s, c);
    return s * rcp_f((float)x);
}

SAS_HD float sph_j1c_d(double x)
{
    const float xf = (float)x;
    if (fabsf(xf) < 1.0f) return sph_j1c_f(xf);
    float s, c;
    sincos_d(x, s, c);
    const float inv = rcp_f(xf);
    return 3.0f * (s * inv - c) * inv * inv;
}

SAS_HD float j1c_d(double x)
{
    const double axd = fabs(x);
    const float ax = (float)axd;
    if (ax <= 2.0f) return j1c_f(ax);
    const float q = rcp_f(ax);
    const float w = sqrt_f(q);
    float sx, cx;
    sincos_d(axd, sx, cx);
    const float j1 = w * 0.70710678118654752f * j1_mc(q, sx, cx);
    return 2.0f * j1 * q;
}

// ---- phases in revolutions (the FP32 path of the 2D kernels).  A phase
// t = x / (2 pi) computed in FP32 from lengths already divided by 2 pi goes
// straight to the hardware's v_sin_f32 / v_cos_f32, which take revolutions
// (valid for |t| < 256): measured on gfx950 their error beyond the rounding
// of t is below 3e-7 absolute, so the phase error is that of computing t in
// FP32 (about 1e-7 of x), and no reduction by pi/2 is needed.  Device only.
constexpr float TWO_PI_F = 6.283185307179586f;
constexpr float INV_TWO_PI_F = 0.15915494309189535f;
constexpr float FOUR_PI2_F = 39.47841760435743f;

__device__ __forceinline__ void sincos_rev(float t, float &s, float &c)
{
    s = __builtin_amdgcn_sinf(t);
    c = __builtin_amdgcn_cosf(t);
}

// 3 j1(x)/x at x = 2 pi t, from t^2: the series below x = 2 (to x^10,
// truncation below 5e-10), the closed form above, where one reciprocal
// square root gives both 1/t and t.
__device__ __forceinline__ float sph_j1c_rev2(float t2)
{
    const float z = t2 * FOUR_PI2_F;
    if (z < 4.0f) {
        float p = fmaf(z, -5.78125e-9f, 7.51563251e-7f);   // -1/172972800, 1/1330560
        p = fmaf(z, p, -6.61375661e-5f);                     // -1/15120
        p = fmaf(z, p, 3.57142857e-3f);                      // 1/280
        p = fmaf(z, p, -0.1f);                               // -1/10
        return fmaf(z, p, 1.0f);
    }
    const float it = rsqrt_f(t2);
    float s, c;
    sincos_rev(t2 * it, s, c);
    const float inv = it * INV_TWO_PI_F;                     // 1/x
    return 3.0f * (s * inv - c) * inv * inv;
}

// 2 J1(x)/x at x = 2 pi t, t >= 0: the Cephes j1f branches of j1c_f with
// sin and cos of x from the phase.
__device__ __forceinline__ float j1c_rev(float t)
{
    const float x = t * TWO_PI_F;
    if (x <= 2.0f) return j1c_small(x);
    float s, c;
    sincos_rev(t, s, c);
    return j1c_big(x, s, c);
}

// sin(x)/x at x = 2 pi t
__device__ __forceinline__ float sinc_rev(float t)
{
    const float x = t * TWO_PI_F;
    if (fabsf(x) <= 2.0f) return sinc_taylor(x);
    return __builtin_amdgcn_sinf(t) * rcp_f(x);
}

// FP64 helpers used where the kernels need double (volumes, weights).
SAS_HD double cube_d(double x) { return x * x * x; }

// 3 j1(x)/x in FP64: series sum_k (-1)^k 3 (2k+2) x^(2k) / (2k+3)! below
// |x| = 0.25 (to x^10: remainder < 1e-17), the closed form above.
SAS_HD double sph_j1c_dd(double x)
{
    const double z = x * x;
    if (z < 0.0625) {
        double p = fma(z, -1.0 / 172972800.0, 1.0 / 1330560.0);
        p = fma(z, p, -1.0 / 15120.0);
        p = fma(z, p, 1.0 / 280.0);
        p = fma(z, p, -0.1);
        return fma(z, p, 1.0);
    }
    double s, c;
    sincos(x, &s, &c);
    return 3.0 * (s - x * c) / (z * x);
}

// A double stored in two consecutive words of an FP32 record.
SAS_HD double get_d(const float *r, int i)
{
    double d;
    __builtin_memcpy(&d, r + i, sizeof d);
    return d;
}
SAS_HD void put_d(float *r, int i, double v) { __builtin_memcpy(r + i, &v, sizeof v); }

}  // namespace sas
