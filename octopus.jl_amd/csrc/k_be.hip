// Bassetti-Erskine field of a transverse Gaussian charge distribution and the Faddeeva function it
// is built on, on the device.
//
// gaussian_beambeam_kick(sigx, sigy, x, y) (strong_beam.jl:1139-1146): the normalised kick of a
// Gaussian of rms sizes (sigx, sigy) at (x, y); the caller multiplies by the beam-beam strength.
// Per point, on the principal axes (sig1 >= sig2, strong_beam.jl:1040-1059):
//   round (eta = 0):          K = (x, y) (1 - exp(-u)) / (u v), u = r^2 / 2v        (:705-709)
//   near round (eta <= eta*/2): the cubic series in eta of the radial moments m0..m6  (:799-895)
//   elliptic (eta >= eta*):   K = Im/Re of A (w(z1) - B w(z2)), or the fifth-degree axis series
//                             when rho^7 <= eps / sqrt(eta)                        (:962-1038)
//   between:                  a quintic smoothstep from the series to the elliptic form  (:785-797)
// with v = (sig1^2 + sig2^2)/2, eta = (sig1^2 - sig2^2)/2v and eta* = (8/3 * 64 eps)^(1/4).
//
// w(z) = exp(-z^2) erfc(-iz) for Im z >= 0 (SpecialMath.jl:77-151): Weideman's rational approximation
// with N = 32 terms (J. A. C. Weideman, SIAM J. Numer. Anal. 31 (1994) 1497-1518) inside
// |x| <= 6, y <= 7 (wider bands near the real axis), the Laplace continued fraction outside, and its
// one- and two-term asymptotic forms beyond |x| + y > 4000 and 1e7; the lower half plane by
// w(z) = 2 exp(-z^2) - w(-z) (:60-67).  The 32 coefficients are computed here from Weideman's
// construction (a 128-point DFT of exp(-t^2)(L^2 + t^2) at t = L tan(theta/2), L = sqrt(N/sqrt 2)),
// in long double, once per process.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "launch.h"

namespace octopic {

namespace {

constexpr double kInvSqrtPi = 0.56418958354775628695;
constexpr double kSqrtPi = 1.7724538509055160273;
constexpr double kSqrt2 = 1.4142135623730950488;

__device__ __forceinline__ void cmul(double ar, double ai, double br, double bi, double &r, double &i) {
    r = ar * br - ai * bi;
    i = ar * bi + ai * br;
}

__device__ __forceinline__ void cdiv(double ar, double ai, double br, double bi, double &r, double &i) {
    const double inv = 1.0 / (br * br + bi * bi);
    r = (ar * br + ai * bi) * inv;
    i = (ai * br - ar * bi) * inv;
}

// Laplace continued fraction of w for large |z| (Im z >= 0), with Gautschi's term-count estimate
// nu = floor(3.9 + 11.398 / (0.08254 x + 0.1421 y + 0.2023)) (SpecialMath.jl:113-151)
__device__ void w_far(double zr, double zi, double &wr, double &wi) {
    const double x = fabs(zr), y = zi;
    if (x + y > 4000.0) {
        if (x + y > 1e7) {   // w ~ i / (sqrt(pi) z)
            if (x > y) {
                const double q = y / zr, d = kInvSqrtPi / (zr + q * y);
                wr = d * q; wi = d;
            } else {
                const double q = zr / y, d = kInvSqrtPi / (q * zr + y);
                wr = d; wi = d * q;
            }
            return;
        }
        // w ~ i z / (sqrt(pi) (z^2 - 1/2))
        const double dr = zr * zr - y * y - 0.5, di = 2.0 * zr * y;
        const double d = kInvSqrtPi / (dr * dr + di * di);
        wr = d * (zr * di - y * dr);
        wi = d * (zr * dr + y * di);
        return;
    }
    double nu = floor(3.9 + 11.398 / (0.08254 * x + 0.1421 * y + 0.2023));
    double tr = zr, ti = y;
    nu = 0.5 * (nu - 1.0);
    while (nu > 0.4) {   // t <- z - nu / t, nu = (n-1)/2, ..., 1/2
        const double s = nu / (tr * tr + ti * ti);
        const double nr = zr - tr * s, ni = y + ti * s;
        tr = nr; ti = ni;
        nu -= 0.5;
    }
    const double s = kInvSqrtPi / (tr * tr + ti * ti);   // i / (sqrt(pi) t)
    wr = s * ti;
    wi = s * tr;
}

__device__ void w_upper(const WeidemanCoeffs &wc, double zr, double zi, double &wr, double &wi) {
    const double x = fabs(zr), y = zi;
    if (y > 7.0 || (x > 6.0 && (y > 0.1 || (x > 8.0 && y > 1e-10) || x > 28.0))) {
        w_far(zr, y, wr, wi);
        return;
    }
    // Z = (L + iz) / (L - iz); w = 2 p(Z) / (L - iz)^2 + 1 / (sqrt(pi) (L - iz))
    const double L = wc.L;
    const double dr = L + zi, di = -zr;
    double Zr, Zi;
    cdiv(L - zi, zr, dr, di, Zr, Zi);
    double pr = 0.0, pi_ = 0.0;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        double r, i;
        cmul(pr, pi_, Zr, Zi, r, i);
        pr = r + wc.c[k];
        pi_ = i;
    }
    double d2r, d2i, tr, ti, ir, ii;
    cmul(dr, di, dr, di, d2r, d2i);
    cdiv(2.0 * pr, 2.0 * pi_, d2r, d2i, tr, ti);
    cdiv(kInvSqrtPi, 0.0, dr, di, ir, ii);
    wr = tr + ir;
    wi = ti + ii;
}

__device__ void w_any(const WeidemanCoeffs &wc, double zr, double zi, double &wr, double &wi) {
    if (zi < 0.0) {   // w(z) = 2 exp(-z^2) - w(-z)
        double ur, ui;
        w_upper(wc, -zr, -zi, ur, ui);
        const double e = exp(-zr * zr + zi * zi), a = -2.0 * zr * zi;
        wr = 2.0 * (e * cos(a)) - ur;
        wi = 2.0 * (e * sin(a)) - ui;
        return;
    }
    w_upper(wc, zr, zi, wr, wi);
}

// eta* of the FP64 near-round blend: (8/3 * 64 * 2^-52)^(1/4) (strong_beam.jl:770-783)
__device__ __forceinline__ double near_round_outer() { return sqrt(sqrt((8.0 / 3.0) * 64.0 * 0x1p-52)); }

// radial moments m_k(q) = int_0^1 s^k exp(-q s) ds, k = 0..6: 25-term series for q <= 2, the upward
// recurrence m_k = (k m_{k-1} - e^-q) / q above (strong_beam.jl:799-833)
__device__ void moments_0_6(double q, double m[7]) {
    if (q <= 2.0) {
        for (int j = 0; j < 7; ++j) m[j] = 0.0;
        double term = 1.0;
        for (int k = 0; k <= 24; ++k) {
            for (int j = 0; j < 7; ++j) m[j] += term / (double)(k + 1 + j);
            term *= -q / (double)(k + 1);
        }
        return;
    }
    const double eq = exp(-q);
    m[0] = -expm1(-q) / q;
    for (int j = 1; j < 7; ++j) m[j] = ((double)j * m[j - 1] - eq) / q;
}

// cubic series in eta about the round beam (strong_beam.jl:878-895)
__device__ void near_round_kick(double v, double eta, double x, double y, double &kx, double &ky) {
    const double q = (x * x + y * y) / (2.0 * v), dr = (x * x - y * y) / (2.0 * v);
    double m[7];
    moments_0_6(q, m);
    const double e2 = eta * eta, e3 = e2 * eta;
    const double c1x = dr * m[2] - m[1], c1y = dr * m[2] + m[1];
    const double c2x = 1.5 * m[2] + (-dr - q) * m[3] + (dr * dr / 2.0) * m[4];
    const double c2y = 1.5 * m[2] + (dr - q) * m[3] + (dr * dr / 2.0) * m[4];
    const double c3x = -1.5 * m[3] + (2.5 * dr + q) * m[4] + (-dr * q - dr * dr / 2.0) * m[5] + dr * dr * dr / 6.0 * m[6];
    const double c3y = 1.5 * m[3] + (2.5 * dr - q) * m[4] + (-dr * q + dr * dr / 2.0) * m[5] + dr * dr * dr / 6.0 * m[6];
    kx = x / v * (m[0] + eta * c1x + e2 * c2x + e3 * c3x);
    ky = y / v * (m[0] + eta * c1y + e2 * c2y + e3 * c3y);
}

// fifth-degree axis series of one component (strong_beam.jl:974-1007): x times the expansion of
// Kx / x in x^2 and y^2
__device__ double axis_component(double sx, double sy, double x, double y) {
    const double s = sx + sy, sx2 = sx * sx, sx3 = sx2 * sx, sx5 = sx3 * sx2, sy2 = sy * sy, sy3 = sy2 * sy;
    const double s2 = s * s, s3 = s2 * s;
    const double j00 = 2.0 / (sx * s);
    const double j10 = 2.0 * (2.0 * sx + sy) / (3.0 * sx3 * s2);
    const double j01 = 2.0 / (sx * sy * s2);
    const double j20 = 2.0 * (8.0 * sx2 + 9.0 * sx * sy + 3.0 * sy2) / (15.0 * sx5 * s3);
    const double j11 = 2.0 * (3.0 * sx + sy) / (3.0 * sx3 * sy * s3);
    const double j02 = 2.0 * (sx + 3.0 * sy) / (3.0 * sx * sy3 * s3);
    const double x2 = x * x, y2 = y * y;
    const double scale = j00 - (x2 * j10 + y2 * j01) / 2.0 + x2 * x2 * j20 / 8.0 + x2 * y2 * j11 / 4.0 + y2 * y2 * j02 / 8.0;
    return x * scale;
}

// the elliptic Bassetti-Erskine form on the principal axes (strong_beam.jl:1016-1038)
__device__ void elliptic_kick(const WeidemanCoeffs &wc, double s1, double s2, double x, double y, double eta,
                              double &kx, double &ky) {
    const double rho2 = x * x / (s1 * s1) + y * y / (s2 * s2);
    if (rho2 * rho2 * rho2 * sqrt(rho2) <= 0x1p-52 / sqrt(eta)) {
        kx = axis_component(s1, s2, x, y);
        ky = axis_component(s2, s1, y, x);
        return;
    }
    const bool negx = x < 0.0, negy = y < 0.0;
    x = fabs(x);
    y = fabs(y);
    const double den = kSqrt2 * sqrt(s1 * s1 - s2 * s2);
    double w1r, w1i, w2r, w2i;
    w_upper(wc, x / den, y / den, w1r, w1i);
    w_upper(wc, s2 / s1 * x / den, s1 / s2 * y / den, w2r, w2i);
    const double A = 2.0 * kSqrtPi / den;
    const double B = exp(-x * x / (2.0 * s1 * s1) - y * y / (2.0 * s2 * s2));
    const double rr = A * (w1r - B * w2r), ri = A * (w1i - B * w2i);
    kx = negx ? -ri : ri;
    ky = negy ? -rr : rr;
}

// sig1 >= sig2 (strong_beam.jl:1040-1059)
__device__ void principal_kick(const WeidemanCoeffs &wc, double s1, double s2, double x, double y, double &kx,
                               double &ky) {
    const double v = (s1 * s1 + s2 * s2) / 2.0;
    const double eta = (s1 * s1 - s2 * s2) / (2.0 * v);
    if (eta == 0.0) {
        const double r2 = x * x + y * y, u = r2 / (2.0 * v);
        const double phi = u == 0.0 ? 1.0 : -expm1(-u) / u;
        kx = phi / v * x;
        ky = phi / v * y;
        return;
    }
    const double outer = near_round_outer(), inner = outer / 2.0;
    if (eta <= inner) {
        near_round_kick(v, eta, x, y, kx, ky);
        return;
    }
    double ex, ey;
    elliptic_kick(wc, s1, s2, x, y, eta, ex, ey);
    if (eta >= outer) {
        kx = ex; ky = ey;
        return;
    }
    const double t = (eta - inner) * (1.0 / (outer - inner));
    const double w = t * t * t * (10.0 + t * (-15.0 + 6.0 * t));
    double sx, sy;
    near_round_kick(v, eta, x, y, sx, sy);
    kx = sx + w * (ex - sx);
    ky = sy + w * (ey - sy);
}

__global__ __launch_bounds__(256) void k_be_kick(WeidemanCoeffs wc, const double *X, const double *Y, int64_t n,
                                                 double sigx, double sigy, double *KX, double *KY) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = X[i], y = Y[i];
    double kx = 0.0, ky = 0.0;
    if (sigx != 0.0 && sigy != 0.0) {
        if (sigx >= sigy) principal_kick(wc, sigx, sigy, x, y, kx, ky);
        else principal_kick(wc, sigy, sigx, y, x, ky, kx);
    }
    KX[i] = kx;
    KY[i] = ky;
}

__global__ __launch_bounds__(256) void k_faddeeva(WeidemanCoeffs wc, const double *ZR, const double *ZI, int64_t n,
                                                  double *WR, double *WI) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wr, wi;
    w_any(wc, ZR[i], ZI[i], wr, wi);
    WR[i] = wr;
    WI[i] = wi;
}

}  // namespace

const WeidemanCoeffs &weideman_coeffs() {
    static WeidemanCoeffs wc;
    static std::once_flag once;
    std::call_once(once, [] {
        constexpr int N = 32, M = 2 * N, M2 = 2 * M;
        const long double pi = 3.141592653589793238462643383279502884L;
        const long double L = sqrtl((long double)N / sqrtl(2.0L));
        // f_k = exp(-t^2)(L^2 + t^2), t = L tan(k pi / 2M), k = -M+1..M-1, placed in DFT order (index k
        // mod 2M; index M, the point at infinity, is zero)
        long double f[M2];
        for (int j = 0; j < M2; ++j) {
            const int k = j < M ? j : j - M2;
            if (j == M) { f[j] = 0.0L; continue; }
            const long double t = L * tanl((long double)k * pi / (2.0L * M));
            f[j] = expl(-t * t) * (L * L + t * t);
        }
        // a_m = Re DFT(f)_m / 2M, m = 1..N; the polynomial's coefficients from the highest power down
        for (int m = 1; m <= N; ++m) {
            long double re = 0.0L;
            for (int j = 0; j < M2; ++j) re += f[j] * cosl(2.0L * pi * (long double)((m * j) % M2) / M2);
            wc.c[N - m] = (double)(re / M2);
        }
        wc.L = (double)L;
    });
    return wc;
}

hipError_t launch_be_kick(const double *x, const double *y, int64_t n, double sigx, double sigy, double *kx, double *ky,
                          hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_be_kick, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, weideman_coeffs(), x, y, n, sigx,
                       sigy, kx, ky);
    return hipGetLastError();
}

hipError_t launch_faddeeva(const double *zr, const double *zi, int64_t n, double *wr, double *wi, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_faddeeva, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, weideman_coeffs(), zr, zi, n, wr,
                       wi);
    return hipGetLastError();
}

}  // namespace octopic
