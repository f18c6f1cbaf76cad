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

// ---- the kick with the diagonal of the potential's Hessian (_gaussian_beambeam_kick_response,
// strong_beam.jl:1061-1128).  K is normalised; h1, h2 carry the factor kbb.  Used by the Gaussian-subtracted
// solver's add-back, whose pz term needs d(potential)/d(sigma^2) = Hessian / 2 (heat equation), with the
// blend's chain term where the near-round series and the elliptic form are blended (:1095-1115).

// round beam Hessian (strong_beam.jl:711-772): f = 2(1 - e)/r^2, fp = df/d(r^2), series below u = 0.01
__device__ void round_hessian(double kbb, double sigma, double x, double y, double expterm, double &hxx, double &hyy) {
    const double r2 = x * x + y * y, sigma2 = sigma * sigma;
    if (r2 == 0.0) { hxx = hyy = -kbb / sigma2; return; }
    const double u = r2 / (2.0 * sigma2);
    double f, fp;
    if (u < 1.0e-2) {
        const double phi = 1.0 + u * (-1.0 / 2 + u * (1.0 / 6 + u * (-1.0 / 24 + u * (1.0 / 120 + u * (-1.0 / 720 + u * (1.0 / 5040))))));
        const double dphi =
            -1.0 / 2 + u * (1.0 / 3 + u * (-1.0 / 8 + u * (1.0 / 30 + u * (-1.0 / 144 + u * (1.0 / 840 - u * (1.0 / 5760))))));
        const double inv = 1.0 / sigma2;
        f = phi * inv;
        fp = dphi * inv * inv / 2.0;
    } else {
        const double ome = 1.0 - expterm;
        f = 2.0 * ome / r2;
        fp = 2.0 * (expterm * u - ome) / (r2 * r2);
    }
    hxx = -kbb * (f + 2.0 * x * x * fp);
    hyy = -kbb * (f + 2.0 * y * y * fp);
}

// near-round series with its Hessian diagonal (strong_beam.jl:897-937)
__device__ void near_round_response(double kbb, double v, double eta, double x, double y, double &kx, double &ky,
                                    double &h1, double &h2) {
    const double q = (x * x + y * y) / (2.0 * v), dr = (x * x - y * y) / (2.0 * v);
    double m[7];
    moments_0_6(q, m);
    const double e2 = eta * eta, e3 = e2 * eta;
    const double u1 = dr * m[1];
    const double u2 = m[1] / 2.0 - q * m[2] + (dr * dr / 2.0) * m[3];
    const double u3 = 1.5 * dr * m[3] - dr * q * m[4] + dr * dr * dr / 6.0 * m[5];
    const double v1 = dr * (q * m[2] - m[1]);
    const double v2 = 1.5 * q * m[2] - (q * q + dr * dr) * m[3] + (q * dr * dr / 2.0) * m[4];
    const double v3 = -1.5 * dr * m[3] + 3.5 * dr * q * m[4] - (dr * q * q + dr * dr * dr / 2.0) * m[5] + dr * dr * dr * q / 6.0 * m[6];
    const double c1x = dr * m[2] - m[1], c1y = dr * m[2] + m[1];
    const double c2x = 1.5 * m[2] + (-dr - q) * m[3] + (dr * dr / 2.0) * m[4];
    const double c2y = 1.5 * m[2] + (dr - q) * m[3] + (dr * dr / 2.0) * m[4];
    const double c3x = -1.5 * m[3] + (2.5 * dr + q) * m[4] + (-dr * q - dr * dr / 2.0) * m[5] + dr * dr * dr / 6.0 * m[6];
    const double c3y = 1.5 * m[3] + (2.5 * dr - q) * m[4] + (-dr * q + dr * dr / 2.0) * m[5] + dr * dr * dr / 6.0 * m[6];
    kx = x / v * (m[0] + eta * c1x + e2 * c2x + e3 * c3x);
    ky = y / v * (m[0] + eta * c1y + e2 * c2y + e3 * c3y);
    const double uv = -exp(-q) + eta * v1 + e2 * v2 + e3 * v3;
    const double ueta = u1 + 2.0 * eta * u2 + 3.0 * e2 * u3;
    h1 = kbb / v * (uv + (1.0 - eta) * ueta);
    h2 = kbb / v * (uv - (1.0 + eta) * ueta);
}

// U_exact - U_series through eta^6 (strong_beam.jl:939-960), from the radial moments m3..m11
__device__ double near_round_potential_residual(double v, double eta, double x, double y) {
    const double q = (x * x + y * y) / (2.0 * v), dr = (x * x - y * y) / (2.0 * v);
    double m[12];
    if (q <= 2.0) {
        for (int j = 3; j < 12; ++j) m[j] = 0.0;
        double term = 1.0;
        for (int k = 0; k <= 24; ++k) {
            for (int j = 3; j < 12; ++j) m[j] += term / (double)(k + 1 + j);
            term *= -q / (double)(k + 1);
        }
    } else {
        const double eq = exp(-q);
        double prev = -expm1(-q) / q;
        for (int j = 1; j < 12; ++j) {
            prev = ((double)j * prev - eq) / q;
            m[j] = prev;
        }
    }
    const double dr2 = dr * dr, q2 = q * q;
    const double u4 = 3.0 / 8.0 * m[3] - 1.5 * q * m[4] + (5.0 * dr2 + 2.0 * q2) / 4.0 * m[5] - dr2 * q / 2.0 * m[6] +
                      dr2 * dr2 / 24.0 * m[7];
    const double u5 = 15.0 / 8.0 * dr * m[5] - 2.5 * dr * q * m[6] + dr * (7.0 * dr2 + 6.0 * q2) / 12.0 * m[7] -
                      dr * dr2 * q / 6.0 * m[8] + dr * dr2 * dr2 / 120.0 * m[9];
    const double u6 = 5.0 / 16.0 * m[5] - 15.0 / 8.0 * q * m[6] + 5.0 * (7.0 * dr2 + 4.0 * q2) / 16.0 * m[7] -
                      q * (21.0 * dr2 + 2.0 * q2) / 12.0 * m[8] + dr2 * (3.0 * dr2 + 4.0 * q2) / 16.0 * m[9] -
                      dr2 * dr2 * q / 24.0 * m[10] + dr2 * dr2 * dr2 / 720.0 * m[11];
    return eta * eta * eta * eta * (u4 + eta * (u5 + eta * u6));
}

// d/dx of the axis series component (strong_beam.jl:974-1007), times -kbb
__device__ double axis_hessian(double kbb, double sx, double sy, double x, double y) {
    const double s = sx + sy, sx2 = sx * sx, sx3 = sx2 * sx, sx5 = sx3 * sx2, sy2 = sy * sy, sy3 = sy2 * sy;
    const double s2 = s * s, s3 = s2 * s;
    const double j00 = 2.0 / (sx * s);
    const double j10 = 2.0 * (2.0 * sx + sy) / (3.0 * sx3 * s2);
    const double j01 = 2.0 / (sx * sy * s2);
    const double j20 = 2.0 * (8.0 * sx2 + 9.0 * sx * sy + 3.0 * sy2) / (15.0 * sx5 * s3);
    const double j11 = 2.0 * (3.0 * sx + sy) / (3.0 * sx3 * sy * s3);
    const double j02 = 2.0 * (sx + 3.0 * sy) / (3.0 * sx * sy3 * s3);
    const double x2 = x * x, y2 = y * y;
    const double derivative = j00 - 1.5 * x2 * j10 - y2 * j01 / 2.0 + 0.625 * x2 * x2 * j20 + 0.75 * x2 * y2 * j11 + y2 * y2 * j02 / 8.0;
    return -kbb * derivative;
}

// sig1 >= sig2 (strong_beam.jl:1061-1117)
__device__ void principal_response(const WeidemanCoeffs &wc, double kbb, double s1, double s2, double x, double y,
                                   double &kx, double &ky, double &h1, double &h2) {
    const double v = (s1 * s1 + s2 * s2) / 2.0;
    const double eta = (s1 * s1 - s2 * s2) / (2.0 * v);
    if (eta == 0.0) {
        const double r2 = x * x + y * y, u = r2 / (2.0 * v);
        const double scale = (u == 0.0 ? 1.0 : -expm1(-u) / u) / v;
        kx = scale * x;
        ky = scale * y;
        round_hessian(kbb, sqrt(v), x, y, exp(-r2 / (2.0 * v)), h1, h2);
        return;
    }
    const double outer = near_round_outer(), inner = outer / 2.0;
    if (eta <= inner) {
        near_round_response(kbb, v, eta, x, y, kx, ky, h1, h2);
        return;
    }
    double ex, ey;
    elliptic_kick(wc, s1, s2, x, y, eta, ex, ey);
    const double expterm = exp(-0.5 * (x * x / (s1 * s1) + y * y / (s2 * s2)));
    double he1, he2;
    const double rho2 = x * x / (s1 * s1) + y * y / (s2 * s2);
    if (rho2 * rho2 * rho2 * sqrt(rho2) <= 0x1p-52 / sqrt(eta)) {   // the axis series' Hessian (:1009-1014)
        he1 = axis_hessian(kbb, s1, s2, x, y);
        he2 = axis_hessian(kbb, s2, s1, y, x);
    } else {   // _elliptic_gaussian_hessian_diagonal (:695-703)
        const double t1 = kbb * (x * ex + y * ey), t2 = s1 * s1 - s2 * s2, t3 = s2 / s1;
        he1 = (t1 - 2.0 * kbb * (1.0 - expterm * t3)) / t2;
        he2 = (-t1 + 2.0 * kbb * (1.0 - expterm / t3)) / t2;
    }
    if (eta >= outer) {
        kx = ex; ky = ey; h1 = he1; h2 = he2;
        return;
    }
    const double invw = 1.0 / (outer - inner), t = (eta - inner) * invw;
    const double w = t * t * t * (10.0 + t * (-15.0 + 6.0 * t));
    const double dw = 30.0 * t * t * (t - 1.0) * (t - 1.0) * invw;
    double sx, sy, hs1, hs2;
    near_round_response(kbb, v, eta, x, y, sx, sy, hs1, hs2);
    const double chain = kbb * dw * near_round_potential_residual(v, eta, x, y) / v;
    kx = sx + w * (ex - sx);
    ky = sy + w * (ey - sy);
    h1 = hs1 + w * (he1 - hs1) + (1.0 - eta) * chain;
    h2 = hs2 + w * (he2 - hs2) - (1.0 + eta) * chain;
}

__device__ void beambeam_response(const WeidemanCoeffs &wc, double kbb, double sigx, double sigy, double x, double y,
                                  double &kx, double &ky, double &hx, double &hy) {
    if (sigx >= sigy) principal_response(wc, kbb, sigx, sigy, x, y, kx, ky, hx, hy);
    else principal_response(wc, kbb, sigy, sigx, y, x, ky, kx, hy, hx);
}

__global__ __launch_bounds__(256) void k_be_response(WeidemanCoeffs wc, const double *X, const double *Y, int64_t n,
                                                     double kbb, double sigx, double sigy, double *OUT) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double kx, ky, hx, hy;
    beambeam_response(wc, kbb, sigx, sigy, X[i], Y[i], kx, ky, hx, hy);
    OUT[4 * i] = kx; OUT[4 * i + 1] = ky; OUT[4 * i + 2] = hx; OUT[4 * i + 3] = hy;
}

// The Gaussian-subtracted solver's add-back per field particle (_gpic_apply_kick_range!, gaussian_pic.jl:
// 818-892, uncoupled): at the particle's forward-drifted position (the kick stage recomputes the same
// drift), the exact field of the source's drifted Gaussians at the two boundaries, interpolated as the
// mesh kick.  Writes (dpx, dpy, pz covariance term, pz centroid term) for the kick stage to add.
__global__ __launch_bounds__(256) void k_gpic_addback(WeidemanCoeffs wc, GpicAddbackArgs a) {
    const int d = blockIdx.y;
    const Inter &it = a.inters[d];
    if (!it.valid || !it.gauss || *((volatile const int32_t *)&a.err->code) != 0) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= it.fld_n) return;
    const int fb = it.fld_beam;
    const int64_t k = it.fld_off + t;
    const double x0 = a.st.c[fb][0][k], px = a.st.c[fb][1][k], y0 = a.st.c[fb][2][k], py = a.st.c[fb][3][k];
    const double z = a.st.c[fb][4][k];
    const double s = 0.5 * (z - it.c_src);
    const double x = x0 + s * px, y = y0 + s * py;
    double hzi, zbias;
    {
        const double denom = it.rb - it.lb;
        if (!isfinite(denom) || denom == 0.0) { hzi = 0.0; zbias = 0.5; }
        else { hzi = 1.0 / denom; zbias = it.rb * hzi; }
    }
    const double zl = -z * hzi + zbias;
    const double zL = zl > 1.0 ? 1.0 : (zl < 0.0 ? 0.0 : zl), zR = 1.0 - zL;
    const double ks = it.kick_scale, half_ns = 0.5 * (double)it.src_n, kbb_eff = ks * half_ns;
    double kxL = 0, kyL = 0, hxL = 0, hyL = 0, kxR = 0, kyR = 0, hxR = 0, hyR = 0;
    if (a.lkick) {
        beambeam_response(wc, kbb_eff, it.g[0].sigx, it.g[0].sigy, x - it.g[0].mux, y - it.g[0].muy, kxL, kyL, hxL, hyL);
        beambeam_response(wc, kbb_eff, it.g[1].sigx, it.g[1].sigy, x - it.g[1].mux, y - it.g[1].muy, kxR, kyR, hxR, hyR);
    } else {
        if (it.g[0].sigx >= it.g[0].sigy) principal_kick(wc, it.g[0].sigx, it.g[0].sigy, x - it.g[0].mux, y - it.g[0].muy, kxL, kyL);
        else principal_kick(wc, it.g[0].sigy, it.g[0].sigx, y - it.g[0].muy, x - it.g[0].mux, kyL, kxL);
        if (it.g[1].sigx >= it.g[1].sigy) principal_kick(wc, it.g[1].sigx, it.g[1].sigy, x - it.g[1].mux, y - it.g[1].muy, kxR, kyR);
        else principal_kick(wc, it.g[1].sigy, it.g[1].sigx, y - it.g[1].muy, x - it.g[1].mux, kyR, kxR);
    }
    const double Kxa = half_ns * (zL * kxL + zR * kxR), Kya = half_ns * (zL * kyL + zR * kyR);
    const double dpxa = ks * Kxa, dpya = ks * Kya;
    const double covL = (hxL * it.g[0].rx + hyL * it.g[0].ry) / 4.0, covR = (hxR * it.g[1].rx + hyR * it.g[1].ry) / 4.0;
    double *o = a.ab + 4 * (a.base[fb] + k);
    o[0] = dpxa;
    o[1] = dpya;
    o[2] = zL * covL + zR * covR;
    o[3] = 0.5 * (dpxa * it.gmpx + dpya * it.gmpy);
}

// ---------------------------------------------------------------------- the soft-Gaussian solver
// GaussianPoissonSolver (gaussian.jl), uncoupled: the virtual drifts (strong_beam.jl:497-617; forward returns
// S = (z - zo)/2 of the drift's rule), and one particle's map between them (_apply_slice_kick_one!,
// gaussian.jl:152-188, with the uncoupled _cp_covariance_kick and _transport_transverse_moments,
// strong_beam.jl:619-632, 679-687).  Returns the luminosity density times 2 pi.
struct SoftMoments { double mx, mpx, my, mpy, a0, d0, bxx, bypy, qxx, qyy; };
constexpr double kTwoPi = 6.283185307179586;

__device__ __forceinline__ double soft_forward_drift(int d, double &x, double px, double &y, double py, double &z, double &pz,
                                                     double zo) {
    if (d == OCTOPIC_DRIFT_EXACT) {
        const double ps = sqrt((1 + pz) * (1 + pz) - px * px - py * py);
        const double H = 1 + pz - ps;
        const double rr = 0.5 * H / ps;
        const double z2 = (z + rr * zo) / (1 + rr);
        const double S = 0.5 * (z2 - zo);
        z -= H / ps * S;
        pz -= 0.5 * H;
        x += px / ps * S;
        y += py / ps * S;
        return S;
    }
    const double S = 0.5 * (z - zo);
    if (d == OCTOPIC_DRIFT_HIRATA) {
        x += S * px; y += S * py;
        pz -= 0.25 * (px * px + py * py);
    } else if (d == OCTOPIC_DRIFT_UNSAFE_PARAXIAL_FROZEN_LONGITUDINAL) {
        x += S * px; y += S * py;
    } else {   // chromatic, and its frozen-energy form
        const double phi = sqrt(1 - 0.5 * (px * px + py * py) / ((1 + pz) * (1 + pz))) - 1;
        x += S * px / (1 + pz);
        y += S * py / (1 + pz);
        z += 2 * S * phi;
        if (d == OCTOPIC_DRIFT_CHROMATIC) pz += (1 + pz) * phi;
    }
    return S;
}

__device__ __forceinline__ void soft_reverse_drift(int d, double &x, double px, double &y, double py, double &z, double &pz,
                                                   double zo) {
    const double S = 0.5 * (z - zo);
    if (d == OCTOPIC_DRIFT_EXACT) {
        const double H0 = 0.5 * (px * px + py * py) / (1 + pz);
        const double ps0 = 1 + pz - 0.5 * H0;
        pz += 0.5 * H0;
        x -= px / ps0 * S;
        y -= py / ps0 * S;
        z += H0 / ps0 * S;
    } else if (d == OCTOPIC_DRIFT_HIRATA) {
        x -= S * px; y -= S * py;
        pz += 0.25 * (px * px + py * py);
    } else if (d == OCTOPIC_DRIFT_UNSAFE_PARAXIAL_FROZEN_LONGITUDINAL) {
        x -= S * px; y -= S * py;
    } else {
        const double psi = sqrt(1 + 0.5 * (px * px + py * py) / ((1 + pz) * (1 + pz))) - 1;
        x -= S * px / (1 + pz);
        y -= S * py / (1 + pz);
        z += 2 * S * psi;
        if (d == OCTOPIC_DRIFT_CHROMATIC) pz += (1 + pz) * psi;
    }
}

__device__ __forceinline__ double soft_kick_one(const WeidemanCoeffs &wc, double &x, double &px, double &y, double &py,
                                                double &z, double &pz, const SoftMoments &m, double center, double kbb,
                                                int drift, int lkick) {
    const int d = lkick ? drift : OCTOPIC_DRIFT_UNSAFE_PARAXIAL_FROZEN_LONGITUDINAL;   // _soft_gaussian_drift
    const double S = soft_forward_drift(d, x, px, y, py, z, pz, center);
    const double xx = x - m.mx + m.mpx * S;
    const double yy = y - m.my + m.mpy * S;
    const double px0 = px, py0 = py, pz0 = pz;
    const double a = m.a0 - 2 * S * m.bxx + S * S * m.qxx;
    const double dd = m.d0 - 2 * S * m.bypy + S * S * m.qyy;
    const double au = 2 * (m.bxx - S * m.qxx);
    const double du = 2 * (m.bypy - S * m.qyy);
    double density = 0.0;
    if (!(a <= 0 || dd <= 0)) {
        const double sigx = sqrt(a), sigy = sqrt(dd);
        double kx, ky, hx, hy;
        beambeam_response(wc, kbb, sigx, sigy, xx, yy, kx, ky, hx, hy);
        const double expterm = exp(-0.5 * (xx * xx / (sigx * sigx) + yy * yy / (sigy * sigy)));
        px += kx * kbb;
        py += ky * kbb;
        pz += 0.25 * (hx * au + hy * du);
        density = expterm / kTwoPi / sigx / sigy;
    }
    if (lkick) pz += 0.5 * ((px - px0) * m.mpx + (py - py0) * m.mpy);
    else pz = pz0;
    soft_reverse_drift(d, x, px, y, py, z, pz, center);
    return density * kTwoPi;
}

__device__ __forceinline__ bool finite_all(const SoftMoments &m) {   // _gaussian_moments_finite (gaussian.jl:126-130)
    return isfinite(m.mx) && isfinite(m.mpx) && isfinite(m.my) && isfinite(m.mpy) && isfinite(sqrt(m.a0)) &&
           isfinite(sqrt(m.d0)) && isfinite(sqrt(m.qxx)) && isfinite(sqrt(m.qyy)) && isfinite(m.bxx) && isfinite(m.bypy);
}

// Block (b, d): interaction d = 2k + dir - 1 of the batch, pair k.  Its source slice (beam 1's slice i for
// dir 1, beam 2's slice j for dir 2) becomes its Gaussian from the moment partials (_slice_transverse_moments,
// slicing.jl:670-782: shifted sums folded over the chunks in order, the shifted second moments as fused
// multiply-adds, the min_sigma floor, ignore_centroid), and the block's share of the field slice's
// particles goes through the map in place.  The sampled direction's block sums of density x 2 pi (fixed
// tree) go to lum_block[p][b] for an ordered fold at the end of the collide.
__global__ __launch_bounds__(256) void k_soft_kick(WeidemanCoeffs wc, SoftKickArgs a) {
    __shared__ double sm[10];
    __shared__ SoftMoments mom;
    __shared__ int ok;
    __shared__ double red[256];
    const int d = blockIdx.y, k = d >> 1, side = d & 1;
    const BatchPair pr = a.pairs[k];
    const SliceTable *ts = side == 0 ? a.tab1 : a.tab2, *tf = side == 0 ? a.tab2 : a.tab1;
    const int si = side == 0 ? pr.i : pr.j, fi = side == 0 ? pr.j : pr.i;
    const int64_t ns = ts->count[si], nf = tf->count[fi], foff = tf->offset[fi];
    {   // the chunk partials folded in a fixed order: 25 strided runs per moment, then the runs in order
        constexpr int kRuns = 25;
        if (threadIdx.x < 10 * kRuns) {
            const int j = threadIdx.x % 10, r = threadIdx.x / 10;
            const double *part = a.mom_part + (int64_t)d * a.mom_chunks * 10;
            double acc = 0.0;
#pragma unroll 4
            for (int c = r; c < a.mom_chunks; c += kRuns) acc += part[c * 10 + j];
            red[threadIdx.x] = acc;
        }
        __syncthreads();
        if (threadIdx.x < 10) {
            double acc = 0.0;
            for (int r = 0; r < kRuns; ++r) acc += red[r * 10 + threadIdx.x];
            sm[threadIdx.x] = acc;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        SoftMoments m;
        const double floor2 = a.min_sigma * a.min_sigma;
        if (ns == 0) {
            m = SoftMoments{0, 0, 0, 0, floor2, floor2, 0, 0, 0, 0};
        } else {
            const double *sh = a.shift + (int64_t)d * 4;
            const double x0 = sh[0], px0 = sh[1], y0 = sh[2], py0 = sh[3];
            const double invn = 1.0 / (double)ns;
            const double dmx = sm[0] * invn, dmpx = sm[1] * invn, dmy = sm[2] * invn, dmpy = sm[3] * invn;
            m.mx = x0 + dmx; m.mpx = px0 + dmpx; m.my = y0 + dmy; m.mpy = py0 + dmpy;
            m.a0 = fmax(fma(-dmx, dmx, sm[4] * invn), floor2);
            m.d0 = fmax(fma(-dmy, dmy, sm[6] * invn), floor2);
            m.qxx = fmax(fma(-dmpx, dmpx, sm[5] * invn), 0.0);
            m.qyy = fmax(fma(-dmpy, dmpy, sm[7] * invn), 0.0);
            m.bxx = fma(-dmx, dmpx, sm[8] * invn);
            m.bypy = fma(-dmy, dmpy, sm[9] * invn);
            if (side == 0 ? a.ignore1 : a.ignore2) { m.mx = 0.0; m.mpx = 0.0; m.my = 0.0; m.mpy = 0.0; }
        }
        mom = m;
        ok = finite_all(m) ? 1 : 0;
        if (!ok && atomicCAS(&a.err->code, 0, OCTOPIC_ERR_NONFINITE) == 0) {
            a.err->p = pr.p; a.err->dir = side + 1; a.err->stage = 6; a.err->a = 0.0; a.err->b = 0.0;
        }
    }
    __syncthreads();
    // dir 1 (side 0) kicks beam 2 with beam 1's slice: luminosity lum1, sampled when gaussian_when_luminosity = 1
    const bool want = a.compute_lum && (side == 0 ? a.sample1 : !a.sample1);
    const double w = ts->weight[si];
    const double kbb = w * (side == 0 ? a.kbb2 : a.kbb1);
    double acc = 0.0;
    if (ok && *((volatile const int32_t *)&a.err->code) == 0) {
        const SoftMoments m = mom;
        const double center = ts->center[si];
        const int fb = side == 0 ? 1 : 0;
        double *X = a.st.c[fb][0] + foff, *PX = a.st.c[fb][1] + foff, *Y = a.st.c[fb][2] + foff,
               *PY = a.st.c[fb][3] + foff, *Z = a.st.c[fb][4] + foff, *PZ = a.st.c[fb][5] + foff;
        for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nf; t += (int64_t)gridDim.x * blockDim.x) {
            double x = X[t], px = PX[t], y = Y[t], py = PY[t], z = Z[t], pz = PZ[t];
            const double dens = soft_kick_one(wc, x, px, y, py, z, pz, m, center, kbb, a.drift, a.lkick);
            if (want) acc += dens;
            X[t] = x; PX[t] = px; Y[t] = y; PY[t] = py; Z[t] = z; PZ[t] = pz;
        }
    }
    if (!want) return;
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.lum_block[(int64_t)pr.p * kSoftBlocks + blockIdx.x] = red[0];
        if (blockIdx.x == 0) a.lum_scale[pr.p] = w * (side == 0 ? a.klum2 : a.klum1);
    }
}

// lum_parts[p] = (sum of the pair's block partials, in block order) / 2 pi * klum_slice
// (_slice_slice_gaussian_kick!, gaussian.jl:97-123); pairs whose kicks did not run read zeros.
__global__ __launch_bounds__(256) void k_soft_lum_fold(const double *lum_block, const double *lum_scale, int npairs,
                                                       double *lum_parts) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    double s = 0.0;
    for (int b = 0; b < kSoftBlocks; ++b) s += lum_block[(int64_t)p * kSoftBlocks + b];
    lum_parts[p] = s / kTwoPi * lum_scale[p];
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

hipError_t launch_be_response(const double *x, const double *y, int64_t n, double kbb, double sigx, double sigy,
                              double *out4, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_be_response, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, weideman_coeffs(), x, y, n, kbb,
                       sigx, sigy, out4);
    return hipGetLastError();
}

hipError_t launch_gpic_addback(const GpicAddbackArgs &a, int64_t max_fld, hipStream_t s) {
    if (a.ninters <= 0 || max_fld <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gpic_addback, dim3((unsigned)((max_fld + 255) / 256), a.ninters), dim3(256), 0, s,
                       weideman_coeffs(), a);
    return hipGetLastError();
}

hipError_t launch_soft_kick(const SoftKickArgs &a, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_soft_kick, dim3(kSoftBlocks, 2 * a.npairs), dim3(256), 0, s, weideman_coeffs(), a);
    return hipGetLastError();
}

hipError_t launch_soft_lum_fold(const double *lum_block, const double *lum_scale, int npairs, double *lum_parts,
                                hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_soft_lum_fold, dim3((npairs + 255) / 256), dim3(256), 0, s, lum_block, lum_scale, npairs, lum_parts);
    return hipGetLastError();
}

hipError_t launch_faddeeva(const double *zr, const double *zi, int64_t n, double *wr, double *wi, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_faddeeva, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, weideman_coeffs(), zr, zi, n, wr,
                       wi);
    return hipGetLastError();
}

}  // namespace octopic
