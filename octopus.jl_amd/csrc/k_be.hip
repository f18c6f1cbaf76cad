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

// The evaluation is one source for both working types, FP64 and FP32, with the per-type constants of the
// reference (strong_beam.jl:770-783, 809-811): the near-round blend's conditioning factor C (64, 8) and
// rounding unit eps (2^-52, 2^-23), which set eta* = (8/3 C eps)^(1/4) and the axis-series switch
// rho^7 <= eps / sqrt(eta), and the last order of the radial-moment series (24, 16).
template <typename T> struct BeType;
template <> struct BeType<double> {
    static constexpr double eps = 0x1p-52, cond = 64.0;
    static constexpr int last_order = 24;
    using Coeffs = WeidemanCoeffs;
};
template <> struct BeType<float> {
    static constexpr float eps = 0x1p-23f, cond = 8.0f;
    static constexpr int last_order = 16;
    using Coeffs = WeidemanCoeffsF;
};
template <typename T> using Coeffs = typename BeType<T>::Coeffs;

constexpr double kInvSqrtPi = 0.56418958354775628695;
constexpr double kSqrtPi = 1.7724538509055160273;
constexpr double kSqrt2 = 1.4142135623730950488;

template <typename T>
__device__ __forceinline__ void cmul(T ar, T ai, T br, T bi, T &r, T &i) {
    r = ar * br - ai * bi;
    i = ar * bi + ai * br;
}

template <typename T>
__device__ __forceinline__ void cdiv(T ar, T ai, T br, T bi, T &r, T &i) {
    const T inv = T(1) / (br * br + bi * bi);
    r = (ar * br + ai * bi) * inv;
    i = (ai * br - ar * bi) * inv;
}

// Laplace continued fraction of w for large |z| (Im z >= 0), with Gautschi's term-count estimate
// nu = floor(3.9 + 11.398 / (0.08254 x + 0.1421 y + 0.2023)) (SpecialMath.jl:113-151)
template <typename T>
__device__ void w_far(T zr, T zi, T &wr, T &wi) {
    const T isp = T(kInvSqrtPi);
    const T x = fabs(zr), y = zi;
    if (x + y > T(4000)) {
        if (x + y > T(1e7)) {   // w ~ i / (sqrt(pi) z)
            if (x > y) {
                const T q = y / zr, d = isp / (zr + q * y);
                wr = d * q; wi = d;
            } else {
                const T q = zr / y, d = isp / (q * zr + y);
                wr = d; wi = d * q;
            }
            return;
        }
        // w ~ i z / (sqrt(pi) (z^2 - 1/2))
        const T dr = zr * zr - y * y - T(0.5), di = T(2) * zr * y;
        const T d = isp / (dr * dr + di * di);
        wr = d * (zr * di - y * dr);
        wi = d * (zr * dr + y * di);
        return;
    }
    T nu = floor(T(3.9) + T(11.398) / (T(0.08254) * x + T(0.1421) * y + T(0.2023)));
    T tr = zr, ti = y;
    nu = T(0.5) * (nu - T(1));
    while (nu > T(0.4)) {   // t <- z - nu / t, nu = (n-1)/2, ..., 1/2
        const T s = nu / (tr * tr + ti * ti);
        const T nr = zr - tr * s, ni = y + ti * s;
        tr = nr; ti = ni;
        nu -= T(0.5);
    }
    const T s = isp / (tr * tr + ti * ti);   // i / (sqrt(pi) t)
    wr = s * ti;
    wi = s * tr;
}

template <typename T>
__device__ void w_upper(const Coeffs<T> &wc, T zr, T zi, T &wr, T &wi) {
    const T x = fabs(zr), y = zi;
    if (y > T(7) || (x > T(6) && (y > T(0.1) || (x > T(8) && y > T(1e-10)) || x > T(28)))) {
        w_far(zr, y, wr, wi);
        return;
    }
    // Z = (L + iz) / (L - iz); w = 2 p(Z) / (L - iz)^2 + 1 / (sqrt(pi) (L - iz))
    const T L = wc.L;
    const T dr = L + zi, di = -zr;
    T Zr, Zi;
    cdiv(L - zi, zr, dr, di, Zr, Zi);
    T pr = T(0), pi_ = T(0);
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        T r, i;
        cmul(pr, pi_, Zr, Zi, r, i);
        pr = r + wc.c[k];
        pi_ = i;
    }
    T d2r, d2i, tr, ti, ir, ii;
    cmul(dr, di, dr, di, d2r, d2i);
    cdiv(T(2) * pr, T(2) * pi_, d2r, d2i, tr, ti);
    cdiv(T(kInvSqrtPi), T(0), dr, di, ir, ii);
    wr = tr + ir;
    wi = ti + ii;
}

template <typename T>
__device__ void w_any(const Coeffs<T> &wc, T zr, T zi, T &wr, T &wi) {
    if (zi < T(0)) {   // w(z) = 2 exp(-z^2) - w(-z)
        T ur, ui;
        w_upper(wc, -zr, -zi, ur, ui);
        const T e = exp(-zr * zr + zi * zi), a = T(-2) * zr * zi;
        wr = T(2) * (e * cos(a)) - ur;
        wi = T(2) * (e * sin(a)) - ui;
        return;
    }
    w_upper(wc, zr, zi, wr, wi);
}

// eta* of the near-round blend: (8/3 C eps)^(1/4) in the working type (strong_beam.jl:779-783)
template <typename T>
__device__ __forceinline__ T near_round_outer() {
    return sqrt(sqrt((T(8) / T(3)) * BeType<T>::cond * BeType<T>::eps));
}

// radial moments m_k(q) = int_0^1 s^k exp(-q s) ds, k = 0..6: the series through the type's last order for
// q <= 2, the upward recurrence m_k = (k m_{k-1} - e^-q) / q above (strong_beam.jl:799-833)
template <typename T>
__device__ void moments_0_6(T q, T m[7]) {
    if (q <= T(2)) {
        for (int j = 0; j < 7; ++j) m[j] = T(0);
        T term = T(1);
        for (int k = 0; k <= BeType<T>::last_order; ++k) {
            for (int j = 0; j < 7; ++j) m[j] += term / (T)(k + 1 + j);
            term *= -q / (T)(k + 1);
        }
        return;
    }
    const T eq = exp(-q);
    m[0] = -expm1(-q) / q;
    for (int j = 1; j < 7; ++j) m[j] = ((T)j * m[j - 1] - eq) / q;
}

// cubic series in eta about the round beam (strong_beam.jl:878-895)
template <typename T>
__device__ void near_round_kick(T v, T eta, T x, T y, T &kx, T &ky) {
    const T q = (x * x + y * y) / (T(2) * v), dr = (x * x - y * y) / (T(2) * v);
    T m[7];
    moments_0_6(q, m);
    const T e2 = eta * eta, e3 = e2 * eta;
    const T c1x = dr * m[2] - m[1], c1y = dr * m[2] + m[1];
    const T c2x = T(1.5) * m[2] + (-dr - q) * m[3] + (dr * dr / T(2)) * m[4];
    const T c2y = T(1.5) * m[2] + (dr - q) * m[3] + (dr * dr / T(2)) * m[4];
    const T c3x = T(-1.5) * m[3] + (T(2.5) * dr + q) * m[4] + (-dr * q - dr * dr / T(2)) * m[5] + dr * dr * dr / T(6) * m[6];
    const T c3y = T(1.5) * m[3] + (T(2.5) * dr - q) * m[4] + (-dr * q + dr * dr / T(2)) * m[5] + dr * dr * dr / T(6) * m[6];
    kx = x / v * (m[0] + eta * c1x + e2 * c2x + e3 * c3x);
    ky = y / v * (m[0] + eta * c1y + e2 * c2y + e3 * c3y);
}

// the fifth-degree axis series' coefficients (strong_beam.jl:974-1007)
template <typename T>
struct AxisSeries {
    T j00, j10, j01, j20, j11, j02;
    __device__ AxisSeries(T sx, T sy) {
        const T s = sx + sy, sx2 = sx * sx, sx3 = sx2 * sx, sx5 = sx3 * sx2, sy2 = sy * sy, sy3 = sy2 * sy;
        const T s2 = s * s, s3 = s2 * s;
        j00 = T(2) / (sx * s);
        j10 = T(2) * (T(2) * sx + sy) / (T(3) * sx3 * s2);
        j01 = T(2) / (sx * sy * s2);
        j20 = T(2) * (T(8) * sx2 + T(9) * sx * sy + T(3) * sy2) / (T(15) * sx5 * s3);
        j11 = T(2) * (T(3) * sx + sy) / (T(3) * sx3 * sy * s3);
        j02 = T(2) * (sx + T(3) * sy) / (T(3) * sx * sy3 * s3);
    }
};

// one component of the axis series: x times the expansion of Kx / x in x^2 and y^2
template <typename T>
__device__ T axis_component(T sx, T sy, T x, T y) {
    const AxisSeries<T> j(sx, sy);
    const T x2 = x * x, y2 = y * y;
    const T scale = j.j00 - (x2 * j.j10 + y2 * j.j01) / T(2) + x2 * x2 * j.j20 / T(8) + x2 * y2 * j.j11 / T(4) +
                    y2 * y2 * j.j02 / T(8);
    return x * scale;
}

// the axis series is used where rho^7 <= eps / sqrt(eta) (strong_beam.jl:962-972)
template <typename T>
__device__ __forceinline__ bool near_axis(T s1, T s2, T x, T y, T eta) {
    const T rho2 = x * x / (s1 * s1) + y * y / (s2 * s2);
    return rho2 * rho2 * rho2 * sqrt(rho2) <= BeType<T>::eps / sqrt(eta);
}

// the elliptic Bassetti-Erskine form on the principal axes (strong_beam.jl:1016-1038)
template <typename T>
__device__ void elliptic_kick(const Coeffs<T> &wc, T s1, T s2, T x, T y, T eta, T &kx, T &ky) {
    if (near_axis(s1, s2, x, y, eta)) {
        kx = axis_component(s1, s2, x, y);
        ky = axis_component(s2, s1, y, x);
        return;
    }
    const bool negx = x < T(0), negy = y < T(0);
    x = fabs(x);
    y = fabs(y);
    const T den = T(kSqrt2) * sqrt(s1 * s1 - s2 * s2);
    T w1r, w1i, w2r, w2i;
    w_upper(wc, x / den, y / den, w1r, w1i);
    w_upper(wc, s2 / s1 * x / den, s1 / s2 * y / den, w2r, w2i);
    const T A = T(2) * T(kSqrtPi) / den;
    const T B = exp(-x * x / (T(2) * s1 * s1) - y * y / (T(2) * s2 * s2));
    const T rr = A * (w1r - B * w2r), ri = A * (w1i - B * w2i);
    kx = negx ? -ri : ri;
    ky = negy ? -rr : rr;
}

// sig1 >= sig2 (strong_beam.jl:1040-1059)
template <typename T>
__device__ void principal_kick(const Coeffs<T> &wc, T s1, T s2, T x, T y, T &kx, T &ky) {
    const T v = (s1 * s1 + s2 * s2) / T(2);
    const T eta = (s1 * s1 - s2 * s2) / (T(2) * v);
    if (eta == T(0)) {
        const T r2 = x * x + y * y, u = r2 / (T(2) * v);
        const T phi = u == T(0) ? T(1) : -expm1(-u) / u;
        kx = phi / v * x;
        ky = phi / v * y;
        return;
    }
    const T outer = near_round_outer<T>(), inner = outer / T(2);
    if (eta <= inner) {
        near_round_kick(v, eta, x, y, kx, ky);
        return;
    }
    T ex, ey;
    elliptic_kick(wc, s1, s2, x, y, eta, ex, ey);
    if (eta >= outer) {
        kx = ex; ky = ey;
        return;
    }
    const T t = (eta - inner) * (T(1) / (outer - inner));
    const T w = t * t * t * (T(10) + t * (T(-15) + T(6) * t));
    T sx, sy;
    near_round_kick(v, eta, x, y, sx, sy);
    kx = sx + w * (ex - sx);
    ky = sy + w * (ey - sy);
}

template <typename T>
__device__ void beambeam_kick(const Coeffs<T> &wc, T sigx, T sigy, T x, T y, T &kx, T &ky) {
    if (sigx >= sigy) principal_kick(wc, sigx, sigy, x, y, kx, ky);
    else principal_kick(wc, sigy, sigx, y, x, ky, kx);
}

// ---- the kick with the diagonal of the potential's Hessian (_gaussian_beambeam_kick_response,
// strong_beam.jl:1061-1128).  K is normalised; h1, h2 carry the factor kbb.  The Gaussian-subtracted solver's
// add-back and the soft-Gaussian map take their pz terms from it: d(potential)/d(sigma^2) = Hessian / 2 (heat
// equation), with the blend's chain term where the near-round series and the elliptic form are blended
// (:1095-1115).

// round beam Hessian (strong_beam.jl:711-772): f = 2(1 - e)/r^2, fp = df/d(r^2), series below u = 0.01
template <typename T>
__device__ void round_hessian(T kbb, T sigma, T x, T y, T expterm, T &hxx, T &hyy) {
    const T r2 = x * x + y * y, sigma2 = sigma * sigma;
    if (r2 == T(0)) { hxx = hyy = -kbb / sigma2; return; }
    const T u = r2 / (T(2) * sigma2);
    T f, fp;
    if (u < T(1.0e-2)) {
        const T phi = T(1) + u * (T(-1) / T(2) + u * (T(1) / T(6) + u * (T(-1) / T(24) + u * (T(1) / T(120) +
                      u * (T(-1) / T(720) + u * (T(1) / T(5040)))))));
        const T dphi = T(-1) / T(2) + u * (T(1) / T(3) + u * (T(-1) / T(8) + u * (T(1) / T(30) + u * (T(-1) / T(144) +
                       u * (T(1) / T(840) - u * (T(1) / T(5760)))))));
        const T inv = T(1) / sigma2;
        f = phi * inv;
        fp = dphi * inv * inv / T(2);
    } else {
        const T ome = T(1) - expterm;
        f = T(2) * ome / r2;
        fp = T(2) * (expterm * u - ome) / (r2 * r2);
    }
    hxx = -kbb * (f + T(2) * x * x * fp);
    hyy = -kbb * (f + T(2) * y * y * fp);
}

// near-round series with its Hessian diagonal (strong_beam.jl:897-937)
template <typename T>
__device__ void near_round_response(T kbb, T v, T eta, T x, T y, T &kx, T &ky, T &h1, T &h2) {
    const T q = (x * x + y * y) / (T(2) * v), dr = (x * x - y * y) / (T(2) * v);
    T m[7];
    moments_0_6(q, m);
    const T e2 = eta * eta, e3 = e2 * eta;
    const T u1 = dr * m[1];
    const T u2 = m[1] / T(2) - q * m[2] + (dr * dr / T(2)) * m[3];
    const T u3 = T(1.5) * dr * m[3] - dr * q * m[4] + dr * dr * dr / T(6) * m[5];
    const T v1 = dr * (q * m[2] - m[1]);
    const T v2 = T(1.5) * q * m[2] - (q * q + dr * dr) * m[3] + (q * dr * dr / T(2)) * m[4];
    const T v3 = T(-1.5) * dr * m[3] + T(3.5) * dr * q * m[4] - (dr * q * q + dr * dr * dr / T(2)) * m[5] +
                 dr * dr * dr * q / T(6) * m[6];
    const T c1x = dr * m[2] - m[1], c1y = dr * m[2] + m[1];
    const T c2x = T(1.5) * m[2] + (-dr - q) * m[3] + (dr * dr / T(2)) * m[4];
    const T c2y = T(1.5) * m[2] + (dr - q) * m[3] + (dr * dr / T(2)) * m[4];
    const T c3x = T(-1.5) * m[3] + (T(2.5) * dr + q) * m[4] + (-dr * q - dr * dr / T(2)) * m[5] + dr * dr * dr / T(6) * m[6];
    const T c3y = T(1.5) * m[3] + (T(2.5) * dr - q) * m[4] + (-dr * q + dr * dr / T(2)) * m[5] + dr * dr * dr / T(6) * m[6];
    kx = x / v * (m[0] + eta * c1x + e2 * c2x + e3 * c3x);
    ky = y / v * (m[0] + eta * c1y + e2 * c2y + e3 * c3y);
    const T uv = -exp(-q) + eta * v1 + e2 * v2 + e3 * v3;
    const T ueta = u1 + T(2) * eta * u2 + T(3) * e2 * u3;
    h1 = kbb / v * (uv + (T(1) - eta) * ueta);
    h2 = kbb / v * (uv - (T(1) + eta) * ueta);
}

// U_exact - U_series through eta^6 (strong_beam.jl:939-960), from the radial moments m3..m11
template <typename T>
__device__ T near_round_potential_residual(T v, T eta, T x, T y) {
    const T q = (x * x + y * y) / (T(2) * v), dr = (x * x - y * y) / (T(2) * v);
    T m[12];
    if (q <= T(2)) {
        for (int j = 3; j < 12; ++j) m[j] = T(0);
        T term = T(1);
        for (int k = 0; k <= BeType<T>::last_order; ++k) {
            for (int j = 3; j < 12; ++j) m[j] += term / (T)(k + 1 + j);
            term *= -q / (T)(k + 1);
        }
    } else {
        const T eq = exp(-q);
        T prev = -expm1(-q) / q;
        for (int j = 1; j < 12; ++j) {
            prev = ((T)j * prev - eq) / q;
            m[j] = prev;
        }
    }
    const T dr2 = dr * dr, q2 = q * q;
    const T u4 = T(3) / T(8) * m[3] - T(1.5) * q * m[4] + (T(5) * dr2 + T(2) * q2) / T(4) * m[5] - dr2 * q / T(2) * m[6] +
                 dr2 * dr2 / T(24) * m[7];
    const T u5 = T(15) / T(8) * dr * m[5] - T(2.5) * dr * q * m[6] + dr * (T(7) * dr2 + T(6) * q2) / T(12) * m[7] -
                 dr * dr2 * q / T(6) * m[8] + dr * dr2 * dr2 / T(120) * m[9];
    const T u6 = T(5) / T(16) * m[5] - T(15) / T(8) * q * m[6] + T(5) * (T(7) * dr2 + T(4) * q2) / T(16) * m[7] -
                 q * (T(21) * dr2 + T(2) * q2) / T(12) * m[8] + dr2 * (T(3) * dr2 + T(4) * q2) / T(16) * m[9] -
                 dr2 * dr2 * q / T(24) * m[10] + dr2 * dr2 * dr2 / T(720) * m[11];
    return eta * eta * eta * eta * (u4 + eta * (u5 + eta * u6));
}

// d/dx of the axis series component (strong_beam.jl:974-1007), times -kbb
template <typename T>
__device__ T axis_hessian(T kbb, T sx, T sy, T x, T y) {
    const AxisSeries<T> j(sx, sy);
    const T x2 = x * x, y2 = y * y;
    const T derivative = j.j00 - T(1.5) * x2 * j.j10 - y2 * j.j01 / T(2) + T(0.625) * x2 * x2 * j.j20 +
                         T(0.75) * x2 * y2 * j.j11 + y2 * y2 * j.j02 / T(8);
    return -kbb * derivative;
}

// sig1 >= sig2 (strong_beam.jl:1061-1117)
template <typename T>
__device__ void principal_response(const Coeffs<T> &wc, T kbb, T s1, T s2, T x, T y, T &kx, T &ky, T &h1, T &h2) {
    const T v = (s1 * s1 + s2 * s2) / T(2);
    const T eta = (s1 * s1 - s2 * s2) / (T(2) * v);
    if (eta == T(0)) {
        const T r2 = x * x + y * y, u = r2 / (T(2) * v);
        const T scale = (u == T(0) ? T(1) : -expm1(-u) / u) / v;
        kx = scale * x;
        ky = scale * y;
        round_hessian(kbb, sqrt(v), x, y, exp(-r2 / (T(2) * v)), h1, h2);
        return;
    }
    const T outer = near_round_outer<T>(), inner = outer / T(2);
    if (eta <= inner) {
        near_round_response(kbb, v, eta, x, y, kx, ky, h1, h2);
        return;
    }
    T ex, ey;
    elliptic_kick(wc, s1, s2, x, y, eta, ex, ey);
    const T expterm = exp(T(-0.5) * (x * x / (s1 * s1) + y * y / (s2 * s2)));
    T he1, he2;
    if (near_axis(s1, s2, x, y, eta)) {   // the axis series' Hessian (:1009-1014)
        he1 = axis_hessian(kbb, s1, s2, x, y);
        he2 = axis_hessian(kbb, s2, s1, y, x);
    } else {   // _elliptic_gaussian_hessian_diagonal (:695-703)
        const T t1 = kbb * (x * ex + y * ey), t2 = s1 * s1 - s2 * s2, t3 = s2 / s1;
        he1 = (t1 - T(2) * kbb * (T(1) - expterm * t3)) / t2;
        he2 = (-t1 + T(2) * kbb * (T(1) - expterm / t3)) / t2;
    }
    if (eta >= outer) {
        kx = ex; ky = ey; h1 = he1; h2 = he2;
        return;
    }
    const T invw = T(1) / (outer - inner), t = (eta - inner) * invw;
    const T w = t * t * t * (T(10) + t * (T(-15) + T(6) * t));
    const T dw = T(30) * t * t * (t - T(1)) * (t - T(1)) * invw;
    T sx, sy, hs1, hs2;
    near_round_response(kbb, v, eta, x, y, sx, sy, hs1, hs2);
    const T chain = kbb * dw * near_round_potential_residual(v, eta, x, y) / v;
    kx = sx + w * (ex - sx);
    ky = sy + w * (ey - sy);
    h1 = hs1 + w * (he1 - hs1) + (T(1) - eta) * chain;
    h2 = hs2 + w * (he2 - hs2) - (T(1) + eta) * chain;
}

template <typename T>
__device__ void beambeam_response(const Coeffs<T> &wc, T kbb, T sigx, T sigy, T x, T y, T &kx, T &ky, T &hx, T &hy) {
    if (sigx >= sigy) principal_response(wc, kbb, sigx, sigy, x, y, kx, ky, hx, hy);
    else principal_response(wc, kbb, sigy, sigx, y, x, ky, kx, hy, hx);
}

// FP32 evaluation on the scale of the larger rms size l: the kick is homogeneous of degree -1 and the
// Hessian of degree -2 in (sigx, sigy, x, y), so K = K(sigx/l, sigy/l, x/l, y/l) / l and H likewise / l^2.
// The series coefficients of the unscaled form reach sig^8 (about 1e-40 for sizes of 1e-5 m), beyond the
// FP32 normal range; scaled, every intermediate is of order one.  (The FP64 path evaluates unscaled.)
__device__ __forceinline__ void scaled_kick_f32(const WeidemanCoeffsF &wc, float sigx, float sigy, float x, float y,
                                                float &kx, float &ky) {
    const float l = fmaxf(sigx, sigy), il = 1.0f / l;
    beambeam_kick(wc, sigx * il, sigy * il, x * il, y * il, kx, ky);
    kx *= il;
    ky *= il;
}

__device__ __forceinline__ void scaled_response_f32(const WeidemanCoeffsF &wc, float kbb, float sigx, float sigy,
                                                    float x, float y, float &kx, float &ky, float &hx, float &hy) {
    const float l = fmaxf(sigx, sigy), il = 1.0f / l;
    beambeam_response(wc, kbb, sigx * il, sigy * il, x * il, y * il, kx, ky, hx, hy);
    kx *= il;
    ky *= il;
    hx *= il * il;
    hy *= il * il;
}

template <typename T>
__global__ __launch_bounds__(256) void k_be_response(Coeffs<T> wc, const T *X, const T *Y, int64_t n, T kbb, T sigx,
                                                     T sigy, T *OUT) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T kx, ky, hx, hy;
    if constexpr (sizeof(T) == 8) beambeam_response(wc, kbb, sigx, sigy, X[i], Y[i], kx, ky, hx, hy);
    else scaled_response_f32(wc, kbb, sigx, sigy, X[i], Y[i], kx, ky, hx, hy);
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
        beambeam_kick(wc, it.g[0].sigx, it.g[0].sigy, x - it.g[0].mux, y - it.g[0].muy, kxL, kyL);
        beambeam_kick(wc, it.g[1].sigx, it.g[1].sigy, x - it.g[1].mux, y - it.g[1].muy, kxR, kyR);
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

// TB: the working type of the Bassetti-Erskine evaluation (the field, its Hessian and the density); the drifts,
// the moment transport and the coordinate updates are FP64 either way.
template <typename TB>
__device__ __forceinline__ double soft_kick_one(const Coeffs<TB> &wc, double &x, double &px, double &y, double &py,
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
        double kx, ky, hx, hy, expterm;
        if constexpr (sizeof(TB) == 8) {
            beambeam_response(wc, kbb, sigx, sigy, xx, yy, kx, ky, hx, hy);
            expterm = exp(-0.5 * (xx * xx / (sigx * sigx) + yy * yy / (sigy * sigy)));
        } else {   // scaled by the larger size in FP64 before the rounding to FP32 (see scaled_response_f32)
            const double l = fmax(sigx, sigy), il = 1.0 / l;
            const float fx = (float)(xx * il), fy = (float)(yy * il), fsx = (float)(sigx * il), fsy = (float)(sigy * il);
            float k1, k2, h1, h2;
            beambeam_response(wc, (float)kbb, fsx, fsy, fx, fy, k1, k2, h1, h2);
            kx = (double)k1 * il;
            ky = (double)k2 * il;
            hx = (double)h1 * (il * il);
            hy = (double)h2 * (il * il);
            expterm = (double)exp(-0.5f * (fx * fx / (fsx * fsx) + fy * fy / (fsy * fsy)));
        }
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
template <typename TB>
__global__ __launch_bounds__(256) void k_soft_kick(Coeffs<TB> wc, SoftKickArgs a) {
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
            const double dens = soft_kick_one<TB>(wc, x, px, y, py, z, pz, m, center, kbb, a.drift, a.lkick);
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

template <typename T>
__global__ __launch_bounds__(256) void k_be_kick(Coeffs<T> wc, const T *X, const T *Y, int64_t n, T sigx, T sigy, T *KX,
                                                 T *KY) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T x = X[i], y = Y[i];
    T kx = T(0), ky = T(0);
    if (sigx != T(0) && sigy != T(0)) {
        if constexpr (sizeof(T) == 8) beambeam_kick(wc, sigx, sigy, x, y, kx, ky);
        else scaled_kick_f32(wc, sigx, sigy, x, y, kx, ky);
    }
    KX[i] = kx;
    KY[i] = ky;
}

template <typename T>
__global__ __launch_bounds__(256) void k_faddeeva(Coeffs<T> wc, const T *ZR, const T *ZI, int64_t n, T *WR, T *WI) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T wr, wi;
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

const WeidemanCoeffsF &weideman_coeffs_f32() {
    static WeidemanCoeffsF wf;
    static std::once_flag once;
    std::call_once(once, [] {
        const WeidemanCoeffs &wc = weideman_coeffs();
        for (int k = 0; k < 32; ++k) wf.c[k] = (float)wc.c[k];   // T(c) of the FP64 constants
        wf.L = (float)wc.L;
    });
    return wf;
}

namespace {
template <typename T> const Coeffs<T> &coeffs();
template <> const WeidemanCoeffs &coeffs<double>() { return weideman_coeffs(); }
template <> const WeidemanCoeffsF &coeffs<float>() { return weideman_coeffs_f32(); }
}  // namespace

template <typename T>
hipError_t launch_be_kick(const T *x, const T *y, int64_t n, T sigx, T sigy, T *kx, T *ky, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_be_kick<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coeffs<T>(), x, y, n, sigx, sigy,
                       kx, ky);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_be_response(const T *x, const T *y, int64_t n, T kbb, T sigx, T sigy, T *out4, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_be_response<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coeffs<T>(), x, y, n, kbb,
                       sigx, sigy, out4);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_faddeeva(const T *zr, const T *zi, int64_t n, T *wr, T *wi, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_faddeeva<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coeffs<T>(), zr, zi, n, wr, wi);
    return hipGetLastError();
}

#define OCTOPIC_BE_INSTANTIATE(T)                                                                                     \
    template hipError_t launch_be_kick<T>(const T *, const T *, int64_t, T, T, T *, T *, hipStream_t);                \
    template hipError_t launch_be_response<T>(const T *, const T *, int64_t, T, T, T, T *, hipStream_t);              \
    template hipError_t launch_faddeeva<T>(const T *, const T *, int64_t, T *, T *, hipStream_t);
OCTOPIC_BE_INSTANTIATE(double)
OCTOPIC_BE_INSTANTIATE(float)
#undef OCTOPIC_BE_INSTANTIATE

hipError_t launch_gpic_addback(const GpicAddbackArgs &a, int64_t max_fld, hipStream_t s) {
    if (a.ninters <= 0 || max_fld <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gpic_addback, dim3((unsigned)((max_fld + 255) / 256), a.ninters), dim3(256), 0, s,
                       weideman_coeffs(), a);
    return hipGetLastError();
}

hipError_t launch_soft_kick(const SoftKickArgs &a, bool fp32, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    if (fp32) hipLaunchKernelGGL(k_soft_kick<float>, dim3(kSoftBlocks, 2 * a.npairs), dim3(256), 0, s, weideman_coeffs_f32(), a);
    else hipLaunchKernelGGL(k_soft_kick<double>, dim3(kSoftBlocks, 2 * a.npairs), dim3(256), 0, s, weideman_coeffs(), a);
    return hipGetLastError();
}

hipError_t launch_soft_lum_fold(const double *lum_block, const double *lum_scale, int npairs, double *lum_parts,
                                hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_soft_lum_fold, dim3((npairs + 255) / 256), dim3(256), 0, s, lum_block, lum_scale, npairs, lum_parts);
    return hipGetLastError();
}

}  // namespace octopic
