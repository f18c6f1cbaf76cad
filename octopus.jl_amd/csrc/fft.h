// fft.h -- in-LDS mixed-radix Stockham FFT for the padded PIC meshes (P = 2g points per axis).
//
// A group of `nthr` lanes transforms one P-point complex FP64 sequence held in LDS, ping-ponging
// between two LDS buffers; every group of a workgroup runs the same stage sequence, so the
// workgroup barrier between stages is uniform.  Stage s of radix r reads r inputs P/r apart,
// twiddles them by W_P^(k*q*P/(Ns*r)) (k = j mod Ns, Ns = product of the earlier radices), applies
// the r-point DFT and writes them Ns apart at (j/Ns)*Ns*r + k: the autosort (Stockham) form, so
// the result is in natural order without a bit-reversal pass.
#pragma once

#include <hip/hip_runtime.h>

namespace octopic {

constexpr int kMaxStages = 12;

struct FFTPlanDev {
    int n;
    int nstages;
    int radix[kMaxStages];
    const double2 *tw;   // tw[t] = exp(-2 pi i t / n), t in [0, n)
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by -i (SIGN = -1) or +i (SIGN = +1)
template <int SIGN>
__device__ __forceinline__ double2 mul_i(double2 a) {
    return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

template <int SIGN>
__device__ __forceinline__ void dft2(double2 *v) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <int SIGN>
__device__ __forceinline__ void dft4(double2 *v) {
    double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    double2 s13 = cadd(v[1], v[3]), d13 = mul_i<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}
template <int SIGN>
__device__ __forceinline__ void dft3(double2 *v) {
    const double c = 0.86602540378443864676;   // sqrt(3)/2
    double2 t1 = cadd(v[1], v[2]);
    double2 t2 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    double2 d = csub(v[1], v[2]);
    double2 t3 = mul_i<SIGN>(make_double2(c * d.x, c * d.y));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

template <int SIGN>
__device__ __forceinline__ double2 twiddle(const double2 *tw, int t) {
    double2 w = tw[t];
    return SIGN < 0 ? w : cconj(w);
}

// Transforms the sequence in `a` (LDS) using `b` as scratch; returns the buffer holding the result.
// Must be called by every lane of the workgroup (it contains workgroup barriers).
template <int SIGN>
__device__ double2 *fft_lds(double2 *a, double2 *b, const FFTPlanDev &P, int lane, int nthr) {
    int Ns = 1;
    const int n = P.n;
    for (int s = 0; s < P.nstages; ++s) {
        const int r = P.radix[s];
        const int m = n / r;
        const int tstep = n / (Ns * r);
        for (int j = lane; j < m; j += nthr) {
            const int q0 = j / Ns;
            const int k = j - q0 * Ns;
            double2 v[4];
            v[0] = a[j];
            for (int q = 1; q < r; ++q) v[q] = cmul(a[j + q * m], twiddle<SIGN>(P.tw, k * q * tstep));
            if (r == 4) dft4<SIGN>(v);
            else if (r == 2) dft2<SIGN>(v);
            else dft3<SIGN>(v);
            const int base = q0 * Ns * r + k;
            for (int q = 0; q < r; ++q) b[base + q * Ns] = v[q];
        }
        __syncthreads();
        double2 *t = a; a = b; b = t;
        Ns *= r;
    }
    return a;
}

}  // namespace octopic
