// fft.h -- in-LDS mixed-radix (4, 2, 3, 5) Stockham FFT for the padded PIC meshes (P = 2g points per axis).
//
// One wave (64 lanes) transforms one P-point complex sequence (double2 in the FP64 mode, float2 in
// the FP32 mode) held in LDS, ping-ponging between two LDS buffers, with the twiddle table staged in
// LDS once per workgroup.  Stage s of radix r reads r inputs P/r apart, twiddles them by
// W_P^(k*q*P/(Ns*r)) (k = j mod Ns, Ns = product of the earlier radices), applies the r-point DFT and
// writes them Ns apart at (j/Ns)*Ns*r + k: the autosort (Stockham) form, so the result is in natural
// order without a bit-reversal pass.
#pragma once

#include <hip/hip_runtime.h>

namespace octopic {

constexpr int kMaxStages = 12;

struct FFTPlanDev {
    int n;
    int nstages;
    int radix[kMaxStages];
    const double2 *tw;     // tw[t] = exp(-2 pi i t / n), t in [0, n), rounded from long double
    const float2 *tw32;    // the same twiddles rounded to float
};

// Complex arithmetic over double2 / float2 (C::x, C::y of scalar type real_of<C>).
template <class C> struct real_of_t;
template <> struct real_of_t<double2> { using type = double; };
template <> struct real_of_t<float2> { using type = float; };
template <class C> using real_of = typename real_of_t<C>::type;

template <class C>
__device__ __forceinline__ C cx(real_of<C> re, real_of<C> im) {
    C r;
    r.x = re;
    r.y = im;
    return r;
}
template <class C>
__device__ __forceinline__ C cmul(C a, C b) {
    return cx<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class C> __device__ __forceinline__ C cadd(C a, C b) { return cx<C>(a.x + b.x, a.y + b.y); }
template <class C> __device__ __forceinline__ C csub(C a, C b) { return cx<C>(a.x - b.x, a.y - b.y); }
template <class C> __device__ __forceinline__ C cconj(C a) { return cx<C>(a.x, -a.y); }
// multiply by -i (SIGN = -1) or +i (SIGN = +1)
template <int SIGN, class C>
__device__ __forceinline__ C mul_i(C a) {
    return SIGN < 0 ? cx<C>(a.y, -a.x) : cx<C>(-a.y, a.x);
}

template <int SIGN, class C>
__device__ __forceinline__ void dft2(C *v) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <int SIGN, class C>
__device__ __forceinline__ void dft4(C *v) {
    const C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const C s13 = cadd(v[1], v[3]), d13 = mul_i<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}
template <int SIGN, class C>
__device__ __forceinline__ void dft3(C *v) {
    using R = real_of<C>;
    const R c = (R)0.86602540378443864676;   // sqrt(3)/2
    const C t1 = cadd(v[1], v[2]);
    const C t2 = cx<C>(v[0].x - (R)0.5 * t1.x, v[0].y - (R)0.5 * t1.y);
    const C d = csub(v[1], v[2]);
    const C t3 = mul_i<SIGN>(cx<C>(c * d.x, c * d.y));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

template <int SIGN, class C>
__device__ __forceinline__ void dft5(C *v) {
    using R = real_of<C>;
    const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;   // cos(2pi/5), cos(4pi/5)
    const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;    // sin(2pi/5), sin(4pi/5)
    const C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const C a1 = cx<C>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const C a2 = cx<C>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    const C b1 = mul_i<SIGN>(cx<C>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    const C b2 = mul_i<SIGN>(cx<C>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

// Orders this wave's earlier LDS writes before its later LDS reads (a transform group is one wave, so
// no workgroup barrier is needed between its stages).
__device__ __forceinline__ void wave_lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

template <int SIGN, class C>
__device__ __forceinline__ C twiddle(const C *tw, int t) {
    const C w = tw[t];
    return SIGN < 0 ? w : cconj(w);
}

// Copies the plan's twiddles of type C into LDS (`tw_lds`, P entries); the caller barriers afterwards.
__device__ __forceinline__ void load_twiddles(const FFTPlanDev &P, double2 *tw_lds) {
    for (int t = threadIdx.x; t < P.n; t += blockDim.x) tw_lds[t] = P.tw[t];
}
__device__ __forceinline__ void load_twiddles(const FFTPlanDev &P, float2 *tw_lds) {
    for (int t = threadIdx.x; t < P.n; t += blockDim.x) tw_lds[t] = P.tw32[t];
}

// Transforms the sequence in `a` (LDS) using `b` as scratch; the result is in `a` when the plan has an
// even number of stages, else in `b` (see fft_result).  Called by the 64 lanes of one wave, which own
// `a` and `b`; `tw` is the LDS twiddle table.  The caller orders its writes of `a` before the call
// (wave_lds_sync or a workgroup barrier).
template <int SIGN, class C>
__device__ void fft_lds(C *a, C *b, const FFTPlanDev &P, const C *tw, int lane) {
    int Ns = 1;
    const int n = P.n;
    for (int s = 0; s < P.nstages; ++s) {
        const int r = P.radix[s];
        const int m = n / r;
        const int tstep = n / (Ns * r);
        if (r == 4) {
            for (int j = lane; j < m; j += 64) {
                const int q0 = j / Ns, k = j - q0 * Ns;
                C v[4];
                v[0] = a[j];
                v[1] = cmul(a[j + m], twiddle<SIGN>(tw, k * tstep));
                v[2] = cmul(a[j + 2 * m], twiddle<SIGN>(tw, 2 * k * tstep));
                v[3] = cmul(a[j + 3 * m], twiddle<SIGN>(tw, 3 * k * tstep));
                dft4<SIGN>(v);
                const int base = q0 * Ns * 4 + k;
                b[base] = v[0]; b[base + Ns] = v[1]; b[base + 2 * Ns] = v[2]; b[base + 3 * Ns] = v[3];
            }
        } else if (r == 2) {
            for (int j = lane; j < m; j += 64) {
                const int q0 = j / Ns, k = j - q0 * Ns;
                C v[2];
                v[0] = a[j];
                v[1] = cmul(a[j + m], twiddle<SIGN>(tw, k * tstep));
                dft2<SIGN>(v);
                const int base = q0 * Ns * 2 + k;
                b[base] = v[0]; b[base + Ns] = v[1];
            }
        } else if (r == 3) {
            for (int j = lane; j < m; j += 64) {
                const int q0 = j / Ns, k = j - q0 * Ns;
                C v[3];
                v[0] = a[j];
                v[1] = cmul(a[j + m], twiddle<SIGN>(tw, k * tstep));
                v[2] = cmul(a[j + 2 * m], twiddle<SIGN>(tw, 2 * k * tstep));
                dft3<SIGN>(v);
                const int base = q0 * Ns * 3 + k;
                b[base] = v[0]; b[base + Ns] = v[1]; b[base + 2 * Ns] = v[2];
            }
        } else {
            for (int j = lane; j < m; j += 64) {
                const int q0 = j / Ns, k = j - q0 * Ns;
                C v[5];
                v[0] = a[j];
#pragma unroll
                for (int q = 1; q < 5; ++q) v[q] = cmul(a[j + q * m], twiddle<SIGN>(tw, q * k * tstep));
                dft5<SIGN>(v);
                const int base = q0 * Ns * 5 + k;
#pragma unroll
                for (int q = 0; q < 5; ++q) b[base + q * Ns] = v[q];
            }
        }
        wave_lds_sync();
        C *t = a; a = b; b = t;
        Ns *= r;
    }
}

template <class C>
__device__ __forceinline__ C *fft_result(C *a, C *b, const FFTPlanDev &P) {
    return (P.nstages & 1) ? b : a;
}

}  // namespace octopic
