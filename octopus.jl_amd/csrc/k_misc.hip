// k_misc.hip -- the turn around the collision: linear one-turn map, synthetic beams, stencil probe.
//
//   linear6d : track_particle(::Symplectic6DMap, ::Linear6D, ...) (src/elements/linear6d.jl:110-118).
//   generate : Beam(N, ...) counter-RNG construction (src/beam/Beam.jl:127-181, 545-553) with the
//              Philox4x32-10 stream of src/math/counter_rng.jl:184-277 and Box-Muller (153-173).
#include <algorithm>
#include <mutex>

#include "common.h"
#include "launch.h"

namespace octopic {

struct Mat36 { double m[36]; };

__global__ __launch_bounds__(256) void k_linear6d(Mat36 M, double *x, double *px, double *y, double *py, double *z,
                                                  double *pz, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v[6] = {x[i], px[i], y[i], py[i], z[i], pz[i]};
        double o[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const double *q = M.m + 6 * r;
            o[r] = q[0] * v[0] + q[1] * v[1] + q[2] * v[2] + q[3] * v[3] + q[4] * v[4] + q[5] * v[5];
        }
        x[i] = o[0]; px[i] = o[1]; y[i] = o[2]; py[i] = o[3]; z[i] = o[4]; pz[i] = o[5];
    }
}

hipError_t launch_linear6d(const double *m36, double *const c[6], int64_t n, hipStream_t s) {
    Mat36 M;
    for (int k = 0; k < 36; ++k) M.m[k] = m36[k];
    int g = (int)((n + 255) / 256);
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(k_linear6d, dim3(g), dim3(256), 0, s, M, c[0], c[1], c[2], c[3], c[4], c[5], n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ crossing-angle boost, crab cavity
// LorentzBoost / RevLorentzBoost (lorentz_boost.jl:128-166), Hirata's map with angle = half the full
// crossing angle; cos, sin, tan = sin/cos computed on the host once (lorentz_boost.jl:168-174).
template <bool REV>
__global__ __launch_bounds__(256) void k_lorentz_boost(double c, double s, double t, double *X, double *PX, double *Y,
                                                       double *PY, double *Z, double *PZ, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x0 = X[i], px0 = PX[i], y0 = Y[i], py0 = PY[i], z0 = Z[i], pz0 = PZ[i];
        const double ps0 = sqrt((1.0 + pz0) * (1.0 + pz0) - px0 * px0 - py0 * py0);
        const double h0 = 1.0 + pz0 - ps0;
        double x1, px1, y1, py1, z1, pz1;
        if (!REV) {
            py1 = py0 / c;
            const double h1 = h0 / (c * c);
            px1 = px0 / c - h1 * s;
            pz1 = pz0 - px1 * s;
            const double ps1 = 1.0 + pz1 - h1;
            const double ds = x0 * s;
            x1 = x0 + z0 * t + px1 / ps1 * ds;
            y1 = y0 + py1 / ps1 * ds;
            z1 = z0 / c - h1 / ps1 * ds;
        } else {
            x1 = x0 - z0 * s;
            x1 = x1 / (1.0 + (px0 + h0 * s) * s / ps0);
            z1 = (z0 + h0 / ps0 * x1 * s) * c;
            y1 = y0 - py0 / ps0 * x1 * s;
            pz1 = pz0 + px0 * s;
            px1 = (px0 + h0 * s) * c;
            py1 = py0 * c;
        }
        X[i] = x1; PX[i] = px1; Y[i] = y1; PY[i] = py1; Z[i] = z1; PZ[i] = pz1;
    }
}

hipError_t launch_lorentz_boost(double angle, int reverse, double *const c[6], int64_t n, hipStream_t s) {
    const double co = cos(angle), si = sin(angle), ta = si / co;
    int g = (int)((n + 255) / 256);
    if (g > 16384) g = 16384;
    if (reverse) hipLaunchKernelGGL(k_lorentz_boost<true>, dim3(g), dim3(256), 0, s, co, si, ta, c[0], c[1], c[2], c[3], c[4], c[5], n);
    else hipLaunchKernelGGL(k_lorentz_boost<false>, dim3(g), dim3(256), 0, s, co, si, ta, c[0], c[1], c[2], c[3], c[4], c[5], n);
    return hipGetLastError();
}

// ThinCrabCavity{N} (crab_cavity.jl:99-164): per harmonic i, theta = i kcc z + phase_i,
// px -= sx_i sin/(i kcc), py -= sy_i sin/(i kcc), pz -= (sx_i x + sy_i y) cos; x, y, z unchanged.
struct CrabHarmonics { double sx[kMaxCrabHarmonics], sy[kMaxCrabHarmonics], ph[kMaxCrabHarmonics]; };

__global__ __launch_bounds__(256) void k_crab_cavity(double kcc, int nh, CrabHarmonics hm, const double *X, double *PX,
                                                     const double *Y, double *PY, const double *Z, double *PZ, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x0 = X[i], y0 = Y[i], z0 = Z[i];
        double dpx = 0.0, dpy = 0.0, dpz = 0.0;
        for (int k = 1; k <= nh; ++k) {
            const double ikcc = (double)k * kcc;
            const double th = ikcc * z0 + hm.ph[k - 1];
            const double sn = sin(th), cs = cos(th);
            dpx -= hm.sx[k - 1] * sn / ikcc;
            dpy -= hm.sy[k - 1] * sn / ikcc;
            dpz -= (hm.sx[k - 1] * x0 + hm.sy[k - 1] * y0) * cs;
        }
        PX[i] = PX[i] + dpx;
        PY[i] = PY[i] + dpy;
        PZ[i] = PZ[i] + dpz;
    }
}

hipError_t launch_crab_cavity(double kcc, int nh, const double *sx, const double *sy, const double *phase, double *const c[6],
                              int64_t n, hipStream_t s) {
    if (nh < 0 || nh > kMaxCrabHarmonics) return hipErrorInvalidValue;
    CrabHarmonics hm{};
    for (int k = 0; k < nh; ++k) { hm.sx[k] = sx[k]; hm.sy[k] = sy[k]; hm.ph[k] = phase[k]; }
    int g = (int)((n + 255) / 256);
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(k_crab_cavity, dim3(g), dim3(256), 0, s, kcc, nh, hm, c[0], c[1], c[2], c[3], c[4], c[5], n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Philox4x32-10 counter RNG
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t counter_uint64(uint64_t seed, uint64_t turn, uint64_t rng_id, uint64_t particle,
                                                   uint64_t component) {
    uint32_t c0 = (uint32_t)particle, c1 = (uint32_t)(particle >> 32), c2 = (uint32_t)turn, c3 = (uint32_t)(turn >> 32);
    const uint64_t key = splitmix64(seed) ^ splitmix64(rng_id + 0x9e3779b97f4a7c15ull) ^
                         splitmix64(component + 0xbf58476d1ce4e5b9ull);
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1,
                       n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return ((uint64_t)c0 << 32) | (uint64_t)c1;
}

__device__ __forceinline__ double uniform_open01(uint64_t v) { return ((double)(v >> 12) + 0.5) * 2.220446049250313e-16; }

// One coordinate d (0..5) of every particle: octopus_normal(seed, philox, 0, rng_id, i, d + 1), clamped.
__global__ __launch_bounds__(256) void k_normals(double *out, int64_t n, uint64_t seed, uint64_t rng_id, int d, double cutoff) {
    const int pair_id = d / 2 + 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t part = (uint64_t)(i + 1);
        const double u1 = uniform_open01(counter_uint64(seed, 0, rng_id, part, 2 * pair_id - 1));
        const double u2 = uniform_open01(counter_uint64(seed, 0, rng_id, part, 2 * pair_id));
        const double r = sqrt(-2.0 * log(u1));
        const double th = 6.283185307179586476925286766559 * u2;
        double v = (d % 2 == 0) ? r * cos(th) : r * sin(th);
        if (__builtin_isfinite(cutoff)) v = jclamp(v, -cutoff, cutoff);
        out[i] = v;
    }
}

// Deterministic two-level sum (fixed partition) of v or (v - mu)^2.
__global__ __launch_bounds__(256) void k_partial_sum(const double *v, int64_t n, const double *mu, double *partial) {
    __shared__ double red[256];
    const double m = mu ? *mu : 0.0;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = mu ? (v[i] - m) * (v[i] - m) : v[i];
        acc += a;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_finish_sum(const double *partial, int nb, int64_t n, double *out, int sqrt_it) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[b];
    s = s / (double)n;
    *out = sqrt_it ? sqrt(s > 0.0 ? s : 0.0) : s;
}

__global__ __launch_bounds__(256) void k_standardize_scale(double *v, int64_t n, const double *mu, const double *sd, double mul) {
    const double m = *mu, s = *sd;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = s == 0.0 ? 0.0 : ((v[i] - m) / s) * mul;
}

__global__ __launch_bounds__(256) void k_alpha(double *q, double *p, int64_t n, double alpha, double fac) {
    // p <- (sigma/beta) * (p_std - alpha * q_std), applied before q is scaled; fac = sigma/beta
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = fac * (p[i] - alpha * q[i]);
}

// Blocks of the generator's partial sums: the partition (and so every sum's rounding) depends only on n.
constexpr int kGenMaxBlocks = 2048;
int64_t generate_workspace_doubles(int64_t n) {
    const int64_t g = std::min<int64_t>((n + 255) / 256, kGenMaxBlocks);
    return g + 2;
}

// Standard-normal coordinates, standardised to mean 0 / rms 1; the Twiss scaling is applied by the host
// wrapper with launch_scale_beam.  ws holds generate_workspace_doubles(n) doubles: the block partial sums,
// then the mean and the rms of the coordinate in hand.  Every entry is written before it is read.
hipError_t launch_generate_ws(double *const c[6], int64_t n, uint64_t seed, uint64_t rng_id, double cutoff, double *ws,
                              hipStream_t s) {
    const int g = (int)(generate_workspace_doubles(n) - 2);
    for (int d = 0; d < 6; ++d) {
        hipLaunchKernelGGL(k_normals, dim3(g), dim3(256), 0, s, c[d], n, seed, rng_id, d, cutoff);
        hipLaunchKernelGGL(k_partial_sum, dim3(g), dim3(256), 0, s, c[d], n, (const double *)nullptr, ws);
        hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(64), 0, s, ws, g, n, ws + g, 0);
        hipLaunchKernelGGL(k_partial_sum, dim3(g), dim3(256), 0, s, c[d], n, (const double *)(ws + g), ws);
        hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(64), 0, s, ws, g, n, ws + g + 1, 1);
        hipLaunchKernelGGL(k_standardize_scale, dim3(g), dim3(256), 0, s, c[d], n, (const double *)(ws + g),
                           (const double *)(ws + g + 1), 1.0);
    }
    return hipGetLastError();
}

hipError_t launch_generate(double *const c[6], int64_t n, uint64_t seed, uint64_t rng_id, double cutoff, hipStream_t s) {
    double *ws = nullptr;
    hipError_t e = hipMallocAsync((void **)&ws, sizeof(double) * generate_workspace_doubles(n), s);
    if (e != hipSuccess) return e;
    e = launch_generate_ws(c, n, seed, rng_id, cutoff, ws, s);
    hipFreeAsync(ws, s);
    return e != hipSuccess ? e : hipGetLastError();
}

// The generator's workspace for callers that cannot allocate: one static array per device, as large as
// the partition's cap needs (2048 block partials + mean + rms = 16400 bytes), so every n fits.  Every
// entry is written before it is read within a call, so nothing is cleared; two calls that overlap on
// one device share the array and race.  In host code `scratch` names a host-side stand-in: its device
// address comes from hipGetSymbolAddress, once per device, and only that address is used.
constexpr int kGenScratchDoubles = kGenMaxBlocks + 2;
static __device__ __align__(256) double scratch[kGenScratchDoubles];
constexpr int kScratchMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kScratchMaxDevices];

int64_t generate_scratch_doubles() { return kGenScratchDoubles; }

hipError_t generate_scratch_init(int device) {
    if (device < 0 || device >= kScratchMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;   // the caller makes `device` current
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = static_cast<double *>(p);
    return hipSuccess;
}

hipError_t launch_generate_noalloc(double *const c[6], int64_t n, uint64_t seed, uint64_t rng_id, double cutoff, int device,
                                   hipStream_t s) {
    if (device < 0 || device >= kScratchMaxDevices) return hipErrorInvalidDevice;
    if (n <= 0 || generate_workspace_doubles(n) > kGenScratchDoubles) return hipErrorInvalidValue;
    double *ws = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        ws = g_scratch_ptr[device];
    }
    if (!ws) return hipErrorNotInitialized;
    return launch_generate_ws(c, n, seed, rng_id, cutoff, ws, s);
}


__global__ __launch_bounds__(256) void k_scale(double *v, int64_t n, double mul) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) v[i] *= mul;
}

hipError_t launch_scale_beam(double *const c[6], int64_t n, const double *beta, const double *alpha, const double *sigma,
                             hipStream_t s) {
    int g = (int)((n + 255) / 256);
    if (g > 4096) g = 4096;
    for (int k = 0; k < 3; ++k) {   // _normalize_sigma! (Beam.jl:545-553)
        hipLaunchKernelGGL(k_alpha, dim3(g), dim3(256), 0, s, c[2 * k], c[2 * k + 1], n, alpha[k], sigma[k] / beta[k]);
        hipLaunchKernelGGL(k_scale, dim3(g), dim3(256), 0, s, c[2 * k], n, sigma[k]);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ Float32 beams
// Exact widening of a float coordinate array into the FP64 collision copy, and the rounding back.
__global__ __launch_bounds__(256) void k_widen(const float *src, double *dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (double)src[i];
}
__global__ __launch_bounds__(256) void k_narrow(const double *src, float *dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

hipError_t launch_widen(const float *src, double *dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int g = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_widen, dim3(g), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}
hipError_t launch_narrow(const double *src, float *dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int g = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_narrow, dim3(g), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ stencil probe
__global__ void k_debug_weights(int method, const double *u, int64_t n, int npts, int *base, double *w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b;
    double ww[3];
    if (method == 0) stencil<0>(u[i], npts, b, ww); else stencil<1>(u[i], npts, b, ww);
    base[i] = b;
    w[3 * i] = ww[0]; w[3 * i + 1] = ww[1]; w[3 * i + 2] = ww[2];
}

hipError_t launch_debug_weights(int method, const double *u, int64_t n, int npts, int *base, double *w, hipStream_t s) {
    hipLaunchKernelGGL(k_debug_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, method, u, n, npts, base, w);
    return hipGetLastError();
}

}  // namespace octopic
