// common.h -- device-side types and scalar helpers shared by the octopic HIP kernels.
//
// Scalar semantics follow the reference's Julia code exactly where a bit-exact result is owed
// (bounds, grid geometry, cell indices): NaN-propagating min/max, Julia's clamp, and the
// CIC/TSC weight closed forms (src/tasks/strongstrong/pic_cpu.jl:2182-2252).  Everything is
// compiled with -ffp-contract=off so that no product is fused into an FMA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "octopic.h"

namespace octopic {

constexpr int kFoldLanes = 4096;          // _SLICE_FOLD_LANES (slicing.jl:223)
constexpr double kTemplateMargin = 1.5;   // _PIC_TEMPLATE_MARGIN_CELLS (interface.jl:674)

// ------------------------------------------------------------------ Julia scalar semantics
__host__ __device__ __forceinline__ double jmin(double x, double y) {
    if (x != x) return x;
    if (y != y) return y;
    if (y < x) return y;
    if (x < y) return x;
    return __builtin_signbit(y) ? y : x;
}
__host__ __device__ __forceinline__ double jmax(double x, double y) {
    if (x != x) return x;
    if (y != y) return y;
    if (y > x) return y;
    if (x > y) return x;
    return __builtin_signbit(y) ? x : y;
}
__host__ __device__ __forceinline__ double jclamp(double x, double lo, double hi) {
    return x > hi ? hi : (x < lo ? lo : x);
}
__host__ __device__ __forceinline__ int iclamp(int x, int lo, int hi) { return x > hi ? hi : (x < lo ? lo : x); }
__host__ __device__ __forceinline__ bool jisless(double a, double b) {
    if (a != a) return false;
    if (b != b) return true;
    if (a < b) return true;
    if (a > b) return false;
    return __builtin_signbit(a) && !__builtin_signbit(b);
}
__host__ __device__ __forceinline__ bool finite4(double a, double b, double c, double d) {
    return __builtin_isfinite(a) && __builtin_isfinite(b) && __builtin_isfinite(c) && __builtin_isfinite(d);
}

// CIC / TSC stencil on an axis of n nodes.  base0 is the 0-based first stencil node; out-of-range
// or non-finite u gives zero weights at node 0, as the reference's (1, zeros) (pic_cpu.jl:2182-2252).
__host__ __device__ __forceinline__ void cic_weights(double u, int n, int &base0, double w[3]) {
    if (!(0.0 <= u && u <= (double)(n - 1))) { base0 = 0; w[0] = 0.0; w[1] = 0.0; w[2] = 0.0; return; }
    int base = iclamp((int)floor(u) + 1, 1, n - 1);
    double f = jclamp(u - (double)(base - 1), 0.0, 1.0);
    w[0] = 1.0 - f; w[1] = f; w[2] = 0.0;
    base0 = base - 1;
}
__host__ __device__ __forceinline__ void tsc_weights(double u, int n, int &base0, double w[3]) {
    if (!(0.0 <= u && u <= (double)(n - 1))) { base0 = 0; w[0] = 0.0; w[1] = 0.0; w[2] = 0.0; return; }
    double fl = floor(u);
    int ix = (int)fl;
    double f = u - fl;
    int base;
    if (f < 0.5) {
        double t = f * f;
        w[0] = 0.125 + 0.5 * (t - f); w[1] = 0.75 - t; w[2] = 0.125 + 0.5 * (t + f);
        base = ix;
    } else {
        double fr = 1.0 - f;
        double t = fr * fr;
        w[0] = 0.125 + 0.5 * (t + fr); w[1] = 0.75 - t; w[2] = 0.125 + 0.5 * (t - fr);
        base = ix + 1;
    }
    base0 = iclamp(base, 1, n - 2) - 1;
}
template <int METHOD>
__host__ __device__ __forceinline__ void stencil(double u, int n, int &base0, double w[3]) {
    if (METHOD == 0) cic_weights(u, n, base0, w); else tsc_weights(u, n, base0, w);
}

// Stencil nodes [lo, hi] on an axis of n nodes used by every position whose normalised coordinate lies
// in [umin, umax] (the first stencil node is monotonic in u over the in-range interval).  A bound that is
// out of range or non-finite opens the window to that end of the axis; positions outside the axis
// (first node 0, zero weights) may still fall outside the window, so users check each node.
template <int METHOD>
__device__ __forceinline__ void stencil_window(double umin, double umax, int n, int &lo, int &hi) {
    constexpr int K = METHOD == 0 ? 2 : 3;
    double w[3];
    int b;
    if (0.0 <= umin && umin <= (double)(n - 1)) { stencil<METHOD>(umin, n, b, w); lo = b; } else lo = 0;
    if (0.0 <= umax && umax <= (double)(n - 1)) { stencil<METHOD>(umax, n, b, w); hi = b + K - 1; } else hi = n - 1;
    if (hi < lo) { lo = 0; hi = n - 1; }
}

// Round-to-nearest-even of a double in [0, 2^52) to uint64 (what __double2ull_rn gives there): adding
// 2^52 leaves the rounded integer in the low mantissa bits, one FP64 add instead of the general
// conversion sequence.
__device__ __forceinline__ unsigned long long fixed_point(double v) {
    const double t = v + 4503599627370496.0;   // 2^52
    return (unsigned long long)__builtin_bit_cast(long long, t) - 0x4330000000000000ull;
}

// round(a * b) to uint64 for a * b in [0, 2^52), in one rounding of the exact product: the fused
// a * b + 2^52 lands the rounded integer in the low 52 mantissa bits, which a mask keeps.  With a the
// weight times the fixed-point scale (a power of two, so that product is exact) this is the deposit's
// quantisation round(wx * wy * 2^k).
__device__ __forceinline__ unsigned long long fixed_point_product(double a, double b) {
    const double t = __builtin_fma(a, b, 4503599627370496.0);   // 2^52
    return (unsigned long long)__builtin_bit_cast(long long, t) & 0x000FFFFFFFFFFFFFull;
}

// Order-preserving uint64 key of a non-NaN double (-0.0 sorts below +0.0, as Julia's min/max), so that
// exact min/max reductions can use 64-bit integer atomics in any order.
__host__ __device__ __forceinline__ unsigned long long order_key(double x) {
    unsigned long long u = (unsigned long long)__builtin_bit_cast(long long, x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double key_value(unsigned long long k) {
    unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, (long long)u);
}

// Bound slots for exact min/max folds with 64-bit atomicMax: a max slot holds order_key(v), a min slot
// ~order_key(v); NaN is all-ones in both, so it wins every max as NaN wins Julia's min/max; zero is
// the empty slot.
__device__ __forceinline__ unsigned long long bound_slot(double v, bool is_min) {
    if (v != v) return ~0ull;
    const unsigned long long k = order_key(v);
    return is_min ? ~k : k;
}
__device__ __forceinline__ double bound_value(unsigned long long k, bool is_min) {
    if (k == ~0ull) return __builtin_nan("");
    if (k == 0ull) return is_min ? INFINITY : -INFINITY;
    return key_value(is_min ? ~k : k);
}


// ------------------------------------------------------------------ descriptors in HBM
struct Grid { double x0, y0, width, height; };

// Per-beam slice table, written by the slicing kernels.
struct SliceTable {
    double boundary[257];
    double center[256];
    double weight[256];
    int64_t count[256];
    int64_t offset[256];
    double zmin, zmax, mean, sigma;
    int32_t ns;
    int32_t degenerate;
};
constexpr int kMaxSlices = 256;

// The source slice's Gaussian drifted to one field-slice boundary (Gaussian-subtracted solver): centroid,
// rms sizes and the drift derivatives of the variances, d(sigma^2)/ds (gaussian_pic.jl:482-523).
struct GaussBnd { double mux, muy, sigx, sigy, rx, ry; };

// One directed interaction of a wavefront batch (a source slice solved on a mesh, kicking a field slice).
struct Inter {
    int32_t p, i, j, dir;          // pair position in collision order, slices, direction (1, 2)
    int32_t src_beam, fld_beam;    // 0 = beam 1, 1 = beam 2
    int64_t src_off, src_n, fld_off, fld_n;
    double c_src, c_fld, lb, rb;   // source centre; field slice centre and boundaries
    double sL, sR, sM;             // source drifts to the field-slice boundaries / luminosity plane
    double kick_scale;             // 2 * kbb
    Grid sg, fg;                   // meshes (slice_pair mode)
    int32_t plane0;                // first field plane (L, R[, M | Z]) of this interaction in the batch
    int32_t nplanes;               // 2 or 3
    int32_t event;                 // 0 none, 1 hit, 2 miss, 3 rebuild
    int32_t valid;
    Grid fgR;                      // node mode: the right node's field mesh
    int32_t count_field;           // 1: count field particles outside fg after the forward drift
    int32_t gauss;                 // 1: Gaussian-subtracted planes with the exact field added back per particle
    GaussBnd g[2];                 // the source's drifted Gaussian at sL and sR (gauss = 1)
    double gmpx, gmpy;             // the source's mean px, py (gauss = 1)
};

// One deposited plane: a source slice drifted by s onto a mesh.
struct Plane {
    int32_t inter;                 // owning interaction (or pair, for luminosity planes)
    int32_t kind;                  // 0 field plane, 1 luminosity plane
    int32_t beam;                  // source beam
    int32_t n1, n2;                // nodes along x and y of the deposit mesh
    int32_t slot;                  // Green spectrum slot (field planes)
    int32_t count_box;             // 1: count particles whose drifted position leaves [x0, bx1] x [y0, by1]
    int32_t reserved;
    int64_t off, n;                // particle range in the slice-sorted state
    double s;                      // drift
    double x0, y0, hxi, hyi;       // deposit mesh origin and inverse spacing
    double hx, hy;                 // spacing (field derivative)
    double bx1, by1;               // x0 + width, y0 + height (dropped-particle box)
};

// A Green table to build into a cache slot.
struct GreenJob {
    int32_t slot;
    int32_t valid;
    double fx0, fy0, sx0, sy0, hx, hy;
};

struct ErrorInfo {
    int32_t code;                  // OCTOPIC_ERR_* or 0
    int32_t p, dir, stage;
    double a, b;
};

// phi and E = -grad(phi) at node (x, y) of an nx x ny plane of spacings 1/hxi, 1/hyi (pic_cpu.jl:2531-2591):
// one-sided second-order differences on the edges, the centred second-order difference one node in when
// `fourth`, the fourth-order one inside (else second order throughout).  F(i, j) returns phi at node
// (i, j); only nodes within two of (x, y) on the mesh are read.
template <class R, class F>
__device__ __forceinline__ void field_node(const F &phi, int x, int y, int nx, int ny, int fourth, R hxi, R hyi, R &v0,
                                           R &v1, R &v2) {
    const R c4 = (R)(1.0 / 12.0);
    const R h5 = (R)0.5, one5 = (R)1.5, two = (R)2.0, eight = (R)8.0;
    R ex, ey;
    if (y == 0) ey = hyi * (one5 * phi(x, 0) - two * phi(x, 1) + h5 * phi(x, 2));
    else if (y == ny - 1) ey = hyi * (-one5 * phi(x, ny - 1) + two * phi(x, ny - 2) - h5 * phi(x, ny - 3));
    else if (fourth && (y == 1 || y == ny - 2)) ey = h5 * hyi * (phi(x, y - 1) - phi(x, y + 1));
    else if (fourth) ey = c4 * hyi * ((phi(x, y + 2) - phi(x, y - 2)) + eight * (phi(x, y - 1) - phi(x, y + 1)));
    else ey = h5 * hyi * (phi(x, y - 1) - phi(x, y + 1));
    if (x == 0) ex = hxi * (one5 * phi(0, y) - two * phi(1, y) + h5 * phi(2, y));
    else if (x == nx - 1) ex = hxi * (-one5 * phi(nx - 1, y) + two * phi(nx - 2, y) - h5 * phi(nx - 3, y));
    else if (fourth && (x == 1 || x == nx - 2)) ex = h5 * hxi * (phi(x - 1, y) - phi(x + 1, y));
    else if (fourth) ex = c4 * hxi * ((phi(x + 2, y) - phi(x - 2, y)) + eight * (phi(x - 1, y) - phi(x + 1, y)));
    else ex = h5 * hxi * (phi(x - 1, y) - phi(x + 1, y));
    v0 = phi(x, y);
    v1 = ex;
    v2 = ey;
}

}  // namespace octopic
