// pic_oracle.cpp -- CPU oracle for the strong-strong PIC collision.
//
// TEST INFRASTRUCTURE ONLY.  This is a single-threaded (optionally OpenMP pair-batched) FP64
// restatement of the reference's CPU algorithm, used as the checker of the HIP library and as the
// timed CPU baseline.  Nothing in the product (octopus.jl_amd/) links, loads or calls it; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
//
// It follows, function by function (file:line cite the reference tree):
//   src/tasks/strongstrong/pic_cpu.jl    the collision (_pic_collide!, _pic_interaction!, ...)
//   src/tasks/strongstrong/slicing.jl    longitudinal slicing and the collision schedule
//   src/math/counter_rng.jl              Philox4x32-10 counter RNG + Box-Muller
//   src/beam/Beam.jl                     synthetic Gaussian beams
//   src/elements/linear6d.jl             the linear one-turn map
//   src/elements/lorentz_boost.jl        crossing-angle boost and its reverse
//   src/elements/crab_cavity.jl          thin crab cavity
//   src/elements/strong_beam.jl          gaussian_beambeam_kick (Bassetti-Erskine), pinned by
//                                        quadrature of the field integral and the reference's KATs
//   src/math/SpecialMath.jl              the Faddeeva function of the device path, pinned by
//                                        scipy.special.wofz vectors (tests/golden/bassetti_erskine.npz)
// Julia evaluation order is kept expression by expression (a*b*c == (a*b)*c, a-b+c == (a-b)+c),
// Julia's NaN-propagating min/max/clamp are restated, and the file is compiled with
// -ffp-contract=off so that no product is fused into an FMA the reference does not perform.
// FFTW (the reference's FFT, third-party FFTW_jll 3.3.12) is restated as a mixed-radix DFT; its
// rounding is not reproduced bitwise (the oracle is pinned against numpy.fft and a direct DFT).
// SpecialFunctions.jl 2.8.0 erfinv is restated as erf-based Newton refinement of a rational seed.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/octopic.h"

namespace {

using cd = std::complex<double>;
using Vec = std::vector<double>;

struct OracleError : std::runtime_error {
    int code;
    OracleError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

thread_local std::string g_last_error;

// ------------------------------------------------------------------------------------------
// Julia scalar semantics
// ------------------------------------------------------------------------------------------
// Base.min / Base.max for floats: NaN-propagating, -0.0 < 0.0.
inline double jmin(double x, double y) {
    if (std::isnan(x)) return x;
    if (std::isnan(y)) return y;
    if (y < x) return y;
    if (x < y) return x;
    return std::signbit(y) ? y : x;
}
inline double jmax(double x, double y) {
    if (std::isnan(x)) return x;
    if (std::isnan(y)) return y;
    if (y > x) return y;
    if (x > y) return x;
    return std::signbit(y) ? x : y;
}
// Base.clamp(x, lo, hi) = ifelse(x > hi, hi, ifelse(x < lo, lo, x))
inline double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }
inline int64_t iclamp(int64_t x, int64_t lo, int64_t hi) { return x > hi ? hi : (x < lo ? lo : x); }
// Base.isless for Float64 (NaN largest, -0.0 < 0.0)
inline bool jisless(double a, double b) {
    if (std::isnan(a)) return false;
    if (std::isnan(b)) return true;
    if (a < b) return true;
    if (a > b) return false;
    return std::signbit(a) && !std::signbit(b);
}
// Base.searchsortedlast(v, x) on a sorted vector, 1-based result (0 = before the first).
inline int64_t searchsortedlast(const Vec &v, double x) {
    int64_t lo = 0, hi = (int64_t)v.size() + 1;
    while (lo < hi - 1) {
        int64_t m = lo + ((hi - lo) >> 1);
        if (jisless(x, v[m - 1])) hi = m; else lo = m;
    }
    return lo;
}
inline int64_t fld(int64_t a, int64_t b) {  // floor division, b > 0
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

// ------------------------------------------------------------------------------------------
// Counter RNG (counter_rng.jl:105-431)
// ------------------------------------------------------------------------------------------
constexpr uint32_t PHILOX_M0 = 0xD2511F53u, PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u, PHILOX_W1 = 0xBB67AE85u;
constexpr double TWO_NEG_52 = 2.220446049250313e-16;
constexpr double COUNTER_RNG_TWO_PI = 6.283185307179586476925286766559;

inline uint64_t splitmix64(uint64_t x) {  // counter_rng.jl:405-410
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

inline void philox_block(uint32_t c[4], uint32_t k0, uint32_t k1) {  // counter_rng.jl:232-240, 420-431
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c[0];
        uint64_t p1 = (uint64_t)PHILOX_M1 * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
}

inline uint64_t counter_uint64(uint64_t seed, uint64_t turn, uint64_t rng_id, uint64_t particle,
                               uint64_t component) {  // counter_rng.jl:184-220, 273-277
    uint32_t c[4] = {(uint32_t)particle, (uint32_t)(particle >> 32), (uint32_t)turn,
                     (uint32_t)(turn >> 32)};
    uint64_t key = splitmix64(seed) ^ splitmix64(rng_id + 0x9e3779b97f4a7c15ull) ^
                   splitmix64(component + 0xbf58476d1ce4e5b9ull);
    philox_block(c, (uint32_t)key, (uint32_t)(key >> 32));
    return ((uint64_t)c[0] << 32) | (uint64_t)c[1];
}

inline double uniform_open01(uint64_t v) {  // counter_rng.jl:337-340
    uint64_t bits = v >> 12;
    return ((double)bits + 0.5) * TWO_NEG_52;
}

inline double octopus_normal(uint64_t seed, uint64_t turn, uint64_t rng_id, uint64_t particle,
                             int64_t component) {  // counter_rng.jl:153-173
    int64_t pair_id = (component + 1) / 2;
    double u1 = uniform_open01(counter_uint64(seed, turn, rng_id, particle, 2 * pair_id - 1));
    double u2 = uniform_open01(counter_uint64(seed, turn, rng_id, particle, 2 * pair_id));
    double r = std::sqrt(-2.0 * std::log(u1));
    double theta = COUNTER_RNG_TWO_PI * u2;
    return (component % 2 == 1) ? r * std::cos(theta) : r * std::sin(theta);
}

// Pairwise sum (Julia's `sum` reduces pairwise; its SIMD blocking is not reproduced).
double pairwise_sum(const double *v, int64_t n) {
    if (n <= 1024) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += v[i];
        return s;
    }
    int64_t h = n >> 1;
    return pairwise_sum(v, h) + pairwise_sum(v + h, n - h);
}

void standardize(Vec &v) {  // Beam.jl:148-160
    int64_t N = (int64_t)v.size();
    double mu = pairwise_sum(v.data(), N) / (double)N;
    for (auto &a : v) a -= mu;
    Vec sq(N);
    for (int64_t i = 0; i < N; ++i) sq[i] = v[i] * v[i];
    double s2 = pairwise_sum(sq.data(), N) / (double)N;
    double s = std::sqrt(jmax(s2, 0.0));
    if (s == 0.0) {
        for (auto &a : v) a = 0.0;
    } else {
        for (auto &a : v) a /= s;
    }
}

// ------------------------------------------------------------------------------------------
// erfinv (SpecialFunctions.erfinv via SpecialMath.jl:4) -- rational seed + Newton on erf/erfc
// ------------------------------------------------------------------------------------------
double erfinv(double y) {
    if (std::isnan(y) || y < -1.0 || y > 1.0) return NAN;
    if (y == 1.0) return INFINITY;
    if (y == -1.0) return -INFINITY;
    if (y == 0.0) return y;
    // Seed: Giles, "Approximating the erfinv function" (GPU Computing Gems, 2011), single-precision form.
    double w = -std::log((1.0 - y) * (1.0 + y));
    double x;
    if (w < 5.0) {
        w = w - 2.5;
        double p = 2.81022636e-08;
        p = 3.43273939e-07 + p * w; p = -3.5233877e-06 + p * w; p = -4.39150654e-06 + p * w;
        p = 0.00021858087 + p * w; p = -0.00125372503 + p * w; p = -0.00417768164 + p * w;
        p = 0.246640727 + p * w; p = 1.50140941 + p * w;
        x = p * y;
    } else {
        w = std::sqrt(w) - 3.0;
        double p = -0.000200214257;
        p = 0.000100950558 + p * w; p = 0.00134934322 + p * w; p = -0.00367342844 + p * w;
        p = 0.00573950773 + p * w; p = -0.0076224613 + p * w; p = 0.00943887047 + p * w;
        p = 1.00167406 + p * w; p = 2.83297682 + p * w;
        x = p * y;
    }
    const double two_over_sqrtpi = 1.1283791670955126;
    double ay = std::fabs(y);
    for (int it = 0; it < 3; ++it) {
        double ax = std::fabs(x);
        double f;
        if (ay > 0.5) f = (1.0 - ay) - std::erfc(ax);   // erf(ax) - ay, without cancellation
        else f = std::erf(ax) - ay;
        double d = two_over_sqrtpi * std::exp(-ax * ax);
        double step = f / d;
        // Halley correction: f'' / f' = -2x
        step = step / (1.0 + ax * step);
        ax = ax - step;
        x = std::copysign(ax, y);
    }
    return x;
}

// ------------------------------------------------------------------------------------------
// FFT: mixed-radix decimation-in-time DFT (restates FFTW's unnormalised forward / normalised inverse).
// Complex products are the algebraic formula (compiled with -fcx-limited-range), as Julia's
// *(::Complex{Float64}, ::Complex{Float64}).
// ------------------------------------------------------------------------------------------
struct FFTPlan {
    int n = 0;
    std::vector<int> factors;
    std::vector<cd> tw_fwd;   // exp(-2 pi i k / n)
    explicit FFTPlan(int n_) : n(n_) {
        int m = n;
        for (int r : {4, 2, 3, 5}) {
            while (m % r == 0) { factors.push_back(r); m /= r; }
        }
        if (m > 1) factors.push_back(m);
        tw_fwd.resize(n);
        for (int k = 0; k < n; ++k) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
            tw_fwd[k] = cd((double)cosl(a), (double)sinl(a));
        }
    }
    cd tw(int64_t k, int sign) const {   // k < n for every call below
        cd t = tw_fwd[(size_t)k];
        return sign < 0 ? t : std::conj(t);
    }
    void rec(const cd *in, cd *out, int len, int stride, size_t fi, int sign) const {
        if (len == 1) { out[0] = in[0]; return; }
        int r = factors[fi];
        int m = len / r;
        for (int q = 0; q < r; ++q) rec(in + (size_t)q * stride, out + (size_t)q * m, m, stride * r, fi + 1, sign);
        int step = n / len;
        cd t[64];
        std::vector<cd> tbig;
        cd *tt = t;
        if (r > 64) { tbig.resize(r); tt = tbig.data(); }
        const cd imag_unit = sign < 0 ? cd(0.0, -1.0) : cd(0.0, 1.0);   // W_4 for this direction
        for (int k = 0; k < m; ++k) {
            for (int q = 0; q < r; ++q) tt[q] = out[(size_t)q * m + k] * tw((int64_t)q * k * step, sign);
            if (r == 2) {
                out[(size_t)k] = tt[0] + tt[1];
                out[(size_t)k + m] = tt[0] - tt[1];
            } else if (r == 4) {
                cd s02 = tt[0] + tt[2], d02 = tt[0] - tt[2];
                cd s13 = tt[1] + tt[3], d13 = (tt[1] - tt[3]) * imag_unit;
                out[(size_t)k] = s02 + s13;
                out[(size_t)k + m] = d02 + d13;
                out[(size_t)k + 2 * m] = s02 - s13;
                out[(size_t)k + 3 * m] = d02 - d13;
            } else {
                for (int s = 0; s < r; ++s) {
                    cd acc = tt[0];
                    for (int q = 1; q < r; ++q) acc += tt[q] * tw((int64_t)((q * s) % r) * (n / r), sign);
                    out[(size_t)k + (size_t)m * s] = acc;
                }
            }
        }
    }
    // In-place transform of `len` = n points with the given element stride.
    void exec(cd *data, int64_t stride, int sign, std::vector<cd> &scratch_in, std::vector<cd> &scratch_out) const {
        scratch_in.resize(n); scratch_out.resize(n);
        for (int k = 0; k < n; ++k) scratch_in[k] = data[k * stride];
        rec(scratch_in.data(), scratch_out.data(), n, 1, 0, sign);
        for (int k = 0; k < n; ++k) data[k * stride] = scratch_out[k];
    }
};

// Column-major P1 x P2 complex array; full 2D DFT.  sign -1 forward (unnormalised), +1 inverse (1/(P1 P2)).
void fft2(cd *a, int P1, int P2, int sign, const FFTPlan &p1, const FFTPlan &p2) {
    std::vector<cd> si, so;
    for (int j = 0; j < P2; ++j) p1.exec(a + (size_t)j * P1, 1, sign, si, so);
    for (int i = 0; i < P1; ++i) p2.exec(a + i, P1, sign, si, so);
    if (sign > 0) {
        double scale = 1.0 / (double)((int64_t)P1 * (int64_t)P2);
        for (int64_t k = 0; k < (int64_t)P1 * P2; ++k) a[k] *= scale;
    }
}

// ------------------------------------------------------------------------------------------
// Options (PICPoissonSolver, interface.jl:1324-1507)
// ------------------------------------------------------------------------------------------
struct Slicing {
    int nslices = 1, method = OCTOPIC_SLICE_EQUAL_AREA, resolution = 100, center = OCTOPIC_CENTER_CENTROID;
    Vec positions;
};

struct Opt {
    int nx = 128, ny = 128, deposit = OCTOPIC_CIC, green_type = OCTOPIC_GREEN_INTEGRATED;
    int green_cache = OCTOPIC_CACHE_SLICE_PAIR, field_derivative = 2, interp = OCTOPIC_INTERP_LINEAR;
    int igrid = OCTOPIC_GRID_SLICE_PAIR, extent = OCTOPIC_EXTENT_EXTREMA, lkick = 1;
    int lum_nx = 128, lum_ny = 128, lum_deposit = OCTOPIC_CIC;
    double extent_sigma = 6.0, min_x = 0.0, min_y = 0.0, quantize = 0.0, min_ratio = 0.5, growth = 0.25;
    double kbb1 = NAN, kbb2 = NAN, lum_scale = NAN;
    Slicing s1, s2;
    // GaussianPICPoissonSolver (gaussian_pic.jl:87-131)
    int solver = OCTOPIC_SOLVER_PIC, neutralize = 1;
    double margin = 5.0, coupling_tol = INFINITY;
    // GaussianPoissonSolver (interface.jl:876-1010)
    double min_sigma = 0.0;
    int vdrift = OCTOPIC_DRIFT_HIRATA, gwl = 2, ignore1 = 0, ignore2 = 0, sigma_xy = 0;
};

Slicing to_slicing(const octopic_slicing &s) {
    Slicing r;
    r.nslices = s.nslices; r.method = s.method; r.resolution = s.resolution; r.center = s.center_position;
    if (s.positions && s.npositions > 0) r.positions.assign(s.positions, s.positions + s.npositions);
    return r;
}

void validate_slicing(const Slicing &s) {  // interface.jl:816-835
    if (s.nslices <= 0) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "nslices must be positive");
    if (s.resolution <= 0) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "slicing resolution must be positive");
    if (s.method < 0 || s.method > 4) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "unsupported longitudinal slicing method");
    if (s.center != OCTOPIC_CENTER_CENTROID && s.center != OCTOPIC_CENTER_MIDPOINT)
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "center_position must be :centroid or :midpoint");
    if (s.method == OCTOPIC_SLICE_SPECIFIED && (int)s.positions.size() != s.nslices - 1)
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "specified slicing requires nslices-1 internal positions");
}

Opt to_opt(const octopic_options &o) {
    Opt r;
    r.nx = o.nx; r.ny = o.ny; r.deposit = o.deposit; r.green_type = o.green_type; r.green_cache = o.green_cache;
    r.field_derivative = o.field_derivative; r.interp = o.slice_interpolation; r.igrid = o.interaction_grid;
    r.extent = o.grid_extent; r.lkick = o.longitudinal_kick ? 1 : 0;
    r.lum_nx = o.lum_nx > 0 ? o.lum_nx : o.nx;
    r.lum_ny = o.lum_ny > 0 ? o.lum_ny : o.ny;
    r.lum_deposit = o.lum_deposit >= 0 ? o.lum_deposit : o.deposit;
    r.extent_sigma = o.grid_extent_sigma; r.min_x = o.min_extent_x; r.min_y = o.min_extent_y;
    r.quantize = o.grid_quantize; r.min_ratio = o.green_min_ratio; r.growth = o.green_growth;
    r.kbb1 = o.kbb1; r.kbb2 = o.kbb2; r.lum_scale = o.luminosity_scale;
    r.s1 = to_slicing(o.slicing1); r.s2 = to_slicing(o.slicing2);
    r.solver = o.solver; r.neutralize = o.neutralize ? 1 : 0; r.margin = o.margin_sigma; r.coupling_tol = o.coupling_tol;
    r.min_sigma = o.min_sigma; r.vdrift = o.virtual_drift; r.gwl = o.gaussian_when_luminosity;
    r.ignore1 = o.ignore_centroid1 ? 1 : 0; r.ignore2 = o.ignore_centroid2 ? 1 : 0; r.sigma_xy = o.include_sigma_xy ? 1 : 0;
    // Constructor checks (interface.jl:1410-1476) and _validate_pic_solver (pic_cpu.jl:517-576).
    if (!(r.min_ratio > 0.0 && r.min_ratio <= 1.0))
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "slice_pair_green_min_ratio must be in (0, 1]");
    if (!(r.growth >= 0.0)) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "slice_pair_green_growth must be non-negative");
    if (r.nx < 5 || r.ny < 5) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "PICPoissonSolver grid must be at least (5, 5)");
    if (r.deposit != OCTOPIC_CIC && r.deposit != OCTOPIC_TSC) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "deposit_method must be :CIC or :TSC");
    if (r.lum_deposit != OCTOPIC_CIC && r.lum_deposit != OCTOPIC_TSC)
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "luminosity_deposit_method must resolve to :CIC or :TSC");
    if (r.green_type < 0 || r.green_type > 2) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "green_type must be :integrated, :standard or :lattice");
    if (r.green_cache != OCTOPIC_CACHE_NONE && r.green_cache != OCTOPIC_CACHE_SLICE_PAIR)
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "green_cache must be :none or :slice_pair");
    if (r.field_derivative != 2 && r.field_derivative != 4)
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "field_derivative must be :second or :fourth");
    if (r.interp != OCTOPIC_INTERP_LINEAR && r.interp != OCTOPIC_INTERP_QUADRATIC)
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "slice_interpolation must be :linear or :quadratic");
    if (r.igrid < 0 || r.igrid > 2) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "interaction_grid must be :slice_pair, :source_slice or :node");
    if (r.extent != OCTOPIC_EXTENT_EXTREMA && r.extent != OCTOPIC_EXTENT_SIGMA)
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "grid_extent must be :extrema or :sigma");
    if (!(r.extent_sigma > 0.0)) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "grid_extent_sigma must be positive");
    if (!(std::isfinite(r.min_x) && std::isfinite(r.min_y) && r.min_x >= 0.0 && r.min_y >= 0.0))
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "min_transverse_extent values must be finite and nonnegative");
    if (!(r.quantize >= 0.0)) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "grid_quantize must be non-negative");
    if (r.lum_nx < 3 || r.lum_ny < 3) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "luminosity_grid dimensions must both be at least 3");
    if (r.igrid != OCTOPIC_GRID_SLICE_PAIR && r.extent != OCTOPIC_EXTENT_EXTREMA)
        throw OracleError(OCTOPIC_ERR_UNSUPPORTED, "grid_extent is not implemented for this interaction_grid");
    if (r.igrid == OCTOPIC_GRID_NODE && r.interp != OCTOPIC_INTERP_LINEAR)
        throw OracleError(OCTOPIC_ERR_UNSUPPORTED,
                          "slice_interpolation = :quadratic is not implemented for interaction_grid = :node");
    if (r.green_type == OCTOPIC_GREEN_LATTICE)
        throw OracleError(OCTOPIC_ERR_UNSUPPORTED, "green_type = :lattice is not restated by the oracle");
    validate_slicing(r.s1);
    validate_slicing(r.s2);
    if (r.solver != OCTOPIC_SOLVER_PIC && r.solver != OCTOPIC_SOLVER_GAUSSIAN_PIC && r.solver != OCTOPIC_SOLVER_SOFT_GAUSSIAN)
        throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "solver must be one of the OCTOPIC_SOLVER_* values");
    if (r.solver == OCTOPIC_SOLVER_SOFT_GAUSSIAN) {   // GaussianPoissonSolver constructor (interface.jl:976-1000)
        if (!(std::isfinite(r.min_sigma) && r.min_sigma >= 0.0))
            throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "min_sigma must be finite and nonnegative");
        if (r.gwl != 1 && r.gwl != 2) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "gaussian_when_luminosity must be 1 or 2");
        if (r.vdrift < OCTOPIC_DRIFT_HIRATA || r.vdrift > OCTOPIC_DRIFT_UNSAFE_CHROMATIC_FROZEN_ENERGY)
            throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "unknown virtual_drift");
        if (r.sigma_xy)
            throw OracleError(OCTOPIC_ERR_UNSUPPORTED, "include_sigma_xy = true (the coupled soft-Gaussian kick) is not restated");
    }
    if (r.solver == OCTOPIC_SOLVER_GAUSSIAN_PIC) {
        // GaussianPICPoissonSolver constructor (gaussian_pic.jl:94-110) and _require_linear_slice_interpolation
        // (pic_cpu.jl:788-809)
        if (!(r.margin >= 0.0)) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "margin_sigma must be non-negative");
        if (!(r.coupling_tol >= 0.0)) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "coupling_tol must be non-negative");
        if (r.interp != OCTOPIC_INTERP_LINEAR)
            throw OracleError(OCTOPIC_ERR_UNSUPPORTED, "slice_interpolation = :quadratic is not implemented by GaussianPICPoissonSolver");
        if (r.igrid != OCTOPIC_GRID_SLICE_PAIR)
            throw OracleError(OCTOPIC_ERR_UNSUPPORTED, "interaction_grid other than :slice_pair is not implemented by GaussianPICPoissonSolver");
        if (r.extent != OCTOPIC_EXTENT_EXTREMA)
            throw OracleError(OCTOPIC_ERR_UNSUPPORTED, "grid_extent = :sigma is not implemented by GaussianPICPoissonSolver");
        if (std::isfinite(r.coupling_tol))
            throw OracleError(OCTOPIC_ERR_UNSUPPORTED, "the coupled (rotated) subtraction (finite coupling_tol) is not restated");
    }
    return r;
}

// ------------------------------------------------------------------------------------------
// Longitudinal slicing (slicing.jl)
// ------------------------------------------------------------------------------------------
constexpr int64_t FOLD_LANES = 4096;   // _SLICE_FOLD_LANES, slicing.jl:223

struct Slices {
    Vec center, weight, boundary;
    std::vector<std::vector<int64_t>> indices;   // 0-based particle indices
};

struct ZStats { int64_t n; double zmin, zmax, mean, sigma; };

double lane_z_moment(const double *z, int64_t n, double mu, int pw) {  // slicing.jl:227-241
    std::vector<double> acc(FOLD_LANES, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        double v = pw == 2 ? (z[i] - mu) * (z[i] - mu) : z[i];
        acc[i % FOLD_LANES] += v;
    }
    double s = 0.0;
    for (int64_t t = 0; t < FOLD_LANES; ++t) s += acc[t];
    return s;
}

double lane_indexed_sum(const double *z, const std::vector<int64_t> &idx) {  // slicing.jl:245-257
    std::vector<double> acc(FOLD_LANES, 0.0);
    for (size_t k = 0; k < idx.size(); ++k) acc[k % FOLD_LANES] += z[idx[k]];
    double s = 0.0;
    for (int64_t t = 0; t < FOLD_LANES; ++t) s += acc[t];
    return s;
}

ZStats live_z_stats(const double *z, int64_t n) {  // slicing.jl:271-289
    double zmin = INFINITY, zmax = -INFINITY;
    for (int64_t i = 0; i < n; ++i) { zmin = jmin(zmin, z[i]); zmax = jmax(zmax, z[i]); }
    double mu = lane_z_moment(z, n, 0.0, 1) / (double)n;
    double s2 = lane_z_moment(z, n, mu, 2);
    return ZStats{n, zmin, zmax, mu, std::sqrt(jmax(s2 / (double)n, 0.0))};
}

inline int64_t slice_bin(double zi, double zmin, double width, int64_t bins) {  // slicing.jl:172-176
    double d = (zi - zmin) / width;
    if (!(d > -INFINITY && d < INFINITY)) return 0;
    return iclamp((int64_t)std::floor(d) + 1, 1, bins);
}

Slices finish_slices(const double *z, int64_t n, const Slicing &sl, std::vector<std::vector<int64_t>> indices,
                     const Vec &boundaries) {  // slicing.jl:523-559
    for (double b : boundaries)
        if (!std::isfinite(b))
            throw OracleError(OCTOPIC_ERR_NONFINITE, "non-finite particle coordinate(s) detected (longitudinal slicing)");
    int64_t total = n;
    size_t ns = indices.size();
    Slices s;
    s.center.resize(ns); s.weight.resize(ns);
    for (size_t k = 0; k < ns; ++k) {
        const auto &idx = indices[k];
        s.weight[k] = (double)idx.size() / (double)total;
        if (sl.center == OCTOPIC_CENTER_CENTROID)
            s.center[k] = idx.empty() ? (boundaries[k] + boundaries[k + 1]) / 2.0
                                      : lane_indexed_sum(z, idx) / (double)idx.size();
        else
            s.center[k] = (boundaries[k] + boundaries[k + 1]) / 2.0;
    }
    s.boundary = boundaries;
    s.indices = std::move(indices);
    return s;
}

Slices slices_from_boundaries(const double *z, int64_t n, const Slicing &sl, const Vec &b) {  // slicing.jl:431-460
    int64_t ns = (int64_t)b.size() - 1;
    std::vector<std::vector<int64_t>> idx(ns);
    if (ns > 0 && b[0] == b[ns]) {
        for (int64_t i = 0; i < n; ++i) idx[0].push_back(i);
        return finish_slices(z, n, sl, std::move(idx), b);
    }
    for (int64_t i = 0; i < n; ++i) {
        int64_t s = iclamp(searchsortedlast(b, z[i]), 1, ns);
        idx[s - 1].push_back(i);
    }
    return finish_slices(z, n, sl, std::move(idx), b);
}

// range(zmin, zmax; length = ns+1).  Julia evaluates the range in TwicePrecision; this restatement
// uses zmin + k*((zmax - zmin)/ns), which can differ from it in the last ulp (parity unpinned).
Vec linspace(double a, double b, int64_t len) {
    Vec r(len);
    double step = (b - a) / (double)(len - 1);
    for (int64_t i = 0; i < len; ++i) r[i] = a + (double)i * step;
    r[0] = a; r[len - 1] = b;
    return r;
}

Slices longitudinal_slices(const double *z, int64_t n, const Slicing &sl) {  // slicing.jl:51-75
    if (sl.nslices <= 0) throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "nslices must be positive");
    if (n <= 0) throw OracleError(OCTOPIC_ERR_INVALID_ARGUMENT, "longitudinal slicing requires at least one particle");
    int64_t ns = sl.nslices;
    if (sl.method == OCTOPIC_SLICE_EQUAL_AREA) {  // slicing.jl:104-156
        int64_t bins = ns * sl.resolution;
        ZStats st = live_z_stats(z, n);
        if (st.zmin == st.zmax) {
            Vec b(ns + 1, st.zmin);
            std::vector<std::vector<int64_t>> idx(ns);
            for (int64_t i = 0; i < n; ++i) idx[0].push_back(i);
            return finish_slices(z, n, sl, std::move(idx), b);
        }
        double width = (st.zmax - st.zmin) / (double)bins;
        std::vector<int64_t> counts(bins, 0);
        for (int64_t i = 0; i < n; ++i) {
            int64_t bin = slice_bin(z[i], st.zmin, width, bins);
            if (bin != 0) counts[bin - 1] += 1;
        }
        Vec cum(bins);
        int64_t run = 0;
        for (int64_t k = 0; k < bins; ++k) { run += counts[k]; cum[k] = (double)run / (double)std::max<int64_t>(st.n, 1); }
        cum[bins - 1] = 1.0;
        Vec centers(bins);
        for (int64_t k = 0; k < bins; ++k) centers[k] = st.zmin + ((double)(k + 1) - 0.5) * width;
        Vec b(ns + 1);
        b[0] = st.zmin; b[ns] = st.zmax;
        int64_t current = 1;
        for (int64_t s = 1; s <= ns - 1; ++s) {
            double target = (double)s / (double)ns;
            while (current <= bins && cum[current - 1] <= target) current += 1;
            double x1, x2, y1, y2;
            if (current <= 1) { x1 = b[0]; x2 = centers[0]; y1 = 0.0; y2 = cum[0]; }
            else if (current > bins) { x1 = centers[bins - 1]; x2 = b[ns]; y1 = cum[bins - 2]; y2 = 1.0; }
            else { x1 = centers[current - 2]; x2 = centers[current - 1]; y1 = cum[current - 2]; y2 = cum[current - 1]; }
            b[s] = (y2 == y1) ? (x1 + x2) / 2.0 : x2 * (target - y1) / (y2 - y1) + x1 * (target - y2) / (y1 - y2);
        }
        return slices_from_boundaries(z, n, sl, b);
    }
    if (sl.method == OCTOPIC_SLICE_EQUAL_COUNT) {  // slicing.jl:330-365
        std::vector<int64_t> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return jisless(z[a], z[b]); });
        std::vector<std::vector<int64_t>> idx(ns);
        for (int64_t s = 1; s <= ns; ++s) {
            int64_t first = (int64_t)std::floor((double)((s - 1) * n) / (double)ns) + 1;
            int64_t last = (int64_t)std::floor((double)(s * n) / (double)ns);
            for (int64_t p = first; p <= last; ++p) idx[s - 1].push_back(order[p - 1]);
        }
        ZStats st = live_z_stats(z, n);
        Vec b(ns + 1);
        b[0] = st.zmin; b[ns] = st.zmax;
        for (int64_t s = 1; s <= ns - 1; ++s) {
            int64_t pos = (int64_t)std::floor((double)(s * n) / (double)ns);
            if (pos == 0) b[s] = z[order[0]];
            else if (pos == n) b[s] = z[order[n - 1]];
            else b[s] = (z[order[pos - 1]] + z[order[pos]]) / 2.0;
        }
        return finish_slices(z, n, sl, std::move(idx), b);
    }
    if (sl.method == OCTOPIC_SLICE_EQUAL_WIDTH) {  // slicing.jl:367-387
        ZStats st = live_z_stats(z, n);
        Vec b = linspace(st.zmin, st.zmax, ns + 1);
        std::vector<std::vector<int64_t>> idx(ns);
        if (st.zmin == st.zmax) {
            for (int64_t i = 0; i < n; ++i) idx[0].push_back(i);
            return finish_slices(z, n, sl, std::move(idx), b);
        }
        double width = (st.zmax - st.zmin) / (double)ns;
        for (int64_t i = 0; i < n; ++i) {
            int64_t s = slice_bin(z[i], st.zmin, width, ns);
            if (s != 0) idx[s - 1].push_back(i);
        }
        return finish_slices(z, n, sl, std::move(idx), b);
    }
    if (sl.method == OCTOPIC_SLICE_NORMAL_QUANTILE) {  // slicing.jl:405-429
        ZStats st = live_z_stats(z, n);
        if (st.sigma == 0.0) return slices_from_boundaries(z, n, sl, Vec(ns + 1, st.mean));
        Vec b(ns + 1);
        b[0] = st.zmin; b[ns] = st.zmax;
        for (int64_t s = 1; s <= ns - 1; ++s) {
            double q = std::sqrt(2.0) * erfinv((double)(2 * s) / (double)ns - 1.0);
            b[s] = jclamp(st.mean + st.sigma * q, b[0], b[ns]);
        }
        return slices_from_boundaries(z, n, sl, b);
    }
    if (sl.method == OCTOPIC_SLICE_SPECIFIED) {  // slicing.jl:389-403
        ZStats st = live_z_stats(z, n);
        Vec internal;
        for (double p : sl.positions) internal.push_back(st.mean + p * st.sigma);
        std::stable_sort(internal.begin(), internal.end(), jisless);
        Vec b(internal.size() + 2);
        b[0] = st.zmin; b[b.size() - 1] = st.zmax;
        for (size_t k = 0; k < internal.size(); ++k) b[k + 1] = jclamp(internal[k], b[0], b[b.size() - 1]);
        return slices_from_boundaries(z, n, sl, b);
    }
    throw OracleError(OCTOPIC_ERR_INVALID_OPTION, "unknown longitudinal slicing method");
}

struct OrderEntry { double t; int i, j; };

std::vector<OrderEntry> collision_order(const Vec &c1, const Vec &c2) {  // slicing.jl:561-578
    std::vector<OrderEntry> order;
    for (int i = 0; i < (int)c1.size(); ++i)
        for (int j = 0; j < (int)c2.size(); ++j) order.push_back({-(c1[i] + c2[j]) / 2.0, i, j});
    std::stable_sort(order.begin(), order.end(), [](const OrderEntry &a, const OrderEntry &b) { return jisless(a.t, b.t); });
    return order;
}

std::vector<std::vector<OrderEntry>> collision_pair_batches(const Vec &c1, const Vec &c2) {  // slicing.jl:601-647
    auto pairs = collision_order(c1, c2);
    std::vector<std::vector<OrderEntry>> batches;
    if (pairs.empty()) return batches;
    int ns1 = (int)c1.size(), ns2 = (int)c2.size();
    std::vector<std::vector<int>> by_i(ns1), by_j(ns2);
    for (int k = 0; k < (int)pairs.size(); ++k) { by_i[pairs[k].i].push_back(k); by_j[pairs[k].j].push_back(k); }
    std::vector<int> next_i(ns1, 0), next_j(ns2, 0);
    std::vector<char> done(pairs.size(), 0);
    int remaining = (int)pairs.size();
    while (remaining > 0) {
        std::vector<OrderEntry> cur;
        std::vector<char> used_i(ns1, 0), used_j(ns2, 0);
        for (int k = 0; k < (int)pairs.size(); ++k) {
            if (done[k]) continue;
            const auto &p = pairs[k];
            if (used_i[p.i] || used_j[p.j]) continue;
            bool ri = next_i[p.i] < (int)by_i[p.i].size() && by_i[p.i][next_i[p.i]] == k;
            bool rj = next_j[p.j] < (int)by_j[p.j].size() && by_j[p.j][next_j[p.j]] == k;
            if (ri && rj) {
                cur.push_back(p); used_i[p.i] = 1; used_j[p.j] = 1; done[k] = 1;
                next_i[p.i] += 1; next_j[p.j] += 1; remaining -= 1;
            }
        }
        if (cur.empty()) throw OracleError(OCTOPIC_ERR_INVALID_ARGUMENT, "internal collision scheduler error");
        batches.push_back(std::move(cur));
    }
    return batches;
}

// ------------------------------------------------------------------------------------------
// Deposition weights (pic_cpu.jl:2182-2252)
// ------------------------------------------------------------------------------------------
// base is the 1-based first stencil node, as in the reference.
inline void cic_weights(double u, int64_t n, int64_t &base, double w[3]) {
    if (!(0.0 <= u && u <= (double)(n - 1))) { base = 1; w[0] = w[1] = 0.0; w[2] = 0.0; return; }
    base = iclamp((int64_t)std::floor(u) + 1, 1, n - 1);
    double f = jclamp(u - (double)(base - 1), 0.0, 1.0);
    w[0] = 1.0 - f; w[1] = f; w[2] = 0.0;
}
inline void tsc_weights(double u, int64_t n, int64_t &base, double w[3]) {
    if (!(0.0 <= u && u <= (double)(n - 1))) { base = 1; w[0] = w[1] = w[2] = 0.0; return; }
    int64_t ix = (int64_t)std::floor(u);
    double f = u - std::floor(u);
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
    base = iclamp(base, 1, n - 2);
}
inline int weights(int method, double u, int64_t n, int64_t &base, double w[3]) {
    if (method == OCTOPIC_CIC) { cic_weights(u, n, base, w); return 2; }
    tsc_weights(u, n, base, w);
    return 3;
}

// ------------------------------------------------------------------------------------------
// Green function (pic_cpu.jl:2254-2272, 2451-2509)
// ------------------------------------------------------------------------------------------
inline double atan_ratio(double num, double den) {
    if (den == 0.0) {
        if (num == 0.0) return 0.0;
        return std::copysign(M_PI / 2.0, num);
    }
    return std::atan(num / den);
}
inline double kernel_integral(double x, double y) {
    double r2 = x * x + y * y;
    if (r2 == 0.0) return 0.0;
    return (std::log(r2) - 3.0) * x * y + atan_ratio(y, x) * x * x + atan_ratio(x, y) * y * y;
}

// green: column-major 2nx x 2ny
void green_table(double *green, int gtype, double fx0, double fy0, double sx0, double sy0, double hx, double hy,
                 int nx, int ny) {
    double half_hx = hx / 2.0, half_hy = hy / 2.0;
    double hxihyi = -0.5 / (hx * hy);
    bool integrated = gtype == OCTOPIC_GREEN_INTEGRATED;
    int64_t P1 = 2 * (int64_t)nx;
    for (int64_t j = 0; j < 2 * (int64_t)ny; ++j) {
        for (int64_t i = 0; i < P1; ++i) {
            int64_t ii = i < nx ? i : i - 2 * (int64_t)nx;
            int64_t jj = j < ny ? j : j - 2 * (int64_t)ny;
            double x = fx0 - sx0 + (double)ii * hx;
            double y = fy0 - sy0 + (double)jj * hy;
            double val;
            if (integrated || (x * x + y * y) == 0.0) {
                val = kernel_integral(x + half_hx, y + half_hy);
                val += kernel_integral(x - half_hx, y - half_hy);
                val -= kernel_integral(x + half_hx, y - half_hy);
                val -= kernel_integral(x - half_hx, y + half_hy);
                green[i + j * P1] = hxihyi * val;
            } else {
                green[i + j * P1] = -0.5 * std::log(x * x + y * y);
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Field derivative (pic_cpu.jl:2531-2591), column-major nx x ny
// ------------------------------------------------------------------------------------------
void field_derivative(double *Ex, double *Ey, const double *phi, int nx, int ny, double hx, double hy, bool fourth) {
    auto P = [&](int i, int j) { return phi[(i - 1) + (int64_t)(j - 1) * nx]; };
    auto EX = [&](int i, int j) -> double & { return Ex[(i - 1) + (int64_t)(j - 1) * nx]; };
    auto EY = [&](int i, int j) -> double & { return Ey[(i - 1) + (int64_t)(j - 1) * nx]; };
    double hxi = 1.0 / hx, hyi = 1.0 / hy;
    double c4 = 1.0 / 12.0;
    for (int i = 1; i <= nx; ++i) {
        EY(i, 1) = hyi * (1.5 * P(i, 1) - 2.0 * P(i, 2) + 0.5 * P(i, 3));
        EY(i, ny) = hyi * (-1.5 * P(i, ny) + 2.0 * P(i, ny - 1) - 0.5 * P(i, ny - 2));
    }
    if (fourth && ny >= 5) {
        for (int i = 1; i <= nx; ++i) {
            EY(i, 2) = 0.5 * hyi * (P(i, 1) - P(i, 3));
            EY(i, ny - 1) = 0.5 * hyi * (P(i, ny - 2) - P(i, ny));
        }
        for (int j = 3; j <= ny - 2; ++j)
            for (int i = 1; i <= nx; ++i)
                EY(i, j) = c4 * hyi * ((P(i, j + 2) - P(i, j - 2)) + 8.0 * (P(i, j - 1) - P(i, j + 1)));
    } else {
        for (int j = 2; j <= ny - 1; ++j)
            for (int i = 1; i <= nx; ++i) EY(i, j) = 0.5 * hyi * (P(i, j - 1) - P(i, j + 1));
    }
    for (int j = 1; j <= ny; ++j) {
        EX(1, j) = hxi * (1.5 * P(1, j) - 2.0 * P(2, j) + 0.5 * P(3, j));
        EX(nx, j) = hxi * (-1.5 * P(nx, j) + 2.0 * P(nx - 1, j) - 0.5 * P(nx - 2, j));
        if (fourth && nx >= 5) {
            EX(2, j) = 0.5 * hxi * (P(1, j) - P(3, j));
            EX(nx - 1, j) = 0.5 * hxi * (P(nx - 2, j) - P(nx, j));
            for (int i = 3; i <= nx - 2; ++i)
                EX(i, j) = c4 * hxi * ((P(i + 2, j) - P(i - 2, j)) + 8.0 * (P(i - 1, j) - P(i + 1, j)));
        } else {
            for (int i = 2; i <= nx - 1; ++i) EX(i, j) = 0.5 * hxi * (P(i - 1, j) - P(i + 1, j));
        }
    }
}

// ------------------------------------------------------------------------------------------
// Grids (pic_cpu.jl:1634-1718, 1780-1905)
// ------------------------------------------------------------------------------------------
struct Grid { double x0, y0, width, height; };
struct Box { double xmin, xmax, ymin, ymax; };

void resolve_extent(const Opt &o, double xspan, double yspan, double &w, double &h, const char *ctx) {
    w = jmax(xspan, o.min_x);
    h = jmax(yspan, o.min_y);
    if (!(std::isfinite(w) && std::isfinite(h) && w > 0.0 && h > 0.0)) {
        char buf[512];
        snprintf(buf, sizeof buf,
                 "%s requires finite, positive transverse extents; observed (x=%g, y=%g) with "
                 "min_transverse_extent=(%g, %g). Supply a positive physical bound for every degenerate axis.",
                 ctx, xspan, yspan, o.min_x, o.min_y);
        throw OracleError(OCTOPIC_ERR_DEGENERATE_EXTENT, buf);
    }
}

inline double quantize_extent(double w, double q) {
    if (!(q > 0.0 && std::isfinite(w) && w > 0.0)) return w;
    return std::exp2(std::ceil(std::log2(w) / q) * q);
}

inline void align_origins(int gtype, double s0, double f0, double h, double &so, double &fo) {
    double f1 = s0 / h - std::floor(s0 / h);
    double f2 = f0 / h - std::floor(f0 / h);
    double t = (f2 - f1) / 2.0;
    if (gtype == OCTOPIC_GREEN_STANDARD) {
        t += t > 0.0 ? -0.25 : 0.25;
    } else {
        if (t > 0.5) t -= 0.5;
        else if (t < -0.5) t += 0.5;
    }
    double shift = t * h;
    so = s0 + shift;
    fo = f0 - shift;
}

void interaction_grids(const Opt &o, double sxmin, double sxmax, double symin, double symax, double fxmin,
                       double fxmax, double fymin, double fymax, Grid &sg, Grid &fg) {
    double width, height;
    resolve_extent(o, jmax(sxmax - sxmin, fxmax - fxmin), jmax(symax - symin, fymax - fymin), width, height,
                   "PIC interaction grid");
    double tx = width / (double)(o.nx - 4);
    double ty = height / (double)(o.ny - 4);
    width += 3.0 * tx;
    height += 3.0 * ty;
    double q = o.quantize;
    if (q > 0.0) {
        width = quantize_extent(width, q);
        height = quantize_extent(height, q);
        double hx = width / (double)(o.nx - 1);
        double hy = height / (double)(o.ny - 1);
        double sx0 = std::nearbyint(((sxmin + sxmax) / 2.0 - width / 2.0) / hx) * hx;
        double sy0 = std::nearbyint(((symin + symax) / 2.0 - height / 2.0) / hy) * hy;
        double fx0 = std::nearbyint(((fxmin + fxmax) / 2.0 - width / 2.0) / hx) * hx;
        double fy0 = std::nearbyint(((fymin + fymax) / 2.0 - height / 2.0) / hy) * hy;
        align_origins(o.green_type, sx0, fx0, hx, sx0, fx0);
        align_origins(o.green_type, sy0, fy0, hy, sy0, fy0);
        sg = {sx0, sy0, width, height};
        fg = {fx0, fy0, width, height};
        return;
    }
    double sx0 = sxmin - 1.5 * tx, sy0 = symin - 1.5 * ty;
    double fx0 = fxmin - 1.5 * tx, fy0 = fymin - 1.5 * ty;
    align_origins(o.green_type, sx0, fx0, tx, sx0, fx0);
    align_origins(o.green_type, sy0, fy0, ty, sy0, fy0);
    sg = {sx0, sy0, width, height};
    fg = {fx0, fy0, width, height};
}

inline bool grid_size_usable(const Grid &c, const Grid &r, double min_ratio) {
    if (!(c.width >= r.width)) return false;
    if (!(c.height >= r.height)) return false;
    if (!(r.width >= min_ratio * c.width)) return false;
    if (!(r.height >= min_ratio * c.height)) return false;
    return true;
}
inline bool grid_covers(const Opt &o, const Grid &g, const Box &b) {
    double hx = g.width / (double)(o.nx - 1);
    double hy = g.height / (double)(o.ny - 1);
    double margin = 1.5;
    if (!(g.x0 + margin * hx <= b.xmin)) return false;
    if (!(g.x0 + g.width - margin * hx >= b.xmax)) return false;
    if (!(g.y0 + margin * hy <= b.ymin)) return false;
    if (!(g.y0 + g.height - margin * hy >= b.ymax)) return false;
    return true;
}
inline Grid expand_grid(const Grid &g, double factor) {
    if (factor <= 1.0) return g;
    double nw = g.width * factor, nh = g.height * factor;
    return Grid{g.x0 - (nw - g.width) / 2.0, g.y0 - (nh - g.height) / 2.0, nw, nh};
}

// ------------------------------------------------------------------------------------------
// Workspace, caches
// ------------------------------------------------------------------------------------------
struct FieldWS { Vec phi, Ex, Ey; };
struct Plans {
    std::unique_ptr<FFTPlan> p1, p2;
    Plans(int P1, int P2) : p1(new FFTPlan(P1)), p2(new FFTPlan(P2)) {}
};

struct Workspace {
    int nx, ny;
    Vec charge;
    std::vector<cd> spectral;
    Vec green;
    std::vector<Vec> local_charge;
    FieldWS left, right, mid;
    Vec q1, q2;
    int64_t dropped = 0;
    Workspace(int nx_, int ny_, int lnx, int lny) : nx(nx_), ny(ny_) {
        size_t P = (size_t)4 * nx * ny;
        charge.assign(P, 0.0);
        spectral.assign(P, cd(0, 0));
        green.assign(P, 0.0);
        for (FieldWS *f : {&left, &right, &mid}) {
            f->phi.assign((size_t)nx * ny, 0.0); f->Ex.assign((size_t)nx * ny, 0.0); f->Ey.assign((size_t)nx * ny, 0.0);
        }
        q1.assign((size_t)(lnx + 1) * (lny + 1), 0.0);
        q2.assign((size_t)(lnx + 1) * (lny + 1), 0.0);
    }
};

struct GreenEntry {
    Grid sg, fg;
    std::vector<cd> gfft;
    int64_t uses = 0, rebuilds = 0;
};

struct GreenCache {
    std::map<std::tuple<int, int, int>, GreenEntry> entries;
    int64_t hits = 0, misses = 0, rebuilds = 0;
    std::mutex lock;
};

struct NodeEntry { Grid sg, fg; std::vector<cd> gfft; };

struct Trace {
    std::vector<octopic_interaction_record> recs;   // 2 per pair, indexed by 2*p + dir - 1
};

// ------------------------------------------------------------------------------------------
// Deposit (pic_cpu.jl:1992-2163)
// ------------------------------------------------------------------------------------------
constexpr int64_t PIC_DEPOSIT_CHUNKS = 16;
inline bool deposit_parallel(int64_t n, int64_t nx, int64_t ny) { return n >= 4096 && n >= 160 * nx * ny; }

// charge: column-major with leading dimension ld
void deposit_range(double *charge, int64_t ld, int method, const double *x, const double *px, const double *y,
                   const double *py, double s, bool drifted, double x0, double y0, double hx, double hy, int64_t nx,
                   int64_t ny, int64_t first, int64_t last) {
    double hxi = 1.0 / hx, hyi = 1.0 / hy;
    for (int64_t i = first; i < last; ++i) {
        double xd = drifted ? x[i] + px[i] * s : x[i];
        double yd = drifted ? y[i] + py[i] * s : y[i];
        int64_t ix, iy;
        double wx[3], wy[3];
        int kx = weights(method, (xd - x0) * hxi, nx, ix, wx);
        int ky = weights(method, (yd - y0) * hyi, ny, iy, wy);
        for (int m = 0; m < kx; ++m)
            for (int q = 0; q < ky; ++q) charge[(ix - 1 + m) + (iy - 1 + q) * ld] += wx[m] * wy[q];
    }
}

void deposit(double *charge, int64_t ld, int64_t ncols, int method, const double *x, const double *px,
             const double *y, const double *py, double s, bool drifted, double x0, double y0, double hx, double hy,
             int64_t nx, int64_t ny, int64_t n, std::vector<Vec> *local) {
    if (!deposit_parallel(n, nx, ny)) {
        deposit_range(charge, ld, method, x, px, y, py, s, drifted, x0, y0, hx, hy, nx, ny, 0, n);
        return;
    }
    size_t sz = (size_t)ld * ncols;
    std::vector<Vec> own;
    std::vector<Vec> &loc = local ? *local : own;
    loc.resize(PIC_DEPOSIT_CHUNKS);
    for (int64_t c = 1; c <= PIC_DEPOSIT_CHUNKS; ++c) {
        Vec &g = loc[c - 1];
        g.assign(sz, 0.0);
        int64_t first = fld((c - 1) * n, PIC_DEPOSIT_CHUNKS);   // 0-based start = first_i - 1
        int64_t last = fld(c * n, PIC_DEPOSIT_CHUNKS);
        deposit_range(g.data(), ld, method, x, px, y, py, s, drifted, x0, y0, hx, hy, nx, ny, first, last);
    }
    for (int64_t c = 0; c < PIC_DEPOSIT_CHUNKS; ++c)
        for (size_t k = 0; k < sz; ++k) charge[k] += loc[c][k];
}

// ------------------------------------------------------------------------------------------
// The collision
// ------------------------------------------------------------------------------------------
struct Coords { Vec x, px, y, py, z, pz; size_t size() const { return x.size(); } };
struct SliceParam { double weight, lb, center, rb; };

struct Context {
    Opt o;
    int gaussian_replace = 0;   // test mode: each field deposit replaced by its neutralised drifted Gaussian
    Plans plans;
    std::vector<std::unique_ptr<Workspace>> ws;
    GreenCache cache;
    Trace *trace = nullptr;
    Context(const Opt &o_) : o(o_), plans(2 * o_.nx, 2 * o_.ny) {}
};

std::vector<cd> green_fft(Context &C, Workspace &w, const Grid &sg, const Grid &fg) {  // pic_cpu.jl:1917-1939
    const Opt &o = C.o;
    double hx = sg.width / (double)(o.nx - 1);
    double hy = sg.height / (double)(o.ny - 1);
    green_table(w.green.data(), o.green_type, fg.x0, fg.y0, sg.x0, sg.y0, hx, hy, o.nx, o.ny);
    std::vector<cd> g(w.green.size());
    for (size_t k = 0; k < g.size(); ++k) g[k] = cd(w.green[k], 0.0);
    fft2(g.data(), 2 * o.nx, 2 * o.ny, -1, *C.plans.p1, *C.plans.p2);
    return g;
}

const std::vector<cd> *slice_pair_green(Context &C, Workspace &w, const std::tuple<int, int, int> *key, Grid &sg, Grid &fg,
                                        const Box &sb, const Box &fb, std::vector<cd> &tmp, int &event) {  // pic_cpu.jl:1729-1778
    const Opt &o = C.o;
    event = 0;
    if (o.green_cache == OCTOPIC_CACHE_SLICE_PAIR && key) {
        bool usable = false;
        bool have = false;
        {
            std::lock_guard<std::mutex> g(C.cache.lock);
            auto it = C.cache.entries.find(*key);
            have = it != C.cache.entries.end();
            if (have) {
                const GreenEntry &e = it->second;
                usable = grid_size_usable(e.sg, sg, o.min_ratio) && grid_size_usable(e.fg, fg, o.min_ratio) &&
                         grid_covers(o, e.sg, sb) && grid_covers(o, e.fg, fb);
                if (usable) {
                    C.cache.hits += 1;
                    it->second.uses += 1;
                    sg = e.sg; fg = e.fg;
                    event = 1;
                    return &it->second.gfft;
                }
            }
        }
        Grid es = expand_grid(sg, 1.0 + o.growth);
        Grid ef = expand_grid(fg, 1.0 + o.growth);
        if (o.green_type != OCTOPIC_GREEN_INTEGRATED) {   // _pic_realign_expanded_grids, pic_cpu.jl:1849-1858
            double hx = es.width / (double)(o.nx - 1);
            double hy = es.height / (double)(o.ny - 1);
            double sx0, fx0, sy0, fy0;
            align_origins(o.green_type, es.x0, ef.x0, hx, sx0, fx0);
            align_origins(o.green_type, es.y0, ef.y0, hy, sy0, fy0);
            es.x0 = sx0; es.y0 = sy0; ef.x0 = fx0; ef.y0 = fy0;
        }
        std::vector<cd> g = green_fft(C, w, es, ef);
        std::lock_guard<std::mutex> gl(C.cache.lock);
        GreenEntry &e = C.cache.entries[*key];
        if (!have) {
            C.cache.misses += 1;
            e.rebuilds = 0;
            event = 2;
        } else {
            C.cache.rebuilds += 1;
            e.rebuilds += 1;
            event = 3;
        }
        e.sg = es; e.fg = ef; e.gfft = std::move(g); e.uses = 1;
        sg = es; fg = ef;
        return &e.gfft;
    }
    tmp = green_fft(C, w, sg, fg);
    return &tmp;
}

// _gpic_gaussian_profile! (gaussian_pic.jl:206-244): node integrals of the unit Gaussian against the
// CIC tent / TSC spline, node i at x0 + i h.
void gaussian_profile(double *g, int n, double x0, double h, double mu, double s, int method) {
    const double invroot = 1.0 / (s * std::sqrt(2.0)), norm = 1.0 / (s * std::sqrt(2.0 * M_PI));
    auto pdf = [&](double x) { return norm * std::exp(-((x - mu) * (x - mu)) / (2 * s * s)); };
    auto m0 = [&](double A, double B) { return 0.5 * (std::erf((B - mu) * invroot) - std::erf((A - mu) * invroot)); };
    auto m1 = [&](double A, double B, double xi) { return (mu - xi) * m0(A, B) - s * s * (pdf(B) - pdf(A)); };
    auto m2 = [&](double A, double B, double xi) {
        return (s * s + (mu - xi) * (mu - xi)) * m0(A, B) - s * s * ((B - mu) * pdf(B) - (A - mu) * pdf(A)) -
               2 * (mu - xi) * s * s * (pdf(B) - pdf(A));
    };
    const double half = h / 2;
    for (int i = 0; i < n; ++i) {
        const double xi = x0 + i * h;
        if (method == OCTOPIC_CIC) {
            g[i] = m0(xi - h, xi + h) + (m1(xi - h, xi, xi) - m1(xi, xi + h, xi)) / h;
        } else {
            const double cL = xi - half, cR = xi + half, lL = xi - 3 * half, lR = xi - half, rL = xi + half, rR = xi + 3 * half;
            g[i] = 0.75 * m0(cL, cR) - m2(cL, cR, xi) / (h * h) + 1.125 * (m0(lL, lR) + m0(rL, rR)) +
                   1.5 * (m1(lL, lR, xi) - m1(rL, rR, xi)) / h + 0.5 * (m2(lL, lR, xi) + m2(rL, rR, xi)) / (h * h);
        }
    }
}

// The test mode's replacement of one deposit: the source's moments (_gpic_source_moments,
// gaussian_pic.jl:374-435, uncoupled), drifted by s (_gpic_drifted_gaussian, :482-491), its profile
// scaled to the deposited total (the neutralised amplitude, :561-566).  Sources with fewer than two
// members or no positive width keep their deposit (the plain-PIC fallback, :525-530).
void replace_by_gaussian(double *charge, int64_t ld, const Coords &src, double s, const Grid &sg, int nx, int ny, int method) {
    const int64_t n = (int64_t)src.size();
    if (n < 2) return;
    const double x0 = src.x[0], px0 = src.px[0], y0 = src.y[0], py0 = src.py[0];
    double sx = 0, spx = 0, sy = 0, spy = 0, sxx = 0, spxpx = 0, syy = 0, spypy = 0, sxpx = 0, sypy = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double dx = src.x[i] - x0, dpx = src.px[i] - px0, dy = src.y[i] - y0, dpy = src.py[i] - py0;
        sx += dx; spx += dpx; sy += dy; spy += dpy;
        sxx += dx * dx; spxpx += dpx * dpx; syy += dy * dy; spypy += dpy * dpy;
        sxpx += dx * dpx; sypy += dy * dpy;
    }
    const double invn = 1.0 / (double)n;
    const double dmx = sx * invn, dmpx = spx * invn, dmy = sy * invn, dmpy = spy * invn;
    const double varx = std::max(sxx * invn - dmx * dmx, 0.0), varpx = std::max(spxpx * invn - dmpx * dmpx, 0.0);
    const double vary = std::max(syy * invn - dmy * dmy, 0.0), varpy = std::max(spypy * invn - dmpy * dmpy, 0.0);
    const double cxpx = sxpx * invn - dmx * dmpx, cypy = sypy * invn - dmy * dmpy;
    const double mux = (x0 + dmx) + (px0 + dmpx) * s, muy = (y0 + dmy) + (py0 + dmpy) * s;
    const double vx = varx + 2 * s * cxpx + s * s * varpx, vy = vary + 2 * s * cypy + s * s * varpy;
    const double sigx = std::sqrt(std::max(vx, 0.0)), sigy = std::sqrt(std::max(vy, 0.0));
    if (!(sigx > 0 && sigy > 0 && std::isfinite(sigx) && std::isfinite(sigy))) return;
    const double hx = sg.width / (double)(nx - 1), hy = sg.height / (double)(ny - 1);
    std::vector<double> gx(nx), gy(ny);
    gaussian_profile(gx.data(), nx, sg.x0, hx, mux, sigx, method);
    gaussian_profile(gy.data(), ny, sg.y0, hy, muy, sigy, method);
    double q = 0, sgx = 0, sgy = 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) q += charge[i + (size_t)j * ld];
    for (double v : gx) sgx += v;
    for (double v : gy) sgy += v;
    if (!(sgx * sgy > 0)) return;
    const double amp = q / (sgx * sgy);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) charge[i + (size_t)j * ld] = amp * gx[i] * gy[j];
}

void solve_drifted_field(Context &C, Workspace &w, FieldWS &f, const Coords &src, double s, const Grid &sg,
                         const std::vector<cd> &gfft) {  // pic_cpu.jl:1965-1990
    const Opt &o = C.o;
    int nx = o.nx, ny = o.ny;
    int64_t P1 = 2 * nx, P2 = 2 * ny;
    double hx = sg.width / (double)(nx - 1);
    double hy = sg.height / (double)(ny - 1);
    std::fill(w.charge.begin(), w.charge.end(), 0.0);
    deposit(w.charge.data(), P1, P2, o.deposit, src.x.data(), src.px.data(), src.y.data(), src.py.data(), s, true,
            sg.x0, sg.y0, hx, hy, nx, ny, (int64_t)src.size(), &w.local_charge);
    if (C.gaussian_replace) replace_by_gaussian(w.charge.data(), P1, src, s, sg, nx, ny, o.deposit);
    for (size_t k = 0; k < w.charge.size(); ++k) w.spectral[k] = cd(w.charge[k], 0.0);
    fft2(w.spectral.data(), (int)P1, (int)P2, -1, *C.plans.p1, *C.plans.p2);
    for (size_t k = 0; k < w.spectral.size(); ++k) w.spectral[k] *= gfft[k];
    fft2(w.spectral.data(), (int)P1, (int)P2, +1, *C.plans.p1, *C.plans.p2);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) f.phi[i + (size_t)j * nx] = w.spectral[i + (size_t)j * P1].real();
    field_derivative(f.Ex.data(), f.Ey.data(), f.phi.data(), nx, ny, hx, hy, o.field_derivative == 4);
}

// _pic_interpolate_kick (pic_cpu.jl:2596-2619); planes are column-major nx x ny
inline void interpolate_kick(const Opt &o, const Grid &g, double x, double y, const double *phiL, const double *ExL,
                             const double *EyL, const double *phiR, const double *ExR, const double *EyR, double zL,
                             double zR, double &Kx, double &Ky, double &Kz) {
    int nx = o.nx, ny = o.ny;
    double hx = g.width / (double)(nx - 1);
    double hy = g.height / (double)(ny - 1);
    int64_t ix, iy;
    double wx[3], wy[3];
    int kx = weights(o.deposit, (x - g.x0) / hx, nx, ix, wx);
    int ky = weights(o.deposit, (y - g.y0) / hy, ny, iy, wy);
    Kx = 0.0; Ky = 0.0; Kz = 0.0;
    for (int m = 0; m < kx; ++m)
        for (int q = 0; q < ky; ++q) {
            size_t c = (size_t)(ix - 1 + m) + (size_t)(iy - 1 + q) * nx;
            double wgt = wx[m] * wy[q];
            Kx += wgt * (zL * ExL[c] + zR * ExR[c]);
            Ky += wgt * (zL * EyL[c] + zR * EyR[c]);
            Kz += wgt * (phiL[c] - phiR[c]);
        }
}
inline void interpolate_kick(const Opt &o, const Grid &g, double x, double y, const FieldWS &L, const FieldWS &R,
                             double zL, double zR, double &Kx, double &Ky, double &Kz) {
    interpolate_kick(o, g, x, y, L.phi.data(), L.Ex.data(), L.Ey.data(), R.phi.data(), R.Ex.data(), R.Ey.data(), zL,
                     zR, Kx, Ky, Kz);
}

// _pic_interpolate_kick_quadratic (pic_cpu.jl:2648-2680)
inline void interpolate_kick_quadratic(const Opt &o, const Grid &g, double x, double y, const FieldWS &L,
                                       const FieldWS &M, const FieldWS &R, double t, double &Kx, double &Ky,
                                       double &Kz) {
    int nx = o.nx, ny = o.ny;
    double hx = g.width / (double)(nx - 1);
    double hy = g.height / (double)(ny - 1);
    int64_t ix, iy;
    double wx[3], wy[3];
    int kx = weights(o.deposit, (x - g.x0) / hx, nx, ix, wx);
    int ky = weights(o.deposit, (y - g.y0) / hy, ny, iy, wy);
    double t2 = t * t;
    double aL = 2.0 * t2 - 3.0 * t + 1.0;
    double aM = 4.0 * t - 4.0 * t2;
    double aR = 2.0 * t2 - t;
    double bL = 3.0 - 4.0 * t;
    double bM = 8.0 * t - 4.0;
    double bR = 1.0 - 4.0 * t;
    Kx = 0.0; Ky = 0.0; Kz = 0.0;
    for (int m = 0; m < kx; ++m)
        for (int q = 0; q < ky; ++q) {
            size_t c = (size_t)(ix - 1 + m) + (size_t)(iy - 1 + q) * nx;
            double wgt = wx[m] * wy[q];
            Kx += wgt * (aL * L.Ex[c] + aM * M.Ex[c] + aR * R.Ex[c]);
            Ky += wgt * (aL * L.Ey[c] + aM * M.Ey[c] + aR * R.Ey[c]);
            Kz += wgt * (bL * L.phi[c] + bM * M.phi[c] + bR * R.phi[c]);
        }
}

inline void slice_interp_params(double lb, double rb, double &hzi, double &zbias) {  // pic_cpu.jl:1540-1548
    double denom = rb - lb;
    if (!std::isfinite(denom) || denom == 0.0) { hzi = 0.0; zbias = 0.5; return; }
    hzi = 1.0 / denom;
    zbias = rb * hzi;
}

inline void axis_extent(int method, double &lo, double &hi, double origin, double ds, double ds2, int64_t n, double k) {
    if (method == OCTOPIC_EXTENT_EXTREMA) return;     // pic_cpu.jl:1581-1593
    if (n < 2) return;
    double invn = 1.0 / (double)n;
    double mean_offset = ds * invn;
    double m = origin + mean_offset;
    double var = jmax(std::fma(-mean_offset, mean_offset, ds2 * invn), 0.0);   // muladd
    double sd = std::sqrt(var);
    double half = k * sd;
    if (!(half > 0.0)) return;
    lo = m - half;
    hi = m + half;
}

int64_t count_outside_box(const Vec &xs, const Vec &ys, double xlo, double xhi, double ylo, double yhi) {
    int64_t c = 0;
    for (size_t i = 0; i < xs.size(); ++i) {
        double x = xs[i], y = ys[i];
        if (!(std::isfinite(x) && xlo <= x && x <= xhi && std::isfinite(y) && ylo <= y && y <= yhi)) c += 1;
    }
    return c;
}
int64_t count_outside_box_drifted(const Coords &s, double sL, double sR, double xlo, double xhi, double ylo, double yhi) {
    int64_t c = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        double xl = s.x[i] + s.px[i] * sL, yl = s.y[i] + s.py[i] * sL;
        double xr = s.x[i] + s.px[i] * sR, yr = s.y[i] + s.py[i] * sR;
        if (!(std::isfinite(xl) && xlo <= xl && xl <= xhi && std::isfinite(yl) && ylo <= yl && yl <= yhi &&
              std::isfinite(xr) && xlo <= xr && xr <= xhi && std::isfinite(yr) && ylo <= yr && yr <= yhi))
            c += 1;
    }
    return c;
}

[[noreturn]] void nonfinite(const char *role, const char *ctx) {
    char buf[512];
    snprintf(buf, sizeof buf, "non-finite particle coordinate(s) detected in the %s particles of a strong-strong collision (%s)",
             role, ctx);
    throw OracleError(OCTOPIC_ERR_NONFINITE, buf);
}

struct UnionBounds { Box s, f; };

// _pic_interaction! (pic_cpu.jl:954-1144).  Returns the virtual source positions (vx, vy).
void interaction(Context &C, Workspace &w, const Coords &src, const SliceParam &ps, Coords &fld, const SliceParam &pf,
                 double kbb, const std::tuple<int, int, int> *key, const UnionBounds *ov, Vec &vx, Vec &vy,
                 octopic_interaction_record *rec) {
    const Opt &o = C.o;
    int64_t nsource = (int64_t)src.size(), nfield = (int64_t)fld.size();
    double sL = 0.5 * (ps.center - pf.lb);
    double sR = 0.5 * (ps.center - pf.rb);
    double sxl = src.x[0] + src.px[0] * sL, syl = src.y[0] + src.py[0] * sL;
    double sxr = src.x[0] + src.px[0] * sR, syr = src.y[0] + src.py[0] * sR;
    double sxmin = jmin(sxl, sxr), sxmax = jmax(sxl, sxr), symin = jmin(syl, syr), symax = jmax(syl, syr);
    int ge = o.extent;
    double sx0 = sxl, sy0 = syl;
    double sxs = 0, sxs2 = 0, sys = 0, sys2 = 0;
    for (int64_t i = 0; i < nsource; ++i) {
        double xl = src.x[i] + src.px[i] * sL, yl = src.y[i] + src.py[i] * sL;
        double xr = src.x[i] + src.px[i] * sR, yr = src.y[i] + src.py[i] * sR;
        sxmin = jmin(jmin(sxmin, xl), xr); sxmax = jmax(jmax(sxmax, xl), xr);
        symin = jmin(jmin(symin, yl), yr); symax = jmax(jmax(symax, yl), yr);
        if (ge != OCTOPIC_EXTENT_EXTREMA) {
            double dxl = xl - sx0, dxr = xr - sx0, dyl = yl - sy0, dyr = yr - sy0;
            sxs += dxl + dxr; sxs2 += dxl * dxl + dxr * dxr;
            sys += dyl + dyr; sys2 += dyl * dyl + dyr * dyr;
        }
    }
    double kext = o.extent_sigma;
    axis_extent(ge, sxmin, sxmax, sx0, sxs, sxs2, 2 * nsource, kext);
    axis_extent(ge, symin, symax, sy0, sys, sys2, 2 * nsource, kext);
    if (!(std::isfinite(sxmin) && std::isfinite(sxmax) && std::isfinite(symin) && std::isfinite(symax)))
        nonfinite("source", "slice pair");

    double c = ps.center;
    double fxmin = fld.x[0] + 0.5 * (fld.z[0] - c) * fld.px[0];
    double fxmax = fxmin;
    double fymin = fld.y[0] + 0.5 * (fld.z[0] - c) * fld.py[0];
    double fymax = fymin;
    double fx0 = fxmin, fy0 = fymin;
    double fxs = 0, fxs2 = 0, fys = 0, fys2 = 0;
    for (int64_t i = 0; i < nfield; ++i) {
        double s = 0.5 * (fld.z[i] - c);
        fld.x[i] += s * fld.px[i];
        fld.y[i] += s * fld.py[i];
        if (o.lkick) fld.pz[i] -= 0.25 * (fld.px[i] * fld.px[i] + fld.py[i] * fld.py[i]);
        fxmin = jmin(fxmin, fld.x[i]); fxmax = jmax(fxmax, fld.x[i]);
        fymin = jmin(fymin, fld.y[i]); fymax = jmax(fymax, fld.y[i]);
        if (ge != OCTOPIC_EXTENT_EXTREMA) {
            double dx = fld.x[i] - fx0, dy = fld.y[i] - fy0;
            fxs += dx; fxs2 += dx * dx; fys += dy; fys2 += dy * dy;
        }
    }
    axis_extent(ge, fxmin, fxmax, fx0, fxs, fxs2, nfield, kext);
    axis_extent(ge, fymin, fymax, fy0, fys, fys2, nfield, kext);
    if (!(std::isfinite(fxmin) && std::isfinite(fxmax) && std::isfinite(fymin) && std::isfinite(fymax)))
        nonfinite("field", "slice pair");

    Grid sg, fg;
    if (!ov) interaction_grids(o, sxmin, sxmax, symin, symax, fxmin, fxmax, fymin, fymax, sg, fg);
    else interaction_grids(o, ov->s.xmin, ov->s.xmax, ov->s.ymin, ov->s.ymax, ov->f.xmin, ov->f.xmax, ov->f.ymin,
                           ov->f.ymax, sg, fg);
    Box sb{sxmin, sxmax, symin, symax}, fb{fxmin, fxmax, fymin, fymax};
    std::vector<cd> tmp;
    int event = 0;
    const std::vector<cd> *gfft = slice_pair_green(C, w, key, sg, fg, sb, fb, tmp, event);
    if (rec) {
        rec->cache_event = event;
        rec->source_grid[0] = sg.x0; rec->source_grid[1] = sg.y0; rec->source_grid[2] = sg.width; rec->source_grid[3] = sg.height;
        rec->field_grid[0] = fg.x0; rec->field_grid[1] = fg.y0; rec->field_grid[2] = fg.width; rec->field_grid[3] = fg.height;
    }
    if (ge != OCTOPIC_EXTENT_EXTREMA || o.igrid != OCTOPIC_GRID_SLICE_PAIR) {
        w.dropped += count_outside_box(fld.x, fld.y, fg.x0, fg.x0 + fg.width, fg.y0, fg.y0 + fg.height);
        w.dropped += count_outside_box_drifted(src, sL, sR, sg.x0, sg.x0 + sg.width, sg.y0, sg.y0 + sg.height);
    }
    solve_drifted_field(C, w, w.left, src, sL, sg, *gfft);
    solve_drifted_field(C, w, w.right, src, sR, sg, *gfft);
    double kick_scale = 2.0 * kbb;
    double hzi, zbias;
    slice_interp_params(pf.lb, pf.rb, hzi, zbias);
    bool quad = o.interp == OCTOPIC_INTERP_QUADRATIC;
    if (quad) {
        double s_mid = 0.5 * (sL + sR);
        solve_drifted_field(C, w, w.mid, src, s_mid, sg, *gfft);
    }
    for (int64_t i = 0; i < nfield; ++i) {
        double zL = jclamp(-fld.z[i] * hzi + zbias, 0.0, 1.0);
        double Kx, Ky, Kz;
        if (quad) {
            interpolate_kick_quadratic(o, fg, fld.x[i], fld.y[i], w.left, w.mid, w.right, 1.0 - zL, Kx, Ky, Kz);
        } else {
            double zR = 1.0 - zL;
            interpolate_kick(o, fg, fld.x[i], fld.y[i], w.left, w.right, zL, zR, Kx, Ky, Kz);
        }
        fld.px[i] += kick_scale * Kx;
        fld.py[i] += kick_scale * Ky;
        if (o.lkick) fld.pz[i] += kick_scale * Kz * hzi;
        double s = 0.5 * (c - fld.z[i]);
        fld.x[i] += s * fld.px[i];
        fld.y[i] += s * fld.py[i];
        if (o.lkick) fld.pz[i] += 0.25 * (fld.px[i] * fld.px[i] + fld.py[i] * fld.py[i]);
    }
    double sM = 0.5 * (ps.center - pf.center);
    vx.resize(nsource); vy.resize(nsource);
    for (int64_t i = 0; i < nsource; ++i) {
        vx[i] = src.x[i] + src.px[i] * sM;
        vy[i] = src.y[i] + src.py[i] * sM;
    }
}

// _pic_build_node_grids! (pic_cpu.jl:1317-1382): meshes for every node of one (source slice, dir).
void build_node_grids(Context &C, Workspace &w, std::map<int, NodeEntry> &cache, const Coords &src, double center,
                      const Coords &fld_rep, const Slices &fsl) {
    const Opt &o = C.o;
    int nb = (int)fsl.boundary.size();
    int ns = nb - 1;
    double c = center;
    Vec drifts(nb);
    for (int b = 0; b < nb; ++b) drifts[b] = 0.5 * (c - fsl.boundary[b]);
    Vec sxlo(nb, INFINITY), sxhi(nb, -INFINITY), sylo(nb, INFINITY), syhi(nb, -INFINITY);
    for (size_t i = 0; i < src.size(); ++i) {
        double x = src.x[i], px = src.px[i], y = src.y[i], py = src.py[i];
        for (int b = 0; b < nb; ++b) {
            double d = drifts[b];
            double xv = x + px * d, yv = y + py * d;
            sxlo[b] = jmin(sxlo[b], xv); sxhi[b] = jmax(sxhi[b], xv);
            sylo[b] = jmin(sylo[b], yv); syhi[b] = jmax(syhi[b], yv);
        }
    }
    Vec fxlo(ns, INFINITY), fxhi(ns, -INFINITY), fylo(ns, INFINITY), fyhi(ns, -INFINITY);
    for (int sl = 0; sl < ns; ++sl)
        for (int64_t i : fsl.indices[sl]) {
            double sh = 0.5 * (fld_rep.z[i] - c);
            double xv = fld_rep.x[i] + sh * fld_rep.px[i];
            double yv = fld_rep.y[i] + sh * fld_rep.py[i];
            fxlo[sl] = jmin(fxlo[sl], xv); fxhi[sl] = jmax(fxhi[sl], xv);
            fylo[sl] = jmin(fylo[sl], yv); fyhi[sl] = jmax(fyhi[sl], yv);
        }
    for (int b = 0; b < nb; ++b)
        if (std::isnan(sxlo[b]) || std::isnan(sxhi[b]) || std::isnan(sylo[b]) || std::isnan(syhi[b]))
            nonfinite("source", "node-mesh build");
    for (int sl = 0; sl < ns; ++sl)
        if (std::isnan(fxlo[sl]) || std::isnan(fxhi[sl]) || std::isnan(fylo[sl]) || std::isnan(fyhi[sl]))
            nonfinite("field", "node-mesh build");
    for (int b = 0; b < nb; ++b) {
        int bn = std::min(b + 1, nb - 1);
        double sxmin = jmin(sxlo[b], sxlo[bn]), sxmax = jmax(sxhi[b], sxhi[bn]);
        double symin = jmin(sylo[b], sylo[bn]), symax = jmax(syhi[b], syhi[bn]);
        double fxmin = INFINITY, fxmax = -INFINITY, fymin = INFINITY, fymax = -INFINITY;
        for (int adj : {b - 1, b}) {
            if (!(0 <= adj && adj < ns)) continue;
            if (!std::isfinite(fxlo[adj])) continue;
            fxmin = jmin(fxmin, fxlo[adj]); fxmax = jmax(fxmax, fxhi[adj]);
            fymin = jmin(fymin, fylo[adj]); fymax = jmax(fymax, fyhi[adj]);
        }
        if (!std::isfinite(fxmin)) continue;
        NodeEntry e;
        interaction_grids(o, sxmin, sxmax, symin, symax, fxmin, fxmax, fymin, fymax, e.sg, e.fg);
        e.gfft = green_fft(C, w, e.sg, e.fg);
        cache[b] = std::move(e);
    }
}

// _pic_interaction_node! (pic_cpu.jl:1448-1538)
void interaction_node(Context &C, Workspace &w, const Coords &src, const SliceParam &ps, Coords &fld,
                      const SliceParam &pf, double kbb, const NodeEntry &gL, const NodeEntry &gR, Vec &vx, Vec &vy,
                      octopic_interaction_record *rec) {
    const Opt &o = C.o;
    int64_t nsource = (int64_t)src.size(), nfield = (int64_t)fld.size();
    double c = ps.center;
    double sL = 0.5 * (c - pf.lb), sR = 0.5 * (c - pf.rb);
    for (int64_t i = 0; i < nfield; ++i) {
        double s = 0.5 * (fld.z[i] - c);
        fld.x[i] += s * fld.px[i];
        fld.y[i] += s * fld.py[i];
        if (o.lkick) fld.pz[i] -= 0.25 * (fld.px[i] * fld.px[i] + fld.py[i] * fld.py[i]);
    }
    w.dropped += count_outside_box_drifted(src, sL, sL, gL.sg.x0, gL.sg.x0 + gL.sg.width, gL.sg.y0, gL.sg.y0 + gL.sg.height);
    w.dropped += count_outside_box_drifted(src, sR, sR, gR.sg.x0, gR.sg.x0 + gR.sg.width, gR.sg.y0, gR.sg.y0 + gR.sg.height);
    w.dropped += count_outside_box(fld.x, fld.y, gL.fg.x0, gL.fg.x0 + gL.fg.width, gL.fg.y0, gL.fg.y0 + gL.fg.height);
    if (rec) {
        rec->cache_event = 0;
        rec->source_grid[0] = gL.sg.x0; rec->source_grid[1] = gL.sg.y0; rec->source_grid[2] = gL.sg.width; rec->source_grid[3] = gL.sg.height;
        rec->field_grid[0] = gL.fg.x0; rec->field_grid[1] = gL.fg.y0; rec->field_grid[2] = gL.fg.width; rec->field_grid[3] = gL.fg.height;
    }
    solve_drifted_field(C, w, w.left, src, sL, gL.sg, gL.gfft);
    solve_drifted_field(C, w, w.right, src, sR, gR.sg, gR.gfft);
    if (o.lkick) solve_drifted_field(C, w, w.mid, src, sR, gL.sg, gL.gfft);
    double kick_scale = 2.0 * kbb;
    double hzi, zbias;
    slice_interp_params(pf.lb, pf.rb, hzi, zbias);
    for (int64_t i = 0; i < nfield; ++i) {
        double zL = jclamp(-fld.z[i] * hzi + zbias, 0.0, 1.0);
        double zR = 1.0 - zL;
        double x = fld.x[i], y = fld.y[i];
        double KxL, KyL, KzL, KxR, KyR, KzR;
        interpolate_kick(o, gL.fg, x, y, w.left, w.left, 1.0, 0.0, KxL, KyL, KzL);
        interpolate_kick(o, gR.fg, x, y, w.right, w.right, 1.0, 0.0, KxR, KyR, KzR);
        fld.px[i] += kick_scale * (zL * KxL + zR * KxR);
        fld.py[i] += kick_scale * (zL * KyL + zR * KyR);
        if (o.lkick) {
            double a, b, Kz;
            // phiL - phiZ on gL; the Z plane's field parts are the L plane's, as in the reference
            interpolate_kick(o, gL.fg, x, y, w.left.phi.data(), w.left.Ex.data(), w.left.Ey.data(), w.mid.phi.data(),
                             w.left.Ex.data(), w.left.Ey.data(), 1.0, 0.0, a, b, Kz);
            fld.pz[i] += kick_scale * Kz * hzi;
        }
        double s = 0.5 * (c - fld.z[i]);
        fld.x[i] += s * fld.px[i];
        fld.y[i] += s * fld.py[i];
        if (o.lkick) fld.pz[i] += 0.25 * (fld.px[i] * fld.px[i] + fld.py[i] * fld.py[i]);
    }
    double sM = 0.5 * (ps.center - pf.center);
    vx.resize(nsource); vy.resize(nsource);
    for (int64_t i = 0; i < nsource; ++i) {
        vx[i] = src.x[i] + src.px[i] * sM;
        vy[i] = src.y[i] + src.py[i] * sM;
    }
}

// _pic_luminosity! (pic_cpu.jl:2700-2736)
double luminosity(const Opt &o, const Vec &x1, const Vec &y1, const Vec &x2, const Vec &y2, double klum, Vec &q1, Vec &q2) {
    int64_t nx = o.lum_nx, ny = o.lum_ny;
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    double a = INFINITY, b = -INFINITY;
    for (double v : x1) { a = jmin(a, v); b = jmax(b, v); }
    double c = INFINITY, d = -INFINITY;
    for (double v : x2) { c = jmin(c, v); d = jmax(d, v); }
    xmin = jmin(a, c); xmax = jmax(b, d);
    a = INFINITY; b = -INFINITY;
    for (double v : y1) { a = jmin(a, v); b = jmax(b, v); }
    c = INFINITY; d = -INFINITY;
    for (double v : y2) { c = jmin(c, v); d = jmax(d, v); }
    ymin = jmin(a, c); ymax = jmax(b, d);
    double width, height;
    resolve_extent(o, xmax - xmin, ymax - ymin, width, height, "PIC luminosity grid");
    double tx = width / ((double)nx - 1.1);
    double ty = height / ((double)ny - 1.1);
    width += 0.1 * tx;
    height += 0.1 * ty;
    xmin -= 0.05 * tx;
    ymin -= 0.05 * ty;
    double hx = width / (double)(nx - 1);
    double hy = height / (double)(ny - 1);
    double hxi = 1.0 / hx, hyi = 1.0 / hy;
    std::fill(q1.begin(), q1.end(), 0.0);
    std::fill(q2.begin(), q2.end(), 0.0);
    deposit(q1.data(), nx + 1, ny + 1, o.lum_deposit, x1.data(), nullptr, y1.data(), nullptr, 0.0, false, xmin, ymin, hx,
            hy, nx + 1, ny + 1, (int64_t)x1.size(), nullptr);
    deposit(q2.data(), nx + 1, ny + 1, o.lum_deposit, x2.data(), nullptr, y2.data(), nullptr, 0.0, false, xmin, ymin, hx,
            hy, nx + 1, ny + 1, (int64_t)x2.size(), nullptr);
    double lum = 0.0;
    for (int64_t j = 0; j < ny + 1; ++j)
        for (int64_t i = 0; i < nx + 1; ++i) lum += q1[i + j * (nx + 1)] * q2[i + j * (nx + 1)];
    return lum * klum * hxi * hyi;
}

Coords extract_slice(double *const b[6], const std::vector<int64_t> &idx) {
    Coords c;
    Vec *dst[6] = {&c.x, &c.px, &c.y, &c.py, &c.z, &c.pz};
    for (int k = 0; k < 6; ++k) {
        dst[k]->resize(idx.size());
        for (size_t j = 0; j < idx.size(); ++j) (*dst[k])[j] = b[k][idx[j]];
    }
    return c;
}
void store_slice(double *const b[6], const std::vector<int64_t> &idx, const Coords &c) {
    const Vec *src[6] = {&c.x, &c.px, &c.y, &c.py, &c.z, &c.pz};
    for (int k = 0; k < 6; ++k)
        for (size_t j = 0; j < idx.size(); ++j) b[k][idx[j]] = (*src[k])[j];
}

// _pic_union_bounds (pic_cpu.jl:665-716), the resident-state form
bool union_bounds(const Coords &src, double center, const std::vector<Coords> &states, UnionBounds &out) {
    double c = center;
    double zmin = INFINITY, zmax = -INFINITY;
    double fxmin = INFINITY, fxmax = -INFINITY, fymin = INFINITY, fymax = -INFINITY;
    for (const Coords &fc : states)
        for (size_t p = 0; p < fc.size(); ++p) {
            double zi = fc.z[p];
            zmin = jmin(zmin, zi); zmax = jmax(zmax, zi);
            double s = 0.5 * (zi - c);
            double xv = fc.x[p] + s * fc.px[p], yv = fc.y[p] + s * fc.py[p];
            fxmin = jmin(fxmin, xv); fxmax = jmax(fxmax, xv); fymin = jmin(fymin, yv); fymax = jmax(fymax, yv);
        }
    if (std::isnan(zmin) || !(std::isfinite(fxmin) && std::isfinite(fxmax) && std::isfinite(fymin) && std::isfinite(fymax))) {
        if (std::isnan(zmin) || std::isfinite(zmin)) nonfinite("field", "source-slice union bounds");
    }
    if (!std::isfinite(zmin)) return false;
    double sA = 0.5 * (c - zmax), sB = 0.5 * (c - zmin);
    double sxmin = INFINITY, sxmax = -INFINITY, symin = INFINITY, symax = -INFINITY;
    for (size_t i = 0; i < src.size(); ++i) {
        double xa = src.x[i] + src.px[i] * sA, xb = src.x[i] + src.px[i] * sB;
        double ya = src.y[i] + src.py[i] * sA, yb = src.y[i] + src.py[i] * sB;
        sxmin = jmin(jmin(sxmin, xa), xb); sxmax = jmax(jmax(sxmax, xa), xb);
        symin = jmin(jmin(symin, ya), yb); symax = jmax(jmax(symax, ya), yb);
    }
    if (!(std::isfinite(sxmin) && std::isfinite(sxmax) && std::isfinite(symin) && std::isfinite(symax)))
        nonfinite("source", "source-slice union bounds");
    out.s = {sxmin, sxmax, symin, symax};
    out.f = {fxmin, fxmax, fymin, fymax};
    return true;
}

void gpic_interaction(Context &C, Workspace &w, const Coords &src, const SliceParam &ps, Coords &fld, const SliceParam &pf,
                      double kbb, const std::tuple<int, int, int> *key, Vec &vx, Vec &vy, octopic_interaction_record *rec);

struct CollideOut {
    double lum = NAN;
    std::vector<int32_t> pi, pj, ran;
    Vec plum;
    int64_t dropped = 0, batches = 0, widest = 0;
};

struct State {
    std::unique_ptr<Context> C;
    std::vector<octopic_interaction_record> records;
    // :node meshes of the last collide: (dir, source slice, node) -> source and field mesh
    std::map<std::tuple<int, int, int>, std::pair<Grid, Grid>> node_meshes;
    Slices last1, last2;
    CollideOut last;
};

double pic_kbb_phys(const Opt &, double given, const double p_self[5], const double p_other[5], bool first) {
    // _strong_strong_kbb1/2 (slicing.jl:77-91); params = charge, mc2, E0, r0, npart
    if (!std::isnan(given)) return given;
    double E0 = p_self[2];
    if (!(std::isfinite(E0) && E0 != 0.0))
        throw OracleError(OCTOPIC_ERR_INVALID_ARGUMENT, first ? "beam1 E0 must be finite and nonzero when kbb1 is not specified"
                                                              : "beam2 E0 must be finite and nonzero when kbb2 is not specified");
    const double *p1 = first ? p_self : p_other;
    const double *p2 = first ? p_other : p_self;
    return p1[0] * p2[0] * p_self[3] * p_other[4] * p_self[1] / p_self[2];
}

// ---- GaussianPoissonSolver: the sliced soft-Gaussian solver (gaussian.jl), uncoupled
constexpr int64_t REDUCTION_CHUNKS = 64;       // _REDUCTION_CHUNKS (interface.jl:664)
constexpr int64_t PARALLEL_MOMENT_MIN = 4096;  // _STRONG_STRONG_PARALLEL_MOMENT_MIN (interface.jl:609)
constexpr int64_t PARALLEL_KICK_MIN = 4096;    // _STRONG_STRONG_PARALLEL_KICK_MIN (interface.jl:610)
constexpr double TWOPI_ = 6.283185307179586;

// The Bassetti-Erskine kick with its Hessian diagonal (defined with the Faddeeva evaluation below):
// out = (Kx, Ky, kbb Hxx, kbb Hyy).
void be_response4(double kbb, double sigx, double sigy, double x, double y, double out[4]);

// A slice's moments (_slice_transverse_moments, slicing.jl:670-782, uncoupled): the centroid (zeroed by
// ignore_centroid) and the StrongTransverseMoments a0 = var x (floored), d0 = var y (floored), bxx =
// cov(x, px), bypy = cov(y, py), qxx = var px, qyy = var py.
struct SoftMom { double mx, mpx, my, mpy, a0, d0, bxx, bypy, qxx, qyy; };

SoftMom soft_moments(double *const b[6], const std::vector<int64_t> &idx, bool ignore_centroid, double min_sigma) {
    const int64_t n = (int64_t)idx.size();
    const double floor2 = min_sigma * min_sigma;
    if (n == 0) return SoftMom{0, 0, 0, 0, floor2, floor2, 0, 0, 0, 0};
    const int64_t f = idx[0];
    const double x0 = b[0][f], px0 = b[1][f], y0 = b[2][f], py0 = b[3][f];
    double s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // dx, dpx, dy, dpy, dx2, dpx2, dy2, dpy2, dxdpx, dydpy
    auto add = [&](double *acc, int64_t i) {
        const double dx = b[0][i] - x0, dpx = b[1][i] - px0, dy = b[2][i] - y0, dpy = b[3][i] - py0;
        acc[0] += dx; acc[1] += dpx; acc[2] += dy; acc[3] += dpy;
        acc[4] += dx * dx; acc[5] += dpx * dpx; acc[6] += dy * dy; acc[7] += dpy * dpy;
        acc[8] += dx * dpx; acc[9] += dy * dpy;
    };
    if (n < PARALLEL_MOMENT_MIN) {
        for (int64_t k = 0; k < n; ++k) add(s, idx[k]);
    } else {   // the fixed chunk grid (_chunk_bounds, slicing.jl:517-521), chunks folded in order
        for (int64_t c = 0; c < REDUCTION_CHUNKS; ++c) {
            double part[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int64_t k = fld(c * n, REDUCTION_CHUNKS); k < fld((c + 1) * n, REDUCTION_CHUNKS); ++k) add(part, idx[k]);
            for (int q = 0; q < 10; ++q) s[q] += part[q];
        }
    }
    const double invn = 1.0 / (double)n;
    const double dmx = s[0] * invn, dmpx = s[1] * invn, dmy = s[2] * invn, dmpy = s[3] * invn;
    auto shifted = [&](double sum, double a, double bb) { return std::fma(-a, bb, sum * invn); };
    SoftMom m;
    m.mx = x0 + dmx; m.mpx = px0 + dmpx; m.my = y0 + dmy; m.mpy = py0 + dmpy;
    m.a0 = std::max(shifted(s[4], dmx, dmx), floor2);
    m.d0 = std::max(shifted(s[6], dmy, dmy), floor2);
    m.qxx = std::max(shifted(s[5], dmpx, dmpx), 0.0);
    m.qyy = std::max(shifted(s[7], dmpy, dmpy), 0.0);
    m.bxx = shifted(s[8], dmx, dmpx);
    m.bypy = shifted(s[9], dmy, dmpy);
    if (ignore_centroid) { m.mx = 0.0; m.mpx = 0.0; m.my = 0.0; m.mpy = 0.0; }
    return m;
}

bool soft_moments_finite(const SoftMom &m) {   // _gaussian_moments_finite (gaussian.jl:126-130)
    return std::isfinite(m.mx) && std::isfinite(m.mpx) && std::isfinite(m.my) && std::isfinite(m.mpy) &&
           std::isfinite(std::sqrt(m.a0)) && std::isfinite(std::sqrt(m.d0)) && std::isfinite(std::sqrt(m.qxx)) &&
           std::isfinite(std::sqrt(m.qyy)) && std::isfinite(m.bxx) && std::isfinite(m.bypy);
}

// The virtual drifts (strong_beam.jl:497-617): forward returns S = (z - zo)/2 of the drift's own rule.
double forward_drift(int d, double &x, double &px, double &y, double &py, double &z, double &pz, double zo) {
    (void)px; (void)py;
    if (d == OCTOPIC_DRIFT_EXACT) {
        const double ps = std::sqrt((1 + pz) * (1 + pz) - px * px - py * py);
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
    } else {   // chromatic (and its frozen-energy form)
        const double phi = std::sqrt(1 - 0.5 * (px * px + py * py) / ((1 + pz) * (1 + pz))) - 1;
        x += S * px / (1 + pz);
        y += S * py / (1 + pz);
        z += 2 * S * phi;
        if (d == OCTOPIC_DRIFT_CHROMATIC) pz += (1 + pz) * phi;
    }
    return S;
}

void reverse_drift(int d, double &x, double &px, double &y, double &py, double &z, double &pz, double zo) {
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
        const double psi = std::sqrt(1 + 0.5 * (px * px + py * py) / ((1 + pz) * (1 + pz))) - 1;
        x -= S * px / (1 + pz);
        y -= S * py / (1 + pz);
        z += 2 * S * psi;
        if (d == OCTOPIC_DRIFT_CHROMATIC) pz += (1 + pz) * psi;
    }
}

// _apply_slice_kick_one! (gaussian.jl:152-188) with the uncoupled _cp_covariance_kick (strong_beam.jl:619-632)
// and _transport_transverse_moments (:679-687); returns the luminosity density times 2 pi (0 if not wanted).
double soft_kick_one(double &x, double &px, double &y, double &py, double &z, double &pz, const SoftMom &m, double center,
                     double kbb, int drift, bool lkick, bool want_lum) {
    const int d = lkick ? drift : OCTOPIC_DRIFT_UNSAFE_PARAXIAL_FROZEN_LONGITUDINAL;   // _soft_gaussian_drift
    const double S = forward_drift(d, x, px, y, py, z, pz, center);
    const double xx = x - m.mx + m.mpx * S;
    const double yy = y - m.my + m.mpy * S;
    const double px0 = px, py0 = py, pz0 = pz;
    const double a = m.a0 - 2 * S * m.bxx + S * S * m.qxx;
    const double dd = m.d0 - 2 * S * m.bypy + S * S * m.qyy;
    const double au = 2 * (m.bxx - S * m.qxx);
    const double du = 2 * (m.bypy - S * m.qyy);
    double density = 0.0;
    if (!(a <= 0 || dd <= 0)) {
        const double sigx = std::sqrt(a), sigy = std::sqrt(dd);
        double r[4];
        be_response4(kbb, sigx, sigy, xx, yy, r);
        const double expterm = std::exp(-0.5 * (xx * xx / (sigx * sigx) + yy * yy / (sigy * sigy)));
        px += r[0] * kbb;
        py += r[1] * kbb;
        pz += 0.25 * (r[2] * au + r[3] * du);
        density = expterm / TWOPI_ / sigx / sigy;
    }
    if (lkick) pz += 0.5 * ((px - px0) * m.mpx + (py - py0) * m.mpy);
    else pz = pz0;
    reverse_drift(d, x, px, y, py, z, pz, center);
    return want_lum ? density * TWOPI_ : 0.0;
}

// _slice_slice_gaussian_kick! (gaussian.jl:59-124): kicks the slice `idx` of beam b; returns its luminosity.
double soft_slice_kick(double *const b[6], const std::vector<int64_t> &idx, const SoftMom &m, double center, double kbb_slice,
                       double klum_slice, int drift, bool lkick, bool want_lum) {
    const int64_t n = (int64_t)idx.size();
    if (n == 0) return 0.0;
    auto one = [&](int64_t i) {
        return soft_kick_one(b[0][i], b[1][i], b[2][i], b[3][i], b[4][i], b[5][i], m, center, kbb_slice, drift, lkick, want_lum);
    };
    double lum = 0.0;
    if (n < PARALLEL_KICK_MIN) {
        for (int64_t k = 0; k < n; ++k) lum += one(idx[k]);
    } else {
        for (int64_t c = 0; c < REDUCTION_CHUNKS; ++c) {
            double part = 0.0;
            for (int64_t k = fld(c * n, REDUCTION_CHUNKS); k < fld((c + 1) * n, REDUCTION_CHUNKS); ++k) part += one(idx[k]);
            lum += part;
        }
    }
    return lum / TWOPI_ * klum_slice;
}

// collide!(::GaussianPoissonSolver, ...) (gaussian.jl:7-57): pairs in collision order (the conflict-free
// batches of the order run concurrently, each pair touching only its two slices); the per-pair luminosity
// is folded in collision order.
void soft_collide(State &S, double *b1[6], int64_t n1, const double p1[5], double *b2[6], int64_t n2, const double p2[5],
                  int compute_lum, int nthreads) {
    const Opt &o = S.C->o;
    if (n1 <= 0 || n2 <= 0) throw OracleError(OCTOPIC_ERR_INVALID_ARGUMENT, "longitudinal slicing requires at least one particle");
    Slices s1 = longitudinal_slices(b1[4], n1, o.s1);
    Slices s2 = longitudinal_slices(b2[4], n2, o.s2);
    const double kbb1 = pic_kbb_phys(o, o.kbb1, p1, p2, true), kbb2 = pic_kbb_phys(o, o.kbb2, p2, p1, false);
    double klum1 = o.lum_scale, klum2 = o.lum_scale;   // _strong_strong_luminosity_scales (slicing.jl:93-102)
    if (std::isnan(o.lum_scale)) { klum1 = p1[4] * p2[4] / (double)n1; klum2 = p1[4] * p2[4] / (double)n2; }
    const bool sample1 = o.gwl == 1;
    auto order = collision_order(s1.center, s2.center);
    const int npairs = (int)order.size();
    Vec lum_parts(npairs, 0.0);
    std::vector<int32_t> ran(npairs, 1);
    std::map<std::pair<int, int>, int> pair_pos;
    for (int p = 0; p < npairs; ++p) pair_pos[{order[p].i, order[p].j}] = p;
    auto pair_fn = [&](int p, int i, int j) {   // _cpu_gaussian_slice_pair! (gaussian.jl:24-57)
        const SoftMom m1 = soft_moments(b1, s1.indices[i], o.ignore1, o.min_sigma);
        const SoftMom m2 = soft_moments(b2, s2.indices[j], o.ignore2, o.min_sigma);
        if (!soft_moments_finite(m1))
            throw OracleError(OCTOPIC_ERR_NONFINITE, "non-finite source coordinate (beam 1, slice " + std::to_string(i + 1) + ")");
        if (!soft_moments_finite(m2))
            throw OracleError(OCTOPIC_ERR_NONFINITE, "non-finite source coordinate (beam 2, slice " + std::to_string(j + 1) + ")");
        const double l2 = soft_slice_kick(b1, s1.indices[i], m2, s2.center[j], s2.weight[j] * kbb1, s2.weight[j] * klum1,
                                          o.vdrift, o.lkick != 0, !sample1);
        const double l1 = soft_slice_kick(b2, s2.indices[j], m1, s1.center[i], s1.weight[i] * kbb2, s1.weight[i] * klum2,
                                          o.vdrift, o.lkick != 0, sample1);
        lum_parts[p] = sample1 ? l1 : l2;
    };
    auto batches = collision_pair_batches(s1.center, s2.center);
    S.last.batches = (int64_t)batches.size();
    S.last.widest = 0;
    for (auto &b : batches) S.last.widest = std::max<int64_t>(S.last.widest, (int64_t)b.size());
    for (auto &batch : batches) {
        const int nb = (int)batch.size();
        std::vector<std::string> errs(nb);
        std::vector<int> codes(nb, 0);
#ifdef _OPENMP
#pragma omp parallel for num_threads(std::max(1, std::min(nthreads, nb))) schedule(static, 1)
#endif
        for (int k = 0; k < nb; ++k) {
            try {
                pair_fn(pair_pos.at({batch[k].i, batch[k].j}), batch[k].i, batch[k].j);
            } catch (OracleError &e) {
                codes[k] = e.code; errs[k] = e.what();
            }
        }
        for (int k = 0; k < nb; ++k)
            if (codes[k]) throw OracleError(codes[k], errs[k]);
    }
    double lum = NAN;
    if (compute_lum) {   // the reported value; the kicks are the same either way
        lum = 0.0;
        for (int p = 0; p < npairs; ++p) lum += lum_parts[p];
    }
    S.records.clear();
    S.node_meshes.clear();
    S.last.lum = lum;
    S.last.pi.resize(npairs); S.last.pj.resize(npairs); S.last.ran = ran; S.last.plum = lum_parts;
    for (int p = 0; p < npairs; ++p) { S.last.pi[p] = order[p].i; S.last.pj[p] = order[p].j; }
    S.last.dropped = 0;
    S.last1 = std::move(s1);
    S.last2 = std::move(s2);
}

void collide(State &S, double *b1[6], int64_t n1, const double p1[5], double *b2[6], int64_t n2, const double p2[5],
             int compute_lum, int nthreads) {
    Context &C = *S.C;
    const Opt &o = C.o;
    if (o.solver == OCTOPIC_SOLVER_SOFT_GAUSSIAN) {
        soft_collide(S, b1, n1, p1, b2, n2, p2, compute_lum, nthreads);
        return;
    }
    if (n1 <= 0 || n2 <= 0) throw OracleError(OCTOPIC_ERR_INVALID_ARGUMENT, "longitudinal slicing requires at least one particle");
    Slices s1 = longitudinal_slices(b1[4], n1, o.s1);
    Slices s2 = longitudinal_slices(b2[4], n2, o.s2);
    double kbb1 = pic_kbb_phys(o, o.kbb1, p1, p2, true) / (double)n2;   // pic_cpu.jl:821-824
    double kbb2 = pic_kbb_phys(o, o.kbb2, p2, p1, false) / (double)n1;
    double klum = !std::isnan(o.lum_scale) ? o.lum_scale : p1[4] * p2[4] / (double)(n1 * n2);
    int ns1 = (int)s1.center.size(), ns2 = (int)s2.center.size();
    int pool = std::max(1, std::min(nthreads, std::min(ns1, ns2)));
    bool node = o.igrid == OCTOPIC_GRID_NODE, share = o.igrid == OCTOPIC_GRID_SOURCE_SLICE;
    while ((int)C.ws.size() < pool) C.ws.emplace_back(new Workspace(o.nx, o.ny, o.lum_nx, o.lum_ny));
    for (auto &w : C.ws) w->dropped = 0;

    std::map<std::pair<int, int>, std::map<int, NodeEntry>> node_cache;
    if (node) {   // _pic_prebuild_node_caches! (pic_cpu.jl:1404-1419)
        Coords rep1{Vec(b1[0], b1[0] + n1), Vec(b1[1], b1[1] + n1), Vec(b1[2], b1[2] + n1), Vec(b1[3], b1[3] + n1),
                    Vec(b1[4], b1[4] + n1), Vec(b1[5], b1[5] + n1)};
        Coords rep2{Vec(b2[0], b2[0] + n2), Vec(b2[1], b2[1] + n2), Vec(b2[2], b2[2] + n2), Vec(b2[3], b2[3] + n2),
                    Vec(b2[4], b2[4] + n2), Vec(b2[5], b2[5] + n2)};
        for (int i = 0; i < ns1; ++i) {
            if (s1.indices[i].empty()) continue;
            Coords src = extract_slice(b1, s1.indices[i]);
            build_node_grids(C, *C.ws[0], node_cache[{i, 1}], src, s1.center[i], rep2, s2);
        }
        for (int j = 0; j < ns2; ++j) {
            if (s2.indices[j].empty()) continue;
            Coords src = extract_slice(b2, s2.indices[j]);
            build_node_grids(C, *C.ws[0], node_cache[{j, 2}], src, s2.center[j], rep1, s1);
        }
    }
    S.node_meshes.clear();
    for (const auto &kv : node_cache)
        for (const auto &nb : kv.second)
            S.node_meshes[std::make_tuple(kv.first.second, kv.first.first, nb.first)] = {nb.second.sg, nb.second.fg};
    auto order = collision_order(s1.center, s2.center);
    int npairs = (int)order.size();
    std::map<std::pair<int, int>, int> pair_pos;
    for (int p = 0; p < npairs; ++p) pair_pos[{order[p].i, order[p].j}] = p;
    Vec lum_parts(npairs, 0.0);
    std::vector<int32_t> ran(npairs, 0);
    std::vector<Coords> st1(ns1), st2(ns2);
    for (int i = 0; i < ns1; ++i) st1[i] = extract_slice(b1, s1.indices[i]);
    for (int j = 0; j < ns2; ++j) st2[j] = extract_slice(b2, s2.indices[j]);
    S.records.assign((size_t)2 * npairs, octopic_interaction_record{});
    std::map<std::pair<int, int>, UnionBounds> ub;
    std::map<std::pair<int, int>, bool> ub_have;

    auto pair_fn = [&](int p, int i, int j, Workspace &w) -> bool {   // _pic_collide_pair! (pic_cpu.jl:292-392)
        if (s1.indices[i].empty() || s2.indices[j].empty()) return false;
        SliceParam q1{s1.weight[i], s1.boundary[i], s1.center[i], s1.boundary[i + 1]};
        SliceParam q2{s2.weight[j], s2.boundary[j], s2.center[j], s2.boundary[j + 1]};
        Coords &coord1 = st1[i];
        Coords field2 = st2[j];
        Vec vx1, vy1, vx2, vy2;
        octopic_interaction_record *r1 = &S.records[(size_t)2 * p], *r2 = &S.records[(size_t)2 * p + 1];
        r1->p = p; r1->i = i; r1->j = j; r1->dir = 1;
        r2->p = p; r2->i = i; r2->j = j; r2->dir = 2;
        if (node) {
            auto it1 = node_cache.find({i, 1}), it2 = node_cache.find({j, 2});
            if (it1 == node_cache.end() || it2 == node_cache.end() || it1->second.empty() || it2->second.empty())
                throw OracleError(OCTOPIC_ERR_INVALID_ARGUMENT, "PIC node mesh cache is missing or empty at collision time");
            auto &nc1 = it1->second;
            auto &nc2 = it2->second;
            auto gL1 = nc1.find(j), gR1 = nc1.find(j + 1), gL2 = nc2.find(i), gR2 = nc2.find(i + 1);
            if (gL1 == nc1.end() || gR1 == nc1.end() || gL2 == nc2.end() || gR2 == nc2.end()) return false;
            interaction_node(C, w, coord1, q1, field2, q2, kbb2, gL1->second, gR1->second, vx1, vy1, r1);
            interaction_node(C, w, st2[j], q2, coord1, q1, kbb1, gL2->second, gR2->second, vx2, vy2, r2);
        } else {
            const UnionBounds *ov1 = nullptr, *ov2 = nullptr;
            UnionBounds u1, u2;
            if (share) {
                auto k1 = std::make_pair(i, 1), k2 = std::make_pair(j, 2);
                if (!ub_have.count(k1)) { ub_have[k1] = union_bounds(coord1, q1.center, st2, ub[k1]); }
                if (!ub_have.count(k2)) { ub_have[k2] = union_bounds(st2[j], q2.center, st1, ub[k2]); }
                if (ub_have[k1]) { u1 = ub[k1]; ov1 = &u1; }
                if (ub_have[k2]) { u2 = ub[k2]; ov2 = &u2; }
            }
            auto key1 = share ? std::make_tuple(i, -1, 1) : std::make_tuple(i, j, 1);
            auto key2 = share ? std::make_tuple(j, -1, 2) : std::make_tuple(i, j, 2);
            if (o.solver == OCTOPIC_SOLVER_GAUSSIAN_PIC) {   // _gpic_collide_pair! (gaussian_pic.jl:1024-1060)
                gpic_interaction(C, w, coord1, q1, field2, q2, kbb2, &key1, vx1, vy1, r1);
                gpic_interaction(C, w, st2[j], q2, coord1, q1, kbb1, &key2, vx2, vy2, r2);
            } else {
                interaction(C, w, coord1, q1, field2, q2, kbb2, &key1, ov1, vx1, vy1, r1);
                interaction(C, w, st2[j], q2, coord1, q1, kbb1, &key2, ov2, vx2, vy2, r2);
            }
        }
        st2[j] = std::move(field2);
        if (compute_lum) lum_parts[p] = luminosity(o, vx1, vy1, vx2, vy2, klum, w.q1, w.q2);
        return true;
    };

    S.last.batches = 0; S.last.widest = 0;
    bool batchable = !share;
    if (batchable && pool > 1 && npairs > 1) {
        auto batches = collision_pair_batches(s1.center, s2.center);
        S.last.batches = (int64_t)batches.size();
        for (auto &b : batches) S.last.widest = std::max<int64_t>(S.last.widest, (int64_t)b.size());
        for (auto &batch : batches) {
            int nb = (int)batch.size();
            std::vector<std::string> errs(nb);
            std::vector<int> codes(nb, 0);
#ifdef _OPENMP
#pragma omp parallel for num_threads(std::min(pool, nb)) schedule(static, 1)
#endif
            for (int k = 0; k < nb; ++k) {
                int tid = 0;
#ifdef _OPENMP
                tid = omp_get_thread_num();
#endif
                try {
                    int p = pair_pos.at({batch[k].i, batch[k].j});
                    ran[p] = pair_fn(p, batch[k].i, batch[k].j, *C.ws[tid]) ? 1 : 0;
                } catch (OracleError &e) {
                    codes[k] = e.code; errs[k] = e.what();
                }
            }
            for (int k = 0; k < nb; ++k)
                if (codes[k]) throw OracleError(codes[k], errs[k]);
        }
    } else {
        for (int p = 0; p < npairs; ++p) ran[p] = pair_fn(p, order[p].i, order[p].j, *C.ws[0]) ? 1 : 0;
    }
    for (int i = 0; i < ns1; ++i) if (!s1.indices[i].empty()) store_slice(b1, s1.indices[i], st1[i]);
    for (int j = 0; j < ns2; ++j) if (!s2.indices[j].empty()) store_slice(b2, s2.indices[j], st2[j]);
    double lum = NAN;
    if (compute_lum) {
        lum = 0.0;
        for (int p = 0; p < npairs; ++p) if (ran[p]) lum += lum_parts[p];
    }
    S.last.lum = lum;
    S.last.pi.resize(npairs); S.last.pj.resize(npairs); S.last.ran = ran; S.last.plum = lum_parts;
    for (int p = 0; p < npairs; ++p) { S.last.pi[p] = order[p].i; S.last.pj[p] = order[p].j; }
    S.last.dropped = 0;
    for (auto &w : C.ws) S.last.dropped += w->dropped;
    S.last1 = std::move(s1);
    S.last2 = std::move(s2);
}

template <class F>
int guarded(F &&f) {
    try {
        f();
        return OCTOPIC_OK;
    } catch (OracleError &e) {
        g_last_error = e.what();
        return e.code;
    } catch (std::exception &e) {
        g_last_error = e.what();
        return OCTOPIC_ERR_INVALID_ARGUMENT;
    }
}

// ==========================================================================================
// Bassetti-Erskine kick of a transverse Gaussian and the Faddeeva function (test infrastructure)
// ==========================================================================================
// w(z) = exp(-z^2) erfc(-iz) as the reference's device path evaluates it (SpecialMath.jl:60-151):
// Weideman's N = 32 rational approximation (SIAM J. Numer. Anal. 31 (1994) 1497-1518) inside the
// band, the Laplace continued fraction with Gautschi's term count outside, the asymptotic forms
// for |x| + y > 4000 and > 1e7, and w(z) = 2 exp(-z^2) - w(-z) below the axis.
struct OracleWeideman {
    double c[32];
    double L;
    OracleWeideman() {
        // a_m = (1/2M) sum_k f(k) cos(pi m k / M) over k = -M+1..M-1, f(k) = exp(-t^2)(L^2 + t^2),
        // t = L tan(k pi / 2M), M = 2N; the polynomial in Z takes a_N first
        const int N = 32, M = 64;
        const long double pi = 3.141592653589793238462643383279502884L;
        const long double LL = std::sqrt((long double)N / std::sqrt(2.0L));
        for (int m = 1; m <= N; ++m) {
            long double acc = 0.0L;
            for (int k = -M + 1; k <= M - 1; ++k) {
                const long double t = LL * std::tan(pi * k / (2.0L * M));
                acc += std::exp(-t * t) * (LL * LL + t * t) * std::cos(pi * (long double)m * k / M);
            }
            c[N - m] = (double)(acc / (2.0L * M));
        }
        L = (double)LL;
    }
};
const OracleWeideman &oracle_weideman() {
    static const OracleWeideman w;
    return w;
}

std::complex<double> faddeeva_far(double zr, double y) {
    const double isp = 0.56418958354775628695, x = std::fabs(zr);
    const std::complex<double> I(0.0, 1.0), z(zr, y);
    if (x + y > 4000.0) {
        if (x + y > 1e7) return I * isp / z;
        return I * isp * z / (z * z - 0.5);
    }
    double nu = std::floor(3.9 + 11.398 / (0.08254 * x + 0.1421 * y + 0.2023));
    std::complex<double> t = z;
    for (nu = 0.5 * (nu - 1.0); nu > 0.4; nu -= 0.5) t = z - nu / t;
    return I * isp / t;
}

std::complex<double> faddeeva_upper(double zr, double zi) {
    const double x = std::fabs(zr), y = zi;
    if (y > 7.0 || (x > 6.0 && (y > 0.1 || (x > 8.0 && y > 1e-10) || x > 28.0))) return faddeeva_far(zr, y);
    const OracleWeideman &wc = oracle_weideman();
    const std::complex<double> iz(-zi, zr), den = wc.L - iz, Z = (wc.L + iz) / den;
    std::complex<double> p(0.0, 0.0);
    for (int k = 0; k < 32; ++k) p = p * Z + wc.c[k];
    return 2.0 * p / (den * den) + 0.56418958354775628695 / den;
}

std::complex<double> faddeeva(double zr, double zi) {
    if (zi < 0.0) {
        const std::complex<double> z(zr, zi);
        return 2.0 * std::exp(-z * z) - faddeeva_upper(-zr, -zi);
    }
    return faddeeva_upper(zr, zi);
}

// gaussian_beambeam_kick (strong_beam.jl:1139-1146) and its principal-axis branches (:705-709 round,
// :779-797 near-round bounds and blend, :799-895 near-round series, :962-1038 elliptic form)
void near_round_series(double v, double eta, double x, double y, double &kx, double &ky) {
    const double q = (x * x + y * y) / (2 * v), d = (x * x - y * y) / (2 * v);
    double m[7];
    if (q <= 2.0) {
        std::fill(m, m + 7, 0.0);
        double term = 1.0;
        for (int k = 0; k <= 24; ++k) {
            for (int j = 0; j < 7; ++j) m[j] += term / (k + 1 + j);
            term *= -q / (k + 1);
        }
    } else {
        const double e = std::exp(-q);
        m[0] = -std::expm1(-q) / q;
        for (int j = 1; j < 7; ++j) m[j] = (j * m[j - 1] - e) / q;
    }
    const double c1x = d * m[2] - m[1], c1y = d * m[2] + m[1];
    const double c2x = 1.5 * m[2] + (-d - q) * m[3] + d * d / 2 * m[4];
    const double c2y = 1.5 * m[2] + (d - q) * m[3] + d * d / 2 * m[4];
    const double c3x = -1.5 * m[3] + (2.5 * d + q) * m[4] + (-d * q - d * d / 2) * m[5] + d * d * d / 6 * m[6];
    const double c3y = 1.5 * m[3] + (2.5 * d - q) * m[4] + (-d * q + d * d / 2) * m[5] + d * d * d / 6 * m[6];
    kx = x / v * (m[0] + eta * (c1x + eta * (c2x + eta * c3x)));
    ky = y / v * (m[0] + eta * (c1y + eta * (c2y + eta * c3y)));
}

double axis_series(double a, double b, double u, double w) {
    const double s = a + b;
    const double j00 = 2 / (a * s), j10 = 2 * (2 * a + b) / (3 * a * a * a * s * s), j01 = 2 / (a * b * s * s);
    const double j20 = 2 * (8 * a * a + 9 * a * b + 3 * b * b) / (15 * std::pow(a, 5) * s * s * s);
    const double j11 = 2 * (3 * a + b) / (3 * a * a * a * b * s * s * s);
    const double j02 = 2 * (a + 3 * b) / (3 * a * b * b * b * s * s * s);
    const double u2 = u * u, w2 = w * w;
    return u * (j00 - (u2 * j10 + w2 * j01) / 2 + u2 * u2 * j20 / 8 + u2 * w2 * j11 / 4 + w2 * w2 * j02 / 8);
}

void elliptic(double s1, double s2, double x, double y, double eta, double &kx, double &ky) {
    const double rho2 = x * x / (s1 * s1) + y * y / (s2 * s2);
    if (std::pow(rho2, 3.5) <= std::ldexp(1.0, -52) / std::sqrt(eta)) {
        kx = axis_series(s1, s2, x, y);
        ky = axis_series(s2, s1, y, x);
        return;
    }
    const double ax = std::fabs(x), ay = std::fabs(y), den = std::sqrt(2.0 * (s1 * s1 - s2 * s2));
    const std::complex<double> r =
        (2.0 * std::sqrt(M_PI) / den) *
        (faddeeva_upper(ax / den, ay / den) -
         std::exp(-ax * ax / (2 * s1 * s1) - ay * ay / (2 * s2 * s2)) * faddeeva_upper(s2 / s1 * ax / den, s1 / s2 * ay / den));
    kx = x < 0 ? -r.imag() : r.imag();
    ky = y < 0 ? -r.real() : r.real();
}

void be_principal(double s1, double s2, double x, double y, double &kx, double &ky) {
    const double v = (s1 * s1 + s2 * s2) / 2, eta = (s1 * s1 - s2 * s2) / (2 * v);
    if (eta == 0.0) {
        const double u = (x * x + y * y) / (2 * v), phi = u == 0.0 ? 1.0 : -std::expm1(-u) / u;
        kx = phi / v * x;
        ky = phi / v * y;
        return;
    }
    const double outer = std::sqrt(std::sqrt(8.0 / 3.0 * 64.0 * std::ldexp(1.0, -52))), inner = outer / 2;
    if (eta <= inner) return near_round_series(v, eta, x, y, kx, ky);
    double ex, ey;
    elliptic(s1, s2, x, y, eta, ex, ey);
    if (eta >= outer) { kx = ex; ky = ey; return; }
    const double t = (eta - inner) / (outer - inner), w = t * t * t * (10 + t * (-15 + 6 * t));
    double sx, sy;
    near_round_series(v, eta, x, y, sx, sy);
    kx = sx + w * (ex - sx);
    ky = sy + w * (ey - sy);
}

// gaussian_beambeam_kick (strong_beam.jl:1139-1146): zero sizes give zero kicks, the larger size is principal
void be_kick(double sigx, double sigy, double x, double y, double &kx, double &ky) {
    kx = ky = 0.0;
    if (sigx == 0.0 || sigy == 0.0) return;
    if (sigx >= sigy) be_principal(sigx, sigy, x, y, kx, ky);
    else be_principal(sigy, sigx, y, x, ky, kx);
}

// ---- the kick with its Hessian diagonal: _gaussian_beambeam_kick_response (strong_beam.jl:1061-1128) and the
// helpers it combines (:695-703 elliptic Hessian, :705-772 round force scale and Hessian, :779-797 bounds and
// blend, :799-876 radial moments, :897-960 near-round series response and potential residual, :962-1014 the
// near-axis series).  K is the normalised kick; H1, H2 carry the factor kbb.
struct Resp { double kx, ky, h1, h2; };

void radial_moments(double q, int first, int count, double *m) {   // m[k] = m_{first+k}
    if (q <= 2.0) {
        std::fill(m, m + count, 0.0);
        double term = 1.0;
        for (int k = 0; k <= 24; ++k) {
            for (int j = 0; j < count; ++j) m[j] += term / (double)(k + 1 + first + j);
            term *= -q / (double)(k + 1);
        }
        return;
    }
    const double eq = std::exp(-q);
    double prev = -std::expm1(-q) / q;   // m0
    if (first == 0) m[0] = prev;
    for (int j = 1; j < first + count; ++j) {
        prev = ((double)j * prev - eq) / q;
        if (j >= first) m[j - first] = prev;
    }
}

void round_hessian(double kbb, double sigma, double x, double y, double expterm, double &hxx, double &hyy) {
    const double r2 = x * x + y * y, sigma2 = sigma * sigma;
    if (r2 == 0.0) { hxx = hyy = -kbb / sigma2; return; }
    const double u = r2 / (2 * sigma2);
    double f, fp;
    if (u < 1.0e-2) {
        const double phi = 1.0 + u * (-1.0 / 2 + u * (1.0 / 6 + u * (-1.0 / 24 + u * (1.0 / 120 + u * (-1.0 / 720 + u * (1.0 / 5040))))));
        const double dphi = -1.0 / 2 + u * (1.0 / 3 + u * (-1.0 / 8 + u * (1.0 / 30 + u * (-1.0 / 144 + u * (1.0 / 840 - u * (1.0 / 5760))))));
        const double inv = 1.0 / sigma2;
        f = phi * inv;
        fp = dphi * inv * inv / 2;
    } else {
        const double ome = 1 - expterm;
        f = 2 * ome / r2;
        fp = 2 * (expterm * u - ome) / (r2 * r2);
    }
    hxx = -kbb * (f + 2 * x * x * fp);
    hyy = -kbb * (f + 2 * y * y * fp);
}

Resp series_response(double kbb, double v, double eta, double x, double y) {
    const double q = (x * x + y * y) / (2 * v), dr = (x * x - y * y) / (2 * v);
    double m[7];
    radial_moments(q, 0, 7, m);
    const double eta2 = eta * eta, eta3 = eta2 * eta;
    const double u1 = dr * m[1];
    const double u2 = m[1] / 2 - q * m[2] + (dr * dr / 2) * m[3];
    const double u3 = 1.5 * dr * m[3] - dr * q * m[4] + dr * dr * dr / 6 * m[5];
    const double v1 = dr * (q * m[2] - m[1]);
    const double v2 = 1.5 * q * m[2] - (q * q + dr * dr) * m[3] + (q * dr * dr / 2) * m[4];
    const double v3 = -1.5 * dr * m[3] + 3.5 * dr * q * m[4] - (dr * q * q + dr * dr * dr / 2) * m[5] + dr * dr * dr * q / 6 * m[6];
    const double c1x = dr * m[2] - m[1], c1y = dr * m[2] + m[1];
    const double c2x = 1.5 * m[2] + (-dr - q) * m[3] + (dr * dr / 2) * m[4];
    const double c2y = 1.5 * m[2] + (dr - q) * m[3] + (dr * dr / 2) * m[4];
    const double c3x = -1.5 * m[3] + (2.5 * dr + q) * m[4] + (-dr * q - dr * dr / 2) * m[5] + dr * dr * dr / 6 * m[6];
    const double c3y = 1.5 * m[3] + (2.5 * dr - q) * m[4] + (-dr * q + dr * dr / 2) * m[5] + dr * dr * dr / 6 * m[6];
    Resp r;
    r.kx = x / v * (m[0] + eta * c1x + eta2 * c2x + eta3 * c3x);
    r.ky = y / v * (m[0] + eta * c1y + eta2 * c2y + eta3 * c3y);
    const double uv = -std::exp(-q) + eta * v1 + eta2 * v2 + eta3 * v3;
    const double ueta = u1 + 2 * eta * u2 + 3 * eta2 * u3;
    r.h1 = kbb / v * (uv + (1 - eta) * ueta);
    r.h2 = kbb / v * (uv - (1 + eta) * ueta);
    return r;
}

double potential_residual(double v, double eta, double x, double y) {   // U_exact - U_series through eta^6
    const double q = (x * x + y * y) / (2 * v), dr = (x * x - y * y) / (2 * v);
    double mm[9];
    radial_moments(q, 3, 9, mm);
    const double m3 = mm[0], m4 = mm[1], m5 = mm[2], m6 = mm[3], m7 = mm[4], m8 = mm[5], m9 = mm[6], m10 = mm[7], m11 = mm[8];
    const double dr2 = dr * dr, q2 = q * q;
    const double u4 = 3.0 / 8 * m3 - 1.5 * q * m4 + (5 * dr2 + 2 * q2) / 4 * m5 - dr2 * q / 2 * m6 + dr2 * dr2 / 24 * m7;
    const double u5 = 15.0 / 8 * dr * m5 - 2.5 * dr * q * m6 + dr * (7 * dr2 + 6 * q2) / 12 * m7 - dr * dr2 * q / 6 * m8 +
                      dr * dr2 * dr2 / 120 * m9;
    const double u6 = 5.0 / 16 * m5 - 15.0 / 8 * q * m6 + 5 * (7 * dr2 + 4 * q2) / 16 * m7 - q * (21 * dr2 + 2 * q2) / 12 * m8 +
                      dr2 * (3 * dr2 + 4 * q2) / 16 * m9 - dr2 * dr2 * q / 24 * m10 + dr2 * dr2 * dr2 / 720 * m11;
    return eta * eta * eta * eta * (u4 + eta * (u5 + eta * u6));
}

void axis_component(double kbb, double sigx, double sigy, double x, double y, double &k, double &h) {
    const double s = sigx + sigy, sx2 = sigx * sigx, sx3 = sx2 * sigx, sx5 = sx3 * sx2, sy2 = sigy * sigy, sy3 = sy2 * sigy;
    const double s2 = s * s, s3 = s2 * s;
    const double j00 = 2 / (sigx * s), j10 = 2 * (2 * sigx + sigy) / (3 * sx3 * s2), j01 = 2 / (sigx * sigy * s2);
    const double j20 = 2 * (8 * sx2 + 9 * sigx * sigy + 3 * sy2) / (15 * sx5 * s3);
    const double j11 = 2 * (3 * sigx + sigy) / (3 * sx3 * sigy * s3), j02 = 2 * (sigx + 3 * sigy) / (3 * sigx * sy3 * s3);
    const double x2 = x * x, y2 = y * y, x4 = x2 * x2, y4 = y2 * y2, x2y2 = x2 * y2;
    const double scale = j00 - (x2 * j10 + y2 * j01) / 2 + x4 * j20 / 8 + x2y2 * j11 / 4 + y4 * j02 / 8;
    const double derivative = j00 - 1.5 * x2 * j10 - y2 * j01 / 2 + 0.625 * x4 * j20 + 0.75 * x2y2 * j11 + y4 * j02 / 8;
    k = x * scale;
    h = -kbb * derivative;
}

bool near_axis(double s1, double s2, double x, double y, double eta) {
    const double rho2 = x * x / (s1 * s1) + y * y / (s2 * s2);
    return rho2 * rho2 * rho2 * std::sqrt(rho2) <= std::ldexp(1.0, -52) / std::sqrt(eta);
}

Resp response_principal(double kbb, double s1, double s2, double x, double y) {
    const double v = (s1 * s1 + s2 * s2) / 2, eta = (s1 * s1 - s2 * s2) / (2 * v);
    Resp r;
    if (eta == 0.0) {
        const double r2 = x * x + y * y, u = r2 / (2 * v);
        const double scale = (u == 0.0 ? 1.0 : -std::expm1(-u) / u) / v;
        const double expterm = std::exp(-r2 / (2 * v));
        r.kx = scale * x;
        r.ky = scale * y;
        round_hessian(kbb, std::sqrt(v), x, y, expterm, r.h1, r.h2);
        return r;
    }
    const double outer = std::sqrt(std::sqrt(8.0 / 3.0 * 64.0 * std::ldexp(1.0, -52))), inner = outer / 2;
    if (eta <= inner) return series_response(kbb, v, eta, x, y);
    double kex, key;
    elliptic(s1, s2, x, y, eta, kex, key);
    const double expterm = std::exp(-0.5 * (x * x / (s1 * s1) + y * y / (s2 * s2)));
    double he1, he2;
    if (near_axis(s1, s2, x, y, eta)) {
        double k;
        axis_component(kbb, s1, s2, x, y, k, he1);
        axis_component(kbb, s2, s1, y, x, k, he2);
    } else {   // _elliptic_gaussian_hessian_diagonal (strong_beam.jl:695-703)
        const double t1 = kbb * (x * kex + y * key), t2 = s1 * s1 - s2 * s2, t3 = s2 / s1;
        he1 = (t1 - 2 * kbb * (1 - expterm * t3)) / t2;
        he2 = (-t1 + 2 * kbb * (1 - expterm / t3)) / t2;
    }
    if (eta >= outer) return Resp{kex, key, he1, he2};
    const double invw = 1.0 / (outer - inner), t = (eta - inner) * invw;
    const double w = t * t * t * (10 + t * (-15 + 6 * t)), dw = 30 * t * t * (t - 1) * (t - 1) * invw;
    const Resp sr = series_response(kbb, v, eta, x, y);
    const double chain = kbb * dw * potential_residual(v, eta, x, y) / v;
    r.kx = sr.kx + w * (kex - sr.kx);
    r.ky = sr.ky + w * (key - sr.ky);
    r.h1 = sr.h1 + w * (he1 - sr.h1) + (1 - eta) * chain;
    r.h2 = sr.h2 + w * (he2 - sr.h2) - (1 + eta) * chain;
    return r;
}

Resp beambeam_response(double kbb, double sigx, double sigy, double x, double y) {
    if (sigx >= sigy) return response_principal(kbb, sigx, sigy, x, y);
    const Resp r = response_principal(kbb, sigy, sigx, y, x);
    return Resp{r.ky, r.kx, r.h2, r.h1};
}

void be_response4(double kbb, double sigx, double sigy, double x, double y, double out[4]) {
    const Resp r = beambeam_response(kbb, sigx, sigy, x, y);
    out[0] = r.kx; out[1] = r.ky; out[2] = r.h1; out[3] = r.h2;
}

// ---- GaussianPICPoissonSolver, the separable (uncoupled) subtraction (gaussian_pic.jl:374-491, 543-595,
// 646-892)
struct GMom { int64_t n; double mx, mpx, varx, cxpx, varpx, my, mpy, vary, cypy, varpy; };
struct GBnd { double mux, muy, sigx, sigy, rx, ry; };

GMom gpic_moments(const Coords &src) {   // _gpic_source_moments, uncoupled
    const int64_t n = (int64_t)src.size();
    const double x0 = src.x[0], px0 = src.px[0], y0 = src.y[0], py0 = src.py[0];
    double sx = 0, spx = 0, sy = 0, spy = 0, sxx = 0, spxpx = 0, syy = 0, spypy = 0, sxpx = 0, sypy = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double dx = src.x[i] - x0, dpx = src.px[i] - px0, dy = src.y[i] - y0, dpy = src.py[i] - py0;
        sx += dx; spx += dpx; sy += dy; spy += dpy;
        sxx += dx * dx; spxpx += dpx * dpx; syy += dy * dy; spypy += dpy * dpy;
        sxpx += dx * dpx; sypy += dy * dpy;
    }
    const double invn = 1.0 / (double)n;
    const double dmx = sx * invn, dmpx = spx * invn, dmy = sy * invn, dmpy = spy * invn;
    // _shifted_second_moment / _shifted_cross_moment (slicing.jl:664-668): muladd(-a, b, sum * invn)
    auto shifted = [&](double sum, double a, double b) { return std::fma(-a, b, sum * invn); };
    GMom m;
    m.n = n;
    m.mx = x0 + dmx; m.mpx = px0 + dmpx; m.my = y0 + dmy; m.mpy = py0 + dmpy;
    m.varx = std::max(shifted(sxx, dmx, dmx), 0.0);
    m.vary = std::max(shifted(syy, dmy, dmy), 0.0);
    m.cxpx = shifted(sxpx, dmx, dmpx);
    m.varpx = std::max(shifted(spxpx, dmpx, dmpx), 0.0);
    m.cypy = shifted(sypy, dmy, dmpy);
    m.varpy = std::max(shifted(spypy, dmpy, dmpy), 0.0);
    return m;
}

GBnd gpic_boundary(const GMom &m, double s) {   // _gpic_drifted_gaussian + _gpic_boundary, uncoupled
    GBnd b;
    b.mux = m.mx + m.mpx * s;
    b.muy = m.my + m.mpy * s;
    const double vx = m.varx + 2 * s * m.cxpx + s * s * m.varpx;
    const double vy = m.vary + 2 * s * m.cypy + s * s * m.varpy;
    b.sigx = std::sqrt(std::max(vx, 0.0));
    b.sigy = std::sqrt(std::max(vy, 0.0));
    b.rx = 2 * (m.cxpx + s * m.varpx);
    b.ry = 2 * (m.cypy + s * m.varpy);
    return b;
}

// _gpic_solve_drifted_field! (gaussian_pic.jl:543-595): deposit, subtract the neutralised erf profile of the
// drifted Gaussian, solve the residual.
void gpic_solve(Context &C, Workspace &w, FieldWS &f, const Coords &src, double s, const Grid &sg,
                const std::vector<cd> &gfft, const GBnd &b) {
    const Opt &o = C.o;
    const int nx = o.nx, ny = o.ny;
    const int64_t P1 = 2 * nx, P2 = 2 * ny;
    const double hx = sg.width / (double)(nx - 1), hy = sg.height / (double)(ny - 1);
    std::fill(w.charge.begin(), w.charge.end(), 0.0);
    deposit(w.charge.data(), P1, P2, o.deposit, src.x.data(), src.px.data(), src.y.data(), src.py.data(), s, true,
            sg.x0, sg.y0, hx, hy, nx, ny, (int64_t)src.size(), &w.local_charge);
    std::vector<double> gx(nx), gy(ny);
    gaussian_profile(gx.data(), nx, sg.x0, hx, b.mux, b.sigx, o.deposit);
    gaussian_profile(gy.data(), ny, sg.y0, hy, b.muy, b.sigy, o.deposit);
    double sgx = 0, sgy = 0, qsum = 0;
    for (double v : gx) sgx += v;
    for (double v : gy) sgy += v;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) qsum += w.charge[i + (size_t)j * P1];
    double amp = (double)src.size();
    if (o.neutralize && sgx * sgy > 0.0) amp = qsum / (sgx * sgy);
    for (int j = 0; j < ny; ++j) {
        const double gj = amp * gy[j];
        for (int i = 0; i < nx; ++i) w.charge[i + (size_t)j * P1] -= gx[i] * gj;
    }
    for (size_t k = 0; k < w.charge.size(); ++k) w.spectral[k] = cd(w.charge[k], 0.0);
    fft2(w.spectral.data(), (int)P1, (int)P2, -1, *C.plans.p1, *C.plans.p2);
    for (size_t k = 0; k < w.spectral.size(); ++k) w.spectral[k] *= gfft[k];
    fft2(w.spectral.data(), (int)P1, (int)P2, +1, *C.plans.p1, *C.plans.p2);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) f.phi[i + (size_t)j * nx] = w.spectral[i + (size_t)j * P1].real();
    field_derivative(f.Ex.data(), f.Ey.data(), f.phi.data(), nx, ny, hx, hy, o.field_derivative == 4);
}

// _gpic_interaction! (gaussian_pic.jl:646-816) with _gpic_apply_kick_range! (:818-892), uncoupled.
void gpic_interaction(Context &C, Workspace &w, const Coords &src, const SliceParam &ps, Coords &fld, const SliceParam &pf,
                      double kbb, const std::tuple<int, int, int> *key, Vec &vx, Vec &vy, octopic_interaction_record *rec) {
    const Opt &o = C.o;
    const int64_t nsource = (int64_t)src.size(), nfield = (int64_t)fld.size();
    const double sL = 0.5 * (ps.center - pf.lb), sR = 0.5 * (ps.center - pf.rb);
    const GMom mom = gpic_moments(src);
    const GBnd bL = gpic_boundary(mom, sL), bR = gpic_boundary(mom, sR);
    const bool resolved = nsource >= 2 && std::isfinite(bL.sigx) && std::isfinite(bL.sigy) && std::isfinite(bR.sigx) &&
                          std::isfinite(bR.sigy) && bL.sigx > 0 && bL.sigy > 0 && bR.sigx > 0 && bR.sigy > 0;
    if (!resolved) {   // _gpic_control_variate_mode -> :pic: the whole interaction is plain PIC
        interaction(C, w, src, ps, fld, pf, kbb, key, nullptr, vx, vy, rec);
        return;
    }
    const double margin = o.margin;
    double sxmin = jmin(src.x[0] + src.px[0] * sL, src.x[0] + src.px[0] * sR);
    double sxmax = jmax(src.x[0] + src.px[0] * sL, src.x[0] + src.px[0] * sR);
    double symin = jmin(src.y[0] + src.py[0] * sL, src.y[0] + src.py[0] * sR);
    double symax = jmax(src.y[0] + src.py[0] * sL, src.y[0] + src.py[0] * sR);
    for (int64_t i = 1; i < nsource; ++i) {
        const double xl = src.x[i] + src.px[i] * sL, yl = src.y[i] + src.py[i] * sL;
        const double xr = src.x[i] + src.px[i] * sR, yr = src.y[i] + src.py[i] * sR;
        sxmin = jmin(jmin(sxmin, xl), xr); sxmax = jmax(jmax(sxmax, xl), xr);
        symin = jmin(jmin(symin, yl), yr); symax = jmax(jmax(symax, yl), yr);
    }
    if (margin > 0) {
        const double gxlo = jmin(bL.mux - margin * bL.sigx, bR.mux - margin * bR.sigx);
        const double gxhi = jmax(bL.mux + margin * bL.sigx, bR.mux + margin * bR.sigx);
        const double gylo = jmin(bL.muy - margin * bL.sigy, bR.muy - margin * bR.sigy);
        const double gyhi = jmax(bL.muy + margin * bL.sigy, bR.muy + margin * bR.sigy);
        sxmin = jmin(sxmin, gxlo); sxmax = jmax(sxmax, gxhi);
        symin = jmin(symin, gylo); symax = jmax(symax, gyhi);
    }
    if (!(std::isfinite(sxmin) && std::isfinite(sxmax) && std::isfinite(symin) && std::isfinite(symax)))
        nonfinite("source", "slice pair");
    const double c = ps.center;
    double fxmin = fld.x[0] + 0.5 * (fld.z[0] - c) * fld.px[0], fxmax = fxmin;
    double fymin = fld.y[0] + 0.5 * (fld.z[0] - c) * fld.py[0], fymax = fymin;
    for (int64_t i = 0; i < nfield; ++i) {
        const double s = 0.5 * (fld.z[i] - c);
        fld.x[i] += s * fld.px[i];
        fld.y[i] += s * fld.py[i];
        if (o.lkick) fld.pz[i] -= 0.25 * (fld.px[i] * fld.px[i] + fld.py[i] * fld.py[i]);
        fxmin = jmin(fxmin, fld.x[i]); fxmax = jmax(fxmax, fld.x[i]);
        fymin = jmin(fymin, fld.y[i]); fymax = jmax(fymax, fld.y[i]);
    }
    if (!(std::isfinite(fxmin) && std::isfinite(fxmax) && std::isfinite(fymin) && std::isfinite(fymax)))
        nonfinite("field", "slice pair");
    Grid sg, fg;
    interaction_grids(o, sxmin, sxmax, symin, symax, fxmin, fxmax, fymin, fymax, sg, fg);
    Box sb{sxmin, sxmax, symin, symax}, fb{fxmin, fxmax, fymin, fymax};
    std::vector<cd> tmp;
    int event = 0;
    const std::vector<cd> *gfft = slice_pair_green(C, w, key, sg, fg, sb, fb, tmp, event);
    if (rec) {
        rec->cache_event = event;
        rec->source_grid[0] = sg.x0; rec->source_grid[1] = sg.y0; rec->source_grid[2] = sg.width; rec->source_grid[3] = sg.height;
        rec->field_grid[0] = fg.x0; rec->field_grid[1] = fg.y0; rec->field_grid[2] = fg.width; rec->field_grid[3] = fg.height;
    }
    gpic_solve(C, w, w.left, src, sL, sg, *gfft, bL);
    gpic_solve(C, w, w.right, src, sR, sg, *gfft, bR);
    const double kick_scale = 2.0 * kbb, half_ns = 0.5 * (double)nsource, kbb_eff = kick_scale * half_ns;
    double hzi, zbias;
    slice_interp_params(pf.lb, pf.rb, hzi, zbias);
    for (int64_t i = 0; i < nfield; ++i) {
        const double x = fld.x[i], y = fld.y[i];
        const double zL = jclamp(-fld.z[i] * hzi + zbias, 0.0, 1.0), zR = 1.0 - zL;
        double Kx, Ky, Kz;
        interpolate_kick(o, fg, x, y, w.left, w.right, zL, zR, Kx, Ky, Kz);
        Resp rl, rr;
        if (o.lkick) {
            rl = beambeam_response(kbb_eff, bL.sigx, bL.sigy, x - bL.mux, y - bL.muy);
            rr = beambeam_response(kbb_eff, bR.sigx, bR.sigy, x - bR.mux, y - bR.muy);
        } else {
            rl = Resp{0, 0, 0, 0};
            rr = Resp{0, 0, 0, 0};
            be_kick(bL.sigx, bL.sigy, x - bL.mux, y - bL.muy, rl.kx, rl.ky);
            be_kick(bR.sigx, bR.sigy, x - bR.mux, y - bR.muy, rr.kx, rr.ky);
        }
        const double Kxa = half_ns * (zL * rl.kx + zR * rr.kx), Kya = half_ns * (zL * rl.ky + zR * rr.ky);
        const double dpxa = kick_scale * Kxa, dpya = kick_scale * Kya;
        fld.px[i] += kick_scale * Kx + dpxa;
        fld.py[i] += kick_scale * Ky + dpya;
        if (o.lkick) {
            const double covL = (rl.h1 * bL.rx + rl.h2 * bL.ry) / 4, covR = (rr.h1 * bR.rx + rr.h2 * bR.ry) / 4;
            fld.pz[i] += kick_scale * Kz * hzi;
            fld.pz[i] += zL * covL + zR * covR;
            fld.pz[i] += 0.5 * (dpxa * mom.mpx + dpya * mom.mpy);
        }
        const double s = 0.5 * (c - fld.z[i]);
        fld.x[i] += s * fld.px[i];
        fld.y[i] += s * fld.py[i];
        if (o.lkick) fld.pz[i] += 0.25 * (fld.px[i] * fld.px[i] + fld.py[i] * fld.py[i]);
    }
    const double sM = 0.5 * (ps.center - pf.center);
    vx.resize(nsource); vy.resize(nsource);
    for (int64_t i = 0; i < nsource; ++i) {
        vx[i] = src.x[i] + src.px[i] * sM;
        vy[i] = src.y[i] + src.py[i] * sM;
    }
}

}  // namespace

// ==========================================================================================
// C API of the oracle (ctypes; test infrastructure)
// ==========================================================================================
extern "C" {

void oracle_faddeeva_w(double zr, double zi, double *wr, double *wi) {
    const std::complex<double> w = faddeeva(zr, zi);
    *wr = w.real();
    *wi = w.imag();
}

void oracle_gaussian_beambeam_kick(double sigx, double sigy, double x, double y, double *kx, double *ky) {
    be_kick(sigx, sigy, x, y, *kx, *ky);
}

// _gaussian_beambeam_kick_response(kbb, sigx, sigy, x, y) (strong_beam.jl:1119-1128): the normalised kick
// and the diagonal of the potential's Hessian times kbb.
void oracle_gaussian_beambeam_kick_response(double kbb, double sigx, double sigy, double x, double y, double out[4]) {
    const Resp r = beambeam_response(kbb, sigx, sigy, x, y);
    out[0] = r.kx; out[1] = r.ky; out[2] = r.h1; out[3] = r.h2;
}

void oracle_weideman_coefficients(double *c, double *L) {
    std::memcpy(c, oracle_weideman().c, sizeof(double) * 32);
    *L = oracle_weideman().L;
}

const char *oracle_last_error(void) { return g_last_error.c_str(); }

int oracle_philox_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox_block(c, key[0], key[1]);
    for (int k = 0; k < 4; ++k) out[k] = c[k];
    return 0;
}
uint64_t oracle_counter_uint64(uint64_t seed, uint64_t turn, uint64_t rng_id, uint64_t particle, uint64_t component) {
    return counter_uint64(seed, turn, rng_id, particle, component);
}
double oracle_octopus_normal(uint64_t seed, uint64_t turn, uint64_t rng_id, uint64_t particle, int64_t component) {
    return octopus_normal(seed, turn, rng_id, particle, component);
}

// Beam(N, ...; beta, alpha, sigma, cutoff, rng_id) with the counter RNG (Beam.jl:127-181, 466-515, 545-553).
// Particle indices are 1-based in the counter, as in the reference.
int oracle_generate_beam(int64_t N, uint64_t seed, uint64_t rng_id, const double beta[3], const double alpha[3],
                         const double sigma[3], double cutoff, double *out[6]) {
    return guarded([&] {
        std::vector<Vec> c(6, Vec(N));
        for (int d = 0; d < 6; ++d) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
            for (int64_t i = 0; i < N; ++i) {
                double x = octopus_normal(seed, 0, rng_id, (uint64_t)(i + 1), d + 1);
                c[d][i] = std::isfinite(cutoff) ? jclamp(x, -cutoff, cutoff) : x;
            }
            standardize(c[d]);
        }
        for (int k = 0; k < 3; ++k) {
            Vec &q = c[2 * k], &p = c[2 * k + 1];
            double f = sigma[k] / beta[k];
            for (int64_t i = 0; i < N; ++i) p[i] = f * (p[i] - alpha[k] * q[i]);
            for (int64_t i = 0; i < N; ++i) q[i] *= sigma[k];
        }
        for (int d = 0; d < 6; ++d) std::memcpy(out[d], c[d].data(), sizeof(double) * N);
    });
}

double oracle_erfinv(double y) { return erfinv(y); }

int oracle_weights(int32_t method, double u, int64_t n, int64_t *base, double w[3]) {
    int64_t b;
    weights(method, u, n, b, w);
    *base = b;
    return 0;
}
double oracle_kernel_integral(double x, double y) { return kernel_integral(x, y); }
double oracle_atan_ratio(double n, double d) { return atan_ratio(n, d); }

int oracle_green_table(int32_t gtype, double fx0, double fy0, double sx0, double sy0, double hx, double hy, int32_t nx,
                       int32_t ny, double *out) {
    green_table(out, gtype, fx0, fy0, sx0, sy0, hx, hy, nx, ny);
    return 0;
}

// Column-major n1 x n2 complex (interleaved re,im) in-place 2D DFT.
int oracle_fft2(double *data, int32_t n1, int32_t n2, int32_t sign) {
    FFTPlan p1(n1), p2(n2);
    fft2(reinterpret_cast<cd *>(data), n1, n2, sign < 0 ? -1 : +1, p1, p2);
    return 0;
}

int oracle_field(const double *phi, int32_t nx, int32_t ny, double hx, double hy, int32_t fourth, double *Ex, double *Ey) {
    field_derivative(Ex, Ey, phi, nx, ny, hx, hy, fourth != 0);
    return 0;
}

int oracle_interaction_grids(const octopic_options *opt, const double b[8], double sg[4], double fg[4]) {
    return guarded([&] {
        Opt o = to_opt(*opt);
        Grid s, f;
        interaction_grids(o, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], s, f);
        sg[0] = s.x0; sg[1] = s.y0; sg[2] = s.width; sg[3] = s.height;
        fg[0] = f.x0; fg[1] = f.y0; fg[2] = f.width; fg[3] = f.height;
    });
}

// Undrifted deposit onto an ld x ncols column-major grid (_pic_deposit!, pic_cpu.jl:1992-2013).
int oracle_deposit(double *charge, int64_t ld, int64_t ncols, int32_t method, const double *x, const double *y, int64_t n,
                   double x0, double y0, double hx, double hy, int64_t nx, int64_t ny) {
    deposit(charge, ld, ncols, method, x, nullptr, y, nullptr, 0.0, false, x0, y0, hx, hy, nx, ny, n, nullptr);
    return 0;
}

// Field solve of one plane on a given source/field grid pair (_pic_solve_field, pic_cpu.jl:1907-1915):
// deposit (x, y) + s*(px, py) on sg, convolve with the Green function of (sg, fg); phi/Ex/Ey nx x ny column-major.
int oracle_solve_field(const octopic_options *opt, const double *x, const double *px, const double *y, const double *py,
                       int64_t n, double s, const double sg_[4], const double fg_[4], double *phi, double *Ex, double *Ey,
                       double *charge_out) {
    return guarded([&] {
        Opt o = to_opt(*opt);
        Context C(o);
        Workspace w(o.nx, o.ny, o.lum_nx, o.lum_ny);
        Grid sg{sg_[0], sg_[1], sg_[2], sg_[3]}, fg{fg_[0], fg_[1], fg_[2], fg_[3]};
        std::vector<cd> g = green_fft(C, w, sg, fg);
        Coords src{Vec(x, x + n), Vec(px, px + n), Vec(y, y + n), Vec(py, py + n), Vec(n, 0.0), Vec(n, 0.0)};
        solve_drifted_field(C, w, w.left, src, s, sg, g);
        std::memcpy(phi, w.left.phi.data(), sizeof(double) * o.nx * o.ny);
        std::memcpy(Ex, w.left.Ex.data(), sizeof(double) * o.nx * o.ny);
        std::memcpy(Ey, w.left.Ey.data(), sizeof(double) * o.nx * o.ny);
        if (charge_out) std::memcpy(charge_out, w.charge.data(), sizeof(double) * 4 * o.nx * o.ny);
    });
}

int oracle_interpolate_kick(const octopic_options *opt, const double grid[4], double x, double y, const double *phiL,
                            const double *ExL, const double *EyL, const double *phiR, const double *ExR,
                            const double *EyR, double zL, double zR, double K[3]) {
    return guarded([&] {
        Opt o = to_opt(*opt);
        size_t m = (size_t)o.nx * o.ny;
        FieldWS L{Vec(phiL, phiL + m), Vec(ExL, ExL + m), Vec(EyL, EyL + m)};
        FieldWS R{Vec(phiR, phiR + m), Vec(ExR, ExR + m), Vec(EyR, EyR + m)};
        interpolate_kick(o, Grid{grid[0], grid[1], grid[2], grid[3]}, x, y, L, R, zL, zR, K[0], K[1], K[2]);
    });
}

int oracle_interpolate_kick_quadratic(const octopic_options *opt, const double grid[4], double x, double y,
                                      const double *const planes[9], double t, double K[3]) {
    return guarded([&] {
        Opt o = to_opt(*opt);
        size_t m = (size_t)o.nx * o.ny;
        FieldWS L{Vec(planes[0], planes[0] + m), Vec(planes[1], planes[1] + m), Vec(planes[2], planes[2] + m)};
        FieldWS M{Vec(planes[3], planes[3] + m), Vec(planes[4], planes[4] + m), Vec(planes[5], planes[5] + m)};
        FieldWS R{Vec(planes[6], planes[6] + m), Vec(planes[7], planes[7] + m), Vec(planes[8], planes[8] + m)};
        interpolate_kick_quadratic(o, Grid{grid[0], grid[1], grid[2], grid[3]}, x, y, L, M, R, t, K[0], K[1], K[2]);
    });
}

int oracle_luminosity(const octopic_options *opt, const double *x1, const double *y1, int64_t n1, const double *x2,
                      const double *y2, int64_t n2, double klum, double *lum, double *q1_out, double *q2_out) {
    return guarded([&] {
        Opt o = to_opt(*opt);
        Vec q1((size_t)(o.lum_nx + 1) * (o.lum_ny + 1)), q2(q1.size());
        *lum = luminosity(o, Vec(x1, x1 + n1), Vec(y1, y1 + n1), Vec(x2, x2 + n2), Vec(y2, y2 + n2), klum, q1, q2);
        if (q1_out) std::memcpy(q1_out, q1.data(), sizeof(double) * q1.size());
        if (q2_out) std::memcpy(q2_out, q2.data(), sizeof(double) * q2.size());
    });
}

// longitudinal_slices: boundaries[ns+1], centers[ns], weights[ns], counts[ns], indices[n] (concatenated, 0-based)
int oracle_slices(const double *z, int64_t n, const octopic_slicing *s, double *boundaries, double *centers,
                  double *weights_out, int64_t *counts, int64_t *indices) {
    return guarded([&] {
        Slicing sl = to_slicing(*s);
        validate_slicing(sl);
        Slices r = longitudinal_slices(z, n, sl);
        std::memcpy(boundaries, r.boundary.data(), sizeof(double) * r.boundary.size());
        std::memcpy(centers, r.center.data(), sizeof(double) * r.center.size());
        std::memcpy(weights_out, r.weight.data(), sizeof(double) * r.weight.size());
        int64_t off = 0;
        for (size_t k = 0; k < r.indices.size(); ++k) {
            counts[k] = (int64_t)r.indices[k].size();
            for (int64_t v : r.indices[k]) indices[off++] = v;
        }
    });
}

// collision_pair_batches: out_i/out_j/out_batch have ns1*ns2 entries (0-based); returns number of batches
int oracle_collision_batches(const double *c1, int32_t ns1, const double *c2, int32_t ns2, int32_t *out_i, int32_t *out_j,
                             int32_t *out_batch, int32_t *nbatches) {
    return guarded([&] {
        auto b = collision_pair_batches(Vec(c1, c1 + ns1), Vec(c2, c2 + ns2));
        int k = 0;
        for (int q = 0; q < (int)b.size(); ++q)
            for (auto &e : b[q]) { out_i[k] = e.i; out_j[k] = e.j; out_batch[k] = q; ++k; }
        *nbatches = (int32_t)b.size();
    });
}
int oracle_collision_order(const double *c1, int32_t ns1, const double *c2, int32_t ns2, int32_t *out_i, int32_t *out_j) {
    auto o = collision_order(Vec(c1, c1 + ns1), Vec(c2, c2 + ns2));
    for (size_t k = 0; k < o.size(); ++k) { out_i[k] = o[k].i; out_j[k] = o[k].j; }
    return 0;
}

// Persistent collider state = the task's runtime cache (Green cache + workspaces).
int oracle_create(const octopic_options *opt, void **out) {
    *out = nullptr;
    return guarded([&] {
        Opt o = to_opt(*opt);
        State *S = new State();
        S->C.reset(new Context(o));
        *out = S;
    });
}
void oracle_destroy(void *s) { delete static_cast<State *>(s); }
// Test mode: every field deposit replaced by its neutralised drifted Gaussian (what the device's
// denoised collide gives with a network that outputs zero).
int oracle_set_gaussian_replace(void *s, int32_t on) {
    static_cast<State *>(s)->C->gaussian_replace = on ? 1 : 0;
    return 0;
}

int oracle_collide(void *s, double *b1[6], int64_t n1, const double p1[5], double *b2[6], int64_t n2, const double p2[5],
                   int32_t compute_lum, int32_t nthreads, double *lum) {
    return guarded([&] {
        State &S = *static_cast<State *>(s);
        collide(S, b1, n1, p1, b2, n2, p2, compute_lum, nthreads);
        *lum = S.last.lum;
    });
}

int oracle_stats(void *s, int64_t out[6]) {
    State &S = *static_cast<State *>(s);
    out[0] = S.C->cache.hits; out[1] = S.C->cache.misses; out[2] = S.C->cache.rebuilds;
    out[3] = S.last.dropped; out[4] = S.last.batches; out[5] = S.last.widest;
    return 0;
}

// The last collide's :node mesh of node `node` (0-based boundary index) for source slice `slice` of
// direction `dir` (1: beam 1 is the source): source grid then field grid, (x0, y0, width, height) each.
// Returns 1, or 0 when that node has no mesh (no field particle in its adjacent slices).
int oracle_node_mesh(void *s, int dir, int slice, int node, double *out8) {
    State &S = *static_cast<State *>(s);
    auto it = S.node_meshes.find(std::make_tuple(dir, slice, node));
    if (it == S.node_meshes.end()) return 0;
    const Grid &a = it->second.first, &b = it->second.second;
    const double v[8] = {a.x0, a.y0, a.width, a.height, b.x0, b.y0, b.width, b.height};
    for (int k = 0; k < 8; ++k) out8[k] = v[k];
    return 1;
}

int64_t oracle_npairs(void *s) { return (int64_t)static_cast<State *>(s)->last.pi.size(); }

int oracle_pair_luminosity(void *s, int32_t *i, int32_t *j, int32_t *ran, double *lum) {
    State &S = *static_cast<State *>(s);
    for (size_t k = 0; k < S.last.pi.size(); ++k) { i[k] = S.last.pi[k]; j[k] = S.last.pj[k]; ran[k] = S.last.ran[k]; lum[k] = S.last.plum[k]; }
    return 0;
}

int oracle_records(void *s, octopic_interaction_record *out, int64_t cap, int64_t *n) {
    State &S = *static_cast<State *>(s);
    *n = (int64_t)S.records.size();
    for (int64_t k = 0; k < std::min<int64_t>(cap, *n); ++k) out[k] = S.records[k];
    return 0;
}

int oracle_last_slices(void *s, int32_t which, double *boundaries, double *centers, double *weights_out, int64_t *counts,
                       int64_t *indices) {
    State &S = *static_cast<State *>(s);
    const Slices &r = which == 1 ? S.last1 : S.last2;
    std::memcpy(boundaries, r.boundary.data(), sizeof(double) * r.boundary.size());
    std::memcpy(centers, r.center.data(), sizeof(double) * r.center.size());
    std::memcpy(weights_out, r.weight.data(), sizeof(double) * r.weight.size());
    int64_t off = 0;
    for (size_t k = 0; k < r.indices.size(); ++k) {
        counts[k] = (int64_t)r.indices[k].size();
        for (int64_t v : r.indices[k]) indices[off++] = v;
    }
    return 0;
}

// Linear6D from optics with zero coupling/dispersion (linear6d.jl:233-276); out is row-major 6x6.
int oracle_linear6d_matrix(const double beta1[3], const double beta2[3], const double alpha1[3], const double alpha2[3],
                           const double dmu[3], double out[36]) {
    for (int k = 0; k < 36; ++k) out[k] = 0.0;
    for (int i = 0; i < 3; ++i) {
        double s = std::sin(dmu[i]), c = std::cos(dmu[i]);
        int r = 2 * i;
        out[r * 6 + r] = std::sqrt(beta2[i] / beta1[i]) * (c + alpha1[i] * s);
        out[r * 6 + r + 1] = std::sqrt(beta2[i] * beta1[i]) * s;
        out[(r + 1) * 6 + r] = (-(1.0 + alpha1[i] * alpha2[i]) * s + (alpha1[i] - alpha2[i]) * c) / std::sqrt(beta1[i] * beta2[i]);
        out[(r + 1) * 6 + r + 1] = std::sqrt(beta1[i] / beta2[i]) * (c - alpha2[i] * s);
    }
    return 0;
}

// track_particle(::Symplectic6DMap, ::Linear6D, ...) (linear6d.jl:110-118)
int oracle_track_linear6d(const double m[36], double *b[6], int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        double v[6] = {b[0][i], b[1][i], b[2][i], b[3][i], b[4][i], b[5][i]};
        for (int r = 0; r < 6; ++r) {
            const double *q = m + 6 * r;
            b[r][i] = q[0] * v[0] + q[1] * v[1] + q[2] * v[2] + q[3] * v[3] + q[4] * v[4] + q[5] * v[5];
        }
    }
    return 0;
}

// track_particle(::NonSymplectic6DMap, ::LorentzBoost / ::RevLorentzBoost, ...) (lorentz_boost.jl:128-166):
// Hirata's crossing-angle transformation with `angle` = half the full crossing angle; cos, sin and
// tan = sin/cos as _lorentz_boost (lorentz_boost.jl:168-174).
int oracle_track_lorentz_boost(double angle, int reverse, double *b[6], int64_t n) {
    const double c = std::cos(angle), s = std::sin(angle), t = s / c;
    for (int64_t i = 0; i < n; ++i) {
        const double x0 = b[0][i], px0 = b[1][i], y0 = b[2][i], py0 = b[3][i], z0 = b[4][i], pz0 = b[5][i];
        const double ps0 = std::sqrt((1.0 + pz0) * (1.0 + pz0) - px0 * px0 - py0 * py0);
        const double h0 = 1.0 + pz0 - ps0;
        double x1, px1, y1, py1, z1, pz1;
        if (!reverse) {
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
        b[0][i] = x1; b[1][i] = px1; b[2][i] = y1; b[3][i] = py1; b[4][i] = z1; b[5][i] = pz1;
    }
    return 0;
}

// track_particle(::Symplectic6DMap, ::ThinCrabCavity{N}, ...) (crab_cavity.jl:147-164) with
// kcc = 2 pi f / c (crab_cavity.jl:99-111): per harmonic i, theta = i kcc z + phase_i,
// px -= sx_i sin/(i kcc), py -= sy_i sin/(i kcc), pz -= (sx_i x + sy_i y) cos.
int oracle_track_crab_cavity(double kcc, int nh, const double *sx, const double *sy, const double *phase, double *b[6],
                             int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const double x0 = b[0][i], y0 = b[2][i], z0 = b[4][i];
        double dpx = 0.0, dpy = 0.0, dpz = 0.0;
        for (int k = 1; k <= nh; ++k) {
            const double ikcc = (double)k * kcc;
            const double th = ikcc * z0 + phase[k - 1];
            const double sn = std::sin(th), cs = std::cos(th);
            dpx -= sx[k - 1] * sn / ikcc;
            dpy -= sy[k - 1] * sn / ikcc;
            dpz -= (sx[k - 1] * x0 + sy[k - 1] * y0) * cs;
        }
        b[1][i] = b[1][i] + dpx;
        b[3][i] = b[3][i] + dpy;
        b[5][i] = b[5][i] + dpz;
    }
    return 0;
}

}  // extern "C"
