// octopic_host.cpp -- host orchestration of one collide! and the C ABI of include/octopic.h.
//
// One collide (reference: _pic_collide!, src/tasks/strongstrong/pic_cpu.jl:82-275):
//   1. slicing of both beams on the device (k_slicing.hip), gathering each beam into slice-contiguous
//      state arrays (the reference's _pic_slice_states, pic_cpu.jl:866-867);
//   2. ONE device->host read of the two slice tables, from which the host forms the collision order
//      and the conflict-free wavefront batches (slicing.jl:561-647);
//   3. per batch, a fixed kernel sequence with no host round trip: bounds -> geometry + Green-cache
//      decisions -> Green builds -> deposits -> row FFT -> column FFT x Green -> inverse row FFT ->
//      field -> kick -> luminosity overlap;
//   4. scatter of the states back into the beams and one read of the per-pair luminosities.
// Because every batch's stage writes only what later stages of the same batch (or later batches)
// read, both directions of a pair kick in place: the deposits of both directions have read their
// pre-pair sources before either kick runs, which is what the reference's scratch copy ensures
// (pic_cpu.jl:307-327).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <numeric>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "octopic.h"
#include "common.h"
#include "fft.h"
#include "launch.h"
#include "denoise.h"

using namespace octopic;

namespace {

struct Fail {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &m) { throw Fail{code, m}; }

#define HIPCHECK(call)                                                                                          \
    do {                                                                                                        \
        hipError_t e_ = (call);                                                                                 \
        if (e_ != hipSuccess) {                                                                                 \
            if (e_ == hipErrorOutOfMemory) fail(OCTOPIC_ERR_OOM, std::string("device allocation failed: ") + #call); \
            fail(OCTOPIC_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));                        \
        }                                                                                                       \
    } while (0)

// Grow-only device buffer.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void ensure(size_t m) {
        if (m <= n) return;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        HIPCHECK(hipMalloc((void **)&p, std::max<size_t>(m, 1) * sizeof(T)));
        n = m;
    }
    void ensure_zero(size_t m, hipStream_t s) {
        if (m <= n) return;
        ensure(m);
        HIPCHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// ------------------------------------------------------------------ erfinv for the normal quantiles
// Same published method as the CPU restatement: Giles' seed (GPU Computing Gems, 2011) refined by
// Halley steps on erf/erfc, so the quantile boundaries agree bit for bit with it.
double erfinv_host(double y) {
    if (std::isnan(y) || y < -1.0 || y > 1.0) return NAN;
    if (y == 1.0) return INFINITY;
    if (y == -1.0) return -INFINITY;
    if (y == 0.0) return y;
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
    const double ay = std::fabs(y);
    for (int it = 0; it < 3; ++it) {
        double ax = std::fabs(x);
        double f = ay > 0.5 ? (1.0 - ay) - std::erfc(ax) : std::erf(ax) - ay;
        double d = two_over_sqrtpi * std::exp(-ax * ax);
        double step = f / d;
        step = step / (1.0 + ax * step);
        ax = ax - step;
        x = std::copysign(ax, y);
    }
    return x;
}

bool jisless_h(double a, double b) { return jisless(a, b); }

struct OrderEntry { double t; int i, j; };

std::vector<OrderEntry> collision_order(const double *c1, int ns1, const double *c2, int ns2) {  // slicing.jl:561-568
    std::vector<OrderEntry> order;
    order.reserve((size_t)ns1 * ns2);
    for (int i = 0; i < ns1; ++i)
        for (int j = 0; j < ns2; ++j) order.push_back({-(c1[i] + c2[j]) / 2.0, i, j});
    std::stable_sort(order.begin(), order.end(),
                     [](const OrderEntry &a, const OrderEntry &b) { return jisless_h(a.t, b.t); });
    return order;
}

// collision_pair_batches (slicing.jl:601-647), as positions into `order`
std::vector<std::vector<int>> pair_batches(const std::vector<OrderEntry> &order, int ns1, int ns2) {
    std::vector<std::vector<int>> batches;
    std::vector<std::vector<int>> by_i(ns1), by_j(ns2);
    for (int k = 0; k < (int)order.size(); ++k) { by_i[order[k].i].push_back(k); by_j[order[k].j].push_back(k); }
    std::vector<int> next_i(ns1, 0), next_j(ns2, 0);
    std::vector<char> done(order.size(), 0);
    int remaining = (int)order.size();
    while (remaining > 0) {
        std::vector<int> cur;
        std::vector<char> ui(ns1, 0), uj(ns2, 0);
        for (int k = 0; k < (int)order.size(); ++k) {
            if (done[k]) continue;
            const auto &p = order[k];
            if (ui[p.i] || uj[p.j]) continue;
            bool ri = next_i[p.i] < (int)by_i[p.i].size() && by_i[p.i][next_i[p.i]] == k;
            bool rj = next_j[p.j] < (int)by_j[p.j].size() && by_j[p.j][next_j[p.j]] == k;
            if (ri && rj) {
                cur.push_back(k); ui[p.i] = 1; uj[p.j] = 1; done[k] = 1;
                next_i[p.i]++; next_j[p.j]++; remaining--;
            }
        }
        if (cur.empty()) fail(OCTOPIC_ERR_INVALID_ARGUMENT, "internal collision scheduler error: no ready slice-pair found");
        batches.push_back(std::move(cur));
    }
    return batches;
}

struct FFTPlanHost {
    FFTPlanDev dev{};
    DBuf<double2> tw;
    DBuf<float2> tw32;
    void build(int n) {
        dev.n = n;
        int m = n, k = 0;
        while (m % 4 == 0) { dev.radix[k++] = 4; m /= 4; }
        while (m % 2 == 0) { dev.radix[k++] = 2; m /= 2; }
        while (m % 3 == 0) { dev.radix[k++] = 3; m /= 3; }
        while (m % 5 == 0) { dev.radix[k++] = 5; m /= 5; }
        if (m != 1 || k > kMaxStages)
            fail(OCTOPIC_ERR_UNSUPPORTED, "padded mesh size 2*grid = " + std::to_string(n) +
                                              " has a prime factor other than 2, 3 and 5; this backend's FFT supports "
                                              "grids whose doubled size factors into 2s, 3s and 5s");
        dev.nstages = k;
        std::vector<double2> h(n);
        for (int t = 0; t < n; ++t) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
            h[t] = make_double2((double)cosl(a), (double)sinl(a));
        }
        tw.ensure(n);
        HIPCHECK(hipMemcpy(tw.p, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
        dev.tw = tw.p;
        std::vector<float2> h32(n);
        for (int t = 0; t < n; ++t) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
            h32[t] = make_float2((float)cosl(a), (float)sinl(a));
        }
        tw32.ensure(n);
        HIPCHECK(hipMemcpy(tw32.p, h32.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
        dev.tw32 = tw32.p;
    }
};

const char *kFamilies[] = {"slicing", "bounds", "geometry", "green", "deposit", "fft_rows_fwd", "fft_cols_conv",
                           "fft_rows_inv", "field", "kick", "lum", "store", "node_prepare", "tile_order", "denoise"};
constexpr int kNFamilies = sizeof(kFamilies) / sizeof(kFamilies[0]);
// collides between timed trials of the two deposit forms (deposit_multi -1)
constexpr int64_t kDepositTrialEvery = 256;
enum Fam { F_SLICING, F_BOUNDS, F_GEOMETRY, F_GREEN, F_DEPOSIT, F_ROWS_FWD, F_COLS, F_ROWS_INV, F_FIELD, F_KICK, F_LUM,
           F_STORE, F_NODE, F_TILE, F_DENOISE };

}  // namespace

struct octopic_handle {
    octopic_options opt{};
    std::vector<double> pos1, pos2;   // owned copies of the slicing positions
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;   // created with the handle; blocking with respect to the null stream
    std::string err;
    int nx = 0, ny = 0, lnx = 0, lny = 0, lum_dep = 0, npl = 2, mode = 0;
    bool gauss = false;         // GaussianPICPoissonSolver: Gaussian-subtracted planes with the exact field added back
    bool soft = false;          // GaussianPoissonSolver: the sliced soft-Gaussian kick, no mesh
    DBuf<double> soft_lum_block;   // [pairs][kSoftBlocks] luminosity partials
    DBuf<double> soft_lum_scale;   // [pairs] the sampled direction's klum_slice
    DBuf<double> soft_shift;       // [2 * widest][4] the moment sums' shifts
    DBuf<double> gmom;          // source moment partials of a batch's interactions [2 * widest][chunks][10]
    DBuf<double> gab;           // per-particle add-back of both beams [N1 + N2][4]
    FFTPlanHost px, py;
    // persistent device state
    DBuf<SliceTable> tab;       // 2
    DBuf<double> q;             // 2 * 256 quantiles / positions
    DBuf<ErrorInfo> errinfo;
    DBuf<unsigned long long> counters;   // hits, misses, rebuilds, builds
    DBuf<CacheGeo> cache;
    DBuf<double2> spec;
    int nkeys = 0;
    int epoch = 0;              // the current (or last) collide's number, 1-based: CacheGeo::epoch
    int64_t slot_stride = 0;
    // per-collide buffers
    DBuf<double> beam_in[2][6];
    DBuf<float> beam_f32[2][6];  // host Float32 beams: device staging of the float arrays
    DBuf<double> st[2][6];
    DBuf<int64_t> perm[2];
    int tiles = 16;             // transverse tiles per axis of the in-slice order (0: slice order only)
    // Kick launch shape: each bounds chunk is split over enough blocks for about kick_target_blocks
    // blocks per launch (0: one block per chunk); per-pass LDS windows for chunks whose window does not
    // fit (-1: in the FP64 form only, where they pay; 0 never; 1 such chunks; 2 every chunk).
    int kick_target_blocks = 1024;
    int kick_pass_windows = -1;
    int deposit_span = 8192;    // source particles per deposit block (the LDS mesh is zeroed and flushed once per block)
    int cus = 0;                // compute units of the device (launch sizing)
    // deposit items take all planes of an interaction (slice-pair meshes): 1 always, 0 never, -1 the
    // faster form, both timed on the device every kDepositTrialEvery collides
    int deposit_multi = -1;
    bool deposit_per_plane = false;   // -1: the slower form at the last trial was all planes
    bool deposit_all_planes_used = false;
    int64_t deposit_trial = 0;        // collides since the first trial
    // Green tables built by the previous collide (-1: none yet).  A collide after one that built at least a
    // table per batch launches its Green builds over every job; otherwise over the few that build.
    int64_t last_builds = -1;
    double deposit_trial_ms[2] = {0.0, 0.0};
    std::vector<hipEvent_t> deposit_ev;   // per batch: before and after the deposit launches (trials)
    int deposit_min_span = kChunkSpan;   // smallest items the two-per-CU rule may make
    int fft_group = 4;                   // transforms per workgroup of the charge-path transform passes (2, 4, 8)
    // Kick launch orders and deposit work queues: per-launch counters and launch orders (larger slices first).
    DBuf<int32_t> korder;
    DBuf<int> dqueue;            // [batch][field planes, luminosity planes]
    DBuf<int32_t> dorder;        // per batch: field planes, then luminosity planes
    DBuf<double> srt[2][6];     // tile-ordered states (tiles > 0)
    DBuf<int64_t> perm_srt[2];
    DBuf<int32_t> inv_srt[2];   // inverse of perm_srt
    DBuf<int32_t> inv_sl[2];    // inverse of perm (slice order), for the gathering store without tiles
    DBuf<unsigned long long> rank_keys[2], rank_sorted[2];   // :equal_count rank sort, per beam
    DBuf<int64_t> rank_vals[2];
    DBuf<unsigned long long> box;   // transverse bound slots of both beams
    std::vector<double> hq[2];      // host quantiles / positions of the slicing
    DBuf<int> thist;
    DBuf<uint16_t> sid[2];
    DBuf<int> hist[2];
    DBuf<double> lanes, mm, clanes;     // slicing scratch: one share per beam (the two beams slice concurrently)
    DBuf<unsigned long long> bins;
    int clanes_share = 0;               // slices per beam share of clanes
    // beam 2's slicing runs on side_stream between two events of the handle's stream
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool in_many = false;               // set while octopic_collide_many runs the handle
    DBuf<double *> ptrs;        // [beam_in 2x6][st 2x6]
    DBuf<BatchPair> pairs;
    DBuf<unsigned long long> keys;   // bounds slots [2 * widest batch][kBoundSlots], zero between batches
    DBuf<double> chunkbox;           // per-chunk boxes of a batch [2 * widest batch][kMaxChunks][8]
    DBuf<Inter> inters;
    DBuf<Plane> planes;
    DBuf<GreenJob> jobs;
    DBuf<double> lumgeo;
    DBuf<unsigned long long> charge;
    DBuf<double2> qh;
    DBuf<double> phi, table;
    // FP32 form: half spectra, potentials, kick tables, and the float copy of every cached Green spectrum
    bool fp32 = false;
    DBuf<float2> qh32, spec32;
    DBuf<float> phi32, table32;
    DBuf<octopic_interaction_record> records;
    DBuf<double> lum_parts;
    DBuf<double> lum_partial;
    // node-mode buffers
    DBuf<CacheGeo> nodes;
    DBuf<GreenJob> node_jobs;
    DBuf<unsigned long long> node_keys;   // node-mesh bound slots of both directions
    DBuf<int32_t> ran_dev;
    // pinned host mirrors
    SliceTable *h_tab = nullptr;
    ErrorInfo *h_err = nullptr;
    unsigned long long *h_counters = nullptr;   // hits, misses, rebuilds, builds, dropped, -, denoised, fallbacks
    // residual denoiser of the field planes (octopic_set_denoiser): not owned
    octopic_denoiser *dn = nullptr;
    double dn_limit = 0.0;
    DBuf<double> dn_mom, dn_meta, dn_prof;
    DBuf<float> dn_in;
    // last-collide results
    int ns[2] = {0, 0};
    int64_t nbeam[2] = {0, 0};
    std::vector<int32_t> pi, pj, pran;
    std::vector<double> plum;
    bool unfinished = false;    // a collide left early (error): consume-and-zero buffers must be cleared
    std::vector<octopic_interaction_record> recs;
    octopic_stats stats{};
    // timing
    bool timing = false;
    unsigned fam_mask = ~0u;     // families timed when timing is on
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;
    double fam_ms[kNFamilies] = {0};
    int64_t fam_n[kNFamilies] = {0};

    bool timed(int fam) const { return timing && ((fam_mask >> fam) & 1u); }
    void tic(int fam, hipEvent_t &a) {
        if (!timed(fam)) return;
        a = next_event();
        hipEventRecord(a, stream);
    }
    void toc(int fam, hipEvent_t a) {
        if (!timed(fam)) return;
        hipEvent_t b = next_event();
        hipEventRecord(b, stream);
        ev.push_back({fam, {a, b}});
    }
    hipEvent_t next_event() {
        if (ev_next >= ev_pool.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            ev_pool.push_back(e);
        }
        return ev_pool[ev_next++];
    }
};

namespace {

struct Timed {
    octopic_handle *h;
    int fam;
    hipEvent_t a = nullptr;
    Timed(octopic_handle *h_, int f) : h(h_), fam(f) { h->tic(fam, a); }
    ~Timed() { h->toc(fam, a); }
};

int set_err(octopic_handle *h, int code, const std::string &m) {
    if (h) h->err = m;
    return code;
}

void validate_slicing(const octopic_slicing &s, int beam) {
    const std::string b = "slicing" + std::to_string(beam) + ": ";
    if (s.nslices <= 0) fail(OCTOPIC_ERR_INVALID_OPTION, b + "nslices must be positive");
    if (s.nslices > kMaxSlices) fail(OCTOPIC_ERR_UNSUPPORTED, b + "more than 256 slices per beam are not supported");
    if (s.resolution <= 0) fail(OCTOPIC_ERR_INVALID_OPTION, b + "slicing resolution must be positive");
    if (s.method < 0 || s.method > 4) fail(OCTOPIC_ERR_INVALID_OPTION, b + "unsupported longitudinal slicing method");
    if (s.center_position != OCTOPIC_CENTER_CENTROID && s.center_position != OCTOPIC_CENTER_MIDPOINT)
        fail(OCTOPIC_ERR_INVALID_OPTION, b + "center_position must be :centroid or :midpoint");
    if (s.method == OCTOPIC_SLICE_SPECIFIED && s.npositions != s.nslices - 1)
        fail(OCTOPIC_ERR_INVALID_OPTION, b + "specified slicing requires nslices-1 internal positions");
    if (s.method == OCTOPIC_SLICE_EQUAL_AREA && (int64_t)s.nslices * s.resolution > 16384)
        fail(OCTOPIC_ERR_UNSUPPORTED, b + ":equal_area with nslices*resolution > 16384 histogram bins is not supported");
}

// PICPoissonSolver constructor checks (interface.jl:1410-1476), _validate_pic_solver (pic_cpu.jl:517-576)
// and the device-route checks of this backend (the reference's _require_cuda_pic_options, 730-777).
void validate(const octopic_options &o) {
    if (o.precision != OCTOPIC_PRECISION_FP64 && o.precision != OCTOPIC_PRECISION_FP32)
        fail(OCTOPIC_ERR_INVALID_OPTION, "precision must be OCTOPIC_PRECISION_FP64 or OCTOPIC_PRECISION_FP32");
    if (!(o.green_min_ratio > 0.0 && o.green_min_ratio <= 1.0))
        fail(OCTOPIC_ERR_INVALID_OPTION, "slice_pair_green_min_ratio must be in (0, 1]");
    if (!(o.green_growth >= 0.0)) fail(OCTOPIC_ERR_INVALID_OPTION, "slice_pair_green_growth must be non-negative");
    if (o.nx < 5 || o.ny < 5) fail(OCTOPIC_ERR_INVALID_OPTION, "PICPoissonSolver grid dimensions must both be at least 5");
    if (o.deposit != OCTOPIC_CIC && o.deposit != OCTOPIC_TSC) fail(OCTOPIC_ERR_INVALID_OPTION, "deposit_method must be :CIC or :TSC");
    if (o.green_type < 0 || o.green_type > 2)
        fail(OCTOPIC_ERR_INVALID_OPTION, "green_type must be :integrated, :standard or :lattice");
    if (o.green_cache != OCTOPIC_CACHE_NONE && o.green_cache != OCTOPIC_CACHE_SLICE_PAIR)
        fail(OCTOPIC_ERR_INVALID_OPTION, "green_cache must be :none or :slice_pair");
    if (o.field_derivative != 2 && o.field_derivative != 4)
        fail(OCTOPIC_ERR_INVALID_OPTION, "field_derivative must be :second or :fourth");
    if (o.slice_interpolation != OCTOPIC_INTERP_LINEAR && o.slice_interpolation != OCTOPIC_INTERP_QUADRATIC)
        fail(OCTOPIC_ERR_INVALID_OPTION, "slice_interpolation must be :linear or :quadratic");
    if (o.grid_extent != OCTOPIC_EXTENT_EXTREMA && o.grid_extent != OCTOPIC_EXTENT_SIGMA)
        fail(OCTOPIC_ERR_INVALID_OPTION, "grid_extent must be :extrema or :sigma");
    if (!(o.grid_extent_sigma > 0.0)) fail(OCTOPIC_ERR_INVALID_OPTION, "grid_extent_sigma must be positive");
    if (!(std::isfinite(o.min_extent_x) && std::isfinite(o.min_extent_y) && o.min_extent_x >= 0.0 && o.min_extent_y >= 0.0))
        fail(OCTOPIC_ERR_INVALID_OPTION, "min_transverse_extent values must be finite and nonnegative");
    if (!(o.grid_quantize >= 0.0)) fail(OCTOPIC_ERR_INVALID_OPTION, "grid_quantize must be non-negative (0 disables)");
    if (o.interaction_grid < 0 || o.interaction_grid > 2)
        fail(OCTOPIC_ERR_INVALID_OPTION, "interaction_grid must be :slice_pair, :source_slice or :node");
    if (o.interaction_grid != OCTOPIC_GRID_SLICE_PAIR && o.grid_extent != OCTOPIC_EXTENT_EXTREMA)
        fail(OCTOPIC_ERR_UNSUPPORTED, "grid_extent is not implemented for this interaction_grid: that mode sizes its mesh "
                                      "from the union/per-node extrema of its own particle sets");
    if (o.batch_mode != OCTOPIC_BATCH_SEQUENTIAL && o.batch_mode != OCTOPIC_BATCH_WAVEFRONT)
        fail(OCTOPIC_ERR_INVALID_OPTION, "batch_mode must be :sequential or :wavefront");
    int lnx = o.lum_nx > 0 ? o.lum_nx : o.nx, lny = o.lum_ny > 0 ? o.lum_ny : o.ny;
    if (lnx < 3 || lny < 3) fail(OCTOPIC_ERR_INVALID_OPTION, "luminosity_grid dimensions must both be at least 3");
    int ld = o.lum_deposit >= 0 ? o.lum_deposit : o.deposit;
    if (ld != OCTOPIC_CIC && ld != OCTOPIC_TSC)
        fail(OCTOPIC_ERR_INVALID_OPTION, "luminosity_deposit_method must be nothing, :CIC, or :TSC");
    validate_slicing(o.slicing1, 1);
    validate_slicing(o.slicing2, 2);
    if (o.interaction_grid == OCTOPIC_GRID_NODE && o.slice_interpolation != OCTOPIC_INTERP_LINEAR)
        fail(OCTOPIC_ERR_UNSUPPORTED,
             "slice_interpolation = :quadratic is not implemented for interaction_grid = :node: the node-plane path "
             "applies its own interpolation. Use interaction_grid = :slice_pair for :quadratic, or "
             "slice_interpolation = :linear.");
    // Device-route checks (as the reference's CUDA backend, pic_cpu.jl:730-777).
    if (o.grid_extent != OCTOPIC_EXTENT_EXTREMA)
        fail(OCTOPIC_ERR_UNSUPPORTED, "grid_extent = :sigma is not implemented by the HIP PIC backend: its bounds are "
                                      "order-free min/max reductions and apply no estimator. Use grid_extent = :extrema.");
    if (o.interaction_grid == OCTOPIC_GRID_SOURCE_SLICE)
        fail(OCTOPIC_ERR_UNSUPPORTED, "interaction_grid = :source_slice is not implemented by the HIP PIC backend; its union "
                                      "mesh depends on the sequential pair order. Use :slice_pair or :node.");
    if (o.green_type == OCTOPIC_GREEN_LATTICE)
        fail(OCTOPIC_ERR_UNSUPPORTED, "green_type = :lattice (experimental in the reference) is not implemented by the HIP "
                                      "PIC backend; use :integrated or :standard.");
    if (o.nx > 512 || o.ny > 512)
        fail(OCTOPIC_ERR_UNSUPPORTED, "meshes above 512 nodes per axis exceed the LDS transforms of this backend "
                                      "(padded size 2*grid <= 1024)");
    if ((int64_t)(lnx + 1) * (lny + 1) > (int64_t)1 << 20)
        fail(OCTOPIC_ERR_UNSUPPORTED, "luminosity meshes above 2^20 nodes are not supported");
    if (o.solver != OCTOPIC_SOLVER_PIC && o.solver != OCTOPIC_SOLVER_GAUSSIAN_PIC && o.solver != OCTOPIC_SOLVER_SOFT_GAUSSIAN)
        fail(OCTOPIC_ERR_INVALID_OPTION,
             "solver must be OCTOPIC_SOLVER_PIC, OCTOPIC_SOLVER_GAUSSIAN_PIC or OCTOPIC_SOLVER_SOFT_GAUSSIAN");
    if (o.solver == OCTOPIC_SOLVER_SOFT_GAUSSIAN) {
        // GaussianPoissonSolver constructor (interface.jl:976-1000); the mesh fields above keep their defaults
        if (!(std::isfinite(o.min_sigma) && o.min_sigma >= 0.0))
            fail(OCTOPIC_ERR_INVALID_OPTION, "min_sigma must be finite and nonnegative");
        if (o.gaussian_when_luminosity != 1 && o.gaussian_when_luminosity != 2)
            fail(OCTOPIC_ERR_INVALID_OPTION, "gaussian_when_luminosity must be 1 or 2");
        if (o.virtual_drift < OCTOPIC_DRIFT_HIRATA || o.virtual_drift > OCTOPIC_DRIFT_UNSAFE_CHROMATIC_FROZEN_ENERGY)
            fail(OCTOPIC_ERR_INVALID_OPTION, "unknown virtual_drift; expected :hirata, :chromatic, :exact or an unsafe drift");
        if (o.include_sigma_xy)
            fail(OCTOPIC_ERR_UNSUPPORTED, "include_sigma_xy = true (the coupled, rotated soft-Gaussian kick) is not "
                                          "implemented by the HIP GaussianPoissonSolver; use include_sigma_xy = false.");
    }
    if (o.solver == OCTOPIC_SOLVER_GAUSSIAN_PIC) {
        // GaussianPICPoissonSolver constructor (gaussian_pic.jl:94-110) and the options it rejects at collide
        // time (_require_linear_slice_interpolation, pic_cpu.jl:788-809)
        if (!(o.margin_sigma >= 0.0)) fail(OCTOPIC_ERR_INVALID_OPTION, "margin_sigma must be non-negative");
        if (!(o.coupling_tol >= 0.0)) fail(OCTOPIC_ERR_INVALID_OPTION, "coupling_tol must be non-negative");
        if (o.slice_interpolation != OCTOPIC_INTERP_LINEAR)
            fail(OCTOPIC_ERR_UNSUPPORTED, "slice_interpolation = :quadratic is not implemented by GaussianPICPoissonSolver; "
                                          "set slice_interpolation = :linear.");
        if (o.interaction_grid != OCTOPIC_GRID_SLICE_PAIR)
            fail(OCTOPIC_ERR_UNSUPPORTED, "interaction_grid other than :slice_pair is not implemented by "
                                          "GaussianPICPoissonSolver; set interaction_grid = :slice_pair.");
        if (std::isfinite(o.coupling_tol))
            fail(OCTOPIC_ERR_UNSUPPORTED, "a finite coupling_tol (the coupled, rotated subtraction) is not implemented by "
                                          "the HIP GaussianPICPoissonSolver; use coupling_tol = Inf (the separable "
                                          "subtraction).");
    }
}

void upload_ptrs(octopic_handle *h, double *const in[2][6]) {
    double *hp[36];
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 6; ++c) {
            hp[b * 6 + c] = in[b][c];
            hp[12 + b * 6 + c] = h->st[b][c].p;
            hp[24 + b * 6 + c] = h->srt[b][c].p;
        }
    h->ptrs.ensure(36);
    HIPCHECK(hipMemcpyAsync(h->ptrs.p, hp, sizeof(hp), hipMemcpyHostToDevice, h->stream));
}

void run_slicing(octopic_handle *h, int b, const octopic_slicing &s, int64_t n) {
    const int ns = s.nslices;
    std::vector<double> &q = h->hq[b];   // kept in the handle: the copy below is asynchronous
    q.assign(kMaxSlices, 0.0);
    if (s.method == OCTOPIC_SLICE_NORMAL_QUANTILE) {
        for (int k = 1; k < ns; ++k) q[k] = std::sqrt(2.0) * erfinv_host((double)(2 * k) / (double)ns - 1.0);
    } else if (s.method == OCTOPIC_SLICE_SPECIFIED) {
        const std::vector<double> &p = b == 0 ? h->pos1 : h->pos2;
        for (int k = 0; k < ns - 1; ++k) q[k] = p[k];
    }
    HIPCHECK(hipMemcpyAsync(h->q.p + (size_t)b * kMaxSlices, q.data(), sizeof(double) * kMaxSlices, hipMemcpyHostToDevice,
                            h->stream));
    h->perm[b].ensure(n);
    h->inv_sl[b].ensure(n);
    h->sid[b].ensure(n);
    // the :equal_count rank sort's passes count 256 digits per block
    h->hist[b].ensure(slicing_hist_entries(n, s.method == OCTOPIC_SLICE_EQUAL_COUNT ? 256 : ns));
    if (s.method == OCTOPIC_SLICE_EQUAL_COUNT) {
        h->rank_keys[b].ensure(n);
        h->rank_sorted[b].ensure(n);
        h->rank_vals[b].ensure(n);
    }
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int octopic_options_default(octopic_options *o) {
    if (!o) return OCTOPIC_ERR_INVALID_ARGUMENT;
    std::memset(o, 0, sizeof(*o));
    o->nx = 128; o->ny = 128;
    o->deposit = OCTOPIC_CIC;
    o->green_type = OCTOPIC_GREEN_INTEGRATED;
    o->green_cache = OCTOPIC_CACHE_SLICE_PAIR;
    o->field_derivative = 2;
    o->slice_interpolation = OCTOPIC_INTERP_LINEAR;
    o->interaction_grid = OCTOPIC_GRID_SLICE_PAIR;
    o->grid_extent = OCTOPIC_EXTENT_EXTREMA;
    o->longitudinal_kick = 1;
    o->lum_nx = 0; o->lum_ny = 0; o->lum_deposit = -1;
    o->batch_mode = OCTOPIC_BATCH_WAVEFRONT;
    o->cuda_flags = 15;
    o->grid_extent_sigma = 6.0;
    o->min_extent_x = 0.0; o->min_extent_y = 0.0;
    o->grid_quantize = 0.0;
    o->green_min_ratio = 0.5;
    o->green_growth = 0.25;
    o->kbb1 = NAN; o->kbb2 = NAN; o->luminosity_scale = NAN;
    for (octopic_slicing *s : {&o->slicing1, &o->slicing2}) {
        s->nslices = 1; s->method = OCTOPIC_SLICE_EQUAL_AREA; s->resolution = 100;
        s->center_position = OCTOPIC_CENTER_CENTROID; s->npositions = 0; s->positions = nullptr;
    }
    o->solver = OCTOPIC_SOLVER_PIC;
    o->neutralize = 1;
    o->margin_sigma = 5.0;
    o->coupling_tol = INFINITY;
    o->min_sigma = 0.0;
    o->virtual_drift = OCTOPIC_DRIFT_HIRATA;
    o->gaussian_when_luminosity = 2;
    o->ignore_centroid1 = 0; o->ignore_centroid2 = 0;
    o->include_sigma_xy = 0;
    o->reserved = 0;
    return OCTOPIC_OK;
}

const char *octopic_last_error(const octopic_handle *h) { return h ? h->err.c_str() : "null handle"; }

int octopic_set_stream(octopic_handle *h, void *s) {
    if (!h) return OCTOPIC_ERR_INVALID_ARGUMENT;
    h->stream = s ? (hipStream_t)s : h->own_stream;
    return OCTOPIC_OK;
}

int octopic_set_denoiser(octopic_handle *h, octopic_denoiser *d, double monitor_limit) {
    if (!h) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (!d) {
        h->dn = nullptr;
        return OCTOPIC_OK;
    }
    if (h->mode == 2) return set_err(h, OCTOPIC_ERR_UNSUPPORTED, "the residual denoiser runs with interaction_grid = :slice_pair");
    if (h->gauss || h->soft)
        return set_err(h, OCTOPIC_ERR_UNSUPPORTED, "the residual denoiser runs with PICPoissonSolver, not the Gaussian solvers");
    int dh = 0, dw = 0, dmb = 0;
    octopic::denoiser_shape(d, &dh, &dw, &dmb);
    if (dh != h->ny || dw != h->nx)
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "the denoiser's planes (" + std::to_string(dh) + " x " + std::to_string(dw) +
                                                            ") must be the grid's (ny x nx = " + std::to_string(h->ny) + " x " +
                                                            std::to_string(h->nx) + ")");
    if (!(monitor_limit > 0.0)) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "monitor_limit must be positive");
    h->dn = d;
    h->dn_limit = monitor_limit;
    return OCTOPIC_OK;
}

void octopic_destroy(octopic_handle *h) {
    if (!h) return;
    hipSetDevice(h->device);
    hipDeviceSynchronize();
    if (h->own_stream) hipStreamDestroy(h->own_stream);
    if (h->side_stream) hipStreamDestroy(h->side_stream);
    if (h->ev_fork) hipEventDestroy(h->ev_fork);
    if (h->ev_join) hipEventDestroy(h->ev_join);
    for (auto e : h->ev_pool) hipEventDestroy(e);
    for (auto e : h->deposit_ev) hipEventDestroy(e);
    if (h->h_tab) hipHostFree(h->h_tab);
    if (h->h_err) hipHostFree(h->h_err);
    if (h->h_counters) hipHostFree(h->h_counters);
    h->tab.release(); h->q.release(); h->errinfo.release(); h->counters.release(); h->cache.release(); h->spec.release();
    for (int b = 0; b < 2; ++b) {
        for (int c = 0; c < 6; ++c) {
            h->beam_in[b][c].release(); h->beam_f32[b][c].release(); h->st[b][c].release(); h->srt[b][c].release();
        }
        h->perm[b].release(); h->perm_srt[b].release(); h->inv_srt[b].release(); h->inv_sl[b].release(); h->sid[b].release(); h->hist[b].release();
    }
    for (int b = 0; b < 2; ++b) { h->rank_keys[b].release(); h->rank_sorted[b].release(); h->rank_vals[b].release(); }
    h->lanes.release(); h->mm.release(); h->clanes.release(); h->bins.release(); h->ptrs.release(); h->pairs.release();
    h->korder.release(); h->dqueue.release(); h->dorder.release();
    h->keys.release(); h->chunkbox.release(); h->box.release(); h->thist.release(); h->inters.release(); h->planes.release(); h->jobs.release(); h->lumgeo.release();
    h->charge.release(); h->qh.release(); h->phi.release(); h->table.release(); h->records.release();
    h->lum_parts.release(); h->lum_partial.release(); h->nodes.release(); h->node_jobs.release(); h->node_keys.release(); h->ran_dev.release();
    h->px.tw.release(); h->py.tw.release(); h->px.tw32.release(); h->py.tw32.release();
    h->spec32.release(); h->qh32.release(); h->phi32.release(); h->table32.release();
    h->dn_mom.release(); h->dn_meta.release(); h->dn_prof.release(); h->dn_in.release();
    h->gmom.release(); h->gab.release();
    h->soft_lum_block.release(); h->soft_lum_scale.release(); h->soft_shift.release();
    delete h;
}

int octopic_create(const octopic_options *opt, int32_t device, octopic_handle **out) {
    if (!opt || !out) return OCTOPIC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    octopic_handle *h = new octopic_handle();
    try {
        validate(*opt);
        h->opt = *opt;
        if (opt->slicing1.positions) h->pos1.assign(opt->slicing1.positions, opt->slicing1.positions + opt->slicing1.npositions);
        if (opt->slicing2.positions) h->pos2.assign(opt->slicing2.positions, opt->slicing2.positions + opt->slicing2.npositions);
        h->opt.slicing1.positions = nullptr;
        h->opt.slicing2.positions = nullptr;
        h->device = device;
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreate(&h->own_stream));   // default priority; orders with the null stream
        HIPCHECK(generate_scratch_init(device));     // the allocation-free generator's per-device workspace
        h->stream = h->own_stream;
        h->nx = opt->nx; h->ny = opt->ny;
        h->fp32 = opt->precision == OCTOPIC_PRECISION_FP32;
        h->lnx = opt->lum_nx > 0 ? opt->lum_nx : opt->nx;
        h->lny = opt->lum_ny > 0 ? opt->lum_ny : opt->ny;
        h->lum_dep = opt->lum_deposit >= 0 ? opt->lum_deposit : opt->deposit;
        h->mode = opt->interaction_grid == OCTOPIC_GRID_NODE ? 2 : (opt->slice_interpolation == OCTOPIC_INTERP_QUADRATIC ? 1 : 0);
        h->gauss = opt->solver == OCTOPIC_SOLVER_GAUSSIAN_PIC;
        h->soft = opt->solver == OCTOPIC_SOLVER_SOFT_GAUSSIAN;
        h->npl = h->mode == 0 ? 2 : 3;
        if (h->mode == 2 && !opt->longitudinal_kick) h->npl = 3;   // Z plane slot kept; unused without the longitudinal kick
        h->px.build(2 * h->nx);
        h->py.build(2 * h->ny);
        h->tab.ensure(2);
        h->q.ensure(2 * kMaxSlices);
        if (const char *e = std::getenv("OCTOPIC_TRANSVERSE_TILES")) {
            const int t = std::atoi(e);
            if (t < 0 || t > 128 || (t & (t - 1)) != 0)
                fail(OCTOPIC_ERR_INVALID_ARGUMENT, "OCTOPIC_TRANSVERSE_TILES must be 0 or a power of two up to 128");
            h->tiles = t;
        }
        if (const char *e = std::getenv("OCTOPIC_KICK_TARGET_BLOCKS")) {
            const int v = std::atoi(e);
            if (v < 0 || v > (1 << 20)) fail(OCTOPIC_ERR_INVALID_ARGUMENT, "OCTOPIC_KICK_TARGET_BLOCKS must be in 0..2^20");
            h->kick_target_blocks = v;
        }
        if (const char *e = std::getenv("OCTOPIC_KICK_PASS_WINDOWS")) {
            const int v = std::atoi(e);
            if (v < -1 || v > 2) fail(OCTOPIC_ERR_INVALID_ARGUMENT, "OCTOPIC_KICK_PASS_WINDOWS must be -1, 0, 1 or 2");
            h->kick_pass_windows = v;
        }
        if (hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || h->cus <= 0)
            h->cus = 256;
        h->box.ensure(8);
        h->errinfo.ensure(1);
        h->counters.ensure(16);   // live counters, then the snapshot a collide takes before its batches
        HIPCHECK(hipMemset(h->errinfo.p, 0, sizeof(ErrorInfo)));
        HIPCHECK(hipMemset(h->counters.p, 0, 16 * sizeof(unsigned long long)));
        h->lanes.ensure(2 * kFoldLanes);
        h->mm.ensure(2 * 2 * kZFoldBlocks);
        h->bins.ensure(2 * 16384);
        HIPCHECK(hipMemset(h->bins.p, 0, 2 * 16384 * sizeof(unsigned long long)));
        const int ns1 = opt->slicing1.nslices, ns2 = opt->slicing2.nslices;
        h->clanes_share = std::max(ns1, ns2);
        h->clanes.ensure((size_t)2 * h->clanes_share * kFoldLanes);
        h->slot_stride = (int64_t)(h->nx + 1) * 2 * h->ny;
        if (opt->green_cache == OCTOPIC_CACHE_SLICE_PAIR && h->mode != 2 && !h->soft) {
            h->nkeys = 2 * ns1 * ns2;
            h->cache.ensure(h->nkeys);
            HIPCHECK(hipMemset(h->cache.p, 0, sizeof(CacheGeo) * h->nkeys));
        }
        HIPCHECK(hipHostMalloc((void **)&h->h_tab, 2 * sizeof(SliceTable)));
        HIPCHECK(hipHostMalloc((void **)&h->h_err, sizeof(ErrorInfo)));
        HIPCHECK(hipHostMalloc((void **)&h->h_counters, 16 * sizeof(unsigned long long)));
        HIPCHECK(hipDeviceSynchronize());
    } catch (Fail &f) {
        int code = f.code;
        std::string msg = f.msg;
        octopic_destroy(h);
        // A failed create has no handle to carry its message; report through a static buffer.
        static thread_local std::string last;
        last = msg;
        (void)last;
        fprintf(stderr, "octopic_create: %s\n", msg.c_str());
        return code;
    }
    *out = h;
    return OCTOPIC_OK;
}

}  // extern "C"

namespace {

void collide_impl(octopic_handle *h, octopic_beam *b1, octopic_beam *b2, int64_t turn, int compute_lum, double *lum_out);

}  // namespace

// ------------------------------------------------------------------ launch shapes and the autotuner
namespace {
struct ShapeKnob {
    const char *name;
    int octopic_handle::*field;
    int64_t lo, hi;
    bool pow2;
};
const ShapeKnob kShapeKnobs[] = {
    {"transverse_tiles", &octopic_handle::tiles, 0, 128, true},
    {"kick_target_blocks", &octopic_handle::kick_target_blocks, 0, 1 << 20, false},
    {"kick_pass_windows", &octopic_handle::kick_pass_windows, -1, 2, false},
    {"deposit_span", &octopic_handle::deposit_span, 1024, 1 << 20, false},
    {"deposit_multi", &octopic_handle::deposit_multi, -1, 1, false},
    {"deposit_min_span", &octopic_handle::deposit_min_span, 1024, 1 << 20, false},
    {"fft_transforms_per_block", &octopic_handle::fft_group, 2, 8, true},
};
const ShapeKnob *find_knob(const char *name) {
    for (const auto &k : kShapeKnobs)
        if (name && std::strcmp(k.name, name) == 0) return &k;
    return nullptr;
}
}  // namespace

extern "C" int octopic_set_launch_shape(octopic_handle *h, const char *name, int64_t value) {
    if (!h || !name) return OCTOPIC_ERR_INVALID_ARGUMENT;
    const ShapeKnob *k = find_knob(name);
    if (!k) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, std::string("unknown launch shape '") + name + "'");
    if (value < k->lo || value > k->hi || (k->pow2 && value > 0 && (value & (value - 1)) != 0))
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, std::string("launch shape '") + name + "' out of range");
    h->*(k->field) = (int)value;
    return OCTOPIC_OK;
}

extern "C" int octopic_get_launch_shape(const octopic_handle *h, const char *name, int64_t *value) {
    if (!h || !name || !value) return OCTOPIC_ERR_INVALID_ARGUMENT;
    const ShapeKnob *k = find_knob(name);
    if (!k) return OCTOPIC_ERR_INVALID_ARGUMENT;
    *value = h->*(k->field);
    return OCTOPIC_OK;
}

extern "C" int octopic_autotune(octopic_handle *h, octopic_beam *beam1, octopic_beam *beam2, int32_t trials, double *best_ms) {
    if (!h || !beam1 || !beam2 || trials < 1 || trials > 64) return OCTOPIC_ERR_INVALID_ARGUMENT;
    octopic_handle *t = nullptr;
    double *buf = nullptr;
    int rc = OCTOPIC_OK;
    try {
        HIPCHECK(hipSetDevice(h->device));
        if (int r = octopic_create(&h->opt, h->device, &t)) fail(r, "autotune: scratch handle: " + std::string(octopic_last_error(nullptr)));
        for (const auto &k : kShapeKnobs) t->*(k.field) = h->*(k.field);
        t->dn = h->dn;
        t->dn_limit = h->dn_limit;
        const int64_t n1 = beam1->n, n2 = beam2->n;
        if (n1 <= 0 || n2 <= 0) fail(OCTOPIC_ERR_INVALID_ARGUMENT, "autotune: empty beam");
        // originals and working copies of both beams, six coordinates each
        for (const octopic_beam *bm : {beam1, beam2})
            if (bm->dtype != OCTOPIC_DTYPE_F64 && bm->dtype != OCTOPIC_DTYPE_F32) fail(OCTOPIC_ERR_INVALID_ARGUMENT, "autotune: beam dtype");
        // element sizes of the two beams (Float32 beams are copied as floats)
        const size_t es[2] = {beam1->dtype == OCTOPIC_DTYPE_F32 ? sizeof(float) : sizeof(double),
                              beam2->dtype == OCTOPIC_DTYPE_F32 ? sizeof(float) : sizeof(double)};
        if (hipMalloc((void **)&buf, sizeof(double) * 12 * (size_t)(n1 + n2)) != hipSuccess) fail(OCTOPIC_ERR_OOM, "autotune: beam copies");
        double *orig[2][6], *work[2][6];
        const octopic_beam *src[2] = {beam1, beam2};
        size_t off = 0;
        for (int b = 0; b < 2; ++b) {
            const int64_t n = b ? n2 : n1;
            const double *const cs[6] = {src[b]->x, src[b]->px, src[b]->y, src[b]->py, src[b]->z, src[b]->pz};
            for (int c = 0; c < 6; ++c) {
                orig[b][c] = buf + off; off += (size_t)n;
                work[b][c] = buf + off; off += (size_t)n;
                HIPCHECK(hipMemcpy(orig[b][c], cs[c], es[b] * n, hipMemcpyDefault));
            }
        }
        octopic_beam wb[2] = {*beam1, *beam2};
        for (int b = 0; b < 2; ++b) {
            wb[b].x = work[b][0]; wb[b].px = work[b][1]; wb[b].y = work[b][2];
            wb[b].py = work[b][3]; wb[b].z = work[b][4]; wb[b].pz = work[b][5];
            wb[b].on_device = 1;
        }
        // the median collide time of `trials` collides from the same start (after one that warms the
        // scratch handle's Green cache)
        auto measure = [&]() {
            std::vector<double> ms;
            for (int r = 0; r <= trials; ++r) {
                for (int b = 0; b < 2; ++b)
                    for (int c = 0; c < 6; ++c)
                        HIPCHECK(hipMemcpyAsync(work[b][c], orig[b][c], es[b] * (b ? n2 : n1), hipMemcpyDeviceToDevice,
                                                t->stream));
                HIPCHECK(hipStreamSynchronize(t->stream));
                const auto t0 = std::chrono::steady_clock::now();
                double lum = 0.0;
                if (int e = octopic_collide(t, &wb[0], &wb[1], 0, 1, &lum)) fail(e, "autotune: " + t->err);
                const auto t1 = std::chrono::steady_clock::now();
                if (r > 0) ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
            }
            std::sort(ms.begin(), ms.end());
            return ms[ms.size() / 2];
        };
        struct Cand { int octopic_handle::*field; std::vector<int> values; };
        const std::vector<Cand> cands = {
            {&octopic_handle::deposit_multi, {-1, 1, 0}},
            {&octopic_handle::tiles, {8, 16, 32}},
            {&octopic_handle::kick_target_blocks, {512, 1024, 2048}},
            {&octopic_handle::deposit_span, {4096, 8192, 16384}},
            {&octopic_handle::kick_pass_windows, {-1, 0, 1, 2}},
            {&octopic_handle::fft_group, {4, 2, 8}},
        };
        double best = measure();
        for (const auto &cd : cands) {   // coordinate descent: one knob at a time, keep the fastest value
            const int keep = t->*(cd.field);
            int choice = keep;
            for (int v : cd.values) {
                if (v == keep) continue;
                t->*(cd.field) = v;
                const double m = measure();
                if (m < best) { best = m; choice = v; }
            }
            t->*(cd.field) = choice;
        }
        for (const auto &k : kShapeKnobs) h->*(k.field) = t->*(k.field);
        if (best_ms) *best_ms = best;
    } catch (Fail &f) {
        rc = set_err(h, f.code, f.msg);
    }
    if (buf) hipFree(buf);
    if (t) octopic_destroy(t);
    if (rc == OCTOPIC_OK) h->err.clear();
    return rc;
}

extern "C" int octopic_collide(octopic_handle *h, octopic_beam *beam1, octopic_beam *beam2, int64_t turn,
                               int32_t compute_luminosity, double *luminosity) {
    if (!h) return OCTOPIC_ERR_INVALID_ARGUMENT;
    try {
        HIPCHECK(hipSetDevice(h->device));
        if (h->unfinished) {
            // kernels of the abandoned collide may have left charge or bounds slots filled (and a failure
            // between the slicing's fork and join may have left work on the side stream)
            if (h->side_stream) HIPCHECK(hipStreamSynchronize(h->side_stream));
            if (h->charge.p) HIPCHECK(hipMemsetAsync(h->charge.p, 0, h->charge.n * sizeof(unsigned long long), h->stream));
            if (h->keys.p) HIPCHECK(hipMemsetAsync(h->keys.p, 0, h->keys.n * sizeof(unsigned long long), h->stream));
            HIPCHECK(launch_cache_invalidate(h->cache.p, h->nkeys, h->epoch, h->stream));
            HIPCHECK(hipMemsetAsync(h->errinfo.p, 0, sizeof(ErrorInfo), h->stream));
            HIPCHECK(hipStreamSynchronize(h->stream));
            h->unfinished = false;
        }
        h->unfinished = true;
        h->epoch = (int)(h->stats.collides % 2147483646) + 1;
        collide_impl(h, beam1, beam2, turn, compute_luminosity, luminosity);
        h->unfinished = false;
    } catch (Fail &f) {
        return set_err(h, f.code, f.msg);
    }
    h->err.clear();
    return OCTOPIC_OK;
}

namespace {

const char *stage_name(int stage) {
    switch (stage) {
        case 0: return "source bounds";
        case 1: return "field bounds";
        case 2: return "luminosity grid";
        case 3: return "interaction grid";
        case 4: return "node-mesh build";
        case 6: return "soft-Gaussian slice moments";
        default: return "slicing";
    }
}

void raise_device_error(octopic_handle *h, const ErrorInfo &e, int64_t turn) {
    char buf[768];
    if (e.code == OCTOPIC_ERR_NONFINITE) {
        snprintf(buf, sizeof buf,
                 "non-finite particle coordinate(s) detected in a strong-strong collision (turn=%lld, %s, collision-order "
                 "pair %d, direction %d): non-finite coordinates cannot be represented on the interaction mesh, so the "
                 "solver fails fast; the beams are left unmodified.",
                 (long long)turn, stage_name(e.stage), e.p, e.dir);
    } else if (e.code == OCTOPIC_ERR_DEGENERATE_EXTENT) {
        snprintf(buf, sizeof buf,
                 "PIC %s requires finite, positive transverse extents; observed (x=%g, y=%g) with min_transverse_extent="
                 "(%g, %g) (turn=%lld, collision-order pair %d, direction %d). Supply a positive physical bound for every "
                 "degenerate axis.",
                 e.stage == 2 ? "luminosity grid" : "interaction grid", e.a, e.b, h->opt.min_extent_x, h->opt.min_extent_y,
                 (long long)turn, e.p, e.dir);
    } else {
        snprintf(buf, sizeof buf, "device-side failure code %d at stage %s", e.code, stage_name(e.stage));
    }
    // Kernels after the failure returned early; clear any charge they left so the next collide starts clean,
    // and drop the cache entries this collide built (their spectra may not have been built).
    if (h->charge.p) hipMemsetAsync(h->charge.p, 0, h->charge.n * sizeof(unsigned long long), h->stream);
    launch_cache_invalidate(h->cache.p, h->nkeys, h->epoch, h->stream);
    hipMemsetAsync(h->errinfo.p, 0, sizeof(ErrorInfo), h->stream);
    hipStreamSynchronize(h->stream);
    fail(e.code, buf);
}

void check_beam(const octopic_beam *b, int k) {
    if (!b) fail(OCTOPIC_ERR_INVALID_ARGUMENT, "beam" + std::to_string(k) + " is null");
    if (b->n <= 0) fail(OCTOPIC_ERR_INVALID_ARGUMENT, "longitudinal slicing requires at least one particle");
    if (!b->x || !b->px || !b->y || !b->py || !b->z || !b->pz)
        fail(OCTOPIC_ERR_INVALID_ARGUMENT, "beam" + std::to_string(k) + " has a null coordinate pointer");
    if (b->n > (int64_t)2147483647) fail(OCTOPIC_ERR_UNSUPPORTED, "more than 2^31-1 macroparticles per beam");
    if (b->dtype != OCTOPIC_DTYPE_F64 && b->dtype != OCTOPIC_DTYPE_F32)
        fail(OCTOPIC_ERR_INVALID_ARGUMENT, "beam" + std::to_string(k) + " dtype must be OCTOPIC_DTYPE_F64 or OCTOPIC_DTYPE_F32");
}

// _strong_strong_kbb1/2 (slicing.jl:77-91) divided by the source macroparticle count (pic_cpu.jl:821-824)
double kbb_of(double given, const octopic_beam *self, const octopic_beam *other, bool first, int64_t nsrc) {
    double v = given;
    if (std::isnan(v)) {
        if (!(std::isfinite(self->E0) && self->E0 != 0.0))
            fail(OCTOPIC_ERR_INVALID_ARGUMENT, first ? "beam1 E0 must be finite and nonzero when kbb1 is not specified"
                                                     : "beam2 E0 must be finite and nonzero when kbb2 is not specified");
        const octopic_beam *p1 = first ? self : other, *p2 = first ? other : self;
        v = p1->charge * p2->charge * self->r0 * other->npart * self->mc2 / self->E0;
    }
    return v / (double)nsrc;
}

}  // namespace

// Independent bunch-pair instances: each handle runs its collide on its own stream from its own host
// thread, so the instances' kernels overlap on the device.
extern "C" int octopic_collide_many(octopic_handle *const *handles, octopic_beam *beams1, octopic_beam *beams2,
                                    const int64_t *turns, const int32_t *compute_luminosity, double *luminosity,
                                    int32_t *status, int32_t count) {
    if (count < 0 || (count > 0 && (!handles || !beams1 || !beams2 || !luminosity))) return OCTOPIC_ERR_INVALID_ARGUMENT;
    std::set<const octopic_handle *> seen;
    std::set<const octopic_denoiser *> denoisers;   // a denoiser's buffers serve one collide at a time
    for (int k = 0; k < count; ++k) {
        if (!handles[k] || !seen.insert(handles[k]).second) return OCTOPIC_ERR_INVALID_ARGUMENT;
        if (handles[k]->dn && !denoisers.insert(handles[k]->dn).second)
            return set_err(handles[k], OCTOPIC_ERR_INVALID_ARGUMENT,
                           "octopic_collide_many: two handles share one denoiser (it holds one collide's activations)");
    }
    // instances update their beams in place from concurrent threads and streams: no coordinate array may
    // belong to two beams
    std::map<const double *, int> owner;
    for (int k = 0; k < count; ++k)
        for (int b = 0; b < 2; ++b) {
            const octopic_beam &bm = b ? beams2[k] : beams1[k];
            const double *cs[6] = {bm.x, bm.px, bm.y, bm.py, bm.z, bm.pz};
            for (const double *c : cs) {
                if (!c) continue;   // reported by the collide itself
                auto ins = owner.emplace(c, 2 * k + b);
                if (!ins.second && ins.first->second != 2 * k + b)
                    return set_err(handles[k], OCTOPIC_ERR_INVALID_ARGUMENT,
                                   "octopic_collide_many: instance " + std::to_string(k) + " beam" + std::to_string(b + 1) +
                                       " shares a coordinate array with instance " + std::to_string(ins.first->second / 2) +
                                       " beam" + std::to_string(ins.first->second % 2 + 1) +
                                       " (instances collide concurrently and update their beams in place)");
            }
        }
    std::vector<int32_t> st(count, OCTOPIC_OK);
    std::vector<std::thread> pool;
    pool.reserve(count);
    for (int k = 0; k < count; ++k)
        pool.emplace_back([&, k]() {
            handles[k]->in_many = true;
            st[k] = octopic_collide(handles[k], beams1 + k, beams2 + k, turns ? turns[k] : -1,
                                    compute_luminosity ? compute_luminosity[k] : 1, luminosity + k);
            handles[k]->in_many = false;
        });
    for (auto &t : pool) t.join();
    int first = OCTOPIC_OK;
    for (int k = 0; k < count; ++k) {
        if (status) status[k] = st[k];
        if (first == OCTOPIC_OK) first = st[k];
    }
    return first;
}

namespace {

// The collided coordinates back to the caller's arrays: Float32 beams rounded, host beams copied (z only
// where the solver moves it: the soft-Gaussian solver's exact and chromatic drifts).
void write_back(octopic_handle *h, octopic_beam *const B[2], double *const bp[2][6], const int64_t n[2], bool with_z) {
    hipStream_t s = h->stream;
    for (int b = 0; b < 2; ++b) {
        double *hp2[6] = {B[b]->x, B[b]->px, B[b]->y, B[b]->py, B[b]->z, B[b]->pz};
        if (B[b]->dtype == OCTOPIC_DTYPE_F32) {
            for (int c = 0; c < 6; ++c) {
                if (c == 4 && !with_z) continue;
                if (B[b]->on_device) {
                    HIPCHECK(launch_narrow(bp[b][c], reinterpret_cast<float *>(hp2[c]), n[b], s));
                } else {
                    HIPCHECK(launch_narrow(bp[b][c], h->beam_f32[b][c].p, n[b], s));
                    HIPCHECK(hipMemcpyAsync(hp2[c], h->beam_f32[b][c].p, sizeof(float) * n[b], hipMemcpyDeviceToHost, s));
                }
            }
            continue;
        }
        if (B[b]->on_device) continue;
        for (int c = 0; c < 6; ++c)
            if (c != 4 || with_z) HIPCHECK(hipMemcpyAsync(hp2[c], bp[b][c], sizeof(double) * n[b], hipMemcpyDeviceToHost, s));
    }
    HIPCHECK(hipStreamSynchronize(s));
}

// collide!(::GaussianPoissonSolver, ...) (gaussian.jl:7-57) after the slicing: per conflict-free batch of
// the collision order, the source moments of both slices of every pair, then every field slice through
// the soft-Gaussian map in place (the slice order, no mesh).  Every pair runs, empty slices included; the
// sampled direction's per-pair luminosity is folded in collision order.
void soft_collide(octopic_handle *h, octopic_beam *const B[2], double *const bp[2][6], const int64_t n[2], int64_t turn,
                  int compute_lum, double *lum_out) {
    hipStream_t s = h->stream;
    const octopic_options &o = h->opt;
    const SliceTable &T1 = h->h_tab[0], &T2 = h->h_tab[1];
    const int ns1 = o.slicing1.nslices, ns2 = o.slicing2.nslices;
    const double kbb1 = kbb_of(o.kbb1, B[0], B[1], true, 1);   // gaussian.jl:12-13: not divided by N
    const double kbb2 = kbb_of(o.kbb2, B[1], B[0], false, 1);
    double klum1 = o.luminosity_scale, klum2 = o.luminosity_scale;   // slicing.jl:93-102
    if (std::isnan(o.luminosity_scale)) {
        klum1 = B[0]->npart * B[1]->npart / (double)n[0];
        klum2 = B[0]->npart * B[1]->npart / (double)n[1];
    }
    auto order = collision_order(T1.center, ns1, T2.center, ns2);
    auto batches = pair_batches(order, ns1, ns2);
    const int npairs = (int)order.size();
    std::vector<BatchPair> hp;
    std::vector<int> boff, bcnt;
    int widest = 0;
    for (auto &bt : batches) {
        boff.push_back((int)hp.size());
        for (int pos : bt) hp.push_back(BatchPair{pos, order[pos].i, order[pos].j, 0});
        bcnt.push_back((int)bt.size());
        widest = std::max(widest, (int)bt.size());
    }
    const int W = std::max(widest, 1);
    h->pairs.ensure(std::max<size_t>(hp.size(), 1));
    if (!hp.empty())
        HIPCHECK(hipMemcpyAsync(h->pairs.p, hp.data(), sizeof(BatchPair) * hp.size(), hipMemcpyHostToDevice, s));
    h->gmom.ensure((size_t)2 * W * kMaxMomentChunks * 10);
    h->soft_shift.ensure((size_t)2 * W * 4);
    h->soft_lum_block.ensure((size_t)std::max(npairs, 1) * kSoftBlocks);
    h->soft_lum_scale.ensure(std::max(npairs, 1));
    h->lum_parts.ensure(std::max(npairs, 1));
    States st{};
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 6; ++c) st.c[b][c] = h->st[b][c].p;
    const SliceTable *dt1 = h->tab.p, *dt2 = h->tab.p + 1;
    for (size_t bi = 0; bi < batches.size(); ++bi) {
        const int Bn = bcnt[bi];
        if (Bn == 0) continue;
        const BatchPair *dp = h->pairs.p + boff[bi];
        int64_t bmax = 0;
        for (int k = 0; k < Bn; ++k) {
            const BatchPair &pr = hp[boff[bi] + k];
            bmax = std::max<int64_t>(bmax, std::max<int64_t>(T1.count[pr.i], T2.count[pr.j]));
        }
        const int mch = moment_chunks(bmax);
        {
            Timed t(h, F_GEOMETRY);
            GpicMomentArgs a{dp, Bn, dt1, dt2, st, mch, h->gmom.p, h->errinfo.p, h->soft_shift.p};
            HIPCHECK(launch_gpic_moments(a, s));
        }
        {
            Timed t(h, F_KICK);
            SoftKickArgs a{};
            a.pairs = dp; a.npairs = Bn; a.tab1 = dt1; a.tab2 = dt2; a.st = st;
            a.mom_chunks = mch; a.mom_part = h->gmom.p; a.shift = h->soft_shift.p;
            a.kbb1 = kbb1; a.kbb2 = kbb2; a.klum1 = klum1; a.klum2 = klum2; a.min_sigma = o.min_sigma;
            a.ignore1 = o.ignore_centroid1 ? 1 : 0; a.ignore2 = o.ignore_centroid2 ? 1 : 0;
            a.drift = o.virtual_drift; a.lkick = o.longitudinal_kick ? 1 : 0;
            a.sample1 = o.gaussian_when_luminosity == 1 ? 1 : 0; a.compute_lum = compute_lum ? 1 : 0;
            a.lum_block = h->soft_lum_block.p; a.lum_scale = h->soft_lum_scale.p; a.err = h->errinfo.p;
            HIPCHECK(launch_soft_kick(a, h->fp32, s));
        }
    }
    if (compute_lum) HIPCHECK(launch_soft_lum_fold(h->soft_lum_block.p, h->soft_lum_scale.p, npairs, h->lum_parts.p, s));
    {
        Timed t(h, F_STORE);
        for (int b = 0; b < 2; ++b)   // gathered through the inverse permutation: the writes run in beam order
            HIPCHECK(launch_store_gather(h->ptrs.p + b * 6, n[b], h->ptrs.p + 12 + b * 6, h->inv_sl[b].p, h->errinfo.p, s, 1));
    }
    HIPCHECK(hipMemcpyAsync(h->h_err, h->errinfo.p, sizeof(ErrorInfo), hipMemcpyDeviceToHost, s));
    h->plum.assign(npairs, 0.0);
    if (npairs && compute_lum)
        HIPCHECK(hipMemcpyAsync(h->plum.data(), h->lum_parts.p, sizeof(double) * npairs, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (h->h_err->code) raise_device_error(h, *h->h_err, turn);
    write_back(h, B, bp, n, true);
    h->pran.assign(npairs, 1);
    h->recs.clear();
    h->pi.resize(npairs); h->pj.resize(npairs);
    for (int p = 0; p < npairs; ++p) { h->pi[p] = order[p].i; h->pj[p] = order[p].j; }
    double lum = NAN;
    if (compute_lum) {   // the fold in collision order (gaussian.jl:50-56)
        lum = 0.0;
        for (int p = 0; p < npairs; ++p) lum += h->plum[p];
    }
    if (lum_out) *lum_out = lum;
    octopic_stats &S = h->stats;
    S.green_builds = 0;
    h->last_builds = 0;
    S.dropped = 0;
    S.denoised_planes = 0;
    S.denoise_fallbacks = 0;
    S.deposit_all_planes = 0;
    S.pairs = npairs;
    S.batches = (int64_t)batches.size();
    S.widest_batch = widest;
    S.planes = 0;
    S.collides += 1;
    if (h->timing) {
        for (auto &e : h->ev) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, e.second.first, e.second.second);
            h->fam_ms[e.first] += ms;
            h->fam_n[e.first] += 1;
        }
    }
}

void collide_impl(octopic_handle *h, octopic_beam *B1, octopic_beam *B2, int64_t turn, int compute_lum, double *lum_out) {
    check_beam(B1, 1);
    check_beam(B2, 2);
    hipStream_t s = h->stream;
    const octopic_options &o = h->opt;
    const int64_t n[2] = {B1->n, B2->n};
    octopic_beam *B[2] = {B1, B2};
    const int tiles = h->soft ? 0 : h->tiles;   // the soft-Gaussian solver keeps the slice order
    h->ev.clear();
    h->ev_next = 0;
    // ---- beams on the device
    double *bp[2][6];
    for (int b = 0; b < 2; ++b) {
        double *hp[6] = {B[b]->x, B[b]->px, B[b]->y, B[b]->py, B[b]->z, B[b]->pz};
        if (tiles > 0) {
            for (int c = 0; c < 6; ++c) h->srt[b][c].ensure(n[b]);
            h->perm_srt[b].ensure(n[b]);
            h->inv_srt[b].ensure(n[b]);
        }
        for (int c = 0; c < 6; ++c) {
            h->st[b][c].ensure(n[b]);
            if (B[b]->dtype == OCTOPIC_DTYPE_F32) {
                // Float32 beam: an exact FP64 copy for the collision, rounded back at the end
                h->beam_in[b][c].ensure(n[b]);
                const float *src = reinterpret_cast<const float *>(hp[c]);
                if (!B[b]->on_device) {
                    h->beam_f32[b][c].ensure(n[b]);
                    HIPCHECK(hipMemcpyAsync(h->beam_f32[b][c].p, hp[c], sizeof(float) * n[b], hipMemcpyHostToDevice, s));
                    src = h->beam_f32[b][c].p;
                }
                HIPCHECK(launch_widen(src, h->beam_in[b][c].p, n[b], s));
                bp[b][c] = h->beam_in[b][c].p;
            } else if (B[b]->on_device) {
                bp[b][c] = hp[c];
            } else {
                h->beam_in[b][c].ensure(n[b]);
                HIPCHECK(hipMemcpyAsync(h->beam_in[b][c].p, hp[c], sizeof(double) * n[b], hipMemcpyHostToDevice, s));
                bp[b][c] = h->beam_in[b][c].p;
            }
        }
    }
    upload_ptrs(h, bp);
    if (tiles > 0) HIPCHECK(hipMemsetAsync(h->box.p, 0, 8 * sizeof(unsigned long long), s));
    // ---- slicing
    {
        Timed t(h, F_SLICING);
        const octopic_slicing *sl[2] = {&o.slicing1, &o.slicing2};
        for (int b = 0; b < 2; ++b) run_slicing(h, b, *sl[b], n[b]);
        // the two beams' slicing chains are independent (own tables, states and scratch shares): beam 2's
        // runs on the side stream, forked from and joined back into the handle's stream -- except under
        // octopic_collide_many, whose instances already fill the GPU side by side (and whose streams share
        // the hardware queues, where an instance's join would wait behind other instances' work)
        const bool split = !h->in_many;
        if (split && !h->side_stream) {
            // created on first use: a stream takes one of the process's hardware queues in creation order, and
            // handles that only run under octopic_collide_many never need it
            HIPCHECK(hipStreamCreate(&h->side_stream));   // default priority
            HIPCHECK(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
            HIPCHECK(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
        }
        if (split) {
            HIPCHECK(hipEventRecord(h->ev_fork, s));
            HIPCHECK(hipStreamWaitEvent(h->side_stream, h->ev_fork, 0));
        }
        for (int b = 0; b < 2; ++b) {
            SlicingArgs a{};
            a.z = bp[b][4]; a.x = bp[b][0]; a.y = bp[b][2];
            a.box = tiles > 0 ? h->box.p + 4 * b : nullptr;
            a.n = n[b];
            a.method = sl[b]->method; a.ns = sl[b]->nslices; a.resolution = sl[b]->resolution;
            a.centroid = sl[b]->center_position == OCTOPIC_CENTER_CENTROID ? 1 : 0;
            a.q = h->q.p + (size_t)b * kMaxSlices;
            a.tab = h->tab.p + b;
            a.lanes = h->lanes.p + (size_t)b * kFoldLanes;
            a.mm_partial = h->mm.p + (size_t)b * 2 * kZFoldBlocks;
            a.centroid_lanes = h->clanes.p + (size_t)b * h->clanes_share * kFoldLanes;
            a.bin_counts = h->bins.p + (size_t)b * 16384;
            a.sid = h->sid[b].p; a.hist = h->hist[b].p;
            a.beam = h->ptrs.p + b * 6;
            a.st = h->ptrs.p + 12 + b * 6;
            a.st_z = h->st[b][4].p;
            a.perm = h->perm[b].p;
            a.inv = tiles > 0 ? nullptr : h->inv_sl[b].p;   // the tile order stores through its own inverse
            a.err = h->errinfo.p;
            if (a.method == OCTOPIC_SLICE_EQUAL_COUNT) {
                a.rank_keys = h->rank_keys[b].p; a.rank_sorted = h->rank_sorted[b].p; a.rank_vals = h->rank_vals[b].p;
            }
            HIPCHECK(launch_slicing(a, b == 0 || !split ? s : h->side_stream));
        }
        if (split) {
            HIPCHECK(hipEventRecord(h->ev_join, h->side_stream));
            HIPCHECK(hipStreamWaitEvent(s, h->ev_join, 0));
        }
    }
    HIPCHECK(hipMemcpyAsync(h->h_tab, h->tab.p, 2 * sizeof(SliceTable), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(h->h_err, h->errinfo.p, sizeof(ErrorInfo), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (h->h_err->code) raise_device_error(h, *h->h_err, turn);
    const SliceTable &T1 = h->h_tab[0], &T2 = h->h_tab[1];
    const int ns1 = o.slicing1.nslices, ns2 = o.slicing2.nslices;
    h->ns[0] = ns1; h->ns[1] = ns2; h->nbeam[0] = n[0]; h->nbeam[1] = n[1];
    for (int k = 0; k <= ns1; ++k)
        if (!std::isfinite(T1.boundary[k])) raise_device_error(h, ErrorInfo{OCTOPIC_ERR_NONFINITE, -1, 0, 5, 0, 0}, turn);
    for (int k = 0; k <= ns2; ++k)
        if (!std::isfinite(T2.boundary[k])) raise_device_error(h, ErrorInfo{OCTOPIC_ERR_NONFINITE, -1, 0, 5, 0, 0}, turn);
    if (h->soft) {
        soft_collide(h, B, bp, n, turn, compute_lum, lum_out);
        return;
    }

    // ---- transverse tile order inside the slices (queued before the host builds the schedule)
    if (tiles > 0) {
        Timed t(h, F_TILE);
        const int TT = tiles * tiles;
        for (int b = 0; b < 2; ++b) {
            const SliceTable &T = h->h_tab[b];
            int64_t mc = 0;
            for (int k = 0; k < T.ns; ++k) mc = std::max<int64_t>(mc, T.count[k]);
            // few blocks per slice: each block's members of one tile land in one contiguous run
            int G = (int)std::min<int64_t>(64, std::max<int64_t>(1, (mc + kTileBlockSpan - 1) / kTileBlockSpan));
            G = std::max(1, std::min<int>(G, (int)((int64_t)(1 << 22) / ((int64_t)T.ns * TT))));
            h->thist.ensure((size_t)T.ns * TT * G);
            TileSortArgs a{};
            a.tab = h->tab.p + b; a.ns = T.ns; a.T = tiles; a.box = h->box.p + 4 * b;
            for (int c = 0; c < 6; ++c) { a.src[c] = h->st[b][c].p; a.dst[c] = h->srt[b][c].p; }
            a.perm_src = h->perm[b].p; a.perm_dst = h->perm_srt[b].p; a.inv = h->inv_srt[b].p;
            a.thist = h->thist.p;
            HIPCHECK(launch_tile_sort(a, G, s));
        }
    }

    // ---- schedule (host): collision order and conflict-free batches
    const double kbb1 = kbb_of(o.kbb1, B1, B2, true, n[1]);
    const double kbb2 = kbb_of(o.kbb2, B2, B1, false, n[0]);
    const double klum = !std::isnan(o.luminosity_scale) ? o.luminosity_scale : B1->npart * B2->npart / (double)(n[0] * n[1]);
    auto order = collision_order(T1.center, ns1, T2.center, ns2);
    auto batches = pair_batches(order, ns1, ns2);
    const int npairs = (int)order.size();
    std::vector<BatchPair> hp;
    std::vector<int> boff, bcnt;
    int widest = 0;
    int64_t maxcount = 0;
    for (int k = 0; k < ns1; ++k) maxcount = std::max<int64_t>(maxcount, T1.count[k]);
    for (int k = 0; k < ns2; ++k) maxcount = std::max<int64_t>(maxcount, T2.count[k]);
    h->pran.assign(npairs, 0);
    for (auto &bt : batches) {
        boff.push_back((int)hp.size());
        int c = 0;
        for (int pos : bt) {
            const auto &e = order[pos];
            if (T1.count[e.i] == 0 || T2.count[e.j] == 0) continue;   // an empty slice skips the pair (pic_cpu.jl:302)
            hp.push_back(BatchPair{pos, e.i, e.j, 0});
            h->pran[pos] = 1;
            ++c;
        }
        bcnt.push_back(c);
        widest = std::max(widest, c);
    }
    const int W = std::max(widest, 1);
    // fixed-point scale 2^k with k <= 40 and (maxcount + 1) * 2^k < 2^62
    int kbits = 40;
    while (kbits > 8 && (double)(maxcount + 1) * std::ldexp(1.0, kbits) >= std::ldexp(1.0, 62)) --kbits;
    const double scale = std::ldexp(1.0, kbits), inv_scale = std::ldexp(1.0, -kbits);

    // ---- buffers sized for the widest batch
    const int nx = h->nx, ny = h->ny, npl = h->npl;
    const int64_t field_cells = (int64_t)nx * ny, lum_cells = (int64_t)(h->lnx + 1) * (h->lny + 1);
    const int64_t charge_stride = std::max(field_cells, lum_cells);
    const int nfield_planes = 2 * W * npl, nplanes_max = nfield_planes + 2 * W;
    if (h->dn) {
        int dh = 0, dw = 0, dmb = 0;
        denoiser_shape(h->dn, &dh, &dw, &dmb);
        if (dmb < nfield_planes)
            fail(OCTOPIC_ERR_INVALID_ARGUMENT, "the denoiser's max_batch (" + std::to_string(dmb) +
                                                   ") is below the widest batch's field planes (" + std::to_string(nfield_planes) + ")");
        h->dn_mom.ensure((size_t)2 * W * kDnMomentChunks * 10);
        h->dn_meta.ensure((size_t)nfield_planes * 4);
        h->dn_prof.ensure((size_t)nfield_planes * (nx + ny));
        h->dn_in.ensure((size_t)nfield_planes * field_cells * 2);
    }
    h->pairs.ensure(std::max<size_t>(hp.size(), 1));
    if (!hp.empty())
        HIPCHECK(hipMemcpyAsync(h->pairs.p, hp.data(), sizeof(BatchPair) * hp.size(), hipMemcpyHostToDevice, s));
    if (!hp.empty()) {
        // per batch: interactions by field-slice population, largest first (stable)
        std::vector<int32_t> ord(2 * hp.size());
        for (size_t bi = 0; bi < batches.size(); ++bi) {
            const int m = 2 * bcnt[bi], o0 = 2 * boff[bi];
            auto fld_count = [&](int d) {   // d = 2k + dir - 1: dir 1 kicks beam 2's slice j, dir 2 beam 1's slice i
                const BatchPair &pr = hp[boff[bi] + d / 2];
                return (d & 1) ? T1.count[pr.i] : T2.count[pr.j];
            };
            std::vector<int32_t> idx(m);
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return fld_count(x) > fld_count(y); });
            std::copy(idx.begin(), idx.end(), ord.begin() + o0);
        }
        h->korder.ensure(ord.size());
        HIPCHECK(hipMemcpyAsync(h->korder.p, ord.data(), sizeof(int32_t) * ord.size(), hipMemcpyHostToDevice, s));
        // deposit planes by source-slice population, largest first (stable): per batch the field planes
        // (interaction d = 2k + dir - 1 deposits beam 1's slice i for dir 1, beam 2's slice j for dir 2),
        // then the luminosity planes (2k + side)
        std::vector<int32_t> dord((size_t)(h->npl + 1) * 2 * hp.size());
        for (size_t bi = 0; bi < batches.size(); ++bi) {
            const int B = bcnt[bi];
            const size_t o0 = (size_t)(h->npl + 1) * 2 * boff[bi];
            auto src_count = [&](int beam_side, int k) {
                const BatchPair &pr = hp[boff[bi] + k];
                return beam_side == 0 ? T1.count[pr.i] : T2.count[pr.j];
            };
            std::vector<int32_t> f(2 * B * h->npl), l(2 * B);
            std::iota(f.begin(), f.end(), 0);
            std::iota(l.begin(), l.end(), 0);
            std::stable_sort(f.begin(), f.end(), [&](int32_t x, int32_t y) {
                const int dx = x / h->npl, dy = y / h->npl;
                return src_count(dx & 1, dx / 2) > src_count(dy & 1, dy / 2);
            });
            std::stable_sort(l.begin(), l.end(), [&](int32_t x, int32_t y) {
                return src_count(x & 1, x / 2) > src_count(y & 1, y / 2);
            });
            std::copy(f.begin(), f.end(), dord.begin() + o0);
            std::copy(l.begin(), l.end(), dord.begin() + o0 + f.size());
        }
        h->dorder.ensure(dord.size());
        HIPCHECK(hipMemcpyAsync(h->dorder.p, dord.data(), sizeof(int32_t) * dord.size(), hipMemcpyHostToDevice, s));
        h->dqueue.ensure(2 * batches.size());
        HIPCHECK(hipMemsetAsync(h->dqueue.p, 0, sizeof(int) * 2 * batches.size(), s));
    }
    h->keys.ensure_zero((size_t)2 * W * kBoundSlots, s);
    h->chunkbox.ensure((size_t)2 * W * kMaxChunks * 8);
    h->inters.ensure_zero(2 * W, s);
    h->planes.ensure_zero(nplanes_max, s);
    h->jobs.ensure_zero(2 * W, s);
    h->lumgeo.ensure(3 * W);
    h->charge.ensure_zero((size_t)nplanes_max * charge_stride, s);
    if (h->fp32) {
        h->qh32.ensure((size_t)nfield_planes * (nx + 1) * ny);
        h->phi32.ensure((size_t)nfield_planes * field_cells);
        h->table32.ensure((size_t)2 * W * field_cells * 3 * npl);
    } else {
        h->qh.ensure((size_t)nfield_planes * (nx + 1) * ny);
        h->phi.ensure((size_t)nfield_planes * field_cells);
        h->table.ensure((size_t)2 * W * field_cells * 3 * npl);
    }
    if (h->gauss) {
        h->gmom.ensure((size_t)2 * W * kDnMomentChunks * 10);
        h->gab.ensure((size_t)4 * (n[0] + n[1]));
    }
    h->records.ensure((size_t)2 * std::max(npairs, 1));
    h->lum_parts.ensure(std::max(npairs, 1));
    h->lum_partial.ensure((size_t)W * kLumBlocks);
    HIPCHECK(hipMemsetAsync(h->records.p, 0, sizeof(octopic_interaction_record) * 2 * npairs, s));
    HIPCHECK(hipMemsetAsync(h->lum_parts.p, 0, sizeof(double) * npairs, s));
    const int temp_slot0 = h->nkeys;
    const int nslots = h->nkeys + (o.green_cache == OCTOPIC_CACHE_SLICE_PAIR ? 0 : 2 * W);
    h->spec.ensure((size_t)std::max(nslots, 1) * h->slot_stride);
    if (h->fp32) h->spec32.ensure((size_t)std::max(nslots, 1) * h->slot_stride);
    // the counters as this collide's batches start, snapshotted on the device (no host wait here)
    HIPCHECK(hipMemcpyAsync(h->counters.p + 8, h->counters.p, 8 * sizeof(unsigned long long), hipMemcpyDeviceToDevice, s));

    GeomOpts go{};
    go.nx = nx; go.ny = ny; go.lum_nx = h->lnx; go.lum_ny = h->lny;
    go.green_type = o.green_type; go.green_cache = o.green_cache; go.quadratic = h->mode == 1;
    go.interaction_grid = o.interaction_grid; go.compute_lum = compute_lum ? 1 : 0;
    go.lkick = o.longitudinal_kick ? 1 : 0;
    go.ns1 = ns1; go.ns2 = ns2; go.npl = npl; go.temp_slot0 = temp_slot0; go.epoch = h->epoch;
    go.gauss = h->gauss ? 1 : 0; go.mom_chunks = kDnMomentChunks; go.margin = o.margin_sigma;
    go.min_x = o.min_extent_x; go.min_y = o.min_extent_y; go.quantize = o.grid_quantize;
    go.min_ratio = o.green_min_ratio; go.growth = o.green_growth;
    go.kbb1 = kbb1; go.kbb2 = kbb2; go.klum = klum;

    States st{};
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 6; ++c) st.c[b][c] = tiles > 0 ? h->srt[b][c].p : h->st[b][c].p;
    const SliceTable *dt1 = h->tab.p, *dt2 = h->tab.p + 1;

    int slot_node1 = 0, slot_node2 = 0;
    if (h->mode == 2) {
        // node meshes of both directions at turn start (pic_cpu.jl:113-119, 1404-1419)
        Timed t(h, F_NODE);
        const int nn1 = ns1 * (ns2 + 1), nn2 = ns2 * (ns1 + 1);
        slot_node1 = 0;
        slot_node2 = nn1;
        h->spec.ensure((size_t)(nn1 + nn2) * h->slot_stride);
        if (h->fp32) h->spec32.ensure((size_t)(nn1 + nn2) * h->slot_stride);
        h->nodes.ensure_zero(nn1 + nn2, s);
        h->node_jobs.ensure_zero(nn1 + nn2, s);
        const int nch = (int)std::min<int64_t>(32, std::max<int64_t>(1, (maxcount + 8191) / 8192));
        const int nsmax = std::max(ns1, ns2);
        const size_t skn = (size_t)nsmax * (nsmax + 1) * 4, fkn = (size_t)nsmax * nsmax * 4;
        h->node_keys.ensure(2 * (skn + fkn));
        HIPCHECK(hipMemsetAsync(h->node_keys.p, 0, 2 * (skn + fkn) * sizeof(unsigned long long), s));
        for (int dir = 1; dir <= 2; ++dir) {
            NodeArgs a{};
            a.tab_src = dir == 1 ? dt1 : dt2;
            a.tab_fld = dir == 1 ? dt2 : dt1;
            a.st = st;
            a.src_beam = dir == 1 ? 0 : 1;
            a.fld_beam = dir == 1 ? 1 : 0;
            a.dir = dir;
            a.nchunks = nch;
            a.skeys = h->node_keys.p + (dir - 1) * (skn + fkn);
            a.fkeys = a.skeys + skn;
            a.o = go;
            a.slot0 = dir == 1 ? slot_node1 : slot_node2;
            a.nodes = h->nodes.p + (dir == 1 ? 0 : nn1);
            a.jobs = h->node_jobs.p + (dir == 1 ? 0 : nn1);
            a.counters = h->counters.p;
            a.err = h->errinfo.p;
            HIPCHECK(launch_node_prepare(a, dir == 1 ? ns1 : ns2, dir == 1 ? ns2 : ns1, s));
        }
        HIPCHECK(launch_green_build(h->node_jobs.p, nn1 + nn2, h->spec.p, h->fp32 ? h->spec32.p : nullptr, h->slot_stride,
                                    nn1 + nn2, nx, ny, o.green_type == OCTOPIC_GREEN_INTEGRATED ? 1 : 0, h->px.dev,
                                    h->py.dev, h->errinfo.p, s, 1));
        h->ran_dev.ensure(std::max(npairs, 1));
        HIPCHECK(hipMemsetAsync(h->ran_dev.p, 0, sizeof(int32_t) * npairs, s));
    }

    int64_t planes_total = 0, builds_guess = 0;
    (void)builds_guess;
    // the automatic deposit form: all-planes items timed in one collide, per-plane items in the next
    const int trial = (h->deposit_multi == -1 && h->mode != 2 && npl > 1 && h->deposit_trial % kDepositTrialEvery < 2)
                          ? (int)(h->deposit_trial % kDepositTrialEvery) : -1;
    if (h->deposit_multi == -1 && h->mode != 2 && npl > 1) ++h->deposit_trial;
    if (trial >= 0)
        while (h->deposit_ev.size() < 2 * batches.size()) {
            hipEvent_t ev;
            HIPCHECK(hipEventCreate(&ev));
            h->deposit_ev.push_back(ev);
        }
    const int green_full = (h->last_builds < 0 || h->last_builds >= (int64_t)batches.size()) ? 1 : 0;
    for (size_t bi = 0; bi < batches.size(); ++bi) {
        const int B = bcnt[bi];
        if (B == 0) continue;
        const BatchPair *dp = h->pairs.p + boff[bi];
        int64_t bmax = 0;
        for (int k = 0; k < B; ++k) {
            const BatchPair &pr = hp[boff[bi] + k];
            bmax = std::max<int64_t>(bmax, std::max<int64_t>(T1.count[pr.i], T2.count[pr.j]));
        }
        // one particle-range chunking for bounds, deposit and kick (chunk boxes describe the blocks)
        const int nchunks = (int)std::min<int64_t>(kMaxChunks, std::max<int64_t>(1, (bmax + kChunkSpan - 1) / kChunkSpan));
        const double *cbox = h->mode == 2 ? nullptr : h->chunkbox.p;   // node mode folds no source/field boxes
        {
            Timed t(h, F_BOUNDS);
            BoundsArgs a{dp, B, dt1, dt2, st, nchunks, h->mode == 2 ? 0 : 1, compute_lum ? 1 : 0, h->keys.p,
                         h->chunkbox.p, h->errinfo.p};
            HIPCHECK(launch_bounds(a, s));
        }
        if (h->gauss) {   // the source moments the geometry stage builds the subtracted Gaussians from
            Timed t(h, F_GEOMETRY);
            GpicMomentArgs a{dp, B, dt1, dt2, st, kDnMomentChunks, h->gmom.p, h->errinfo.p};
            HIPCHECK(launch_gpic_moments(a, s));
        }
        if (h->mode == 2) {
            Timed t(h, F_GEOMETRY);
            NodeGeometryArgs a{};
            a.pairs = dp; a.npairs = B; a.tab1 = dt1; a.tab2 = dt2; a.keys = h->keys.p;
            a.o = go; a.nodes1 = h->nodes.p; a.nodes2 = h->nodes.p + ns1 * (ns2 + 1);
            a.slot1 = slot_node1; a.slot2 = slot_node2;
            a.inters = h->inters.p; a.planes = h->planes.p; a.lumgeo = h->lumgeo.p; a.records = h->records.p;
            a.ran = h->ran_dev.p; a.err = h->errinfo.p;
            HIPCHECK(launch_node_geometry(a, s));
        } else {
            Timed t(h, F_GEOMETRY);
            GeometryArgs a{};
            a.pairs = dp; a.npairs = B; a.tab1 = dt1; a.tab2 = dt2; a.keys = h->keys.p;
            a.o = go; a.cache = h->cache.p; a.inters = h->inters.p; a.planes = h->planes.p; a.jobs = h->jobs.p;
            a.lumgeo = h->lumgeo.p; a.records = h->records.p; a.counters = h->counters.p; a.err = h->errinfo.p;
            a.mom_part = h->gmom.p; a.st = st;
            HIPCHECK(launch_geometry(a, s));
        }
        if (h->mode != 2) {
            Timed t(h, F_GREEN);
            HIPCHECK(launch_green_build(h->jobs.p, 2 * B, h->spec.p, h->fp32 ? h->spec32.p : nullptr, h->slot_stride, nslots, nx,
                                        ny, o.green_type == OCTOPIC_GREEN_INTEGRATED ? 1 : 0, h->px.dev, h->py.dev,
                                        h->errinfo.p, s, green_full));
        }
        const int nfp = 2 * B * npl;
        {
            Timed t(h, F_DEPOSIT);
            DepositArgs a{};
            a.planes = h->planes.p; a.nplanes = nfp; a.st = st; a.method = o.deposit; a.ncells = (int)field_cells;
            const bool all_planes = h->deposit_multi == 1 ||
                                    (h->deposit_multi == -1 && (trial >= 0 ? trial == 0 : !h->deposit_per_plane));
            a.multi = (h->mode == 2 || !all_planes) ? 1 : npl;
            // ~8k particles per block: the LDS mesh is zeroed and flushed once per block; when a block takes
            // all planes of an interaction and the slices are small, fewer particles per block, down to one
            // chunk, until the launch has two items per compute unit
            a.chunks = nchunks;
            auto group_of = [&](int64_t span) {
                return std::max(1, (int)std::min<int64_t>(nchunks, span * nchunks / std::max<int64_t>(bmax, 1)));
            };
            int64_t span = h->deposit_span;
            if (a.multi > 1) {
                const int units = nfp / a.multi;
                while (span > std::max<int64_t>(kChunkSpan, h->deposit_min_span)) {
                    const int g = group_of(span);
                    if ((int64_t)units * ((nchunks + g - 1) / g) >= 2 * (int64_t)std::max(h->cus, 1)) break;
                    span /= 2;
                }
            }
            a.group = group_of(span);
            a.scale = scale; a.charge = h->charge.p; a.charge_stride = charge_stride;
            a.dropped = h->counters.p + 4; a.chunkbox = cbox; a.err = h->errinfo.p;
            a.fp32 = h->fp32 ? 1 : 0;
            a.queue = h->dqueue.p + 2 * bi;
            a.order = h->dorder.p + (size_t)(npl + 1) * 2 * boff[bi];
            // slice-pair meshes: the planes of an interaction share mesh and particles, one read feeds them all
            // the luminosity planes ride along when they use the field deposit's stencil
            const bool lum_fold = compute_lum && a.multi > 1 && h->lum_dep == o.deposit;
            a.lum_plane0 = lum_fold ? nfp : -1;
            h->deposit_all_planes_used = a.multi > 1;
            if (trial >= 0) HIPCHECK(hipEventRecord(h->deposit_ev[2 * bi], s));
            HIPCHECK(launch_deposit(a, s));
            if (compute_lum && !lum_fold) {
                DepositArgs l = a;
                l.fp32 = 0;   // luminosity deposits stay FP64
                l.planes = h->planes.p + nfp; l.nplanes = 2 * B; l.method = h->lum_dep; l.ncells = (int)lum_cells;
                l.charge = h->charge.p + (size_t)nfp * charge_stride;
                l.queue = h->dqueue.p + 2 * bi + 1;
                l.order = a.order + nfp;
                l.multi = 1;
                l.lum_plane0 = -1;
                HIPCHECK(launch_deposit(l, s));
            }
            if (trial >= 0) HIPCHECK(hipEventRecord(h->deposit_ev[2 * bi + 1], s));
        }
        if (h->gauss) {   // the Gaussian subtracted from each field plane before the convolution
            Timed t(h, F_DEPOSIT);
            GpicResidualArgs a{h->inters.p, h->planes.p, nfp, h->charge.p, charge_stride, scale, inv_scale, o.deposit,
                               nx, ny, o.neutralize ? 1 : 0, h->errinfo.p};
            HIPCHECK(launch_gpic_residual(a, s));
        }
        if (h->dn) {
            // field planes through the residual denoiser before the convolution
            Timed t(h, F_DENOISE);
            DnCollideArgs a{};
            a.inters = h->inters.p; a.ninters = 2 * B; a.planes = h->planes.p; a.nplanes = nfp; a.st = st;
            a.charge = h->charge.p; a.charge_stride = charge_stride; a.scale = scale; a.inv_scale = inv_scale;
            a.method = o.deposit; a.nx = nx; a.ny = ny; a.chunks = kDnMomentChunks;
            a.mom_part = h->dn_mom.p; a.meta = h->dn_meta.p; a.prof = h->dn_prof.p; a.input = h->dn_in.p;
            a.limit = h->dn_limit; a.counters = h->counters.p + 6; a.err = h->errinfo.p;
            HIPCHECK(launch_dn_prepare(a, s));
            hipError_t e = hipSuccess;
            const float *out = denoiser_run(h->dn, h->dn_in.p, nfp, s, &e);
            HIPCHECK(e);
            HIPCHECK(launch_dn_apply(a, out, s));
        }
        const int64_t qh_stride = (int64_t)(nx + 1) * ny;
        // the batch's luminosity overlap sums run as extra blocks of the charge row transform
        LumArgs lum{};
        if (compute_lum) {
            lum.pairs = dp; lum.npairs = B; lum.charge = h->charge.p + (size_t)nfp * charge_stride;
            lum.charge_stride = charge_stride; lum.ncells = (int)lum_cells; lum.inv_scale = inv_scale;
            lum.lumgeo = h->lumgeo.p; lum.lum_parts = h->lum_parts.p; lum.lum_partial = h->lum_partial.p;
            lum.err = h->errinfo.p;
        }
        auto solve = [&](auto *qh, const auto *spec, auto *phi, auto *table) {
            {
                Timed t(h, F_ROWS_FWD);
                HIPCHECK(launch_rows_fwd(h->planes.p, nfp, h->charge.p, charge_stride, qh, qh_stride, nx, ny, inv_scale,
                                         h->px.dev, h->errinfo.p, lum, s, h->fft_group));
            }
            {
                Timed t(h, F_COLS);
                HIPCHECK(launch_cols_conv(h->planes.p, nfp, qh, qh_stride, spec, h->slot_stride, nx, ny, h->py.dev,
                                          h->errinfo.p, lum, s, h->fft_group));
            }
            {
                Timed t(h, F_ROWS_INV);
                HIPCHECK(launch_rows_inv(h->planes.p, nfp, (const decltype(qh))qh, qh_stride, phi, nx, ny, h->px.dev,
                                         h->errinfo.p, s, h->fft_group));
            }
            {
                Timed t(h, F_FIELD);
                HIPCHECK(launch_field(h->inters.p, 2 * B, h->planes.p, (const decltype(phi))phi, table, nx, ny,
                                      o.field_derivative == 4 ? 1 : 0, npl, h->errinfo.p, s));
            }
        };
        if (h->fp32) solve(h->qh32.p, h->spec32.p, h->phi32.p, h->table32.p);
        else solve(h->qh.p, h->spec.p, h->phi.p, h->table.p);
        if (h->gauss) {   // the exact field of the subtracted Gaussians at every field particle
            Timed t(h, F_KICK);
            GpicAddbackArgs a{h->inters.p, 2 * B, st, h->gab.p, {0, n[0]}, o.longitudinal_kick ? 1 : 0, h->errinfo.p};
            HIPCHECK(launch_gpic_addback(a, bmax, s));
        }
        {
            Timed t(h, F_KICK);
            KickArgs a{};
            a.inters = h->inters.p; a.ninters = 2 * B; a.st = st; a.nx = nx; a.ny = ny; a.npl = npl;
            a.table = h->fp32 ? (const void *)h->table32.p : (const void *)h->table.p;
            a.fp32 = h->fp32 ? 1 : 0;
            a.method = o.deposit; a.mode = h->gauss ? 3 : h->mode; a.lkick = o.longitudinal_kick ? 1 : 0;
            if (h->gauss) { a.ab = h->gab.p; a.ab_base[0] = 0; a.ab_base[1] = n[0]; }
            a.chunks = nchunks;
            a.subdiv = 1;
            if (h->kick_target_blocks > 0) {
                const int per = nchunks * 2 * B;
                // parts of chunks for about kick_target_blocks items, with items of at least a block's
                // threads in particles on average (each item stages its chunk's LDS window)
                int sd = (h->kick_target_blocks + per - 1) / per;
                int64_t tot = 0;
                for (int k = 0; k < B; ++k) { const BatchPair &pr = hp[boff[bi] + k]; tot += T1.count[pr.i] + T2.count[pr.j]; }
                const int64_t cap = std::max<int64_t>(1, tot / ((int64_t)per * 512));   // 512: the kick's block size
                if (sd > cap) sd = (int)cap;
                a.subdiv = sd < 1 ? 1 : (sd > 64 ? 64 : sd);
            }
            a.pass_windows = h->kick_pass_windows >= 0 ? h->kick_pass_windows : (h->fp32 ? 0 : 1);
            a.order = h->korder.p + (size_t)2 * boff[bi];
            a.dropped = h->counters.p + 4;
            a.chunkbox = cbox;
            a.err = h->errinfo.p;
            HIPCHECK(launch_kick(a, s));
        }
        planes_total += nfp;
    }
    {
        Timed t(h, F_STORE);
        for (int b = 0; b < 2; ++b) {
            if (tiles > 0)
                HIPCHECK(launch_store_gather(h->ptrs.p + b * 6, n[b], h->ptrs.p + 24 + b * 6, h->inv_srt[b].p, h->errinfo.p, s));
            else
                HIPCHECK(launch_store_gather(h->ptrs.p + b * 6, n[b], h->ptrs.p + 12 + b * 6, h->inv_sl[b].p, h->errinfo.p, s));
        }
    }
    // ---- results
    HIPCHECK(hipMemcpyAsync(h->h_err, h->errinfo.p, sizeof(ErrorInfo), hipMemcpyDeviceToHost, s));
    h->plum.assign(npairs, 0.0);
    if (npairs) HIPCHECK(hipMemcpyAsync(h->plum.data(), h->lum_parts.p, sizeof(double) * npairs, hipMemcpyDeviceToHost, s));
    h->recs.assign((size_t)2 * npairs, octopic_interaction_record{});
    if (npairs)
        HIPCHECK(hipMemcpyAsync(h->recs.data(), h->records.p, sizeof(octopic_interaction_record) * 2 * npairs,
                                hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(h->h_counters, h->counters.p, 16 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    std::vector<int32_t> ran_node;
    if (h->mode == 2 && npairs) {
        ran_node.resize(npairs);
        HIPCHECK(hipMemcpyAsync(ran_node.data(), h->ran_dev.p, sizeof(int32_t) * npairs, hipMemcpyDeviceToHost, s));
    }
    HIPCHECK(hipStreamSynchronize(s));
    if (h->h_err->code) raise_device_error(h, *h->h_err, turn);
    if (h->mode == 2)
        for (int p = 0; p < npairs; ++p) h->pran[p] = h->pran[p] && ran_node[p];
    if (trial >= 0) {
        double ms = 0.0;
        for (size_t bi = 0; bi < batches.size(); ++bi) {
            if (bcnt[bi] == 0) continue;
            float e = 0.f;
            HIPCHECK(hipEventElapsedTime(&e, h->deposit_ev[2 * bi], h->deposit_ev[2 * bi + 1]));
            ms += e;
        }
        h->deposit_trial_ms[trial] = ms;
        if (trial == 1) h->deposit_per_plane = h->deposit_trial_ms[1] < h->deposit_trial_ms[0];
    }
    write_back(h, B, bp, n, false);
    h->pi.resize(npairs); h->pj.resize(npairs);
    for (int p = 0; p < npairs; ++p) { h->pi[p] = order[p].i; h->pj[p] = order[p].j; }
    double lum = NAN;
    if (compute_lum) {   // the fold in collision order (pic_cpu.jl:258-271)
        lum = 0.0;
        for (int p = 0; p < npairs; ++p) if (h->pran[p]) lum += h->plum[p];
    }
    if (lum_out) *lum_out = lum;
    octopic_stats &S = h->stats;
    const unsigned long long *counters_before = h->h_counters + 8;
    S.green_hits = (int64_t)h->h_counters[0];
    S.green_misses = (int64_t)h->h_counters[1];
    S.green_rebuilds = (int64_t)h->h_counters[2];
    S.green_builds = (int64_t)(h->h_counters[3] - counters_before[3]);
    h->last_builds = S.green_builds;
    S.dropped = (int64_t)(h->h_counters[4] - counters_before[4]);
    S.denoised_planes = (int64_t)(h->h_counters[6] - counters_before[6]);
    S.denoise_fallbacks = (int64_t)(h->h_counters[7] - counters_before[7]);
    S.deposit_all_planes = h->deposit_all_planes_used ? 1 : 0;
    S.pairs = npairs;
    S.batches = (int64_t)batches.size();
    S.widest_batch = widest;
    S.planes = planes_total;
    S.collides += 1;
    if (h->timing) {
        for (auto &e : h->ev) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, e.second.first, e.second.second);
            h->fam_ms[e.first] += ms;
            h->fam_n[e.first] += 1;
        }
    }
}

}  // namespace

extern "C" {

int octopic_get_stats(const octopic_handle *h, octopic_stats *out) {
    if (!h || !out) return OCTOPIC_ERR_INVALID_ARGUMENT;
    *out = h->stats;
    return OCTOPIC_OK;
}

int octopic_pair_luminosity(const octopic_handle *h, int32_t *i, int32_t *j, int32_t *ran, double *lum, int64_t cap,
                            int64_t *n) {
    if (!h || !n) return OCTOPIC_ERR_INVALID_ARGUMENT;
    *n = (int64_t)h->pi.size();
    for (int64_t k = 0; k < std::min<int64_t>(cap, *n); ++k) {
        if (i) i[k] = h->pi[k];
        if (j) j[k] = h->pj[k];
        if (ran) ran[k] = h->pran[k];
        if (lum) lum[k] = h->plum[k];
    }
    return OCTOPIC_OK;
}

int octopic_interaction_records(const octopic_handle *h, octopic_interaction_record *out, int64_t cap, int64_t *n) {
    if (!h || !n) return OCTOPIC_ERR_INVALID_ARGUMENT;
    *n = (int64_t)h->recs.size();
    for (int64_t k = 0; k < std::min<int64_t>(cap, *n); ++k) out[k] = h->recs[k];
    return OCTOPIC_OK;
}

int octopic_last_slices(const octopic_handle *h, int32_t which, double *boundaries, double *centers, double *weights,
                        int64_t *counts, int64_t *indices, int64_t n) {
    if (!h || (which != 1 && which != 2)) return OCTOPIC_ERR_INVALID_ARGUMENT;
    const int b = which - 1;
    const SliceTable &T = h->h_tab[b];
    const int ns = h->ns[b];
    if (ns == 0) return set_err(const_cast<octopic_handle *>(h), OCTOPIC_ERR_INVALID_ARGUMENT, "no collide has run");
    SliceTable full;
    hipSetDevice(h->device);
    if (hipMemcpy(&full, h->tab.p + b, sizeof(SliceTable), hipMemcpyDeviceToHost) != hipSuccess) return OCTOPIC_ERR_DEVICE;
    (void)T;
    if (boundaries) std::memcpy(boundaries, full.boundary, sizeof(double) * (ns + 1));
    if (centers) std::memcpy(centers, full.center, sizeof(double) * ns);
    if (weights) std::memcpy(weights, full.weight, sizeof(double) * ns);
    if (counts) std::memcpy(counts, full.count, sizeof(int64_t) * ns);
    if (indices) {
        if (n != h->nbeam[b]) return set_err(const_cast<octopic_handle *>(h), OCTOPIC_ERR_INVALID_ARGUMENT, "index buffer size");
        if (hipMemcpy(indices, h->perm[b].p, sizeof(int64_t) * n, hipMemcpyDeviceToHost) != hipSuccess) return OCTOPIC_ERR_DEVICE;
    }
    return OCTOPIC_OK;
}

int octopic_track_linear6d(octopic_handle *h, const double *m, octopic_beam *b) {
    if (!h || !m || !b) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (!b->on_device) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "octopic_track_linear6d needs device beam pointers");
    double *c[6] = {b->x, b->px, b->y, b->py, b->z, b->pz};
    hipSetDevice(h->device);
    if (launch_linear6d(m, c, b->n, h->stream) != hipSuccess) return set_err(h, OCTOPIC_ERR_DEVICE, "linear6d launch failed");
    return OCTOPIC_OK;
}

extern "C++" {
namespace {
template <typename T>
int be_kick(octopic_handle *h, const T *x, const T *y, int64_t n, T sigx, T sigy, T *kx, T *ky) {
    if (!h) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (n < 0 || (n > 0 && (!x || !y || !kx || !ky)))
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "gaussian_beambeam_kick: null array or negative count");
    if (!(std::isfinite(sigx) && std::isfinite(sigy) && sigx >= T(0) && sigy >= T(0)))
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "gaussian_beambeam_kick: sizes must be finite and non-negative");
    hipSetDevice(h->device);
    if (launch_be_kick<T>(x, y, n, sigx, sigy, kx, ky, h->stream) != hipSuccess)
        return set_err(h, OCTOPIC_ERR_DEVICE, "gaussian_beambeam_kick launch failed");
    return OCTOPIC_OK;
}

template <typename T>
int be_response(octopic_handle *h, const T *x, const T *y, int64_t n, T kbb, T sigx, T sigy, T *out4) {
    if (!h) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (n < 0 || (n > 0 && (!x || !y || !out4)))
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "gaussian_beambeam_kick_response: null array or negative count");
    if (!(std::isfinite(sigx) && std::isfinite(sigy) && sigx > T(0) && sigy > T(0) && std::isfinite(kbb)))
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "gaussian_beambeam_kick_response: sizes must be finite and positive");
    hipSetDevice(h->device);
    if (launch_be_response<T>(x, y, n, kbb, sigx, sigy, out4, h->stream) != hipSuccess)
        return set_err(h, OCTOPIC_ERR_DEVICE, "gaussian_beambeam_kick_response launch failed");
    return OCTOPIC_OK;
}

template <typename T>
int faddeeva(octopic_handle *h, const T *zr, const T *zi, int64_t n, T *wr, T *wi) {
    if (!h) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (n < 0 || (n > 0 && (!zr || !zi || !wr || !wi)))
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "faddeeva_w: null array or negative count");
    hipSetDevice(h->device);
    if (launch_faddeeva<T>(zr, zi, n, wr, wi, h->stream) != hipSuccess)
        return set_err(h, OCTOPIC_ERR_DEVICE, "faddeeva_w launch failed");
    return OCTOPIC_OK;
}
}  // namespace
}  // extern "C++"

int octopic_gaussian_beambeam_kick(octopic_handle *h, const double *x, const double *y, int64_t n, double sigx,
                                   double sigy, double *kx, double *ky) {
    return be_kick<double>(h, x, y, n, sigx, sigy, kx, ky);
}

int octopic_gaussian_beambeam_kick_f32(octopic_handle *h, const float *x, const float *y, int64_t n, float sigx,
                                       float sigy, float *kx, float *ky) {
    return be_kick<float>(h, x, y, n, sigx, sigy, kx, ky);
}

int octopic_gaussian_beambeam_kick_response(octopic_handle *h, const double *x, const double *y, int64_t n, double kbb,
                                            double sigx, double sigy, double *out4) {
    return be_response<double>(h, x, y, n, kbb, sigx, sigy, out4);
}

int octopic_gaussian_beambeam_kick_response_f32(octopic_handle *h, const float *x, const float *y, int64_t n, float kbb,
                                                float sigx, float sigy, float *out4) {
    return be_response<float>(h, x, y, n, kbb, sigx, sigy, out4);
}

int octopic_faddeeva_w(octopic_handle *h, const double *zr, const double *zi, int64_t n, double *wr, double *wi) {
    return faddeeva<double>(h, zr, zi, n, wr, wi);
}

int octopic_faddeeva_w_f32(octopic_handle *h, const float *zr, const float *zi, int64_t n, float *wr, float *wi) {
    return faddeeva<float>(h, zr, zi, n, wr, wi);
}

int octopic_track_lorentz_boost(octopic_handle *h, double angle, int32_t reverse, octopic_beam *b) {
    if (!h || !b) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (!b->on_device) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "octopic_track_lorentz_boost needs device beam pointers");
    if (!std::isfinite(angle)) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "Lorentz boost angle must be finite");
    double *c[6] = {b->x, b->px, b->y, b->py, b->z, b->pz};
    hipSetDevice(h->device);
    if (launch_lorentz_boost(angle, reverse ? 1 : 0, c, b->n, h->stream) != hipSuccess)
        return set_err(h, OCTOPIC_ERR_DEVICE, "Lorentz boost launch failed");
    return OCTOPIC_OK;
}

int octopic_track_crab_cavity(octopic_handle *h, double frequency, int32_t nharmonics, const double *strength_x,
                              const double *strength_y, const double *phase, octopic_beam *b) {
    if (!h || !b || (nharmonics > 0 && (!strength_x || !strength_y || !phase))) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (!b->on_device) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "octopic_track_crab_cavity needs device beam pointers");
    // _validate_crab_frequency (crab_cavity.jl:113-117)
    if (!(std::isfinite(frequency) && frequency != 0.0))
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "thin crab-cavity frequency must be finite and nonzero");
    if (nharmonics < 0 || nharmonics > kMaxCrabHarmonics)
        return set_err(h, OCTOPIC_ERR_UNSUPPORTED, "thin crab cavities with more than 16 harmonics are not supported");
    const double kcc = 2.0 * 3.141592653589793 * frequency / 299792458.0;   // 2 pi f / CLIGHT (Constants.jl:8)
    double *c[6] = {b->x, b->px, b->y, b->py, b->z, b->pz};
    hipSetDevice(h->device);
    if (launch_crab_cavity(kcc, nharmonics, strength_x, strength_y, phase, c, b->n, h->stream) != hipSuccess)
        return set_err(h, OCTOPIC_ERR_DEVICE, "crab cavity launch failed");
    return OCTOPIC_OK;
}

int octopic_generate_gaussian_beam(octopic_handle *h, uint64_t seed, uint64_t rng_id, const double *beta3,
                                   const double *alpha3, const double *sigma3, double cutoff, octopic_beam *b) {
    if (!h || !b || !beta3 || !alpha3 || !sigma3) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (!b->on_device) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "octopic_generate_gaussian_beam needs device pointers");
    if (b->n <= 0) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "random Beam construction requires a positive particle count");
    if (b->dtype != OCTOPIC_DTYPE_F64) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "the beam generator writes FP64 beams");
    double *c[6] = {b->x, b->px, b->y, b->py, b->z, b->pz};
    hipSetDevice(h->device);
    if (launch_generate(c, b->n, seed, rng_id, cutoff, h->stream) != hipSuccess)
        return set_err(h, OCTOPIC_ERR_DEVICE, "beam generation failed");
    if (launch_scale_beam(c, b->n, beta3, alpha3, sigma3, h->stream) != hipSuccess)
        return set_err(h, OCTOPIC_ERR_DEVICE, "beam scaling failed");
    return OCTOPIC_OK;
}

int octopic_generate_gaussian_beam_noalloc(octopic_handle *h, uint64_t seed, uint64_t rng_id, const double *beta3,
                                           const double *alpha3, const double *sigma3, double cutoff, octopic_beam *b) {
    if (!h || !b || !beta3 || !alpha3 || !sigma3) return OCTOPIC_ERR_INVALID_ARGUMENT;
    if (!b->on_device)
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "octopic_generate_gaussian_beam_noalloc needs device pointers");
    if (b->n <= 0) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "random Beam construction requires a positive particle count");
    if (b->dtype != OCTOPIC_DTYPE_F64) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "the beam generator writes FP64 beams");
    if (generate_workspace_doubles(b->n) > generate_scratch_doubles())
        return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT,
                       "octopic_generate_gaussian_beam_noalloc: the workspace for n = " + std::to_string(b->n) + " needs " +
                           std::to_string(generate_workspace_doubles(b->n)) + " doubles; the static workspace holds " +
                           std::to_string(generate_scratch_doubles()));
    double *c[6] = {b->x, b->px, b->y, b->py, b->z, b->pz};
    if (hipSetDevice(h->device) != hipSuccess) return set_err(h, OCTOPIC_ERR_DEVICE, "hipSetDevice failed");
    const hipError_t e = launch_generate_noalloc(c, b->n, seed, rng_id, cutoff, h->device, h->stream);
    if (e == hipErrorNotInitialized)
        return set_err(h, OCTOPIC_ERR_DEVICE, "the static generator workspace of this device was not set up");
    if (e != hipSuccess) return set_err(h, OCTOPIC_ERR_DEVICE, "beam generation failed");
    if (launch_scale_beam(c, b->n, beta3, alpha3, sigma3, h->stream) != hipSuccess)
        return set_err(h, OCTOPIC_ERR_DEVICE, "beam scaling failed");
    return OCTOPIC_OK;
}

int octopic_debug_weights(octopic_handle *h, int32_t method, const double *u, int64_t n, int32_t npts, int32_t *base,
                          double *w) {
    if (!h || !u || !base || !w || n <= 0) return OCTOPIC_ERR_INVALID_ARGUMENT;
    hipSetDevice(h->device);
    double *du = nullptr, *dw = nullptr;
    int *db = nullptr;
    int rc = OCTOPIC_OK;
    if (hipMalloc((void **)&du, sizeof(double) * n) != hipSuccess || hipMalloc((void **)&dw, sizeof(double) * 3 * n) != hipSuccess ||
        hipMalloc((void **)&db, sizeof(int) * n) != hipSuccess) {
        rc = set_err(h, OCTOPIC_ERR_OOM, "debug buffers");
    } else {
        hipMemcpy(du, u, sizeof(double) * n, hipMemcpyHostToDevice);
        launch_debug_weights(method, du, n, npts, db, dw, h->stream);
        hipStreamSynchronize(h->stream);
        hipMemcpy(base, db, sizeof(int) * n, hipMemcpyDeviceToHost);
        hipMemcpy(w, dw, sizeof(double) * 3 * n, hipMemcpyDeviceToHost);
    }
    if (du) hipFree(du);
    if (dw) hipFree(dw);
    if (db) hipFree(db);
    return rc;
}

int octopic_set_timing(octopic_handle *h, int32_t enabled) {
    if (!h) return OCTOPIC_ERR_INVALID_ARGUMENT;
    h->timing = enabled != 0;
    h->fam_mask = ~0u;
    for (int k = 0; k < kNFamilies; ++k) { h->fam_ms[k] = 0.0; h->fam_n[k] = 0; }
    return OCTOPIC_OK;
}

int octopic_set_timing_families(octopic_handle *h, const char *families) {
    if (!h || !families) return OCTOPIC_ERR_INVALID_ARGUMENT;
    unsigned mask = 0u;
    std::string list(families);
    size_t pos = 0;
    while (pos <= list.size()) {
        size_t end = list.find(',', pos);
        if (end == std::string::npos) end = list.size();
        const std::string name = list.substr(pos, end - pos);
        if (!name.empty()) {
            int k = 0;
            while (k < kNFamilies && name != kFamilies[k]) ++k;
            if (k == kNFamilies) return set_err(h, OCTOPIC_ERR_INVALID_ARGUMENT, "unknown kernel family '" + name + "'");
            mask |= 1u << k;
        }
        pos = end + 1;
    }
    h->timing = mask != 0u;
    h->fam_mask = mask;
    for (int k = 0; k < kNFamilies; ++k) { h->fam_ms[k] = 0.0; h->fam_n[k] = 0; }
    return OCTOPIC_OK;
}

int octopic_kernel_times(const octopic_handle *h, double *ms_total, int64_t *launches, int64_t cap, int64_t *n, char *names,
                         int64_t names_cap) {
    if (!h || !n) return OCTOPIC_ERR_INVALID_ARGUMENT;
    *n = kNFamilies;
    int64_t off = 0;
    for (int k = 0; k < kNFamilies && k < cap; ++k) {
        if (ms_total) ms_total[k] = h->fam_ms[k];
        if (launches) launches[k] = h->fam_n[k];
        if (names) {
            size_t len = std::strlen(kFamilies[k]) + 1;
            if (off + (int64_t)len <= names_cap) { std::memcpy(names + off, kFamilies[k], len); off += (int64_t)len; }
        }
    }
    return OCTOPIC_OK;
}

}  // extern "C"
