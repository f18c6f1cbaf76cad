// launch.h -- host-side launchers of the octopic kernels (one translation unit per stage family).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "fft.h"

namespace octopic {

// Slice-contiguous particle states of both beams: st.c[beam][coord], coord = x, px, y, py, z, pz.
struct States {
    double *c[2][6];
};

struct SlicingArgs {
    const double *z, *x, *y;         // input beam coordinates (x, y only for the transverse box)
    int64_t n;
    int method, ns, resolution, centroid;
    const double *q;                 // quantiles (normal_quantile) or positions (specified)
    SliceTable *tab;
    double *lanes, *mm_partial, *centroid_lanes;
    unsigned long long *bin_counts;  // equal_area bins (zero on entry, reset on exit)
    uint16_t *sid;
    int *hist;
    double *const *beam;             // device array of 6 device pointers (input beam)
    double *const *st;               // device array of 6 device pointers (slice-contiguous state)
    const double *st_z;
    int64_t *perm;
    unsigned long long *box;         // 4 bound slots (x min/max, y min/max), zero on entry; null: not folded
    ErrorInfo *err;
    // :equal_count rank sort: n keys and n indices each way (the sorted keys end in rank_sorted, the indices
    // in perm); hist then holds 256 digit counts per block
    unsigned long long *rank_keys, *rank_sorted;
    int64_t *rank_vals;
    int32_t *inv = nullptr;          // when set: inv[i] = the slice-order position of particle i
};
hipError_t launch_slicing(const SlicingArgs &a, hipStream_t s);
// blocks of the z lane fold (64 lanes each); mm_partial holds 2 values per block
constexpr int kZFoldBlocks = kFoldLanes / 64;

constexpr int kTileChunk = 4096;     // particles per block visit; block c of a slice takes visits c, c+G, ...
// particles per block and slice the tile order aims for: one visit per block (4096 measured fastest of
// 2048..32768; longer spans give longer tile runs but too few blocks)
constexpr int kTileBlockSpan = 4096;

// Transverse tile order inside each slice (after slicing, with the slice table on the device).
struct TileSortArgs {
    const SliceTable *tab;
    int ns, T;                       // slices, tiles per axis
    const unsigned long long *box;   // the beam's transverse bound slots
    const double *src[6];
    const int64_t *perm_src;
    double *dst[6];
    int64_t *perm_dst;
    int32_t *inv;                    // inv[original index] = tile-ordered position
    int *thist;                      // [ns][T*T][G]
};
hipError_t launch_tile_sort(const TileSortArgs &a, int G, hipStream_t s);
int64_t slicing_hist_entries(int64_t n, int ns);
// The slice-ordered (or tile-ordered) states back into the beam through the inverse permutation; with_z:
// z written back too (solvers whose drifts move it)
hipError_t launch_store_gather(double *const *beam, int64_t n, double *const *st, const int32_t *inv, const ErrorInfo *err,
                               hipStream_t s, int with_z = 0);

struct BatchPair { int32_t p, i, j, reserved; };

// A cached slice-pair mesh pair.  `epoch` is the collide (numbered from 1) that built the entry, so that
// a collide that fails can invalidate every entry it built (their Green spectra may be unbuilt).
struct CacheGeo { int32_t valid, epoch; Grid sg, fg; };

// Static options the geometry kernel needs.
struct GeomOpts {
    int nx, ny, lum_nx, lum_ny;
    int green_type, green_cache, quadratic, interaction_grid, compute_lum, lkick;
    int ns1, ns2;
    int npl;                     // planes per interaction (2 or 3)
    int temp_slot0;              // first per-batch temporary Green slot (green_cache = none)
    int epoch;                   // this collide's number (CacheGeo::epoch of the entries it builds)
    int gauss;                   // Gaussian-subtracted solver: fold the source moments, widen the source box
    int mom_chunks;              // moment chunks per interaction (gauss)
    double margin;               // margin_sigma (gauss)
    double min_x, min_y, quantize, min_ratio, growth;
    double kbb1, kbb2, klum;
};

constexpr int kBoundSlots = 12;   // order-key slots per (pair, side) segment

struct BoundsArgs {
    const BatchPair *pairs;
    int npairs;
    const SliceTable *tab1, *tab2;
    States st;
    int nchunks;
    int srcfld, lum;             // which boxes to fold: source/field (slots 0-7), luminosity (8-11)
    unsigned long long *keys;    // [2*npairs][kBoundSlots], zero on entry, zeroed again by the geometry stage
    double *chunkbox;            // [2*npairs][nchunks][8]: per chunk, source box (sL, sR) and field box (srcfld only)
    const ErrorInfo *err;
};

// Particle ranges of a batch: slice members [n*c/nchunks, n*(c+1)/nchunks) for chunk c, the same in
// the bounds, deposit and kick launches, so that a chunk's box describes the block that takes it.
constexpr int kChunkSpan = 2048;
constexpr int kMaxChunks = 128;
hipError_t launch_bounds(const BoundsArgs &a, hipStream_t s);

struct GeometryArgs {
    const BatchPair *pairs;
    int npairs;
    const SliceTable *tab1, *tab2;
    unsigned long long *keys;    // bounds slots of this batch, consumed (zeroed)
    GeomOpts o;
    CacheGeo *cache;             // [2*ns1*ns2]
    Inter *inters;               // [2*npairs]
    Plane *planes;               // field planes [2*npairs*npl], then luminosity planes [2*npairs]
    GreenJob *jobs;              // [2*npairs]
    double *lumgeo;              // [npairs][3] = hxi, hyi, klum
    octopic_interaction_record *records;
    unsigned long long *counters;   // hits, misses, rebuilds, builds
    const double *mom_part;         // gauss: [2*npairs][mom_chunks][10] source moment partials
    States st;                      // gauss: the states (first member of each source slice)
    ErrorInfo *err;
};
hipError_t launch_geometry(const GeometryArgs &a, hipStream_t s);
// Invalidate the cache entries built by collide `epoch` (after that collide failed).
hipError_t launch_cache_invalidate(CacheGeo *cache, int nkeys, int epoch, hipStream_t s);

struct DepositArgs {
    const Plane *planes;
    int nplanes;
    States st;
    int method;
    int ncells;                  // n1*n2 of these planes
    int chunks;                  // the batch's chunking of each slice
    int group;                   // chunks per block
    int fp32;                    // weight product in FP32
    double scale;                // fixed-point scale 2^k
    unsigned long long *charge;  // [plane][charge_stride]
    int64_t charge_stride;
    unsigned long long *dropped; // dropped-particle counter
    const double *chunkbox;      // chunk boxes of the batch's bounds launch, or null (whole mesh windows)
    const ErrorInfo *err;
    int lds_cells;               // set by the launcher: LDS window capacity in cells
    int *queue;                  // work-queue counter of this launch (zero on entry)
    const int32_t *order;        // launch order of the planes (position -> plane)
    int multi;                   // > 1: items take this many consecutive planes of one interaction (same
                                 // mesh and particles, 2 or 3), which the order lists together
    int lum_plane0;              // >= 0 (with multi > 1): interaction d's luminosity plane lum_plane0 + d is
                                 // deposited by the same items (same particles at sM, same stencil method)
};
constexpr int kDepositMaxLdsCells = 20480 - 2;   // 160 KB per block, less 16 bytes of the block's own LDS
hipError_t launch_deposit(DepositArgs a, hipStream_t s);

// Bassetti-Erskine kick and the Faddeeva function (k_be.hip)
struct WeidemanCoeffs {
    double c[32];   // polynomial coefficients, highest power first
    double L;       // sqrt(32 / sqrt 2)
};
struct WeidemanCoeffsF {   // the same, rounded to FP32 (T(c), SpecialMath.jl:98-100)
    float c[32];
    float L;
};
const WeidemanCoeffs &weideman_coeffs();
const WeidemanCoeffsF &weideman_coeffs_f32();
// T = double or float: the FP64 evaluation or the FP32 one with the reference's Float32 constants
template <typename T>
hipError_t launch_be_kick(const T *x, const T *y, int64_t n, T sigx, T sigy, T *kx, T *ky, hipStream_t s);
// the kick and kbb times the Hessian diagonal, (kx, ky, hxx, hyy) per point (strong_beam.jl:1119-1128)
template <typename T>
hipError_t launch_be_response(const T *x, const T *y, int64_t n, T kbb, T sigx, T sigy, T *out4, hipStream_t s);
template <typename T>
hipError_t launch_faddeeva(const T *zr, const T *zi, int64_t n, T *wr, T *wi, hipStream_t s);

// Gaussian-subtracted solver (gaussian_pic.jl): source moments, residual planes, per-particle add-back.
struct GpicMomentArgs {
    const BatchPair *pairs;
    int npairs;
    const SliceTable *tab1, *tab2;
    States st;
    int chunks;                  // chunks per source slice
    double *mom_part;            // [2*npairs][chunks][10], interaction d = 2k + dir - 1
    const ErrorInfo *err;
    double *shift = nullptr;     // when set: [2*npairs][4], the sums' shift (the slice's first x, px, y, py)
};
hipError_t launch_gpic_moments(const GpicMomentArgs &a, hipStream_t s);
// Chunks of the source moment sums for a batch whose largest slice has `bmax` members: about 1024
// members per chunk (a few loads per thread, so the launch is not a chain of dependent reads), 16 to
// kMaxMomentChunks.
constexpr int kMaxMomentChunks = 256;
inline int moment_chunks(int64_t bmax) {
    const int64_t c = (bmax + 1023) / 1024;
    return (int)(c < 16 ? 16 : (c > kMaxMomentChunks ? kMaxMomentChunks : c));
}
struct GpicResidualArgs {
    const Inter *inters;
    const Plane *planes;
    int nplanes;                 // field planes of the batch (L, R per interaction)
    unsigned long long *charge;  // fixed-point planes, the residual replaces the deposit
    int64_t charge_stride;
    double scale, inv_scale;
    int method, nx, ny, neutralize;
    const ErrorInfo *err;
};
hipError_t launch_gpic_residual(const GpicResidualArgs &a, hipStream_t s);
struct GpicAddbackArgs {
    const Inter *inters;
    int ninters;
    States st;
    double *ab;                  // [N1 + N2][4]: dpx, dpy, pz covariance term, pz centroid term
    int64_t base[2];             // first row of each beam's particles in ab
    int lkick;
    const ErrorInfo *err;
};
hipError_t launch_gpic_addback(const GpicAddbackArgs &a, int64_t max_fld, hipStream_t s);

// The soft-Gaussian solver (GaussianPoissonSolver, gaussian.jl): per batch, after launch_gpic_moments over
// the batch's source slices, every interaction's field slice through the map; kSoftBlocks blocks per
// interaction, whose luminosity partials a collide folds per pair at its end.
constexpr int kSoftBlocks = 128;
struct SoftKickArgs {
    const BatchPair *pairs;
    int npairs;
    const SliceTable *tab1, *tab2;
    States st;
    int mom_chunks;              // chunks of the moment partials
    const double *mom_part;      // [2*npairs][mom_chunks][10], interaction d = 2k + dir - 1
    const double *shift;         // [2*npairs][4], the shift of those sums (read here: a launch kicks in place
                                 // the slices other interactions of it take as sources)
    double kbb1, kbb2;           // kicks on beam 1 and beam 2 (times the source slice's weight)
    double klum1, klum2;         // luminosity scales sampling beam 1 and beam 2 (times the weight)
    double min_sigma;
    int ignore1, ignore2;        // ignore_centroid of each beam's slices
    int drift, lkick, sample1, compute_lum;
    double *lum_block;           // [npairs of the collide][kSoftBlocks]
    double *lum_scale;           // [npairs of the collide]
    ErrorInfo *err;
};
// fp32: the Bassetti-Erskine evaluation in FP32 (GaussianPoissonSolver{Float32}); FP64 otherwise
hipError_t launch_soft_kick(const SoftKickArgs &a, bool fp32, hipStream_t s);
hipError_t launch_soft_lum_fold(const double *lum_block, const double *lum_scale, int npairs, double *lum_parts, hipStream_t s);

struct KickArgs {
    const Inter *inters;
    int ninters;
    States st;
    const void *table;           // double (FP64 mode) or float (fp32) kick table
    int fp32;
    int nx, ny, npl, method, mode;   // mode: 0 linear, 1 quadratic, 2 node
    int lkick;
    int chunks;
    int subdiv;                  // kick blocks per chunk (each takes an equal share of the chunk's members)
    int pass_windows;            // stage a window per pass when the chunk's window does not fit
    unsigned long long *dropped;
    const double *chunkbox;      // chunk boxes of the batch's bounds launch, or null (no LDS windows)
    const ErrorInfo *err;
    const int32_t *order;        // launch order of the interactions (position -> interaction)
    const double *ab;            // Gaussian-subtracted add-back [N1 + N2][4] (mode 0 only), or null
    int64_t ab_base[2];
};
hipError_t launch_kick(const KickArgs &a, hipStream_t s);

struct LumArgs {
    const BatchPair *pairs;
    int npairs;
    unsigned long long *charge;  // luminosity planes, pair k at planes 2k, 2k+1
    int64_t charge_stride;
    int ncells;
    double inv_scale;
    const double *lumgeo;
    double *lum_parts;           // indexed by pair position p
    double *lum_partial;         // [npairs][kLumBlocks]
    const ErrorInfo *err;
};
constexpr int kLumBlocks = 16;   // blocks per pair of the luminosity overlap sum (in the charge row launch)

// FFT stages (k_fft.hip)
// full = 1: a block per (job, row block) of every job, for launches where most jobs build; 0: a small
// persistent grid over the jobs that build.  Results do not depend on it.
// spec32: the FP32 mode's float copy of the slots (written by the column pass), or nullptr.
hipError_t launch_green_build(const GreenJob *jobs, int njobs, double2 *spec, float2 *spec32, int64_t slot_stride, int nslots,
                              int nx, int ny, int integrated, const FFTPlanDev &px, const FFTPlanDev &py, const ErrorInfo *err,
                              hipStream_t s, int full);
// The charge-path transform passes.  group: transforms per workgroup, 2, 4 or 8 (halved while its LDS does
// not fit); results do not depend on it.
template <class C>   // C = double2 (FP64 mode) or float2 (FP32 mode)
hipError_t launch_rows_fwd(const Plane *planes, int nplanes, unsigned long long *charge, int64_t charge_stride, C *qh,
                           int64_t qh_stride, int nx, int ny, double inv_scale, const FFTPlanDev &px, const ErrorInfo *err,
                           const LumArgs &lum, hipStream_t s, int group);   // lum.npairs = 0: no luminosity sums
template <class C>
// lum.npairs > 0: the launch also folds the luminosity partials of the row launch before it
hipError_t launch_cols_conv(const Plane *planes, int nplanes, C *qh, int64_t qh_stride, const C *gspec, int64_t slot_stride,
                            int nx, int ny, const FFTPlanDev &py, const ErrorInfo *err, const LumArgs &lum, hipStream_t s,
                            int group);
template <class C>
hipError_t launch_rows_inv(const Plane *planes, int nplanes, const C *qh, int64_t qh_stride, real_of<C> *phi, int nx,
                           int ny, const FFTPlanDev &px, const ErrorInfo *err, hipStream_t s, int group);
template <class R>
hipError_t launch_field(const Inter *inters, int ninters, const Plane *planes, const R *phi, R *table, int nx, int ny,
                        int fourth, int npl, const ErrorInfo *err, hipStream_t s);

// Node-mode mesh preparation (k_node.hip)
struct NodeArgs {
    const SliceTable *tab_src, *tab_fld;
    States st;
    int src_beam, fld_beam;
    int dir;                     // 1 or 2
    int nchunks;
    unsigned long long *skeys;   // source boxes [ns_src][nb][4] as bound slots, zero on entry
    unsigned long long *fkeys;   // field boxes [ns_src][ns_fld][4] as bound slots, zero on entry
    GeomOpts o;
    int slot0;                   // first Green slot of this direction's node meshes
    CacheGeo *nodes;             // [ns_src][nb] node meshes of this direction
    GreenJob *jobs;              // [ns_src][nb]
    unsigned long long *counters;
    ErrorInfo *err;
};
hipError_t launch_node_prepare(const NodeArgs &a, int ns_src, int ns_fld, hipStream_t s);

struct NodeGeometryArgs {
    const BatchPair *pairs;
    int npairs;
    const SliceTable *tab1, *tab2;
    unsigned long long *keys;    // bounds slots of this batch (luminosity boxes), consumed (zeroed)
    GeomOpts o;
    const CacheGeo *nodes1, *nodes2;   // dir-1 meshes [ns1][ns2+1], dir-2 meshes [ns2][ns1+1]
    int slot1, slot2;                  // first Green slot of each direction
    Inter *inters;
    Plane *planes;
    double *lumgeo;
    octopic_interaction_record *records;
    int32_t *ran;                      // per pair position: 0 when a node mesh is missing
    ErrorInfo *err;
};
hipError_t launch_node_geometry(const NodeGeometryArgs &a, hipStream_t s);

// misc (k_misc.hip)
hipError_t launch_linear6d(const double *m36, double *const c[6], int64_t n, hipStream_t s);
hipError_t launch_lorentz_boost(double angle, int reverse, double *const c[6], int64_t n, hipStream_t s);
constexpr int kMaxCrabHarmonics = 16;
hipError_t launch_crab_cavity(double kcc, int nh, const double *sx, const double *sy, const double *phase, double *const c[6],
                              int64_t n, hipStream_t s);
hipError_t launch_generate(double *const c[6], int64_t n, uint64_t seed, uint64_t rng_id, double cutoff, hipStream_t s);
// The same with a caller-provided workspace of generate_workspace_doubles(n) doubles (no allocation).
int64_t generate_workspace_doubles(int64_t n);
hipError_t launch_generate_ws(double *const c[6], int64_t n, uint64_t seed, uint64_t rng_id, double cutoff, double *ws,
                              hipStream_t s);
// The same on the library's static per-device workspace (no allocation): generate_scratch_init(device)
// once per device with that device current (octopic_create does it), then launch_generate_noalloc.
int64_t generate_scratch_doubles();
hipError_t generate_scratch_init(int device);
hipError_t launch_generate_noalloc(double *const c[6], int64_t n, uint64_t seed, uint64_t rng_id, double cutoff, int device,
                                   hipStream_t s);
hipError_t launch_scale_beam(double *const c[6], int64_t n, const double *beta, const double *alpha, const double *sigma,
                             hipStream_t s);
hipError_t launch_widen(const float *src, double *dst, int64_t n, hipStream_t s);
hipError_t launch_narrow(const double *src, float *dst, int64_t n, hipStream_t s);
hipError_t launch_debug_weights(int method, const double *u, int64_t n, int npts, int *base, double *w, hipStream_t s);

}  // namespace octopic
