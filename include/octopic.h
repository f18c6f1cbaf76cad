/*
 * octopic.h -- C ABI of the MI355X strong-strong PIC collision library.
 *
 * This library is a drop-in for one code path of Octopus.jl: the collision
 *     collide!(solver::PICPoissonSolver, beam1::Beam, beam2::Beam, backend[, ctx]) -> Float64
 * (src/tasks/strongstrong/pic_cpu.jl:1-32, pic_cuda.jl:94-143).  Julia host code binds it with
 * `ccall` under a new backend tag; the Python package octopic (octopus.jl_amd/octopic) binds it with
 * ctypes.  Every signature takes plain pointers, sizes and scalars.
 *
 * Conventions
 *   - Beams are structure-of-arrays FP64 (x, px, y, py, z, pz), as Phase6DRep (src/beam/Beam.jl:45-56).
 *     Coordinates are modified in place, as collide! does.
 *   - Indices are 0-based in this ABI; the reference is 1-based.
 *   - Every entry point returns an OCTOPIC_* status.  A nonzero status leaves a message that
 *     octopic_last_error() returns.  The status codes mirror the reference's ArgumentError throws.
 *   - A handle owns its device workspaces and the persistent slice-pair Green cache, i.e. what the
 *     reference keeps in StrongStrongTask.runtime_cache (pic_cpu.jl:23-32, 488-515).  A handle is not
 *     thread-safe: one caller per handle, as one task per collision label.
 */
#ifndef OCTOPIC_H
#define OCTOPIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCTOPIC_ABI_VERSION 1

/* Status codes (reference: the ArgumentError sites in pic_cpu.jl:517-576, 730-777, interface.jl:541-599). */
enum {
    OCTOPIC_OK = 0,
    OCTOPIC_ERR_INVALID_OPTION = 1,       /* an option value outside its schema (interface.jl:1410-1476) */
    OCTOPIC_ERR_UNSUPPORTED = 2,          /* a combination the reference rejects, e.g. :node + :quadratic (pic_cpu.jl:568-574) */
    OCTOPIC_ERR_NONFINITE = 3,            /* non-finite coordinates (interface.jl:541-599) */
    OCTOPIC_ERR_DEGENERATE_EXTENT = 4,    /* zero transverse extent with no floor (pic_cpu.jl:1646-1672) */
    OCTOPIC_ERR_OOM = 5,                  /* device allocation failed */
    OCTOPIC_ERR_DEVICE = 6,               /* a HIP runtime call failed */
    OCTOPIC_ERR_INVALID_ARGUMENT = 7      /* null pointer, empty beam, size mismatch */
};

/* Enumerations of the PICPoissonSolver symbol options (interface.jl:1324-1356). */
enum { OCTOPIC_CIC = 0, OCTOPIC_TSC = 1 };
enum { OCTOPIC_GREEN_INTEGRATED = 0, OCTOPIC_GREEN_STANDARD = 1, OCTOPIC_GREEN_LATTICE = 2 };
enum { OCTOPIC_CACHE_NONE = 0, OCTOPIC_CACHE_SLICE_PAIR = 1 };
enum { OCTOPIC_INTERP_LINEAR = 0, OCTOPIC_INTERP_QUADRATIC = 1 };
enum { OCTOPIC_GRID_SLICE_PAIR = 0, OCTOPIC_GRID_SOURCE_SLICE = 1, OCTOPIC_GRID_NODE = 2 };
enum { OCTOPIC_EXTENT_EXTREMA = 0, OCTOPIC_EXTENT_SIGMA = 1 };
enum { OCTOPIC_BATCH_SEQUENTIAL = 0, OCTOPIC_BATCH_WAVEFRONT = 1 };
/* LongitudinalSlicing methods (interface.jl:800-835; slicing.jl:51-75). */
enum {
    OCTOPIC_SLICE_EQUAL_AREA = 0,
    OCTOPIC_SLICE_EQUAL_COUNT = 1,
    OCTOPIC_SLICE_EQUAL_WIDTH = 2,
    OCTOPIC_SLICE_NORMAL_QUANTILE = 3,
    OCTOPIC_SLICE_SPECIFIED = 4
};
enum { OCTOPIC_CENTER_CENTROID = 0, OCTOPIC_CENTER_MIDPOINT = 1 };
enum { OCTOPIC_PRECISION_FP64 = 0, OCTOPIC_PRECISION_FP32 = 1 };
/* The solver a handle runs: PICPoissonSolver, GaussianPICPoissonSolver (gaussian_pic.jl:87-131), the
 * Gaussian-subtracted hybrid that composes it, or GaussianPoissonSolver (interface.jl:876-1010,
 * gaussian.jl), the sliced soft-Gaussian solver (no mesh: every slice pair kicks with the exact
 * Bassetti-Erskine field of the other slice's Gaussian moments). */
enum { OCTOPIC_SOLVER_PIC = 0, OCTOPIC_SOLVER_GAUSSIAN_PIC = 1, OCTOPIC_SOLVER_SOFT_GAUSSIAN = 2 };
/* The soft-Gaussian solver's virtual drift to and from the collision point (strong_beam.jl:11-37,
 * 497-617): HirataParaxialDrift (the default), ChromaticDrift, ExactHamiltonianDrift, and the two
 * non-symplectic UnsafeVirtualDrift models. */
enum {
    OCTOPIC_DRIFT_HIRATA = 0,
    OCTOPIC_DRIFT_CHROMATIC = 1,
    OCTOPIC_DRIFT_EXACT = 2,
    OCTOPIC_DRIFT_UNSAFE_PARAXIAL_FROZEN_LONGITUDINAL = 3,
    OCTOPIC_DRIFT_UNSAFE_CHROMATIC_FROZEN_ENERGY = 4
};

/* LongitudinalSlicing (interface.jl:800-835).  `positions` holds nslices-1 internal boundaries in
 * units of the bunch rms (method SPECIFIED only); it may be NULL otherwise. */
typedef struct octopic_slicing {
    int32_t nslices;
    int32_t method;
    int32_t resolution;
    int32_t center_position;
    int32_t npositions;
    int32_t reserved;
    const double *positions;
} octopic_slicing;

/* PICPoissonSolver{Float64} fields (interface.jl:1324-1356).  `kbb1`, `kbb2`, `luminosity_scale`
 * are NaN for the reference's `nothing` (derive from the beam parameters).  `lum_nx`/`lum_ny` = 0
 * and `lum_deposit` = -1 inherit `grid` / `deposit` (interface.jl:1513-1517).  The CUDA route flags
 * (batch_mode, cuda_*) are accepted and recorded; they select nothing on this backend. */
typedef struct octopic_options {
    int32_t nx, ny;
    int32_t deposit;               /* OCTOPIC_CIC | OCTOPIC_TSC */
    int32_t green_type;            /* OCTOPIC_GREEN_* */
    int32_t green_cache;           /* OCTOPIC_CACHE_* */
    int32_t field_derivative;      /* 2 | 4 */
    int32_t slice_interpolation;   /* OCTOPIC_INTERP_* */
    int32_t interaction_grid;      /* OCTOPIC_GRID_* */
    int32_t grid_extent;           /* OCTOPIC_EXTENT_* */
    int32_t longitudinal_kick;     /* 0 | 1 */
    int32_t lum_nx, lum_ny;        /* 0 = inherit grid */
    int32_t lum_deposit;           /* -1 = inherit deposit */
    int32_t batch_mode;            /* OCTOPIC_BATCH_*; recorded only */
    int32_t cuda_flags;            /* bit0 cuda_async, bit1 cuda_batch_fft, bit2 cuda_wavefront_fft,
                                      bit3 cuda_indexed_wavefront; recorded only */
    int32_t precision;             /* OCTOPIC_PRECISION_*: FP64 (the reference's Float64 path) or the FP32
                                      form: deposit weights, FFT convolution, field and kick gather in FP32
                                      on FP64 beams; Green evaluation, slice moments, luminosity in FP64.
                                      The FP32 form of the PIC solvers has no counterpart in the reference.
                                      For OCTOPIC_SOLVER_SOFT_GAUSSIAN, FP32 is GaussianPoissonSolver{Float32}'s
                                      Bassetti-Erskine evaluation (field, Hessian, density) in FP32 with the
                                      Float32 constants; drifts, slice moments, coordinates and luminosity
                                      sums stay FP64. */
    double grid_extent_sigma;
    double min_extent_x, min_extent_y;
    double grid_quantize;
    double green_min_ratio;
    double green_growth;
    double kbb1, kbb2, luminosity_scale;
    octopic_slicing slicing1, slicing2;
    /* GaussianPICPoissonSolver (gaussian_pic.jl:87-131), with every field above its embedded
     * PICPoissonSolver: solver = OCTOPIC_SOLVER_GAUSSIAN_PIC subtracts each source slice's drifted
     * Gaussian (its own moments) from the deposits before the solve and adds the exact Bassetti-Erskine
     * field of that Gaussian back per field particle.  margin_sigma >= 0 widens the source mesh to
     * +-margin_sigma of the Gaussian; neutralize scales the subtracted Gaussian to the deposited total;
     * coupling_tol is the correlation threshold of the rotated subtraction (only +Inf, the reference's
     * default, the separable subtraction, is built here).  Ignored for OCTOPIC_SOLVER_PIC. */
    int32_t solver;                /* OCTOPIC_SOLVER_* */
    int32_t neutralize;            /* 0 | 1 */
    double margin_sigma;
    double coupling_tol;
    /* GaussianPoissonSolver (interface.jl:876-1010), with kbb1/kbb2/luminosity_scale, slicing1/2,
     * longitudinal_kick and batch_mode above (the mesh fields are ignored): min_sigma >= 0 floors each
     * slice's transverse rms; gaussian_when_luminosity 1 | 2 names the beam whose particles sample the
     * other beam's Gaussian for the luminosity; ignore_centroid1/2 zero a beam's slice centroids;
     * include_sigma_xy = 1 (the coupled, rotated kick) is not implemented here (UNSUPPORTED).  Ignored
     * for the other solvers. */
    double min_sigma;
    int32_t virtual_drift;         /* OCTOPIC_DRIFT_* */
    int32_t gaussian_when_luminosity;
    int32_t ignore_centroid1, ignore_centroid2;
    int32_t include_sigma_xy;
    int32_t reserved;
} octopic_options;

/* Element type of a beam's coordinate arrays: Phase6DRep{Float64} or Phase6DRep{Float32}
 * (src/beam/Beam.jl:45-56). */
enum { OCTOPIC_DTYPE_F64 = 0, OCTOPIC_DTYPE_F32 = 1 };

/* One beam: Phase6DRep + BeamParams (src/beam/Beam.jl:29-56).  With on_device = 1 the six pointers
 * are device (HBM) pointers and the collision runs on them in place; with on_device = 0 they are
 * host pointers that are copied in and out around the collision.
 * dtype = OCTOPIC_DTYPE_F32: the six pointers address float arrays (cast to double * here).  Such a
 * beam is widened exactly to FP64 for the collision and its x, px, y, py, pz are rounded back to
 * float (round to nearest) at the end; the arithmetic in between is the handle's precision option.
 * The reference's Phase6DRep{Float32} runs its whole collision in Float32 (pic_cpu.jl:1646-1660);
 * here a Float32 beam gives the Float32 rounding of this library's FP64 (or FP32-form) result. */
typedef struct octopic_beam {
    double *x, *px, *y, *py, *z, *pz;
    int64_t n;
    double charge, mc2, E0, r0, npart;
    int32_t on_device;
    int32_t dtype;                 /* OCTOPIC_DTYPE_F64 | OCTOPIC_DTYPE_F32 */
} octopic_beam;

/* Counters of the last collide and of the handle's lifetime. */
typedef struct octopic_stats {
    int64_t green_hits, green_misses, green_rebuilds;   /* cumulative, _PICSlicePairGreenCache (interface.jl:716-742) */
    int64_t dropped;                                    /* last collide (pic_cpu.jl:394-457) */
    int64_t pairs, batches, widest_batch;               /* last collide's schedule */
    int64_t green_builds;                               /* Green tables built in the last collide */
    int64_t planes;                                     /* field planes solved in the last collide */
    int64_t collides;
    int64_t denoised_planes;                            /* last collide: field planes rebuilt by the denoiser */
    int64_t denoise_fallbacks;                          /* last collide: field planes left to plain PIC */
    int64_t deposit_all_planes;                         /* last collide: 1 if deposit items took all planes of
                                                           an interaction, 0 if one plane each */
} octopic_stats;

/* One directed interaction of the last collide, in execution order (pair position p in collision
 * order, then direction 1, 2).  cache_event: 0 = no cache, 1 = hit, 2 = miss, 3 = rebuild.
 * Grids are (x0, y0, width, height) of the source and field meshes actually used. */
typedef struct octopic_interaction_record {
    int32_t p, i, j, dir;
    int32_t cache_event;
    int32_t reserved;
    double source_grid[4];
    double field_grid[4];
} octopic_interaction_record;

typedef struct octopic_handle octopic_handle;

/* Fill `opt` with the PICPoissonSolver() defaults (interface.jl:1384-1407). */
int octopic_options_default(octopic_options *opt);

/* PICPoissonSolver{Float64}(...) validation + _validate_pic_solver (interface.jl:1384-1507,
 * pic_cpu.jl:517-576) and the workspace/cache set-up of _pic_cpu_workspace_pool! / _pic_green_cache!
 * (pic_cpu.jl:467-515).  `device` selects the GPU.  The handle enqueues on a stream of its own, created
 * with it (blocking: ordered with the null stream, concurrent with other handles), until
 * octopic_set_stream names another (NULL restores its own). */
int octopic_create(const octopic_options *opt, int32_t device, octopic_handle **out);
void octopic_destroy(octopic_handle *h);
const char *octopic_last_error(const octopic_handle *h);
int octopic_set_stream(octopic_handle *h, void *hip_stream);

/* collide!(solver, beam1, beam2, backend, ctx) -> luminosity (pic_cpu.jl:82-275).
 * `turn` is ctx.turn (-1 for ctx = nothing); `compute_luminosity` is the evaluated
 * luminosity_schedule (interface.jl:2596-2620).  *luminosity is NaN when it is 0. */
int octopic_collide(octopic_handle *h, octopic_beam *beam1, octopic_beam *beam2, int64_t turn,
                    int32_t compute_luminosity, double *luminosity);

/* `count` independent collides (bunch-pair instances: distinct handles, beams, turns), each on its own
 * handle's stream from its own host thread, so that small instances fill the GPU together.  Returns the
 * first non-OK status; per-instance statuses go to `status` (may be NULL), luminosities to
 * luminosity[k]; `turns` and `compute_luminosity` may be NULL (-1 and 1).  A handle may appear once.
 * No counterpart in the reference (its instances are separate tasks). */
int octopic_collide_many(octopic_handle *const *handles, octopic_beam *beams1, octopic_beam *beams2,
                         const int64_t *turns, const int32_t *compute_luminosity, double *luminosity,
                         int32_t *status, int32_t count);

int octopic_get_stats(const octopic_handle *h, octopic_stats *out);

/* Per-pair luminosity of the last collide in collision order (the pair sink, interface.jl:389-410).
 * Writes min(cap, npairs) entries; *n receives npairs.  ran[k] = 0 marks a skipped pair. */
int octopic_pair_luminosity(const octopic_handle *h, int32_t *i, int32_t *j, int32_t *ran,
                            double *lum, int64_t cap, int64_t *n);

/* The last collide's longitudinal slices of beam `which` (1 or 2): LongitudinalSlices
 * (slicing.jl:1-6).  boundaries has nslices+1 entries; indices receives the concatenated ascending
 * member lists (n entries), counts the list lengths. */
int octopic_last_slices(const octopic_handle *h, int32_t which, double *boundaries, double *centers,
                        double *weights, int64_t *counts, int64_t *indices, int64_t n);

/* The last collide's directed-interaction records (grid geometry + Green-cache decision). */
int octopic_interaction_records(const octopic_handle *h, octopic_interaction_record *out, int64_t cap,
                                int64_t *n);

/* Linear 6D one-turn map (src/elements/linear6d.jl:110-118): coordinates <- M * coordinates, with M a
 * row-major 6x6 matrix.  Runs on the handle's stream; beam pointers must be device pointers. */
int octopic_track_linear6d(octopic_handle *h, const double *matrix36, octopic_beam *beam);

/* track_particle(::NonSymplectic6DMap, ::LorentzBoost | ::RevLorentzBoost, ...) (lorentz_boost.jl:128-166):
 * Hirata's crossing-angle transformation, `angle` = half the full crossing angle, `reverse` = 1 for
 * the reverse map; device beam, in place on the handle's stream. */
int octopic_track_lorentz_boost(octopic_handle *h, double angle, int32_t reverse, octopic_beam *beam);

/* track_particle(::Symplectic6DMap, ::ThinCrabCavity{N}, ...) (crab_cavity.jl:99-164): frequency in Hz
 * (finite, nonzero), nharmonics <= 16 harmonics of (strength_x, strength_y, phase) host arrays; device
 * beam, in place on the handle's stream. */
int octopic_track_crab_cavity(octopic_handle *h, double frequency, int32_t nharmonics, const double *strength_x,
                              const double *strength_y, const double *phase, octopic_beam *beam);

/* gaussian_beambeam_kick(sigx, sigy, x, y) (src/elements/strong_beam.jl:1139-1146, principal-axis
 * branches :1040-1059): the normalised Bassetti-Erskine kick (kx[i], ky[i]) of a transverse Gaussian
 * of rms sizes (sigx, sigy) at (x[i], y[i]), i < n; callers multiply by the beam-beam strength.  A zero
 * size gives zero kicks.  FP64; device pointers; on the handle's stream. */
int octopic_gaussian_beambeam_kick(octopic_handle *h, const double *x, const double *y, int64_t n, double sigx,
                                   double sigy, double *kx, double *ky);

/* gaussian_beambeam_kick(sigx::Float32, sigy, x, y) (the same, strong_beam.jl:770-783, 809-811 for the
 * Float32 constants: near-round conditioning factor 8, eps = 2^-23, radial series through order 16, and
 * the Weideman coefficients rounded to Float32, SpecialMath.jl:98-100): FP32 arithmetic throughout,
 * evaluated on the scale of max(sigx, sigy) (the kick is homogeneous of degree -1), so that sizes of
 * any magnitude stay within the FP32 range.  Device float pointers. */
int octopic_gaussian_beambeam_kick_f32(octopic_handle *h, const float *x, const float *y, int64_t n, float sigx,
                                       float sigy, float *kx, float *ky);

/* _gaussian_beambeam_kick_response(kbb, sigx, sigy, x, y) (src/elements/strong_beam.jl:1061-1128): per point
 * out4[4i..4i+3] = (Kx, Ky, Hxx, Hyy), the normalised Bassetti-Erskine kick and kbb times the diagonal of
 * the Gaussian potential's Hessian (with the near-round blend's chain term), as the Gaussian-subtracted
 * solver's longitudinal add-back uses them.  Sizes must be positive.  FP64; device pointers; on the
 * handle's stream. */
int octopic_gaussian_beambeam_kick_response(octopic_handle *h, const double *x, const double *y, int64_t n, double kbb,
                                            double sigx, double sigy, double *out4);

/* _gaussian_beambeam_kick_response for Float32 arguments: as above, in FP32 with the Float32 constants,
 * on the scale of max(sigx, sigy) (Hessian homogeneous of degree -2).  Device float pointers. */
int octopic_gaussian_beambeam_kick_response_f32(octopic_handle *h, const float *x, const float *y, int64_t n, float kbb,
                                                float sigx, float sigy, float *out4);

/* faddeeva_w_approx_reim(zr, zi) (src/math/SpecialMath.jl:60-151): w(z) = exp(-z^2) erfc(-iz) by
 * Weideman's 32-term rational approximation, the Laplace continued fraction for large |z| and the
 * reflection w(z) = 2 exp(-z^2) - w(-z) below the real axis.  FP64; device pointers; on the handle's
 * stream. */
int octopic_faddeeva_w(octopic_handle *h, const double *zr, const double *zi, int64_t n, double *wr, double *wi);

/* faddeeva_w_approx_reim(zr::Float32, zi::Float32) (SpecialMath.jl:60-151 in T = Float32): the same
 * algorithm in FP32 arithmetic with T(c) coefficients.  Device float pointers. */
int octopic_faddeeva_w_f32(octopic_handle *h, const float *zr, const float *zi, int64_t n, float *wr, float *wi);

/* Beam(N, ...) synthetic Gaussian construction on the device (Beam.jl:127-181, 545-553): Philox4x32-10
 * counter normals, clamp to +-cutoff, standardise, scale by sigma/beta with alpha.  beta, alpha,
 * sigma are (x, y, z) triples.  Writes the six device arrays of `beam` (n = beam->n). */
int octopic_generate_gaussian_beam(octopic_handle *h, uint64_t seed, uint64_t rng_id,
                                   const double *beta3, const double *alpha3, const double *sigma3,
                                   double cutoff, octopic_beam *beam);

/* The same beam as octopic_generate_gaussian_beam, bit for bit, with no device allocation: the partial
 * sums of the standardisation live in a static per-device workspace of 2050 doubles (16400 bytes, the
 * partition's cap of 2048 blocks plus mean and rms, so every n fits), set up by octopic_create.  No
 * counterpart in the reference: for callers that cannot allocate (HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Calls that overlap on one device share the workspace and
 * race; calls in order on one stream, or on streams the caller orders, are safe. */
int octopic_generate_gaussian_beam_noalloc(octopic_handle *h, uint64_t seed, uint64_t rng_id,
                                           const double *beta3, const double *alpha3, const double *sigma3,
                                           double cutoff, octopic_beam *beam);

/* Device-side CIC/TSC cell index and weights (pic_cpu.jl:2182-2252) for n normalised coordinates u
 * on an axis of `npts` nodes: base (0-based first stencil node) and w (2 or 3 per point).
 * u, base, w are host pointers.  Used by the parity tests of the cell indices. */
int octopic_debug_weights(octopic_handle *h, int32_t method, const double *u, int64_t n, int32_t npts,
                          int32_t *base, double *w);

/* Per-kernel-family device time, from HIP events recorded on the handle's stream around each launch
 * group.  octopic_set_timing(h, 1) times every family, (h, 0) none; octopic_set_timing_families
 * times only the comma-separated families named (e.g. "kick"), which keeps the event overhead out of
 * the other stages.  Either call resets the accumulators.  octopic_kernel_times returns, per family,
 * the total ms and the number of timed launch groups since then; names: NUL-separated family names.
 * Diagnostics of this library (no counterpart in the reference). */
int octopic_set_timing(octopic_handle *h, int32_t enabled);
int octopic_set_timing_families(octopic_handle *h, const char *families);
int octopic_kernel_times(const octopic_handle *h, double *ms_total, int64_t *launches, int64_t cap,
                         int64_t *n, char *names, int64_t names_cap);

/* Launch shapes (no counterpart in the reference; only the speed changes, and the results of the plain
 * PIC solver are bit-identical for every value; "transverse_tiles" sets the order in which the Gaussian-
 * subtracted solver and the denoiser sum their slice moments, so their results follow it in the last
 * bits): "transverse_tiles" (0 or a power of two <= 128: in-slice tile order, default 16), "kick_target_blocks" (0..2^20 work items per kick launch, default 1024), "kick_pass_windows"
 * (-1..2, default -1), "deposit_span" (1024..2^20 source particles per deposit block, default 8192),
 * "deposit_multi" (slice-pair meshes: 1 = each deposit item takes all planes of an interaction, one
 * read of the particles; 0 = one plane per item; -1 = the faster of the two, timed on the device in
 * two consecutive collides every 256; default -1),
 * "deposit_min_span" (1024..2^20: smallest item the all-planes items are split to when a launch would
 * have fewer than two items per compute unit, default 2048) and "fft_transforms_per_block" (2, 4 or 8
 * transforms per workgroup of the charge rows, the convolution columns and the inverse rows, halved
 * while a workgroup's LDS would not fit; default 4).  octopic_autotune times candidate shapes by coordinate descent (median of `trials` collides per
 * candidate, after a warming one) on copies of the two beams in a scratch handle with the same
 * options, so the handle's own Green cache, counters and the beams are untouched, and sets the
 * fastest shape on the handle; best_ms (optional) receives its collide time. */
int octopic_set_launch_shape(octopic_handle *h, const char *name, int64_t value);
int octopic_get_launch_shape(const octopic_handle *h, const char *name, int64_t *value);
int octopic_autotune(octopic_handle *h, octopic_beam *beam1, octopic_beam *beam2, int32_t trials, double *best_ms);

/* ---------------------------------------------------------------------------------------------
 * Residual denoiser (no counterpart in the reference: the library's learned extension of the plain
 * PIC path).  A three-level U-Net (3x3 convolutions with 16/32/64 channels, ReLU, 2x2 average
 * pooling, nearest upsampling with skip concatenation, 1x1 head; 118129 parameters) that maps a
 * Gaussian-subtracted residual charge plane to its converged form.  bf16 MFMA products with fp32
 * accumulation in both the forward and the backward pass; fp32 parameters and Adam.
 *
 * Planes are NHWC: the input is [batch][height][width][2] (the residual and the shot-noise standard
 * deviation, both divided by the subtracted Gaussian's peak node value), the output and the target
 * [batch][height][width].  height and width: multiples of 16.  All plane pointers are device
 * memory; `loss` is a device float (the mean squared error over the batch); `grad_out` a device
 * float[num_params] or null.  Parameters are the flat float vector: per layer (e1a, e1b, e2a, e2b,
 * ba, bb, d2a, d2b, d1a, d1b) W[cout][3][3][cin] then b[cout]; the head's w[16] and b.  Calls on one
 * denoiser run in order on its stream (octopic_denoiser_set_stream; default: the null stream).
 * --------------------------------------------------------------------------------------------- */
typedef struct octopic_denoiser octopic_denoiser;

int64_t octopic_denoiser_num_params(void);
/* He-normal weights from `seed` (zero biases); batches of up to max_batch planes. */
int octopic_denoiser_create(int32_t device, int32_t height, int32_t width, int32_t max_batch, uint64_t seed,
                            octopic_denoiser **out);
void octopic_denoiser_destroy(octopic_denoiser *d);
const char *octopic_denoiser_last_error(const octopic_denoiser *d);
int octopic_denoiser_set_stream(octopic_denoiser *d, void *hip_stream);
/* Host float[n] copies of the parameters; set_params also resets the optimiser state. */
int octopic_denoiser_get_params(const octopic_denoiser *d, float *params, int64_t n);
int octopic_denoiser_set_params(octopic_denoiser *d, const float *params, int64_t n);
int octopic_denoiser_forward(octopic_denoiser *d, const float *input, int32_t batch, float *output);
/* Loss and gradient without an update. */
int octopic_denoiser_gradient(octopic_denoiser *d, const float *input, const float *target, int32_t batch,
                              float *loss, float *grad_out);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean squared error against target. */
int octopic_denoiser_train_step(octopic_denoiser *d, const float *input, const float *target, int32_t batch,
                                float lr, float *loss, float *grad_out);
/* Training pairs generated by the method itself: for samples first_sample .. first_sample+batch-1, a
 * two-component Gaussian-mixture slice (core with dipole and quadrupole modes, centred halo; one
 * sample in three a plain Gaussian; the distribution of each sample follows from seed and its id)
 * is drawn twice with Philox counter
 * streams, n_noisy in [16384, 262144] and n_target macroparticles, and deposited (method 0 CIC,
 * 1 TSC) on the height x width mesh in 2^-32 fixed point.  Each deposit has its own Gaussian (its
 * particles' moments) subtracted by the neutralised erf profile (gaussian_pic.jl:206-244, 543-576).
 * input: the noisy sample's network input; target: the converged residual at the noisy sample's
 * particle count, over the same peak; monitor (optional, device float[batch]): rms of the input
 * residual.  raw_planes ([batch][2][height*width] uint64) and raw_moments ([batch][2][5]: count,
 * sum u, sum v, sum u^2, sum v^2, u/v in cells, sums in 2^-24 fixed point), optional device
 * pointers, receive the deposits.  height * width <= 20480. */
/* Attach a denoiser to a collision handle (NULL detaches; the handle does not own it).  Every field
 * plane of every collide is then deposited as usual, its source slice's drifted Gaussian (the
 * source's own moments, gaussian_pic.jl:374-491) subtracted with the neutralised erf profile
 * (gaussian_pic.jl:206-244, 543-576), the residual passed through the network, and the plane
 * rebuilt as Gaussian + denoised residual with its deposited total kept, before the convolution.
 * A plane stays plain PIC when its source has fewer than two members or no positive width, or when
 * its monitor (rms input residual) exceeds monitor_limit (the library's fallback: three times the
 * largest monitor seen in training).  Needs interaction_grid = :slice_pair (else UNSUPPORTED), a
 * denoiser of height = ny and width = nx, and max_batch >= the field planes of the widest batch
 * (checked at collide).  The luminosity planes are not touched.  A denoiser holds the activations of
 * one collide: handles collided together by octopic_collide_many need denoisers of their own. */
int octopic_set_denoiser(octopic_handle *h, octopic_denoiser *d, double monitor_limit);

int octopic_denoiser_make_pairs(octopic_denoiser *d, uint64_t seed, int64_t first_sample, int32_t batch,
                                int64_t n_target, int32_t method, float *input, float *target, float *monitor,
                                uint64_t *raw_planes, uint64_t *raw_moments);

#ifdef __cplusplus
}
#endif

#endif /* OCTOPIC_H */
