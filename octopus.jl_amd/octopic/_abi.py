"""ctypes mirror of include/octopic.h (struct layouts and status codes).

The layouts here must match the header field for field; tests/test_abi.py checks the sizes against
a C compile of the header.
"""
import ctypes as C

OK = 0
ERR_INVALID_OPTION = 1
ERR_UNSUPPORTED = 2
ERR_NONFINITE = 3
ERR_DEGENERATE_EXTENT = 4
ERR_OOM = 5
ERR_DEVICE = 6
ERR_INVALID_ARGUMENT = 7

STATUS_NAMES = {
    OK: "OK",
    ERR_INVALID_OPTION: "INVALID_OPTION",
    ERR_UNSUPPORTED: "UNSUPPORTED",
    ERR_NONFINITE: "NONFINITE",
    ERR_DEGENERATE_EXTENT: "DEGENERATE_EXTENT",
    ERR_OOM: "OOM",
    ERR_DEVICE: "DEVICE",
    ERR_INVALID_ARGUMENT: "INVALID_ARGUMENT",
}

DEPOSIT = {"CIC": 0, "TSC": 1}
GREEN_TYPE = {"integrated": 0, "standard": 1, "lattice": 2}
GREEN_CACHE = {"none": 0, "slice_pair": 1}
SLICE_INTERPOLATION = {"linear": 0, "quadratic": 1}
INTERACTION_GRID = {"slice_pair": 0, "source_slice": 1, "node": 2}
GRID_EXTENT = {"extrema": 0, "sigma": 1}
BATCH_MODE = {"sequential": 0, "wavefront": 1}
SLICE_METHOD = {
    "equal_area": 0,
    "equal_count": 1,
    "equal_width": 2,
    "equal_spaced": 2,
    "normal_quantile": 3,
    "gaussian": 3,
    "Gaussian": 3,
    "specified": 4,
}
CENTER_POSITION = {"centroid": 0, "midpoint": 1}
OCTOPIC_DTYPE_F64 = 0
OCTOPIC_DTYPE_F32 = 1
FIELD_DERIVATIVE = {"second": 2, "fourth": 4}


class Slicing(C.Structure):
    _fields_ = [
        ("nslices", C.c_int32),
        ("method", C.c_int32),
        ("resolution", C.c_int32),
        ("center_position", C.c_int32),
        ("npositions", C.c_int32),
        ("reserved", C.c_int32),
        ("positions", C.POINTER(C.c_double)),
    ]


PRECISION = {"fp64": 0, "fp32": 1}
SOLVER_PIC = 0
SOLVER_GAUSSIAN_PIC = 1
SOLVER_SOFT_GAUSSIAN = 2
VIRTUAL_DRIFT = {"hirata": 0, "chromatic": 1, "exact": 2, "paraxial_frozen_longitudinal": 3, "chromatic_frozen_energy": 4}


class Options(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("deposit", C.c_int32),
        ("green_type", C.c_int32),
        ("green_cache", C.c_int32),
        ("field_derivative", C.c_int32),
        ("slice_interpolation", C.c_int32),
        ("interaction_grid", C.c_int32),
        ("grid_extent", C.c_int32),
        ("longitudinal_kick", C.c_int32),
        ("lum_nx", C.c_int32),
        ("lum_ny", C.c_int32),
        ("lum_deposit", C.c_int32),
        ("batch_mode", C.c_int32),
        ("cuda_flags", C.c_int32),
        ("precision", C.c_int32),
        ("grid_extent_sigma", C.c_double),
        ("min_extent_x", C.c_double),
        ("min_extent_y", C.c_double),
        ("grid_quantize", C.c_double),
        ("green_min_ratio", C.c_double),
        ("green_growth", C.c_double),
        ("kbb1", C.c_double),
        ("kbb2", C.c_double),
        ("luminosity_scale", C.c_double),
        ("slicing1", Slicing),
        ("slicing2", Slicing),
        ("solver", C.c_int32),
        ("neutralize", C.c_int32),
        ("margin_sigma", C.c_double),
        ("coupling_tol", C.c_double),
        ("min_sigma", C.c_double),
        ("virtual_drift", C.c_int32),
        ("gaussian_when_luminosity", C.c_int32),
        ("ignore_centroid1", C.c_int32),
        ("ignore_centroid2", C.c_int32),
        ("include_sigma_xy", C.c_int32),
        ("reserved", C.c_int32),
    ]


DPTR = C.POINTER(C.c_double)
FPTR = C.POINTER(C.c_float)


class BeamStruct(C.Structure):
    _fields_ = [
        ("x", C.c_void_p),
        ("px", C.c_void_p),
        ("y", C.c_void_p),
        ("py", C.c_void_p),
        ("z", C.c_void_p),
        ("pz", C.c_void_p),
        ("n", C.c_int64),
        ("charge", C.c_double),
        ("mc2", C.c_double),
        ("E0", C.c_double),
        ("r0", C.c_double),
        ("npart", C.c_double),
        ("on_device", C.c_int32),
        ("dtype", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("green_hits", C.c_int64),
        ("green_misses", C.c_int64),
        ("green_rebuilds", C.c_int64),
        ("dropped", C.c_int64),
        ("pairs", C.c_int64),
        ("batches", C.c_int64),
        ("widest_batch", C.c_int64),
        ("green_builds", C.c_int64),
        ("planes", C.c_int64),
        ("collides", C.c_int64),
        ("denoised_planes", C.c_int64),
        ("denoise_fallbacks", C.c_int64),
        ("deposit_all_planes", C.c_int64),
    ]


class InteractionRecord(C.Structure):
    _fields_ = [
        ("p", C.c_int32),
        ("i", C.c_int32),
        ("j", C.c_int32),
        ("dir", C.c_int32),
        ("cache_event", C.c_int32),
        ("reserved", C.c_int32),
        ("source_grid", C.c_double * 4),
        ("field_grid", C.c_double * 4),
    ]


# Every entry point the header declares, with its ctypes signature.
def declare(lib):
    """Attach argtypes/restype to the exported functions of liboctopic."""
    vp = C.c_void_p
    i32, i64 = C.c_int32, C.c_int64
    sigs = {
        "octopic_options_default": (i32, [C.POINTER(Options)]),
        "octopic_create": (i32, [C.POINTER(Options), i32, C.POINTER(vp)]),
        "octopic_destroy": (None, [vp]),
        "octopic_last_error": (C.c_char_p, [vp]),
        "octopic_set_stream": (i32, [vp, vp]),
        "octopic_collide": (i32, [vp, C.POINTER(BeamStruct), C.POINTER(BeamStruct), i64, i32, DPTR]),
        "octopic_get_stats": (i32, [vp, C.POINTER(Stats)]),
        "octopic_pair_luminosity": (
            i32,
            [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), DPTR, i64, C.POINTER(i64)],
        ),
        "octopic_last_slices": (i32, [vp, i32, DPTR, DPTR, DPTR, C.POINTER(i64), C.POINTER(i64), i64]),
        "octopic_interaction_records": (i32, [vp, C.POINTER(InteractionRecord), i64, C.POINTER(i64)]),
        "octopic_track_linear6d": (i32, [vp, DPTR, C.POINTER(BeamStruct)]),
        "octopic_track_lorentz_boost": (i32, [vp, C.c_double, i32, C.POINTER(BeamStruct)]),
        "octopic_track_crab_cavity": (i32, [vp, C.c_double, i32, DPTR, DPTR, DPTR, C.POINTER(BeamStruct)]),
        "octopic_gaussian_beambeam_kick": (i32, [vp, vp, vp, i64, C.c_double, C.c_double, vp, vp]),
        "octopic_faddeeva_w": (i32, [vp, vp, vp, i64, vp, vp]),
        "octopic_gaussian_beambeam_kick_response": (i32, [vp, vp, vp, i64, C.c_double, C.c_double, C.c_double, vp]),
        "octopic_gaussian_beambeam_kick_f32": (i32, [vp, vp, vp, i64, C.c_float, C.c_float, vp, vp]),
        "octopic_faddeeva_w_f32": (i32, [vp, vp, vp, i64, vp, vp]),
        "octopic_gaussian_beambeam_kick_response_f32": (i32, [vp, vp, vp, i64, C.c_float, C.c_float, C.c_float, vp]),
        "octopic_generate_gaussian_beam": (
            i32,
            [vp, C.c_uint64, C.c_uint64, DPTR, DPTR, DPTR, C.c_double, C.POINTER(BeamStruct)],
        ),
        "octopic_generate_gaussian_beam_noalloc": (
            i32,
            [vp, C.c_uint64, C.c_uint64, DPTR, DPTR, DPTR, C.c_double, C.POINTER(BeamStruct)],
        ),
        "octopic_debug_weights": (i32, [vp, i32, DPTR, i64, i32, C.POINTER(i32), DPTR]),
        "octopic_collide_many": (i32, [C.POINTER(vp), C.POINTER(BeamStruct), C.POINTER(BeamStruct), C.POINTER(i64),
                                       C.POINTER(i32), DPTR, C.POINTER(i32), i32]),
        "octopic_set_timing": (i32, [vp, i32]),
        "octopic_set_timing_families": (i32, [vp, C.c_char_p]),
        "octopic_kernel_times": (
            i32,
            [vp, DPTR, C.POINTER(i64), i64, C.POINTER(i64), C.c_char_p, i64],
        ),
        "octopic_denoiser_num_params": (i64, []),
        "octopic_denoiser_create": (i32, [i32, i32, i32, i32, C.c_uint64, C.POINTER(vp)]),
        "octopic_denoiser_destroy": (None, [vp]),
        "octopic_denoiser_last_error": (C.c_char_p, [vp]),
        "octopic_denoiser_set_stream": (i32, [vp, vp]),
        "octopic_denoiser_get_params": (i32, [vp, FPTR, i64]),
        "octopic_denoiser_set_params": (i32, [vp, FPTR, i64]),
        "octopic_denoiser_forward": (i32, [vp, vp, i32, vp]),
        "octopic_denoiser_gradient": (i32, [vp, vp, vp, i32, vp, vp]),
        "octopic_denoiser_train_step": (i32, [vp, vp, vp, i32, C.c_float, vp, vp]),
        "octopic_denoiser_make_pairs": (i32, [vp, C.c_uint64, i64, i32, i64, i32, vp, vp, vp, vp, vp]),
        "octopic_set_denoiser": (i32, [vp, vp, C.c_double]),
        "octopic_set_launch_shape": (i32, [vp, C.c_char_p, i64]),
        "octopic_get_launch_shape": (i32, [vp, C.c_char_p, C.POINTER(i64)]),
        "octopic_autotune": (i32, [vp, C.POINTER(BeamStruct), C.POINTER(BeamStruct), i32, DPTR]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return sigs


EXPORTED = [
    "octopic_options_default",
    "octopic_create",
    "octopic_destroy",
    "octopic_last_error",
    "octopic_set_stream",
    "octopic_collide",
    "octopic_get_stats",
    "octopic_pair_luminosity",
    "octopic_last_slices",
    "octopic_interaction_records",
    "octopic_track_linear6d",
    "octopic_track_lorentz_boost",
    "octopic_track_crab_cavity",
    "octopic_gaussian_beambeam_kick",
    "octopic_faddeeva_w",
    "octopic_gaussian_beambeam_kick_response",
    "octopic_gaussian_beambeam_kick_f32",
    "octopic_faddeeva_w_f32",
    "octopic_gaussian_beambeam_kick_response_f32",
    "octopic_generate_gaussian_beam",
    "octopic_generate_gaussian_beam_noalloc",
    "octopic_debug_weights",
    "octopic_collide_many",
    "octopic_set_timing",
    "octopic_set_timing_families",
    "octopic_kernel_times",
    "octopic_denoiser_num_params",
    "octopic_denoiser_create",
    "octopic_denoiser_destroy",
    "octopic_denoiser_last_error",
    "octopic_denoiser_set_stream",
    "octopic_denoiser_get_params",
    "octopic_denoiser_set_params",
    "octopic_denoiser_forward",
    "octopic_denoiser_gradient",
    "octopic_denoiser_train_step",
    "octopic_denoiser_make_pairs",
    "octopic_set_denoiser",
    "octopic_set_launch_shape",
    "octopic_get_launch_shape",
    "octopic_autotune",
]
