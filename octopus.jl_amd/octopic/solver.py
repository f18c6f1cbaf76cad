"""Host-side mirror of the reference's PIC solver configuration.

`PICPoissonSolver`, `LongitudinalSlicing`, `EveryNSteps` and `AtTurns` keep the reference's keyword
names, defaults, validation order and error type, so that a caller of
    PICPoissonSolver(; grid=(128,128), deposit_method=:TSC, ...)
(src/tasks/strongstrong/interface.jl:1103-1507) writes
    PICPoissonSolver(grid=(128, 128), deposit_method="TSC", ...)
here.  Symbols are strings.  `to_options()` lowers a solver to the C ABI struct of include/octopic.h.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

from . import _abi

# Physical constants (src/constants/Constants.jl)
RE = 2.8179403205e-15
EMASS_EV = 0.51099895069e6
ME0 = EMASS_EV
PMASS_EV = 938.27208943e6
TWOPI = 6.283185307179586476925286766559005768394338


class ArgumentError(ValueError):
    """The reference's ArgumentError: an invalid or unsupported option, or invalid input."""


@dataclass(frozen=True)
class LongitudinalSlicing:
    """LongitudinalSlicing(; nslices=1, method=:equal_area, resolution=100, center_position=:centroid,
    positions=Float64[]) (interface.jl:800-835)."""

    nslices: int = 1
    method: str = "equal_area"
    resolution: int = 100
    center_position: str = "centroid"
    positions: Tuple[float, ...] = ()

    def __post_init__(self):
        if self.nslices <= 0:
            raise ArgumentError("nslices must be positive")
        if self.resolution <= 0:
            raise ArgumentError("slicing resolution must be positive")
        if self.method not in _abi.SLICE_METHOD:
            raise ArgumentError(f"unsupported longitudinal slicing method :{self.method}")
        if self.center_position not in _abi.CENTER_POSITION:
            raise ArgumentError("center_position must be :centroid or :midpoint")
        object.__setattr__(self, "positions", tuple(float(p) for p in self.positions))
        if self.method == "specified" and len(self.positions) != self.nslices - 1:
            raise ArgumentError("specified slicing requires nslices-1 internal positions")

    def _lower(self, keep):
        s = _abi.Slicing()
        s.nslices = self.nslices
        s.method = _abi.SLICE_METHOD[self.method]
        s.resolution = self.resolution
        s.center_position = _abi.CENTER_POSITION[self.center_position]
        s.npositions = len(self.positions)
        if self.positions:
            arr = (C.c_double * len(self.positions))(*self.positions)
            keep.append(arr)
            s.positions = C.cast(arr, C.POINTER(C.c_double))
        return s


@dataclass(frozen=True)
class EveryNSteps:
    """EveryNSteps(; start=0, stop=typemax(Int), step=1) (src/tasks/BeamObservers.jl:40-55, 74-78)."""

    start: int = 0
    stop: int = 2**63 - 1
    step: int = 1

    def __post_init__(self):
        if self.step <= 0:
            raise ArgumentError("step must be positive")

    def should_run(self, turn: int) -> bool:
        return turn >= self.start and turn < self.stop and (turn - self.start) % self.step == 0


@dataclass(frozen=True)
class AtTurns:
    """AtTurns(turns) (src/tasks/BeamObservers.jl:57-60)."""

    turns: frozenset = field(default_factory=frozenset)

    def __init__(self, turns):
        object.__setattr__(self, "turns", frozenset(int(t) for t in turns))

    def should_run(self, turn: int) -> bool:
        return turn in self.turns


def _extent_pair(value):
    """_solver_transverse_extent (interface.jl:1087-1101)."""
    if isinstance(value, (int, float)):
        req = (float(value), float(value))
    elif isinstance(value, (tuple, list)) and len(value) == 2:
        req = (float(value[0]), float(value[1]))
    else:
        raise ArgumentError(f"min_transverse_extent must be a real scalar or a two-real tuple; got {value!r}.")
    if not all(math.isfinite(v) and v >= 0.0 for v in req):
        raise ArgumentError(f"min_transverse_extent values must be finite and nonnegative; got {value!r}.")
    return req


@dataclass(frozen=True)
class PICPoissonSolver:
    """PICPoissonSolver{Float64} (interface.jl:1324-1507): keyword constructor with eager validation."""

    kbb1: Optional[float] = None
    kbb2: Optional[float] = None
    luminosity_scale: Optional[float] = None
    grid: Tuple[int, int] = (128, 128)
    deposit_method: str = "CIC"
    green_type: str = "integrated"
    green_cache: str = "slice_pair"
    field_derivative: str = "second"
    slice_interpolation: str = "linear"
    interaction_grid: str = "slice_pair"
    grid_extent: str = "extrema"
    grid_extent_sigma: float = 6.0
    min_transverse_extent: Tuple[float, float] = (0.0, 0.0)
    grid_quantize: float = 0.0
    slice_pair_green_min_ratio: float = 0.50
    slice_pair_green_growth: float = 0.25
    longitudinal_kick: bool = True
    batch_mode: str = "wavefront"
    cuda_async: bool = True
    cuda_batch_fft: bool = True
    cuda_wavefront_fft: bool = True
    cuda_indexed_wavefront: bool = True
    luminosity_grid: Optional[Tuple[int, int]] = None
    luminosity_deposit_method: Optional[str] = None
    luminosity_schedule: object = None
    slicing: LongitudinalSlicing = LongitudinalSlicing()
    slicing1: Optional[LongitudinalSlicing] = None
    slicing2: Optional[LongitudinalSlicing] = None
    # "fp64": the reference's PICPoissonSolver{Float64} arithmetic.  "fp32": this backend's FP32 form
    # (deposit weights, FFT convolution, field and kick gather in FP32 on FP64 beams; Green function,
    # slice moments and luminosity in FP64); no counterpart in the reference.
    precision: str = "fp64"

    def __post_init__(self):
        # Same order of checks as the reference constructor (interface.jl:1410-1476).
        if not (0.0 < float(self.slice_pair_green_min_ratio) <= 1.0):
            raise ArgumentError(
                f"slice_pair_green_min_ratio must be in (0, 1]; got {self.slice_pair_green_min_ratio}.")
        if not (float(self.slice_pair_green_growth) >= 0.0):
            raise ArgumentError(f"slice_pair_green_growth must be non-negative; got {self.slice_pair_green_growth}.")
        g = (int(self.grid[0]), int(self.grid[1]))
        if not all(v >= 5 for v in g):
            raise ArgumentError(f"PICPoissonSolver grid dimensions must both be at least 5; got {self.grid}.")
        object.__setattr__(self, "grid", g)
        if self.deposit_method not in _abi.DEPOSIT:
            raise ArgumentError(f"deposit_method must be :CIC or :TSC; got :{self.deposit_method}.")
        if self.green_type not in _abi.GREEN_TYPE:
            raise ArgumentError(f"green_type must be :integrated, :standard or :lattice; got :{self.green_type}.")
        if self.green_cache not in _abi.GREEN_CACHE:
            raise ArgumentError(f"green_cache must be :none or :slice_pair; got :{self.green_cache}.")
        if self.field_derivative not in _abi.FIELD_DERIVATIVE:
            raise ArgumentError(f"field_derivative must be :second or :fourth; got :{self.field_derivative}.")
        if self.slice_interpolation not in _abi.SLICE_INTERPOLATION:
            raise ArgumentError(f"slice_interpolation must be :linear or :quadratic; got :{self.slice_interpolation}.")
        if self.grid_extent not in _abi.GRID_EXTENT:
            raise ArgumentError(f"grid_extent must be :extrema or :sigma; got :{self.grid_extent}.")
        if not (float(self.grid_extent_sigma) > 0.0):
            raise ArgumentError(f"grid_extent_sigma must be positive; got {self.grid_extent_sigma}.")
        object.__setattr__(self, "min_transverse_extent", _extent_pair(self.min_transverse_extent))
        if not (float(self.grid_quantize) >= 0.0):
            raise ArgumentError(f"grid_quantize must be non-negative (0 disables); got {self.grid_quantize}.")
        if self.interaction_grid not in _abi.INTERACTION_GRID:
            raise ArgumentError(
                f"interaction_grid must be :slice_pair, :source_slice or :node; got :{self.interaction_grid}.")
        if not (self.interaction_grid == "slice_pair" or self.grid_extent == "extrema"):
            raise ArgumentError(
                f"grid_extent = :{self.grid_extent} is not implemented for interaction_grid = "
                f":{self.interaction_grid}: that mode sizes its mesh from the union/per-node extrema of its own "
                "particle sets and applies no extent estimator, so a non-default value would be silently ignored.")
        if self.batch_mode not in _abi.BATCH_MODE:
            raise ArgumentError(f"batch_mode must be :sequential or :wavefront; got :{self.batch_mode}.")
        if self.luminosity_grid is not None:
            lg = (int(self.luminosity_grid[0]), int(self.luminosity_grid[1]))
            if not all(v >= 3 for v in lg):
                raise ArgumentError(f"luminosity_grid dimensions must both be at least 3; got {self.luminosity_grid}.")
            object.__setattr__(self, "luminosity_grid", lg)
        if self.precision not in _abi.PRECISION:
            raise ArgumentError(f"precision must be fp64 or fp32; got {self.precision!r}.")
        if self.luminosity_deposit_method not in (None, "CIC", "TSC"):
            raise ArgumentError(
                f"luminosity_deposit_method must be nothing, :CIC, or :TSC; got :{self.luminosity_deposit_method}.")

    # --- resolved views (interface.jl:1513-1517) ---
    @property
    def resolved_slicing1(self) -> LongitudinalSlicing:
        return self.slicing1 if self.slicing1 is not None else self.slicing

    @property
    def resolved_slicing2(self) -> LongitudinalSlicing:
        return self.slicing2 if self.slicing2 is not None else self.slicing

    @property
    def resolved_luminosity_grid(self):
        return self.grid if self.luminosity_grid is None else self.luminosity_grid

    @property
    def resolved_luminosity_deposit_method(self):
        return self.deposit_method if self.luminosity_deposit_method is None else self.luminosity_deposit_method

    def compute_luminosity(self, turn: Optional[int]) -> bool:
        """_pic_compute_luminosity (interface.jl:2596-2620): every turn unless a schedule says otherwise."""
        if turn is None or self.luminosity_schedule is None:
            return True
        return bool(self.luminosity_schedule.should_run(turn))

    def to_options(self):
        """Lower to the C ABI options struct.  Returns (struct, keepalive)."""
        keep = []
        o = _abi.Options()
        o.nx, o.ny = self.grid
        o.deposit = _abi.DEPOSIT[self.deposit_method]
        o.green_type = _abi.GREEN_TYPE[self.green_type]
        o.green_cache = _abi.GREEN_CACHE[self.green_cache]
        o.field_derivative = _abi.FIELD_DERIVATIVE[self.field_derivative]
        o.slice_interpolation = _abi.SLICE_INTERPOLATION[self.slice_interpolation]
        o.interaction_grid = _abi.INTERACTION_GRID[self.interaction_grid]
        o.grid_extent = _abi.GRID_EXTENT[self.grid_extent]
        o.longitudinal_kick = 1 if self.longitudinal_kick else 0
        lg = self.resolved_luminosity_grid
        o.lum_nx, o.lum_ny = lg
        o.lum_deposit = _abi.DEPOSIT[self.resolved_luminosity_deposit_method]
        o.batch_mode = _abi.BATCH_MODE[self.batch_mode]
        o.cuda_flags = (int(self.cuda_async) | (int(self.cuda_batch_fft) << 1) |
                        (int(self.cuda_wavefront_fft) << 2) | (int(self.cuda_indexed_wavefront) << 3))
        o.precision = _abi.PRECISION[self.precision]
        o.grid_extent_sigma = float(self.grid_extent_sigma)
        o.min_extent_x, o.min_extent_y = self.min_transverse_extent
        o.grid_quantize = float(self.grid_quantize)
        o.green_min_ratio = float(self.slice_pair_green_min_ratio)
        o.green_growth = float(self.slice_pair_green_growth)
        nan = float("nan")
        o.kbb1 = nan if self.kbb1 is None else float(self.kbb1)
        o.kbb2 = nan if self.kbb2 is None else float(self.kbb2)
        o.luminosity_scale = nan if self.luminosity_scale is None else float(self.luminosity_scale)
        o.slicing1 = self.resolved_slicing1._lower(keep)
        o.slicing2 = self.resolved_slicing2._lower(keep)
        o.solver = _abi.SOLVER_PIC
        o.neutralize = 1
        o.margin_sigma = 5.0
        o.coupling_tol = math.inf
        o.min_sigma = 0.0
        o.virtual_drift = _abi.VIRTUAL_DRIFT["hirata"]
        o.gaussian_when_luminosity = 2
        return o, keep


class GaussianPICPoissonSolver:
    """GaussianPICPoissonSolver{Float64}(; margin_sigma=5.0, neutralize=true, coupling_tol=Inf, kwargs...)
    (src/tasks/strongstrong/gaussian_pic.jl:87-131): the Gaussian-subtracted PIC hybrid.  Every other keyword
    builds the embedded PICPoissonSolver (`self.pic`), with its own validation.  The hybrid deposits each
    source slice, subtracts the erf-integrated drifted Gaussian of the slice's own moments, solves the
    residual with the PIC machinery and adds the exact Bassetti-Erskine field of that Gaussian back per field
    particle.  It runs with interaction_grid = :slice_pair, slice_interpolation = :linear and grid_extent =
    :extrema only (the reference rejects the others at collide time, gaussian_pic.jl:33-41)."""

    def __init__(self, margin_sigma: float = 5.0, neutralize: bool = True, coupling_tol: float = math.inf, **kwargs):
        self.pic = PICPoissonSolver(**kwargs)
        ms, ct = float(margin_sigma), float(coupling_tol)
        if not ms >= 0.0:
            raise ArgumentError(f"margin_sigma must be non-negative; got {margin_sigma}.")
        if not ct >= 0.0:
            raise ArgumentError(f"coupling_tol must be non-negative; got {coupling_tol}.")
        self.margin_sigma = ms
        self.neutralize = bool(neutralize)
        self.coupling_tol = ct

    def __getattr__(self, name):
        # the embedded PIC solver's options and resolved views (slicing, luminosity grid, schedule, ...)
        if name == "pic":
            raise AttributeError(name)
        return getattr(self.pic, name)

    def compute_luminosity(self, turn):
        return self.pic.compute_luminosity(turn)

    def to_options(self):
        o, keep = self.pic.to_options()
        o.solver = _abi.SOLVER_GAUSSIAN_PIC
        o.neutralize = 1 if self.neutralize else 0
        o.margin_sigma = self.margin_sigma
        o.coupling_tol = self.coupling_tol
        return o, keep


@dataclass(frozen=True)
class UnsafeVirtualDrift:
    """UnsafeVirtualDrift(name) (strong_beam.jl:41-59): the non-symplectic diagnostic drifts
    :paraxial_frozen_longitudinal and :chromatic_frozen_energy."""

    name: str

    def __post_init__(self):
        if self.name not in ("paraxial_frozen_longitudinal", "chromatic_frozen_energy"):
            raise ArgumentError(f"unknown unsafe virtual drift :{self.name}; expected one of "
                                "(:chromatic_frozen_energy, :paraxial_frozen_longitudinal)")


@dataclass(frozen=True)
class GaussianPoissonSolver:
    """GaussianPoissonSolver{Float64}(; kbb1=nothing, kbb2=nothing, luminosity_scale=nothing, slicing, slicing1,
    slicing2, min_sigma=0, gaussian_when_luminosity=2, ignore_centroid1=false, ignore_centroid2=false,
    longitudinal_kick=true, virtual_drift=:hirata, include_sigma_xy=false, batch_mode=:wavefront,
    luminosity_schedule=nothing) (interface.jl:876-1010): the sliced soft-Gaussian solver (gaussian.jl).  Every
    slice pair, in collision order, kicks each beam's slice with the exact Bassetti-Erskine field of the other
    slice's Gaussian moments, transported to each particle's collision point, between the virtual drifts.
    include_sigma_xy = true (the coupled, rotated kick) is accepted here as by the reference's constructor and
    rejected by the backends with UnsupportedOptionError.  precision = "fp32" is GaussianPoissonSolver{Float32}'s
    Bassetti-Erskine evaluation (field, Hessian and density in FP32 with the reference's Float32 constants,
    strong_beam.jl:770-783, 809-811); the drifts, slice moments, coordinates and luminosity sums stay FP64."""

    kbb1: Optional[float] = None
    kbb2: Optional[float] = None
    luminosity_scale: Optional[float] = None
    slicing: LongitudinalSlicing = LongitudinalSlicing()
    slicing1: Optional[LongitudinalSlicing] = None
    slicing2: Optional[LongitudinalSlicing] = None
    min_sigma: float = 0.0
    gaussian_when_luminosity: int = 2
    ignore_centroid1: bool = False
    ignore_centroid2: bool = False
    longitudinal_kick: bool = True
    virtual_drift: object = "hirata"
    include_sigma_xy: bool = False
    batch_mode: str = "wavefront"
    luminosity_schedule: object = None
    precision: str = "fp64"

    def __post_init__(self):
        # The reference constructor's order of checks (interface.jl:976-1000).
        ms = float(self.min_sigma)
        if not (math.isfinite(ms) and ms >= 0.0):
            raise ArgumentError(f"min_sigma must be finite and nonnegative; got {self.min_sigma}.")
        object.__setattr__(self, "min_sigma", ms)
        if self.gaussian_when_luminosity not in (1, 2):
            raise ArgumentError("gaussian_when_luminosity must be 1 or 2")
        if self.batch_mode not in _abi.BATCH_MODE:
            raise ArgumentError("batch_mode must be :sequential or :wavefront")
        vd = self.virtual_drift
        if isinstance(vd, str) and vd not in ("hirata", "chromatic", "exact"):
            raise ArgumentError(f"unknown virtual_drift :{vd}; expected one of (:hirata, :chromatic, :exact)")
        if not isinstance(vd, (str, UnsafeVirtualDrift)):
            raise ArgumentError(f"virtual_drift must be a supported Symbol or AbstractVirtualDrift, got {type(vd).__name__}")
        for k in ("kbb1", "kbb2", "luminosity_scale"):
            v = getattr(self, k)
            if v is not None:
                object.__setattr__(self, k, float(v))
        if self.precision not in _abi.PRECISION:
            raise ArgumentError(f"precision must be fp64 or fp32; got {self.precision!r}.")

    @property
    def resolved_slicing1(self) -> LongitudinalSlicing:
        return self.slicing1 if self.slicing1 is not None else self.slicing

    @property
    def resolved_slicing2(self) -> LongitudinalSlicing:
        return self.slicing2 if self.slicing2 is not None else self.slicing

    def compute_luminosity(self, turn: Optional[int]) -> bool:
        """The schedule gates only the reported value (the kicks are the same on every turn): unscheduled
        turns return NaN (interface.jl:906-912)."""
        if self.luminosity_schedule is None or turn is None:
            return True
        return bool(self.luminosity_schedule.should_run(int(turn)))

    def to_options(self):
        o, keep = PICPoissonSolver(kbb1=self.kbb1, kbb2=self.kbb2, luminosity_scale=self.luminosity_scale,
                                   longitudinal_kick=self.longitudinal_kick, batch_mode=self.batch_mode,
                                   slicing=self.slicing, slicing1=self.slicing1, slicing2=self.slicing2).to_options()
        o.solver = _abi.SOLVER_SOFT_GAUSSIAN
        o.min_sigma = self.min_sigma
        vd = self.virtual_drift
        o.virtual_drift = _abi.VIRTUAL_DRIFT[vd.name if isinstance(vd, UnsafeVirtualDrift) else vd]
        o.gaussian_when_luminosity = int(self.gaussian_when_luminosity)
        o.ignore_centroid1 = 1 if self.ignore_centroid1 else 0
        o.ignore_centroid2 = 1 if self.ignore_centroid2 else 0
        o.include_sigma_xy = 1 if self.include_sigma_xy else 0
        o.precision = _abi.PRECISION[self.precision]
        return o, keep


@dataclass
class BeamParams:
    """BeamParams(; charge, mc2, E0, r0, npart) (src/beam/Beam.jl:29-35)."""

    charge: float = 0.0
    mc2: float = 0.0
    E0: float = 0.0
    r0: float = 0.0
    npart: float = 0.0


def beam_params(charge, mass_ev, energy_ev, npart):
    """The production recipe: r0 = RE * ME0 / mass (examples/strong_strong_tracking.jl:158-166)."""
    return BeamParams(charge=float(charge), mc2=float(mass_ev), E0=float(energy_ev),
                      r0=RE * ME0 / float(mass_ev), npart=float(npart))


def linear6d_matrix(beta1: Sequence[float], dmu: Sequence[float], beta2: Optional[Sequence[float]] = None,
                    alpha1: Sequence[float] = (0.0, 0.0, 0.0), alpha2: Optional[Sequence[float]] = None):
    """Linear6DSpec(; beta1, dmu, beta2=beta1, alpha1, alpha2=alpha1) with zero coupling, dispersion
    and crab dispersion: the block-diagonal Twiss map (src/elements/linear6d.jl:233-276).
    Returns a row-major list of 36 floats."""
    beta2 = beta1 if beta2 is None else beta2
    alpha2 = alpha1 if alpha2 is None else alpha2
    m = [0.0] * 36
    for i in range(3):
        s, c = math.sin(dmu[i]), math.cos(dmu[i])
        r = 2 * i
        b1, b2, a1, a2 = float(beta1[i]), float(beta2[i]), float(alpha1[i]), float(alpha2[i])
        m[r * 6 + r] = math.sqrt(b2 / b1) * (c + a1 * s)
        m[r * 6 + r + 1] = math.sqrt(b2 * b1) * s
        m[(r + 1) * 6 + r] = (-(1.0 + a1 * a2) * s + (a1 - a2) * c) / math.sqrt(b1 * b2)
        m[(r + 1) * 6 + r + 1] = math.sqrt(b1 / b2) * (c - a2 * s)
    return m
