"""octopic -- MI355X-native strong-strong PIC beam-beam collision, a drop-in for Octopus.jl's
collide!(::PICPoissonSolver, beam1, beam2, backend[, ctx]) path.

The compute runs in liboctopic.so (hand-written HIP kernels for gfx950, C ABI in include/octopic.h);
this package mirrors the reference's solver options and entry points on the host.
"""
from .solver import (AtTurns, ArgumentError, BeamParams, EveryNSteps, GaussianPICPoissonSolver, GaussianPoissonSolver,
                     LongitudinalSlicing, PICPoissonSolver, UnsafeVirtualDrift, EMASS_EV, ME0, PMASS_EV, RE, TWOPI,
                     beam_params, linear6d_matrix)
from .backend import (Beam, BackendUnavailable, DegenerateExtentError, HIPBackend, NonFiniteCoordinateError,
                      OctopicError, PICCollider, UnsupportedOptionError, collide_, collide_many, generate_beam, lib)
from .artifact import RunArtifact, read_luminosity

__all__ = [
    "AtTurns", "ArgumentError", "BeamParams", "EveryNSteps", "GaussianPICPoissonSolver", "GaussianPoissonSolver",
    "LongitudinalSlicing", "PICPoissonSolver", "UnsafeVirtualDrift",
    "EMASS_EV", "ME0", "PMASS_EV", "RE", "TWOPI", "beam_params", "linear6d_matrix",
    "Beam", "BackendUnavailable", "DegenerateExtentError", "HIPBackend", "NonFiniteCoordinateError",
    "OctopicError", "PICCollider", "UnsupportedOptionError", "collide_", "collide_many", "generate_beam", "lib",
    "RunArtifact", "read_luminosity",
]
