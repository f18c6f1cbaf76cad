"""GPU parity of the soft-Gaussian solver (GaussianPoissonSolver, gaussian.jl) against the oracle, through the
C ABI.

- Collides against the oracle on identical inputs, the reference's backend-consistency bars: slices
  bit-identical, pair schedule identical, coordinates (z included: the exact and chromatic drifts move it)
  and luminosity per pair and in total within 1e-10.  The device sums each source slice's moments in chunk
  partials and the oracle serially below 4096 members and over 64 chunks above, which is the reference's
  CPU/CUDA difference too; over the three safe drifts and the two unsafe ones, with and without the
  longitudinal kick, either sampled beam, ignore_centroid, a min_sigma floor, slices beyond one pass of
  the kick grid, and two collides with a one-turn map between them.
- Float32 beams: the FP64 collide of the widened beams rounded back, z included, on device and host arrays.
- A non-finite source coordinate fails the collide and leaves the beams as they were; the unscheduled
  turns of a luminosity schedule return NaN with the same kicks; rejected combinations.
- Empty slices (every pair still runs, with and without a min_sigma floor) and beams smaller than a block.
- precision = "fp32" (GaussianPoissonSolver{Float32}'s Bassetti-Erskine evaluation in FP32) against the FP64
  oracle within the FP32 evaluator's error bounds, and on Float32 beams.
"""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from octopic import (EMASS_EV, PMASS_EV, Beam, EveryNSteps, GaussianPoissonSolver, LongitudinalSlicing,  # noqa: E402
                     PICCollider, UnsafeVirtualDrift, beam_params, linear6d_matrix)
from octopic import _abi  # noqa: E402
from octopic.backend import NonFiniteCoordinateError, UnsupportedOptionError, lib  # noqa: E402

from test_gpu_boundary import host_struct  # noqa: E402
from test_gpu_parity import compare_collide, flat_beams, round_beams  # noqa: E402

PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
PP = beam_params(1.0, PMASS_EV, 275e9, 6.9e10)


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _offset(e, p):
    """Offsets and tilts so that every centroid and covariance term of the map is exercised."""
    e, p = e.copy(), p.copy()
    e[0] += 1.5e-5
    e[1] += 2e-5 * e[4] / 0.01
    e[3] -= 1e-5
    p[2] += 2e-6
    p[1] -= 1e-5
    return np.ascontiguousarray(e), np.ascontiguousarray(p)


CASES = [
    dict(),
    dict(virtual_drift="exact", gaussian_when_luminosity=1),
    dict(virtual_drift="chromatic", ignore_centroid1=True),
    dict(virtual_drift="exact", longitudinal_kick=False, min_sigma=3e-5),
    dict(virtual_drift=UnsafeVirtualDrift("paraxial_frozen_longitudinal")),
    dict(virtual_drift=UnsafeVirtualDrift("chromatic_frozen_energy"), ignore_centroid2=True),
]


@pytest.mark.parametrize("beams", ["round", "flat"])
@pytest.mark.parametrize("kw", CASES)
def test_soft_collide_matches_oracle(oracle, gpu, beams, kw):
    e, p = (round_beams if beams == "round" else flat_beams)(oracle)
    e, p = _offset(e, p)
    sl1 = LongitudinalSlicing(nslices=7, method="normal_quantile", center_position="centroid")
    sl2 = LongitudinalSlicing(nslices=5, method="equal_area")
    solver = GaussianPoissonSolver(slicing1=sl1, slicing2=sl2, **kw)
    m = linear6d_matrix((1.0, 1.0, 20.0), (0.31, 0.29, 0.01))
    compare_collide(oracle, gpu, solver, e, p, collides=2, turn_map=(m, m))


def test_soft_collide_large_slices(oracle, gpu):
    # slices far above one pass of the kick grid and the 4096-member chunked moment sums
    e, p = flat_beams(oracle, 300000, 120000, seed=4)
    e, p = _offset(e, p)
    solver = GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=3, method="equal_count"), virtual_drift="exact")
    compare_collide(oracle, gpu, solver, e, p, collides=1)


@pytest.mark.parametrize("drift", ["hirata", "exact"])
def test_float32_beams_are_the_rounded_fp64_collide(oracle, gpu, drift):
    e, p = flat_beams(oracle, 20000, 9000, seed=6)
    e32, p32 = e.astype(np.float32), p.astype(np.float32)
    solver = GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=5, method="normal_quantile"), virtual_drift=drift)
    ref = PICCollider(solver)
    w1 = Beam.from_host(e32.astype(np.float64), PE, device=gpu)
    w2 = Beam.from_host(p32.astype(np.float64), PP, device=gpu)
    lum_ref = ref.collide(w1, w2, turn=0)
    if drift == "exact":
        assert not np.array_equal(w1.to_host()[4], e32[4].astype(np.float64))   # the exact drift moves z
    x1, x2 = w1.to_host().astype(np.float32), w2.to_host().astype(np.float32)
    col = PICCollider(solver)
    f1 = Beam.from_host(e32, PE, device=gpu, dtype=torch.float32)
    f2 = Beam.from_host(p32, PP, device=gpu, dtype=torch.float32)
    assert col.collide(f1, f2, turn=0) == lum_ref
    assert np.array_equal(f1.to_host(), x1) and np.array_equal(f2.to_host(), x2)
    h1, h2 = e32.copy(), p32.copy()
    hc = PICCollider(solver)
    s1, s2 = host_struct(h1, PE, _abi.OCTOPIC_DTYPE_F32), host_struct(h2, PP, _abi.OCTOPIC_DTYPE_F32)
    lum_h = C.c_double()
    hc._check(lib().octopic_collide(hc.h, C.byref(s1), C.byref(s2), 0, 1, C.byref(lum_h)))
    torch.cuda.synchronize()
    assert lum_h.value == lum_ref
    assert np.array_equal(h1, x1) and np.array_equal(h2, x2)


def test_nonfinite_source_fails_and_leaves_beams(oracle, gpu):
    e, p = flat_beams(oracle, 20000, 8000, seed=2)
    e[1, 123] = math.nan
    solver = GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=5, method="normal_quantile"))
    with pytest.raises(oracle.OracleError) as info:
        oracle.Collider(solver).collide(e.copy(), PE, p.copy(), PP)
    assert info.value.code == _abi.ERR_NONFINITE
    col = PICCollider(solver)
    be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    with pytest.raises(NonFiniteCoordinateError):
        col.collide(be, bp, turn=0)
    assert np.array_equal(be.to_host(), e, equal_nan=True) and np.array_equal(bp.to_host(), p)
    # the handle collides again once the beam is finite
    e[1, 123] = 0.0
    be2, bp2 = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    assert math.isfinite(col.collide(be2, bp2, turn=1))


def test_luminosity_schedule(oracle, gpu):
    e, p = round_beams(oracle, 8000, 4000)
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile")
    sched = GaussianPoissonSolver(slicing=sl, luminosity_schedule=EveryNSteps(step=2))
    plain = GaussianPoissonSolver(slicing=sl)
    a = PICCollider(sched)
    b = PICCollider(plain)
    a1, a2 = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    b1, b2 = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    for turn in range(3):
        la = a.collide(a1, a2, turn=turn)
        lb = b.collide(b1, b2, turn=turn)
        if sched.compute_luminosity(turn):
            assert la == lb
        else:
            assert math.isnan(la)
    assert np.array_equal(a1.to_host(), b1.to_host()) and np.array_equal(a2.to_host(), b2.to_host())


def test_rejected_combinations(gpu):
    with pytest.raises(UnsupportedOptionError):
        PICCollider(GaussianPoissonSolver(include_sigma_xy=True))
    o, keep = GaussianPoissonSolver().to_options()
    o.precision = 7
    h = C.c_void_p()
    assert lib().octopic_create(C.byref(o), 0, C.byref(h)) == _abi.ERR_INVALID_OPTION


@pytest.mark.parametrize("min_sigma", [0.0, 2e-5])
def test_empty_slices_run_every_pair(oracle, gpu, min_sigma):
    """`:specified` boundaries at z = -9 m and -8.5 m leave the first two slices of both beams empty. The
    soft-Gaussian solver runs every pair (gaussian.jl:7-22): an empty source is a zero-width Gaussian at
    the origin (floored by min_sigma), whose kick on the other slice is zero without the floor."""
    e, p = round_beams(oracle, 12000, 6000, seed=4)
    sl = LongitudinalSlicing(nslices=4, method="specified", center_position="centroid", positions=(-9.0, -8.5, 0.0))
    gcol = compare_collide(oracle, gpu, GaussianPoissonSolver(slicing=sl, min_sigma=min_sigma), e, p)
    _, _, ran, _ = gcol.pair_luminosity()
    assert all(ran)
    assert gcol.stats()["pairs"] == 16


def test_tiny_ragged_beams(oracle, gpu):
    """Beams far smaller than one block, with sizes that fill no block evenly."""
    e, p = round_beams(oracle, 61, 23, seed=9)
    compare_collide(oracle, gpu, GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=3, method="equal_count"),
                                                       virtual_drift="exact"), e, p, collides=2)


@pytest.mark.parametrize("kw", [dict(), dict(virtual_drift="exact", gaussian_when_luminosity=1),
                                dict(virtual_drift="chromatic", min_sigma=3e-5, ignore_centroid2=True)])
def test_fp32_evaluation_against_fp64_oracle(oracle, gpu, kw):
    """GaussianPoissonSolver{Float32}'s Bassetti-Erskine evaluation (precision="fp32": field, Hessian and density
    in FP32 with the Float32 constants) against the FP64 oracle on the same inputs.  The FP32 force's natural-scale
    error is below C eps / sqrt(eta*/2) = 6.8e-6 of 1/sigma (the near-round transition sweep), its response's below
    6e-5 of 1/sigma^2, and every coordinate change is linear in them: each coordinate's difference from the
    oracle stays within 1e-4 of that coordinate's largest change in the collide (px, py within 2e-5), and the
    luminosity within 1e-5 relative.  The FP32 path must differ from the FP64 one (it ran)."""
    e, p = flat_beams(oracle, 20000, 9000, seed=5)
    e, p = _offset(e, p)
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile")
    s64 = GaussianPoissonSolver(slicing=sl, **kw)
    s32 = GaussianPoissonSolver(slicing=sl, precision="fp32", **kw)
    oe, op = e.copy(), p.copy()
    lum_ref = oracle.Collider(s64).collide(oe, PE, op, PP)
    col = PICCollider(s32)
    b1, b2 = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    lum = col.collide(b1, b2, turn=0)
    g1, g2 = b1.to_host(), b2.to_host()
    assert abs(lum - lum_ref) <= 1e-5 * abs(lum_ref), (lum, lum_ref)
    differs = False
    for g, o, start in ((g1, oe, e), (g2, op, p)):
        for c in range(6):
            change = np.abs(o[c] - start[c]).max()
            d = np.abs(g[c] - o[c]).max()
            tol = (2e-5 if c in (1, 3) else 1e-4) * change + 1e-300
            assert d <= tol, (c, d, change)
            differs |= d > 0
    assert differs
    # the FP64 handle on the same inputs stays within the path's 1e-10 of the oracle
    c64 = PICCollider(s64)
    h1, h2 = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    assert abs(c64.collide(h1, h2, turn=0) - lum_ref) <= 1e-10 * abs(lum_ref)


def test_fp32_evaluation_float32_beams(oracle, gpu):
    """Float32 beams under the FP32 evaluation: the FP32-evaluation collide of the widened beams, rounded back."""
    e, p = flat_beams(oracle, 12000, 6000, seed=8)
    e32, p32 = e.astype(np.float32), p.astype(np.float32)
    solver = GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=4, method="normal_quantile"), precision="fp32")
    w1 = Beam.from_host(e32.astype(np.float64), PE, device=gpu)
    w2 = Beam.from_host(p32.astype(np.float64), PP, device=gpu)
    lum_w = PICCollider(solver).collide(w1, w2, turn=0)
    f1 = Beam.from_host(e32, PE, device=gpu, dtype=torch.float32)
    f2 = Beam.from_host(p32, PP, device=gpu, dtype=torch.float32)
    assert PICCollider(solver).collide(f1, f2, turn=0) == lum_w
    assert np.array_equal(f1.to_host(), w1.to_host().astype(np.float32))
    assert np.array_equal(f2.to_host(), w2.to_host().astype(np.float32))
