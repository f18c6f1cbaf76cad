"""GPU parity of the Gaussian-subtracted PIC solver (GaussianPICPoissonSolver, gaussian_pic.jl) against the
oracle, through the C ABI.

- The device kick response (Bassetti-Erskine kick and kbb times the potential's Hessian diagonal,
  strong_beam.jl:1061-1128) against the oracle in every branch: <= 1e-12 of the largest value.
- Collides against the oracle on identical inputs, the PIC contract's bars: slices bit-identical, cache
  decisions identical, coordinates and luminosity within 1e-10; mesh geometry within 1e-12 relative.  The
  source mesh includes the subtracted Gaussian's margin box, mean +- margin_sigma * sigma from the source
  moments, which the device sums in chunk partials and the oracle serially, as the reference's CUDA path
  (block partials, gaussian_pic_cuda.jl:191-223) and CPU path (a serial loop, gaussian_pic.jl:374-405)
  do; the reference names that order as the only backend difference (test/runtests.jl:11384-11390); over CIC/TSC, 2nd/4th order, with and
  without the longitudinal kick, neutralize on/off, margin_sigma 5 and 0, round and flat beams, a
  one-member slice (its interactions fall back to plain PIC), and two collides (cache hits).
- The device solver reproduces the reference's published hybrid field-error medians through the collide
  (poisson_solver_review_2026_07_24.md:843-854), as the oracle test does.
- The FP32 form of the hybrid against the FP64 oracle within the FP32 form's bars.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from octopic import (EMASS_EV, PMASS_EV, Beam, GaussianPICPoissonSolver, LongitudinalSlicing, PICCollider,  # noqa: E402
                     PICPoissonSolver, beam_params)

from test_gpu_parity import compare_collide  # noqa: E402
from test_oracle_solver import be_kick  # noqa: E402

PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
PP = beam_params(1.0, PMASS_EV, 275e9, 6.9e10)
ETA_OUTER = (8.0 / 3.0 * 64.0 * 2.0 ** -52) ** 0.25


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def test_device_response_matches_oracle(oracle, gpu):
    col = PICCollider(PICPoissonSolver(grid=(32, 32)))
    rng = np.random.default_rng(5)
    s1 = 1e-3
    for eta in (0.0, 1e-5, 0.3 * ETA_OUTER, 0.75 * ETA_OUTER, 2 * ETA_OUTER, 0.3, 0.99):
        s2 = s1 * math.sqrt((1 - eta) / (1 + eta))
        for sx, sy in ((s1, s2), (s2, s1)):
            x = np.concatenate([rng.uniform(-5, 5, 300) * sx, [0.0, 1e-9 * sx, -2e-8 * sx]])
            y = np.concatenate([rng.uniform(-5, 5, 300) * sy, [0.0, 3e-9 * sy, 1e-8 * sy]])
            out = col.gaussian_beambeam_kick_response(torch.tensor(x, device=gpu), torch.tensor(y, device=gpu), 2.0,
                                                      sx, sy).cpu().numpy()
            ref = np.array([oracle.gaussian_beambeam_kick_response(2.0, sx, sy, a, b) for a, b in zip(x, y)])
            for k in range(4):
                scale = np.max(np.abs(ref[:, k]))
                assert np.max(np.abs(out[:, k] - ref[:, k])) <= 1e-12 * scale, (eta, k)


def round_beams(oracle, n1=20000, n2=8000, seed=0):
    e = oracle.generate_beam(n1, seed, 1, (1.0, 1.0, 20.0), (0, 0, 0), (100e-6, 100e-6, 0.01))
    p = oracle.generate_beam(n2, seed, 2, (1.0, 1.0, 90.909), (0, 0, 0), (100e-6, 100e-6, 0.06))
    return e, p


def flat_beams(oracle, n1=30000, n2=12000, seed=3):
    e = oracle.generate_beam(n1, seed, 1, (0.8, 0.072, 12.727), (0, 0, 0), (95e-6, 8.5e-6, 0.007))
    p = oracle.generate_beam(n2, seed, 2, (0.8, 0.072, 90.909), (0, 0, 0), (95e-6, 8.5e-6, 0.06))
    return e, p


CASES = [
    ("round", dict()),
    ("round", dict(deposit_method="TSC", field_derivative="fourth")),
    ("flat", dict(deposit_method="TSC", field_derivative="fourth")),
    ("flat", dict(deposit_method="TSC", neutralize=False, margin_sigma=0.0)),
    ("flat", dict(longitudinal_kick=False, margin_sigma=3.0)),
    ("round", dict(green_cache="none", deposit_method="TSC", luminosity_grid=(48, 40))),
]


@pytest.mark.parametrize("beams,kw", CASES, ids=[str(i) for i in range(len(CASES))])
def test_gpic_collide_matches_oracle(oracle, gpu, beams, kw):
    e, p = (round_beams if beams == "round" else flat_beams)(oracle)
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile", center_position="centroid")
    grid = (64, 64) if beams == "round" else (64, 48)
    compare_collide(oracle, gpu, GaussianPICPoissonSolver(grid=grid, slicing=sl, **kw), e, p, collides=2,
                    first_use_exact=False)


def test_gpic_one_member_slice_falls_back_to_pic(oracle, gpu):
    e, p = flat_beams(oracle, 20000, 7)
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile")
    compare_collide(oracle, gpu, GaussianPICPoissonSolver(grid=(32, 32), deposit_method="TSC", slicing=sl), e, p,
                    first_use_exact=False)


def test_gpic_published_medians_on_the_device(oracle, gpu):
    sy, n = 8.5e-6, 400
    u = (np.arange(1, n + 1) - 0.5) / n
    q = np.array([oracle.erfinv(2 * v - 1) for v in u]) * math.sqrt(2.0)
    for a, ref in zip((1.0, 5.0, 11.0, 25.0), (1.6e-4, 1.9e-4, 1.7e-4, 1.6e-4)):   # hybrid TSC 64^2
        sx = a * sy
        yy, xx = np.meshgrid(sy * q, sx * q, indexing="ij")
        src = np.zeros((6, xx.size))
        src[0], src[2] = xx.ravel(), yy.ravel()
        xg, yg = np.linspace(-4 * sx, 4 * sx, 81), np.linspace(-4 * sy, 4 * sy, 81)
        PY, PX = np.meshgrid(yg, xg, indexing="ij")
        m = PX.size
        pr = np.zeros((6, m + 2))
        pr[0, :m], pr[2, :m] = PX.ravel(), PY.ravel()
        ex, ey = math.sqrt(np.mean((src[0] - src[0].mean()) ** 2)), math.sqrt(np.mean((src[2] - src[2].mean()) ** 2))
        pr[0, m], pr[2, m] = src[0].mean() - 5 * ex, src[2].mean() - 5 * ey
        pr[0, m + 1], pr[2, m + 1] = src[0].mean() + 5 * ex, src[2].mean() + 5 * ey
        solver = GaussianPICPoissonSolver(grid=(64, 64), deposit_method="TSC", kbb1=1.0, kbb2=1.0,
                                          longitudinal_kick=False, green_cache="none",
                                          slicing=LongitudinalSlicing(nslices=1, method="normal_quantile"))
        b2 = Beam.from_host(pr, PP, device=gpu)
        PICCollider(solver).collide(Beam.from_host(src, PE, device=gpu), b2)
        out = b2.to_host()
        kx, ky = be_kick(sx, sy, PX.ravel(), PY.ravel())
        med = float(np.median(np.hypot(out[1, :m] - kx, out[3, :m] - ky)) / np.max(np.hypot(kx, ky)))
        assert abs(med / ref - 1.0) < 0.06, (a, med, ref)


def test_gpic_fp32_form_against_fp64_oracle(oracle, gpu):
    e, p = flat_beams(oracle)
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile", center_position="centroid")
    kw = dict(grid=(64, 64), deposit_method="TSC", field_derivative="fourth", slicing=sl)
    ce, cp = e.copy(), p.copy()
    lo = oracle.Collider(GaussianPICPoissonSolver(**kw)).collide(ce, PE, cp, PP)
    col = PICCollider(GaussianPICPoissonSolver(precision="fp32", **kw))
    be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    lg = col.collide(be, bp)
    assert math.isclose(lo, lg, rel_tol=1e-5)
    for a, g, a0 in ((ce, be.to_host(), e), (cp, bp.to_host(), p)):
        for k, bar in ((1, 1e-4), (3, 1e-4), (5, 3e-3)):
            dk = a[k] - a0[k]
            assert np.max(np.abs((g[k] - a0[k]) - dk)) <= bar * np.max(np.abs(dk)), k


def test_gpic_rejected_combinations(gpu):
    from octopic import UnsupportedOptionError
    for kw in (dict(slice_interpolation="quadratic"), dict(interaction_grid="node"), dict(coupling_tol=0.2)):
        with pytest.raises(UnsupportedOptionError):
            PICCollider(GaussianPICPoissonSolver(grid=(32, 32), **kw))
