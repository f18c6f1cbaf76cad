"""The oracle's Faddeeva function and Bassetti-Erskine kick, pinned by:
- scipy.special.wofz on points in every branch (tests/golden/bassetti_erskine.npz);
- adaptive quadrature of the field integral for round, near-round (inside, across and outside the
  blend band), elliptic and near-axis points, both signs (same fixture);
- the reference's own known answers: the round-beam closed form 2x(1 - exp(-r^2/2s^2))/r^2 at rtol
  16 eps and a zero kick at the origin (test/runtests.jl:624-647), and the core gradient
  Kx/x -> 2/(s1(s1 + s2)) at rtol 32 eps for eta in {eta*, 0.001, 0.1, 0.9} (runtests.jl:741-752);
- the symmetries w(-conj z) = conj w(z), w(iy) real, and Kx odd in x, even in y.
"""
import math
import os

import numpy as np
import pytest

import oracle as O

FIX = os.path.join(os.path.dirname(__file__), "golden", "bassetti_erskine.npz")


@pytest.fixture(scope="module")
def fix():
    with np.load(FIX, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_faddeeva_against_wofz(fix):
    for z, w in zip(fix["z"], fix["w"]):
        got = O.faddeeva_w(complex(z))
        assert abs(got - w) <= 1e-12 * abs(w), (z, got, w)


def test_faddeeva_symmetries():
    rng = np.random.default_rng(5)
    for z in rng.uniform(-9, 9, 50) + 1j * rng.uniform(0, 9, 50):
        a, b = O.faddeeva_w(complex(-z.real, z.imag)), O.faddeeva_w(complex(z))
        assert abs(a - b.conjugate()) <= 1e-14 * abs(b)
    for y in (0.0, 0.3, 2.0, 6.5, 7.5, 100.0):
        assert O.faddeeva_w(complex(0.0, y)).imag == 0.0
    assert abs(O.faddeeva_w(0j) - 1.0) <= 1e-13


def test_weideman_coefficients_reproduce_w_at_the_origin():
    # w(0) = 2 sum(c) / L^2 + 1 / (sqrt(pi) L): Z = 1 at z = 0
    c, L = O.weideman_coefficients()
    assert L == pytest.approx(math.sqrt(32 / math.sqrt(2)), rel=1e-15)
    assert 2 * math.fsum(c) / L ** 2 + 1 / (math.sqrt(math.pi) * L) == pytest.approx(1.0, abs=1e-13)


def test_kick_against_quadrature(fix):
    for sx, sy, x, y, kx, ky in fix["kick"]:
        gx, gy = O.gaussian_beambeam_kick(sx, sy, x, y)
        assert math.hypot(gx - kx, gy - ky) <= 1e-10 * math.hypot(kx, ky), (sx, sy, x, y)


@pytest.mark.parametrize("x,y", [(1e-8, -0.5e-8), (0.1, -0.05), (0.2, -0.1), (2.0, -1.0)])
def test_round_beam_known_answer(x, y):
    r2 = x * x + y * y
    k = 2 * (1 - math.exp(-r2 / 2)) / r2 if r2 > 1e-12 else 1.0 - r2 / 4
    kx, ky = O.gaussian_beambeam_kick(1.0, 1.0, x, y)
    assert kx == pytest.approx(k * x, rel=16 * 2.0 ** -52) and ky == pytest.approx(k * y, rel=16 * 2.0 ** -52)
    assert O.gaussian_beambeam_kick(1.0, 1.0, 0.0, 0.0) == (0.0, 0.0)
    assert O.gaussian_beambeam_kick(0.0, 1.0, x, y) == (0.0, 0.0)


def test_core_gradient_known_answer():
    outer = math.sqrt(math.sqrt(8 / 3 * 64 * 2.0 ** -52))
    for eta in (outer, 0.001, 0.1, 0.9):
        s1, s2 = math.sqrt(1 + eta), math.sqrt(1 - eta)
        kx, _ = O.gaussian_beambeam_kick(s1, s2, 1e-8, 0.0)
        _, ky = O.gaussian_beambeam_kick(s1, s2, 0.0, 1e-8)
        assert kx / 1e-8 == pytest.approx(2 / (s1 * (s1 + s2)), rel=32 * 2.0 ** -52)
        assert ky / 1e-8 == pytest.approx(2 / (s2 * (s1 + s2)), rel=32 * 2.0 ** -52)


def test_kick_parities_and_axis_swap():
    for sx, sy in ((1.0, 1.0), (1.0, 1.00005), (2.0, 1.0), (1.0, 7.0)):
        for x, y in ((0.4, 0.9), (2.2, -0.3)):
            kx, ky = O.gaussian_beambeam_kick(sx, sy, x, y)
            assert O.gaussian_beambeam_kick(sx, sy, -x, y) == (-kx, ky)
            assert O.gaussian_beambeam_kick(sx, sy, x, -y) == (kx, -ky)
            # swapping the axes swaps the components
            sy_, sx_ = O.gaussian_beambeam_kick(sy, sx, y, x)
            assert sx_ == pytest.approx(kx, rel=1e-13) and sy_ == pytest.approx(ky, rel=1e-13)


def test_a_dropped_series_term_is_caught(fix):
    """The quadrature fixture separates the near-round series from its first-order truncation: the
    eta^2 and eta^3 terms matter at the 1e-10 bar inside the blend band."""
    sx, sy = 1.0 + 5e-5, 1.0
    rows = [r for r in fix["kick"] if r[0] == sx and r[1] == sy]
    assert rows
    v = (sx * sx + sy * sy) / 2
    eta = (sx * sx - sy * sy) / (2 * v)
    worst = 0.0
    for _, _, x, y, kx, ky in rows:
        # first order in eta only (the round kick plus the c1 term)
        q = (x * x + y * y) / (2 * v)
        m0 = -math.expm1(-q) / q
        m1 = (m0 - math.exp(-q)) / q
        m2 = (2 * m1 - math.exp(-q)) / q
        d = (x * x - y * y) / (2 * v)
        gx = x / v * (m0 + eta * (d * m2 - m1))
        worst = max(worst, abs(gx - kx) / math.hypot(kx, ky))
    assert worst > 1e-10
