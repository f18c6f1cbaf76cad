"""The turn around the collision: the crossing-angle boost and the thin crab cavity of the oracle.

- LorentzBoost / RevLorentzBoost (lorentz_boost.jl:128-166): the reverse map undoes the forward map
  (both orders); the forward map's phase-space Jacobian determinant is sec(angle)^3 (its documented
  quasi-symplectic factor, lorentz_boost.jl:46-50), the reverse map's cos(angle)^3; angle 0 is the
  identity; to first order in the angle, x picks up z tan(angle) and z picks up nothing.
- ThinCrabCavity (crab_cavity.jl:99-164): the kick is minus the gradient of
  V(x, y, z) = sum_i (sx_i x + sy_i y) sin(i k z + phase_i) / (i k) (finite differences), x, y, z are
  unchanged, the map is symplectic (its Jacobian J satisfies J^T S J = S), and it is periodic in z with
  period 2 pi / k.
"""
import math

import numpy as np
import pytest


def _beam(oracle, n=64, seed=3):
    c = oracle.generate_beam(n, seed, 1, (0.8, 0.072, 12.7), (0, 0, 0), (95e-6, 8.5e-6, 7e-3))
    c[5] *= 50.0     # larger momentum deviations exercise the square roots
    return c


def _jacobian(f, v, eps):
    J = np.zeros((6, 6))
    for k in range(6):
        vp, vm = v.copy(), v.copy()
        vp[k] += eps[k]
        vm[k] -= eps[k]
        J[:, k] = (f(vp) - f(vm)) / (2 * eps[k])
    return J


def _single(fn):
    def g(v):
        c = v.reshape(6, 1).copy()
        fn(c)
        return c[:, 0]
    return g


@pytest.mark.parametrize("angle", [0.0125, 0.0, -0.03])
def test_lorentz_boost_round_trip(oracle, angle):
    c0 = _beam(oracle)
    c = c0.copy()
    oracle.track_lorentz_boost(angle, c)
    if angle == 0.0:
        assert np.array_equal(c, c0)
    oracle.track_lorentz_boost(angle, c, reverse=True)
    scale = np.max(np.abs(c0), axis=1, keepdims=True)
    assert np.max(np.abs(c - c0) / scale) < 1e-12
    c = c0.copy()
    oracle.track_lorentz_boost(angle, c, reverse=True)
    oracle.track_lorentz_boost(angle, c)
    assert np.max(np.abs(c - c0) / scale) < 1e-12


def test_lorentz_boost_jacobian_and_first_order(oracle):
    angle = 0.0125
    c0 = _beam(oracle, n=8)
    eps = np.array([1e-7, 1e-8, 1e-8, 1e-8, 1e-6, 1e-7])
    for p in range(c0.shape[1]):
        v = c0[:, p].copy()
        Jf = _jacobian(_single(lambda c: oracle.track_lorentz_boost(angle, c)), v, eps)
        Jr = _jacobian(_single(lambda c: oracle.track_lorentz_boost(angle, c, reverse=True)), v, eps)
        assert math.isclose(np.linalg.det(Jf), 1.0 / math.cos(angle) ** 3, rel_tol=1e-6)
        assert math.isclose(np.linalg.det(Jr), math.cos(angle) ** 3, rel_tol=1e-6)
    # first order in a small angle on a particle at rest in the transverse plane
    a = 1e-6
    c = np.zeros((6, 1))
    c[4, 0] = 0.01
    oracle.track_lorentz_boost(a, c)
    assert math.isclose(c[0, 0], 0.01 * math.tan(a), rel_tol=1e-9)
    assert math.isclose(c[4, 0], 0.01 / math.cos(a), rel_tol=1e-12)


def test_crab_cavity_kick_is_gradient_and_symplectic(oracle):
    f = 394.0e6
    k = 2 * math.pi * f / oracle.CLIGHT
    sx, sy, ph = [1.3e-3, -2.0e-4], [5.0e-5, 3.0e-4], [0.3, -1.1]

    def V(x, y, z):
        return sum((sx[i] * x + sy[i] * y) * math.sin((i + 1) * k * z + ph[i]) / ((i + 1) * k) for i in range(2))

    c0 = _beam(oracle, n=16)
    c = c0.copy()
    oracle.track_crab_cavity(f, sx, sy, ph, c)
    for r in (0, 2, 4):
        assert np.array_equal(c[r], c0[r])
    for p in range(c0.shape[1]):
        x, y, z = c0[0, p], c0[2, p], c0[4, p]
        h = (1e-6, 1e-7, 1e-5)
        gx = (V(x + h[0], y, z) - V(x - h[0], y, z)) / (2 * h[0])
        gy = (V(x, y + h[1], z) - V(x, y - h[1], z)) / (2 * h[1])
        gz = (V(x, y, z + h[2]) - V(x, y, z - h[2])) / (2 * h[2])
        for r, g in ((1, gx), (3, gy), (5, gz)):
            assert abs((c[r, p] - c0[r, p]) + g) <= 1e-8 * max(abs(g), 1e-12)
    S = np.zeros((6, 6))
    for i in range(3):
        S[2 * i, 2 * i + 1], S[2 * i + 1, 2 * i] = 1.0, -1.0
    J = _jacobian(_single(lambda cc: oracle.track_crab_cavity(f, sx, sy, ph, cc)), c0[:, 0].copy(),
                  np.array([1e-6, 1e-6, 1e-7, 1e-6, 1e-5, 1e-6]))
    assert np.max(np.abs(J.T @ S @ J - S)) < 1e-6
    # periodic in z with the fundamental wavelength
    c1, c2 = c0.copy(), c0.copy()
    c2[4] += 2 * math.pi / k
    oracle.track_crab_cavity(f, sx, sy, ph, c1)
    oracle.track_crab_cavity(f, sx, sy, ph, c2)
    assert np.allclose(c1[[1, 3, 5]], c2[[1, 3, 5]], rtol=1e-8, atol=1e-15)
