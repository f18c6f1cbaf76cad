"""Oracle pins: counter RNG and the CIC/TSC deposition stencils.

Known answers come from the reference's own tests and from the mathematics:
- Philox4x32-10 Random123 KAT vectors (src/math/counter_rng.jl:250-261);
- CIC/TSC known answers (test/runtests.jl:9923-9958);
- the TSC weights equal the cardinal quadratic B-spline evaluated at the node offsets, and the CIC
  weights the linear (hat) B-spline, an independent definition of the same stencils.
"""
import math

import numpy as np
import pytest


def test_philox_kat(oracle):
    # Random123 kat_vectors for philox4x32-10
    assert oracle.philox_block((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert oracle.philox_block((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert oracle.philox_block((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == (
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_counter_stream_properties(oracle):
    # uniform_open01 keeps 52 bits plus a half-bin offset: strictly inside (0, 1)
    vals = [oracle.counter_uint64(123456789, 0, 1, i, 1) for i in range(1, 2001)]
    u = np.array([((v >> 12) + 0.5) * 2.0 ** -52 for v in vals])
    assert np.all(u > 0.0) and np.all(u < 1.0)
    assert abs(u.mean() - 0.5) < 0.02
    # components 1 and 2 are the two halves of one Box-Muller pair: n1^2 + n2^2 = -2 ln(u1)
    for i in range(1, 50):
        n1 = oracle.octopus_normal(7, 0, 3, i, 1)
        n2 = oracle.octopus_normal(7, 0, 3, i, 2)
        u1 = ((oracle.counter_uint64(7, 0, 3, i, 1) >> 12) + 0.5) * 2.0 ** -52
        assert math.isclose(n1 * n1 + n2 * n2, -2.0 * math.log(u1), rel_tol=1e-13)
    # a different rng_id or seed gives a different stream
    assert oracle.counter_uint64(1, 0, 1, 5, 1) != oracle.counter_uint64(1, 0, 2, 5, 1)
    assert oracle.counter_uint64(1, 0, 1, 5, 1) != oracle.counter_uint64(2, 0, 1, 5, 1)


def test_generated_beam_moments(oracle):
    beta, alpha, sigma = (0.8, 0.072, 90.909), (0.0, 0.0, 0.0), (95e-6, 8.5e-6, 0.06)
    c = oracle.generate_beam(20000, 0, 2, beta, alpha, sigma, cutoff=5.0)
    for k in range(3):
        q, p = c[2 * k], c[2 * k + 1]
        assert abs(q.mean()) < 1e-12 * sigma[k] * 1e3
        assert math.isclose(np.sqrt(np.mean(q * q)), sigma[k], rel_tol=1e-12)
        assert math.isclose(np.sqrt(np.mean(p * p)), sigma[k] / beta[k], rel_tol=1e-12)
        assert np.max(np.abs(q)) <= 5.5 * sigma[k]
    c2 = oracle.generate_beam(20000, 0, 2, beta, alpha, sigma, cutoff=5.0)
    assert np.array_equal(c, c2)


# ---------------------------------------------------------------- stencils
def bspline1(t):
    t = abs(t)
    return 1.0 - t if t < 1.0 else 0.0


def bspline2(t):
    t = abs(t)
    if t <= 0.5:
        return 0.75 - t * t
    if t <= 1.5:
        return 0.5 * (1.5 - t) ** 2
    return 0.0


def test_weights_known_answers(oracle):
    for u in (-2.0, 70.0, float("nan"), float("inf"), float("-inf")):
        assert sum(oracle.weights("CIC", u, 64)[1]) == 0.0
        assert sum(oracle.weights("TSC", u, 64)[1]) == 0.0
    assert math.isclose(sum(oracle.weights("CIC", 5.3, 64)[1]), 1.0)
    assert math.isclose(sum(oracle.weights("TSC", 5.3, 64)[1]), 1.0)
    for n in (5, 16, 64):
        assert oracle.weights("CIC", float(n - 1), n) == (n - 1, (0.0, 1.0))
    assert oracle.weights("CIC", 2.25, 5) == (3, (0.75, 0.25))


def test_weights_are_bsplines(oracle):
    rng = np.random.default_rng(20260805)
    n = 33
    us = np.concatenate([np.linspace(1.0, n - 2.5, 257), 1.0 + (n - 3.5) * rng.random(512)])
    for u in us:
        base, w = oracle.weights("CIC", u, n)
        for m, wm in enumerate(w):
            node = base - 1 + m
            assert abs(wm - bspline1(node - u)) < 1e-15
        base, w = oracle.weights("TSC", u, n)
        for m, wm in enumerate(w):
            node = base - 1 + m
            assert abs(wm - bspline2(node - u)) < 4e-16
        # first moment reproduces the particle position (charge conservation of the dipole)
        assert abs(sum(wm * (base - 1 + m) for m, wm in enumerate(w)) - u) < 1e-13


def test_tsc_edge_stencil_is_clamped(oracle):
    # inside [0, 1/2) the stencil is displaced one cell (documented behaviour, pic_cpu.jl:2202-2224)
    base, w = oracle.weights("TSC", 0.2, 16)
    assert base == 1 and math.isclose(sum(w), 1.0)


def test_deposit_drops_nan_and_keeps_edge_charge(oracle):
    q = oracle.deposit("CIC", [-5.0e-4, float("nan"), -6.0e-4], [-5.0e-4, -5.0e-4, float("nan")],
                       -1.0e-3, -1.0e-3, 1.0e-5, 1.0e-5, 64, 64, ld=128, ncols=128)
    assert np.all(np.isfinite(q))
    assert abs(q.sum() - 1.0) < 1e-12
    q = oracle.deposit("CIC", [4.0], [4.0], 0.0, 0.0, 1.0, 1.0, 5, 5)
    assert q.sum() == 1.0 and q[4, 4] == 1.0
