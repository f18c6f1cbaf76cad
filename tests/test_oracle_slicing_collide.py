"""Oracle pins: slicing, collision schedule, luminosity, and whole-collide invariants.

- erfinv == scipy.special.erfinv; normal-quantile boundaries from the scipy quantiles;
- 4096-lane fold == plain left fold for n <= 4096 (test/runtests.jl:13183-13189);
- membership == brute-force linear scan; degenerate beams go to slice 1; :equal_count rank contract
  ([2,2,3] for seven co-located particles at ns = 3, slicing.jl:437-447);
- collision order == a stable sort by -(c1+c2)/2; batches conflict-free and order-preserving;
- luminosity overlap sums the full (n+1)^2 node extent (test/runtests.jl:10047-10111) and matches the
  analytic overlap of two Gaussians;
- a collide is bitwise invariant under the worker count (test/runtests.jl:4898-5170); options reach
  their consumers; the reference's rejected combinations are rejected.
"""
import math

import numpy as np
import pytest
from scipy import special

from octopic import (EMASS_EV, PMASS_EV, LongitudinalSlicing, PICPoissonSolver, beam_params, linear6d_matrix)


def test_erfinv_matches_scipy(oracle):
    ys = np.concatenate([np.linspace(-0.999, 0.999, 4001), [1e-300, -1e-20, 0.5, -0.9999999]])
    for y in ys:
        ref = special.erfinv(y)
        assert abs(oracle.erfinv(y) - ref) <= 4e-16 * max(abs(ref), 1e-300), y


def left_fold_mean(z):
    s = 0.0
    for v in z:
        s += float(v)
    return s / len(z)


def test_normal_quantile_boundaries_and_membership(oracle):
    rng = np.random.default_rng(5)
    z = rng.normal(0.0, 0.01, 4000)
    sl = LongitudinalSlicing(nslices=7, method="normal_quantile", center_position="centroid")
    b, c, w, lists = oracle.slices(z, sl)
    mu = left_fold_mean(z)                      # lane fold == left fold for n <= 4096
    s2 = 0.0
    for v in z:
        s2 += (float(v) - mu) * (float(v) - mu)
    sig = math.sqrt(s2 / len(z))
    assert b[0] == z.min() and b[-1] == z.max()
    for s in range(1, 7):
        q = math.sqrt(2.0) * special.erfinv(2 * s / 7 - 1)
        assert abs(b[s] - min(max(mu + sig * q, z.min()), z.max())) <= 2e-16 * 0.05
    # brute-force membership: clamp(#(b <= z), 1, ns)
    for s in range(7):
        assert np.all(np.diff(lists[s]) > 0)
    sid = np.empty(len(z), dtype=int)
    for s in range(7):
        sid[lists[s]] = s
    brute = np.clip(np.array([np.sum(b <= v) for v in z]), 1, 7) - 1
    assert np.array_equal(sid, brute)
    assert math.isclose(sum(w), 1.0)
    for s in range(7):
        assert c[s] == left_fold_mean(z[lists[s]])


def test_lane_fold_large_n_close_to_exact(oracle):
    rng = np.random.default_rng(6)
    z = rng.normal(0.003, 0.01, 100_000)
    b, c, w, lists = oracle.slices(z, LongitudinalSlicing(nslices=1, method="normal_quantile"))
    assert abs(c[0] - math.fsum(z) / len(z)) <= 1e-13 * 0.01


def test_degenerate_and_rank_contracts(oracle):
    z = np.full(7, 0.25)
    for method in ("normal_quantile", "equal_area", "equal_width"):
        b, c, w, lists = oracle.slices(z, LongitudinalSlicing(nslices=3, method=method))
        assert [len(l) for l in lists] == [7, 0, 0]
    b, c, w, lists = oracle.slices(z, LongitudinalSlicing(nslices=3, method="equal_count"))
    assert [len(l) for l in lists] == [2, 2, 3]
    b, c, w, lists = oracle.slices(np.array([0.1]), LongitudinalSlicing(nslices=3, method="equal_count"))
    assert [len(l) for l in lists] == [0, 0, 1]


def test_equal_area_and_width(oracle):
    rng = np.random.default_rng(8)
    z = rng.normal(0, 1, 50_000)
    b, c, w, lists = oracle.slices(z, LongitudinalSlicing(nslices=10, method="equal_area"))
    counts = np.array([len(l) for l in lists])
    assert np.all(np.diff(b) >= 0) and np.all(np.abs(counts - 5000) < 250)
    b, c, w, lists = oracle.slices(z, LongitudinalSlicing(nslices=10, method="equal_width", center_position="midpoint"))
    assert np.allclose(np.diff(b), (z.max() - z.min()) / 10, rtol=1e-12)
    assert np.allclose(c, 0.5 * (b[:-1] + b[1:]), rtol=0, atol=1e-15)
    b, c, w, lists = oracle.slices(z, LongitudinalSlicing(nslices=3, method="specified", positions=(1.0, -1.0)))
    assert b[1] < b[2] and abs(b[2] - (np.mean(z) + np.std(z))) < 1e-12


def test_collision_order_and_batches(oracle):
    rng = np.random.default_rng(9)
    c1 = np.sort(rng.normal(0, 0.007, 15))
    c2 = np.sort(rng.normal(0, 0.06, 15))
    c2[3] = -c1[5]            # a tie in time with another pair
    order = oracle.collision_order(c1, c2)
    brute = sorted([(-(c1[i] + c2[j]) / 2, i, j) for i in range(15) for j in range(15)], key=lambda t: t[0])
    assert order == [(i, j) for _, i, j in brute]
    batches = oracle.collision_pair_batches(c1, c2)
    flat = [p for bt in batches for p in bt]
    assert sorted(flat) == sorted(order)
    for bt in batches:
        assert len({i for i, _ in bt}) == len(bt) and len({j for _, j in bt}) == len(bt)
    pos = {p: k for k, p in enumerate(order)}
    seen = {}
    for bt in batches:
        for i, j in bt:
            for key in (("i", i), ("j", j)):
                assert seen.get(key, -1) < pos[(i, j)]
                seen[key] = pos[(i, j)]
    # equal centres: anti-diagonal wavefronts
    assert len(oracle.collision_pair_batches(np.zeros(4), np.zeros(5))) == 8


def test_luminosity_full_extent(oracle):
    n = 200
    x1 = np.array([1.0e-4 * (i - 1) / (n - 1) for i in range(1, n + 1)])
    y1 = np.array([1.0e-4 * ((i * 7) % n) / n for i in range(1, n + 1)])
    x2, y2 = x1.copy(), y1[::-1].copy()
    res = {}
    for method in ("CIC", "TSC"):
        s = PICPoissonSolver(kbb1=1e-6, kbb2=1e-6, grid=(16, 16), deposit_method=method)
        lum, q1, q2 = oracle.luminosity(s, x1, y1, x2, y2, 1.0)
        excluded = q1[16, :].sum() + q1[:, 16].sum() - q1[16, 16]
        res[method] = (lum, float((q1 * q2).sum()), excluded)
    assert res["CIC"][2] == 0.0
    assert res["TSC"][2] > 0.4
    assert math.isclose(res["CIC"][0] / res["CIC"][1], res["TSC"][0] / res["TSC"][1], rel_tol=1e-12)


def test_luminosity_matches_gaussian_overlap(oracle):
    rng = np.random.default_rng(12)
    n = 200_000
    s1, s2 = (100e-6, 30e-6), (80e-6, 40e-6)
    x1, y1 = rng.normal(0, s1[0], n), rng.normal(0, s1[1], n)
    x2, y2 = rng.normal(0, s2[0], n), rng.normal(0, s2[1], n)
    s = PICPoissonSolver(grid=(128, 128))
    lum, _, _ = oracle.luminosity(s, x1, y1, x2, y2, 1.0 / (n * n))
    ref = 1.0 / (2 * math.pi * math.sqrt((s1[0] ** 2 + s2[0] ** 2) * (s1[1] ** 2 + s2[1] ** 2)))
    assert abs(lum / ref - 1) < 0.01


# ---------------------------------------------------------------- whole collide
E_PARAMS = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
P_PARAMS = beam_params(1.0, PMASS_EV, 275e9, 0.7e11)


def beams(oracle, n1=6000, n2=6000, seed=77):
    e = oracle.generate_beam(n1, seed, 1, (0.55, 0.056, 12.0), (0, 0, 0), (106e-6, 9.5e-6, 7e-3))
    p = oracle.generate_beam(n2, seed, 2, (0.8, 0.072, 90.0), (0, 0, 0), (95e-6, 8.5e-6, 6e-2))
    return e, p


def run(oracle, nthreads=1, seed=77, **kw):
    sl = LongitudinalSlicing(nslices=kw.pop("nslices", 3), method="normal_quantile", center_position="centroid")
    s = PICPoissonSolver(slicing=sl, grid=kw.pop("grid", (64, 64)), **kw)
    e, p = beams(oracle, seed=seed)
    col = oracle.Collider(s)
    lum = col.collide(e, E_PARAMS, p, P_PARAMS, nthreads=nthreads)
    return lum, e, p, col


def test_worker_count_invariance(oracle):
    l1, e1, p1, _ = run(oracle, nthreads=1, nslices=5)
    l4, e4, p4, c4 = run(oracle, nthreads=4, nslices=5)
    assert c4.stats()["batches"] > 0          # the batched schedule actually ran
    assert l1 == l4 and np.array_equal(e1, e4) and np.array_equal(p1, p4)


def test_options_reach_consumers(oracle):
    l_def, e_def, p_def, _ = run(oracle)
    l_2, e_2, p_2, _ = run(oracle, field_derivative="second")
    l_4, e_4, p_4, _ = run(oracle, field_derivative="fourth")
    l_q, e_q, p_q, _ = run(oracle, slice_interpolation="quadratic")
    assert l_def == l_2 and np.array_equal(e_def, e_2)
    assert not np.array_equal(e_def, e_4) and not np.array_equal(e_def, e_q)
    assert math.isclose(l_4, l_def, rel_tol=1e-3) and math.isclose(l_q, l_def, rel_tol=1e-3)
    assert np.all(np.isfinite(e_q)) and np.all(np.isfinite(p_q))


def test_quadratic_interpolation_identities(oracle):
    s = PICPoissonSolver(grid=(8, 8), deposit_method="CIC")
    sg, fg = oracle.interaction_grids(s, [-1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0])
    I, J = np.meshgrid(np.arange(1, 9), np.arange(1, 9), indexing="ij")
    rv = lambda k: np.sin(k * I + 0.3 * J)
    L, M, R = (rv(1.0), rv(1.1), rv(1.2)), (rv(2.0), rv(2.1), rv(2.2)), (rv(3.0), rv(3.1), rv(3.2))
    quad = lambda t, L=L, M=M, R=R: oracle.interpolate_kick_quadratic(s, fg, 0.13, -0.07, L, M, R, t)
    lin = lambda zL: oracle.interpolate_kick(s, fg, 0.13, -0.07, L, R, zL, 1.0 - zL)
    assert np.allclose(quad(0.0)[:2], lin(1.0)[:2], rtol=1e-12)
    assert np.allclose(quad(1.0)[:2], lin(0.0)[:2], rtol=1e-12)
    assert math.isclose(quad(0.5)[2], lin(0.5)[2], rel_tol=1e-12)
    flat = np.ones((8, 8))
    F = (flat, flat, flat)
    for t in (0.0, 0.25, 0.5, 0.75, 1.0):
        kx, ky, _ = oracle.interpolate_kick_quadratic(s, fg, 0.13, -0.07, F, F, F, t)
        assert math.isclose(kx, 1.0, rel_tol=1e-12) and math.isclose(ky, 1.0, rel_tol=1e-12)
        _, _, kz0 = quad(t)
        sh = lambda P: (P[0] + 17.0, P[1], P[2])
        _, _, kz1 = oracle.interpolate_kick_quadratic(s, fg, 0.13, -0.07, sh(L), sh(M), sh(R), t)
        assert abs(kz0 - kz1) < 1e-12


def test_rejected_combinations(oracle):
    with pytest.raises(oracle.OracleError) as ei:
        oracle.Collider(PICPoissonSolver(interaction_grid="node", slice_interpolation="quadratic"))
    assert ei.value.code == 2


def test_green_cache_history_and_luminosity(oracle):
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile", center_position="centroid")
    s = PICPoissonSolver(slicing=sl, grid=(64, 64))
    e = oracle.generate_beam(100_000, 0, 1, (1.0, 1.0, 20.0), (0, 0, 0), (100e-6, 100e-6, 0.01))
    p = oracle.generate_beam(40_000, 0, 2, (1.0, 1.0, 90.909), (0, 0, 0), (100e-6, 100e-6, 0.06))
    pe, pp = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11), beam_params(1.0, PMASS_EV, 275e9, 6.9e10)
    col = oracle.Collider(s)
    lum = col.collide(e, pe, p, pp, nthreads=4)
    st = col.stats()
    assert st["misses"] == 50 and st["hits"] == 0
    lum2 = col.collide(e, pe, p, pp, nthreads=4)
    st2 = col.stats()
    assert st2["hits"] + st2["misses"] + st2["rebuilds"] == 100 and st2["hits"] > 0
    analytic = 1.7e11 * 6.9e10 / (2 * math.pi * math.sqrt((2 * 100e-6 ** 2) * (2 * 100e-6 ** 2)))
    assert abs(lum / analytic - 1) < 0.02 and abs(lum2 / analytic - 1) < 0.02


def test_node_mode_runs_and_agrees(oracle):
    l_sp, e_sp, p_sp, _ = run(oracle, nslices=5)
    l_nd, e_nd, p_nd, c = run(oracle, nslices=5, interaction_grid="node")
    assert np.all(np.isfinite(e_nd)) and math.isclose(l_nd, l_sp, rel_tol=2e-3)
    assert not np.array_equal(e_nd, e_sp)


def test_nonfinite_fails_fast_and_leaves_beams(oracle):
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile")
    s = PICPoissonSolver(slicing=sl, grid=(32, 32))
    e, p = beams(oracle)
    e[0, 17] = np.nan
    before = e.copy(), p.copy()
    with pytest.raises(oracle.OracleError) as ei:
        oracle.Collider(s).collide(e, E_PARAMS, p, P_PARAMS)
    assert ei.value.code == 3
    assert np.array_equal(e, before[0], equal_nan=True) and np.array_equal(p, before[1])


def test_degenerate_extent_and_floor(oracle):
    sl = LongitudinalSlicing(nslices=2, method="normal_quantile")
    e, p = beams(oracle, 500, 500)
    e[0, :] = 0.0; e[1, :] = 0.0
    p[0, :] = 0.0; p[1, :] = 0.0
    with pytest.raises(oracle.OracleError) as ei:
        oracle.Collider(PICPoissonSolver(slicing=sl, grid=(16, 16))).collide(e.copy(), E_PARAMS, p.copy(), P_PARAMS)
    assert ei.value.code == 4
    lum = oracle.Collider(PICPoissonSolver(slicing=sl, grid=(16, 16), min_transverse_extent=(1e-4, 0.0))).collide(
        e, E_PARAMS, p, P_PARAMS)
    assert math.isfinite(lum)


def test_sigma_extent_counts_dropped(oracle):
    # forcing an under-covering estimator from the host: particles fall outside and are counted
    _, _, _, col = run(oracle, grid_extent="sigma", grid_extent_sigma=1.0)
    assert col.stats()["dropped"] > 0
    _, _, _, col = run(oracle)
    assert col.stats()["dropped"] == 0


def test_linear6d_is_symplectic_rotation(oracle):
    m = np.array(linear6d_matrix((0.8, 0.072, 90.9), [2 * math.pi * t for t in (0.228, 0.21, 0.01)])).reshape(6, 6)
    J = np.zeros((6, 6))
    for k in range(3):
        J[2 * k, 2 * k + 1], J[2 * k + 1, 2 * k] = 1.0, -1.0
    assert np.allclose(m.T @ J @ m, J, atol=1e-12)
    for k, tune in enumerate((0.228, 0.21, 0.01)):
        blk = m[2 * k:2 * k + 2, 2 * k:2 * k + 2]
        assert math.isclose(np.trace(blk) / 2, math.cos(2 * math.pi * tune), rel_tol=1e-12)
    mo = oracle.linear6d_matrix((0.8, 0.072, 90.9), (0.8, 0.072, 90.9), (0, 0, 0), (0, 0, 0),
                                [2 * math.pi * t for t in (0.228, 0.21, 0.01)])
    assert np.array_equal(mo, m.ravel())
