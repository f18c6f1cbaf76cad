"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on identical inputs.

Bars (the reference's CPU/CUDA backend-consistency contract, src/contracts/Contracts.jl:831-1050):
- slice boundaries, centres, weights and member lists: bit-identical;
- cell indices (device stencil function) and the mesh geometry of every interaction whose two slices
  enter it with identical coordinates (first use in collision order): bit-identical;
- Green-cache decisions (hit / miss / rebuild) and cumulative counters: identical;
- coordinates: max|GPU - CPU| / max|CPU| <= 1e-10 per component, and the reference's per-component
  form |a - b| <= atol + rtol max(|a|, |b|) with atol = rtol = 1e-10;
- luminosity per pair and in total: rtol 1e-10.
Later pairs see coordinates already kicked by both implementations, which differ at round-off
(~1e-14): their meshes agree to 1e-12 relative instead of bitwise.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from octopic import (EMASS_EV, PMASS_EV, Beam, LongitudinalSlicing, PICCollider, PICPoissonSolver,  # noqa: E402
                     beam_params, linear6d_matrix)
from octopic.backend import DegenerateExtentError, NonFiniteCoordinateError, UnsupportedOptionError  # noqa: E402

PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
PP = beam_params(1.0, PMASS_EV, 275e9, 6.9e10)


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def round_beams(oracle, n1=20000, n2=8000, seed=0):
    e = oracle.generate_beam(n1, seed, 1, (1.0, 1.0, 20.0), (0, 0, 0), (100e-6, 100e-6, 0.01))
    p = oracle.generate_beam(n2, seed, 2, (1.0, 1.0, 90.909), (0, 0, 0), (100e-6, 100e-6, 0.06))
    return e, p


def flat_beams(oracle, n1=30000, n2=12000, seed=3):
    e = oracle.generate_beam(n1, seed, 1, (0.8, 0.072, 12.727), (0, 0, 0), (95e-6, 8.5e-6, 0.007))
    p = oracle.generate_beam(n2, seed, 2, (0.8, 0.072, 90.909), (0, 0, 0), (95e-6, 8.5e-6, 0.06))
    return e, p


def coords_close(a, b, tol=1e-10, floor=None):
    for k in range(6):
        scale = np.max(np.abs(a[k]))
        if floor is not None:
            scale = max(scale, floor[k])
        err = np.max(np.abs(a[k] - b[k]))
        assert err <= tol * scale, (k, err, scale)
        # the reference contract's per-component form, atol = rtol = 1e-10
        assert np.all(np.abs(a[k] - b[k]) <= 1e-10 + 1e-10 * np.maximum(np.abs(a[k]), np.abs(b[k])))


def first_use_pairs(order):
    seen_i, seen_j, first = set(), set(), set()
    for p, (i, j) in enumerate(order):
        if i not in seen_i and j not in seen_j:
            first.add(p)
        seen_i.add(i)
        seen_j.add(j)
    return first


def compare_collide(oracle, gpu, solver, e, p, collides=1, turn_map=None, floor=None):
    ce, cp = e.copy(), p.copy()
    ocol = oracle.Collider(solver)
    gcol = PICCollider(solver)
    be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    for c in range(collides):
        if c > 0:
            # re-seed the oracle from the GPU's turn-start state (identical inputs every collide)
            ce[:], cp[:] = be.to_host(), bp.to_host()
        lo = ocol.collide(ce, PE, cp, PP)
        lg = gcol.collide(be, bp, turn=c)
        torch.cuda.synchronize()
        # slicing: bit-identical
        for which, arr in ((1, ce), (2, cp)):
            n = arr.shape[1]
            bo, co, wo, no, io = ocol.last_slices(which, n)
            bg, cg, wg, ng, ig = gcol.last_slices(which, n)
            assert np.array_equal(bo, bg) and np.array_equal(co, cg) and np.array_equal(wo, wg)
            assert np.array_equal(no, ng) and np.array_equal(io, ig)
        # schedule, cache decisions, geometry
        oi, oj, orr, ol = ocol.pair_luminosity()
        gi, gj, gr, gl = gcol.pair_luminosity()
        assert list(oi) == list(gi) and list(oj) == list(gj) and list(orr) == list(gr)
        first = first_use_pairs(list(zip(oi, oj)))
        ro, rg = ocol.records(), gcol.interaction_records()
        assert len(ro) == len(rg)
        for a, b in zip(ro, rg):
            assert (a.p, a.i, a.j, a.dir, a.cache_event) == (b.p, b.i, b.j, b.dir, b.cache_event)
            ga = np.array(list(a.source_grid) + list(a.field_grid))
            gb = np.array(list(b.source_grid) + list(b.field_grid))
            if a.p in first:
                assert np.array_equal(ga, gb), (a.p, a.dir)
            else:
                assert np.allclose(ga, gb, rtol=1e-12, atol=0)
        ol, gl = np.array(ol), np.array(gl)
        assert np.max(np.abs(ol - gl)) <= 1e-10 * np.max(np.abs(ol))
        so, sg = ocol.stats(), gcol.stats()
        assert (so["hits"], so["misses"], so["rebuilds"]) == (sg["green_hits"], sg["green_misses"], sg["green_rebuilds"])
        assert so["dropped"] == sg["dropped"]
        assert math.isclose(lo, lg, rel_tol=1e-10)
        coords_close(ce, be.to_host(), floor=floor)
        coords_close(cp, bp.to_host(), floor=floor)
        if turn_map is not None:
            me, mp = turn_map
            gcol.track_linear6d(be, me)
            gcol.track_linear6d(bp, mp)
    return gcol


CASES = [
    dict(),                                                               # reference defaults (CIC, linear, 2nd)
    dict(deposit_method="TSC", field_derivative="fourth"),
    dict(deposit_method="TSC", field_derivative="fourth", slice_interpolation="quadratic"),
    dict(slice_interpolation="quadratic", green_type="standard", green_cache="none"),
    dict(green_cache="none", deposit_method="TSC"),
    dict(longitudinal_kick=False, luminosity_grid=(48, 40), luminosity_deposit_method="TSC"),
]


@pytest.mark.parametrize("kw", CASES, ids=[str(i) for i in range(len(CASES))])
def test_collide_parity_round(oracle, gpu, kw):
    e, p = round_beams(oracle)
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile", center_position="centroid")
    compare_collide(oracle, gpu, PICPoissonSolver(grid=(64, 64), slicing=sl, **kw), e, p)


@pytest.mark.parametrize("kw", [dict(deposit_method="TSC", field_derivative="fourth", slice_interpolation="quadratic"),
                                dict(grid=(96, 64))], ids=["quad128", "rect"])
def test_collide_parity_flat_multi_turn(oracle, gpu, kw):
    e, p = flat_beams(oracle)
    kw = dict(kw)
    grid = kw.pop("grid", (128, 128))
    sl = LongitudinalSlicing(nslices=7, method="normal_quantile", center_position="centroid")
    solver = PICPoissonSolver(grid=grid, slicing=sl, **kw)
    me = linear6d_matrix((0.8, 0.072, 12.727), [2 * math.pi * t for t in (0.08, 0.06, 0.01)])
    mp = linear6d_matrix((0.8, 0.072, 90.909), [2 * math.pi * t for t in (0.228, 0.21, 0.01)])
    gcol = compare_collide(oracle, gpu, solver, e, p, collides=3, turn_map=(me, mp))
    assert gcol.stats()["green_hits"] > 0


@pytest.mark.parametrize("kw", [dict(), dict(deposit_method="TSC", field_derivative="fourth"),
                                dict(longitudinal_kick=False, green_type="standard")], ids=["cic", "tsc4", "nolk"])
def test_node_mode_parity(oracle, gpu, kw):
    """interaction_grid = :node (pic_cpu.jl:1317-1538): node meshes at turn start, three solves per
    direction, transverse kick blended across node meshes, longitudinal kick on one mesh."""
    e, p = flat_beams(oracle, 24000, 10000, seed=5)
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile", center_position="centroid")
    solver = PICPoissonSolver(grid=(64, 64), slicing=sl, interaction_grid="node", **kw)
    me = linear6d_matrix((0.8, 0.072, 12.727), [2 * math.pi * t for t in (0.08, 0.06, 0.01)])
    mp = linear6d_matrix((0.8, 0.072, 90.909), [2 * math.pi * t for t in (0.228, 0.21, 0.01)])
    compare_collide(oracle, gpu, solver, e, p, collides=2, turn_map=(me, mp))


def test_standard_green_cache_divergence_is_a_half_cell_realign(oracle, gpu):
    """With green_type = :standard and the slice-pair cache, an expanded grid is re-aligned with
    t += t > 0 ? -1/4 : 1/4 (pic_cpu.jl:1877-1905).  When the expanded origin separation is a whole number
    of cells (the reference's k = 2 mod 5 case, pic_cpu.jl:1835-1846) t is rounding noise around 0, so
    1e-16 coordinate differences between two correct implementations can pick opposite half-cell shifts,
    after which the two runs legitimately differ at the discretisation level.  Every interaction up to the
    first divergence must match the oracle, and that first divergence must be exactly such a flip:
    origins shifted by +-h/2 in opposite directions along one axis, extents equal, same cache event."""
    e, p = round_beams(oracle)
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile", center_position="centroid")
    solver = PICPoissonSolver(grid=(64, 64), slicing=sl, green_type="standard")
    ocol = oracle.Collider(solver)
    gcol = PICCollider(solver)
    ce, cp = e.copy(), p.copy()
    ocol.collide(ce, PE, cp, PP)
    be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    gcol.collide(be, bp)
    oi, oj, _, _ = ocol.pair_luminosity()
    first = first_use_pairs(list(zip(oi, oj)))
    for a, b in zip(ocol.records(), gcol.interaction_records()):
        assert (a.p, a.dir, a.cache_event) == (b.p, b.dir, b.cache_event)
        sa, sb = np.array(a.source_grid), np.array(b.source_grid)
        fa, fb = np.array(a.field_grid), np.array(b.field_grid)
        if a.p in first:
            assert np.array_equal(sa, sb) and np.array_equal(fa, fb)
            continue
        if np.allclose(sa, sb, rtol=1e-12, atol=0) and np.allclose(fa, fb, rtol=1e-12, atol=0):
            continue
        # the first divergence: a half-cell realign flip on one axis
        assert np.allclose(sa[2:], sb[2:], rtol=1e-12, atol=0)
        flipped = 0
        for ax in (0, 1):
            h = sa[2 + ax] / 63.0
            ds, df = sb[ax] - sa[ax], fb[ax] - fa[ax]
            if abs(ds) <= 1e-12 * abs(sa[ax]) and abs(df) <= 1e-12 * abs(fa[ax]):
                continue
            assert math.isclose(abs(ds), 0.5 * h, rel_tol=1e-6) and math.isclose(ds, -df, rel_tol=1e-6)
            flipped += 1
        assert flipped == 1
        break


@pytest.mark.parametrize("method", ["equal_area", "equal_width", "specified"])
def test_slicing_methods_parity(oracle, gpu, method):
    e, p = round_beams(oracle, 12000, 6000, seed=4)
    kw = dict(nslices=4, method=method, center_position="centroid")
    if method == "specified":
        kw["positions"] = (0.8, -0.5, 0.0)
    if method == "equal_width":
        kw["center_position"] = "midpoint"
    sl = LongitudinalSlicing(**kw)
    compare_collide(oracle, gpu, PICPoissonSolver(grid=(32, 32), slicing=sl), e, p)


def test_stencil_cell_indices_bitwise(oracle, gpu):
    solver = PICPoissonSolver(grid=(16, 16))
    col = PICCollider(solver)
    rng = np.random.default_rng(20260805)
    for n in (16, 33):
        u = np.concatenate([np.linspace(0.0, n - 1.0, 257), [0.5, 1.5, 7.25, 7.5, 7.75, n - 1.5, n - 1.0],
                            1.0 + (n - 2.0) * rng.random(512), [-2.0, n + 5.0, np.nan, np.inf]])
        for method in ("CIC", "TSC"):
            base, w = col.debug_weights(method, u, n)
            k = 2 if method == "CIC" else 3
            for t, ut in enumerate(u):
                ob, ow = oracle.weights(method, ut, n)
                if sum(ow) == 0.0:
                    assert np.all(w[t, :k] == 0.0)
                    continue
                assert base[t] == ob - 1 and tuple(w[t, :k]) == ow, (method, ut)


def test_bitwise_reproducible(oracle, gpu):
    e, p = flat_beams(oracle, 40000, 16000)
    sl = LongitudinalSlicing(nslices=9, method="normal_quantile")
    solver = PICPoissonSolver(grid=(64, 64), slicing=sl, deposit_method="TSC")
    outs = []
    for _ in range(2):
        col = PICCollider(solver)
        be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
        lum = col.collide(be, bp)
        outs.append((lum, be.to_host(), bp.to_host()))
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])


@pytest.mark.parametrize("tiles", ["0", "4", "64"])
def test_in_slice_order_does_not_change_results(oracle, gpu, monkeypatch, tiles):
    """The transverse tile order inside slices is a layout choice only: any tiling (0 = slice order)
    gives bit-identical coordinates, luminosities, records and cache history to the default one."""
    e, p = round_beams(oracle, 30000, 12000)
    sl = LongitudinalSlicing(nslices=7, method="normal_quantile")
    solver = PICPoissonSolver(grid=(64, 64), slicing=sl, deposit_method="TSC", field_derivative="fourth",
                              slice_interpolation="quadratic")
    outs = []
    for env in (None, tiles):
        if env is None:
            monkeypatch.delenv("OCTOPIC_TRANSVERSE_TILES", raising=False)
        else:
            monkeypatch.setenv("OCTOPIC_TRANSVERSE_TILES", env)
        col = PICCollider(solver)
        be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
        lums = [col.collide(be, bp) for _ in range(2)]
        outs.append((lums, be.to_host(), bp.to_host(), col.pair_luminosity(), col.stats(),
                     [(r.p, r.dir, r.cache_event, tuple(r.source_grid), tuple(r.field_grid))
                      for r in col.interaction_records()]))
    monkeypatch.delenv("OCTOPIC_TRANSVERSE_TILES", raising=False)
    a, b = outs
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert np.array_equal(np.asarray(a[3]), np.asarray(b[3]))
    assert a[4] == b[4] and a[5] == b[5]


def test_nonfinite_fails_fast_and_leaves_beams(oracle, gpu):
    e, p = round_beams(oracle, 5000, 3000)
    e[2, 123] = np.nan
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile")
    col = PICCollider(PICPoissonSolver(grid=(32, 32), slicing=sl))
    be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    with pytest.raises(NonFiniteCoordinateError):
        col.collide(be, bp)
    assert np.array_equal(be.to_host(), e, equal_nan=True) and np.array_equal(bp.to_host(), p)
    # the handle stays usable after a failure
    e[2, 123] = 0.0
    be2 = Beam.from_host(e, PE, device=gpu)
    assert math.isfinite(col.collide(be2, bp))
    # a non-finite z poisons the slice boundaries: rejected at slicing
    e[4, 7] = np.inf
    with pytest.raises(NonFiniteCoordinateError):
        col.collide(Beam.from_host(e, PE, device=gpu), bp)


def test_degenerate_extent(oracle, gpu):
    e, p = round_beams(oracle, 2000, 2000)
    e[0, :] = 0.0; e[1, :] = 0.0; p[0, :] = 0.0; p[1, :] = 0.0
    sl = LongitudinalSlicing(nslices=2, method="normal_quantile")
    with pytest.raises(DegenerateExtentError):
        PICCollider(PICPoissonSolver(grid=(16, 16), slicing=sl)).collide(Beam.from_host(e, PE, device=gpu),
                                                                         Beam.from_host(p, PP, device=gpu))
    # x and px start at exactly 0: what the kick leaves there (~1e-20) is round-off of a field that vanishes by
    # symmetry, so that axis is compared against the transverse scale of the other axis
    compare_collide(oracle, gpu, PICPoissonSolver(grid=(16, 16), slicing=sl, min_transverse_extent=(1e-4, 0.0)), e, p,
                    floor=(1e-4, 1e-4, 0, 0, 0, 0))


def test_unsupported_combinations_are_rejected(gpu):
    for kw in (dict(interaction_grid="node", slice_interpolation="quadratic"), dict(grid_extent="sigma"),
               dict(interaction_grid="source_slice"), dict(green_type="lattice"),
               dict(slicing=LongitudinalSlicing(nslices=3, method="equal_count"))):
        with pytest.raises(UnsupportedOptionError):
            PICCollider(PICPoissonSolver(**kw))


def test_luminosity_schedule_skips(oracle, gpu):
    from octopic import EveryNSteps
    e, p = round_beams(oracle, 5000, 3000)
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile")
    solver = PICPoissonSolver(grid=(32, 32), slicing=sl, luminosity_schedule=EveryNSteps(step=2))
    col = PICCollider(solver)
    be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    assert math.isfinite(col.collide(be, bp, turn=0))
    assert math.isnan(col.collide(be, bp, turn=1))


def test_production_shape_properties(gpu):
    """At the full production shape (2.56M / 1.024M, 15 slices, 128^2) the result is bitwise reproducible,
    finite, and its luminosity is within a few percent of the analytic Gaussian overlap."""
    from octopic import generate_beam
    sl = LongitudinalSlicing(nslices=15, method="normal_quantile", center_position="centroid")
    solver = PICPoissonSolver(grid=(128, 128), deposit_method="TSC", field_derivative="fourth",
                              slice_interpolation="quadratic", slicing=sl)
    res = []
    for _ in range(2):
        col = PICCollider(solver)
        pe = beam_params(-1.0, EMASS_EV, 10e9, 1.7203e11)
        pp = beam_params(1.0, PMASS_EV, 275e9, 0.6881e11)
        e = generate_beam(col, 2_560_000, pe, (0.8, 0.072, 12.727), (0, 0, 0), (95e-6, 8.5e-6, 0.007), 0, 1, device=gpu)
        p = generate_beam(col, 1_024_000, pp, (0.8, 0.072, 90.909), (0, 0, 0), (95e-6, 8.5e-6, 0.06), 0, 2, device=gpu)
        lum = col.collide(e, p)
        res.append((lum, e.coords.clone(), p.coords.clone()))
    assert res[0][0] == res[1][0]
    assert torch.equal(res[0][1], res[1][1]) and torch.equal(res[0][2], res[1][2])
    assert torch.isfinite(res[0][1]).all() and torch.isfinite(res[0][2]).all()
    analytic = 1.7203e11 * 0.6881e11 / (2 * math.pi * math.sqrt(2 * 95e-6 ** 2 * 2 * 8.5e-6 ** 2))
    assert 0.9 < res[0][0] / analytic < 1.02     # hourglass reduces it below the thin-slice overlap


def test_reference_production_luminosity(gpu):
    """The reference's recorded production collide (profiling/benchmark_collide_cpu.jl:58-132,
    docs/history/cpu_benchmark_history.tsv:2): EIC pair at seed 123456789, 15 :normal_quantile centroid
    slices, 128^2 CIC :integrated, slice-pair cache 0.5 / 0.25, one persistent task, beams restored
    before each collide.  Luminosity 1.0254361754701657e+30; 450 cache misses, then 450 hits."""
    from octopic import generate_beam
    from test_oracle_production import ELE, PE, PP, PRO, REFERENCE_LUMINOSITY, SEED, production_solver
    col = PICCollider(production_solver())
    e = generate_beam(col, ELE["n"], PE, ELE["beta"], (0, 0, 0), ELE["sigma"], SEED, 1, device=gpu)
    p = generate_beam(col, PRO["n"], PP, PRO["beta"], (0, 0, 0), PRO["sigma"], SEED, 2, device=gpu)
    e0, p0 = e.coords.clone(), p.coords.clone()
    lums, stats = [], []
    for turn in range(2):
        e.coords.copy_(e0)
        p.coords.copy_(p0)
        lums.append(col.collide(e, p, turn=turn))
        stats.append(col.stats())
    assert (stats[0]["green_misses"], stats[0]["green_hits"], stats[0]["green_rebuilds"]) == (450, 0, 0)
    assert (stats[1]["green_misses"], stats[1]["green_hits"], stats[1]["green_rebuilds"]) == (450, 450, 0)
    for lum in lums:
        assert lum == pytest.approx(REFERENCE_LUMINOSITY, rel=1e-10)


def test_device_beam_generator_matches_oracle(oracle, gpu):
    """Beam(N, ...) on the device (Philox counter RNG, Box-Muller, +-5 sigma clamp, standardise, Twiss
    scale; Beam.jl:127-181, 545-553) against the oracle's restatement: same Philox integers, so every
    coordinate agrees to the last bits of the device/host libm and of the standardisation sums."""
    from octopic import generate_beam
    col = PICCollider(PICPoissonSolver(grid=(32, 32)))
    for n, rng_id, beta, sigma in ((200_000, 1, (0.55, 0.056, 12.727), (106e-6, 9.5e-6, 0.007)),
                                   (65_537, 2, (0.8, 0.072, 90.909), (95e-6, 8.5e-6, 0.06))):
        g = generate_beam(col, n, PE, beta, (0, 0, 0), sigma, 123456789, rng_id, device=gpu).to_host()
        o = oracle.generate_beam(n, 123456789, rng_id, beta, (0, 0, 0), sigma)
        for k in range(6):
            scale = np.max(np.abs(o[k]))
            assert np.max(np.abs(g[k] - o[k])) <= 1e-13 * scale, (n, k)
