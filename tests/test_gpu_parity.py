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


def compare_collide(oracle, gpu, solver, e, p, collides=1, turn_map=None, floor=None, first_use_exact=True):
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
            if a.p in first and first_use_exact:
                assert np.array_equal(ga, gb), (a.p, a.dir)
            else:
                # origins are differences of nearly equal coordinates and can sit near zero: the bar is
                # relative to the meshes' extent
                assert np.allclose(ga, gb, rtol=1e-12, atol=1e-12 * np.max(np.abs(ga))), (a.p, a.dir)
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


@pytest.mark.parametrize("method", ["equal_area", "equal_width", "specified", "equal_count"])
def test_slicing_methods_parity(oracle, gpu, method):
    e, p = round_beams(oracle, 12000, 6000, seed=4)
    kw = dict(nslices=4, method=method, center_position="centroid")
    if method == "specified":
        kw["positions"] = (0.8, -0.5, 0.0)
    if method == "equal_width":
        kw["center_position"] = "midpoint"
    sl = LongitudinalSlicing(**kw)
    compare_collide(oracle, gpu, PICPoissonSolver(grid=(32, 32), slicing=sl), e, p)


def test_equal_count_ties_follow_rank_order(oracle, gpu):
    """:equal_count (slicing.jl:330-365) splits by rank, not by boundary comparison: with z quantised to
    a few values (ties everywhere, and a -0.0 / +0.0 pair) the populations stay equal and each slice's
    members are listed in stable rank order (ascending index within ties), bit-identical to the
    oracle's stable sortperm; the whole collide then matches."""
    e, p = round_beams(oracle, 9001, 4003, seed=11)
    e[4] = np.round(e[4] / 0.004) * 0.004
    e[4, 5], e[4, 6] = -0.0, 0.0
    p[4] = np.round(p[4] / 0.02) * 0.02
    sl = LongitudinalSlicing(nslices=6, method="equal_count", center_position="centroid")
    gcol = compare_collide(oracle, gpu, PICPoissonSolver(grid=(32, 32), slicing=sl, deposit_method="TSC"), e, p)
    b, c, w, cnt, idx = gcol.last_slices(1, e.shape[1])
    assert np.all(np.abs(np.diff(cnt)) <= 1)
    # Julia's isless order: -0.0 before +0.0, then ascending index within ties
    order = np.lexsort((np.arange(e.shape[1]), ~np.signbit(e[4]), e[4]))
    assert np.array_equal(np.asarray(idx), order)


def test_empty_slices_skip_their_pairs(oracle, gpu):
    """`:specified` boundaries at z = -9 m and -8.5 m leave the first two slices of both beams empty:
    every pair with an empty slice is skipped (pic_cpu.jl:302), the others match the oracle."""
    e, p = round_beams(oracle, 12000, 6000, seed=4)
    sl = LongitudinalSlicing(nslices=4, method="specified", center_position="centroid", positions=(-9.0, -8.5, 0.0))
    gcol = compare_collide(oracle, gpu, PICPoissonSolver(grid=(32, 32), slicing=sl, deposit_method="TSC"), e, p)
    _, _, ran, _ = gcol.pair_luminosity()
    assert 0 < sum(ran) < len(ran)


def test_tiny_ragged_beams(oracle, gpu):
    """Beams far smaller than one particle chunk, with sizes that fill no block evenly."""
    e, p = round_beams(oracle, 61, 23, seed=9)
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile", center_position="centroid")
    compare_collide(oracle, gpu, PICPoissonSolver(grid=(16, 16), slicing=sl, deposit_method="TSC",
                                                  field_derivative="fourth", slice_interpolation="quadratic"), e, p)


@pytest.mark.parametrize("interaction_grid", ["slice_pair", "node"])
def test_large_mesh_parity(oracle, gpu, interaction_grid):
    """Meshes above 143 x 143 nodes (here 192 x 162, luminosity 160 x 150): deposit windows that fit
    160 KB of LDS stay there, larger ones (the whole-mesh windows of the node mode) add straight into
    HBM; the charge is the same integer sum either way."""
    e, p = round_beams(oracle, 12000, 5000, seed=6)
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile", center_position="centroid")
    kw = dict(deposit_method="TSC", field_derivative="fourth")
    if interaction_grid == "slice_pair":
        kw["slice_interpolation"] = "quadratic"
    solver = PICPoissonSolver(grid=(192, 162), slicing=sl, interaction_grid=interaction_grid,
                              luminosity_grid=(160, 150), **kw)
    compare_collide(oracle, gpu, solver, e, p)


@pytest.mark.parametrize("grid", [(80, 80), (100, 120)], ids=["80", "100x120"])
def test_radix5_meshes_parity(oracle, gpu, grid):
    """Padded sizes 2g with factors of 5 (160 = 2^5 5, 200 = 2^3 5^2, 240 = 2^4 3 5) run the radix-5
    stage of the LDS transforms."""
    e, p = round_beams(oracle, 15000, 6000, seed=8)
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile", center_position="centroid")
    compare_collide(oracle, gpu, PICPoissonSolver(grid=grid, slicing=sl, deposit_method="TSC", field_derivative="fourth",
                                                  slice_interpolation="quadratic"), e, p)


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
        st = col.stats()
        st.pop("deposit_all_planes")   # a launch-shape figure, not a result
        outs.append((lums, be.to_host(), bp.to_host(), col.pair_luminosity(), st,
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
               dict(grid=(70, 70)), dict(grid=(576, 576))):   # 2*70 has a factor 7; 576 > 512 nodes
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


@pytest.mark.parametrize("kw", [dict(deposit_method="TSC", field_derivative="fourth", slice_interpolation="quadratic"),
                                dict(interaction_grid="node", deposit_method="TSC", field_derivative="fourth"),
                                dict()], ids=["quad", "node", "defaults"])
def test_fp32_form_against_fp64_oracle(oracle, gpu, kw):
    """precision = fp32 (deposit weights, FFT convolution, field table and kick gather in FP32 on FP64
    beams; Green function, moments and luminosity in FP64): the per-particle transverse kick (change of
    px, py over one collide) stays within 1e-4 of the largest FP64 kick and the longitudinal one (pz,
    from potential differences) within 3e-3, the luminosity within 1e-5, the slicing (FP64) is
    bit-identical, and the result is bitwise reproducible."""
    e, p = flat_beams(oracle, 40000, 16000, seed=11)
    sl = LongitudinalSlicing(nslices=7, method="normal_quantile", center_position="centroid")
    s64 = PICPoissonSolver(grid=(128, 128), slicing=sl, **kw)
    s32 = PICPoissonSolver(grid=(128, 128), slicing=sl, precision="fp32", **kw)
    ce, cp = e.copy(), p.copy()
    lo = oracle.Collider(s64).collide(ce, PE, cp, PP)
    outs = []
    for _ in range(2):
        g = PICCollider(s32)
        be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
        lg = g.collide(be, bp)
        outs.append((lg, be.to_host(), bp.to_host(), g.last_slices(1, e.shape[1])))
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])
    lg, ge, gp, gsl = outs[0]
    assert math.isclose(lg, lo, rel_tol=1e-5)
    osl = oracle.Collider(s64)
    osl.collide(e.copy(), PE, p.copy(), PP)
    for a, b in zip(osl.last_slices(1, e.shape[1]), gsl):
        assert np.array_equal(a, b)
    for init, ref, got in ((e, ce, ge), (p, cp, gp)):
        for k, tol in ((1, 1e-4), (3, 1e-4), (5, 3e-3)):
            dref, dgot = ref[k] - init[k], got[k] - init[k]
            assert np.max(np.abs(dgot - dref)) <= tol * np.max(np.abs(dref)), k
        for k in (0, 2):
            assert np.max(np.abs(got[k] - ref[k])) <= 1e-4 * np.max(np.abs(ref[k] - init[k]) + 1e-30) + 1e-12 * np.max(np.abs(ref[k]))


def _be_kick(sx, sy, x, y):
    from test_oracle_solver import be_kick
    return be_kick(sx, sy, x, y)


def _gpu_field_error(oracle, gpu, sx, sy, precision, npts=160, grid=128):
    """Median |kick - exact Bassetti-Erskine kick| / max |BE kick| on a 41 x 41 probe grid within
    +-2 sigma, for a deterministic quantile Gaussian source (npts^2 particles, one slice, no drifts),
    through the whole device collide: deposit, convolution, 4th-order field, TSC gather, kick."""
    u = (np.arange(1, npts + 1) - 0.5) / npts
    q = np.array([oracle.erfinv(2 * v - 1) for v in u]) * math.sqrt(2.0)
    X, Y = np.meshgrid(sx * q, sy * q, indexing="ij")
    src = np.zeros((6, X.size))
    src[0], src[2] = X.ravel(), Y.ravel()
    xg, yg = np.linspace(-2 * sx, 2 * sx, 41), np.linspace(-2 * sy, 2 * sy, 41)
    PX, PY = np.meshgrid(xg, yg, indexing="ij")
    probes = np.zeros((6, PX.size))
    probes[0], probes[2] = PX.ravel(), PY.ravel()
    sl = LongitudinalSlicing(nslices=1, method="normal_quantile", center_position="centroid")
    solver = PICPoissonSolver(grid=(grid, grid), deposit_method="TSC", field_derivative="fourth", kbb1=1.0, kbb2=1.0,
                              longitudinal_kick=False, slicing=sl, precision=precision)
    col = PICCollider(solver)
    b1, b2 = Beam.from_host(src, PE, device=gpu), Beam.from_host(probes, PP, device=gpu)
    col.collide(b1, b2)
    out = b2.to_host()
    kx, ky = _be_kick(sx, sy, PX.ravel(), PY.ravel())
    err = np.hypot(out[1] - kx, out[3] - ky)
    return float(np.median(err) / np.max(np.hypot(kx, ky))), out


@pytest.mark.parametrize("aspect", [1.0, 4.0, 11.0, 25.0])
def test_field_vs_bassetti_erskine_both_precisions(oracle, gpu, aspect):
    """The FP64 and FP32 device fields against the exact Bassetti-Erskine field at sigma_x/sigma_y =
    1, 4, 11, 25 (sigma_y = 8.5 um, 128^2, TSC, 4th order): the FP64 error is the discretisation error
    (the reference quotes medians of 4.1e-4 round to 1.2e-3 at 25:1 at 128^2), and the FP32 form adds
    at most 1e-5 of the largest kick to it."""
    sy = 8.5e-6
    e64, k64 = _gpu_field_error(oracle, gpu, aspect * sy, sy, "fp64")
    e32, k32 = _gpu_field_error(oracle, gpu, aspect * sy, sy, "fp32")
    assert e64 < 3e-3
    assert abs(e32 - e64) <= 1e-5
    scale = np.max(np.hypot(k64[1], k64[3]))
    assert np.max(np.hypot(k32[1] - k64[1], k32[3] - k64[3])) <= 1e-4 * scale


def test_reference_cpu_configuration_ten_turns(oracle, gpu):
    """The reference's CPU-path configuration: round beams (sigma 100 um, sigma_p 1e-4, sigma_z 1 cm /
    6 cm, sigma_dp 5e-4 / 6.6e-4) of 100k e- and 40k p from seed 0, 1.7e11 / 6.9e10, 10 / 275 GeV;
    5 slices, 64^2, CIC, integrated Green, 2nd order, linear interpolation, slice-pair cache; 10 turns
    through the linear 6D maps with tunes (0.08, 0.06, 0.01) and (0.228, 0.21, 0.01).  Every turn the
    oracle starts from the GPU's turn-start state: slicing bit-identical, coordinates and luminosity
    within 1e-10, cache history identical."""
    e = oracle.generate_beam(100_000, 0, 1, (1.0, 1.0, 20.0), (0, 0, 0), (100e-6, 100e-6, 0.01))
    p = oracle.generate_beam(40_000, 0, 2, (1.0, 1.0, 90.909), (0, 0, 0), (100e-6, 100e-6, 0.06))
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile", center_position="centroid")
    solver = PICPoissonSolver(grid=(64, 64), deposit_method="CIC", green_type="integrated",
                              field_derivative="second", slice_interpolation="linear", slicing=sl)
    me = linear6d_matrix((1.0, 1.0, 20.0), [2 * math.pi * t for t in (0.08, 0.06, 0.01)])
    mp = linear6d_matrix((1.0, 1.0, 90.909), [2 * math.pi * t for t in (0.228, 0.21, 0.01)])
    gcol = compare_collide(oracle, gpu, solver, e, p, collides=10, turn_map=(me, mp))
    st = gcol.stats()
    assert st["green_misses"] == 50 and st["green_hits"] + st["green_rebuilds"] == 9 * 50


def test_collide_many_equals_one_by_one(oracle, gpu):
    """Independent instances of uneven shape (slices 7/11/15, meshes 64/96/128, FP64 and FP32) collided
    together through octopic_collide_many give bit-identical results to the same instances collided one
    by one, over two turns."""
    from octopic import collide_many
    shapes = [(30000, 12000, 7, 64, "fp64"), (50000, 9000, 11, 96, "fp32"), (20000, 16000, 15, 128, "fp64"),
              (41000, 7000, 7, 96, "fp32")]
    insts = []
    for k, (n1, n2, ns, g, prec) in enumerate(shapes):
        e, p = flat_beams(oracle, n1, n2, seed=20 + k)
        sl = LongitudinalSlicing(nslices=ns, method="normal_quantile", center_position="centroid")
        solver = PICPoissonSolver(grid=(g, g), slicing=sl, deposit_method="TSC", field_derivative="fourth",
                                  slice_interpolation="quadratic", precision=prec)
        insts.append((solver, e, p))
    results = []
    for together in (False, True):
        cols = [PICCollider(s) for s, _, _ in insts]
        b1 = [Beam.from_host(e, PE, device=gpu) for _, e, _ in insts]
        b2 = [Beam.from_host(p, PP, device=gpu) for _, _, p in insts]
        lums = []
        for turn in range(2):
            if together:
                lums.append(collide_many(cols, b1, b2, [turn] * len(cols)))
            else:
                lums.append([c.collide(x, y, turn=turn) for c, x, y in zip(cols, b1, b2)])
        torch.cuda.synchronize()
        results.append((lums, [b.to_host() for b in b1], [b.to_host() for b in b2], [c.stats() for c in cols]))
    a, b = results
    assert a[0] == b[0] and a[3] == b[3]
    for x, y in zip(a[1] + a[2], b[1] + b[2]):
        assert np.array_equal(x, y)


def test_turn_elements_match_oracle(oracle, gpu):
    """LorentzBoost / RevLorentzBoost and ThinCrabCavity on the device against the oracle on the same
    coordinates: agreement to the last bits of the device and host sin/cos/sqrt."""
    e = oracle.generate_beam(50000, 4, 1, (0.8, 0.072, 12.7), (0, 0, 0), (95e-6, 8.5e-6, 7e-3))
    e[5] *= 50.0
    col = PICCollider(PICPoissonSolver(grid=(32, 32)))
    steps = [("boost", 0.0125), ("crab", (394.0e6, [1.3e-3, -2e-4], [5e-5, 3e-4], [0.3, -1.1])), ("rev", 0.0125)]
    b = Beam.from_host(e, PE, device=gpu)
    ref = e.copy()
    for kind, arg in steps:
        if kind == "boost":
            col.track_lorentz_boost(b, arg)
            oracle.track_lorentz_boost(arg, ref)
        elif kind == "rev":
            col.track_lorentz_boost(b, arg, reverse=True)
            oracle.track_lorentz_boost(arg, ref, reverse=True)
        else:
            f, sx, sy, ph = arg
            col.track_crab_cavity(b, f, sx, sy, ph)
            oracle.track_crab_cavity(f, sx, sy, ph, ref)
        torch.cuda.synchronize()
        got = b.to_host()
        for k in range(6):
            assert np.max(np.abs(got[k] - ref[k])) <= 1e-13 * np.max(np.abs(ref[k])), (kind, k)


def test_autotune_keeps_results_and_cache(oracle, gpu):
    """The autotuner picks launch shapes from its candidate sets on a scratch handle: the tuned
    handle's Green-cache counters and the beams are untouched by tuning, and its next collide is
    bit-identical to a collide with the default shapes."""
    solver = PICPoissonSolver(grid=(64, 64), deposit_method="TSC", field_derivative="fourth",
                              slicing=LongitudinalSlicing(5, method="normal_quantile"))
    e, p = flat_beams(oracle)
    tuned = PICCollider(solver)
    be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    shape = tuned.autotune(be, bp, trials=1)
    assert shape["transverse_tiles"] in (8, 16, 32) and shape["kick_target_blocks"] in (512, 1024, 2048)
    assert shape["deposit_span"] in (4096, 8192, 16384) and shape["kick_pass_windows"] in (-1, 0, 1, 2)
    assert shape["deposit_multi"] in (-1, 0, 1) and shape["fft_transforms_per_block"] in (2, 4, 8)
    assert shape["collide_ms"] > 0
    assert np.array_equal(be.to_host(), e) and np.array_equal(bp.to_host(), p)
    st = tuned.stats()
    assert st["collides"] == 0 and st["green_misses"] == 0
    plain = PICCollider(solver)
    ce, cp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    assert tuned.collide(be, bp) == plain.collide(ce, cp)
    assert torch.equal(be.coords, ce.coords) and torch.equal(bp.coords, cp.coords)
    with pytest.raises(Exception):
        tuned.set_launch_shape("transverse_tiles", 3)


@pytest.mark.parametrize("kw", [dict(grid=(128, 128)), dict(grid=(128, 128), precision="fp32"),
                                dict(grid=(64, 48), interaction_grid="node", slice_interpolation="linear"),
                                dict(grid=(320, 40))])
def test_fft_group_shapes_bit_identical(oracle, gpu, kw):
    """Two, four or eight transforms per workgroup in the charge-row, convolution-column and inverse-row passes
    give the same transforms, so coordinates, luminosities and cache history are bit-identical; a 320-node
    axis (640-point transforms) is past the LDS of eight FP64 transforms per workgroup and runs with four."""
    e, p = flat_beams(oracle, 20000, 8000)
    solver = PICPoissonSolver(slicing=LongitudinalSlicing(nslices=5, method="normal_quantile"), deposit_method="TSC",
                              field_derivative="fourth", **kw)
    outs = []
    for g in (4, 2, 8):
        col = PICCollider(solver)
        col.set_launch_shape("fft_transforms_per_block", g)
        assert col.launch_shape()["fft_transforms_per_block"] == g
        be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
        lums = [col.collide(be, bp) for _ in range(2)]
        outs.append((lums, be.to_host(), bp.to_host(), col.stats()))
    for g, o in zip((2, 8), outs[1:]):
        assert o[0] == outs[0][0], g
        assert np.array_equal(o[1], outs[0][1]) and np.array_equal(o[2], outs[0][2]), g
        assert o[3] == outs[0][3], g
    col = PICCollider(solver)
    for bad in (1, 3, 16):
        with pytest.raises(Exception):
            col.set_launch_shape("fft_transforms_per_block", bad)


@pytest.mark.parametrize("beams", ["round", "flat"])
def test_deposit_shapes_bit_identical(oracle, gpu, monkeypatch, beams):
    """Per-plane deposit items, items that take all planes of an interaction (with the luminosity plane
    folded in), their split into smaller items, their plane-by-plane fallback for windows too large
    for one block (slice order instead of transverse tiles makes every chunk box span the beam), and
    the timed choice between the two forms all give the same integer charge sums, so coordinates,
    luminosities and cache history are bit-identical."""
    e, p = (round_beams if beams == "round" else flat_beams)(oracle, 30000, 12000)
    sl = LongitudinalSlicing(nslices=7, method="normal_quantile")
    solver = PICPoissonSolver(grid=(128, 128), slicing=sl, deposit_method="TSC", field_derivative="fourth",
                              slice_interpolation="quadratic")
    shapes = [dict(deposit_multi=0), dict(deposit_multi=1), dict(),
              dict(deposit_multi=1, deposit_span=16384, deposit_min_span=1024),
              dict(deposit_multi=1, deposit_span=4096, deposit_min_span=4096),
              dict(deposit_multi=1, tiles="0"), dict(tiles="0"), dict(deposit_multi=1, tiles="0", deposit_span=16384)]
    outs, figures = [], []
    for shape in shapes:
        shape = dict(shape)
        tiles = shape.pop("tiles", None)
        if tiles is None:
            monkeypatch.delenv("OCTOPIC_TRANSVERSE_TILES", raising=False)
        else:
            monkeypatch.setenv("OCTOPIC_TRANSVERSE_TILES", tiles)
        col = PICCollider(solver)
        for k, v in shape.items():
            col.set_launch_shape(k, v)
        be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
        lums, used = [], []
        for _ in range(3):
            lums.append(col.collide(be, bp))
            st = col.stats()
            used.append(st.pop("deposit_all_planes"))
        figures.append(used)
        outs.append((lums, be.to_host(), bp.to_host(), np.asarray(col.pair_luminosity()), st))
    monkeypatch.delenv("OCTOPIC_TRANSVERSE_TILES", raising=False)
    for shape, o in zip(shapes[1:], outs[1:]):
        assert o[0] == outs[0][0], shape
        assert np.array_equal(o[1], outs[0][1]) and np.array_equal(o[2], outs[0][2]), shape
        assert np.array_equal(o[3], outs[0][3]), shape
        assert o[4] == outs[0][4], shape
    assert figures[0] == [0, 0, 0] and figures[1] == [1, 1, 1]
    # the automatic form times all-planes items, then per-plane items, then keeps the faster
    assert figures[2][:2] == [1, 0] and figures[6][:2] == [1, 0]
    # the default shape against the oracle (one collide)
    compare_collide(oracle, gpu, solver, e, p)
