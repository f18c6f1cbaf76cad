"""GPU against the CPU oracle at the production shape, and the batched-instance configuration.

- The linear 6D one-turn map on the device (k_linear6d, linear6d.jl:110-118) against the oracle's
  restatement: the same products summed in the same order without fused multiply-adds, so every
  coordinate is bit-identical.
- One full-size collide of each production arm against the oracle (the oracle on the host's cores
  over the conflict-free pair batches, bitwise independent of the thread count): 2.56M e- / 1.024M p
  flat beams, 15 :normal_quantile centroid slices, 128^2, TSC, 4th order,
    C1a  interaction_grid = :node, slice_interpolation = :linear
    C1b  interaction_grid = :slice_pair, slice_interpolation = :quadratic
  Slices bit-identical, cache decisions (C1b) identical, mesh geometry bit-identical on first use,
  coordinates and per-pair luminosity within 1e-10 (the reference's CPU/CUDA contract).
- The field against the exact Bassetti-Erskine field through the whole device collide, on the
  reference's published protocol (docs/history/poisson_solver_review_2026_07_24.md:823-852: 400 x 400
  Gaussian quantile lattice, 81 x 81 probes over +-4 sigma, CIC, :integrated, second order, no
  constant removed): medians within 6 % of the published 128^2 values 4.1e-4 / 7.1e-4 / 9.7e-4 /
  1.2e-3 (round, 5:1, 11:1, 25:1); the FP32 form within 1e-5 of the FP64 median.
- The batched configuration: 16 independent bunch pairs (N e- ~ U[0.3M, 2.56M], N p ~ U[0.12M, 1.024M],
  slices {7, 11, 15}, meshes {64, 96, 128}, per-instance seeds and tunes, FP32 form) collided together
  by octopic_collide_many.  At a tenth of the particle counts every instance is compared with an oracle
  collide: FP64 instances within 1e-10, FP32-form instances within the FP32 form's stated bars (1e-4 of
  the largest transverse kick; 5e-4 on a field slice whose mesh the two precisions aligned a cell apart).  At the
  full counts the 16 instances collided together are bit-identical to the same instances collided one
  by one, finite, and each luminosity lies within the hourglass band of the thin-slice overlap.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from octopic import (EMASS_EV, PMASS_EV, Beam, LongitudinalSlicing, PICCollider, PICPoissonSolver,  # noqa: E402
                     beam_params, collide_many, linear6d_matrix)

from test_gpu_parity import coords_close, first_use_pairs  # noqa: E402
from test_oracle_solver import be_kick  # noqa: E402

PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7203e11)
PP = beam_params(1.0, PMASS_EV, 275e9, 0.6881e11)
E_SHAPE = dict(beta=(0.8, 0.072, 0.7e-2 / 5.5e-4), sigma=(95e-6, 8.5e-6, 0.7e-2))
P_SHAPE = dict(beta=(0.8, 0.072, 6.0e-2 / 6.6e-4), sigma=(95e-6, 8.5e-6, 6.0e-2))
THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def production_beams(oracle, n1=2_560_000, n2=1_024_000, seed=0):
    e = oracle.generate_beam(n1, seed, 1, E_SHAPE["beta"], (0, 0, 0), E_SHAPE["sigma"])
    p = oracle.generate_beam(n2, seed, 2, P_SHAPE["beta"], (0, 0, 0), P_SHAPE["sigma"])
    return e, p


def test_linear6d_matches_oracle(oracle, gpu):
    e, _ = production_beams(oracle, 300_001, 10)
    m = linear6d_matrix(E_SHAPE["beta"], [2 * math.pi * t for t in (0.08, 0.06, 0.01)])
    col = PICCollider(PICPoissonSolver(grid=(32, 32)))
    b = Beam.from_host(e, PE, device=gpu)
    ref = e.copy()
    for _ in range(3):
        col.track_linear6d(b, m)
        oracle.track_linear6d(np.asarray(m, dtype=np.float64), ref)
    torch.cuda.synchronize()
    assert np.array_equal(b.to_host(), ref)
    assert not np.array_equal(ref, e)


def compare_full(oracle, gpu, solver, e, p, cache_history=True):
    ce, cp = e.copy(), p.copy()
    ocol = oracle.Collider(solver)
    gcol = PICCollider(solver)
    be, bp = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    lo = ocol.collide(ce, PE, cp, PP, nthreads=THREADS)
    lg = gcol.collide(be, bp, turn=0)
    torch.cuda.synchronize()
    for which, arr in ((1, ce), (2, cp)):
        n = arr.shape[1]
        so, sg = ocol.last_slices(which, n), gcol.last_slices(which, n)
        for a, b in zip(so, sg):
            assert np.array_equal(a, b)
    oi, oj, orr, ol = ocol.pair_luminosity()
    gi, gj, gr, gl = gcol.pair_luminosity()
    assert list(oi) == list(gi) and list(oj) == list(gj) and list(orr) == list(gr)
    first = first_use_pairs(list(zip(oi, oj)))
    ro, rg = ocol.records(), gcol.interaction_records()
    assert len(ro) == len(rg) == 2 * len(oi)
    for a, b in zip(ro, rg):
        assert (a.p, a.i, a.j, a.dir, a.cache_event) == (b.p, b.i, b.j, b.dir, b.cache_event)
        ga = np.array(list(a.source_grid) + list(a.field_grid))
        gb = np.array(list(b.source_grid) + list(b.field_grid))
        if a.p in first:
            assert np.array_equal(ga, gb), (a.p, a.dir)
        else:
            assert np.allclose(ga, gb, rtol=1e-12, atol=1e-12 * np.max(np.abs(ga))), (a.p, a.dir)
    ol, gl = np.array(ol), np.array(gl)
    assert np.max(np.abs(ol - gl)) <= 1e-10 * np.max(np.abs(ol))
    if cache_history:
        so, sg = ocol.stats(), gcol.stats()
        assert (so["hits"], so["misses"], so["rebuilds"]) == (sg["green_hits"], sg["green_misses"], sg["green_rebuilds"])
    assert math.isclose(lo, lg, rel_tol=1e-10)
    coords_close(ce, be.to_host())
    coords_close(cp, bp.to_host())
    return lo


def production_solver(arm):
    sl = LongitudinalSlicing(nslices=15, method="normal_quantile", center_position="centroid")
    if arm == "node":
        return PICPoissonSolver(grid=(128, 128), deposit_method="TSC", field_derivative="fourth",
                                interaction_grid="node", slice_interpolation="linear", slicing=sl)
    return PICPoissonSolver(grid=(128, 128), deposit_method="TSC", field_derivative="fourth",
                            interaction_grid="slice_pair", slice_interpolation="quadratic", slicing=sl)


@pytest.mark.parametrize("arm", ["quadratic", "node"], ids=["C1b_slice_pair_quadratic", "C1a_node_linear"])
def test_production_shape_collide_matches_oracle(oracle, gpu, arm):
    e, p = production_beams(oracle)
    lum = compare_full(oracle, gpu, production_solver(arm), e, p, cache_history=(arm == "quadratic"))
    analytic = PE.npart * PP.npart / (2 * math.pi * math.sqrt(2 * 95e-6 ** 2 * 2 * 8.5e-6 ** 2))
    assert 0.9 < lum / analytic < 1.02


PUBLISHED_128 = (4.1e-4, 7.1e-4, 9.7e-4, 1.2e-3)   # poisson_solver_review_2026_07_24.md:850


def gpu_field_error(oracle, gpu, sigx, sigy, precision, n=400, probes=81, extent=4.0):
    u = (np.arange(1, n + 1) - 0.5) / n
    q = np.array([oracle.erfinv(2 * v - 1) for v in u]) * math.sqrt(2.0)
    y, x = np.meshgrid(sigy * q, sigx * q, indexing="ij")
    src = np.zeros((6, x.size))
    src[0], src[2] = x.ravel(), y.ravel()
    xg, yg = np.linspace(-extent * sigx, extent * sigx, probes), np.linspace(-extent * sigy, extent * sigy, probes)
    PY, PX = np.meshgrid(yg, xg, indexing="ij")
    pr = np.zeros((6, PX.size))
    pr[0], pr[2] = PX.ravel(), PY.ravel()
    sl = LongitudinalSlicing(nslices=1, method="normal_quantile", center_position="centroid")
    # kbb = 1: the probes' px, py after the collide are 2 K / N_source, the normalised kick.  No Green
    # cache: a cache miss widens the meshes by the growth factor (1.25), a coarser spacing than the
    # published protocol's (the cached collide gives 4.68e-4 on the round beam).
    solver = PICPoissonSolver(grid=(128, 128), deposit_method="CIC", green_type="integrated", field_derivative="second",
                              kbb1=1.0, kbb2=1.0, longitudinal_kick=False, slicing=sl, precision=precision,
                              green_cache="none")
    col = PICCollider(solver)
    b1, b2 = Beam.from_host(src, PE, device=gpu), Beam.from_host(pr, PP, device=gpu)
    col.collide(b1, b2)
    out = b2.to_host()
    kx, ky = be_kick(sigx, sigy, PX.ravel(), PY.ravel())
    err = np.hypot(out[1] - kx, out[3] - ky)
    return float(np.median(err) / np.max(np.hypot(kx, ky)))


@pytest.mark.parametrize("k", range(4), ids=["round", "5to1", "11to1", "25to1"])
def test_device_field_error_reproduces_published_medians(oracle, gpu, k):
    aspect = (1.0, 5.0, 11.0, 25.0)[k]
    sy = 8.5e-6
    e64 = gpu_field_error(oracle, gpu, aspect * sy, sy, "fp64")
    assert abs(e64 / PUBLISHED_128[k] - 1.0) < 0.06, (aspect, e64)
    e32 = gpu_field_error(oracle, gpu, aspect * sy, sy, "fp32")
    assert abs(e32 - e64) <= 1e-5, (aspect, e32, e64)


def instance_configs(k=16, seed=7):
    """The batched configuration's instances (the same draw as bench.py --instances)."""
    rs = np.random.RandomState(seed)
    out = []
    for i in range(k):
        out.append(dict(n1=int(rs.uniform(0.3e6, 2.56e6)), n2=int(rs.uniform(0.12e6, 1.024e6)),
                        ns=int(rs.choice([7, 11, 15])), grid=int(rs.choice([64, 96, 128])), seed=i + 1,
                        dtune=0.002 * rs.uniform(-1, 1)))
    return out


def instance_solver(c, precision):
    sl = LongitudinalSlicing(nslices=c["ns"], method="normal_quantile", center_position="centroid")
    return PICPoissonSolver(grid=(c["grid"], c["grid"]), deposit_method="TSC", field_derivative="fourth",
                            slice_interpolation="quadratic", slicing=sl, precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_instances_match_oracle(oracle, gpu, precision):
    cfgs = instance_configs()
    assert len({c["ns"] for c in cfgs}) == 3 and len({c["grid"] for c in cfgs}) == 3
    cols, b1, b2, host = [], [], [], []
    for c in cfgs:
        e, p = production_beams(oracle, c["n1"] // 10, c["n2"] // 10, seed=c["seed"])
        host.append((e, p))
        cols.append(PICCollider(instance_solver(c, precision)))
        b1.append(Beam.from_host(e, PE, device=gpu))
        b2.append(Beam.from_host(p, PP, device=gpu))
    lums = collide_many(cols, b1, b2, [0] * len(cols))
    torch.cuda.synchronize()
    worst = []
    for c, (e, p), col, x, y, lg in zip(cfgs, host, cols, b1, b2, lums):
        ce, cp = e.copy(), p.copy()
        ocol = oracle.Collider(instance_solver(c, "fp64"))
        lo = ocol.collide(ce, PE, cp, PP, nthreads=THREADS)
        ge, gp = x.to_host(), y.to_host()
        # slicing and schedule are FP64 in both forms
        for which, arr in ((1, ce), (2, cp)):
            so, sg = ocol.last_slices(which, arr.shape[1]), col.last_slices(which, arr.shape[1])
            for a, b in zip(so, sg):
                assert np.array_equal(a, b), c
        so, sg = ocol.stats(), col.stats()
        assert (so["hits"], so["misses"], so["rebuilds"]) == (sg["green_hits"], sg["green_misses"], sg["green_rebuilds"])
        if precision == "fp64":
            assert math.isclose(lo, lg, rel_tol=1e-10), c
            coords_close(ce, ge)
            coords_close(cp, gp)
        else:
            # The FP32 form's bars: transverse kick 1e-4 and longitudinal 3e-3 of the largest FP64 kick,
            # luminosity 1e-5.  Kicks update the slices in place, so later bounds of the FP32 collide differ
            # from the FP64 ones by the FP32 kick error, and where a source-field origin separation sits
            # near a rounding tie of the mesh alignment (integer cells, pic_cpu.jl:1877-1905) the two
            # collides can align that interaction's meshes one cell apart: its field slice then differs by
            # the mesh's discretization error, not by rounding.  Those field slices (found by comparing
            # the interaction records with the oracle's) are held to 5e-4 transversely.
            assert math.isclose(lo, lg, rel_tol=1e-5), c
            g1 = c["grid"] - 1
            moved = set()
            for ra, rb in zip(ocol.records(), col.interaction_records()):
                sa, fa = np.array(ra.source_grid), np.array(ra.field_grid)
                sb, fb = np.array(rb.source_grid), np.array(rb.field_grid)
                offa = (fa[:2] - sa[:2]) / (sa[2:] / g1)
                offb = (fb[:2] - sb[:2]) / (sb[2:] / g1)
                if np.max(np.abs(offa - offb)) > 1e-3:
                    moved.add((2, ra.j) if ra.dir == 1 else (1, ra.i))
            rel = []
            for which, (a, g, a0) in ((1, (ce, ge, e)), (2, (cp, gp, p))):
                sl = ocol.last_slices(which, a.shape[1])
                flip = np.zeros(a.shape[1], dtype=bool)
                off = 0
                for k_, n_ in enumerate(sl[3]):
                    if (which, k_) in moved:
                        flip[np.asarray(sl[4][off:off + n_])] = True
                    off += n_
                for k in (1, 3, 5):
                    dk = a[k] - a0[k]
                    d = np.abs((g[k] - a0[k]) - dk) / np.max(np.abs(dk))
                    rel.append((float(np.max(d[~flip], initial=0.0)), float(np.max(d[flip], initial=0.0))))
                assert np.array_equal(g[4], a[4])
            print(f"fp32 instance grid {c['grid']} ns {c['ns']} n {c['n1'] // 10}/{c['n2'] // 10}: max kick error "
                  f"(same meshes / realigned meshes {sorted(moved)}) e- px {rel[0]} py {rel[1]} pz {rel[2]}  "
                  f"p px {rel[3]} py {rel[4]} pz {rel[5]}", flush=True)
            worst.append((rel, c))
    for rel, c in worst:
        for k, (same, realigned) in enumerate(rel):
            longitudinal = k % 3 == 2
            assert same <= (3e-3 if longitudinal else 1e-4), (c, k, rel)
            assert realigned <= (3e-3 if longitudinal else 5e-4), (c, k, rel)


def test_batched_instances_full_size(gpu):
    from octopic import generate_beam
    cfgs = instance_configs()
    runs = []
    for together in (True, False):
        cols, b1, b2, maps = [], [], [], []
        for c in cfgs:
            col = PICCollider(instance_solver(c, "fp32"))
            cols.append(col)
            b1.append(generate_beam(col, c["n1"], PE, E_SHAPE["beta"], (0, 0, 0), E_SHAPE["sigma"], c["seed"], 1,
                                    device=gpu))
            b2.append(generate_beam(col, c["n2"], PP, P_SHAPE["beta"], (0, 0, 0), P_SHAPE["sigma"], c["seed"], 2,
                                    device=gpu))
            maps.append((linear6d_matrix(E_SHAPE["beta"], [2 * math.pi * (t + c["dtune"]) for t in (0.08, 0.06, 0.01)]),
                         linear6d_matrix(P_SHAPE["beta"], [2 * math.pi * (t + c["dtune"]) for t in (0.228, 0.21, 0.01)])))
        lums = []
        for turn in range(2):
            if together:
                lums.append(collide_many(cols, b1, b2, [turn] * len(cols)))
            else:
                lums.append([col.collide(x, y, turn=turn) for col, x, y in zip(cols, b1, b2)])
            for col, x, y, (mx, my) in zip(cols, b1, b2, maps):
                col.track_linear6d(x, mx)
                col.track_linear6d(y, my)
        torch.cuda.synchronize()
        runs.append((lums, [x.coords.clone() for x in b1], [y.coords.clone() for y in b2],
                     [col.stats() for col in cols]))
        del cols, b1, b2
    a, b = runs
    assert a[0] == b[0] and a[3] == b[3]
    for x, y in zip(a[1] + a[2], b[1] + b[2]):
        assert torch.equal(x, y)
        assert torch.isfinite(x).all()
    # turn 0 collides the generated Gaussians: the hourglass keeps the luminosity below the thin-slice overlap;
    # by turn 1 the e- beam has taken a kick (beam-beam parameter ~0.1) and a tune rotation, which pinches it
    # above that overlap, so turn 1 is only required to be finite and positive
    analytic = PE.npart * PP.npart / (2 * math.pi * math.sqrt(2 * 95e-6 ** 2 * 2 * 8.5e-6 ** 2))
    print("luminosity / thin-slice overlap per turn:", [[round(v / analytic, 4) for v in t] for t in a[0]], flush=True)
    for lum in a[0][0]:
        assert 0.85 < lum / analytic < 1.02
    for lum in a[0][1]:
        assert math.isfinite(lum) and lum > 0
