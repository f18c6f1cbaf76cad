"""Oracle pins fixed by the reference's published numbers and by the mathematics of the method.

- Field accuracy against the exact Bassetti-Erskine field: the reference's published medians of the
  "as used" transverse-kick error (docs/history/poisson_solver_review_2026_07_24.md:823-852) for plain
  PIC, CIC, :integrated, second order, on the deterministic Gaussian quantile lattice (400 x 400
  sources, 81 x 81 field points over +-4 sigma, error |K - K_exact| per point over max |K_exact|, no
  constant removed): 64^2: 8.0e-4 / 1.8e-3 / 2.5e-3 / 2.9e-3 and 128^2: 4.1e-4 / 7.1e-4 / 9.7e-4 /
  1.2e-3 for round / 5:1 / 11:1 / 25:1 beams.  The published values carry two digits; the bar is 6 %.
  A factor two in the Green scale, the spacing or the derivative moves them by far more.  Also the
  published second- against fourth-order medians on a round beam over +-2 sigma
  (docs/theory/pic_free_space_kernels.md:74-82: 3.03e-3 / 1.83e-3 at 64^2, 7.50e-4 / 4.66e-4 at 128^2).
- :node meshes (_pic_build_node_grids!, pic_cpu.jl:1317-1382): the reference's coverage invariant
  (test/runtests.jl:9860-9901): the field mesh of node b contains every particle of the field slices
  b-1 and b at its drift 1/2(z - c_src) from the turn-start state; likewise the source mesh of node b
  contains the source slice at the boundary drifts of nodes b and b+1.
- :node longitudinal kick (_pic_interaction_node!, pic_cpu.jl:1448-1538): with px = py = 0 in every
  source slice the drifts move no source particle, so the L plane (drift sL on mesh G_s) and the Z
  plane (drift sR on the same mesh) carry identical charge and K_z = sum w (phi_L - phi_Z) is exactly
  zero; what remains of the pz change is the slingshot -1/4(px^2+py^2) + 1/4(px'^2+py'^2).  A Z plane
  solved on the wrong mesh breaks the zero.
- :equal_area slicing (_longitudinal_slices_equal_area, slicing.jl:104-156): on a source whose
  histogram of ns*resolution bins is flat, the cumulative fractions are c/bins exactly and the target
  s/ns is reached exactly at bin s*resolution, so the interpolation returns that bin's centre,
  zmin + (s*resolution - 1/2)*width, bit for bit; on any histogram with no empty bin it is the linear
  interpolation of the cumulative distribution through the bin centres (numpy.interp).
"""
import math

import numpy as np
import pytest

from octopic import EMASS_EV, PMASS_EV, LongitudinalSlicing, PICPoissonSolver, beam_params

from test_oracle_solver import be_kick

PUBLISHED_MEDIANS = {   # poisson_solver_review_2026_07_24.md:849-850, PIC CIC rows
    64: (8.0e-4, 1.8e-3, 2.5e-3, 2.9e-3),
    128: (4.1e-4, 7.1e-4, 9.7e-4, 1.2e-3),
}
ASPECTS = (1.0, 5.0, 11.0, 25.0)


def quantile_lattice(oracle, sigx, sigy, n):
    u = (np.arange(1, n + 1) - 0.5) / n
    q = np.array([oracle.erfinv(2 * v - 1) for v in u]) * math.sqrt(2.0)
    y, x = np.meshgrid(sigy * q, sigx * q, indexing="ij")
    return x.ravel(), y.ravel()


def field_error_median(oracle, grid, sigx, sigy, n=400, field_axis=81, extent=4.0, order="second"):
    solver = PICPoissonSolver(grid=(grid, grid), deposit_method="CIC", green_type="integrated",
                              field_derivative=order)
    sx, sy = quantile_lattice(oracle, sigx, sigy, n)
    xg = np.linspace(-extent * sigx, extent * sigx, field_axis)
    yg = np.linspace(-extent * sigy, extent * sigy, field_axis)
    box = [sx.min(), sx.max(), sy.min(), sy.max(), min(xg.min(), sx.min()), max(xg.max(), sx.max()),
           min(yg.min(), sy.min()), max(yg.max(), sy.max())]
    sg, fg = oracle.interaction_grids(solver, box)
    zero = np.zeros_like(sx)
    phi, ex, ey, _ = oracle.solve_field(solver, sx, zero, sy, zero, 0.0, sg, fg)
    ns = len(sx)
    errs, norms = [], []
    for yy in yg:
        for xx in xg:
            kx, ky = be_kick(sigx, sigy, xx, yy)
            ax, ay, _ = oracle.interpolate_kick(solver, fg, xx, yy, (phi, ex, ey), (phi, ex, ey), 1.0, 0.0)
            errs.append(math.hypot(2 * ax / ns - kx, 2 * ay / ns - ky))
            norms.append(math.hypot(kx, ky))
    return float(np.median(errs) / np.max(norms))


@pytest.mark.parametrize("grid", [64, 128])
def test_published_pic_field_error_medians(oracle, grid):
    sy = 8.5e-6
    got = [field_error_median(oracle, grid, a * sy, sy) for a in ASPECTS]
    for a, g, ref in zip(ASPECTS, got, PUBLISHED_MEDIANS[grid]):
        assert abs(g / ref - 1.0) < 0.06, (grid, a, g, ref)


PUBLISHED_ORDER_MEDIANS = {   # docs/theory/pic_free_space_kernels.md:74-82, round beam, +-2 sigma
    64: (3.03e-3, 1.83e-3),
    128: (7.50e-4, 4.66e-4),
}


@pytest.mark.parametrize("grid", [64, 128])
def test_published_second_and_fourth_order_medians(oracle, grid):
    """The reference's measured gain of the fourth-order derivative (round beam, field points over +-2 sigma,
    deterministic source): three published digits, a 1 % bar."""
    for order, ref in zip(("second", "fourth"), PUBLISHED_ORDER_MEDIANS[grid]):
        got = field_error_median(oracle, grid, 2e-3, 2e-3, extent=2.0, order=order)
        assert abs(got / ref - 1.0) < 0.01, (grid, order, got, ref)


PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
PP = beam_params(1.0, PMASS_EV, 275e9, 6.9e10)


def flat_beams(oracle, n1, n2, seed=3):
    e = oracle.generate_beam(n1, seed, 1, (0.8, 0.072, 12.727), (0, 0, 0), (95e-6, 8.5e-6, 0.007))
    p = oracle.generate_beam(n2, seed, 2, (0.8, 0.072, 90.909), (0, 0, 0), (95e-6, 8.5e-6, 0.06))
    return e, p


def node_solver(ns=4, **kw):
    sl = LongitudinalSlicing(nslices=ns, method="normal_quantile", center_position="centroid")
    return PICPoissonSolver(grid=(64, 64), interaction_grid="node", deposit_method="TSC",
                            field_derivative="fourth", slicing=sl, **kw)


def inside(grid, x, y):
    x0, y0, w, h = grid
    return bool(np.all((x0 <= x) & (x <= x0 + w) & (y0 <= y) & (y <= y0 + h)))


def test_node_mesh_coverage(oracle):
    e, p = flat_beams(oracle, 24000, 12000)
    e0, p0 = e.copy(), p.copy()
    solver = node_solver(ns=4)
    col = oracle.Collider(solver)
    col.collide(e, PE, p, PP)
    checked = 0
    for direction, (src0, fld0) in ((1, (e0, p0)), (2, (p0, e0))):
        bs, cs, _, cnt_s, idx_s = col.last_slices(direction, src0.shape[1])
        bf, cf, _, cnt_f, idx_f = col.last_slices(3 - direction, fld0.shape[1])
        off_s = np.concatenate([[0], np.cumsum(cnt_s)])
        off_f = np.concatenate([[0], np.cumsum(cnt_f)])
        ns_f = len(cf)
        for i in range(len(cs)):
            members = idx_s[off_s[i]:off_s[i + 1]]
            c = cs[i]
            for b in range(ns_f + 1):
                mesh = col.node_mesh(direction, i, b)
                assert mesh is not None, (direction, i, b)
                sg, fg = mesh
                # field slices adjacent to node b, drifted by 1/2(z - c_src)
                for adj in (b - 1, b):
                    if not 0 <= adj < ns_f:
                        continue
                    f = idx_f[off_f[adj]:off_f[adj + 1]]
                    sh = 0.5 * (fld0[4, f] - c)
                    assert inside(fg, fld0[0, f] + sh * fld0[1, f], fld0[2, f] + sh * fld0[3, f]), (direction, i, b, adj)
                # the source slice at the boundary drifts of nodes b and b + 1
                for bb in (b, min(b + 1, ns_f)):
                    d = 0.5 * (c - bf[bb])
                    xs = src0[0, members] + src0[1, members] * d
                    ys = src0[2, members] + src0[3, members] * d
                    assert inside(sg, xs, ys), (direction, i, b, bb)
                checked += 1
    assert checked == 2 * 4 * 5


def test_node_mesh_coverage_fails_for_a_reversed_drift(oracle):
    """The coverage check sees the drift's sign: the field particles drifted the other way, 1/2(c_src - z),
    leave some node meshes (the extreme slices drift by up to ~1e-2 m x 1e-4 rad against a mesh whose
    margin is a cell of ~1e-7 m in y)."""
    e, p = flat_beams(oracle, 24000, 12000)
    e0, p0 = e.copy(), p.copy()
    col = oracle.Collider(node_solver(ns=4))
    col.collide(e, PE, p, PP)
    bs, cs, _, cnt_s, idx_s = col.last_slices(1, e0.shape[1])
    bf, cf, _, cnt_f, idx_f = col.last_slices(2, p0.shape[1])
    off_f = np.concatenate([[0], np.cumsum(cnt_f)])
    violations = 0
    for i in range(len(cs)):
        for b in range(len(cf) + 1):
            _, fg = col.node_mesh(1, i, b)
            for adj in (b - 1, b):
                if not 0 <= adj < len(cf):
                    continue
                f = idx_f[off_f[adj]:off_f[adj + 1]]
                sh = 0.5 * (cs[i] - p0[4, f])
                violations += not inside(fg, p0[0, f] + sh * p0[1, f], p0[2, f] + sh * p0[3, f])
    assert violations > 0


def test_node_longitudinal_kick_vanishes_for_parallel_sources(oracle):
    # one slice per beam: both directions see pre-pair sources with px = py = 0, so the pz change of every
    # particle is exactly the slingshot term 1/4(px'^2 + py'^2) of its final angles
    e, p = flat_beams(oracle, 20000, 10000)
    for c in (e, p):
        c[1, :] = 0.0
        c[3, :] = 0.0
    e0, p0 = e.copy(), p.copy()
    col = oracle.Collider(node_solver(ns=1))
    col.collide(e, PE, p, PP)
    for after, before in ((e, e0), (p, p0)):
        assert np.any(after[1] != 0.0)                      # the transverse kick did act
        expect = before[5] + 0.25 * (after[1] * after[1] + after[3] * after[3])
        assert np.array_equal(after[5], expect)


def test_node_longitudinal_kick_vanishes_over_many_slices(oracle):
    # beam 1 has px = py = 0 and no kick acts on it (kbb1 = 0), so every source slice of direction 1 is
    # parallel in every pair: beam 2's pz changes by its slingshot terms alone, to round-off over the pairs
    e, p = flat_beams(oracle, 30000, 15000)
    e[1, :] = 0.0
    e[3, :] = 0.0
    p0 = p.copy()
    kbb = 2.0e-9
    col = oracle.Collider(node_solver(ns=5, kbb1=0.0, kbb2=kbb))
    col.collide(e, PE, p, PP)
    slingshot = 0.25 * (p[1] ** 2 + p[3] ** 2) - 0.25 * (p0[1] ** 2 + p0[3] ** 2)
    dev = np.max(np.abs(p[5] - p0[5] - slingshot))
    assert dev <= 1e-17, dev
    # with angles in the source the longitudinal kick is there, many orders above that round-off
    e2, p2 = flat_beams(oracle, 30000, 15000)
    p20 = p2.copy()
    col2 = oracle.Collider(node_solver(ns=5, kbb1=0.0, kbb2=kbb))
    col2.collide(e2, PE, p2, PP)
    slingshot2 = 0.25 * (p2[1] ** 2 + p2[3] ** 2) - 0.25 * (p20[1] ** 2 + p20[3] ** 2)
    assert np.max(np.abs(p2[5] - p20[5] - slingshot2)) > 1e4 * max(dev, 1e-20)


def flat_histogram_z(ns, resolution, per_bin):
    """z in [0, 1] with exactly per_bin values in every one of ns*resolution bins: each bin's centre, and the
    end points 0 and 1 in the first and last bin."""
    bins = ns * resolution
    centres = (np.arange(1, bins + 1) - 0.5) / bins
    z = np.repeat(centres, per_bin)
    z[0] = 0.0
    z[-1] = 1.0
    return z


def test_equal_area_flat_histogram_known_answer(oracle):
    for ns, res, per in ((5, 100, 3), (7, 20, 4), (3, 1, 5)):
        z = flat_histogram_z(ns, res, per)
        rng = np.random.default_rng(ns)
        z = z[rng.permutation(len(z))]
        b, c, w, lists = oracle.slices(z, LongitudinalSlicing(nslices=ns, method="equal_area", resolution=res))
        width = (1.0 - 0.0) / (ns * res)
        assert b[0] == 0.0 and b[-1] == 1.0
        for s in range(1, ns):
            assert b[s] == 0.0 + (s * res - 0.5) * width, (ns, res, s, b[s])
        # membership of the boundaries themselves: lb <= z < rb
        sid = np.empty(len(z), dtype=int)
        for k, l in enumerate(lists):
            sid[l] = k
        assert np.array_equal(sid, np.clip(np.searchsorted(b, z, side="right"), 1, ns) - 1)


def test_equal_area_is_the_interpolated_cdf_through_bin_centres(oracle):
    ns, res = 6, 10
    bins = ns * res
    rng = np.random.default_rng(42)
    counts = rng.integers(1, 9, bins)            # no empty bin, uneven
    zmin, zmax = -0.3, 0.5
    width = (zmax - zmin) / bins
    centres = zmin + (np.arange(1, bins + 1) - 0.5) * width
    z = np.repeat(centres, counts)
    z[np.argmin(z)] = zmin                       # end points inside the first and last bins
    z[np.argmax(z)] = zmax
    b, _, _, _ = oracle.slices(z, LongitudinalSlicing(nslices=ns, method="equal_area", resolution=res))
    cum = np.cumsum(counts) / len(z)
    for s in range(1, ns):
        t = s / ns
        expect = np.interp(t, np.concatenate([[0.0], cum]), np.concatenate([[zmin], centres]))
        assert abs(b[s] - expect) <= 1e-12 * (zmax - zmin), (s, b[s], expect)
    # a one-bin shift of the interpolation nodes would be off by about a bin width
    assert width > 1e6 * 1e-12 * (zmax - zmin)
