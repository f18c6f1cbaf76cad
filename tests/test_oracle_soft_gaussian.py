"""Oracle pins for the soft-Gaussian solver (GaussianPoissonSolver, gaussian.jl; the uncoupled kick of
strong_beam.jl:619-687 between the virtual drifts of strong_beam.jl:497-617).

- The map one particle goes through (forward drift, transported moments, Bassetti-Erskine kick with the
  Hessian and centroid terms of the energy change, reverse drift) is symplectic with the longitudinal kick
  on, for the three safe drifts: J^T Omega J = Omega for a finite-difference Jacobian of a 13-probe stencil
  around an off-axis, off-momentum particle, against a correlated, offset source.  The correct map is at
  4.6e-8 (the differences' own error); a flipped Hessian term gives 4.1e-5, the centroid term on the wrong
  axes 4.3e-4 and the frozen-longitudinal map 1.4e-4, all above the 1e-6 bar.
- At the collision point (every z on the slice centre, so S = 0) the kick is kbb * w * K, K the scipy
  Bassetti-Erskine kick of the source's population moments (numpy, two-pass), both for a serial (< 4096)
  and a chunked moment sum; x and y are unchanged.
- With kbb = 0 the drifts telescope: the map is the identity to rounding.
- The luminosity of a quantile lattice of Gaussian particles against the other beam's Gaussian is the
  analytic overlap N1 N2 / (2 pi sqrt((sx1^2 + sx2^2)(sy1^2 + sy2^2))) exp(-dx^2/(2 Sx^2) - dy^2/(2 Sy^2)),
  for either sampled beam (gaussian_when_luminosity), with n1 != n2.
- min_sigma floors the variances: a zero-width source at min_sigma = s gives density 1 / (2 pi s^2) at
  its centre; at min_sigma = 0 the zero-width beams of runtests.jl:10614-10628 give luminosity 0 and stay
  finite.
- The reference's "Hybrid reproduces the analytic kick" test (runtests.jl:9470-9500): the Gaussian-
  subtracted PIC kick is within 3 % rms of this solver's and better than 0.95 of plain PIC's error at a
  48 x 48 mesh, which pins the kick's physical scale and sign against the PIC solvers.
- Constructor and collide errors as the reference's (runtests.jl:10607-10635, 12086-12094).
"""
import math

import numpy as np
import pytest

from octopic import (EMASS_EV, PMASS_EV, GaussianPICPoissonSolver, GaussianPoissonSolver, LongitudinalSlicing,
                     PICPoissonSolver, UnsafeVirtualDrift, beam_params)
from octopic.solver import ArgumentError, BeamParams

from test_oracle_solver import be_kick

PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
PP = beam_params(1.0, PMASS_EV, 275e9, 6.9e10)
ONE = LongitudinalSlicing(nslices=1, method="normal_quantile")


def _omega():
    om = np.zeros((6, 6))
    for k in range(3):
        om[2 * k, 2 * k + 1] = 1.0
        om[2 * k + 1, 2 * k] = -1.0
    return om


def _source(oracle):
    src = oracle.generate_beam(20000, 3, 1, (0.8, 0.3, 12.727), (0.6, -0.8, 0.0), (95e-6, 20e-6, 0.007))
    src[0] += 2e-5
    src[1] += 3e-5
    src[2] -= 3e-6
    src[3] -= 2e-5
    return np.ascontiguousarray(src)


@pytest.mark.parametrize("drift", ["hirata", "chromatic", "exact"])
def test_map_is_symplectic(oracle, drift):
    src = _source(oracle)
    q0 = np.array([3e-5, 2e-5, -4e-6, 5e-5, 2e-3, 3e-4])
    h = np.array([3e-7, 3e-7, 5e-8, 3e-7, 3e-5, 3e-7])
    pts = [q0.copy()]
    for k in range(6):
        for sgn in (1.0, -1.0):
            q = q0.copy()
            q[k] += sgn * h[k]
            pts.append(q)
    probes = np.ascontiguousarray(np.array(pts).T)
    s = GaussianPoissonSolver(slicing=ONE, virtual_drift=drift, longitudinal_kick=True, kbb2=2e-8)
    oracle.Collider(s).collide(src, PE, probes, PP)
    J = np.zeros((6, 6))
    for k in range(6):
        J[:, k] = (probes[:, 1 + 2 * k] - probes[:, 2 + 2 * k]) / (2 * h[k])
    om = _omega()
    assert np.max(np.abs(J.T @ om @ J - om)) < 1e-6
    assert abs(probes[1, 0] - q0[1]) > 1e-6   # the probe was kicked


@pytest.mark.parametrize("n_src", [300, 6000])
def test_kick_at_the_collision_point(oracle, n_src):
    rng = np.random.default_rng(5)
    src = np.zeros((6, n_src))
    src[0] = 1e-4 * rng.standard_normal(n_src) + 1e-5
    src[1] = 3e-5 * rng.standard_normal(n_src)
    src[2] = 2e-5 * rng.standard_normal(n_src) - 4e-6
    src[3] = 1e-5 * rng.standard_normal(n_src)
    src[5] = 1e-4 * rng.standard_normal(n_src)
    src = np.ascontiguousarray(src)
    npr = 200
    prb = np.zeros((6, npr))
    prb[0] = rng.uniform(-4e-4, 4e-4, npr)
    prb[2] = rng.uniform(-8e-5, 8e-5, npr)
    prb[1] = 1e-6 * rng.standard_normal(npr)
    prb[3] = 1e-6 * rng.standard_normal(npr)
    prb = np.ascontiguousarray(prb)
    before = prb.copy()
    kbb = 3e-9
    s = GaussianPoissonSolver(slicing=ONE, kbb1=0.0, kbb2=kbb, longitudinal_kick=False)
    oracle.Collider(s).collide(src, PE, prb, PP)
    mx, my = src[0].mean(), src[2].mean()
    sx, sy = math.sqrt(np.mean((src[0] - mx) ** 2)), math.sqrt(np.mean((src[2] - my) ** 2))
    kx, ky = be_kick(sx, sy, before[0] - mx, before[2] - my)
    scale = max(np.max(np.abs(kx)), np.max(np.abs(ky))) * kbb
    assert np.max(np.abs((prb[1] - before[1]) - kbb * kx)) <= 1e-10 * scale
    assert np.max(np.abs((prb[3] - before[3]) - kbb * ky)) <= 1e-10 * scale
    assert np.array_equal(prb[0], before[0]) and np.array_equal(prb[2], before[2])
    assert np.array_equal(prb[5], before[5])   # no longitudinal kick


@pytest.mark.parametrize("drift", ["hirata", "chromatic", "exact"])
def test_drifts_telescope_without_a_kick(oracle, drift):
    rng = np.random.default_rng(8)
    b1 = oracle.generate_beam(3000, 4, 1, (0.5, 0.2, 10.0), (0.3, -0.2, 0.0), (1e-4, 2e-5, 0.01))
    b2 = oracle.generate_beam(2000, 4, 2, (0.6, 0.3, 10.0), (0.1, 0.2, 0.0), (1e-4, 2e-5, 0.01))
    b2[1] += 1e-4 * rng.standard_normal(2000)
    b2[5] += 1e-3 * rng.standard_normal(2000)
    b1, b2 = np.ascontiguousarray(b1), np.ascontiguousarray(b2)
    a1, a2 = b1.copy(), b2.copy()
    s = GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=5, method="equal_count"), kbb1=0.0, kbb2=0.0,
                              virtual_drift=drift)
    oracle.Collider(s).collide(b1, PE, b2, PP)
    for got, want in ((b1, a1), (b2, a2)):
        for c in range(6):
            # pz's rounding scale is 1 + pz (the drifts change it through (1 + pz) phi)
            scale = 1.0 if c == 5 else np.max(np.abs(want[c]))
            assert np.max(np.abs(got[c] - want[c])) <= 1e-14 * scale, (drift, c)


def _lattice(m, sx, sy, ox=0.0, oy=0.0):
    from scipy.special import ndtri
    q = ndtri((np.arange(m) + 0.5) / m)
    X, Y = np.meshgrid(q * sx + ox, q * sy + oy, indexing="ij")
    c = np.zeros((6, m * m))
    c[0], c[2] = X.ravel(), Y.ravel()
    return np.ascontiguousarray(c)


@pytest.mark.parametrize("gwl", [1, 2])
def test_luminosity_is_the_gaussian_overlap(oracle, gwl):
    rng = np.random.default_rng(3)
    n_g = 40000
    g = np.zeros((6, n_g))
    g[0] = 1.2e-4 * rng.standard_normal(n_g) + 2e-5
    g[2] = 3e-5 * rng.standard_normal(n_g) - 1e-5
    g = np.ascontiguousarray(g)
    lat = _lattice(300, 0.9e-4, 2e-5)
    # gaussian_when_luminosity names the beam taken as the Gaussian; the other one's particles (here the
    # lattice, a deterministic quadrature of its Gaussian) sample its density
    b1, b2 = (g, lat) if gwl == 1 else (lat, g)
    n1, n2 = b1.shape[1], b2.shape[1]
    p1 = BeamParams(charge=-1.0, mc2=EMASS_EV, E0=10e9, r0=1e-15, npart=1.3e11)
    p2 = BeamParams(charge=1.0, mc2=PMASS_EV, E0=275e9, r0=1e-15, npart=0.7e11)
    s = GaussianPoissonSolver(slicing=ONE, kbb1=0.0, kbb2=0.0, gaussian_when_luminosity=gwl)
    lum = oracle.Collider(s).collide(b1, p1, b2, p2)
    mx, my = g[0].mean(), g[2].mean()
    sgx2, sgy2 = np.mean((g[0] - mx) ** 2), np.mean((g[2] - my) ** 2)
    Sx2, Sy2 = sgx2 + 0.9e-4 ** 2, sgy2 + 2e-5 ** 2   # the lattice is a quadrature of its nominal Gaussian
    want = p1.npart * p2.npart / (2 * math.pi * math.sqrt(Sx2 * Sy2)) * math.exp(-mx * mx / (2 * Sx2) - my * my / (2 * Sy2))
    assert n1 != n2
    assert abs(lum / want - 1) < 2e-4


def test_min_sigma_floor(oracle):
    n = 16
    s0 = 2e-5
    b1, b2 = np.zeros((6, n)), np.zeros((6, n))
    s = GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=1, method="equal_count"), kbb1=1e-4, kbb2=1e-4,
                              luminosity_scale=1.0, min_sigma=s0)
    lum = oracle.Collider(s).collide(b1, PE, b2, PP)
    assert math.isclose(lum, n / (2 * math.pi * s0 * s0), rel_tol=1e-14)


def test_zero_width_slices_stay_finite(oracle):
    # runtests.jl:10614-10628
    n = 16
    b1, b2 = np.zeros((6, n)), np.zeros((6, n))
    s = GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=1, method="equal_count"), kbb1=1e-4, kbb2=1e-4,
                              luminosity_scale=1.0, min_sigma=0.0)
    lum = oracle.Collider(s).collide(b1, PE, b2, PP)
    assert lum == 0.0
    assert np.all(np.isfinite(b1)) and np.all(np.isfinite(b2))


def test_hybrid_reproduces_the_soft_gaussian_kick(oracle):
    # runtests.jl:9470-9500, with the oracle's generator for the two 8000-particle round beams
    def pair():
        e = oracle.generate_beam(8000, 11, 1, (1.0, 1.0, 10.0), (0.0, 0.0, 0.0), (1e-4, 1e-4, 1e-2))
        p = oracle.generate_beam(8000, 11, 2, (1.0, 1.0, 10.0), (0.0, 0.0, 0.0), (1e-4, 1e-4, 1e-2))
        return np.ascontiguousarray(e), np.ascontiguousarray(p)
    PPr = beam_params(1.0, PMASS_EV, 275e9, 1.7e11)
    sl = LongitudinalSlicing(nslices=1, method="normal_quantile", center_position="centroid")
    e0, _ = pair()
    px0 = e0[1].copy()
    kicks = {}
    for name, solver in (
            ("soft", GaussianPoissonSolver(slicing=sl, longitudinal_kick=False)),
            ("hybrid", GaussianPICPoissonSolver(slicing=sl, grid=(48, 48), green_cache="none", longitudinal_kick=False)),
            ("pic", PICPoissonSolver(slicing=sl, grid=(48, 48), green_cache="none", longitudinal_kick=False))):
        e, p = pair()
        oracle.Collider(solver).collide(e, PE, p, PPr)
        kicks[name] = e[1] - px0
    rms = lambda v: math.sqrt(float(np.mean(v * v)))
    err_h = rms(kicks["hybrid"] - kicks["soft"]) / rms(kicks["soft"])
    err_p = rms(kicks["pic"] - kicks["soft"]) / rms(kicks["soft"])
    assert err_h < 0.03
    assert err_h < 0.95 * err_p


def test_constructor_checks():
    # runtests.jl:10607-10612, 10632-10635
    assert GaussianPoissonSolver().batch_mode == "wavefront"
    assert not GaussianPoissonSolver().include_sigma_xy
    assert GaussianPoissonSolver(include_sigma_xy=True).include_sigma_xy
    assert GaussianPoissonSolver().longitudinal_kick
    assert GaussianPoissonSolver().min_sigma == 0.0
    for bad in (-1.0, math.inf, math.nan):
        with pytest.raises(ArgumentError):
            GaussianPoissonSolver(min_sigma=bad)
    with pytest.raises(ArgumentError):
        GaussianPoissonSolver(batch_mode="invalid")
    with pytest.raises(ArgumentError):
        GaussianPoissonSolver(virtual_drift="invalid")
    with pytest.raises(ArgumentError):
        GaussianPoissonSolver(gaussian_when_luminosity=3)
    o, _ = GaussianPoissonSolver(virtual_drift=UnsafeVirtualDrift("chromatic_frozen_energy")).to_options()
    assert o.virtual_drift == 4
    # GaussianPoissonSolver{Float32}: the working precision is an option of the solver
    assert GaussianPoissonSolver().to_options()[0].precision == 0
    assert GaussianPoissonSolver(precision="fp32").to_options()[0].precision == 1
    with pytest.raises(ArgumentError):
        GaussianPoissonSolver(precision="fp16")


def test_collide_errors(oracle):
    # the coupled kick is not restated (the HIP library rejects it too)
    with pytest.raises(oracle.OracleError) as info:
        oracle.Collider(GaussianPoissonSolver(include_sigma_xy=True))
    assert info.value.code == 2   # OCTOPIC_ERR_UNSUPPORTED
    # runtests.jl:12086-12094: one particle per beam collides; E0 = 0 without kbb is an argument error
    one = lambda: np.zeros((6, 1))
    assert math.isfinite(oracle.Collider(GaussianPoissonSolver(kbb1=0.0, kbb2=0.0)).collide(one(), PE, one(), PP))
    bad = BeamParams(charge=1.0, mc2=1.0, E0=0.0, r0=1.0, npart=1.0)
    with pytest.raises(oracle.OracleError) as info:
        oracle.Collider(GaussianPoissonSolver()).collide(one(), bad, one(), PP)
    assert info.value.code == 7   # OCTOPIC_ERR_INVALID_ARGUMENT
    # a non-finite source coordinate fails the collide (gaussian.jl:126-130)
    n = 32
    b1 = np.ascontiguousarray(np.stack([1e-4 * np.sin(0.7 * np.arange(1, n + 1) + ph) for ph in range(6)]))
    b2 = b1.copy()
    b1[0, 4] = math.nan
    with pytest.raises(oracle.OracleError) as info:
        oracle.Collider(GaussianPoissonSolver(slicing=LongitudinalSlicing(nslices=2, method="equal_count"))).collide(
            b1, PE, b2, PP)
    assert info.value.code == 3   # OCTOPIC_ERR_NONFINITE
