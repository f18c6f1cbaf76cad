"""The residual denoiser on the device against the CPU oracle (oracle/denoise_oracle.py).

The oracle rounds to bfloat16 exactly where the device does (activations, product weights, the
gradients entering the products) and sums in float64 where the device sums in float32, so the
comparisons allow for float32 accumulation (a few 1e-3 relative, stated per test), nothing more.
"""
import math

import denoise_oracle as D
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def _make(torch, h=16, w=16, b=4, seed=3):
    from octopic.denoise import Denoiser
    return Denoiser(height=h, width=w, max_batch=b, seed=seed)


def _inputs(rng, b, h, w):
    x = rng.standard_normal((b, h, w, 2)).astype(np.float32)
    x[..., 1] = np.abs(x[..., 1]) * 0.2
    t = (0.5 * x[..., 0] + 0.1 * rng.standard_normal((b, h, w))).astype(np.float32)
    return x, t


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def test_param_count_and_roundtrip(torch):
    d = _make(torch)
    assert d.n_params == D.n_params() == 118129
    p = d.get_params()
    assert np.isfinite(p).all() and np.abs(p).max() > 0
    q = p.copy()
    q[::7] *= -0.5
    d.set_params(q)
    assert np.array_equal(d.get_params(), q)


# small batches take the 8 x 8-pixel tile kernel at every level; 64 planes of 64 x 64 take the 16 x 16
# one at full resolution (1024 tiles and more)
@pytest.mark.parametrize("hw,b", [((16, 16), 2), ((32, 48), 3), ((64, 64), 64)])
def test_forward_matches_oracle(torch, hw, b):
    h, w = hw
    d = _make(torch, h, w, b, seed=11)
    rng = np.random.default_rng(5)
    x, _ = _inputs(rng, b, h, w)
    out = d.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ref, _ = D.forward(D.unpack(d.get_params()), x.astype(np.float64), emulate=True)
    # float32 sums against float64 sums of the same bf16 operands; an occasional activation lands on the
    # other side of a bf16 rounding boundary, which moves the output by about one bf16 ulp (4e-3 rel.)
    assert _rel(out, ref) <= 5e-3, _rel(out, ref)


def test_gradient_matches_oracle(torch):
    h, w, b = 16, 16, 2
    d = _make(torch, h, w, b, seed=12)
    rng = np.random.default_rng(6)
    x, t = _inputs(rng, b, h, w)
    loss, g = d.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    p = D.unpack(d.get_params())
    ref_loss, ref_g = D.loss_and_grad(p, x.astype(np.float64), t.astype(np.float64), emulate=True)
    assert abs(float(loss.item()) - ref_loss) <= 1e-3 * ref_loss
    gd = D.unpack(g.cpu().numpy())
    for name in ref_g:
        err = _rel(gd[name], ref_g[name])
        assert err <= 2e-2, (name, err)
    # the whole vector, the form the optimiser sees
    assert _rel(g.cpu().numpy(), D.pack(ref_g)) <= 1e-2


def test_adam_step_matches_oracle(torch):
    h, w, b = 16, 16, 2
    d = _make(torch, h, w, b, seed=13)
    rng = np.random.default_rng(7)
    x, t = _inputs(rng, b, h, w)
    p0 = d.get_params()
    _, g = d.train_step(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(), lr=1e-3, want_grad=True)
    p1 = d.get_params()
    ref, _, _ = D.adam(p0, g.cpu().numpy(), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    np.testing.assert_allclose(p1, ref, rtol=1e-6, atol=1e-9)


def test_training_is_deterministic_and_learns(torch):
    """A smoothing task the network can learn: noisy plane -> its 3x3 box-filtered noise-free plane.
    Two trainings from the same seed give identical parameters (fixed-order reductions); the loss
    falls by more than 5x in 150 steps."""
    h, w, b = 32, 32, 8
    rng = np.random.default_rng(8)
    yy, xx = np.mgrid[0:h, 0:w]
    clean = np.stack([np.exp(-((xx - 16 - 4 * np.cos(k)) ** 2 + (yy - 16 - 4 * np.sin(k)) ** 2) / (2 * (3 + k % 3) ** 2))
                      for k in range(b)]).astype(np.float32)
    noisy = clean + 0.1 * rng.standard_normal(clean.shape).astype(np.float32)
    x = np.stack([noisy, np.full_like(noisy, 0.1)], axis=-1)
    xt, tt = torch.from_numpy(x).cuda(), torch.from_numpy(clean).cuda()
    finals = []
    for _ in range(2):
        d = _make(torch, h, w, b, seed=21)
        losses = [float(d.train_step(xt, tt, lr=2e-3).item()) for _ in range(150)]
        finals.append(d.get_params())
        assert losses[-1] < losses[0] / 5.0, (losses[0], losses[-1])
    assert np.array_equal(finals[0], finals[1])


def test_rejects_bad_sizes(torch):
    from octopic.backend import ArgumentError
    from octopic.denoise import Denoiser
    with pytest.raises(ArgumentError):
        Denoiser(height=18, width=16, max_batch=1)
    d = _make(torch, 16, 16, 2)
    x = torch.zeros((3, 16, 16, 2), dtype=torch.float32, device="cuda")
    with pytest.raises(ArgumentError):
        d.forward(x)


# ------------------------------------------------------------------------------ training pairs
@pytest.mark.parametrize("method", ["CIC", "TSC"])
def test_pairs_match_oracle_finish(torch, method):
    """The device residuals of both deposits against the oracle's neutralised subtraction
    (gaussian_pic.jl:543-576) applied to the same deposits and moment sums."""
    h = w = 32
    d = _make(torch, h, w, 3)
    x, y, mon, planes, mom = d.make_pairs(seed=5, first=10, batch=3, n_target=1_000_000, method=method, raw=True)
    x, y, mon = x.cpu().numpy(), y.cpu().numpy(), mon.cpu().numpy()
    planes, mom = planes.cpu().numpy().view(np.uint64), mom.cpu().numpy().view(np.uint64)
    m = {"CIC": 0, "TSC": 1}[method]
    for k in range(3):
        inp, tgt, monitor = D.pair_finish(planes[k], mom[k], m, h, w)
        np.testing.assert_allclose(x[k], inp, rtol=0, atol=2e-6 * np.abs(inp).max())
        np.testing.assert_allclose(y[k], tgt, rtol=0, atol=2e-6 * np.abs(tgt).max())
        assert abs(mon[k] - monitor) <= 1e-5 * monitor


def test_pair_deposits_conserve_moments(torch):
    """Each TSC deposit holds its accepted particles exactly (partition of unity), its centroid is the
    particles' mean (first-moment exactness) and its variance the particles' variance + 1/4 cell^2
    (the TSC spline's second moment); noisy and converged draws of one sample agree in width."""
    h = w = 48
    d = _make(torch, h, w, 4)
    _, _, _, planes, mom = d.make_pairs(seed=9, first=0, batch=4, n_target=2_000_000, method="TSC", raw=True)
    planes, mom = planes.cpu().numpy().view(np.uint64), mom.cpu().numpy()
    v, u = np.mgrid[0:h, 0:w]
    for k in range(4):
        widths = []
        for st in range(2):
            q = planes[k, st].astype(np.float64).reshape(h, w) / 2.0 ** 32
            n = float(mom[k, st, 0])
            mu, mv = mom[k, st, 1] / 2.0 ** 24 / n, mom[k, st, 2] / 2.0 ** 24 / n
            vu = mom[k, st, 3] / 2.0 ** 24 / n - mu * mu
            assert abs(q.sum() - n) <= 1e-8 * n
            assert abs((q * u).sum() / q.sum() - mu) <= 1e-6
            assert abs((q * v).sum() / q.sum() - mv) <= 1e-6
            assert abs((q * (u - mu) ** 2).sum() / q.sum() - (vu + 0.25)) <= 1e-5 * vu
            widths.append(math.sqrt(vu))
        n_noisy = float(mom[k, 0, 0])
        assert abs(widths[0] / widths[1] - 1.0) <= 6.0 / math.sqrt(2.0 * n_noisy)


def test_pairs_are_deterministic(torch):
    d = _make(torch, 32, 32, 2)
    a = d.make_pairs(seed=3, first=4, batch=2, n_target=500_000, raw=True)
    b = d.make_pairs(seed=3, first=4, batch=2, n_target=500_000, raw=True)
    for s, t in zip(a, b):
        assert torch.equal(s, t)
    c = d.make_pairs(seed=4, first=4, batch=2, n_target=500_000, raw=True)
    assert not torch.equal(a[3], c[3])


def test_trained_network_beats_raw_residual(torch):
    """What the learned part gains: after 300 Adam steps on fresh pairs (16 a step, converged deposits
    of 4M macroparticles), the denoised residual of held-out slices (another seed) is at least 3x
    closer, in mean squared error, to the converged residual than the raw noisy residual is."""
    from octopic import train_denoiser as TD
    d = _make(torch, 64, 64, 16, seed=0)
    losses, mon = TD.train(d, steps=300, batch=16, seed=1, n_target=4_000_000)
    assert np.isfinite(losses).all() and mon > 0
    assert losses[-20:].mean() < 0.5 * losses[:20].mean()
    raw, den = TD.evaluate(d, seed=1001, samples=64, batch=16, n_target=4_000_000)
    assert den < raw / 3.0, (raw, den)


# ------------------------------------------------------------------------- in the collision
def _collide_beams(oracle, n1=30000, n2=12000, seed=3):
    e = oracle.generate_beam(n1, seed, 1, (0.8, 0.072, 12.727), (0, 0, 0), (95e-6, 8.5e-6, 0.007))
    p = oracle.generate_beam(n2, seed, 2, (0.8, 0.072, 90.909), (0, 0, 0), (95e-6, 8.5e-6, 0.06))
    return e, p


def _params():
    from octopic import EMASS_EV, PMASS_EV, beam_params
    return beam_params(-1.0, EMASS_EV, 10e9, 1.7e11), beam_params(1.0, PMASS_EV, 275e9, 6.9e10)


def _solver(**kw):
    from octopic import LongitudinalSlicing, PICPoissonSolver
    sl = LongitudinalSlicing(5, method="normal_quantile", center_position="centroid")
    base = dict(grid=(32, 32), deposit_method="TSC", field_derivative="fourth", slicing1=sl, slicing2=sl)
    base.update(kw)
    return PICPoissonSolver(**base)


@pytest.mark.parametrize("kw", [dict(), dict(slice_interpolation="quadratic"), dict(deposit_method="CIC", green_cache="none")],
                         ids=["linear", "quadratic", "cic"])
def test_zero_network_collide_matches_gaussian_oracle(oracle, torch, kw):
    """A network with all parameters zero outputs zero, so the denoised collide deposits each field plane
    as the source's neutralised drifted Gaussian: the oracle's replacement mode restates exactly that
    from the reference (gaussian_pic.jl:206-244, 374-491, 543-576). Coordinates within 1e-10 of the
    largest, luminosity (untouched planes) within 1e-10, the same cache decisions, every plane denoised."""
    from octopic import Beam, PICCollider
    PE, PP = _params()
    solver = _solver(**kw)
    e, p = _collide_beams(oracle)
    ce, cp = e.copy(), p.copy()
    ocol = oracle.Collider(solver, gaussian_replace=True)
    lo = ocol.collide(ce, PE, cp, PP)
    d = _make(torch, 32, 32, 32)
    d.set_params(np.zeros(d.n_params, np.float32))
    g = PICCollider(solver)
    g.set_denoiser(d, monitor_limit=1e30)
    be, bp = Beam.from_host(e, PE), Beam.from_host(p, PP)
    lg = g.collide(be, bp)
    torch.cuda.synchronize()
    st = g.stats()
    assert st["denoised_planes"] == st["planes"] and st["denoise_fallbacks"] == 0
    so = ocol.stats()
    assert (so["hits"], so["misses"], so["rebuilds"]) == (st["green_hits"], st["green_misses"], st["green_rebuilds"])
    assert math.isclose(lo, lg, rel_tol=1e-10)
    for a, b in ((ce, be.to_host()), (cp, bp.to_host())):
        for k in range(6):
            assert np.max(np.abs(a[k] - b[k])) <= 1e-10 * np.max(np.abs(a[k])), k
    # and it is not the plain collide: the Gaussian planes move the kicks
    e2, p2 = e.copy(), p.copy()
    oracle.Collider(solver).collide(e2, PE, p2, PP)
    assert np.max(np.abs(e2[1] - ce[1])) > 1e-6 * np.max(np.abs(e2[1] - e[1]))


def test_monitor_fallback_is_plain_pic(oracle, torch):
    """With a monitor limit below every plane's monitor, each plane stays plain PIC: the collide is
    bit-identical to one without a denoiser, and every plane counts as a fallback."""
    from octopic import Beam, PICCollider
    PE, PP = _params()
    solver = _solver()
    e, p = _collide_beams(oracle)
    d = _make(torch, 32, 32, 32)
    ref = PICCollider(solver)
    b1, b2 = Beam.from_host(e, PE), Beam.from_host(p, PP)
    l1 = ref.collide(b1, b2)
    g = PICCollider(solver)
    g.set_denoiser(d, monitor_limit=1e-30)
    c1, c2 = Beam.from_host(e, PE), Beam.from_host(p, PP)
    l2 = g.collide(c1, c2)
    torch.cuda.synchronize()
    st = g.stats()
    assert st["denoised_planes"] == 0 and st["denoise_fallbacks"] == st["planes"] > 0
    assert l1 == l2
    assert torch.equal(b1.coords, c1.coords) and torch.equal(b2.coords, c2.coords)


def test_denoiser_attachment_is_validated(torch):
    from octopic import PICCollider
    from octopic.backend import ArgumentError, UnsupportedOptionError
    d = _make(torch, 32, 32, 8)
    with pytest.raises(ArgumentError):
        PICCollider(_solver(grid=(64, 64))).set_denoiser(d, 1.0)
    with pytest.raises(UnsupportedOptionError):
        PICCollider(_solver(interaction_grid="node")).set_denoiser(d, 1.0)
    with pytest.raises(ArgumentError):
        PICCollider(_solver()).set_denoiser(d, 0.0)


def test_committed_weights_reduce_collide_kick_error(torch):
    """The committed 64 x 64 TSC weights inside the collision (production flat beams, 15 slices,
    :slice_pair + :quadratic, FP32 form, 640k e- / 256k p): the e- kicks from the 256k p beam are
    closer to the kicks from a 4x (1.024M) p beam in FP64 with the denoiser than without it, in x and
    in y."""
    import os
    from octopic import denoise_study as S
    from octopic import train_denoiser as TD
    from octopic.denoise import Denoiser
    w = os.path.join(os.path.dirname(__file__), "..", "octopus.jl_amd", "octopic", "weights", "denoiser_tsc_64.npz")
    d = Denoiser(64, 64, 96, seed=0)
    limit = 3.0 * TD.load(d, w)
    err = S.kick_error(64, d, limit, 640_000, 256_000, 4)
    assert err["denoised"]["x"] < 0.85 * err["plain"]["x"], err
    assert err["denoised"]["y"] < 0.9 * err["plain"]["y"], err
