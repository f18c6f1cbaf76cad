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


@pytest.mark.parametrize("hw,b", [((16, 16), 2), ((32, 48), 3)])
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
