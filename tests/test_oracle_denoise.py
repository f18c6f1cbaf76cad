"""Pins of the residual-deposit and denoiser oracle (CPU only).

- bfloat16 rounding against torch's float32 -> bfloat16 conversion (round to nearest even);
- the 3x3 convolution against scipy.signal.correlate;
- the network gradient (unrounded form) against central finite differences, and the three
  convolution products against each other through the adjoint identity <dgrad(dy), x> = <dy, conv(x)>;
- the erf profile (gaussian_pic.jl:206-244) against brute-force quadrature of G(x) W(x - x_i) and
  the closed forms sum = 1, mean = mu, central second moment = sigma^2 + h^2/6 (CIC) or h^2/4 (TSC);
- the source moments (gaussian_pic.jl:374-435) against numpy mean / variance / covariance;
- the neutralised residual (gaussian_pic.jl:543-576): zero for a pure profile, zero sum always.
"""
import math

import denoise_oracle as D
import numpy as np
import pytest
from scipy import integrate, signal


def test_param_count():
    assert D.n_params() == 118129


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-8, 8, 4000),
                        [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(D.bf16(v), ref)
    assert D.bf16(np.float32(1.0 + 2 ** -8)) == 1.0          # tie -> even
    assert D.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def test_conv_matches_scipy():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 7, 9, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = D.conv3x3(x, w, b)
    for bi in range(2):
        for co in range(4):
            ref = b[co] + sum(signal.correlate(x[bi, :, :, ci], w[co, :, :, ci], mode="same") for ci in range(3))
            np.testing.assert_allclose(y[bi, :, :, co], ref, rtol=1e-12, atol=1e-12)


def test_conv_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 8, 6, 5))
    w = rng.standard_normal((7, 3, 3, 5))
    dy = rng.standard_normal((2, 8, 6, 7))
    lhs = np.sum(D.conv3x3_dgrad(dy, w) * x)
    rhs = np.sum(dy * D.conv3x3(x, w))
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs)
    dw = D.conv3x3_wgrad(x, dy)
    assert abs(np.sum(dw * w) - rhs) <= 1e-10 * abs(rhs)


def _random_params(rng, scale=1.0):
    p = {}
    for name, ci, co in D.LAYERS:
        p[name + ".w"] = rng.standard_normal((co, 3, 3, ci)) * scale * math.sqrt(2.0 / (9 * ci))
        p[name + ".b"] = rng.standard_normal(co) * 0.05
    p["head.w"] = rng.standard_normal(D.HEAD_IN) * 0.3
    p["head.b"] = np.array([0.01])
    return p


def test_gradient_finite_differences():
    rng = np.random.default_rng(4)
    p = _random_params(rng)
    x = rng.standard_normal((2, 8, 8, 2))
    t = rng.standard_normal((2, 8, 8))
    _, g = D.loss_and_grad(p, x, t, emulate=False)
    eps = 1e-6
    checked = 0
    for name in list(p):
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = D.loss_and_grad(p, x, t, emulate=False)[0]
            flat[idx] = old - eps
            lm = D.loss_and_grad(p, x, t, emulate=False)[0]
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = g[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (name, idx, fd, an)
            checked += 1
    assert checked >= 40


def test_pack_roundtrip():
    rng = np.random.default_rng(5)
    p = _random_params(rng)
    q = D.unpack(D.pack(p))
    for k in p:
        np.testing.assert_allclose(q[k], p[k].astype(np.float32), rtol=0, atol=0)


def test_adam_first_step_is_sign_step():
    g = np.array([1e-3, -2.0, 0.5, 0.0], np.float32)
    f = np.zeros(4, np.float32)
    f1, m, v = D.adam(f, g, np.zeros(4, np.float32), np.zeros(4, np.float32), 1, 0.01)
    # bias-corrected first step: -lr * g / (|g| + eps)
    np.testing.assert_allclose(f1, [-0.01 * 1e-3 / (1e-3 + 1e-8), 0.01, -0.01, 0.0], rtol=1e-5)


# ------------------------------------------------------------------------- residual deposit
def _tent(u):
    return np.maximum(1.0 - np.abs(u), 0.0)


def _tsc(u):
    a = np.abs(u)
    return np.where(a <= 0.5, 0.75 - a * a, np.where(a <= 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


@pytest.mark.parametrize("method", [0, 1])
def test_profile_quadrature(method):
    mu, s, h, x0 = 0.37, 0.9, 0.31, -4.0
    g = D.gaussian_profile(26, x0, h, mu, s, method)
    W = _tent if method == 0 else _tsc
    for i in (3, 11, 14, 20):
        xi = x0 + i * h
        G = lambda x: math.exp(-(x - mu) ** 2 / (2 * s * s)) / (s * math.sqrt(2 * math.pi)) * float(W((x - xi) / h))
        br = (1.0 if method == 0 else 1.5) * h
        pts = [xi - br, xi - h / 2, xi, xi + h / 2, xi + br] if method else [xi - br, xi, xi + br]
        ref = sum(integrate.quad(G, a, b, epsabs=1e-15, epsrel=1e-13)[0] for a, b in zip(pts[:-1], pts[1:]))
        assert abs(g[i] - ref) <= 1e-12


@pytest.mark.parametrize("method,var_w", [(0, 1.0 / 6.0), (1, 0.25)])
def test_profile_closed_forms(method, var_w):
    mu, s, h = 0.123, 0.75, 0.25
    n = 101
    x0 = mu - 50 * h - 0.07
    g = D.gaussian_profile(n, x0, h, mu, s, method)
    xs = x0 + h * np.arange(n)
    assert abs(g.sum() - 1.0) <= 1e-13
    assert abs((g * xs).sum() - mu) <= 1e-13
    assert abs((g * (xs - mu) ** 2).sum() - (s * s + var_w * h * h)) <= 1e-12


def test_source_moments_vs_numpy():
    rng = np.random.default_rng(6)
    n = 5000
    x = 3e-4 + 9.5e-5 * rng.standard_normal(n)
    px = 1e-4 * rng.standard_normal(n) + 0.3 * (x - 3e-4) / 0.8
    y = -2e-5 + 8.5e-6 * rng.standard_normal(n)
    py = 1.2e-4 * rng.standard_normal(n)
    m = D.source_moments(x, px, y, py)
    assert m["n"] == n
    for key, ref in (("mx", x.mean()), ("mpx", px.mean()), ("my", y.mean()), ("mpy", py.mean())):
        assert abs(m[key] - ref) <= 1e-12 * (abs(ref) + np.abs(x).max() if key == "mx" else abs(ref) + 1e-4)
    for key, ref in (("varx", x.var()), ("varpx", px.var()), ("vary", y.var()), ("varpy", py.var()),
                     ("cxpx", np.cov(x, px, bias=True)[0, 1]), ("cypy", np.cov(y, py, bias=True)[0, 1])):
        assert abs(m[key] - ref) <= 1e-9 * abs(ref), key
    mux, muy, sx, sy = D.drifted_gaussian(m, 0.2)
    assert abs(mux - (x + 0.2 * px).mean()) <= 1e-15
    assert abs(sx - (x + 0.2 * px).std()) <= 1e-9 * sx
    assert abs(sy - (y + 0.2 * py).std()) <= 1e-9 * sy


@pytest.mark.parametrize("method", [0, 1])
def test_residual_neutralised(method):
    nx, ny = 24, 20
    gx = D.gaussian_profile(nx, -1.0, 0.1, 0.1, 0.25, method)
    gy = D.gaussian_profile(ny, -0.5, 0.06, -0.02, 0.1, method)
    pure = 1234.0 * np.outer(gx, gy)
    r, _, _, amp = D.residual_plane(pure, -1.0, -0.5, 0.1, 0.06, 0.1, -0.02, 0.25, 0.1, method, 1000)
    assert np.abs(r).max() <= 1e-12 * pure.max()
    assert abs(amp - 1234.0) <= 1e-9
    rng = np.random.default_rng(7)
    noisy = pure + rng.standard_normal(pure.shape)
    r2, *_ = D.residual_plane(noisy, -1.0, -0.5, 0.1, 0.06, 0.1, -0.02, 0.25, 0.1, method, 1000)
    assert abs(r2.sum()) <= 1e-9 * noisy.sum()
