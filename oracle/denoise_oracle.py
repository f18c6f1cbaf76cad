"""CPU oracle of the Gaussian-subtracted residual deposit and of the residual denoiser.

Test infrastructure only: tests, the smoke check and the bench's CPU leg may call it as the
checker; the library never does.

Residual deposit (restates the reference):

- ``source_moments``: ``_gpic_source_moments`` (src/tasks/strongstrong/gaussian_pic.jl:374-435),
  uncoupled sums shifted by the first member, ``_shifted_second_moment`` /
  ``_shifted_cross_moment`` (src/tasks/strongstrong/slicing.jl:664-668);
- ``drifted_gaussian``: ``_gpic_drifted_gaussian`` (gaussian_pic.jl:482-491);
- ``gaussian_profile``: ``_gpic_gaussian_profile!`` (gaussian_pic.jl:206-244), the node integral
  of the unit Gaussian against the CIC tent or the TSC quadratic spline, by erf moments;
- ``residual_plane``: the neutralised subtraction of ``_gpic_solve_drifted_field!``
  (gaussian_pic.jl:543-576): amp = sum(charge) / (sum(gx) sum(gy)), residual = charge - amp gx gy.

Pins (tests/test_oracle_denoise.py): the profile against brute-force quadrature of G(x) W(x - x_i)
and against the closed forms sum = 1, first moment = mu, second central moment = sigma^2 + h^2/6
(CIC) or + h^2/4 (TSC); the moments against numpy mean/var.

Denoiser (no counterpart in the reference; the network is this library's own): a three-level
U-Net on NHWC planes, 3x3 convolutions 16/32/64 channels, ReLU, 2x2 average pooling, nearest
upsampling, skip concatenation (deep features first), 1x1 head to one channel, 118129 parameters.
``forward``/``backward`` emulate the device's mixed precision exactly where it rounds: activations,
convolution weights and the gradients fed to the weight and input-gradient products are bfloat16
(round to nearest even), every sum is taken in float64 here (float32 on the device), the head and
the optimiser are float32. With ``emulate=False`` nothing is rounded, which is the form the
finite-difference gradient check pins. The convolution is pinned against scipy.signal.correlate.
"""
import math

import numpy as np

# name, input channels, output channels (3x3, pad 1); the head is a 1x1 convolution 16 -> 1
LAYERS = [("e1a", 2, 16), ("e1b", 16, 16), ("e2a", 16, 32), ("e2b", 32, 32), ("ba", 32, 64), ("bb", 64, 64),
          ("d2a", 96, 32), ("d2b", 32, 32), ("d1a", 48, 16), ("d1b", 16, 16)]
HEAD_IN = 16
IN_CHANNELS = 2


def n_params():
    return sum(co * 9 * ci + co for _, ci, co in LAYERS) + HEAD_IN + 1


def unpack(flat):
    """Flat float32 parameter vector -> dict. Per layer W[co][ky][kx][ci] then b[co]; head w[16], b."""
    p, o = {}, 0
    for name, ci, co in LAYERS:
        p[name + ".w"] = np.asarray(flat[o:o + co * 9 * ci], dtype=np.float64).reshape(co, 3, 3, ci)
        o += co * 9 * ci
        p[name + ".b"] = np.asarray(flat[o:o + co], dtype=np.float64)
        o += co
    p["head.w"] = np.asarray(flat[o:o + HEAD_IN], dtype=np.float64)
    o += HEAD_IN
    p["head.b"] = np.asarray(flat[o:o + 1], dtype=np.float64)
    o += 1
    assert o == len(flat)
    return p


def pack(p):
    parts = []
    for name, _, _ in LAYERS:
        parts += [p[name + ".w"].ravel(), p[name + ".b"].ravel()]
    parts += [p["head.w"].ravel(), p["head.b"].ravel()]
    return np.concatenate(parts).astype(np.float32)


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(a.shape)


def conv3x3(x, w, b=None):
    """Zero-padded 3x3 correlation, x [B,H,W,Ci], w [Co,3,3,Ci] -> [B,H,W,Co] (float64 sums)."""
    B, H, W, _ = x.shape
    xp = np.zeros((B, H + 2, W + 2, x.shape[3]))
    xp[:, 1:-1, 1:-1, :] = x
    y = np.zeros((B, H, W, w.shape[0]))
    for ky in range(3):
        for kx in range(3):
            y += xp[:, ky:ky + H, kx:kx + W, :] @ w[:, ky, kx, :].T
    if b is not None:
        y += b
    return y


def conv3x3_wgrad(x, dy):
    """dW[co,ky,kx,ci] = sum_p dy[p,co] x[p + (ky-1, kx-1), ci]."""
    B, H, W, ci = x.shape
    xp = np.zeros((B, H + 2, W + 2, ci))
    xp[:, 1:-1, 1:-1, :] = x
    dw = np.zeros((dy.shape[3], 3, 3, ci))
    d2 = dy.reshape(-1, dy.shape[3])
    for ky in range(3):
        for kx in range(3):
            dw[:, ky, kx, :] = d2.T @ xp[:, ky:ky + H, kx:kx + W, :].reshape(-1, ci)
    return dw


def conv3x3_dgrad(dy, w):
    """Input gradient: the correlation of dy with the flipped, channel-transposed kernel."""
    wt = np.transpose(w[:, ::-1, ::-1, :], (3, 1, 2, 0))
    return conv3x3(dy, wt)


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_back(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(d):
    B, H, W, C = d.shape
    return d.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward(p, x, emulate=True):
    """x [B,H,W,2] -> (out [B,H,W], cache). Mirrors the device forward pass."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    cache = {"x": r(x)}

    def layer(name, inp):
        a = r(np.maximum(conv3x3(inp, r(p[name + ".w"]), p[name + ".b"]), 0.0))
        cache[name] = a
        return a

    s1 = layer("e1b", layer("e1a", cache["x"]))
    q1 = r(pool(s1))
    cache["q1"] = q1
    s2 = layer("e2b", layer("e2a", q1))
    q2 = r(pool(s2))
    cache["q2"] = q2
    bt = layer("bb", layer("ba", q2))
    c2 = np.concatenate([up(bt), s2], axis=3)
    cache["c2"] = c2
    d2 = layer("d2b", layer("d2a", c2))
    c1 = np.concatenate([up(d2), s1], axis=3)
    cache["c1"] = c1
    d1 = layer("d1b", layer("d1a", c1))
    out = d1 @ p["head.w"] + p["head.b"][0]
    if emulate:
        out = out.astype(np.float32).astype(np.float64)
    return out, cache


def loss_and_grad(p, x, target, emulate=True):
    """Mean squared error over all pixels and its gradient (dict like p)."""
    out, c = forward(p, x, emulate)
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    npix = out.size
    diff = out - target
    loss = float(np.mean(diff * diff))
    dout = 2.0 * diff / npix
    g = {"head.w": np.tensordot(dout, c["d1b"], axes=([0, 1, 2], [0, 1, 2])), "head.b": np.array([dout.sum()])}
    da = dout[..., None] * p["head.w"]

    def back(name, inp, da):
        mask = c[name] > 0
        dpre = da * mask
        g[name + ".b"] = dpre.sum(axis=(0, 1, 2))
        dq = r(dpre)
        g[name + ".w"] = conv3x3_wgrad(inp, dq)
        return conv3x3_dgrad(dq, r(p[name + ".w"]))

    d = back("d1b", c["d1a"], da)
    d = back("d1a", c["c1"], d)
    dup, ds1 = d[..., :32], d[..., 32:]
    d = back("d2b", c["d2a"], up_back(dup))
    d = back("d2a", c["c2"], d)
    dup, ds2 = d[..., :64], d[..., 64:]
    d = back("bb", c["ba"], up_back(dup))
    d = back("ba", c["q2"], d)
    d = back("e2b", c["e2a"], ds2 + pool_back(d))
    d = back("e2a", c["q1"], d)
    d = back("e1b", c["e1a"], ds1 + pool_back(d))
    back("e1a", c["x"], d)
    return loss, g


def adam(flat, grad, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One float32 Adam step (t counts from 1); returns new (flat, m, v)."""
    f32 = np.float32
    g = grad.astype(f32)
    m = (f32(beta1) * m + f32(1 - beta1) * g).astype(f32)
    v = (f32(beta2) * v + f32(1 - beta2) * g * g).astype(f32)
    mh = m / f32(1 - beta1 ** t)
    vh = v / f32(1 - beta2 ** t)
    return (flat - f32(lr) * mh / (np.sqrt(vh) + f32(eps))).astype(f32), m, v


# ----------------------------------------------------------------------------- residual deposit
def source_moments(x, px, y, py):
    """gaussian_pic.jl:374-435 (uncoupled): sums shifted by the first member, serial order."""
    n = len(x)
    x0, px0, y0, py0 = x[0], px[0], y[0], py[0]
    sx = spx = sy = spy = sxx = spxpx = syy = spypy = sxpx = sypy = 0.0
    for i in range(n):
        dx, dpx, dy, dpy = x[i] - x0, px[i] - px0, y[i] - y0, py[i] - py0
        sx += dx; spx += dpx; sy += dy; spy += dpy
        sxx += dx * dx; spxpx += dpx * dpx; syy += dy * dy; spypy += dpy * dpy
        sxpx += dx * dpx; sypy += dy * dpy
    invn = 1.0 / n
    dmx, dmpx, dmy, dmpy = sx * invn, spx * invn, sy * invn, spy * invn
    sec = lambda s2, d: s2 * invn - d * d      # slicing.jl:664-665 (muladd(-d, d, s2*invn))
    crs = lambda s12, a, b: s12 * invn - a * b  # slicing.jl:667-668
    return dict(n=n, mx=x0 + dmx, mpx=px0 + dmpx, my=y0 + dmy, mpy=py0 + dmpy,
                varx=max(sec(sxx, dmx), 0.0), cxpx=crs(sxpx, dmx, dmpx), varpx=max(sec(spxpx, dmpx), 0.0),
                vary=max(sec(syy, dmy), 0.0), cypy=crs(sypy, dmy, dmpy), varpy=max(sec(spypy, dmpy), 0.0))


def drifted_gaussian(m, s):
    """gaussian_pic.jl:482-491: centroid and rms of the source drifted by s."""
    mux = m["mx"] + m["mpx"] * s
    muy = m["my"] + m["mpy"] * s
    varx = m["varx"] + 2 * s * m["cxpx"] + s * s * m["varpx"]
    vary = m["vary"] + 2 * s * m["cypy"] + s * s * m["varpy"]
    return mux, muy, math.sqrt(max(varx, 0.0)), math.sqrt(max(vary, 0.0))


def gaussian_profile(n, x0, h, mu, sigma, method):
    """gaussian_pic.jl:206-244. method 0 = CIC, 1 = TSC. Node i sits at x0 + i h (0-based)."""
    s = sigma
    invroot = 1.0 / (s * math.sqrt(2.0))
    norm = 1.0 / (s * math.sqrt(2.0 * math.pi))
    pdf = lambda x: norm * math.exp(-((x - mu) * (x - mu)) / (2 * s * s))
    m0 = lambda A, B: 0.5 * (math.erf((B - mu) * invroot) - math.erf((A - mu) * invroot))
    m1 = lambda A, B, xi: (mu - xi) * m0(A, B) - s * s * (pdf(B) - pdf(A))
    m2 = lambda A, B, xi: ((s * s + (mu - xi) ** 2) * m0(A, B) - s * s * ((B - mu) * pdf(B) - (A - mu) * pdf(A))
                           - 2 * (mu - xi) * s * s * (pdf(B) - pdf(A)))
    g = np.zeros(n)
    half = h / 2
    for i in range(n):
        xi = x0 + i * h
        if method == 0:
            g[i] = m0(xi - h, xi + h) + (m1(xi - h, xi, xi) - m1(xi, xi + h, xi)) / h
        else:
            cL, cR = xi - half, xi + half
            lL, lR = xi - 3 * half, xi - half
            rL, rR = xi + half, xi + 3 * half
            g[i] = (0.75 * m0(cL, cR) - m2(cL, cR, xi) / (h * h) + 1.125 * (m0(lL, lR) + m0(rL, rR))
                    + 1.5 * (m1(lL, lR, xi) - m1(rL, rR, xi)) / h + 0.5 * (m2(lL, lR, xi) + m2(rL, rR, xi)) / (h * h))
    return g


def residual_plane(charge, x0, y0, hx, hy, mux, muy, sigx, sigy, method, ns, neutralize=True):
    """gaussian_pic.jl:543-576. charge [nx, ny] (x index first). Returns (residual, gx, gy, amp)."""
    nx, ny = charge.shape
    gx = gaussian_profile(nx, x0, hx, mux, sigx, method)
    gy = gaussian_profile(ny, y0, hy, muy, sigy, method)
    sgx, sgy = gx.sum(), gy.sum()
    amp = float(ns)
    if neutralize and sgx * sgy > 0:
        amp = charge.sum() / (sgx * sgy)
    return charge - amp * np.outer(gx, gy), gx, gy, amp


# ----------------------------------------------------------------------------- denoiser inputs
def network_input(residual, gx, gy, amp, n_particles):
    """The two input planes the library feeds the network, from one residual plane [nx, ny]:
    residual / peak and the shot-noise standard deviation sqrt(expected count) / peak, where
    peak = amp max(gx) max(gy) is the subtracted Gaussian's largest node value (in particles)."""
    peak = amp * gx.max() * gy.max()
    noise = np.sqrt(np.maximum(amp * np.outer(gx, gy), 0.0) * 1.0)
    return np.stack([residual / peak, noise / peak], axis=-1), peak


# ----------------------------------------------------------------------------- training pairs
def pair_finish(planes, mom, method, H, W):
    """The library's training-pair residuals from its two deposits of one sample: planes [2, H*W]
    (uint64, 2^-32 units, row = v, column = u), mom [2, 5] (count, sum u, sum v, sum u^2, sum v^2 in
    2^-24 units). Each deposit gets its own neutralised Gaussian (residual_plane) from its moments;
    the target is the converged residual at the noisy sample's count, both over the noisy peak.
    Returns (input [H, W, 2] float32, target [H, W] float32, monitor)."""
    ms, ws = 2.0 ** 24, 2.0 ** 32
    res = []
    for st in range(2):
        n = float(mom[st][0])
        mu = float(np.int64(mom[st][1])) / ms / n
        mv = float(np.int64(mom[st][2])) / ms / n
        vu = float(np.int64(mom[st][3])) / ms / n - mu * mu
        vv = float(np.int64(mom[st][4])) / ms / n - mv * mv
        q = np.asarray(planes[st], dtype=np.uint64).astype(np.float64).reshape(H, W) / ws
        r, gx, gy, amp = residual_plane(q.T, 0.0, 0.0, 1.0, 1.0, mu, mv, math.sqrt(max(vu, 0.0)),
                                        math.sqrt(max(vv, 0.0)), method, int(n))
        res.append((n, r.T, gx, gy, amp))
    n0, r0, gx, gy, amp = res[0]
    n1, r1 = res[1][0], res[1][1]
    peak = amp * gx.max() * gy.max()
    g = amp * np.outer(gy, gx)
    inp = np.stack([r0 / peak, np.sqrt(np.maximum(g, 0.0)) / peak], axis=-1).astype(np.float32)
    tgt = (r1 * (n0 / n1) / peak).astype(np.float32)
    monitor = math.sqrt(float(np.mean(inp[..., 0].astype(np.float64) ** 2)))
    return inp, tgt, monitor
