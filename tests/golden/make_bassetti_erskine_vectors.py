"""Writes bassetti_erskine.npz: w(z) from scipy.special.wofz on points in every branch of the
device evaluation (Weideman band, continued fraction, both asymptotic forms, near the real axis,
below it), and Bassetti-Erskine kicks of transverse Gaussians by adaptive quadrature of the field
integral

    Kx = 2x int_0^inf exp(-x^2/(2sx^2+q) - y^2/(2sy^2+q)) / ((2sx^2+q)^(3/2) (2sy^2+q)^(1/2)) dq

(and Ky with the roles of x and y swapped), whose round-beam value is 2x(1 - exp(-r^2/2s^2))/r^2.
Run from the repository root: python tests/golden/make_bassetti_erskine_vectors.py
"""
import os

import numpy as np
from scipy.integrate import quad
from scipy.special import wofz


def kick_quad(sx, sy, x, y):
    def fx(q):
        return np.exp(-x * x / (2 * sx * sx + q) - y * y / (2 * sy * sy + q)) / (
            (2 * sx * sx + q) ** 1.5 * (2 * sy * sy + q) ** 0.5)

    def fy(q):
        return np.exp(-x * x / (2 * sx * sx + q) - y * y / (2 * sy * sy + q)) / (
            (2 * sy * sy + q) ** 1.5 * (2 * sx * sx + q) ** 0.5)

    s2 = max(sx, sy) ** 2
    out = []
    for f, c in ((fx, x), (fy, y)):
        # split at a few multiples of the larger variance, where the integrand turns over
        pts = [0.0, 0.5 * s2, 2 * s2, 8 * s2, 64 * s2]
        v = sum(quad(f, a, b, epsabs=0, epsrel=1e-13, limit=400)[0] for a, b in zip(pts[:-1], pts[1:]))
        v += quad(f, pts[-1], np.inf, epsabs=0, epsrel=1e-13, limit=400)[0]
        out.append(2 * c * v)
    return out


def main():
    rng = np.random.default_rng(20261203)
    z = [0j, 1j, 3j, 1 + 0j, 5.9 + 0.05j, 6.5 + 0.05j, 8.5 + 1e-11j, 30 + 0j, 2 + 6.9j, 2 + 7.1j, 50 + 50j,
         3000 + 1500j, 1e6 + 1e5j, 5e7 + 1e3j, 1e3 + 2e7j, -2.5 + 0.3j, 0.7 - 0.4j, -3 - 1j, 1e-8 + 1e-8j]
    z += list(rng.uniform(-8, 8, 120) + 1j * rng.uniform(0, 9, 120))
    z += list(rng.uniform(-3, 3, 30) - 1j * rng.uniform(0, 2, 30))
    z = np.array(z, dtype=np.complex128)
    w = wofz(z)
    cases = []
    for sx, sy in ((1.0, 1.0), (1.0, 1.0 - 1e-6), (1.0 + 5e-5, 1.0), (1.0, 1.0 + 1.5e-4), (1.0003, 1.0),
                   (1.0, 1.0006), (1.3, 1.0), (1.0, 3.0),
                   (11.0, 1.0), (95e-6, 8.5e-6)):
        s = max(sx, sy)
        pts = [(0.3, 0.2), (-1.1, 0.4), (2.5, -1.7), (-0.05, -3.0), (4.0, 6.0), (1e-5, 2e-6), (0.0, 0.7), (0.9, 0.0)]
        pts += [tuple(p) for p in rng.normal(0, 1.5, (4, 2))]
        for px, py in pts:
            x, y = px * s, py * s
            kx, ky = kick_quad(sx, sy, x, y)
            cases.append((sx, sy, x, y, kx, ky))
    cases = np.array(cases)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bassetti_erskine.npz")
    np.savez(out, z=z, w=w, kick=cases)
    print(f"wrote {out}: {len(z)} w(z) points, {len(cases)} kicks")


if __name__ == "__main__":
    main()
