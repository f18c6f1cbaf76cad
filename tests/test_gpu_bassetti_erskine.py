"""GPU parity of the Bassetti-Erskine kick and the Faddeeva function (octopic_gaussian_beambeam_kick,
octopic_faddeeva_w) against the oracle's restatement of strong_beam.jl:1139-1146 and
SpecialMath.jl:60-151 on the same inputs, and against scipy's wofz and the quadrature fixture.
FP64 bar: 1e-13 relative to the oracle on the fixture points and 1e-10 (the path's kick bar) on
random points, where the elliptic form's cancellation near round beams amplifies last-bit
differences (the device's exp/expm1/cos/sin and complex divisions round differently); 1e-12 to wofz
and 1e-10 to the quadrature."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle as O  # noqa: E402
from octopic import PICCollider, PICPoissonSolver  # noqa: E402
from octopic.solver import ArgumentError  # noqa: E402

FIX = os.path.join(os.path.dirname(__file__), "golden", "bassetti_erskine.npz")


@pytest.fixture(scope="module")
def col():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return PICCollider(PICPoissonSolver(grid=(16, 16)))


@pytest.fixture(scope="module")
def fix():
    with np.load(FIX, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def test_faddeeva_device(col, fix):
    z = fix["z"]
    w = col.faddeeva_w(torch.as_tensor(z, device="cuda:0")).cpu().numpy()
    for zi, wi, ref in zip(z, w, fix["w"]):
        o = O.faddeeva_w(complex(zi))
        assert abs(wi - o) <= 1e-13 * abs(o), (zi, wi, o)
        assert abs(wi - ref) <= 1e-12 * abs(ref), (zi, wi, ref)


def test_kick_device_on_fixture(col, fix):
    k = fix["kick"]
    for sx, sy in sorted(set(map(tuple, k[:, :2]))):
        rows = k[(k[:, 0] == sx) & (k[:, 1] == sy)]
        kx, ky = col.gaussian_beambeam_kick(dev(rows[:, 2]), dev(rows[:, 3]), sx, sy)
        kx, ky = kx.cpu().numpy(), ky.cpu().numpy()
        for (_, _, x, y, qx, qy), gx, gy in zip(rows, kx, ky):
            ox, oy = O.gaussian_beambeam_kick(sx, sy, x, y)
            assert math.hypot(gx - ox, gy - oy) <= 1e-13 * math.hypot(ox, oy) + 1e-300, (sx, sy, x, y)
            assert math.hypot(gx - qx, gy - qy) <= 1e-10 * math.hypot(qx, qy), (sx, sy, x, y)


@pytest.mark.parametrize("ratio", [1.0, 1.0 + 1e-5, 1.0003, 1.002, 0.5, 11.0, 1 / 25.0])
def test_kick_device_against_oracle_random(col, ratio):
    """Gaussian-distributed points (with tails to 6 sigma) over every branch the size ratio reaches."""
    rng = np.random.default_rng(int(ratio * 1000))
    sx, sy = 1.3e-5 * ratio, 1.3e-5
    x = rng.normal(0, sx, 4000) * rng.choice([0.001, 1.0, 3.0], 4000)
    y = rng.normal(0, sy, 4000) * rng.choice([0.001, 1.0, 3.0], 4000)
    kx, ky = col.gaussian_beambeam_kick(dev(x), dev(y), sx, sy)
    kx, ky = kx.cpu().numpy(), ky.cpu().numpy()
    ref = np.array([O.gaussian_beambeam_kick(sx, sy, a, b) for a, b in zip(x, y)])
    err = np.hypot(kx - ref[:, 0], ky - ref[:, 1]) / np.hypot(ref[:, 0], ref[:, 1])
    # the elliptic form subtracts two Faddeeva terms that nearly cancel for small eta and small rho
    # (relative conditioning ~ 1/(rho sqrt(eta))): last-bit differences of the two evaluations grow
    # there to ~1e-12; the kick bar of the path is 1e-10
    assert err.max() <= 1e-10, err.max()
    assert np.median(err) <= 1e-15


def test_kick_edge_cases(col):
    x, y = dev([0.0, 1e-6, -2e-6]), dev([0.0, -3e-6, 0.0])
    kx, ky = col.gaussian_beambeam_kick(x, y, 0.0, 1e-5)
    assert torch.count_nonzero(kx) == 0 and torch.count_nonzero(ky) == 0
    kx, ky = col.gaussian_beambeam_kick(x, y, 1e-5, 1e-5)
    assert kx[0].item() == 0.0 and ky[0].item() == 0.0
    e = torch.empty(0, dtype=torch.float64, device="cuda:0")
    kx, ky = col.gaussian_beambeam_kick(e, e, 1e-5, 1e-5)
    assert kx.numel() == 0
    with pytest.raises(ArgumentError):
        col.gaussian_beambeam_kick(x, y, -1.0, 1e-5)
    with pytest.raises(ArgumentError):
        col.gaussian_beambeam_kick(x, y, float("nan"), 1e-5)
