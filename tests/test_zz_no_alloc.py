"""The allocation-free beam generator (octopic_generate_gaussian_beam_noalloc) against the allocating one
(octopic_generate_gaussian_beam), through the C ABI: bit-identical beams on two calls in a row, the second
with a smaller beam (the static workspace keeps the first call's partial sums, which the second must not
read), over the partition's cap of 2048 blocks and below it; a non-positive count is rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from octopic import EMASS_EV, Beam, PICCollider, PICPoissonSolver, beam_params, lib  # noqa: E402
from octopic import _abi  # noqa: E402

PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
BETA, ALPHA, SIGMA = (0.8, 0.072, 12.727), (0.1, -0.2, 0.0), (95e-6, 8.5e-6, 0.007)


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def generate(col, n, seed, rng_id, gpu, noalloc):
    beam = Beam(torch.full((6, n), float("nan"), dtype=torch.float64, device=gpu), PE)
    s = beam.struct()
    D3 = C.c_double * 3
    fn = lib().octopic_generate_gaussian_beam_noalloc if noalloc else lib().octopic_generate_gaussian_beam
    rc = fn(col.h, seed, rng_id, D3(*BETA), D3(*ALPHA), D3(*SIGMA), 5.0, C.byref(s))
    return rc, beam


def test_noalloc_generator_matches_allocating_generator(gpu):
    col = PICCollider(PICPoissonSolver(grid=(32, 32)))
    for n, seed, rng_id in ((1_000_003, 11, 1), (5_001, 12, 2)):
        print(f"noalloc generator: n {n} seed {seed} rng_id {rng_id}", flush=True)
        rc, got = generate(col, n, seed, rng_id, gpu, noalloc=True)
        assert rc == _abi.OK, lib().octopic_last_error(col.h).decode()
        rc, ref = generate(col, n, seed, rng_id, gpu, noalloc=False)
        assert rc == _abi.OK, lib().octopic_last_error(col.h).decode()
        torch.cuda.synchronize()
        g, r = got.to_host(), ref.to_host()
        assert np.isfinite(g).all()
        assert np.array_equal(g, r)
        # standardised before the Twiss scaling: z has rms sigma_z exactly up to rounding
        assert abs(np.sqrt(np.mean((g[4] - g[4].mean()) ** 2)) / SIGMA[2] - 1.0) < 1e-12


def test_noalloc_generator_rejects_empty_beam(gpu):
    col = PICCollider(PICPoissonSolver(grid=(32, 32)))
    beam = Beam(torch.zeros((6, 1), dtype=torch.float64, device=gpu), PE)
    s = beam.struct()
    s.n = 0
    D3 = C.c_double * 3
    rc = lib().octopic_generate_gaussian_beam_noalloc(col.h, 1, 1, D3(*BETA), D3(*ALPHA), D3(*SIGMA), 5.0, C.byref(s))
    assert rc == _abi.ERR_INVALID_ARGUMENT
    assert "positive particle count" in lib().octopic_last_error(col.h).decode()
