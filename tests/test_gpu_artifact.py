"""The run artifact's luminosity channel fed by the device collides (octopic.artifact, RunArtifact.jl): two
collisions on their own schedules, a plain-PIC one every turn and a soft-Gaussian one every other turn,
write exactly their evaluated turns; the rows hold the values the collides returned, and the first
turn's rows agree with the oracle's collides to 1e-10."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from octopic import (EMASS_EV, PMASS_EV, Beam, EveryNSteps, GaussianPoissonSolver, LongitudinalSlicing,  # noqa: E402
                     PICCollider, PICPoissonSolver, beam_params)
from octopic.artifact import RunArtifact, read_execution, read_luminosity  # noqa: E402

PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
PP = beam_params(1.0, PMASS_EV, 275e9, 6.9e10)


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def test_device_rows_follow_each_collision_schedule(oracle, gpu, tmp_path):
    e = oracle.generate_beam(20000, 9, 1, (0.8, 0.072, 12.727), (0, 0, 0), (95e-6, 8.5e-6, 0.007))
    p = oracle.generate_beam(8000, 9, 2, (0.8, 0.072, 90.909), (0, 0, 0), (95e-6, 8.5e-6, 0.06))
    sl = LongitudinalSlicing(nslices=5, method="normal_quantile")
    collisions = [("ip_pic", PICPoissonSolver(grid=(64, 64), slicing=sl)),
                  ("ip_soft", GaussianPoissonSolver(slicing=sl, luminosity_schedule=EveryNSteps(step=2)))]
    cols = {label: PICCollider(s) for label, s in collisions}
    beams = {label: (Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)) for label, _ in collisions}
    # turn 0 of each collision on the oracle, from the same inputs
    first = {label: oracle.Collider(s).collide(e.copy(), PE, p.copy(), PP) for label, s in collisions}
    pa = str(tmp_path / "gpu.artifact")
    art = RunArtifact(pa, capacity=3)
    art.prepare([label for label, _ in collisions], 0, 5)
    returned = {label: [] for label, _ in collisions}
    for t in range(5):
        for label, _ in collisions:
            lum = cols[label].collide(*beams[label], turn=t)
            returned[label].append(lum)
            if not math.isnan(lum):
                art.push_luminosity(label, t, lum)
        art.reached(t)
    art.finalize()
    torch.cuda.synchronize()
    tp, vp = read_luminosity(pa, "ip_pic")
    ts, vs = read_luminosity(pa, "ip_soft")
    assert list(tp) == [0, 1, 2, 3, 4] and list(ts) == [0, 2, 4]
    assert np.array_equal(vp, np.array(returned["ip_pic"]))
    assert np.array_equal(vs, np.array(returned["ip_soft"])[[0, 2, 4]])
    assert all(math.isnan(returned["ip_soft"][t]) for t in (1, 3))   # unscheduled turns report NaN
    for label, v in (("ip_pic", vp), ("ip_soft", vs)):
        assert math.isclose(v[0], first[label], rel_tol=1e-10)
    assert list(read_execution(pa)["current_turn"]) == [4]
