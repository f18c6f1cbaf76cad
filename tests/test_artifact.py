"""The luminosity channel and execution ledger of the run artifact (octopic.artifact, RunArtifact.jl), as
the reference's own artifact tests pin them (test/runtests.jl:5645-5745): every evaluated turn is a
row and the cursor trails the rows; two collisions with disagreeing schedules keep their own turns;
append continues per group, a rewind drops per group, and every run keeps its ledger row; a torn tail
beyond the cursor is truncated on the next open.  The luminosity values come from collides of the CPU
oracle (the device path is tests/test_gpu_artifact.py) on the reference fixture's shape: 300
macroparticles per beam, 2 :equal_count slices."""
import math
import os

import numpy as np
import pytest

from octopic import EMASS_EV, PMASS_EV, EveryNSteps, LongitudinalSlicing, PICPoissonSolver, beam_params
from octopic.artifact import RunArtifact, attributes, read_execution, read_luminosity
from octopic.solver import ArgumentError

PE = beam_params(-1.0, EMASS_EV, 10e9, 1e10)
PP = beam_params(1.0, PMASS_EV, 275e9, 1e10)


def beams(oracle):
    e = oracle.generate_beam(300, 5, 1, (0.55, 0.056, 12.7), (0, 0, 0), (106e-6, 9.5e-6, 7e-3))
    p = oracle.generate_beam(300, 5, 2, (0.55, 0.056, 12.7), (0, 0, 0), (106e-6, 9.5e-6, 7e-3))
    return np.ascontiguousarray(e), np.ascontiguousarray(p)


def run(oracle, art, collisions, turns, start_turn=0):
    """One run of `turns` turns: every collision collides each turn and writes a row when its own
    schedule evaluated (interface.jl:2503-2511)."""
    e, p = beams(oracle)
    art.prepare([label for label, _ in collisions], start_turn, turns)
    cols = {label: oracle.Collider(s) for label, s in collisions}
    try:
        for t in range(start_turn, start_turn + turns):
            for label, s in collisions:
                lum = cols[label].collide(e, PE, p, PP)
                if s.compute_luminosity(t):
                    art.push_luminosity(label, t, lum)
            art.reached(t)
    finally:
        art.finalize()


def solver(**kw):
    return PICPoissonSolver(grid=(32, 32), slicing=LongitudinalSlicing(nslices=2, method="equal_count"), **kw)


def test_rows_cursor_and_ledger(oracle, tmp_path):
    pa = str(tmp_path / "a.artifact")
    run(oracle, RunArtifact(pa), [("ip", solver())], 3)
    turns, values = read_luminosity(pa, name="ip")
    assert list(turns) == [0, 1, 2]
    assert np.all(np.isfinite(values)) and np.all(values > 0)
    assert attributes(pa, "ip")["rows_valid_through_turn"] == 2
    ex = read_execution(pa)
    assert list(ex["start_turn"]) == [0] and list(ex["turns"]) == [3] and list(ex["current_turn"]) == [2]
    assert np.all(ex["elapsed"] > 0)


def test_disagreeing_schedules_keep_their_own_turns(oracle, tmp_path):
    pm = str(tmp_path / "m.artifact")
    run(oracle, RunArtifact(pm), [("ip1", solver()), ("ip2", solver(luminosity_schedule=EveryNSteps(step=3)))], 7)
    assert list(read_luminosity(pm, "ip1")[0]) == list(range(7))
    assert list(read_luminosity(pm, "ip2")[0]) == [0, 3, 6]


def test_append_rewind_and_torn_tail(oracle, tmp_path):
    pa2 = str(tmp_path / "a2.artifact")
    ip = [("ip", solver())]
    run(oracle, RunArtifact(pa2, append=True), ip, 3)
    run(oracle, RunArtifact(pa2, append=True), ip, 2, start_turn=3)
    run(oracle, RunArtifact(pa2, append=True), ip, 1, start_turn=2)   # rewind
    assert list(read_luminosity(pa2, "ip")[0]) == [0, 1, 2]
    ex = read_execution(pa2)
    assert list(ex["start_turn"]) == [0, 3, 2] and list(ex["current_turn"]) == [2, 4, 2]
    # rows beyond the cursor (written without it, as a crash between rows and cursor leaves them)
    g = os.path.join(pa2, "luminosity", "ip")
    with open(os.path.join(g, "turns.i64"), "ab") as f:
        f.write(np.array([98, 99], dtype=np.int64).tobytes())
    with open(os.path.join(g, "values.f64"), "ab") as f:
        f.write(np.array([1.0, 2.0]).tobytes())
    run(oracle, RunArtifact(pa2, append=True), ip, 1, start_turn=3)
    assert list(read_luminosity(pa2, "ip")[0]) == [0, 1, 2, 3]


def test_capacity_flushes_and_nan_rows(tmp_path):
    pa = str(tmp_path / "c.artifact")
    art = RunArtifact(pa, capacity=2)
    art.prepare(["ip"], 0, 5)
    art.push_luminosity("ip", 0, 1.0)
    assert attributes(pa, "ip")["rows_valid_through_turn"] == -(2 ** 31)   # buffered, nothing blessed yet
    art.push_luminosity("ip", 1, math.nan)                                  # evaluated and failed
    assert attributes(pa, "ip")["rows_valid_through_turn"] == 1            # flushed at capacity
    art.push_luminosity("ip", 2, 3.0)
    art.finalize()
    t, v = read_luminosity(pa, "ip")
    assert list(t) == [0, 1, 2] and v[0] == 1.0 and math.isnan(v[1]) and v[2] == 3.0


def test_refusals(tmp_path):
    pa = str(tmp_path / "r.artifact")
    art = RunArtifact(pa)
    art.prepare(["ip"], 0, 1)
    with pytest.raises(ArgumentError):
        art.prepare(["ip"], 0, 1)   # already open
    with pytest.raises(ArgumentError):
        art.push_luminosity("other", 0, 1.0)
    art.finalize()
    with pytest.raises(ArgumentError):   # append to an artifact whose labels differ
        RunArtifact(pa, append=True).prepare(["ip", "ip2"], 1, 1)
    other = tmp_path / "not_an_artifact"
    other.mkdir()
    (other / "data.txt").write_text("x")
    with pytest.raises(ArgumentError):   # a non-empty directory that is no artifact is never removed
        RunArtifact(str(other)).prepare(["ip"], 0, 1)
    assert (other / "data.txt").exists()
    with pytest.raises(ArgumentError):
        RunArtifact(pa, capacity=0)
