"""The luminosity channel and execution ledger of the run artifact (RunArtifact, src/tasks/RunArtifact.jl),
for Python callers of the collide path.

The reference keeps them in one HDF5 file; neither the HDF5 library nor h5py is available to this
package, so the same groups, datasets and attributes are kept as a directory tree:

    <path>/attrs.json                                  {"octopus_run_artifact": 1}
    <path>/luminosity/<label>/turns.i64, values.f64   one row per evaluated turn of that collision
    <path>/luminosity/<label>/attrs.json               {"label": ..., "rows_valid_through_turn": ...}
    <path>/execution/start_turn.i64, turns.i64, current_turn.i64, elapsed.f64   one row per run

Semantics follow RunArtifact.jl:
- each label has its own turn axis: a collision writes a row exactly when its own luminosity
  schedule evaluated (push_luminosity!, RunArtifact.jl:227-239); NaN keeps its meaning (evaluated
  and failed);
- rows are buffered per label and flushed every `capacity` rows and at finalize; a flush writes the
  rows first and the `rows_valid_through_turn` cursor after them (_ra_flush_luminosity!, :241-258);
- prepare (prepare_run_artifact!, :141-225): `append=False` recreates the artifact; `append=True`
  continues it, refuses a directory whose label set differs, truncates every group to the rows at
  or before min(cursor, first_turn - 1) (a torn tail, then the replay rule), and appends a ledger
  row (start_turn, planned turns, current_turn = first_turn - 1, elapsed 0);
- finalize (finalize_run_artifact!, :296-310) flushes every label and stamps the ledger row.
"""
import json
import os
import shutil
import time

import numpy as np

from .solver import ArgumentError

NOTHING_KNOWN = -(2 ** 31)   # typemin(Int32), the cursor of a group with no blessed rows


def _write_json(path, obj):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(obj, f)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)


def _read(path, dtype):
    return np.fromfile(path, dtype=dtype) if os.path.exists(path) else np.zeros(0, dtype=dtype)


def _append(path, values, dtype):
    with open(path, "ab") as f:
        f.write(np.asarray(values, dtype=dtype).tobytes())
        f.flush()
        os.fsync(f.fileno())


def _truncate(path, nrows, dtype):
    with open(path, "r+b") as f:
        f.truncate(nrows * np.dtype(dtype).itemsize)


class RunArtifact:
    """RunArtifact(path; append=false, capacity=64), luminosity channel and execution ledger."""

    def __init__(self, path, append=False, capacity=64):
        if int(capacity) < 1:
            raise ArgumentError("capacity must be positive")
        self.path = str(path)
        self.append = bool(append)
        self.capacity = int(capacity)
        self.open = False
        self.labels = []
        self.pending = {}
        self.slot = -1
        self.current_turn = 0
        self.t0 = 0.0

    # ------------------------------------------------------------------ prepare / finalize
    def prepare(self, labels, first_turn, planned_turns):
        if self.open:
            raise ArgumentError(f"RunArtifact at {self.path} is already open; one artifact serves one run at a time")
        labels = [str(x) for x in labels]
        if len(set(labels)) != len(labels):
            raise ArgumentError("collision labels must be unique")
        lum = os.path.join(self.path, "luminosity")
        fresh = not self.append or not os.path.isdir(self.path) or not os.listdir(self.path)
        if fresh:
            if os.path.isdir(self.path) and os.listdir(self.path):
                if not os.path.exists(os.path.join(self.path, "attrs.json")):
                    raise ArgumentError(f"{self.path} exists and is not a run artifact; use a new path")
                shutil.rmtree(self.path)
            os.makedirs(lum)
            os.makedirs(os.path.join(self.path, "execution"))
            _write_json(os.path.join(self.path, "attrs.json"), {"octopus_run_artifact": 1})
            for label in labels:
                g = os.path.join(lum, label)
                os.makedirs(g)
                for name in ("turns.i64", "values.f64"):
                    open(os.path.join(g, name), "wb").close()
                _write_json(os.path.join(g, "attrs.json"), {"label": label, "rows_valid_through_turn": NOTHING_KNOWN})
        else:
            if not os.path.isdir(lum):
                raise ArgumentError(f"RunArtifact(append=true): {self.path} is not a run artifact "
                                    "(no luminosity group). Use a new path.")
            existing = sorted(os.listdir(lum))
            if existing != sorted(labels):
                raise ArgumentError(f"RunArtifact(append=true): the luminosity groups at {self.path} ({existing}) do "
                                    f"not match this task's collision labels ({sorted(labels)}). Use a new path.")
            for label in labels:
                g = os.path.join(lum, label)
                cursor = int(json.load(open(os.path.join(g, "attrs.json")))["rows_valid_through_turn"])
                self._truncate_group(g, label, min(cursor, int(first_turn) - 1))
        ex = os.path.join(self.path, "execution")
        n = len(_read(os.path.join(ex, "start_turn.i64"), np.int64))
        _append(os.path.join(ex, "start_turn.i64"), [first_turn], np.int64)
        _append(os.path.join(ex, "turns.i64"), [planned_turns], np.int64)
        _append(os.path.join(ex, "current_turn.i64"), [int(first_turn) - 1], np.int64)
        _append(os.path.join(ex, "elapsed.f64"), [0.0], np.float64)
        self.slot = n
        self.current_turn = int(first_turn) - 1
        self.t0 = time.perf_counter()
        self.labels = labels
        self.pending = {label: ([], []) for label in labels}
        self.open = True

    @staticmethod
    def _truncate_group(g, label, keep_through):
        """_ra_truncate_group!: rows are in ascending turn order; keep those with turn <= keep_through."""
        turns = _read(os.path.join(g, "turns.i64"), np.int64)
        keep = int(np.searchsorted(turns, keep_through, side="right"))
        if keep < len(turns):
            _truncate(os.path.join(g, "turns.i64"), keep, np.int64)
            _truncate(os.path.join(g, "values.f64"), keep, np.float64)
        _write_json(os.path.join(g, "attrs.json"),
                    {"label": label, "rows_valid_through_turn": int(turns[keep - 1]) if keep else NOTHING_KNOWN})

    def finalize(self):
        if not self.open:
            return
        try:
            for label in self.labels:
                self._flush(label)
            self._stamp()
        finally:
            self.open = False

    # ------------------------------------------------------------------ rows
    def push_luminosity(self, label, turn, value):
        """push_luminosity!: one collision's luminosity for one turn, on that label's own turn axis."""
        if not self.open:
            raise ArgumentError("the run artifact is not prepared")
        label = str(label)
        if label not in self.pending:
            raise ArgumentError(f"unknown collision label {label!r}")
        t, v = self.pending[label]
        t.append(int(turn))
        v.append(float(value))
        if len(t) >= self.capacity:
            self._flush(label)

    def reached(self, turn):
        """The run has reached `turn` (the ledger's current_turn at the next flush)."""
        self.current_turn = int(turn)

    def _flush(self, label):
        t, v = self.pending[label]
        if not t:
            return
        g = os.path.join(self.path, "luminosity", label)
        _append(os.path.join(g, "turns.i64"), t, np.int64)
        _append(os.path.join(g, "values.f64"), v, np.float64)
        # rows first, the cursor after: a crash in between leaves rows the next prepare truncates
        _write_json(os.path.join(g, "attrs.json"), {"label": label, "rows_valid_through_turn": int(t[-1])})
        self._stamp()
        t.clear()
        v.clear()

    def _stamp(self):
        ex = os.path.join(self.path, "execution")
        for name, val, dt in (("current_turn.i64", self.current_turn, np.int64),
                              ("elapsed.f64", time.perf_counter() - self.t0, np.float64)):
            p = os.path.join(ex, name)
            arr = _read(p, dt)
            arr[self.slot] = val
            with open(p, "r+b") as f:
                f.write(arr.tobytes())
                f.flush()
                os.fsync(f.fileno())


def read_luminosity(path, name=None):
    """read(TaskOutput(path), :luminosity[; name]): {label: (turn, value)} or one label's (turn, value)."""
    lum = os.path.join(str(path), "luminosity")
    out = {}
    for label in sorted(os.listdir(lum)):
        g = os.path.join(lum, label)
        out[label] = (_read(os.path.join(g, "turns.i64"), np.int64), _read(os.path.join(g, "values.f64"), np.float64))
    if name is not None:
        if name not in out:
            raise ArgumentError(f"no luminosity series named {name!r}")
        return out[name]
    return out


def read_execution(path):
    """The execution ledger: start_turn, turns, current_turn, elapsed (one row per prepared run)."""
    ex = os.path.join(str(path), "execution")
    return {name: _read(os.path.join(ex, name + ext), dt)
            for name, ext, dt in (("start_turn", ".i64", np.int64), ("turns", ".i64", np.int64),
                                  ("current_turn", ".i64", np.int64), ("elapsed", ".f64", np.float64))}


def attributes(path, label):
    """The attributes of one luminosity group (label, rows_valid_through_turn)."""
    return json.load(open(os.path.join(str(path), "luminosity", str(label), "attrs.json")))
