"""The C ABI boundary without a GPU: struct layouts, exported symbols, host-side option validation.

- every octopic_* function declared in include/octopic.h is exported by liboctopic.so, and the library
  loads without a GPU (no compute call is made here);
- the ctypes mirrors in octopic._abi have the C layouts (compared against a gcc compile of the header);
- PICPoissonSolver rejects invalid options with ArgumentError, as the reference constructor
  (src/tasks/strongstrong/interface.jl:1410-1476).
"""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "octopic.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(octopic_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from octopic import backend
    if not os.path.exists(backend.LIB_PATH):
        subprocess.run(["make", "-s", "-j4", "-C", os.path.join(ROOT, "octopus.jl_amd", "csrc")], check=True)
    lib = backend.lib()   # loads without a GPU
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", backend.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (octopic_\w+)", out))
    assert set(names) <= exported
    from octopic import _abi
    assert set(_abi.EXPORTED) == set(names)


def test_struct_layouts_match_header():
    from octopic import _abi
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "octopic.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(octopic_slicing), sizeof(octopic_options), sizeof(octopic_beam),
         sizeof(octopic_stats), sizeof(octopic_interaction_record), offsetof(octopic_options, slicing2));
  printf("%zu %zu %zu %zu %zu\n", offsetof(octopic_options, grid_extent_sigma), offsetof(octopic_beam, n),
         offsetof(octopic_interaction_record, source_grid), offsetof(octopic_options, min_sigma),
         offsetof(octopic_options, include_sigma_xy));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    a = [int(v) for v in lines[0].split()]
    b = [int(v) for v in lines[1].split()]
    assert a == [C.sizeof(_abi.Slicing), C.sizeof(_abi.Options), C.sizeof(_abi.BeamStruct), C.sizeof(_abi.Stats),
                 C.sizeof(_abi.InteractionRecord), _abi.Options.slicing2.offset]
    assert b == [_abi.Options.grid_extent_sigma.offset, _abi.BeamStruct.n.offset,
                 _abi.InteractionRecord.source_grid.offset, _abi.Options.min_sigma.offset,
                 _abi.Options.include_sigma_xy.offset]


@pytest.mark.parametrize("kw", [
    dict(slice_pair_green_min_ratio=0.0), dict(slice_pair_green_min_ratio=1.5), dict(slice_pair_green_growth=-0.1),
    dict(grid=(4, 64)), dict(deposit_method="NGP"), dict(green_type="log"), dict(green_cache="always"),
    dict(field_derivative="sixth"), dict(slice_interpolation="cubic"), dict(grid_extent="quantile"),
    dict(grid_extent_sigma=0.0), dict(min_transverse_extent=-1.0), dict(grid_quantize=-0.5),
    dict(interaction_grid="global"), dict(interaction_grid="node", grid_extent="sigma"),
    dict(batch_mode="parallel"), dict(luminosity_grid=(2, 2)), dict(luminosity_deposit_method="NGP"),
])
def test_solver_validation(kw):
    from octopic import ArgumentError, PICPoissonSolver
    with pytest.raises(ArgumentError):
        PICPoissonSolver(**kw)


def test_solver_defaults_and_lowering():
    from octopic import LongitudinalSlicing, PICPoissonSolver, EveryNSteps, AtTurns
    s = PICPoissonSolver()
    assert s.grid == (128, 128) and s.deposit_method == "CIC" and s.green_cache == "slice_pair"
    assert s.resolved_luminosity_grid == (128, 128) and s.resolved_luminosity_deposit_method == "CIC"
    o, keep = PICPoissonSolver(grid=(64, 96), luminosity_grid=(32, 40), deposit_method="TSC",
                               slicing1=LongitudinalSlicing(nslices=3, method="specified", positions=(-1.0, 1.0))).to_options()
    assert (o.nx, o.ny, o.lum_nx, o.lum_ny, o.deposit, o.lum_deposit) == (64, 96, 32, 40, 1, 1)
    assert o.slicing1.npositions == 2 and o.slicing1.positions[1] == 1.0 and o.slicing2.nslices == 1
    sch = PICPoissonSolver(luminosity_schedule=EveryNSteps(start=2, step=3))
    assert [sch.compute_luminosity(t) for t in range(7)] == [False, False, True, False, False, True, False]
    assert PICPoissonSolver(luminosity_schedule=AtTurns([0, 4])).compute_luminosity(4)
    assert sch.compute_luminosity(None)
