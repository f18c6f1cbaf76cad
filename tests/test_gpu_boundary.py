"""GPU tests of the C ABI boundary: Float32 beams, the Green cache after a failed collide, and the
argument checks of the batched-instance entry point.

- A Float32 beam (Phase6DRep{Float32}, src/beam/Beam.jl:45-56) is widened exactly to FP64 for the
  collision and rounded back at the end, so its result is bit-identical to the float32 rounding of the
  FP64 collide of the widened beam; z is untouched.  Checked for device and host (copied) storage.
- A collide that fails on the device (here a non-finite coordinate in a slice of a middle wavefront
  batch, pic_cpu.jl:954-1035 / interface.jl:541-599) leaves no Green-cache entry it built: the next
  collide of repaired beams on the same handle is bit-identical, cache decisions included, to the same
  collide on a fresh handle.
- octopic_collide_many rejects two instances that share a coordinate array (they collide concurrently
  and update their beams in place).
"""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from octopic import (EMASS_EV, PMASS_EV, Beam, LongitudinalSlicing, PICCollider, PICPoissonSolver,  # noqa: E402
                     beam_params, collide_many)
from octopic import _abi  # noqa: E402
from octopic.backend import ArgumentError, NonFiniteCoordinateError, lib  # noqa: E402

PE = beam_params(-1.0, EMASS_EV, 10e9, 1.7e11)
PP = beam_params(1.0, PMASS_EV, 275e9, 6.9e10)


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def flat_beams(oracle, n1=30000, n2=12000, seed=3):
    e = oracle.generate_beam(n1, seed, 1, (0.8, 0.072, 12.727), (0, 0, 0), (95e-6, 8.5e-6, 0.007))
    p = oracle.generate_beam(n2, seed, 2, (0.8, 0.072, 90.909), (0, 0, 0), (95e-6, 8.5e-6, 0.06))
    return e, p


def host_struct(arr, params, dtype):
    """octopic_beam over host arrays (on_device = 0)."""
    b = _abi.BeamStruct()
    b.x, b.px, b.y, b.py, b.z, b.pz = (arr[k].ctypes.data for k in range(6))
    b.n = arr.shape[1]
    b.charge, b.mc2, b.E0, b.r0, b.npart = (float(params.charge), float(params.mc2), float(params.E0),
                                           float(params.r0), float(params.npart))
    b.on_device = 0
    b.dtype = dtype
    return b


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_float32_beams_are_the_rounded_fp64_collide(oracle, gpu, precision):
    e, p = flat_beams(oracle)
    e32, p32 = e.astype(np.float32), p.astype(np.float32)
    sl = LongitudinalSlicing(nslices=7, method="normal_quantile")
    solver = PICPoissonSolver(grid=(64, 64), slicing=sl, deposit_method="TSC", field_derivative="fourth",
                              slice_interpolation="quadratic", precision=precision)
    # the FP64 collide of the exactly widened beams, rounded to float32
    ref = PICCollider(solver)
    w1 = Beam.from_host(e32.astype(np.float64), PE, device=gpu)
    w2 = Beam.from_host(p32.astype(np.float64), PP, device=gpu)
    lum_ref = ref.collide(w1, w2, turn=0)
    x1, x2 = w1.to_host().astype(np.float32), w2.to_host().astype(np.float32)
    # float32 device beams
    col = PICCollider(solver)
    f1 = Beam.from_host(e32, PE, device=gpu, dtype=torch.float32)
    f2 = Beam.from_host(p32, PP, device=gpu, dtype=torch.float32)
    assert col.collide(f1, f2, turn=0) == lum_ref
    g1, g2 = f1.to_host(), f2.to_host()
    assert g1.dtype == np.float32 and g2.dtype == np.float32
    assert np.array_equal(g1, x1) and np.array_equal(g2, x2)
    assert np.array_equal(g1[4], e32[4]) and np.array_equal(g2[4], p32[4])
    assert not np.array_equal(g1[1], e32[1])   # the collide did kick
    st, sr = col.stats(), ref.stats()
    assert (st["green_misses"], st["green_hits"]) == (sr["green_misses"], sr["green_hits"])
    # float32 host arrays (copied in and out by the library)
    h1, h2 = e32.copy(), p32.copy()
    hc = PICCollider(solver)
    s1, s2 = host_struct(h1, PE, _abi.OCTOPIC_DTYPE_F32), host_struct(h2, PP, _abi.OCTOPIC_DTYPE_F32)
    lum_h = C.c_double()
    hc._check(lib().octopic_collide(hc.h, C.byref(s1), C.byref(s2), 0, 1, C.byref(lum_h)))
    torch.cuda.synchronize()
    assert lum_h.value == lum_ref
    assert np.array_equal(h1, x1) and np.array_equal(h2, x2)


def test_beam_dtype_is_checked(oracle, gpu):
    e, p = flat_beams(oracle, 2000, 1000)
    col = PICCollider(PICPoissonSolver(grid=(32, 32), slicing=LongitudinalSlicing(3, method="normal_quantile")))
    s1, s2 = host_struct(e, PE, 2), host_struct(p, PP, _abi.OCTOPIC_DTYPE_F64)
    lum = C.c_double()
    rc = lib().octopic_collide(col.h, C.byref(s1), C.byref(s2), 0, 1, C.byref(lum))
    assert rc == _abi.ERR_INVALID_ARGUMENT
    assert "dtype" in lib().octopic_last_error(col.h).decode()


def test_failed_collide_leaves_no_cache_entry_it_built(oracle, gpu):
    e, p = flat_beams(oracle, 40000, 16000, seed=11)
    sl = LongitudinalSlicing(nslices=7, method="normal_quantile", center_position="centroid")
    solver = PICPoissonSolver(grid=(64, 64), slicing=sl, deposit_method="TSC", field_derivative="fourth",
                              slice_interpolation="quadratic")
    # a non-finite px in beam 1's middle slice: its pairs sit in the middle wavefront batches, which hold
    # several pairs, and the batches before them build and commit their entries
    k = int(np.argmin(np.abs(e[4])))
    bad = e.copy()
    bad[1, k] = np.nan
    col = PICCollider(solver)
    b1, b2 = Beam.from_host(bad, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
    with pytest.raises(NonFiniteCoordinateError) as info:
        col.collide(b1, b2, turn=0)
    assert "collision-order pair" in str(info.value)
    pair = int(str(info.value).split("collision-order pair ")[1].split(",")[0])
    assert pair > 0   # not the first pair: earlier batches ran
    assert np.array_equal(b1.to_host(), bad, equal_nan=True) and np.array_equal(b2.to_host(), p)
    # repaired beams on the same handle and on a fresh one: bit-identical results and cache decisions
    outs = []
    for c in (col, PICCollider(solver)):
        x1, x2 = Beam.from_host(e, PE, device=gpu), Beam.from_host(p, PP, device=gpu)
        lums = [c.collide(x1, x2, turn=t) for t in range(2)]
        outs.append((lums, x1.to_host(), x2.to_host(), c.stats()["green_builds"],
                     [(r.p, r.dir, r.cache_event, tuple(r.source_grid), tuple(r.field_grid))
                      for r in c.interaction_records()]))
    a, b = outs
    assert a[0] == b[0] and all(math.isfinite(v) for v in a[0])
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3] and a[4] == b[4]


def test_collide_many_rejects_shared_beams(oracle, gpu):
    e, p = flat_beams(oracle, 4000, 2000)
    sl = LongitudinalSlicing(nslices=3, method="normal_quantile")
    solver = PICPoissonSolver(grid=(32, 32), slicing=sl)
    cols = [PICCollider(solver), PICCollider(solver)]
    shared = Beam.from_host(p, PP, device=gpu)
    b1 = [Beam.from_host(e, PE, device=gpu), Beam.from_host(e, PE, device=gpu)]
    with pytest.raises(ArgumentError, match="shares a coordinate array"):
        collide_many(cols, b1, [shared, shared])
    assert np.array_equal(shared.to_host(), p) and np.array_equal(b1[0].to_host(), e)
    # distinct beams run
    lums = collide_many(cols, b1, [shared, Beam.from_host(p, PP, device=gpu)])
    assert lums[0] == lums[1]
