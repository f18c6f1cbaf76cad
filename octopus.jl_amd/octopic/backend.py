"""HIP backend: ctypes binding of liboctopic.so and the collide! entry points.

Mirrors the reference's dispatch (src/tasks/strongstrong/pic_cpu.jl:1-32, pic_cuda.jl:94-143):

    collide_(solver, beam1, beam2, HIPBackend, ctx=None) -> luminosity   # cold Green cache, as bare collide!
    PICCollider(solver).collide(beam1, beam2, turn)       -> luminosity  # persistent cache, as the task path

The library is required: every entry point raises BackendUnavailable when liboctopic.so is missing
or no GPU is visible.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from typing import Optional, Sequence

from . import _abi
from .solver import ArgumentError, BeamParams, PICPoissonSolver

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboctopic.so")
_lib = None


class BackendUnavailable(RuntimeError):
    """liboctopic.so is not built or cannot run here."""


class OctopicError(RuntimeError):
    """A device-side failure (HIP runtime, allocation)."""

    def __init__(self, code, msg):
        super().__init__(f"{_abi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class NonFiniteCoordinateError(ArgumentError):
    code = _abi.ERR_NONFINITE


class DegenerateExtentError(ArgumentError):
    code = _abi.ERR_DEGENERATE_EXTENT


class UnsupportedOptionError(ArgumentError):
    code = _abi.ERR_UNSUPPORTED


def lib():
    """Load liboctopic.so (raises BackendUnavailable if it is not built).  OCTOPIC_LIB, when set, names
    the library file to load instead of the in-tree build, as on the Julia side."""
    global _lib
    if _lib is None:
        path = os.environ.get("OCTOPIC_LIB") or LIB_PATH
        if not os.path.exists(path):
            raise BackendUnavailable(
                f"{path} is not built; run __graft_entry__.build() (or make -C octopus.jl_amd/csrc)")
        L = C.CDLL(path)
        _abi.declare(L)
        _lib = L
    return _lib


def _raise(code, msg):
    if code == _abi.ERR_NONFINITE:
        raise NonFiniteCoordinateError(msg)
    if code == _abi.ERR_DEGENERATE_EXTENT:
        raise DegenerateExtentError(msg)
    if code == _abi.ERR_UNSUPPORTED:
        raise UnsupportedOptionError(msg)
    if code in (_abi.ERR_INVALID_OPTION, _abi.ERR_INVALID_ARGUMENT):
        raise ArgumentError(msg)
    raise OctopicError(code, msg)


class HIPBackend:
    """Backend tag, the analogue of the reference's CPUThreadsBackend / CUDABackend (src/Octopus.jl:16-19)."""


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise BackendUnavailable("no GPU is visible to torch; the HIP backend needs one")
    return torch


@dataclass
class Beam:
    """Beam{HIPBackend}: Phase6DRep coordinates as one [6, n] float64 device tensor + BeamParams.

    Row order x, px, y, py, z, pz (src/beam/Beam.jl:45-56)."""

    coords: object
    params: BeamParams

    @property
    def n(self) -> int:
        return int(self.coords.shape[1])

    def __len__(self):
        return self.n

    def struct(self):
        b = _abi.BeamStruct()
        c = self.coords
        torch = _torch()
        if c.dtype not in (torch.float64, torch.float32) or not c.is_contiguous():
            raise ValueError("beam coordinates must be one contiguous float64 or float32 [6, n] tensor")
        base = c.data_ptr()
        stride = c.stride(0) * c.element_size()
        b.x, b.px, b.y, b.py, b.z, b.pz = (base + k * stride for k in range(6))
        b.n = self.n
        b.charge, b.mc2, b.E0, b.r0, b.npart = (float(self.params.charge), float(self.params.mc2),
                                               float(self.params.E0), float(self.params.r0),
                                               float(self.params.npart))
        b.on_device = 1
        b.dtype = _abi.OCTOPIC_DTYPE_F32 if c.dtype == torch.float32 else _abi.OCTOPIC_DTYPE_F64
        return b

    @staticmethod
    def from_host(coords, params: BeamParams, device="cuda", dtype=None):
        """dtype: torch.float64 (default; Phase6DRep{Float64}) or torch.float32 (Phase6DRep{Float32})."""
        torch = _torch()
        t = torch.as_tensor(coords, dtype=dtype or torch.float64).contiguous().to(device)
        assert t.ndim == 2 and t.shape[0] == 6
        return Beam(t, params)

    def to_host(self):
        return self.coords.detach().cpu().numpy()


class PICCollider:
    """A collide! handle with persistent device state: workspaces and the slice-pair Green cache in HBM,
    the reference's StrongStrongTask.runtime_cache (pic_cpu.jl:23-32, 467-515)."""

    def __init__(self, solver: PICPoissonSolver, device: int = 0):
        torch = _torch()
        self.solver = solver
        self.device = device
        o, keep = solver.to_options()
        self._keep = keep
        h = C.c_void_p()
        rc = lib().octopic_create(C.byref(o), int(device), C.byref(h))
        if rc != 0:
            _raise(rc, "octopic_create rejected the solver options (see stderr for the message)")
        self.h = h
        self.set_stream(torch.cuda.current_stream(device).cuda_stream)

    def close(self):
        if getattr(self, "h", None):
            lib().octopic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            _raise(rc, lib().octopic_last_error(self.h).decode())

    def set_stream(self, stream_ptr):
        """Enqueue on the given HIP stream (None: the handle's own stream)."""
        self._check(lib().octopic_set_stream(self.h, C.c_void_p(stream_ptr)))

    def collide(self, beam1: Beam, beam2: Beam, turn: Optional[int] = None) -> float:
        """collide!(solver, beam1, beam2, HIPBackend, ctx) with ctx.turn = turn (None for ctx = nothing)."""
        if beam1.coords.device != beam2.coords.device:
            raise ArgumentError("beam1 and beam2 must live on the same device")
        s1, s2 = beam1.struct(), beam2.struct()
        lum = C.c_double()
        cl = self.solver.compute_luminosity(turn)
        self._check(lib().octopic_collide(self.h, C.byref(s1), C.byref(s2), -1 if turn is None else int(turn),
                                          1 if cl else 0, C.byref(lum)))
        return lum.value

    LAUNCH_SHAPES = ("transverse_tiles", "kick_target_blocks", "kick_pass_windows", "deposit_span", "deposit_multi",
                     "deposit_min_span", "fft_transforms_per_block")

    def set_launch_shape(self, name: str, value: int):
        """A launch shape of this handle (results do not depend on it; see octopic_set_launch_shape)."""
        self._check(lib().octopic_set_launch_shape(self.h, name.encode(), int(value)))

    def launch_shape(self) -> dict:
        out = {}
        for n in self.LAUNCH_SHAPES:
            v = C.c_int64()
            self._check(lib().octopic_get_launch_shape(self.h, n.encode(), C.byref(v)))
            out[n] = v.value
        return out

    def autotune(self, beam1: Beam, beam2: Beam, trials: int = 3) -> dict:
        """Time candidate launch shapes on copies of the beams (octopic_autotune) and keep the fastest;
        returns the chosen shape and its collide time in ms. The beams and the Green cache are untouched."""
        s1, s2 = beam1.struct(), beam2.struct()
        ms = C.c_double()
        self._check(lib().octopic_autotune(self.h, C.byref(s1), C.byref(s2), int(trials), C.byref(ms)))
        out = self.launch_shape()
        out["collide_ms"] = ms.value
        return out

    def set_denoiser(self, denoiser, monitor_limit: float = 0.0):
        """Attach an octopic.denoise.Denoiser to this collision (None detaches): the field planes are
        then Gaussian-subtracted, denoised and rebuilt before the convolution (octopic_set_denoiser)."""
        if denoiser is None:
            self._denoiser = None
            self._check(lib().octopic_set_denoiser(self.h, None, 0.0))
            return
        self._check(lib().octopic_set_denoiser(self.h, denoiser._h, float(monitor_limit)))
        self._denoiser = denoiser   # the handle does not own it

    def stats(self):
        s = _abi.Stats()
        self._check(lib().octopic_get_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _abi.Stats._fields_}

    def pair_luminosity(self):
        n = C.c_int64()
        self._check(lib().octopic_pair_luminosity(self.h, None, None, None, None, 0, C.byref(n)))
        m = max(n.value, 1)
        i = (C.c_int32 * m)()
        j = (C.c_int32 * m)()
        r = (C.c_int32 * m)()
        v = (C.c_double * m)()
        self._check(lib().octopic_pair_luminosity(self.h, i, j, r, v, m, C.byref(n)))
        k = n.value
        return list(i)[:k], list(j)[:k], list(r)[:k], list(v)[:k]

    def interaction_records(self):
        n = C.c_int64()
        self._check(lib().octopic_interaction_records(self.h, None, 0, C.byref(n)))
        arr = (_abi.InteractionRecord * max(n.value, 1))()
        self._check(lib().octopic_interaction_records(self.h, arr, n.value, C.byref(n)))
        return list(arr)[: n.value]

    def last_slices(self, which: int, n: int):
        """(boundaries, centers, weights, counts, indices) of the last collide's slicing of beam `which`."""
        import numpy as np
        sl = self.solver.resolved_slicing1 if which == 1 else self.solver.resolved_slicing2
        ns = sl.nslices
        b = np.zeros(ns + 1)
        c = np.zeros(ns)
        w = np.zeros(ns)
        counts = np.zeros(ns, dtype=np.int64)
        idx = np.zeros(n, dtype=np.int64)
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int64)
        self._check(lib().octopic_last_slices(self.h, which, b.ctypes.data_as(D), c.ctypes.data_as(D),
                                              w.ctypes.data_as(D), counts.ctypes.data_as(I),
                                              idx.ctypes.data_as(I), n))
        return b, c, w, counts, idx

    def track_linear6d(self, beam: Beam, matrix36: Sequence[float]):
        m = (C.c_double * 36)(*[float(v) for v in matrix36])
        s = beam.struct()
        self._check(lib().octopic_track_linear6d(self.h, m, C.byref(s)))

    def track_lorentz_boost(self, beam: Beam, angle: float, reverse: bool = False):
        """LorentzBoost(angle) / RevLorentzBoost(angle) (lorentz_boost.jl:128-166) in place on the device;
        angle = half the full crossing angle."""
        s = beam.struct()
        self._check(lib().octopic_track_lorentz_boost(self.h, float(angle), 1 if reverse else 0, C.byref(s)))

    def track_crab_cavity(self, beam: Beam, frequency: float, strength_x: Sequence[float] = (),
                          strength_y: Sequence[float] = (), phase: Sequence[float] = ()):
        """ThinCrabCavity{N}(frequency; strengthX, strengthY, phase) (crab_cavity.jl:99-164) in place on the
        device; empty tuples are zeros as _harmonic_tuple (crab_cavity.jl:166-172)."""
        nh = max(len(strength_x), len(strength_y), len(phase))
        vals = []
        for v in (strength_x, strength_y, phase):
            if len(v) not in (0, nh):
                raise ArgumentError(f"expected {nh} harmonic values, got {len(v)}")
            vals.append([float(u) for u in v] if len(v) else [0.0] * nh)
        arr = [(C.c_double * max(nh, 1))(*v) for v in vals]
        s = beam.struct()
        self._check(lib().octopic_track_crab_cavity(self.h, float(frequency), nh, arr[0], arr[1], arr[2], C.byref(s)))

    @staticmethod
    def _be_args(x, y):
        import torch
        if (x.shape != y.shape or x.dtype != y.dtype or x.dtype not in (torch.float64, torch.float32)
                or not (x.is_cuda and y.is_cuda)):
            raise ArgumentError("x and y must be float64 or float32 device tensors of one shape and dtype")
        return x.contiguous(), y.contiguous(), x.dtype == torch.float32

    def gaussian_beambeam_kick(self, x, y, sigx: float, sigy: float):
        """gaussian_beambeam_kick(sigx, sigy, x, y) (strong_beam.jl:1139-1146) elementwise over device tensors
        x, y of one shape; returns (kx, ky), the normalised Bassetti-Erskine kick.  float32 tensors take the
        Float32 evaluation (the reference's Float32 constants, FP32 arithmetic)."""
        import torch
        x, y, f32 = self._be_args(x, y)
        kx, ky = torch.empty_like(x), torch.empty_like(y)
        fn = lib().octopic_gaussian_beambeam_kick_f32 if f32 else lib().octopic_gaussian_beambeam_kick
        self._check(fn(self.h, x.data_ptr(), y.data_ptr(), x.numel(), float(sigx), float(sigy), kx.data_ptr(),
                       ky.data_ptr()))
        return kx, ky

    def gaussian_beambeam_kick_response(self, x, y, kbb: float, sigx: float, sigy: float):
        """_gaussian_beambeam_kick_response(kbb, sigx, sigy, x, y) (strong_beam.jl:1119-1128) elementwise over
        device tensors x, y: a [n, 4] tensor of (Kx, Ky, Hxx, Hyy), the normalised kick and kbb times the
        potential's Hessian diagonal, of the inputs' dtype (float32: the Float32 evaluation)."""
        import torch
        x, y, f32 = self._be_args(x, y)
        x, y = x.reshape(-1), y.reshape(-1)
        out = torch.empty((x.numel(), 4), dtype=x.dtype, device=x.device)
        fn = lib().octopic_gaussian_beambeam_kick_response_f32 if f32 else lib().octopic_gaussian_beambeam_kick_response
        self._check(fn(self.h, x.data_ptr(), y.data_ptr(), x.numel(), float(kbb), float(sigx), float(sigy),
                       out.data_ptr()))
        return out

    def faddeeva_w(self, z):
        """w(z) = exp(-z^2) erfc(-iz) (SpecialMath.jl:60-151) elementwise over a complex128 device tensor, or a
        complex64 one in FP32 arithmetic."""
        import torch
        if z.dtype not in (torch.complex128, torch.complex64) or not z.is_cuda:
            raise ArgumentError("z must be a complex128 or complex64 device tensor")
        zr, zi = z.real.contiguous(), z.imag.contiguous()
        wr, wi = torch.empty_like(zr), torch.empty_like(zi)
        fn = lib().octopic_faddeeva_w_f32 if z.dtype == torch.complex64 else lib().octopic_faddeeva_w
        self._check(fn(self.h, zr.data_ptr(), zi.data_ptr(), zr.numel(), wr.data_ptr(), wi.data_ptr()))
        return torch.complex(wr, wi)

    def set_timing(self, enabled: bool, families: Optional[str] = None):
        """Per-family HIP-event timing: every family, or only the comma-separated `families`."""
        if enabled and families:
            self._check(lib().octopic_set_timing_families(self.h, families.encode()))
        else:
            self._check(lib().octopic_set_timing(self.h, 1 if enabled else 0))

    def kernel_times(self):
        """{family: (total_ms, launches)} accumulated since set_timing(True)."""
        n = C.c_int64()
        ms = (C.c_double * 32)()
        cnt = (C.c_int64 * 32)()
        names = C.create_string_buffer(1024)
        self._check(lib().octopic_kernel_times(self.h, ms, cnt, 32, C.byref(n), names, 1024))
        keys = names.raw.split(b"\0")[: n.value]
        return {k.decode(): (ms[i], cnt[i]) for i, k in enumerate(keys)}

    def debug_weights(self, method: str, u, npts: int):
        """Device CIC/TSC stencil (0-based base, weights) for normalised coordinates u."""
        import numpy as np
        u = np.ascontiguousarray(u, dtype=np.float64)
        base = np.zeros(len(u), dtype=np.int32)
        w = np.zeros((len(u), 3))
        self._check(lib().octopic_debug_weights(self.h, _abi.DEPOSIT[method], u.ctypes.data_as(C.POINTER(C.c_double)),
                                                len(u), int(npts), base.ctypes.data_as(C.POINTER(C.c_int32)),
                                                w.ctypes.data_as(C.POINTER(C.c_double))))
        return base, w


def generate_beam(collider: PICCollider, n: int, params: BeamParams, beta, alpha, sigma, seed: int, rng_id: int,
                  cutoff: float = 5.0, device="cuda") -> Beam:
    """Beam(N, HIPBackend; beta, alpha, sigma, cutoff, rng_id) built on the device from the Philox counter RNG
    (Beam.jl:127-181, 466-515, 545-553)."""
    torch = _torch()
    t = torch.empty((6, n), dtype=torch.float64, device=device)
    beam = Beam(t, params)
    s = beam.struct()
    D3 = C.c_double * 3
    rc = lib().octopic_generate_gaussian_beam(collider.h, int(seed), int(rng_id), D3(*beta), D3(*alpha), D3(*sigma),
                                              float(cutoff), C.byref(s))
    collider._check(rc)
    return beam


def collide_many(colliders: Sequence["PICCollider"], beams1: Sequence[Beam], beams2: Sequence[Beam],
                 turns: Optional[Sequence[Optional[int]]] = None):
    """Independent bunch-pair instances collided together: instance k is colliders[k].collide(beams1[k],
    beams2[k], turns[k]), each on its handle's own stream from its own host thread, so that the
    instances' kernels overlap on the GPU.  Returns the luminosities; raises the first instance's error."""
    torch = _torch()
    n = len(colliders)
    if not (len(beams1) == len(beams2) == n) or len(set(id(c) for c in colliders)) != n:
        raise ArgumentError("collide_many needs one distinct collider and two beams per instance")
    turns = list(turns) if turns is not None else [None] * n
    s1 = (_abi.BeamStruct * n)(*[b.struct() for b in beams1])
    s2 = (_abi.BeamStruct * n)(*[b.struct() for b in beams2])
    hs = (C.c_void_p * n)(*[c.h for c in colliders])
    tv = (C.c_int64 * n)(*[-1 if t is None else int(t) for t in turns])
    cl = (C.c_int32 * n)(*[1 if c.solver.compute_luminosity(t) else 0 for c, t in zip(colliders, turns)])
    lums = (C.c_double * n)()
    st = (C.c_int32 * n)()
    torch.cuda.synchronize()      # the handles' streams run beside torch's
    for c in colliders:
        c.set_stream(None)
    try:
        rc = lib().octopic_collide_many(hs, s1, s2, tv, cl, lums, st, n)
    finally:
        for c in colliders:
            c.set_stream(torch.cuda.current_stream(c.device).cuda_stream)
    for k in range(n):
        if st[k] != 0:
            colliders[k]._check(st[k])
    if rc != 0:   # rejected before any instance ran: the message is on the instance that was checked
        msgs = [lib().octopic_last_error(c.h).decode() for c in colliders]
        _raise(rc, "; ".join(m for m in msgs if m) or "octopic_collide_many rejected its arguments")
    return list(lums)


def collide_(solver: PICPoissonSolver, beam1: Beam, beam2: Beam, backend=HIPBackend, ctx=None) -> float:
    """collide!(solver, beam1, beam2, HIPBackend[, ctx]): a bare call uses a cold Green cache, as the
    reference's bare collide! does (pic_cpu.jl:15-21).  ctx may be None or an int turn."""
    if backend is not HIPBackend:
        raise ArgumentError("octopic.collide_ runs the HIP backend only")
    col = PICCollider(solver, device=beam1.coords.device.index or 0)
    try:
        return col.collide(beam1, beam2, turn=ctx)
    finally:
        col.close()
