"""Python handle of the residual denoiser in liboctopic (octopic_denoiser_* in include/octopic.h).

The network and its training run on the device through the C ABI; torch tensors are only the
device buffers handed to it. There is no host fallback: without the library this raises.
"""
import ctypes as C

import numpy as np

from . import _abi
from .backend import BackendUnavailable, _raise, _torch, lib


class Denoiser:
    """U-Net residual denoiser on [batch, height, width, 2] planes (see include/octopic.h)."""

    def __init__(self, height: int = 64, width: int = 64, max_batch: int = 32, seed: int = 0, device: int = 0):
        self._torch = _torch()
        L = lib()
        h = C.c_void_p()
        rc = L.octopic_denoiser_create(device, height, width, max_batch, seed, C.byref(h))
        if rc != _abi.OK:
            _raise(rc, f"octopic_denoiser_create({height}, {width}, {max_batch}) failed")
        self._h = h
        self.height, self.width, self.max_batch, self.device = height, width, max_batch, device
        self.n_params = int(L.octopic_denoiser_num_params())
        self._loss = self._torch.zeros(1, dtype=self._torch.float32, device=f"cuda:{device}")

    def close(self):
        if getattr(self, "_h", None):
            lib().octopic_denoiser_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != _abi.OK:
            msg = lib().octopic_denoiser_last_error(self._h)
            _raise(rc, msg.decode() if msg else "denoiser error")

    def _plane_args(self, x, channels):
        t = self._torch
        if not (x.is_cuda and x.dtype == t.float32 and x.is_contiguous()):
            raise ValueError("planes must be contiguous float32 device tensors")
        want = (self.height, self.width) + ((channels,) if channels > 1 else ())
        if x.dim() != 1 + len(want) or tuple(x.shape[1:]) != want:
            raise ValueError(f"expected [batch, {', '.join(map(str, want))}], got {tuple(x.shape)}")
        return x.shape[0]

    def set_stream(self, stream_ptr: int):
        self._check(lib().octopic_denoiser_set_stream(self._h, C.c_void_p(stream_ptr)))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self._check(lib().octopic_denoiser_get_params(self._h, out.ctypes.data_as(_abi.FPTR), self.n_params))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        self._check(lib().octopic_denoiser_set_params(self._h, p.ctypes.data_as(_abi.FPTR), self.n_params))

    def forward(self, x):
        b = self._plane_args(x, 2)
        out = self._torch.empty((b, self.height, self.width), dtype=self._torch.float32, device=x.device)
        self._check(lib().octopic_denoiser_forward(self._h, x.data_ptr(), b, out.data_ptr()))
        return out

    def gradient(self, x, target):
        b = self._plane_args(x, 2)
        self._plane_args(target, 1)
        g = self._torch.empty(self.n_params, dtype=self._torch.float32, device=x.device)
        self._check(lib().octopic_denoiser_gradient(self._h, x.data_ptr(), target.data_ptr(), b, self._loss.data_ptr(),
                                                    g.data_ptr()))
        return self._loss.clone(), g

    def train_step(self, x, target, lr: float, want_grad: bool = False):
        """One Adam step; returns the loss as a device tensor (no host synchronisation)."""
        b = self._plane_args(x, 2)
        self._plane_args(target, 1)
        g = self._torch.empty(self.n_params, dtype=self._torch.float32, device=x.device) if want_grad else None
        self._check(lib().octopic_denoiser_train_step(self._h, x.data_ptr(), target.data_ptr(), b, float(lr),
                                                      self._loss.data_ptr(), g.data_ptr() if g is not None else None))
        return (self._loss.clone(), g) if want_grad else self._loss.clone()

    def make_pairs(self, seed: int, first: int, batch: int, n_target: int = 16_000_000, method: str = "TSC",
                   raw: bool = False):
        """Training pairs generated on the device (octopic_denoiser_make_pairs): returns
        (input [batch, H, W, 2], target [batch, H, W], monitor [batch]) and, with raw=True, also the
        deposits ([batch, 2, H*W] int64 in 2^-32 units) and moment sums ([batch, 2, 5] int64)."""
        t = self._torch
        dev = f"cuda:{self.device}"
        x = t.empty((batch, self.height, self.width, 2), dtype=t.float32, device=dev)
        y = t.empty((batch, self.height, self.width), dtype=t.float32, device=dev)
        mon = t.empty(batch, dtype=t.float32, device=dev)
        planes = t.empty((batch, 2, self.height * self.width), dtype=t.int64, device=dev) if raw else None
        mom = t.empty((batch, 2, 5), dtype=t.int64, device=dev) if raw else None
        self._check(lib().octopic_denoiser_make_pairs(
            self._h, seed, first, batch, n_target, {"CIC": 0, "TSC": 1}[method], x.data_ptr(), y.data_ptr(),
            mon.data_ptr(), planes.data_ptr() if raw else None, mom.data_ptr() if raw else None))
        return (x, y, mon, planes, mom) if raw else (x, y, mon)


__all__ = ["Denoiser", "BackendUnavailable"]
