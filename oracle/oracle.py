"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker.  The product (octopus.jl_amd/) never imports this module.

Arrays follow the reference's layout: grids are column-major (x fastest), index lists 0-based.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "octopus.jl_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from octopic import _abi  # noqa: E402  (struct layouts of include/octopic.h)

LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_abi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        D = C.POINTER(C.c_double)
        I64 = C.POINTER(C.c_int64)
        I32 = C.POINTER(C.c_int32)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_counter_uint64.restype = C.c_uint64
        L.oracle_counter_uint64.argtypes = [C.c_uint64] * 5
        L.oracle_octopus_normal.restype = C.c_double
        L.oracle_octopus_normal.argtypes = [C.c_uint64] * 4 + [C.c_int64]
        L.oracle_erfinv.restype = C.c_double
        L.oracle_erfinv.argtypes = [C.c_double]
        L.oracle_faddeeva_w.restype = None
        L.oracle_faddeeva_w.argtypes = [C.c_double, C.c_double, D, D]
        L.oracle_gaussian_beambeam_kick.restype = None
        L.oracle_gaussian_beambeam_kick.argtypes = [C.c_double] * 4 + [D, D]
        L.oracle_weideman_coefficients.restype = None
        L.oracle_weideman_coefficients.argtypes = [D, D]
        L.oracle_kernel_integral.restype = C.c_double
        L.oracle_kernel_integral.argtypes = [C.c_double, C.c_double]
        L.oracle_atan_ratio.restype = C.c_double
        L.oracle_atan_ratio.argtypes = [C.c_double, C.c_double]
        L.oracle_weights.argtypes = [C.c_int32, C.c_double, C.c_int64, I64, D]
        L.oracle_create.argtypes = [C.POINTER(_abi.Options), C.POINTER(C.c_void_p)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_set_gaussian_replace.argtypes = [C.c_void_p, C.c_int32]
        L.oracle_collide.argtypes = [C.c_void_p, C.POINTER(D), C.c_int64, D, C.POINTER(D), C.c_int64, D,
                                     C.c_int32, C.c_int32, D]
        L.oracle_stats.argtypes = [C.c_void_p, I64]
        L.oracle_npairs.restype = C.c_int64
        L.oracle_npairs.argtypes = [C.c_void_p]
        L.oracle_pair_luminosity.argtypes = [C.c_void_p, I32, I32, I32, D]
        L.oracle_records.argtypes = [C.c_void_p, C.POINTER(_abi.InteractionRecord), C.c_int64, I64]
        L.oracle_gaussian_beambeam_kick_response.argtypes = [C.c_double] * 5 + [D]
        L.oracle_gaussian_beambeam_kick_response.restype = None
        L.oracle_node_mesh.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, D]
        L.oracle_node_mesh.restype = C.c_int32
        L.oracle_last_slices.argtypes = [C.c_void_p, C.c_int32, D, D, D, I64, I64]
        L.oracle_slices.argtypes = [D, C.c_int64, C.POINTER(_abi.Slicing), D, D, D, I64, I64]
        L.oracle_generate_beam.argtypes = [C.c_int64, C.c_uint64, C.c_uint64, D, D, D, C.c_double, C.POINTER(D)]
        L.oracle_green_table.argtypes = [C.c_int32] + [C.c_double] * 6 + [C.c_int32, C.c_int32, D]
        L.oracle_fft2.argtypes = [D, C.c_int32, C.c_int32, C.c_int32]
        L.oracle_field.argtypes = [D, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32, D, D]
        L.oracle_interaction_grids.argtypes = [C.POINTER(_abi.Options), D, D, D]
        L.oracle_deposit.argtypes = [D, C.c_int64, C.c_int64, C.c_int32, D, D, C.c_int64, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int64, C.c_int64]
        L.oracle_solve_field.argtypes = [C.POINTER(_abi.Options), D, D, D, D, C.c_int64, C.c_double, D, D, D, D, D, D]
        L.oracle_interpolate_kick.argtypes = [C.POINTER(_abi.Options), D, C.c_double, C.c_double, D, D, D, D, D, D,
                                              C.c_double, C.c_double, D]
        L.oracle_interpolate_kick_quadratic.argtypes = [C.POINTER(_abi.Options), D, C.c_double, C.c_double,
                                                        C.POINTER(D), C.c_double, D]
        L.oracle_luminosity.argtypes = [C.POINTER(_abi.Options), D, D, C.c_int64, D, D, C.c_int64, C.c_double, D, D, D]
        L.oracle_collision_batches.argtypes = [D, C.c_int32, D, C.c_int32, I32, I32, I32, I32]
        L.oracle_collision_order.argtypes = [D, C.c_int32, D, C.c_int32, I32, I32]
        L.oracle_linear6d_matrix.argtypes = [D, D, D, D, D, D]
        L.oracle_track_linear6d.argtypes = [D, C.POINTER(D), C.c_int64]
        L.oracle_track_lorentz_boost.argtypes = [C.c_double, C.c_int, C.POINTER(D), C.c_int64]
        L.oracle_track_crab_cavity.argtypes = [C.c_double, C.c_int, D, D, D, C.POINTER(D), C.c_int64]
        L.oracle_philox_block.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib().oracle_last_error().decode())


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _i32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- scalar kernels
def philox_block(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox_block(c, k, o)
    return tuple(o)


def counter_uint64(seed, turn, rng_id, particle, component):
    return lib().oracle_counter_uint64(seed, turn, rng_id, particle, component)


def octopus_normal(seed, turn, rng_id, particle, component):
    return lib().oracle_octopus_normal(seed, turn, rng_id, particle, component)


def erfinv(y):
    return lib().oracle_erfinv(float(y))


def weights(method, u, n):
    b = C.c_int64()
    w = (C.c_double * 3)()
    lib().oracle_weights(_abi.DEPOSIT[method] if isinstance(method, str) else method, float(u), n, C.byref(b), w)
    k = 2 if (method == "CIC" or method == 0) else 3
    return int(b.value), tuple(w[:k])


def faddeeva_w(z):
    """w(z) as SpecialMath.jl:60-151 evaluates it (Weideman N = 32 / continued fraction)."""
    wr, wi = C.c_double(), C.c_double()
    lib().oracle_faddeeva_w(float(z.real), float(z.imag), C.byref(wr), C.byref(wi))
    return complex(wr.value, wi.value)


def gaussian_beambeam_kick(sigx, sigy, x, y):
    """(Kx, Ky) of strong_beam.jl:1139-1146."""
    kx, ky = C.c_double(), C.c_double()
    lib().oracle_gaussian_beambeam_kick(float(sigx), float(sigy), float(x), float(y), C.byref(kx), C.byref(ky))
    return kx.value, ky.value


def gaussian_beambeam_kick_response(kbb, sigx, sigy, x, y):
    """_gaussian_beambeam_kick_response(kbb, sigx, sigy, x, y) (strong_beam.jl:1119-1128): (Kx, Ky), the
    normalised kick, and (Hxx, Hyy), the diagonal of the potential's Hessian times kbb."""
    out = np.zeros(4)
    lib().oracle_gaussian_beambeam_kick_response(float(kbb), float(sigx), float(sigy), float(x), float(y), _d(out))
    return tuple(out)


def weideman_coefficients():
    c = (C.c_double * 32)()
    L = C.c_double()
    lib().oracle_weideman_coefficients(c, C.byref(L))
    return list(c), L.value


def kernel_integral(x, y):
    return lib().oracle_kernel_integral(float(x), float(y))


def atan_ratio(n, d):
    return lib().oracle_atan_ratio(float(n), float(d))


def green_table(green_type, fx0, fy0, sx0, sy0, hx, hy, nx, ny):
    out = np.zeros((2 * ny, 2 * nx))  # C order (j, i) == column-major (i, j)
    lib().oracle_green_table(_abi.GREEN_TYPE[green_type], fx0, fy0, sx0, sy0, hx, hy, nx, ny, _d(out))
    return out.T.copy()  # [i, j]


def fft2(a, inverse=False):
    """2D DFT of a [n1, n2] complex array (FFTW convention: forward unnormalised, inverse 1/(n1 n2))."""
    a = np.asarray(a, dtype=np.complex128)
    n1, n2 = a.shape
    buf = np.ascontiguousarray(a.T)  # column-major storage of a
    lib().oracle_fft2(buf.ctypes.data_as(C.POINTER(C.c_double)), n1, n2, 1 if inverse else -1)
    return buf.T.copy()


def field(phi, hx, hy, fourth):
    phi = np.asarray(phi, dtype=np.float64)
    nx, ny = phi.shape
    p = np.ascontiguousarray(phi.T)
    ex = np.zeros_like(p)
    ey = np.zeros_like(p)
    lib().oracle_field(_d(p), nx, ny, hx, hy, 1 if fourth else 0, _d(ex), _d(ey))
    return ex.T.copy(), ey.T.copy()


def interaction_grids(solver, b):
    o, keep = solver.to_options()
    bb = _f64(b)
    sg = np.zeros(4)
    fg = np.zeros(4)
    _check(lib().oracle_interaction_grids(C.byref(o), _d(bb), _d(sg), _d(fg)))
    return tuple(sg), tuple(fg)


def deposit(method, x, y, x0, y0, hx, hy, nx, ny, ld=None, ncols=None):
    """_pic_deposit! onto an (ld x ncols) grid; returns array [i, j]."""
    ld = nx if ld is None else ld
    ncols = ny if ncols is None else ncols
    x, y = _f64(x), _f64(y)
    q = np.zeros((ncols, ld))
    lib().oracle_deposit(_d(q), ld, ncols, _abi.DEPOSIT[method], _d(x), _d(y), len(x), x0, y0, hx, hy, nx, ny)
    return q.T.copy()


def solve_field(solver, x, px, y, py, s, sg, fg):
    o, keep = solver.to_options()
    nx, ny = solver.grid
    x, px, y, py = (_f64(v) for v in (x, px, y, py))
    phi = np.zeros((ny, nx))
    ex = np.zeros_like(phi)
    ey = np.zeros_like(phi)
    q = np.zeros((2 * ny, 2 * nx))
    sgv, fgv = _f64(sg), _f64(fg)
    _check(lib().oracle_solve_field(C.byref(o), _d(x), _d(px), _d(y), _d(py), len(x), float(s), _d(sgv), _d(fgv),
                                    _d(phi), _d(ex), _d(ey), _d(q)))
    return phi.T.copy(), ex.T.copy(), ey.T.copy(), q.T.copy()


def interpolate_kick(solver, grid, x, y, L, R, zL, zR):
    o, keep = solver.to_options()
    g = _f64(grid)
    planes = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).T) for a in (*L, *R)]
    K = np.zeros(3)
    _check(lib().oracle_interpolate_kick(C.byref(o), _d(g), float(x), float(y), *[_d(p) for p in planes],
                                         float(zL), float(zR), _d(K)))
    return tuple(K)


def interpolate_kick_quadratic(solver, grid, x, y, L, M, R, t):
    o, keep = solver.to_options()
    g = _f64(grid)
    planes = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).T) for a in (*L, *M, *R)]
    arr = (C.POINTER(C.c_double) * 9)(*[_d(p) for p in planes])
    K = np.zeros(3)
    _check(lib().oracle_interpolate_kick_quadratic(C.byref(o), _d(g), float(x), float(y), arr, float(t), _d(K)))
    return tuple(K)


def luminosity(solver, x1, y1, x2, y2, klum):
    o, keep = solver.to_options()
    lnx, lny = solver.resolved_luminosity_grid
    x1, y1, x2, y2 = (_f64(v) for v in (x1, y1, x2, y2))
    q1 = np.zeros((lny + 1, lnx + 1))
    q2 = np.zeros_like(q1)
    lum = C.c_double()
    _check(lib().oracle_luminosity(C.byref(o), _d(x1), _d(y1), len(x1), _d(x2), _d(y2), len(x2), float(klum),
                                   C.byref(lum), _d(q1), _d(q2)))
    return lum.value, q1.T.copy(), q2.T.copy()


def slices(z, slicing):
    """longitudinal_slices(rep, slicing) -> (boundaries, centers, weights, [index arrays])."""
    z = _f64(z)
    keep = []
    s = slicing._lower(keep)
    ns = slicing.nslices
    b = np.zeros(ns + 1)
    c = np.zeros(ns)
    w = np.zeros(ns)
    counts = np.zeros(ns, dtype=np.int64)
    idx = np.zeros(len(z), dtype=np.int64)
    _check(lib().oracle_slices(_d(z), len(z), C.byref(s), _d(b), _d(c), _d(w), _i64(counts), _i64(idx)))
    lists, off = [], 0
    for k in range(ns):
        lists.append(idx[off:off + counts[k]].copy())
        off += counts[k]
    return b, c, w, lists


def collision_order(c1, c2):
    c1, c2 = _f64(c1), _f64(c2)
    n = len(c1) * len(c2)
    oi = np.zeros(n, dtype=np.int32)
    oj = np.zeros(n, dtype=np.int32)
    lib().oracle_collision_order(_d(c1), len(c1), _d(c2), len(c2), _i32(oi), _i32(oj))
    return list(zip(oi.tolist(), oj.tolist()))


def collision_pair_batches(c1, c2):
    c1, c2 = _f64(c1), _f64(c2)
    n = len(c1) * len(c2)
    oi = np.zeros(n, dtype=np.int32)
    oj = np.zeros(n, dtype=np.int32)
    ob = np.zeros(n, dtype=np.int32)
    nb = C.c_int32()
    _check(lib().oracle_collision_batches(_d(c1), len(c1), _d(c2), len(c2), _i32(oi), _i32(oj), _i32(ob),
                                          C.byref(nb)))
    batches = [[] for _ in range(nb.value)]
    for i, j, b in zip(oi, oj, ob):
        batches[b].append((int(i), int(j)))
    return batches


def generate_beam(n, seed, rng_id, beta, alpha, sigma, cutoff=5.0):
    """Beam(N, ...) coordinates [6, n] from the Philox counter RNG (Beam.jl:127-181, 545-553)."""
    out = np.zeros((6, n))
    ptrs = (C.POINTER(C.c_double) * 6)(*[_d(out[k]) for k in range(6)])
    _check(lib().oracle_generate_beam(n, seed, rng_id, _d(_f64(beta)), _d(_f64(alpha)), _d(_f64(sigma)),
                                      float(cutoff), ptrs))
    return out


def linear6d_matrix(beta1, beta2, alpha1, alpha2, dmu):
    out = np.zeros(36)
    lib().oracle_linear6d_matrix(*[_d(_f64(v)) for v in (beta1, beta2, alpha1, alpha2, dmu)], _d(out))
    return out


CLIGHT = 299_792_458.0   # src/constants/Constants.jl:8


def track_lorentz_boost(angle, coords, reverse=False):
    """LorentzBoost / RevLorentzBoost (lorentz_boost.jl:128-166) of coords [6, n] in place; angle =
    half the full crossing angle."""
    ptrs = (C.POINTER(C.c_double) * 6)(*[_d(coords[k]) for k in range(6)])
    lib().oracle_track_lorentz_boost(C.c_double(angle), 1 if reverse else 0, ptrs, coords.shape[1])


def track_crab_cavity(frequency, strength_x, strength_y, phase, coords):
    """ThinCrabCavity{N} (crab_cavity.jl:99-164) of coords [6, n] in place."""
    kcc = 2.0 * math.pi * float(frequency) / CLIGHT
    nh = len(phase)
    ptrs = (C.POINTER(C.c_double) * 6)(*[_d(coords[k]) for k in range(6)])
    lib().oracle_track_crab_cavity(C.c_double(kcc), nh, _d(_f64(strength_x)), _d(_f64(strength_y)), _d(_f64(phase)),
                                   ptrs, coords.shape[1])


def track_linear6d(m, coords):
    """coords: float64 [6, n] C-contiguous, modified in place."""
    m = _f64(m)
    ptrs = (C.POINTER(C.c_double) * 6)(*[_d(coords[k]) for k in range(6)])
    lib().oracle_track_linear6d(_d(m), ptrs, coords.shape[1])


class Collider:
    """A collide! with the persistent task state (Green cache + workspaces)."""

    def __init__(self, solver, gaussian_replace=False):
        """gaussian_replace: test mode, every field deposit replaced by its neutralised drifted Gaussian."""
        self.solver = solver
        o, keep = solver.to_options()
        self._keep = keep
        h = C.c_void_p()
        self.h = None
        _check(lib().oracle_create(C.byref(o), C.byref(h)))
        self.h = h
        if gaussian_replace:
            lib().oracle_set_gaussian_replace(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def collide(self, c1, p1, c2, p2, compute_luminosity=True, nthreads=1):
        """c1/c2: float64 [6, n] C-contiguous coordinate arrays (modified in place); p1/p2 BeamParams."""
        assert c1.dtype == np.float64 and c1.flags.c_contiguous and c1.shape[0] == 6
        assert c2.dtype == np.float64 and c2.flags.c_contiguous and c2.shape[0] == 6
        P = C.POINTER(C.c_double) * 6
        b1 = P(*[_d(c1[k]) for k in range(6)])
        b2 = P(*[_d(c2[k]) for k in range(6)])
        pa = np.array([p1.charge, p1.mc2, p1.E0, p1.r0, p1.npart])
        pb = np.array([p2.charge, p2.mc2, p2.E0, p2.r0, p2.npart])
        lum = C.c_double()
        _check(lib().oracle_collide(self.h, b1, c1.shape[1], _d(pa), b2, c2.shape[1], _d(pb),
                                    1 if compute_luminosity else 0, int(nthreads), C.byref(lum)))
        return lum.value

    def stats(self):
        out = np.zeros(6, dtype=np.int64)
        lib().oracle_stats(self.h, _i64(out))
        return dict(hits=int(out[0]), misses=int(out[1]), rebuilds=int(out[2]), dropped=int(out[3]),
                    batches=int(out[4]), widest_batch=int(out[5]))

    def pair_luminosity(self):
        n = lib().oracle_npairs(self.h)
        i = np.zeros(n, dtype=np.int32)
        j = np.zeros(n, dtype=np.int32)
        r = np.zeros(n, dtype=np.int32)
        lum = np.zeros(n)
        lib().oracle_pair_luminosity(self.h, _i32(i), _i32(j), _i32(r), _d(lum))
        return i, j, r, lum

    def records(self):
        n = C.c_int64()
        lib().oracle_records(self.h, None, 0, C.byref(n))
        arr = (_abi.InteractionRecord * max(n.value, 1))()
        lib().oracle_records(self.h, arr, n.value, C.byref(n))
        return list(arr)[: n.value]

    def node_mesh(self, direction, slice_index, node):
        """The last collide's :node mesh of `node` (0-based boundary of the field beam's slicing) for source
        slice `slice_index` of direction `direction` (1: beam 1 is the source): (source grid, field grid) as
        (x0, y0, width, height) each, or None when the node has no mesh (_pic_build_node_grids!,
        pic_cpu.jl:1317-1382)."""
        out = np.zeros(8)
        if not lib().oracle_node_mesh(self.h, int(direction), int(slice_index), int(node), _d(out)):
            return None
        return tuple(out[:4]), tuple(out[4:])

    def last_slices(self, which, n):
        sl = self.solver.resolved_slicing1 if which == 1 else self.solver.resolved_slicing2
        ns = sl.nslices
        b = np.zeros(ns + 1)
        c = np.zeros(ns)
        w = np.zeros(ns)
        counts = np.zeros(ns, dtype=np.int64)
        idx = np.zeros(n, dtype=np.int64)
        lib().oracle_last_slices(self.h, which, _d(b), _d(c), _d(w), _i64(counts), _i64(idx))
        return b, c, w, counts, idx
