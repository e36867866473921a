"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (strict FP64 parity build) or liboracle_fast.so (timing build).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use this module,
and only as the checker or the timed baseline; the product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from cil_cuda_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
RNG_COUNTER, RNG_EXTERNAL, RNG_DSFMT = 0, 1, 2
_libs = {}


def lib(fast: bool = False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle not built: {path} (make -C oracle)")
    L = C.CDLL(path)
    vp = C.c_void_p
    PV = C.POINTER(abi.SpeciesView)
    L.or_create.argtypes = [C.POINTER(abi.Params), C.POINTER(abi.Box)]
    L.or_create.restype = vp
    L.or_destroy.argtypes = [vp]
    L.or_set_rng.argtypes = [vp, C.c_int, C.c_uint32]
    L.or_rng_draw.argtypes = [vp, C.c_int]
    L.or_rng_draw.restype = C.c_double
    L.or_set_external_noise.argtypes = [vp, abi.PD, C.c_size_t, C.c_int]
    L.or_set_traversal.argtypes = [vp, C.c_int]
    L.or_set_threads.argtypes = [vp, C.c_int]
    L.or_set_sample.argtypes = [vp, C.c_int]
    L.or_set_division_log.argtypes = [vp, C.c_int]
    L.or_division_log.argtypes = [vp, C.c_int, abi.PD, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int]
    L.or_forces.argtypes = [vp, PV, C.c_int, C.c_int]
    L.or_step.argtypes = [vp, PV, C.c_int, C.c_int, abi.PI]
    L.or_update.argtypes = [vp, PV, C.c_int, C.c_int]
    L.or_coms.argtypes = [vp, PV, C.c_int, abi.PD]
    L.or_divisions.argtypes = [vp, PV, C.c_int, C.c_int]
    L.or_check_positions.argtypes = [vp, PV, C.c_int, abi.PI]
    L.or_pairs.argtypes = [vp, PV, C.c_int, C.c_int, C.POINTER(C.c_int32), C.c_long]
    L.or_pairs.restype = C.c_long
    L.or_force_scale.argtypes = [vp, PV, C.c_int, abi.PD]
    L.or_force_scale_gather.argtypes = [vp, PV, C.c_int, abi.PD]
    L.or_pair_f0.argtypes = [C.POINTER(abi.Params), C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                             abi.PD]
    L.or_linklist_info.argtypes = [vp, abi.PI, abi.PD]
    L.or_last_error.argtypes = [vp]
    L.or_last_error.restype = C.c_char_p
    L.or_sizeof.argtypes = [C.c_int]
    L.or_sizeof.restype = C.c_size_t
    L.or_dsfmt_fill.argtypes = [C.c_uint32, C.c_int, abi.PD, C.c_size_t]
    L.or_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.or_philox_normal2.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, abi.PD]
    L.or_philox_uniform2.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, abi.PD]
    for k, cls in {0: abi.Params, 1: abi.Species, 2: abi.SpeciesView, 3: abi.Box}.items():
        if L.or_sizeof(k) != C.sizeof(cls):
            raise RuntimeError(f"oracle ABI layout mismatch for {cls.__name__}")
    _libs[name] = L
    return L


class OracleError(RuntimeError):
    pass


class Oracle:
    """CPU restatement of the reference Euler step on a host Colony."""

    def __init__(self, params: abi.Params, box: abi.Box, fast=False, traversal=0, threads=1):
        self.L = lib(fast)
        self.params = params
        self.box = box
        self.w = self.L.or_create(C.byref(params), C.byref(box))
        self.L.or_set_traversal(self.w, traversal)
        self.L.or_set_threads(self.w, threads)
        if params.rng == abi.RNG_EXTERNAL:
            self.L.or_set_rng(self.w, RNG_EXTERNAL, 0)

    def close(self):
        if self.w:
            self.L.or_destroy(self.w)
            self.w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise OracleError(f"{what}: {rc} {self.L.or_last_error(self.w).decode()}")

    def set_sample(self, every):
        """Threaded gather only: pair forces of every `every`-th cell (by UID) only."""
        self.L.or_set_sample(self.w, int(every))

    def set_division_log(self, on=True):
        self.L.or_set_division_log(self.w, 1 if on else 0)

    def division_log(self, s, dim, nelem, clear=True):
        """(positions [n][nelem][dim], steps [n]) of species s's recorded divisions."""
        n = C.c_int()
        self.L.or_division_log(self.w, s, abi.PD(), None, 0, C.byref(n), 0)
        pos = np.zeros((max(n.value, 1), nelem, dim))
        ts = np.zeros(max(n.value, 1), dtype=np.int32)
        self.L.or_division_log(self.w, s, pos.ctypes.data_as(abi.PD), ts.ctypes.data_as(C.POINTER(C.c_int)), n.value,
                               C.byref(n), 1 if clear else 0)
        return pos[:n.value], ts[:n.value]

    def set_rng(self, mode, seed=0):
        self.L.or_set_rng(self.w, mode, seed)

    def rng_draw(self, kind):
        """One draw from the sequential dSFMT stream: 0 uniform in [0, 1), 1 normal."""
        return self.L.or_rng_draw(self.w, kind)

    def set_external_noise(self, normals):
        normals = np.ascontiguousarray(normals, dtype=np.float64)
        self._chk(self.L.or_set_external_noise(self.w, normals.ctypes.data_as(abi.PD), normals.shape[1],
                                               normals.shape[0]), "external noise")

    def forces(self, col, ts=0):
        v = col.views()
        self._chk(self.L.or_forces(self.w, v, col.nspec, ts), "forces")
        col.pull_counts(v)
        return col

    def step(self, col, nsteps, ts):
        v = col.views()
        t = C.c_int(ts)
        self._chk(self.L.or_step(self.w, v, col.nspec, nsteps, C.byref(t)), "step")
        col.pull_counts(v)
        return t.value

    def pairs(self, col, mode=0):
        v = col.views()
        n = self.L.or_pairs(self.w, v, col.nspec, mode, None, 0)
        buf = np.zeros(2 * max(n, 1), dtype=np.int32)
        self.L.or_pairs(self.w, v, col.nspec, mode, buf.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return buf[: 2 * n].reshape(-1, 2)

    def force_scale(self, col):
        v = col.views()
        n = col.total_cells() * col.nelem
        out = np.zeros(max(n, 1))
        self.L.or_force_scale(self.w, v, col.nspec, out.ctypes.data_as(abi.PD))
        return out[:n]

    def force_scale_gather(self, col):
        """force_scale over the link list, threaded, for the sample of set_sample (others 0)."""
        v = col.views()
        n = col.total_cells() * col.nelem
        out = np.zeros(max(n, 1))
        self._chk(self.L.or_force_scale_gather(self.w, v, col.nspec, out.ctypes.data_as(abi.PD)), "force_scale_gather")
        return out[:n]

    def check_positions(self, col):
        v = col.views()
        r = C.c_int()
        self._chk(self.L.or_check_positions(self.w, v, col.nspec, C.byref(r)), "check_positions")
        col.pull_counts(v)
        return r.value

    def coms(self, col):
        v = col.views()
        out = np.zeros((col.total_cells(), col.dim))
        self.L.or_coms(self.w, v, col.nspec, out.ctypes.data_as(abi.PD))
        return out

    def linklist_info(self):
        ns = (C.c_int * 3)()
        w = C.c_double()
        self.L.or_linklist_info(self.w, ns, C.byref(w))
        return list(ns), w.value


def dsfmt_fill(seed, kind, n):
    out = np.zeros(n)
    lib().or_dsfmt_fill(seed, kind, out.ctypes.data_as(abi.PD), n)
    return out


def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.or_philox4x32(c, k, o)
    return [o[i] for i in range(4)]


def philox_normal2(seed, key, ctr, stream, sub):
    """The two standard normals of one counter-based draw (the FDT noise of a step takes sub 0
    and 1 of (seed, particle key, step, stream 2))."""
    z = (C.c_double * 2)()
    lib().or_philox_normal2(int(seed), int(key), int(ctr), int(stream), int(sub), z)
    return z[0], z[1]


def pair_f0(params, dim, ti, tj, si, sj, r2):
    f = C.c_double()
    hit = lib().or_pair_f0(C.byref(params), dim, ti, tj, si, sj, r2, C.byref(f))
    return bool(hit), f.value
