"""Host-side mirror of the reference's colony interfaces over the C ABI.

`Colony` holds per-species arrays in the reference's layout (CellColony, Colony.hpp:81-113:
pos(cell, elem, d), sigmaCore(cell, elem), per-cell state), and `HipMD` is the drop-in
Euler step (`mdStep = EulerHIP`, MD.cuh:450-452) running on the GPU through
include/cil_hip.h.  Everything here is plumbing; the step itself runs in the HIP library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class Colony:
    """Per-species cell arrays (numpy), reference layout, with a fixed capacity."""

    DOUBLE_CELL = ("sigma_t", "pressure", "theta", "rho", "ang_vel", "rest_len")
    INT_CELL = ("state", "t0", "tswitch", "tpressure")

    def __init__(self, dim: int, nelem: int, capacities, ncells=None):
        self.dim = dim
        self.nelem = nelem
        self.species = []
        for k, cap in enumerate(capacities):
            cap = int(cap)
            sp = {
                "ncells": 0 if ncells is None else int(ncells[k]),
                "capacity": cap,
                "pos": np.zeros((cap, nelem, dim)),
                "frc": np.zeros((cap, nelem, dim)),
                "vel": np.zeros((cap, nelem, dim)),
                "sigma": np.ones((cap, nelem)),
            }
            for f in self.DOUBLE_CELL:
                sp[f] = np.zeros(cap)
            for f in self.INT_CELL:
                sp[f] = np.zeros(cap, dtype=np.int32)
            sp["state"][:] = abi.CS_UNDEFINED
            sp["tswitch"][:] = -1
            self.species.append(sp)

    @property
    def nspec(self):
        return len(self.species)

    def ncells(self):
        return [sp["ncells"] for sp in self.species]

    def total_cells(self):
        return sum(self.ncells())

    def copy(self):
        c = Colony(self.dim, self.nelem, [sp["capacity"] for sp in self.species])
        for a, b in zip(c.species, self.species):
            for k, v in b.items():
                a[k] = v.copy() if isinstance(v, np.ndarray) else v
        return c

    def views(self, with_frc=True):
        arr = (abi.SpeciesView * self.nspec)()
        for k, sp in enumerate(self.species):
            v = arr[k]
            v.ncells = sp["ncells"]
            v.capacity = sp["capacity"]
            for f in abi.VIEW_DOUBLE_FIELDS:
                if f == "frc" and not with_frc:
                    setattr(v, f, abi.PD())
                    continue
                a = sp[f]
                assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
                setattr(v, f, a.ctypes.data_as(abi.PD))
            for f in abi.VIEW_INT_FIELDS:
                a = sp[f]
                assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
                setattr(v, f, a.ctypes.data_as(abi.PI))
        return arr

    def pull_counts(self, views):
        for k, sp in enumerate(self.species):
            sp["ncells"] = int(views[k].ncells)

    def flat(self, field):
        """Concatenate the active part of a per-element field over species (UID order)."""
        return np.concatenate([sp[field][: sp["ncells"]] for sp in self.species], axis=0)


def _check(L, ctx, rc, what):
    if rc != 0:
        msg = L.cil_hip_last_error(ctx).decode() if ctx else ""
        raise abi.CilError(rc, f"{what}: {msg}")


class HipMD:
    """GPU Euler step behind the C ABI (the drop-in for the reference's mdStep)."""

    def __init__(self, params: abi.Params, box: abi.Box, capacity):
        self.L = abi.lib()
        abi.check_layout(self.L)
        caps = (C.c_int * abi.MAX_SPECIES)(*([int(c) for c in capacity] + [0] * (abi.MAX_SPECIES - len(capacity))))
        self.ctx = C.c_void_p()
        rc = self.L.cil_hip_create(C.byref(params), C.byref(box), caps, C.byref(self.ctx))
        if rc != 0:
            raise abi.CilError(rc, "cil_hip_create failed")
        self.params = params
        self.nspec = params.nspecies
        self.dim = params.dim
        self.nelem = params.species[0].nelem

    def close(self):
        if self.ctx:
            self.L.cil_hip_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, colony: Colony, with_frc=False):
        v = colony.views(with_frc=with_frc)
        _check(self.L, self.ctx, self.L.cil_hip_upload(self.ctx, v, colony.nspec), "upload")

    def download(self, colony: Colony):
        v = colony.views()
        req = (C.c_int * abi.MAX_SPECIES)()
        _check(self.L, self.ctx, self.L.cil_hip_download(self.ctx, v, colony.nspec, req), "download")
        colony.pull_counts(v)
        return colony

    def forces(self, want=True):
        """calculForcesPressure at the current state; returns (frc[Np,D], pressure[Nc])."""
        if not want:
            _check(self.L, self.ctx, self.L.cil_hip_forces(self.ctx, abi.PD(), abi.PD()), "forces")
            return None
        nc = self.num_cells()
        ntot = sum(nc)
        frc = np.zeros((ntot * self.nelem, self.dim))
        pr = np.zeros(max(ntot, 1))
        _check(self.L, self.ctx, self.L.cil_hip_forces(self.ctx, frc.ctypes.data_as(abi.PD), pr.ctypes.data_as(abi.PD)),
               "forces")
        return frc, pr[:ntot]

    def step(self, nsteps: int, ts: int) -> int:
        t = C.c_int(ts)
        _check(self.L, self.ctx, self.L.cil_hip_step(self.ctx, int(nsteps), C.byref(t)), "step")
        return t.value

    def pairs(self, cap=None):
        if cap is None:
            cap = 1 << 16
        while True:
            buf = np.zeros(2 * cap, dtype=np.int32)
            n = C.c_size_t()
            _check(self.L, self.ctx,
                   self.L.cil_hip_pairs(self.ctx, buf.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(n)), "pairs")
            if n.value <= cap:
                return buf[: 2 * n.value].reshape(-1, 2)
            cap = int(n.value)

    def check_positions(self) -> int:
        r = C.c_int()
        _check(self.L, self.ctx, self.L.cil_hip_check_positions(self.ctx, C.byref(r)), "check_positions")
        return r.value

    def set_external_noise(self, normals: np.ndarray):
        normals = np.ascontiguousarray(normals, dtype=np.float64)
        nsteps, per = normals.shape
        _check(self.L, self.ctx,
               self.L.cil_hip_set_external_noise(self.ctx, normals.ctypes.data_as(abi.PD), per, nsteps),
               "set_external_noise")

    def num_cells(self):
        nc = (C.c_int * abi.MAX_SPECIES)()
        _check(self.L, self.ctx, self.L.cil_hip_num_cells(self.ctx, nc), "num_cells")
        return [nc[k] for k in range(self.nspec)]

    def bin_info(self):
        nb = (C.c_int * 3)()
        w = (C.c_double * 3)()
        _check(self.L, self.ctx, self.L.cil_hip_bin_info(self.ctx, nb, w), "bin_info")
        return list(nb), list(w)

    def set_tuning(self, bins_per_cutoff=0, block=0):
        _check(self.L, self.ctx, self.L.cil_hip_set_tuning(self.ctx, int(bins_per_cutoff), int(block)), "set_tuning")

    def autotune(self, reps=5):
        """Pick bins per cut-off, pair block size and kernel variant by timing force passes."""
        b, k, v = C.c_int(), C.c_int(), C.c_int()
        us = C.c_double()
        _check(self.L, self.ctx, self.L.cil_hip_autotune(self.ctx, int(reps), C.byref(b), C.byref(k), C.byref(v),
                                                         C.byref(us)), "autotune")
        return {"bins_per_cutoff": b.value, "block": k.value, "variant": v.value, "force_pass_us": us.value}

    def set_kernel(self, variant):
        _check(self.L, self.ctx, self.L.cil_hip_set_kernel(self.ctx, int(variant)), "set_kernel")

    def set_division_log(self, on=True):
        """Record each division's new cell (positions right after the division, step)."""
        _check(self.L, self.ctx, self.L.cil_hip_set_division_log(self.ctx, 1 if on else 0), "set_division_log")

    def division_log(self, s, clear=True):
        """(positions [n][nelem][dim], steps [n]) of species s's recorded divisions."""
        n = C.c_int()
        _check(self.L, self.ctx, self.L.cil_hip_division_log(self.ctx, s, abi.PD(), None, 0, C.byref(n), 0),
               "division_log")
        d, ne = self.params.dim, self.params.species[0].nelem
        pos = np.zeros((max(n.value, 1), ne, d))
        ts = np.zeros(max(n.value, 1), dtype=np.int32)
        _check(self.L, self.ctx, self.L.cil_hip_division_log(self.ctx, s, pos.ctypes.data_as(abi.PD),
                                                             ts.ctypes.data_as(C.POINTER(C.c_int)), n.value,
                                                             C.byref(n), 1 if clear else 0), "division_log")
        return pos[:n.value], ts[:n.value]

    def set_neighbour_list(self, on=True, skin=0.3, max_neighbours=256, inner_skin=0.0):
        """Neighbour-list mode; inner_skin > 0 adds the inner list of two-level lists."""
        _check(self.L, self.ctx, self.L.cil_hip_set_neighbour_list(self.ctx, 1 if on else 0, float(skin),
                                                                   int(max_neighbours)), "set_neighbour_list")
        if on and inner_skin:
            _check(self.L, self.ctx, self.L.cil_hip_set_neighbour_list_inner(self.ctx, float(inner_skin)),
                   "set_neighbour_list_inner")

    FAMILIES = ("update", "division", "sort", "pair", "noise", "guard", "build", "knn", "mlp", "macro")

    def set_timing(self, on=True, families=None):
        """Per-family HIP-event timing; `families` (names) restricts it, None times all."""
        v = 0
        if on:
            v = 1 if families is None else sum(2 << self.FAMILIES.index(f) for f in families)
        _check(self.L, self.ctx, self.L.cil_hip_set_timing(self.ctx, v), "set_timing")

    def timings(self):
        us = (C.c_double * 16)()
        n = C.c_int()
        names = C.create_string_buffer(256)
        _check(self.L, self.ctx, self.L.cil_hip_timings(self.ctx, us, 16, C.byref(n), names, 256), "timings")
        keys = names.value.decode().split(";")
        return {keys[k]: us[k] for k in range(n.value)}

    # ---------------------------------------------------------------- drift surrogate
    def _f32(self, a, shape=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if shape is not None and a.shape != shape:
            raise ValueError(f"expected shape {shape}, got {a.shape}")
        return a

    def mlp_set_params(self, flat):
        p = self._f32(flat, (abi.MLP_NPARAM,))
        _check(self.L, self.ctx, self.L.cil_hip_mlp_set_params(self.ctx, _fp(p), p.size), "mlp_set_params")

    def mlp_get_params(self):
        p = np.zeros(abi.MLP_NPARAM, dtype=np.float32)
        _check(self.L, self.ctx, self.L.cil_hip_mlp_get_params(self.ctx, _fp(p), p.size), "mlp_get_params")
        return p

    def mlp_init(self, seed=0):
        _check(self.L, self.ctx, self.L.cil_hip_mlp_init(self.ctx, int(seed)), "mlp_init")

    def mlp_forward(self, X):
        X = self._f32(X)
        n = X.shape[0]
        Y = np.zeros((n, 3), dtype=np.float32)
        _check(self.L, self.ctx, self.L.cil_hip_mlp_forward(self.ctx, _fp(X), n, _fp(Y)), "mlp_forward")
        return Y

    def mlp_grad(self, X, T):
        X = self._f32(X)
        T = self._f32(T)
        g = np.zeros(abi.MLP_NPARAM, dtype=np.float32)
        loss = C.c_float()
        _check(self.L, self.ctx, self.L.cil_hip_mlp_grad(self.ctx, _fp(X), _fp(T), X.shape[0], _fp(g), C.byref(loss)),
               "mlp_grad")
        return g, float(loss.value)

    def knn_features(self, nrows):
        F = np.zeros((nrows, abi.MLP_NFEAT), dtype=np.float32)
        _check(self.L, self.ctx, self.L.cil_hip_knn_features(self.ctx, _fp(F), nrows), "knn_features")
        return F

    def surrogate_collect(self, nmacro, stride, cap, ts):
        t = C.c_int(ts)
        n = C.c_long()
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_collect(self.ctx, int(nmacro), int(stride), int(cap),
                                                                  C.byref(t), C.byref(n)), "surrogate_collect")
        return t.value, n.value

    def surrogate_set_samples(self, X, T):
        X = self._f32(X)
        T = self._f32(T)
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_set_samples(self.ctx, _fp(X), _fp(T), X.shape[0]),
               "surrogate_set_samples")

    def surrogate_get_samples(self, with_x=True):
        n = C.c_long()
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_get_samples(self.ctx, None, None, 0, C.byref(n)),
               "surrogate_get_samples")
        X = np.zeros((n.value if with_x else 0, abi.MLP_NFEAT), dtype=np.float32)
        T = np.zeros((n.value, 3), dtype=np.float32)
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_get_samples(self.ctx, _fp(X) if with_x else None, _fp(T),
                                                                      n.value, C.byref(n)), "surrogate_get_samples")
        return (X if with_x else None), T

    def surrogate_clear_samples(self):
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_clear_samples(self.ctx), "surrogate_clear_samples")

    def surrogate_train(self, nsteps, batch=4096, lr=1e-3, seed=0):
        losses = np.zeros(max(nsteps, 1), dtype=np.float32)
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_train(self.ctx, int(nsteps), int(batch), float(lr), int(seed),
                                                                _fp(losses)), "surrogate_train")
        return losses[:nsteps]

    def surrogate_eval(self, ts):
        """One judged macro-step: RMS/max errors of (network, zero drift, one 8dt Euler step)."""
        t = C.c_int(ts)
        rms = (C.c_double * 3)()
        mx = (C.c_double * 3)()
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_eval(self.ctx, C.byref(t), rms, mx), "surrogate_eval")
        return t.value, list(rms), list(mx)

    def surrogate_step(self, nmacro, ts, judge_every=16, tol=1e-3):
        t = C.c_int(ts)
        st = abi.SurrogateStats()
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_step(self.ctx, int(nmacro), int(judge_every), float(tol),
                                                               C.byref(t), C.byref(st)), "surrogate_step")
        stats = {k: getattr(st, k) for k, _ in abi.SurrogateStats._fields_}
        return t.value, stats

    def surrogate_stats_reset(self):
        _check(self.L, self.ctx, self.L.cil_hip_surrogate_stats_reset(self.ctx), "surrogate_stats_reset")


class HipBatch:
    """Many HipMD colonies advanced together, one launch per kernel (cil_hip_batch_*).
    The member contexts stay usable on their own (upload, download, check_positions)."""

    def __init__(self, members):
        self.L = abi.lib()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        self.b = C.c_void_p()
        rc = self.L.cil_hip_batch_create(arr, len(self.members), C.byref(self.b))
        if rc != 0:
            raise abi.CilError(rc, "cil_hip_batch_create: the colonies differ in a launch-shaping setting "
                                   "(dimension, precision, division, noise kernel, LJ, pair kernel, block, rng)")

    def _check(self, rc, what):
        if rc != 0:
            raise abi.CilError(rc, f"{what}: {self.L.cil_hip_batch_last_error(self.b).decode()}")

    def step(self, nsteps, ts):
        t = (C.c_int * len(self.members))(*[int(x) for x in ts])
        self._check(self.L.cil_hip_batch_step(self.b, int(nsteps), t), "batch_step")
        return [t[k] for k in range(len(self.members))]

    def autotune(self, reps=3):
        b, k, v = C.c_int(), C.c_int(), C.c_int()
        us = C.c_double()
        self._check(self.L.cil_hip_batch_autotune(self.b, int(reps), C.byref(b), C.byref(k), C.byref(v), C.byref(us)),
                    "batch_autotune")
        return {"bins_per_cutoff": b.value, "block": k.value, "variant": v.value, "force_pass_us": us.value}

    def close(self):
        if self.b:
            self.L.cil_hip_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))
