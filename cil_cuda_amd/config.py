"""The reference's JSON run configurations through the C ABI (cil_hip_config_read / _init).

    params, box, run = read_config(path_or_text)
    colony = initial_colony(path_or_text, params, capacity)   # sets params.rng_offset

The parsing, the parameter derivations of readSpecs / readMD / initializeTimeStep and the
initial colony (readInit, initializeCellStatus on the reference's random stream) all run in the
library's C++ code; this module only carries the arrays.
"""
from __future__ import annotations

import ctypes as C
import os

from . import abi
from .colony import Colony


def _text(src) -> bytes:
    if isinstance(src, bytes):
        return src
    if os.path.exists(src):
        with open(src, "rb") as f:
            return f.read()
    return src.encode()


class ConfigError(RuntimeError):
    pass


def read_config(src):
    """(params, box, run) of a reference configuration (file path or JSON text)."""
    L = abi.lib()
    p, box, run = abi.Params(), abi.Box(), abi.Run()
    err = C.create_string_buffer(512)
    rc = L.cil_hip_config_read(_text(src), C.byref(p), C.byref(box), C.byref(run), err, len(err))
    if rc != 0:
        raise ConfigError(err.value.decode())
    return p, box, run


def initial_colony(src, params: abi.Params, capacity=None):
    """The initial colony of the configuration, with capacity[s] cells per species (default:
    the initial count, at least 16).  Updates params.rng_offset to the draws the
    initialisation consumed from the reference's stream."""
    _, box, run = read_config(src)
    n = [run.ncells[s] for s in range(run.nspecies)]
    cap = capacity or [max(16, k) for k in n]
    col = Colony(params.dim, params.species[0].nelem, cap)
    views = col.views()
    err = C.create_string_buffer(512)
    rc = abi.lib().cil_hip_config_init(_text(src), C.byref(params), views, col.nspec, err, len(err))
    if rc != 0:
        raise ConfigError(err.value.decode())
    col.pull_counts(views)
    return col


def dsfmt_fill(seed: int, kind: int, n: int, skip: int = 0):
    """n draws of the reference's random stream (0 uniform [0,1), 1 normal, 2 open (0,1))."""
    import numpy as np

    out = np.zeros(n)
    rc = abi.lib().cil_hip_dsfmt_fill(int(seed), int(skip), int(kind), out.ctypes.data_as(abi.PD), int(n))
    if rc != 0:
        raise ConfigError("dsfmt_fill: bad arguments")
    return out
