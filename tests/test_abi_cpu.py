"""CPU-side checks of the C-ABI boundary: the in-tree library loads, exports every entry
point include/cil_hip.h declares, and its struct layouts match the ctypes mirror."""
import ctypes
import os
import re

from cil_cuda_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cil_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cil_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(abi.EXPORTS)


def test_struct_layouts_match():
    assert abi.check_layout()


def test_version_string():
    assert b"gfx950" in abi.lib().cil_hip_version()


def test_create_rejects_bad_params_without_touching_device():
    from cil_cuda_amd import configs

    p, box, col = configs.plain_2d_sc(n_cells=10, L=10.0)
    p.abi_version = 99
    ctx = ctypes.c_void_p()
    caps = (ctypes.c_int * 4)(10, 0, 0, 0)
    assert abi.lib().cil_hip_create(ctypes.byref(p), ctypes.byref(box), caps, ctypes.byref(ctx)) == abi.E_INVALID
    p.abi_version = abi.ABI_VERSION
    p.species[0].zeta[0] = 0.0  # Euler requires zeta > 0 (CIL.cu:925)
    assert abi.lib().cil_hip_create(ctypes.byref(p), ctypes.byref(box), caps, ctypes.byref(ctx)) == abi.E_INVALID
