"""cil_hip_pairs_noalloc (the pair hook staged in the library's fixed device array) against
cil_hip_pairs (its own allocated buffer), through the C ABI: identical sorted pair sets on two
calls in a row, the second on a smaller colony (the shared array keeps the first call's
contents, so its counter must be cleared), and a capacity beyond the array is rejected."""
import ctypes as C

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import HipMD

pytestmark = pytest.mark.gpu


def _pairs(md, fn, cap):
    L = md.L
    buf = np.zeros(2 * max(cap, 1), dtype=np.int32)
    n = C.c_size_t()
    rc = getattr(L, fn)(md.ctx, buf.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(n))
    if rc != 0:
        raise abi.CilError(rc, L.cil_hip_last_error(md.ctx).decode())
    return buf[: 2 * min(n.value, cap)].reshape(-1, 2), n.value


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_FP32])
def test_pairs_noalloc_matches_pairs(gpu, precision):
    counts = []
    for n_cells, L, seed in ((900, 30.0, 3), (200, 14.0, 4)):  # second call: smaller input
        print("pairs_noalloc", precision, n_cells, flush=True)
        p, box, col = configs.plain_2d_sc(n_cells=n_cells, L=L, seed=seed, precision=precision)
        md = HipMD(p, box, [sp["capacity"] for sp in col.species])
        md.upload(col)
        ref, nref = _pairs(md, "cil_hip_pairs", 1 << 16)
        got, ngot = _pairs(md, "cil_hip_pairs_noalloc", 1 << 16)
        assert nref == ngot > 0
        assert np.array_equal(ref, got)
        # a capacity below the count: the count is still exact, the first cap pairs are written
        part, npart = _pairs(md, "cil_hip_pairs_noalloc", 10)
        assert npart == nref and part.shape == (10, 2)
        md.close()
        counts.append(nref)
    assert counts[1] < counts[0]


def test_pairs_noalloc_rejects_capacity_beyond_scratch(gpu):
    p, box, col = configs.plain_2d_sc(n_cells=50, L=12.0, seed=2)
    md = HipMD(p, box, [sp["capacity"] for sp in col.species])
    md.upload(col)
    with pytest.raises(abi.CilError) as ei:
        _pairs(md, "cil_hip_pairs_noalloc", (1 << 20) + 1)
    assert "scratch" in str(ei.value)
    md.close()
