"""Pins of the oracle's random-number restatements.

- Philox4x32-10: the published Random123 known-answer vectors.
- dSFMT-19937: bit-for-bit against the reference's own mt/dSFMT.c compiled into oracle/_ref
  (when that build exists), over several blocks of the state recursion.
- Leva normal + Nint + draw order: the reference's InitConf known answers
  (output/RunParams/z0.001t3.16_v2.InitConf.dat:3 -> tswitch 47049, seed 1, dt 1e-5;
   output/RunParams/cil.InitConf.dat:3 -> tswitch 139509, seed 408, dt 2.5e-6), which
  initializeCellStatus (CIL.cuh:270-298) produces from one uniform, one normal, two uniforms.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

REF_DSFMT = os.path.join(os.path.dirname(O.HERE), "oracle", "_ref", "libdsfmt_ref.so")


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 10 rounds
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_normals_are_standard():
    L = O.lib()
    z = np.zeros(2)
    vals = []
    for i in range(20000):
        L.or_philox_normal2(12345, i, 7, 1, 0, z.ctypes.data_as(C.POINTER(C.c_double)))
        vals.extend(z.tolist())
    v = np.array(vals)
    assert abs(v.mean()) < 0.02
    assert abs(v.std() - 1.0) < 0.02
    # 4th moment of N(0,1) is 3
    assert abs(np.mean(v ** 4) - 3.0) < 0.15


def _ref_stream(seed, n):
    L = C.CDLL(REF_DSFMT)
    buf = C.create_string_buffer(8192)
    L.dsfmt_chk_init_gen_rand(buf, C.c_uint32(seed), C.c_int(19937))
    out = []
    while len(out) < n:
        L.dsfmt_gen_rand_all(buf)
        words = np.frombuffer(buf.raw[: 382 * 8], dtype=np.uint64)
        out.extend(words.tolist())
    return np.array(out[:n], dtype=np.uint64)


@pytest.mark.skipif(not os.path.exists(REF_DSFMT), reason="reference dSFMT build absent")
@pytest.mark.parametrize("seed", [1, 408, 4357, 0xDEADBEEF])
def test_dsfmt_matches_reference_build(seed):
    n = 382 * 5 + 17
    ref = _ref_stream(seed, n).view(np.float64)
    mine = O.dsfmt_fill(seed, 0, n) + 1.0  # close_open + 1 == close1_open2 exactly
    assert np.array_equal(mine.view(np.uint64), ref.view(np.uint64))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dsfmt_ref.npz")


@pytest.mark.parametrize("seed", [1, 408, 4357, 0xDEADBEEF])
def test_dsfmt_matches_reference_golden_vectors(seed):
    """The same check against the committed vectors of the reference build
    (tests/golden/make_dsfmt_golden.py), so it holds where the reference tree is absent."""
    ref = np.load(GOLDEN)[f"seed_{seed}"]
    mine = O.dsfmt_fill(seed, 0, len(ref)) + 1.0
    assert np.array_equal(mine.view(np.uint64), ref)


def _nint(a):
    return int(a + 0.5) if a >= 0.0 else int(a - 0.5)


class _Stream:
    def __init__(self, seed):
        self.L = O.lib()
        self.g = C.create_string_buffer(4096)
        self.L.or_dsfmt_init.argtypes = [C.c_void_p, C.c_uint32]
        self.L.or_dsfmt_close_open.argtypes = [C.c_void_p]
        self.L.or_dsfmt_close_open.restype = C.c_double
        self.L.or_dsfmt_normal.argtypes = [C.c_void_p]
        self.L.or_dsfmt_normal.restype = C.c_double
        self.L.or_dsfmt_init(self.g, seed)

    def uniform0x(self, x):
        return self.L.or_dsfmt_close_open(self.g) * x

    def normal(self, mu=0.0, sigma=1.0):
        return mu + sigma * self.L.or_dsfmt_normal(self.g)


@pytest.mark.parametrize("seed,dt,expected", [(1, 1e-5, 47049), (408, 2.5e-6, 139509)])
def test_initconf_tswitch_known_answer(seed, dt, expected):
    # initializeCellStatus, CIL.cuh:270-298, with TAU_DIVISION (1, 0.1), TAU_MIGRATION (0, 0)
    r = _Stream(seed)
    prob = 0.0 / (0.0 + 1.0)
    u = r.uniform0x(1.0)
    assert not (u <= prob)
    time_span = max(_nint(r.normal(1.0, 0.1) / dt), 1)
    r.uniform0x(1)  # theta
    time_span -= _nint(r.uniform0x(float(time_span))) - 1
    assert time_span == expected


def test_leva_normal_distribution():
    z = O.dsfmt_fill(7, 1, 200000)
    assert abs(z.mean()) < 0.01
    assert abs(z.std() - 1.0) < 0.01
    # tail probability P(|z|>2) = 0.0455
    assert abs(np.mean(np.abs(z) > 2.0) - 0.0455) < 0.003


def test_open_open_range():
    u = O.dsfmt_fill(3, 2, 10000)
    assert np.all(u > 0.0) and np.all(u < 1.0)
    c = O.dsfmt_fill(3, 0, 10000)
    assert np.all(c >= 0.0) and np.all(c < 1.0)
    assert math.isclose(float(c.mean()), 0.5, abs_tol=0.02)
