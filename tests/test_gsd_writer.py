"""The library's GSD writer (cil_hip_gsd_*) against the reference's own GSD output.

Pins, each independent of the writer:
  - the reference's `z0.001t3.16_v2.Last.gsd` (tests/golden, written by the reference's
    outputGSDFrame at the end of its seeded run): the colony it holds, written again by the
    library at ts = 0 with the run's own configuration, gives the same file byte for byte;
  - the reference's gsd.c, compiled from its sources into oracle/_ref by oracle/Makefile: the
    same chunk sequence over 300 frames (the index grows three times) gives the same file byte
    for byte;
  - a reader of the file format written here from its published description (header, index,
    name list) finds every chunk of 2D two-disc and 3D one-disc frames with the values computed
    from the colony in numpy.
"""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from cil_cuda_amd import abi, config, configs
from cil_cuda_amd.colony import Colony
from cil_cuda_amd.output import GsdWriter

HERE = os.path.dirname(os.path.abspath(__file__))
LAST = os.path.join(HERE, "golden", "z0.001t3.16_v2.Last.gsd")
SEEDED = os.path.join(HERE, "golden", "configs", "z0.001t3.16_v2.RunParams.json")
REF_GSD = os.path.join(HERE, "..", "oracle", "_ref", "libgsd_ref.so")
DTYPES = {1: np.uint8, 2: np.uint16, 3: np.uint32, 4: np.uint64, 5: np.int8, 6: np.int16, 7: np.int32, 8: np.int64,
          9: np.float32, 10: np.float64}


def read_gsd(path):
    """{frame: {name: array}} and the header fields of a GSD 1.0 file."""
    b = open(path, "rb").read()
    magic, iloc, ialloc, nloc, nalloc, schv, gsdv = struct.unpack_from("<QQQQQII", b, 0)
    assert magic == 0x65DF65DF65DF65DF and gsdv == 1 << 16
    names = [b[nloc + 64 * k:nloc + 64 * k + 64].split(b"\0")[0].decode() for k in range(nalloc)]
    frames = {}
    for k in range(ialloc):
        fr, n, loc, m, nid, ty, _ = struct.unpack_from("<QQqIHBB", b, iloc + 32 * k)
        if loc == 0:
            break
        a = np.frombuffer(b, DTYPES[ty], n * m, loc)
        frames.setdefault(fr, {})[names[nid]] = a.reshape(n, m) if m > 1 else a
    hdr = dict(schema_version=schv, application=b[48:112].split(b"\0")[0], schema=b[112:176].split(b"\0")[0])
    return frames, hdr


def test_reference_last_frame_written_byte_for_byte(tmp_path):
    frames, hdr = read_gsd(LAST)
    f = frames[0]
    assert hdr["application"] == b"CIL" and hdr["schema"] == b"hoomd" and hdr["schema_version"] == (1 << 16) | 1
    p, box, run = config.read_config(SEEDED)
    n = int(f["particles/N"][0])
    col = Colony(2, 1, [n], ncells=[n])
    s = col.species[0]
    half = np.array(box.length[:2]) / 2
    # float32 + L/2 - L/2 is exact in double, so the file's floats come back unchanged
    s["pos"][:, 0] = f["particles/position"][:, :2].astype(np.float64) + half
    s["vel"][:, 0] = f["particles/velocity"][:, :2]
    s["sigma"][:, 0] = f["particles/Z"]
    s["rho"][:] = f["particles/charge"]
    s["pressure"][:] = f["particles/mass"]
    out = tmp_path / "last.gsd"
    with GsdWriter(out) as g:
        g.frame(p, box, col, 0, run)
    assert out.read_bytes() == open(LAST, "rb").read()


def _ref_lib():
    if not os.path.exists(REF_GSD):
        pytest.skip("oracle/_ref/libgsd_ref.so not built (the reference tree is absent)")
    L = C.CDLL(REF_GSD)
    L.gsd_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32]
    L.gsd_open.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    L.gsd_write_chunk.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_uint64, C.c_uint32, C.c_uint8, C.c_void_p]
    L.gsd_end_frame.argtypes = [C.c_void_p]
    L.gsd_close.argtypes = [C.c_void_p]
    return L


def test_chunk_sequence_matches_reference_library(tmp_path):
    L = _ref_lib()
    rng = np.random.default_rng(5)
    seq = []
    for fr in range(300):
        n = int(rng.integers(1, 50))
        seq.append([("configuration/step", 4, np.array([fr * 1000], np.uint64), 1, 1),
                    ("particles/position", 9, rng.standard_normal((n, 3)).astype(np.float32), n, 3),
                    ("particles/typeid", 3, rng.integers(0, 4, n).astype(np.uint32), n, 1)])
        if fr % 20 == 0:  # a name longer than 63 characters is cut to 63 (and takes a new slot)
            seq[-1].append(("x" * 80, 10, rng.standard_normal(2), 2, 1))
    mine, ref = tmp_path / "mine.gsd", tmp_path / "ref.gsd"
    with GsdWriter(mine) as g:
        for frame in seq:
            for name, ty, a, n, m in frame:
                g.chunk(name, ty, np.ascontiguousarray(a), n, m)
            g.end_frame()
    assert L.gsd_create(str(ref).encode(), b"CIL", b"hoomd", (1 << 16) | 1) == 0
    h = C.create_string_buffer(4096)
    assert L.gsd_open(h, str(ref).encode(), 3) == 0  # GSD_OPEN_APPEND
    for frame in seq:
        for name, ty, a, n, m in frame:
            a = np.ascontiguousarray(a)
            assert L.gsd_write_chunk(h, name.encode(), ty, n, m, 0, a.ctypes.data) == 0
        assert L.gsd_end_frame(h) == 0
    assert L.gsd_close(h) == 0
    assert mine.read_bytes() == ref.read_bytes()
    frames, _ = read_gsd(mine)
    assert len(frames) == 300 and np.array_equal(frames[299]["particles/position"], seq[299][1][2])


def _run_names(nspec, species, elements):
    run = abi.Run()
    run.nspecies = nspec
    for s, nm in enumerate(species):
        run.species[s].value = nm.encode()
    for s, els in enumerate(elements):
        for e, nm in enumerate(els):
            run.elements[s][e].value = nm.encode()
    return run


def test_two_disc_and_3d_frames_read_back(tmp_path):
    # 2D two-disc colonies of two species, dividing (linear growth)
    p, box, col = configs.channel_2d_hz(width=12.0, length=30.0, seed=2, capacity=16)
    q = configs.base_params(2, 2, 2)
    q.division = abi.DIV_LINEAR_GROWTH
    for s in range(2):
        q.species[s].sigma[0], q.species[s].sigma[1] = 1.0 + 0.1 * s, 0.9
    b2 = configs.make_box([12.0, 30.0], [1, 1])
    c2 = Colony(2, 2, [8, 8], ncells=[3, 5])
    rng = np.random.default_rng(1)
    for s, sp in enumerate(c2.species):
        n = sp["ncells"]
        sp["pos"][:n] = rng.uniform(0, 12, (n, 2, 2))
        sp["vel"][:n] = rng.standard_normal((n, 2, 2))
        sp["sigma"][:n] = rng.uniform(0.8, 1.2, (n, 2))
        sp["rho"][:n] = rng.uniform(0, 1, n)
        sp["pressure"][:n] = rng.standard_normal(n)
    run = _run_names(2, ["A", "B"], [["BACK", "FRONT"], ["BACK", "FRONT"]])
    out = tmp_path / "two.gsd"
    with GsdWriter(out) as g:
        g.frame(q, b2, c2, 0, run)
        g.frame(q, b2, c2, 500, run)
    frames, _ = read_gsd(out)
    pos = np.concatenate([sp["pos"][:sp["ncells"]].reshape(-1, 2) for sp in c2.species])
    for fr, ts in ((0, 0), (1, 500)):
        f = frames[fr]
        assert f["configuration/step"][0] == ts and f["configuration/dimensions"][0] == 3
        assert np.array_equal(f["configuration/box"], np.array([12, 30, 1.1, 0, 0, 0], np.float32))
        assert f["particles/N"][0] == 16
        want = np.zeros((16, 3), np.float32)
        want[:, :2] = pos - np.array([6.0, 15.0])
        want[1::2, 2] = np.float32(-0.01)
        assert np.array_equal(f["particles/position"], want)
        assert np.array_equal(np.signbit(f["particles/position"][0::2, 2]), np.ones(8, bool))  # -0.0
        assert np.array_equal(f["particles/typeid"], np.array([0, 1] * 3 + [2, 3] * 5, np.uint32))
        rho = np.concatenate([np.repeat(sp["rho"][:sp["ncells"]], 2) for sp in c2.species]).astype(np.float32)
        assert np.array_equal(f["particles/charge"], rho)
        pr = np.concatenate([np.repeat(sp["pressure"][:sp["ncells"]], 2) for sp in c2.species]).astype(np.float32)
        assert np.array_equal(f["particles/mass"], pr)
        assert ("particles/types" in f) == (ts == 0)
    types = [bytes(r).split(b"\0")[0] for r in frames[0]["particles/types"].view(np.uint8)]
    assert types == [b"A_BACK", b"A_FRONT", b"B_BACK", b"B_FRONT"]
    # 3D single discs, not dividing: Z, diameter and typeid only in the ts == 0 frame
    p3, box3, c3 = configs.bulk_3d_mixed(n=300, seed=4)
    p3.division = abi.DIV_NONE
    out3 = tmp_path / "three.gsd"
    with GsdWriter(out3) as g:
        g.frame(p3, box3, c3, 0)
        g.frame(p3, box3, c3, 10)
    frames, _ = read_gsd(out3)
    x = np.concatenate([sp["pos"][:sp["ncells"], 0] for sp in c3.species])
    L3 = np.array(box3.length)
    for fr in (0, 1):
        f = frames[fr]
        assert f["configuration/box"][2] == np.float32(L3[2])
        assert np.array_equal(f["particles/position"], (x - L3 / 2).astype(np.float32))
        assert ("particles/Z" in f) == (fr == 0) and "particles/charge" not in f
    types = [bytes(r).split(b"\0")[0] for r in frames[0]["particles/types"].view(np.uint8)]
    assert types == [b"C1", b"C2"]


def test_writer_errors(tmp_path):
    with pytest.raises(abi.CilError):
        GsdWriter(tmp_path / "no_such_dir" / "x.gsd")
    with GsdWriter(tmp_path / "e.gsd") as g:
        with pytest.raises(abi.CilError):
            g.chunk("a", 99, np.zeros(1), 1, 1)  # unknown type
        with pytest.raises(abi.CilError):
            g.chunk("a", 9, np.zeros(0, np.float32), 0, 1)  # empty chunk


@pytest.mark.gpu
def test_frame_from_device_equals_frame_from_host(gpu, tmp_path):
    """cil_hip_gsd_write_frame_ctx (the context's device colony) writes the same bytes as the
    host writer on the downloaded colony, after the first division of a two-disc channel."""
    from cil_cuda_amd.colony import HipMD

    p, box, col = configs.channel_2d_hz(width=12.0, length=40.0, seed=3, dt=1e-3, capacity=64)
    md = HipMD(p, box, [64])
    md.upload(col)
    md.forces(want=False)
    ts = md.step(1700, 0)
    a, b = tmp_path / "dev.gsd", tmp_path / "host.gsd"
    with GsdWriter(a) as g:
        g.frame_ctx(md, 0)
        g.frame_ctx(md, ts)
    host = md.download(col.copy())
    md.close()
    assert host.ncells()[0] > 1
    with GsdWriter(b) as g:
        g.frame(p, box, host, 0)
        g.frame(p, box, host, ts)
    assert a.read_bytes() == b.read_bytes()
    frames, _ = read_gsd(a)
    assert frames[1]["particles/N"][0] == 2 * host.ncells()[0]
