"""Writes tests/golden/dsfmt_ref.npz: raw dSFMT-19937 output words from the reference's own
mt/dSFMT.c (built into oracle/_ref by `make -C oracle ref`), for seeds 1, 408, 4357 and
0xDEADBEEF: the first 382*5+17 words of each stream (close1_open2 doubles as uint64).
Run from the repository root in a container where /root/reference is present."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref", "libdsfmt_ref.so")
SEEDS = (1, 408, 4357, 0xDEADBEEF)
N = 382 * 5 + 17


def stream(L, seed, n):
    buf = C.create_string_buffer(8192)
    L.dsfmt_chk_init_gen_rand(buf, C.c_uint32(seed), C.c_int(19937))
    out = []
    while len(out) < n:
        L.dsfmt_gen_rand_all(buf)
        out.extend(np.frombuffer(buf.raw[: 382 * 8], dtype=np.uint64).tolist())
    return np.array(out[:n], dtype=np.uint64)


def main():
    L = C.CDLL(REF)
    data = {f"seed_{s}": stream(L, s, N) for s in SEEDS}
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "dsfmt_ref.npz"), **data)
    print("wrote", list(data))


if __name__ == "__main__":
    main()
