"""The oracle's sampled gather (or_set_sample, or_force_scale_gather), the checker of the 1M-particle
force test: on a 16,384-particle bulk its sampled forces equal the full pass's and its gathered
force scale equals the all-pairs one."""
import numpy as np

from cil_cuda_amd import abi, configs
from oracle import oracle as O


def test_sampled_gather_equals_full_pass():
    p, box, col = configs.bulk_3d_mixed(n=1 << 14, seed=3, precision=abi.PREC_FP64)
    every = 4
    orc = O.Oracle(p, box, traversal=0, threads=4)
    orc.set_sample(every)
    ref = col.copy()
    orc.forces(ref, 0)
    scale = orc.force_scale_gather(ref)
    orc.close()
    full = O.Oracle(p, box, traversal=0, threads=4)
    got = col.copy()
    full.forces(got, 0)
    full.close()
    brute = O.Oracle(p, box, traversal=1)
    sc_all = brute.force_scale(got)
    brute.close()
    f1, f2 = ref.flat("frc").reshape(-1, 3), got.flat("frc").reshape(-1, 3)
    assert np.array_equal(f1[::every], f2[::every])
    assert np.max(np.abs(f1[::every])) > 0.0
    assert np.allclose(scale[::every], sc_all[::every], rtol=1e-12, atol=0)
    assert not np.any(scale[1::every])
