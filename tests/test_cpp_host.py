"""The C++ host layer (include/cil_hip_md.hpp: CellColony, Session, EulerHIP as an
MDIntegrator, RunConfig, GsdFile) builds against the C ABI and, on a GPU, matches the oracle
over 200 steps (FP64 within 1e-9 sigma; FP32 over neighbour lists within 1e-5 sigma) and over
two frames of the reference's seeded run read from its JSON file, written as GSD frames."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "euler_hip_check.cpp")
EXE = os.path.join(ROOT, "tests", "cpp", "euler_hip_check")


def _build():
    cmd = ["g++", "-std=c++17", "-O2", SRC, "-o", EXE, "-I", os.path.join(ROOT, "include"),
           "-L", os.path.join(ROOT, "cil_cuda_amd"), "-lcil_hip", "-L", os.path.join(ROOT, "oracle"), "-loracle",
           "-Wl,-rpath," + os.path.join(ROOT, "cil_cuda_amd") + ":" + os.path.join(ROOT, "oracle")]
    subprocess.run(cmd, check=True)
    return EXE


def test_cpp_host_layer_builds_and_reports_missing_device():
    from conftest import gpu_available

    exe = _build()
    if gpu_available():
        pytest.skip("GPU present: covered by the gpu test")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 3, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_euler_hip_matches_oracle(gpu, tmp_path):
    exe = EXE if os.path.exists(EXE) else _build()
    env = dict(os.environ, CIL_CHECK_CONFIG=os.path.join(ROOT, "tests", "golden", "configs",
                                                         "z0.001t3.16_v2.RunParams.json"),
               CIL_CHECK_OUT=str(tmp_path))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
