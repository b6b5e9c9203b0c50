"""The C++ interface (include/pvfmm.hpp) in the reference's example1 call sequence
(examples/src/example1.cpp:59-100), built as tests/cpp/api_example.  On a machine without a
GPU the program checks the host-side steps and that SetupFMM reports the missing device as
pvfmm::Error; on the GPU it checks each kernel against Kernel::ker_poten."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "api_example")


def _built():
    if not os.path.exists(EXE):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp")])
    return EXE


def test_cpp_api_host_side():
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")  # hide any GPU: host-side path only
    r = subprocess.run([_built()], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "api_example: OK (gpu=0)" in r.stdout


@pytest.mark.gpu
def test_cpp_api_on_gpu(gpu):
    r = subprocess.run([_built()], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "api_example: OK (gpu=1)" in r.stdout
