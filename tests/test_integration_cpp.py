"""The C++ drop-in demo (tests/cpp/reference_api_demo.cpp) builds against
include/ and libsbslpm.so; host mode runs here, the GPU mode on the GPU."""
import os
import subprocess

import pytest

from conftest import REPO

DEMO = os.path.join(REPO, "tests", "cpp", "_build", "reference_api_demo")


def _build():
    os.makedirs(os.path.dirname(DEMO), exist_ok=True)
    lib = os.path.join(REPO, "spectral-bslpm_amd")
    cmd = ["hipcc", "-std=c++17", "-O2", "-I", os.path.join(REPO, "include"),
           os.path.join(REPO, "tests", "cpp", "reference_api_demo.cpp"), "-L", lib, "-lsbslpm",
           "-Wl,-rpath," + lib, "-o", DEMO]
    if not os.path.exists(DEMO) or os.path.getmtime(DEMO) < os.path.getmtime(cmd[5]):
        subprocess.check_call(cmd)
    return DEMO


def test_cpp_demo_host_mode(sb):
    sb.load()
    out = subprocess.run([_build()], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "interactions 17910" in out.stdout


@pytest.mark.gpu
def test_cpp_demo_reference_sequence_on_gpu(sb):
    demo = DEMO if os.path.exists(DEMO) else _build()
    out = subprocess.run([demo, "gpu"], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
