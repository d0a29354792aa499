"""The FP64 logarithm of the log pair kernel (csrc/fastlog.hpp: atanh series with a
Cody-Waite split of ln 2) against the long-double logarithm on the host, over the
range of 1 - x.y the kernel sees: at most 2.5 units in the last place."""
import os
import subprocess

from conftest import REPO


def test_log_pos_accuracy(tmp_path):
    exe = str(tmp_path / "fastlog_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-DSBSLPM_HD=", "-I",
                           os.path.join(REPO, "spectral-bslpm_amd", "csrc"),
                           os.path.join(REPO, "tests", "cpp", "fastlog_check.cpp"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
