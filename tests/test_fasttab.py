"""The FP64 logarithm and exponentials of the FP64 pair kernels (csrc/fasttab.hpp: Tang's
table-driven method with 1,024 logarithm cells, a 128-entry general exponential table and a
3,840-entry table of 2^(n/256) for arguments in [-10.375, 0], built in long double) against the
long-double logarithm and exponential on the host, over the ranges the kernels see
(1 - x.y in [1e-15, 2], near 1 and near 2, every cell boundary; (t^2 - 1)/eps in [-10, 0]) and
beyond: at most 2 units in the last place."""
import os
import subprocess

from conftest import REPO


def test_log_and_exp_accuracy(tmp_path):
    exe = str(tmp_path / "fasttab_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-DSBSLPM_HD=", "-I",
                           os.path.join(REPO, "spectral-bslpm_amd", "csrc"),
                           os.path.join(REPO, "tests", "cpp", "fasttab_check.cpp"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr


def test_folded_laplacian_kernel(tmp_path):
    """lapkernel.hpp's folded coefficients (FP64 in powers of t, FP32 in powers of the
    exponent u) against the reference's K(t) in long double and its central-difference
    derivative, in the saturated form (t >= 0.6) and the erf form over [-1, 1]."""
    exe = str(tmp_path / "lapkernel_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(REPO, "spectral-bslpm_amd", "csrc"),
                           os.path.join(REPO, "tests", "cpp", "lapkernel_check.cpp"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
