"""Build recipe of libsbslpm.so (gfx950) with hipcc, in-tree.

Every HIP translation unit is compiled with packed-FP32 instruction
generation switched off; the host set-up (tree.cpp) is compiled with
-ffp-contract=off so its geometry rounds exactly as written.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsbslpm.so")
PROBE_LIB = os.path.join(HERE, "libsbslpm_probe.so")
ARCH = os.environ.get("SBSLPM_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["fmm_passes.hip", "fmm_interact.hip", "remap.hip", "direct.hip", "targets.hip", "swe.hip", "adjoint.hip", "remap_adjoint.hip", "output.hip", "irreg.hip", "swe_step.hip", "swe_adjoint.hip", "swe_ll.hip", "capi.hip"]
# MFMA results in VGPRs (gfx950 reads them from the unified register file), so
# the vector unit can take the reciprocal without accumulator-register copies.
EXTRA = {"fmm_interact.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
HOST_SOURCES = ["tree.cpp", "irreg_tree.cpp", "tree_api.cpp"]

COMMON = ["-std=c++17", "-O3", "-fPIC", "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        jobs.append(["hipcc", "-x", "hip", "-c", os.path.join(CSRC, s), "-o", obj] + COMMON + DEVICE + EXTRA.get(s, []))
    for s in HOST_SOURCES:
        obj = os.path.join(BUILD, s + ".o")
        jobs.append(["hipcc", "-c", os.path.join(CSRC, s), "-o", obj, "-ffp-contract=off"] + COMMON)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [j[j.index("-o") + 1] for j in jobs]
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # register-only issue-rate probes (bench.py --config probe), a library of their own
    src = os.path.join(CSRC, "probe", "issue_rates.hip")
    ptmp = PROBE_LIB + ".tmp"
    _run(["hipcc", "-x", "hip", src, "-shared", "-o", ptmp] + COMMON + DEVICE)
    os.replace(ptmp, PROBE_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
