"""Build liblss_amd.so (HIP, gfx950) in-tree with hipcc.

Every HIP source is compiled with -ffp-contract=off (no fused multiply-add: the kernels keep the
reference's rounding) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "liblss_amd.so")
# the heaviest translation units first, so that a parallel build ends soonest
SOURCES = ([f"lss_adi_partitioned_{t}_{p}.hip" for t in ("f64", "f32") for p in ("generic", "heston", "sabr")] +
           [f"lss_adi_adjoint_{t}_{p}.hip" for t in ("f64", "f32") for p in ("generic", "heston", "sabr")] +
           ["lss_adi_kernels.hip", "lss_tridiag.hip", "lss_adi_capi.hip", "lss_heat1d_wave.hip", "lss_heat1d.hip"])
ARCH = os.environ.get("LSS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-result",
            f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]


def build(verbose: bool = False) -> str:
    import concurrent.futures as cf
    cc = hipcc()
    objs = [os.path.join(CSRC, src.replace(".hip", ".o")) for src in SOURCES]
    cmds = [[cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj] for src, obj in zip(SOURCES, objs)]
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8")), os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(jobs) as ex:
        for r in ex.map(lambda c: subprocess.run(c), cmds):
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, r.args)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
