"""Build libmtlp.so (the gfx950 kernels + C++ engine + C ABI) in-tree.

    python -m mtlphys_amd.build        -> mtlphys_amd/libmtlp.so

hipcc cross-compiles for gfx950 without a GPU.  Objects go to
mtlphys_amd/_build/, the shared library next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmtlp.so")
INCLUDE = os.path.join(ROOT, "include")

SOURCES = ["engine.cpp", "adjoint.cpp", "state_io.cpp", "scenes.cpp", "scan.hip", "hash.hip", "solver.hip", "adjoint.hip"]
HEADERS = ["mtlp_kernels.h", "mtlp_device.h", "mtlp_stencil.h", "mtlp_host.h", "engine_impl.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every device compile: some of
# their forms have produced wrong results on this hardware.  The host pass
# warns that it does not know the feature; that is expected.
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-x", "hip",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function",
]


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for (o, err), src in zip(results, SOURCES):
            lines = [l for l in err.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                print(f"[{src}]\n" + "\n".join(lines))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
