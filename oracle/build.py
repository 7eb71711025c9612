"""Build recipe for the CPU oracle (test infrastructure only).

    python oracle/build.py        -> oracle/liboracle.so

The reference (rakdcolon/mtlphys) is Metal Shading Language plus metal-cpp
host code; neither compiles without Apple's toolchain, so there is no
oracle/_ref build.  This builds only the project's own C restatement.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle.so")
SRC = [os.path.join(HERE, "pbf_oracle.c")]
DEPS = SRC + [os.path.join(HERE, n) for n in ("pbf_oracle.h", "pbf_solver_impl.h", "pbf_adjoint_impl.h")]

# -ffp-contract=off: predict and the cell hash are separately rounded FP32
# operations (no FMA), which the GPU kernels reproduce bit for bit.  -mfma:
# the explicit fmaf of the neighbour test becomes the (correctly rounded)
# hardware instruction instead of a slow library call; contraction stays off.
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-mfma",
          "-fno-fast-math", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build(force=True, verbose=False):
    if not force and os.path.exists(OUT):
        newest = max(os.path.getmtime(p) for p in DEPS)
        if os.path.getmtime(OUT) >= newest:
            return OUT
    tmp = OUT + ".tmp"
    cmd = ["gcc", *CFLAGS, *SRC, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
