"""Build recipe of the oracle's C++ restatement (the CPU baseline): g++ -> oracle/libfss_cpu.so.

Test infrastructure only.  Two translation units: the pair-term kernel (cpu_kernels.cpp) with
-ffast-math, so its exp calls vectorise through glibc's libmvec and its sums may be reassociated
(the numba kernels' fastmath), and the optimiser tail (cpu_baseline.cpp) without it.  Base ISA
x86-64-v3 (AVX2 + FMA), because the library is built here and run on the GPU host as well; the
kernel also carries an AVX-512 clone that the loader selects where the host has it.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
KERNEL = os.path.join(HERE, "cpu_kernels.cpp")
SRC = os.path.join(HERE, "cpu_baseline.cpp")
OUT = os.path.join(HERE, "libfss_cpu.so")
COMMON = ["-O3", "-std=c++17", "-march=x86-64-v3", "-fPIC"]
KERNEL_FLAGS = COMMON + ["-ffast-math", "-fopenmp-simd"]
TAIL_FLAGS = COMMON + ["-pthread", "-fno-math-errno", "-ffp-contract=fast"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{res.stderr[-4000:]}")


def build_oracle(verbose: bool = False) -> str:
    with tempfile.TemporaryDirectory(dir=HERE, prefix=".build.") as d:
        ko, to = os.path.join(d, "kernels.o"), os.path.join(d, "tail.o")
        _run(["g++", *KERNEL_FLAGS, "-c", "-o", ko, KERNEL], verbose)
        _run(["g++", *TAIL_FLAGS, "-c", "-o", to, SRC], verbose)
        tmp = os.path.join(d, "libfss_cpu.so")
        _run(["g++", "-shared", "-pthread", "-o", tmp, ko, to, "-lm"], verbose)
        os.chmod(tmp, 0o755)
        os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_oracle(verbose=True))
