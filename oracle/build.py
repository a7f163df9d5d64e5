"""Build recipe of the oracle's C++ restatement (the CPU baseline): g++ -> oracle/libfss_cpu.so.

Test infrastructure only.  Portable ISA level (x86-64-v3) because the library is built here and
run on the GPU host as well.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_baseline.cpp")
OUT = os.path.join(HERE, "libfss_cpu.so")
FLAGS = ["-O3", "-std=c++17", "-march=x86-64-v3", "-pthread", "-fPIC", "-shared",
         "-fno-math-errno", "-ffp-contract=fast", "-fopenmp-simd"]


def build_oracle(verbose: bool = False) -> str:
    fd, tmp = tempfile.mkstemp(suffix=".so", dir=HERE, prefix=".libfss_cpu.")
    os.close(fd)
    cmd = ["g++", *FLAGS, "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        os.unlink(tmp)
        raise RuntimeError(f"g++ failed:\n{res.stderr[-4000:]}")
    os.chmod(tmp, 0o755)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_oracle(verbose=True))
