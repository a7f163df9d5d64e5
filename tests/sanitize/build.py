"""Sanitizer builds of the library's host C++ and of the oracle's CPU baseline (test infrastructure).

g++ builds, no GPU and no HIP compiler needed:
  * screen_asan / screen_tsan: the screening service's host code
    (fast_shepherd_score_amd/csrc/fss_screen.cpp, fss_common.cpp; it includes only the HIP
    runtime API header) linked with the host stand-ins of hip_host.cpp and driven by
    screen_stress.cpp, under AddressSanitizer + UndefinedBehaviorSanitizer and under
    ThreadSanitizer (with tsan_clockwait.cpp: the condition-variable timed wait made visible to
    GCC 11's TSan runtime);
  * oracle_asan: oracle/cpu_baseline.cpp + oracle/cpu_kernels.cpp (with the flags of
    oracle/build.py) driven by oracle_driver.cpp, under AddressSanitizer + UndefinedBehaviorSanitizer.
Binaries go to tests/sanitize/_build/ (not committed).

    python tests/sanitize/build.py [--run]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
CSRC = os.path.join(ROOT, "fast_shepherd_score_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
ROCM_INCLUDE = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "include")

COMMON = ["-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-pthread"]
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
TSAN = ["-fsanitize=thread"]
SERVICE = ["-D__HIP_PLATFORM_AMD__", "-I", ROCM_INCLUDE]
SERVICE_SRC = [os.path.join(CSRC, "fss_screen.cpp"), os.path.join(CSRC, "fss_common.cpp"),
               os.path.join(HERE, "hip_host.cpp"), os.path.join(HERE, "screen_stress.cpp")]
# the oracle's own flags (oracle/build.py): fast-math only in the pair-term kernel unit
ORACLE_ISA = ["-march=x86-64-v3"]
ORACLE_KERNEL_FLAGS = ["-ffast-math", "-fopenmp-simd"]
ORACLE_TAIL_FLAGS = ["-fno-math-errno", "-ffp-contract=fast"]

ENV = {
    "ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:abort_on_error=0",
    "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
    "TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1",
}


def _cc(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{res.stderr[-4000:]}")


def build(verbose: bool = False) -> dict:
    os.makedirs(OUT, exist_ok=True)
    out = {}
    for name, san, extra in (("screen_asan", ASAN, []),
                             ("screen_tsan", TSAN, [os.path.join(HERE, "tsan_clockwait.cpp")])):
        exe = os.path.join(OUT, name)
        _cc(["g++", *COMMON, *san, *SERVICE, "-o", exe, *SERVICE_SRC, *extra], verbose)
        out[name] = exe
    exe = os.path.join(OUT, "oracle_asan")
    ko = os.path.join(OUT, "oracle_kernels_asan.o")
    _cc(["g++", *COMMON, *ASAN, *ORACLE_ISA, *ORACLE_KERNEL_FLAGS, "-c", "-o", ko,
         os.path.join(ORACLE, "cpu_kernels.cpp")], verbose)
    _cc(["g++", *COMMON, *ASAN, *ORACLE_ISA, *ORACLE_TAIL_FLAGS, "-o", exe, ko,
         os.path.join(ORACLE, "cpu_baseline.cpp"), os.path.join(HERE, "oracle_driver.cpp")], verbose)
    out["oracle_asan"] = exe
    return out


def run(exe: str, timeout: float = 600.0) -> subprocess.CompletedProcess:
    env = dict(os.environ, **ENV)
    return subprocess.run([exe], capture_output=True, text=True, timeout=timeout, env=env)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--run", action="store_true")
    a = ap.parse_args()
    exes = build(verbose=True)
    if a.run:
        rc = 0
        for name, exe in exes.items():
            res = run(exe)
            print(f"== {name}: exit {res.returncode}\n{res.stdout}{res.stderr[-3000:]}", flush=True)
            rc = rc or res.returncode
        sys.exit(rc)


if __name__ == "__main__":
    main()
