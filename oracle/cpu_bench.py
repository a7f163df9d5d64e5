"""Time the CPU baseline (oracle/libfss_cpu.so) on saved workload samples, in its own process.

Test infrastructure: bench.py's cpu_baseline leg runs this as a child process so the CPU
library never shares a process with the HIP runtime.  Every sample is aligned once untimed
(warm-up: page-in, thread start, the kernel's ISA dispatch), then ``--runs`` times; the median
run is the figure.  Usage:
    python -m oracle.cpu_bench [--runs 3] sample.npz [more.npz ...]
prints one JSON line {"cores", "cpu_model", "vector_bits", "samples": [{"seconds", "runs",
"pairs", "score"}, ...]}.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np


def cpu_model() -> str:
    """The host CPU's model name, as lscpu prints it (else /proc/cpuinfo)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_sample(path: str, runs: int, cores: int) -> dict:
    from fast_shepherd_score_amd.align import OptimSpec, ScoreSpec
    from oracle import cpu_ref
    d = np.load(path, allow_pickle=False)
    split = lambda flat, cnt: np.split(flat, np.cumsum(cnt)[:-1]) if len(cnt) else []
    ref, fit = split(d["ref"], d["ref_n"]), split(d["fit"], d["fit_n"])
    rw = split(d["ref_w"], d["ref_n"]) if "ref_w" in d else None
    fw = split(d["fit_w"], d["fit_n"]) if "fit_w" in d else None
    spec = ScoreSpec(mode=int(d["mode"]), alpha=float(d["alpha"]), lam=float(d["lam"]))
    opt = OptimSpec(max_steps=int(d["steps"]), lr=float(d["lr"]), early_stop=bool(d["early_stop"]))
    hb = cpu_ref.HostBatch(ref, fit, rw, fw)
    q, t = d["q"], d["t"]
    cpu_ref.align(hb, spec, q, t, opt, cores)              # warm-up
    times, out = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        out = cpu_ref.align(hb, spec, q, t, opt, cores)
        times.append(time.perf_counter() - t0)
    return {"seconds": float(np.median(times)), "runs": times, "pairs": int(d["n"]),
            "score": out["score"].tolist()}


def main(argv=None) -> None:
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    ap = argparse.ArgumentParser()
    ap.add_argument("samples", nargs="+")
    ap.add_argument("--runs", type=int, default=3)
    a = ap.parse_args(argv)
    from oracle import cpu_ref
    cores = len(os.sched_getaffinity(0))
    samples = [run_sample(p, max(1, a.runs), cores) for p in a.samples]
    bits = cpu_ref.load().fss_cpu_vector_bits
    bits.restype = ctypes.c_int
    print(json.dumps({"cores": cores, "cpu_model": cpu_model(), "vector_bits": int(bits()),
                      "samples": samples}))


if __name__ == "__main__":
    main()
