"""Time the CPU baseline (oracle/libfss_cpu.so) on a saved workload sample, in its own process.

Test infrastructure: bench.py's cpu_baseline leg runs this as a child process so the CPU
library never shares a process with the HIP runtime.  Usage:
    python -m oracle.cpu_bench sample.npz   -> one JSON line {"seconds", "pairs", "score", ...}
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np


def main(path: str) -> None:
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from fast_shepherd_score_amd.align import OptimSpec, ScoreSpec
    from oracle import cpu_ref
    d = np.load(path, allow_pickle=False)
    n = int(d["n"])
    split = lambda flat, cnt: np.split(flat, np.cumsum(cnt)[:-1]) if len(cnt) else []
    ref = split(d["ref"], d["ref_n"])
    fit = split(d["fit"], d["fit_n"])
    rw = split(d["ref_w"], d["ref_n"]) if "ref_w" in d else None
    fw = split(d["fit_w"], d["fit_n"]) if "fit_w" in d else None
    spec = ScoreSpec(mode=int(d["mode"]), alpha=float(d["alpha"]), lam=float(d["lam"]))
    opt = OptimSpec(max_steps=int(d["steps"]), lr=float(d["lr"]), early_stop=bool(d["early_stop"]))
    cores = len(os.sched_getaffinity(0))
    hb = cpu_ref.HostBatch(ref, fit, rw, fw)
    q, t = d["q"], d["t"]
    warm = OptimSpec(max_steps=2, lr=opt.lr, early_stop=False)
    cpu_ref.align(cpu_ref.HostBatch(ref[:1], fit[:1], rw[:1] if rw else None, fw[:1] if fw else None),
                  spec, q[:1], t[:1], warm, 1)
    t0 = time.perf_counter()
    out = cpu_ref.align(hb, spec, q, t, opt, cores)
    dt = time.perf_counter() - t0
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    print(json.dumps({"seconds": dt, "pairs": n, "cores": cores, "cpu_model": model,
                      "score": out["score"].tolist()}))


if __name__ == "__main__":
    main(sys.argv[1])
