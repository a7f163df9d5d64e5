"""In-tree build of libfss_hip.so for gfx950 (hipcc; no JIT cache, so the .so travels with the tree)."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
# Device code: the alignment kernels and the screening top-k merge.  Host-only C++: the error
# helpers and the screening service (HIP runtime API only, so it also builds with g++ against a
# host stand-in of that API for the sanitizer runs in tests/sanitize).
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("fss_kernels.hip", "fss_topk.hip", "fss_common.cpp",
                                                   "fss_screen.cpp")]
DEPS = [os.path.join(HERE, "csrc", f) for f in ("fss_device.h", "fss_common.h", "fss_topk.h")] + \
    [os.path.join(os.path.dirname(HERE), "include", "fss.h")]
OUT = os.path.join(HERE, "libfss_hip.so")
ARCH = os.environ.get("FSS_OFFLOAD_ARCH", "gfx950")
# Packed-FP32 instructions are disabled for every device pass.
HIPCC_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libfss_hip.so)")


def build_hip(verbose: bool = False) -> str:
    """Compile the HIP sources into fast_shepherd_score_amd/libfss_hip.so (always rebuilds)."""
    fd, tmp = tempfile.mkstemp(suffix=".so", dir=HERE, prefix=".libfss_hip.")
    os.close(fd)
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        os.unlink(tmp)
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-4000:]}")
    os.chmod(tmp, 0o755)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_hip(verbose=True))
