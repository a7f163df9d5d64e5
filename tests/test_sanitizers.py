"""Sanitizer runs of the host C++ (CPU only, no GPU).

tests/sanitize/build.py builds with g++:
  * the screening service's host code (fss_screen.cpp over the HIP runtime API, linked with host
    stand-ins of the runtime and of the device entry points) under AddressSanitizer +
    UndefinedBehaviorSanitizer and under ThreadSanitizer, driven by several submitting, polling
    and library-appending threads, memory waits and a close with queries queued
    (screen_stress.cpp: every query's hits checked against the stand-in scores);
  * the oracle's C++ CPU baseline under AddressSanitizer + UndefinedBehaviorSanitizer
    (oracle_driver.cpp: four modes, 1 and 4 threads give the same results).
A sanitizer report makes the binary exit nonzero.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def binaries():
    spec = importlib.util.spec_from_file_location("sanitize_build", os.path.join(ROOT, "tests", "sanitize", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod, mod.build()


@pytest.mark.parametrize("name", ["screen_asan", "screen_tsan", "oracle_asan"])
def test_sanitized_run_is_clean(binaries, name):
    mod, exes = binaries
    res = mod.run(exes[name], timeout=600)
    out = res.stdout + res.stderr
    print(out[-3000:])
    assert res.returncode == 0, out[-6000:]
    assert "Sanitizer" not in res.stderr
    assert ("screen stress: OK" if name.startswith("screen") else "oracle driver: OK") in res.stdout
