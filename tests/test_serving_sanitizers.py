"""The serving runtime (tpa_serve.cpp) with the CPU stand-in pool (tests/cpu_pool/mock_pool.cpp)
under AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer: the C++ harness
tests/cpu_pool/serve_harness.cpp replays the host serving scenarios (NLL and gradient, fits
alone and in a mix, the Hessian start, destroy with fits running, a rejected map, starts outside
one-sided bounds) and every variant must pass them with no sanitizer report. Host code only
(g++); no GPU code is built."""

import os
import subprocess

import pytest

from conftest import ROOT

SRC = [os.path.join(ROOT, "tests", "cpu_pool", "serve_harness.cpp"),
       os.path.join(ROOT, "tf-pwa_amd", "csrc", "tpa_serve.cpp"),
       os.path.join(ROOT, "tests", "cpu_pool", "mock_pool.cpp")]
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "tf-pwa_amd", "csrc")]

VARIANTS = {
    "plain": ([], {}),
    "asan_ubsan": (["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-fno-omit-frame-pointer"],
                   {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:exitcode=77",
                    "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}),
    # the batcher's timed wait on the system clock here: this toolchain's ThreadSanitizer does
    # not intercept the clock-wait libstdc++ uses for steady_clock and would report every
    # such wait as a double lock (tpa_serve.cpp, TPA_SERVE_WAIT_SYSTEM_CLOCK)
    "tsan": (["-fsanitize=thread", "-DTPA_SERVE_WAIT_SYSTEM_CLOCK"],
             {"TSAN_OPTIONS": "halt_on_error=1:exitcode=66"}),
}


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_serving_scenarios_under_sanitizers(variant, tmp_path):
    flags, env_extra = VARIANTS[variant]
    exe = str(tmp_path / ("serve_harness_" + variant))
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread"] + flags + INC + ["-o", exe] +
                   SRC, check=True)
    env = dict(os.environ)
    env.pop("LD_PRELOAD", None)  # the sanitizer runtime must come first in the process
    env.update(env_extra)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ALL PASSED" in r.stdout, out[-4000:]
    for marker in ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer",
                   "ERROR: LeakSanitizer"):
        assert marker not in out, out[-4000:]
