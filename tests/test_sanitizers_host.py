"""The C host layer under ThreadSanitizer and under AddressSanitizer + UndefinedBehaviorSanitizer (CPU):
tests/native/threads_host.c drives the configuration registry, the learned-predictor slot (host inference)
and the per-thread error state from 8 threads at once; built with gcc over the host sources
(tests/native/Makefile). Any sanitizer report (data race, out-of-bounds access, use after free, leak,
undefined behaviour) or wrong read-back fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
MODEL = os.path.join(ROOT, "libxstream_amd", "params", "mlp_MI355X.txt")

SANITIZERS = {
    "tsan": ("ThreadSanitizer", dict(TSAN_OPTIONS="halt_on_error=0 exitcode=66")),
    "asan": ("Sanitizer", dict(ASAN_OPTIONS="detect_leaks=1 exitcode=66", UBSAN_OPTIONS="print_stacktrace=1")),
}


@pytest.mark.parametrize("san", sorted(SANITIZERS))
def test_host_layer_under_sanitizer(san):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    if not os.path.exists(os.path.join(ROOT, "libxstream_amd", "_build", "lxs_launch.o")):
        pytest.skip("libxstream_amd not built (its HIP objects are linked in)")
    subprocess.run(["make", "-C", NATIVE, "-j8", san], check=True, stdout=subprocess.DEVNULL)
    marker, env = SANITIZERS[san]
    run = subprocess.run([os.path.join(NATIVE, "_build", san, "threads_host"), MODEL], capture_output=True,
                         text=True, timeout=600, env=dict(os.environ, **env))
    out = run.stdout + run.stderr
    assert marker not in out and "runtime error" not in out, out[-4000:]
    assert run.returncode == 0, out[-2000:]
    assert " 0 failures" in out and " 0 learned lookups" not in out, out[-500:]
