"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer and ThreadSanitizer.

tests/sanitize_check.cpp drives the library's KD-line permutation
(fpsample_amd/csrc/kdline.cpp, threaded for large subtrees) and every entry of
the oracle's C restatement (oracle/fps_oracle.c) over random, lattice,
duplicate and NaN/inf clouds, and compares the permutation with the oracle's.
It is compiled here with gcc/g++ only (no device code is involved), once with
``-fsanitize=address,undefined`` and once with ``-fsanitize=thread``; any
sanitizer report halts the run with a non-zero status.  A planted mistake per
build (a heap read one past the end, a racy counter) must be reported, which
shows the instrumentation is live rather than silently absent.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fpsample_amd", "csrc")
DRIVER = os.path.join(ROOT, "tests", "sanitize_check.cpp")
KDLINE = os.path.join(CSRC, "kdline.cpp")
ORACLE_C = os.path.join(ROOT, "oracle", "fps_oracle.c")

FLAVOURS = {
    "asan_ubsan": (["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer"],
                   {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:abort_on_error=0",
                    "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"},
                   "--plant-overflow", "AddressSanitizer"),
    "tsan": (["-fsanitize=thread"],
             {"TSAN_OPTIONS": "halt_on_error=1"},
             "--plant-race", "ThreadSanitizer: data race"),
}

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="gcc/g++ not available")


def _build(tmp_path, flags):
    obj = tmp_path / "fps_oracle.o"
    exe = tmp_path / "sanitize_check"
    common = ["-O1", "-g"] + flags
    subprocess.run(["gcc", "-std=c11", "-ffp-contract=off", *common, "-c", ORACLE_C,
                    "-o", str(obj)], check=True, capture_output=True, text=True)
    subprocess.run(["g++", "-std=c++17", *common, "-I", CSRC, DRIVER, KDLINE, str(obj),
                    "-o", str(exe), "-pthread"], check=True, capture_output=True, text=True)
    return exe


@pytest.mark.parametrize("flavour", sorted(FLAVOURS))
def test_host_code_is_sanitizer_clean(tmp_path, flavour):
    flags, opts, plant, report = FLAVOURS[flavour]
    exe = _build(tmp_path, flags)
    env = dict(os.environ, **opts)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0 and out.stdout.strip() == "OK", (out.stdout + out.stderr)[-4000:]
    assert "Sanitizer" not in out.stderr and "runtime error" not in out.stderr, out.stderr[-4000:]
    # The planted mistake is reported: the instrumentation is really there.
    bad = subprocess.run([str(exe), plant], capture_output=True, text=True, timeout=120, env=env)
    assert bad.returncode != 0 and report in bad.stderr, (bad.stdout + bad.stderr)[-2000:]
