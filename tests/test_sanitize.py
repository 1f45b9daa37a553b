"""Host sanitizers over the oracle and the library's host-only code.

oracle/san_driver.cpp runs the oracle's entry points (every scheme switch),
the initial-condition generator (threaded transforms, radix-2 and Bluestein
lengths) through a host stand-in handle, and the readers and writers of the
reference's file formats (write, read back, compare; error paths).  It is built
twice by oracle/Makefile: with AddressSanitizer + UndefinedBehaviorSanitizer
(every report fatal) and, for the threaded generator alone, with
ThreadSanitizer.  A report makes the driver exit non-zero; its last line is
"san_driver: ok" otherwise.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle"


def _build(target: str) -> Path:
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    r = subprocess.run(["make", "-C", str(ORACLE), target], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        if "sanitize" in r.stderr and ("cannot find" in r.stderr or "not supported" in r.stderr):
            pytest.skip(f"the toolchain has no {target} runtime")
        raise AssertionError(r.stderr[-2000:])
    return ORACLE / "_build" / ("san_driver" if target == "san" else "tsan_driver")


def _run(exe: Path, tmp_path, env_extra):
    env = dict(os.environ, **env_extra)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip().splitlines()[-1] == "san_driver: ok"
    assert "runtime error" not in r.stderr  # UBSan
    return r


def test_asan_ubsan(tmp_path):
    exe = _build("san")
    _run(exe, tmp_path, {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
                         "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1", "OMP_NUM_THREADS": "2"})


def test_tsan_generator(tmp_path):
    exe = _build("tsan")
    r = _run(exe, tmp_path, {"TSAN_OPTIONS": "halt_on_error=1"})
    assert "WARNING: ThreadSanitizer" not in r.stderr
