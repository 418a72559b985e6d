"""CPU: the host C++ under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize/Makefile builds the oracle and the product's file readers
(scene_io.cpp: glTF JSON / GLB / accessors; image_io.cpp: PNG in and out,
Radiance .hdr) with -fsanitize=address,undefined (g++, host code only), and
tests/sanitize/run_sanitized.py drives them in a child process with libasan
preloaded: renders, traversal, light sampling and cube lookups of the
oracle; exported and malformed glTF files (negative and overflowing
offsets, bad strides, non-integer counts, broken PNGs); flat, run-length and
truncated .hdr files. Any sanitizer report ends the child with a non-zero
status. (The reader bug this guards against, an accessor byteOffset that
wrapped the range check, is reported as a heap-buffer-overflow there.)
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "sanitize"


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if out and os.path.isabs(out) and os.path.exists(out) else None


def test_host_code_is_clean_under_asan_and_ubsan():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("g++ has no AddressSanitizer runtime here")
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, str(HERE / "run_sanitized.py")], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0 and "SANITIZED OK" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]
