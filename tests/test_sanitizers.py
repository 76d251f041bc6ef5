"""The CPU suite once more with the oracle and the C++ host library built
with AddressSanitizer + UndefinedBehaviorSanitizer (tests/sanitize_host.sh);
a sanitizer report aborts the inner run and fails this test."""
import ctypes
import os
import subprocess

import pytest

from staccato_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))


def _have_asan():
    try:
        out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return False
    return os.path.isabs(out) and os.path.exists(out)


def test_overridden_libraries_are_the_ones_loaded():
    import oracle
    oracle.load()
    ctypes.CDLL(_lib.HOST_LIB_PATH)
    maps = open("/proc/self/maps").read()
    assert os.path.realpath(oracle.LIB) in maps
    assert os.path.realpath(_lib.HOST_LIB_PATH) in maps


@pytest.mark.skipif("STC_ORACLE_LIB" in os.environ, reason="already the sanitizer run")
@pytest.mark.skipif(not _have_asan(), reason="gcc's libasan is not installed")
def test_cpu_suite_under_asan_and_ubsan():
    r = subprocess.run(["bash", os.path.join(HERE, "sanitize_host.sh"), "-x"],
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-4000:], r.stderr[-4000:])
    assert " passed" in r.stdout
