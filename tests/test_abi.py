"""The C-ABI library loads and exports every symbol include/staccato_sweep.h
declares; host-only entry points behave without a GPU.  No compute calls."""
import ctypes
import os
import re
import subprocess

import pytest

from staccato_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "staccato_sweep.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(stc_[a-z_]+)\s*\(", txt)))


def test_header_lists_every_binding():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_declared_symbols():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (stc_[a-z_]+)$", out, flags=re.M))
    assert set(declared_symbols()) <= exported


def test_library_is_built_for_gfx950():
    out = subprocess.run(["strings", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_strerror_and_invalid_arguments():
    lib = _lib.load()
    assert _lib.strerror(0) == "ok"
    assert "singular" in _lib.strerror(1)
    assert lib.stc_create(None, 0) == _lib.STC_EINVAL
    assert lib.stc_destroy(None) == _lib.STC_EINVAL
    assert lib.stc_device_ok(None) == 0
    assert lib.stc_rom_set(None, 0, 4, None, None, None, 4, 0.0) == _lib.STC_EINVAL
    assert lib.stc_plan_destroy(None) == _lib.STC_EINVAL
    pol = _lib.Policy(2, 80 * 1024, 0)
    assert lib.stc_set_policy(None, 256, 1, ctypes.byref(pol)) == _lib.STC_EINVAL
    assert lib.stc_clear_policies(None) == _lib.STC_EINVAL
    assert lib.stc_sweep_dev_noalloc(None, 0, 1, None, 1, None, 1, None, 1, None, None, 0, 0.0, 0,
                                     None) == _lib.STC_EINVAL
    assert lib.stc_gesv_batched(None, 4, 1, None, 4, 16, 1, None, 4, 4, None, 4, 4, None, None,
                                0, 0.0, 0, None) == _lib.STC_EINVAL
    assert lib.stc_gesv_batched_dev(None, 4, 1, None, 16, 1, None, 4, 4, None, 4, None, None,
                                    0, 0.0, 0, None) == _lib.STC_EINVAL


def test_autotune_cache_round_trip(tmp_path):
    from staccato_amd import autotune
    path = str(tmp_path / "t.json")
    assert autotune.load_cache(path) == {}
    c = {autotune._key(256, 4, "refined"): {"name": "x", "kernel": 2, "lds_budget": 81920,
                                            "jb_max": 0, "pts_per_s": 1.0}}
    autotune.save_cache(c, path)
    assert autotune.load_cache(path) == c
    assert [k for k, *_ in autotune.CANDIDATES][0] == "register-panel"


def test_null_stream_is_refused():
    """NULL at the ABI selects the handle's stream, so the Python wrappers
    refuse stream 0 (the null stream) instead of silently changing streams."""
    from staccato_amd.sweep import _stream_arg
    assert _stream_arg(None) is None
    assert _stream_arg(0x1234).value == 0x1234
    with pytest.raises(ValueError):
        _stream_arg(0)


def test_missing_library_fails_loudly(tmp_path):
    saved = _lib._lib
    try:
        _lib._lib = None
        with pytest.raises(OSError):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved
