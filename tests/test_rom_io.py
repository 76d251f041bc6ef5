"""Dense MTX ROM files (the ROM export of KrylovROMSubstructure::exportROMToFiles,
AuxiliaryFunctions.cpp:80-110, read like srcPerformanceTest/io/io.cpp:17-73)
through the C++ host mirror.  File I/O only: no GPU call."""
import ctypes

import numpy as np
import pytest

from staccato_amd import _lib


@pytest.fixture(scope="module")
def host():
    _lib.load()
    h = ctypes.CDLL(_lib.HOST_LIB_PATH)
    h.stc_host_read_mtx_dense.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.c_void_p, ctypes.c_longlong]
    h.stc_host_write_mtx_dense.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return h


def _read(host, path, cap=10000, is_complex=1):
    r, c = ctypes.c_int(), ctypes.c_int()
    out = np.zeros(cap, dtype=np.complex128)
    rc = host.stc_host_read_mtx_dense(str(path).encode(), is_complex, ctypes.byref(r),
                                      ctypes.byref(c), out.ctypes.data, cap)
    return rc, r.value, c.value, out


def test_reads_reference_layout(host, tmp_path):
    A = np.arange(12).reshape(3, 4) + 1j * (100 + np.arange(12).reshape(3, 4))
    p = tmp_path / "x_myKR.dat"
    with open(p, "w") as f:   # the ROM export: header, "rows\tcols", column-major pairs
        f.write("%%MatrixMarket matrix coordinate complex dense column major\n3\t4\n")
        for v in A.flatten(order="F"):
            f.write(f"{v.real:.25e}\t{v.imag:.25e}\n")
        f.write("\n")
    rc, r, c, out = _read(host, p)
    assert rc == 0 and (r, c) == (3, 4)
    np.testing.assert_array_equal(out[:12].reshape(4, 3).T, A)
    rc, r, c, out = _read(host, p, is_complex=0)
    np.testing.assert_array_equal(out[:12].reshape(4, 3).T, A.real)


def test_write_read_round_trip_is_exact(host, tmp_path):
    rng = np.random.default_rng(1)
    A = np.asfortranarray(rng.standard_normal((5, 7)) + 1j * rng.standard_normal((5, 7)))
    p = tmp_path / "y.dat"
    host.stc_host_write_mtx_dense(str(p).encode(), 5, 7, A.ctypes.data)
    rc, r, c, out = _read(host, p)
    assert rc == 0 and (r, c) == (5, 7)
    np.testing.assert_array_equal(out[:35].reshape(7, 5).T, A)
    assert open(p).readline().startswith("%%MatrixMarket matrix coordinate complex dense column major")


def test_short_or_missing_file_is_an_error(host, tmp_path):
    p = tmp_path / "short.dat"
    p.write_text("%%MatrixMarket\n2\t2\n1.0\t0.0\n2.0\t0.0\n")
    assert _read(host, p)[0] == _lib.STC_EINVAL
    assert _read(host, tmp_path / "none.dat")[0] == _lib.STC_EINVAL
