"""Python host interface of the MI355X frequency sweep (over the C ABI).

Mirrors the reference's sweep entry (KrylovROMSubstructure::backTransformKMOR,
src/analysis/KrylovROMSubstructure.cpp:1393-1460): the ROM operators are set
once, then every call solves A(w) X = F for a list of frequencies in Hz
(w = 2 pi f, :1408).  Arrays are numpy complex128; matrices are passed to the
library column-major (the layout of MKL_Complex16 arrays in the reference).
Errors come back as StcError with the library status, never as exit().
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import STC_PLAIN_FP64, STC_REFINED_FP32, Stats, SysDesc

MODES = {"plain": STC_PLAIN_FP64, "fp64": STC_PLAIN_FP64,
         "refined": STC_REFINED_FP32, "fp32": STC_REFINED_FP32}


class StcError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {_lib.strerror(code)} ({code})")
        self.code = code


def _mode(mode):
    if isinstance(mode, int):
        return mode
    return MODES[mode]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _fortran_c128(a, ndim=2):
    a = np.asarray(a, dtype=np.complex128)
    if a.ndim == 1 and ndim == 2:
        a = a[:, None]
    return np.asfortranarray(a)


def _stream_arg(stream):
    """None selects the handle's own stream (NULL at the C ABI).  The null
    stream (0, torch's default) cannot be named through the ABI, and work on
    it is not ordered with the handle's non-blocking stream: refuse it."""
    if stream is None:
        return None
    if int(stream) == 0:
        raise ValueError("stream 0 (the null stream) cannot be passed: NULL selects the "
                         "handle's stream; run the torch work on a torch.cuda.Stream")
    return ctypes.c_void_p(int(stream))


def _check_tensor(t, name, dtype, shape, device=None):
    """Refuse a tensor the C ABI would read or write out of bounds: wrong
    dtype, shape, layout (the ABI takes packed column-major blocks) or device."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if t.device.type != "cuda" or (device is not None and t.device.index != device):
        raise ValueError(f"{name}: must live on cuda:{device}, not {t.device}")
    return t


def _check(rc, what, soft=(0,)):
    if rc not in soft:
        raise StcError(rc, what)
    return rc


class Handle:
    """One device, one stream, a set of ROMs (stc_create / stc_destroy)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _check(self.lib.stc_create(ctypes.byref(h), int(device)), "stc_create")
        self.h = h
        self.n = {}
        self.device = int(device)

    def close(self):
        if self.h:
            self.lib.stc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_ok(self):
        return bool(self.lib.stc_device_ok(self.h))

    # ---- launch policy ------------------------------------------------------------
    def set_policy(self, n, mode, kernel=0, lds_budget=0, jb_max=0):
        """Launch policy for systems of order n (or the nearest order with an
        entry) in `mode`: kernel 0 automatic, 1 register panel, 2 wide LDS
        panel, 3 wide LDS panel with Gauss 3M products (complex64), 4 wide LDS
        panel with the two-level update; lds_budget
        bytes of LDS per workgroup (0 automatic); jb_max
        widest panel (0 no cap).  None for kernel removes the entry."""
        if kernel is None:
            _check(self.lib.stc_set_policy(self.h, int(n), _mode(mode), None), "stc_set_policy")
            return
        p = _lib.Policy(int(kernel), int(lds_budget), int(jb_max))
        _check(self.lib.stc_set_policy(self.h, int(n), _mode(mode), ctypes.byref(p)), "stc_set_policy")

    def clear_policies(self):
        _check(self.lib.stc_clear_policies(self.h), "stc_clear_policies")

    # ---- operators --------------------------------------------------------------
    def rom_set(self, rom_id, K, M, D=None, eta=0.0):
        K = _fortran_c128(K)
        M = _fortran_c128(M)
        Dd = _fortran_c128(D) if D is not None else None
        n = K.shape[0]
        _check(self.lib.stc_rom_set(self.h, rom_id, n, _ptr(K), _ptr(M), _ptr(Dd), n,
                                    float(eta)), "stc_rom_set")
        self.n[rom_id] = n

    def rom_set_io(self, rom_id, B, C):
        B = _fortran_c128(B)
        C = _fortran_c128(C)
        _check(self.lib.stc_rom_set_io(self.h, rom_id, B.shape[1], _ptr(B), B.shape[0],
                                       C.shape[0], _ptr(C), C.shape[0]), "stc_rom_set_io")

    def rom_clear(self, rom_id):
        _check(self.lib.stc_rom_clear(self.h, rom_id), "stc_rom_clear")
        self.n.pop(rom_id, None)

    # ---- sweep ------------------------------------------------------------------
    def sweep(self, rom_id, freq, F, mode="plain", tol=0.0, max_iter=0, berr=True):
        """X[j] = A(2 pi freq[j])^{-1} F.  Returns (X (nfreq, n, nrhs), info, berr, stats)."""
        n = self.n[rom_id]
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        F = _fortran_c128(F)
        nf, nrhs = freq.size, F.shape[1]
        X = np.empty((nf, nrhs, n), dtype=np.complex128)
        info = np.zeros(nf, dtype=np.int32)
        be = np.empty((nf, nrhs), dtype=np.float64) if berr else None
        st = Stats()
        rc = self.lib.stc_sweep(self.h, rom_id, nf, _ptr(freq), nrhs, _ptr(F), F.shape[0],
                                _ptr(X), n, _ptr(info), _ptr(be), _mode(mode), float(tol),
                                int(max_iter), ctypes.byref(st))
        _check(rc, "stc_sweep", soft=(0, 1, 2))
        return X.transpose(0, 2, 1), info, be, st

    def sweep_dev(self, rom_id, freq_t, F_t, X_t, info_t, berr_t=None, mode="plain", tol=0.0,
                  max_iter=0, stream=None, noalloc=False):
        """Device-resident sweep on torch tensors (no copies, no synchronisation).
        F_t: (nrhs, n) complex128 contiguous (column-major n x nrhs), X_t:
        (nfreq, nrhs, n) complex128, info_t int32 (nfreq), berr_t float64 or None.
        noalloc: run through stc_sweep_dev_noalloc (the library's fixed per-GPU
        scratch array, no device allocation; calls overlapping on one GPU race)."""
        import torch
        n = self.n[rom_id]
        if F_t.dim() != 2:
            raise ValueError("F_t: expected (nrhs, n)")
        nrhs = F_t.shape[0]
        nf = freq_t.numel()
        _check_tensor(freq_t, "freq_t", torch.float64, (nf,), self.device)
        _check_tensor(F_t, "F_t", torch.complex128, (nrhs, n), self.device)
        _check_tensor(X_t, "X_t", torch.complex128, (nf, nrhs, n), self.device)
        _check_tensor(info_t, "info_t", torch.int32, (nf,), self.device)
        if berr_t is not None:
            _check_tensor(berr_t, "berr_t", torch.float64, (nf, nrhs), self.device)
        s = _stream_arg(stream)
        fn = self.lib.stc_sweep_dev_noalloc if noalloc else self.lib.stc_sweep_dev
        rc = fn(
            self.h, rom_id, freq_t.numel(), ctypes.c_void_p(freq_t.data_ptr()), nrhs,
            ctypes.c_void_p(F_t.data_ptr()), n, ctypes.c_void_p(X_t.data_ptr()), n,
            ctypes.c_void_p(info_t.data_ptr()),
            ctypes.c_void_p(berr_t.data_ptr()) if berr_t is not None else None,
            _mode(mode), float(tol), int(max_iter), s)
        _check(rc, "stc_sweep_dev_noalloc" if noalloc else "stc_sweep_dev")

    def getrf(self, rom_id, freq, mode="plain"):
        """LU factors (LAPACK layout, nfreq x n x n), 1-based pivots, info."""
        n = self.n[rom_id]
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        nf = freq.size
        dt = np.complex128 if _mode(mode) == STC_PLAIN_FP64 else np.complex64
        LU = np.empty((nf, n, n), dtype=dt)       # [f, col, row] in memory
        ipiv = np.empty((nf, n), dtype=np.int32)
        info = np.zeros(nf, dtype=np.int32)
        rc = self.lib.stc_getrf(self.h, rom_id, nf, _ptr(freq), _mode(mode), _ptr(LU),
                                _ptr(ipiv), _ptr(info))
        _check(rc, "stc_getrf", soft=(0, 1))
        return LU.transpose(0, 2, 1), ipiv, info

    def sweep_batch(self, systems, mode="refined", tol=0.0, max_iter=0):
        """systems: list of (rom_id, freq, F).  One persistent launch over all
        points.  Returns list of (X, info, berr) and the stats."""
        keep, descs, outs = [], (SysDesc * len(systems))(), []
        for i, (rid, freq, F) in enumerate(systems):
            n = self.n[rid]
            freq = np.ascontiguousarray(freq, dtype=np.float64)
            F = _fortran_c128(F)
            nf, nrhs = freq.size, F.shape[1]
            X = np.empty((nf, nrhs, n), dtype=np.complex128)
            info = np.zeros(nf, dtype=np.int32)
            be = np.empty((nf, nrhs), dtype=np.float64)
            keep += [freq, F]
            outs.append((X, info, be))
            d = descs[i]
            d.rom_id, d.nfreq, d.freq_hz, d.nrhs = rid, nf, freq.ctypes.data, nrhs
            d.F, d.ldf, d.X, d.ldx = F.ctypes.data, F.shape[0], X.ctypes.data, n
            d.info, d.berr = info.ctypes.data, be.ctypes.data
        st = Stats()
        rc = self.lib.stc_sweep_batch(self.h, len(systems), descs, _mode(mode), float(tol),
                                      int(max_iter), ctypes.byref(st))
        _check(rc, "stc_sweep_batch", soft=(0, 1, 2))
        return [(X.transpose(0, 2, 1), info, be) for X, info, be in outs], st

    def adjoint(self, rom_id, freq, F, c, Phi, kdiag, mode="plain", tol=0.0, max_iter=0):
        """J = sum_w sum_r |c^H X_r(w)|^2 and dJ/ds (modal stiffness scales), dJ/deta."""
        n = self.n[rom_id]
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        F = _fortran_c128(F)
        c = np.ascontiguousarray(c, dtype=np.complex128)
        Phi = np.asfortranarray(Phi, dtype=np.float64)
        kdiag = np.ascontiguousarray(kdiag, dtype=np.float64)
        p = Phi.shape[1]
        J, de = ctypes.c_double(), ctypes.c_double()
        ds = np.zeros(p, dtype=np.float64)
        info = np.zeros(freq.size, dtype=np.int32)
        rc = self.lib.stc_adjoint(self.h, rom_id, freq.size, _ptr(freq), F.shape[1], _ptr(F),
                                  n, _ptr(c), p, _ptr(Phi), n, _ptr(kdiag), _mode(mode),
                                  float(tol), int(max_iter), ctypes.byref(J), _ptr(ds),
                                  ctypes.byref(de), _ptr(info))
        _check(rc, "stc_adjoint", soft=(0, 1, 2))
        return J.value, ds, de.value, info

    def back_transform(self, rom_id, freq, input_load, nin, nout, transfer=True, mode="plain",
                       tol=0.0, max_iter=0):
        """y(w) = C_R A(w)^{-1} B_R inputLoad (nfreq, nout, nrhs) and, when
        transfer, H(w) = C_R A(w)^{-1} B_R (nfreq, nout, nin)."""
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        L = _fortran_c128(input_load)
        nrhs = L.shape[1]
        nf = freq.size
        Y = np.empty((nf, nrhs, nout), dtype=np.complex128)
        H = np.empty((nf, nin, nout), dtype=np.complex128) if transfer else None
        info = np.zeros(nf, dtype=np.int32)
        st = Stats()
        rc = self.lib.stc_back_transform(self.h, rom_id, nf, _ptr(freq), _ptr(L), nrhs, _ptr(Y),
                                         _ptr(H), _ptr(info), _mode(mode), float(tol),
                                         int(max_iter), ctypes.byref(st))
        _check(rc, "stc_back_transform", soft=(0, 1, 2))
        return (Y.transpose(0, 2, 1), H.transpose(0, 2, 1) if transfer else None, info, st)

    # ---- explicit matrices ------------------------------------------------------
    def gesv_batched(self, A, B, mode="plain", tol=0.0, max_iter=0, berr=True):
        """X[j] = A[j]^{-1} B[j] (or B shared: 2-D) for explicit matrices A
        (batch, n, n), with pivoted LU as zgetrf/zgetrs.  Returns (X (batch,
        n, nrhs), info, berr, stats)."""
        A = np.asarray(A, dtype=np.complex128)
        nb, n = A.shape[0], A.shape[1]
        Ac = np.ascontiguousarray(A.transpose(0, 2, 1))      # column-major blocks
        B = np.asarray(B, dtype=np.complex128)
        if B.ndim == 1:
            B = B[:, None]
        shared = B.ndim == 2
        Bc = np.ascontiguousarray(B.T if shared else B.transpose(0, 2, 1))
        nrhs = Bc.shape[-2]
        X = np.empty((nb, nrhs, n), dtype=np.complex128)
        info = np.zeros(nb, dtype=np.int32)
        be = np.empty((nb, nrhs), dtype=np.float64) if berr else None
        st = Stats()
        rc = self.lib.stc_gesv_batched(self.h, n, nb, _ptr(Ac), n, n * n, nrhs, _ptr(Bc), n,
                                       0 if shared else n * nrhs, _ptr(X), n, n * nrhs,
                                       _ptr(info), _ptr(be), _mode(mode), float(tol),
                                       int(max_iter), ctypes.byref(st))
        _check(rc, "stc_gesv_batched", soft=(0, 1, 2))
        return X.transpose(0, 2, 1), info, be, st

    def gesv_batched_dev(self, A_t, B_t, X_t, info_t, berr_t=None, mode="plain", tol=0.0,
                         max_iter=0, stream=None):
        """Device-resident batched solve on torch tensors (no copies, no
        synchronisation).  A_t: (batch, n, n) complex128 holding each matrix
        column-major (A[j] transposed in memory), B_t: (batch, nrhs, n)
        column-major blocks, X_t: (batch, nrhs, n), info_t int32 (batch)."""
        import torch
        if A_t.dim() != 3 or B_t.dim() != 3:
            raise ValueError("A_t: expected (batch, n, n), B_t: (batch, nrhs, n)")
        nb, n = A_t.shape[0], A_t.shape[1]
        nrhs = B_t.shape[1]
        _check_tensor(A_t, "A_t", torch.complex128, (nb, n, n), self.device)
        _check_tensor(B_t, "B_t", torch.complex128, (nb, nrhs, n), self.device)
        _check_tensor(X_t, "X_t", torch.complex128, (nb, nrhs, n), self.device)
        _check_tensor(info_t, "info_t", torch.int32, (nb,), self.device)
        if berr_t is not None:
            _check_tensor(berr_t, "berr_t", torch.float64, (nb, nrhs), self.device)
        s = _stream_arg(stream)
        rc = self.lib.stc_gesv_batched_dev(
            self.h, n, nb, ctypes.c_void_p(A_t.data_ptr()), n * n, nrhs,
            ctypes.c_void_p(B_t.data_ptr()), n, n * nrhs, ctypes.c_void_p(X_t.data_ptr()), n,
            ctypes.c_void_p(info_t.data_ptr()),
            ctypes.c_void_p(berr_t.data_ptr()) if berr_t is not None else None,
            _mode(mode), float(tol), int(max_iter), s)
        _check(rc, "stc_gesv_batched_dev")


class Plan:
    """A captured (hipGraph) persistent launch over device-resident systems."""

    def __init__(self, handle, descs, mode="refined", tol=0.0, max_iter=0):
        self.handle = handle
        self.descs = descs
        p = ctypes.c_void_p()
        _check(handle.lib.stc_plan_create(handle.h, len(descs), descs, _mode(mode), float(tol),
                                          int(max_iter), ctypes.byref(p)), "stc_plan_create")
        self.p = p

    def execute(self, stream=None):
        _check(self.handle.lib.stc_plan_execute(self.p, _stream_arg(stream)), "stc_plan_execute")

    def close(self):
        if self.p:
            self.handle.lib.stc_plan_destroy(self.p)
            self.p = None
