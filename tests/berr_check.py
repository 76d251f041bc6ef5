"""Independent evaluation of the normwise backward error of a computed solution.

    berr(x) = ||f - A x||_inf / (||A||_inf ||x||_inf + ||f||_inf)

with |z| the complex modulus, ||A||_inf the largest row sum of moduli and
||v||_inf the largest modulus of a vector.  This is the gate the sweep reports
(and the refined mode iterates on).  Here it is evaluated in numpy's extended
precision (80-bit long double, unit roundoff 5.4e-20) from A(w) formed exactly
from the FP64 operators, so it is independent of the GPU's FP64 evaluation and
of the oracle's C code.

`berr_bound` is how far an FP64 evaluation of the same quantity may stray: the
residual's rounding is at most gamma_{n+4} times sum_j (|K| + w|D| + w^2|M| +
eta|K|)_{ij} |x_j| (the assembly's terms before cancellation), over the same
denominator.
"""
import numpy as np

U64 = 2.0 ** -53


def assemble_ld(K, M, D, eta, f_hz):
    """A(w) = (1 + i eta) K + i w D - w^2 M in long double, w = 2 pi f formed in
    FP64 as the kernel and the oracle form it."""
    w = np.longdouble(2.0 * np.pi * float(f_hz))
    Kl = np.asarray(K).astype(np.clongdouble)
    A = Kl * (1 + 1j * np.longdouble(eta)) - (w * w) * np.asarray(M).astype(np.clongdouble)
    if D is not None:
        A = A + 1j * w * np.asarray(D).astype(np.clongdouble)
    return A


def abs_terms(K, M, D, eta, f_hz):
    w = 2.0 * np.pi * float(f_hz)
    T = np.abs(K) * (1 + abs(eta)) + w * w * np.abs(M)
    if D is not None:
        T = T + w * np.abs(D)
    return T


def berr(A, X, F, vec_norm="inf", elem="modulus", drop_f=False):
    """Per-column backward error of X (n x k) for A X = F, in long double.
    The keyword arguments plant mistakes for the tests that show the check
    discriminates: vec_norm="one" (sum instead of max), elem="cabs1"
    (|re| + |im| instead of the modulus), drop_f (no ||f|| term)."""
    A = np.asarray(A).astype(np.clongdouble)
    X = np.asarray(X).astype(np.clongdouble)
    F = np.asarray(F).astype(np.clongdouble)
    if X.ndim == 1:
        X, F = X[:, None], F[:, None]
    R = F - A @ X

    def mag(z):
        return np.abs(z.real) + np.abs(z.imag) if elem == "cabs1" else np.abs(z)

    def vn(v):
        return mag(v).sum(axis=0) if vec_norm == "one" else mag(v).max(axis=0)

    anorm = mag(A).sum(axis=1).max()
    den = anorm * vn(X) + (0 if drop_f else vn(F))
    return np.asarray(vn(R) / den, dtype=np.float64)


def berr_bound(T, X, F, A_norm):
    """Largest deviation of an FP64 evaluation of berr from the exact value."""
    n = T.shape[0]
    X = np.asarray(X)
    F = np.asarray(F)
    if X.ndim == 1:
        X, F = X[:, None], F[:, None]
    num = (n + 4) * U64 * ((T @ np.abs(X)).max(axis=0) + np.abs(F).max(axis=0))
    den = A_norm * np.abs(X).max(axis=0) + np.abs(F).max(axis=0)
    return 2.0 * num / den
