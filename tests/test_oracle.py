"""Pin the CPU oracle (oracle/stc_oracle.c) before trusting it as the GPU checker.

The reference ships no fixtures for the sweep, so each oracle function is
checked against something independent of its own formula:
  - LAPACK golden vectors written by tests/golden/gen_golden.py (scipy's
    OpenBLAS zgetrf/zgetrs/cgetrf/cgetrs): pivots, factors, 'N'/'T'/'C'
    solves in both precisions, the zero-pivot info;
  - the normwise backward error against an independent long-double
    evaluation (tests/berr_check.py);
  - brute force for n = 1, 2, 3 (Cramer's rule / explicit inverse);
  - the closed-form modal solution X = Phi diag(1/lam) Phi^T F of the
    synthetic ROMs, and the exact condition number from lam;
  - the invariant ||P A - L U|| <= c n u ||A||;
  - the closed-form adjoint gradient and central finite differences;
  - planted mistakes (wrong sign of the damping term, i f instead of i w)
    that the closed-form check must reject.
"""
import os

import numpy as np
import pytest

import oracle
from staccato_amd import synth

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "lapack_small.npz"))
SIZES = (1, 2, 3, 5, 8, 17, 33, 60)


@pytest.mark.parametrize("n", SIZES)
def test_zgetrf_matches_lapack(n):
    LU, ipiv, info = oracle.zgetrf(GOLD[f"A_{n}"])
    assert info == 0
    np.testing.assert_array_equal(ipiv, GOLD[f"ipiv_{n}"])
    ref = GOLD[f"LU_{n}"]
    assert np.abs(LU - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * n


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_zgetrs_matches_lapack(n, trans):
    LU, ipiv, _ = oracle.zgetrf(GOLD[f"A_{n}"])
    X = oracle.zgetrs(LU, ipiv, GOLD[f"B_{n}"], trans)
    ref = GOLD[f"X{trans}_{n}"]
    assert np.abs(X - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * n


@pytest.mark.parametrize("n", SIZES)
def test_cgetrf_matches_lapack(n):
    LU, ipiv, info = oracle.cgetrf(GOLD[f"A_{n}"].astype(np.complex64))
    assert info == 0
    np.testing.assert_array_equal(ipiv, GOLD[f"ipiv32_{n}"])
    ref = GOLD[f"LU32_{n}"]
    assert np.abs(LU - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max()) * n


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_cgetrs_matches_lapack(n, trans):
    """cgetrs on the cgetrf fixture factors against LAPACK's cgetrs."""
    X = oracle.cgetrs(GOLD[f"LU32_{n}"], GOLD[f"ipiv32_{n}"], GOLD[f"B_{n}"], trans)
    ref = GOLD[f"X32{trans}_{n}"]
    assert np.abs(X - ref).max() <= 2e-6 * max(1.0, np.abs(ref).max()) * n


def test_cgetrs_rejects_planted_transpose():
    """The fixture tells 'T' from 'C' (a dropped conjugation fails)."""
    n = 8
    X = oracle.cgetrs(GOLD[f"LU32_{n}"], GOLD[f"ipiv32_{n}"], GOLD[f"B_{n}"], "T")
    ref = GOLD[f"X32C_{n}"]
    assert np.abs(X - ref).max() > 1e-2 * np.abs(ref).max()


@pytest.mark.parametrize("n", [1, 3, 17, 60])
def test_backward_error_matches_extended_precision(n):
    """orc_backward_error against the long-double evaluation of
    ||f - A x||_inf / (||A||_inf ||x||_inf + ||f||_inf) (tests/berr_check.py),
    on solutions perturbed to berr ~ 1e-8, and planted mistakes in the
    long-double checker (1-norm of the vectors, |re|+|im| instead of the
    modulus, ||f|| dropped) disagree with the oracle."""
    from berr_check import berr
    rng = np.random.default_rng(40 + n)
    for trial in range(5):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        x = np.linalg.solve(A, f)
        x = x * (1 + 1e-8 * (rng.standard_normal(n) + 1j * rng.standard_normal(n)))
        ex = float(berr(A, x, f)[0])
        got = oracle.backward_error(A, x, f)
        assert abs(got - ex) <= 2 * (n + 2) * 2.0 ** -53 + 1e-9 * ex, (got, ex)
        assert ex > 1e-11
        for kw in ({"vec_norm": "one"}, {"elem": "cabs1"}, {"drop_f": True}):
            if n == 1 and "vec_norm" in kw:
                continue     # one entry: both vector norms agree
            bad = float(berr(A, x, f, **kw)[0])
            assert abs(bad - got) > 1e-6 * ex, kw


def test_zero_pivot_info_and_ipiv():
    LU, ipiv, info = oracle.zgetrf(GOLD["S_sing"])
    assert info == int(GOLD["info_sing"])
    np.testing.assert_array_equal(ipiv, GOLD["ipiv_sing"])
    np.testing.assert_allclose(LU, GOLD["LU_sing"], atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_brute_force_small(n):
    rng = np.random.default_rng(7 + n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    det = np.linalg.det(A)
    x_cramer = np.empty(n, dtype=np.complex128)
    for k in range(n):
        Ak = A.copy()
        Ak[:, k] = b
        x_cramer[k] = np.linalg.det(Ak) / det
    LU, ipiv, _ = oracle.zgetrf(A)
    x = oracle.zgetrs(LU, ipiv, b[:, None])[:, 0]
    np.testing.assert_allclose(x, x_cramer, rtol=1e-12)


@pytest.mark.parametrize("n", [8, 60, 128])
def test_lu_backward_invariant(n):
    rom, _ = synth.modal_rom(n, 11, 0.02, alpha=1e-4, beta=0.3)
    A = synth.assemble(rom, 123.4)
    LU, ipiv, _ = oracle.zgetrf(A)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for k in range(n):
        p = ipiv[k] - 1
        PA[[k, p]] = PA[[p, k]]
    assert np.abs(PA - L @ U).max() <= 4 * n * 2.2e-16 * np.abs(A).max()
    # pivoting on |re|+|im| (izamax) bounds the multipliers by sqrt(2) in modulus
    assert np.abs(np.tril(LU, -1)).max() <= np.sqrt(2.0) * (1 + 1e-12)


def test_assembly_matches_definition():
    """A(w) = (1+i eta) K + i w D - w^2 M (KrylovROMSubstructure.cpp:1412-1420)."""
    rom, _ = synth.modal_rom(9, 5, 0.03, alpha=2e-4, beta=0.7)
    w = 2 * np.pi * 77.0
    A = oracle.assemble(rom.K, rom.M, rom.D, rom.eta, w)
    lam = (1 + 1j * rom.eta) * rom.kdiag + 1j * w * rom.ddiag - w * w * rom.mdiag
    np.testing.assert_allclose(A, (rom.Phi * lam) @ rom.Phi.T, rtol=0, atol=1e-9 * np.abs(lam).max())


def test_sweep_n60_closed_form():
    c = synth.single_n60()
    X, info, be, _ = oracle.sweep(c.rom.K, c.rom.M, c.rom.D, c.rom.eta, c.freq, c.F, nthreads=4)
    assert (info == 0).all()
    assert be.max() <= 1e-14
    Xc = synth.closed_form(c.rom, c.freq, c.F)
    kappa = synth.cond2(c.rom, c.freq)
    fe = np.abs(X - Xc).max(axis=(1, 2)) / np.abs(Xc).max(axis=(1, 2))
    assert (fe <= 1e-14 * kappa + 1e-14).all()


def test_sweep_closed_form_rejects_planted_mistakes():
    """Planted in the oracle's inputs: a flipped damping sign and the i*f (Hz)
    damping term of test/IJCSA/src/solver/assembly.cu:67-74 both fail."""
    rom, rng = synth.modal_rom(20, 3, 0.02, alpha=1e-4, beta=0.5)
    F = synth.complex_normal(rng, (20, 2))
    f = np.linspace(5.0, 900.0, 40)
    Xc = synth.closed_form(rom, f, F)
    ok, _, _, _ = oracle.sweep(rom.K, rom.M, rom.D, rom.eta, f, F)
    assert np.abs(ok - Xc).max() / np.abs(Xc).max() < 1e-10
    bad, _, _, _ = oracle.sweep(rom.K, rom.M, -rom.D, rom.eta, f, F)
    assert np.abs(bad - Xc).max() / np.abs(Xc).max() > 1e-3
    # i f D instead of i w D: emulate by scaling D by 1/(2 pi)
    bad2, _, _, _ = oracle.sweep(rom.K, rom.M, rom.D / (2 * np.pi), rom.eta, f, F)
    assert np.abs(bad2 - Xc).max() / np.abs(Xc).max() > 1e-3


def test_special_case_static_limit():
    """w -> 0, eta = 0: A = K, X = K^{-1} F."""
    rom, rng = synth.modal_rom(16, 2, 0.0)
    F = synth.complex_normal(rng, (16, 1))
    X, info, _, _ = oracle.sweep(rom.K, rom.M, None, 0.0, np.array([0.0]), F)
    np.testing.assert_allclose(X[0], np.linalg.solve(rom.K, F), rtol=1e-10)


def test_refined_oracle_converges():
    c = synth.uniform_n256(12)
    X, info, iters, be, _ = oracle.sweep_refined(c.rom.K, c.rom.M, c.rom.D, c.rom.eta, c.freq, c.F,
                                                 tol=1e-10, max_iter=10, nthreads=4)
    assert (info == 0).all()
    assert be.max() <= 1e-10
    assert 1 <= iters.max() <= 4
    Xc = synth.closed_form(c.rom, c.freq, c.F)
    kappa = synth.cond2(c.rom, c.freq)
    fe = np.abs(X - Xc).max(axis=(1, 2)) / np.abs(Xc).max(axis=(1, 2))
    assert (fe <= 4e-10 * kappa).all()


def test_adjoint_oracle_closed_form_and_fd():
    c = synth.adjoint_n128(24)
    rom, cv = c.rom, c.extra["c"]
    J, ds, de, ns = oracle.adjoint(rom.K, rom.M, rom.D, rom.eta, c.freq, c.F, cv, rom.Phi,
                                   rom.kdiag, nthreads=4)
    assert ns == 0
    Jc, dsc, dec = synth.adjoint_closed_form(rom, c.freq, c.F, cv)
    assert abs(J / Jc - 1) < 1e-10
    assert np.abs(ds - dsc).max() <= 1e-8 * np.abs(dsc).max()
    assert abs(de / dec - 1) < 1e-8
    # central finite differences of the oracle sweep itself (h = 1e-6)
    h = 1e-6

    def Jof(s, eta):
        K = (rom.Phi * (s * rom.kdiag)) @ rom.Phi.T
        K = 0.5 * (K + K.T)
        X, _, _, _ = oracle.sweep(K, rom.M, rom.D, eta, c.freq, c.F, nthreads=4, berr=False)
        return float(np.sum(np.abs(np.einsum("i,fir->fr", cv.conj(), X)) ** 2))

    # FD at h = 1e-6 carries the rounding noise of J (~ kappa u J / h), so it is
    # gated norm-wise; the closed form above is the tight pin
    s1 = np.ones(rom.n)
    gmax = np.abs(ds).max()
    for k in (0, 17, 101):
        e = np.zeros(rom.n)
        e[k] = h
        fd = (Jof(s1 + e, rom.eta) - Jof(s1 - e, rom.eta)) / (2 * h)
        assert abs(fd - ds[k]) <= 1e-5 * gmax, (k, fd, ds[k])
    fd = (Jof(s1, rom.eta + h) - Jof(s1, rom.eta - h)) / (2 * h)
    assert abs(fd - de) <= 1e-6 * abs(de)


def test_lapack_baseline_matches_oracle():
    if oracle.openblas_path() is None:
        pytest.skip("no OpenBLAS shared object next to scipy")
    c = synth.uniform_n256(8)
    r = oracle.sweep_lapack(c.rom.K, c.rom.M, c.rom.D, c.rom.eta, c.freq, c.F, nthreads=2)
    assert r is not None
    X, info, _ = r
    Xo, _, _, _ = oracle.sweep(c.rom.K, c.rom.M, c.rom.D, c.rom.eta, c.freq, c.F, nthreads=2)
    assert (info == 0).all()
    assert np.abs(X - Xo).max() / np.abs(Xo).max() < 1e-10
