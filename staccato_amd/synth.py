"""Synthetic Krylov ROMs with a known modal structure.

Every system is built the same way: Phi = Q from the QR factorisation of an
n x n N(0,1) matrix (sign of diag(R) folded into Q), modal masses
m_k ~ U[0.5, 2], modal frequencies f_k = 10**U[1, 3] Hz (log-uniform in
10..1000 Hz), w_k = 2 pi f_k, and

    M = Phi diag(m) Phi^T,   K = Phi diag(m w_k^2) Phi^T,
    D = alpha K + beta M     (Rayleigh; alpha = beta = 0: undamped).

K, M, D are real symmetric, so A(w) = (1 + i eta) K + i w D - w^2 M is complex
symmetric with the exact eigen-decomposition A = Phi diag(lam) Phi^T,
lam_k = (1 + i eta) m_k w_k^2 + i w (alpha m_k w_k^2 + beta m_k) - w^2 m_k.
Draws come from numpy.random.default_rng(seed) in this order: the N(0,1)
matrix, m, the exponents of f, then the load (real parts, imaginary parts).
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Rom:
    n: int
    K: np.ndarray
    M: np.ndarray
    D: Optional[np.ndarray]
    eta: float
    Phi: np.ndarray
    mdiag: np.ndarray
    kdiag: np.ndarray
    ddiag: np.ndarray
    seed: int


@dataclass
class Case:
    """One system of a workload: a ROM, its frequencies (Hz) and its load."""
    rom: Rom
    freq: np.ndarray
    F: np.ndarray
    extra: dict = field(default_factory=dict)


def modal_rom(n, seed, eta, alpha=0.0, beta=0.0, rng=None):
    rng = np.random.default_rng(seed) if rng is None else rng
    G = rng.standard_normal((n, n))
    Q, R = np.linalg.qr(G)
    sgn = np.sign(np.diag(R))
    sgn[sgn == 0] = 1.0
    Q = Q * sgn
    m = rng.uniform(0.5, 2.0, n)
    f = 10.0 ** rng.uniform(1.0, 3.0, n)
    wk = 2.0 * np.pi * f
    kd = m * wk ** 2
    dd = alpha * kd + beta * m

    def form(d):
        A = (Q * d) @ Q.T
        return 0.5 * (A + A.T)

    K = form(kd)
    M = form(m)
    D = form(dd) if (alpha != 0.0 or beta != 0.0) else None
    return Rom(n, K, M, D, float(eta), Q, m, kd, dd, seed), rng


def complex_normal(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def closed_form(rom, freq, F):
    """X(w) = Phi diag(1/lam(w)) Phi^T F for every frequency; (nfreq, n, nrhs)."""
    w = 2.0 * np.pi * np.asarray(freq, dtype=np.float64)
    lam = ((1.0 + 1j * rom.eta) * rom.kdiag[None, :] + 1j * w[:, None] * rom.ddiag[None, :]
           - (w ** 2)[:, None] * rom.mdiag[None, :])
    PF = rom.Phi.T @ F
    return np.einsum("ik,fk,kr->fir", rom.Phi, 1.0 / lam, PF)


def cond2(rom, freq):
    """Exact 2-norm condition number of A(w) per frequency."""
    w = 2.0 * np.pi * np.asarray(freq, dtype=np.float64)
    lam = ((1.0 + 1j * rom.eta) * rom.kdiag[None, :] + 1j * w[:, None] * rom.ddiag[None, :]
           - (w ** 2)[:, None] * rom.mdiag[None, :])
    a = np.abs(lam)
    return a.max(axis=1) / a.min(axis=1)


def assemble(rom, f_hz):
    """A(w) in complex128 (the dense matrix the sweep factors)."""
    w = 2.0 * np.pi * f_hz
    A = (1.0 + 1j * rom.eta) * rom.K - w * w * rom.M
    if rom.D is not None:
        A = A + 1j * w * rom.D
    return A


# ---- the five workloads -------------------------------------------------------------

def single_n60():
    """One reduced system n=60 (seed 0), undamped operators, eta=0.02, 1000
    frequencies linear in 1..1200 Hz, one unit force at DOF 0."""
    rom, _ = modal_rom(60, 0, 0.02)
    F = np.zeros((60, 1), dtype=np.complex128)
    F[0, 0] = 1.0
    return Case(rom, np.linspace(1.0, 1200.0, 1000), F)


def uniform_n256(nfreq=4096):
    """n=256 (seed 1), D = 1e-4 K + 0.5 M, eta=0.01, 4096 frequencies linear in
    1..2000 Hz, 4 complex N(0,1) right-hand sides."""
    rom, rng = modal_rom(256, 1, 0.01, alpha=1e-4, beta=0.5)
    F = complex_normal(rng, (256, 4))
    return Case(rom, np.linspace(1.0, 2000.0, nfreq), F)


UNEVEN_SIZES = (32, 48, 64, 96, 128, 192, 256, 384, 512)


def uneven_batch(nsys=64, seed=2, eta=0.02, scale=1.0):
    """nsys independent ROMs (seed 1000+i) built like single_n60 (undamped
    operators, eta=0.02): n drawn from UNEVEN_SIZES, 256..2048 frequencies over
    1..1500 Hz, 1..8 complex N(0,1) right-hand sides.  scale < 1 shrinks the
    frequency counts (for quick tests)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nsys):
        n = int(rng.choice(UNEVEN_SIZES))
        nf = int(rng.integers(256, 2049))
        nrhs = int(rng.integers(1, 9))
        nf = max(1, int(round(nf * scale)))
        rom, r2 = modal_rom(n, 1000 + i, eta)
        F = complex_normal(r2, (n, nrhs))
        out.append(Case(rom, np.linspace(1.0, 1500.0, nf), F))
    return out


def adjoint_n128(nfreq=2048):
    """n=128 (seed 3), undamped operators, eta=0.02, 2048 frequencies in
    1..1000 Hz, one complex N(0,1) right-hand side and objective vector c."""
    rom, rng = modal_rom(128, 3, 0.02)
    F = complex_normal(rng, (128, 1))
    c = complex_normal(rng, (128,))
    return Case(rom, np.linspace(1.0, 1000.0, nfreq), F, {"c": c})


def dense_n1024(nfreq=512):
    """n=1024 (seed 4), eta=0.03, D = 2e-5 K, 512 frequencies in 1..3000 Hz,
    16 complex N(0,1) right-hand sides."""
    rom, rng = modal_rom(1024, 4, 0.03, alpha=2e-5)
    F = complex_normal(rng, (1024, 16))
    return Case(rom, np.linspace(1.0, 3000.0, nfreq), F)


def adjoint_closed_form(rom, freq, F, c):
    """J = sum_w sum_r |c^H X_r|^2 and its exact gradient w.r.t. the modal
    stiffness scales s_k (dlam_k/ds_k = (1+i eta) kdiag_k) and eta
    (dlam_k/deta = i kdiag_k), from the modal decomposition."""
    w = 2.0 * np.pi * np.asarray(freq, dtype=np.float64)
    lam = ((1.0 + 1j * rom.eta) * rom.kdiag[None, :] + 1j * w[:, None] * rom.ddiag[None, :]
           - (w ** 2)[:, None] * rom.mdiag[None, :])
    cphi = c.conj() @ rom.Phi            # (n,)  c^H phi_k
    phiF = rom.Phi.T @ F                 # (n, nrhs)
    a = cphi[None, :, None] * phiF[None, :, :]        # (1, n, nrhs)
    g = np.sum(a / lam[:, :, None], axis=1)           # (nfreq, nrhs)
    J = float(np.sum(np.abs(g) ** 2))
    dg_ds = -a * ((1.0 + 1j * rom.eta) * rom.kdiag)[None, :, None] / lam[:, :, None] ** 2
    dJ_ds = np.sum(2.0 * np.real(np.conj(g)[:, None, :] * dg_ds), axis=(0, 2))
    dg_de = np.sum(-a * (1j * rom.kdiag)[None, :, None] / lam[:, :, None] ** 2, axis=1)
    dJ_de = float(np.sum(2.0 * np.real(np.conj(g) * dg_de)))
    return J, dJ_ds, dJ_de
