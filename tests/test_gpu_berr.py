"""The backward error behind the gate, recomputed from the GPU's own X.

The sweep reports berr = ||F - A X||_inf / (||A||_inf ||X||_inf + ||F||_inf)
per point and column, computed by the kernel in FP64.  Here the GPU's X is
judged independently: the exact backward error (tests/berr_check.py, long
double, A(w) from the FP64 operators) must meet the gate (1e-12 plain,
1e-10 refined) and must equal the kernel's reported value to within the
rounding an FP64 evaluation can incur.  The refined mode with a loose tol
(one complex64 solve, no refinement, berr ~ 1e-7) makes that comparison
sharp: there, checks with a planted mistake (1-norm of the vectors,
|re| + |im| instead of the modulus, the ||F|| term dropped) disagree with the
kernel.  Sizes: the n=60, n=256 and n=1024 workloads (sampled frequencies)
and a mixed batch of the uneven workload's sizes.
"""
import numpy as np
import pytest

from berr_check import abs_terms, assemble_ld, berr, berr_bound
from staccato_amd import synth

pytestmark = pytest.mark.gpu


def _cases():
    c0 = synth.single_n60()
    c1 = synth.uniform_n256()
    c4 = synth.dense_n1024()
    return [("n60", c0, c0.freq[::67][:15]), ("n256", c1, c1.freq[::257][:16]),
            ("n1024", c4, c4.freq[::170][:3])]


def _check_points(rom, freq, F, X, be, gate, rel):
    """Every point: exact berr <= gate and |kernel berr - exact| within the
    FP64 evaluation bound (+ rel of the value)."""
    worst = 0.0
    for j, f in enumerate(freq):
        A = assemble_ld(rom.K, rom.M, rom.D, rom.eta, f)
        ex = berr(A, X[j], F)
        assert (ex <= gate).all(), (f, ex.max())
        T = abs_terms(rom.K, rom.M, rom.D, rom.eta, f)
        bd = berr_bound(T, X[j], F, float(np.abs(A).sum(axis=1).max()))
        dev = np.abs(be[j] - ex)
        assert (dev <= bd + rel * ex).all(), (f, be[j], ex, bd)
        worst = max(worst, float((dev / np.maximum(bd + rel * ex, 1e-300)).max()))
    return worst


@pytest.mark.parametrize("mode,tol", [("plain", 1e-12), ("refined", 1e-10)])
def test_gate_holds_on_gpu_solution(handle, mode, tol):
    for name, c, freq in _cases():
        handle.rom_set(60, c.rom.K, c.rom.M, c.rom.D, c.rom.eta)
        X, info, be, st = handle.sweep(60, freq, c.F, mode=mode, tol=tol)
        assert (info == 0).all(), (name, info)
        _check_points(c.rom, freq, c.F, X, be, tol, 1e-6)


def test_reported_berr_matches_exact_and_planted_checks_fail(handle):
    """Unrefined complex64 solutions (tol = 1: the first solve is accepted):
    berr ~ 1e-7, so the kernel's value must agree with the exact one to
    1e-6 relative, and a checker with a planted mistake must not."""
    n_bad = {"one": 0, "cabs1": 0, "drop_f": 0}
    n_pts = 0
    for name, c, freq in _cases():
        handle.rom_set(61, c.rom.K, c.rom.M, c.rom.D, c.rom.eta)
        X, info, be, st = handle.sweep(61, freq, c.F, mode="refined", tol=1.0)
        assert st.max_refine_steps == 0
        assert be.min() > 1e-12, (name, be.min())      # the comparison is sharp
        _check_points(c.rom, freq, c.F, X, be, 1e-3, 1e-6)
        for j, f in enumerate(freq):
            A = assemble_ld(c.rom.K, c.rom.M, c.rom.D, c.rom.eta, f)
            ex = berr(A, X[j], c.F)
            for plant, kw in (("one", {"vec_norm": "one"}), ("cabs1", {"elem": "cabs1"}),
                              ("drop_f", {"drop_f": True})):
                bad = berr(A, X[j], c.F, **kw)
                if (np.abs(bad - be[j]) > 1e-6 * ex).any():
                    n_bad[plant] += 1
            n_pts += 1
    # each planted mistake is caught at (nearly) every point
    for plant, k in n_bad.items():
        assert k >= 0.9 * n_pts, (plant, k, n_pts)


def test_uneven_batch_berr(handle):
    cases = synth.uneven_batch(nsys=64, scale=0.004)
    # one system of each size
    pick, seen = [], set()
    for i, c in enumerate(cases):
        if c.rom.n not in seen:
            seen.add(c.rom.n)
            pick.append((i, c))
    systems = []
    for i, c in pick:
        handle.rom_set(600 + i, c.rom.K, c.rom.M, c.rom.D, c.rom.eta)
        systems.append((600 + i, c.freq[:3], c.F))
    for mode, tol in (("refined", 1e-10), ("plain", 1e-12)):
        outs, st = handle.sweep_batch(systems, mode=mode, tol=tol)
        for (X, info, be), (i, c) in zip(outs, pick):
            assert (info == 0).all()
            _check_points(c.rom, c.freq[:3], c.F, X, be, tol, 1e-6)
