"""Regression tests of the launch plumbing and of failure handling on the GPU.

- a hipGraph plan owns its factor slots and queue head: a larger sweep on the
  same handle between replays (which grows the handle's slots) leaves the
  replays correct, and the plan's ROMs cannot be replaced or cleared while it
  lives;
- mixed-size batches whose small systems share the large system's rows per
  lane get a panel region that holds the register panel's multiplier buffers;
- a matrix full of NaN pivots like LAPACK izamax (first row) without
  disturbing the other matrices of the batch;
- launches of one handle on two streams are ordered;
- the tensor checks of the device wrappers;
- the two-level update (the default above n = 256) with partial steps and
  bands, many right-hand sides and the adjoint.
Checked against the closed-form modal solution (staccato_amd/synth.py) and
numpy's LAPACK solve.
"""
import ctypes

import numpy as np
import pytest

from staccato_amd import Plan, StcError, synth
from staccato_amd._lib import STC_EBUSY, SysDesc

pytestmark = pytest.mark.gpu


def _closed_ok(X, c, gate):
    Xc = synth.closed_form(c.rom, c.freq, c.F)
    kappa = synth.cond2(c.rom, c.freq)
    fe = np.abs(X - Xc).max(axis=(1, 2)) / np.abs(Xc).max(axis=(1, 2))
    return (fe <= gate * kappa + 1e-13).all(), float((fe / kappa).max())


def _device_batch(handle, cases, base):
    import torch
    dev = torch.device("cuda:0")
    keep, descs = [], (SysDesc * len(cases))()
    for i, c in enumerate(cases):
        handle.rom_set(base + i, c.rom.K, c.rom.M, c.rom.D, c.rom.eta)
        n, nf, nr = c.rom.n, c.freq.size, c.F.shape[1]
        t = (torch.tensor(c.freq, dtype=torch.float64, device=dev),
             torch.tensor(np.ascontiguousarray(c.F.T), dtype=torch.complex128, device=dev),
             torch.zeros((nf, nr, n), dtype=torch.complex128, device=dev),
             torch.zeros(nf, dtype=torch.int32, device=dev),
             torch.zeros((nf, nr), dtype=torch.float64, device=dev))
        keep.append(t)
        d = descs[i]
        d.rom_id, d.nfreq, d.freq_hz, d.nrhs = base + i, nf, t[0].data_ptr(), nr
        d.F, d.ldf, d.X, d.ldx = t[1].data_ptr(), n, t[2].data_ptr(), n
        d.info, d.berr = t[3].data_ptr(), t[4].data_ptr()
    return keep, descs


def test_plan_survives_larger_sweep_and_keeps_its_roms(handle):
    import torch
    cases = [synth.Case(*synth.modal_rom(n, 300 + n, 0.02)[:1], np.linspace(5.0, 900.0, 9),
                        synth.complex_normal(np.random.default_rng(n), (n, 2)))
             for n in (96, 200, 256)]
    keep, descs = _device_batch(handle, cases, 400)
    plan = Plan(handle, descs, mode="refined", tol=1e-10)
    try:
        big = synth.dense_n1024(4)
        for rep in range(2):
            for t in keep:
                t[2].zero_()
            torch.cuda.synchronize()
            plan.execute()
            torch.cuda.synchronize()
            for t, c in zip(keep, cases):
                assert (t[3].cpu().numpy() == 0).all()
                ok, worst = _closed_ok(t[2].cpu().numpy().transpose(0, 2, 1), c, 4e-10)
                assert ok, (rep, c.rom.n, worst)
            # a larger system on the same handle grows (and frees) the handle's
            # factor slots between the replays
            handle.rom_set(499, big.rom.K, big.rom.M, big.rom.D, big.rom.eta)
            X, info, be, _ = handle.sweep(499, big.freq, big.F[:, :2], mode="plain")
            assert (info == 0).all() and be.max() <= 1e-12
            handle.rom_clear(499)
        # the plan's ROMs are pinned while it lives
        c0 = cases[0]
        with pytest.raises(StcError) as e:
            handle.rom_set(400, c0.rom.K, c0.rom.M, c0.rom.D, c0.rom.eta)
        assert e.value.code == STC_EBUSY
        with pytest.raises(StcError) as e:
            handle.rom_clear(401)
        assert e.value.code == STC_EBUSY
        assert handle.lib.stc_destroy(handle.h) == STC_EBUSY
    finally:
        plan.close()
    c0 = cases[0]
    handle.rom_set(400, c0.rom.K, c0.rom.M, c0.rom.D, c0.rom.eta)   # free again


@pytest.mark.parametrize("mode,sizes", [("plain", (110, 300)), ("refined", (200, 1024)),
                                        ("plain", (70, 520)), ("refined", (40, 600))])
def test_mixed_batch_small_and_large(handle, mode, sizes):
    """A small blocked system shares the rows per lane of the large one; its
    panel region must still hold the register panel's multiplier buffers."""
    rng = np.random.default_rng(sum(sizes))
    systems, cases = [], []
    for i, n in enumerate(sizes):
        rom, r2 = synth.modal_rom(n, 700 + n, 0.02, alpha=1e-4, beta=0.2)
        c = synth.Case(rom, np.linspace(3.0, 1400.0, 6 if n > 512 else 16),
                       synth.complex_normal(r2, (n, 1 + i)))
        handle.rom_set(450 + i, rom.K, rom.M, rom.D, rom.eta)
        systems.append((450 + i, c.freq, c.F))
        cases.append(c)
    tol = 1e-10 if mode == "refined" else 1e-12
    outs, st = handle.sweep_batch(systems, mode=mode, tol=tol)
    for (X, info, be), c in zip(outs, cases):
        assert (info == 0).all(), (c.rom.n, info)
        assert be.max() <= tol, (c.rom.n, be.max())
        ok, worst = _closed_ok(X, c, 4 * tol if mode == "refined" else 1e-13)
        assert ok, (c.rom.n, worst)


@pytest.mark.parametrize("mode", ["plain", "refined"])
@pytest.mark.parametrize("n", [60, 200])
def test_nan_matrix_does_not_disturb_batch(handle, mode, n):
    """izamax keeps the first row when the column holds NaN there; the NaN
    matrix yields NaN (plain: info 0, as zgetrf; refined: info -1, its backward
    error is NaN), every other matrix of the batch solves as numpy (plain: to
    1e-9; refined: to its backward-error contract)."""
    rng = np.random.default_rng(n)
    nb = 6
    A = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n)) + 4 * np.eye(n)
    A[2] = np.nan
    A[4, :, 0] = np.nan         # NaN in the first pivot column only
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    X, info, be, _ = handle.gesv_batched(A, B, mode=mode, tol=1e-10 if mode == "refined" else 0.0)
    for j in range(nb):
        if j in (2, 4):
            assert np.isnan(X[j]).any()
            assert info[j] == (0 if mode == "plain" else -1), (j, info[j])
            continue
        assert info[j] == 0
        Xr = np.linalg.solve(A[j], B)
        if mode == "plain":
            assert np.abs(X[j] - Xr).max() <= 1e-9 * np.abs(Xr).max(), j
            continue
        # refined: the contract is the normwise backward error (FP64 residual,
        # inf-norms with |z|); the forward error then stays within
        # cond_inf(A) times it (cond_inf is about 5e3 to 9e3 here)
        R = B - A[j] @ X[j]
        anorm = np.abs(A[j]).sum(axis=1).max()
        berr = (np.abs(R).max(axis=0) / (anorm * np.abs(X[j]).max(axis=0) + np.abs(B).max(axis=0))).max()
        assert berr <= 2e-10, (j, berr)
        cond = float(np.abs(np.linalg.cond(A[j], np.inf)))
        assert np.abs(X[j] - Xr).max() <= 4 * cond * 1e-10 * np.abs(Xr).max(), j


def test_launches_on_two_streams_are_ordered(handle):
    """Two device sweeps of different sizes on two streams, back to back: they
    share the handle's slots and queue head, so the second waits for the first."""
    import torch
    dev = torch.device("cuda:0")
    s1, s2 = torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)
    res = []
    for rid, n, nf in ((460, 256, 64), (461, 300, 48)):
        rom, rng = synth.modal_rom(n, 900 + n, 0.01, alpha=1e-4, beta=0.5)
        c = synth.Case(rom, np.linspace(2.0, 1800.0, nf), synth.complex_normal(rng, (n, 3)))
        handle.rom_set(rid, rom.K, rom.M, rom.D, rom.eta)
        fr = torch.tensor(c.freq, dtype=torch.float64, device=dev)
        F = torch.tensor(np.ascontiguousarray(c.F.T), dtype=torch.complex128, device=dev)
        X = torch.zeros((nf, 3, n), dtype=torch.complex128, device=dev)
        info = torch.full((nf,), -7, dtype=torch.int32, device=dev)
        res.append((rid, c, fr, F, X, info))
    torch.cuda.synchronize()
    for rep in range(3):
        for (rid, c, fr, F, X, info), st in zip(res, (s1, s2)):
            handle.sweep_dev(rid, fr, F, X, info, None, mode="refined", tol=1e-10,
                             stream=st.cuda_stream)
    torch.cuda.synchronize()
    for rid, c, fr, F, X, info in res:
        assert (info.cpu().numpy() == 0).all()
        ok, worst = _closed_ok(X.cpu().numpy().transpose(0, 2, 1), c, 4e-10)
        assert ok, (rid, worst)


def test_rom_set_rejects_non_finite(handle):
    rom, _ = synth.modal_rom(8, 5, 0.02)
    K = rom.K.copy()
    K[3, 4] = np.inf
    with pytest.raises(StcError):
        handle.rom_set(470, K, rom.M, None, 0.02)
    M = rom.M.astype(np.complex128)
    M[1, 1] = complex(0.0, np.nan)
    with pytest.raises(StcError):
        handle.rom_set(470, rom.K, M, None, 0.02)


def test_rom_set_new_order_drops_projections(handle):
    rom, rng = synth.modal_rom(12, 6, 0.02)
    handle.rom_set(471, rom.K, rom.M, None, 0.02)
    handle.rom_set_io(471, synth.complex_normal(rng, (12, 2)), synth.complex_normal(rng, (3, 12)))
    rom2, _ = synth.modal_rom(20, 7, 0.02)
    handle.rom_set(471, rom2.K, rom2.M, None, 0.02)
    # B_R / C_R belonged to n = 12: the back-transformation must refuse
    with pytest.raises(StcError):
        handle.back_transform(471, np.array([10.0]), np.ones((2, 1)), 2, 3)


def test_device_wrappers_check_tensors(handle):
    import torch
    dev = torch.device("cuda:0")
    rom, rng = synth.modal_rom(16, 8, 0.02)
    handle.rom_set(472, rom.K, rom.M, None, 0.02)
    fr = torch.tensor([10.0, 20.0], dtype=torch.float64, device=dev)
    F = torch.ones((2, 16), dtype=torch.complex128, device=dev)
    X = torch.empty((2, 2, 16), dtype=torch.complex128, device=dev)
    info = torch.empty(2, dtype=torch.int32, device=dev)
    handle.sweep_dev(472, fr, F, X, info, mode="plain", stream=torch.cuda.Stream().cuda_stream)
    torch.cuda.synchronize()
    bad = [(fr.float(), F, X, info), (fr, F.to(torch.complex64), X, info),
           (fr, F, X[:, :1], info), (fr, F, X.transpose(1, 2), info), (fr, F, X, info.long()),
           (fr, F.cpu(), X, info), (fr, torch.ones((2, 15), dtype=torch.complex128, device=dev),
                                    X, info)]
    for args in bad:
        with pytest.raises((ValueError, TypeError)):
            handle.sweep_dev(472, *args, mode="plain")
    A = torch.eye(8, dtype=torch.complex128, device=dev).expand(3, 8, 8).contiguous()
    B = torch.ones((3, 1, 8), dtype=torch.complex128, device=dev)
    Xb = torch.empty((3, 1, 8), dtype=torch.complex128, device=dev)
    ib = torch.empty(3, dtype=torch.int32, device=dev)
    for args in [(A.to(torch.complex64), B, Xb, ib), (A, B, Xb[:2], ib), (A, B[:, :, :7], Xb, ib),
                 (A.transpose(1, 2), B, Xb, ib)]:
        with pytest.raises((ValueError, TypeError)):
            handle.gesv_batched_dev(*args)


# ---- the two-level update (the default above n = 256) on its edge cases ----------------

@pytest.mark.parametrize("mode", ["plain", "refined"])
def test_two_level_partial_steps_many_rhs_and_mixed_batch(handle, mode):
    """n = 300 (last step and last band partial) and n = 512 with 16
    right-hand sides (complex64: two workgroups per CU, groups of four with
    the FP64 residuals in the factor slot), next to small systems in the same
    launch."""
    rng = np.random.default_rng(17)
    systems, cases = [], []
    for i, (n, nrhs) in enumerate(((512, 16), (300, 3), (64, 2), (130, 5))):
        rom, r2 = synth.modal_rom(n, 900 + n, 0.02, alpha=1e-4, beta=0.3)
        c = synth.Case(rom, np.linspace(2.0, 1600.0, 11), synth.complex_normal(r2, (n, nrhs)))
        handle.rom_set(470 + i, rom.K, rom.M, rom.D, rom.eta)
        systems.append((470 + i, c.freq, c.F))
        cases.append(c)
    tol = 1e-10 if mode == "refined" else 1e-12
    outs, st = handle.sweep_batch(systems, mode=mode, tol=tol)
    for (X, info, be), c in zip(outs, cases):
        assert (info == 0).all(), (c.rom.n, info)
        assert be.max() <= tol, (c.rom.n, be.max())
        ok, worst = _closed_ok(X, c, 4 * tol if mode == "refined" else 1e-13)
        assert ok, (c.rom.n, worst)
    # one system alone (its own launch) agrees with the batch to the tolerance
    X1, info1, be1, _ = handle.sweep(471, cases[1].freq, cases[1].F, mode=mode, tol=tol)
    assert (info1 == 0).all()
    Xb = outs[1][0]
    assert np.abs(X1 - Xb).max() <= 1e3 * tol * np.abs(Xb).max()


@pytest.mark.parametrize("mode,tol", [("plain", 0.0), ("refined", 1e-12)])
def test_adjoint_on_the_two_level_path(handle, mode, tol):
    """The fused adjoint of an n = 300 ROM (two-level factorisation, residuals
    in LDS) against the closed-form modal gradient."""
    rom, rng = synth.modal_rom(300, 31, 0.02)
    freq = np.linspace(1.0, 900.0, 64)
    F = synth.complex_normal(rng, (300, 1))
    cv = synth.complex_normal(rng, (300,))
    handle.rom_set(480, rom.K, rom.M, rom.D, rom.eta)
    J, ds, de, info = handle.adjoint(480, freq, F, cv, rom.Phi, rom.kdiag, mode=mode, tol=tol)
    assert (info == 0).all()
    Jc, dsc, dec = synth.adjoint_closed_form(rom, freq, F, cv)
    gate = 1e-8 if mode == "plain" else 1e-6
    assert abs(J / Jc - 1) < gate
    assert np.abs(ds - dsc).max() <= gate * np.abs(dsc).max()
    assert abs(de / dec - 1) < gate
