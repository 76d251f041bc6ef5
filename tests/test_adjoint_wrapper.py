"""Host logic of the torch.autograd.Function over the adjoint sweep
(staccato_amd/adjoint.py), on the CPU: the GPU handle is replaced by a stand-in
whose adjoint is the oracle's (oracle/stc_oracle.c, pinned in test_oracle.py).

Checked: backward() returns the adjoint gradient scaled by the incoming
gradient; the gradient matches central differences (h = 1e-6) of the wrapper's
own forward to 1e-6 relative on the dominant components; K(s) is formed as
Phi diag(s kdiag) Phi^T; incomplete sweeps raise.
"""
import numpy as np
import pytest
import torch

import oracle
from staccato_amd import synth
from staccato_amd.adjoint import ModalObjective


class OracleHandle:
    """The two calls ModalObjective makes, answered by the CPU oracle."""

    def __init__(self, fail=False):
        self.roms, self.fail, self.calls = {}, fail, 0

    def rom_set(self, rom_id, K, M, D=None, eta=0.0):
        self.roms[rom_id] = (np.array(K), M, D, eta)

    def adjoint(self, rom_id, freq, F, c, Phi, kdiag, mode="plain", tol=0.0, max_iter=0):
        K, M, D, eta = self.roms[rom_id]
        self.calls += 1
        J, ds, de, _ = oracle.adjoint(K, M, D, eta, freq, F, c, Phi, kdiag, nthreads=4)
        info = np.zeros(len(freq), dtype=np.int32)
        if self.fail:
            info[3] = -1
        return J, ds, de, info


def _objective(handle, nfreq=24):
    c = synth.adjoint_n128(nfreq)
    rom = c.rom
    return c, ModalObjective(handle, 7, rom.Phi, rom.kdiag, rom.M, rom.D, c.freq, c.F,
                             c.extra["c"])


def test_backward_is_the_adjoint_gradient():
    h = OracleHandle()
    c, obj = _objective(h)
    rom = c.rom
    s = torch.ones(rom.n, dtype=torch.float64, requires_grad=True)
    eta = torch.tensor(rom.eta, dtype=torch.float64, requires_grad=True)
    J = obj(s, eta)
    (3.0 * J).backward()
    # K(1) is the ROM's own stiffness, so the wrapper reproduces the closed form
    Jc, dsc, dec = synth.adjoint_closed_form(rom, c.freq, c.F, c.extra["c"])
    np.testing.assert_allclose(h.roms[7][0], rom.K, rtol=0, atol=1e-9 * np.abs(rom.K).max())
    assert abs(float(J.detach()) / Jc - 1) < 1e-10
    assert np.abs(s.grad.numpy() - 3.0 * dsc).max() <= 1e-8 * 3.0 * np.abs(dsc).max()
    assert abs(float(eta.grad) / (3.0 * dec) - 1) < 1e-8


def test_gradient_matches_central_differences():
    h = OracleHandle()
    c, obj = _objective(h)
    rom = c.rom
    s0 = torch.ones(rom.n, dtype=torch.float64, requires_grad=True)
    e0 = torch.tensor(rom.eta, dtype=torch.float64, requires_grad=True)
    obj(s0, e0).backward()
    g = s0.grad.numpy()
    step = 1e-6
    # dominant components: FD noise is ~ u J / step, so only gradients within
    # 1/10 of the largest are resolved to 1e-6 relative
    for k in np.argsort(-np.abs(g))[:5]:
        e = torch.zeros(rom.n, dtype=torch.float64)
        e[k] = step
        fd = (float(obj(s0.detach() + e, e0.detach())) - float(obj(s0.detach() - e, e0.detach()))) \
            / (2 * step)
        assert abs(fd / g[k] - 1) < 1e-6, (k, fd, g[k])
    fd = (float(obj(s0.detach(), e0.detach() + step)) - float(obj(s0.detach(), e0.detach() - step))) \
        / (2 * step)
    assert abs(fd / float(e0.grad) - 1) < 1e-6


def test_scaled_stiffness_and_failures():
    h = OracleHandle()
    c, obj = _objective(h, 4)
    rom = c.rom
    s = np.linspace(0.5, 1.5, rom.n)
    K = obj.stiffness(s)
    np.testing.assert_allclose(K @ rom.Phi[:, 5], s[5] * rom.kdiag[5] * rom.Phi[:, 5],
                               rtol=0, atol=1e-9 * rom.kdiag.max())
    with pytest.raises(ValueError):
        obj.evaluate(np.ones(rom.n - 1), rom.eta)
    _, bad = _objective(OracleHandle(fail=True), 4)
    with pytest.raises(RuntimeError, match="not converged"):
        bad(torch.ones(rom.n, dtype=torch.float64), torch.tensor(rom.eta, dtype=torch.float64))
