"""torch.autograd.Function over the fused adjoint sweep (stc_adjoint), for
gradient checks against central finite differences and the FP64 gradient.

The objective is the one stc_adjoint evaluates,

    J(s, eta) = sum_w sum_r |c^H X_r(w)|^2,   A(w; s, eta) X_r(w) = F_r,
    A(w; s, eta) = (1 + i eta) K(s) + i w D - w^2 M,
    K(s) = Phi diag(s_k kdiag_k) Phi^T,

with per-mode stiffness scales s (length p = Phi.shape[1]) and the loss factor
eta as the differentiable inputs.  forward() forms K(s) on the host, loads it as
the ROM `rom_id` of the handle and runs one adjoint sweep, which returns J
together with dJ/ds and dJ/deta from the same LU factors; backward() scales
those by the incoming gradient.  The reference has no adjoint.

Accuracy: in refined mode (complex64 factors, FP64 refinement) both solves are
refined to the given tol; at tol = 1e-12 the gradient of the n=128 workload is
within 1e-7 (relative, on its dominant components) of the FP64 gradient, at
tol = 1e-10 only within ~1e-5.
"""
import numpy as np
import torch


class _AdjointSweep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s, eta, objective):
        J, ds, de = objective.evaluate(s.detach().cpu().numpy(), float(eta.detach()))
        ctx.save_for_backward(torch.as_tensor(ds, dtype=s.dtype, device=s.device),
                              torch.as_tensor(de, dtype=eta.dtype, device=eta.device))
        return torch.as_tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        ds, de = ctx.saved_tensors
        return gJ.to(ds) * ds, gJ.to(de) * de, None


class ModalObjective:
    """J(s, eta) above for one ROM, frequency grid, load F (n x nrhs) and
    objective vector c (n).  Call it with float64 tensors s (p) and eta (0-d);
    the result is a 0-d float64 tensor that back-propagates into both."""

    def __init__(self, handle, rom_id, Phi, kdiag, M, D, freq, F, c, mode="plain", tol=0.0,
                 max_iter=0):
        self.handle, self.rom_id = handle, rom_id
        self.Phi = np.asarray(Phi, dtype=np.float64)
        self.kdiag = np.asarray(kdiag, dtype=np.float64)
        n, p = self.Phi.shape
        if self.kdiag.shape != (p,):
            raise ValueError(f"kdiag must have {p} entries")
        self.M = np.asarray(M, dtype=np.float64)
        self.D = None if D is None else np.asarray(D, dtype=np.float64)
        if self.M.shape != (n, n) or (self.D is not None and self.D.shape != (n, n)):
            raise ValueError(f"M and D must be {n} x {n}")
        self.freq = np.asarray(freq, dtype=np.float64)
        self.F = np.asarray(F, dtype=np.complex128).reshape(n, -1)
        self.c = np.asarray(c, dtype=np.complex128)
        if self.c.shape != (n,):
            raise ValueError(f"c must have {n} entries")
        self.mode, self.tol, self.max_iter = mode, tol, max_iter

    def stiffness(self, s):
        K = (self.Phi * (np.asarray(s, dtype=np.float64) * self.kdiag)) @ self.Phi.T
        return 0.5 * (K + K.T)

    def evaluate(self, s, eta):
        """(J, dJ/ds, dJ/deta) in FP64; raises when a point is singular or a
        refined solve missed tol (the sums would then be incomplete)."""
        if np.shape(s) != self.kdiag.shape:
            raise ValueError(f"s must have {self.kdiag.size} entries")
        h = self.handle
        h.rom_set(self.rom_id, self.stiffness(s), self.M, self.D, eta)
        J, ds, de, info = h.adjoint(self.rom_id, self.freq, self.F, self.c, self.Phi, self.kdiag,
                                    mode=self.mode, tol=self.tol, max_iter=self.max_iter)
        bad = np.flatnonzero(np.asarray(info) != 0)
        if bad.size:
            raise RuntimeError(f"adjoint: {bad.size} frequencies singular or not converged "
                               f"(first: index {bad[0]}, info {int(info[bad[0]])})")
        return J, np.asarray(ds, dtype=np.float64), float(de)

    def __call__(self, s, eta):
        return _AdjointSweep.apply(s, eta, self)
