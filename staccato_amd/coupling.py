"""Interface-Jacobian coupled substructure analysis (IJCSA) on the GPU.

Mirrors the simulation-model path of test/IJCSA/src/IJSA.cu, one Newton step
of the interface residual per frequency, with every array device resident:

* FRFs of every substructure at its interface, load and internal-output DoFs,
  H_s(w) = C_s A_s(w)^{-1} B_s (IJSA.cu:830-917): one sweep of the ROM
  (Handle.sweep_dev) with right-hand sides [B_iface | B_load], then C_iface X
  and C_internal X as batched device GEMMs;
* initial interface solution Y_s = H_ii u_s + H_il f_s (IJSA.cu:920-946),
  residual r = sum_s dRdY_s Y_s + dRdU_s u_s (IJSA.cu:948-968) and interface
  Jacobian J = dRdU + sum_s dRdY_s H_ii (IJSA.cu:523-527, 970-981): batched
  device GEMMs;
* correction delta = J^{-1} r with the pivoted LU kernel in FP64
  (Handle.gesv_batched_dev; IJSA.cu:1199-1208 factors J unpivoted);
* update u_s -= delta_s (IJSA.cu:1349-1357) and the responses
  Y_s = H_ii u_s + H_il f_s, Y_int,s = H_ni u_s + H_nl f_s (IJSA.cu:1359-1411).

dRdY and dRdU are Nu x Nu, Nu = the sum of the interface sizes, with the
column blocks of the substructures in order (IJSA.cu:537-553).  Experimental
models (measured FRFs, IJSA.cu:1011-1187) enter the residual and Jacobian the
same way and are not read here.
"""
import numpy as np

from .sweep import StcError


class Substructure:
    """A simulation model: the id of a ROM already set on the handle (rom_set),
    its input matrices B_iface (n x ni) and B_load (n x nl), its output matrices
    C_iface (ni x n) and C_internal (nn x n) (the interface/load columns of the
    ROM's B_R and the interface/internal rows of its C_R, IJSA.cu:869-917), its
    initial interface forces u0 (ni) and its load values f (nl)."""

    def __init__(self, rom_id, B_iface, C_iface, B_load=None, C_internal=None, u0=None, f=None):
        self.rom_id = rom_id
        self.B_iface = np.asarray(B_iface, np.complex128)
        n, ni = self.B_iface.shape
        self.C_iface = np.asarray(C_iface, np.complex128)
        self.B_load = (np.zeros((n, 0), np.complex128) if B_load is None
                       else np.asarray(B_load, np.complex128).reshape(n, -1))
        self.C_internal = (np.zeros((0, n), np.complex128) if C_internal is None
                           else np.asarray(C_internal, np.complex128).reshape(-1, n))
        nl = self.B_load.shape[1]
        self.u0 = np.zeros(ni, np.complex128) if u0 is None else np.asarray(u0, np.complex128)
        self.f = np.zeros(nl, np.complex128) if f is None else np.asarray(f, np.complex128)
        if ni < 1 or self.C_iface.shape != (ni, n) or self.u0.shape != (ni,) \
                or self.f.shape != (nl,):
            raise ValueError("substructure needs B_iface (n x ni), C_iface (ni x n), ni >= 1, "
                             "u0 (ni,) and f (nl,)")

    @classmethod
    def from_dofs(cls, rom_id, n, iface, load=(), internal=(), u0=None, f=None):
        """Interface, load and output DoFs given as 0-based indices of the ROM."""
        E = np.eye(n, dtype=np.complex128)
        for idx in (iface, load, internal):
            if len(idx) and (min(idx) < 0 or max(idx) >= n):
                raise ValueError("DoF index outside the ROM")
        return cls(rom_id, E[:, list(iface)], E[list(iface), :], E[:, list(load)],
                   E[list(internal), :], u0, f)

    @property
    def n(self):
        return self.B_iface.shape[0]


class MeasuredSubstructure:
    """An experimental model (IJSA.cu:1011-1187): its FRFs at the coupling
    frequencies are given instead of a ROM, H_ii (nf, ni, ni) interface to
    interface, H_il (nf, ni, nl) load to interface, H_ni (nf, nn, ni) and
    H_nl (nf, nn, nl) to its internal outputs; u0 (ni) and f (nl) as for
    Substructure."""

    def __init__(self, H_ii, H_il=None, H_ni=None, H_nl=None, u0=None, f=None):
        self.H_ii = np.asarray(H_ii, np.complex128)
        nf, ni = self.H_ii.shape[0], self.H_ii.shape[1]
        self.H_il = (np.zeros((nf, ni, 0), np.complex128) if H_il is None
                     else np.asarray(H_il, np.complex128))
        nl = self.H_il.shape[2]
        self.H_ni = (np.zeros((nf, 0, ni), np.complex128) if H_ni is None
                     else np.asarray(H_ni, np.complex128))
        nn = self.H_ni.shape[1]
        self.H_nl = (np.zeros((nf, nn, nl), np.complex128) if H_nl is None
                     else np.asarray(H_nl, np.complex128))
        self.u0 = np.zeros(ni, np.complex128) if u0 is None else np.asarray(u0, np.complex128)
        self.f = np.zeros(nl, np.complex128) if f is None else np.asarray(f, np.complex128)
        if self.H_ii.shape != (nf, ni, ni) or self.H_il.shape != (nf, ni, nl) or \
                self.H_ni.shape != (nf, nn, ni) or self.H_nl.shape != (nf, nn, nl) or \
                self.u0.shape != (ni,) or self.f.shape != (nl,):
            raise ValueError("measured FRF blocks, u0 and f do not agree in shape")


def couple(h, subs, dRdY, dRdU, freq, mode="plain", tol=0.0, max_iter=0):
    """Coupled responses at the frequencies freq (Hz); mode, tol and max_iter
    apply to the substructure sweeps (the Jacobians are solved in FP64).
    Returns a dict with
    u (nfreq, Nu) interface forces, Y (nfreq, Nu) interface responses,
    Y_internal [per substructure (nfreq, nn_s)], info_sub [per substructure
    (nfreq,): 0 = ok, k > 0 = zero pivot k, -1 = not refined to tol] and
    info_J (nfreq,) (0 = ok, k > 0 = zero pivot k)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("couple() runs on the GPU; no device is visible")
    dev = torch.device("cuda", h.device)
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    Nu = int(sum(s.u0.size for s in subs))
    dRdY = np.asarray(dRdY, np.complex128)
    dRdU = np.asarray(dRdU, np.complex128)
    if dRdY.shape != (Nu, Nu) or dRdU.shape != (Nu, Nu):
        raise ValueError(f"dRdY and dRdU must be {Nu} x {Nu}")
    for s in subs:
        if isinstance(s, MeasuredSubstructure):
            if s.H_ii.shape[0] != freq.size:
                raise ValueError("measured FRFs must be given at every coupling frequency")
        elif h.n.get(s.rom_id) != s.n:
            raise ValueError(f"ROM {s.rom_id} is not set with n = {s.n}")
    # all device work of the step on one stream of its own, which the library's
    # kernels are enqueued on too (the null stream cannot be named through the ABI)
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        return _couple_on(h, subs, dRdY, dRdU, freq, mode, tol, max_iter, dev, side)


def _couple_on(h, subs, dRdY, dRdU, freq, mode, tol, max_iter, dev, side):
    import torch

    c128 = torch.complex128
    nf = freq.size
    Nu = dRdY.shape[0]
    freq_t = torch.as_tensor(freq, device=dev)
    dY_t = torch.as_tensor(dRdY, device=dev)
    dU_t = torch.as_tensor(dRdU, device=dev)
    J = dU_t.expand(nf, Nu, Nu).clone()
    r = torch.zeros((nf, Nu, 1), dtype=c128, device=dev)
    parts, infos = [], []
    off = 0
    for s in subs:
        ni, nl = s.u0.size, s.f.size
        u0 = torch.as_tensor(s.u0[:, None], device=dev)
        f = torch.as_tensor(s.f[:, None], device=dev)
        # Hi = [H_ii | H_il] (nf, ni, ni+nl), Hn = [H_ni | H_nl] (nf, nn, ni+nl) or None
        if isinstance(s, MeasuredSubstructure):
            Hi = torch.cat([torch.as_tensor(s.H_ii, device=dev),
                            torch.as_tensor(s.H_il, device=dev)], dim=2)
            Hn = None
            if s.H_ni.shape[1]:
                Hn = torch.cat([torch.as_tensor(s.H_ni, device=dev),
                                torch.as_tensor(s.H_nl, device=dev)], dim=2)
            infos.append(torch.zeros(nf, dtype=torch.int32, device=dev))
        else:
            Hi, Hn, info_t = _simulated_frfs(h, s, freq_t, nf, mode, tol, max_iter, dev, side)
            infos.append(info_t)
        uf0 = torch.cat([u0, f], dim=0)
        y0 = torch.matmul(Hi, uf0)                          # H_ii u0 + H_il f (IJSA.cu:920-946)
        dYs = dY_t[:, off:off + ni].contiguous()
        dUs = dU_t[:, off:off + ni].contiguous()
        r += torch.matmul(dYs, y0) + torch.matmul(dUs, u0)
        J[:, :, off:off + ni] += torch.matmul(dYs, Hi[:, :, :ni].contiguous())
        parts.append((Hi, Hn, u0, f, ni, nl, off))
        off += ni
    return _correct(h, J, r, parts, infos, nf, Nu, dev, side)


def _simulated_frfs(h, s, freq_t, nf, mode, tol, max_iter, dev, side):
    """(C_iface X, C_internal X or None, info) for X = A^{-1} [B_iface | B_load]
    at every frequency: one sweep of the ROM, then device GEMMs."""
    import torch

    # right-hand sides: the columns of B_iface and B_load (column-major n x ncol,
    # stored as (ncol, n))
    rhs = np.hstack([s.B_iface, s.B_load])
    F_t = torch.as_tensor(np.ascontiguousarray(rhs.T), device=dev)
    X_t = torch.empty((nf, rhs.shape[1], s.n), dtype=torch.complex128, device=dev)
    info_t = torch.empty(nf, dtype=torch.int32, device=dev)
    h.sweep_dev(s.rom_id, freq_t, F_t, X_t, info_t, None, mode=mode, tol=tol,
                max_iter=max_iter, stream=side.cuda_stream)
    # H = C X as (X^T C^T)^T: both operands of the batched product contiguous
    CiT = torch.as_tensor(np.ascontiguousarray(s.C_iface.T), device=dev)
    Hi = torch.matmul(X_t, CiT).transpose(1, 2).contiguous()    # (nf, ni, ncol)
    Hn = None
    if s.C_internal.shape[0]:
        CnT = torch.as_tensor(np.ascontiguousarray(s.C_internal.T), device=dev)
        Hn = torch.matmul(X_t, CnT).transpose(1, 2).contiguous()    # (nf, nn, ncol)
    return Hi, Hn, info_t


def _correct(h, J, r, parts, infos, nf, Nu, dev, side):
    import torch

    c128 = torch.complex128
    # delta = J^{-1} r by the LU kernel, on the same stream as the GEMMs above
    Jc = J.transpose(1, 2).contiguous()
    rc = r.transpose(1, 2).contiguous()
    delta = torch.empty_like(rc)
    info_J = torch.empty(nf, dtype=torch.int32, device=dev)
    # the interface Jacobians are solved in FP64 in either mode, as IJSA.cu does
    # (cuDoubleComplex): their rows mix compliances (~1/k) with unit equilibrium
    # rows, so a normwise complex64 tolerance would not hold compatibility
    h.gesv_batched_dev(Jc, rc, delta, info_J, mode="plain", stream=side.cuda_stream)
    delta = delta.transpose(1, 2)                           # (nf, Nu, 1)
    U = torch.empty((nf, Nu, 1), dtype=c128, device=dev)
    Y = torch.empty((nf, Nu, 1), dtype=c128, device=dev)
    Yint = []
    for Hi, Hn, u0, f, ni, nl, off in parts:
        u = u0 - delta[:, off:off + ni, :]
        U[:, off:off + ni] = u
        # [H_ii | H_il] [u; f] and [H_ni | H_nl] [u; f]
        uf = torch.cat([u, f.expand(nf, nl, 1)], dim=1)
        Y[:, off:off + ni] = torch.matmul(Hi, uf)
        if Hn is not None:
            Yint.append(torch.matmul(Hn, uf)[:, :, 0])
        else:
            Yint.append(torch.zeros((nf, 0), dtype=c128, device=dev))
    side.synchronize()
    return {"u": U[:, :, 0].cpu().numpy(), "Y": Y[:, :, 0].cpu().numpy(),
            "Y_internal": [y.cpu().numpy() for y in Yint],
            "info_sub": [i.cpu().numpy() for i in infos], "info_J": info_J.cpu().numpy()}


__all__ = ["Substructure", "MeasuredSubstructure", "couple", "StcError"]
