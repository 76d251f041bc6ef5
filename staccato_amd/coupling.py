"""Interface-Jacobian coupled substructure analysis (IJCSA) on the GPU.

Mirrors the simulation-model path of test/IJCSA/src/IJSA.cu, one Newton step
of the interface residual per frequency:

* FRFs of every substructure at its interface, load and internal-output DoFs,
  H_s(w) = C_s A_s(w)^{-1} B_s (IJSA.cu:830-917), from the sweep kernel
  (Handle.back_transform with B_s = [E_iface | E_load], C_s = [E_iface ; E_internal]);
* initial interface solution Y_s = H_ii u_s + H_il f_s (IJSA.cu:920-946),
  residual r = sum_s dRdY_s Y_s + dRdU_s u_s (IJSA.cu:948-968) and interface
  Jacobian J = dRdU + sum_s dRdY_s H_ii (IJSA.cu:523-527, 970-981): batched
  device GEMMs;
* correction delta = J^{-1} r with the pivoted LU kernel
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
    """A simulation model: the id of a ROM already set on the handle
    (rom_set), 0-based local indices of its interface, load and internal
    output DoFs, its initial interface forces u0 and its load values f.
    couple() sets the ROM's input/output maps (rom_set_io)."""

    def __init__(self, rom_id, iface, load=(), internal=(), u0=None, f=None):
        self.rom_id = rom_id
        self.iface = np.asarray(iface, dtype=np.int64)
        self.load = np.asarray(load, dtype=np.int64)
        self.internal = np.asarray(internal, dtype=np.int64)
        ni, nl = self.iface.size, self.load.size
        self.u0 = np.zeros(ni, np.complex128) if u0 is None else np.asarray(u0, np.complex128)
        self.f = np.zeros(nl, np.complex128) if f is None else np.asarray(f, np.complex128)
        if ni < 1 or self.u0.shape != (ni,) or self.f.shape != (nl,):
            raise ValueError("substructure needs >= 1 interface DoF, u0 (ni,) and f (nl,)")


def _io_maps(n, s):
    ni, nl, nn = s.iface.size, s.load.size, s.internal.size
    for idx in (s.iface, s.load, s.internal):
        if idx.size and (idx.min() < 0 or idx.max() >= n):
            raise ValueError("DoF index outside the ROM")
    B = np.zeros((n, ni + nl), np.complex128)
    B[s.iface, np.arange(ni)] = 1.0
    B[s.load, ni + np.arange(nl)] = 1.0
    C = np.zeros((ni + nn, n), np.complex128)
    C[np.arange(ni), s.iface] = 1.0
    C[ni + np.arange(nn), s.internal] = 1.0
    return B, C


def couple(h, subs, dRdY, dRdU, freq, mode="plain", tol=0.0, max_iter=0):
    """Coupled responses at the frequencies freq (Hz).  Returns a dict with
    u (nfreq, Nu) interface forces, Y (nfreq, Nu) interface responses,
    Y_internal [per substructure (nfreq, nn_s)], info_sub [per substructure
    (nfreq,)] and info_J (nfreq,) (0 = ok, k > 0 = zero pivot k)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("couple() runs on the GPU; no device is visible")
    dev = torch.device("cuda", h.device)
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    Nu = int(sum(s.iface.size for s in subs))
    dRdY = np.asarray(dRdY, np.complex128)
    dRdU = np.asarray(dRdU, np.complex128)
    if dRdY.shape != (Nu, Nu) or dRdU.shape != (Nu, Nu):
        raise ValueError(f"dRdY and dRdU must be {Nu} x {Nu}")
    # all device work of the step on one stream of its own, which the library's
    # kernel is enqueued on too (the null stream cannot be named through the ABI)
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        return _couple_on(h, subs, dRdY, dRdU, freq, mode, tol, max_iter, dev, side)


def _couple_on(h, subs, dRdY, dRdU, freq, mode, tol, max_iter, dev, side):
    import torch

    c128 = torch.complex128
    nf = freq.size
    Nu = dRdY.shape[0]
    dY_t = torch.as_tensor(dRdY, device=dev)
    dU_t = torch.as_tensor(dRdU, device=dev)
    J = dU_t.expand(nf, Nu, Nu).clone()
    r = torch.zeros((nf, Nu, 1), dtype=c128, device=dev)
    parts, info_sub = [], []
    off = 0
    for s in subs:
        n = h.n[s.rom_id]
        ni, nl, nn = s.iface.size, s.load.size, s.internal.size
        B, C = _io_maps(n, s)
        h.rom_set_io(s.rom_id, B, C)
        load = np.concatenate([s.u0, s.f])[:, None]
        Y0, H, info, _ = h.back_transform(s.rom_id, freq, load, ni + nl, ni + nn, transfer=True,
                                          mode=mode, tol=tol, max_iter=max_iter)
        info_sub.append(info)
        H_t = torch.as_tensor(np.ascontiguousarray(H), device=dev)  # (nf, ni+nn, ni+nl)
        y0 = torch.as_tensor(np.ascontiguousarray(Y0[:, :ni, :]), device=dev)
        u0 = torch.as_tensor(s.u0[:, None], device=dev)
        f = torch.as_tensor(s.f[:, None], device=dev)
        dYs = dY_t[:, off:off + ni].contiguous()
        dUs = dU_t[:, off:off + ni].contiguous()
        r += torch.matmul(dYs, y0) + torch.matmul(dUs, u0)
        J[:, :, off:off + ni] += torch.matmul(dYs, H_t[:, :ni, :ni].contiguous())
        parts.append((H_t, u0, f, ni, nl, nn, off))
        off += ni
    # delta = J^{-1} r on the handle's kernel, ordered after the GEMMs above
    Jc = J.transpose(1, 2).contiguous()
    rc = r.transpose(1, 2).contiguous()
    delta = torch.empty_like(rc)
    info_J = torch.zeros(nf, dtype=torch.int32, device=dev)
    h.gesv_batched_dev(Jc, rc, delta, info_J, mode=mode, tol=tol, max_iter=max_iter,
                       stream=side.cuda_stream)
    delta = delta.transpose(1, 2)  # (nf, Nu, 1)
    U = torch.empty((nf, Nu, 1), dtype=c128, device=dev)
    Y = torch.empty((nf, Nu, 1), dtype=c128, device=dev)
    Yint = []
    for H_t, u0, f, ni, nl, nn, off in parts:
        u = u0 - delta[:, off:off + ni, :]
        U[:, off:off + ni] = u
        y = torch.matmul(H_t[:, :ni, :ni].contiguous(), u)
        if nl:
            y = y + torch.matmul(H_t[:, :ni, ni:].contiguous(), f)
        Y[:, off:off + ni] = y
        if nn:
            yi = torch.matmul(H_t[:, ni:, :ni].contiguous(), u)
            if nl:
                yi = yi + torch.matmul(H_t[:, ni:, ni:].contiguous(), f)
            Yint.append(yi[:, :, 0].cpu().numpy())
        else:
            Yint.append(np.zeros((nf, 0), np.complex128))
    side.synchronize()
    out = {"u": U[:, :, 0].cpu().numpy(), "Y": Y[:, :, 0].cpu().numpy(), "Y_internal": Yint,
           "info_sub": info_sub, "info_J": info_J.cpu().numpy()}
    return out


__all__ = ["Substructure", "couple", "StcError"]
