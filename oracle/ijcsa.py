"""CPU restatement of the interface-Jacobian coupled substructure analysis.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the staccato_amd
package.  Follows test/IJCSA/src/IJSA.cu for simulation models, one Newton
step of the interface residual per frequency, in complex128 numpy:

  per substructure s (IJSA.cu:830-917): A_s(w) = (1+i eta)K_s + i w D_s - w^2 M_s,
      H_s = C_s A_s^{-1} B_s with B_s = [B_iface | B_load] and
      C_s = [C_iface ; C_internal] (zgetrf + two zgetrs + four GEMMs; the
      matrices are given, or unit-vector selections of index lists)
  initial interface solution (IJSA.cu:920-946): Y_s = H_ii u_s + H_il f_s
  residual (IJSA.cu:948-968):  r = sum_s dRdY_s Y_s + dRdU_s u_s
  Jacobian (IJSA.cu:523-527, 970-981): J = dRdU + sum_s dRdY_s H_ii (column
      block of s)
  correction (IJSA.cu:1199-1208): delta = J^{-1} r
  update (IJSA.cu:1349-1357, solver/assembly.cu:119-137): u_s -= delta_s
  solution (IJSA.cu:1359-1411): Y_s = H_ii u_s + H_il f_s,
      Y_int,s = H_ni u_s + H_nl f_s

dRdY and dRdU are Nu x Nu with the column blocks of the substructures in
order (Nu = sum of the interface sizes), as the reference's
dRdY_interface / dRdU blocks (IJSA.cu:537-553).  The reference's batched
solves are unpivoted (pivots NULL); numpy.linalg.solve pivots.

Pinned by the frequency-based substructuring identity: for two
substructures joined at shared interface DoFs with compatibility
(Y_a = Y_b) and equilibrium (u_a + u_b = 0) rows, the coupled response
equals the solve of the assembled (monolithic) system (tests/test_coupling.py).
"""
import numpy as np


def assemble(K, M, D, eta, w):
    A = (1.0 + 1j * eta) * K - w * w * M
    if D is not None:
        A = A + 1j * w * D
    return A


def io_matrices(s):
    """(B_iface, B_load, C_iface, C_internal) of a substructure dict: the
    matrices when given (a ROM's B_R / C_R blocks, IJSA.cu:869-917 multiplies
    by them), else unit-vector selections of the index lists."""
    n = s["K"].shape[0]
    if "B_iface" in s:
        return (np.asarray(s["B_iface"], np.complex128), np.asarray(s["B_load"], np.complex128),
                np.asarray(s["C_iface"], np.complex128), np.asarray(s["C_internal"], np.complex128))
    E = np.eye(n, dtype=np.complex128)
    return (E[:, list(s["iface"])], E[:, list(s["load"])], E[list(s["iface"]), :],
            E[list(s["internal"]), :])


def frf_blocks(s, w):
    """(H_ii, H_il, H_ni, H_nl) of one substructure at angular frequency w:
    X = A^{-1} [B_iface | B_load] (zgetrf + zgetrs, IJSA.cu:836-866), then
    C_iface X and C_internal X (IJSA.cu:869-917)."""
    Bi, Bl, Ci, Cn = io_matrices(s)
    ni = Bi.shape[1]
    X = np.linalg.solve(assemble(s["K"], s["M"], s.get("D"), s["eta"], w), np.hstack([Bi, Bl]))
    Hi, Hn = Ci @ X, Cn @ X
    return Hi[:, :ni], Hi[:, ni:], Hn[:, :ni], Hn[:, ni:]


def measured_blocks(s, j):
    """Experimental model (IJSA.cu:1011-1187): its FRF blocks at frequency j
    are given, H_ii (nf, ni, ni), H_il (nf, ni, nl), H_ni (nf, nn, ni),
    H_nl (nf, nn, nl)."""
    return s["H_ii"][j], s["H_il"][j], s["H_ni"][j], s["H_nl"][j]


def couple(subs, dRdY, dRdU, freq):
    """subs: list of dicts with u0 (ni), f (nl) and either a simulation model
    (K, M, D (or None), eta, and index lists iface, load, internal or the
    matrices B_iface (n x ni), B_load (n x nl), C_iface (ni x n), C_internal
    (nn x n)) or an experimental one (H_ii, H_il, H_ni, H_nl).  Returns
    (u (nfreq, Nu), Y (nfreq, Nu), [Y_int_s (nfreq, nn_s)])."""
    Nu = sum(len(s["u0"]) for s in subs)
    assert dRdY.shape == (Nu, Nu) and dRdU.shape == (Nu, Nu)
    nf = len(freq)
    U = np.zeros((nf, Nu), np.complex128)
    Y = np.zeros((nf, Nu), np.complex128)
    Yint = [np.zeros((nf, s["H_ni"].shape[1] if "H_ii" in s else io_matrices(s)[3].shape[0]),
                     np.complex128) for s in subs]
    for j, fz in enumerate(freq):
        w = 2.0 * np.pi * fz
        H, u, Ys = [], [], []
        J = np.array(dRdU, dtype=np.complex128)
        r = np.zeros(Nu, np.complex128)
        off = 0
        for s in subs:
            ni = len(s["u0"])
            Hb = measured_blocks(s, j) if "H_ii" in s else frf_blocks(s, w)
            us = np.array(s["u0"], dtype=np.complex128)
            fs = np.array(s["f"], dtype=np.complex128)
            ys = Hb[0] @ us + (Hb[1] @ fs if fs.size else 0)
            cols = slice(off, off + ni)
            r += dRdY[:, cols] @ ys + dRdU[:, cols] @ us
            J[:, cols] += dRdY[:, cols] @ Hb[0]
            H.append(Hb)
            u.append(us)
            off += ni
        delta = np.linalg.solve(J, r)
        off = 0
        for k, s in enumerate(subs):
            ni = len(s["u0"])
            us = u[k] - delta[off:off + ni]
            fs = np.array(s["f"], dtype=np.complex128)
            Hii, Hil, Hni, Hnl = H[k]
            U[j, off:off + ni] = us
            Y[j, off:off + ni] = Hii @ us + (Hil @ fs if fs.size else 0)
            if Hni.shape[0]:
                Yint[k][j] = Hni @ us + (Hnl @ fs if fs.size else 0)
            off += ni
    return U, Y, Yint
