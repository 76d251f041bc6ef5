"""Interface-Jacobian coupling (test/IJCSA/src/IJSA.cu) and the batched solve of
explicit matrices it needs (IJSA.cu:1199-1208).

CPU: the oracle restatement (oracle/ijcsa.py) is pinned by the
frequency-based substructuring identity: substructures joined at shared
interface DoFs, with compatibility rows (Y_a - Y_b = 0) and equilibrium rows
(u_a + u_b = 0), respond as the assembled (monolithic) system does.  A
planted mistake (the dRdU term left out of the Jacobian) breaks it.
GPU: stc_gesv_batched against numpy / the LAPACK-semantics oracle, and the
coupled GPU path against the oracle and the monolithic solve.
"""
import numpy as np
import pytest

import ijcsa
import oracle
from staccato_amd.coupling import Substructure


def _spd(rng, n, scale):
    G = rng.standard_normal((n, n))
    return scale * (G @ G.T / n + 0.5 * np.eye(n))


def chain_case(sizes, ni, seed=0, nl=2, eta=0.02, damping=True):
    """Substructures s = 0..S-1 with n_s DoFs; substructure s shares its last
    ni DoFs with the first ni of s+1.  Returns (subs for the oracle, the
    monolithic K, M, D, global DoF maps, dRdY, dRdU, load vector)."""
    rng = np.random.default_rng(seed)
    S = len(sizes)
    starts = np.cumsum([0] + [n - ni for n in sizes[:-1]])
    N = int(starts[-1] + sizes[-1])
    Kg, Mg, Dg = (np.zeros((N, N)) for _ in range(3))
    subs, gmaps = [], []
    for s, n in enumerate(sizes):
        K = _spd(rng, n, 1e4)
        M = _spd(rng, n, 1.0)
        D = 1e-4 * K + 0.3 * M if damping else None
        g = np.arange(starts[s], starts[s] + n)
        Kg[np.ix_(g, g)] += K
        Mg[np.ix_(g, g)] += M
        if damping:
            Dg[np.ix_(g, g)] += D
        # interface DoFs of s: shared with s-1 (first ni) and with s+1 (last ni)
        iface = ([] if s == 0 else list(range(ni))) + ([] if s == S - 1 else list(range(n - ni, n)))
        inner = [i for i in range(n) if i not in iface]
        load = inner[:nl] if s == 0 else []
        internal = inner
        f = (rng.standard_normal(len(load)) + 1j * rng.standard_normal(len(load)))
        subs.append(dict(K=K, M=M, D=D, eta=eta, iface=iface, load=load, internal=internal,
                         u0=np.zeros(len(iface), np.complex128), f=f))
        gmaps.append(g)
    # interface unknowns: per substructure, its interface DoFs in order
    offs = np.cumsum([0] + [len(s["iface"]) for s in subs])
    Nu = int(offs[-1])
    dRdY = np.zeros((Nu, Nu))
    dRdU = np.zeros((Nu, Nu))
    row = 0
    for s in range(S - 1):
        a_cols = offs[s] + len(subs[s]["iface"]) - ni + np.arange(ni)   # last ni of s
        b_cols = offs[s + 1] + np.arange(ni)                             # first ni of s+1
        for k in range(ni):
            dRdY[row, a_cols[k]], dRdY[row, b_cols[k]] = 1.0, -1.0       # compatibility
            dRdU[row + ni, a_cols[k]], dRdU[row + ni, b_cols[k]] = 1.0, 1.0  # equilibrium
            row += 1
        row += ni
    assert row == Nu
    F = np.zeros(N, np.complex128)
    F[gmaps[0][subs[0]["load"]]] = subs[0]["f"]
    return subs, (Kg, Mg, Dg if damping else None), gmaps, dRdY, dRdU, F


def monolithic(glob, eta, freq, F):
    K, M, D = glob
    return np.stack([np.linalg.solve(ijcsa.assemble(K, M, D, eta, 2 * np.pi * f), F) for f in freq])


def check_against_monolithic(subs, gmaps, Y, Yint, Xm, tol):
    """Interface and internal responses of every substructure equal the
    monolithic solution at the same global DoFs."""
    scale = np.abs(Xm).max()
    off = 0
    for k, (s, g) in enumerate(zip(subs, gmaps)):
        ni = len(s["iface"])
        assert np.abs(Y[:, off:off + ni] - Xm[:, g[s["iface"]]]).max() <= tol * scale
        if len(s["internal"]):
            assert np.abs(Yint[k] - Xm[:, g[s["internal"]]]).max() <= tol * scale
        off += ni


FREQ = np.array([3.0, 47.5, 120.0, 333.3, 810.0])


@pytest.mark.parametrize("sizes,ni", [((12, 10), 3), ((9, 14, 11), 2), ((6, 6), 1)])
def test_oracle_coupling_equals_monolithic(sizes, ni):
    subs, glob, gmaps, dRdY, dRdU, F = chain_case(sizes, ni, seed=len(sizes) + ni)
    U, Y, Yint = ijcsa.couple(subs, dRdY, dRdU, FREQ)
    Xm = monolithic(glob, subs[0]["eta"], FREQ, F)
    check_against_monolithic(subs, gmaps, Y, Yint, Xm, 1e-10)


def measured_version(s, freq):
    """The experimental-model form of a simulated substructure: its FRF
    blocks at every frequency (as a test rig would measure them)."""
    blocks = [ijcsa.frf_blocks(s, 2 * np.pi * fz) for fz in freq]
    m = {key: np.stack([b[q] for b in blocks]) for q, key in enumerate(("H_ii", "H_il", "H_ni", "H_nl"))}
    m.update(u0=s["u0"], f=s["f"])
    return m


def test_oracle_measured_member_equals_monolithic():
    """A chain whose middle member enters only through its FRFs (an
    experimental model, IJSA.cu:1011-1187) still gives the monolithic solution."""
    subs, glob, gmaps, dRdY, dRdU, F = chain_case((9, 14, 11), 2, seed=9)
    mixed = [subs[0], measured_version(subs[1], FREQ), subs[2]]
    U, Y, Yint = ijcsa.couple(mixed, dRdY, dRdU, FREQ)
    Xm = monolithic(glob, subs[0]["eta"], FREQ, F)
    check_against_monolithic(subs, gmaps, Y, Yint, Xm, 1e-10)


def _couple_without_dRdU_in_J(subs, dRdY, dRdU, freq):
    """ijcsa.couple's steps with the dRdU term dropped from the Jacobian
    (a small shift keeps J invertible); returns substructure 0's internal
    responses."""
    Nu = dRdU.shape[0]
    s0 = subs[0]
    Yi = np.zeros((len(freq), len(s0["internal"])), np.complex128)
    for j, fz in enumerate(freq):
        w = 2 * np.pi * fz
        J = 1e-3 * np.eye(Nu, dtype=np.complex128)
        r = np.zeros(Nu, np.complex128)
        off, H = 0, []
        for s in subs:
            ni = len(s["iface"])
            Hb = ijcsa.frf_blocks(s, w)
            ys = Hb[0] @ s["u0"] + (Hb[1] @ s["f"] if len(s["f"]) else 0)
            r += dRdY[:, off:off + ni] @ ys + dRdU[:, off:off + ni] @ s["u0"]
            J[:, off:off + ni] += dRdY[:, off:off + ni] @ Hb[0]
            H.append(Hb)
            off += ni
        d = np.linalg.solve(J, r)
        u = s0["u0"] - d[:len(s0["iface"])]
        Yi[j] = H[0][2] @ u + H[0][3] @ s0["f"]
    return Yi


def test_oracle_planted_mistake_breaks_identity():
    """A dropped term (dRdU left out of the Jacobian) fails the identity."""
    subs, glob, gmaps, dRdY, dRdU, F = chain_case((12, 10), 3, seed=5)
    Xm = monolithic(glob, subs[0]["eta"], FREQ, F)
    ref = Xm[:, gmaps[0][subs[0]["internal"]]]
    _, _, Yint = ijcsa.couple(subs, dRdY, dRdU, FREQ)
    assert np.abs(Yint[0] - ref).max() <= 1e-10 * np.abs(Xm).max()
    bad = _couple_without_dRdU_in_J(subs, dRdY, dRdU, FREQ)
    assert np.abs(bad - ref).max() > 1e-3 * np.abs(Xm).max()


# ---- GPU ---------------------------------------------------------------------------------

def _batch(rng, nb, n, cond_shift=2.0):
    A = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n))
    return A + cond_shift * np.sqrt(n) * np.eye(n)


@pytest.mark.gpu
@pytest.mark.parametrize("n,nb,nrhs,shared", [(1, 7, 1, False), (3, 5, 2, True), (17, 33, 3, False),
                                              (60, 64, 1, True), (200, 40, 4, False),
                                              (300, 9, 2, False)])
@pytest.mark.parametrize("mode", ["plain", "refined"])
def test_gesv_batched_matches_numpy(handle, n, nb, nrhs, shared, mode):
    rng = np.random.default_rng(n * 7 + nrhs)
    A = _batch(rng, nb, n)
    B = (rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))) if shared else \
        (rng.standard_normal((nb, n, nrhs)) + 1j * rng.standard_normal((nb, n, nrhs)))
    print(f"gesv_batched n={n} nb={nb} nrhs={nrhs} shared={shared} mode={mode}", flush=True)
    X, info, be, st = handle.gesv_batched(A, B, mode=mode, tol=1e-12 if mode == "refined" else 0.0)
    assert (info == 0).all()
    Xr = np.linalg.solve(A, np.broadcast_to(B, (nb, n, nrhs)))
    err = np.abs(X - Xr).max() / np.abs(Xr).max()
    assert err < (1e-11 if mode == "plain" else 1e-9), err
    assert be.max() <= (1e-13 if mode == "plain" else 1e-12)


@pytest.mark.gpu
def test_gesv_batched_pivots_and_singular(handle):
    """A zero column in one matrix gives info = its first zero pivot (as the
    LAPACK-semantics oracle's zgetrf); the others solve normally, and the
    plain solution of a pivoting-heavy matrix matches the oracle's zgetrs."""
    rng = np.random.default_rng(3)
    n, nb = 24, 6
    A = _batch(rng, nb, n, cond_shift=0.0)
    A[2][:, 5] = 0.0
    B = rng.standard_normal((n, 1)) + 0j
    X, info, be, _ = handle.gesv_batched(A, B, mode="plain")
    for j in range(nb):
        LU, ipiv, inf = oracle.zgetrf(A[j])
        assert info[j] == inf, (j, info[j], inf)
        if inf == 0:
            Xo = oracle.zgetrs(LU, ipiv, B)
            assert np.abs(X[j] - Xo).max() <= 1e-10 * np.abs(Xo).max()
    assert info[2] > 0


@pytest.mark.gpu
def test_gesv_batched_dev_matches_host(handle):
    import torch
    rng = np.random.default_rng(4)
    n, nb, nrhs = 40, 16, 2
    A = _batch(rng, nb, n)
    B = rng.standard_normal((nb, n, nrhs)) + 1j * rng.standard_normal((nb, n, nrhs))
    Xh, info_h, _, _ = handle.gesv_batched(A, B, mode="refined", tol=1e-12)
    dev = torch.device("cuda", 0)
    A_t = torch.as_tensor(np.ascontiguousarray(A.transpose(0, 2, 1)), device=dev)
    B_t = torch.as_tensor(np.ascontiguousarray(B.transpose(0, 2, 1)), device=dev)
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        X_t = torch.full_like(B_t, float("nan"))     # a kernel on the same stream first
        info_t = torch.full((nb,), -7, dtype=torch.int32, device=dev)
        handle.gesv_batched_dev(A_t, B_t, X_t, info_t, mode="refined", tol=1e-12,
                                stream=side.cuda_stream)
        Xd_t = X_t * 1.0                               # read back on the same stream
    side.synchronize()
    with pytest.raises(ValueError):
        handle.gesv_batched_dev(A_t, B_t, X_t, info_t, stream=0)
    Xd = Xd_t.cpu().numpy().transpose(0, 2, 1)
    assert (info_t.cpu().numpy() == 0).all() and (info_h == 0).all()
    assert np.abs(Xd - Xh).max() <= 1e-12 * np.abs(Xh).max()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "refined"])
@pytest.mark.parametrize("sizes,ni", [((12, 10), 3), ((40, 70, 33), 4)])
def test_gpu_coupling_matches_oracle_and_monolithic(handle, mode, sizes, ni):
    from staccato_amd.coupling import couple
    subs, glob, gmaps, dRdY, dRdU, F = chain_case(sizes, ni, seed=11 + ni)
    freq = np.linspace(1.0, 900.0, 37)
    print(f"coupling sizes={sizes} ni={ni} mode={mode}", flush=True)
    gsubs = []
    for k, s in enumerate(subs):
        rid = 40 + k
        handle.rom_set(rid, s["K"], s["M"], s["D"], s["eta"])
        gsubs.append(Substructure.from_dofs(rid, s["K"].shape[0], s["iface"], s["load"],
                                            s["internal"], s["u0"], s["f"]))
    tol = 1e-12 if mode == "refined" else 0.0
    out = couple(handle, gsubs, dRdY, dRdU, freq, mode=mode, tol=tol)
    assert (out["info_J"] == 0).all() and all((i == 0).all() for i in out["info_sub"])
    U, Y, Yint = ijcsa.couple(subs, dRdY, dRdU, freq)
    # the interface forces solve J delta = r, and cond(J) reaches 4e7 on these
    # chains (compliance rows ~1e-4 beside unit equilibrium rows), so their
    # error bound is cond(J) times the solve's backward error: u (1.1e-16,
    # plain) or 1e-12 (refined); the responses are much better conditioned
    gate = 1e-10 if mode == "plain" else 1e-8
    gate_u = 1e-9 if mode == "plain" else 1e-6
    assert np.abs(out["Y"] - Y).max() <= gate * np.abs(Y).max()
    assert np.abs(out["u"] - U).max() <= gate_u * np.abs(U).max()
    Xm = monolithic(glob, subs[0]["eta"], freq, F)
    check_against_monolithic(subs, gmaps, out["Y"], out["Y_internal"], Xm, gate)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "refined"])
def test_gpu_coupling_dense_io_matches_oracle(handle, mode):
    """General B/C blocks (a ROM's B_R / C_R, not DoF selections), nonzero
    initial interface forces: the GPU step against the oracle's."""
    from staccato_amd.coupling import couple
    subs, glob, gmaps, dRdY, dRdU, F = chain_case((30, 24), 3, seed=21)
    rng = np.random.default_rng(22)
    freq = np.linspace(5.0, 700.0, 29)
    gsubs = []
    for k, s in enumerate(subs):
        n = s["K"].shape[0]
        E = np.eye(n)
        cn = lambda *shape: 0.1 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        s["B_iface"] = E[:, s["iface"]] + cn(n, len(s["iface"]))
        s["C_iface"] = E[s["iface"], :] + cn(len(s["iface"]), n)
        s["B_load"] = E[:, s["load"]] + cn(n, len(s["load"]))
        s["C_internal"] = cn(5, n)
        s["u0"] = cn(len(s["iface"]))
        rid = 50 + k
        handle.rom_set(rid, s["K"], s["M"], s["D"], s["eta"])
        gsubs.append(Substructure(rid, s["B_iface"], s["C_iface"], s["B_load"], s["C_internal"],
                                  s["u0"], s["f"]))
    print(f"coupling dense io mode={mode}", flush=True)
    out = couple(handle, gsubs, dRdY, dRdU, freq, mode=mode, tol=1e-12 if mode == "refined" else 0)
    U, Y, Yint = ijcsa.couple(subs, dRdY, dRdU, freq)
    gate = 1e-10 if mode == "plain" else 1e-8
    gate_u = 1e-9 if mode == "plain" else 1e-6
    assert (out["info_J"] == 0).all() and all((i == 0).all() for i in out["info_sub"])
    assert np.abs(out["Y"] - Y).max() <= gate * np.abs(Y).max()
    assert np.abs(out["u"] - U).max() <= gate_u * np.abs(U).max()
    for a, b in zip(out["Y_internal"], Yint):
        assert np.abs(a - b).max() <= gate * np.abs(b).max()


@pytest.mark.gpu
def test_gpu_coupling_with_measured_member(handle):
    """Simulated ends, measured middle member: GPU step against the oracle and
    the monolithic solve."""
    from staccato_amd.coupling import MeasuredSubstructure, couple
    subs, glob, gmaps, dRdY, dRdU, F = chain_case((20, 26, 18), 3, seed=31)
    freq = np.linspace(2.0, 800.0, 23)
    m = measured_version(subs[1], freq)
    for k in (0, 2):
        s = subs[k]
        handle.rom_set(70 + k, s["K"], s["M"], s["D"], s["eta"])
    gsubs = [Substructure.from_dofs(70, 20, subs[0]["iface"], subs[0]["load"], subs[0]["internal"],
                                    subs[0]["u0"], subs[0]["f"]),
             MeasuredSubstructure(m["H_ii"], m["H_il"], m["H_ni"], m["H_nl"], m["u0"], m["f"]),
             Substructure.from_dofs(72, 18, subs[2]["iface"], subs[2]["load"], subs[2]["internal"],
                                    subs[2]["u0"], subs[2]["f"])]
    print("coupling with a measured member", flush=True)
    out = couple(handle, gsubs, dRdY, dRdU, freq)
    U, Y, Yint = ijcsa.couple([subs[0], m, subs[2]], dRdY, dRdU, freq)
    assert (out["info_J"] == 0).all() and all((i == 0).all() for i in out["info_sub"])
    assert np.abs(out["Y"] - Y).max() <= 1e-10 * np.abs(Y).max()
    assert np.abs(out["u"] - U).max() <= 1e-9 * np.abs(U).max()
    Xm = monolithic(glob, subs[0]["eta"], freq, F)
    check_against_monolithic(subs, gmaps, out["Y"], out["Y_internal"], Xm, 1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "refined"])
def test_gesv_batched_largest_size_and_limits(handle, mode):
    """n = 1024 (the largest system the kernel takes), an empty batch, and
    n = 1025 rejected before any launch."""
    from staccato_amd._lib import STC_EINVAL
    rng = np.random.default_rng(8)
    n, nb = 1024, 3
    A = _batch(rng, nb, n)
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    print(f"gesv_batched n=1024 mode={mode}", flush=True)
    X, info, be, _ = handle.gesv_batched(A, B, mode=mode, tol=1e-12 if mode == "refined" else 0.0)
    assert (info == 0).all()
    Xr = np.linalg.solve(A, np.broadcast_to(B, (nb, n, 2)))
    assert np.abs(X - Xr).max() <= 1e-9 * np.abs(Xr).max()
    assert be.max() <= 1e-12
    lib = handle.lib
    z = np.zeros(4, np.complex128)
    info1 = np.zeros(1, np.int32)
    assert lib.stc_gesv_batched(handle.h, 1025, 1, z.ctypes.data, 1025, 1025 * 1025, 1,
                                z.ctypes.data, 1025, 1025, z.ctypes.data, 1025, 1025,
                                info1.ctypes.data, None, 0, 0.0, 0, None) == STC_EINVAL
    assert lib.stc_gesv_batched(handle.h, 4, 0, None, 4, 16, 1, None, 4, 4, None, 4, 4, None,
                                None, 0, 0.0, 0, None) == 0


@pytest.mark.gpu
def test_gesv_batched_padded_layouts(handle):
    """Leading dimensions and batch strides larger than the packed ones (the
    caller's storage of IJSA.cu-style pointer batches), through the C ABI."""
    import ctypes
    rng = np.random.default_rng(12)
    n, nb, nrhs = 37, 5, 3
    lda, sA = n + 3, (n + 3) * n + 5
    ldb, sB = n + 2, (n + 2) * nrhs + 1
    ldx, sX = n + 1, (n + 1) * nrhs + 2
    A = _batch(rng, nb, n)
    B = rng.standard_normal((nb, n, nrhs)) + 1j * rng.standard_normal((nb, n, nrhs))
    Abuf = np.full(sA * nb, np.nan, np.complex128)
    Bbuf = np.full(sB * nb, np.nan, np.complex128)
    Xbuf = np.full(sX * nb, 7.0 + 7j, np.complex128)
    for j in range(nb):
        for c in range(n):
            Abuf[j * sA + c * lda: j * sA + c * lda + n] = A[j][:, c]
        for c in range(nrhs):
            Bbuf[j * sB + c * ldb: j * sB + c * ldb + n] = B[j][:, c]
    info = np.zeros(nb, np.int32)
    print("gesv_batched padded layouts", flush=True)
    rc = handle.lib.stc_gesv_batched(handle.h, n, nb, Abuf.ctypes.data, lda, sA, nrhs,
                                     Bbuf.ctypes.data, ldb, sB, Xbuf.ctypes.data, ldx, sX,
                                     info.ctypes.data, None, 0, 0.0, 0, None)
    assert rc == 0 and (info == 0).all()
    Xr = np.linalg.solve(A, B)
    for j in range(nb):
        for c in range(nrhs):
            col = Xbuf[j * sX + c * ldx: j * sX + c * ldx + n]
            assert np.abs(col - Xr[j][:, c]).max() <= 1e-11 * np.abs(Xr).max()
            # padding between columns and blocks untouched
            assert (Xbuf[j * sX + c * ldx + n: j * sX + (c + 1) * ldx] == 7.0 + 7j).all()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "refined"])
def test_cpp_host_coupling_matches_oracle(mode):
    """The C++ host mirror (IJCSACoupling, through stc_host_ijcsa_couple)
    against the oracle and the monolithic solve."""
    import ctypes
    from staccato_amd import _lib
    _lib.load()
    host = ctypes.CDLL(_lib.HOST_LIB_PATH)
    subs, glob, gmaps, dRdY, dRdU, F = chain_case((20, 26, 18), 3, seed=41)
    freq = np.linspace(3.0, 850.0, 31)
    dims, Ks, Ms, Ds, Bs, Cs, us, fs, etas = [], [], [], [], [], [], [], [], []
    for s in subs:
        n = s["K"].shape[0]
        Bi, Bl, Ci, Cn = ijcsa.io_matrices(s)
        dims += [n, Bi.shape[1], Bl.shape[1], Cn.shape[0]]
        Ks.append(s["K"].astype(np.complex128).ravel("F"))
        Ms.append(s["M"].astype(np.complex128).ravel("F"))
        Ds.append(s["D"].astype(np.complex128).ravel("F"))
        Bs.append(np.hstack([Bi, Bl]).ravel("F"))
        Cs.append(np.vstack([Ci, Cn]).ravel("F"))
        us.append(s["u0"])
        fs.append(np.asarray(s["f"], np.complex128))
        etas.append(s["eta"])
    cat = lambda xs: np.ascontiguousarray(np.concatenate(xs), dtype=np.complex128)
    dims = np.array(dims, np.int32)
    K, M, D, B, C, u0, f = map(cat, (Ks, Ms, Ds, Bs, Cs, us, fs))
    eta = np.array(etas, np.float64)
    Nu = dRdY.shape[0]
    nint = sum(dims[3::4])
    U = np.zeros((freq.size, Nu), np.complex128)
    Y = np.zeros((freq.size, Nu), np.complex128)
    Yint = np.zeros(freq.size * nint, np.complex128)
    infoJ = np.zeros(freq.size, np.int32)
    dY = np.asfortranarray(dRdY, dtype=np.complex128)
    dU = np.asfortranarray(dRdU, dtype=np.complex128)
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    fn = host.stc_host_ijcsa_couple
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 11 + \
        [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double] + [ctypes.c_void_p] * 4
    print(f"C++ coupling mode={mode}", flush=True)
    rc = fn(0, len(subs), p(dims), p(K), p(M), p(D), p(eta), p(B), p(C), p(u0), p(f), p(dY), p(dU),
            freq.size, p(freq), 1 if mode == "refined" else 0, 1e-12 if mode == "refined" else 0.0,
            p(U), p(Y), p(Yint), p(infoJ))
    assert rc == 0 and (infoJ == 0).all()
    Uo, Yo, Yinto = ijcsa.couple(subs, dRdY, dRdU, freq)
    gate = 1e-10 if mode == "plain" else 1e-8
    assert np.abs(Y - Yo).max() <= gate * np.abs(Yo).max()
    assert np.abs(U - Uo).max() <= (1e-9 if mode == "plain" else 1e-6) * np.abs(Uo).max()
    o, Yint_l = 0, []
    for k, s in enumerate(subs):
        nn = dims[4 * k + 3]
        Yint_l.append(Yint[o:o + freq.size * nn].reshape(freq.size, nn))
        o += freq.size * nn
    Xm = monolithic(glob, subs[0]["eta"], freq, F)
    check_against_monolithic(subs, gmaps, Y, Yint_l, Xm, gate)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "refined"])
def test_gpu_coupling_full_size_residual_vanishes(handle, mode):
    """Three n=256 substructures (the sweep's size), 16 shared DoFs per joint:
    after the Newton step the interface residual R(u) = dRdY Y + dRdU u is zero
    (compatibility and equilibrium hold) at every frequency."""
    from staccato_amd.coupling import couple
    subs, glob, gmaps, dRdY, dRdU, F = chain_case((256, 256, 256), 16, seed=51, nl=4)
    freq = np.linspace(1.0, 2000.0, 512)
    gsubs = []
    for k, s in enumerate(subs):
        handle.rom_set(80 + k, s["K"], s["M"], s["D"], s["eta"])
        gsubs.append(Substructure.from_dofs(80 + k, 256, s["iface"], s["load"], s["internal"][-8:],
                                            s["u0"], s["f"]))
    print(f"coupling 3 x n=256 mode={mode}", flush=True)
    out = couple(handle, gsubs, dRdY, dRdU, freq, mode=mode, tol=1e-12 if mode == "refined" else 0)
    assert (out["info_J"] == 0).all() and all((i == 0).all() for i in out["info_sub"])
    Y, U = out["Y"], out["u"]
    R = Y @ dRdY.T + U @ dRdU.T                     # (nf, Nu)
    Nu, nj = dRdY.shape[0], 16
    compat = np.concatenate([R[:, 0:nj], R[:, 2 * nj:3 * nj]], axis=1)
    equil = np.concatenate([R[:, nj:2 * nj], R[:, 3 * nj:4 * nj]], axis=1)
    # R(u) is formed from the same FRFs as J, and J is solved in FP64 in both
    # modes: the residual vanishes to FP64 accuracy whatever the sweeps' mode
    gate_c, gate_e = 1e-10, 1e-12
    assert np.abs(compat).max() <= gate_c * np.abs(Y).max()
    assert np.abs(equil).max() <= gate_e * np.abs(U).max()
