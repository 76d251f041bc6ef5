"""Benchmark: frequency points solved per second at n=256 (refined FP32 vs plain FP64).

Workload (one "step" = one full sweep): a Krylov ROM of order n=256 (seed 1,
Rayleigh damping D = 1e-4 K + 0.5 M, eta = 0.01), 4096 frequencies linear in
1..2000 Hz, 4 complex N(0,1) right-hand sides.  For every frequency the GPU
assembles A(w) = (1 + i eta) K + i w D - w^2 M, factors it with partial
pivoting and solves; the refined mode factors in complex64 and refines in FP64
to a normwise backward error <= 1e-10.  Inputs (K, M, D, frequencies, F) are
resident in HBM before the timed region; outputs stay on the device.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  The plain FP64 sweep (the denominator of the metric),
a vendor batched LU (rocSOLVER zgetrf/zgetrs strided-batched on the
materialised A(w), complex128) and the
CPU reference loop (zgetrf + zgetrs per frequency over OpenBLAS, OpenMP over
frequencies) are measured beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, NFREQ, NRHS = 256, 4096, 4
TOL_REFINED, TOL_PLAIN, TOL_ADJOINT = 1e-10, 1e-12, 1e-12
PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 matrix
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the sweep kernel on this workload, per launch policy
# (the autotune candidate the run uses): rocprofv3 --pmc passes (FETCH_SIZE, then
# WRITE_SIZE; KiB per dispatch, mean over the launches) scaled by what the same
# counters read for a 256 MiB device-to-device copy (FETCH_SIZE reads half the
# bytes copied: x2.0; WRITE_SIZE reads them exactly: x1.0).  The summary and the
# commands are in profiles/sweep_kernel_n256_pmc.txt; re-measure when the kernel changes.
TRAFFIC_KIB = {
    ("refined", "wide-2perCU"): (7.035e6, 8.095e6),
    ("refined", "wide-3M-2perCU"): (7.025e6, 7.991e6),
    ("refined", "wide-1perCU-jb32"): (6.900e6, 8.069e6),
    ("refined", "register-panel"): (11.23e6, 14.23e6),
    ("plain", "wide-2perCU"): (15.12e6, 26.92e6),
    ("plain", "wide-2level-2perCU"): (13.98e6, 16.01e6),
    ("plain", "wide-1perCU"): (7.217e6, 14.84e6),
    ("plain", "register-panel"): (19.79e6, 34.96e6),
}
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0


def work_per_point(n, nrhs):
    """Real flops of LU + solve per frequency point: 8/3 n^3 + 8 n^2 nrhs."""
    return 8.0 / 3.0 * n ** 3 + 8.0 * n * n * nrhs


def time_gpu(run, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        run()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_time(e1) * 1e-3


def vendor_fp64(case, steps, warmup):
    """Plain FP64 with the vendor stack: A(w) materialised per frequency (torch
    elementwise ops), rocsolver_zgetrf_strided_batched + zgetrs_strided_batched
    (librocsolver, called through ctypes), in chunks of 512 frequencies."""
    import ctypes
    import torch
    # the copies bundled with torch match the HIP runtime torch has loaded
    tlib = os.path.join(os.path.dirname(torch.__file__), "lib")
    rb = ctypes.CDLL(os.path.join(tlib, "librocblas.so"))
    rs = ctypes.CDLL(os.path.join(tlib, "librocsolver.so"))
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
    rb.rocblas_create_handle.argtypes = [ctypes.POINTER(vp)]
    rb.rocblas_set_stream.argtypes = [vp, vp]
    rb.rocblas_destroy_handle.argtypes = [vp]
    rs.rocsolver_zgetrf_strided_batched.argtypes = [vp, i32, i32, vp, i32, i64, vp, i64, vp, i32]
    rs.rocsolver_zgetrs_strided_batched.argtypes = [vp, i32, i32, i32, vp, i32, i64, vp, i64, vp,
                                                    i32, i64, i32]
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    handle = vp()
    if rb.rocblas_create_handle(ctypes.byref(handle)) != 0:
        raise RuntimeError("rocblas_create_handle failed")
    rb.rocblas_set_stream(handle, vp(stream.cuda_stream))
    rom = case.rom
    K = torch.tensor(rom.K, dtype=torch.complex128, device=dev)
    M = torch.tensor(rom.M, dtype=torch.complex128, device=dev)
    D = torch.tensor(rom.D, dtype=torch.complex128, device=dev)
    w = torch.tensor(2 * np.pi * case.freq, dtype=torch.float64, device=dev)
    chunk = 512
    # column-major A: build A^T in row-major torch storage (A is symmetric here, but
    # transpose anyway so the layout is right for any operator)
    Kt, Mt, Dt = K.T.contiguous(), M.T.contiguous(), D.T.contiguous()
    A = torch.empty((chunk, N, N), dtype=torch.complex128, device=dev)
    B = torch.empty((chunk, NRHS, N), dtype=torch.complex128, device=dev)
    Ft = torch.tensor(np.ascontiguousarray(case.F.T), dtype=torch.complex128, device=dev)
    ipiv = torch.empty((chunk, N), dtype=torch.int32, device=dev)
    info = torch.empty(chunk, dtype=torch.int32, device=dev)
    status = []

    def run():
        for s0 in range(0, NFREQ, chunk):
            ww = w[s0:s0 + chunk].to(torch.complex128)[:, None, None]
            torch.add((1 + 1j * rom.eta) * Kt, 1j * ww * Dt - ww * ww * Mt, out=A)
            B.copy_(Ft.expand(chunk, -1, -1))
            status.append(rs.rocsolver_zgetrf_strided_batched(handle, N, N, vp(A.data_ptr()), N, N * N,
                                                              vp(ipiv.data_ptr()), N,
                                                              vp(info.data_ptr()), chunk))
            status.append(rs.rocsolver_zgetrs_strided_batched(handle, 111, N, NRHS, vp(A.data_ptr()), N,
                                                              N * N, vp(ipiv.data_ptr()), N,
                                                              vp(B.data_ptr()), N, N * NRHS, chunk))

    wall, _ = time_gpu(run, steps, warmup, stream)
    # correctness of the vendor result on the last chunk
    X = B.cpu().numpy().transpose(0, 2, 1)
    Xc = synth_closed_last_chunk(case, chunk)
    err = float((np.abs(X - Xc).max(axis=(1, 2)) / np.abs(Xc).max(axis=(1, 2))).max())
    rb.rocblas_destroy_handle(handle)
    if any(st != 0 for st in status):
        raise RuntimeError(f"rocsolver status {set(status)}")
    return NFREQ * steps / wall, err


def synth_closed_last_chunk(case, chunk):
    from staccato_amd import synth
    return synth.closed_form(case.rom, case.freq[NFREQ - chunk:], case.F)


def host_cpu_info():
    """Logical CPUs of the host, those this process may run on, and the CPU
    share it uses (OMP_NUM_THREADS when set: the share a GPU job is given)."""
    total = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = total
    share = avail
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        share = max(1, min(avail, int(os.environ["OMP_NUM_THREADS"])))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return total, avail, share, model


def cpu_baseline(case, budget_s=20.0):
    """Reference loop on the host cores (OpenBLAS zgetrf/zgetrs per frequency),
    OpenMP over frequencies with the process's CPU share, plus one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    total, avail, cores, model = host_cpu_info()
    rom = case.rom
    lib = oracle.openblas_path()
    # one full 4096-frequency sweep calibrates; whole sweeps then repeat for ~budget_s
    if lib is not None and oracle.sweep_lapack(rom.K, rom.M, rom.D, rom.eta, case.freq[:8], case.F,
                                               nthreads=cores) is not None:
        kind_lib = "OpenBLAS (scipy-bundled) zgetrf_/zgetrs_"
        fn = lambda fr: oracle.sweep_lapack(rom.K, rom.M, rom.D, rom.eta, fr, case.F,
                                            nthreads=cores)[2]
    else:
        kind_lib = "oracle C zgetf2/zgetrs"
        fn = lambda fr: oracle.sweep(rom.K, rom.M, rom.D, rom.eta, fr, case.F, nthreads=cores,
                                     berr=False)[3]
    per_sweep = max(fn(case.freq), 1e-6)
    reps = int(max(1, min(200, budget_s / per_sweep)))
    t = 0.0
    for _ in range(reps):
        t += fn(case.freq)
    # one thread: a quarter sweep (every 4th frequency)
    if lib is not None and "OpenBLAS" in kind_lib:
        fn1 = lambda fr: oracle.sweep_lapack(rom.K, rom.M, rom.D, rom.eta, fr, case.F, nthreads=1)[2]
    else:
        fn1 = lambda fr: oracle.sweep(rom.K, rom.M, rom.D, rom.eta, fr, case.F, nthreads=1,
                                      berr=False)[3]
    fr1 = case.freq[::4]
    t1 = fn1(fr1)
    return {"value": reps * NFREQ / t, "unit": "freq points/s", "cores": cores, "kind": "port",
            "sample": f"{reps} x the full {NFREQ}-frequency sweep (n=256, 4 RHS), {kind_lib}, "
                      f"OpenMP over frequencies, {cores} threads (the process's CPU share; the host "
                      f"has {total} logical CPUs, {avail} available: {model}), 1 LAPACK thread each, "
                      f"{t:.1f} s",
            "one_thread": {"value": fr1.size / t1, "unit": "freq points/s", "cores": 1,
                           "sample": f"{fr1.size} frequencies (every 4th), {t1:.1f} s"},
            "host_logical_cpus": total, "available_cpus": avail}


def pts_total(cases):
    return sum(c.freq.size for c in cases)


def other_workload(name, steps, warmup, policy=""):
    """Throughput of the other configurations (reported beside the headline,
    never as its value): the uneven 64-ROM batch (hipGraph plan), the n=128
    adjoint and the n=1024 dense sweep, all refined FP32 unless stated."""
    import torch
    from staccato_amd import Handle, Plan, synth
    from staccato_amd._lib import SysDesc
    dev = torch.device("cuda:0")
    h = Handle(0)
    out = {"workload": name}
    if name == "uneven":
        cases = synth.uneven_batch(nsys=64)
        keep, descs = [], (SysDesc * len(cases))()
        W = 0.0
        for i, c in enumerate(cases):
            h.rom_set(i, c.rom.K, c.rom.M, c.rom.D, c.rom.eta)
            n, nf, nr = c.rom.n, c.freq.size, c.F.shape[1]
            t = (torch.tensor(c.freq, dtype=torch.float64, device=dev),
                 torch.tensor(np.ascontiguousarray(c.F.T), dtype=torch.complex128, device=dev),
                 torch.empty((nf, nr, n), dtype=torch.complex128, device=dev),
                 torch.empty(nf, dtype=torch.int32, device=dev),
                 torch.empty((nf, nr), dtype=torch.float64, device=dev))
            keep.append(t)
            d = descs[i]
            d.rom_id, d.nfreq, d.freq_hz, d.nrhs = i, nf, t[0].data_ptr(), nr
            d.F, d.ldf, d.X, d.ldx = t[1].data_ptr(), n, t[2].data_ptr(), n
            d.info, d.berr = t[3].data_ptr(), t[4].data_ptr()
            W += nf * work_per_point(n, nr)
        # launch policy by measurement: the plan's launch is keyed by its largest order
        from staccato_amd import autotune
        nmax = max(c.rom.n for c in cases)
        stream = torch.cuda.current_stream()
        trial = {}
        for name, kernel, budget, jb in autotune.CANDIDATES:
            h.set_policy(nmax, "refined", kernel, budget, jb)
            p = Plan(h, descs, mode="refined", tol=TOL_REFINED)
            w_, _ = time_gpu(lambda: p.execute(), 2, 1, stream)
            ok = all(int((t[3] != 0).sum()) == 0 for t in keep)
            trial[name] = (pts_total(cases) * 2 / w_ if ok else 0.0, kernel, budget, jb)
            p.close()
        best = max(trial, key=lambda k: trial[k][0])
        h.set_policy(nmax, "refined", *trial[best][1:])
        out["autotune"] = {"chosen": best, "candidates_pts_per_s": {k: round(v[0]) for k, v in trial.items()}}
        plan = Plan(h, descs, mode="refined", tol=TOL_REFINED)
        wall, _ = time_gpu(lambda: plan.execute(), steps, warmup, stream)
        pts = sum(c.freq.size for c in cases)
        bmax = max(float(t[4].max()) for t in keep)
        bad = sum(int((t[3] != 0).sum()) for t in keep)
        # per-system error against the FP64 closed-form solution (every 16th frequency)
        errs = []
        for c, t in zip(cases, keep):
            sel = np.arange(0, c.freq.size, 16)
            Xh = t[2].cpu().numpy().transpose(0, 2, 1)[sel]
            Xc = synth.closed_form(c.rom, c.freq[sel], c.F)
            errs.append(float((np.abs(Xh - Xc).max(axis=(1, 2)) / np.abs(Xc).max(axis=(1, 2))).max()))
        worst = int(np.argmax(errs))
        out.update(points=pts, pts_per_s=pts * steps / wall, tflops=W * steps / wall / 1e12,
                   max_backward_error=bmax, points_failed=bad,
                   per_system_max_rel_error_vs_fp64_closed_form=errs,
                   worst_system={"index": worst, "n": int(cases[worst].rom.n),
                                 "rel_error": errs[worst]})
        plan.close()
    elif name == "adjoint":
        # refined: both solves refined to a backward error of TOL_ADJOINT (at the
        # sweep's 1e-10 the gradient carries ~kappa * 1e-10 ~ 1e-5 error)
        import torch
        from staccato_amd.adjoint import ModalObjective
        c = synth.adjoint_n128()
        rom, cv = c.rom, c.extra["c"]
        h.rom_set(0, rom.K, rom.M, rom.D, rom.eta)
        Jc, dsc, dec = synth.adjoint_closed_form(rom, c.freq, c.F, cv)
        comps = [int(k) for k in np.argsort(-np.abs(dsc))[:4]]
        # central differences (h = 1e-6) of the FP64 objective through the autograd wrapper
        fdobj = ModalObjective(h, 1, rom.Phi, rom.kdiag, rom.M, rom.D, c.freq, c.F, cv)
        s0 = torch.ones(rom.n, dtype=torch.float64)
        e0 = torch.tensor(rom.eta, dtype=torch.float64)
        step = 1e-6
        fd = {}
        for k in comps:
            e = torch.zeros(rom.n, dtype=torch.float64)
            e[k] = step
            fd[k] = (float(fdobj(s0 + e, e0)) - float(fdobj(s0 - e, e0))) / (2 * step)
        fde = (float(fdobj(s0, e0 + step)) - float(fdobj(s0, e0 - step))) / (2 * step)
        out["fd"] = {"h": step, "components": comps, "of": "plain complex128 objective"}
        for mode in ("plain", "refined"):
            tol = TOL_ADJOINT if mode == "refined" else 0.0
            for _ in range(warmup):
                h.adjoint(0, c.freq, c.F, cv, rom.Phi, rom.kdiag, mode=mode, tol=tol)
            t0 = time.perf_counter()
            for _ in range(steps):
                J, ds, de, info = h.adjoint(0, c.freq, c.F, cv, rom.Phi, rom.kdiag, mode=mode,
                                            tol=tol)
            dt = (time.perf_counter() - t0) / steps
            out[mode] = {"tol": tol if mode == "refined" else None,
                         "pts_per_s_incl_host_copies": c.freq.size / dt,
                         "points_failed": int((info != 0).sum()),
                         "J_rel_err_vs_closed_form": float(abs(J / Jc - 1)),
                         "grad_rel_err_vs_closed_form_normwise":
                             float(np.abs(ds - dsc).max() / np.abs(dsc).max()),
                         "grad_rel_err_vs_closed_form_fd_components":
                             float(max(abs(ds[k] / dsc[k] - 1) for k in comps)),
                         "deta_rel_err_vs_closed_form": float(abs(de / dec - 1)),
                         "grad_rel_err_vs_fd": float(max(max(abs(ds[k] / fd[k] - 1) for k in comps),
                                                         abs(de / fde - 1)))}
    elif name == "dense1024":
        c = synth.dense_n1024()
        rom = c.rom
        h.rom_set(0, rom.K, rom.M, rom.D, rom.eta)
        fr = torch.tensor(c.freq, dtype=torch.float64, device=dev)
        F = torch.tensor(np.ascontiguousarray(c.F.T), dtype=torch.complex128, device=dev)
        X = torch.empty((c.freq.size, 16, 1024), dtype=torch.complex128, device=dev)
        info = torch.empty(c.freq.size, dtype=torch.int32, device=dev)
        be = torch.empty((c.freq.size, 16), dtype=torch.float64, device=dev)
        stream = torch.cuda.Stream(device=dev)
        from staccato_amd import autotune
        out["autotune"] = {}
        fixed = dict(kv.split("=", 1) for kv in policy.split(",") if "=" in kv)
        cands = {q[0]: q for q in autotune.CANDIDATES}
        for mode in ("refined", "plain"):
            tol = TOL_REFINED if mode == "refined" else TOL_PLAIN
            if mode in fixed:
                _, kernel, budget, jb = cands[fixed[mode]]
                h.set_policy(1024, mode, kernel, budget, jb)
                out["autotune"][mode] = {"chosen": fixed[mode], "given": True}
            else:
                best, cand = autotune.tune(h, 0, c.freq, c.F, mode, tol, sample=c.freq.size, reps=1,
                                           cache=False)
                out["autotune"][mode] = {"chosen": best["name"] if best else "automatic",
                                         "candidates_pts_per_s": {q["name"]: round(q["pts_per_s"]) for q in cand}}
            run = lambda: h.sweep_dev(0, fr, F, X, info, be if mode == "refined" else None,
                                      mode=mode, tol=TOL_REFINED, max_iter=10,
                                      stream=stream.cuda_stream)
            wall, ev = time_gpu(run, steps, warmup, stream)
            h.sweep_dev(0, fr, F, X, info, be, mode=mode, tol=tol, max_iter=10, stream=stream.cuda_stream)
            torch.cuda.synchronize()
            Xh = X.cpu().numpy().transpose(0, 2, 1)
            sel = np.arange(0, c.freq.size, 8)  # closed form on every 8th frequency (host FP64)
            Xc = synth.closed_form(rom, c.freq[sel], c.F)
            fe = np.abs(Xh[sel] - Xc).max(axis=(1, 2)) / np.abs(Xc).max(axis=(1, 2))
            tfl = c.freq.size * work_per_point(1024, 16) * steps / ev / 1e12
            peak = PEAK_FP32_TFLOPS if mode == "refined" else PEAK_FP64_TFLOPS
            out[mode] = {"pts_per_s": c.freq.size * steps / wall, "tflops": tfl,
                         "mfma_peak_frac": tfl / peak,
                         "max_backward_error": float(be.max()),
                         "points_failed": int((info != 0).sum()),
                         "max_rel_error_vs_closed_form": float(fe.max())}
    elif name == "small60":
        # the reference's CPU-runnable case: one n=60 ROM, 1000 frequencies, a unit
        # force at DoF 0; the whole matrix is factored in LDS (no factor slot)
        c = synth.single_n60()
        rom = c.rom
        h.rom_set(0, rom.K, rom.M, rom.D, rom.eta)
        fr = torch.tensor(c.freq, dtype=torch.float64, device=dev)
        F = torch.tensor(np.ascontiguousarray(c.F.T), dtype=torch.complex128, device=dev)
        nf, nr = c.freq.size, c.F.shape[1]
        X = torch.empty((nf, nr, 60), dtype=torch.complex128, device=dev)
        info = torch.empty(nf, dtype=torch.int32, device=dev)
        be = torch.empty((nf, nr), dtype=torch.float64, device=dev)
        stream = torch.cuda.Stream(device=dev)
        for mode in ("plain", "refined"):
            tol = TOL_REFINED if mode == "refined" else TOL_PLAIN
            run = lambda: h.sweep_dev(0, fr, F, X, info, be, mode=mode, tol=tol, max_iter=10,
                                      stream=stream.cuda_stream)
            wall, ev = time_gpu(run, steps, warmup, stream)
            torch.cuda.synchronize()
            Xh = X.cpu().numpy().transpose(0, 2, 1)
            Xc = synth.closed_form(rom, c.freq, c.F)
            fe = np.abs(Xh - Xc).max(axis=(1, 2)) / np.abs(Xc).max(axis=(1, 2))
            # algorithmic HBM bytes per sweep: operators once, F in, X out, info and berr
            byts = 3 * 60 * 60 * 8 + nr * 60 * 16 + nf * (nr * 60 * 16 + 4 + nr * 8)
            out[mode] = {"pts_per_s": nf * steps / wall, "kernel_ms": 1e3 * ev / steps,
                         "algorithmic_hbm_gbs": byts * steps / ev / 1e9,
                         "tflops": nf * work_per_point(60, nr) * steps / ev / 1e12,
                         "max_backward_error": float(be.max()), "points_failed": int((info != 0).sum()),
                         "max_rel_error_vs_closed_form": float(fe.max())}
    elif name == "ijcsa":
        # interface-Jacobian coupling: a chain of three n=256 substructures
        # (16 shared DoFs per joint, 4 loads, 8 outputs each), 4096 frequencies
        from staccato_amd.coupling import Substructure, couple
        rng = np.random.default_rng(7)
        nfq, n, nj = 4096, 256, 16
        freq = np.linspace(1.0, 2000.0, nfq)
        subs = []
        for k in range(3):
            G = rng.standard_normal((n, n))
            K = 1e4 * (G @ G.T / n + 0.5 * np.eye(n))
            G = rng.standard_normal((n, n))
            M = G @ G.T / n + 0.5 * np.eye(n)
            h.rom_set(k, K, M, 1e-4 * K + 0.3 * M, 0.02)
            iface = ([] if k == 0 else list(range(nj))) + ([] if k == 2 else list(range(n - nj, n)))
            inner = [i for i in range(n) if i not in iface]
            load = inner[:4] if k == 0 else []
            f = rng.standard_normal(len(load)) + 0j
            subs.append(Substructure.from_dofs(k, n, iface, load, inner[-8:], None, f))
        Nu = sum(s.u0.size for s in subs)
        dRdY, dRdU = np.zeros((Nu, Nu)), np.zeros((Nu, Nu))
        offs = np.cumsum([0] + [s.u0.size for s in subs])
        row = 0
        for k in range(2):
            a_cols = offs[k] + subs[k].u0.size - nj + np.arange(nj)
            b_cols = offs[k + 1] + np.arange(nj)
            for q in range(nj):
                dRdY[row, a_cols[q]], dRdY[row, b_cols[q]] = 1.0, -1.0
                dRdU[row + nj, a_cols[q]], dRdU[row + nj, b_cols[q]] = 1.0, 1.0
                row += 1
            row += nj
        # the interface-Jacobian solves alone, device resident (Nu = 64)
        Js = torch.as_tensor(rng.standard_normal((nfq, Nu, Nu)) + 8.0 * np.eye(Nu) + 0j, device=dev)
        rs = torch.as_tensor(rng.standard_normal((nfq, 1, Nu)) + 0j, device=dev)
        ds = torch.empty_like(rs)
        inf = torch.empty(nfq, dtype=torch.int32, device=dev)
        stream = torch.cuda.Stream(device=dev)
        for mode in ("refined", "plain"):
            tol = TOL_REFINED if mode == "refined" else 0.0
            for _ in range(warmup):
                couple(h, subs, dRdY, dRdU, freq, mode=mode, tol=tol)
            t0 = time.perf_counter()
            for _ in range(steps):
                res = couple(h, subs, dRdY, dRdU, freq, mode=mode, tol=tol)
            dt = (time.perf_counter() - t0) / steps
            bad = int((res["info_J"] != 0).sum()) + sum(int((i != 0).sum()) for i in res["info_sub"])
            # the interface residual after the Newton step (compatibility rows
            # relative to the responses, equilibrium rows to the forces)
            R = res["Y"] @ dRdY.T + res["u"] @ dRdU.T
            comp = max(float(np.abs(R[:, row0:row0 + nj]).max()) for row0 in (0, 2 * nj))
            equi = max(float(np.abs(R[:, row0:row0 + nj]).max()) for row0 in (nj, 3 * nj))
            comp /= float(np.abs(res["Y"]).max())
            equi /= float(np.abs(res["u"]).max())
            run = lambda: h.gesv_batched_dev(Js, rs, ds, inf, mode=mode, tol=tol,
                                             stream=stream.cuda_stream)
            wall, ev = time_gpu(run, steps, warmup, stream)
            out[mode] = {"coupled_pts_per_s": nfq / dt, "points_failed": bad,
                         "compatibility_residual_rel": comp, "equilibrium_residual_rel": equi,
                         "interface_solve_pts_per_s": nfq * steps / wall,
                         "interface_solve_kernel_ms": 1e3 * ev / steps}
        out["config"] = {"substructures": 3, "n": n, "shared_dofs_per_joint": nj, "Nu": int(Nu),
                         "nfreq": nfq, "loads": 4, "outputs_per_substructure": 8}
    h.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-vendor", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", default="both", choices=["both", "refined", "plain"])
    ap.add_argument("--workload", default="uniform",
                    choices=["uniform", "uneven", "adjoint", "dense1024", "small60", "ijcsa"],
                    help="uniform = the headline n=256 sweep; the others print their own line")
    ap.add_argument("--policy", default="", metavar="MODE=NAME[,MODE=NAME]",
                    help="use these autotune candidates instead of tuning (e.g. for a profile "
                         "of exactly the kernels a tuned run uses)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run 3 device-to-device copies of MIB MiB (a kernel of known "
                         "traffic, to scale rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the sweep shards by frequency with no exchange; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload != "uniform":
        other_workload(args.workload, args.steps, args.warmup, args.policy)
        return

    import torch
    from staccato_amd import Handle, autotune, synth

    case = synth.uniform_n256(NFREQ)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    h = Handle(0)
    assert h.device_ok, "not a gfx950 device"
    rom = case.rom
    h.rom_set(0, rom.K, rom.M, rom.D, rom.eta)
    freq_t = torch.tensor(case.freq, dtype=torch.float64, device=dev)
    F_t = torch.tensor(np.ascontiguousarray(case.F.T), dtype=torch.complex128, device=dev)
    X_t = torch.empty((NFREQ, NRHS, N), dtype=torch.complex128, device=dev)
    info_t = torch.empty(NFREQ, dtype=torch.int32, device=dev)
    berr_t = torch.empty((NFREQ, NRHS), dtype=torch.float64, device=dev)
    Xc = synth.closed_form(rom, case.freq, case.F)
    W = work_per_point(N, NRHS)

    def runner(mode, tol, berr):
        def run():
            h.sweep_dev(0, freq_t, F_t, X_t, info_t, berr, mode=mode, tol=tol,
                        max_iter=10, stream=stream.cuda_stream)
        return run

    def check(mode, tol):
        runner(mode, tol, berr_t)()
        torch.cuda.synchronize()
        info = info_t.cpu().numpy()
        be = berr_t.cpu().numpy()
        X = X_t.cpu().numpy().transpose(0, 2, 1)
        fe = np.abs(X - Xc).max(axis=(1, 2)) / np.abs(Xc).max(axis=(1, 2))
        return {"max_backward_error": float(be.max()), "points_failed": int((info != 0).sum()),
                "max_rel_error_vs_closed_form": float(fe.max())}

    if args.copy_calibration > 0:
        nb = args.copy_calibration * 1024 * 1024
        a = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    # launch policy per mode picked by measurement (staccato_amd.autotune): every
    # candidate timed on the full sweep, the fastest whose points all meet tol
    tuned = {}
    fixed = dict(kv.split("=", 1) for kv in args.policy.split(",") if "=" in kv)
    cands = {c[0]: c for c in autotune.CANDIDATES}
    for mode, tol in (("refined", TOL_REFINED), ("plain", TOL_PLAIN)):
        if args.mode not in ("both", mode):
            continue
        if mode in fixed:
            _, kernel, budget, jb = cands[fixed[mode]]
            h.set_policy(N, mode, kernel, budget, jb)
            tuned[mode] = {"chosen": fixed[mode], "given": True}
            continue
        best, cand = autotune.tune(h, 0, case.freq, case.F, mode, tol, sample=NFREQ, reps=2,
                                   cache=False)
        tuned[mode] = {"chosen": best["name"] if best else "automatic",
                       "candidates_pts_per_s": {c["name"]: round(c["pts_per_s"]) for c in cand}}

    res = {}
    for mode, tol in (("refined", TOL_REFINED), ("plain", TOL_PLAIN)):
        if args.mode not in ("both", mode):
            continue
        # refined mode computes its backward error as part of the algorithm;
        # the plain mode is timed as zgetrf + zgetrs only and checked afterwards
        run = runner(mode, tol, berr_t if mode == "refined" else None)
        wall, ev = time_gpu(run, args.steps, args.warmup, stream)
        chk = check(mode, tol)
        res[mode] = {"pts_per_s": NFREQ * args.steps / wall, "ms_per_step": 1e3 * wall / args.steps,
                     "kernel_ms": 1e3 * ev / args.steps, **chk}

    main_mode = "refined" if "refined" in res else "plain"
    r = res[main_mode]
    peak = PEAK_FP32_TFLOPS if main_mode == "refined" else PEAK_FP64_TFLOPS
    achieved = W * NFREQ / (r["kernel_ms"] * 1e-3) / 1e12
    chosen = tuned.get(main_mode, {}).get("chosen", "")
    if (main_mode, chosen) in TRAFFIC_KIB:
        fk, wk = TRAFFIC_KIB[(main_mode, chosen)]
        traffic = (fk * FETCH_SCALE + wk * WRITE_SCALE) * 1024.0
    else:
        traffic = None
    out = {
        "metric": "freq points solved/s at n=256 vs plain FP64 GPU LU (rel. residual <=1e-10)",
        "value": r["pts_per_s"],
        "unit": "freq points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 LU + fp64 refinement" if main_mode == "refined" else "complex128",
        "data": "synthetic (modal ROM, seed 1)",
        "config": {"workload": "uniform batch: one ROM n=256, 4096 frequencies 1-2000 Hz, "
                               "4 RHS, Rayleigh D=1e-4K+0.5M, eta=0.01",
                   "n": N, "nfreq": NFREQ, "nrhs": NRHS, "mode": main_mode,
                   "tol": TOL_REFINED if main_mode == "refined" else TOL_PLAIN},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "traffic_unit": "bytes per launch (HBM, rocprof FETCH_SIZE/WRITE_SIZE) of "
                                     "the launch policy the run uses",
                     "work_per_point_flop": W, "points_per_launch": NFREQ},
        "results": res,
        "autotune": tuned,
    }
    if "plain" in res and "refined" in res:
        out["speedup_vs_plain_fp64"] = res["refined"]["pts_per_s"] / res["plain"]["pts_per_s"]
    # host buffers in and out (stc_sweep: PCIe copies of F and X included); reported, never value
    if "refined" in res:
        h.sweep(0, case.freq, case.F, mode="refined", tol=TOL_REFINED, max_iter=10, berr=True)
        t0 = time.perf_counter()
        for _ in range(3):
            h.sweep(0, case.freq, case.F, mode="refined", tol=TOL_REFINED, max_iter=10, berr=True)
        out["refined_host_buffers_pts_per_s"] = 3 * NFREQ / (time.perf_counter() - t0)
    if not args.no_vendor:
        try:
            v, verr = vendor_fp64(case, max(1, args.steps // 5), 1)
            out["vendor_fp64_pts_per_s"] = v
            out["vendor_fp64"] = "rocsolver_zgetrf/zgetrs_strided_batched on materialised A(w)"
            out["vendor_fp64_rel_error_vs_closed_form"] = verr
        except Exception as e:  # vendor stack unavailable: report, do not fail
            out["vendor_fp64_pts_per_s"] = None
            out["vendor_error"] = repr(e)[:200]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case)
    h.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
