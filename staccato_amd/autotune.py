"""Autotuner of the sweep kernel's launch policy, by measurement on the device.

The blocked factorisation (systems that do not fit the LDS whole) can run as
  - the 16-column register panel (kernel 1: three workgroups per CU for
    complex64, two for complex128), or
  - the wide LDS panel (kernel 2) with 160 KiB of LDS per workgroup (one per
    CU, panels up to 64 columns) or 80 KiB (two per CU, up to 32), with the
    panel width capped or not,
and which is fastest depends on the order n, the right-hand sides and the
mode.  tune() times every candidate on a sample of the caller's frequencies
(inputs and outputs resident on the device, HIP events on the launch stream),
keeps only candidates whose every point meets the tolerance, installs the
fastest with Handle.set_policy and records it in a JSON cache keyed by
(arch, n, nrhs, mode).  apply_cached() installs recorded policies without
measuring.  The reference has no counterpart: its only tuning knob is a fixed
batch size (test/krylovMidsizeTest/gpu/concurrent/batchedLU/src/config/config.cu:29-30).
"""
import json
import os

import numpy as np

ARCH = "gfx950"

# (name, kernel, lds_budget bytes, jb_max)
CANDIDATES = [
    ("register-panel", 1, 0, 0),
    # the register-panel kernel at two workgroups per CU (80 KB of LDS each):
    # systems whose whole matrix fits then stay LDS-resident (complex64 n <= 96).
    # One workgroup per CU (n = 128 complex64 whole in LDS) measured 4x slower
    # than three register-panel workgroups per CU and is not a candidate.
    ("lds-2perCU", 1, 80 * 1024, 0),
    ("wide-2perCU", 2, 80 * 1024, 0),
    ("wide-2perCU-jb16", 2, 80 * 1024, 16),
    ("wide-1perCU", 2, 160 * 1024 - 256, 0),
    ("wide-1perCU-jb32", 2, 160 * 1024 - 256, 32),
    # two-level update: several LDS panels per 64-column step (32 complex128),
    # one deferred update of the columns right of the step
    ("wide-2level-2perCU", 4, 80 * 1024, 0),
    ("wide-2level-1perCU", 4, 160 * 1024 - 256, 0),
    # Gauss 3M complex products in the trailing update (complex64 only)
    ("wide-3M-2perCU", 3, 80 * 1024, 0),
    ("wide-3M-1perCU", 3, 160 * 1024 - 256, 0),
]


def candidates_for(mode):
    """The candidates that differ in `mode` (the 3M kernel exists for complex64 only)."""
    complex64 = mode in ("refined", "fp32", 1)
    return [c for c in CANDIDATES if complex64 or c[1] != 3]


def default_cache_path():
    return os.environ.get("STC_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "staccato_amd", "tune.json"))


def _key(n, nrhs, mode):
    return f"{ARCH}/n={int(n)}/nrhs={int(nrhs)}/{mode}"


def load_cache(path=None):
    path = path or default_cache_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=None):
    path = path or default_cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply_cached(handle, n, nrhs, mode, cache=None):
    """Install the recorded policy for (n, nrhs, mode); True when one exists."""
    cache = load_cache() if cache is None else cache
    e = cache.get(_key(n, nrhs, mode))
    if not e:
        return False
    handle.set_policy(n, mode, e["kernel"], e["lds_budget"], e["jb_max"])
    return True


# The automatic policy of the library for an order n and mode (kept unless a
# candidate is faster by more than `margin`: near-ties are noise).
def automatic_name(n, mode):
    complex64 = mode in ("refined", "fp32", 1)
    if n <= 128:
        return "register-panel"
    one_per_cu = n > (512 if complex64 else 256)
    if n > 256:  # the two-level update (complex128 beyond n = 512: register panel either way)
        return "wide-2level-1perCU" if one_per_cu else "wide-2level-2perCU"
    if complex64:  # Gauss 3M products up to n = 256
        return "wide-3M-2perCU"
    return "wide-1perCU" if one_per_cu else "wide-2perCU"


def tune(handle, rom_id, freq_hz, F, mode, tol, sample=1024, reps=3, candidates=None,
         cache=None, cache_path=None, verbose=False, margin=0.02):
    """Time each candidate policy on `sample` of the frequencies (evenly
    spaced) with load F (n x nrhs), install the fastest that meets tol at every
    point (the automatic policy unless another is faster by more than margin),
    record it (cache_path, or the default cache; cache=False: do not record)
    and return (best_entry, all_results)."""
    import torch
    n = handle.n[rom_id]
    F = np.asarray(F, dtype=np.complex128)
    if F.ndim == 1:
        F = F[:, None]
    nrhs = F.shape[1]
    freq = np.asarray(freq_hz, dtype=np.float64)
    if freq.size > sample:
        freq = freq[np.linspace(0, freq.size - 1, sample).round().astype(int)]
    dev = torch.device("cuda", handle.device)
    stream = torch.cuda.Stream(device=dev)
    freq_t = torch.from_numpy(np.ascontiguousarray(freq)).to(dev)
    F_t = torch.from_numpy(np.ascontiguousarray(F.T)).to(dev)
    X_t = torch.empty((freq.size, nrhs, n), dtype=torch.complex128, device=dev)
    info_t = torch.empty(freq.size, dtype=torch.int32, device=dev)
    be_t = torch.empty((freq.size, nrhs), dtype=torch.float64, device=dev)
    results = []
    for name, kernel, budget, jb in (candidates or candidates_for(mode)):
        handle.set_policy(n, mode, kernel, budget, jb)
        with torch.cuda.stream(stream):
            handle.sweep_dev(rom_id, freq_t, F_t, X_t, info_t, be_t, mode=mode, tol=tol,
                             stream=stream.cuda_stream)  # warm-up (and first-call set-up)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                handle.sweep_dev(rom_id, freq_t, F_t, X_t, info_t, be_t, mode=mode, tol=tol,
                                 stream=stream.cuda_stream)
            e1.record(stream)
        torch.cuda.synchronize(dev)
        ms = e0.elapsed_time(e1) / reps
        ok = bool((info_t == 0).all().item()) and float(be_t.max().item()) <= tol
        r = {"name": name, "kernel": kernel, "lds_budget": budget, "jb_max": jb,
             "pts_per_s": freq.size / (ms * 1e-3), "ok": ok}
        results.append(r)
        if verbose:
            print(f"autotune n={n} nrhs={nrhs} {mode}: {name:18s} {r['pts_per_s']:12.0f} pts/s "
                  f"{'ok' if ok else 'FAILS tol'}", flush=True)
    good = [r for r in results if r["ok"]]
    if not good:
        handle.set_policy(n, mode, None)
        return None, results
    best = max(good, key=lambda r: r["pts_per_s"])
    auto = [r for r in good if r["name"] == automatic_name(n, mode)]
    if auto and best["pts_per_s"] <= auto[0]["pts_per_s"] * (1.0 + margin):
        best = auto[0]
    handle.set_policy(n, mode, best["kernel"], best["lds_budget"], best["jb_max"])
    if cache is not False:
        c = load_cache(cache_path) if cache is None else cache
        c[_key(n, nrhs, mode)] = {k: best[k] for k in ("name", "kernel", "lds_budget", "jb_max", "pts_per_s")}
        save_cache(c, cache_path)
    return best, results
