"""Every launch policy the autotuner chooses among gives the same answers: the
oracle's zgetf2 pivots and factors for the plain mode, and the oracle's
solutions within the gates for both modes, on blocked systems of several
orders; the autotuner installs a policy that meets tol and records it."""
import numpy as np
import pytest

import oracle
from staccato_amd import autotune, synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n", [100, 256, 300, 520])
def test_every_policy_matches_oracle(handle, n):
    rom, _ = synth.modal_rom(n, 500 + n, 0.02, alpha=1e-4, beta=0.5)
    handle.rom_set(3, rom.K, rom.M, rom.D, rom.eta)
    freq = np.array([2.0, 333.0, 1500.0])
    rng = np.random.default_rng(n)
    F = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    Xo, io, _, _ = oracle.sweep(rom.K, rom.M, rom.D, rom.eta, freq, F, nthreads=8)
    assert (io == 0).all()
    try:
        for name, kernel, budget, jb in autotune.CANDIDATES:
            for mode, tol in (("plain", 1e-12), ("refined", 1e-10)):
                print(f"policy {name} n={n} {mode}", flush=True)
                handle.set_policy(n, mode, kernel, budget, jb)
                X, info, be, _ = handle.sweep(3, freq, F, mode=mode, tol=tol)
                assert (info == 0).all(), (name, mode, info)
                assert be.max() <= tol, (name, mode, be.max())
                assert _rel(X, Xo) < (1e-10 if mode == "plain" else 1e-5), (name, mode, _rel(X, Xo))
            if n <= 300:
                LU, ipiv, info = handle.getrf(3, freq[:2], mode="plain")
                for j in range(2):
                    LUo, ipo, _ = oracle.zgetrf(synth.assemble(rom, freq[j]))
                    np.testing.assert_array_equal(ipiv[j], ipo)
                    assert _rel(LU[j], LUo) < 1e-11 * max(1, n / 64)
    finally:
        handle.clear_policies()


def test_autotune_installs_and_records(handle, tmp_path):
    rom, _ = synth.modal_rom(200, 77, 0.02, alpha=1e-4, beta=0.5)
    handle.rom_set(4, rom.K, rom.M, rom.D, rom.eta)
    rng = np.random.default_rng(3)
    F = rng.standard_normal((200, 2)) + 0j
    freq = np.linspace(1.0, 1500.0, 600)
    path = str(tmp_path / "tune.json")
    try:
        best, res = autotune.tune(handle, 4, freq, F, "refined", 1e-10, sample=256, reps=1,
                                  cache_path=path)
        assert best is not None and best["ok"]
        assert len(res) == len(autotune.CANDIDATES) and all(r["ok"] for r in res)
        cache = autotune.load_cache(path)
        e = cache["gfx950/n=200/nrhs=2/refined"]
        assert e["name"] == best["name"]
        handle.clear_policies()
        assert autotune.apply_cached(handle, 200, 2, "refined", cache)
        X, info, be, _ = handle.sweep(4, freq[:5], F, mode="refined", tol=1e-10)
        assert (info == 0).all() and be.max() <= 1e-10
    finally:
        handle.clear_policies()
