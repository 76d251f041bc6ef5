"""stc_sweep_dev_noalloc (fixed per-GPU scratch, no allocation in the call)
against stc_sweep_dev on the same inputs: identical X, info and backward
errors, on two calls in a row (the second on a smaller input), in both modes,
for a blocked system (factor slots in the scratch array) and a small one (LDS
only).  Runs alone in a fresh process or after the other tests."""
import numpy as np
import pytest

from staccato_amd import Handle, synth

pytestmark = pytest.mark.gpu


def _run(h, rom_id, n, freq, F, mode, tol, noalloc):
    import torch
    dev = torch.device("cuda", 0)
    nf, nrhs = freq.size, F.shape[1]
    freq_t = torch.from_numpy(np.ascontiguousarray(freq)).to(dev)
    F_t = torch.from_numpy(np.ascontiguousarray(F.T)).to(dev)
    X_t = torch.full((nf, nrhs, n), complex("nan"), dtype=torch.complex128, device=dev)
    info_t = torch.full((nf,), 7, dtype=torch.int32, device=dev)
    be_t = torch.full((nf, nrhs), -1.0, dtype=torch.float64, device=dev)
    h.sweep_dev(rom_id, freq_t, F_t, X_t, info_t, be_t, mode=mode, tol=tol, noalloc=noalloc)
    torch.cuda.synchronize()
    return X_t.cpu().numpy(), info_t.cpu().numpy(), be_t.cpu().numpy()


@pytest.mark.parametrize("mode,tol", [("plain", 1e-12), ("refined", 1e-10)])
def test_noalloc_matches_sweep_dev_two_calls(mode, tol):
    big = synth.uniform_n256()
    small = synth.single_n60()
    rng = np.random.default_rng(11)
    F_small = (rng.standard_normal((60, 3)) + 1j * rng.standard_normal((60, 3)))
    with Handle(0) as h:
        assert h.device_ok
        h.rom_set(0, big.rom.K, big.rom.M, big.rom.D, big.rom.eta)
        h.rom_set(1, small.rom.K, small.rom.M, small.rom.D, small.rom.eta)
        calls = [
            (0, 256, big.freq[::8], big.F),          # 512 points, blocked path, slot reuse
            (0, 256, big.freq[3::97][:40], big.F[:, :2]),  # smaller input on the same ROM
            (1, 60, small.freq[::25], F_small),      # small path: only the queue head
        ]
        for rom_id, n, freq, F in calls:
            print(f"noalloc vs sweep_dev: mode={mode} n={n} nfreq={freq.size} nrhs={F.shape[1]}",
                  flush=True)
            Xr, ir, br = _run(h, rom_id, n, freq, F, mode, tol, noalloc=False)
            Xn, inn, bn = _run(h, rom_id, n, freq, F, mode, tol, noalloc=True)
            assert (ir == 0).all() and (inn == 0).all()
            assert np.nanmax(bn) <= tol
            # same kernels, same layout, one workgroup per point: bit-identical
            np.testing.assert_array_equal(Xn, Xr)
            np.testing.assert_array_equal(bn, br)
