"""The allocation-free entry point qnd_observables_noalloc against
qnd_observables (the entry point whose kernels it shares), through the C ABI:
two calls in a row on one GPU, the second on a smaller grid (the device
scratch keeps the first call's contents), both precisions."""
import numpy as np
import pytest

import tdse_nd_mpi_amd as q

pytestmark = pytest.mark.gpu

NAMES = ["norm", "dipole_r", "dipole_a"]


def _case(torch, nr, dr, prec, pot, seed, Ef):
    rng = np.random.default_rng(seed)
    x = [(-0.5 * n + np.arange(n)) * dr for n in nr]
    env = np.exp(-sum(np.meshgrid(*[xi ** 2 for xi in x], indexing="ij")) / 30.0)
    psi = env * (rng.standard_normal(nr) + 1j * rng.standard_normal(nr))
    g = q.Grid(nr, dr, prec, n_slots=1)
    g.set_potential(pot)
    g.set_integrator("BM4", 0.05)
    # the existing entry point on the same state
    g.upload(0, psi)
    ref = g.observables(0, NAMES, Efield=Ef)
    # the state the library stores (complex64 rounds on upload), on the device
    dt = torch.complex64 if prec == "c64" else torch.complex128
    dev = torch.from_numpy(psi.astype(np.complex64 if prec == "c64" else np.complex128)).to("cuda")
    out = torch.full((len(ref),), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.observables_noalloc(dev.data_ptr(), NAMES, out.data_ptr(), Efield=Ef)
    g.synchronize()
    got = out.cpu().numpy()
    g.close()
    return got, ref


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_noalloc_matches_observables_two_calls(gpu, prec):
    import torch
    print(f"case noalloc {prec}: 256x512 then 64x32", flush=True)
    got1, ref1 = _case(torch, (256, 512), 0.3, prec, "Ne_2D", 1, [0.03, -0.01])
    assert np.all(np.isfinite(got1))
    assert np.allclose(got1, ref1, rtol=1e-13, atol=1e-16), (got1, ref1)
    # second call: fewer partial rows than the first left in the scratch
    got2, ref2 = _case(torch, (64, 32), 0.4, prec, "H_2D", 2, [-0.02, 0.05])
    assert np.all(np.isfinite(got2))
    assert np.allclose(got2, ref2, rtol=1e-13, atol=1e-16), (got2, ref2)


def test_noalloc_rejects_unsupported_codes(gpu):
    import torch
    g = q.Grid((16, 16), 0.5, "c128", n_slots=1)
    g.set_potential("H_2D")
    dev = torch.zeros((16, 16), dtype=torch.complex128, device="cuda")
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(q.QNDError, match="norm, dipole_r and dipole_a only"):
        g.observables_noalloc(dev.data_ptr(), ["energy"], out.data_ptr())
    g.close()
