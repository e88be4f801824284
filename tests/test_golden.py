"""Regression fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the FP64 oracle; the reference itself ships no golden vectors).

CPU: the oracle reproduces every fixture (to FP64 summation order).  GPU: libqnd (complex128)
matches them to the parity bar without running the oracle.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as MG  # noqa: E402


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name", list(MG.CASES))
def test_oracle_reproduces_fixture(name):
    ref = load(name)
    got = MG.make(name)
    for k in ref:   # OpenMP reductions may sum in another order on another core count
        assert np.allclose(got[k], ref[k], rtol=1e-12, atol=1e-14), k


@pytest.mark.gpu
def test_gpu_matches_imag_fixture(gpu):
    import tdse_nd_mpi_amd as q
    c, ref = MG.CASES["imag_h1d_256"], load("imag_h1d_256")
    print("case golden imag", flush=True)
    g = q.Grid(c["nr"], c["dr"], "c128", n_slots=c["nstates"])
    g.set_potential(c["pot"])
    g.set_integrator(c["integ"], c["dt"])
    for j in range(c["nstates"]):
        g.upload(j, ref["guess"][j])
        g.imag_steps(j, c["nsteps"], gram_schmidt=True)
    for j in range(c["nstates"]):
        got = g.download(j)
        assert np.linalg.norm(got - ref["psi"][j]) / np.linalg.norm(ref["psi"][j]) < 1e-12
        assert np.allclose(g.energy_moments(j, 2), ref["E"][j], rtol=0, atol=1e-11)
    g.close()


@pytest.mark.gpu
def test_gpu_matches_real_fixture(gpu):
    import tdse_nd_mpi_amd as q
    c, ref = MG.CASES["real_ne2d_32x64"], load("real_ne2d_32x64")
    print("case golden real", flush=True)
    g = q.Grid(c["nr"], c["dr"], "c128", n_slots=1)
    g.set_potential(c["pot"])
    g.set_integrator(c["integ"], c["dt"])
    g.set_mask(0.1)
    g.set_field(ref["Et"], ref["At"], ref["Eobs"], ref["Aobs"])
    g.upload(0, ref["psi0"])
    obs = g.real_steps(0, "length", 0, c["nsteps"], ["dipole_r", "dipole_a", "norm"])
    got = g.download(0)
    assert np.linalg.norm(got - ref["psi"]) / np.linalg.norm(ref["psi"]) < 1e-12
    assert np.allclose(obs, ref["obs"], rtol=1e-11, atol=1e-12)
    g.close()
