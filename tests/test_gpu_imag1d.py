"""1-D imaginary time in one launch (k_imag1d: one workgroup keeps the line
for every step of a call) against the oracle and against the form with one
launch per pass (QND_IMAG1D=0), including Gram-Schmidt against previous
states and its collapse error (src/integrators.c:135-163,
src/QuantumND_imag.c:182-185, 265-311)."""
import numpy as np
import pytest

import oracle_lib as O
import tdse_nd_mpi_amd as q

from test_gpu_parity import rel

pytestmark = pytest.mark.gpu


def _relax(monkeypatch, one_launch, nr, dr, pot, integ, dt, prec, guesses, nsteps):
    monkeypatch.setenv("QND_IMAG1D", "1" if one_launch else "0")
    g = q.Grid(nr, dr, prec, n_slots=len(guesses))
    g.set_potential(pot)
    g.set_integrator(integ, dt)
    for j, psi in enumerate(guesses):
        g.upload(j, psi)
        g.imag_steps(j, nsteps)
    out = np.array([g.download(j) for j in range(len(guesses))])
    E = np.array([g.energy_moments(j, 2) for j in range(len(guesses))])
    g.close()
    return out, E


@pytest.mark.parametrize("n,dr,prec,integ", [(4096, 0.1, "c128", "BM4"), (2048, 0.1, "c64", "BM4"),
                                             (512, 0.1, "c128", "BM6"), (256, 0.3, "c64", "RKN4"),
                                             (1024, 0.2, "c128", "Verlet")])
def test_one_launch_matches_passes_and_oracle(gpu, monkeypatch, n, dr, prec, integ):
    """Three states with Gram-Schmidt after every step: the one-launch form
    equals the per-pass form (same operators, reductions in another order)
    and the FP64 oracle (complex128 to 1e-12; complex64 within its rounding,
    no worse than the per-pass form)."""
    print(f"case imag1d one launch n={n} dr={dr} {prec} {integ}", flush=True)
    dt, nsteps = 0.1, 40
    o = O.Oracle((n,), dr, "H_1D", integ, dt)
    guesses = o.initial_guess(3, guess_mode=1, seed=11)
    one, E1 = _relax(monkeypatch, True, (n,), dr, "H_1D", integ, dt, prec, guesses, nsteps)
    per, E2 = _relax(monkeypatch, False, (n,), dr, "H_1D", integ, dt, prec, guesses, nsteps)
    ref = guesses.copy()
    for j in range(3):
        for _ in range(nsteps):
            ref[j] = o.step_imag(ref[j])
            rc, ref = o.gram_schmidt(ref, j)
            assert rc == 0
    e_one = [rel(one[j], ref[j]) for j in range(3)]
    e_per = [rel(per[j], ref[j]) for j in range(3)]
    d = [rel(one[j], per[j]) for j in range(3)]
    print(f"  vs oracle: one launch {e_one}, per pass {e_per}; between {d}", flush=True)
    if prec == "c128":
        assert max(d) < 1e-13 and max(e_one) < 1e-12, (d, e_one)
        assert np.allclose(E1, E2, rtol=0, atol=1e-12), (E1, E2)
    else:
        # two complex64 rounding sequences of the same steps: each within
        # the complex64 bound of the oracle, the one-launch form no worse
        assert max(e_one) < 5e-5 and max(e_one) <= 2 * max(e_per) + 1e-6, (e_one, e_per)
        assert np.allclose(E1, E2, rtol=0, atol=1e-5), (E1, E2)
    gram = one.conj() @ one.T * dr
    assert np.allclose(gram, np.eye(3), atol=1e-12 if prec == "c128" else 1e-6)


def test_one_launch_without_gram_schmidt(gpu, monkeypatch):
    """Steps alone (no Gram-Schmidt, no normalisation) in both forms."""
    print("case imag1d one launch without Gram-Schmidt", flush=True)
    n, dr, dt = 1024, 0.1, 0.05
    o = O.Oracle((n,), dr, "H_1D", "BM4", dt)
    psi = o.initial_guess(1, guess_mode=1, seed=3)[0]
    res = []
    for one_launch in (True, False):
        monkeypatch.setenv("QND_IMAG1D", "1" if one_launch else "0")
        g = q.Grid((n,), dr, "c128", n_slots=1)
        g.set_potential("H_1D")
        g.set_integrator("BM4", dt)
        g.upload(0, psi)
        g.imag_steps(0, 25, gram_schmidt=False)
        res.append(g.download(0))
        g.close()
    ref = psi.copy()
    for _ in range(25):
        ref = o.step_imag(ref)
    assert rel(res[0], res[1]) < 1e-13 and rel(res[0], ref) < 1e-12


def test_one_launch_collapse_error(gpu, monkeypatch):
    """A state with nothing left after the projection (here zero) collapses
    at the first step: the call reports that step (the reference aborts
    there)."""
    print("case imag1d one launch collapse", flush=True)
    monkeypatch.setenv("QND_IMAG1D", "1")
    n, dr = 512, 0.1
    o = O.Oracle((n,), dr, "H_1D", "BM4", 0.1)
    s = o.initial_guess(1, guess_mode=1, seed=5)[0]
    g = q.Grid((n,), dr, "c128", n_slots=2)
    g.set_potential("H_1D")
    g.set_integrator("BM4", 0.1)
    g.upload(0, s)
    g.imag_steps(0, 5)
    g.upload(1, np.zeros(n, dtype=complex))
    with pytest.raises(q.QNDError, match=r"norm decayed to zero during Gram-Schmidt for state 1 \(step 0 "):
        g.imag_steps(1, 100)
    g.close()
