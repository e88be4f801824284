"""Write the regression fixtures of tests/golden/ from the FP64 oracle.

The reference ships no golden vectors and cannot be built here (FFTW is
absent), so these fixtures are the oracle's own outputs on small seeded
cases: a regression anchor for the oracle (CPU test) and a fixed target for
the GPU path (GPU test) that does not need the oracle at test time.

    python tests/golden/make_golden.py      # rewrites tests/golden/*.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

CASES = {
    # imaginary time, 2 states with Gram-Schmidt, seeded Gaussian-random guesses
    "imag_h1d_256": dict(nr=(256,), dr=0.2, pot="H_1D", integ="BM4", dt=0.05, nstates=2, nsteps=120, seed=21),
    # real time, length gauge, mask, 2-D, dipole/acceleration/norm every step
    "real_ne2d_32x64": dict(nr=(32, 64), dr=0.4, pot="Ne_2D", integ="BM4", dt=0.05, nsteps=16, seed=22),
}


def field(o, nt, dim):
    w = 45.5637639947958 / 800.0
    modes = dict(freq=[w], amp=[5.33803e-9 * np.sqrt(1e14)], ellip=[0.2], duration=[2 * 2 * np.pi / w],
                 cep=[0.3], delay=[0.0], env=[O.ENV["cos2"]])
    a, asc = o.coefficients()
    time = np.arange(nt) * o.dt
    A, E = O.field_tables(dim, modes, asc, o.s, time)
    Ao, Eo = np.zeros((nt, dim)), np.zeros((nt, dim))
    for i, t in enumerate(time):
        Ao[i], Eo[i] = O.field_vectors(dim, modes, t)
    return E, A, Eo, Ao


def make(name):
    c = CASES[name]
    o = O.Oracle(c["nr"], c["dr"], c["pot"], c["integ"], c["dt"])
    if name.startswith("imag"):
        psi = o.initial_guess(c["nstates"], guess_mode=1, seed=c["seed"])
        guess = psi.copy()
        for j in range(c["nstates"]):
            for _ in range(c["nsteps"]):
                psi[j] = o.step_imag(psi[j])
                rc, psi = o.gram_schmidt(psi, j)
                assert rc == 0
        E = np.array([o.energies(psi[j], 2) for j in range(c["nstates"])])
        return dict(guess=guess, psi=psi, E=E)
    o.set_mask(0.1)
    psi0 = o.initial_guess(1, guess_mode=1, seed=c["seed"])[0]
    E, A, Eo, Ao = field(o, c["nsteps"], len(c["nr"]))
    psi, obs = o.real_run(psi0, "length", E, A, Eo, Ao, ["dipole_r", "dipole_a", "norm"])
    return dict(psi0=psi0, Et=E, At=A, Eobs=Eo, Aobs=Ao, psi=psi, obs=obs)


if __name__ == "__main__":
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **make(name))
        print("wrote", name)
