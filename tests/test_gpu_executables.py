"""The drop-in executables end to end on the GPU, against the FP64 oracle.

QuantumND_imag <params.imag> writes output.imag; QuantumND_real
<params.real> reads it (centred embedding into a larger grid, superposition
with normalised COEFFS_STATES) and writes observables.real, PED.real,
PMD.real and wavefunction.real.  Each file is compared with what the oracle
computes on the CPU from the same inputs (src/QuantumND_imag.c:169-263,
src/QuantumND_real.c:47-264, src/io_functions.c:57-106,136-287,292-499).

The executables run as child processes (subprocess), never by exec from
this process.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import tdse_nd_mpi_amd as q

pytestmark = pytest.mark.gpu

IMAG = os.path.join(q.BIN_DIR, "QuantumND_imag")
REAL = os.path.join(q.BIN_DIR, "QuantumND_real")


def run(exe, args, cwd, timeout=240):
    print(f"case executable {os.path.basename(exe)} {' '.join(args)}", flush=True)
    p = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-3000:]
    return out


def read_output_imag(path):
    """Header keys and the FP64 complex states after the PSI line."""
    raw = open(path, "rb").read()
    k = raw.index(b"\nPSI\n") + len(b"\nPSI\n")
    hdr = {}
    for line in raw[:k].decode().splitlines():
        parts = line.split()
        if parts and parts[0] != "PSI":
            hdr[parts[0]] = parts[1:]
    data = np.frombuffer(raw[k:], dtype=np.complex128)
    return hdr, data


def read_binary_after(path, marker):
    raw = open(path, "rb").read()
    k = raw.index(marker) + len(marker)
    hdr = {}
    for line in raw[:k].decode().splitlines():
        parts = line.split()
        if parts:
            hdr[parts[0]] = parts[1:]
    return hdr, np.frombuffer(raw[k:], dtype=np.float64)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def check_phase_convention(psi):
    """Q7: the phase at the first strict maximum of |psi| is removed.  Near-ties
    (symmetric states, maxima equal to ~1e-12) may move the maximum by the
    rounding of the rotation itself, so any point within 1e-9 of the maximum
    may carry the zero phase."""
    a = np.abs(psi)
    near = np.flatnonzero(a >= a.max() * (1 - 1e-9))
    ok = [abs(psi[i].imag) <= 1e-12 * a[i] and psi[i].real > 0 for i in near]
    assert any(ok), psi[near]


def same_up_to_phase(got, ref, tol):
    """got = e^{i phi} ref within tol (rel L2).  The phase convention picks the
    first maximum of |psi|; a symmetric eigenstate has two maxima equal to
    rounding, so the two programs may pick different ones."""
    ph = np.vdot(ref, got)
    ph /= abs(ph)
    return rel(got, ph * ref) < tol


IMAG_1D = """OUTPUT_IMAG \toutput.imag

DIMENSION\t1
N_STATES \t4

POTENTIAL\tH_1D

DR\t\t.1
NR\t\t4096

INTEGRATOR \tBM4
DT \t\t.1
TINTEGRATION \t100.

GUESS\tgaussian
SEED\t20261204
"""


def test_imag_executable_1d_four_states(tmp_path):
    """The benchmark's first configuration through the executable: H_1D 4096,
    BM4, DT 0.1, T 100 (1001 steps per state), four seeded states with
    Gram-Schmidt after every step.  Energies within 1e-10 Ha of the oracle,
    states within 1e-12 relative L2, the global phase convention kept."""
    (tmp_path / "params.imag").write_text(IMAG_1D)
    out = run(IMAG, ["params.imag"], tmp_path)
    hdr, data = read_output_imag(tmp_path / "output.imag")
    assert hdr["DIMENSION"] == ["1"] and hdr["N_STATES"] == ["4"] and hdr["NR"] == ["4096"]
    assert hdr["POTENTIAL"] == ["H_1D"] and abs(float(hdr["DR"][0]) - 0.1) < 1e-15
    assert not (tmp_path / "output.imag.softcore").exists()
    o = O.Oracle((4096,), 0.1, "H_1D", "BM4", 0.1)
    psi, E, _ = o.imag_run(4, 100.0, guess_mode=1, seed=20261204)
    E_file = np.array([float(v) for v in hdr["E"]])
    DE_file = np.array([float(v) for v in hdr["DE"]])
    assert np.allclose(E_file, E[:, 0], rtol=0, atol=1e-10), (E_file, E[:, 0])
    assert np.allclose(DE_file, E[:, 1] - E[:, 0] ** 2, rtol=0, atol=1e-10)
    assert abs(E_file[0] + 0.5) < 1e-3        # the soft-core a^2 = 2 design value
    states = data.reshape(4, 4096)
    for j in range(4):
        check_phase_convention(states[j])
        assert same_up_to_phase(states[j], psi[j], 1e-12), (j, rel(states[j], psi[j]))
    gram = states.conj() @ states.T * 0.1
    assert np.allclose(gram, np.eye(4), atol=1e-12)
    assert "eigenfunction 4" in out and "The code ended successfully" in out


IMAG_2D = """OUTPUT_IMAG \toutput.imag

DIMENSION\t2
N_STATES \t2

POTENTIAL\tH_2D

DR\t\t.4\t.4
NR\t\t32\t16

INTEGRATOR \tBM4
DT \t\t.1
TINTEGRATION \t20.

GUESS\tgaussian
SEED\t77
SOFTCORE\t1.
"""

REAL_2D = """OUTPUT_IMAG\toutput.imag

NR \t\t64\t32

N_STATES \t2
J_STATES \t1 \t2
COEFFS_STATES\t1.\t.5+.5j

CAP \t\t.1\t4.

GAUGE \t\t{gauge}
N_MODES \t2
INTENSITY \t1.e14 \t1.e13
LAMBDA \t\t800 400
ELLIPTICITY \t0. \t.3
CEP \t\t0. \t1.
DELAY \t\t0. \t2.
ENVELOPE \tcos2 \ttrapz
DURATION \t40.\t30.

INTEGRATOR \tBM4
DT \t\t.05
TINTEGRATION \t3.

DT_SAVE \t1.
N_OBSERVABLES \t8
OBSERVABLES \tdipole_r \tPMD\tdipole_v \tenergy \tPED \tLz \twavefunction \tnorm

P_MAX \t\t2.\t1.5
R_MASK \t\t3. \t6.
DE \t\t.02
NE \t\t60
STEPS_PER_CALL\t16
"""


def embed(states, nr_real, coeffs):
    """Read_wavefunction (src/io_functions.c:292-459): each state centred in
    the larger grid at offset (NR_real - NR_imag) / 2, superposed with the
    normalised coefficients."""
    c = np.asarray(coeffs, dtype=complex)
    c = c / np.sqrt(np.sum(np.abs(c) ** 2))
    out = np.zeros(nr_real, dtype=complex)
    ni = states.shape[1:]
    sl = tuple(slice((a - b) // 2, (a - b) // 2 + b) for a, b in zip(nr_real, ni))
    for k in range(len(c)):
        out[sl] += c[k] * states[k]
    return out


# real-time precision of the chain and its tolerances against the FP64
# oracle: observables (of the largest value), final state (rel. L2), PED
# (relative), PMD phase (radians, where the density is above `big` of its
# maximum: a complex64 amplitude carries an absolute error, so its phase is
# compared only where the amplitude is large)
PREC_TOL = {"c128": dict(obs=5e-12, wf=1e-12, ped=1e-10, phase=1e-8, big=1e-10),
            "c64f64": dict(obs=2e-5, wf=1e-5, ped=1e-4, phase=1e-3, big=1e-4)}


@pytest.mark.parametrize("gauge,prec", [("length", "c128"), ("velocity", "c128"), ("length", "c64f64")])
def test_imag_then_real_executables_2d(tmp_path, gauge, prec):
    """A 2-D chain: QuantumND_imag (32 x 16, two seeded states, soft-core
    extension in its side file) then QuantumND_real on a larger 64 x 32 grid
    (centred embedding, a complex superposition), CAP, a two-mode field with
    two envelopes, every observable kind recorded every step, then PED, PMD
    and the final wavefunction.  Every file against the oracle.  The real-time
    leg runs in complex128 or, through the PRECISION key, in c64f64."""
    tol = PREC_TOL[prec]
    (tmp_path / "params.imag").write_text(IMAG_2D)
    run(IMAG, ["params.imag"], tmp_path)
    hdr, data = read_output_imag(tmp_path / "output.imag")
    assert "SOFTCORE" not in hdr      # the reference reader stops at unknown keys
    assert (tmp_path / "output.imag.softcore").read_text().split() == ["SOFTCORE", "1.000000000000000"]
    oi = O.Oracle((32, 16), 0.4, "H_2D", "BM4", 0.1, softcore=1.0)
    psi_i, E_i, _ = oi.imag_run(2, 20.0, guess_mode=1, seed=77)
    states = data.reshape(2, 32, 16)
    for j in range(2):
        check_phase_convention(states[j].ravel())
        assert same_up_to_phase(states[j], psi_i[j], 1e-12), (j, rel(states[j], psi_i[j]))
    assert np.allclose([float(v) for v in hdr["E"]], E_i[:, 0], rtol=0, atol=1e-10)

    (tmp_path / "params.real").write_text(REAL_2D.format(gauge=gauge) + f"PRECISION\t{prec}\n")
    out = run(REAL, ["params.real"], tmp_path)
    nr = (64, 32)
    dr = 0.4
    # the oracle's run starts from the states the imag executable wrote
    psi0 = embed(states, nr, [1.0, 0.5 + 0.5j])
    o = O.Oracle(nr, dr, "H_2D", "BM4", 0.05, cap=(0.1, 4.0), softcore=1.0)
    nrm = np.sqrt(np.sum(np.abs(psi0) ** 2) * dr * dr)
    E2 = o.energies(psi0, 2)
    m = re.search(r"norm\t+([-0-9.e]+)\n\s*E\t+([-0-9.e]+)", out)
    assert m, out[-2000:]
    etol = 1e-12 if prec == "c128" else 1e-6
    assert abs(float(m.group(1)) - nrm) < etol and abs(float(m.group(2)) - E2[0]) < etol
    # field tables at the stage times (src/integrators.c:74-129), field at t_n
    dt, T = 0.05, 3.0
    nt = int(T / dt)
    s, _, asc = O.integrator_coefficients("BM4", dt)
    lam = np.array([800.0, 400.0])
    modes = dict(freq=45.5637639947958 / lam, amp=5.33803e-9 * np.sqrt([1e14, 1e13]), ellip=[0.0, 0.3],
                 duration=[40.0, 30.0], cep=[0.0, 1.0], delay=[0.0, 2.0], env=[O.ENV["cos2"], O.ENV["trapz"]])
    time = np.arange(nt) * dt
    At, Et = O.field_tables(2, modes, asc, s, time)
    Ao = np.zeros((nt, 2))
    Eo = np.zeros((nt, 2))
    for i, t in enumerate(time):
        Ao[i], Eo[i] = O.field_vectors(2, modes, t)
    names = ["dipole_r", "dipole_v", "energy", "Lz", "norm"]
    psi_T, rec = o.real_run(psi0, gauge, Et, At, Eo, Ao, names)
    # observables.real: time, then the values in the order of OBSERVABLES
    obs = np.loadtxt(tmp_path / "observables.real", skiprows=1)
    assert obs.shape == (nt, 1 + rec.shape[1]), obs.shape
    assert np.allclose(obs[:, 0], time, atol=1e-15)
    assert np.allclose(obs[:, 1:], rec, rtol=0, atol=tol["obs"] * max(1.0, np.abs(rec).max())), \
        np.abs(obs[:, 1:] - rec).max()
    # wavefunction.real: the final state before the PMD mask
    whdr, wv = read_binary_after(tmp_path / "wavefunction.real", b"PSI_BINARY_START\n")
    assert whdr["DIMENSION"] == ["2"] and whdr["NR"] == ["64", "32"]
    wf = wv.view(np.complex128).reshape(nr)
    assert rel(wf, psi_T) < tol["wf"], rel(wf, psi_T)
    # PED.real and PMD.real
    Rm, pm, Ne, de = (3.0, 6.0), (2.0, 1.5), 60, 0.02
    psi_p, ped, pmd, npr = o.final_observables(psi_T, Rm, pm, Ne, de)
    P = np.loadtxt(tmp_path / "PED.real")
    assert P.shape == (Ne, 2) and np.allclose(P[:, 0], np.arange(Ne) * de, atol=1e-15)
    assert np.allclose(P[:, 1], ped, rtol=tol["ped"], atol=1e-4 * tol["ped"] * max(ped.max(), 1e-300))
    phdr, pv = read_binary_after(tmp_path / "PMD.real", b"DATA_START\n")
    assert phdr["DIMENSION"] == ["2"] and [int(v) for v in phdr["NPR"]] == list(npr)
    assert np.allclose([float(v) for v in phdr["DP"]], [2 * np.pi / (dr * n) for n in nr], rtol=1e-15)
    got = pv.reshape(-1, 2)
    assert got.shape == pmd.shape
    dens = pmd[:, 0]
    assert np.allclose(got[:, 0], dens, rtol=tol["ped"], atol=1e-4 * tol["ped"] * dens.max())
    big = dens > tol["big"] * dens.max()
    dphi = np.angle(np.exp(1j * (got[big, 1] - pmd[big, 1])))
    assert np.abs(dphi).max() < tol["phase"], np.abs(dphi).max()
