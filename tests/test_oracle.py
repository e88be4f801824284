"""Pins of the FP64 CPU oracle (oracle/qnd_oracle.c) against checks that do
not restate its code: numpy's FFT, a brute-force DFT, dense-matrix
Hamiltonians, closed-form dynamics (free packet, harmonic oscillator),
observed convergence orders, Ehrenfest's theorem, gauge equivalence and the
literal constants of the reference (src/integrators.c:22-56,
src/meshgrid.c:140-153, src/parameters.c:807-810).
"""
import numpy as np
import pytest

import oracle_lib as O


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(1,), (2,), (8,), (64,), (384,), (4096,), (12, 6), (16, 32), (8, 4, 16),
                                   (3, 5, 7), (2, 2, 2, 8)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert rel(O.fftnd(x, -1), np.fft.fftn(x)) < 1e-15 * 8 * np.log2(max(x.size, 2))
    assert rel(O.fftnd(x, +1), np.fft.ifftn(x) * x.size) < 1e-15 * 8 * np.log2(max(x.size, 2))


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 16, 24, 64])
def test_fft_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    F = np.exp(-2j * np.pi * (np.outer(k, k) % n) / n)
    assert np.abs(O.fftnd(x, -1) - F @ x).max() < 2e-15 * n
    assert np.abs(O.fftnd(x, +1) - F.conj() @ x).max() < 2e-15 * n


def test_fft_round_trip_and_parseval():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((32, 64)) + 1j * rng.standard_normal((32, 64))
    X = O.fftnd(x, -1)
    assert rel(O.fftnd(X, +1) / x.size, x) < 1e-15
    assert abs(np.sum(np.abs(X) ** 2) / x.size - np.sum(np.abs(x) ** 2)) / np.sum(np.abs(x) ** 2) < 1e-14


# ---------------------------------------------------------------- integrators
def test_integrator_literals_and_symmetry():
    lit = {"Verlet": [.5], "BM4": [.0792036964311957, .1303114101821663, .2228614958676077,
                                  -.3667132690474257, .3246481886897062, .1096884778767498],
           "BM6": [.050262764400392, .098553683500650, .314960616927694, -.447346482695478,
                   .492426372489876, -.425118767797691, .237063913978122, .195602488600053,
                   .346358189850727, -.362762779254345],
           "RKN4": [.082984406417405, .162314550766866, .233995250731502, .370877414979578,
                    -.409933719901926, .059762097006575]}
    for name, first in lit.items():
        s, a, asc = O.integrator_coefficients(name, 0.1)
        assert s == len(first)
        assert np.array_equal(a[:s], first)
        assert np.array_equal(a[s:2 * s], first[::-1])
        assert abs(a[:s].sum() - 0.5) < 1e-12 and abs(a[:2 * s].sum() - 1.0) < 1e-12
        assert np.allclose(asc[:2 * s], a[:2 * s] * 0.1, rtol=0, atol=0)
    assert O.integrator_coefficients("Euler", 0.1)[0] == -1


def free_packet(nr, dr, x0=0.0, k0=0.0, w=1.0):
    dim = len(nr)
    axes = [-0.5 * n * d + np.arange(n) * d for n, d in zip(nr, dr)]
    X = np.meshgrid(*axes, indexing="ij")
    r2 = (X[0] - x0) ** 2 + sum(x ** 2 for x in X[1:])
    psi = np.exp(-r2 / (2 * w ** 2)) * np.exp(1j * k0 * X[0])
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * np.prod(dr))
    return psi, X


def kinetic_grid(nr, dr):
    ps = [2 * np.pi * np.fft.fftfreq(n, d) for n, d in zip(nr, dr)]
    P = np.meshgrid(*ps, indexing="ij")
    return 0.5 * sum(p ** 2 for p in P), P


@pytest.mark.parametrize("integ", ["Verlet", "BM4", "BM6", "RKN4"])
def test_free_steps_closed_form(integ):
    nr, dr, dt = (64, 32), (0.4, 0.5), 0.07
    o = O.Oracle(nr, dr, "zero_test", integ, dt)
    psi, _ = free_packet(nr, dr, x0=1.0, k0=0.5)
    K, P = kinetic_grid(nr, dr)
    # length gauge, E = 0: psi_p -> exp(-i K dt) psi_p exactly
    s = o.s
    got = o.step_length(np.fft.fftn(psi), np.zeros((s, 2)))
    assert rel(got, np.exp(-1j * K * dt) * np.fft.fftn(psi)) < 1e-13
    # velocity gauge, constant A: kinetic phase exp(-i (K + (p+A/2).A) dt)
    A = np.array([0.3, -0.2])
    W = sum((P[j] + 0.5 * A[j]) * A[j] for j in range(2))
    got = o.step_velocity(psi, np.tile(A, (s, 1)))
    ref = np.fft.ifftn(np.exp(-1j * (K + W) * dt) * np.fft.fftn(psi))
    assert rel(got, ref) < 1e-13
    # imaginary time, V = 0: exp(-K dt)
    got = o.step_imag(psi)
    assert rel(got, np.fft.ifftn(np.exp(-K * dt) * np.fft.fftn(psi))) < 1e-13


@pytest.mark.parametrize("integ,order", [("Verlet", 2), ("BM4", 4), ("RKN4", 4), ("BM6", 6)])
def test_convergence_order_harmonic(integ, order):
    # coherent state of V = x^2/2: <x>(t) = x0 cos t
    nr, dr = (128,), (0.2,)
    x0, T = 1.5, 4.0
    errs = []
    for dt in (0.2, 0.1):
        o = O.Oracle(nr, dr, "harmonic_test", integ, dt)
        psi, _ = free_packet(nr, dr, x0=x0)
        nt = int(round(T / dt))
        s = o.s
        out_psi, obs = o.real_run(psi, "length", np.zeros((nt, s, 1)), np.zeros((nt, s, 1)),
                                  names=["dipole_r"])
        d_final = o.observables(out_psi, ["dipole_r"])[0]
        errs.append(abs(d_final - x0 * np.cos(T)))
    observed = np.log2(errs[0] / errs[1])
    assert observed > order - 0.6, (errs, observed)


def test_imag_harmonic_spectrum():
    # 1D oscillator: E_n = n + 1/2 with imaginary time + Gram-Schmidt
    o = O.Oracle((128,), 0.15, "harmonic_test", "BM4", 0.05)
    psi, E, Lz = o.imag_run(3, 30.0, guess_mode=1, seed=3)
    assert np.allclose(E[:, 0], [0.5, 1.5, 2.5], atol=1e-6), E
    var = E[:, 1] - E[:, 0] ** 2
    assert np.all(np.abs(var) < 1e-6)
    G = np.einsum("ix,jx->ij", psi.conj(), psi) * 0.15
    assert np.allclose(G, np.eye(3), atol=1e-13)


def test_imag_step_count_quirk():
    assert O.imag_step_count(100.0, 0.1) == 1001
    assert O.imag_step_count(1000.0, 0.1) == 10000
    assert O.imag_step_count(1.0, 0.25) == 4


# ---------------------------------------------------------------- tables
def test_grid_tables_closed_form():
    nr, dr = (8, 4), (0.5, 0.25)
    o = O.Oracle(nr, dr, "Ne_2D", "BM4", 0.1, cap=(0.2, 0.5))
    r, p, K, Vc, V = o.tables()
    assert np.allclose(r[:8], -2.0 + 0.5 * np.arange(8))
    assert np.allclose(p[:8], 2 * np.pi * np.fft.fftfreq(8, 0.5))   # FFT order, p[4] = -pi/dr
    assert p[4] == -np.pi / 0.5
    # V(0) of Ne_2D: -(1+9)/sqrt(2.88172) at the grid point r = (0, 0)
    assert np.isclose(V[4, 2], -10.0 / np.sqrt(2.88172), rtol=1e-15)
    assert np.isclose(K[1, 1], 0.5 * ((2 * np.pi / 4.0) ** 2 + (2 * np.pi / 1.0) ** 2))
    # CAP starts 0.5 inside the edge: |x| >= L/2 - 0.5 = 1.5 on axis 0
    assert np.isclose(Vc[0, 2], 0.2 * (2.0 - 1.5) ** 2)
    dNr, dNp, fs = o.scalings()
    assert np.isclose(dNr, 0.125) and np.isclose(fs, 1 / 32)
    assert np.isclose(dNp, (2 * np.pi / 4.0) * (2 * np.pi / 1.0))


@pytest.mark.parametrize("name,val0", [("H_1D", -1 / np.sqrt(2.)), ("H_2D", -1 / np.sqrt(.7396)),
                                       ("H_3D", -1 / np.sqrt(1e-4)), ("Ar_2D", -18 / np.sqrt(.91205)),
                                       ("Kr_2D", -36 / np.sqrt(.7875)), ("Xe_2D", -54 / np.sqrt(.52592)),
                                       ("Ne_short_2D", -2.7459)])
def test_potential_values_at_origin(name, val0):
    dim = int(name.split("_")[-1][0])
    nr = (4,) * dim
    o = O.Oracle(nr, 1.0, name, "BM4", 0.1)
    V = o.tables()[4]
    assert np.isclose(V[(2,) * dim], val0, rtol=1e-14)


def test_softcore_extension():
    o = O.Oracle((4, 4), 1.0, "H_2D", "BM4", 0.1, softcore=1.0)
    assert np.isclose(o.tables()[4][2, 2], -1.0)


def test_unknown_names():
    with pytest.raises(ValueError, match="Unknown potential"):
        O.Oracle((8,), 1.0, "He_1D", "BM4", 0.1)
    with pytest.raises(ValueError, match="Unknown integrator"):
        O.Oracle((8,), 1.0, "H_1D", "Euler", 0.1)


# ---------------------------------------------------------------- observables
def dense_H(nr, dr, V):
    N = int(np.prod(nr))
    K, _ = kinetic_grid(nr, dr)
    I = np.eye(N).reshape((N,) + tuple(nr))
    cols = [np.fft.ifftn(K * np.fft.fftn(e)).ravel() for e in I]
    return np.array(cols).T + np.diag(V.ravel())


@pytest.mark.parametrize("nr,dr,pot", [((32,), (0.5,), "H_1D"), ((8, 8), (0.7, 0.6), "Ne_2D")])
def test_energies_dense_matrix(nr, dr, pot):
    o = O.Oracle(nr, dr, pot, "BM4", 0.1)
    V = o.tables()[4]
    H = dense_H(nr, dr, V)
    rng = np.random.default_rng(1)
    psi = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
    dN = np.prod(dr)
    v = psi.ravel()
    E = o.energies(psi, 3)
    ref = [np.real(v.conj() @ (np.linalg.matrix_power(H, m) @ v)) * dN for m in (1, 2, 3)]
    assert np.allclose(E, ref, rtol=1e-11)


def test_imag_ground_state_matches_dense_eigenvalue():
    nr, dr = (48,), (0.5,)
    o = O.Oracle(nr, dr, "H_1D", "BM4", 0.02)
    H = dense_H(nr, dr, o.tables()[4])
    ev = np.linalg.eigvalsh(0.5 * (H + H.conj().T))
    psi, E, _ = o.imag_run(2, 300.0, guess_mode=1, seed=5)
    # the split-operator fixed point differs from the exact eigenvector by O(dt^4)
    assert np.allclose(E[:, 0], ev[:2], atol=2e-7), (E[:, 0], ev[:2])


def test_dipole_and_dipole_v_scaling():
    nr, dr = (128, 64), (0.25, 0.3)
    o = O.Oracle(nr, dr, "H_2D", "BM4", 0.1)
    x0, k0 = 1.25, 0.7
    psi, _ = free_packet(nr, dr, x0=x0, k0=k0)
    out = o.observables(psi, ["norm", "dipole_r", "dipole_v"], alpha=[0.1, 0.0])
    assert np.isclose(out[0], 1.0, atol=1e-12)
    assert np.isclose(out[1], x0, atol=1e-10) and abs(out[2]) < 1e-10
    # reference scaling of dipole_v (src/calculate_observables.c:263): <p+A> dNp/dNr
    fac = np.prod([2 * np.pi / (n * d * d) for n, d in zip(nr, dr)])
    assert np.isclose(out[3], (k0 + 0.1) * fac, rtol=1e-8), (out[3], (k0 + 0.1) * fac)


def test_lz_vortex_and_real_state():
    nr, dr = (64, 64), (0.25, 0.25)
    o = O.Oracle(nr, dr, "H_2D", "BM4", 0.1)
    psi, X = free_packet(nr, dr)
    vort = (X[0] + 1j * X[1]) * psi
    vort /= np.sqrt(np.sum(np.abs(vort) ** 2) * 0.0625)
    assert np.isclose(o.observables(vort, ["Lz"])[0], 1.0, atol=1e-8)
    assert abs(o.observables(psi.real + 0j, ["Lz"])[0]) < 1e-13


def test_ehrenfest_dipole_a():
    # d^2<x>/dt^2 = -<dV/dx> - E <psi|psi>  (length gauge, weak field)
    nr, dr, dt = (256,), (0.2,), 0.01
    o = O.Oracle(nr, dr, "H_1D", "BM4", dt)
    psi, _ = free_packet(nr, dr, x0=0.8, w=1.2)
    nt = 300
    E0 = 0.02
    time = np.arange(nt) * dt
    s = o.s
    _, asc = o.coefficients()
    offs = np.concatenate([[asc[0]], asc[0] + np.cumsum(asc[1:2 * s - 1:2] + asc[2:2 * s:2])])[:s]
    Et = (E0 * np.sin(0.5 * (time[:, None] + offs[None, :])))[:, :, None]
    Eobs = (E0 * np.sin(0.5 * time))[:, None]
    _, obs = o.real_run(psi, "length", Et, np.zeros_like(Et), Eobs=Eobs, names=["dipole_r", "dipole_a"])
    x = obs[:, 0]
    acc_fd = (x[2:] - 2 * x[1:-1] + x[:-2]) / dt ** 2
    acc = obs[1:-1, 1]
    assert np.abs(acc_fd - acc).max() < 2e-4 * np.abs(acc).max() + 1e-6, np.abs(acc_fd - acc).max()


def test_gauge_equivalence_dipole():
    nr, dr, dt = (512,), (0.2,), 0.05
    modes = dict(freq=[0.0569547], amp=[0.01], ellip=[0.0], duration=[220.0], cep=[0.0], delay=[0.0],
                 env=[O.ENV["cos2"]])
    traces = {}
    for gauge in ("length", "velocity"):
        o = O.Oracle(nr, dr, "H_1D", "BM4", dt)
        psi, _, _ = o.imag_run(1, 20.0, guess_mode=0)
        nt = 400
        time = np.arange(nt) * dt
        _, asc = o.coefficients()
        A, E = O.field_tables(1, modes, asc, o.s, time)
        Ao = np.array([O.field_vectors(1, modes, t)[0] for t in time])
        _, obs = o.real_run(psi[0], gauge, E, A, Aobs=Ao, names=["dipole_r"])
        traces[gauge] = obs[:, 0]
    d = np.abs(traces["length"] - traces["velocity"]).max()
    assert d < 1e-4 * np.abs(traces["length"]).max() + 1e-7, d


def test_norm_conservation_and_cap():
    nr, dr, dt = (64, 64), (0.4, 0.4), 0.05
    o = O.Oracle(nr, dr, "H_2D", "BM4", dt)
    psi, _ = free_packet(nr, dr, x0=0.5, k0=1.0)
    nt = 40
    z = np.zeros((nt, o.s, 2))
    _, obs = o.real_run(psi, "velocity", z, z, names=["norm"])
    assert np.abs(obs[:, 0] - 1).max() < 1e-12
    oc = O.Oracle(nr, dr, "H_2D", "BM4", dt, cap=(0.5, 4.0))
    _, obs = oc.real_run(psi, "velocity", z, z, names=["norm"])
    assert np.all(np.diff(obs[:, 0]) <= 1e-15) and obs[-1, 0] < obs[0, 0]


def test_gram_schmidt_properties():
    nr, dr = (64,), 0.3
    o = O.Oracle(nr, dr, "H_1D", "BM4", 0.1)
    rng = np.random.default_rng(2)
    st = rng.standard_normal((3, 64)) + 1j * rng.standard_normal((3, 64))
    for j in range(3):
        rc, st = o.gram_schmidt(st, j)
        assert rc == 0
    G = np.einsum("ix,jx->ij", st.conj(), st) * 0.3
    assert np.allclose(G, np.eye(3), atol=1e-14)
    st[1] = st[0]
    rc, _ = o.gram_schmidt(st, 1)
    assert rc == 1


def test_mask_extension():
    o = O.Oracle((40,), 0.5, "H_1D", "BM4", 0.1)
    o.set_mask(0.1)
    m = o.apply_mask(np.ones(40, dtype=complex)).real
    x = -10 + 0.5 * np.arange(40)
    assert np.all(m[np.abs(x) <= 9.0] == 1.0)
    assert m[0] < 1e-2 and np.all(m > 0 - 1e-300)
    expect = np.where(np.abs(x) > 9.0, np.cos(0.5 * np.pi * (np.abs(x) - 9.0) / 1.0) ** 0.125, 1.0)
    assert np.allclose(m, expect, rtol=1e-14)


# ---------------------------------------------------------------- field
def test_field_closed_forms():
    w, I = 0.0569547, 1e14
    F0 = 5.33803e-9 * np.sqrt(I)
    tau = 4 * 2 * np.pi / w
    for env in ("constant", "trapz", "cos2", "cos4"):
        modes = dict(freq=[w], amp=[F0], ellip=[0.5], duration=[tau], cep=[0.3], delay=[10.0],
                     env=[O.ENV[env]])
        ts = np.linspace(20.0, tau - 20.0, 11)
        for t in ts:
            A, E = O.field_vectors(2, modes, t)
            h = 1e-4
            Ap, _ = O.field_vectors(2, modes, t + h)
            Am, _ = O.field_vectors(2, modes, t - h)
            dA = (Ap - Am) / (2 * h)
            assert np.allclose(E, -dA, atol=1e-9 * F0 + 1e-12), (env, t, E, -dA)
    modes = dict(freq=[w], amp=[F0], ellip=[0.0], duration=[tau], cep=[0.0], delay=[0.0],
                 env=[O.ENV["cos2"]])
    A, E = O.field_vectors(2, modes, tau / 2)
    assert np.isclose(A[0], F0 * np.cos(w * tau / 2) / w) and A[1] == 0.0
    A, E = O.field_vectors(2, modes, tau + 1.0)
    assert np.all(A == 0) and np.all(E == 0)


def test_field_stage_times():
    # stage times t_n + a0 h + sum_{j'<j}(a_{2j'+1}+a_{2j'+2}) h (src/integrators.c:94-101)
    modes = dict(freq=[0.0], amp=[0.0], ellip=[0.0], duration=[1.0], cep=[0.0], delay=[0.0],
                 env=[O.ENV["constant"]])
    s, a, asc = O.integrator_coefficients("BM4", 0.1)
    # a linear-in-time probe: a "mode" with w -> use constant envelope and tiny frequency
    w = 1e-3
    modes.update(freq=[w], amp=[w], ellip=[0.0])   # A_x = cos(w t), E_x = -w sin(w t)... scaled
    time = np.array([0.0, 0.1])
    A, E = O.field_tables(1, modes, asc, s, time)
    tt = np.arccos(np.clip(A[:, :, 0], -1, 1)) / w
    expect = time[:, None] + np.concatenate([[asc[0]], asc[0] + np.cumsum(asc[1:2 * s - 1:2] + asc[2:2 * s:2])])[None, :s]
    assert np.allclose(tt[1:], expect[1:], atol=1e-6)


# ---------------------------------------------------------------- drivers
def test_real_run_composition_length():
    nr, dr, dt = (64,), (0.3,), 0.05
    o = O.Oracle(nr, dr, "H_1D", "BM4", dt)
    psi, _ = free_packet(nr, dr, x0=0.5)
    nt, s = 5, o.s
    rng = np.random.default_rng(0)
    Et = rng.standard_normal((nt, s, 1)) * 0.01
    got, obs = o.real_run(psi, "length", Et, np.zeros_like(Et), names=["dipole_r"])
    N = 64
    p = np.fft.fft(psi) * np.sqrt(1 / N)
    for it in range(nt):
        r = np.fft.ifft(p) * N * np.sqrt(1 / N)
        assert np.isclose(o.observables(r, ["dipole_r"])[0], obs[it, 0], atol=1e-12)
        p = np.fft.fft(r) * np.sqrt(1 / N)
        p = o.step_length(p, Et[it])
    assert rel(got, np.fft.ifft(p) * N * np.sqrt(1 / N)) < 1e-13


def test_final_observables_restated():
    nr, dr = (16, 8), (0.7, 0.9)
    o = O.Oracle(nr, dr, "H_2D", "BM4", 0.1)
    psi, X = free_packet(nr, dr, x0=0.3, k0=0.9, w=2.0)
    Rm, pm, Ne, de = (0.5, 3.0), (2.0, 1.5), 20, 0.1
    psi_p, ped, pmd, npr = o.final_observables(psi, Rm, pm, Ne, de)
    R = np.sqrt(X[0] ** 2 + X[1] ** 2)
    fm = np.where((R >= Rm[0]) & (R <= Rm[1]), np.sin(0.5 * np.pi * (R - Rm[0]) / (Rm[1] - Rm[0])) ** 2,
                  np.where(R > Rm[1], 1.0, 0.0))
    ref_p = np.fft.fftn(fm * psi) / np.sqrt(psi.size)
    assert rel(psi_p, ref_p) < 1e-14
    K, P = kinetic_grid(nr, dr)
    inside = (np.abs(P[0]) <= pm[0]) & (np.abs(P[1]) <= pm[1])
    dens = np.abs(ref_p) ** 2
    assert list(npr) == [int(np.sum(np.abs(2 * np.pi * np.fft.fftfreq(n, d)) <= m)) for n, d, m in zip(nr, dr, pm)]
    sel = np.stack([dens[inside], np.angle(ref_p[inside])], axis=1)
    assert np.allclose(pmd, sel, rtol=1e-12, atol=1e-15)
    E = np.arange(Ne) * de
    ref_ped = np.array([np.sum(dens[inside] * np.exp(-0.5 * (K[inside] - e) ** 2 / de ** 2))
                        for e in E]) / (2 * np.pi * de * de)
    assert np.allclose(ped, ref_ped, rtol=1e-12)


def test_gaussian_stream_splitmix_boxmuller():
    def splitmix(state):
        M = (1 << 64) - 1
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return state, z ^ (z >> 31)
    st, ref = 12345, []
    for _ in range(4):
        st, a = splitmix(st)
        st, b = splitmix(st)
        u1 = ((a >> 11) + 1) * 2.0 ** -53
        u2 = ((b >> 11) + 1) * 2.0 ** -53
        r = np.sqrt(-2 * np.log(u1))
        ref += [r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)]
    got = O.gaussian(12345, 8)
    assert np.allclose(got, ref, rtol=1e-14, atol=1e-15)
    big = O.gaussian(7, 200000)
    assert abs(big.mean()) < 0.01 and abs(big.std() - 1) < 0.01


# ---------------------------------------------------------------- PED pinned by mathematics
def _boosted_packet(n, dr, w, k0):
    x = -0.5 * n * dr + np.arange(n) * dr
    X, Y = np.meshgrid(x, x, indexing="ij")
    return np.exp(-(X ** 2 + Y ** 2) / (2 * w * w)) * np.exp(1j * k0 * X)


def test_ped_sum_rule_and_moments():
    """The PED (src/io_functions.c:172-199) is the unitary p-space density
    smoothed by a normalised Gaussian of width sigma = DE and sampled at
    E_j = j DE.  For a boosted Gaussian packet whose energies lie well inside
    [0, (Ne-1) DE], the Riemann sums of that kernel are exact up to a Poisson
    term ~ e^{-2 pi^2}, which fixes three properties independently of the
    oracle's formula:
      sum_j S_j DE         = (2 pi sigma^2)^((1-d)/2) sum |psi|^2   (normalisation, exponent d/2)
      sum_j E_j S_j / sum  = <K>                                     (kernel centred at j DE)
      variance of S        = Var(K) + sigma^2                        (kernel width sigma, not sigma^2)
    <K> and Var(K) come from numpy's FFT of the packet and, to the grid's
    accuracy, from the closed forms of a Gaussian in k-space."""
    n, dr, w, k0, d = 256, 0.5, 4.0, 1.5, 2
    o = O.Oracle((n, n), dr, "H_2D", "BM4", 0.1)
    psi = _boosted_packet(n, dr, w, k0)
    Ne, de = 400, 0.01
    # fm = 1 everywhere (every radius exceeds R1 < 0); every momentum inside
    _, ped, _, _ = o.final_observables(psi, (-2.0, -1.0), (100.0, 100.0), Ne, de)
    sigma = de
    E = np.arange(Ne) * de
    total = np.sum(np.abs(psi) ** 2)            # Parseval: the unitary p-space state has the same sum
    lhs = np.sum(ped) * de
    rhs = (2 * np.pi * sigma * sigma) ** ((1 - d) / 2) * total
    assert abs(lhs / rhs - 1) < 1e-8, (lhs, rhs)
    # moments of K from numpy's FFT
    p = 2 * np.pi * np.fft.fftfreq(n, dr)
    K = 0.5 * (p[:, None] ** 2 + p[None, :] ** 2)
    dens = np.abs(np.fft.fftn(psi)) ** 2
    mK = np.sum(dens * K) / np.sum(dens)
    vK = np.sum(dens * (K - mK) ** 2) / np.sum(dens)
    mS = np.sum(E * ped) / np.sum(ped)
    vS = np.sum((E - mS) ** 2 * ped) / np.sum(ped)
    assert abs(mS - mK) < 1e-8 * mK, (mS, mK)
    assert abs(vS - (vK + sigma ** 2)) < 1e-6 * vK, (vS, vK + sigma ** 2)
    # closed forms: |psi_p(k)|^2 ~ exp(-(k - k0)^2 w^2), variance s2 = 1/(2 w^2) per axis
    s2 = 1.0 / (2 * w * w)
    assert abs(mK - 0.5 * (k0 * k0 + d * s2)) < 1e-9, mK
    vK_closed = (4 * k0 * k0 * s2 + 2 * s2 * s2) / 4 + (d - 1) * (2 * s2 * s2) / 4
    assert abs(vK - vK_closed) < 1e-9, (vK, vK_closed)


def test_ped_restricted_momenta_and_mask():
    """P_MAX restricts the PED to the points of the PMD selection, and R_MASK
    zeroes the inner region: with both the sum rule holds on the selected,
    masked density (numpy FFT of the masked packet, selected by |p_j|)."""
    n, dr, w, k0 = 128, 0.5, 3.0, 1.2
    o = O.Oracle((n, n), dr, "H_2D", "BM4", 0.1)
    psi = _boosted_packet(n, dr, w, k0)
    Rm, pm, Ne, de = (1.0, 4.0), (1.6, 0.3), 500, 0.01
    _, ped, _, _ = o.final_observables(psi, Rm, pm, Ne, de)
    x = -0.5 * n * dr + np.arange(n) * dr
    R = np.sqrt(x[:, None] ** 2 + x[None, :] ** 2)
    fm = np.where((R >= Rm[0]) & (R <= Rm[1]), np.sin(0.5 * np.pi * (R - Rm[0]) / (Rm[1] - Rm[0])) ** 2,
                  np.where(R > Rm[1], 1.0, 0.0))
    pp = np.abs(np.fft.fftn(fm * psi)) ** 2 / psi.size
    p = 2 * np.pi * np.fft.fftfreq(n, dr)
    inside = (np.abs(p)[:, None] <= pm[0]) & (np.abs(p)[None, :] <= pm[1])
    K = 0.5 * (p[:, None] ** 2 + p[None, :] ** 2)
    assert K[inside].max() < (Ne - 1) * de - 8 * de and K[inside].min() >= 0
    lhs = np.sum(ped) * de
    # points with K < 8 sigma lose part of the kernel below E = 0: sum them exactly
    kern = np.exp(-0.5 * (K[inside][:, None] - np.arange(Ne)[None, :] * de) ** 2 / de ** 2).sum(axis=1)
    rhs = np.sum(pp[inside] * kern) * de / (2 * np.pi * de * de)
    assert abs(lhs / rhs - 1) < 1e-12, (lhs, rhs)
    far = K[inside] > 8 * de
    rhs_far = (2 * np.pi * de * de) ** (-0.5) * np.sum(pp[inside][far])
    lhs_far = np.sum(pp[inside][far] * kern[far]) * de / (2 * np.pi * de * de)
    assert abs(lhs_far / rhs_far - 1) < 1e-8, (lhs_far, rhs_far)


# ---------------------------------------------------------------- polarisation angle (extension)
def _two_colour(angle2=None):
    w1 = 45.5637639947958 / 800.0
    m = dict(freq=[w1, 2 * w1], amp=[5.33803e-9 * np.sqrt(1e14), 5.33803e-9 * np.sqrt(1e13)],
             ellip=[0.0, 0.3], duration=[4 * 2 * np.pi / w1, 8 * 2 * np.pi / (2 * w1)], cep=[0.0, 1.1],
             delay=[0.0, 5.0], env=[O.ENV["cos2"], O.ENV["cos2"]])
    if angle2 is not None:
        m["angle"] = [0.0, angle2]
    return m


def test_field_angle_zero_is_reference():
    for t in np.linspace(0.0, 450.0, 37):
        A0, E0 = O.field_vectors(3, _two_colour(), t)
        A1, E1 = O.field_vectors(3, _two_colour(0.0), t)
        assert np.allclose(A1, A0, rtol=1e-15, atol=1e-300) and np.allclose(E1, E0, rtol=1e-15, atol=1e-300)
        assert A1[2] == 0.0 and E1[2] == 0.0     # no z component (src/parameters.c:942-947)


def test_field_angle_is_a_rotation():
    """A rotated mode is the rotation matrix applied to the unrotated mode's
    (x, y) vectors; modes add linearly; E = -dA/dt still holds."""
    th = np.deg2rad(20.0)
    Rot = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    single = lambda m, k: {key: [v[k]] for key, v in m.items()}
    m = _two_colour(th)
    for t in np.linspace(1.0, 440.0, 23):
        A, E = O.field_vectors(2, m, t)
        A1, E1 = O.field_vectors(2, single(_two_colour(), 0), t)
        A2, E2 = O.field_vectors(2, single(_two_colour(), 1), t)
        assert np.allclose(A, A1 + Rot @ A2, rtol=1e-13, atol=1e-15)
        assert np.allclose(E, E1 + Rot @ E2, rtol=1e-13, atol=1e-15)
        A2r, E2r = O.field_vectors(2, single(m, 1), t)
        assert abs(np.linalg.norm(E2r) - np.linalg.norm(E2)) < 1e-15 + 1e-13 * np.linalg.norm(E2)
        h = 1e-4
        Ap, _ = O.field_vectors(2, m, t + h)
        Am, _ = O.field_vectors(2, m, t - h)
        assert np.allclose(E, -(Ap - Am) / (2 * h), atol=1e-9)


def test_field_py_matches_oracle():
    """tdse_nd_mpi_amd.field (the tables the benchmark and several GPU tests
    propagate under) against the oracle's stage tables, every envelope, with
    and without the polarisation angle."""
    from tdse_nd_mpi_amd import field as F
    dt, nt = 0.05, 900
    for env in ("constant", "trapz", "cos2", "cos4"):
        for integ in ("BM4", "BM6"):
            modes = [F.Mode(intensity=1e14, wavelength=800.0, ellipticity=0.2, cep=0.4, delay=3.0,
                            envelope=env, cycles=2),
                     F.Mode(intensity=1e13, wavelength=400.0, ellipticity=0.0, cep=1.1, delay=0.0,
                            envelope=env, cycles=3, angle=np.deg2rad(20.0))]
            Et, At, Eo, Ao = F.tables(modes, integ, dt, nt, 2)
            om = dict(freq=[m.w for m in modes], amp=[m.F0 for m in modes], ellip=[m.xi for m in modes],
                      duration=[m.tau for m in modes], cep=[m.phi for m in modes],
                      delay=[m.delay for m in modes], env=[O.ENV[env]] * 2, angle=[m.angle for m in modes])
            s, _, asc = O.integrator_coefficients(integ, dt)
            time = np.arange(nt) * dt
            A, E = O.field_tables(2, om, asc, s, time)
            scale = max(np.abs(E).max(), 1e-300)
            assert np.allclose(Et, E, rtol=0, atol=1e-13 * scale), (env, integ, np.abs(Et - E).max())
            assert np.allclose(At, A, rtol=0, atol=1e-13 * np.abs(A).max()), (env, integ)
            Eo_ref = np.array([O.field_vectors(2, om, t)[1] for t in time[::50]])
            assert np.allclose(Eo[::50], Eo_ref, rtol=0, atol=1e-13 * scale)
