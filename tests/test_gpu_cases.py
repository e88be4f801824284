"""GPU parity, further cases: the named potentials with their own device
branches (the Gaussian well Ne_short_2D, the screened Ar_2D), the large 3-D
configuration's imaginary-time leg (BM6, eight states, one-pass Gram matrix of
seven previous states, both precisions) and its angled two-colour field,
size-independent properties at 512^3, and the set-up sequences that the
library's lazily paid operators must get right.  All through the C ABI,
against the FP64 oracle on identical inputs.
"""
import numpy as np
import pytest

import oracle_lib as O
import tdse_nd_mpi_amd as q
from tdse_nd_mpi_amd import field as F

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 2e-5


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def oracle_field(o, nt, dim, modes):
    """Stage and step-time tables from the oracle for F.Mode objects."""
    om = dict(freq=[m.w for m in modes], amp=[m.F0 for m in modes], ellip=[m.xi for m in modes],
              duration=[m.tau for m in modes], cep=[m.phi for m in modes], delay=[m.delay for m in modes],
              env=[O.ENV[m.env] for m in modes], angle=[m.angle for m in modes])
    _, asc = o.coefficients()
    time = np.arange(nt) * o.dt
    A, E = O.field_tables(dim, om, asc, o.s, time)
    Ao = np.zeros((nt, dim))
    Eo = np.zeros((nt, dim))
    for i, t in enumerate(time):
        Ao[i], Eo[i] = O.field_vectors(dim, om, t)
    return E, A, Eo, Ao


# ---------------------------------------------------------------- named potentials
# complex64-storage bounds of the acceleration records per named potential
# (see the test below)
ACCEL_TOL = {"Ne_short_2D": 1e-4, "Ar_2D": 1e-4, "Kr_2D": 2e-3, "Xe_2D": 2e-2}


@pytest.mark.parametrize("prec", ["c128", "c64", "c64f64"])
@pytest.mark.parametrize("pot", ["Ne_short_2D", "Ar_2D", "Kr_2D", "Xe_2D"])
def test_named_potential_imag_and_real(gpu, pot, prec):
    """Imaginary time with Gram-Schmidt (two states), then length-gauge real
    time with CAP, mask and field, on the potential's own device branch."""
    nr, dr, dt = (48, 32), 0.4, 0.05
    print(f"case named potential {pot} {prec}", flush=True)
    o = O.Oracle(nr, dr, pot, "BM4", dt, cap=(0.05, 2.0))
    o.set_mask(0.1)
    g = q.Grid(nr, dr, prec, n_slots=2)
    g.set_potential(pot, cap=(0.05, 2.0))
    g.set_integrator("BM4", dt)
    g.set_mask(0.1)
    states = o.initial_guess(2, guess_mode=1, seed=31)
    ref = states.copy()
    for j in range(2):
        g.upload(j, states[j])
        for _ in range(15):
            ref[j] = o.step_imag(ref[j])
            rc, ref = o.gram_schmidt(ref, j)
            assert rc == 0
        g.imag_steps(j, 15)
    tol = TOL128 if prec == "c128" else TOL64
    for j in range(2):
        assert rel(g.download(j), ref[j]) < tol, (j, rel(g.download(j), ref[j]))
    E = g.energy_moments(0, 2)
    Eo = o.energies(g.download(0), 2)
    if prec == "c128":
        assert np.allclose(E, Eo, rtol=0, atol=1e-11), (E, Eo)
    else:   # FP32 transforms inside the FP64-accumulated moments
        assert np.allclose(E, Eo, rtol=1e-6, atol=1e-7), (E, Eo)
    nt = 30
    modes = [F.Mode(intensity=1e14, wavelength=800.0, ellipticity=0.2, cep=0.3, envelope="cos2", cycles=1)]
    Et, At, Eob, Aob = oracle_field(o, nt, 2, modes)
    g.set_field(Et, At, Eob, Aob)
    names = ["dipole_r", "dipole_a", "norm"]
    psi0 = ref[0]
    g.upload(0, psi0)
    got_obs = g.real_steps(0, "length", 0, nt, names)
    ref_psi, ref_obs = o.real_run(psi0, "length", Et, At, Eob, Aob, names)
    got = g.download(0)
    g.close()
    if prec == "c128":
        assert rel(got, ref_psi) < TOL128, rel(got, ref_psi)
        assert np.allclose(got_obs, ref_obs, rtol=1e-11, atol=1e-12), np.abs(got_obs - ref_obs).max()
    else:
        assert rel(got, ref_psi) < 1e-4, rel(got, ref_psi)
        scale = np.abs(ref_obs).max(axis=0) + 1e-3
        err = (np.abs(got_obs - ref_obs) / scale).max(axis=0)
        print(f"named potential {pot} {prec}: state {rel(got, ref_psi):.2e}, records {err}", flush=True)
        # columns: dipole_r (2), dipole_a (2), norm (1), each against its largest value plus
        # 1e-3.  dipole_a = -<grad V> - E<psi|psi> is a small difference of large terms for
        # the deep potentials (|V(0)| = 19, 41 and 74 Ha for Ar, Kr and Xe), which magnifies
        # the complex64 rounding of the stored state: measured 1.6e-5, 6.3e-4 and 7.7e-3 in
        # complex64, 1.6e-6, 2.9e-4 and 5.7e-3 with FP64 pass arithmetic (c64f64), whose
        # state error is ten times smaller; complex128 matches to 1e-11 (above).  The bounds
        # catch a wrong potential branch or constant, which moves the records by O(1).
        assert err[[0, 1, 4]].max() < 3e-4, err
        assert err[[2, 3]].max() < ACCEL_TOL[pot], (err, ACCEL_TOL[pot])


# ---------------------------------------------------------------- the 3-D configuration's imaginary leg
@pytest.mark.parametrize("prec", ["c128", "c64", "c64f64"])
def test_3d_bm6_eight_states(gpu, prec):
    """H_3D with soft-core length 0.05, BM6, eight states relaxed one after
    another with Gram-Schmidt after every step: the eighth state's projection
    uses the one-pass Gram matrix of seven previous states (k_gram<R, 7>),
    in complex64 (and c64f64) the FP32-state path with FP64 reductions."""
    nr, dr, dt, nsteps = (32, 32, 32), 0.25, 0.05, 6
    print(f"case 3-D BM6 eight states {prec}", flush=True)
    o = O.Oracle(nr, dr, "H_3D", "BM6", dt, softcore=0.05)
    g = q.Grid(nr, dr, prec, n_slots=8)
    g.set_potential("H_3D", softcore=0.05)
    g.set_integrator("BM6", dt)
    states = o.initial_guess(8, guess_mode=1, seed=5)
    ref = states.copy()
    for j in range(8):
        g.upload(j, states[j])
        for _ in range(nsteps):
            ref[j] = o.step_imag(ref[j])
            rc, ref = o.gram_schmidt(ref, j)
            assert rc == 0
        g.imag_steps(j, nsteps)
    got = np.array([g.download(j) for j in range(8)])
    tol = TOL128 if prec == "c128" else 5e-5
    for j in range(8):
        assert rel(got[j], ref[j]) < tol, (j, rel(got[j], ref[j]))
    dV = dr ** 3
    gram = np.einsum("ixyz,jxyz->ij", got.conj(), got) * dV
    assert np.allclose(gram, np.eye(8), atol=1e-12 if prec == "c128" else 1e-6), np.abs(gram - np.eye(8)).max()
    E = g.energy_moments(7, 2)
    Eo = o.energies(ref[7], 2)
    assert np.allclose(E, Eo, rtol=1e-10 if prec == "c128" else 1e-4), (E, Eo)
    g.close()


@pytest.mark.parametrize("prec", ["c128", "c64", "c64f64"])
@pytest.mark.parametrize("gauge", ["length", "velocity"])
def test_3d_angled_two_colour_field(gpu, gauge, prec):
    """The 3-D configuration's field: 800 nm at 1e14 W/cm^2 plus its second
    harmonic at 1e13 W/cm^2 polarised 20 degrees away, seeded relative phase.
    The GPU propagates under tdse_nd_mpi_amd.field's tables, the oracle under
    its own angle tables (the two are compared on the CPU)."""
    nr, dr, dt, nt = (16, 32, 16), 0.5, 0.02, 40
    print(f"case 3-D angled two-colour field {gauge} {prec}", flush=True)
    ph = float(np.random.default_rng(5).uniform(0, 2 * np.pi))
    # constant envelopes: the field is at full strength from the first step
    modes = [F.Mode(intensity=1e14, wavelength=800.0, cep=0.0, envelope="constant", cycles=1),
             F.Mode(intensity=1e13, wavelength=400.0, cep=ph, envelope="constant", cycles=2,
                    angle=np.deg2rad(20.0))]
    o = O.Oracle(nr, dr, "H_3D", "BM4", dt, softcore=0.05, cap=(0.05, 1.5))
    o.set_mask(0.1)
    Et, At, Eob, Aob = oracle_field(o, nt, 3, modes)
    assert np.abs(Et[:, :, 1]).max() > 1e-3 and np.all(Et[:, :, 2] == 0.0)
    g = q.Grid(nr, dr, prec, n_slots=1)
    g.set_potential("H_3D", softcore=0.05, cap=(0.05, 1.5))
    g.set_integrator("BM4", dt)
    g.set_mask(0.1)
    g.set_field(*F.tables(modes, "BM4", dt, nt, 3))
    psi0 = o.initial_guess(1, guess_mode=1, seed=9)[0]
    names = ["dipole_r", "dipole_a", "norm"]
    g.upload(0, psi0)
    got_obs = g.real_steps(0, gauge, 0, nt, names)
    got = g.download(0)
    g.close()
    ref_psi, ref_obs = o.real_run(psi0, gauge, Et, At, Eob, Aob, names)
    if prec == "c128":
        assert rel(got, ref_psi) < TOL128, rel(got, ref_psi)
        assert np.allclose(got_obs, ref_obs, rtol=1e-11, atol=1e-12), np.abs(got_obs - ref_obs).max()
    else:
        assert rel(got, ref_psi) < 1e-4, rel(got, ref_psi)
        scale = np.abs(ref_obs).max(axis=0) + 1e-3
        assert (np.abs(got_obs - ref_obs) / scale).max() < 1e-4


# ---------------------------------------------------------------- full size: 512^3
def _dot64(a, b, chunk=1 << 23):
    """sum conj(a) b accumulated in FP64 over chunks (a, b complex64, flat)."""
    s = 0j
    for k in range(0, a.size, chunk):
        s += np.vdot(a[k:k + chunk].astype(np.complex128), b[k:k + chunk].astype(np.complex128))
    return s


def test_512_cube_properties(gpu):
    """The 3-D configuration at its full size, 512^3 complex64, through
    properties that hold at any size: eight states after two BM6
    imaginary-time steps each with Gram-Schmidt are orthonormal (Gram matrix
    = I to FP32 rounding), and without absorbers three real-time BM4 steps
    under the angled two-colour field conserve the norm."""
    n, dr = 512, 0.2
    print("case 512^3 c64 eight states + three real steps", flush=True)
    x = -0.5 * n * dr + np.arange(n) * dr
    g = q.Grid((n, n, n), dr, "c64", n_slots=8)
    g.set_potential("H_3D", softcore=0.05)
    g.set_integrator("BM6", 0.05)
    rng = np.random.default_rng(512)
    for j in range(8):
        # separable smooth guesses: different low-order profiles per state
        f = [np.exp(-x * x / rng.uniform(6, 12)) * np.polynomial.hermite.hermval(x / 2.0, rng.standard_normal(4))
             for _ in range(3)]
        psi = (f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :]).astype(np.complex64)
        psi *= np.complex64(np.exp(1j * rng.uniform(0, 2 * np.pi)))
        g.upload(j, psi)
        del psi
        g.imag_steps(j, 2)
    flat = []
    for j in range(8):
        flat.append(g.download(j).astype(np.complex64).ravel())
    dV = dr ** 3
    gram = np.zeros((8, 8), dtype=complex)
    for i in range(8):
        for j in range(i, 8):
            gram[i, j] = _dot64(flat[i], flat[j]) * dV
            gram[j, i] = np.conj(gram[i, j])
    err = np.abs(gram - np.eye(8)).max()
    print(f"512^3 Gram matrix max |G - I| = {err:.3e}", flush=True)
    assert err < 1e-8, err
    del flat
    nt = 3
    modes = [F.Mode(intensity=1e14, wavelength=800.0, envelope="constant", cycles=4),
             F.Mode(intensity=1e13, wavelength=400.0, cep=1.0, envelope="constant", cycles=8,
                    angle=np.deg2rad(20.0))]
    g.set_integrator("BM4", 0.02)
    g.set_field(*F.tables(modes, "BM4", 0.02, nt, 3))
    rec = g.real_steps(0, "length", 0, nt, ["norm", "dipole_r"])
    g.close()
    print(f"512^3 norms {rec[:, 0]}", flush=True)
    assert abs(rec[0, 0] - 1.0) < 1e-5, rec[0, 0]
    assert np.abs(rec[:, 0] - rec[0, 0]).max() < 1e-6, rec[:, 0]
    assert np.all(np.isfinite(rec))


# ---------------------------------------------------------------- set-up sequences
def _setup(nr, dr, prec, cap=(0.05, 2.0), mask=0.1, integ="BM4", dt=0.05):
    o = O.Oracle(nr, dr, "H_2D", integ, dt, cap=cap)
    o.set_mask(mask)
    g = q.Grid(nr, dr, prec, n_slots=1)
    g.set_potential("H_2D", cap=cap)
    g.set_integrator(integ, dt)
    g.set_mask(mask)
    return o, g


def test_velocity_then_length_steps(gpu):
    """Velocity-gauge steps leave their last V half-step owed; a following
    length-gauge call pays it first, at its own time, so the sequence equals
    the oracle's velocity run followed by its length run."""
    nr, dr, nt = (32, 32), 0.4, 10
    print("case velocity then length", flush=True)
    o, g = _setup(nr, dr, "c128")
    modes = [F.Mode(intensity=1e14, wavelength=800.0, ellipticity=0.3, cep=0.2, envelope="cos2", cycles=1)]
    Et, At, Eob, Aob = oracle_field(o, 2 * nt, 2, modes)
    g.set_field(Et, At, Eob, Aob)
    psi0 = o.initial_guess(1, guess_mode=1, seed=3)[0]
    g.upload(0, psi0)
    g.real_steps(0, "velocity", 0, nt)
    g.real_steps(0, "length", nt, nt)
    got = g.download(0)
    g.close()
    mid, _ = o.real_run(psi0, "velocity", Et[:nt], At[:nt], Eob[:nt], Aob[:nt])
    ref, _ = o.real_run(mid, "length", Et[nt:], At[nt:], Eob[nt:], Aob[nt:])
    assert rel(got, ref) < TOL128, rel(got, ref)


@pytest.mark.parametrize("gauge", ["length", "velocity"])
def test_setup_change_between_calls(gauge, gpu):
    """A mask and (velocity gauge) a V half-step owed by the last call are paid
    with the set-up of that call when the mask, potential or integrator
    changes afterwards."""
    nr, dr, nt = (32, 32), 0.4, 8
    print(f"case set-up change between calls {gauge}", flush=True)
    o, g = _setup(nr, dr, "c128", mask=0.1)
    modes = [F.Mode(intensity=1e14, wavelength=800.0, ellipticity=0.2, cep=0.5, envelope="cos2", cycles=1)]
    Et, At, Eob, Aob = oracle_field(o, 2 * nt, 2, modes)
    g.set_field(Et, At, Eob, Aob)
    psi0 = o.initial_guess(1, guess_mode=1, seed=8)[0]
    g.upload(0, psi0)
    g.real_steps(0, gauge, 0, nt)
    g.set_mask(0.2)                          # the owed mask is the 0.1 one
    g.set_potential("H_2D", cap=(0.08, 3.0))  # the owed V half-step uses the old potential and CAP
    g.real_steps(0, gauge, nt, nt)
    got = g.download(0)
    g.close()
    mid, _ = o.real_run(psi0, gauge, Et[:nt], At[:nt], Eob[:nt], Aob[:nt])
    o2 = O.Oracle(nr, dr, "H_2D", "BM4", 0.05, cap=(0.08, 3.0))
    o2.set_mask(0.2)
    ref, _ = o2.real_run(mid, gauge, Et[nt:], At[nt:], Eob[nt:], Aob[nt:])
    assert rel(got, ref) < TOL128, rel(got, ref)


def test_field_tables_must_match_the_integrator(gpu):
    """Tables laid out for one stage count are refused after the integrator
    changes to another, by the Python check and by the library."""
    nr, dr = (16, 16), 0.5
    g = q.Grid(nr, dr, "c64", n_slots=1)
    g.set_potential("H_2D")
    g.set_integrator("Verlet", 0.05)
    with pytest.raises(q.QNDError, match="shape"):
        g.set_field(np.zeros((4, 6, 2)), np.zeros((4, 6, 2)))
    g.set_field(np.zeros((4, 1, 2)), np.zeros((4, 1, 2)))
    g.set_integrator("BM6", 0.05)
    g.upload(0, np.ones(nr, dtype=complex))
    with pytest.raises(q.QNDError, match="stages"):
        g.real_steps(0, "length", 0, 2)
    with pytest.raises(q.QNDError, match="stages"):
        g.forward_checkpointed(0, 2, 1)
    g.close()


def test_adjoint_rejects_wrong_step_count(gpu):
    nr, dr = (16, 16), 0.5
    g = q.Grid(nr, dr, "c64", n_slots=1)
    g.set_potential("H_2D")
    g.set_integrator("BM4", 0.05)
    g.set_field(np.zeros((12, 6, 2)), np.zeros((12, 6, 2)))
    g.upload(0, np.ones(nr, dtype=complex) * 0.1)
    g.forward_checkpointed(0, 10, 4)
    with pytest.raises(q.QNDError, match="10"):
        g.adjoint(np.zeros((12, 2)), 0.0)
    dEt, dEo, _ = g.adjoint(np.zeros((10, 2)), 1.0)
    assert dEt.shape == (10, 6, 2) and np.all(np.isfinite(dEt))
    g.close()


def test_two_applications_before_backward(gpu):
    """Two applications of the autograd function on one grid, then the
    backward of the first: its gradient is the first field's, as a separate
    run gives it."""
    import torch
    from tdse_nd_mpi_amd import adjoint as AD
    nr, dr, nt = (32, 32), 0.4, 24
    print("case two applications before backward", flush=True)
    o = O.Oracle(nr, dr, "H_2D", "BM4", 0.05)
    psi0 = o.initial_guess(1, guess_mode=1, seed=2)[0]
    psi0 = psi0 / np.sqrt(np.sum(np.abs(psi0) ** 2) * dr * dr)
    rng = np.random.default_rng(4)
    E1 = torch.tensor(0.02 * rng.standard_normal(nt), requires_grad=True)
    E2 = torch.tensor(0.05 * rng.standard_normal(nt), requires_grad=True)
    w = torch.from_numpy(rng.standard_normal((nt, 2)))

    def grad_of(spec, E):
        a, nf = AD.Propagate.apply(E, torch.tensor(-1.0, dtype=torch.float64), spec)
        return a, nf

    spec = AD.Spec(nr, dr, "H_2D", "BM4", 0.05, psi0, precision="c128", mask=0.1, every=8)
    a1, n1 = grad_of(spec, E1)
    a2, n2 = grad_of(spec, E2)
    ((a1 * w).sum() + 0.3 * n1).backward()
    g_first = E1.grad.clone()
    spec.close()
    spec2 = AD.Spec(nr, dr, "H_2D", "BM4", 0.05, psi0, precision="c128", mask=0.1, every=8)
    E1b = E1.detach().clone().requires_grad_(True)
    a, nf = grad_of(spec2, E1b)
    ((a * w).sum() + 0.3 * nf).backward()
    spec2.close()
    assert torch.allclose(g_first, E1b.grad, rtol=1e-10, atol=1e-14), (g_first - E1b.grad).abs().max()


def test_gram_schmidt_collapse_stops_the_call(gpu):
    """A state that Gram-Schmidt reduces to zero is reported at the step of
    the collapse, and the call stops within two check intervals (the
    reference aborts at that step, src/QuantumND_imag.c:297-304)."""
    import time
    nr, dr = (256, 256), 0.3
    g = q.Grid(nr, dr, "c128", n_slots=2)
    g.set_potential("H_2D")
    g.set_integrator("BM4", 0.05)
    x = -0.5 * nr[0] * dr + np.arange(nr[0]) * dr
    psi = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / 4.0).astype(complex)
    g.upload(0, psi / np.sqrt(np.sum(np.abs(psi) ** 2) * dr * dr))
    g.imag_steps(0, 10)
    g.upload(1, np.zeros(nr, dtype=complex))
    t0 = time.perf_counter()
    g.imag_steps(0, 1, gram_schmidt=False)       # warm the launch path
    t_one = time.perf_counter() - t0
    t0 = time.perf_counter()
    with pytest.raises(q.QNDError, match=r"norm decayed to zero during Gram-Schmidt for state 1 \(step 0 "):
        g.imag_steps(1, 5000)
    dt_fail = time.perf_counter() - t0
    g.close()
    print(f"collapse reported after {dt_fail:.3f} s (one step {t_one * 1e3:.2f} ms)", flush=True)
    assert dt_fail < 400 * max(t_one, 1e-4), (dt_fail, t_one)
