"""Field synthesis on the device (qnd_set_field_modes) against the FP64
oracle's Field_vectors / stage tables (src/parameters.c:899-949,
src/integrators.c:74-129), every envelope, with and without the polarisation
angle, 1-D to 3-D, several instances with their own modes; and propagation
under the device tables against propagation under host tables."""
import numpy as np
import pytest

import oracle_lib as O
import tdse_nd_mpi_amd as q
from tdse_nd_mpi_amd import field as F

pytestmark = pytest.mark.gpu


def instance_modes(b, env):
    """Two modes per instance, instance b with its own intensity, CEP,
    delay, ellipticity and angle."""
    return [F.Mode(intensity=1e14 * (1 + 0.3 * b), wavelength=800.0, ellipticity=0.2 * b, cep=0.4 + 0.7 * b,
                   delay=3.0 * b, envelope=env, cycles=2, angle=0.0 if b == 0 else np.deg2rad(20.0 * b)),
            F.Mode(intensity=1e13, wavelength=400.0, ellipticity=0.0, cep=1.1 - 0.2 * b, delay=0.0,
                   envelope=env, cycles=3)]


def oracle_modes(ms, env):
    return dict(freq=[m.w for m in ms], amp=[m.F0 for m in ms], ellip=[m.xi for m in ms],
                duration=[m.tau for m in ms], cep=[m.phi for m in ms], delay=[m.delay for m in ms],
                env=[O.ENV[env]] * len(ms), angle=[m.angle for m in ms])


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("env", ["constant", "trapz", "cos2", "cos4"])
def test_device_field_matches_oracle(gpu, dim, env):
    print(f"case device field dim={dim} env={env}", flush=True)
    nr = {1: (64,), 2: (16, 16), 3: (8, 8, 8)}[dim]
    integ, dt, nt, nb = ("BM4" if dim != 2 else "BM6"), 0.05, 700, 3
    g = q.Grid(nr, 0.3, "c128", n_slots=nb)
    g.set_potential({1: "H_1D", 2: "H_2D", 3: "H_3D"}[dim])
    g.set_integrator(integ, dt)
    modes = [instance_modes(b, env) for b in range(nb)]
    g.set_field_modes(modes, nt)
    Et, At, Eo = g.field_batch()
    s, _, asc = O.integrator_coefficients(integ, dt)
    time = np.arange(nt) * dt
    for b in range(nb):
        om = oracle_modes(modes[b], env)
        A, E = O.field_tables(dim, om, asc, s, time)
        scale_e, scale_a = max(np.abs(E).max(), 1e-300), max(np.abs(A).max(), 1e-300)
        assert np.abs(E).max() > 0
        assert np.allclose(Et[b], E, rtol=0, atol=1e-13 * scale_e), (b, np.abs(Et[b] - E).max() / scale_e)
        assert np.allclose(At[b], A, rtol=0, atol=1e-13 * scale_a), (b, np.abs(At[b] - A).max() / scale_a)
        Eo_ref = np.array([O.field_vectors(dim, om, t)[1] for t in time[::25]])
        assert np.allclose(Eo[b][::25], Eo_ref, rtol=0, atol=1e-13 * scale_e), b
        if dim == 3:
            assert not Et[b][..., 2].any() and not Eo[b][..., 2].any()
    g.close()


def test_device_field_angle_zero_is_reference(gpu):
    """angle 0 gives exactly the reference's components: the same tables as
    the oracle's Field_vectors without the extension."""
    g = q.Grid((16, 16), 0.3, "c128", n_slots=1)
    g.set_potential("H_2D")
    g.set_integrator("BM4", 0.05)
    ms = [F.Mode(intensity=1e14, wavelength=800.0, ellipticity=0.5, cep=0.3, envelope="cos4", cycles=2)]
    g.set_field_modes([ms], 400)
    Et, At, _ = g.field_batch()
    s, _, asc = O.integrator_coefficients("BM4", 0.05)
    om = oracle_modes(ms, "cos4")
    del om["angle"]
    A, E = O.field_tables(2, om, asc, s, np.arange(400) * 0.05)
    assert np.allclose(Et[0], E, rtol=0, atol=1e-13 * np.abs(E).max())
    assert np.allclose(At[0], A, rtol=0, atol=1e-13 * np.abs(A).max())
    g.close()


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_batched_run_under_device_field(gpu, prec):
    """A batched length-gauge run under the device tables equals the same run
    under the host tables of field.py (complex128: 1e-12; complex64: the
    FP32 noise floor), and instance 0 equals the FP64 oracle."""
    print(f"case batched run under device field {prec}", flush=True)
    nr, dr, dt, nt, nb = (32, 32), 0.4, 0.05, 30, 3
    o = O.Oracle(nr, dr, "H_2D", "BM4", dt, softcore=1.0)
    psi = o.initial_guess(1, guess_mode=1, seed=3)[0]
    modes = [instance_modes(b, "cos2") for b in range(nb)]
    names = ["dipole_r", "dipole_a", "norm"]
    runs = []
    for source in ("device", "host"):
        g = q.Grid(nr, dr, prec, n_slots=nb)
        g.set_potential("H_2D", softcore=1.0)
        g.set_integrator("BM4", dt)
        g.set_mask(0.1)
        if source == "device":
            g.set_field_modes(modes, nt)
        else:
            tabs = [F.tables(m, "BM4", dt, nt, 2) for m in modes]
            g.set_field_batch(np.stack([t[0] for t in tabs]), np.stack([t[1] for t in tabs]),
                              np.stack([t[2] for t in tabs]))
        for b in range(nb):
            g.upload(b, psi)
        obs = g.real_steps_batch(0, nb, "length", 0, nt, names)
        runs.append((obs, [g.download(b) for b in range(nb)]))
        g.close()
    (od, sd), (oh, sh) = runs
    tol = 1e-12 if prec == "c128" else 2e-5
    scale = np.abs(oh).max(axis=(0, 1)) + 1e-300
    assert (np.abs(od - oh) / scale).max() < tol
    for b in range(nb):
        assert np.linalg.norm(sd[b] - sh[b]) / np.linalg.norm(sh[b]) < tol
    if prec == "c128":
        Et, At, Eo, _ = F.tables(modes[0], "BM4", dt, nt, 2)
        o.set_mask(0.1)
        rp, ro = o.real_run(psi, "length", Et, At, Eo, np.zeros_like(Eo), names)
        assert np.allclose(od[0], ro, rtol=1e-11, atol=1e-12)


def test_single_instance_field_sets_both_tables(gpu):
    """nb == 1 sets the single-instance tables too: qnd_real_steps (velocity
    gauge, which reads A(t_n) as well) under the device field equals the run
    under host tables."""
    nr, dr, dt, nt = (32, 32), 0.4, 0.05, 20
    o = O.Oracle(nr, dr, "H_2D", "BM4", dt, softcore=1.0)
    psi = o.initial_guess(1, guess_mode=1, seed=9)[0]
    ms = instance_modes(1, "cos2")
    out = []
    for source in ("device", "host"):
        g = q.Grid(nr, dr, "c128", n_slots=1)
        g.set_potential("H_2D", softcore=1.0)
        g.set_integrator("BM4", dt)
        if source == "device":
            g.set_field_modes([ms], nt)
        else:
            g.set_field(*F.tables(ms, "BM4", dt, nt, 2))
        g.upload(0, psi)
        out.append(g.real_steps(0, "velocity", 0, nt, ["dipole_v", "norm"]))
        out.append(g.download(0))
        g.close()
    assert np.allclose(out[0], out[2], rtol=1e-12, atol=1e-14)
    assert np.linalg.norm(out[1] - out[3]) / np.linalg.norm(out[3]) < 1e-12


def test_field_modes_errors(gpu):
    g = q.Grid((16, 16), 0.3, "c64", n_slots=2)
    g.set_potential("H_2D")
    with pytest.raises(q.QNDError, match="integrator"):
        g.set_field_modes([[F.Mode()]], 10)
    g.set_integrator("BM4", 0.05)
    with pytest.raises(q.QNDError, match="slots"):
        g.set_field_modes([[F.Mode()]] * 3, 10)
    with pytest.raises(q.QNDError, match="wavelength"):
        g.set_field_modes([[F.Mode(wavelength=-800.0)]], 10)
    with pytest.raises(q.QNDError, match="Unknown envelope"):
        g.set_field_modes([[F.Mode(envelope="gauss")]], 10)
    with pytest.raises(q.QNDError, match="same number"):
        g.set_field_modes([[F.Mode()], [F.Mode(), F.Mode()]], 10)
    g.close()
