"""GPU parity: libqnd (HIP, gfx950) against the FP64 CPU oracle on identical
seeded inputs, through the C ABI.

Tolerances: complex128 path rel-L2 1e-12 on states and 1e-11 absolute on
energies/observables (reference parity bar); complex64 path 2e-5 relative
(FP32 storage and phases, FP64 reductions).
"""
import numpy as np
import pytest

import oracle_lib as O
import tdse_nd_mpi_amd as q

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 2e-5


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def rand_state(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return x


def smooth_state(o, seed, n_states=1):
    return o.initial_guess(n_states, guess_mode=1, seed=seed)


@pytest.mark.parametrize("prec,tol", [("c128", 1e-14), ("c64", 1e-6)])
@pytest.mark.parametrize("shape", [(8,), (16,), (32,), (64,), (128,), (256,), (512,), (1024,), (2048,),
                                   (4096,), (16, 32), (512, 512), (8, 16, 32), (32, 8, 64),
                                   (12,), (24,), (48,), (96,), (192,), (384,), (768,), (1536,), (3072,),
                                   (48, 96), (384, 384), (12, 32, 24)])
def test_transform_matches_numpy(gpu, shape, prec, tol):
    if prec == "c128" and shape[-1] > 4096:
        pytest.skip("line too long for c128")
    print(f"case transform shape={shape} prec={prec}", flush=True)
    x = rand_state(shape, 1)
    g = q.Grid(shape, 0.1, prec, n_slots=1)
    for sign in (-1, +1):
        g.upload(0, x)
        g.transform(0, sign)
        got = g.download(0)
        ref = np.fft.fftn(x) if sign < 0 else np.fft.ifftn(x) * x.size
        assert rel(got, ref) < tol, (sign, rel(got, ref))
    g.close()


IMAG_CASES = [
    ((256,), 0.2, "H_1D", "BM4"),
    ((128,), 0.3, "H_1D", "Verlet"),
    ((128,), 0.3, "H_1D", "RKN4"),
    ((64, 32), 0.4, "H_2D", "BM4"),
    ((32, 64), 0.4, "Ne_2D", "BM6"),
    ((16, 16, 16), 0.5, "H_3D", "BM4"),
    ((8, 16, 8, 8), 0.8, "H_3D", "Verlet"),
    ((96,), 0.3, "H_1D", "BM4"),                 # 3 * 2^k lines: radix-3 stage
    ((48, 24), 0.4, "H_2D", "BM4"),
]


@pytest.mark.parametrize("nr,dr,pot,integ", IMAG_CASES)
def test_imag_steps_gram_schmidt(gpu, nr, dr, pot, integ):
    print(f"case imag nr={nr} pot={pot} integ={integ}", flush=True)
    dt, nst, nsteps = 0.05, 3, 12
    o = O.Oracle(nr, dr, pot, integ, dt)
    g = q.Grid(nr, dr, "c128", n_slots=nst)
    g.set_potential(pot)
    g.set_integrator(integ, dt)
    states = smooth_state(o, 11, nst)
    ref = states.copy()
    for j in range(nst):
        g.upload(j, states[j])
        for _ in range(nsteps):
            ref[j] = o.step_imag(ref[j])
            rc, ref = o.gram_schmidt(ref, j)
            assert rc == 0
        g.imag_steps(j, nsteps, gram_schmidt=True)
    for j in range(nst):
        got = g.download(j)
        assert rel(got, ref[j]) < TOL128, (j, rel(got, ref[j]))
    E = g.energy_moments(nst - 1, 2)
    Eo = o.energies(ref[nst - 1], 2)
    assert np.allclose(E, Eo, rtol=0, atol=1e-11), (E, Eo)
    g.close()


@pytest.mark.parametrize("prec,nr,pot", [("c128", (48, 24), "H_2D"), ("c128", (16, 16, 16), "H_3D"),
                                         ("c64", (32, 16, 32), "H_3D")])
def test_gram_schmidt_split_form(gpu, monkeypatch, prec, nr, pot):
    """The Gram-Schmidt of large grids (QND_GS_SPLIT=1 forces it here): the
    dots and the update as contiguous kernels of their own (k_gs_dots,
    k_gs_update) instead of fused into the first and last pass of a step.
    Against the oracle (complex128) and against the fused form."""
    print(f"case imag GS split prec={prec} nr={nr}", flush=True)
    dr, dt, nst, nsteps = 0.5, 0.05, 4, 10
    o = O.Oracle(nr, dr, pot, "BM4", dt)
    states = smooth_state(o, 17, nst)
    out = {}
    for split in ("1", "0"):
        monkeypatch.setenv("QND_GS_SPLIT", split)
        g = q.Grid(nr, dr, prec, n_slots=nst)
        g.set_potential(pot)
        g.set_integrator("BM4", dt)
        for j in range(nst):
            g.upload(j, states[j])
            g.imag_steps(j, nsteps // 2, gram_schmidt=True)
            g.imag_steps(j, nsteps - nsteps // 2, gram_schmidt=True)
        out[split] = [g.download(j) for j in range(nst)]
        g.close()
    tol = TOL128 if prec == "c128" else TOL64
    for j in range(nst):
        assert rel(out["1"][j], out["0"][j]) < tol, (j, rel(out["1"][j], out["0"][j]))
    if prec == "c128":
        ref = states.copy()
        for j in range(nst):
            for _ in range(nsteps):
                ref[j] = o.step_imag(ref[j])
                rc, ref = o.gram_schmidt(ref, j)
                assert rc == 0
            assert rel(out["1"][j], ref[j]) < TOL128, (j, rel(out["1"][j], ref[j]))
    # orthonormal to the FP32 noise floor in complex64
    M = np.array([[np.vdot(a, b) * dr ** len(nr) for b in out["1"]] for a in out["1"]])
    assert np.abs(M - np.eye(nst)).max() < (1e-10 if prec == "c128" else 1e-5), np.abs(M - np.eye(nst)).max()


def test_gram_schmidt_ten_states(gpu):
    """Ten states: the Gram matrix of up to eight previous states comes from one
    pass over them, of nine from one reduction per pair; both against the oracle."""
    nr, dr, pot, nst, nsteps = (16, 32), 0.5, "H_2D", 10, 4
    print("case imag ten states", flush=True)
    o = O.Oracle(nr, dr, pot, "BM4", 0.05)
    g = q.Grid(nr, dr, "c128", n_slots=nst)
    g.set_potential(pot)
    g.set_integrator("BM4", 0.05)
    states = smooth_state(o, 13, nst)
    ref = states.copy()
    for j in range(nst):
        g.upload(j, states[j])
        for _ in range(nsteps):
            ref[j] = o.step_imag(ref[j])
            rc, ref = o.gram_schmidt(ref, j)
            assert rc == 0
        g.imag_steps(j, nsteps, gram_schmidt=True)
    for j in range(nst):
        assert rel(g.download(j), ref[j]) < TOL128, (j, rel(g.download(j), ref[j]))
    g.close()


@pytest.mark.parametrize("nr,dr,pot", [((128,), 0.3, "H_1D"), ((32, 64), 0.4, "Ne_2D"), ((16, 8, 16), 0.5, "H_3D")])
def test_observables_match_oracle(gpu, nr, dr, pot):
    print(f"case observables nr={nr} pot={pot}", flush=True)
    o = O.Oracle(nr, dr, pot, "BM4", 0.05)
    g = q.Grid(nr, dr, "c128", n_slots=1)
    g.set_potential(pot)
    g.set_integrator("BM4", 0.05)
    psi = smooth_state(o, 5)[0] * np.exp(0.3j * np.arange(nr[-1]) * dr)
    g.upload(0, psi)
    names = ["norm", "dipole_r", "dipole_a", "dipole_v", "energy", "Lz"]
    dim = len(nr)
    alpha = np.linspace(0.1, 0.2, dim)
    Ef = np.linspace(-0.05, 0.03, dim)
    got = g.observables(0, names, alpha=alpha, Efield=Ef)
    ref = o.observables(psi, names, alpha=alpha, Efield=Ef)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (got, ref)
    E = g.energy_moments(0, 3)
    assert np.allclose(E, o.energies(psi, 3), rtol=1e-12, atol=1e-12)
    g.close()


def field_tables(o, nt, dim, I=1e14, lam=800.0, cycles=2, ellip=0.3):
    w = 45.5637639947958 / lam
    tau = cycles * 2 * np.pi / w
    modes = dict(freq=[w], amp=[5.33803e-9 * np.sqrt(I)], ellip=[ellip], duration=[tau],
                 cep=[0.4], delay=[0.0], env=[O.ENV["cos2"]])
    a, asc = o.coefficients()
    time = np.arange(nt) * o.dt
    A, E = O.field_tables(dim, modes, asc, o.s, time)
    Aobs = np.zeros((nt, dim))
    Eobs = np.zeros((nt, dim))
    for i, t in enumerate(time):
        Aobs[i], Eobs[i] = O.field_vectors(dim, modes, t)
    return E, A, Eobs, Aobs


REAL_CASES = [
    ((256,), 0.2, "H_1D", "BM4", "length", 0.0, (0.0, 0.0)),
    ((256,), 0.2, "H_1D", "BM4", "velocity", 0.0, (0.0, 0.0)),
    ((64, 32), 0.4, "Ne_2D", "BM4", "length", 0.1, (0.0, 0.0)),
    ((64, 32), 0.4, "Ne_2D", "BM4", "velocity", 0.1, (0.0, 0.0)),
    ((32, 64), 0.4, "H_2D", "RKN4", "length", 0.0, (0.1, 3.0)),
    ((32, 64), 0.4, "H_2D", "BM6", "velocity", 0.0, (0.1, 3.0)),
    ((16, 16, 16), 0.5, "H_3D", "BM4", "length", 0.1, (0.05, 1.0)),
    ((16, 16, 16), 0.5, "H_3D", "Verlet", "velocity", 0.1, (0.05, 1.0)),
    ((384,), 0.2, "H_1D", "BM4", "velocity", 0.0, (0.1, 3.0)),
    ((48, 96), 0.4, "Ne_2D", "BM4", "length", 0.1, (0.0, 0.0)),
]


@pytest.mark.parametrize("nr,dr,pot,integ,gauge,mask,cap", REAL_CASES)
def test_real_steps_match_oracle(gpu, nr, dr, pot, integ, gauge, mask, cap):
    print(f"case real nr={nr} pot={pot} integ={integ} gauge={gauge} mask={mask} cap={cap}", flush=True)
    dt, nt = 0.05, 24
    dim = len(nr)
    o = O.Oracle(nr, dr, pot, integ, dt, cap=cap)
    if mask:
        o.set_mask(mask)
    g = q.Grid(nr, dr, "c128", n_slots=1)
    g.set_potential(pot, cap=cap)
    g.set_integrator(integ, dt)
    g.set_mask(mask)
    E, A, Eobs, Aobs = field_tables(o, nt, dim)
    g.set_field(E, A, Eobs, Aobs)
    psi0 = smooth_state(o, 3)[0]
    names = ["dipole_r", "norm", "dipole_a"]
    ref_psi, ref_obs = o.real_run(psi0, gauge, E, A, Eobs, Aobs, names)
    g.upload(0, psi0)
    obs1 = g.real_steps(0, gauge, 0, nt // 2, names)
    obs2 = g.real_steps(0, gauge, nt // 2, nt - nt // 2, names)
    got_obs = np.vstack([obs1, obs2])
    got = g.download(0)
    assert rel(got, ref_psi) < TOL128, rel(got, ref_psi)
    assert np.allclose(got_obs, ref_obs, rtol=1e-11, atol=1e-12), np.abs(got_obs - ref_obs).max()
    g.close()


@pytest.mark.parametrize("prec,gauge,tun", [("c128", "length", (8, 0, 0)), ("c128", "velocity", (8, 0, 4)),
                                            ("c64", "length", (8, 0, 0)), ("c128", "length", (4, 0, 16))])
def test_real_steps_3d_geometries(gpu, prec, gauge, tun):
    """3-D steps with eight values per thread and wide strided tiles (the
    launch shapes of large grids): the kinetic block is one pass per axis,
    the pass along axis 1 strided, with the observable record finalized by
    workgroups appended to it."""
    nr, dr, pot = (16, 32, 16), 0.5, "H_3D"
    print(f"case real-3d prec={prec} gauge={gauge} tune={tun}", flush=True)
    dt, nt = 0.05, 10
    o = O.Oracle(nr, dr, pot, "BM4", dt, cap=(0.05, 1.0))
    o.set_mask(0.1)
    g = q.Grid(nr, dr, prec, n_slots=1)
    g.set_potential(pot, cap=(0.05, 1.0))
    g.set_integrator("BM4", dt)
    g.set_mask(0.1)
    g.set_tuning(*tun)
    E, A, Eobs, Aobs = field_tables(o, nt, 3)
    g.set_field(E, A, Eobs, Aobs)
    psi0 = smooth_state(o, 4)[0]
    names = ["dipole_r", "norm", "dipole_a"]
    ref_psi, ref_obs = o.real_run(psi0, gauge, E, A, Eobs, Aobs, names)
    g.upload(0, psi0)
    got_obs = g.real_steps(0, gauge, 0, nt, names)
    got = g.download(0)
    tol = TOL128 if prec == "c128" else TOL64
    assert rel(got, ref_psi) < tol, rel(got, ref_psi)
    assert np.allclose(got_obs, ref_obs, rtol=tol * 10, atol=tol), np.abs(got_obs - ref_obs).max()
    g.close()


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_merged_boundary_matches_single_steps(gpu, prec):
    """One call of many length-gauge steps merges each step boundary (the
    record's state saved to a ring, the mask and the first kinetic factor
    split by axis over the neighbouring passes, records from one batched pass
    per ring fill; 150 steps wrap the 64-record ring twice).  One-step calls
    take the unmerged boundary pass.  Both give the same states and records."""
    nr, dr, pot = (32, 48), 0.4, "Ne_2D"
    print(f"case merged-boundary prec={prec}", flush=True)
    dt, nt = 0.05, 150
    o = O.Oracle(nr, dr, pot, "BM4", dt, cap=(0.05, 1.0))
    E, A, Eobs, Aobs = field_tables(o, nt, 2)
    psi0 = smooth_state(o, 6)[0]
    names = ["dipole_r", "norm", "dipole_a"]
    runs = []
    for split in (False, True):
        g = q.Grid(nr, dr, prec, n_slots=1)
        g.set_potential(pot, cap=(0.05, 1.0))
        g.set_integrator("BM4", dt)
        g.set_mask(0.1)
        g.set_field(E, A, Eobs, Aobs)
        g.upload(0, psi0)
        if split:
            obs = np.vstack([g.real_steps(0, "length", k, 1, names) for k in range(nt)])
        else:
            obs = g.real_steps(0, "length", 0, nt, names)
        runs.append((g.download(0), obs))
        g.close()
    (p1, o1), (p2, o2) = runs
    o.set_mask(0.1)
    ref_psi, ref_obs = o.real_run(psi0, "length", E, A, Eobs, Aobs, names)
    if prec == "c128":
        assert rel(p1, p2) < TOL128, rel(p1, p2)
        assert np.allclose(o1, o2, rtol=1e-11, atol=1e-12), np.abs(o1 - o2).max()
        assert rel(p1, ref_psi) < TOL128, rel(p1, ref_psi)
        assert np.allclose(o1, ref_obs, rtol=1e-11, atol=1e-12), np.abs(o1 - ref_obs).max()
    else:
        # two FP32 rounding sequences of the same operators: each within the
        # complex64 error of the FP64 oracle, the merged one no worse
        e1, e2 = rel(p1, ref_psi), rel(p2, ref_psi)
        r1 = np.abs(o1 - ref_obs).max()
        r2 = np.abs(o2 - ref_obs).max()
        print(f"c64 vs oracle: merged {e1:.3e} (records {r1:.3e}), one-step calls {e2:.3e} (records {r2:.3e})")
        assert e1 < 1e-4 and e2 < 1e-4, (e1, e2)
        assert e1 < 2.0 * e2 + 1e-6, (e1, e2)
        scale = np.abs(ref_obs).max()
        assert r1 < 1e-3 * scale and r1 < 2.0 * r2 + 1e-7, (r1, r2, scale)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_full_size_512_matches_oracle(gpu, prec):
    """The benchmark's grid and operators at full size: Ne_2D 512^2, DR 0.2,
    BM4 length gauge with the 800 nm field and the outer-10 % mask, 6 steps in
    one call (merged boundaries) from a smooth state, against the FP64 oracle."""
    nr, dr, pot, dt, nt = (512, 512), 0.2, "Ne_2D", 0.05, 6
    print(f"case full-size 512^2 prec={prec}", flush=True)
    o = O.Oracle(nr, dr, pot, "BM4", dt)
    o.set_mask(0.1)
    E, A, Eobs, Aobs = field_tables(o, nt, 2, ellip=0.0)
    psi0 = smooth_state(o, 21)[0]
    names = ["dipole_r", "dipole_a", "norm"]
    ref_psi, ref_obs = o.real_run(psi0, "length", E, A, Eobs, Aobs, names)
    g = q.Grid(nr, dr, prec, n_slots=1)
    g.set_potential(pot)
    g.set_integrator("BM4", dt)
    g.set_mask(0.1)
    g.set_field(E, A, Eobs, Aobs)
    g.upload(0, psi0)
    got_obs = g.real_steps(0, "length", 0, nt, names)
    got = g.download(0)
    g.close()
    if prec == "c128":
        assert rel(got, ref_psi) < TOL128, rel(got, ref_psi)
        assert np.allclose(got_obs, ref_obs, rtol=1e-11, atol=1e-12), np.abs(got_obs - ref_obs).max()
    else:
        assert rel(got, ref_psi) < TOL64, rel(got, ref_psi)
        assert np.allclose(got_obs, ref_obs, rtol=TOL64, atol=TOL64 * np.abs(ref_obs).max())


def test_unitarity_256_cube(gpu):
    """A size-independent property at 3-D size: without CAP or mask the
    complex64 real-time step is unitary, so the norm record stays constant
    (H_3D 256^3, BM4 length gauge under a field, 8 steps)."""
    n, dr, dt, nt = 256, 0.3, 0.05, 8
    print("case unitarity 256^3 c64", flush=True)
    x = -0.5 * n * dr + np.arange(n) * dr
    gx = np.exp(-(x * x) / 8.0)
    psi0 = (gx[:, None, None] * gx[None, :, None] * gx[None, None, :]).astype(np.complex128)
    psi0 = psi0 * np.exp(0.2j * x)[None, None, :]
    psi0 /= np.sqrt(np.sum(np.abs(psi0) ** 2) * dr ** 3)
    from tdse_nd_mpi_amd import field as F
    g = q.Grid((n, n, n), dr, "c64", n_slots=1)
    g.set_potential("H_3D")
    g.set_integrator("BM4", dt)
    g.set_field(*F.tables([F.Mode(intensity=1e14, cycles=2)], "BM4", dt, nt + 2, 3))
    g.upload(0, psi0)
    rec = g.real_steps(0, "length", 0, nt, ["norm"])[:, 0]
    g.close()
    assert abs(rec[0] - 1.0) < 1e-6, rec[0]
    assert np.abs(rec - rec[0]).max() < 1e-6, np.abs(rec - rec[0]).max()


@pytest.mark.parametrize("gauge", ["length", "velocity"])
def test_real_unfused_observables(gpu, gauge):
    nr, dr, pot = (32, 32), 0.4, "H_2D"
    print(f"case real-unfused gauge={gauge}", flush=True)
    dt, nt = 0.05, 6
    o = O.Oracle(nr, dr, pot, "BM4", dt)
    g = q.Grid(nr, dr, "c128", n_slots=1)
    g.set_potential(pot)
    g.set_integrator("BM4", dt)
    E, A, Eobs, Aobs = field_tables(o, nt, 2)
    g.set_field(E, A, Eobs, Aobs)
    psi0 = smooth_state(o, 9)[0]
    names = ["Lz", "energy", "dipole_r", "dipole_v", "PMD", "norm"]
    names = [n for n in names if n in O.OBS]
    ref_psi, ref_obs = o.real_run(psi0, gauge, E, A, Eobs, Aobs, names)
    g.upload(0, psi0)
    got_obs = g.real_steps(0, gauge, 0, nt, names)
    assert np.allclose(got_obs, ref_obs, rtol=1e-11, atol=1e-12), np.abs(got_obs - ref_obs).max()
    assert rel(g.download(0), ref_psi) < TOL128
    g.close()


def test_final_observables(gpu):
    nr, dr = (64, 64), 0.5
    print("case final-observables", flush=True)
    o = O.Oracle(nr, dr, "H_2D", "BM4", 0.05)
    g = q.Grid(nr, dr, "c128", n_slots=1)
    g.set_potential("H_2D")
    g.set_integrator("BM4", 0.05)
    psi = smooth_state(o, 2)[0] * np.exp(1j * 0.8 * (np.arange(64) * dr))[None, :]
    g.upload(0, psi)
    Rm, pm, Ne, de = (2.0, 6.0), (2.0, 1.5), 64, 0.02
    ped, psi_p = g.final_observables(0, Rm, pm, Ne, de)
    ref_p, ref_ped, ref_pmd, npr = o.final_observables(psi, Rm, pm, Ne, de)
    assert rel(psi_p, ref_p) < TOL128
    assert np.allclose(ped, ref_ped, rtol=1e-11, atol=1e-13)
    g.close()


@pytest.mark.parametrize("nr,dr,pot,gauge", [((256,), 0.2, "H_1D", "length"), ((64, 64), 0.4, "Ne_2D", "length"),
                                             ((64, 64), 0.4, "Ne_2D", "velocity"),
                                             ((16, 32, 16), 0.5, "H_3D", "length")])
def test_complex64_path(gpu, nr, dr, pot, gauge):
    print(f"case c64 nr={nr} gauge={gauge}", flush=True)
    dt, nt = 0.05, 40
    dim = len(nr)
    o = O.Oracle(nr, dr, pot, "BM4", dt)
    g = q.Grid(nr, dr, "c64", n_slots=2)
    g.set_potential(pot)
    g.set_integrator("BM4", dt)
    psi0 = smooth_state(o, 4, 2)
    # imaginary time with Gram-Schmidt
    ref = psi0.copy()
    for j in range(2):
        g.upload(j, psi0[j])
        for _ in range(10):
            ref[j] = o.step_imag(ref[j])
            rc, ref = o.gram_schmidt(ref, j)
        g.imag_steps(j, 10)
        assert rel(g.download(j), ref[j]) < TOL64
    # real time
    E, A, Eobs, Aobs = field_tables(o, nt, dim)
    g.set_field(E, A, Eobs, Aobs)
    names = ["dipole_r", "norm", "dipole_a"]
    ref_psi, ref_obs = o.real_run(ref[0], gauge, E, A, Eobs, Aobs, names)
    got_obs = g.real_steps(0, gauge, 0, nt, names)
    assert rel(g.download(0), ref_psi) < 1e-4
    scale = np.abs(ref_obs).max(axis=0) + 1e-3
    assert (np.abs(got_obs - ref_obs) / scale).max() < 1e-4
    g.close()


def test_gram_schmidt_collapse_error(gpu):
    nr, dr = (64,), 0.3
    g = q.Grid(nr, dr, "c128", n_slots=2)
    g.set_potential("H_1D")
    g.set_integrator("BM4", 0.05)
    o = O.Oracle(nr, dr, "H_1D", "BM4", 0.05)
    s = smooth_state(o, 1)[0]
    s = s / np.sqrt(np.sum(np.abs(s) ** 2) * dr)
    g.upload(0, s)
    g.upload(1, s)   # identical state: the projection leaves nothing
    with pytest.raises(q.QNDError, match="norm decayed"):
        g.gram_schmidt(1)
    g.close()


@pytest.mark.parametrize("tun", [(4, 0, 0), (8, 0, 0)])
def test_complex64_norm_has_no_drift(gpu, tun):
    """Unitary complex64 propagation keeps the norm: a flat potential and a small
    time step make most phase angles small (cos rounds to 1 in FP32, the
    second-order term sits in the factor's lo part), the regime in which a
    compensated product that rounds a large partial sum first loses that term
    and grows the norm by about theta^2/2 per application (3e-5 over this run
    before the product added its larger term last).  Rounding noise and the
    twiddles' fixed product order stay within a few 1e-6 here."""
    print(f"case c64 norm drift tune={tun}", flush=True)
    n, dr, dt, nt = 128, 0.2, 0.01, 3000
    x = -0.5 * n * dr + np.arange(n) * dr
    psi0 = np.exp(-(x[:, None] ** 2 + (x[None, :] - 1.0) ** 2) / 2.0) * np.exp(0.3j * x[:, None])
    psi0 /= np.sqrt(np.sum(np.abs(psi0) ** 2) * dr * dr)
    g = q.Grid((n, n), dr, "c64", n_slots=1)
    g.set_potential("Ne_2D", softcore=100.0)
    g.set_integrator("BM4", dt)
    from tdse_nd_mpi_amd import field as F
    g.set_field(*[np.zeros_like(t) for t in F.tables([F.Mode(cycles=4)], "BM4", dt, nt + 2, 2)])
    g.set_tuning(*tun)
    g.upload(0, psi0)
    norm = g.real_steps(0, "length", 0, nt, ["norm"])[:, 0]
    g.close()
    assert abs(norm[-1] - norm[0]) < 1e-5, (norm[0], norm[-1])


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("nr,gauge", [((32, 48), "length"), ((64, 32), "velocity"), ((96,), "length")])
def test_batched_real_steps(gpu, nr, gauge, prec):
    """nb instances under different fields in one launch per pass: each equals
    its own single-instance run, and instance 0 (complex128) the oracle."""
    print(f"case batch nr={nr} gauge={gauge} prec={prec}", flush=True)
    dt, nt, nb, dr, pot = 0.05, 10, 3, 0.4, "H_2D" if len(nr) == 2 else "H_1D"
    dim = len(nr)
    o = O.Oracle(nr, dr, pot, "BM4", dt)
    fields = [field_tables(o, nt, dim, I=1e14 * (1 + b), ellip=0.1 * b) for b in range(nb)]
    psi0 = smooth_state(o, 5)[0]
    names = ["dipole_r", "dipole_a", "norm"]
    g = q.Grid(nr, dr, prec, n_slots=nb)
    g.set_potential(pot)
    g.set_integrator("BM4", dt)
    g.set_mask(0.1)
    g.set_field_batch(np.stack([f[0] for f in fields]), np.stack([f[1] for f in fields]),
                      np.stack([f[2] for f in fields]))
    for b in range(nb):
        g.upload(b, psi0)
    obs = np.concatenate([g.real_steps_batch(0, nb, gauge, 0, nt // 2, names),
                          g.real_steps_batch(0, nb, gauge, nt // 2, nt - nt // 2, names)], axis=1)
    got = [g.download(b) for b in range(nb)]
    g.close()
    for b in range(nb):
        s = q.Grid(nr, dr, prec, n_slots=1)
        s.set_potential(pot)
        s.set_integrator("BM4", dt)
        s.set_mask(0.1)
        s.set_field(*fields[b])
        s.upload(0, psi0)
        one = s.real_steps(0, gauge, 0, nt, names)
        # the batch may pick another launch geometry (values per thread), so
        # complex64 agrees to its own rounding, complex128 to FP64 rounding
        tol = 1e-12 if prec == "c128" else TOL64
        assert rel(got[b], s.download(0)) < tol, (b, rel(got[b], s.download(0)))
        if prec == "c128":
            assert np.allclose(obs[b], one, rtol=tol, atol=tol * 1e-2), (b, np.abs(obs[b] - one).max())
        else:
            # two FP32 rounding sequences: compared at each record's scale (a
            # component crossing zero has no relative accuracy of its own)
            err = np.abs(obs[b] - one) / (np.abs(one).max(axis=0) + 1e-300)
            assert err.max() < tol, (b, err.max(axis=0))
        s.close()
    if prec == "c128":
        o.set_mask(0.1)
        ref_psi, ref_obs = o.real_run(psi0, gauge, *fields[0], names)
        assert rel(got[0], ref_psi) < TOL128, rel(got[0], ref_psi)
        assert np.allclose(obs[0], ref_obs, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("prec,nr", [("c128", (64, 128)), ("c128", (48, 96)), ("c64", (512, 32))])
def test_tuning_geometries_agree(gpu, prec, nr):
    """Every launch geometry the autotuner may pick gives the same step: up to
    FP64 rounding in complex128; in complex64 bit for bit among geometries with
    the same values per thread (same radix sequence), and the tuner restores
    the state it timed on."""
    from tdse_nd_mpi_amd import tuning
    print(f"case tuning {prec} nr={nr}", flush=True)
    dt, nt, dr = 0.05, 8, 0.4
    o = O.Oracle(nr, dr, "H_2D", "BM4", dt)
    E, A, Eo, Ao = field_tables(o, nt + 8, 2)
    psi0 = smooth_state(o, 2)[0]
    g = q.Grid(nr, dr, prec, n_slots=1)
    g.set_potential("H_2D")
    g.set_integrator("BM4", dt)
    g.set_field(E, A, Eo, Ao)
    g.set_mask(0.1)
    ref = {}
    for cand in tuning.candidates(nr, prec):
        g.set_tuning(*cand)
        g.upload(0, psi0)
        obs = g.real_steps(0, "length", 0, nt, ["dipole_r", "dipole_a", "norm"])
        got = g.download(0)
        key = cand[0] if prec == "c64" else 0
        if key not in ref:
            ref[key] = (got, obs)
        r_got, r_obs = ref[key]
        if prec == "c64":
            assert np.array_equal(got, r_got), (cand, rel(got, r_got))
        else:
            assert rel(got, r_got) < 1e-13, (cand, rel(got, r_got))
        assert np.allclose(obs, r_obs, rtol=1e-11, atol=1e-12), cand
    g.upload(0, psi0)
    best, table = tuning.autotune_batch(g, 0, 1, "length", 0, nsteps=2, cands=tuning.candidates(nr, prec)[:3])
    stored = psi0.astype(np.complex64).astype(np.complex128) if prec == "c64" else psi0
    assert rel(g.download(0), stored) < 1e-15 and len(table) == 3
    g.close()


def _adjoint_problem(prec, nr=(64, 64), nsteps=40, every=8, dr=0.4, pot="H_2D"):
    from tdse_nd_mpi_amd import adjoint as adj
    rng = np.random.default_rng(7)
    o = O.Oracle(nr, dr, pot, "BM4", 0.05)
    psi0 = smooth_state(o, 4)[0]
    psi0 = psi0 / np.sqrt(np.sum(np.abs(psi0) ** 2) * dr ** len(nr))
    t = np.arange(nsteps) * 0.05
    E = 0.05 * np.sin(0.5 * t) * np.exp(-((t - 1.0) / 1.2) ** 2) + 0.01 * rng.standard_normal(nsteps)
    spec = adj.Spec(nr, dr, pot, "BM4", 0.05, psi0, precision=prec, mask=0.1, every=every)
    wa = rng.standard_normal((nsteps, len(nr)))
    return adj, spec, E, wa


def _linear_loss(adj, spec, E, soft, wa, gnorm=0.7):
    a, _, nf = adj.forward_numpy(spec, E, soft)
    return float(np.sum(wa * a) + gnorm * nf)


def test_adjoint_matches_finite_differences(gpu):
    """dL/dE(t_n) and dL/da (soft-core) of L = sum w.a + 0.7 norm_final,
    complex128, against central differences of the forward run."""
    print("case adjoint fd", flush=True)
    adj, spec, E, wa = _adjoint_problem("c128")
    soft = 1.0
    adj.forward_numpy(spec, E, soft)
    dEt, dEo, ds = spec.grid.adjoint(wa, 0.7)
    dE = dEt[:, :, 0].sum(axis=1) + dEo[:, 0]
    h = 1e-4
    for k in [0, 3, 11, 20, 27, 39]:
        Ep, Em = E.copy(), E.copy()
        Ep[k] += h
        Em[k] -= h
        fd = (_linear_loss(adj, spec, Ep, soft, wa) - _linear_loss(adj, spec, Em, soft, wa)) / (2 * h)
        assert abs(fd - dE[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, dE[k])
    hs = 1e-4
    fd = (_linear_loss(adj, spec, E, soft + hs, wa) - _linear_loss(adj, spec, E, soft - hs, wa)) / (2 * hs)
    assert abs(fd - ds) <= 1e-6 * max(1.0, abs(fd)), (fd, ds)
    spec.close()


def test_adjoint_3d_matches_finite_differences(gpu):
    """The 3-D adjoint (kinetic block of one conjugated pass per axis) against
    central differences of the forward run, complex128."""
    print("case adjoint fd 3-D", flush=True)
    adj, spec, E, wa = _adjoint_problem("c128", nr=(16, 16, 16), nsteps=16, every=4, dr=0.5, pot="H_3D")
    soft = 1.0
    adj.forward_numpy(spec, E, soft)
    dEt, dEo, ds = spec.grid.adjoint(wa, 0.7)
    dE = dEt[:, :, 0].sum(axis=1) + dEo[:, 0]
    h = 1e-4
    for k in [0, 5, 15]:
        Ep, Em = E.copy(), E.copy()
        Ep[k] += h
        Em[k] -= h
        fd = (_linear_loss(adj, spec, Ep, soft, wa) - _linear_loss(adj, spec, Em, soft, wa)) / (2 * h)
        assert abs(fd - dE[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, dE[k])
    fd = (_linear_loss(adj, spec, E, soft + h, wa) - _linear_loss(adj, spec, E, soft - h, wa)) / (2 * h)
    assert abs(fd - ds) <= 1e-6 * max(1.0, abs(fd)), (fd, ds)
    spec.close()


def test_adjoint_complex64_tracks_complex128(gpu):
    print("case adjoint c64", flush=True)
    grads = {}
    for prec in ("c128", "c64"):
        adj, spec, E, wa = _adjoint_problem(prec)
        adj.forward_numpy(spec, E, 1.0)
        dEt, dEo, ds = spec.grid.adjoint(wa, 0.7)
        grads[prec] = (dEt[:, :, 0].sum(axis=1) + dEo[:, 0], ds)
        spec.close()
    assert rel(grads["c64"][0], grads["c128"][0]) < 1e-4, rel(grads["c64"][0], grads["c128"][0])
    assert abs(grads["c64"][1] - grads["c128"][1]) < 1e-4 * max(1.0, abs(grads["c128"][1]))


def test_autograd_function_harmonic_loss(gpu):
    """torch.autograd through Propagate with the harmonic-window loss equals
    central differences of the same loss (complex128)."""
    import torch
    print("case adjoint autograd", flush=True)
    adj, spec, E, _ = _adjoint_problem("c128", nsteps=48, every=16)
    Et = torch.tensor(E, requires_grad=True)
    soft = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    a, nf = adj.Propagate.apply(Et, soft, spec)
    L = adj.harmonic_loss(a, nf, 0.05, 0.8, orders=(2, 5))
    L.backward()

    def loss(Ev, sv):
        a2, nf2 = adj.Propagate.apply(torch.tensor(Ev), torch.tensor(sv, dtype=torch.float64), spec)
        return float(adj.harmonic_loss(a2, nf2, 0.05, 0.8, orders=(2, 5)))

    h = 1e-4
    for k in [5, 24, 40]:
        Ep, Em = E.copy(), E.copy()
        Ep[k] += h
        Em[k] -= h
        fd = (loss(Ep, 1.0) - loss(Em, 1.0)) / (2 * h)
        assert abs(fd - Et.grad[k].item()) <= 1e-6 * max(1e-3, abs(fd)), (k, fd, Et.grad[k].item())
    fd = (loss(E, 1.0 + h) - loss(E, 1.0 - h)) / (2 * h)
    assert abs(fd - soft.grad.item()) <= 1e-6 * max(1e-3, abs(fd)), (fd, soft.grad.item())
    spec.close()


@pytest.mark.parametrize("wave_lines", ["1", "0"])
def test_wavefront_local_exchanges(gpu, wave_lines, monkeypatch):
    """Eight values per thread on 512-point lines: each contiguous line is one
    wavefront and exchanges without workgroup barriers (or with them when
    disabled); transforms and steps match numpy and the oracle."""
    print(f"case wave-lines={wave_lines}", flush=True)
    monkeypatch.setenv("QND_WAVE_LINES", wave_lines)
    for shape in [(512,), (8, 512), (8, 16, 512)]:
        x = rand_state(shape, 3)
        g = q.Grid(shape, 0.1, "c128", n_slots=1)
        g.set_tuning(8, 0, 0)
        for sign in (-1, +1):
            g.upload(0, x)
            g.transform(0, sign)
            ref = np.fft.fftn(x) if sign < 0 else np.fft.ifftn(x) * x.size
            assert rel(g.download(0), ref) < 1e-14, (shape, sign)
        g.close()
    nr, dt, nt = (64, 512), 0.05, 12
    o = O.Oracle(nr, 0.2, "Ne_2D", "BM4", dt)
    E, A, Eo, Ao = field_tables(o, nt, 2)
    psi0 = smooth_state(o, 6)[0]
    g = q.Grid(nr, 0.2, "c128", n_slots=1)
    g.set_tuning(8, 0, 0)
    g.set_potential("Ne_2D")
    g.set_integrator("BM4", dt)
    g.set_field(E, A, Eo, Ao)
    g.upload(0, psi0)
    obs = g.real_steps(0, "length", 0, nt, ["dipole_r", "norm"])
    ref_psi, ref_obs = o.real_run(psi0, "length", E, A, Eo, Ao, ["dipole_r", "norm"])
    assert rel(g.download(0), ref_psi) < TOL128
    assert np.allclose(obs, ref_obs, rtol=1e-11, atol=1e-12)
    g.close()


def test_adjoint_rejects_changed_setup(gpu):
    """The costate needs the forward run's operators: a field or mask change in
    between is an error, not a silently wrong gradient."""
    print("case adjoint setup change", flush=True)
    adj, spec, E, wa = _adjoint_problem("c128", nr=(32, 32), nsteps=8, every=4)
    adj.forward_numpy(spec, E, 1.0)
    spec.grid.set_mask(0.2)
    with pytest.raises(q.QNDError, match="changed since"):
        spec.grid.adjoint(wa, 0.7)
    adj.forward_numpy(spec, E, 1.0)
    dEt, dEo, ds = spec.grid.adjoint(wa, 0.7)
    assert np.isfinite(dEt).all() and np.isfinite(ds)
    spec.close()
