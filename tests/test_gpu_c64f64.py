"""complex64 storage with FP64 arithmetic in the passes (precision "c64f64",
QND_C64F64): the state is stored as complex64 and each pass holds it in FP64
registers, so a value is rounded to complex64 once per pass, at its store.
Checked against the FP64 oracle, against the complex128 path, and against
the complex64 path (its error must be the smaller one)."""
import numpy as np
import pytest

import oracle_lib as O
import tdse_nd_mpi_amd as q

from test_gpu_parity import field_tables, rel, smooth_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(512,), (64, 128), (384,), (8, 16, 32)])
def test_transform(gpu, shape):
    print(f"case c64f64 transform {shape}", flush=True)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    g = q.Grid(shape, 0.1, "c64f64", n_slots=1)
    g.upload(0, x)
    g.transform(0, -1)
    got = g.download(0)
    ref = np.fft.fftn(x)
    # one rounding to complex64 of the upload and one of the stored result
    assert rel(got, ref) < 2e-7, rel(got, ref)
    g.close()


@pytest.mark.parametrize("nr,dr,pot,gauge", [((64, 64), 0.4, "Ne_2D", "length"), ((64, 64), 0.4, "H_2D", "velocity"),
                                           ((16, 32, 16), 0.5, "H_3D", "length"), ((256,), 0.2, "H_1D", "length"),
                                           ((48, 96), 0.4, "Ne_2D", "length"), ((96,), 0.3, "H_1D", "velocity")])
def test_steps_match_oracle(gpu, nr, dr, pot, gauge):
    """Imaginary time with Gram-Schmidt, then real time with CAP, mask and a
    field, against the FP64 oracle; the complex64 path's error on the same
    run is the upper bound of this one's."""
    print(f"case c64f64 steps nr={nr} pot={pot} gauge={gauge}", flush=True)
    dt, nt = 0.05, 60
    dim = len(nr)
    o = O.Oracle(nr, dr, pot, "BM4", dt, cap=(0.05, 1.0))
    o.set_mask(0.1)
    psi0 = smooth_state(o, 6, 2)
    ref = psi0.copy()
    for j in range(2):
        for _ in range(10):
            ref[j] = o.step_imag(ref[j])
            rc, ref = o.gram_schmidt(ref, j)
    E, A, Eobs, Aobs = field_tables(o, nt, dim)
    names = ["dipole_r", "norm", "dipole_a"]
    ref_psi, ref_obs = o.real_run(ref[0], gauge, E, A, Eobs, Aobs, names)
    errs = {}
    for prec in ("c64f64", "c64"):
        g = q.Grid(nr, dr, prec, n_slots=2)
        g.set_potential(pot, cap=(0.05, 1.0))
        g.set_integrator("BM4", dt)
        g.set_mask(0.1)
        for j in range(2):
            g.upload(j, psi0[j])
            g.imag_steps(j, 10)
        e_imag = max(rel(g.download(j), ref[j]) for j in range(2))
        g.set_field(E, A, Eobs, Aobs)
        g.upload(0, ref[0])
        got_obs = g.real_steps(0, gauge, 0, nt, names)
        scale = np.abs(ref_obs).max(axis=0) + 1e-3
        errs[prec] = (e_imag, rel(g.download(0), ref_psi), (np.abs(got_obs - ref_obs) / scale).max())
        g.close()
    e_imag, e_psi, e_obs = errs["c64f64"]
    # records: the acceleration a = -<grad V> - E<psi|psi> magnifies the
    # state's error (a difference of two terms much larger than a)
    assert e_imag < 5e-6 and e_psi < 5e-6 and e_obs < 1e-4, errs
    assert e_psi <= errs["c64"][1] and e_obs <= errs["c64"][2], errs


@pytest.mark.parametrize("gauge", ["length", "velocity"])
def test_batched_and_merged_paths(gpu, gauge):
    """The batched step (one launch over instances, per-instance fields) and,
    in the length gauge, the merged step boundary of long 2-D calls, against
    complex128."""
    print(f"case c64f64 batched + merged boundary {gauge}", flush=True)
    nr, dr, dt, nt, nb = (48, 64), 0.4, 0.05, 150, 3
    o = O.Oracle(nr, dr, "H_2D", "BM4", dt, cap=(0.05, 1.0))
    psi = smooth_state(o, 9)[0]
    rng = np.random.default_rng(3)
    tabs = [field_tables(o, nt, 2, I=float(10 ** rng.uniform(13.7, 14.3))) for _ in range(nb)]
    names = ["dipole_r", "norm", "dipole_a"]
    out = {}
    for prec in ("c64f64", "c128"):
        g = q.Grid(nr, dr, prec, n_slots=nb)
        g.set_potential("H_2D", cap=(0.05, 1.0))
        g.set_integrator("BM4", dt)
        g.set_mask(0.1)
        g.set_field_batch(np.stack([t[0] for t in tabs]), np.stack([t[1] for t in tabs]),
                          np.stack([t[2] for t in tabs]))
        for b in range(nb):
            g.upload(b, psi)
        ob = g.real_steps_batch(0, nb, gauge, 0, nt, names)
        # one instance alone through the single-instance path (merged boundary)
        g.set_field(*tabs[1])
        g.upload(0, psi)
        o1 = g.real_steps(0, gauge, 0, nt, names)
        out[prec] = (ob, o1, g.download(0))
        g.close()
    (ob, o1, p1), (rb, r1, rp) = out["c64f64"], out["c128"]
    scale = np.abs(rb).max(axis=(0, 1)) + 1e-3
    assert (np.abs(ob - rb) / scale).max() < 1e-5
    assert (np.abs(o1 - r1) / (np.abs(r1).max(axis=0) + 1e-3)).max() < 1e-5
    assert rel(p1, rp) < 1e-5


def test_group_refuses_c64f64(gpu):
    a = q.Grid((32, 32), 0.4, "c64f64", n_slots=1)
    a.set_potential("H_2D")
    a.set_integrator("BM4", 0.05)
    a.set_field_batch(np.zeros((1, 4, 6, 2)), np.zeros((1, 4, 6, 2)))
    with pytest.raises(q.QNDError, match="FP64-arithmetic"):
        q.group_real_steps([a], "length", 0, 2)
    a.close()


def test_adjoint_tracks_complex128(gpu):
    """The adjoint gradient (checkpointed forward, recompute, costate sweep)
    in c64f64 against complex128; its error is at most the complex64 one."""
    from test_gpu_parity import _adjoint_problem
    print("case c64f64 adjoint", flush=True)
    grads = {}
    for prec in ("c128", "c64f64", "c64"):
        adj, spec, E, wa = _adjoint_problem(prec)
        adj.forward_numpy(spec, E, 1.0)
        dEt, dEo, ds = spec.grid.adjoint(wa, 0.7)
        grads[prec] = (dEt[:, :, 0].sum(axis=1) + dEo[:, 0], ds)
        spec.close()
    ref = grads["c128"]
    e = {p: (rel(grads[p][0], ref[0]), abs(grads[p][1] - ref[1]) / max(1.0, abs(ref[1]))) for p in ("c64f64", "c64")}
    print(e, flush=True)
    assert e["c64f64"][0] < 2e-5 and e["c64f64"][1] < 2e-5, e
    assert e["c64f64"][0] <= e["c64"][0], e
