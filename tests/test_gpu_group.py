"""Several grids in one launch per pass (qnd_group_real_steps) against each
grid's own batched run (qnd_real_steps_batch) and, for one instance, the FP64
oracle.  The group's steps after the first run from one launch sequence (a
device step counter advances the field rows and records); with use_graph that
sequence is a captured HIP graph, replayed once per step.  In the branch
form each grid runs its own kernels on its own stream, forked and joined every
step, and one captured graph holds up to 32 steps (a call of 70 steps runs
one step eagerly, replays a 32-step graph twice, then runs a 5-step one).  Each call's records (every replay)
and final states are checked against the eager results, with the program's
own eager work and synchronizations between the calls."""
import numpy as np
import pytest

import oracle_lib as O
import tdse_nd_mpi_amd as q
from tdse_nd_mpi_amd import field as F

pytestmark = pytest.mark.gpu

# (grid, instances): a power-of-two grid with four values per thread, one with
# eight, and a 3 * 2^k grid (twelve values per thread, radix-3 stage)
MEMBERS = [((32, 32), 3), ((48, 48), 2), ((128, 64), 2)]
DR, DT, NT = 0.4, 0.05, 140
NAMES = ["dipole_r", "dipole_a", "norm"]


def member_inputs(k, nr, nb):
    rng = np.random.default_rng(100 + k)
    o = O.Oracle(nr, DR, "H_2D", "BM4", DT, softcore=1.0)
    psi = o.initial_guess(1, guess_mode=1, seed=40 + k)[0]
    tabs = [F.tables([F.Mode(intensity=float(10 ** rng.uniform(13.7, 14.7)), wavelength=800.0,
                             cep=float(rng.uniform(0, 2 * np.pi)), envelope="cos2", cycles=1)],
                     "BM4", DT, NT, 2) for _ in range(nb)]
    return o, psi, tabs


def make_grid(nr, nb, prec, psi, tabs):
    g = q.Grid(nr, DR, prec, n_slots=nb)
    g.set_potential("H_2D", softcore=1.0)
    g.set_integrator("BM4", DT)
    g.set_mask(0.1)
    g.set_field_batch(np.stack([t[0] for t in tabs]), np.stack([t[1] for t in tabs]),
                      np.stack([t[2] for t in tabs]))
    for b in range(nb):
        g.upload(b, psi)
    return g


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("branches", [False, True])
def test_group_matches_member_batches(gpu, prec, use_graph, branches):
    print(f"case group prec={prec} graph={use_graph} branches={branches}", flush=True)
    inputs = [member_inputs(k, nr, nb) for k, (nr, nb) in enumerate(MEMBERS)]
    # each member alone, its batch in one launch per pass
    ref_obs, ref_psi = [], []
    for (nr, nb), (o, psi, tabs) in zip(MEMBERS, inputs):
        g = make_grid(nr, nb, prec, psi, tabs)
        ref_obs.append(g.real_steps_batch(0, nb, "length", 0, NT, NAMES))
        ref_psi.append([g.download(b) for b in range(nb)])
        g.close()
    # the group: two calls, eager work in between
    grids = [make_grid(nr, nb, prec, psi, tabs) for (nr, nb), (o, psi, tabs) in zip(MEMBERS, inputs)]
    half = NT // 2
    o1 = q.group_real_steps(grids, "length", 0, half, NAMES, use_graph=use_graph, branches=branches)
    _ = grids[0].observables(1, ["norm"])          # eager work on a member between the calls
    grids[1].synchronize()
    o2 = q.group_real_steps(grids, "length", half, NT - half, NAMES, use_graph=use_graph, branches=branches)
    # complex64: the eager observable call sends member 0 through one more
    # inverse / forward transform pair than its reference run, another FP32
    # rounding sequence
    tol = 1e-12 if prec == "c128" else 2e-5
    for k, ((nr, nb), g) in enumerate(zip(MEMBERS, grids)):
        got = np.concatenate([o1[k], o2[k]], axis=1)
        scale = np.abs(ref_obs[k]).max(axis=(0, 1)) + 1e-300
        err = (np.abs(got - ref_obs[k]) / scale).max()
        assert err < tol, (k, err)
        for b in range(nb):
            e = rel(g.download(b), ref_psi[k][b])
            assert e < tol, (k, b, e)
        g.close()
    if prec == "c128":
        # instance 0 of the 3 * 2^k member against the FP64 oracle
        (nr, nb), (o, psi, tabs) = MEMBERS[1], inputs[1]
        o.set_mask(0.1)
        rp, ro = o.real_run(psi, "length", *tabs[0][:3], tabs[0][3], NAMES)
        got = np.concatenate([o1[1][0], o2[1][0]], axis=0)
        assert np.allclose(got, ro, rtol=1e-11, atol=1e-12), np.abs(got - ro).max()
        assert rel(ref_psi[1][0], rp) < 1e-12


def test_group_rejects_mismatched_members(gpu):
    a = q.Grid((32, 32), DR, "c64", n_slots=1)
    b = q.Grid((32, 32), DR, "c128", n_slots=1)
    for g in (a, b):
        g.set_potential("H_2D")
        g.set_integrator("BM4", DT)
        g.set_field_batch(np.zeros((1, 4, 6, 2)), np.zeros((1, 4, 6, 2)))
    with pytest.raises(q.QNDError, match="one precision"):
        q.group_real_steps([a, b], "length", 0, 2)
    b.close()
    c = q.Grid((32, 32), DR, "c64", n_slots=1)
    c.set_potential("H_2D")
    c.set_integrator("BM4", 0.1)
    c.set_field_batch(np.zeros((1, 4, 6, 2)), np.zeros((1, 4, 6, 2)))
    with pytest.raises(q.QNDError, match="integrator and DT"):
        q.group_real_steps([a, c], "length", 0, 2)
    with pytest.raises(q.QNDError, match="outside"):
        q.group_real_steps([a], "length", 2, 4)
    a.close()
    c.close()
