"""Benchmark of the split-operator hot path on one MI355X.

Workload (the headline configuration): Ne_2D on a 512x512 grid, DR = 0.2;
initial state = the 2-D ground state from imaginary time (BM4, DT 0.1,
T 100 -> 1001 steps); real time with BM4, DT 0.05 over a 4-cycle sin^2
(cos2) 800 nm pulse at 1e14 W/cm^2, CEP 0, polarised along x, length gauge,
cos^(1/8) absorbing mask on the outer 10 % of each half-width; dipole <x>,<y>
and dipole acceleration recorded every step (8825 steps).

A "step" is one real-time step of the whole grid.  `value` is the
mixed-precision path (complex64 state, FP32 phases from FP64-reduced angles,
FP64 reductions) in grid-point-steps/s with the state resident in HBM; the
FP64 path, their ratio, the roofline of the dominant kernel and the CPU
baseline (FP64 oracle on the host cores) are reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, AMD datasheet
NR = (512, 512)
DR = 0.2
DT = 0.05
MASK = 0.1
POT = "Ne_2D"
METRIC = "grid-point-steps/s and % HBM roofline, FP32 vs FP64; adjoint/forward time ratio"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def reference_guess(nr, dr):
    axes = [-0.5 * n * dr + np.arange(n) * dr for n in nr]
    X = np.meshgrid(*axes, indexing="ij")
    rad = np.sqrt(sum(x * x for x in X))
    idx = np.arange(int(np.prod(nr))).reshape(nr)
    return ((rad + (idx % 3)) * np.exp(-rad)).astype(np.complex128)


def ground_state(q):
    g = q.Grid(NR, DR, "c128", n_slots=1)
    g.set_potential(POT)
    g.set_integrator("BM4", 0.1)
    g.upload(0, reference_guess(NR, DR))
    nsteps = q.imag_step_count(100.0, 0.1)
    t0 = time.perf_counter()
    g.imag_steps(0, nsteps, gram_schmidt=True)
    g.synchronize()
    t_imag = time.perf_counter() - t0
    E = g.energy_moments(0, 2)
    psi = g.download(0)
    g.close()
    return psi, E, nsteps, t_imag


def real_run(q, field, prec, psi0, warmup, steps, Et, At, Eo, Ao, profile_steps=0, tune=True, fixed=None):
    g = q.Grid(NR, DR, prec, n_slots=1)
    g.set_potential(POT)
    g.set_integrator("BM4", DT)
    g.set_mask(MASK)
    g.set_field(Et, At, Eo, Ao)
    g.upload(0, psi0)
    from tdse_nd_mpi_amd import tuning
    if fixed:
        # a fixed launch geometry: the same kernels with and without a profiler
        g.set_tuning(*fixed)
        best, table = tuple(fixed), []
    else:
        best, table = (tuning.autotune_batch(g, 0, 1, "length", 0, nsteps=8)   # restores psi0
                       if tune else ((0, 0, 0), []))
    names = ["dipole_r", "dipole_a", "norm"]
    obs_w = g.real_steps(0, "length", 0, warmup, names) if warmup else np.zeros((0, 5))
    import torch
    torch.cuda.synchronize()
    g.synchronize()
    g.event_record(0)
    t0 = time.perf_counter()
    obs_t = g.real_steps(0, "length", warmup, steps, names)
    g.event_record(1)
    g.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = g.event_elapsed(0, 1)
    prof = None
    if profile_steps:
        g.profile(True)
        g.upload(0, psi0)
        g.real_steps(0, "length", 0, profile_steps, names)
        cnt, ms = g.profile_read()
        g.profile(False)
        prof = (cnt, ms)
    g.close()
    return np.vstack([obs_w, obs_t]), wall, gpu_ms, prof, (best, table)


# ---------------------------------------------------------------------
# CEP / intensity sweep: 96 H_2D instances (soft-core a = 1) on uneven grids,
# 32 each at 128^2, 256^2, 384^2, DR 0.25; CEP uniform in [0, 2 pi) and
# intensity log-uniform in [5e13, 5e14] W/cm^2 from a seed; 800 nm 3-cycle
# cos2 pulse, BM4, DT 0.05, complex64 with FP64 reductions.  One batched
# context per grid size (every pass is one launch over its 32 instances),
# launch geometry autotuned, the three classes running concurrently.
# ---------------------------------------------------------------------
SWEEP_SIZES = (128, 256, 384)
SWEEP_PER = 32
SWEEP_DR = 0.25
SWEEP_SEED = 20261203


def sweep_fields(F):
    """The sweep's per-instance laser modes (random CEP and intensity); their
    field tables are formed on the device (qnd_set_field_modes)."""
    rng = np.random.default_rng(SWEEP_SEED)
    nb = SWEEP_PER * len(SWEEP_SIZES)
    ceps = rng.uniform(0.0, 2 * np.pi, nb)
    ints = 10.0 ** rng.uniform(np.log10(5e13), np.log10(5e14), nb)
    modes = [[F.Mode(intensity=ints[i], wavelength=800.0, ellipticity=0.0, cep=ceps[i], envelope="cos2",
                     cycles=3)] for i in range(nb)]
    ntime_full = int(F.Mode(cycles=3).tau / DT)
    return ceps, ints, modes, ntime_full


def sweep_ground_state(q, n):
    g = q.Grid((n, n), SWEEP_DR, "c128", n_slots=1)
    g.set_potential("H_2D", softcore=1.0)
    g.set_integrator("BM4", 0.1)
    g.upload(0, reference_guess((n, n), SWEEP_DR))
    g.imag_steps(0, q.imag_step_count(100.0, 0.1), gram_schmidt=True)
    E = g.energy_moments(0, 1)[0]
    psi = g.download(0)
    g.close()
    return psi, E


def sweep_run(q, F, steps, warmup, compare_steps=1000):
    """The 96-instance sweep, timed over `steps` steps from the ground states
    (steps <= 0: the whole pulse, 6619 steps) in three forms:
      one_launch: every pass of a step is one kernel launch over all three
        size classes (qnd_group_real_steps), the steps after the first a
        replayed HIP graph (also timed without the graph over compare_steps);
      graph_branches: qnd_group_real_steps in its branch form: each size
        class's own kernels on its own stream, forked and joined once per
        chunk, 32 steps of all 96 instances per launch of a captured HIP graph;
      three_contexts: one batched context per size class (one launch per
        pass per class), the three on three host threads.
    `value` is the fastest form."""
    import threading
    import torch
    ceps, ints, modes, ntime_full = sweep_fields(F)
    nsteps = ntime_full if steps <= 0 else steps
    ntime = max(nsteps, compare_steps) + warmup + 8
    names = ["dipole_r", "dipole_a"]
    classes = []
    for k, n in enumerate(SWEEP_SIZES):
        psi0, E0 = sweep_ground_state(q, n)
        g = q.Grid((n, n), SWEEP_DR, "c64", n_slots=SWEEP_PER)
        g.set_potential("H_2D", softcore=1.0)
        g.set_integrator("BM4", DT)
        sl = slice(k * SWEEP_PER, (k + 1) * SWEEP_PER)
        g.set_field_modes(modes[sl], ntime)
        for b in range(SWEEP_PER):
            g.upload(b, psi0)
        from tdse_nd_mpi_amd import tuning
        best, table = tuning.autotune_batch(g, 0, SWEEP_PER, "length", 0, nsteps=4)
        classes.append(dict(n=n, grid=g, psi0=psi0, E0=E0, best=best, table=table,
                            cached=getattr(g, "tuning_cached", False)))
    grids = [c["grid"] for c in classes]

    def reset():
        for c in classes:
            for b in range(SWEEP_PER):
                c["grid"].upload(b, c["psi0"])

    def three_contexts(n):
        res3 = [None] * len(classes)

        def work(i):
            res3[i] = classes[i]["grid"].real_steps_batch(0, SWEEP_PER, "length", 0, n, names)

        th = [threading.Thread(target=work, args=(i,)) for i in range(len(classes))]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        return time.perf_counter() - t0, res3

    pts = sum(SWEEP_PER * c["n"] ** 2 for c in classes)
    q.group_real_steps(grids, "length", 0, warmup, names, use_graph=True)      # warm-up, then from t = 0
    reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out1 = q.group_real_steps(grids, "length", 0, nsteps, names, use_graph=True)
    wall1 = time.perf_counter() - t0
    reset()
    t0 = time.perf_counter()
    q.group_real_steps(grids, "length", 0, compare_steps, names, use_graph=False)
    wall_ng = time.perf_counter() - t0
    reset()
    q.group_real_steps(grids, "length", 0, warmup, names, use_graph=True, branches=True)
    reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outb = q.group_real_steps(grids, "length", 0, nsteps, names, use_graph=True, branches=True)
    wallb = time.perf_counter() - t0
    reset()
    three_contexts(warmup)
    reset()
    wall3, out3 = three_contexts(nsteps)
    v1, vb, v3 = pts * nsteps / wall1, pts * nsteps / wallb, pts * nsteps / wall3
    forms = {"one_launch": v1, "graph_branches": vb, "three_contexts": v3}
    value = max(forms.values())
    agree = max(float(np.abs(a - b).max() / (np.abs(b).max() + 1e-300))
                for o in (out1, outb) for a, b in zip(o, out3))
    res = {"metric": "grid-point-steps/s (all 96 instances)", "value": value,
           "form": max(forms, key=forms.get), "steps": nsteps,
           "steps_full_pulse": ntime_full, "points_per_step": pts, "seed": SWEEP_SEED,
           "step_roofline_frac": 14 * 2 * 8 * value / (HBM_PEAK_GBS * 1e9),
           "one_launch": {"value": v1, "wall_s": wall1,
                          "launch": "one kernel launch per pass over all 96 instances of the three sizes "
                                    "(qnd_group_real_steps); steps after the first: one captured HIP graph "
                                    "replayed per step",
                          "without_graph": {"steps": compare_steps, "value": pts * compare_steps / wall_ng}},
           "graph_branches": {"value": vb, "wall_s": wallb,
                              "launch": "each size class's own kernels on its own stream, forked and joined once per "
                                        "chunk (qnd_group_real_steps, branch form); 32 steps of all three classes "
                                        "per launch of a captured HIP graph"},
           "three_contexts": {"value": v3, "wall_s": wall3,
                              "launch": "one batched context per size (one launch per pass per size), three "
                                        "host threads"},
           "records_max_rel_diff_between_forms": agree,
           "classes": [{"grid": [c["n"], c["n"]], "instances": SWEEP_PER, "E0": c["E0"],
                        "tuned_vpt_twc_tws": list(c["best"]), "tuning_cached": bool(c["cached"]),
                        "tuning_ms_per_step": {str(list(k)): round(v, 4) for k, v in c["table"]}}
                       for c in classes],
           "finite": bool(all(np.isfinite(o).all() for o in out1 + outb + out3))}
    for c in classes:
        c["grid"].close()
    return res


# ---------------------------------------------------------------------
# Adjoint gradient: H_2D (a = 1) on 256^2, DR 0.3, 4000 BM4 steps at DT 0.05
# under a seeded smooth random pulse (8 Gaussian-windowed cosines, peak 0.05
# a.u., x-polarized, held over the stages of a step); loss = power of the
# dipole-acceleration spectrum in windows around harmonics 11..21 of 800 nm
# plus the norm lost to the absorber (cos^(1/8) mask); dLoss/dE(t_n) and
# dLoss/d(soft-core a) in complex64 and complex128 with checkpoints every 64
# steps, checked against central differences on 16 random E samples (FP64).
# ---------------------------------------------------------------------
ADJ_NR = (256, 256)
ADJ_DR = 0.3
ADJ_STEPS = 4000
ADJ_SEED = 4000


def adjoint_field(nt):
    rng = np.random.default_rng(ADJ_SEED)
    t = np.arange(nt) * DT
    E = np.zeros(nt)
    for _ in range(8):
        A, tk, sk = rng.uniform(0.5, 1.0), rng.uniform(0.2, 0.8) * t[-1], rng.uniform(10.0, 40.0)
        wk, pk = rng.uniform(0.04, 0.08), rng.uniform(0.0, 2 * np.pi)
        E += A * np.exp(-(t - tk) ** 2 / (2 * sk * sk)) * np.cos(wk * t + pk)
    return 0.05 * E / np.abs(E).max()


def adjoint_run(q, steps, fd_samples):
    import torch
    from tdse_nd_mpi_amd import adjoint as adj
    g = q.Grid(ADJ_NR, ADJ_DR, "c128", n_slots=1)
    g.set_potential("H_2D", softcore=1.0)
    g.set_integrator("BM4", 0.1)
    g.upload(0, reference_guess(ADJ_NR, ADJ_DR))
    g.imag_steps(0, q.imag_step_count(100.0, 0.1), gram_schmidt=True)
    psi0 = g.download(0)
    g.close()
    E = adjoint_field(steps)
    w0 = 45.5637639947958 / 800.0
    res = {"grid": list(ADJ_NR), "steps": steps, "checkpoint_every": 64, "seed": ADJ_SEED}
    grads = {}
    for prec in ("c64", "c128"):
        spec = adj.Spec(ADJ_NR, ADJ_DR, "H_2D", "BM4", DT, psi0, precision=prec, mask=0.1, every=64)
        gr = spec.ensure(1.0)
        gr.set_field(*adj.field_tables(E, gr.s, 2))
        gr.upload(0, psi0)
        gr.real_steps(0, "length", 0, 50, ["dipole_a", "norm"])          # warm-up
        gr.upload(0, psi0)
        gr.synchronize()
        t0 = time.perf_counter()
        gr.real_steps(0, "length", 0, steps, ["dipole_a", "norm"])
        t_fwd = time.perf_counter() - t0
        Et = torch.tensor(E, requires_grad=True)
        soft = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
        t0 = time.perf_counter()
        a, nf = adj.Propagate.apply(Et, soft, spec)
        L = adj.harmonic_loss(a, nf, DT, w0, orders=(11, 21))
        L.backward()
        t_tot = time.perf_counter() - t0
        grads[prec] = (Et.grad.numpy().copy(), float(soft.grad), float(L.detach()))
        res[prec] = {"forward_s": t_fwd, "forward_checkpointed_plus_adjoint_s": t_tot,
                     "adjoint_over_forward": t_tot / t_fwd, "loss": float(L.detach()),
                     "dL_dsoftcore": float(soft.grad)}
        if prec == "c128" and fd_samples:
            rng = np.random.default_rng(ADJ_SEED + 1)
            ks = np.sort(rng.choice(steps, fd_samples, replace=False))
            h = 1e-5
            errs = []

            def loss(Ev):
                a2, _, nf2 = adj.forward_numpy(spec, Ev, 1.0)
                return float(adj.harmonic_loss(torch.from_numpy(a2), torch.tensor(nf2, dtype=torch.float64), DT,
                                                 w0, orders=(11, 21)))
            gscale = np.abs(grads["c128"][0]).max()
            for k in ks:
                Ep, Em = E.copy(), E.copy()
                Ep[k] += h
                Em[k] -= h
                fd = (loss(Ep) - loss(Em)) / (2 * h)
                errs.append(abs(fd - grads["c128"][0][k]) / gscale)
            res["fd_check"] = {"samples": [int(k) for k in ks], "h": h,
                               "max_abs_err_over_max_grad": float(max(errs))}
        spec.close()
    g64, g128 = grads["c64"][0], grads["c128"][0]
    res["c64_vs_c128_grad_rel_l2"] = float(np.linalg.norm(g64 - g128) / np.linalg.norm(g128))
    res["metric"] = "adjoint/forward time ratio (forward with checkpoints + recompute + backward over plain forward)"
    res["value"] = res["c64"]["adjoint_over_forward"]
    return res


# ---------------------------------------------------------------------
# 1-D eigenstates: H_1D, NR 4096, DR 0.1, BM4 imaginary time DT 0.1, T 100
# (1001 steps per state), 4 states with Gram-Schmidt after every step,
# complex128; guesses: seeded complex Gaussian-random vectors times
# exp(-x^2/20), and the reference's deterministic guess.  Eigenenergies are
# compared with the FP64 oracle run on the same guesses.
# ---------------------------------------------------------------------
def eig1d_run(q):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    nr, dr, nst, dt = (4096,), 0.1, 4, 0.1
    nsteps = q.imag_step_count(100.0, dt)
    res = {"grid": [4096], "states": nst, "steps_per_state": nsteps}
    o = O.Oracle(nr, dr, "H_1D", "BM4", dt)
    for mode, name in ((1, "seeded_gaussian"), (0, "reference_guess")):
        guess = o.initial_guess(nst, guess_mode=mode, seed=1234)
        g = q.Grid(nr, dr, "c128", n_slots=nst)
        g.set_potential("H_1D")
        g.set_integrator("BM4", dt)
        t0 = time.perf_counter()
        E = []
        for j in range(nst):
            g.upload(j, guess[j])
            g.imag_steps(j, nsteps, gram_schmidt=True)
            E.append(g.energy_moments(j, 1)[0])
        g.synchronize()
        wall = time.perf_counter() - t0
        g.close()
        t0 = time.perf_counter()
        Eo = o.imag_run(nst, 100.0, guess_mode=mode, seed=1234)[1][:, 0]
        t_cpu = time.perf_counter() - t0
        res[name] = {"E_gpu": [float(e) for e in E], "E_oracle": [float(e) for e in Eo],
                     "max_abs_diff_Ha": float(np.max(np.abs(np.array(E) - np.array(Eo)))),
                     "gpu_seconds": wall, "grid_point_steps_per_s": 4096 * nst * nsteps / wall,
                     "oracle_seconds": t_cpu}
    return res


# ---------------------------------------------------------------------
# Large 3-D: H_3D with soft-core length a = 0.05 (V = -1/sqrt(r^2 + a^2)) on
# 512^3, DR 0.2, complex64 (1 GiB per state).  Timed: imaginary-time BM6
# steps (DT 0.05) of state 7 with FP64 Gram-Schmidt against 7 states, then
# real-time BM4 steps (DT 0.02) under 800 nm 1e14 W/cm^2 plus 400 nm 1e13
# W/cm^2 at 20 degrees, seeded relative phase, with dipole, acceleration and
# the cos^(1/8) mask every step.  Few steps (the per-step cost is stationary).
# ---------------------------------------------------------------------
def big3d_run(q, F, imag_steps, real_steps, fixed=None, prec="c64"):
    nr, dr = (512, 512, 512), 0.2
    N = 512 ** 3
    g = q.Grid(nr, dr, prec, n_slots=8)
    g.set_potential("H_3D", softcore=0.05)
    g.set_integrator("BM6", 0.05)
    x = -0.5 * 512 * dr + np.arange(512) * dr
    gauss = np.exp(-(x * x) / 20.0)
    psi = (gauss[:, None, None] * gauss[None, :, None] * gauss[None, None, :]).astype(np.complex128)
    for j in range(8):
        g.upload(j, psi * (1.0 + 0.1 * j * (x[:, None, None] / 10.0)))
    g.imag_steps(7, 1, gram_schmidt=True)
    g.synchronize()
    # launch geometry of the imaginary-time step, timed over 2 steps each (the
    # state relaxes a little further meanwhile; `fixed`: the automatic one)
    cands = [(0, c, s) for c in (0, 1, 2) for s in (0, 4, 8)]
    itable = []
    for cand in ([] if fixed else cands):
        g.set_tuning(*cand)
        g.imag_steps(7, 1, gram_schmidt=True)
        g.synchronize()
        t0 = time.perf_counter()
        g.imag_steps(7, 2, gram_schmidt=True)
        g.synchronize()
        itable.append((cand, (time.perf_counter() - t0) / 2 * 1e3))
    ibest = min(itable, key=lambda r: r[1])[0] if itable else (0, 0, 0)
    g.set_tuning(*ibest)
    t0 = time.perf_counter()
    g.imag_steps(7, imag_steps, gram_schmidt=True)
    g.synchronize()
    t_imag = (time.perf_counter() - t0) / imag_steps
    rng = np.random.default_rng(512)
    ph = rng.uniform(0, 2 * np.pi)
    modes = [F.Mode(intensity=1e14, wavelength=800.0, cep=0.0, envelope="cos2", cycles=4),
             F.Mode(intensity=1e13, wavelength=400.0, cep=ph, envelope="cos2", cycles=8, angle=np.deg2rad(20.0))]
    g.set_integrator("BM4", 0.02)
    g.set_mask(0.1)
    nt = max(real_steps, 8) + 2
    g.set_field(*F.tables(modes, "BM4", 0.02, nt, 3))
    # launch geometry of the real-time step: contiguous and strided lines per
    # workgroup timed over 6 steps each (the tuner restores the state)
    # (`fixed`: that geometry, untimed, for profiled runs)
    from tdse_nd_mpi_amd import tuning
    if fixed:
        best, table = tuple(fixed), []
        g.set_tuning(*best)
    else:
        cands = [(0, c, s) for c in (0, 1, 2) for s in (0, 4, 8)]
        best, table = tuning.autotune_batch(g, 0, 1, "length", 0, nsteps=6, cands=cands)
    g.real_steps(0, "length", 0, 1, ["dipole_r", "dipole_a", "norm"])
    g.synchronize()
    t0 = time.perf_counter()
    obs = g.real_steps(0, "length", 1, real_steps, ["dipole_r", "dipole_a", "norm"])
    g.synchronize()
    t_real = (time.perf_counter() - t0) / real_steps
    g.close()
    return {"grid": list(nr), "precision": prec, "softcore_a": 0.05, "imag_BM6_ms_per_step_state7": t_imag * 1e3,
            "imag_grid_point_steps_per_s": N / t_imag,
            "real_BM4_ms_per_step": t_real * 1e3, "real_grid_point_steps_per_s": N / t_real,
            "real_step_roofline_frac": 14 * 2 * 8 * (N / t_real) / (HBM_PEAK_GBS * 1e9),
            "imag_tuned_vpt_twc_tws": list(ibest),
            "imag_tuning_ms_per_step": {str(list(k)): round(v, 3) for k, v in itable},
            "real_tuned_vpt_twc_tws": list(best),
            "real_tuning_ms_per_step": {str(list(k)): round(v, 3) for k, v in table},
            "second_mode_phase": ph, "finite": bool(np.isfinite(obs).all()),
            "steps_timed": {"imag": imag_steps, "real": real_steps}}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_passes_ne2d_512x512_c64.txt")


def measured_traffic(kernel="void qnd::k_pass<float, 6, 1, 4>"):
    """HBM bytes per launch of `kernel` from the committed counter profile
    (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of `bench.py --traffic-probe`):
    each counter's KiB per dispatch scaled by its reading for the calibration
    copy in the same run (an FP64 -> FP32 conversion of 32 Mi elements: 256 MiB
    read, 128 MiB written).  None when the profile is absent."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    vals = {}
    for line in open(TRAFFIC_PROFILE):
        f = line.split()
        if len(f) < 4 or f[-3] not in ("FETCH_SIZE", "WRITE_SIZE"):
            continue
        ctr, mean = f[-3], float(f[-2])
        if "elementwise_kernel" in line and f[-4] == "8388608":
            vals[("cal", ctr)] = mean
        elif line.strip().startswith(kernel + " ") or (kernel + "  ") in line:
            vals[("k", ctr)] = mean
    try:
        fetch = vals[("k", "FETCH_SIZE")] * (256 * 1024) / vals[("cal", "FETCH_SIZE")]
        write = vals[("k", "WRITE_SIZE")] * (128 * 1024) / vals[("cal", "WRITE_SIZE")]
    except KeyError:
        return None
    return {"bytes": (fetch + write) * 1024, "fetch_bytes": fetch * 1024, "write_bytes": write * 1024,
            "source": os.path.relpath(TRAFFIC_PROFILE, ROOT) + " (default launch geometry)"}


BARRIER_FORMS = {0: "one counter, acquire polls", 1: "one counter, relaxed polls + one acquire fence",
                 2: "per-XCD counters then a top counter, relaxed polls + one acquire fence",
                 3: "per-workgroup arrival flags polled by one workgroup, then a release flag"}


def latency_probes():
    """Grid barriers (one cooperative launch, device-scope release/acquire
    atomics, bounded waits; three forms, BARRIER_FORMS) against the kernel
    boundary (back-to-back empty launches), microseconds each, for the
    workgroup shapes of the 512^2 passes (tdse_nd_mpi_amd/libqnd_probe.so)."""
    import ctypes as C
    path = os.path.join(ROOT, "tdse_nd_mpi_amd", "libqnd_probe.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.qnd_probe_barrier.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.qnd_probe_boundary.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    out = []
    for blocks, threads in ((256, 256), (512, 128), (1024, 64)):
        k = C.c_double()
        L.qnd_probe_boundary(blocks, threads, 2000, C.byref(k))
        rec = {"workgroups": blocks, "threads": threads, "kernel_boundary_us": k.value, "grid_barrier_us": {}}
        for form in BARRIER_FORMS:
            b = C.c_double()
            rb = L.qnd_probe_barrier(blocks, threads, 2000, form, C.byref(b))
            rec["grid_barrier_us"][str(form)] = b.value if rb == 0 else None
            if rb:
                rec.setdefault("barrier_status", {})[str(form)] = rb
        rec["best_grid_barrier_us"] = min((v for v in rec["grid_barrier_us"].values() if v is not None),
                                          default=None)
        out.append(rec)
    return out


def tile_probes():
    """The memory side of a 512^3 complex64 line pass alone (an in-place
    scaled copy read and written as tiles of TW lines of 512 points, the
    passes' lane mapping and XCD tile order): microseconds per pass and GB/s
    of its 2 GiB, per axis and tile width (tdse_nd_mpi_amd/libqnd_probe.so)."""
    import ctypes as C
    path = os.path.join(ROOT, "tdse_nd_mpi_amd", "libqnd_probe.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.qnd_probe_tiles.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    out = []
    for axis, tw, vpt in ((2, 4, 8), (1, 8, 8), (0, 4, 8), (0, 8, 8), (0, 16, 8), (0, 16, 16), (0, 32, 16)):
        us = C.c_double()
        rc = L.qnd_probe_tiles(axis, tw, vpt, 10, C.byref(us))
        out.append({"axis": axis, "lines_per_tile": tw, "values_per_thread": vpt,
                    "us": us.value if rc == 0 else None,
                    "GBs": 2 * 8 * 512 ** 3 / (us.value * 1e3) if rc == 0 else None})
    return out


def copy_bandwidth(reps=10):
    """Measured device-copy bandwidth (GB/s, read + write) of a 1 GiB FP32
    copy, timed with events on torch's stream, where the copies run."""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def spectrum(a, dt):
    w = np.hanning(len(a))
    return np.abs(np.fft.rfft((a - a.mean()) * w)) ** 2


def spectra_errors(obs, ref):
    """Errors of one run's records [nt][dipole x, y, accel x, y, norm] against
    a reference run's: relative L2 of the power spectra of the dipole and of
    the acceleration along the field (x, Hann window), relative L2 of the
    dipole trace, and the largest norm difference."""
    Sd, Sd0 = spectrum(obs[:, 0], DT), spectrum(ref[:, 0], DT)
    Sa, Sa0 = spectrum(obs[:, 2], DT), spectrum(ref[:, 2], DT)
    return {"dipole_spectrum_rel_err": float(np.linalg.norm(Sd - Sd0) / np.linalg.norm(Sd0)),
            "accel_spectrum_rel_err": float(np.linalg.norm(Sa - Sa0) / np.linalg.norm(Sa0)),
            "dipole_rel_l2": float(np.linalg.norm(obs[:, :2] - ref[:, :2]) / np.linalg.norm(ref[:, :2])),
            "norm_max_abs_diff": float(np.abs(obs[:, 4] - ref[:, 4]).max())}


PRECISION_PHASES = 6


def precision_leg(q, psi0, Et, At, Eo, Ao, nt, nphase=PRECISION_PHASES):
    """The whole pulse (nt steps), untimed, in complex64, in complex64
    storage with FP64 pass arithmetic (c64f64), and in complex128 with the
    state rounded to complex64 at every store of a pass (QND_ROUND_STATE: the
    arithmetic of c64f64 in complex128 kernels), each against complex128.  complex128 stands for the FP64 oracle here: it matches the
    oracle to ~1e-12 (GPU tests and parity_vs_oracle_512), and a full-pulse
    oracle run takes minutes of CPU.
    The rounding errors of one run are one realization of a random process:
    a change of V by 1e-16 already gives another.  So each variant runs from
    nphase copies of the initial state multiplied by a global phase e^{i th}
    (th = 2 pi k / nphase), which changes every FP32 rounding and none of the
    observables, and the errors are reported per realization with their
    median and maximum."""
    names = ["dipole_r", "dipole_a", "norm"]

    def run(prec, rnd, psi):
        if rnd:
            os.environ["QND_ROUND_STATE"] = "1"
        try:
            g = q.Grid(NR, DR, prec, n_slots=1)
        finally:
            os.environ.pop("QND_ROUND_STATE", None)
        g.set_potential(POT)
        g.set_integrator("BM4", DT)
        g.set_mask(MASK)
        g.set_field(Et[:nt], At[:nt], Eo[:nt], Ao[:nt])
        g.upload(0, psi)
        rec = g.real_steps(0, "length", 0, nt, names)
        g.close()
        return rec

    ref = run("c128", False, psi0)
    out = {"steps": int(nt), "reference": "complex128 GPU path (matches the FP64 oracle to ~1e-12)",
           "realizations": f"{nphase} global phases of the initial state"}
    recs = {"c128": ref}
    finite = bool(np.isfinite(ref).all())
    for label, prec, rnd in (("c64", "c64", False), ("c64f64", "c64f64", False),
                             ("c128_state_rounded_to_c64", "c128", True)):
        errs = []
        for k in range(nphase):
            rec = run(prec, rnd, psi0 * np.exp(2j * np.pi * k / nphase))
            finite &= bool(np.isfinite(rec).all())
            errs.append(spectra_errors(rec, ref))
            if k == 0:
                recs[label] = rec
        d = {key: [e[key] for e in errs] for key in errs[0]}
        out[label] = {"dipole_spectrum_rel_err_median": float(np.median(d["dipole_spectrum_rel_err"])),
                      "dipole_spectrum_rel_err_max": float(np.max(d["dipole_spectrum_rel_err"])),
                      "accel_spectrum_rel_err_median": float(np.median(d["accel_spectrum_rel_err"])),
                      "dipole_rel_l2_median": float(np.median(d["dipole_rel_l2"])),
                      "norm_max_abs_diff_max": float(np.max(d["norm_max_abs_diff"])),
                      "per_realization": d}
    out["finite"] = finite
    return out, recs


def cpu_info():
    """Model, sockets and physical cores of this host (from /proc/cpuinfo),
    and the CPUs this process may run on."""
    model, phys = "unknown", set()
    cur = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if cur:
                    phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor", "0"))))
                cur = {}
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            cur[k] = v
            if k == "model name":
                model = v
        if cur:
            phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor", "0"))))
    except OSError:
        pass
    sockets = len({p for p, _ in phys}) or 1
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "sockets": sockets, "physical_cores": len(phys) or None,
            "logical_cpus": os.cpu_count(), "cpus_available_to_process": avail}


def cpu_baseline(psi0, Et, At, Eo, Ao, budget_s=15.0):
    """The FP64 oracle on the host cores, two variants of one bounded sample
    (each about budget_s of CPU): "best_case", every loop OpenMP-parallel
    (approximating one MPI rank per core without transposes), and
    "faithful", the FFT threaded and the pointwise loops serial (what one
    reference rank with OMP_NUM_THREADS threads runs: its only threads are
    FFTW's, src/meshgrid.c:159-161).  The best-case run's records are the
    full-size parity reference."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.lib().or_set_threads(threads)
    o = O.Oracle(NR, DR, POT, "BM4", DT)
    o.set_mask(MASK)
    names = ["dipole_r", "dipole_a", "norm"]
    out = {}
    obs_ref = None
    for variant, par in (("best_case", 1), ("faithful", 0)):
        O.lib().or_set_pointwise_parallel(par)
        # a short call to size the sample, then one call of that many steps
        # from the same state (one call: the owed mask is applied as the GPU does)
        o.real_run(psi0, "length", Et[:1], At[:1], Eo[:1], Ao[:1], names=names)   # first-call set-up
        t0 = time.perf_counter()
        o.real_run(psi0, "length", Et[:4], At[:4], Eo[:4], Ao[:4], names=names)
        t1 = (time.perf_counter() - t0) / 4
        done = int(max(1, min(2000, budget_s / max(t1, 1e-6))))
        t0 = time.perf_counter()
        _, obs = o.real_run(psi0, "length", Et[:done], At[:done], Eo[:done], Ao[:done], names=names)
        t_tot = time.perf_counter() - t0
        out[variant] = {"steps": done, "seconds": t_tot, "value": int(np.prod(NR)) * done / t_tot}
        if variant == "best_case":
            obs_ref = obs
    O.lib().or_set_pointwise_parallel(1)
    return out, threads, obs_ref


def full_size_parity(gpu_obs, ref_obs):
    """Records of a GPU run against the FP64 oracle's for the same steps of the
    512^2 workload: max abs difference and relative L2 of the dipole and
    acceleration traces, max abs difference of the norm."""
    n = min(len(gpu_obs), len(ref_obs))
    g, r = np.asarray(gpu_obs[:n]), np.asarray(ref_obs[:n])
    out = {"steps": int(n)}
    for name, cols in (("dipole", slice(0, 2)), ("acceleration", slice(2, 4))):
        d = g[:, cols] - r[:, cols]
        out[name + "_max_abs_diff"] = float(np.abs(d).max())
        out[name + "_rel_l2"] = float(np.linalg.norm(d) / max(np.linalg.norm(r[:, cols]), 1e-300))
    out["norm_max_abs_diff"] = float(np.abs(g[:, 4] - r[:, 4]).max())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8800)
    ap.add_argument("--warmup", type=int, default=25)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-precision", action="store_true",
                    help="skip the untimed whole-pulse precision leg (complex64 against complex128)")
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--sweep-steps", type=int, default=-1,
                    help="timed steps of the 96-instance CEP/intensity sweep (-1: the whole 6619-step pulse; "
                         "0 skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=ADJ_STEPS,
                    help="steps of the adjoint-gradient workload (0 skips it)")
    ap.add_argument("--fd-samples", type=int, default=16)
    ap.add_argument("--skip-1d", action="store_true", help="skip the 1-D eigenstate workload")
    ap.add_argument("--big3d-steps", type=int, default=6, help="timed steps of the 512^3 workload (0 skips it)")
    ap.add_argument("--big3d-prec", choices=["c64", "c64f64"], default="c64",
                    help="precision of the 512^3 workload")
    ap.add_argument("--tuning3d", default="",
                    help="vpt,tw_contig,tw_strided of the 512^3 real-time step instead of the autotuner")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for counter runs: a 256 MiB device copy (calibration), then 20 c64 steps; no JSON")
    ap.add_argument("--dump", default="", help="write the observable traces to this .npz")
    ap.add_argument("--tuning", default="",
                    help="fixed launch geometry 'vpt,tw_contig,tw_strided' for both precisions instead of the "
                         "autotuner (profiled runs: the profiler distorts the autotuner's timings)")
    ap.add_argument("--only", choices=["eig1d", "big3d", "adjoint", "sweep"], default=None,
                    help="run one secondary workload alone and print its JSON (for kernel-trace profiles)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build holds the whole grid in one MI355X's HBM "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import torch
    import tdse_nd_mpi_amd as q
    from tdse_nd_mpi_amd import field as F

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    N = int(np.prod(NR))
    fixed3d = tuple(int(v) for v in args.tuning3d.split(",")) if args.tuning3d else None
    if args.only:
        runs = {"eig1d": lambda: eig1d_run(q),
                "big3d": lambda: big3d_run(q, F, max(args.big3d_steps, 1), max(args.big3d_steps, 1), fixed3d,
                                           args.big3d_prec),
                "adjoint": lambda: adjoint_run(q, args.adjoint_steps, args.fd_samples),
                "sweep": lambda: sweep_run(q, F, args.sweep_steps, 10, compare_steps=200)}
        print(json.dumps({"workload": args.only, "result": runs[args.only]()}))
        return
    if args.traffic_probe:
        # calibration: an FP64 -> FP32 conversion copy reads 256 MiB and writes 128 MiB
        src = torch.empty(32 << 20, dtype=torch.float64, device="cuda").uniform_()
        dst = torch.empty(32 << 20, dtype=torch.float32, device="cuda")
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        mode = F.Mode(intensity=1e14, wavelength=800.0, ellipticity=0.0, cep=0.0, envelope="cos2", cycles=4)
        Et, At, Eo, Ao = F.tables([mode], "BM4", DT, 100, 2)
        psi0 = reference_guess(NR, DR)
        real_run(q, F, "c64", psi0 / np.linalg.norm(psi0), 0, 20, Et, At, Eo, Ao, tune=False)
        log("traffic probe done: 3 conversion copies (256 MiB read, 128 MiB written), 20 c64 steps")
        return

    log("ground state (imaginary time, FP64) ...")
    psi0, E0, n_imag, t_imag = ground_state(q)
    log(f"  E0 = {E0[0]:.12f}  <H^2>-E^2 = {E0[1] - E0[0] ** 2:.3e}  {n_imag} steps in {t_imag:.3f}s")

    mode = F.Mode(intensity=1e14, wavelength=800.0, ellipticity=0.0, cep=0.0, envelope="cos2", cycles=4)
    ntime_full = int(mode.tau / DT)
    ntime = max(ntime_full, args.warmup + args.steps)
    Et, At, Eo, Ao = F.tables([mode], "BM4", DT, ntime, 2)

    res = {}
    for prec in (["c64"] if args.skip_fp64 else ["c64", "c128", "c64f64"]):
        log(f"real time {prec}: warmup {args.warmup}, timed {args.steps} ...")
        fixed = tuple(int(v) for v in args.tuning.split(",")) if args.tuning else None
        obs, wall, gpu_ms, prof, tuned = real_run(q, F, prec, psi0, args.warmup, args.steps, Et, At, Eo, Ao,
                                                  profile_steps=args.profile_steps, fixed=fixed)
        res[prec] = dict(obs=obs, wall=wall, gpu_ms=gpu_ms, prof=prof, tuned=tuned)
        log(f"  {prec}: {wall * 1e3 / args.steps:.4f} ms/step wall, {gpu_ms / args.steps:.4f} ms/step gpu")

    if args.dump:
        np.savez(args.dump, **{f"obs_{k}": v["obs"] for k, v in res.items()}, Eobs=Eo, psi0=psi0)
    mixed = res["c64"]
    ms_step = mixed["wall"] * 1e3 / args.steps
    value = N * args.steps / mixed["wall"]
    s = 6
    n_fft = 2 * s + 2                          # length gauge with observables every step
    b_alg_c64 = n_fft * 2 * 8                  # bytes per grid point and step
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (FP32 FFT/phases, FP64 angle reduction and reductions)",
        "data": "synthetic: the reference's deterministic guess relaxed in imaginary time; analytic laser field",
        "config": {"workload": "Ne_2D 512x512 DR=0.2, real-time BM4 DT=0.05 length gauge, "
                               "4-cycle 800nm 1e14 W/cm2 cos2 pulse, cos^(1/8) mask outer 10%, "
                               "dipole + acceleration every step",
                   "grid": list(NR), "steps_full_pulse": ntime_full, "parallelism": "single GPU"},
        "step_roofline_frac": b_alg_c64 * value / (HBM_PEAK_GBS * 1e9),
        "gpu_ms_per_step": mixed["gpu_ms"] / args.steps,
        "tuned_vpt_twc_tws": list(mixed["tuned"][0]),
    }
    if "c128" in res:
        fp64 = res["c128"]
        v64 = N * args.steps / fp64["wall"]
        out["fp64"] = {"value": v64, "ms_per_step": fp64["wall"] * 1e3 / args.steps,
                       "gpu_ms_per_step": fp64["gpu_ms"] / args.steps,
                       "step_roofline_frac": n_fft * 2 * 16 * v64 / (HBM_PEAK_GBS * 1e9),
                       "tuned_vpt_twc_tws": list(fp64["tuned"][0])}
        out["mixed_over_fp64"] = value / v64
        xm = res["c64f64"]
        vx = N * args.steps / xm["wall"]
        out["c64f64"] = {"value": vx, "ms_per_step": xm["wall"] * 1e3 / args.steps,
                         "gpu_ms_per_step": xm["gpu_ms"] / args.steps,
                         "over_fp64": vx / v64, "over_c64": vx / value,
                         "what": "complex64 state in memory, FP64 arithmetic in the passes (one rounding to "
                                 "complex64 per value and pass, at its store); phases as for c64",
                         "tuned_vpt_twc_tws": list(xm["tuned"][0])}
        a64 = mixed["obs"][:, 2]
        a128 = fp64["obs"][:, 2]
        S64, S128 = spectrum(a64, DT), spectrum(a128, DT)
        # over the timed steps only (the whole pulse: precision_full_pulse)
        out["timed_steps_c64_vs_c128"] = {
            "steps": int(len(a64)),
            "accel_spectrum_rel_err": float(np.linalg.norm(S64 - S128) / np.linalg.norm(S128)),
            "dipole_rel_l2": float(np.linalg.norm(mixed["obs"][:, :2] - fp64["obs"][:, :2]) /
                                   np.linalg.norm(fp64["obs"][:, :2]))}
        Sx = spectrum(xm["obs"][:, 2], DT)
        out["timed_steps_c64f64_vs_c128"] = {
            "accel_spectrum_rel_err": float(np.linalg.norm(Sx - S128) / np.linalg.norm(S128)),
            "dipole_rel_l2": float(np.linalg.norm(xm["obs"][:, :2] - fp64["obs"][:, :2]) /
                                   np.linalg.norm(fp64["obs"][:, :2]))}
    # roofline of the dominant kernel class (HIP events around its launches)
    cnt, ms = mixed["prof"]
    names = ["axis0_potential_pass", "kinetic_row_pass", "plain_fft_pass", "reductions", "boundary_passes"]
    k = int(np.argmax(ms[:2]))
    avg_ms = ms[k] / max(cnt[k], 1)
    bytes_launch = 2 * N * 8
    ach = bytes_launch / (avg_ms * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "kernel": names[k], "achieved": ach, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                       "traffic": (tr["bytes"] if (tr := measured_traffic()) else None),
                       "traffic_detail": tr,
                       "avg_launch_us": avg_ms * 1e3, "bytes_per_launch": bytes_launch,
                       "copy_GBs": (cbw := copy_bandwidth()), "frac_of_copy": ach / cbw,
                       "launches_per_step": {names[i]: float(cnt[i]) / args.profile_steps for i in range(5)},
                       "class_ms_per_step": {names[i]: float(ms[i]) / args.profile_steps for i in range(5)}}
    out["latency_probes"] = latency_probes()
    if out["latency_probes"]:
        out["latency_probes_forms"] = BARRIER_FORMS
        log("  grid barrier (best form) vs kernel boundary (us): " + ", ".join(
            f"{p['workgroups']}x{p['threads']}: {p['best_grid_barrier_us']} vs {p['kernel_boundary_us']:.2f}"
            for p in out["latency_probes"]))
    if not args.skip_cpu:
        log("cpu baseline (FP64 oracle on the host cores, best-case and faithful threading) ...")
        var, threads, ref_obs = cpu_baseline(psi0, Et, At, Eo, Ao, args.cpu_budget)
        bc, fa = var["best_case"], var["faithful"]
        info = cpu_info()
        out["cpu_baseline"] = {
            "value": bc["value"], "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
            "sample": f"{bc['steps']} steps ({bc['seconds']:.1f}s) of the same 512x512 length-gauge run from "
                      f"the same ground state: the FP64 C oracle (own radix-2 FFT; FFTW is absent) with "
                      f"{threads} OpenMP threads on FFT lines and pointwise loops (best case; speedups are "
                      f"quoted against it)",
            "faithful": {"value": fa["value"], "steps": fa["steps"], "seconds": fa["seconds"],
                         "threading": f"FFT lines on {threads} threads, pointwise loops serial "
                                      f"(one reference rank: FFTW's threads only)"},
            "fft_backend": "oracle's own FP64 radix-2 FFT (FFTW3 not installed), OpenMP over lines",
            "host": info}
        out["speedup_vs_cpu"] = value / bc["value"]
        out["speedup_vs_cpu_faithful"] = value / fa["value"]
        # the GPU records of the same steps against the oracle's (full grid size)
        out["parity_vs_oracle_512"] = {prec: full_size_parity(res[prec]["obs"], ref_obs) for prec in res}
    if not args.skip_precision:
        log(f"precision leg: the whole pulse ({ntime_full} steps) in c64, c128 and c128 with c64 storage ...")
        prec_res, prec_recs = precision_leg(q, psi0, Et, At, Eo, Ao, ntime_full)
        out["precision_full_pulse"] = prec_res
        c64r, rr_, xr_ = prec_res["c64"], prec_res["c128_state_rounded_to_c64"], prec_res["c64f64"]
        log(f"  c64 vs c128: dipole spectrum median {c64r['dipole_spectrum_rel_err_median']:.3e} "
            f"max {c64r['dipole_spectrum_rel_err_max']:.3e}, acceleration spectrum median "
            f"{c64r['accel_spectrum_rel_err_median']:.3e}; c64f64: dipole spectrum median "
            f"{xr_['dipole_spectrum_rel_err_median']:.3e} max {xr_['dipole_spectrum_rel_err_max']:.3e}; "
            f"c128 with c64 storage: dipole spectrum median "
            f"{rr_['dipole_spectrum_rel_err_median']:.3e} max {rr_['dipole_spectrum_rel_err_max']:.3e}")
        if args.dump:
            np.savez(args.dump.replace(".npz", "") + "_full_pulse.npz", **prec_recs)
    out["imag_ground_state"] = {"E0": float(E0[0]), "steps": n_imag, "seconds": t_imag,
                                "grid_point_steps_per_s_fp64": N * n_imag / t_imag}
    if not args.skip_1d:
        log("1-D eigenstates: 4096 points, 4 states x 1001 steps, GPU and oracle ...")
        out["eig1d_4096"] = eig1d_run(q)
        log(f"  max |E_gpu - E_oracle| = {out['eig1d_4096']['seeded_gaussian']['max_abs_diff_Ha']:.2e} Ha")
    if args.big3d_steps > 0:
        log(f"3-D 512^3: {args.big3d_steps} imaginary and real steps ...")
        out["big3d_512"] = big3d_run(q, F, args.big3d_steps, args.big3d_steps, fixed3d, args.big3d_prec)
        out["big3d_512"]["tile_probes"] = tile_probes()
        log(f"  real {out['big3d_512']['real_BM4_ms_per_step']:.2f} ms/step")
    if args.adjoint_steps > 0:
        log(f"adjoint gradient: 256^2, {args.adjoint_steps} steps, FD on {args.fd_samples} samples ...")
        out["adjoint_256"] = adjoint_run(q, args.adjoint_steps, args.fd_samples)
        log(f"  adjoint/forward = {out['adjoint_256']['value']:.2f}")
    if args.sweep_steps != 0:
        log(f"CEP/intensity sweep: 96 instances, {'the whole pulse' if args.sweep_steps < 0 else args.sweep_steps} "
            f"steps, one launch per pass ...")
        out["sweep_96"] = sweep_run(q, F, args.sweep_steps, 10)
        log(f"  sweep: {out['sweep_96']['value']:.3e} grid-point-steps/s")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
