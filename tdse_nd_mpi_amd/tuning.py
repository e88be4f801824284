"""Launch autotuner: picks values per thread and lines per workgroup for a
context by timing a few real-time steps of the live state.

FFTW's planner (FFTW_MEASURE, src/meshgrid.c:166-167) plays this role in the
reference.  Here the candidates are launch geometries of the same kernels, so
every candidate gives the same results up to rounding; the state is saved
before the trials and restored after them.
"""
import itertools
import time

import numpy as np


def candidates(nr, precision):
    """Geometries worth timing for a grid: values per thread (power-of-two
    lines only; 3 * 2^k lines always take twelve) and lines per workgroup of
    the contiguous and strided passes."""
    pow2 = all(n & (n - 1) == 0 for n in nr)
    vpts = [4, 8] if pow2 else [0]
    tw_c = [0, 1, 2, 4]
    tw_s = [0, 1, 2, 4, 8, 16] if precision in ("c64", 1) else [0, 1, 2, 4, 8]
    if len(nr) == 1:
        tw_s = [0]
    return list(itertools.product(vpts, tw_c, tw_s))


def autotune_batch(grid, slot0, nb, gauge, step0, nsteps=6, cands=None):
    """Time nsteps batched real-time steps per candidate; set and return the
    fastest (vpt, tw_contig, tw_strided) with the table of timings (ms/step)."""
    saved = [grid.download(slot0 + b) for b in range(nb)]
    table = []
    for cand in (cands or candidates(grid.nr, grid.precision)):
        grid.set_tuning(*cand)
        for b in range(nb):
            grid.upload(slot0 + b, saved[b])
        grid.real_steps_batch(slot0, nb, gauge, step0, 1)      # first launch of the geometry
        t0 = time.perf_counter()
        grid.real_steps_batch(slot0, nb, gauge, step0 + 1, nsteps)
        table.append((cand, (time.perf_counter() - t0) / nsteps * 1e3))
    best = min(table, key=lambda r: r[1])[0]
    grid.set_tuning(*best)
    for b in range(nb):
        grid.upload(slot0 + b, saved[b])
    return best, table
