"""Launch autotuner: picks values per thread and lines per workgroup for a
context by timing a few real-time steps of the live state.

FFTW's planner (FFTW_MEASURE, src/meshgrid.c:166-167) plays this role in the
reference, and FFTW's wisdom its memory.  Here the candidates are launch
geometries of the same kernels, so every candidate gives the same results up
to rounding; the state is saved before the trials and restored after them.
Choices are cached in a JSON file keyed by device, grid shape, precision,
batch size, gauge and candidate set (QND_TUNING_CACHE names the file; the
default is tuning_cache.json next to this module; QND_TUNING_CACHE=off
disables the cache).
"""
import itertools
import json
import os
import time

import numpy as np

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")


def candidates(nr, precision):
    """Geometries worth timing for a grid: values per thread (power-of-two
    lines only; 3 * 2^k lines always take twelve) and lines per workgroup of
    the contiguous and strided passes."""
    pow2 = all(n & (n - 1) == 0 for n in nr)
    vpts = [4, 8] if pow2 else [0]
    tw_c = [0, 1, 2, 4]
    tw_s = [0, 1, 2, 4, 8, 16] if precision in ("c64", "c64f64", 1, 2) else [0, 1, 2, 4, 8]
    if len(nr) == 1:
        tw_s = [0]
    return list(itertools.product(vpts, tw_c, tw_s))


def cache_path():
    p = os.environ.get("QND_TUNING_CACHE", DEFAULT_CACHE)
    return None if p in ("", "off", "0") else p


def device_key():
    try:
        import torch
        if torch.cuda.is_available():
            pr = torch.cuda.get_device_properties(torch.cuda.current_device())
            return f"{getattr(pr, 'gcnArchName', pr.name)}/{pr.multi_processor_count}cu"
    except Exception:
        pass
    return "unknown"


def cache_key(grid, nb, gauge, cands):
    prec = {0: "c128", 1: "c64", 2: "c64f64"}.get(grid.precision, str(grid.precision))
    return json.dumps({"device": device_key(), "nr": list(grid.nr), "dr": [float(d) for d in grid.dr],
                       "precision": prec, "nb": int(nb), "gauge": gauge,
                       "candidates": [list(c) for c in cands]}, sort_keys=True)


def load_cache(path):
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def store_cache(path, key, best, table):
    if not path:
        return
    data = load_cache(path)
    data[key] = {"best": list(best), "table": [[list(c), round(ms, 5)] for c, ms in table]}
    tmp = path + ".tmp"
    try:
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def autotune_batch(grid, slot0, nb, gauge, step0, nsteps=6, cands=None, use_cache=True):
    """Time nsteps batched real-time steps per candidate; set and return the
    fastest (vpt, tw_contig, tw_strided) with the table of timings (ms/step).
    A cached choice for the same key is set without timing (its table is the
    one measured when it was cached)."""
    cands = list(cands or candidates(grid.nr, grid.precision))
    path = cache_path() if use_cache else None
    key = cache_key(grid, nb, gauge, cands)
    hit = load_cache(path).get(key) if path else None
    if hit:
        best = tuple(hit["best"])
        grid.set_tuning(*best)
        grid.tuning_cached = True
        return best, [(tuple(c), ms) for c, ms in hit["table"]]
    saved = [grid.download(slot0 + b) for b in range(nb)]
    table = []
    for cand in cands:
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
    grid.tuning_cached = False
    store_cache(path, key, best, table)
    return best, table
