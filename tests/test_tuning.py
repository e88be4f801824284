"""The launch autotuner's cache (host logic, no GPU): a second tuning of the
same key sets the cached geometry without timing; another key, another
precision or a disabled cache times again."""
import numpy as np

from tdse_nd_mpi_amd import tuning


class FakeGrid:
    def __init__(self, nr=(64, 64), precision=1):
        self.nr, self.dr, self.precision = nr, np.array([0.2] * len(nr)), precision
        self.calls, self.tuned = 0, None
        self.state = {0: np.ones(nr)}

    def set_tuning(self, *t):
        self.tuned = t

    def download(self, slot):
        return self.state[slot]

    def upload(self, slot, psi):
        self.state[slot] = psi

    def real_steps_batch(self, slot0, nb, gauge, step0, nsteps):
        self.calls += 1
        # a geometry-dependent cost: (4, 1, 2) is the fastest
        import time
        time.sleep(0.0002 * (1 + abs(self.tuned[0] - 4) + abs(self.tuned[1] - 1) + abs(self.tuned[2] - 2)))


def test_cache_hit_skips_timing(tmp_path, monkeypatch):
    monkeypatch.setenv("QND_TUNING_CACHE", str(tmp_path / "cache.json"))
    cands = [(4, 1, 2), (8, 0, 0), (4, 2, 4)]
    g = FakeGrid()
    best, table = tuning.autotune_batch(g, 0, 1, "length", 0, nsteps=2, cands=cands)
    assert best == (4, 1, 2) and g.tuned == best and len(table) == 3 and not g.tuning_cached
    timed = g.calls
    assert timed == 2 * len(cands)
    g2 = FakeGrid()
    best2, table2 = tuning.autotune_batch(g2, 0, 1, "length", 0, nsteps=2, cands=cands)
    assert best2 == best and g2.tuned == best and g2.calls == 0 and g2.tuning_cached
    assert [tuple(c) for c, _ in table2] == [tuple(c) for c, _ in table]
    # another precision, shape or batch size is another key
    g3 = FakeGrid(precision=0)
    tuning.autotune_batch(g3, 0, 1, "length", 0, nsteps=2, cands=cands)
    assert g3.calls == 2 * len(cands)
    g4 = FakeGrid(nr=(128, 128))
    tuning.autotune_batch(g4, 0, 1, "length", 0, nsteps=2, cands=cands)
    assert g4.calls == 2 * len(cands)


def test_cache_disabled(tmp_path, monkeypatch):
    monkeypatch.setenv("QND_TUNING_CACHE", "off")
    cands = [(4, 1, 2), (8, 0, 0)]
    for _ in range(2):
        g = FakeGrid()
        tuning.autotune_batch(g, 0, 1, "length", 0, nsteps=1, cands=cands)
        assert g.calls == 2 * len(cands)
    assert not (tmp_path / "cache.json").exists()
