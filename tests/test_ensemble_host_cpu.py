"""Host-side logic of the ensemble runner and its tuner (no GPU).

* ``ensemble.phase_rates`` splits a run's timeline where a grid size finishes
  and prices each phase by the cells of the scenarios still running: checked
  on a hand-built timeline whose rates follow from the arithmetic.
* The tuner's table is keyed by the mix of grid sizes; ``lookup_ensemble``
  returns the rows stored for exactly that mix and nothing for another.
"""
import numpy as np
import pytest

from diffswe2d_amd import autotune
from diffswe2d_amd.ensemble import phase_rates


def test_phase_rates_by_hand():
    # scenario A: 1e9 cells, 4 steps; B: 1e10 cells, 8 steps; one replay = 2 global steps = 1 s
    tl = [(0, 0.0, 0), (2, 1.0, 0), (4, 2.0, 1), (6, 3.0, 1), (8, 4.0, 2)]
    ph = phase_rates(tl, [1e9, 1e10], [4, 8])
    assert [p["global_steps"] for p in ph] == [(0, 4), (4, 8)]
    assert [p["running"] for p in ph] == [2, 1]
    assert [p["seconds"] for p in ph] == [2.0, 2.0]
    # phase 1: 4 steps x 1.1e10 cells in 2 s; phase 2: 4 steps x 1e10 in 2 s
    assert [p["gcell_per_s"] for p in ph] == [22.0, 20.0]


def test_phase_rates_count_every_cell_update_once():
    rng = np.random.default_rng(0)
    cells = rng.integers(10**8, 10**9, size=12).astype(float)  # rates of ~1-100 Gcell/s (rounded to 0.01)
    steps = rng.choice([64, 128, 256], size=12)
    tl, t = [(0, 0.0, 0)], 0.0
    for s in range(64, 257, 64):
        t += float(rng.uniform(0.5, 2.0))
        tl.append((s, t, int((steps <= s).sum())))
    ph = phase_rates(tl, cells, steps)
    total = sum(p["gcell_per_s"] * 1e9 * p["seconds"] for p in ph)
    assert total == pytest.approx(float(np.dot(cells, steps)), rel=2e-3)
    assert sum(p["seconds"] for p in ph) == pytest.approx(t, abs=1e-3)


def test_tuned_rows_are_keyed_by_the_mix(monkeypatch):
    scen = [dict(bed=np.zeros((172, 272)))] * 3 + [dict(bed=np.zeros((344, 544)))] * 2
    mix = autotune.ensemble_mix(scen)
    assert mix == {(172, 272): 3, (344, 544): 2}
    table = {autotune._ens_key("cpu", mix): {"rows": {"172x272": 16, "344x544": 32}}}
    monkeypatch.setattr(autotune, "_cache", table)
    assert autotune.lookup_ensemble("cpu", mix) == {(172, 272): 16, (344, 544): 32}
    other = dict(mix)
    other[(344, 544)] = 3
    assert autotune.lookup_ensemble("cpu", other) == {}
