"""Multi-scenario ensemble against batch-1 SWE2D.forward runs of each scenario."""
import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import scenarios
from diffswe2d_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu

GRIDS = ((22, 34, 0.16), (43, 68, 0.08), (86, 136, 0.04))


def make(n=6, n_chunks=12, chunk=0.05):
    sc, ctrl = scenarios.ensemble_scenarios(n=n, seed=7, n_controls=8, n_chunks=n_chunks, sample_dt=chunk,
                                           grids=GRIDS)
    # scale the waveforms up so the short run sees the boundary move
    for s in sc:
        s["levels"] = 5.0 * s["levels"] + 0.01
    return sc


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_matches_single_runs(cuda_device, graph):
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    sc = make()
    ens = Ensemble(cfg, sc, cuda_device, graph=graph, steps_per_graph=16)
    ens.run()
    res = ens.results()
    steps = [r["steps"] for r in res]
    print("per-scenario steps", steps)
    assert all(r["done"] for r in res)
    assert len(set(steps)) > 1  # scenarios really took their own dt
    for i, s in enumerate(sc):
        m = swe.SWE2D(torch.tensor(s["bed"], dtype=torch.float32)[None, None],
                      torch.tensor(s["rough"], dtype=torch.float32)[None, None], s["dx"], s["dy"],
                      config=cfg, train_dem=False).to(cuda_device)
        U = torch.tensor(s["U0"], dtype=torch.float32, device=cuda_device)[None]
        bed = torch.tensor(s["bed"], dtype=torch.float32)
        obs, n_steps = [], 0
        for k, lvl in enumerate(s["levels"]):
            U, _ = m(U, t_end=s["chunk"], start_time=k * s["chunk"], boundary_level=float(np.float32(lvl)))
            n_steps += len(m.last_dt_series)
            obs.append([float(U[0, 0, iy, ix]) + float(bed[iy, ix]) for iy, ix in s["gauges"]])
        assert n_steps == res[i]["steps"], (i, n_steps, res[i]["steps"])
        assert torch.equal(U[0], res[i]["U"]), f"scenario {i}: final state differs"
        np.testing.assert_array_equal(np.float32(obs), res[i]["gauges"].cpu().numpy())


def test_rows_per_strip_change_nothing_but_time(cuda_device, monkeypatch, tmp_path):
    """The tuner's choice (read back by Ensemble from its table) and two fixed strip
    heights give bit-identical results: tiling only changes which wavefront computes a cell."""
    from diffswe2d_amd import autotune

    monkeypatch.setattr(autotune, "TABLE", tmp_path / "tuned_rows.json")
    monkeypatch.setattr(autotune, "_cache", {})
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    sc = make()
    best, rep = autotune.tune_ensemble_rows(cfg, sc, cuda_device, candidates=(4, 8, 16), steps=4)
    print("tuned rows", best, rep["trials"])
    assert set(best) == set(autotune.ensemble_mix(sc)) and all(v in (4, 8, 16) for v in best.values())
    outs = []
    for rows in (None, 4, 16):  # None: whatever the table holds for this mix
        ens = Ensemble(cfg, sc if rows is None else [dict(s, rows=rows) for s in sc], cuda_device, graph=True,
                       steps_per_graph=16)
        if rows is None:
            assert ens.rows == [best[tuple(s["bed"].shape)] for s in sc]
        ens.run()
        outs.append(ens.results())
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert a["steps"] == b["steps"]
            assert torch.equal(a["U"], b["U"])
            assert torch.equal(a["gauges"], b["gauges"])


def test_fast_arithmetic_matches_fast_single_runs(cuda_device):
    """arithmetic="fast": every scenario equals its batch-1 SWE2D.forward run made with
    SWE2D.arithmetic = "fast", bit for bit, and stays within the parity bars of the
    reference-rounding ensemble."""
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    sc = make()
    runs = {}
    for arith in ("reference", "fast"):
        ens = Ensemble(cfg, sc, cuda_device, graph=True, steps_per_graph=16, arithmetic=arith)
        ens.run()
        runs[arith] = ens.results()
    for i, s in enumerate(sc):
        m = swe.SWE2D(torch.tensor(s["bed"], dtype=torch.float32)[None, None],
                      torch.tensor(s["rough"], dtype=torch.float32)[None, None], s["dx"], s["dy"],
                      config=cfg, train_dem=False).to(cuda_device)
        m.arithmetic = "fast"
        U = torch.tensor(s["U0"], dtype=torch.float32, device=cuda_device)[None]
        with torch.no_grad():
            for k, lvl in enumerate(s["levels"]):
                U, _ = m(U, t_end=s["chunk"], start_time=k * s["chunk"], boundary_level=float(np.float32(lvl)))
        fast, ref = runs["fast"][i], runs["reference"][i]
        assert torch.equal(U[0], fast["U"]), f"scenario {i}: fast ensemble differs from the fast single run"
        rel = float(torch.linalg.norm((fast["U"] - ref["U"]).double()) / torch.linalg.norm(ref["U"].double()))
        print(f"scenario {i}: fast vs reference rounding rel-L2 {rel:.2e}, steps {fast['steps']}/{ref['steps']}")
        assert rel <= 1e-5
        assert torch.equal(fast["U"][0] > 1e-6, ref["U"][0] > 1e-6)


def test_bad_arithmetic_is_refused(cuda_device):
    cfg = swe.SWEConfig(boundary_left="wall", boundary_right="wall", boundary_top="wall", boundary_bottom="wall")
    with pytest.raises(ValueError, match="arithmetic"):
        Ensemble(cfg, make(n=1, n_chunks=4), cuda_device, arithmetic="approximate")
