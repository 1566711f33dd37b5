"""bench.py's host-side contract: one GPU only, and the per-workload summary."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 2
    msg = json.loads(res.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_per_workload_summary():
    small = {"dam_break_256x256": {"steps": 500, "gcell_updates_per_s": 2.0, "hbm_frac_32B": 0.008}}
    cpu = {"value": 0.0035}
    adj = {"value": 13.0, "taped_backward_over_forward": 2.8, "checkpointed_backward_over_forward": 3.9}
    inv = {"seconds_per_epoch": 0.21, "speedup_vs_published": 400.0}
    warm = {"plain": {"seconds": 55.0, "epochs": 40}, "ratio_warm_over_plain": 0.05, "ratio_with_one_off": 0.8,
            "one_off": {"ensemble_gcell_per_s": 40.0, "ensemble_arithmetic": "fast"}}
    out = bench.per_workload(small, 38.5, 1232.0, cpu, inv, warm, adj)
    assert out["dam_break_256x256"] == small["dam_break_256x256"]
    basin = out["tsunami_basin_8192x8192"]
    assert basin["gcell_updates_per_s"] == 38.5 and basin["cpu_gcell_updates_per_s"] == 0.0035
    assert abs(basin["hbm_frac_32B"] - 1232.0 / 8000.0) < 1e-4
    assert abs(out["ensemble_256_scenarios"]["hbm_frac_32B"] - 40.0 * 32 / 8000.0) < 1e-4
    assert out["runup_boundary_inversion_544x344"]["seconds_per_epoch"] == 1.375
    assert out["tsunami_basin_8192x8192_adjoint"]["checkpointed_backward_over_forward"] == 3.9
    # legs that did not run are absent, not zero
    bare = bench.per_workload(None, 38.5, 1232.0, None, None, None, None)
    assert set(bare) == {"tsunami_basin_8192x8192"}
    assert bare["tsunami_basin_8192x8192"]["cpu_gcell_updates_per_s"] is None
