"""Extract the 3-D solver's recorded outputs into tests/golden/dem3d_fixtures.json.

Sources (data only, read as JSON text):
  docs/dem3d_ci_regression_metrics.json: the FIXED state of the three-point
    bending load-transfer fixture (24 steps, rows of applied load, support
    reaction and midspan deflection), the load-convergence campaign (8 and
    16 steps at 0.01 N) and the bottom-platen damper probe;
  benchmarks/dem3d/three_point_bending_cases.json: the case configuration.
The protocols that produced them are docs/ci_regression_evidence.py:45-168
and scripts/benchmark_dem3d_three_point_bending.py:12-47.

    python tests/golden/make_dem3d_fixtures.py /root/reference
"""
import json
import sys
from pathlib import Path

root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
metrics = json.loads((root / "docs/dem3d_ci_regression_metrics.json").read_text())
cases = json.loads((root / "benchmarks/dem3d/three_point_bending_cases.json").read_text())
fixed = next(s for s in metrics["states"] if s["state"] == "FIXED")
out = {
    "source": ["docs/dem3d_ci_regression_metrics.json",
               "benchmarks/dem3d/three_point_bending_cases.json"],
    "bending_case": cases["cases"][0]["config"],
    "bending_fixture": {k: fixed[k] for k in ("dt_s", "steps", "particle_count", "bond_count",
                                               "support_particles", "loaded_particles",
                                               "peak_support_reaction_N",
                                               "peak_midspan_deflection_m", "rows")},
    "load_convergence": {"peak_load_N": 0.01, "runs": metrics["load_convergence_campaign_fixed"]["runs"]},
    "bottom_platen_damper": {
        "config": dict(nx=5, ny=2, nz=2, radius=0.01, density=917.0, drag=0.0, fix_x_edges=False,
                       bottom_platen_enabled=True, platen_stiffness=4000.0, platen_damping=50.0),
        "bottom_shift_m": -0.002, "speed_m_s": 0.03,
        "result": metrics["bottom_platen_damper"]["FIXED"],
    },
}
dest = Path(__file__).resolve().parent / "dem3d_fixtures.json"
dest.write_text(json.dumps(out, indent=1) + "\n")
print(dest)
