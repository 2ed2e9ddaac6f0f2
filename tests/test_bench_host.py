"""The benchmark's host-side legs on the CPU: its helpers exist and the reference CPU
baseline (oracle/_ref/cor, built from the reference's sources where they are present)
runs on one site of the benchmark grid and returns the reference's cutoffs."""
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO))


def test_bench_helpers_and_cpu_baseline():
    import bench
    for name in ("build_scans", "flop_model", "time_plan", "traffic_per_launch", "copy_calibration",
                 "cpu_baseline", "main"):
        assert callable(getattr(bench, name)), name
    fm = bench.flop_model()
    assert fm["t96_igrf"]["flops_per_step"] > 1e4
    assert bench.traffic_per_launch(5328000) > 0
    if not (REPO / "oracle" / "_ref" / "cor").exists():
        pytest.skip("reference build absent")
    sites = [(1.0031392868, 0.0, 0.0)]
    cpu, cuts = bench.cpu_baseline(sites, 0.0, 400, 2000)
    assert cpu["kind"] == "reference" and cpu["value"] > 0 and cpu["unit"] == "trajectories/s"
    rl, ru, rc = cuts[0]
    assert 0.0 < rl <= rc <= ru <= 20.01
