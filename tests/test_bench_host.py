"""The benchmark's host-side legs on the CPU: its helpers exist, the reference CPU
baseline (oracle/_ref/cor, built from the reference's sources where they are present)
runs on one site of the benchmark grid and returns the reference's cutoffs and allowed
rigidities, and the accuracy comparison counts what it claims to on planted
differences."""
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO))


def test_bench_helpers_and_cpu_baseline():
    import bench
    for name in ("build_scans", "flop_model", "time_plan", "traffic_per_launch", "copy_calibration",
                 "CpuBaseline", "compare_sets", "main"):
        assert callable(getattr(bench, name)), name
    fm = bench.flop_model()
    assert fm["t96_igrf"]["flops_per_step"] > 1e4
    assert bench.traffic_per_launch(5328000) > 0
    if not (REPO / "oracle" / "_ref" / "cor").exists():
        pytest.skip("reference build absent")
    sites = [(1.0031392868, 0.0, 0.0)]
    cpu = bench.CpuBaseline(sites, 1, 400, 2000)
    cpu.run()
    assert cpu.error is None and cpu.t_total > 0
    allowed, (rl, ru, rc) = cpu.results[0]
    assert 0.0 < rl <= rc <= ru <= 20.01
    # every rigidity above Ru is allowed and listed with its asymptotic direction
    per = 2000
    i_hi = int(round((ru - 0.01) / 0.01))
    assert all(i in allowed for i in range(i_hi + 1, per))


def test_compare_sets_counts_planted_differences():
    """Two synthetic results: a flag flipped inside the band changes Rc only; one flipped
    outside the band counts as a mismatch; a shifted asymptotic longitude shows as the
    angle."""
    import bench
    per = 2000
    rig = 0.01 + 0.01 * np.arange(per)
    f = rig > 5.0
    ast = np.full(per, 10.0)
    asf = np.full(per, 100.0)
    a = {0: (f.copy(), ast.copy(), asf.copy(), (5.0, 5.5, 5.2)), 1: (f.copy(), ast.copy(), asf.copy(), (5.0, 5.5, 5.2))}
    f2 = f.copy()
    f2[1500] = False                       # 15.01 GV: outside the band
    asf2 = asf.copy()
    asf2[1800] += 2.0                      # 2 degrees of longitude at latitude 10
    b = {0: (f.copy(), ast.copy(), asf.copy(), (5.0, 5.5, 5.23)), 1: (f2, ast.copy(), asf2, (5.0, 5.5, 5.2))}
    r = bench.compare_sets(a, b, [0, 1], per)
    assert r["sites"] == 2 and r["abs_dnza_hist"]["3"] == 1 and r["abs_dnza_hist"]["0"] == 1
    assert r["sites_within_0.01GV"] == 0.5
    outside = 2 * int(((rig < 5.0 - 1e-9) | (rig > 5.5 + 1e-9)).sum())
    assert r["trajectories_outside_band"] == outside
    assert abs(r["flag_mismatch_rate_outside_band"] - 1.0 / outside) < 1e-15
    expect = np.degrees(np.arccos(np.sin(np.radians(10)) ** 2 + np.cos(np.radians(10)) ** 2 * np.cos(np.radians(2))))
    assert abs(r["max_asymptotic_angle_deg"] - expect) < 1e-9
