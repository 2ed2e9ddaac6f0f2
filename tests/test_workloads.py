"""Host-side pieces of the workload runner (cor-model_amd/run_workload.py), on the CPU:
the seeded inputs are reproducible and have the documented ranges, the degree-20 SH
table follows IGRF-14 2020.0 to within its +-1 % perturbation, and the mixed-vs-FP64
comparison counts what it says on hand-made cutoffs and records."""
import sys
from pathlib import Path

import numpy as np
import pytest

PKG = Path(__file__).resolve().parent.parent / "cor-model_amd"
sys.path.insert(0, str(PKG))
import cor_amd as A  # noqa: E402
import run_workload as W  # noqa: E402


def test_seeded_sites_reproducible_and_in_range():
    a = W.seeded_sites(50, 7)
    b = W.seeded_sites(50, 7)
    assert a == b and len(a) == 50
    lat = np.array([s[1] for s in W.seeded_sites(400, 3, -60.0, 60.0)])
    lon = np.array([s[2] for s in W.seeded_sites(400, 3, -60.0, 60.0)])
    # geocentric latitude of a geodetic one in [-60, 60] stays within a fraction of a degree
    assert lat.min() > -60.5 and lat.max() < 60.5 and lon.min() >= 0.0 and lon.max() < 360.0
    r0 = np.array([s[0] for s in a])
    assert np.all((r0 > 1.0) & (r0 < 1.01))     # 20 km above the WGS84 ellipsoid, in Re


def test_storm_drivers_ranges():
    d = np.array(W.storm_drivers(288, 11))
    assert d.shape == (288, 10)
    assert np.all((d[:, 0] >= 0.5) & (d[:, 0] <= 10.0))       # Pdyn
    assert np.all((d[:, 1] >= -300.0) & (d[:, 1] <= 0.0))     # Dst
    assert np.all((d[:, 4:] >= 0.0) & (d[:, 4:] <= 2.0))      # W1..W6
    assert np.array_equal(d, np.array(W.storm_drivers(288, 11)))


def test_degree20_table_follows_igrf2020():
    G, H = W.igrf2020_sh_table(21, 2020)
    G2, H2 = W.igrf2020_sh_table(21, 2020)
    assert np.array_equal(G, G2) and np.array_equal(H, H2) and G.shape == (21, 21)
    # IGRF-14 2020.0 g10, g11, h11 (nT) within the +-1 % perturbation
    assert abs(G[1, 0] / -29404.8 - 1.0) <= 0.0101
    assert abs(G[1, 1] / -1450.9 - 1.0) <= 0.0101
    assert abs(H[1, 1] / 4652.5 - 1.0) <= 0.0101
    # no order above its degree, no h for m = 0
    for n in range(21):
        assert np.all(G[n, n + 1:] == 0.0) and np.all(H[n, n + 1:] == 0.0) and H[n, 0] == 0.0
    # degrees 14..20 filled, with the per-coefficient rms still falling
    rms = [np.sqrt(np.mean(np.concatenate([G[n, :n + 1], H[n, 1:n + 1]]) ** 2)) for n in range(9, 21)]
    assert all(r > 0 for r in rms) and rms[-1] < rms[0]


def traj(status, steps, attempts):
    t = np.zeros(len(status), dtype=A.TRAJ_DTYPE)
    t["status"], t["steps"], t["attempts"] = status, steps, attempts
    return t


def test_axis_ended_counts_only_nan_ended():
    t = traj([2, 2, 0, 1], [1000001, 500, 300, 200], [3, 800, 400, 250])
    assert W.axis_ended(t).tolist() == [True, False, False, False]


def test_compare_cuts_known_answer():
    n = 10
    rows = A.scans_array([(1.003, 10.0, 20.0, 10.0, 20.0, 1.0, 1.0, 10.0, -1.0, 100)] * 2)
    cut = np.zeros(2, dtype=A.CUT_DTYPE)
    cut["n"] = n
    cm, cd = cut.copy(), cut.copy()
    # scan 0: identical; scan 1: Rc differs by 0.03 GV, band [3, 5] in one run, [4, 6] in the other
    cm["rl"], cm["ru"], cm["rc"] = [2.0, 3.0], [4.0, 5.0], [3.0, 4.00]
    cd["rl"], cd["ru"], cd["rc"] = [2.0, 4.0], [4.0, 6.0], [3.0, 4.03]
    st = np.ones(2 * n, np.int32)
    st[4:n] = 0
    st[n + 6:] = 0
    tm = traj(st, np.ones(2 * n), np.ones(2 * n))
    td = tm.copy()
    # a flag differing inside the union band (rigidity 4) does not count, one outside (9) does
    td["status"][n + 3] = 0
    td["status"][n + 8] = 1
    out = W.compare_cuts(cm, cd, tm, td, rows)
    assert out["scans"] == 2
    assert out["within_0.01GV"] == 0.5 and out["within_0.05GV"] == 1.0
    assert out["max_abs_dRc_GV"] == pytest.approx(0.03)
    # outside the union band: scan 0 rigidities 1 and 5..10 (7), scan 1 rigidities 1, 2 and 7..10 (6)
    assert out["trajectories_outside_band"] == 13
    assert out["flag_mismatch_rate_outside_band"] == pytest.approx(1 / 13)
