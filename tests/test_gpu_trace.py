"""GPU cutoff tracer against the oracle (trajectory_simulation_cutoff,
src/trajectory.c:6-161) on the same inputs.

Bar (FP64 path): allowed/forbidden flags identical outside the penumbra band (the
union of the two results' [Rl, Ru]), |dRc| <= one rigidity step, and for allowed
trajectories outside the band the exit point and asymptotic direction within 0.05
degree (great-circle angle).  Mixed precision: |dRc| <= 0.05 GV, asymptotic
directions within 0.5 degree and exit points within 5 degrees (grazing exits), from
1 GV above the band.  Inside the band the trajectories are chaotic: last-bit differences between
OCML and glibc transcendentals flip individual flags there, including isolated allowed
windows below the other result's Rl.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as L

GOLD = Path(__file__).resolve().parent / "golden"

pytestmark = pytest.mark.gpu


def angle_deg(lat1, lon1, lat2, lon2):
    """Great-circle angle between two (latitude, longitude) directions, degrees."""
    p1, p2, dl = np.radians(lat1), np.radians(lat2), np.radians(lon2 - lon1)
    c = np.sin(p1) * np.sin(p2) + np.cos(p1) * np.cos(p2) * np.cos(dl)
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


def compare(scan_row, traj_gpu, cut_gpu, traj_cpu, cut_cpu, rc_tol=0.01, dir_tol=0.05, dir_margin=0.0,
            exit_tol=None):
    """Flags outside the union penumbra band, |dRc|, and the asymptotic direction and
    exit point (great-circle angles) of trajectories allowed in both runs more than
    dir_margin GV above the band: just above Ru the exit geometry is still chaotic, so
    reduced-precision runs are compared from there on.  The exit point on the escape
    surface is ill-conditioned for grazing exits (a small change of the path slides it
    along the surface while the asymptotic direction stays put), so it gets its own
    tolerance exit_tol (default dir_tol)."""
    exit_tol = dir_tol if exit_tol is None else exit_tol
    n = len(traj_cpu)
    rig = scan_row["rig"] + np.arange(n) * scan_row["del_"]
    rl, ru = min(cut_cpu[0], cut_gpu["rl"]), max(cut_cpu[1], cut_gpu["ru"])
    outside = (rig < rl - 1e-9) | (rig > ru + 1e-9)
    a_g = traj_gpu["status"] == 0
    a_c = traj_cpu["status"] == 0
    mism = np.nonzero((a_g != a_c) & outside)[0]
    assert mism.size == 0, f"flag mismatch outside penumbra at rigidities {rig[mism][:10]}"
    assert abs(cut_gpu["rc"] - cut_cpu[2]) <= rc_tol + 1e-9, (cut_gpu, cut_cpu)
    both = a_g & a_c & (rig > ru + dir_margin + 1e-9)
    exit_ang = angle_deg(traj_gpu["td"][both], traj_gpu["fei"][both], traj_cpu["td"][both], traj_cpu["fei"][both])
    asym_ang = angle_deg(traj_gpu["ast"][both], traj_gpu["asf"][both], traj_cpu["ast"][both], traj_cpu["asf"][both])
    if both.any():
        print(f"exit-point angle max {exit_ang.max():.3g} p50 {np.median(exit_ang):.3g}; "
              f"asymptotic angle max {asym_ang.max():.3g} p50 {np.median(asym_ang):.3g} deg over {both.sum()}")
        assert asym_ang.max() < dir_tol and exit_ang.max() < exit_tol, (asym_ang.max(), exit_ang.max())
    return mism


def scan(the0, fi0, r0=1.002439629, rig=0.01, dl=0.01, rk=20.0, nk1=100):
    return L.OScan(r0, the0, fi0, the0, fi0, rig, dl, rk, -1.0, nk1, 0)


def run_both(cor, oracle, osc, kinds, parmod=None, step_limit=1000000, precision=0, **kw):
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=parmod)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    if parmod:
        for k, v in enumerate(parmod):
            po.PARMOD[k] = v
    row = cor.scans_array([(osc.r0, osc.the0, osc.fi0, osc.the1, osc.fi1, osc.rig, osc.del_, osc.rk,
                            osc.zn, osc.nk1)])
    tg, cg = cor.trace(cor.model_set(**kinds), p, row, step_limit=step_limit, precision=precision, **kw)
    assert np.all(np.isin(tg["status"], (0, 1, 2))) and np.all(tg["steps"] >= 1), "unfinished trajectories"
    tc, cc = oracle.trace_scan(osc, L.models(**kinds), po, step_limit)
    return row[0], tg, cg[0], tc, cc


def test_config1_igrf_fp64(cor, oracle):
    """47.0N 8.0E geodetic at 20 km, vertical, 0.01..20 GV, IGRF 2020, 1e6 steps."""
    osc = scan(46.808601777, 8.0)
    row, tg, cg, tc, cc = run_both(cor, oracle, osc, dict(internal=1, external=0))
    print("gpu", cg, "cpu", cc, "steps", tg["steps"].sum(), tc["steps"].sum())
    compare(row, tg, cg, tc, cc)
    # the same trajectories: statuses and accepted-step counts agree almost everywhere
    assert (tg["status"] == tc["status"]).mean() > 0.97
    assert (tg["steps"] == tc["steps"]).mean() > 0.9


@pytest.mark.parametrize("lat,lon", [(0.0, 0.0), (50.0, 120.0), (-35.0, 290.0)])
def test_t96_sites_fp64(cor, oracle, lat, lon):
    osc = scan(lat, lon, rig=0.5, dl=0.05, rk=20.0)
    row, tg, cg, tc, cc = run_both(cor, oracle, osc, dict(internal=1, external=1), parmod=[2, -20, 0, -2])
    print(lat, lon, "gpu", cg, "cpu", cc)
    compare(row, tg, cg, tc, cc, rc_tol=0.05)


def test_step_limit_and_nan_classification(cor, oracle):
    """-s 50: long trajectories end forbidden at the limit exactly like the oracle;
    a negative Pdyn makes T96 NaN everywhere, which the reference classifies as
    forbidden by step limit."""
    osc = scan(46.808601777, 8.0, rig=1.0, dl=0.5, rk=20.0)
    row, tg, cg, tc, cc = run_both(cor, oracle, osc, dict(internal=1, external=0), step_limit=50)
    assert np.array_equal(tg["status"], tc["status"])
    assert np.array_equal(tg["steps"], tc["steps"])
    row, tg, cg, tc, cc = run_both(cor, oracle, osc, dict(internal=1, external=1), parmod=[-1, -20, 0, -2],
                                   step_limit=200)
    assert np.all(tc["status"] == 2) and np.all(tg["status"] == 2)
    assert np.array_equal(tg["steps"], tc["steps"])
    assert abs(cg["rc"] - cc[2]) < 1e-12 and abs(cg["ru"] - cc[1]) < 1e-12


@pytest.mark.parametrize("coop,handover", [(-1, 0), (0, 0), (0, 1)],
                         ids=["per-lane", "default", "all-cooperative"])
def test_mixed_config1_close(cor, oracle, coop, handover):
    """Mixed precision on config 1: the effective cutoff within 0.05 GV of the FP64
    oracle, flags identical away from the (union) penumbra band, exit and asymptotic
    directions of the allowed trajectories there within 0.5 degree -- in the per-lane
    mode only, with the default handover, and with every trajectory in the cooperative
    mode from its first step."""
    osc = scan(46.808601777, 8.0)
    row, tg, cg, tc, cc = run_both(cor, oracle, osc, dict(internal=1, external=0), precision=1, coop=coop,
                                   handover_steps=handover)
    print("mixed", coop, handover, cg, "fp64 cpu", cc)
    compare(row, tg, cg, tc, cc, rc_tol=0.05, dir_tol=0.5, dir_margin=1.0, exit_tol=5.0)


@pytest.mark.parametrize("lat,lon", [(0.0, 0.0), (50.0, 120.0)])
def test_t96_sites_cooperative(cor, oracle, lat, lon):
    """IGRF + T96 with every trajectory in the cooperative mode, against the FP64 oracle."""
    osc = scan(lat, lon, rig=0.5, dl=0.05, rk=20.0)
    row, tg, cg, tc, cc = run_both(cor, oracle, osc, dict(internal=1, external=1), parmod=[2, -20, 0, -2],
                                   precision=1, handover_steps=1)
    print(lat, lon, "cooperative", cg, "cpu", cc)
    # reduced precision moves individual flags only inside the chaotic band, so Rc can
    # move by up to the band's width; half of it is the bar (at least two steps)
    rc_tol = max(0.1, 0.5 * (cc[1] - cc[0]))
    compare(row, tg, cg, tc, cc, rc_tol=rc_tol, dir_tol=0.5, dir_margin=1.0, exit_tol=5.0)


@pytest.mark.parametrize("lat,lon", [(0.0, 0.0), (50.0, 120.0), (-35.0, 290.0)])
def test_t96_sites_fp64_cooperative(cor, oracle, lat, lon):
    """The FP64 tracer with its cooperative tail (opt-in), every trajectory cooperative
    from its first step: the FP64 bar of the per-lane FP64 test at the same sites (flags
    identical away from the band, Rc within 0.05 GV of the oracle)."""
    osc = scan(lat, lon, rig=0.5, dl=0.05, rk=20.0)
    row, tg, cg, tc, cc = run_both(cor, oracle, osc, dict(internal=1, external=1), parmod=[2, -20, 0, -2],
                                   coop=1, handover_steps=1)
    print(lat, lon, "fp64 cooperative", cg, "cpu", cc)
    # handed over after the first accepted step (recorded as 2), unless it ended there
    assert np.all(np.isin(tg["handover"], (0, 2))) and (tg["handover"] == 2).mean() > 0.9
    compare(row, tg, cg, tc, cc, rc_tol=0.05)


def test_fp64_cooperative_is_deterministic(cor):
    """The FP64 tracer with the cooperative tail: bit-identical records and cutoffs in
    repeated runs and both queue orders (handover after 300 steps)."""
    sites = cor.grid_sites()[::37]
    rows = cor.scans_array([(r0, lat, lon, lat, lon, 0.05, 0.05, 20.0, -1.0, 100) for r0, lat, lon in sites])
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    out = [cor.trace(ms, p, rows, step_limit=30000, precision=cor.PREC_FP64, coop=1, handover_steps=300,
                     queue_order=q) for q in (1, 1, 0)]
    print(len(rows), "scans,", int((out[0][0]["handover"] > 0).sum()), "trajectories cooperative")
    assert int((out[0][0]["handover"] > 0).sum()) > 0
    for t, c in out[1:]:
        assert np.array_equal(t, out[0][0]) and np.array_equal(c, out[0][1])


def test_cooperative_handover_many_scans(cor):
    """Many scans through the handover: every trajectory finishes exactly once (status
    set, steps counted) whichever mode ran it, a handed-over trajectory left the
    per-lane mode exactly at the handover step, and the per-scan cutoffs of the per-lane
    and cooperative runs agree within the penumbra noise.  Cases: per-lane only; the
    default handover; handover after 2 000 and after 50 steps."""
    sites = cor.grid_sites()[::97]
    rows = cor.scans_array([(r0, lat, lon, lat, lon, 0.05, 0.05, 20.0, -1.0, 100) for r0, lat, lon in sites])
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    out = {}
    cases = [(-1, 0), (0, 0), (0, 2000), (0, 50)]
    for coop, hs in cases:
        tr, cut = cor.trace(ms, p, rows, step_limit=200000, precision=cor.PREC_MIXED, coop=coop, handover_steps=hs)
        assert np.all(np.isin(tr["status"], (0, 1, 2))) and np.all(tr["steps"] >= 1)
        ho = tr["handover"][tr["handover"] > 0]
        if coop < 0:
            assert len(ho) == 0
        elif hs > 0:
            # handed over exactly after hs per-lane steps (recorded as 1 + that step)
            print("handed over", len(ho), "at", np.unique(ho))
            assert len(ho) > 0 and np.all(ho == hs + 1)
            assert np.all(tr["steps"][tr["handover"] > 0] > hs)
        out[(coop, hs)] = cut
    for key in cases[1:]:
        d = np.abs(out[key]["rc"] - out[(-1, 0)]["rc"])
        print(key, "max |dRc|", d.max(), "frac <= 0.1 GV", (d <= 0.1 + 1e-9).mean())
        assert (d <= 0.1 + 1e-9).mean() >= 0.8


def test_mixed_default_is_deterministic(cor):
    """The default mixed path (per-lane up to the handover step, then the cooperative
    mode with one summation order for every group width) gives bit-identical records and
    cutoffs in repeated runs and in both queue orders, over a few hundred grid scans:
    scheduling decides only when a trajectory runs, not what it computes
    (the reference computes each trajectory sequentially in one thread,
    src/trajectory.c:12-153).  A short handover puts most long trajectories through
    the cooperative mode."""
    sites = cor.grid_sites()[::9]
    rows = cor.scans_array([(r0, lat, lon, lat, lon, 0.05, 0.05, 20.0, -1.0, 100) for r0, lat, lon in sites])
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    runs = [(0, 1), (0, 1), (0, 0), (300, 1), (300, 0), (300, 1)]
    ref = {}
    for hs, q in runs:
        tr, cut = cor.trace(ms, p, rows, step_limit=100000, precision=cor.PREC_MIXED, handover_steps=hs,
                            queue_order=q)
        if hs not in ref:
            ref[hs] = (tr, cut)
            print(len(rows), "scans, handover", hs, ":", int((tr["handover"] > 0).sum()), "trajectories cooperative")
            continue
        t0, c0 = ref[hs]
        assert np.array_equal(tr, t0), (hs, q, int((tr != t0).sum()))
        assert np.array_equal(cut, c0), (hs, q)


def test_per_lane_compaction_changes_no_result(cor):
    """With the cooperative mode on but a handover step no trajectory reaches, every
    trajectory stays in the per-lane mode, and after the queue drains the sparse
    per-lane wavefronts hand their trajectories to the per-lane pool, from which
    wavefronts with idle lanes (and idle workers coming back) continue them.  The
    records and cutoffs equal the per-lane-only run bit for bit, in FP64 and mixed
    precision, and a second compacting run equals the first."""
    sites = cor.grid_sites()[::7]
    rows = cor.scans_array([(r0, lat, lon, lat, lon, 0.05, 0.05, 20.0, -1.0, 100) for r0, lat, lon in sites])
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    for prec, coop, reps in ((cor.PREC_MIXED, 0, 2), (cor.PREC_FP64, 1, 1)):
        t0, c0 = cor.trace(ms, p, rows, step_limit=100000 if prec else 30000, precision=prec, coop=-1)
        for _ in range(reps):
            t1, c1 = cor.trace(ms, p, rows, step_limit=100000 if prec else 30000, precision=prec, coop=coop,
                               handover_steps=2000000000)
            assert (t1["handover"] == 0).all()
            assert np.array_equal(t1, t0), (prec, int((t1 != t0).sum()))
            assert np.array_equal(c1, c0), prec
        print(len(rows), "scans, precision", prec, ": compacting runs equal the per-lane-only run")


TS05_STORM = [3, -100, 2, -5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_ts05_site(cor, oracle, precision):
    """IGRF + TS05 at the golden TS05 site (52N 30E, storm drivers, 1..20 GV in 0.25,
    2e4 steps) against the oracle: FP64 at the FP64 bar; mixed with Rc within half the
    penumbra width."""
    osc = scan(52.0, 30.0, r0=1.00280, rig=1.0, dl=0.25, rk=20.0, nk1=100)
    row, tg, cg, tc, cc = run_both(cor, oracle, osc, dict(internal=1, external=2), parmod=TS05_STORM,
                                   step_limit=20000, precision=precision)
    print("ts05", precision, cg, "cpu", cc)
    if precision == 0:
        compare(row, tg, cg, tc, cc, rc_tol=0.25)
    else:
        compare(row, tg, cg, tc, cc, rc_tol=max(0.25, 0.5 * (cc[1] - cc[0])), dir_tol=0.5, dir_margin=1.0,
                exit_tol=5.0)


def test_sh_ts05_site_fp64(cor, oracle):
    """The 1-degree map's field (degree-20 SH core field + TS05 storm drivers, GSM frame
    of 2020.0, no IGRF) at one site, FP64 against the oracle."""
    rng = np.random.default_rng(11)
    nm = 21
    g = np.zeros((nm, nm))
    h = np.zeros((nm, nm))
    for n in range(1, nm):
        sc = 30000.0 * 0.3 ** n
        for m in range(n + 1):
            g[n, m] = sc * rng.normal()
            h[n, m] = sc * rng.normal() if m else 0.0
    osc = scan(-20.0, 150.0, r0=1.00310, rig=1.0, dl=0.25, rk=20.0, nk1=100)
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=TS05_STORM)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    for k, v in enumerate(TS05_STORM):
        po.PARMOD[k] = v
    row = cor.scans_array([(osc.r0, osc.the0, osc.fi0, osc.the1, osc.fi1, osc.rig, osc.del_, osc.rk, osc.zn, osc.nk1)])
    tg, cg = cor.trace(cor.model_set(cor.INTERNAL_NONE, cor.EXTERNAL_TS05, sh=(g, h)), p, row, step_limit=20000,
                       precision=cor.PREC_FP64)
    tc, cc = oracle.trace_scan(osc, L.models(internal=0, external=2, crustal=1, chaos=(g, h)), po, 20000)
    print("sh+ts05", cg[0], "cpu", cc)
    compare(row[0], tg, cg[0], tc, cc, rc_tol=0.25)


@pytest.mark.parametrize("precision,coop", [(0, 0), (1, -1)], ids=["fp64", "mixed-per-lane"])
def test_queue_order_leaves_results_unchanged(cor, precision, coop):
    """The queue order only decides when a trajectory runs, not what it computes: FP64 and
    the per-lane mixed path give bit-identical per-trajectory records and cutoffs in
    rigidity order (0) and longest-expected-first order (1, the benchmark's)."""
    sites = cor.grid_sites()[::211]
    rows = cor.scans_array([(r0, lat, lon, lat, lon, 0.05, 0.05, 20.0, -1.0, 100) for r0, lat, lon in sites])
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    out = [cor.trace(ms, p, rows, step_limit=100000 if precision else 30000, precision=precision, coop=coop,
                     queue_order=q) for q in (0, 1)]
    (t0, c0), (t1, c1) = out
    assert np.array_equal(t0, t1) and np.array_equal(c0, c1)
    print(len(rows), "scans,", len(t0), "trajectories identical in both orders")


GRID5 = ["grid5_+60_000", "grid5_+40_120", "grid5_+00_090", "grid5_-30_300", "grid5_-60_150"]


def grid5_reference():
    """The reference's own text for five grid sites at the benchmark's resolution (IGRF-14
    2020 + T96 quiet, 0.01..20 GV in 0.01 GV steps, nk1 400, 1e6 steps) -> scans, and per
    site the allowed rigidity indices and (Rl, Ru, Rc)."""
    rows, refs = [], []
    for name in GRID5:
        h = L.read_cutoff_header(GOLD / f"{name}.in")
        rows.append((h["r0"], h["the0"], h["fi0"], h["the1"], h["fi1"], h["rig"], h["del"], h["rk"], h["zn"], h["nk1"]))
        allowed, cut = L.parse_cutoff_output((GOLD / f"{name}.out").read_text())
        refs.append(({int(round((r - h["rig"]) / h["del"])) for r in allowed["rig"]}, cut))
    return rows, refs


@pytest.mark.parametrize("precision,coop", [(0, 0), (0, 1), (1, 0)], ids=["fp64", "fp64-cooperative", "mixed"])
def test_grid_sites_at_benchmark_resolution_against_reference(cor, precision, coop):
    """Five sites of the benchmark grid at its own resolution against the reference's
    text.  The north star's bar: allowed/forbidden flags identical to the reference's
    everywhere outside the union penumbra band (for every path), and Rc within 0.01 GV.
    Inside the band the trajectories are chaotic: two FP64 implementations whose
    transcendentals differ in the last bits (OCML here, glibc in the reference) already
    flip some of them, so the FP64 bar on Rc is 0.01 GV at four of the five sites and
    the penumbra's half width at the fifth; the mixed path is held to the same."""
    rows, refs = grid5_reference()
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    tr, cut = cor.trace(ms, p, cor.scans_array(rows), step_limit=1000000, precision=precision, coop=coop)
    # the cooperative tail takes over at the library's default handover step (98 304
    # accepted steps in mixed precision, 65 536 in FP64), recorded as 1 + that step
    ho = tr["handover"][tr["handover"] > 0]
    if precision == 1 or coop > 0:
        assert len(ho) > 0 and np.all(ho == (98304 if precision == 1 else 65536) + 1), np.unique(ho)
    else:
        assert len(ho) == 0
    per = 2000
    within = 0
    for k, (name, (allowed, rcut)) in enumerate(zip(GRID5, refs)):
        t = tr[k * per:(k + 1) * per]
        g = cut[k]
        lo, hi = min(g["rl"], rcut[0]), max(g["ru"], rcut[1])
        rig = 0.01 + 0.01 * np.arange(per)
        out = (rig < lo - 1e-9) | (rig > hi + 1e-9)
        mine = t["status"] == 0
        ref = np.zeros(per, bool)
        ref[list(allowed)] = True
        drc = abs(g["rc"] - rcut[2])
        print(name, "gpu", (g["rl"], g["ru"], g["rc"]), "reference", rcut, "|dRc|", round(drc, 3),
              "flags outside band", int(out.sum()), "mismatches", int((mine[out] != ref[out]).sum()))
        assert np.array_equal(mine[out], ref[out]), name
        within += drc <= 0.01 + 1e-9
        assert drc <= max(0.01, 0.5 * (rcut[1] - rcut[0])) + 1e-9, name
    assert within >= 4


def test_autotuner_picks_result_neutral_launch_parameters(cor):
    """cor_trace_autotune on a seeded sample of 24 grid scans (IGRF + T96, mixed): the
    chosen waves per SIMD, block size and queue order are valid, and a plan launched
    with them gives the default launch's results bit for bit (none of the tuned
    parameters changes a trajectory's arithmetic)."""
    sites = cor.grid_sites(15.0, 30.0, 20.0)
    rows = cor.scans_array([(r0, la, lo, la, lo, 0.2, 0.2, 20.0, -1.0, 100) for (r0, la, lo) in sites])
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    tuned = cor.autotune(ms, p, rows, precision=cor.PREC_MIXED, step_limit=50000, sample_scans=24, seed=7)
    print("autotuned", tuned)
    assert tuned["waves_per_simd"] >= 1 and tuned["block_threads"] in (64, 256) and tuned["queue_order"] in (0, 1)
    assert tuned["sample_ms"] > 0
    kw = dict(step_limit=50000, precision=cor.PREC_MIXED)
    t0, c0 = cor.trace(ms, p, rows, **kw)
    t1, c1 = cor.trace(ms, p, rows, waves_per_simd=tuned["waves_per_simd"], block_threads=tuned["block_threads"],
                       queue_order=tuned["queue_order"], **kw)
    assert np.array_equal(t0, t1) and np.array_equal(c0, c1)
