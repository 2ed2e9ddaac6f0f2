#!/usr/bin/env python3
"""The larger cutoff workloads on one GPU, each printing one JSON line (progress on
stderr).  All inputs are synthetic and seeded; the field models are the library's.

  storm       TS05 storm series: 288 five-minute epochs of 2020-01-01 with seeded
              drivers (Pdyn U(0.5, 10) nPa; Dst a smooth random walk clamped to
              [-300, 0] nT; By, Bz ~ N(0, 5) nT; W1..W6 U(0, 2); the dipole tilt from
              each epoch's time), IGRF-14 + TS05, 50 seeded sites (area-uniform) at
              20 km, vertical, 0.01-20 GV in 0.01 GV steps: 28.8 M trajectories in one
              multi-epoch launch in mixed precision; mixed and FP64 on the first
              epochs (at a reduced step limit), dRc per site per epoch.
  map         high-resolution global map: a degree-20 spherical-harmonic core field
              (IGRF-14 2020.0 Gauss coefficients with a seeded +-1 % perturbation for
              degrees 1..13, seeded coefficients continuing the spectrum's decay for
              14..20) as the SH model, plus TS05 (Pdyn 3, Dst -100, By 2, Bz -5,
              W1..W6 0.5); 1-degree grid (181 x 360 = 65 160 sites) at 20 km, vertical,
              0.02-20 GV in 0.02 GV steps (65.2 M trajectories), mixed precision with
              autotuned launch parameters; mixed and FP64 on a 1 % seeded subsample.
  directions  apparent cutoffs and asymptotic directions: IGRF-14 + T96 (Pdyn 2,
              Dst -20, By 0, Bz -2), 64 seeded sites (latitude U(-60, 60), longitude
              U(0, 360)) at 20 km, 193 directions each (vertical plus zenith 10..80 x
              azimuth 0..345 in 15-degree steps), 0.01-20 GV in 0.01 GV steps, then each
              scan refined to 0.001 GV over its penumbra band; mixed precision, FP64 on a
              seeded subsample of scans.

    python cor-model_amd/run_workload.py storm [--epochs 288] [--sites 50]
    python cor-model_amd/run_workload.py map [--lat-step 1] [--fp64-frac 0.01]
    python cor-model_amd/run_workload.py directions [--sites 64]
    python cor-model_amd/run_workload.py modes

  modes       the benchmark grid (5-degree global grid, IGRF-14 + T96, 0.01-20 GV)
              split into the tracer's two modes, one launch each, so that a counter
              run sees each mode apart: the per-lane mode (the grid at step limit
              98 304, the mixed tracer's handover step, cooperative mode off), then the
              cooperative mode (the trajectories that reach 98 304 steps, each a
              single-rigidity scan traced cooperatively from its first step to the 1e6
              step limit).

Throughput is trajectories/s and accepted steps/s over the tracer kernel's own time
(HIP events on its stream); the roofline fraction uses the op-count flop model of the
workload's model set where profiles/flop_model.json has one.
"""
import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
sys.path.insert(0, str(PKG))
import cor_amd as A  # noqa: E402

PEAK_FP32 = 157.3   # TFLOP/s, MI355X vector FP32


def log(msg):
    print(f"[workload] {msg}", file=sys.stderr, flush=True)


def flop_model(key):
    f = PKG.parent / "profiles" / "flop_model.json"
    if f.exists():
        m = json.loads(f.read_text())
        if key in m:
            return m[key]["flops_per_step"]
    return None


def seeded_sites(n, seed, lat_lo=-90.0, lat_hi=90.0, h_km=20.0):
    """n sites, area-uniform in geodetic latitude [lat_lo, lat_hi] and uniform in
    longitude, converted to the reference's geocentric (r0, latitude)."""
    rng = np.random.default_rng(seed)
    s = rng.uniform(np.sin(np.radians(lat_lo)), np.sin(np.radians(lat_hi)), n)
    lats = np.degrees(np.arcsin(s))
    lons = rng.uniform(0.0, 360.0, n)
    out = []
    for la, lo in zip(lats, lons):
        r0, lc = A.geodetic_to_geocentric(float(la), h_km)
        out.append((r0, lc, float(lo)))
    return out


def vertical_rows(sites, rig, dl, rk, nk1):
    return A.scans_array([(r0, la, lo, la, lo, rig, dl, rk, -1.0, nk1) for (r0, la, lo) in sites])


def storm_drivers(n_epochs, seed):
    """TS05 drivers per five-minute epoch (reference PARMOD order: Pdyn, Dst, By, Bz,
    W1..W6)."""
    rng = np.random.default_rng(seed)
    dst, out = -20.0, []
    for _ in range(n_epochs):
        dst = float(np.clip(dst + rng.normal(0.0, 8.0), -300.0, 0.0))
        out.append([rng.uniform(0.5, 10.0), dst, rng.normal(0.0, 5.0), rng.normal(0.0, 5.0)] +
                   list(rng.uniform(0.0, 2.0, 6)))
    return out


def timed_run(plan, label, every=20.0):
    """One run of the plan; while it runs, the count of trajectories it has finished is
    logged every `every` seconds (read from the kernel's own counter)."""
    t0 = time.perf_counter()
    plan.run()
    t_log = t0
    while True:
        fin, done = plan.progress()
        if done:
            break
        now = time.perf_counter()
        if now - t_log >= every:
            log(f"{label}: {fin} of {plan.n} trajectories finished after {now - t0:.0f} s")
            t_log = now
        time.sleep(0.5)
    plan.sync()
    wall = time.perf_counter() - t0
    st = plan.stats()
    log(f"{label}: {st.kernel_ms / 1e3:.1f} s kernel, {wall:.1f} s wall, {st.n_traj} trajectories, "
        f"{st.steps} steps")
    return st, wall


def axis_ended(tr):
    """Trajectories whose state turned NaN (a start exactly on the polar axis, where the
    SH synthesis divides by sin(theta) = 0, as the reference's does): they end at once
    and count the step limit as their steps, by the reference's convention, but
    compute only their attempts."""
    return (tr["status"] == 2) & (tr["attempts"] < tr["steps"])


def rate_block(st, flops_key, tr=None):
    """Rates over the kernel's time.  With the per-trajectory records, the steps that
    were computed: the NaN-ended trajectories count their attempts, not the step limit."""
    k = st.kernel_ms * 1e-3
    steps = int(st.steps)
    out = {"trajectories": int(st.n_traj), "kernel_s": k, "trajectories_per_s": st.n_traj / k,
           "steps": steps, "attempts": int(st.attempts)}
    if tr is not None:
        nan = axis_ended(tr)
        if nan.any():
            steps = int(tr["steps"][~nan].sum(dtype=np.int64) + tr["attempts"][nan].sum(dtype=np.int64))
            out["nan_ended_trajectories"] = int(nan.sum())
            out["steps_computed"] = steps
    out["steps_per_s"] = steps / k
    f = flop_model(flops_key) if flops_key else None
    if f is not None:
        ach = f * steps / k / 1e12
        out["roofline"] = {"bound": "valu", "achieved": ach, "peak": PEAK_FP32, "unit": "TFLOP/s",
                           "frac": ach / PEAK_FP32, "flops_per_step": f, "flop_model": flops_key}
    return out


def compare_cuts(cm, cd, tm, td, rows_per_scan):
    """Mixed against FP64 on the same scans: |dRc| statistics and the per-trajectory
    flag mismatch rate outside each scan's union penumbra band."""
    d = np.abs(cm["rc"] - cd["rc"]).ravel()
    mism = outside = 0
    tm = tm.reshape(len(cm.ravel()), -1)
    td = td.reshape(len(cd.ravel()), -1)
    for k, (a, b) in enumerate(zip(cm.ravel(), cd.ravel())):
        n = int(a["n"])
        rig = rows_per_scan[k]["rig"] + rows_per_scan[k]["del_"] * np.arange(n)
        lo, hi = min(a["rl"], b["rl"]), max(a["ru"], b["ru"])
        out = (rig < lo - 1e-9) | (rig > hi + 1e-9)
        outside += int(out.sum())
        mism += int(((tm[k][:n]["status"] == 0) != (td[k][:n]["status"] == 0))[out].sum())
    return {"scans": int(d.size), "within_0.01GV": float(np.mean(d <= 0.01 + 1e-9)),
            "within_0.05GV": float(np.mean(d <= 0.05 + 1e-9)), "max_abs_dRc_GV": float(d.max()),
            "mean_abs_dRc_GV": float(d.mean()), "p99_abs_dRc_GV": float(np.quantile(d, 0.99)),
            "flag_mismatch_rate_outside_band": mism / max(outside, 1), "trajectories_outside_band": outside}


def run_storm(args):
    sites = seeded_sites(args.sites, args.seed)
    rows = vertical_rows(sites, 0.01, 0.01, 20.0, args.nk1)
    drivers = storm_drivers(args.epochs, args.seed + 1)
    params = []
    for k, pm in enumerate(drivers):
        m = 5 * k
        params.append(A.recalc(2020, 1, 1, 1, m // 60, m % 60, 0, parmod=pm))
    ms = A.model_set(A.INTERNAL_IGRF, A.EXTERNAL_TS05)
    log(f"storm: {args.epochs} epochs x {args.sites} sites x 2000 rigidities")
    pm = A.Plan(ms, params, rows, precision=A.PREC_MIXED, step_limit=args.step_limit, queue_order=1)
    st_m, wall_m = timed_run(pm, "mixed, all epochs")
    _, cm = pm.results(traj=False)
    pm.close()
    out = {"workload": "TS05 storm series (IGRF-14 + TS05), one multi-epoch launch",
           "epochs": args.epochs, "sites": args.sites, "nk1": args.nk1, "step_limit": args.step_limit,
           "mixed": rate_block(st_m, "ts05_igrf"), "mixed_wall_s": wall_m,
           "rc_range_GV": [float(cm["rc"].min()), float(cm["rc"].max())]}
    if not args.no_fp64:
        # mixed against FP64 on the first cmp_epochs epochs, both at cmp_step_limit: the
        # FP64 per-lane TS05 tracer's trapped trajectories take minutes to reach 1e6 steps
        ne, sl = min(args.cmp_epochs, args.epochs), args.cmp_step_limit
        res = {}
        for prec, name in ((A.PREC_MIXED, "mixed"), (A.PREC_FP64, "fp64")):
            p = A.Plan(ms, params[:ne], rows, precision=prec, step_limit=sl, queue_order=1)
            st, _ = timed_run(p, f"{name} ({ne} epochs, step limit {sl})")
            res[name] = (st,) + p.results()
            p.close()
        (st_a, ta, ca), (st_d, td, cd) = res["mixed"], res["fp64"]
        out["comparison"] = {"epochs": ne, "step_limit": sl, "mixed": rate_block(st_a, "ts05_igrf"),
                             "fp64": rate_block(st_d, "ts05_igrf")}
        out["comparison"]["speedup_mixed_vs_fp64"] = st_d.kernel_ms / st_a.kernel_ms
        per_scan = np.concatenate([rows] * ne)
        out["comparison"]["mixed_vs_fp64"] = compare_cuts(ca, cd, ta, td, per_scan)
        d = np.abs(ca["rc"] - cd["rc"])
        out["comparison"]["abs_dRc_GV_per_site_per_epoch"] = [[round(float(v), 3) for v in row] for row in d]
    return out


def igrf2020_sh_table(nm, seed):
    """[nm][nm] Gauss-coefficient table (row n, column m, nT) of a degree nm-1 core
    field: IGRF-14 2020.0 for n <= 13 with a seeded +-1 % perturbation, and for
    14 <= n < nm seeded N(0, s_n) coefficients, s_n continuing the geometric decay of
    the per-coefficient rms from degree 9 to degree 13."""

    class Coefs(C.Structure):
        _fields_ = [("coefs", C.POINTER(C.POINTER(C.c_double))), ("coefs_size", C.c_uint32),
                    ("years_step", C.c_uint32), ("min_year", C.c_uint32), ("max_year", C.c_uint32),
                    ("extrapolate", C.c_bool)]

    d = Coefs.in_dll(A.lib(), "IGRF_14_COEFS")
    k = (2020 - d.min_year) // d.years_step
    g_row, h_row = d.coefs[2 * k], d.coefs[2 * k + 1]
    rng = np.random.default_rng(seed)
    G = np.zeros((nm, nm))
    H = np.zeros((nm, nm))
    for n in range(1, min(13, nm - 1) + 1):
        for m in range(n + 1):
            i = n * (n + 1) // 2 + m
            G[n, m] = g_row[i] * (1.0 + 0.01 * rng.uniform(-1, 1))
            H[n, m] = h_row[i] * (1.0 + 0.01 * rng.uniform(-1, 1)) if m > 0 else 0.0

    def rms(n):
        v = np.concatenate([G[n, :n + 1], H[n, 1:n + 1]])
        return float(np.sqrt(np.mean(v * v)))

    q = (rms(13) / rms(9)) ** 0.25
    for n in range(14, nm):
        s = rms(13) * q ** (n - 13)
        G[n, :n + 1] = rng.normal(0.0, s, n + 1)
        H[n, 1:n + 1] = rng.normal(0.0, s, n)
    return G, H


def grid_sites_step(step):
    return A.grid_sites(step, step, 20.0)


def autotune(ms, params, rows, args):
    """Pick the launch parameters on a seeded sample of the scans (cor_trace_autotune)."""
    opts = A.autotune(ms, params, rows, precision=A.PREC_MIXED, step_limit=args.step_limit,
                      sample_scans=args.tune_sample, seed=args.seed)
    log(f"autotuned: {opts}")
    return opts


def run_map(args):
    G, H = igrf2020_sh_table(21, args.seed)
    ms = A.model_set(A.INTERNAL_NONE, A.EXTERNAL_TS05, sh=(G, H))
    params = A.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[3.0, -100.0, 2.0, -5.0] + [0.5] * 6)
    sites = grid_sites_step(args.lat_step)
    rows = vertical_rows(sites, 0.02, 0.02, 20.0, args.nk1)
    log(f"map: {len(rows)} sites x 1000 rigidities")
    out = {"workload": f"{args.lat_step:g}-degree global map, SH degree-20 core field (IGRF-2020-scaled) + TS05",
           "sites": len(rows), "nk1": args.nk1, "step_limit": args.step_limit}
    if not args.cmp_only:
        out.update(map_full(ms, params, rows, args))
    if not args.no_fp64:
        out.update(map_comparison(ms, params, rows, args))
    return out


def map_full(ms, params, rows, args):
    tuned = autotune(ms, params, rows, args) if not args.no_tune else {}
    pm = A.Plan(ms, params, rows, precision=A.PREC_MIXED, step_limit=args.step_limit,
                queue_order=tuned.get("queue_order", 1), waves_per_simd=tuned.get("waves_per_simd", 0),
                block_threads=tuned.get("block_threads", 0))
    st_m, wall_m = timed_run(pm, "mixed")
    tm, cm = pm.results()
    pm.close()
    per = 1000
    axis_sites = np.unique(np.nonzero(axis_ended(tm))[0] // per)
    out = {"autotuned": tuned, "mixed": rate_block(st_m, "ts05_sh20", tm), "mixed_wall_s": wall_m,
           "axis_sites": [int(k) for k in axis_sites], "rc_range_GV": [float(cm["rc"].min()), float(cm["rc"].max())]}
    return out


def map_comparison(ms, params, rows, args):
    """FP64 rerun on a seeded subsample of the sites, compared with mixed on the same
    sites, both at cmp_step_limit (the FP64 per-lane TS05 tracer's trapped trajectories
    take minutes to reach 1e6 steps)."""
    out = {}
    rng = np.random.default_rng(args.seed + 2)
    idx = np.sort(rng.choice(len(rows), max(1, int(round(args.fp64_frac * len(rows)))), replace=False))
    sl = args.cmp_step_limit
    res = {}
    for prec, name in ((A.PREC_MIXED, "mixed"), (A.PREC_FP64, "fp64")):
        p = A.Plan(ms, params, rows[idx], precision=prec, step_limit=sl, queue_order=1)
        st, _ = timed_run(p, f"{name} ({len(idx)} sites, step limit {sl})")
        res[name] = (st,) + p.results()
        p.close()
    (st_a, ta, ca), (st_d, td, cd) = res["mixed"], res["fp64"]
    # scans starting on the polar axis, where the SH synthesis divides by sin(colatitude)
    # = 0 and a trajectory whose state turns NaN ends at once (the reference's
    # convention), in either precision: they are reported apart
    nan_d = axis_ended(td).reshape(len(idx), -1).any(axis=1)
    nan_a = axis_ended(ta).reshape(len(idx), -1).any(axis=1)
    keep = ~(nan_d | nan_a)
    out["comparison"] = {"sites": int(len(idx)), "step_limit": sl, "mixed": rate_block(st_a, "ts05_sh20", ta),
                         "fp64": rate_block(st_d, "ts05_sh20", td),
                         "speedup_mixed_vs_fp64": st_d.kernel_ms / st_a.kernel_ms,
                         "axis_sites_excluded": int((~keep).sum()),
                         "mixed_vs_fp64": compare_cuts(ca[keep], cd[keep],
                                                       ta.reshape(len(idx), -1)[keep],
                                                       td.reshape(len(idx), -1)[keep], rows[idx][keep])}
    return out


def run_directions(args):
    sites = seeded_sites(args.sites, args.seed, -60.0, 60.0)
    dirs = [(0.0, 0.0)] + [(float(z), float(a)) for z in range(10, 81, 10) for a in range(0, 346, 15)]
    rows = []
    for (r0, la, lo) in sites:
        for (ze, az) in dirs:
            t1, f1 = (la, lo) if ze == 0.0 else A.direction_to_geo(la, lo, ze, az)
            rows.append((r0, la, lo, t1, f1, 0.01, 0.01, 20.0, -1.0, args.nk1))
    rows = A.scans_array(rows)
    ms = A.model_set(A.INTERNAL_IGRF, A.EXTERNAL_T96)
    params = A.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2.0, -20.0, 0.0, -2.0])
    log(f"directions: {args.sites} sites x {len(dirs)} directions x 2000 rigidities")
    pm = A.Plan(ms, params, rows, precision=A.PREC_MIXED, step_limit=args.step_limit, queue_order=1,
                handover_steps=args.handover_steps)
    st_c, wall_c = timed_run(pm, "mixed coarse")
    tm, cm = pm.results()
    pm.close()
    fine = A.refine_scans(rows, cm, 10)
    pf = A.Plan(ms, params, fine, precision=A.PREC_MIXED, step_limit=args.step_limit, queue_order=1,
                handover_steps=args.handover_steps)
    st_f, wall_f = timed_run(pf, "mixed refinement (0.001 GV over each penumbra band)")
    tf, cf = pf.results()
    pf.close()
    allowed = tm["status"] == 0
    out = {"workload": "apparent cutoffs and asymptotic directions, IGRF-14 + T96, refined penumbra",
           "sites": args.sites, "directions": len(dirs), "scans": len(rows), "nk1": args.nk1,
           "handover_steps": args.handover_steps,
           "mixed_coarse": rate_block(st_c, "t96_igrf"), "mixed_refine": rate_block(st_f, "t96_igrf"),
           "refined_trajectories": int(st_f.n_traj), "allowed_with_asymptotic_direction": int(allowed.sum()),
           "wall_s": wall_c + wall_f,
           "trajectories_per_s_both_passes": (st_c.n_traj + st_f.n_traj) / ((st_c.kernel_ms + st_f.kernel_ms) * 1e-3),
           "rc_coarse_range_GV": [float(cm["rc"].min()), float(cm["rc"].max())],
           "rc_refined_range_GV": [float(cf["rc"].min()), float(cf["rc"].max())]}
    if not args.no_fp64:
        rng = np.random.default_rng(args.seed + 3)
        idx = np.sort(rng.choice(len(rows), args.fp64_scans, replace=False))
        pd = A.Plan(ms, params, rows[idx], precision=A.PREC_FP64, step_limit=args.step_limit, queue_order=1, coop=1)
        st_d, _ = timed_run(pd, f"fp64 ({len(idx)} scans, cooperative tail)")
        td, cd = pd.results()
        pd.close()
        per = 2000
        tm_sub = np.concatenate([tm[i * per:(i + 1) * per] for i in idx])
        out["fp64_subsample"] = rate_block(st_d, "t96_igrf")
        out["mixed_vs_fp64"] = compare_cuts(cm[idx], cd, tm_sub, td, rows[idx])
        both = (tm_sub["status"] == 0) & (td["status"] == 0)
        p1, p2 = np.radians(tm_sub["ast"][both]), np.radians(td["ast"][both])
        dl = np.radians(tm_sub["asf"][both] - td["asf"][both])
        c = np.clip(np.sin(p1) * np.sin(p2) + np.cos(p1) * np.cos(p2) * np.cos(dl), -1.0, 1.0)
        ang = np.degrees(np.arccos(c))
        out["mixed_vs_fp64"]["asymptotic_angle_deg_p50_p99_max"] = [float(np.median(ang)),
                                                                   float(np.quantile(ang, 0.99)),
                                                                   float(ang.max())]
    return out


def run_modes(args):
    sites = A.grid_sites(5.0, 5.0, 20.0)
    rows = vertical_rows(sites, 0.01, 0.01, 20.0, args.nk1)
    ms = A.model_set(A.INTERNAL_IGRF, A.EXTERNAL_T96)
    params = A.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2.0, -20.0, 0.0, -2.0])
    ho = 98304
    log(f"modes: {len(rows)} sites x 2000 rigidities; per-lane launch at step limit {ho}")
    pl = A.Plan(ms, params, rows, precision=A.PREC_MIXED, step_limit=ho, queue_order=1, coop=-1)
    st_l, _ = timed_run(pl, "per-lane mode")
    tl, _ = pl.results()
    pl.close()
    per = 2000
    hit = np.flatnonzero(tl["steps"] >= ho)
    scan_of, rig_i = hit // per, hit % per
    one = [(rows[k]["r0"], rows[k]["the0"], rows[k]["fi0"], rows[k]["the1"], rows[k]["fi1"],
            rows[k]["rig"] + i * rows[k]["del_"], rows[k]["del_"], rows[k]["rig"] + i * rows[k]["del_"],
            rows[k]["zn"], args.nk1) for k, i in zip(scan_of, rig_i)]
    crow = A.scans_array(one)
    log(f"cooperative launch: {len(crow)} trajectories from their first step, step limit {args.step_limit}")
    pc = A.Plan(ms, params, crow, precision=A.PREC_MIXED, step_limit=args.step_limit, queue_order=1, coop=1,
                handover_steps=1)
    st_c, _ = timed_run(pc, "cooperative mode")
    tc, _ = pc.results()
    pc.close()

    def block(st):
        b = rate_block(st, "t96_igrf")
        b.update({"wave_iters": int(st.wave_iters), "lane_iters": int(st.lane_iters),
                  "lane_fill": st.lane_iters / max(64 * st.wave_iters, 1),
                  "coop_wavefront_steps": int(st.coop_steps), "grid_blocks": int(st.grid_blocks),
                  "block_threads": int(st.block_threads), "drain_ms": st.drain_ms,
                  "lane_end_ms": st.lane_end_ms})
        return b
    return {"workload": "benchmark grid split by tracer mode (IGRF-14 + T96, mixed)",
            "per_lane": block(st_l), "cooperative": block(st_c),
            "reached_handover": int(len(hit)),
            "cooperative_step_limit_ended": int((tc["steps"] >= args.step_limit).sum())}


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("storm")
    s.add_argument("--epochs", type=int, default=288)
    s.add_argument("--sites", type=int, default=50)
    s.add_argument("--cmp-epochs", type=int, default=8, help="epochs of the mixed-vs-FP64 comparison")
    s.add_argument("--cmp-step-limit", type=int, default=100000, help="step limit of the comparison runs")
    m = sub.add_parser("map")
    m.add_argument("--lat-step", type=float, default=1.0)
    m.add_argument("--fp64-frac", type=float, default=0.01)
    m.add_argument("--tune-sample", type=int, default=256)
    m.add_argument("--no-tune", action="store_true")
    m.add_argument("--cmp-step-limit", type=int, default=100000, help="step limit of the FP64 comparison runs")
    m.add_argument("--cmp-only", action="store_true", help="only the mixed-vs-FP64 comparison on the subsample")
    d = sub.add_parser("directions")
    d.add_argument("--sites", type=int, default=64)
    d.add_argument("--fp64-scans", type=int, default=96)
    d.add_argument("--handover-steps", type=int, default=0,
                   help="mixed tracer: per-lane steps before the cooperative mode (0: library default)")
    o = sub.add_parser("modes")
    for p in (s, m, d, o):
        p.add_argument("--seed", type=int, default=2020)
        p.add_argument("--nk1", type=int, default=400)
        p.add_argument("--step-limit", type=int, default=1000000)
        p.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    A.init(0)
    t0 = time.perf_counter()
    out = {"storm": run_storm, "map": run_map, "directions": run_directions, "modes": run_modes}[args.cmd](args)
    out["total_wall_s"] = time.perf_counter() - t0
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
