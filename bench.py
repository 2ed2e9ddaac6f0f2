#!/usr/bin/env python3
"""Headline benchmark: trajectories/s of the mixed-precision persistent tracer on one
MI355X, against the plain FP64 GPU tracer, on the global 5-degree vertical-cutoff grid.

Workload (one "step" = one pass over the whole batch, inputs resident in HBM):
  IGRF-14 at 2020-01-01 00:00 plus T96 (Pdyn 2 nPa, Dst -20 nT, By 0, Bz -2 nT);
  sites at geodetic latitude -90..90 x longitude 0..355 in 5-degree steps (2664) at
  20 km (WGS84 -> geocentric), vertical incidence; protons back-traced at 0.01..20 GV
  in 0.01 GV steps (2000 per site, 5.328 M trajectories); tu = 0.01 rad, at most 1e6
  steps per trajectory, nk1 = 400 (the input's starting steps per gyro-period: with
  tu = 0.01 rad most steps are then accepted at the first or second attempt; the CPU
  baseline runs with the same value).

Printed: one JSON line (value = mixed-precision trajectories/s over the whole batch),
with the FP64 GPU figures on every 16th site (the mixed path re-run on the same sites
for a like-for-like time ratio; the cutoff agreement of the full-grid mixed run on
those sites), the roofline of the tracer kernel and the reference CPU path timed on
this host.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "cor-model_amd"))

import numpy as np  # noqa: E402

import cor_amd as A  # noqa: E402

PARMOD = [2.0, -20.0, 0.0, -2.0]        # Pdyn, Dst, By, Bz (reference PARMOD order)
EPOCH = (2020, 1, 1, 1, 0, 0, 0)
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}


def build_scans(stride, nk1, rig0=0.01, dl=0.01, rk=20.0):
    sites = A.grid_sites(5.0, 5.0, 20.0)[::stride]
    rows = [(r0, lat, lon, lat, lon, rig0, dl, rk, -1.0, nk1) for (r0, lat, lon) in sites]
    return A.scans_array(rows), sites


def flop_model():
    """Per-accepted-step FP operation counts measured with the op-counting build of the
    oracle (profiles/flop_model.json, counting rules there)."""
    f = REPO / "profiles" / "flop_model.json"
    if f.exists():
        return json.loads(f.read_text())
    return None


def time_plan(plan, steps, warmup):
    """Wall time per step over `steps` timed passes, and the tracer kernel's average
    duration over the same passes from the plan's HIP events (recorded on the stream
    the kernel is launched on, around the kernel alone)."""
    for _ in range(warmup):
        plan.run()
        plan.sync()
    A.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
        plan.sync()
        kernel_ms.append(plan.stats().kernel_ms)
    A.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / max(steps, 1), float(np.mean(kernel_ms)) if kernel_ms else 0.0


def traffic_per_launch(n_traj):
    """HBM bytes per tracer launch from the committed FETCH_SIZE / WRITE_SIZE counter
    runs (rocprofv3 --pmc), each scaled by the same counter's reading for a plain
    device-to-device copy of known size; stored per trajectory and scaled to this launch."""
    f = REPO / "profiles" / "traffic_trace_kernel_mixed_t96_grid5.json"
    if not f.exists():
        return None
    t = json.loads(f.read_text())
    return t["hbm_bytes_per_trajectory"] * n_traj


def copy_calibration(gib=1.0, reps=3):
    """A plain device-to-device copy of known size (hipMemcpy, the HIP runtime's copy
    kernel), run a few times so that a counter run sees its FETCH_SIZE / WRITE_SIZE
    next to the tracer's."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")          # the runtime libcor_amd.so links
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    n = int(gib * (1 << 30))
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), n) != 0:
            raise RuntimeError("hipMalloc failed")
    ok = hip.hipMemset(src, 1, n) == 0
    for _ in range(reps):
        ok = ok and hip.hipMemcpy(dst, src, n, 3) == 0       # hipMemcpyDeviceToDevice
    ok = ok and hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    if not ok:
        raise RuntimeError("copy calibration failed")
    print(f"[bench] copy calibration: {reps} x {n} bytes copied device to device",
          file=sys.stderr, flush=True)


def cpu_baseline(sites, budget_s, nk1, step_limit):
    """The reference itself (oracle/_ref/cor, built from its sources) on a bounded
    sample of the same sites: whole sites, -p (OpenMP over the rigidity scan) on this
    host's cores, until the time budget is spent.  Returns the rate and, per sampled
    site, the reference's (Rl, Ru, Rc)."""
    ref = REPO / "oracle" / "_ref" / "cor"
    if not ref.exists():
        return None, {}
    rng = np.random.default_rng(12345)
    order = rng.permutation(len(sites))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    done, n_traj, t_total, cuts = 0, 0, 0.0, {}
    with tempfile.TemporaryDirectory() as td:
        for k in order:
            r0, lat, lon = sites[k]
            inp = Path(td) / "in.txt"
            inp.write_text(f"0.01 -1 20.00\n{r0!r} {lat!r} {lon!r}\n{lat!r} {lon!r}\n"
                           f"2020 1 1 1 0 0 0\n{nk1} 0 0 0.01\n-20 2 0 -2\n")
            t0 = time.perf_counter()
            res = subprocess.run([str(ref), "-p", "-i", "igrf14", "-e", "t96", "-s", str(step_limit), str(inp)],
                                 check=True, capture_output=True, timeout=600, text=True)
            t_total += time.perf_counter() - t0
            lines = res.stdout.splitlines()
            for j, ln in enumerate(lines):
                if "CUTOFF with rigidities" in ln:
                    cuts[int(k)] = tuple(float(v) for v in lines[j + 1].split())
            done += 1
            n_traj += 2000
            if t_total >= budget_s and done >= 2:
                break
    return ({"value": n_traj / t_total, "unit": "trajectories/s", "cores": threads, "kind": "reference",
             "sample": f"{done} seeded sites of the benchmark grid x 2000 rigidities "
                       f"({n_traj} trajectories) in {t_total:.1f} s, oracle/_ref/cor -p"}, cuts)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stride", type=int, default=1, help="every stride-th site of the grid")
    ap.add_argument("--nk1", type=int, default=400)
    ap.add_argument("--step-limit", type=int, default=1000000)
    ap.add_argument("--queue-order", type=int, default=1)
    ap.add_argument("--coop", type=int, default=0,
                    help="cooperative mode (0 = on, the library default; -1 = per-lane mode only)")
    ap.add_argument("--handover-steps", type=int, default=0,
                    help="accepted per-lane steps before a trajectory continues in the cooperative "
                         "mode (0: library default)")
    ap.add_argument("--fp64-stride", type=int, default=32,
                    help="FP64 comparison run on every fp64-stride-th site of the benchmarked sites")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (for counter calibration)")
    ap.add_argument("--dump", default=None,
                    help="write the mixed run's per-trajectory status, steps and handover step of "
                         "every 4th site to this .npz file")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the tracer shards by "
              f"station across GPUs with no exchange inside the trace, so one GPU's share is "
              f"the measured unit", file=sys.stderr)
        sys.exit(2)

    A.init(0)
    if args.calibrate_copy:
        copy_calibration()
    scans, sites = build_scans(args.stride, args.nk1)
    params = A.recalc(*EPOCH, parmod=PARMOD)
    models = A.model_set(A.INTERNAL_IGRF, A.EXTERNAL_T96)
    mixed = A.Plan(models, params, scans, step_limit=args.step_limit, precision=A.PREC_MIXED,
                   queue_order=args.queue_order, coop=args.coop,
                   handover_steps=args.handover_steps)
    n = mixed.n
    print(f"[bench] {len(scans)} sites, {n} trajectories; warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    t_mixed, k_mixed = time_plan(mixed, args.steps, args.warmup)
    st_m = mixed.stats()
    tr_m, cut_m = mixed.results(traj=args.dump is not None)
    mixed.close()
    if args.dump:
        per = n // len(scans)
        keep = (np.arange(n) // per) % 4 == 0
        np.savez_compressed(args.dump, status=tr_m["status"][keep].astype(np.int8), steps=tr_m["steps"][keep],
                            handover=tr_m["handover"][keep], per_site=per)
        del tr_m
    print(f"[bench] mixed: {t_mixed*1e3:.1f} ms/step, kernel {k_mixed:.1f} ms (HIP events, mean of "
          f"{args.steps}), {st_m.steps} steps, {st_m.attempts} attempts; lane fill "
          f"{st_m.lane_iters / max(64 * st_m.wave_iters, 1):.3f}, queue drained at {st_m.drain_ms:.0f} ms, "
          f"per-lane mode ended at {st_m.lane_end_ms:.0f} ms, {st_m.coop_steps} cooperative steps",
          file=sys.stderr, flush=True)

    fp64 = None
    cut_d = None
    sub = np.arange(0, len(scans), max(args.fp64_stride, 1))
    if not args.no_fp64:
        p64 = A.Plan(models, params, scans[sub], step_limit=args.step_limit, precision=A.PREC_FP64,
                     queue_order=args.queue_order)
        print(f"[bench] fp64 comparison on {len(sub)} sites", file=sys.stderr, flush=True)
        t64, _ = time_plan(p64, 1, 0)
        st_d = p64.stats()
        _, cut_d = p64.results(traj=False)
        p64.close()
        n64 = len(sub) * (n // len(scans))
        # the mixed path on the same sites, for a like-for-like speed ratio (both runs
        # of a site subset end with the same slowest trajectories)
        msub = A.Plan(models, params, scans[sub], step_limit=args.step_limit, precision=A.PREC_MIXED,
                      queue_order=args.queue_order, coop=args.coop,
                      handover_steps=args.handover_steps)
        t_msub, _ = time_plan(msub, 1, 0)
        msub.close()
        drc = np.abs(cut_m["rc"][sub] - cut_d["rc"])
        fp64 = {"sites": int(len(sub)), "trajectories": int(n64), "value": n64 / t64,
                "ms_per_step": t64 * 1e3, "kernel_ms": st_d.kernel_ms,
                "steps": int(st_d.steps), "attempts": int(st_d.attempts),
                "mixed_ms_same_sites": t_msub * 1e3,
                "speedup_mixed_vs_fp64_same_sites": t64 / t_msub,
                "sites_within_0.01GV": float(np.mean(drc <= 0.01 + 1e-9)),
                "max_abs_dRc_GV": float(drc.max()),
                "mean_abs_dRl_GV": float(np.mean(np.abs(cut_m["rl"][sub] - cut_d["rl"]))),
                "mean_abs_dRu_GV": float(np.mean(np.abs(cut_m["ru"][sub] - cut_d["ru"])))}
        print(f"[bench] fp64: {t64*1e3:.1f} ms/step, kernel {st_d.kernel_ms:.1f} ms (mixed on the same "
              f"sites {t_msub*1e3:.1f} ms); sites within 0.01 GV: {fp64['sites_within_0.01GV']:.4f}",
              file=sys.stderr, flush=True)

    roof = None
    fm = flop_model()
    if fm is not None:
        # algorithmic flops of the launch (accepted steps x the counted flops of one
        # accepted step of the reference's algorithm) over the kernel's mean duration
        f_step = fm["t96_igrf"]["flops_per_step"]
        flops_launch = f_step * st_m.steps
        achieved = flops_launch / (k_mixed * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS["fp32"], "traffic": traffic_per_launch(n),
                "kernel": "trace_kernel (mixed, IGRF+T96)", "kernel_ms": k_mixed,
                "flops_per_step": f_step, "flops_per_launch": flops_launch,
                "steps_per_launch": int(st_m.steps)}

    cpu, ref_cuts = (None, {}) if args.no_cpu else cpu_baseline(sites, args.cpu_budget, args.nk1, args.step_limit)
    if cpu is not None and ref_cuts:
        # how far two FP64 implementations (the reference on the CPU, this tracer's FP64
        # path on the GPU) already disagree on these sites -- the chaotic-penumbra floor
        # for any cutoff comparison -- and where the mixed path sits against the reference
        ks = sorted(ref_cuts)
        rc_ref = np.array([ref_cuts[k][2] for k in ks])
        cpu["sites_compared"] = len(ks)
        cpu["mixed_vs_reference_within_0.01GV"] = float(np.mean(np.abs(cut_m["rc"][ks] - rc_ref) <= 0.01 + 1e-9))
    out = {
        "metric": "trajectories/s on 1 MI355X, mixed vs FP64 GPU, at |dRc|<=0.01 GV vs FP64",
        "value": n / t_mixed,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 field+RK increments, f64 state",
        "data": "synthetic",
        "config": {"workload": "global 5-degree vertical cutoff grid, IGRF-14 2020.0 + T96 "
                               "(Pdyn 2, Dst -20, By 0, Bz -2), 0.01-20 GV step 0.01",
                   "sites": len(scans), "site_stride": args.stride, "trajectories": int(n),
                   "nk1": args.nk1, "step_limit": args.step_limit, "tu_angle": 0.01,
                   "queue_order": args.queue_order, "coop": args.coop,
                   "handover_steps": args.handover_steps},
        "kernel_ms": k_mixed,
        "integration_steps_per_s": st_m.steps / (k_mixed * 1e-3),
        # where the tracer's time goes: mean live lanes per per-lane wavefront step,
        # the time until the queue's last trajectory was started and until the last
        # wavefront left the per-lane mode, and the cooperative-mode steps
        "phases": {"lane_fill": st_m.lane_iters / max(64 * st_m.wave_iters, 1),
                   "wave_steps": int(st_m.wave_iters),
                   "lane_fill_queue_open": st_m.lane_iters_open / max(64 * st_m.wave_iters_open, 1),
                   "wave_steps_queue_open": int(st_m.wave_iters_open),
                   "lane_steps_queue_open": int(st_m.lane_iters_open), "drain_ms": st_m.drain_ms,
                   "lane_end_ms": st_m.lane_end_ms, "coop_steps": int(st_m.coop_steps)},
        "fp64": fp64,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
