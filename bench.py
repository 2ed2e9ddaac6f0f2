#!/usr/bin/env python3
"""Headline benchmark: trajectories/s of the mixed-precision persistent tracer on one
MI355X, against the FP64 GPU tracer, on the global 5-degree vertical-cutoff grid.

Workload (one "step" = one pass over the whole batch, inputs resident in HBM):
  IGRF-14 at 2020-01-01 00:00 plus T96 (Pdyn 2 nPa, Dst -20 nT, By 0, Bz -2 nT);
  sites at geodetic latitude -90..90 x longitude 0..355 in 5-degree steps (2664) at
  20 km (WGS84 -> geocentric), vertical incidence; protons back-traced at 0.01..20 GV
  in 0.01 GV steps (2000 per site, 5.328 M trajectories); tu = 0.01 rad, at most 1e6
  steps per trajectory, nk1 = 400 (the input's starting steps per gyro-period: with
  tu = 0.01 rad most steps are then accepted at the first or second attempt; the CPU
  baseline runs with the same value).

Printed: one JSON line.  value = mixed-precision trajectories/s over the whole batch.
Beside it: the FP64 GPU tracer on the same whole grid (with its cooperative tail), the
roofline of the mixed tracer kernel, the reference CPU path (oracle/_ref/cor, built
from the reference's sources) timed on this host on seeded sites of the grid, run in
the background while the GPU legs run, and an accuracy block that puts FP64-GPU vs
the reference, mixed vs FP64-GPU and mixed vs the reference side by side on those
sites (and mixed vs FP64-GPU on the whole grid).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import threading
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "cor-model_amd"))

import numpy as np  # noqa: E402

import cor_amd as A  # noqa: E402

PARMOD = [2.0, -20.0, 0.0, -2.0]        # Pdyn, Dst, By, Bz (reference PARMOD order)
EPOCH = (2020, 1, 1, 1, 0, 0, 0)
RIG0, DEL, RK = 0.01, 0.01, 20.0
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}
METRIC = "trajectories/s on 1 MI355X, mixed vs FP64 GPU, at |ΔRc|≤0.01 GV vs FP64"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_scans(stride, nk1):
    sites = A.grid_sites(5.0, 5.0, 20.0)[::stride]
    rows = [(r0, lat, lon, lat, lon, RIG0, DEL, RK, -1.0, nk1) for (r0, lat, lon) in sites]
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
    for k in range(warmup):
        plan.run()
        plan.sync()
        log(f"warmup pass {k + 1}/{warmup}: {plan.stats().kernel_ms:.0f} ms")
    A.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for k in range(steps):
        plan.run()
        plan.sync()
        kernel_ms.append(plan.stats().kernel_ms)
        log(f"timed pass {k + 1}/{steps}: {kernel_ms[-1]:.0f} ms")
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
    log(f"copy calibration: {reps} x {n} bytes copied device to device")


def cpu_model():
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def parse_reference(text):
    """The reference's cutoff text: per allowed rigidity index (rig = 0.01 + 0.01 i) its
    asymptotic latitude and longitude, and (Rl, Ru, Rc)."""
    allowed, cut = {}, None
    lines = text.splitlines()
    for j, ln in enumerate(lines):
        if "CUTOFF with rigidities" in ln:
            cut = tuple(float(v) for v in lines[j + 1].split())
            break
        parts = ln.split()
        if len(parts) == 9:
            try:
                v = [float(x) for x in parts]
            except ValueError:
                continue
            allowed[int(round((v[0] - RIG0) / DEL))] = (v[5], v[6])
    return allowed, cut


class CpuBaseline(threading.Thread):
    """The reference itself (oracle/_ref/cor, built from its sources) on n_sites seeded
    sites of the grid, whole sites one after another, -p (OpenMP over the rigidity scan)
    on all but one of this host's cores (the GPU legs run beside it), collecting each
    site's text output for the accuracy comparison."""

    def __init__(self, sites, n_sites, nk1, step_limit):
        super().__init__(daemon=True)
        self.sites, self.n_sites, self.nk1, self.step_limit = sites, n_sites, nk1, step_limit
        self.threads = max(1, (os.cpu_count() or 2) - 1)
        self.results, self.t_total, self.error = {}, 0.0, None

    def run(self):
        ref = REPO / "oracle" / "_ref" / "cor"
        rng = np.random.default_rng(12345)
        order = rng.permutation(len(self.sites))[:self.n_sites]
        env = dict(os.environ, OMP_NUM_THREADS=str(self.threads))
        try:
            with tempfile.TemporaryDirectory() as td:
                for k in order:
                    r0, lat, lon = self.sites[k]
                    inp = Path(td) / "in.txt"
                    inp.write_text(f"{RIG0} -1 {RK:.2f}\n{r0!r} {lat!r} {lon!r}\n{lat!r} {lon!r}\n"
                                   f"2020 1 1 1 0 0 0\n{self.nk1} 0 0 {DEL}\n-20 2 0 -2\n")
                    t0 = time.perf_counter()
                    res = subprocess.run([str(ref), "-p", "-i", "igrf14", "-e", "t96", "-s", str(self.step_limit),
                                          str(inp)], check=True, capture_output=True, timeout=900, text=True, env=env)
                    self.t_total += time.perf_counter() - t0
                    self.results[int(k)] = parse_reference(res.stdout)
        except Exception as e:   # reported in the output, never fatal to the GPU legs
            self.error = repr(e)


def angle_deg(lat1, lon1, lat2, lon2):
    """Great-circle angle between two directions given as latitude/longitude [deg]."""
    p1, p2 = np.radians(lat1), np.radians(lat2)
    dl = np.radians(lon2 - lon1)
    c = np.sin(p1) * np.sin(p2) + np.cos(p1) * np.cos(p2) * np.cos(dl)
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


def compare_sets(a, b, sites, per):
    """Two results on the same sites.  a and b map site -> (allowed flags [per], ast [per],
    asf [per], (Rl, Ru, Rc)).  Returns the site fraction with |dRc| <= 0.01 GV, the
    histogram of |dRc| in rigidity steps (|d nza|, Rc = rig + nza del), the per-trajectory
    flag mismatch rate away from the union penumbra band, and the asymptotic-direction
    angle between trajectories allowed in both away from the band."""
    d_steps, mism, outside, angles = [], 0, 0, []
    rig = RIG0 + DEL * np.arange(per)
    for k in sites:
        fa, ta, pa, ca = a[k]
        fb, tb, pb, cb = b[k]
        d_steps.append(int(round(abs(ca[2] - cb[2]) / DEL)))
        lo, hi = min(ca[0], cb[0]), max(ca[1], cb[1])
        out = (rig < lo - 1e-9) | (rig > hi + 1e-9)
        outside += int(out.sum())
        mism += int((fa[out] != fb[out]).sum())
        both = out & fa & fb
        if both.any():
            angles.append(angle_deg(ta[both], pa[both], tb[both], pb[both]))
    d = np.array(d_steps)
    ang = np.concatenate(angles) if angles else np.zeros(0)
    hist = {str(v): int((d == v).sum()) for v in range(5)}
    hist["5+"] = int((d >= 5).sum())
    return {"sites": int(len(d)), "sites_within_0.01GV": float(np.mean(d <= 1)), "abs_dnza_hist": hist,
            "max_abs_dRc_GV": float(d.max() * DEL) if len(d) else 0.0,
            "flag_mismatch_rate_outside_band": mism / max(outside, 1), "trajectories_outside_band": outside,
            "max_asymptotic_angle_deg": float(ang.max()) if len(ang) else 0.0,
            "p99_asymptotic_angle_deg": float(np.quantile(ang, 0.99)) if len(ang) else 0.0}


def gpu_sets(tr, cut, per, sites):
    out = {}
    for k in sites:
        t = tr[k * per:(k + 1) * per]
        out[k] = (t["status"] == 0, t["ast"], t["asf"], (cut[k]["rl"], cut[k]["ru"], cut[k]["rc"]))
    return out


def ref_sets(results, per):
    out = {}
    for k, (allowed, cut) in results.items():
        f = np.zeros(per, bool)
        ast = np.zeros(per)
        asf = np.zeros(per)
        for i, (a_t, a_f) in allowed.items():
            if 0 <= i < per:
                f[i] = True
                ast[i] = a_t
                asf[i] = a_f
        out[k] = (f, ast, asf, cut)
    return out


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
                    help="mixed tracer's cooperative mode (0 = on, the library default; -1 = per-lane mode only)")
    ap.add_argument("--handover-steps", type=int, default=0,
                    help="accepted per-lane steps before a trajectory continues in the cooperative "
                         "mode (0: library default)")
    ap.add_argument("--handover-low-steps", type=int, default=0,
                    help="per-lane steps before the handover for the lowest rigidities of each scan "
                         "(0: library default)")
    ap.add_argument("--handover-low-count", type=int, default=0,
                    help="how many of each scan's lowest rigidities --handover-low-steps covers (0: library "
                         "default, negative: none)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-sites", type=int, default=20, help="seeded grid sites the reference traces")
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
    per = int(round((RK - RIG0) / DEL)) + 1
    cpu = None
    if not args.no_cpu and (REPO / "oracle" / "_ref" / "cor").exists():
        cpu = CpuBaseline(sites, args.cpu_sites, args.nk1, args.step_limit)
        cpu.start()
        log(f"reference CPU leg started in the background: {args.cpu_sites} sites, {cpu.threads} threads")
    params = A.recalc(*EPOCH, parmod=PARMOD)
    models = A.model_set(A.INTERNAL_IGRF, A.EXTERNAL_T96)
    mixed = A.Plan(models, params, scans, step_limit=args.step_limit, precision=A.PREC_MIXED,
                   queue_order=args.queue_order, coop=args.coop, handover_steps=args.handover_steps,
                   handover_low_steps=args.handover_low_steps, handover_low_count=args.handover_low_count)
    n = mixed.n
    log(f"{len(scans)} sites, {n} trajectories; warmup {args.warmup}, steps {args.steps}")
    t_mixed, k_mixed = time_plan(mixed, args.steps, args.warmup)
    st_m = mixed.stats()
    tr_m, cut_m = mixed.results(traj=True)
    mixed.close()
    if args.dump:
        keep = (np.arange(n) // per) % 4 == 0
        np.savez_compressed(args.dump, status=tr_m["status"][keep].astype(np.int8), steps=tr_m["steps"][keep],
                            handover=tr_m["handover"][keep], per_site=per)
    log(f"mixed: {t_mixed*1e3:.1f} ms/step, kernel {k_mixed:.1f} ms (HIP events, mean of {args.steps}), "
        f"{st_m.steps} steps, {st_m.attempts} attempts; lane fill "
        f"{st_m.lane_iters / max(64 * st_m.wave_iters, 1):.3f}, queue drained at {st_m.drain_ms:.0f} ms, "
        f"per-lane mode ended at {st_m.lane_end_ms:.0f} ms, {st_m.coop_steps} cooperative steps")

    fp64 = None
    tr_d = cut_d = None
    if not args.no_fp64:
        # the FP64 tracer on the same whole grid, with its cooperative tail (the FP64
        # cooperative evaluator sums the same terms in another order than the per-lane
        # FP64 path, which keeps the reference's; both are FP64 throughout)
        p64 = A.Plan(models, params, scans, step_limit=args.step_limit, precision=A.PREC_FP64,
                     queue_order=args.queue_order, coop=1, handover_steps=args.handover_steps,
                     handover_low_steps=args.handover_low_steps, handover_low_count=args.handover_low_count)
        log(f"fp64 leg: the whole grid, {n} trajectories")
        t64, k64 = time_plan(p64, 1, 0)
        st_d = p64.stats()
        tr_d, cut_d = p64.results(traj=True)
        p64.close()
        fp64 = {"value": n / t64, "unit": "trajectories/s", "ms_per_step": t64 * 1e3, "kernel_ms": k64,
                "steps": int(st_d.steps), "attempts": int(st_d.attempts), "coop_steps": int(st_d.coop_steps),
                "speedup_mixed_vs_fp64": t64 / t_mixed, "target_speedup": 3.0,
                "speedup_target_met": bool(t64 / t_mixed >= 3.0)}
        log(f"fp64: {t64*1e3:.1f} ms/step; mixed is {t64 / t_mixed:.2f}x faster")

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

    cpu_out = None
    accuracy = {}
    all_sites = list(range(len(scans)))
    gm = gpu_sets(tr_m, cut_m, per, all_sites)
    gd = gpu_sets(tr_d, cut_d, per, all_sites) if tr_d is not None else None
    if gd is not None:
        accuracy["mixed_vs_fp64_gpu_all_sites"] = compare_sets(gm, gd, all_sites, per)
    if cpu is not None:
        cpu.join()
        if cpu.error or not cpu.results:
            cpu_out = {"error": cpu.error or "no site finished"}
        else:
            n_cpu = per * len(cpu.results)
            cpu_out = {"value": n_cpu / cpu.t_total, "unit": "trajectories/s", "cores": cpu.threads,
                       "kind": "reference", "cpu_model": cpu_model(),
                       "sample": f"{len(cpu.results)} seeded sites of the benchmark grid x {per} rigidities "
                                 f"({n_cpu} trajectories) in {cpu.t_total:.1f} s, oracle/_ref/cor -p with "
                                 f"{cpu.threads} OpenMP threads, beside the GPU legs"}
            rs = ref_sets(cpu.results, per)
            shared = sorted(rs)
            if gd is not None:
                accuracy["fp64_gpu_vs_reference"] = compare_sets(gd, rs, shared, per)
                accuracy["mixed_vs_fp64_gpu"] = compare_sets(gm, gd, shared, per)
            accuracy["mixed_vs_reference"] = compare_sets(gm, rs, shared, per)
            accuracy["shared_sites"] = shared
    if accuracy:
        ref_key = "mixed_vs_fp64_gpu_all_sites" if "mixed_vs_fp64_gpu_all_sites" in accuracy else None
        frac = accuracy[ref_key]["sites_within_0.01GV"] if ref_key else None
        accuracy["qualifier"] = "|dRc| <= 0.01 GV mixed vs FP64 GPU for >= 99.9 % of sites"
        accuracy["qualifier_met"] = bool(frac is not None and frac >= 0.999)
        accuracy["note"] = ("Rc = rig + nza del moves by one rigidity step per flipped penumbra trajectory; "
                            "the FP64-GPU vs reference block is how far two FP64 implementations "
                            "(OCML vs glibc transcendentals) already disagree on the same sites")
    out = {
        "metric": METRIC,
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
                   "handover_steps": args.handover_steps, "handover_low_steps": args.handover_low_steps,
                   "handover_low_count": args.handover_low_count},
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
        "accuracy": accuracy or None,
        "roofline": roof,
        "cpu_baseline": cpu_out,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
