"""Measure the algorithmic work per accepted integration step with the op-counting
build of the oracle (opcount.cpp) and write profiles/flop_model.json.

Measurement infrastructure: run here, on the CPU; the benchmark only reads the JSON.
Sample: a seeded set of sites of the 5-degree vertical grid (WGS84 geodetic ->
geocentric at 20 km), 2000 rigidities each (0.01..20 GV), so >= 1e4 trajectories.

    python oracle/opcount/measure.py [--sites 8] [--nk1 400]
"""
import argparse
import json
import math
import os
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]


def geodetic_to_geocentric(lat, h_km):
    a, f = 6378.137, 1 / 298.257223563
    e2 = f * (2 - f)
    p = math.radians(lat)
    n = a / math.sqrt(1 - e2 * math.sin(p) ** 2)
    x = (n + h_km) * math.cos(p)
    z = (n * (1 - e2) + h_km) * math.sin(p)
    return math.hypot(x, z) / 6371.2, math.degrees(math.atan2(z, x))


def grid():
    out = []
    for i in range(37):
        lat = -90.0 + 5.0 * i
        r0, lc = geodetic_to_geocentric(lat, 20.0)
        for j in range(72):
            out.append((r0, lc, 5.0 * j))
    return out


def run(binary, sites, nk1, internal, external, parmod, sh=None):
    """Op counts of the oracle over the sites' scans; sh = (G, H) [nm][nm] adds that
    SH model as the crustal model of the set."""
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        for r0, lat, lon in sites:
            f.write(f"{r0!r} {lat!r} {lon!r}\n")
        name = f.name
    env = dict(os.environ)
    sh_name = None
    if sh is not None:
        import numpy as np
        g, h = (np.ascontiguousarray(a, dtype=np.float64) for a in sh)
        with tempfile.NamedTemporaryFile("wb", suffix=".bin", delete=False) as f:
            f.write(np.int32(g.shape[0]).tobytes() + g.tobytes() + h.tobytes())
            sh_name = f.name
        env["COR_OPCOUNT_SH"] = sh_name
    try:
        out = subprocess.run([str(binary), name, str(nk1), "1000000", str(internal), str(external)] +
                             [repr(float(v)) for v in parmod], check=True, capture_output=True, text=True, env=env)
    finally:
        os.unlink(name)
        if sh_name:
            os.unlink(sh_name)
    last = out.stdout.strip().splitlines()[-1].split()
    vals = dict(zip(last[1::2], last[2::2]))
    return {"trajectories": int(vals["traj"]), "steps": int(vals["steps"]), "attempts": int(vals["attempts"]),
            "ops": int(vals["ops"]), "flops_per_step": int(vals["ops"]) / int(vals["steps"]),
            "attempts_per_step": int(vals["attempts"]) / int(vals["steps"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sites", type=int, default=8)
    ap.add_argument("--nk1", type=int, default=400)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-ts05", action="store_true")
    args = ap.parse_args()
    binary = Path(tempfile.gettempdir()) / "cor_opcount"
    subprocess.run(["g++", "-std=c++17", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-math-errno",
                    f"-I{REPO / 'oracle'}", "-o", str(binary), str(HERE / "opcount.cpp"), "-lm"], check=True)
    import random
    rng = random.Random(args.seed)
    sites = rng.sample(grid(), args.sites)
    res = {
        "convention": "add/sub/mul/div/sqrt = 1, FMA = 2, exp/log/sin/cos/tan/atan/atan2/asin = 20, "
                      "pow = 40; counted on the oracle (the reference's algorithm) over whole "
                      "trajectories, divided by accepted steps",
        "sample": f"{args.sites} seeded sites (seed {args.seed}) of the 5-degree grid x 2000 rigidities, "
                  f"nk1 {args.nk1}, tu 0.01, step limit 1e6",
        "t96_igrf": run(binary, sites, args.nk1, 1, 1, [2, -20, 0, -2]),
        "igrf": run(binary, sites, args.nk1, 1, 0, []),
    }
    if not args.no_ts05:
        # TS05 at the high-resolution map's drivers (Pdyn 3, Dst -100, By 2, Bz -5,
        # W1..W6 0.5): with IGRF-14 (the storm series' model set, one representative
        # epoch) and with the map's degree-20 SH core field in place of IGRF
        import sys
        sys.path.insert(0, str(REPO / "cor-model_amd"))
        import run_workload
        ts05 = [3, -100, 2, -5] + [0.5] * 6
        res["ts05_igrf"] = run(binary, sites, args.nk1, 1, 2, ts05)
        res["ts05_sh20"] = run(binary, sites, args.nk1, 0, 2, ts05, sh=run_workload.igrf2020_sh_table(21, 2020))
        res["ts05_note"] = ("ts05_igrf and ts05_sh20: TS05 with Pdyn 3, Dst -100, By 2, Bz -5, W1..W6 0.5; "
                            "ts05_sh20 replaces IGRF by the map's degree-20 SH core field "
                            "(cor-model_amd/run_workload.py igrf2020_sh_table(21, 2020))")
    out = REPO / "profiles" / "flop_model.json"
    out.write_text(json.dumps(res, indent=2) + "\n")
    print(json.dumps(res, indent=2))


if __name__ == "__main__":
    main()
