"""Write the golden fixtures of the CPU parity tests from the reference, compiled from
its own sources by oracle/Makefile (oracle/_ref/libcorref.so and oracle/_ref/cor).

The reference ships no tests or fixtures of its own, so these vectors are the pins:
inputs and the reference's outputs, nothing else.  Run in the container that has the
reference sources (make -C oracle ref first):

    python tests/golden/make_golden.py          (everything)
    python tests/golden/make_golden.py ts05     (the TS05 fixtures only)
    python tests/golden/make_golden.py chaos    (the SH synthesis fixtures only)
    python tests/golden/make_golden.py field    (the field-map run only)
    python tests/golden/make_golden.py track    (the trajectory-mode runs only)
    python tests/golden/make_golden.py internal (every internal model's epoch setup only)
    python tests/golden/make_golden.py grid5    (five grid sites at the benchmark's resolution only)
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent / "cor-model_amd"))
import oracle_lib as L  # noqa: E402

EPOCHS = [(2020, 1, 1, 1, 0, 0, 0), (2020, 3, 15, 75, 12, 30, 5), (1965, 6, 1, 152, 3, 0, 0),
          (1900, 1, 1, 1, 0, 0, 0), (2027, 7, 1, 182, 0, 0, 0), (2029, 12, 31, 365, 23, 59, 59),
          (1903, 2, 19, 50, 6, 0, 0)]
T96_PARAMS = [[2, -20, 0, -2], [3, -100, 2, -5], [0.5, 0, 5, 5], [10, -300, -3, 4]]


def sh_points(rng, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = rng.uniform(1.0, 30.0, size=(n, 1))
    pts = d * r
    # poles, the z axis, and radii on both sides of the integer truncation radii
    extra = [[0, 0, 1.5], [0, 0, -2.5], [1e-7, 0, 3.0], [0, 0, 1.0], [1e-6, 1e-6, -7.0]]
    for rr in (2.0, 3.0, 4.0, 9.0, 14.0, 29.0):
        for eps in (-1e-9, 1e-9):
            extra.append([0.3 * (rr + eps), 0.4 * (rr + eps), np.sqrt(0.75) * (rr + eps)])
    return np.vstack([pts, np.array(extra)])


def t96_points(rng, n):
    p = np.column_stack([rng.uniform(-40, 15, n), rng.uniform(-20, 20, n), rng.uniform(-20, 20, n)])
    p = p[np.linalg.norm(p, axis=1) > 1.0]
    # along the noon-midnight meridian and the equator: Region-1 band edges, Region-2
    # xi bands, the magnetopause blend, the tail sheet
    x = np.linspace(-40, 12, 400)
    lines = [np.column_stack([x, np.zeros_like(x), np.full_like(x, z)]) for z in (0.3, 2.0, 5.0, -3.0)]
    lines += [np.column_stack([np.full_like(x, xx), np.linspace(-15, 15, 400), np.full_like(x, 1.0)])
              for xx in (-8.0, 0.5, 6.0)]
    th = np.linspace(0.01, np.pi - 0.01, 400)
    for r in (2.0, 3.5, 6.0):
        lines.append(np.column_stack([r * np.sin(th), np.zeros_like(th), r * np.cos(th)]))
        lines.append(np.column_stack([-r * np.sin(th), np.zeros_like(th), r * np.cos(th)]))
        lines.append(np.column_stack([r * np.sin(th) * 0.6, r * np.sin(th) * 0.8, r * np.cos(th)]))
    return np.vstack([p] + lines)


TS05_PARAMS = [[2, -20, 0, -2, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5], [3, -100, 2, -5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5],
               [8, -250, -4, 6, 1.5, 0.2, 1.9, 0.7, 1.1, 0.3], [0.6, 10, 1, 1, 0, 0, 0, 0, 0, 0]]


def ts05_fixtures():
    """TS05 field vectors and one TS05 cutoff run (lib/ts05/ts05.c, -e ts05)."""
    ref = L.RefLib()
    rng = np.random.default_rng(20240505)
    out = {"ts05_pts": t96_points(rng, 3000)}
    for k, pm in enumerate(TS05_PARAMS):
        out[f"ts05_params{k}"] = np.array(pm, dtype=float)
        out[f"ts05_b{k}"] = ref.ts05(pm, out["ts05_pts"])
    np.savez_compressed(HERE / "ts05_vectors.npz", **out)
    name = "ts05_storm_nk100"
    inp = HERE / f"{name}.in"
    # file order: Dst Pdyn By Bz, then W1..W6
    L.write_cutoff_input(inp, 1.0, -1.0, 20.0, 1.00280, 52.0, 30.0, 52.0, 30.0, nk1=100, dl=0.25,
                         ext_params=[-100, 3, 2, -5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5])
    txt = L.run_ref_cutoff(inp, internal="igrf14", external="ts05", steps=20000)
    (HERE / f"{name}.out").write_text(txt)
    print(name, txt.strip().splitlines()[-1])


def seeded_sh(seed, nm=21):
    """A degree nm-1 table with an IGRF-like spectrum (row n, column m, nT)."""
    rng = np.random.default_rng(seed)
    g = np.zeros((nm, nm))
    h = np.zeros((nm, nm))
    for n in range(1, nm):
        sc = 30000.0 * 0.3 ** n
        for m in range(n + 1):
            g[n, m] = sc * rng.normal()
            h[n, m] = sc * rng.normal() if m else 0.0
    return g, h


def chaos_fixtures():
    """chaos_geo (lib/chaos/chaos.c:120-227) with two seeded tables (degree 20 and 23,
    the largest the device takes) at SH test points: the reference's own outputs."""
    ref = L.RefLib()
    rng = np.random.default_rng(20240507)
    pts = sh_points(rng, 2000)
    out = {"chaos_pts": pts}
    for k, (seed, nm) in enumerate([(11, 21), (12, 24)]):
        g, h = seeded_sh(seed, nm)
        out[f"chaos_g{k}"] = g
        out[f"chaos_h{k}"] = h
        out[f"chaos_b{k}"] = ref.chaos(g, h, pts)
    # the shipped CHAOS-7.16 table (nm_max 186) at a subset of the points
    out["chaos716_pts"] = pts[::8]
    out["chaos716_b"] = ref.chaos716(out["chaos716_pts"])
    np.savez_compressed(HERE / "chaos_vectors.npz", **out)
    print("chaos vectors", {k: v.shape for k, v in out.items()})


# every internal model (lib/internal/internal.h:22-29) with epochs at and around its
# table's ends, inside an interpolation interval, in the extrapolation range, and out of
# range on both sides (the reference returns -1 there)
INTERNAL_EPOCHS = {
    "IGRF_14_COEFS": [1900, 1912, 2020, 2024, 2025, 2027, 2030, 2031, 1899],
    "IGRF_13_COEFS": [1900, 1987, 2015, 2019, 2020, 2022, 2025, 2026, 1899],
    "IGRF_12_COEFS": [1900, 1955, 2010, 2014, 2015, 2018, 2020, 2021, 1899],
    "IGRF_11_COEFS": [1900, 1944, 2005, 2009, 2010, 2013, 2015, 2016, 1899],
    "IGRF_10_COEFS": [1900, 1933, 2000, 2004, 2005, 2008, 2010, 2011, 1899],
    "IGRF_9_COEFS": [1900, 1922, 1995, 1999, 2000, 2003, 2005, 2006, 1899],
    "HIST_COEFS": [0, 1, 49, 50, 777, 1490, 1949, 1950, 1999, 2000, -1],
    "CALS10K_2_COEFS": [0, 1, 49, 50, 333, 1010, 1949, 1950, 1999, 2000, -1],
}
INTERNAL_DAYS = [(1, 1, 1, 0, 0, 0), (7, 2, 183, 12, 30, 15)]


def internal_fixtures():
    """internal_recalc for every internal model of the reference at the epochs above:
    return code, G, H, REC and the GEO->GSM rows; and two cutoff runs of the reference
    CLI with other internal models (IGRF-13 at 2017, CALS10k.2 at 1000 CE)."""
    for name, model, date, site in [("igrf13_2017_nk100", "igrf13", (2017, 5, 20, 140, 6, 0, 0),
                                     (1.00314, 40.0, 250.0)),
                                    ("cals10k2_1000_nk100", "cals10k.2", (1000, 1, 1, 1, 0, 0, 0),
                                     (1.00314, 30.0, 15.0))]:
        inp = HERE / f"{name}.in"
        r0, lat, lon = site
        L.write_cutoff_input(inp, 0.5, -1.0, 20.0, r0, lat, lon, lat, lon, date=date, nk1=100, dl=0.05)
        txt = L.run_ref_cutoff(inp, internal=model, steps=200000)
        (HERE / f"{name}.out").write_text(txt)
        print(name, txt.strip().splitlines()[-1])
    ref = L.RefLib()
    out = {}
    for tab, years in INTERNAL_EPOCHS.items():
        rows = []
        for y in years:
            for d in INTERNAL_DAYS:
                e = (y,) + d
                rc, p = ref.recalc(*e, table=tab)
                rows.append(e)
                k = len(rows) - 1
                out[f"{tab}_{k}_rc"] = np.array(rc)
                if rc == 0:
                    for f in ("G", "H", "REC", "A1", "A2", "A3"):
                        out[f"{tab}_{k}_{f}"] = np.array(getattr(p, f))
        out[f"{tab}_epochs"] = np.array(rows)
    np.savez_compressed(HERE / "internal_vectors.npz", **out)
    print("internal vectors", len(out), "arrays")


# five sites of the 5-degree grid at the benchmark's resolution (config 2: IGRF-14 2020
# + T96 quiet, 0.01..20 GV in 0.01 GV steps, nk1 400, 1e6 steps), geodetic latitude and
# longitude -> the reference's geocentric input (WGS84, 20 km)
GRID5_SITES = [(60.0, 0.0), (40.0, 120.0), (0.0, 90.0), (-30.0, 300.0), (-60.0, 150.0)]


def grid5_fixtures():
    """Reference runs of five grid sites at the benchmark's resolution."""
    import cor_amd as A
    for lat, lon in GRID5_SITES:
        r0, latc = A.geodetic_to_geocentric(lat, 20.0)
        name = f"grid5_{int(lat):+03d}_{int(lon):03d}"
        inp = HERE / f"{name}.in"
        L.write_cutoff_input(inp, 0.01, -1.0, 20.0, r0, latc, lon, latc, lon, nk1=400, dl=0.01,
                             ext_params=[-20, 2, 0, -2])
        txt = L.run_ref_cutoff(inp, internal="igrf14", external="t96", steps=1000000)
        (HERE / f"{name}.out").write_text(txt)
        print(name, txt.strip().splitlines()[-1])


def field_fixtures():
    """Field-map runs of the reference CLI (-m field, src/field.c:6-54), input and output
    text: IGRF-14 + T96 on a 20 x 30 degree grid at 1.5 Re, and IGRF-14 + the
    CHAOS-7.16 crustal model on a 10 x 15 degree patch at 1.01 Re."""
    import subprocess
    runs = {"field_igrf_t96": ("1.5 -80.0 0.0\n20.0 30.0 80.0 360.0\n2020 1 1 1 0 0 0\n-20 2 0 -2\n",
                               ["-e", "t96"]),
            "field_igrf_chaos": ("1.01 -30.0 10.0\n10.0 15.0 30.0 100.0\n2020 1 1 1 0 0 0\n",
                                 ["-c", "chaos7.16"])}
    for name, (text, extra) in runs.items():
        inp = HERE / f"{name}.in"
        inp.write_text(text)
        txt = subprocess.run([str(L.REF_BIN), "-m", "field", "-i", "igrf14"] + extra + [str(inp)],
                             capture_output=True, text=True, timeout=300, check=True).stdout
        (HERE / f"{name}.out").write_text(txt)
        print(name, len(txt.splitlines()), "lines")


TRACK_RUNS = {
    # name: (site r0, lat, lon), rigidity [GV], nk1, external model and its lines, step limit
    "track_t96_5gv": ((1.00260, 49.8074, 120.0), 5.0, 400, "t96", [-20, 2, 0, -2], 20000),
    "track_t96_trapped": ((1.00260, 49.8074, 120.0), 0.3, 100, "t96", [-20, 2, 0, -2], 3000),
    "track_ts05_3gv": ((1.00280, 52.0, 30.0), 3.0, 100, "ts05", [-100, 3, 2, -5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5],
                       20000),
}


def track_fixtures():
    """Trajectory-mode runs of the reference CLI (-m trajectory, src/trajectory.c:163-273):
    one rigidity each, the tragsm_* and trasph_* files it writes."""
    import subprocess
    import tempfile
    for name, (site, rig, nk1, ext, params, steps) in TRACK_RUNS.items():
        r0, lat, lon = site
        inp = HERE / f"{name}.in"
        L.write_cutoff_input(inp, rig, -1.0, rig, r0, lat, lon, lat, lon, nk1=nk1, dl=0.01, ext_params=params)
        with tempfile.TemporaryDirectory() as td:
            subprocess.run([str(L.REF_BIN), "-m", "trajectory", "-i", "igrf14", "-e", ext, "-s", str(steps),
                            "-o", name, str(inp)], cwd=td, check=True, timeout=300, capture_output=True)
            for kind in ("tragsm", "trasph"):
                txt = (Path(td) / f"{kind}_{name}").read_text()
                (HERE / f"{name}.{kind}").write_text(txt)
        print(name, len(txt.splitlines()), "lines")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "track":
        track_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "field":
        field_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "ts05":
        ts05_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "chaos":
        chaos_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "internal":
        internal_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "grid5":
        grid5_fixtures()
        return
    ref = L.RefLib()
    rng = np.random.default_rng(20240501)
    out = {}
    for k, e in enumerate(EPOCHS):
        rc, p = ref.recalc(*e)
        assert rc == 0
        for f in ("G", "H", "REC", "A1", "A2", "A3"):
            out[f"recalc{k}_{f}"] = np.array(getattr(p, f))
    out["recalc_epochs"] = np.array(EPOCHS)
    _, p2020 = ref.recalc(*EPOCHS[0])
    pts = sh_points(rng, 3000)
    out["igrf_pts"] = pts
    out["igrf_b"] = ref.igrf_geo(p2020, pts)
    tp = t96_points(rng, 4000)
    out["t96_pts"] = tp
    for k, pm in enumerate(T96_PARAMS):
        out[f"t96_params{k}"] = np.array(pm, dtype=float)
        out[f"t96_b{k}"] = ref.t96(pm, tp)
    np.savez_compressed(HERE / "field_vectors.npz", **out)

    # cutoff runs of the reference CLI
    runs = {
        "config1_igrf_nk100": dict(site=(1.002439629, 46.808601777, 8.0), nk1=100, rig=0.01, dl=0.01,
                                    rk=20.0, external=None, params=None),
        "t96_equator_nk400": dict(site=(1.0031392868, 0.0, 0.0), nk1=400, rig=0.05, dl=0.05, rk=20.0,
                                  external="t96", params=[-20, 2, 0, -2]),
        "t96_mid_nk400": dict(site=(1.00260, 49.8074, 120.0), nk1=400, rig=0.05, dl=0.05, rk=20.0,
                              external="t96", params=[-20, 2, 0, -2]),
        "t96_storm_limit_nk100": dict(site=(1.00280, -35.0, 290.0), nk1=100, rig=1.0, dl=0.25, rk=20.0,
                                      external="t96", params=[-150, 6, 4, -8], steps=20000),
    }
    for name, r in runs.items():
        r0, lat, lon = r["site"]
        inp = HERE / f"{name}.in"
        L.write_cutoff_input(inp, r["rig"], -1.0, r["rk"], r0, lat, lon, lat, lon, nk1=r["nk1"], dl=r["dl"],
                             ext_params=r["params"])
        txt = L.run_ref_cutoff(inp, internal="igrf14", external=r["external"], steps=r.get("steps", 1000000))
        (HERE / f"{name}.out").write_text(txt)
        print(name, txt.strip().splitlines()[-1])
    ts05_fixtures()
    chaos_fixtures()
    field_fixtures()
    track_fixtures()
    internal_fixtures()
    grid5_fixtures()


if __name__ == "__main__":
    main()
