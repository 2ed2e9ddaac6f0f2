"""The reference's own C API on the GPU, driven as the reference's command line drives
it (src/main.c:12-93 for cutoffs, :95-141 for field maps): open the input file, read
the header, find the models in the registries (get_model), let each model's setup_fun
read its lines, write the output header, then trajectory_simulation_cutoff or
field_simulation into a FILE -- and the text is compared with the reference's own
output of the same input (tests/golden/*.out, written by the reference built from its
sources).

Bar, cutoffs: the header is identical; outside the union penumbra band the allowed
rigidities are identical; from 1 GV above the band every printed column agrees (v/c to
1e-9, exit radius to 1e-4 Re, angles to 0.01 degree, flight time and path length to
1e-5 relative); the summary's Rl, Ru, Rc within the test's rigidity tolerance.  Field
map: header identical, every value within 0.01 nT of the printed reference value.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cor_amd as A
import oracle_lib as L

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parents[1]
GOLD = Path(__file__).resolve().parent / "golden"
libc = C.CDLL(None)
libc.fopen.restype = C.c_void_p
libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
libc.fclose.argtypes = [C.c_void_p]

SETUP = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)


class InfileHeader(C.Structure):
    _fields_ = [("r0", C.c_double), ("the0", C.c_double), ("fi0", C.c_double), ("the1", C.c_double),
                ("fi1", C.c_double), ("iy", C.c_int), ("mes", C.c_int), ("ide", C.c_int), ("id", C.c_int),
                ("ih", C.c_int), ("min", C.c_int), ("is_", C.c_int), ("zn", C.c_double), ("nk1", C.c_int),
                ("iopt", C.c_int), ("ist", C.c_int), ("rig", C.c_double), ("del_", C.c_double),
                ("rk", C.c_double)]


class Model(C.Structure):
    _fields_ = [("name", C.c_char_p), ("version", C.c_char_p), ("setup_data", C.c_void_p),
                ("setup_fun", SETUP), ("calc_fun", C.c_void_p)]


def lib():
    L_ = A.lib()
    L_.get_model.restype = C.c_void_p
    L_.get_model.argtypes = [C.c_void_p, C.c_char_p]
    L_.trajectory_simulation_cutoff.restype = None
    L_.trajectory_simulation_cutoff.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                                C.c_double, C.c_double, C.c_uint32]
    L_.field_simulation.restype = None
    L_.field_simulation.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
    return L_


def registry_models(L_, internal, external, crustal=None):
    """NULL-terminated model list as src/main.c builds it from -i / -e / -c."""
    models = (C.c_void_p * 4)()
    k = 0
    for reg, ver in (("internal_models", internal), ("external_models", external), ("crustal_models", crustal)):
        if ver:
            m = L_.get_model(C.addressof(C.c_void_p.in_dll(L_, reg)), ver.encode())
            assert m, ver
            models[k] = m
            k += 1
    return models


def setup_models(models, f, header, params):
    """src/main.c:51-62: every model's setup_fun reads its lines of the input file
    (models without one, like the crustal model, are skipped)."""
    for m in models:
        if not m:
            break
        md = Model.from_address(m)
        if not C.cast(md.setup_fun, C.c_void_p).value:
            continue
        assert md.setup_fun(f, C.addressof(header), C.addressof(params), md.setup_data) == 0


def run_cutoff(L_, tmp_path, name, internal, external, steps):
    inp = GOLD / f"{name}.in"
    h = InfileHeader()
    p = A.ModelsParams()
    models = registry_models(L_, internal, external)
    f = libc.fopen(str(inp).encode(), b"r")
    assert f
    try:
        assert L_.read_cutoff_infile_header(C.c_void_p(f), C.byref(h)) == 0
        setup_models(models, f, h, p)
    finally:
        libc.fclose(f)
    out = tmp_path / f"{name}.out"
    g = libc.fopen(str(out).encode(), b"w")
    try:
        assert L_.write_cutoff_outfile_header(C.c_void_p(g), C.byref(h), models, C.c_uint64(steps)) == 0
        L_.trajectory_simulation_cutoff(C.byref(h), C.cast(models, C.c_void_p), C.byref(p), g, steps, 0.01,
                                        A.M_PROTON, 1)
    finally:
        libc.fclose(g)
    return out.read_text(), h


def header_part(text):
    lines = text.splitlines()
    for k, ln in enumerate(lines):
        if len(ln.split()) == 9:
            try:
                [float(v) for v in ln.split()]
                return lines[:k]
            except ValueError:
                pass
    return lines


CUTOFF_RUNS = [
    ("config1_igrf_nk100", "igrf14", None, 1000000, 0.01),
    ("t96_equator_nk400", "igrf14", "t96", 1000000, 0.05),
    ("t96_storm_limit_nk100", "igrf14", "t96", 20000, 0.25),
    ("ts05_storm_nk100", "igrf14", "ts05", 20000, 0.25),
]


def compare_cutoff_text(name, mine, ref, rig, dl, rc_tol):
    """The reference's cutoff text against this build's: identical headers, cutoffs
    within rc_tol, the same allowed rigidities outside the union penumbra band, and the
    per-rigidity columns of the allowed lines above the band."""
    assert header_part(mine) == header_part(ref)
    got, cut_g = L.parse_cutoff_output(mine)
    want, cut_r = L.parse_cutoff_output(ref)
    print(name, "gpu", cut_g, "reference", cut_r)
    assert all(abs(a - b) <= rc_tol + 1e-9 for a, b in zip(cut_g, cut_r)), (cut_g, cut_r)
    lo, hi = min(cut_g[0], cut_r[0]), max(cut_g[1], cut_r[1])
    key = lambda r: int(round((r - rig) / dl))        # noqa: E731
    g_out = {key(r["rig"]) for r in got if r["rig"] < lo - 1e-9 or r["rig"] > hi + 1e-9}
    r_out = {key(r["rig"]) for r in want if r["rig"] < lo - 1e-9 or r["rig"] > hi + 1e-9}
    assert g_out == r_out, sorted(g_out ^ r_out)[:10]
    gi = {key(r["rig"]): r for r in got}
    n = 0
    for r in want:
        if r["rig"] <= hi + 1.0:
            continue
        g = gi[key(r["rig"])]
        assert abs(g["vc"] - r["vc"]) <= 1e-9
        assert abs(g["r"] - r["r"]) <= 1e-4
        for c in ("td", "fei", "ast", "asf"):
            d = abs(g[c] - r[c])
            assert min(d, 360.0 - d) <= 0.01, (r["rig"], c, g[c], r[c])
        for c in ("time", "length_km"):
            assert abs(g[c] - r[c]) <= 1e-5 * abs(r[c]) + 0.011, (r["rig"], c, g[c], r[c])
        n += 1
    print(name, n, "lines compared column by column")


@pytest.mark.parametrize("name,internal,external,steps,rc_tol", CUTOFF_RUNS, ids=[r[0] for r in CUTOFF_RUNS])
def test_trajectory_simulation_cutoff_text(tmp_path, name, internal, external, steps, rc_tol):
    A.init(0)
    L_ = lib()
    mine, h = run_cutoff(L_, tmp_path, name, internal, external, steps)
    compare_cutoff_text(name, mine, (GOLD / f"{name}.out").read_text(), h.rig, h.del_, rc_tol)


RELINKED = REPO / "oracle" / "_ref" / "cor_relinked"


@pytest.mark.skipif(not RELINKED.exists(), reason="oracle/_ref/cor_relinked is not built (make -C oracle relink)")
@pytest.mark.parametrize("name,internal,external,steps,rc_tol", CUTOFF_RUNS[:2], ids=[r[0] for r in CUTOFF_RUNS[:2]])
def test_reference_cli_relinked_on_this_library(name, internal, external, steps, rc_tol):
    """The reference's own command line (src/main.c, src/args.c, compiled from its
    sources and linked against libcor_amd.so instead of its own library) on the
    golden inputs, as the golden outputs were made: -p -i igrf14 [-e t96] -s STEPS."""
    cmd = [str(RELINKED), "-p", "-i", internal] + (["-e", external] if external else []) + ["-s", str(steps),
                                                                                          str(GOLD / f"{name}.in")]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    h = L.read_cutoff_header(GOLD / f"{name}.in")
    compare_cutoff_text(name, res.stdout, (GOLD / f"{name}.out").read_text(), h["rig"], h["del"], rc_tol)


FIELD_RUNS = [("field_igrf_t96", "t96", None, 9 * 12), ("field_igrf_chaos", None, "chaos7.16", 7 * 6)]


@pytest.mark.parametrize("name,external,crustal,n_points", FIELD_RUNS, ids=[r[0] for r in FIELD_RUNS])
def test_field_simulation_text(tmp_path, name, external, crustal, n_points):
    A.init(0)
    L_ = lib()
    inp = GOLD / f"{name}.in"
    h = InfileHeader()
    p = A.ModelsParams()
    models = registry_models(L_, "igrf14", external, crustal)
    f = libc.fopen(str(inp).encode(), b"r")
    try:
        assert L_.read_field_infile_header(C.c_void_p(f), C.byref(h)) == 0
        setup_models(models, f, h, p)
    finally:
        libc.fclose(f)
    out = tmp_path / "field.out"
    g = libc.fopen(str(out).encode(), b"w")
    try:
        assert L_.write_field_outfile_header(C.c_void_p(g), C.byref(h), models) == 0
        L_.field_simulation(C.byref(h), C.byref(p), C.cast(models, C.c_void_p), g, 1)
    finally:
        libc.fclose(g)
    mine = out.read_text().splitlines()
    ref = (GOLD / f"{name}.out").read_text().splitlines()
    assert len(mine) == len(ref)
    nums = 0
    for a, b in zip(mine, ref):
        fa, fb = a.split(), b.split()
        if not fa or not fb:
            assert a.strip() == b.strip()
            continue
        try:
            va, vb = np.array(fa, dtype=float), np.array(fb, dtype=float)
        except ValueError:
            assert a == b
            continue
        assert va.shape == vb.shape and np.all(np.abs(va - vb) <= 0.0100001), (a, b)
        nums += 1
    assert nums == n_points


TRACK_RUNS = [("track_t96_5gv", "t96", 20000), ("track_t96_trapped", "t96", 3000), ("track_ts05_3gv", "ts05", 20000)]


@pytest.mark.parametrize("name,external,steps", TRACK_RUNS, ids=[r[0] for r in TRACK_RUNS])
def test_trajectory_simulation_tracks(tmp_path, name, external, steps):
    """The trajectory mode (src/main.c:143-232, src/trajectory.c:163-273): one
    trajectory's track in GSM and spherical coordinates, line for line against the
    reference's files.  Same number of lines and step numbers; the values within 1e-9
    of their scale over the first half of the track and 1e-6 over all of it (last-bit
    differences of the transcendentals grow along a path; the trapped track runs to the
    step limit)."""
    A.init(0)
    L_ = lib()
    L_.trajectory_simulation.restype = None
    L_.trajectory_simulation.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                         C.c_double, C.c_double]
    inp = GOLD / f"{name}.in"
    h = InfileHeader()
    p = A.ModelsParams()
    models = registry_models(L_, "igrf14", external)
    f = libc.fopen(str(inp).encode(), b"r")
    try:
        assert L_.read_cutoff_infile_header(C.c_void_p(f), C.byref(h)) == 0
        setup_models(models, f, h, p)
    finally:
        libc.fclose(f)
    outs = {k: tmp_path / f"{k}_{name}" for k in ("tragsm", "trasph")}
    g1 = libc.fopen(str(outs["tragsm"]).encode(), b"w")
    g2 = libc.fopen(str(outs["trasph"]).encode(), b"w")
    try:
        L_.trajectory_simulation(C.byref(h), C.cast(models, C.c_void_p), C.byref(p), g1, g2, steps, 0.01, A.M_PROTON)
    finally:
        libc.fclose(g1)
        libc.fclose(g2)
    compare_tracks(name, outs["tragsm"], outs["trasph"])


def compare_tracks(name, gsm_file, sph_file):
    for kind, path in (("tragsm", gsm_file), ("trasph", sph_file)):
        mine = np.loadtxt(path, ndmin=2)
        ref = np.loadtxt(GOLD / f"{name}.{kind}", ndmin=2)
        assert mine.shape == ref.shape, (kind, mine.shape, ref.shape)
        assert np.array_equal(mine[:, 0], ref[:, 0])
        # one unit in the last printed digit of each column ("%17.5f" x3, "%14.2f",
        # "%15.10f", "%18.2f") is always allowed
        unit = np.array([1e-5, 1e-5, 1e-5, 1e-2, 1e-10, 1e-2])
        scale = np.maximum(np.abs(ref[:, 1:]).max(axis=0), 1e-30)
        excess = np.maximum(np.abs(mine[:, 1:] - ref[:, 1:]) - 1.0001 * unit, 0.0) / scale
        half = len(ref) // 2
        print(name, kind, len(ref), "lines; max deviation beyond the printed unit / scale: first half",
              excess[:half].max(), "all", excess.max())
        assert excess[:half].max() <= 1e-9 and excess.max() <= 1e-6


# ---------------------------------------------------------------- the `cor` command line

COR = REPO / "cor-model_amd" / "cor"
needs_cli = pytest.mark.skipif(not COR.exists(), reason="cor-model_amd/cor is not built")


def cor(*args, cwd=None):
    res = subprocess.run([str(COR), *map(str, args)], capture_output=True, text=True, timeout=300, cwd=cwd)
    assert res.returncode == 0, res.stderr
    return res.stdout


@needs_cli
def test_cli_cutoff_matches_reference_text():
    """cor -p -i igrf14 -s 1000000 INFILE: the reference's command line and output for
    config 1 (FP64 IGRF is bit-identical, so the cutoffs are equal)."""
    name = "config1_igrf_nk100"
    h = L.read_cutoff_header(GOLD / f"{name}.in")
    out = cor("-p", "-i", "igrf14", "-s", 1000000, GOLD / f"{name}.in")
    compare_cutoff_text(name, out, (GOLD / f"{name}.out").read_text(), h["rig"], h["del"], 0.0)


INTERNAL_RUNS = [("igrf13_2017_nk100", "igrf13"), ("cals10k2_1000_nk100", "cals10k.2")]


@needs_cli
@pytest.mark.parametrize("name,model", INTERNAL_RUNS, ids=[r[0] for r in INTERNAL_RUNS])
def test_cli_other_internal_models(name, model):
    """-i igrf13 at 2017 and -i cals10k.2 at 1000 CE against the reference's runs of the
    same inputs: the other internal models' epoch setup (bit-exact on the host) drives
    the same GPU synthesis, so the FP64 cutoffs are equal."""
    h = L.read_cutoff_header(GOLD / f"{name}.in")
    out = cor("-p", "-i", model, "-s", 200000, GOLD / f"{name}.in")
    compare_cutoff_text(name, out, (GOLD / f"{name}.out").read_text(), h["rig"], h["del"], 0.0)


@needs_cli
def test_cli_batch_of_input_files(tmp_path):
    """Several INFILEs in one invocation (one GPU launch for the files that share the
    epoch and T96 inputs): each output equals the single-file run's bar against the
    reference's text."""
    names = ["t96_equator_nk400", "t96_mid_nk400"]
    cor("-p", "-i", "igrf14", "-e", "t96", "-s", 1000000, "-o", tmp_path, *[GOLD / f"{n}.in" for n in names])
    for name in names:
        h = L.read_cutoff_header(GOLD / f"{name}.in")
        mine = (tmp_path / f"{name}.in.out").read_text()
        compare_cutoff_text(name, mine, (GOLD / f"{name}.out").read_text(), h["rig"], h["del"], 0.05)


@needs_cli
def test_cli_field_and_trajectory_modes(tmp_path):
    """-m field and -m trajectory through the command line, against the reference's
    field map and track files."""
    out = cor("-m", "field", "-i", "igrf14", "-e", "t96", GOLD / "field_igrf_t96.in").splitlines()
    ref = (GOLD / "field_igrf_t96.out").read_text().splitlines()
    assert len(out) == len(ref)
    for a, b in zip(out, ref):
        try:
            va, vb = np.array(a.split(), dtype=float), np.array(b.split(), dtype=float)
        except ValueError:
            assert a == b
            continue
        assert va.shape == vb.shape and np.all(np.abs(va - vb) <= 0.0100001), (a, b)
    cor("-m", "trajectory", "-i", "igrf14", "-e", "t96", "-s", 20000, "-o", "t.txt", GOLD / "track_t96_5gv.in",
        cwd=tmp_path)
    compare_tracks("track_t96_5gv", tmp_path / "tragsm_t.txt", tmp_path / "trasph_t.txt")


def test_calculate_position_and_speed_is_the_reference_step():
    """calculate_position_and_speed (src/trajectory.c:275-356) through the C API, on the
    GPU, bit for bit against the oracle's restatement of the step (itself pinned by the
    reference's golden cutoff runs and the 5th-order convergence closed form), over
    seeded states of the size the tracer meets (Earth radii, MeV..GeV protons, fields
    of tens of nT to tens of microtesla)."""
    L_ = A.lib()
    L_.calculate_position_and_speed.restype = None
    L_.calculate_position_and_speed.argtypes = [C.c_double, C.c_double, A.Vector, C.POINTER(A.Vector),
                                                C.POINTER(A.Vector)]
    o = L.Oracle()
    o.lib.o_rk_step.argtypes = [C.c_double, C.c_double, L.DP, L.DP, L.DP]
    o.lib.o_rk_step.restype = None
    rng = np.random.default_rng(275)
    A.init(0)
    for _ in range(200):
        h = 10.0 ** rng.uniform(-6, -2)
        d = -(10.0 ** rng.uniform(-1.5, 0.5))
        f = rng.normal(size=3) * 10.0 ** rng.uniform(1, 4.5)
        v = rng.normal(size=3) * 1e8
        x = rng.normal(size=3) * 2e7
        vg, xg = A.Vector(*v), A.Vector(*x)
        L_.calculate_position_and_speed(h, d, A.Vector(*f), C.byref(vg), C.byref(xg))
        vo, xo = v.copy(), x.copy()
        o.lib.o_rk_step(h, d, f.ctypes.data_as(L.DP), vo.ctypes.data_as(L.DP), xo.ctypes.data_as(L.DP))
        assert (vg.x, vg.y, vg.z) == tuple(vo) and (xg.x, xg.y, xg.z) == tuple(xo)
