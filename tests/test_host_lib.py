"""The C-ABI library on the CPU: it loads, exports every symbol include/*.h declares,
and its host-side parts (IGRF-14 epoch setup, input parsing and output headers, scan
sizing, site geometry, argument validation) behave like the reference's -- no GPU
compute is called here.
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import cor_amd as A
import oracle_lib as L

REPO = Path(__file__).resolve().parents[1]
GOLD = REPO / "tests" / "golden"
libc = C.CDLL(None)
libc.fopen.restype = C.c_void_p
libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
libc.fclose.argtypes = [C.c_void_p]
libc.fflush.argtypes = [C.c_void_p]


def declared_functions():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
                names.add(m.group(1))
        for m in re.finditer(r"^extern const Model \*(\w+)\[\]", text, re.M):
            names.add(m.group(1))
        for m in re.finditer(r"^extern const InternalModelCoefs (\w+);", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    names = declared_functions()
    assert {"trajectory_simulation_cutoff", "cor_gpu_trace", "cor_gpu_field", "internal_models"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(A.EXPORTED) <= names


def test_internal_recalc_every_model_matches_reference_fixtures():
    """internal_recalc (lib/internal/internal.c:89-180) for IGRF-9..14, the historical
    model and CALS10k.2, bit for bit against the reference at epochs around each table's
    ends, inside interpolation intervals and in the extrapolation range; out-of-range
    years fail as the reference's do."""
    vec = np.load(GOLD / "internal_vectors.npz")
    lib = A.lib()
    n_ok = 0
    for tab in A.INTERNAL_TABLES.values():
        addr = C.addressof(C.c_char.in_dll(lib, tab))
        for k, e in enumerate(vec[f"{tab}_epochs"]):
            p = A.ModelsParams()
            t = A.RecalcTime(*[int(v) for v in e])
            rc = lib.internal_recalc(C.c_void_p(addr), C.byref(t), p.G, p.H, p.REC, p.A1, p.A2, p.A3)
            assert rc == int(vec[f"{tab}_{k}_rc"]), (tab, tuple(e))
            if rc == 0:
                n_ok += 1
                for f in ("G", "H", "REC", "A1", "A2", "A3"):
                    assert np.array_equal(np.array(getattr(p, f)), vec[f"{tab}_{k}_{f}"]), (tab, tuple(e), f)
    assert n_ok >= 100


def test_internal_models_registry_matches_reference():
    """src/models.c:22-112: the eight internal models by version, with their names."""
    lib = A.lib()
    get = lib.get_model
    get.restype = C.c_void_p
    get.argtypes = [C.c_void_p, C.c_char_p]
    reg = C.addressof(C.c_char.in_dll(lib, "internal_models"))
    names = {"igrf14": "IGRF-14", "igrf13": "IGRF-13", "igrf12": "IGRF-12", "igrf11": "IGRF-11",
             "igrf10": "IGRF-10", "igrf9": "IGRF-9", "hist": "Historical (0-1968 CE)",
             "cals10k.2": "Historical (CALS10k.2)"}
    for ver, name in names.items():
        m = get(reg, ver.encode())
        assert m, ver
        assert C.cast(C.c_void_p.from_address(m).value, C.c_char_p).value.decode() == name
    assert not get(reg, b"igrf15")


def test_string_parsers_follow_reference():
    """src/utils.c:8-26: whole-string parses; the command line's -s and -u values."""
    lib = A.lib()
    lib.string_to_uint64.argtypes = [C.c_char_p, C.POINTER(C.c_uint64)]
    lib.string_to_double.argtypes = [C.c_char_p, C.POINTER(C.c_double)]
    u = C.c_uint64()
    d = C.c_double()
    for txt, rc, val in [(b"1000000", 0, 1000000), (b"25000", 0, 25000), (b"12x", -1, None), (b"", -1, None),
                         (b" 7", 0, 7), (b"-1", 0, 2 ** 64 - 1), (b"18446744073709551615", 0, 2 ** 64 - 1)]:
        assert lib.string_to_uint64(txt, C.byref(u)) == rc, txt
        if val is not None:
            assert u.value == val, txt
    for txt, rc, val in [(b"0.01", 0, 0.01), (b"1e-3", 0, 1e-3), (b"0.01rad", -1, None), (b"", -1, None)]:
        assert lib.string_to_double(txt, C.byref(d)) == rc, txt
        if val is not None:
            assert d.value == val, txt


def test_caller_defined_model_is_refused():
    """A Model the registries do not hold has no GPU evaluator: the model list is
    refused before any device work (the reference's five-field struct is all that is
    read)."""
    lib = A.lib()

    class Model(C.Structure):
        _fields_ = [("name", C.c_char_p), ("version", C.c_char_p), ("setup_data", C.c_void_p),
                    ("setup_fun", C.c_void_p), ("calc_fun", C.c_void_p)]

    CALC = C.CFUNCTYPE(None, A.Vector, C.POINTER(A.Vector), C.c_void_p)
    fn = CALC(lambda pos, tot, par: None)
    m = Model(b"my model", b"mine", None, None, C.cast(fn, C.c_void_p))
    models = (C.c_void_p * 2)(C.addressof(m), None)
    lib.calculate_field.argtypes = [A.Vector, C.POINTER(A.Vector), C.c_void_p, C.POINTER(A.ModelsParams)]
    lib.calculate_field.restype = None
    out = A.Vector()
    p = A.recalc(2020)
    lib.calculate_field(A.Vector(1e7, 0, 0), C.byref(out), models, C.byref(p))
    assert np.isnan(out.x) and np.isnan(out.y) and np.isnan(out.z)
    assert b"my model" in lib.cor_gpu_last_hip_error()


def test_igrf14_recalc_matches_reference_fixtures():
    vec = np.load(GOLD / "field_vectors.npz")
    for k, e in enumerate(vec["recalc_epochs"]):
        p = A.recalc(*[int(v) for v in e])
        for f in ("G", "H", "REC", "A1", "A2", "A3"):
            assert np.array_equal(np.array(getattr(p, f)), vec[f"recalc{k}_{f}"]), (tuple(e), f)
    with pytest.raises(A.CorError):
        A.recalc(2031)


def _with_file(path, mode, fn):
    f = libc.fopen(str(path).encode(), mode.encode())
    assert f
    try:
        return fn(f)
    finally:
        libc.fclose(f)


class InfileHeader(C.Structure):
    _fields_ = [("r0", C.c_double), ("the0", C.c_double), ("fi0", C.c_double), ("the1", C.c_double),
                ("fi1", C.c_double), ("iy", C.c_int), ("mes", C.c_int), ("ide", C.c_int), ("id", C.c_int),
                ("ih", C.c_int), ("min", C.c_int), ("is_", C.c_int), ("zn", C.c_double), ("nk1", C.c_int),
                ("iopt", C.c_int), ("ist", C.c_int), ("rig", C.c_double), ("del_", C.c_double),
                ("rk", C.c_double)]


def test_input_parsing_and_output_header(tmp_path):
    """src/io.c:4-50,115-127: the reference's input layout in, its header out (the
    header of a golden run, byte for byte)."""
    lib = A.lib()
    h = InfileHeader()
    p = A.ModelsParams()
    inp = GOLD / "t96_mid_nk400.in"
    rc = _with_file(inp, "r", lambda f: (lib.read_cutoff_infile_header(C.c_void_p(f), C.byref(h)),
                                        lib.read_t96_params(C.c_void_p(f), C.byref(p))))
    assert rc == (0, 0)
    assert (h.rig, h.zn, h.rk, h.nk1, h.del_) == (0.05, -1.0, 20.0, 400, 0.05)
    assert (h.iy, h.mes, h.ide, h.id) == (2020, 1, 1, 1)
    assert list(p.PARMOD[:4]) == [2.0, -20.0, 0.0, -2.0]          # Dst first in the file
    models = (C.c_void_p * 4)()
    lib.get_model.restype = C.c_void_p
    lib.get_model.argtypes = [C.c_void_p, C.c_char_p]
    internal = C.addressof(C.c_void_p.in_dll(lib, "internal_models"))
    external = C.addressof(C.c_void_p.in_dll(lib, "external_models"))
    models[0] = lib.get_model(internal, b"igrf14")
    models[1] = lib.get_model(external, b"t96")
    assert models[0] and models[1]
    assert lib.get_model(external, b"t01") is None
    out = tmp_path / "hdr.txt"
    _with_file(out, "w", lambda f: lib.write_cutoff_outfile_header(C.c_void_p(f), C.byref(h), models,
                                                                   C.c_uint64(1000000)))
    ref = (GOLD / "t96_mid_nk400.out").read_text()
    mine = out.read_text()
    assert ref.startswith(mine)


def test_short_input_is_rejected(tmp_path):
    lib = A.lib()
    bad = tmp_path / "bad.in"
    bad.write_text("")
    h = InfileHeader()
    assert _with_file(bad, "r", lambda f: lib.read_cutoff_infile_header(C.c_void_p(f), C.byref(h))) == -1
    bad.write_text("-20 2 0\n")
    p = A.ModelsParams()
    assert _with_file(bad, "r", lambda f: lib.read_t96_params(C.c_void_p(f), C.byref(p))) == -1


def test_scan_count_like_the_reference():
    row = A.scans_array([(1.0, 0, 0, 0, 0, 0.01, 0.01, 20.0, -1.0, 100)])[0]
    assert A.scan_count(row) == 2000                      # round(1999) + 1
    row["rk"] = 0.01
    assert A.scan_count(row) == 1
    row["rk"] = 0.005
    assert A.scan_count(row) == 0                         # rk below rig: no trajectories
    row["del_"] = 0.0
    assert A.scan_count(row) == A.COR_E_INVALID           # src/main.c:45 rejects del <= 0


def test_site_geometry():
    r0, lat = A.geodetic_to_geocentric(47.0, 20.0)
    assert abs(r0 - 1.002439629) < 1e-9 and abs(lat - 46.808601777) < 1e-9
    r0, lat = A.geodetic_to_geocentric(90.0, 20.0)
    assert lat == pytest.approx(90.0) and r0 == pytest.approx((6356.7523142 + 20.0) / 6371.2, rel=1e-9)
    assert A.direction_to_geo(46.8, 8.0, 0.0, 123.0) == pytest.approx((46.8, 8.0))
    t, f = A.direction_to_geo(0.0, 0.0, 90.0, 90.0)       # horizontal, due east at (0, 0)
    assert t == pytest.approx(0.0, abs=1e-12) and f == pytest.approx(90.0)
    t, f = A.direction_to_geo(0.0, 0.0, 90.0, 0.0)        # horizontal, due north
    assert t == pytest.approx(90.0)
    assert len(A.grid_sites()) == 37 * 72


def test_gpu_calls_fail_cleanly_without_a_gpu():
    """On a host with no GPU the library reports it; nothing falls back to the CPU."""
    import os
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = A.lib()
    assert lib.cor_gpu_init(0) == A.COR_E_NO_DEVICE
    p = A.recalc(2020)
    out = np.zeros(3)
    rc = lib.cor_gpu_field(C.byref(A.model_set()), C.byref(p), np.zeros(3).ctypes.data_as(A.DP),
                           out.ctypes.data_as(A.DP), 1, 0)
    assert rc == A.COR_E_NO_DEVICE


def test_field_precision_checked_first():
    """cor_gpu_field rejects a precision outside FP64, mixed, the five mixed cooperative
    evaluators (64, 8, 16, 32, 4 lanes) and the two FP64 ones (64, 8) before any device
    work."""
    lib = A.lib()
    p = A.recalc(2020)
    out = np.zeros(3)
    pts = np.zeros(3)
    for bad in (-1, 9, 99):
        rc = lib.cor_gpu_field(C.byref(A.model_set()), C.byref(p), pts.ctypes.data_as(A.DP),
                               out.ctypes.data_as(A.DP), 1, bad)
        assert rc == A.COR_E_INVALID, bad
    assert (A.PREC_MIXED_GROUP, A.PREC_MIXED_GROUP16, A.PREC_MIXED_GROUP32) == (3, 4, 5)
    assert (A.PREC_FP64_WAVE, A.PREC_FP64_GROUP, A.PREC_MIXED_GROUP4) == (6, 7, 8)


def test_epoch_series_argument_checks():
    """Invalid arguments of the epoch-series entry points are rejected before any
    device work (the same checks run on a GPU host)."""
    lib = A.lib()
    p = A.recalc(2020)
    ms = A.model_set()
    assert lib.cor_trace_plan_set_epoch(None, C.byref(ms), C.byref(p), None) == A.COR_E_INVALID
    scans = A.scans_array([(1.0, 45.0, 0.0, 45.0, 0.0, 1.0, 1.0, 3.0, -1.0, 100)])
    opts = A.TraceOpts(1000, 0.01, A.M_PROTON, A.PREC_FP64, 0, 0, 0)
    arr = (A.ModelsParams * 1)(p)
    # no epochs, no scans, negative in-flight count
    assert lib.cor_gpu_trace_series(C.byref(ms), C.cast(arr, C.c_void_p), 0, scans.ctypes.data, 1,
                                    C.byref(opts), 0, None, None) == A.COR_E_INVALID
    assert lib.cor_gpu_trace_series(C.byref(ms), C.cast(arr, C.c_void_p), 1, scans.ctypes.data, 0,
                                    C.byref(opts), 0, None, None) == A.COR_E_INVALID
    assert lib.cor_gpu_trace_series(C.byref(ms), C.cast(arr, C.c_void_p), 1, scans.ctypes.data, 1,
                                    C.byref(opts), -1, None, None) == A.COR_E_INVALID


def test_refine_scan():
    """[Rl - del, Ru] at del / factor, clamped to the coarse scan's range."""
    row = A.scans_array([(1.0, 45.0, 0.0, 45.0, 0.0, 0.01, 0.01, 20.0, -1.0, 100)])
    cut = np.zeros(1, dtype=A.CUT_DTYPE)
    cut[0] = (3.72, 4.77, 4.2, 0, 2000)
    f = A.refine_scans(row, cut, 10)[0]
    assert np.isclose(f["rig"], 3.71) and f["rk"] == 4.77 and np.isclose(f["del_"], 0.001)
    assert A.scan_count(f) == 1061
    assert (f["r0"], f["the0"], f["fi0"], f["zn"], f["nk1"]) == (1.0, 45.0, 0.0, -1.0, 100)
    cut[0] = (0.01, 0.01, 0.01, 0, 2000)                     # everything allowed
    f = A.refine_scans(row, cut, 10)[0]
    assert f["rig"] == 0.01 and f["rk"] == 0.01 and A.scan_count(f) == 1
    cut[0] = (20.0, 20.01, 20.0, 2000, 2000)                 # everything forbidden
    f = A.refine_scans(row, cut, 10)[0]
    assert np.isclose(f["rig"], 19.99) and f["rk"] == 20.0 and A.scan_count(f) == 11
    with pytest.raises(A.CorError):
        A.refine_scans(row, cut, 0)


def chaos716_tables():
    """The crustal registry's CHAOS-7.16 tables as the library holds them: the stored
    triangle (m <= n, row after row) expanded to [186][186] as the GPU receives it."""
    lib = A.lib()

    class Model(C.Structure):
        _fields_ = [("name", C.c_char_p), ("version", C.c_char_p), ("setup_data", C.c_void_p),
                    ("setup_fun", C.c_void_p), ("calc_fun", C.c_void_p)]

    lib.get_model.restype = C.c_void_p
    lib.get_model.argtypes = [C.c_void_p, C.c_char_p]
    m = lib.get_model(C.addressof(C.c_void_p.in_dll(lib, "crustal_models")), b"chaos7.16")
    assert m
    md = Model.from_address(m)
    nm = 186
    ntri = nm * (nm + 1) // 2
    tri = np.ctypeslib.as_array((C.c_double * (2 * ntri)).in_dll(lib, "cor_chaos716_tri")).reshape(2, ntri)
    g = np.zeros((nm, nm))
    h = np.zeros((nm, nm))
    rows, cols = np.tril_indices(nm)
    g[rows, cols] = tri[0]
    h[rows, cols] = tri[1]
    return md.name.decode(), g, h


def test_chaos716_table_matches_the_reference():
    """The CHAOS-7.16 coefficients the library carries (extracted by gen_chaos_table.py)
    reproduce the reference's chaos_geo with its own table bit for bit, through the
    oracle's synthesis (itself bit-identical to chaos_geo on seeded tables)."""
    name, g, h = chaos716_tables()
    assert name == "CHAOS-7.16" and g.shape == (186, 186)
    assert np.all(np.triu(g, 1) == 0) and np.all(np.triu(h, 1) == 0) and np.all(h[:, 0] == 0)
    cv = np.load(GOLD / "chaos_vectors.npz")
    got = L.Oracle().chaos(g, h, cv["chaos716_pts"])
    want = cv["chaos716_b"]
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    assert same.all(), np.argwhere(~same)[:5]


def test_t96_derived_tables_match_generator():
    """t96_derived.h (FP32 reciprocal scales and decay rates of the T96 shield
    harmonics) is what gen_t96_derived.py computes from t96_dev.hpp's coefficient
    tables; and a decay rate is sqrt(1/p^2 + 1/r^2) of the ring-current set's first
    y and z scales (12.75434981, 1.752483754, lib/t96/t96.c:274-299)."""
    import importlib.util
    gpu = Path(__file__).resolve().parents[1] / "cor-model_amd" / "csrc" / "gpu"
    spec = importlib.util.spec_from_file_location("gen_t96_derived", gpu / "gen_t96_derived.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    assert gen.render() == (gpu / "t96_derived.h").read_text()
    tabs = dict(gen.derived())
    assert abs(tabs["k_shl_sq"][0] - (1 / 12.75434981 ** 2 + 1 / 1.752483754 ** 2) ** 0.5) < 1e-15
    assert abs(tabs["k_shl_rp"][0] * 12.75434981 - 1) < 1e-15 and abs(tabs["k_shl_rk"][0] * 1.752483754 - 1) < 1e-15


REF_SRC = Path("/root/reference/src")


@pytest.mark.skipif(not (REF_SRC / "main.c").exists(), reason="reference sources not present")
def test_reference_cli_links_against_library(tmp_path):
    """The reference's command line (src/main.c + src/args.c, with its own headers,
    unedited) compiles and links against libcor_amd.so alone: every symbol its callers
    use (string_to_uint64, string_to_double, the registries, get_model, the io
    functions, the three simulation entry points) is exported with the reference's
    signatures.  Only --help and --version are run here (no GPU)."""
    import subprocess
    lib_dir = REPO / "cor-model_amd"
    exe = tmp_path / "cor_relinked"
    cmd = ["gcc", "-std=c17", "-O1", "-w", f"-I{REF_SRC}", f"-I{REF_SRC.parent / 'lib'}", "-o", str(exe),
           str(REF_SRC / "main.c"), str(REF_SRC / "args.c"), f"-L{lib_dir}", "-lcor_amd", f"-Wl,-rpath,{lib_dir}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    out = subprocess.run([str(exe), "--version"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "cor-model 3.0" in out.stdout
    out = subprocess.run([str(exe), "--help"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "--internal-field" in out.stdout
