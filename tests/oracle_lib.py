"""ctypes bindings of the CPU checkers: the oracle restatement (oracle/liboracle.so)
and, where it was built, the reference itself compiled from its sources
(oracle/_ref/libcorref.so, oracle/_ref/cor).

Test infrastructure only: imported by tests/, the smoke check and the CPU-baseline
leg of the benchmark, never by the product.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
ORACLE_DIR = REPO / "oracle"
ORACLE_SO = ORACLE_DIR / "liboracle.so"
REF_SO = ORACLE_DIR / "_ref" / "libcorref.so"
REF_BIN = ORACLE_DIR / "_ref" / "cor"

NCOEF = 105
DP = C.POINTER(C.c_double)


class OParams(C.Structure):
    """Layout of the reference ModelsParams (src/models.h:9-17)."""
    _fields_ = [("PARMOD", C.c_double * 10), ("G", C.c_double * NCOEF),
                ("H", C.c_double * NCOEF), ("REC", C.c_double * NCOEF),
                ("A1", C.c_double * 3), ("A2", C.c_double * 3), ("A3", C.c_double * 3)]


class OModels(C.Structure):
    _fields_ = [("internal", C.c_int32), ("external", C.c_int32), ("crustal", C.c_int32),
                ("chaos_nm_max", C.c_int32), ("chaos_g", DP), ("chaos_h", DP)]


class OScan(C.Structure):
    _fields_ = [("r0", C.c_double), ("the0", C.c_double), ("fi0", C.c_double),
                ("the1", C.c_double), ("fi1", C.c_double), ("rig", C.c_double),
                ("del_", C.c_double), ("rk", C.c_double), ("zn", C.c_double),
                ("nk1", C.c_int32), ("pad", C.c_int32)]


class OTraj(C.Structure):
    _fields_ = [("status", C.c_int32), ("steps", C.c_uint32), ("attempts", C.c_uint32),
                ("pad", C.c_uint32), ("vc", C.c_double), ("r", C.c_double),
                ("td", C.c_double), ("fei", C.c_double), ("ast", C.c_double),
                ("asf", C.c_double), ("time", C.c_double), ("length_m", C.c_double)]


OTRAJ_DTYPE = np.dtype([("status", "<i4"), ("steps", "<u4"), ("attempts", "<u4"),
                        ("pad", "<u4"), ("vc", "<f8"), ("r", "<f8"), ("td", "<f8"),
                        ("fei", "<f8"), ("ast", "<f8"), ("asf", "<f8"), ("time", "<f8"),
                        ("length_m", "<f8")])

INT_NONE, INT_IGRF = 0, 1
EXT_NONE, EXT_T96, EXT_TS05 = 0, 1, 2
CRU_NONE, CRU_CHAOS = 0, 1
ALLOWED, FORBIDDEN_REENTRY, FORBIDDEN_LIMIT = 0, 1, 2
M_PROTON = 1.6725e-27


def _dp(a):
    return a.ctypes.data_as(DP)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} is not built (make -C oracle)")
        lib = C.CDLL(str(path))
        self.lib = lib
        lib.o_igrf14_recalc.argtypes = [C.c_int] * 7 + [C.POINTER(OParams)]
        lib.o_igrf14_recalc.restype = C.c_int
        lib.o_igrf_geo.argtypes = [C.c_double] * 3 + [DP] * 4
        lib.o_t96_gsm.argtypes = [DP] + [C.c_double] * 3 + [DP]
        lib.o_ts05_gsm.argtypes = [DP] + [C.c_double] * 3 + [DP]
        lib.o_chaos_geo.argtypes = [C.c_double] * 3 + [C.c_int, DP, DP, DP]
        lib.o_calculate_field.argtypes = [DP, DP, C.POINTER(OModels), C.POINTER(OParams)]
        lib.o_field_batch.argtypes = [C.POINTER(OModels), C.POINTER(OParams), DP, DP, C.c_int64]
        lib.o_scan_count.argtypes = [C.POINTER(OScan)]
        lib.o_scan_count.restype = C.c_int32
        lib.o_trace_one.argtypes = [C.POINTER(OScan), C.c_int32, C.POINTER(OModels),
                                    C.POINTER(OParams), C.c_uint64, C.c_double, C.c_double,
                                    C.POINTER(OTraj)]
        lib.o_trace_scan.argtypes = [C.POINTER(OScan), C.POINTER(OModels), C.POINTER(OParams),
                                     C.c_uint64, C.c_double, C.c_double, C.c_int32, C.c_void_p]
        lib.o_penumbra.argtypes = [C.POINTER(OScan), C.c_void_p, C.c_int32, DP]
        lib.o_rk_step.argtypes = [C.c_double, C.c_double, DP, DP, DP]
        lib.o_sun.argtypes = [C.c_int] * 5 + [DP]

    def recalc(self, iy, mes, ide, id_, ih, mi, is_):
        p = OParams()
        rc = self.lib.o_igrf14_recalc(iy, mes, ide, id_, ih, mi, is_, C.byref(p))
        return rc, p

    def igrf_geo(self, p, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(xyz)
        b = np.empty(3)
        for i, (x, y, z) in enumerate(xyz):
            self.lib.o_igrf_geo(x, y, z, p.G, p.H, p.REC, _dp(b))
            out[i] = b
        return out

    def t96(self, parmod, xyz):
        pm = np.zeros(10)
        pm[:len(parmod)] = parmod
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(xyz)
        b = np.empty(3)
        for i, (x, y, z) in enumerate(xyz):
            self.lib.o_t96_gsm(_dp(pm), x, y, z, _dp(b))
            out[i] = b
        return out

    def ts05(self, parmod, xyz):
        pm = np.zeros(10)
        pm[:len(parmod)] = parmod
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(xyz)
        b = np.empty(3)
        for i, (x, y, z) in enumerate(xyz):
            self.lib.o_ts05_gsm(_dp(pm), x, y, z, _dp(b))
            out[i] = b
        return out

    def chaos(self, g, h, xyz):
        g = np.ascontiguousarray(g, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        nm = g.shape[0]
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(xyz)
        b = np.empty(3)
        for i, (x, y, z) in enumerate(xyz):
            self.lib.o_chaos_geo(x, y, z, nm, _dp(g), _dp(h), _dp(b))
            out[i] = b
        return out

    def field(self, models, p, gsm_m):
        gsm_m = np.ascontiguousarray(gsm_m, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(gsm_m)
        self.lib.o_field_batch(C.byref(models), C.byref(p), _dp(gsm_m), _dp(out), len(gsm_m))
        return out

    def trace_scan(self, scan, models, p, step_limit, tu=0.01, mass=M_PROTON, threads=None):
        n = self.lib.o_scan_count(C.byref(scan))
        out = np.zeros(n, dtype=OTRAJ_DTYPE)
        threads = threads or os.cpu_count() or 1
        self.lib.o_trace_scan(C.byref(scan), C.byref(models), C.byref(p), step_limit, tu,
                              mass, threads, out.ctypes.data)
        cut = np.empty(3)
        self.lib.o_penumbra(C.byref(scan), out.ctypes.data, n, _dp(cut))
        return out, cut

    def trace_one(self, scan, i, models, p, step_limit, tu=0.01, mass=M_PROTON):
        t = OTraj()
        self.lib.o_trace_one(C.byref(scan), i, C.byref(models), C.byref(p), step_limit, tu,
                             mass, C.byref(t))
        return t


def models(internal=INT_IGRF, external=EXT_NONE, crustal=CRU_NONE, chaos=None):
    m = OModels(internal, external, crustal, 0, None, None)
    if chaos is not None:
        g, h = chaos
        m._keep = (np.ascontiguousarray(g, dtype=np.float64), np.ascontiguousarray(h, dtype=np.float64))
        m.chaos_nm_max = g.shape[0]
        m.chaos_g = _dp(m._keep[0])
        m.chaos_h = _dp(m._keep[1])
    return m


# ----------------------------------------------------------------- reference

class RecalcTime(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("iy", "mes", "ide", "id", "ih", "min", "is_")]


class RefLib:
    """The reference's own field functions, compiled from its sources here."""

    def __init__(self, path=REF_SO):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} is not built (make -C oracle ref)")
        lib = C.CDLL(str(path))
        self.lib = lib
        lib.internal_recalc.argtypes = [C.c_void_p, C.POINTER(RecalcTime)] + [DP] * 6
        lib.internal_recalc.restype = C.c_int
        lib.internal_geo.argtypes = [C.c_double] * 3 + [DP] * 6
        lib.t96_gsm.argtypes = [DP] + [C.c_double] * 3 + [DP] * 3
        lib.ts05_gsm.argtypes = [DP] + [C.c_double] * 3 + [DP] * 3
        lib.chaos_geo.argtypes = [C.c_double] * 3 + [C.c_void_p] + [DP] * 3
        self.igrf14 = C.c_void_p.in_dll(lib, "IGRF_14_COEFS")

    def recalc(self, iy, mes, ide, id_, ih, mi, is_, table="IGRF_14_COEFS"):
        """internal_recalc (lib/internal/internal.c:89-180) with one of the reference's
        coefficient descriptors (lib/internal/internal.h:22-29)."""
        p = OParams()
        t = RecalcTime(iy, mes, ide, id_, ih, mi, is_)
        addr = C.addressof(C.c_char.in_dll(self.lib, table))
        rc = self.lib.internal_recalc(C.c_void_p(addr), C.byref(t), p.G, p.H, p.REC,
                                      p.A1, p.A2, p.A3)
        return rc, p

    def _pointwise(self, fn, xyz, *pre):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(xyz)
        bx, by, bz = C.c_double(), C.c_double(), C.c_double()
        for i, (x, y, z) in enumerate(xyz):
            fn(*pre[:1], x, y, z, *pre[1:], C.byref(bx), C.byref(by), C.byref(bz))
            out[i] = (bx.value, by.value, bz.value)
        return out

    def igrf_geo(self, p, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(xyz)
        bx, by, bz = C.c_double(), C.c_double(), C.c_double()
        for i, (x, y, z) in enumerate(xyz):
            self.lib.internal_geo(x, y, z, p.G, p.H, p.REC, C.byref(bx), C.byref(by), C.byref(bz))
            out[i] = (bx.value, by.value, bz.value)
        return out

    def t96(self, parmod, xyz):
        pm = np.zeros(10)
        pm[:len(parmod)] = parmod
        return self._pointwise(self.lib.t96_gsm, xyz, _dp(pm))

    def ts05(self, parmod, xyz):
        pm = np.zeros(10)
        pm[:len(parmod)] = parmod
        return self._pointwise(self.lib.ts05_gsm, xyz, _dp(pm))

    def chaos716(self, xyz):
        """chaos_geo with the reference's own CHAOS-7.16 table (CHAOS_7_16_COEFS)."""
        co = C.addressof(C.c_char.in_dll(self.lib, "CHAOS_7_16_COEFS"))
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(xyz)
        bx, by, bz = C.c_double(), C.c_double(), C.c_double()
        for i, (x, y, z) in enumerate(xyz):
            self.lib.chaos_geo(x, y, z, co, C.byref(bx), C.byref(by), C.byref(bz))
            out[i] = (bx.value, by.value, bz.value)
        return out

    def chaos(self, g, h, xyz):
        """chaos_geo (lib/chaos/chaos.c:120-227) with a caller's [nm][nm] G/H table."""

        class Coefs(C.Structure):
            _fields_ = [("G", C.c_void_p), ("H", C.c_void_p), ("nm_max", C.c_uint32)]

        g = np.ascontiguousarray(g, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        co = Coefs(g.ctypes.data, h.ctypes.data, g.shape[0])
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty_like(xyz)
        bx, by, bz = C.c_double(), C.c_double(), C.c_double()
        for i, (x, y, z) in enumerate(xyz):
            self.lib.chaos_geo(x, y, z, C.addressof(co), C.byref(bx), C.byref(by), C.byref(bz))
            out[i] = (bx.value, by.value, bz.value)
        return out


def have_ref():
    return REF_SO.exists() and REF_BIN.exists()


def read_cutoff_header(path):
    """The numbers of a cutoff input header (src/io.c:7: rig zn rk / r0 the0 fi0 /
    the1 fi1 / date / nk1 iopt ist del)."""
    v = Path(path).read_text().split()
    return {"rig": float(v[0]), "zn": float(v[1]), "rk": float(v[2]), "r0": float(v[3]), "the0": float(v[4]),
            "fi0": float(v[5]), "the1": float(v[6]), "fi1": float(v[7]), "date": tuple(int(x) for x in v[8:15]),
            "nk1": int(v[15]), "del": float(v[18])}


def write_cutoff_input(path, rig, zn, rk, r0, the0, fi0, the1, fi1, date=(2020, 1, 1, 1, 0, 0, 0),
                       nk1=100, dl=0.01, ext_params=None):
    """Input file in the layout src/io.c:7 reads (plus the external-model lines)."""
    lines = [f"{rig!r} {zn!r} {rk!r}", f"{r0!r} {the0!r} {fi0!r}", f"{the1!r} {fi1!r}",
             " ".join(str(v) for v in date), f"{nk1} 0 0 {dl!r}"]
    if ext_params is not None:
        lines.append(" ".join(repr(float(v)) for v in ext_params[:4]))
        if len(ext_params) > 4:
            lines.append(" ".join(repr(float(v)) for v in ext_params[4:10]))
    Path(path).write_text("\n".join(lines) + "\n")


def run_ref_cutoff(infile, internal="igrf14", external=None, steps=1000000, tu=None,
                   parallel=True, timeout=600):
    cmd = [str(REF_BIN)]
    if parallel:
        cmd.append("-p")
    if internal:
        cmd += ["-i", internal]
    if external:
        cmd += ["-e", external]
    cmd += ["-s", str(steps)]
    if tu is not None:
        cmd += ["-u", repr(tu)]
    cmd.append(str(infile))
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, check=True).stdout


def parse_cutoff_output(text):
    """Allowed lines -> structured array; summary -> (Rl, Ru, Rc)."""
    rows, cut = [], None
    lines = text.splitlines()
    for k, ln in enumerate(lines):
        if "CUTOFF with rigidities" in ln:
            cut = tuple(float(v) for v in lines[k + 1].split())
            break
        parts = ln.split()
        if len(parts) == 9:
            try:
                rows.append(tuple(float(v) for v in parts))
            except ValueError:
                pass
    arr = np.array(rows, dtype=[(n, "f8") for n in
                               ("rig", "vc", "r", "td", "fei", "ast", "asf", "time", "length_km")])
    return arr, cut
