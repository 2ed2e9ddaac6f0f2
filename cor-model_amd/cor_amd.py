"""ctypes binding of libcor_amd.so (include/cor.h, include/cor_gpu.h).

This is the Python face of the C-ABI the reference's C callers link against: the
tests, the smoke check and the benchmark drive the GPU tracer through it.  It loads
the in-tree library and fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcor_amd.so"

NCOEF = 105
DP = C.POINTER(C.c_double)

COR_OK, COR_E_INVALID, COR_E_NO_DEVICE, COR_E_HIP, COR_E_UNSUPPORTED, COR_E_NOMEM = 0, -1, -2, -3, -4, -5
PREC_FP64, PREC_MIXED, PREC_MIXED_WAVE, PREC_MIXED_GROUP, PREC_MIXED_GROUP16, PREC_MIXED_GROUP32 = 0, 1, 2, 3, 4, 5
PREC_FP64_WAVE, PREC_FP64_GROUP, PREC_MIXED_GROUP4 = 6, 7, 8
INTERNAL_NONE, INTERNAL_IGRF = 0, 1
EXTERNAL_NONE, EXTERNAL_T96, EXTERNAL_TS05 = 0, 1, 2
CRUSTAL_NONE, CRUSTAL_SH = 0, 1
ALLOWED, FORBIDDEN_REENTRY, FORBIDDEN_LIMIT = 0, 1, 2
M_PROTON = 1.6725e-27
M_ELECTRON = 9.1093826e-31
RE_M = 6.3712e06

# every symbol include/cor.h and include/cor_gpu.h declare
EXPORTED = [
    "get_model", "calculate_field", "sph_to_car", "car_to_sph", "geo_to_gsm", "gsm_to_geo",
    "internal_recalc", "read_cutoff_infile_header", "write_cutoff_outfile_header",
    "read_field_infile_header", "write_field_outfile_header", "read_ts05_params",
    "read_t96_params", "trajectory_simulation_cutoff", "field_simulation",
    "internal_models", "external_models", "crustal_models",
    "cor_gpu_init", "cor_gpu_error_string", "cor_gpu_last_hip_error", "cor_trace_plan_create",
    "cor_trace_plan_run", "cor_trace_plan_sync", "cor_trace_plan_count", "cor_trace_plan_results",
    "cor_trace_plan_stats", "cor_trace_plan_destroy", "cor_gpu_trace", "cor_gpu_field",
    "cor_scan_count", "cor_geodetic_to_geocentric", "cor_direction_to_geo", "cor_gpu_synchronize",
    "cor_trace_plan_set_epoch", "cor_gpu_trace_series", "cor_refine_scan", "cor_gpu_trace_tracks",
    "trajectory_simulation", "calculate_position_and_speed", "string_to_uint64", "string_to_double",
    "cor_gpu_rk_step", "cor_trace_plan_create_epochs", "cor_gpu_trace_epochs",
    "cor_trace_autotune", "cor_trace_plan_progress", "IGRF_9_COEFS", "IGRF_10_COEFS", "IGRF_11_COEFS", "IGRF_12_COEFS",
    "IGRF_13_COEFS", "IGRF_14_COEFS", "HIST_COEFS", "CALS10K_2_COEFS",
]

# internal model versions (src/models.c:110) -> descriptor symbols
INTERNAL_TABLES = {"igrf14": "IGRF_14_COEFS", "igrf13": "IGRF_13_COEFS", "igrf12": "IGRF_12_COEFS",
                   "igrf11": "IGRF_11_COEFS", "igrf10": "IGRF_10_COEFS", "igrf9": "IGRF_9_COEFS",
                   "hist": "HIST_COEFS", "cals10k.2": "CALS10K_2_COEFS"}


class Vector(C.Structure):
    """src/common.h:13-24 (x, y, z; or r, theta, phi)"""
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class ModelsParams(C.Structure):
    _fields_ = [("PARMOD", C.c_double * 10), ("G", C.c_double * NCOEF), ("H", C.c_double * NCOEF),
                ("REC", C.c_double * NCOEF), ("A1", C.c_double * 3), ("A2", C.c_double * 3),
                ("A3", C.c_double * 3)]


class RecalcTime(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("iy", "mes", "ide", "id", "ih", "min", "is_")]


class ModelSet(C.Structure):
    _fields_ = [("internal", C.c_int32), ("external", C.c_int32), ("crustal", C.c_int32),
                ("sh_nm_max", C.c_int32), ("sh_g", DP), ("sh_h", DP)]


class Scan(C.Structure):
    _fields_ = [("r0", C.c_double), ("the0", C.c_double), ("fi0", C.c_double),
                ("the1", C.c_double), ("fi1", C.c_double), ("rig", C.c_double),
                ("del_", C.c_double), ("rk", C.c_double), ("zn", C.c_double),
                ("nk1", C.c_int32), ("pad", C.c_int32)]


class TraceOpts(C.Structure):
    _fields_ = [("step_limit", C.c_uint64), ("tu_angle", C.c_double), ("particle_mass", C.c_double),
                ("precision", C.c_int32), ("queue_order", C.c_int32), ("coop", C.c_int32),
                ("handover_steps", C.c_int32), ("waves_per_simd", C.c_int32), ("block_threads", C.c_int32),
                ("handover_low_steps", C.c_int32), ("handover_low_count", C.c_int32)]


class Cutoff(C.Structure):
    _fields_ = [("rl", C.c_double), ("ru", C.c_double), ("rc", C.c_double),
                ("n_forbidden", C.c_int32), ("n", C.c_int32)]


class TraceStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("steps", C.c_uint64),
                ("attempts", C.c_uint64), ("n_traj", C.c_int64), ("grid_blocks", C.c_int32),
                ("block_threads", C.c_int32), ("wave_iters", C.c_uint64), ("lane_iters", C.c_uint64),
                ("coop_steps", C.c_uint64), ("wave_iters_open", C.c_uint64), ("lane_iters_open", C.c_uint64),
                ("drain_ms", C.c_double), ("lane_end_ms", C.c_double)]


TRAJ_DTYPE = np.dtype([("status", "<i4"), ("steps", "<u4"), ("attempts", "<u4"), ("handover", "<u4"),
                       ("vc", "<f8"), ("r", "<f8"), ("td", "<f8"), ("fei", "<f8"), ("ast", "<f8"),
                       ("asf", "<f8"), ("time", "<f8"), ("length_m", "<f8")])
SCAN_DTYPE = np.dtype([("r0", "<f8"), ("the0", "<f8"), ("fi0", "<f8"), ("the1", "<f8"),
                       ("fi1", "<f8"), ("rig", "<f8"), ("del_", "<f8"), ("rk", "<f8"),
                       ("zn", "<f8"), ("nk1", "<i4"), ("pad", "<i4")])
TRACK_DTYPE = np.dtype([("step", "<u4"), ("pad", "<u4"), ("gsm", "<f8", (3,)), ("r", "<f8"), ("td", "<f8"),
                        ("fei", "<f8"), ("v", "<f8"), ("time", "<f8"), ("length_m", "<f8")])
CUT_DTYPE = np.dtype([("rl", "<f8"), ("ru", "<f8"), ("rc", "<f8"), ("n_forbidden", "<i4"),
                      ("n", "<i4")])


class CorError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libcor_amd.so (built in-tree by __graft_entry__.build / make)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise CorError(f"{LIB_PATH} is not built: run `make -C cor-model_amd/csrc`")
    L = C.CDLL(str(LIB_PATH))
    L.cor_gpu_init.argtypes = [C.c_int]
    L.cor_gpu_error_string.restype = C.c_char_p
    L.cor_gpu_error_string.argtypes = [C.c_int]
    L.cor_gpu_last_hip_error.restype = C.c_char_p
    L.internal_recalc.argtypes = [C.c_void_p, C.POINTER(RecalcTime)] + [DP] * 6
    L.internal_recalc.restype = C.c_int
    L.cor_gpu_rk_step.argtypes = [DP, DP, DP, DP, DP, C.c_int64]
    L.cor_scan_count.argtypes = [C.POINTER(Scan)]
    L.cor_scan_count.restype = C.c_int64
    L.cor_trace_plan_create.argtypes = [C.POINTER(ModelSet), C.POINTER(ModelsParams), C.c_void_p,
                                        C.c_int32, C.POINTER(TraceOpts), C.POINTER(C.c_void_p)]
    L.cor_trace_plan_run.argtypes = [C.c_void_p, C.c_void_p]
    L.cor_trace_plan_sync.argtypes = [C.c_void_p]
    L.cor_trace_plan_progress.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    L.cor_trace_plan_count.argtypes = [C.c_void_p]
    L.cor_trace_plan_count.restype = C.c_int64
    L.cor_trace_plan_results.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.cor_trace_plan_stats.argtypes = [C.c_void_p, C.POINTER(TraceStats)]
    L.cor_trace_plan_destroy.argtypes = [C.c_void_p]
    L.cor_trace_plan_destroy.restype = None
    L.cor_gpu_trace.argtypes = [C.POINTER(ModelSet), C.POINTER(ModelsParams), C.c_void_p, C.c_int32,
                                C.POINTER(TraceOpts), C.c_void_p, C.c_void_p]
    L.cor_trace_plan_set_epoch.argtypes = [C.c_void_p, C.POINTER(ModelSet), C.POINTER(ModelsParams), C.c_void_p]
    L.cor_gpu_trace_series.argtypes = [C.POINTER(ModelSet), C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                       C.POINTER(TraceOpts), C.c_int32, C.c_void_p, C.c_void_p]
    L.cor_trace_autotune.argtypes = [C.POINTER(ModelSet), C.POINTER(ModelsParams), C.c_void_p, C.c_int32,
                                     C.POINTER(TraceOpts), C.c_int32, C.c_uint64, C.POINTER(TraceOpts),
                                     C.POINTER(C.c_double)]
    L.cor_trace_plan_create_epochs.argtypes = [C.POINTER(ModelSet), C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                               C.POINTER(TraceOpts), C.c_void_p]
    L.cor_gpu_trace_epochs.argtypes = [C.POINTER(ModelSet), C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                       C.POINTER(TraceOpts), C.c_void_p, C.c_void_p]
    L.cor_gpu_trace_tracks.argtypes = [C.POINTER(ModelSet), C.POINTER(ModelsParams), C.c_void_p, C.c_int32,
                                       C.POINTER(TraceOpts), C.c_void_p, C.POINTER(C.c_int64), C.c_void_p,
                                       C.c_int64]
    L.cor_gpu_field.argtypes = [C.POINTER(ModelSet), C.POINTER(ModelsParams), DP, DP, C.c_int64, C.c_int32]
    L.cor_geodetic_to_geocentric.argtypes = [C.c_double, C.c_double, DP, DP]
    L.cor_direction_to_geo.argtypes = [C.c_double] * 4 + [DP, DP]
    L.cor_refine_scan.argtypes = [C.POINTER(Scan), C.POINTER(Cutoff), C.c_int32, C.POINTER(Scan)]
    _lib = L
    return L


def check(rc, what=""):
    if rc != COR_OK:
        L = lib()
        msg = L.cor_gpu_error_string(rc).decode()
        hip = L.cor_gpu_last_hip_error().decode()
        raise CorError(f"{what}: {msg} ({rc}) {hip}")


def init(device=0):
    check(lib().cor_gpu_init(device), "cor_gpu_init")


def synchronize():
    check(lib().cor_gpu_synchronize(), "cor_gpu_synchronize")


def recalc(iy, mes=1, ide=1, id_=1, ih=0, mi=0, is_=0, parmod=None, model="igrf14"):
    """internal_recalc of an internal model (default IGRF-14) on the host ->
    ModelsParams (PARMOD filled from `parmod`, in the reference order Pdyn, Dst, By,
    Bz, W1..W6)."""
    p = ModelsParams()
    t = RecalcTime(iy, mes, ide, id_, ih, mi, is_)
    tab = C.addressof(C.c_char.in_dll(lib(), INTERNAL_TABLES[model]))
    rc = lib().internal_recalc(C.c_void_p(tab), C.byref(t), p.G, p.H, p.REC, p.A1, p.A2, p.A3)
    if rc != 0:
        raise CorError(f"internal_recalc({model}) rejected epoch {iy}")
    if parmod is not None:
        for k, v in enumerate(parmod):
            p.PARMOD[k] = float(v)
    return p


def model_set(internal=INTERNAL_IGRF, external=EXTERNAL_NONE, sh=None):
    m = ModelSet(internal, external, CRUSTAL_NONE, 0, None, None)
    if sh is not None:
        g = np.ascontiguousarray(sh[0], dtype=np.float64)
        h = np.ascontiguousarray(sh[1], dtype=np.float64)
        m._keep = (g, h)
        m.crustal = CRUSTAL_SH
        m.sh_nm_max = g.shape[0]
        m.sh_g = g.ctypes.data_as(DP)
        m.sh_h = h.ctypes.data_as(DP)
    return m


def scans_array(rows):
    """rows: iterable of dicts / tuples (r0, the0, fi0, the1, fi1, rig, del, rk, zn, nk1)."""
    a = np.zeros(len(rows), dtype=SCAN_DTYPE)
    for i, r in enumerate(rows):
        if isinstance(r, dict):
            for k, v in r.items():
                a[i][k] = v
        else:
            (a[i]["r0"], a[i]["the0"], a[i]["fi0"], a[i]["the1"], a[i]["fi1"], a[i]["rig"],
             a[i]["del_"], a[i]["rk"], a[i]["zn"], a[i]["nk1"]) = r
    return a


def scan_count(row):
    s = Scan(*[float(row[k]) for k in ("r0", "the0", "fi0", "the1", "fi1", "rig", "del_", "rk", "zn")],
             int(row["nk1"]), 0)
    return lib().cor_scan_count(C.byref(s))


class Plan:
    """cor_trace_plan: epoch + scans uploaded to HBM; run() traces them all.  With a list
    of epochs (params as a list of ModelsParams) it is the multi-epoch plan
    (cor_trace_plan_create_epochs): one launch over every (epoch, scan) pair, whose
    results() come back epoch-major as (n_epochs, count) and (n_epochs, n_scans)."""

    def __init__(self, models, params, scans, step_limit=1000000, tu_angle=0.01, mass=M_PROTON,
                 precision=PREC_FP64, queue_order=0, coop=0, handover_steps=0, waves_per_simd=0, block_threads=0,
                 handover_low_steps=0, handover_low_count=0):
        self.scans = np.ascontiguousarray(scans, dtype=SCAN_DTYPE)
        opts = TraceOpts(step_limit, tu_angle, mass, precision, queue_order, coop, handover_steps, waves_per_simd,
                         block_threads, handover_low_steps, handover_low_count)
        h = C.c_void_p()
        self._models = models
        self.multi = isinstance(params, (list, tuple))
        if self.multi:
            arr = (ModelsParams * len(params))(*params)
            self.n_epochs = len(params)
            check(lib().cor_trace_plan_create_epochs(C.byref(models), C.cast(arr, C.c_void_p), len(params),
                                                     self.scans.ctypes.data, len(self.scans), C.byref(opts),
                                                     C.byref(h)), "plan_create_epochs")
        else:
            self.n_epochs = 1
            check(lib().cor_trace_plan_create(C.byref(models), C.byref(params), self.scans.ctypes.data,
                                              len(self.scans), C.byref(opts), C.byref(h)), "plan_create")
        self.h = h
        self.n = lib().cor_trace_plan_count(h)

    def run(self, stream=None):
        check(lib().cor_trace_plan_run(self.h, C.c_void_p(stream) if stream else None), "plan_run")

    def sync(self):
        check(lib().cor_trace_plan_sync(self.h), "plan_sync")

    def progress(self):
        """(trajectories ended so far, run complete) of the last run, without waiting."""
        f, d = C.c_int64(), C.c_int32()
        check(lib().cor_trace_plan_progress(self.h, C.byref(f), C.byref(d)), "plan_progress")
        return f.value, bool(d.value)

    def set_epoch(self, params, stream=None):
        """Re-point the plan at another epoch (same model kinds) before the next run."""
        check(lib().cor_trace_plan_set_epoch(self.h, C.byref(self._models), C.byref(params),
                                             C.c_void_p(stream) if stream else None), "plan_set_epoch")

    def results(self, traj=True):
        cut = np.zeros(len(self.scans) * self.n_epochs, dtype=CUT_DTYPE)
        tr = np.zeros(self.n, dtype=TRAJ_DTYPE) if traj else None
        check(lib().cor_trace_plan_results(self.h, tr.ctypes.data if traj else None, cut.ctypes.data),
              "plan_results")
        if self.multi:
            cut = cut.reshape(self.n_epochs, len(self.scans))
            tr = tr.reshape(self.n_epochs, -1) if traj else None
        return tr, cut

    def stats(self):
        st = TraceStats()
        check(lib().cor_trace_plan_stats(self.h, C.byref(st)), "plan_stats")
        return st

    def close(self):
        if self.h:
            lib().cor_trace_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trace(models, params, scans, **kw):
    p = Plan(models, params, scans, **kw)
    try:
        p.run()
        return p.results()
    finally:
        p.close()


def trace_series(models, epochs, scans, in_flight=0, step_limit=1000000, tu_angle=0.01, mass=M_PROTON,
                 precision=PREC_FP64, queue_order=0, coop=0, handover_steps=0, traj=True, handover_low_steps=0,
                 handover_low_count=0):
    """The scans traced at every epoch (a list of ModelsParams): per-trajectory results
    of shape (n_epochs, count) and cutoffs of shape (n_epochs, n_scans)."""
    scans = np.ascontiguousarray(scans, dtype=SCAN_DTYPE)
    count = sum(scan_count(r) for r in scans)
    arr = (ModelsParams * len(epochs))(*epochs)
    opts = TraceOpts(step_limit, tu_angle, mass, precision, queue_order, coop, handover_steps, 0, 0,
                     handover_low_steps, handover_low_count)
    cut = np.zeros((len(epochs), len(scans)), dtype=CUT_DTYPE)
    tr = np.zeros((len(epochs), count), dtype=TRAJ_DTYPE) if traj else None
    check(lib().cor_gpu_trace_series(C.byref(models), C.cast(arr, C.c_void_p), len(epochs), scans.ctypes.data,
                                     len(scans), C.byref(opts), in_flight, tr.ctypes.data if traj else None,
                                     cut.ctypes.data), "cor_gpu_trace_series")
    return tr, cut


def trace_epochs(models, epochs, scans, step_limit=1000000, tu_angle=0.01, mass=M_PROTON, precision=PREC_FP64,
                 queue_order=0, traj=True):
    """cor_gpu_trace_epochs: the scans at every epoch in one launch; results shaped
    (n_epochs, count) and (n_epochs, n_scans)."""
    scans = np.ascontiguousarray(scans, dtype=SCAN_DTYPE)
    count = sum(scan_count(r) for r in scans)
    arr = (ModelsParams * len(epochs))(*epochs)
    opts = TraceOpts(step_limit, tu_angle, mass, precision, queue_order, 0, 0)
    cut = np.zeros((len(epochs), len(scans)), dtype=CUT_DTYPE)
    tr = np.zeros((len(epochs), count), dtype=TRAJ_DTYPE) if traj else None
    check(lib().cor_gpu_trace_epochs(C.byref(models), C.cast(arr, C.c_void_p), len(epochs), scans.ctypes.data,
                                     len(scans), C.byref(opts), tr.ctypes.data if traj else None,
                                     cut.ctypes.data), "cor_gpu_trace_epochs")
    return tr, cut


def autotune(models, params, scans, precision=PREC_MIXED, step_limit=1000000, tu_angle=0.01, mass=M_PROTON,
             coop=0, handover_steps=0, sample_scans=256, seed=0):
    """cor_trace_autotune: the launch parameters (waves per SIMD, threads per block, queue
    order) that trace a seeded sample of the scans fastest; returned as Plan keywords
    plus the sample's tracer time in ms."""
    scans = np.ascontiguousarray(scans, dtype=SCAN_DTYPE)
    base = TraceOpts(step_limit, tu_angle, mass, precision, 1, coop, handover_steps, 0, 0)
    best = TraceOpts()
    ms = C.c_double()
    check(lib().cor_trace_autotune(C.byref(models), C.byref(params), scans.ctypes.data, len(scans), C.byref(base),
                                   sample_scans, seed, C.byref(best), C.byref(ms)), "cor_trace_autotune")
    return {"waves_per_simd": best.waves_per_simd, "block_threads": best.block_threads,
            "queue_order": best.queue_order, "sample_ms": ms.value}


def trace_tracks(models, params, scans, step_limit=1000000, tu_angle=0.01, mass=M_PROTON):
    """cor_gpu_trace_tracks: the per-step tracks of every trajectory (FP64), as a list
    of TRACK_DTYPE arrays in scan order."""
    scans = np.ascontiguousarray(scans, dtype=SCAN_DTYPE)
    n = sum(scan_count(r) for r in scans)
    opts = TraceOpts(step_limit, tu_angle, mass, PREC_FP64, 0, 0, 0)
    counts = np.zeros(max(n, 1), dtype=np.int64)
    total = C.c_int64()
    check(lib().cor_gpu_trace_tracks(C.byref(models), C.byref(params), scans.ctypes.data, len(scans), C.byref(opts),
                                     counts.ctypes.data, C.byref(total), None, 0), "cor_gpu_trace_tracks")
    pts = np.zeros(max(total.value, 1), dtype=TRACK_DTYPE)
    check(lib().cor_gpu_trace_tracks(C.byref(models), C.byref(params), scans.ctypes.data, len(scans), C.byref(opts),
                                     counts.ctypes.data, C.byref(total), pts.ctypes.data, len(pts)),
          "cor_gpu_trace_tracks")
    out, k = [], 0
    for c in counts[:n]:
        out.append(pts[k:k + c])
        k += c
    return out


def field(models, params, gsm_m, precision=PREC_FP64):
    x = np.ascontiguousarray(gsm_m, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(x)
    check(lib().cor_gpu_field(C.byref(models), C.byref(params), x.ctypes.data_as(DP),
                              out.ctypes.data_as(DP), len(x), precision), "cor_gpu_field")
    return out


def geodetic_to_geocentric(lat_deg, h_km):
    r0, lat = C.c_double(), C.c_double()
    check(lib().cor_geodetic_to_geocentric(lat_deg, h_km, C.byref(r0), C.byref(lat)), "geodetic")
    return r0.value, lat.value


def direction_to_geo(lat_deg, lon_deg, zenith_deg, azimuth_deg):
    t, f = C.c_double(), C.c_double()
    check(lib().cor_direction_to_geo(lat_deg, lon_deg, zenith_deg, azimuth_deg, C.byref(t), C.byref(f)),
          "direction")
    return t.value, f.value


def refine_scans(rows, cuts, factor=10):
    """cor_refine_scan per row: the scans re-tracing [Rl - del, Ru] at del / factor."""
    out = np.zeros(len(rows), dtype=SCAN_DTYPE)
    for i, (r, c) in enumerate(zip(rows, cuts)):
        s = Scan(*[float(r[k]) for k in ("r0", "the0", "fi0", "the1", "fi1", "rig", "del_", "rk", "zn")],
                 int(r["nk1"]), 0)
        cc = Cutoff(float(c["rl"]), float(c["ru"]), float(c["rc"]), int(c["n_forbidden"]), int(c["n"]))
        f = Scan()
        check(lib().cor_refine_scan(C.byref(s), C.byref(cc), factor, C.byref(f)), "cor_refine_scan")
        out[i] = (f.r0, f.the0, f.fi0, f.the1, f.fi1, f.rig, f.del_, f.rk, f.zn, f.nk1, 0)
    return out


def grid_sites(lat_step=5.0, lon_step=5.0, h_km=20.0):
    """Global vertical-incidence grid: geodetic latitudes -90..90, longitudes 0..360-step."""
    lats = np.arange(-90.0, 90.0 + 1e-9, lat_step)
    lons = np.arange(0.0, 360.0 - 1e-9, lon_step)
    out = []
    for la in lats:
        r0, lc = geodetic_to_geocentric(float(la), h_km)
        for lo in lons:
            out.append((r0, lc, float(lo)))
    return out


def format_cutoff_lines(scan, traj):
    """The per-rigidity lines src/trajectory.c:118-122 prints for one scan."""
    out = []
    for i, t in enumerate(traj):
        if t["status"] != ALLOWED:
            continue
        rig = scan["rig"] + i * scan["del_"]
        out.append("%10.6f%15.10f%12.6f%10.3f%10.3f%10.3f%10.3f%12.6f%16.2f" % (
            rig, t["vc"], t["r"], t["td"], t["fei"], t["ast"], t["asf"], t["time"], t["length_m"] / 1000.))
    return out
