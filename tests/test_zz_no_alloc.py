"""cor_gpu_trace_noalloc, the tracer for callers that cannot allocate, through the C-ABI:
two calls in a row (the second with a smaller batch, so stale scratch contents would
show) give bit for bit the records and cutoffs of cor_gpu_trace in the per-lane mode,
in mixed precision and in FP64.  The device buffers for the results are allocated here,
by the caller, as such a caller would own them."""
import ctypes as C

import numpy as np
import pytest

import cor_amd as A

pytestmark = pytest.mark.gpu

class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), n) == 0
        return p

    def to_host(self, p, nbytes):
        out = np.empty(nbytes, np.uint8)
        assert self.h.hipMemcpy(out.ctypes.data, p, nbytes, 2) == 0     # device to host
        return out


def noalloc_trace(hip, ms, p, rows, opts):
    lib = A.lib()
    lib.cor_gpu_trace_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]
    lib.cor_gpu_trace_noalloc.restype = C.c_int
    n = sum(int(round((r["rk"] - r["rig"]) / r["del_"])) + 1 for r in rows)
    rec = A.TRAJ_DTYPE.itemsize         # one cor_traj record (80 B) per trajectory
    d_traj = hip.malloc(n * rec)
    d_cut = hip.malloc(len(rows) * A.CUT_DTYPE.itemsize)
    # fill with garbage so that anything the call fails to write shows
    assert hip.h.hipMemset(d_traj, 0x5a, n * rec) == 0
    try:
        rc = lib.cor_gpu_trace_noalloc(C.addressof(ms), C.addressof(p), rows.ctypes.data, len(rows),
                                       C.addressof(opts), d_traj, d_cut, None)
        assert rc == 0, (rc, lib.cor_gpu_last_hip_error())
        raw = hip.to_host(d_traj, n * rec)
        cut = hip.to_host(d_cut, len(rows) * A.CUT_DTYPE.itemsize).view(A.CUT_DTYPE)
    finally:
        hip.h.hipFree(d_traj)
        hip.h.hipFree(d_cut)
    return raw.view(A.TRAJ_DTYPE), cut


@pytest.mark.parametrize("precision", [A.PREC_MIXED, A.PREC_FP64], ids=["mixed", "fp64"])
def test_noalloc_equals_cor_gpu_trace(precision):
    A.init(0)
    hip = Hip()
    p = A.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    ms = A.model_set(A.INTERNAL_IGRF, A.EXTERNAL_T96)
    sites = A.grid_sites()[::211]
    for k, sub in enumerate((sites, sites[3:5])):
        rows = A.scans_array([(r0, lat, lon, lat, lon, 0.05, 0.05, 20.0, -1.0, 100) for r0, lat, lon in sub])
        limit = 100000 if precision == A.PREC_MIXED else 30000
        opts = A.TraceOpts(limit, 0.01, A.M_PROTON, precision, 1, -1, 0)
        got_t, got_c = noalloc_trace(hip, ms, p, rows, opts)
        want_t, want_c = A.trace(ms, p, rows, step_limit=limit, precision=precision, coop=-1)
        print("call", k + 1, len(rows), "scans,", len(got_t), "trajectories; cutoffs", got_c["rc"][:3])
        assert np.array_equal(got_t, want_t), int((got_t != want_t).sum())
        assert np.array_equal(got_c, want_c)
