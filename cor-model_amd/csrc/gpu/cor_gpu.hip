// cor_gpu.hip -- the C-ABI of the MI355X tracer (include/cor_gpu.h): per-epoch
// parameter blocks, scan upload, plan lifetime, kernel launches, the penumbra
// reduction and the batch field evaluation.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "cor_gpu.h"
#include "kernels.hpp"
#include "ts05_host.hpp"

using namespace cor;

namespace {

thread_local char g_hip_error[256] = "";
int g_device = -1;
hipStream_t g_stream = nullptr;
int g_num_cus = 0;
double g_wall_khz = 0.;       // wall_clock64 rate of the device

int fail_hip(hipError_t e, const char *what) {
    snprintf(g_hip_error, sizeof g_hip_error, "%s: %s", what, hipGetErrorString(e));
    return COR_E_HIP;
}

#define HIP_TRY(call)                                        \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return fail_hip(e_, #call);    \
    } while (0)

int ensure_init() {
    if (g_device >= 0) return COR_OK;
    return cor_gpu_init(0);
}

// lib/t96/t96.c:5-21,28-32,54-59: the parameter-only terms of t96_gsm, in the
// reference's arithmetic (computed here with the host libm, as the reference does).
template <class T>
void t96_params(const double *parmod, T96Params<T> &P) {
    const double pdyn = parmod[0];
    const double depr = 0.8 * parmod[1] - 13.0 * sqrt(pdyn);
    const double byimf = parmod[2], bzimf = parmod[3];
    const double bt = sqrt(byimf * byimf + bzimf * bzimf);
    double theta = 0.0;
    if (byimf != 0.0 || bzimf != 0.0) {
        theta = atan2(byimf, bzimf);
        if (theta <= 0.0) theta += 6.2831853;
    }
    const double ct = cos(theta), st = sin(theta);
    const double eps = 718.5 * sqrt(pdyn) * bt * sin(theta / 2.);
    const double xappa = pow(pdyn / 2.0, 0.14);
    const double am = 70.0 / xappa;
    const double asq = am * am;
    const double facteps = eps / 3630.7 - 1.0;
    const double factpd = sqrt(pdyn / 2.0) - 1.0;
    const double b1ampl = 0.5790 + 0.4462 * facteps;
    P.ct = (T)ct;
    P.st = (T)st;
    P.oimfy0 = (T)(0.7850 * byimf);
    P.oimfz0 = (T)(0.7850 * bzimf);
    P.xappa = (T)xappa;
    P.xappa3 = (T)(xappa * xappa * xappa);
    P.am = (T)am;
    P.asq = (T)asq;
    P.asq2 = (T)(2.0 * asq);
    P.asq4 = (T)(4.0 * asq);
    P.c548 = (T)(5.48 / xappa);
    P.rcampl = (T)(-1.162 * depr);
    P.tampl2 = (T)(22.344 + 18.50 * factpd + 2.602 * facteps);
    P.tampl3 = (T)(6.903 + 5.287 * factpd);
    P.b1ampl = (T)b1ampl;
    P.b2ampl = (T)(20.0 * b1ampl);
    P.rimfampl = (T)(0.7850 * bt);
}

int model_kinds(const cor_model_set *m, int *intk, int *ext, int *cru) {
    if (!m) return COR_E_INVALID;
    if (m->internal != COR_INTERNAL_NONE && m->internal != COR_INTERNAL_IGRF) return COR_E_INVALID;
    if (m->external < COR_EXTERNAL_NONE || m->external > COR_EXTERNAL_TS05) return COR_E_INVALID;
    if (m->crustal != COR_CRUSTAL_NONE && m->crustal != COR_CRUSTAL_SH) return COR_E_INVALID;
    if (m->crustal == COR_CRUSTAL_SH &&
        (m->sh_nm_max < 2 || m->sh_nm_max > COR_SH_NM_MAX || !m->sh_g || !m->sh_h))
        return COR_E_INVALID;
    *intk = m->internal == COR_INTERNAL_IGRF ? INT_IGRF : INT_NONE;
    *ext = m->external == COR_EXTERNAL_T96 ? EXT_T96 : (m->external == COR_EXTERNAL_TS05 ? EXT_TS05 : EXT_NONE);
    *cru = m->crustal == COR_CRUSTAL_SH ? CRU_SH : CRU_NONE;
    return COR_OK;
}

void build_epoch(const cor_model_set *m, const ModelsParams *p, EpochDev &E) {
    memset(&E, 0, sizeof E);
    for (int k = 0; k < NCOEF; k++) {
        E.G[k] = p->G[k];
        E.H[k] = p->H[k];
        E.REC[k] = p->REC[k];
        E.Gf[k] = (float)p->G[k];
        E.Hf[k] = (float)p->H[k];
        E.RECf[k] = (float)p->REC[k];
    }
    for (int k = 0; k < 3; k++) {
        E.A[k] = p->A1[k];
        E.A[3 + k] = p->A2[k];
        E.A[6 + k] = p->A3[k];
    }
    for (int k = 0; k < 9; k++) E.Af[k] = (float)E.A[k];
    for (int k = 0; k < 10; k++) E.parmod[k] = p->PARMOD[k];
    if (m->external == COR_EXTERNAL_TS05) {
        ts05_build_params(p->PARMOD, E.ts05d);
        ts05_params_to_float(E.ts05d, E.ts05f);
    }
    if (m->external == COR_EXTERNAL_T96) {
        t96_params<double>(p->PARMOD, E.t96d);
        t96_params<float>(p->PARMOD, E.t96f);
    }
}

// The SH tables of a model set in device memory, [nm][nm] as the caller gives them:
// FP64 G and H, then FP32 G and H, in one block (reused when large enough), pointed to
// by the epoch.  Synchronous: no kernel may be reading the block.
hipError_t upload_sh(const cor_model_set *m, void **buf, size_t *cap, EpochDev &E) {
    if (m->crustal != COR_CRUSTAL_SH) return hipSuccess;
    const int nm = m->sh_nm_max;
    const size_t nn = (size_t)nm * (size_t)nm;
    const size_t bytes = nn * (2 * sizeof(double) + 2 * sizeof(float));
    if (*cap < bytes) {
        (void)hipFree(*buf);
        *buf = nullptr;
        *cap = 0;
        const hipError_t e = hipMalloc(buf, bytes);
        if (e != hipSuccess) return e;
        *cap = bytes;
    }
    std::vector<double> d(2 * nn);
    std::vector<float> f(2 * nn);
    for (size_t k = 0; k < nn; k++) {
        d[k] = m->sh_g[k];
        d[nn + k] = m->sh_h[k];
        f[k] = (float)m->sh_g[k];
        f[nn + k] = (float)m->sh_h[k];
    }
    double *dd = (double *)*buf;
    float *ff = (float *)(dd + 2 * nn);
    hipError_t e = hipMemcpy(dd, d.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ff, f.data(), sizeof(float) * 2 * nn, hipMemcpyHostToDevice);
    E.sh_nm = nm;
    E.shG = dd;
    E.shH = dd + nn;
    E.shGf = ff;
    E.shHf = ff + nn;
    return e;
}

// The SH table pointers of an epoch block built on the host, for tables already on the
// device (the model set of a series is fixed, so its tables are uploaded once).
void point_sh(const cor_model_set *m, void *d_sh, EpochDev &E) {
    if (m->crustal != COR_CRUSTAL_SH || !d_sh) return;
    const size_t nn = (size_t)m->sh_nm_max * (size_t)m->sh_nm_max;
    double *dd = (double *)d_sh;
    float *ff = (float *)(dd + 2 * nn);
    E.sh_nm = m->sh_nm_max;
    E.shG = dd;
    E.shH = dd + nn;
    E.shGf = ff;
    E.shHf = ff + nn;
}

// Station position and arrival-direction trig as src/trajectory.c:18-31 and
// src/utils.c:28-41 compute them (host libm, like the reference).
void build_scan(const cor_scan &s, int64_t first, int32_t n, ScanDev &d) {
    const double r = s.r0 * DEF_RE;
    const double th = (90. - s.the0) * DEF_TO_RAD;
    const double ph = s.fi0 * DEF_TO_RAD;
    const double sq = r * sin(th);
    d.pos0[0] = sq * cos(ph);
    d.pos0[1] = sq * sin(ph);
    d.pos0[2] = r * cos(th);
    const double th1 = (90. - s.the1) * DEF_TO_RAD;
    const double ph1 = s.fi1 * DEF_TO_RAD;
    d.st1 = sin(th1);
    d.ct1 = cos(th1);
    d.sp1 = sin(ph1);
    d.cp1 = cos(ph1);
    d.rig = s.rig;
    d.del = s.del;
    d.rk = s.rk;
    d.zn = s.zn;
    d.nk1 = s.nk1;
    d.n = n;
    d.first = first;
    d.epoch = 0;
    d.pad = 0;
}

// Per scan: src/trajectory.c:49-54,114-115,133-145,156-160.  rmx1/rmx2/rmi are the
// extreme forbidden/allowed rigidities; since rig + i*del grows with i they are taken
// at the extreme indices and then evaluated with the reference's double expressions.
struct ScanStat {
    unsigned long long steps, attempts;
};

__global__ void __launch_bounds__(64) penumbra_kernel(const ScanDev *__restrict__ scans,
                                                      const cor_traj *__restrict__ rec,
                                                      cor_cutoff *__restrict__ cut,
                                                      ScanStat *__restrict__ stat) {
    const ScanDev &S = scans[blockIdx.x];
    const int lane = threadIdx.x;
    int nza = 0, imax1 = -1, imax2 = -1, imin = 0x7fffffff;
    unsigned long long st = 0, at = 0;
    for (int i = lane; i < S.n; i += 64) {
        const int64_t g = S.first + i;
        const uint4 h = *reinterpret_cast<const uint4 *>(rec + g);   // status, steps, attempts, handover
        const int s = (int)h.x;
        st += h.y;
        at += h.z;
        if (s == ST_ALLOWED) {
            imin = min(imin, i);
        } else {
            nza++;
            if (s == ST_FORBIDDEN_REENTRY) imax1 = max(imax1, i); else imax2 = max(imax2, i);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        nza += __shfl_xor(nza, off);
        imax1 = max(imax1, __shfl_xor(imax1, off));
        imax2 = max(imax2, __shfl_xor(imax2, off));
        imin = min(imin, __shfl_xor(imin, off));
        st += __shfl_xor(st, off);
        at += __shfl_xor(at, off);
    }
    if (lane == 0) {
        const double rmx1 = imax1 >= 0 ? fmax(0., S.rig + imax1 * S.del) : 0.;
        const double rmx2 = imax2 >= 0 ? fmax(0., S.rig + imax2 * S.del) : 0.;
        const double rmi = imin != 0x7fffffff ? fmin(S.rk, S.rig + imin * S.del) : S.rk;
        double rmx = fmax(rmx1, rmx2);
        const double zan = nza - (rmi - S.rig) / S.del;
        const double rms = rmi + zan * S.del;
        rmx = rmx + S.del;
        cor_cutoff c;
        c.rl = rmi;
        c.ru = rmx;
        c.rc = rms;
        c.n_forbidden = nza;
        c.n = S.n;
        cut[blockIdx.x] = c;
        stat[blockIdx.x].steps = st;
        stat[blockIdx.x].attempts = at;
    }
}

__global__ void reset_head_kernel(unsigned long long *head, CoopCtl *coop, RunStats *st) {
    *head = 0ull;
    st->wave_iters = 0ull;
    st->lane_iters = 0ull;
    st->coop_steps = 0ull;
    st->wave_iters_open = 0ull;
    st->lane_iters_open = 0ull;
    st->t_start = ~0ull;
    st->t_drain = 0ull;
    st->t_lane_end = 0ull;
    st->finished = 0ull;
    if (coop) {
        coop->res = 0ull;
        coop->head = 0ull;
        coop->running = 0;
        coop->reparks = 0ull;
        coop->done = 0ull;
        coop->converting = 0;
        coop->feeders = 0;
        coop->lres = 0ull;
        coop->lhead = 0ull;
        coop->run = coop->run + 1u;   // this run's id for the park records
    }
}

// Device memory of cor_gpu_trace_noalloc: the epoch block, the scan descriptors, the
// queue head, the run counters and the per-scan step sums of one call, in a file-scope
// array that every GPU holds once for the life of the process.  Host code uses the
// device address each GPU reports for it (noalloc_scratch[device], taken in
// cor_gpu_init), never the array's own name, which is a host-side stand-in.
struct NoAllocScratch {
    EpochDev epoch;
    ScanDev scans[COR_NOALLOC_MAX_SCANS];
    ScanStat stat[COR_NOALLOC_MAX_SCANS];
    unsigned long long head;
    RunStats run;
};
static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];
constexpr int NOALLOC_MAX_DEVICES = 64;
NoAllocScratch *g_noalloc[NOALLOC_MAX_DEVICES] = {};
// host-side staging of the same blocks, one per device (static storage: the call copies
// from it and returns only when the copies are done)
struct NoAllocHost {
    EpochDev epoch;
    ScanDev scans[COR_NOALLOC_MAX_SCANS];
};
NoAllocHost g_noalloc_host[NOALLOC_MAX_DEVICES];

// Accepted per-lane steps before a trajectory moves to the cooperative mode.  On the
// 5-degree IGRF + T96 grid the median trajectory runs 435 steps, 0.8 % run past 65536 and
// hold 40 % of all steps; about 1 in 3000 reaches the 10^6-step limit.  A per-lane step
// costs about a quarter of a cooperative one per trajectory, but a cooperative step has
// a quarter of its latency, which the trapped trajectories need.  Mixed precision, with
// the per-lane compaction keeping the drain's wavefronts full, full grid on one host:
// 15.5-15.8 s (65536), 15.3 s (81920), 15.4 s (90112), 14.8-15.0 s (98304), 15.0 s
// (106496), 15.3 s (114688), 15.5 s (131072); the 64-site x 193-direction workload
// 56.1 s + 61.1 s (98304) against 60.0 s + 66.8 s (65536).  FP64 (its cooperative tail
// is opt-in): 68.2 s (65536) against 68.8 s (98304).
constexpr uint32_t COOP_HANDOVER_DEFAULT_MIXED = 98304u;
constexpr uint32_t COOP_HANDOVER_DEFAULT_FP64 = 65536u;
// An earlier handover for the lowest rigidities of each scan, which hold most of the
// trajectories trapped until the step limit (on the 5-degree grid 1468 of 1852 lie among
// the first 20 of 2000), is off by default: it moves more work into the cooperative mode
// than it saves on the trapped trajectories' path.  Full grid, measured (before the
// per-lane compaction) against 15.9 s with every trajectory at 65536: 16.8 s (32768 for
// the lowest 20), 16.8 s (16384, 20), 16.6 s (32768, 50), 16.8 s (98304 for the others,
// 32768 for the lowest 20).
constexpr uint32_t COOP_HANDOVER_LOW_DEFAULT = 32768u;
constexpr int32_t COOP_HANDOVER_LOW_COUNT_DEFAULT = 0;
// Parked trajectories waiting for workers, beyond eight per converting wavefront, that
// make a feeding wavefront turn worker; and the share of wavefronts that keep feeding
// until the queue is drained (three quarters: 16.2 s on the full grid, against 17.9 s
// with one eighth, whose queue then drained at 10.9 s instead of 3.0 s).
constexpr int32_t COOP_CONVERT_BACKLOG = 64;

// the mixed T96 evaluator reads its interconnection series from the lane tables
bool needs_coop_tables(int precision, int ext) { return precision != COR_PREC_FP64 && ext == EXT_T96; }

hipError_t build_coop_tables(EpochDev *d_epoch, hipStream_t s) {
    void *args[] = {&d_epoch};
    return hipLaunchKernel(coop_build_kernel(), dim3(1), dim3(64), args, 0, s);
}

}  // namespace

struct cor_trace_plan {
    int intk, ext, cru, prec;
    int32_t n_epochs;            // > 1: the multi-epoch tracer (d_epoch holds that many blocks)
    const void *kernel;
    EpochDev *d_epoch;
    ScanDev *d_scans;
    std::vector<ScanDev> h_scans;
    int32_t n_scans;
    int64_t n_items;
    uint32_t *d_order;
    unsigned long long *d_head;
    cor_traj *out = nullptr;        // result records (one per trajectory)
    cor_cutoff *d_cut;
    ScanStat *d_stat;
    uint64_t step_limit;
    double tu_angle, cos_tu, mass;
    int grid, block;
    int coop_on;                 // 0: per-lane mode only
    uint32_t handover_steps;     // per-lane accepted steps before the cooperative mode
    uint32_t handover_low_steps; // ... for the lowest handover_low_count rigidities of a scan
    int32_t handover_low_count;
    size_t park_slots;
    size_t pool_slots;           // per-lane pool records, after the park slots
    CoopCtl *d_coop;
    ParkDev *d_park;
    RunStats *d_run;
    hipEvent_t ev[3];
    bool ran;
    void *d_sh;                  // SH tables of the crustal model (upload_sh)
    size_t sh_cap;
    cor_track_point *d_trk;      // track recording of the next runs (NULL: off)
    int64_t *d_trk_off;
    EpochDev *h_epoch;           // pinned staging of set_epoch uploads
    hipEvent_t ev_up;            // the last set_epoch upload has left h_epoch
    bool up_pending;
    // the epoch graph (cor_gpu_trace_series): upload of h_epoch, lane tables, reset,
    // trace, penumbra and the copies of every result into h_res, captured once and
    // replayed per epoch
    hipStream_t s_progress;      // cor_trace_plan_progress reads the run's counter on it
    unsigned long long *h_progress;
    hipGraph_t graph;
    hipGraphExec_t gexec;
    unsigned char *h_res;        // pinned: CoopCtl, cut[n_scans], then the trajectory arrays
    hipEvent_t ev_graph;         // the last replay is complete
};

extern "C" {

int cor_gpu_init(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        snprintf(g_hip_error, sizeof g_hip_error, "no HIP device");
        return COR_E_NO_DEVICE;
    }
    if (device < 0 || device >= n) return COR_E_INVALID;
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        snprintf(g_hip_error, sizeof g_hip_error, "device %d is %s; this library is built for gfx950",
                 device, prop.gcnArchName);
        return COR_E_NO_DEVICE;
    }
    g_num_cus = prop.multiProcessorCount;
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    g_wall_khz = khz > 0 ? (double)khz : 0.;
    if (!g_stream) HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    // this device's copy of the allocation-free entry point's scratch array
    if (device < NOALLOC_MAX_DEVICES && !g_noalloc[device]) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_noalloc[device] = (NoAllocScratch *)ptr;
    }
    g_device = device;
    return COR_OK;
}

const char *cor_gpu_error_string(int code) {
    switch (code) {
        case COR_OK: return "ok";
        case COR_E_INVALID: return "invalid argument";
        case COR_E_NO_DEVICE: return "no usable gfx950 device";
        case COR_E_HIP: return "HIP runtime error";
        case COR_E_UNSUPPORTED: return "model combination not compiled into the tracer";
        case COR_E_NOMEM: return "out of memory";
        default: return "unknown error";
    }
}

const char *cor_gpu_last_hip_error(void) { return g_hip_error; }

void cor_set_error(const char *msg) { snprintf(g_hip_error, sizeof g_hip_error, "%s", msg ? msg : ""); }

int cor_gpu_synchronize(void) {
    int rc = ensure_init();
    if (rc) return rc;
    HIP_TRY(hipDeviceSynchronize());
    return COR_OK;
}

int64_t cor_scan_count(const cor_scan *s) {
    if (!s || !(s->del > 0.) || !isfinite(s->rig) || !isfinite(s->rk)) return COR_E_INVALID;
    const double q = round((s->rk - s->rig) / s->del);
    if (!(q < 2.0e9)) return COR_E_INVALID;
    const int size = (int)q;
    return size < 0 ? 0 : (int64_t)size + 1;
}

void cor_trace_plan_destroy(cor_trace_plan *p) {
    if (!p) return;
    (void)hipFree(p->d_epoch);
    (void)hipFree(p->d_scans);
    (void)hipFree(p->d_order);
    (void)hipFree(p->d_head);
    (void)hipFree(p->out);
    (void)hipFree(p->d_cut);
    (void)hipFree(p->d_stat);
    (void)hipFree(p->d_coop);
    (void)hipFree(p->d_park);
    (void)hipFree(p->d_run);
    (void)hipFree(p->d_trk);
    (void)hipFree(p->d_sh);
    (void)hipFree(p->d_trk_off);
    for (int k = 0; k < 3; k++)
        if (p->ev[k]) (void)hipEventDestroy(p->ev[k]);
    if (p->ev_graph) {
        (void)hipEventSynchronize(p->ev_graph);
        (void)hipEventDestroy(p->ev_graph);
    }
    if (p->s_progress) {
        (void)hipStreamSynchronize(p->s_progress);
        (void)hipStreamDestroy(p->s_progress);
    }
    if (p->h_progress) (void)hipHostFree(p->h_progress);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->h_res) (void)hipHostFree(p->h_res);
    if (p->up_pending) (void)hipEventSynchronize(p->ev_up);
    if (p->ev_up) (void)hipEventDestroy(p->ev_up);
    if (p->h_epoch) (void)hipHostFree(p->h_epoch);
    delete p;
}

}  // extern "C"

namespace {
// A plan over n_epochs epochs of the same scans.  One epoch: the tracer kernels with the
// epoch staged per workgroup.  Several: the multi-epoch tracer, whose scan list is every
// (epoch, scan) pair, epoch-major, each naming its epoch block.
int plan_create_impl(const cor_model_set *models, const ModelsParams *epochs, int32_t n_epochs,
                     const cor_scan *scans, int32_t n_scans_in, const cor_trace_opts *opts, cor_trace_plan **out) {
    if (!out) return COR_E_INVALID;
    *out = nullptr;
    if (!epochs || n_epochs < 1 || !opts || n_scans_in < 1 || !scans) return COR_E_INVALID;
    if (opts->precision != COR_PREC_FP64 && opts->precision != COR_PREC_MIXED) return COR_E_INVALID;
    if (!(opts->particle_mass > 0.) || !(opts->tu_angle > 0.)) return COR_E_INVALID;
    if ((int64_t)n_epochs * n_scans_in > 0x7fffffffll) return COR_E_INVALID;
    if (opts->waves_per_simd < 0 || (opts->block_threads != 0 && opts->block_threads != 64 &&
                                     opts->block_threads != 256))
        return COR_E_INVALID;
    int rc = ensure_init();
    if (rc) return rc;
    int intk, ext, cru;
    if ((rc = model_kinds(models, &intk, &ext, &cru))) return rc;
    const bool multi = n_epochs > 1;
    // FP64 with the cooperative tail (opt-in) has a kernel of its own; model sets the
    // cooperative evaluator does not cover run per-lane
    const bool fp64_coop = !multi && opts->precision == COR_PREC_FP64 && opts->coop > 0 &&
                           trace_kernel_fp64_coop(intk, ext, cru);
    const void *kern;
    if (multi) {
        kern = opts->precision == COR_PREC_FP64 ? trace_kernel_epochs_fp64(intk, ext, cru)
                                                : trace_kernel_epochs_mixed(intk, ext, cru);
    } else {
        kern = opts->precision == COR_PREC_FP64 ? (fp64_coop ? trace_kernel_fp64_coop(intk, ext, cru)
                                                             : trace_kernel_fp64(intk, ext, cru))
                                                : trace_kernel_mixed(intk, ext, cru);
    }
    if (!kern) return COR_E_UNSUPPORTED;
    const int32_t n_scans = n_epochs * n_scans_in;

    cor_trace_plan *p = new cor_trace_plan();   // value-initialised: every pointer null
    p->intk = intk;
    p->ext = ext;
    p->cru = cru;
    p->prec = opts->precision;
    p->kernel = kern;
    p->n_scans = n_scans;
    p->n_epochs = n_epochs;
    p->step_limit = opts->step_limit;
    p->tu_angle = opts->tu_angle;
    p->cos_tu = cos(opts->tu_angle);
    p->mass = opts->particle_mass;
    p->h_scans.resize(n_scans);
    int64_t total = 0;
    for (int32_t e = 0; e < n_epochs; e++) {
        for (int32_t s = 0; s < n_scans_in; s++) {
            const int64_t n = cor_scan_count(&scans[s]);
            if (n < 0 || scans[s].nk1 <= 0 || !isfinite(scans[s].zn) || scans[s].zn == 0.) {
                delete p;
                return COR_E_INVALID;
            }
            ScanDev &d = p->h_scans[(size_t)e * n_scans_in + s];
            build_scan(scans[s], total, (int32_t)n, d);
            d.epoch = e;
            total += n;
        }
    }
    if (total >= (int64_t)0xffffffffll) {
        delete p;
        return COR_E_INVALID;
    }
    p->n_items = total;

    std::vector<EpochDev> E(n_epochs);
    for (int32_t e = 0; e < n_epochs; e++) build_epoch(models, &epochs[e], E[e]);
    const size_t nt = (size_t)std::max<int64_t>(total, 1);
    {
        const hipError_t e = upload_sh(models, &p->d_sh, &p->sh_cap, E[0]);
        if (e != hipSuccess) {
            cor_trace_plan_destroy(p);
            return fail_hip(e, "upload_sh");
        }
        for (int32_t k = 1; k < n_epochs; k++) point_sh(models, p->d_sh, E[k]);
    }
#define PLAN_TRY(call)                                    \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) {                           \
            cor_trace_plan_destroy(p);                    \
            return fail_hip(e_, #call);                   \
        }                                                 \
    } while (0)
    PLAN_TRY(hipMalloc(&p->d_epoch, sizeof(EpochDev) * n_epochs));
    PLAN_TRY(hipMemcpy(p->d_epoch, E.data(), sizeof(EpochDev) * n_epochs, hipMemcpyHostToDevice));
    PLAN_TRY(hipMalloc(&p->d_scans, sizeof(ScanDev) * n_scans));
    PLAN_TRY(hipMemcpy(p->d_scans, p->h_scans.data(), sizeof(ScanDev) * n_scans, hipMemcpyHostToDevice));
    PLAN_TRY(hipMalloc(&p->d_head, sizeof(unsigned long long)));
    PLAN_TRY(hipMalloc(&p->out, sizeof(cor_traj) * nt));
    PLAN_TRY(hipMemset(p->out, 0, sizeof(cor_traj) * nt));
    PLAN_TRY(hipMalloc(&p->d_cut, sizeof(cor_cutoff) * n_scans));
    PLAN_TRY(hipMalloc(&p->d_stat, sizeof(ScanStat) * n_scans));
    PLAN_TRY(hipMalloc(&p->d_run, sizeof(RunStats)));

    if (opts->queue_order == 1 && total > 0) {
        // longest-expected first: trajectories near the lower end of each scan (low
        // rigidity, often trapped) are issued before the fast-escaping high ones,
        // interleaved across scans so every scan makes progress.
        std::vector<uint32_t> order(total);
        int32_t maxn = 0;
        for (const ScanDev &d : p->h_scans) maxn = std::max(maxn, d.n);
        int64_t k = 0;
        for (int32_t i = 0; i < maxn; i++)
            for (const ScanDev &d : p->h_scans)
                if (i < d.n) order[k++] = (uint32_t)(d.first + i);
        PLAN_TRY(hipMalloc(&p->d_order, sizeof(uint32_t) * total));
        PLAN_TRY(hipMemcpy(p->d_order, order.data(), sizeof(uint32_t) * total, hipMemcpyHostToDevice));
    }

    // the cooperative evaluator covers IGRF and T96 (TS05 and SH traces stay per-lane, and
    // so does the multi-epoch tracer)
    p->coop_on = !multi && (p->prec == COR_PREC_MIXED ? opts->coop >= 0 : opts->coop > 0) && ext != EXT_TS05 &&
                 cru == CRU_NONE;
    p->handover_steps = opts->handover_steps > 0 ? (uint32_t)opts->handover_steps
                        : p->prec == COR_PREC_FP64 ? COOP_HANDOVER_DEFAULT_FP64
                                                   : COOP_HANDOVER_DEFAULT_MIXED;
    // the lowest rigidities hand over no later than the others
    p->handover_low_steps = opts->handover_low_steps > 0   ? (uint32_t)opts->handover_low_steps
                            : opts->handover_low_steps < 0 ? p->handover_steps
                                                           : std::min(COOP_HANDOVER_LOW_DEFAULT, p->handover_steps);
    p->handover_low_count = opts->handover_low_count > 0   ? opts->handover_low_count
                            : opts->handover_low_count < 0 ? 0
                                                           : COOP_HANDOVER_LOW_COUNT_DEFAULT;

    // persistent grid: as many resident blocks as the occupancy allows, but no more
    // blocks than are needed to give every trajectory a lane at the start (or, with the
    // cooperative tail, a wavefront to finish in)
    // (a 256-thread block is four wavefronts, one per SIMD, so blocks per CU = resident
    // wavefronts per SIMD; the caller may ask for fewer, and for a block size)
    p->block = 256;
    int per_cu = 0;
    PLAN_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, p->block, 0));
    if (per_cu < 1) per_cu = 1;
    if (opts->waves_per_simd > 0) per_cu = std::min(per_cu, (int)opts->waves_per_simd);
    int64_t grid = (int64_t)per_cu * g_num_cus;
    const int64_t need = (total + 63) / 64;   // one wavefront per 64 trajectories
    if (opts->block_threads == 64 || (opts->block_threads == 0 && need * 64 < (int64_t)p->block * grid)) {
        // single-wavefront blocks spread over all CUs (by default for small launches)
        p->block = 64;
        grid = (int64_t)per_cu * 4 * g_num_cus;
        if (need * 64 < (int64_t)p->block * grid) grid = std::min<int64_t>(p->coop_on ? total : need, grid);
    }
    p->grid = (int)std::max<int64_t>(grid, 1);
    if (p->coop_on || needs_coop_tables(p->prec, ext)) {
        // lane tables of the epoch: the cooperative evaluator's, whose T96
        // interconnection series the per-lane mixed path also reads
        PLAN_TRY(build_coop_tables(p->d_epoch, g_stream));
        PLAN_TRY(hipStreamSynchronize(g_stream));
    }
    if (p->coop_on) {
        // the handover counters and the park slots: every trajectory parks at most once
        // from the per-lane mode, and workers hand back at most one per lane of the grid
        p->park_slots = (size_t)total + (size_t)p->grid * p->block;
        // the per-lane pool after them: four records per lane of the grid per run (the
        // compaction stops for the rest of a run that has used them up)
        p->pool_slots = (size_t)p->grid * p->block * 4;
        PLAN_TRY(hipMalloc(&p->d_coop, sizeof(CoopCtl)));
        PLAN_TRY(hipMalloc(&p->d_park, sizeof(ParkDev) * (p->park_slots + p->pool_slots)));
        PLAN_TRY(hipMemsetAsync(p->d_coop, 0, sizeof(CoopCtl), g_stream));
        PLAN_TRY(hipMemsetAsync(p->d_park, 0, sizeof(ParkDev) * (p->park_slots + p->pool_slots), g_stream));
        PLAN_TRY(hipStreamSynchronize(g_stream));
    }
    for (int k = 0; k < 3; k++) PLAN_TRY(hipEventCreate(&p->ev[k]));
#undef PLAN_TRY
    *out = p;
    return COR_OK;
}
}  // namespace

extern "C" {

int cor_trace_plan_create(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                          int32_t n_scans, const cor_trace_opts *opts, cor_trace_plan **out) {
    return plan_create_impl(models, epoch, 1, scans, n_scans, opts, out);
}

int cor_trace_plan_create_epochs(const cor_model_set *models, const ModelsParams *epochs, int32_t n_epochs,
                                 const cor_scan *scans, int32_t n_scans, const cor_trace_opts *opts,
                                 cor_trace_plan **out) {
    return plan_create_impl(models, epochs, n_epochs, scans, n_scans, opts, out);
}

int cor_gpu_trace_epochs(const cor_model_set *models, const ModelsParams *epochs, int32_t n_epochs,
                         const cor_scan *scans, int32_t n_scans, const cor_trace_opts *opts, cor_traj *traj,
                         cor_cutoff *cut) {
    cor_trace_plan *p = nullptr;
    int rc = plan_create_impl(models, epochs, n_epochs, scans, n_scans, opts, &p);
    if (rc) return rc;
    rc = cor_trace_plan_run(p, nullptr);
    if (!rc) rc = cor_trace_plan_results(p, traj, cut);
    cor_trace_plan_destroy(p);
    return rc;
}

int cor_trace_autotune(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                       int32_t n_scans, const cor_trace_opts *base, int32_t sample_scans, uint64_t seed,
                       cor_trace_opts *best, double *best_ms) {
    if (!models || !epoch || !scans || n_scans < 1 || !base || !best || sample_scans < 1) return COR_E_INVALID;
    int rc = ensure_init();
    if (rc) return rc;
    int intk, ext, cru;
    if ((rc = model_kinds(models, &intk, &ext, &cru))) return rc;
    // the sample: a seeded choice of scans (partial Fisher-Yates with a 64-bit LCG)
    std::vector<int32_t> idx(n_scans);
    for (int32_t k = 0; k < n_scans; k++) idx[k] = k;
    const int32_t ns = std::min(sample_scans, n_scans);
    uint64_t x = seed * 6364136223846793005ull + 1442695040888963407ull;
    for (int32_t k = 0; k < ns; k++) {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        const int32_t j = k + (int32_t)((x >> 33) % (uint64_t)(n_scans - k));
        std::swap(idx[k], idx[j]);
    }
    std::sort(idx.begin(), idx.begin() + ns);
    std::vector<cor_scan> sample(ns);
    for (int32_t k = 0; k < ns; k++) sample[k] = scans[idx[k]];
    // the register limit of this workload's kernel
    const void *kern = base->precision == COR_PREC_FP64 ? trace_kernel_fp64(intk, ext, cru)
                                                        : trace_kernel_mixed(intk, ext, cru);
    if (!kern) return COR_E_UNSUPPORTED;
    int wps_max = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&wps_max, kern, 256, 0));
    wps_max = std::max(1, std::min(wps_max, 8));
    double t_best = 0.;
    bool have = false;
    for (int wps = 1; wps <= wps_max; wps++) {
        for (int bt : {256, 64}) {
            for (int qo : {1, 0}) {
                cor_trace_opts o = *base;
                o.step_limit = std::min<uint64_t>(base->step_limit, 100000ull);
                o.waves_per_simd = wps;
                o.block_threads = bt;
                o.queue_order = qo;
                cor_trace_plan *p = nullptr;
                rc = cor_trace_plan_create(models, epoch, sample.data(), ns, &o, &p);
                if (rc) return rc;
                cor_trace_stats st;
                rc = cor_trace_plan_run(p, nullptr);
                if (!rc) rc = cor_trace_plan_stats(p, &st);
                cor_trace_plan_destroy(p);
                if (rc) return rc;
                if (!have || st.kernel_ms < t_best) {
                    have = true;
                    t_best = st.kernel_ms;
                    *best = *base;
                    best->waves_per_simd = wps;
                    best->block_threads = bt;
                    best->queue_order = qo;
                }
            }
        }
    }
    if (best_ms) *best_ms = t_best;
    return COR_OK;
}

int64_t cor_trace_plan_count(const cor_trace_plan *p) { return p ? p->n_items : COR_E_INVALID; }

}  // extern "C"

namespace {
// The launches of one run on stream s: queue and counter reset, the tracer, the
// penumbra reduction; with timing, the plan's events around them.  No memset nodes, so
// the sequence can be captured into a graph and replayed.
hipError_t enqueue_run(cor_trace_plan *p, hipStream_t s, bool timing) {
    TraceArgs a;
    memset(&a, 0, sizeof a);
    a.epoch = p->d_epoch;
    a.scans = p->d_scans;
    a.n_scans = p->n_scans;
    a.order = p->d_order;
    a.id_base = 0;
    a.n_items = p->n_items;
    a.head = p->d_head;
    a.step_limit = p->step_limit;
    a.tu_angle = p->tu_angle;
    a.cos_tu = p->cos_tu;
    a.mass = p->mass;
    a.inv_mass = 1.0 / p->mass;
    a.out = p->out;
    a.coop_on = p->coop_on;
    a.coop = p->d_coop;
    a.park = p->d_park;
    a.stats = p->d_run;
    a.repark_limit = (long long)p->grid * p->block;
    a.pool = p->d_park ? p->d_park + p->park_slots : nullptr;
    a.pool_cap = p->d_park ? (unsigned long long)p->pool_slots : 0ull;
    a.trk = p->d_trk;
    a.trk_off = p->d_trk_off;
    a.handover_steps = p->handover_steps;
    a.handover_low_steps = p->handover_low_steps;
    a.handover_low_count = p->handover_low_count;
    a.convert_backlog = COOP_CONVERT_BACKLOG;
    a.min_feeders = std::max(1, (int)((int64_t)p->grid * (p->block / 64) * 3 / 4));
    {
        // longer than any trajectory can run: 60 s + 1 ms per allowed step
        const double secs = 60.0 + 1e-3 * (double)p->step_limit;
        const double ticks = secs * (g_wall_khz > 0. ? g_wall_khz * 1e3 : 1e8);
        a.idle_guard = ticks < 1e18 ? (unsigned long long)ticks : 1000000000000000000ull;
    }
    hipError_t e;
    reset_head_kernel<<<1, 1, 0, s>>>(p->d_head, p->d_coop, p->d_run);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (timing && (e = hipEventRecord(p->ev[0], s)) != hipSuccess) return e;
    if (p->n_items > 0) {
        void *args[] = {&a};
        if ((e = hipLaunchKernel(p->kernel, dim3(p->grid), dim3(p->block), args, 0, s)) != hipSuccess) return e;
    }
    if (timing && (e = hipEventRecord(p->ev[1], s)) != hipSuccess) return e;
    penumbra_kernel<<<p->n_scans, 64, 0, s>>>(p->d_scans, p->out,
                                              p->d_cut, p->d_stat);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (timing && (e = hipEventRecord(p->ev[2], s)) != hipSuccess) return e;
    return hipSuccess;
}
}  // namespace

extern "C" {

int cor_trace_plan_run(cor_trace_plan *p, void *stream) {
    if (!p) return COR_E_INVALID;
    hipStream_t s = stream ? (hipStream_t)stream : g_stream;
    HIP_TRY(enqueue_run(p, s, true));
    p->ran = true;
    return COR_OK;
}

int cor_trace_plan_progress(cor_trace_plan *p, int64_t *finished, int32_t *done) {
    if (!p || !finished || !done || !p->ran) return COR_E_INVALID;
    const hipError_t q = hipEventQuery(p->ev[2]);
    if (q != hipSuccess && q != hipErrorNotReady) return fail_hip(q, "hipEventQuery");
    *done = q == hipSuccess;
    // a snapshot of the counter the running kernel advances, on a stream of its own
    if (!p->s_progress) HIP_TRY(hipStreamCreateWithFlags(&p->s_progress, hipStreamNonBlocking));
    if (!p->h_progress) HIP_TRY(hipHostMalloc((void **)&p->h_progress, sizeof(unsigned long long), hipHostMallocDefault));
    HIP_TRY(hipMemcpyAsync(p->h_progress, &p->d_run->finished, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           p->s_progress));
    HIP_TRY(hipStreamSynchronize(p->s_progress));
    *finished = (int64_t)*p->h_progress;
    return COR_OK;
}

int cor_trace_plan_sync(cor_trace_plan *p) {
    if (!p) return COR_E_INVALID;
    HIP_TRY(hipEventSynchronize(p->ev[2]));
    return COR_OK;
}

int cor_trace_plan_results(cor_trace_plan *p, cor_traj *traj, cor_cutoff *cut) {
    if (!p || !p->ran) return COR_E_INVALID;
    HIP_TRY(hipEventSynchronize(p->ev[2]));
    if (p->coop_on) {
        // every trajectory is counted once when it ends, in either mode
        CoopCtl c;
        HIP_TRY(hipMemcpy(&c, p->d_coop, sizeof c, hipMemcpyDeviceToHost));
        if (c.done != (unsigned long long)p->n_items) {
            snprintf(g_hip_error, sizeof g_hip_error, "cooperative tail finished %llu of %lld trajectories",
                     c.done, (long long)p->n_items);
            return COR_E_HIP;
        }
    }
    if (cut) HIP_TRY(hipMemcpy(cut, p->d_cut, sizeof(cor_cutoff) * p->n_scans, hipMemcpyDeviceToHost));
    if (traj && p->n_items > 0)
        HIP_TRY(hipMemcpy(traj, p->out, sizeof(cor_traj) * (size_t)p->n_items, hipMemcpyDeviceToHost));
    return COR_OK;
}

int cor_trace_plan_stats(cor_trace_plan *p, cor_trace_stats *st) {
    if (!p || !st || !p->ran) return COR_E_INVALID;
    HIP_TRY(hipEventSynchronize(p->ev[2]));
    float k = 0.f, t = 0.f;
    HIP_TRY(hipEventElapsedTime(&k, p->ev[0], p->ev[1]));
    HIP_TRY(hipEventElapsedTime(&t, p->ev[0], p->ev[2]));
    std::vector<ScanStat> ss(p->n_scans);
    HIP_TRY(hipMemcpy(ss.data(), p->d_stat, sizeof(ScanStat) * p->n_scans, hipMemcpyDeviceToHost));
    st->kernel_ms = k;
    st->total_ms = t;
    st->steps = 0;
    st->attempts = 0;
    for (const ScanStat &x : ss) {
        st->steps += x.steps;
        st->attempts += x.attempts;
    }
    st->n_traj = p->n_items;
    st->grid_blocks = p->grid;
    st->block_threads = p->block;
    RunStats r;
    HIP_TRY(hipMemcpy(&r, p->d_run, sizeof r, hipMemcpyDeviceToHost));
    st->wave_iters = r.wave_iters;
    st->lane_iters = r.lane_iters;
    st->coop_steps = r.coop_steps;
    st->wave_iters_open = r.wave_iters_open;
    st->lane_iters_open = r.lane_iters_open;
    const double tick_ms = g_wall_khz > 0. ? 1. / g_wall_khz : 0.;
    const bool started = r.t_start != ~0ull;
    st->drain_ms = started && r.t_drain >= r.t_start ? (double)(r.t_drain - r.t_start) * tick_ms : 0.;
    st->lane_end_ms = started && r.t_lane_end >= r.t_start ? (double)(r.t_lane_end - r.t_start) * tick_ms : 0.;
    return COR_OK;
}

int cor_trace_plan_set_epoch(cor_trace_plan *p, const cor_model_set *models, const ModelsParams *epoch,
                             void *stream) {
    if (!p || !epoch || p->n_epochs != 1) return COR_E_INVALID;
    int intk, ext, cru, rc;
    if ((rc = model_kinds(models, &intk, &ext, &cru))) return rc;
    if (intk != p->intk || ext != p->ext || cru != p->cru) return COR_E_INVALID;
    hipStream_t s = stream ? (hipStream_t)stream : g_stream;
    if (!p->h_epoch) {
        HIP_TRY(hipHostMalloc((void **)&p->h_epoch, sizeof(EpochDev), hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&p->ev_up, hipEventDisableTiming));
    }
    // the staging block may still feed the previous upload, and the SH block the
    // previous run
    if (p->up_pending) HIP_TRY(hipEventSynchronize(p->ev_up));
    p->up_pending = false;
    if (p->ran) HIP_TRY(hipEventSynchronize(p->ev[2]));
    build_epoch(models, epoch, *p->h_epoch);
    HIP_TRY(upload_sh(models, &p->d_sh, &p->sh_cap, *p->h_epoch));
    HIP_TRY(hipMemcpyAsync(p->d_epoch, p->h_epoch, sizeof(EpochDev), hipMemcpyHostToDevice, s));
    if (p->coop_on || needs_coop_tables(p->prec, p->ext)) HIP_TRY(build_coop_tables(p->d_epoch, s));
    HIP_TRY(hipEventRecord(p->ev_up, s));
    p->up_pending = true;
    return COR_OK;
}

namespace {
// streams of the epoch pipeline (slot 0 is the library stream); default priority
constexpr int SERIES_MAX_IN_FLIGHT = 4;
hipStream_t g_series_stream[SERIES_MAX_IN_FLIGHT] = {};

// Offsets of the pinned result block of an epoch graph: the CoopCtl counters, the
// per-scan cutoffs, then the trajectories' cor_traj records.
struct ResLayout {
    size_t cut, rec, total;
};
ResLayout res_layout(const cor_trace_plan *p) {
    ResLayout L;
    L.cut = sizeof(CoopCtl);
    L.rec = (L.cut + sizeof(cor_cutoff) * (size_t)p->n_scans + 15) & ~(size_t)15;
    L.total = L.rec + sizeof(cor_traj) * (size_t)std::max<int64_t>(p->n_items, 1);
    return L;
}

// Capture one epoch of the plan as a graph on stream s: the upload of the pinned epoch
// block, the lane tables, the run (reset, trace, penumbra; no memset nodes) and the
// copies of every result into the pinned result block.  Replayed once per epoch after
// the host has rebuilt the epoch block.
hipError_t capture_epoch_graph(cor_trace_plan *p, hipStream_t s) {
    hipError_t e;
    const ResLayout L = res_layout(p);
    if (!p->h_epoch) {
        if ((e = hipHostMalloc((void **)&p->h_epoch, sizeof(EpochDev), hipHostMallocDefault)) != hipSuccess) return e;
    }
    if ((e = hipHostMalloc((void **)&p->h_res, L.total, hipHostMallocDefault)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&p->ev_graph, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
    const size_t n = (size_t)p->n_items;
    e = hipMemcpyAsync(p->d_epoch, p->h_epoch, sizeof(EpochDev), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && (p->coop_on || needs_coop_tables(p->prec, p->ext))) e = build_coop_tables(p->d_epoch, s);
    if (e == hipSuccess) e = enqueue_run(p, s, false);
    if (e == hipSuccess && p->d_coop)
        e = hipMemcpyAsync(p->h_res, p->d_coop, sizeof(CoopCtl), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->h_res + L.cut, p->d_cut, sizeof(cor_cutoff) * p->n_scans, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && n > 0)
        e = hipMemcpyAsync(p->h_res + L.rec, p->out, sizeof(cor_traj) * n, hipMemcpyDeviceToHost, s);
    // always leave capture mode, also after a failed enqueue
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return e;
    }
    if (e_end != hipSuccess) return e_end;
    p->graph = g;
    return hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
}

// Results of the plan's last graph replay, from the pinned block (waits for it).
int collect_graph(cor_trace_plan *p, cor_traj *traj, cor_cutoff *cut) {
    HIP_TRY(hipEventSynchronize(p->ev_graph));
    const ResLayout L = res_layout(p);
    if (p->coop_on) {
        const CoopCtl *c = (const CoopCtl *)p->h_res;
        if (c->done != (unsigned long long)p->n_items) {
            snprintf(g_hip_error, sizeof g_hip_error, "cooperative tail finished %llu of %lld trajectories",
                     c->done, (long long)p->n_items);
            return COR_E_HIP;
        }
    }
    if (cut) memcpy(cut, p->h_res + L.cut, sizeof(cor_cutoff) * p->n_scans);
    if (traj && p->n_items > 0) memcpy(traj, p->h_res + L.rec, sizeof(cor_traj) * (size_t)p->n_items);
    return COR_OK;
}
}  // namespace

int cor_gpu_trace_series(const cor_model_set *models, const ModelsParams *epochs, int32_t n_epochs,
                         const cor_scan *scans, int32_t n_scans, const cor_trace_opts *opts,
                         int32_t in_flight, cor_traj *traj, cor_cutoff *cut) {
    if (!epochs || n_epochs < 1 || !scans || n_scans < 1 || !opts || in_flight < 0) return COR_E_INVALID;
    int rc = ensure_init();
    if (rc) return rc;
    int k_max = in_flight == 0 ? 3 : std::min(in_flight, SERIES_MAX_IN_FLIGHT);
    k_max = std::min(k_max, n_epochs);
    g_series_stream[0] = g_stream;
    for (int k = 1; k < k_max; k++)
        if (!g_series_stream[k]) HIP_TRY(hipStreamCreateWithFlags(&g_series_stream[k], hipStreamNonBlocking));
    cor_trace_plan *plan[SERIES_MAX_IN_FLIGHT] = {};
    int64_t count = 0;
    // Each slot holds one plan and its epoch graph; epoch e replays slot e % k_max's
    // graph after the host has collected that slot's previous epoch and rebuilt the
    // pinned epoch block, so up to k_max epochs are on the device at once and the tail
    // of one overlaps the bulk of the next.
    auto collect = [&](int slot, int e) -> int {
        cor_traj *t = traj ? traj + (size_t)e * (size_t)count : nullptr;
        cor_cutoff *c = cut ? cut + (size_t)e * (size_t)n_scans : nullptr;
        return collect_graph(plan[slot], t, c);
    };
    for (int e = 0; e < n_epochs && rc == COR_OK; e++) {
        const int slot = e % k_max;
        hipStream_t s = g_series_stream[slot];
        if (!plan[slot]) {
            rc = cor_trace_plan_create(models, &epochs[e], scans, n_scans, opts, &plan[slot]);
            if (rc != COR_OK) break;
            count = plan[slot]->n_items;
            const hipError_t he = capture_epoch_graph(plan[slot], s);
            if (he != hipSuccess) rc = fail_hip(he, "capture_epoch_graph");
        } else {
            rc = collect(slot, e - k_max);
        }
        if (rc != COR_OK) break;
        cor_trace_plan *p = plan[slot];
        build_epoch(models, &epochs[e], *p->h_epoch);
        point_sh(models, p->d_sh, *p->h_epoch);
        hipError_t he = hipGraphLaunch(p->gexec, s);
        if (he == hipSuccess) he = hipEventRecord(p->ev_graph, s);
        if (he != hipSuccess) rc = fail_hip(he, "hipGraphLaunch");
    }
    for (int e = std::max(0, n_epochs - k_max); e < n_epochs && rc == COR_OK; e++) rc = collect(e % k_max, e);
    for (int k = 0; k < k_max; k++) cor_trace_plan_destroy(plan[k]);
    return rc;
}

int cor_gpu_trace(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                  int32_t n_scans, const cor_trace_opts *opts, cor_traj *traj, cor_cutoff *cut) {
    cor_trace_plan *p = nullptr;
    int rc = cor_trace_plan_create(models, epoch, scans, n_scans, opts, &p);
    if (rc) return rc;
    rc = cor_trace_plan_run(p, nullptr);
    if (!rc) rc = cor_trace_plan_results(p, traj, cut);
    cor_trace_plan_destroy(p);
    return rc;
}

int cor_gpu_trace_noalloc(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                          int32_t n_scans, const cor_trace_opts *opts, cor_traj *d_traj, cor_cutoff *d_cut,
                          void *stream) {
    if (!epoch || !opts || !scans || !d_traj || !d_cut) return COR_E_INVALID;
    if (((uintptr_t)d_traj & 15u) != 0) {
        snprintf(g_hip_error, sizeof g_hip_error, "cor_gpu_trace_noalloc: d_traj must be 16-byte aligned");
        return COR_E_INVALID;
    }
    if (n_scans < 1 || n_scans > COR_NOALLOC_MAX_SCANS) {
        snprintf(g_hip_error, sizeof g_hip_error, "cor_gpu_trace_noalloc takes 1 to %d scans, not %d",
                 COR_NOALLOC_MAX_SCANS, (int)n_scans);
        return COR_E_INVALID;
    }
    if (opts->precision != COR_PREC_FP64 && opts->precision != COR_PREC_MIXED) return COR_E_INVALID;
    if (!(opts->particle_mass > 0.) || !(opts->tu_angle > 0.)) return COR_E_INVALID;
    int rc = ensure_init();
    if (rc) return rc;
    if (g_device >= NOALLOC_MAX_DEVICES || !g_noalloc[g_device]) return COR_E_NO_DEVICE;
    NoAllocScratch *sc = g_noalloc[g_device];
    int intk, ext, cru;
    if ((rc = model_kinds(models, &intk, &ext, &cru))) return rc;
    if (cru != CRU_NONE) {
        snprintf(g_hip_error, sizeof g_hip_error, "cor_gpu_trace_noalloc: crustal tables are not held in its scratch");
        return COR_E_UNSUPPORTED;
    }
    const void *kern = opts->precision == COR_PREC_FP64 ? trace_kernel_fp64(intk, ext, cru)
                                                       : trace_kernel_mixed(intk, ext, cru);
    if (!kern) return COR_E_UNSUPPORTED;
    hipStream_t s = stream ? (hipStream_t)stream : g_stream;
    EpochDev &h_epoch = g_noalloc_host[g_device].epoch;
    ScanDev *h_scans = g_noalloc_host[g_device].scans;
    int64_t total = 0;
    for (int32_t k = 0; k < n_scans; k++) {
        const int64_t n = cor_scan_count(&scans[k]);
        if (n < 0 || scans[k].nk1 <= 0 || !isfinite(scans[k].zn) || scans[k].zn == 0.) return COR_E_INVALID;
        build_scan(scans[k], total, (int32_t)n, h_scans[k]);
        total += n;
    }
    if (total >= (int64_t)0x7fffffffll) return COR_E_INVALID;
    build_epoch(models, epoch, h_epoch);
    HIP_TRY(hipMemcpyAsync(&sc->epoch, &h_epoch, sizeof(EpochDev), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sc->scans, h_scans, sizeof(ScanDev) * n_scans, hipMemcpyHostToDevice, s));
    if (needs_coop_tables(opts->precision, ext)) HIP_TRY(build_coop_tables(&sc->epoch, s));
    TraceArgs a;
    memset(&a, 0, sizeof a);
    a.out = d_traj;                        // the caller's records, one per trajectory
    a.epoch = &sc->epoch;
    a.scans = sc->scans;
    a.n_scans = n_scans;
    a.order = nullptr;                     // rigidity order: the results do not depend on it
    a.n_items = total;
    a.head = &sc->head;
    a.step_limit = opts->step_limit;
    a.tu_angle = opts->tu_angle;
    a.cos_tu = cos(opts->tu_angle);
    a.mass = opts->particle_mass;
    a.inv_mass = 1.0 / opts->particle_mass;
    a.coop_on = 0;                         // per-lane mode: its park slots would scale with the batch
    a.stats = &sc->run;
    // zero the queue head and the run counters (a kernel, also under graph capture)
    reset_head_kernel<<<1, 1, 0, s>>>(&sc->head, nullptr, &sc->run);
    HIP_TRY(hipGetLastError());
    if (total > 0) {
        int per_cu = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
        const int64_t need = (total + 63) / 64;
        int64_t grid = (int64_t)std::max(per_cu, 1) * g_num_cus;
        int block = 256;
        if (need * 64 < (int64_t)block * grid) {
            block = 64;
            grid = std::min<int64_t>(need, (int64_t)std::max(per_cu, 1) * 4 * g_num_cus);
        }
        void *args[] = {&a};
        HIP_TRY(hipLaunchKernel(kern, dim3((unsigned)std::max<int64_t>(grid, 1)), dim3(block), args, 0, s));
    }
    penumbra_kernel<<<n_scans, 64, 0, s>>>(sc->scans, a.out, d_cut, sc->stat);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return COR_OK;
}

int cor_gpu_trace_tracks(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                         int32_t n_scans, const cor_trace_opts *opts, int64_t *counts, int64_t *total,
                         cor_track_point *points, int64_t capacity) {
    if (!opts || !total || opts->precision != COR_PREC_FP64 || (points && capacity < 0)) return COR_E_INVALID;
    cor_trace_plan *p = nullptr;
    int rc = cor_trace_plan_create(models, epoch, scans, n_scans, opts, &p);
    if (rc) return rc;
    const int64_t n = p->n_items;
    std::vector<cor_traj> tr((size_t)std::max<int64_t>(n, 1));
    // pass 1: the trace itself; a trajectory ending at step S has S - 1 points
    rc = cor_trace_plan_run(p, nullptr);
    if (!rc) rc = cor_trace_plan_results(p, tr.data(), nullptr);
    std::vector<int64_t> off((size_t)n + 1, 0);
    if (!rc) {
        for (int64_t i = 0; i < n; i++) off[i + 1] = off[i] + (tr[i].steps > 0 ? (int64_t)tr[i].steps - 1 : 0);
        *total = off[n];
        if (counts)
            for (int64_t i = 0; i < n; i++) counts[i] = off[i + 1] - off[i];
        if (points && capacity < off[n]) rc = COR_E_NOMEM;
    }
    if (!rc && points && off[n] > 0) {
        // pass 2: the same trace (FP64 is deterministic) recording every point
        const size_t bytes = sizeof(cor_track_point) * (size_t)off[n];
        hipError_t e = hipMalloc(&p->d_trk, bytes);
        if (e == hipSuccess) e = hipMalloc(&p->d_trk_off, sizeof(int64_t) * ((size_t)n + 1));
        if (e == hipSuccess) e = hipMemset(p->d_trk, 0, bytes);
        if (e == hipSuccess)
            e = hipMemcpy(p->d_trk_off, off.data(), sizeof(int64_t) * ((size_t)n + 1), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            rc = fail_hip(e, "cor_gpu_trace_tracks");
        } else {
            rc = cor_trace_plan_run(p, nullptr);
            if (!rc) rc = cor_trace_plan_sync(p);
            if (!rc) {
                e = hipMemcpy(points, p->d_trk, bytes, hipMemcpyDeviceToHost);
                if (e != hipSuccess) rc = fail_hip(e, "cor_gpu_trace_tracks");
            }
        }
        if (!rc) {
            // a trajectory whose state turned NaN wrote its points up to there only:
            // the written points are the leading ones with a step number
            int64_t t = 0;
            for (int64_t i = 0; i < n; i++) {
                int64_t m = 0;
                while (m < off[i + 1] - off[i] && points[off[i] + m].step != 0u) m++;
                if (counts) counts[i] = m;
                for (int64_t j = 0; j < m && t != off[i]; j++) points[t + j] = points[off[i] + j];
                t += m;
            }
            *total = t;
        }
    }
    cor_trace_plan_destroy(p);
    return rc;
}

int cor_gpu_rk_step(const double *h, const double *d, const double *field, double *vel, double *pos, int64_t n) {
    if (n < 0 || (n > 0 && (!h || !d || !field || !vel || !pos))) return COR_E_INVALID;
    int rc = ensure_init();
    if (rc) return rc;
    if (n == 0) return COR_OK;
    // one block: h, d (n each), field, vel, pos (3n each)
    const size_t nn = (size_t)n;
    double *blk = nullptr;
    hipError_t e = hipMalloc(&blk, sizeof(double) * 11 * nn);
    RkStepArgs a;
    a.h = blk;
    a.d = blk + nn;
    a.field = blk + 2 * nn;
    a.vel = blk + 5 * nn;
    a.pos = blk + 8 * nn;
    a.n = n;
    if (e == hipSuccess) e = hipMemcpyAsync(blk, h, sizeof(double) * nn, hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(blk + nn, d, sizeof(double) * nn, hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(blk + 2 * nn, field, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(a.vel, vel, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(a.pos, pos, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) {
        void *args[] = {&a};
        e = hipLaunchKernel(rk_step_kernel_fp64(), dim3((unsigned)((n + 255) / 256)), dim3(256), args, 0, g_stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(vel, a.vel, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pos, a.pos, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    (void)hipFree(blk);
    if (e != hipSuccess) return fail_hip(e, "cor_gpu_rk_step");
    return COR_OK;
}

int cor_gpu_field(const cor_model_set *models, const ModelsParams *epoch, const double *gsm_m,
                  double *geo_nT, int64_t n, int32_t precision) {
    if (!epoch || n < 0 || (n > 0 && (!gsm_m || !geo_nT))) return COR_E_INVALID;
    if (precision < COR_PREC_FP64 || precision > COR_PREC_MIXED_GROUP4) return COR_E_INVALID;
    const bool coop = precision >= COR_PREC_MIXED_WAVE;
    const bool coop64 = precision == COR_PREC_FP64_WAVE || precision == COR_PREC_FP64_GROUP;
    const int lanes = precision == COR_PREC_MIXED_GROUP || precision == COR_PREC_FP64_GROUP ? 8
                      : precision == COR_PREC_MIXED_GROUP4                                   ? 4
                      : precision == COR_PREC_MIXED_GROUP16                                  ? 16
                      : precision == COR_PREC_MIXED_GROUP32                                  ? 32
                                                                                             : 64;
    int rc = ensure_init();
    if (rc) return rc;
    int intk, ext, cru;
    if ((rc = model_kinds(models, &intk, &ext, &cru))) return rc;
    const void *kern = precision == COR_PREC_FP64    ? field_kernel_fp64(intk, ext, cru)
                       : precision == COR_PREC_MIXED ? field_kernel_mixed(intk, ext, cru)
                       : coop64                      ? field_kernel_coop_fp64(intk, ext, cru, lanes)
                                                     : field_kernel_coop(intk, ext, cru, lanes);
    if (!kern) return COR_E_UNSUPPORTED;
    if (n == 0) return COR_OK;
    EpochDev E;
    build_epoch(models, epoch, E);
    EpochDev *d_e = nullptr;
    double *d_in = nullptr, *d_out = nullptr;
    void *d_sh = nullptr;
    size_t sh_cap = 0;
    hipError_t e = upload_sh(models, &d_sh, &sh_cap, E);
    if (e == hipSuccess) e = hipMalloc(&d_e, sizeof(EpochDev));
    if (e == hipSuccess) e = hipMalloc(&d_in, sizeof(double) * 3 * n);
    if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * 3 * n);
    if (e == hipSuccess) e = hipMemcpyAsync(d_e, &E, sizeof(EpochDev), hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in, gsm_m, sizeof(double) * 3 * n, hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess && (coop || needs_coop_tables(precision, ext)))
        e = build_coop_tables(d_e, g_stream);
    if (e == hipSuccess) {
        FieldBatchArgs a;
        a.epoch = d_e;
        a.gsm_m = d_in;
        a.geo_nT = d_out;
        a.n = n;
        void *args[] = {&a};
        // one lane per point, or one cooperative group per point (strided)
        const int64_t per_block = 256 / lanes;
        const int64_t blocks = coop ? std::min<int64_t>((n + per_block - 1) / per_block, 4096) : (n + 255) / 256;
        e = hipLaunchKernel(kern, dim3((unsigned)blocks), dim3(256), args, 0, g_stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(geo_nT, d_out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    (void)hipFree(d_e);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_sh);
    if (e != hipSuccess) return fail_hip(e, "cor_gpu_field");
    return COR_OK;
}

}  // extern "C"
