// dev_types.hpp -- device-side data layout of the cutoff tracer (shared by host launch
// code and kernels).
//
// HBM layout (one launch = one epoch, so every epoch constant is wavefront-uniform):
//   EpochDev   one block per epoch: IGRF G/H/REC and the GEO->GSM rows in FP64 and
//              FP32, the hoisted T96 / TS05 constants, the SH (CHAOS-form) tables.
//   ScanDev[]  one per (station, direction, rigidity scan): station position and
//              arrival-direction trig computed on the host exactly as the reference's
//              sph_to_car does, rigidity grid, charge, nk1, first global index.
//   order[]    queue order: a permutation of the launch's global trajectory ids.
//   results    structure of arrays indexed by global trajectory id (coalesced-free,
//              written once per trajectory).
#pragma once
#include <stdint.h>
#include "t96_dev.hpp"

namespace cor {

enum { PREC_FP64 = 0, PREC_MIXED = 1 };
enum { INT_NONE = 0, INT_IGRF = 1 };
enum { EXT_NONE = 0, EXT_T96 = 1, EXT_TS05 = 2 };
enum { CRU_NONE = 0, CRU_SH = 1 };
enum { ST_ALLOWED = 0, ST_FORBIDDEN_REENTRY = 1, ST_FORBIDDEN_LIMIT = 2 };

constexpr int NCOEF = 105;
constexpr int SH_NM_MAX = 24;    // SH (CHAOS-form) tables up to degree 23 on the device

// Physical constants exactly as src/common.h:4-11 writes them.
constexpr double K_PI = 3.141592654;
constexpr double K_2PI = 6.28318531;
constexpr double K_RE = 6.3712e+06;
constexpr double K_TO_RAD = K_PI / 180.;
constexpr double K_Q = 1.6021e-19;
constexpr double K_C = 2.99725e+08;

struct EpochDev {
    double G[NCOEF], H[NCOEF], REC[NCOEF];
    double A[9];                 // GEO->GSM rows A1, A2, A3
    float Gf[NCOEF], Hf[NCOEF], RECf[NCOEF];
    float Af[9];
    T96Params<double> t96d;
    T96Params<float> t96f;
    double parmod[10];
    int32_t sh_nm;               // SH table size (0 = none)
    int32_t pad;
    double shG[SH_NM_MAX * SH_NM_MAX], shH[SH_NM_MAX * SH_NM_MAX];
    float shGf[SH_NM_MAX * SH_NM_MAX], shHf[SH_NM_MAX * SH_NM_MAX];
};

struct ScanDev {
    double pos0[3];              // station, GEO Cartesian [m]
    double st1, ct1, sp1, cp1;   // sin/cos of the arrival direction's theta and phi
    double rig, del, rk, zn;
    int32_t nk1, n;
    int64_t first;               // global id of rigidity index 0
};

struct TraceOut {
    int32_t *status;
    uint32_t *steps;
    uint32_t *attempts;
    double *vc, *r, *td, *fei, *ast, *asf, *time, *length;
};

struct TraceArgs {
    const EpochDev *epoch;
    const ScanDev *scans;
    int32_t n_scans;
    int32_t pad;
    const uint32_t *order;       // queue slot -> global id (NULL: identity)
    int64_t id_base;             // global id of queue slot 0's range
    int64_t n_items;
    unsigned long long *head;    // queue head
    uint64_t step_limit;
    double tu_angle;
    double cos_tu;
    double mass;
    TraceOut out;
};

}  // namespace cor
