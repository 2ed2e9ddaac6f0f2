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
#include <stddef.h>
#include <stdint.h>
#include "cor_gpu.h"
#include "t96_dev.hpp"
#include "ts05_dev.hpp"

namespace cor {

enum { PREC_FP64 = 0, PREC_MIXED = 1 };
enum { INT_NONE = 0, INT_IGRF = 1 };
enum { EXT_NONE = 0, EXT_T96 = 1, EXT_TS05 = 2 };
enum { CRU_NONE = 0, CRU_SH = 1 };
enum { ST_ALLOWED = 0, ST_FORBIDDEN_REENTRY = 1, ST_FORBIDDEN_LIMIT = 2 };

constexpr int NCOEF = 105;

// Physical constants exactly as src/common.h:4-11 writes them.
constexpr double K_PI = 3.141592654;
constexpr double K_2PI = 6.28318531;
constexpr double K_RE = 6.3712e+06;
constexpr double K_INV_RE = 1.0 / K_RE;   // mixed path: positions in Re for the FP32 field
constexpr double K_TO_RAD = K_PI / 180.;
constexpr double K_Q = 1.6021e-19;
constexpr double K_C = 2.99725e+08;

// Lane tables of the cooperative (one wavefront per trajectory) T96 evaluation: the
// terms of each family of identical form, one term per lane, with the epoch's
// amplitudes folded in.  Built on the device once per epoch (build_coop_kernel) and
// staged in LDS by every workgroup that runs cooperative steps.
constexpr int CO_NHARM = 47;       // Cartesian harmonics: 3 shield sets x 9, birk1shld 16, birk2shl 4
constexpr int CO_NICON = 3;        // interconnection harmonics, one y harmonic per lane
constexpr int CO_NCONE = 6;        // conical sheets: ring current 2, tail disk 4
constexpr int CO_NCYL = 6;         // cylindrical harmonics
constexpr int CO_NT87 = 3;         // tail87: central sheet and its two images
constexpr int CO_NDIP_LOOP = 21;   // Region-1 loop-basis dipoles
constexpr int CO_NDIP = 67;        // ... followed by 46 conical-basis dipoles
constexpr int CO_NCIRC = 18;       // current loops: Region-1 loop basis 3, Region-2 outer 11, inner 4
constexpr int CO_NCONIC = 10;      // conical harmonics: Region-1 5, Region-2 inner 5
constexpr int CO_NSHEET = 15;      // Region-2 sheet: 3 components x 5 latitude terms

template <class F>
struct alignas(16) CoopTabT {  // rows are read 16 bytes at a time
    F harm[64][4];           // 1/p, 1/r, sqrt(1/p^2 + 1/r^2), coefficient x amplitude
    F cone[8][4];            // f x amplitude, beta, x offset, dd slope (0 ring current, 10 disk)
    F cyl[8][4];             // a x xappa^3, scale, kind (0: j1 terms, 1: mixed terms), 0
    F t87[4][4];             // z offset, weight x amplitude, 0, 0
    F dip[68][8];            // position (3), 0, moment x amplitude (3), 0
    F circ[24][16];          // rotation rows (9), centre (3), radius, coefficient x amplitude, group, 0
    F conic[16][4];          // order, coefficient x amplitude, x offset, group
    F sheet[16][24];         // alpha, beta, gamma, component, d1, d2, d3, d3^4, 16 coefficients
    F icon[4][24];           // per y harmonic: the k = 0 term and the series of the k = 1, 2 pair
};
typedef CoopTabT<float> CoopTab;

struct EpochDev {
    double G[NCOEF], H[NCOEF], REC[NCOEF];
    double A[9];                 // GEO->GSM rows A1, A2, A3
    float Gf[NCOEF], Hf[NCOEF], RECf[NCOEF];
    float Af[9];
    T96Params<double> t96d;
    T96Params<float> t96f;
    TS05Params<double> ts05d;
    TS05Params<float> ts05f;
    double parmod[10];
    int32_t sh_nm;               // SH table size (0 = none)
    int32_t pad;
    const double *shG, *shH;     // SH tables [sh_nm][sh_nm] in their own device block
    const float *shGf, *shHf;
    CoopTab coop;                // mixed precision
    CoopTabT<double> coopd;      // FP64 cooperative tail
};

// State of a trajectory handed from the per-lane mode to the cooperative mode, taken
// at the top of a step (the field is re-evaluated from the position).
struct ParkDev {
    double pos[3], vel[3];
    double v, time, alength, zn;
    int64_t gid;
    uint32_t step, attempts;
    int32_t nk1;
    uint32_t handover;           // 1 + step of the first handover (0: this one is the first)
    uint32_t ready;              // = run id once the record is complete
};

// Lanes per trajectory when a cooperative worker runs several trajectories: groups of 8
// (eight to a wavefront); a lone trajectory gets all 64.  The evaluator also runs with
// groups of 4 and 16 (cor_gpu_field's COR_PREC_MIXED_GROUP4 / _GROUP16), and every width
// gives the same bits, but the tracer is slower with them as the bookkeeping width: on
// the 5-degree grid 16.3 s (4) and 16.6 s (16) against 14.9 s (8), with the handover at
// 98304 steps.
#ifndef COR_COOP_GROUP_MIXED
#define COR_COOP_GROUP_MIXED 8
#endif
constexpr int COOP_GROUP = 8;
constexpr int COOP_GROUP_MIXED = COR_COOP_GROUP_MIXED;
template <int PREC>
constexpr int coop_group() { return PREC == PREC_MIXED ? COOP_GROUP_MIXED : COOP_GROUP; }

// Work-handover counters of one run (cleared by the reset kernel).
// The counters each sit in a cache line of their own: workers poll them while others
// update them.
struct CoopCtl {
    alignas(128) unsigned long long res;      // park slots reserved
    alignas(128) unsigned long long head;     // tickets taken by cooperative workers
    alignas(128) int32_t running;             // wavefronts still in the per-lane mode
    alignas(128) unsigned long long reparks;  // trajectories handed back by cooperative workers
    alignas(128) unsigned long long done;     // trajectories finished (both modes)
    alignas(128) int32_t converting;          // wavefronts running down their lanes to turn worker
    alignas(128) int32_t feeders;             // wavefronts taking new trajectories from the queue
    alignas(128) unsigned long long lres;     // per-lane pool: records reserved by donating wavefronts
    alignas(128) unsigned long long lhead;    // per-lane pool: records claimed by taking wavefronts
    alignas(128) uint32_t run;                // run id: advanced by every reset, never cleared, so
                                              // park records of earlier runs never match (a value
                                              // the device keeps, so a captured run replays right)
};

struct ScanDev {
    double pos0[3];              // station, GEO Cartesian [m]
    double st1, ct1, sp1, cp1;   // sin/cos of the arrival direction's theta and phi
    double rig, del, rk, zn;
    int32_t nk1, n;
    int64_t first;               // global id of rigidity index 0
    int32_t epoch;               // multi-epoch launches: index of the scan's epoch block
    int32_t pad;
};

// Results: one cor_traj record per trajectory (80 B: status, steps, attempts, handover,
// then the eight exit doubles), so that a trajectory's end writes five aligned 16-byte
// stores to neighbouring addresses and the host copies the array to the caller as it is.
static_assert(sizeof(cor_traj) == 80 && offsetof(cor_traj, vc) == 16, "cor_traj record layout");

// Per-run counters of the tracer: how full the per-lane wavefronts were and when the
// phases ended (times in wall_clock64 ticks; t_start starts at all ones).
struct RunStats {
    unsigned long long wave_iters;   // per-lane mode: loop iterations summed over wavefronts
    unsigned long long lane_iters;   // per-lane mode: live lanes summed over those iterations
    unsigned long long coop_steps;   // cooperative mode: field evaluations (one per wavefront step)
    unsigned long long wave_iters_open;   // the part of wave_iters while the queue was open
    unsigned long long lane_iters_open;   // the part of lane_iters while the queue was open
    unsigned long long t_start;      // earliest wavefront start
    unsigned long long t_drain;      // the queue's last slot taken
    unsigned long long t_lane_end;   // last wavefront to leave the per-lane mode
    unsigned long long finished;     // trajectories ended so far (both modes; read by the
                                     // host while the kernel runs, as its progress)
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
    double inv_mass;             // 1 / mass
    cor_traj *out;               // result records, indexed by global trajectory id
    // cooperative mode (mixed precision, IGRF/T96 model sets): a trajectory that has
    // run handover_steps accepted steps in the per-lane mode is parked and continued by
    // cooperative workers, a group of lanes per trajectory (coop_on = 0: per-lane only)
    int32_t coop_on;
    uint32_t pad2;
    CoopCtl *coop;
    ParkDev *park;
    RunStats *stats;
    long long repark_limit;      // park slots beyond the per-lane ones (one per lane of the grid)
    ParkDev *pool;               // per-lane pool: trajectories of sparse per-lane wavefronts after
                                 // the queue drained, continued per-lane by wavefronts with idle lanes
    unsigned long long pool_cap; // records the pool holds over one run (0: no compaction)
    cor_track_point *trk;        // FP64 track recording (NULL: off)
    const int64_t *trk_off;      // first point of each trajectory (n_items + 1 entries)
    uint32_t handover_steps;     // per-lane accepted steps before the handover
    uint32_t handover_low_steps; // ... for the first handover_low_count rigidities of a scan
    int32_t handover_low_count;
    int32_t convert_backlog;     // parked trajectories waiting, beyond what converting
                                 // wavefronts will take, that turn a feeding wavefront
                                 // into a worker
    int32_t min_feeders;         // wavefronts that keep taking new trajectories until the
                                 // queue is drained (so it always drains)
    unsigned long long idle_guard;   // wall-clock ticks without a finished trajectory after
                                     // which an idle worker gives up (a lost-record guard
                                     // longer than any trajectory can legitimately run)
};

}  // namespace cor
