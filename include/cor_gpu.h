/*
 * cor_gpu.h -- batch C-ABI of the MI355X cutoff-rigidity tracer (gfx950).
 *
 * Plain pointers and sizes only.  Host buffers are caller-owned; the library owns its
 * device buffers.  Every call returns 0 or a negative COR_E* code
 * (cor_gpu_error_string).  Calls on one plan are single-threaded.
 *
 * Counterparts in the reference:
 *   cor_gpu_trace / cor_trace_plan_*  replace the per-rigidity loop of
 *       trajectory_simulation_cutoff (src/trajectory.c:6-161, declared
 *       src/trajectory.h:10), batched over many (station, direction, scan) items;
 *   cor_gpu_field  replaces per-point calculate_field (src/models.c:194-207,
 *       declared src/models.h:32) over a batch of points;
 *   cor_cutoff  is the summary line of src/trajectory.c:156-160;
 *   cor_traj  carries the per-rigidity line of src/trajectory.c:118-122.
 */
#ifndef COR_GPU_H
#define COR_GPU_H

#include <stdint.h>
#include "cor.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    COR_OK = 0,
    COR_E_INVALID = -1,       /* bad argument (sizes, pointers, model set, scan) */
    COR_E_NO_DEVICE = -2,     /* no GPU, or the library was not built for it */
    COR_E_HIP = -3,           /* a HIP runtime call failed (cor_gpu_last_hip_error) */
    COR_E_UNSUPPORTED = -4,   /* model combination not compiled into the tracer */
    COR_E_NOMEM = -5
};

/* COR_PREC_MIXED_WAVE / _GROUP4 / _GROUP / _GROUP16 / _GROUP32: cor_gpu_field only --
 * the mixed-precision model evaluated as the tracer's cooperative mode evaluates it:
 * one wavefront (64 lanes) per point, one group of 4 lanes (sixteen trajectories to a
 * wavefront), of 8 lanes (eight), or 16 or 32 lanes (four or two to a wavefront).
 * COR_PREC_FP64_WAVE / _GROUP: the same evaluator in FP64 (the FP64 tracer's
 * cooperative tail), 64 or 8 lanes per point. */
enum { COR_PREC_FP64 = 0, COR_PREC_MIXED = 1, COR_PREC_MIXED_WAVE = 2, COR_PREC_MIXED_GROUP = 3,
       COR_PREC_MIXED_GROUP16 = 4, COR_PREC_MIXED_GROUP32 = 5, COR_PREC_FP64_WAVE = 6,
       COR_PREC_FP64_GROUP = 7, COR_PREC_MIXED_GROUP4 = 8 };
enum { COR_INTERNAL_NONE = 0, COR_INTERNAL_IGRF = 1 };
enum { COR_EXTERNAL_NONE = 0, COR_EXTERNAL_T96 = 1, COR_EXTERNAL_TS05 = 2 };
enum { COR_CRUSTAL_NONE = 0, COR_CRUSTAL_SH = 1 };
enum { COR_ALLOWED = 0, COR_FORBIDDEN_REENTRY = 1, COR_FORBIDDEN_LIMIT = 2 };

#define COR_SH_NM_MAX 200  /* largest SH table (rows = cols) the device accepts */

/* Which evaluators the field sums, in the reference's order (internal, external,
 * crustal; src/main.c:268-299).  The SH table is [sh_nm_max][sh_nm_max], row n,
 * column m, nT, as lib/chaos/chaos.c:198-199 indexes it. */
typedef struct {
    int32_t internal;
    int32_t external;
    int32_t crustal;
    int32_t sh_nm_max;
    const double *sh_g;
    const double *sh_h;
} cor_model_set;

/* One station x one arrival direction x one rigidity scan: the content of one
 * reference input header (src/io.c:4-20, src/common.h:26-54).  The scan is
 * rig + i*del for i = 0 .. round((rk - rig)/del). */
typedef struct {
    double r0, the0, fi0;     /* station radius [Re], geocentric latitude, longitude [deg] */
    double the1, fi1;         /* arrival direction as GEO latitude, longitude [deg] */
    double rig, del, rk;      /* first rigidity, step, last rigidity [GV] */
    double zn;                /* charge number; -1 back-traces a proton */
    int32_t nk1;              /* steps per gyro-period each step starts from */
    int32_t pad;
} cor_scan;

/* Tracer options: the reference's cutoff() arguments (src/main.c:12-30). */
typedef struct {
    uint64_t step_limit;      /* -s (reference default 25000) */
    double tu_angle;          /* -u [rad] (default 0.01) */
    double particle_mass;     /* DEF_HM0_P or DEF_HM0_E */
    int32_t precision;        /* COR_PREC_FP64 or COR_PREC_MIXED */
    int32_t queue_order;      /* 0 = rigidity order, 1 = longest-expected first */
    int32_t coop;             /* the cooperative mode (IGRF and/or T96 model sets): 0 = the
                               * library default (on for COR_PREC_MIXED; off for COR_PREC_FP64,
                               * which then keeps the reference's operation order throughout),
                               * positive = on, negative = per-lane mode only */
    int32_t handover_steps;   /* with the cooperative mode: accepted steps a trajectory runs in
                               * the per-lane mode (one lane each) before it continues in the
                               * cooperative mode (a group of lanes each); 0 = the library
                               * default (98304 in mixed precision, 65536 in FP64).  A fixed
                               * step count: the results do not depend on scheduling */
    int32_t waves_per_simd;   /* resident wavefronts per SIMD of the persistent grid: 0 = as many
                               * as the kernel's registers allow (the default), otherwise at most
                               * that many.  Scheduling only: the results do not change */
    int32_t block_threads;    /* threads per workgroup of the persistent grid: 0 = the library's
                               * choice (256, or 64 for launches with fewer trajectories than
                               * lanes), else 64 or 256.  Scheduling only */
    int32_t handover_low_steps;  /* with the cooperative mode: the handover step count of the
                                  * first handover_low_count rigidities of every scan (the
                                  * lowest, where trapped trajectories concentrate); 0 = the
                                  * library default (32768, or handover_steps if smaller),
                                  * negative = handover_steps */
    int32_t handover_low_count;  /* how many rigidities of each scan handover_low_steps covers:
                                  * 0 = the library default (none), negative = none.  A property
                                  * of the trajectory, so the results do not depend on scheduling */
} cor_trace_opts;

/* Per trajectory.  The doubles are set only for allowed trajectories: exit v/c, exit
 * radius [Re], exit latitude/longitude, asymptotic latitude/longitude [deg], flight
 * time [s] and path length [m]. */
typedef struct {
    int32_t status;           /* COR_ALLOWED / COR_FORBIDDEN_* */
    uint32_t steps;           /* accepted integration steps */
    uint32_t attempts;        /* RK attempts including rejected ones */
    uint32_t handover;        /* mixed precision: 1 + the step at which the trajectory moved to
                               * the cooperative mode; 0 if it never did */
    double vc, r, td, fei, ast, asf, time, length_m;
} cor_traj;

/* One step of a trajectory's track: the values of one line of the reference's
 * trajectory mode files (src/trajectory.c:265-269, tragsm_* and trasph_*). */
typedef struct {
    uint32_t step;            /* accepted step number, from 1 */
    uint32_t pad;
    double gsm[3];            /* GSM position [m] */
    double r, td, fei;        /* radius [Re], GEO latitude and longitude [deg] */
    double v;                 /* speed [m/s] */
    double time;              /* flight time [s] */
    double length_m;          /* path length [m] */
} cor_track_point;

/* Per scan: the reference's P(S), P(C), P(M) = Rl, Ru, Rc (src/trajectory.c:156-160). */
typedef struct {
    double rl, ru, rc;
    int32_t n_forbidden;
    int32_t n;
} cor_cutoff;

typedef struct {
    double kernel_ms;         /* tracer kernel time of the last run (HIP events) */
    double total_ms;          /* queue reset + tracer + reduction of the last run */
    uint64_t steps;           /* accepted steps of the last run */
    uint64_t attempts;        /* RK attempts of the last run */
    int64_t n_traj;
    int32_t grid_blocks;
    int32_t block_threads;
    uint64_t wave_iters;      /* per-lane mode: wavefront loop iterations (steps), summed */
    uint64_t lane_iters;      /* per-lane mode: live lanes summed over those iterations */
    uint64_t coop_steps;      /* cooperative mode: wavefront steps (one field each) */
    uint64_t wave_iters_open; /* the part of wave_iters while the queue still had work */
    uint64_t lane_iters_open; /* the part of lane_iters while the queue still had work */
    double drain_ms;          /* first wavefront start -> the queue's last slot taken */
    double lane_end_ms;       /* first wavefront start -> last wavefront leaves per-lane mode */
} cor_trace_stats;

typedef struct cor_trace_plan cor_trace_plan;

/* Device to run on (default 0).  Returns COR_E_NO_DEVICE when there is none. */
int cor_gpu_init(int device);
const char *cor_gpu_error_string(int code);
const char *cor_gpu_last_hip_error(void);
/* Wait for all work the library queued on the device. */
int cor_gpu_synchronize(void);

/* Upload one epoch's parameters and a list of scans; allocate the result arrays. */
int cor_trace_plan_create(const cor_model_set *models, const ModelsParams *epoch,
                          const cor_scan *scans, int32_t n_scans, const cor_trace_opts *opts,
                          cor_trace_plan **plan);
/* Trace every trajectory of the plan and reduce the cutoffs, on `stream`
 * (a hipStream_t; NULL = the library's own stream).  Asynchronous. */
int cor_trace_plan_run(cor_trace_plan *plan, void *stream);
int cor_trace_plan_sync(cor_trace_plan *plan);
/* Progress of the plan's last run without waiting for it: *finished = trajectories
 * ended so far (a snapshot of a counter the running kernel advances), *done = 1 once
 * the run (tracer and reduction) is complete.  No counterpart in the reference. */
int cor_trace_plan_progress(cor_trace_plan *plan, int64_t *finished, int32_t *done);
int64_t cor_trace_plan_count(const cor_trace_plan *plan);
/* Copy results to the host (either pointer may be NULL).  traj has
 * cor_trace_plan_count() entries in scan order; cut has n_scans. */
int cor_trace_plan_results(cor_trace_plan *plan, cor_traj *traj, cor_cutoff *cut);
int cor_trace_plan_stats(cor_trace_plan *plan, cor_trace_stats *stats);
void cor_trace_plan_destroy(cor_trace_plan *plan);

/* Replace the plan's epoch (what igrf14_recalc and the model input lines fill for
 * one run of the reference, src/main.c:38-75) with the same model kinds, so the scans
 * can be re-traced at another time or for other drivers without rebuilding the plan.
 * The upload is ordered on `stream` (NULL = the library's stream): run the plan on
 * the same stream.  Results of the previous run must be copied out first. */
int cor_trace_plan_set_epoch(cor_trace_plan *plan, const cor_model_set *models,
                             const ModelsParams *epoch, void *stream);

/* create + run + results + destroy */
int cor_gpu_trace(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                  int32_t n_scans, const cor_trace_opts *opts, cor_traj *traj,
                  cor_cutoff *cut);

/* The same scans traced at each of n_epochs epochs (the reference runs its whole
 * cutoff path once per epoch: src/main.c:38-86 per input file).  Up to in_flight
 * epochs (0 = 3, at most 4) run at once on separate streams, so the slow tail of one
 * epoch overlaps the next.  Each stream's epoch is a HIP graph captured once (epoch
 * upload, lane tables, queue reset, tracer, penumbra reduction, result copies) and
 * replayed for every later epoch of that stream.  traj (nullable) receives
 * n_epochs x count entries, epoch major; cut receives n_epochs x n_scans. */
int cor_gpu_trace_series(const cor_model_set *models, const ModelsParams *epochs, int32_t n_epochs,
                         const cor_scan *scans, int32_t n_scans, const cor_trace_opts *opts,
                         int32_t in_flight, cor_traj *traj, cor_cutoff *cut);

/* The same scans at n_epochs epochs in ONE launch: the persistent tracer drains a single
 * queue of (epoch, scan, rigidity) trajectories, each lane reading its epoch's constants
 * from device memory, so the slow trajectories of every epoch run side by side instead
 * of one epoch's tail after another's (the reference runs its whole cutoff path once
 * per epoch: src/main.c:38-86 per input file).  Per-lane mode; model sets without T96
 * (IGRF, TS05, IGRF + TS05; COR_E_UNSUPPORTED otherwise).  Results are epoch-major:
 * traj has n_epochs x count entries (count = the scans' cor_scan_count sum), cut has
 * n_epochs x n_scans; each epoch's results equal a single-epoch trace of it bit for
 * bit.  The plan form (cor_trace_plan_create_epochs) runs, times and reads back like
 * any plan; cor_trace_plan_set_epoch does not apply to it. */
int cor_trace_plan_create_epochs(const cor_model_set *models, const ModelsParams *epochs, int32_t n_epochs,
                                 const cor_scan *scans, int32_t n_scans, const cor_trace_opts *opts,
                                 cor_trace_plan **plan);
int cor_gpu_trace_epochs(const cor_model_set *models, const ModelsParams *epochs, int32_t n_epochs,
                         const cor_scan *scans, int32_t n_scans, const cor_trace_opts *opts, cor_traj *traj,
                         cor_cutoff *cut);
/* cor_gpu_trace for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  No counterpart in
 * the reference.  Nothing is allocated: the epoch block, the scan descriptors and the
 * run's counters live in a fixed scratch array of about 0.5 MB that each GPU holds for
 * the life of the process; the caller passes device buffers for the results:
 *   d_traj  cor_scan_count summed over the scans cor_traj records, 16-byte aligned
 *           (COR_E_INVALID otherwise), in scan order, rigidity ascending;
 *   d_cut   n_scans cor_cutoff.
 * At most COR_NOALLOC_MAX_SCANS scans (COR_E_INVALID beyond); no crustal model
 * (COR_E_UNSUPPORTED).  Always the per-lane mode, whose results equal cor_gpu_trace's
 * with opts->coop < 0 bit for bit; opts->queue_order is ignored (it changes no result).
 * Runs on `stream` (NULL = the library's stream) and returns when the results are in
 * place.  Every call on a GPU uses that GPU's one scratch array: calls that overlap on
 * one GPU race. */
#define COR_NOALLOC_MAX_SCANS 4096
int cor_gpu_trace_noalloc(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                          int32_t n_scans, const cor_trace_opts *opts, cor_traj *d_traj, cor_cutoff *d_cut,
                          void *stream);

/* Per-step tracks of every trajectory of the scans in FP64, as the reference's
 * trajectory mode writes them (trajectory_simulation, src/trajectory.c:163-273): one
 * point per accepted step that does not end the trajectory.  counts (nullable) receives
 * the number of points of each trajectory (cor_scan_count entries per scan, scan
 * order) and *total their sum; points receives them, trajectory after trajectory.  With
 * points == NULL only the counts (upper bounds: a trajectory whose state turns NaN stops
 * there, where the reference would print NaN lines up to the step limit) are returned;
 * COR_E_NOMEM when capacity < *total.  opts->precision must be COR_PREC_FP64. */
int cor_gpu_trace_tracks(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                         int32_t n_scans, const cor_trace_opts *opts, int64_t *counts, int64_t *total,
                         cor_track_point *points, int64_t capacity);

/* Autotuner (no counterpart in the reference, which has no launch parameters): picks
 * the persistent tracer's launch parameters for a workload by timing candidate launches
 * on a seeded sample of its scans (sample_scans of them, at most min(step_limit, 1e5)
 * steps per trajectory so that the throughput phase dominates the timing):
 *   waves_per_simd   1 .. the kernel's register limit (resident wavefronts per SIMD);
 *   block_threads    64 or 256 threads per workgroup (one or four wavefronts sharing
 *                    a workgroup's LDS-staged epoch tables);
 *   queue_order      rigidity order or longest-expected first.
 * Only parameters that change no result are tuned (handover_steps, which moves the mixed
 * path's switch to the cooperative arithmetic, is left as given).  *best receives *base
 * with the fastest candidate's values; *best_ms (nullable) its tracer time on the sample. */
int cor_trace_autotune(const cor_model_set *models, const ModelsParams *epoch, const cor_scan *scans,
                       int32_t n_scans, const cor_trace_opts *base, int32_t sample_scans, uint64_t seed,
                       cor_trace_opts *best, double *best_ms);
/* calculate_field over n points: GSM position [m] (n x 3) -> GEO field [nT] (n x 3). */
int cor_gpu_field(const cor_model_set *models, const ModelsParams *epoch, const double *gsm_m,
                  double *geo_nT, int64_t n, int32_t precision);

/* calculate_position_and_speed (src/trajectory.c:275-356, declared src/trajectory.h:9)
 * over n states: one RK step each with the tracer's FP64 step.  h[n] [s], d[n]
 * (1e-9 zn q / hm), field n x 3 GEO [nT]; vel and pos n x 3 [m/s, m] advanced in
 * place. */
int cor_gpu_rk_step(const double *h, const double *d, const double *field, double *vel, double *pos,
                    int64_t n);

/* round((rk - rig)/del) + 1 as src/trajectory.c:7 computes it, or COR_E_INVALID. */
int64_t cor_scan_count(const cor_scan *scan);

/* Site-list helpers (FP64, host).  No counterpart in the reference, which takes a
 * geocentric station and a GEO-angle direction; they convert WGS84 geodetic
 * latitude/height [deg, km] to the reference's r0 [Re = 6371.2 km] and geocentric
 * latitude, and a local zenith/azimuth [deg, azimuth clockwise from north] to the
 * reference's (the1, fi1). */
int cor_geodetic_to_geocentric(double lat_deg, double h_km, double *r0_re, double *lat_c_deg);
int cor_direction_to_geo(double lat_deg, double lon_deg, double zenith_deg, double azimuth_deg,
                         double *the1_deg, double *fi1_deg);

/* Penumbra refinement (no counterpart in the reference, which is re-run with a finer
 * del by hand): the scan covering [Rl - del, Ru] of a traced scan at del / factor,
 * clamped to the coarse scan's [rig, rk]. */
int cor_refine_scan(const cor_scan *coarse, const cor_cutoff *cut, int32_t factor, cor_scan *fine);

#ifdef __cplusplus
}
#endif

#endif /* COR_GPU_H */
