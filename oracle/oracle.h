/*
 * oracle.h -- CPU restatement of the reference cutoff-rigidity tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is compared
 * with: only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it.  The product never links, includes or calls anything declared here.
 *
 * Every function restates one reference function (file:line given beside it) in
 * plain C17, compiled with the reference's flags (no FMA contraction), so that on the
 * same libm its FP64 results are meant to be bit-identical to the reference's.
 */
#ifndef COR_ORACLE_H
#define COR_ORACLE_H

#include <stdint.h>

#define O_NCOEF 105          /* lib/internal/coefs/coefs.h:7 */
#define O_IGRF14_NROWS 54    /* lib/internal/coefs/igrf_14_coefs.h:90 */

/* Physical constants exactly as src/common.h:4-11 writes them. */
#define O_PI 3.141592654
#define O_2PI 6.28318531
#define O_RE 6.3712e+06
#define O_TO_RAD (O_PI / 180.)
#define O_Q 1.6021e-19
#define O_C 2.99725e+08
#define O_M_E 9.1093826e-31
#define O_M_P 1.6725e-27
#define O_TO_DEG 57.295779513   /* lib/internal/internal.h:6 */

/* Same field order as the reference ModelsParams (src/models.h:9-17). */
typedef struct {
    double PARMOD[10];
    double G[O_NCOEF];
    double H[O_NCOEF];
    double REC[O_NCOEF];
    double A1[3];
    double A2[3];
    double A3[3];
} o_params;

enum { O_INT_NONE = 0, O_INT_IGRF = 1 };
enum { O_EXT_NONE = 0, O_EXT_T96 = 1, O_EXT_TS05 = 2 };
enum { O_CRU_NONE = 0, O_CRU_CHAOS = 1 };

/* Which evaluators calculate_field sums (internal, external, crustal order as
 * src/main.c:268-299 builds the model list). */
typedef struct {
    int32_t internal;
    int32_t external;
    int32_t crustal;
    int32_t chaos_nm_max;          /* rows/cols of the CHAOS G/H tables */
    const double *chaos_g;         /* [nm_max][nm_max], nT */
    const double *chaos_h;
} o_models;

/* One station x one direction x one rigidity scan, the content of the reference
 * input header (src/io.c:4-20). */
typedef struct {
    double r0, the0, fi0;          /* station radius [Re], geocentric lat, lon [deg] */
    double the1, fi1;              /* arrival direction as GEO lat, lon [deg] */
    double rig, del, rk;           /* first, step and last rigidity [GV] */
    double zn;                     /* charge number, -1 back-traces a proton */
    int32_t nk1;                   /* steps per gyro-period each step starts from */
    int32_t pad;
} o_scan;

enum { O_ALLOWED = 0, O_FORBIDDEN_REENTRY = 1, O_FORBIDDEN_LIMIT = 2 };

typedef struct {
    int32_t status;                /* O_ALLOWED / O_FORBIDDEN_* */
    uint32_t steps;                /* accepted steps when the trajectory ended */
    uint32_t attempts;             /* RK attempts, rejected ones included */
    uint32_t pad;
    double vc, r, td, fei, ast, asf, time, length_m;   /* valid when allowed */
} o_traj;

/* ---- transforms: src/utils.c:28-101 ---- */
void o_sph_to_car(const double sph[3], double car[3]);
void o_car_to_sph(const double car[3], double sph[3]);
void o_geo_to_gsm(const double geo[3], double gsm[3], const o_params *p);
void o_gsm_to_geo(const double gsm[3], double geo[3], const o_params *p);

/* ---- internal field ---- */
extern const double o_igrf14_rows[O_IGRF14_NROWS][O_NCOEF];
int o_igrf14_recalc(int iy, int mes, int ide, int id, int ih, int min, int is, o_params *p);
void o_sun(int iy, int iday, int ih, int min, int is, double out[4]);
void o_igrf_geo(double x, double y, double z, const double *G, const double *H,
                const double *REC, double b[3]);

/* ---- external fields ---- */
void o_t96_gsm(const double *parmod, double x, double y, double z, double b[3]);
void o_ts05_gsm(const double *parmod, double x, double y, double z, double b[3]);

/* ---- crustal / generic SH ---- */
void o_chaos_geo(double x, double y, double z, int nm_max, const double *g,
                 const double *h, double b[3]);

/* ---- aggregate: src/models.c:194-207 ---- */
void o_calculate_field(const double gsm_pos_m[3], double geo_field_nT[3],
                       const o_models *m, const o_params *p);
void o_field_batch(const o_models *m, const o_params *p, const double *gsm_pos_m,
                   double *geo_field_nT, int64_t n);

/* ---- tracer: src/trajectory.c:6-161 ---- */
int32_t o_scan_count(const o_scan *s);
void o_trace_one(const o_scan *s, int32_t i, const o_models *m, const o_params *p,
                 uint64_t step_limit, double tu_angle, double mass, o_traj *out);
void o_trace_scan(const o_scan *s, const o_models *m, const o_params *p,
                  uint64_t step_limit, double tu_angle, double mass, int32_t threads,
                  o_traj *out);
void o_penumbra(const o_scan *s, const o_traj *t, int32_t n, double cut[3]);
void o_rk_step(double h, double d, const double field[3], double vel[3], double pos[3]);

#endif /* COR_ORACLE_H */
