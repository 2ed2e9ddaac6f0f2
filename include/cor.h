/*
 * cor.h -- the reference cutoff-rigidity tool's C API, served by the MI355X build.
 *
 * Same types, names, argument meaning and error behaviour as the reference's
 * src/common.h, src/utils.h, src/models.h, src/io.h, src/trajectory.h and
 * src/field.h, so a caller written against the reference links against
 * libcor_amd.so unchanged.  The compute behind trajectory_simulation_cutoff,
 * field_simulation and calculate_field runs on the GPU (cor_gpu.h); the host keeps
 * the input parsing, per-epoch model setup and the text output.
 */
#ifndef COR_H
#define COR_H

#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/common.h:4-11 (values as the reference writes them) */
#define DEF_PI 3.141592654
#define DEF_2PI 6.28318531
#define DEF_RE 6.3712e+06
#define DEF_TO_RAD (DEF_PI / 180.)
#define DEF_Q 1.6021e-19
#define DEF_C 2.99725e+08
#define DEF_HM0_E 9.1093826e-31
#define DEF_HM0_P 1.6725e-27

/* lib/internal/coefs/coefs.h:7 */
#define INTERNAL_COEFS_SIZE 105

/* src/common.h:13-24 */
typedef union {
    struct {
        double x;
        double y;
        double z;
    };
    struct {
        double r;
        double theta;
        double phi;
    };
} Vector;

/* src/common.h:26-54 */
typedef struct {
    double r0;
    double the0;
    double fi0;
    double the1;
    double fi1;
    int iy;
    int mes;
    int ide;
    int id;
    int ih;
    int min;
    int is;
    union {
        struct {
            double zn;
            int nk1;
            int iopt;
            int ist;
            double rig;
            double del;
            double rk;
        };
        struct {
            double lat_step;
            double lon_step;
        };
    };
} InfileHeader;

/* src/models.h:9-17 */
typedef struct {
    double PARMOD[10];
    double G[INTERNAL_COEFS_SIZE];
    double H[INTERNAL_COEFS_SIZE];
    double REC[INTERNAL_COEFS_SIZE];
    double A1[3];
    double A2[3];
    double A3[3];
} ModelsParams;

/* src/models.h:19-25.  The registries' models are routed to the GPU evaluators by
 * identity (their calc_fun); a caller-defined model has no GPU evaluator and makes
 * the GPU entry points return COR_E_UNSUPPORTED. */
typedef struct {
    const char *name;
    const char *version;
    const void *setup_data;
    int (*setup_fun)(FILE *input_file, const InfileHeader *infile_header,
                     ModelsParams *model_params, const void *setup_data);
    void (*calc_fun)(Vector gsm_pos, Vector *total_gsm_field, const ModelsParams *params);
} Model;

/* src/models.c:110-112: NULL-terminated registries */
extern const Model *external_models[];
extern const Model *internal_models[];
extern const Model *crustal_models[];

/* src/models.c:114-122 */
const Model *get_model(const Model **models, const char *version);

/* src/models.c:194-207.  One point on the GPU (a batch of one); use
 * cor_gpu_field for many points. */
void calculate_field(Vector gsm_pos, Vector *total_geo_field, const Model **models,
                     ModelsParams *params);

/* src/utils.h:9-12 */
void sph_to_car(Vector *sph, Vector *car);
void car_to_sph(Vector *car, Vector *sph);
void geo_to_gsm(Vector *geo, Vector *gsm, const double *A1, const double *A2, const double *A3);
void gsm_to_geo(Vector *gsm, Vector *geo, const double *A1, const double *A2, const double *A3);

/* lib/internal/coefs/coefs.h:9-16: a model's chain of Gauss-coefficient rows (G and H
 * alternating per epoch, the secular-variation pair last where it extrapolates). */
typedef struct {
    const double **coefs;
    uint32_t coefs_size;
    uint32_t years_step;
    uint32_t min_year;
    uint32_t max_year;
    bool extrapolate;
} InternalModelCoefs;

/* lib/internal/internal.h:12-20 */
typedef struct {
    int iy;
    int mes;
    int ide;
    int id;
    int ih;
    int min;
    int is;
} RecalcTime;

/* lib/internal/internal.h:22-29: the internal models' tables */
extern const InternalModelCoefs IGRF_9_COEFS;
extern const InternalModelCoefs IGRF_10_COEFS;
extern const InternalModelCoefs IGRF_11_COEFS;
extern const InternalModelCoefs IGRF_12_COEFS;
extern const InternalModelCoefs IGRF_13_COEFS;
extern const InternalModelCoefs IGRF_14_COEFS;
extern const InternalModelCoefs HIST_COEFS;
extern const InternalModelCoefs CALS10K_2_COEFS;

/* lib/internal/internal.h:31.  Epoch setup on the host: recursion-scaled G/H, REC and
 * the GEO->GSM rows A1..A3; -1 for an epoch outside the model's years. */
int internal_recalc(const InternalModelCoefs *coefs, const RecalcTime *time, double *G, double *H,
                    double *REC, double *A1, double *A2, double *A3);

/* src/utils.h:7-8: strtoul / strtod of a whole string, 0 or -1 */
int string_to_uint64(char *str_num, uint64_t *num);
int string_to_double(char *str_num, double *num);

/* src/io.h:8-13 */
int read_cutoff_infile_header(FILE *input_file, InfileHeader *infile_header);
int write_cutoff_outfile_header(FILE *output_file, InfileHeader *infile_header,
                                const Model **models, uint64_t steps_count);
int read_field_infile_header(FILE *input_file, InfileHeader *infile_header);
int write_field_outfile_header(FILE *output_file, InfileHeader *infile_header,
                               const Model **models);
int read_ts05_params(FILE *input_file, ModelsParams *params);
int read_t96_params(FILE *input_file, ModelsParams *params);

/* src/trajectory.h:9.  One step of the reference's 5th-order Runge-Kutta scheme with
 * the field frozen (src/trajectory.c:275-356), on the GPU with the tracer's own FP64
 * step: h [s], d = 1e-9 zn q / hm, GEO field [nT]; velocity [m/s] and position [m]
 * are advanced in place.  GPU errors are reported on stderr and leave both unchanged. */
void calculate_position_and_speed(double h, double d, Vector geo_field, Vector *car_vel, Vector *geo_pos);

/* src/trajectory.h:10.  The whole rigidity scan is traced on the GPU in FP64 and the
 * output is written exactly as the reference writes it.  `parallel` is accepted and
 * ignored (the GPU is always parallel).  GPU errors are reported on stderr. */
void trajectory_simulation_cutoff(InfileHeader *header, const Model **models,
                                  ModelsParams *models_params, FILE *outfile,
                                  uint64_t step_limit, double tu_angle, double particle_mass,
                                  uint32_t parallel);

/* src/trajectory.h:11.  Every trajectory of the header's scan (the reference's
 * trajectory mode has rig == rk: one) is traced on the GPU in FP64 and its track is
 * written to tragsm (GSM position) and trasph (radius, latitude, longitude) line by line
 * as the reference writes them.  GPU errors are reported on stderr. */
void trajectory_simulation(InfileHeader *header, const Model **models, ModelsParams *models_params,
                           FILE *tragsm, FILE *trasph, uint64_t step_limit, double tu_angle,
                           double particle_mass);

/* src/field.h:7 */
void field_simulation(InfileHeader *header, ModelsParams *params, const Model **models,
                      FILE *outfile, uint32_t parallel);

#ifdef __cplusplus
}
#endif

#endif /* COR_H */
