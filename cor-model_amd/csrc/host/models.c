/*
 * models.c -- the field-model plugin registry and the field aggregate of the reference
 * API (src/models.c:22-207), with every evaluation routed to the GPU.
 *
 * setup_fun keeps the reference's contract (read the model's input lines / compute the
 * epoch block on the host; 0 = OK).  calc_fun and calculate_field evaluate one point
 * through the batch kernel; batch callers use cor_gpu_field directly.  A model list
 * is mapped to the GPU evaluators by the identity of each model's calc_fun: the
 * registries' models (and copies of them) have a GPU evaluator, a caller-defined one
 * does not, and the GPU entry points then refuse the list (COR_E_UNSUPPORTED).
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "cor_host.h"

/* src/models.c:124-136: the model's own table at the header's epoch */
static int internal_setup(FILE *input_file, const InfileHeader *h, ModelsParams *p, const void *data) {
    (void)input_file;
    const RecalcTime t = {h->iy, h->mes, h->ide, h->id, h->ih, h->min, h->is};
    return internal_recalc((const InternalModelCoefs *)data, &t, p->G, p->H, p->REC, p->A1, p->A2, p->A3);
}

static int ts05_setup(FILE *input_file, const InfileHeader *h, ModelsParams *p, const void *data) {
    (void)h;
    (void)data;
    return read_ts05_params(input_file, p);
}

static int t96_setup(FILE *input_file, const InfileHeader *h, ModelsParams *p, const void *data) {
    (void)h;
    (void)data;
    return read_t96_params(input_file, p);
}

static void internal_calc(Vector gsm_pos, Vector *total, const ModelsParams *params);
static void ts05_calc(Vector gsm_pos, Vector *total, const ModelsParams *params);
static void t96_calc(Vector gsm_pos, Vector *total, const ModelsParams *params);
static void chaos716_calc(Vector gsm_pos, Vector *total, const ModelsParams *params);

/* The CHAOS-7.16 tables in the [186][186] layout of a model set (row n, column m, zero
 * for m > n), expanded from the stored triangle when the library is loaded. */
static double chaos716_g[COR_CHAOS716_NM * COR_CHAOS716_NM];
static double chaos716_h[COR_CHAOS716_NM * COR_CHAOS716_NM];

__attribute__((constructor)) static void chaos716_expand(void) {
    int k = 0;
    for (int n = 0; n < COR_CHAOS716_NM; n++)
        for (int m = 0; m <= n; m++, k++) {
            chaos716_g[n * COR_CHAOS716_NM + m] = cor_chaos716_tri[0][k];
            chaos716_h[n * COR_CHAOS716_NM + m] = cor_chaos716_tri[1][k];
        }
}

/* src/models.c:22-108 */
#define INTERNAL_MODEL(var, nm, ver, tab) \
    static const Model var = {.name = nm, .version = ver, .setup_data = &tab, .setup_fun = internal_setup, .calc_fun = internal_calc}
INTERNAL_MODEL(igrf_9, "IGRF-9", "igrf9", IGRF_9_COEFS);
INTERNAL_MODEL(igrf_10, "IGRF-10", "igrf10", IGRF_10_COEFS);
INTERNAL_MODEL(igrf_11, "IGRF-11", "igrf11", IGRF_11_COEFS);
INTERNAL_MODEL(igrf_12, "IGRF-12", "igrf12", IGRF_12_COEFS);
INTERNAL_MODEL(igrf_13, "IGRF-13", "igrf13", IGRF_13_COEFS);
INTERNAL_MODEL(igrf_14, "IGRF-14", "igrf14", IGRF_14_COEFS);
INTERNAL_MODEL(hist, "Historical (0-1968 CE)", "hist", HIST_COEFS);
INTERNAL_MODEL(cals10k_2, "Historical (CALS10k.2)", "cals10k.2", CALS10K_2_COEFS);
#undef INTERNAL_MODEL

static const Model ts_05 = {
    .name = "Tsyganenko-Sitnov 05",
    .version = "ts05",
    .setup_data = NULL,
    .setup_fun = ts05_setup,
    .calc_fun = ts05_calc,
};

static const Model t_96 = {
    .name = "Tsyganenko 96",
    .version = "t96",
    .setup_data = NULL,
    .setup_fun = t96_setup,
    .calc_fun = t96_calc,
};

/* src/models.c:102-108: no input lines (setup_fun NULL) */
static const Model chaos_7_16 = {
    .name = "CHAOS-7.16",
    .version = "chaos7.16",
    .setup_data = NULL,
    .setup_fun = NULL,
    .calc_fun = chaos716_calc,
};

/* src/models.c:110-112 */
const Model *internal_models[] = {&igrf_14, &igrf_13, &igrf_12, &igrf_11, &igrf_10, &igrf_9, &hist, &cals10k_2, NULL};
const Model *external_models[] = {&ts_05, &t_96, NULL};
const Model *crustal_models[] = {&chaos_7_16, NULL};

/* src/models.c:114-122 */
const Model *get_model(const Model **models, const char *version) {
    for (int i = 0; models[i] != NULL; i++) {
        if (strcmp(models[i]->version, version) == 0) return models[i];
    }
    return NULL;
}

int cor_models_to_set(const Model **models, cor_model_set *set) {
    memset(set, 0, sizeof *set);
    for (int i = 0; models && models[i]; i++) {
        void (*f)(Vector, Vector *, const ModelsParams *) = models[i]->calc_fun;
        if (f == internal_calc) {
            set->internal = COR_INTERNAL_IGRF;
        } else if (f == t96_calc) {
            set->external = COR_EXTERNAL_T96;
        } else if (f == ts05_calc) {
            set->external = COR_EXTERNAL_TS05;
        } else if (f == chaos716_calc) {
            set->crustal = COR_CRUSTAL_SH;
            set->sh_nm_max = COR_CHAOS716_NM;
            set->sh_g = chaos716_g;
            set->sh_h = chaos716_h;
        } else {
            char msg[200];
            snprintf(msg, sizeof msg, "model '%s' has no GPU evaluator (only the registries' models are served)",
                     models[i]->name ? models[i]->name : "(unnamed)");
            cor_set_error(msg);
            return COR_E_UNSUPPORTED;
        }
    }
    return COR_OK;
}

/* One model at one point on the GPU, added into the GSM accumulator
 * (src/models.c:146-192).  A GPU failure yields NaN, which the tracer semantics
 * already handle (the reference's only failure signal is a NaN field). */
static void calc_one(const Model *self, Vector gsm_re, Vector *total, const ModelsParams *params) {
    const Model *one[2] = {self, NULL};
    Vector pos_m = {{gsm_re.x * DEF_RE, gsm_re.y * DEF_RE, gsm_re.z * DEF_RE}};
    Vector geo;
    calculate_field(pos_m, &geo, one, (ModelsParams *)params);
    Vector gsm;
    geo_to_gsm(&geo, &gsm, params->A1, params->A2, params->A3);
    total->x += gsm.x;
    total->y += gsm.y;
    total->z += gsm.z;
}

/* every internal model's field is the degree-13 synthesis of the epoch's G/H/REC */
static void internal_calc(Vector g, Vector *t, const ModelsParams *p) { calc_one(&igrf_14, g, t, p); }
static void ts05_calc(Vector g, Vector *t, const ModelsParams *p) { calc_one(&ts_05, g, t, p); }
static void t96_calc(Vector g, Vector *t, const ModelsParams *p) { calc_one(&t_96, g, t, p); }
static void chaos716_calc(Vector g, Vector *t, const ModelsParams *p) { calc_one(&chaos_7_16, g, t, p); }

/* src/models.c:194-207: GSM position [m] -> GEO field [nT], on the GPU (FP64). */
void calculate_field(Vector gsm_pos, Vector *total_geo_field, const Model **models, ModelsParams *params) {
    cor_model_set set;
    double in[3] = {gsm_pos.x, gsm_pos.y, gsm_pos.z}, out[3];
    int rc = cor_models_to_set(models, &set);
    if (rc == COR_OK) rc = cor_gpu_field(&set, params, in, out, 1, COR_PREC_FP64);
    if (rc != COR_OK) {
        fprintf(stderr, "Error: GPU field evaluation failed: %s (%s)\n", cor_gpu_error_string(rc),
                cor_gpu_last_hip_error());
        out[0] = out[1] = out[2] = NAN;
    }
    total_geo_field->x = out[0];
    total_geo_field->y = out[1];
    total_geo_field->z = out[2];
}
