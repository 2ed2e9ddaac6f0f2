/*
 * models.c -- the field-model plugin registry and the field aggregate of the reference
 * API (src/models.c:22-207), with every evaluation routed to the GPU.
 *
 * setup_fun keeps the reference's contract (read the model's input lines / compute the
 * epoch block on the host; 0 = OK).  calc_fun and calculate_field evaluate one point
 * through the batch kernel; batch callers use cor_gpu_field directly.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "cor_host.h"

static int internal_setup(FILE *input_file, const InfileHeader *h, ModelsParams *p, const void *data) {
    (void)input_file;
    (void)data;
    const RecalcTime t = {h->iy, h->mes, h->ide, h->id, h->ih, h->min, h->is};
    return igrf14_recalc(&t, p->G, p->H, p->REC, p->A1, p->A2, p->A3);
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

/* One model at one point on the GPU, added into the GSM accumulator
 * (src/models.c:146-192).  A GPU failure yields NaN, which the tracer semantics
 * already handle (the reference's only failure signal is a NaN field). */
static void calc_one(CorModelKind kind, const Model *self, Vector gsm_re, Vector *total,
                     const ModelsParams *params) {
    const Model *one[2] = {self, NULL};
    (void)kind;
    Vector pos_m = {{gsm_re.x * DEF_RE, gsm_re.y * DEF_RE, gsm_re.z * DEF_RE}};
    Vector geo;
    calculate_field(pos_m, &geo, one, (ModelsParams *)params);
    Vector gsm;
    geo_to_gsm(&geo, &gsm, params->A1, params->A2, params->A3);
    total->x += gsm.x;
    total->y += gsm.y;
    total->z += gsm.z;
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

static const Model igrf_14 = {
    .name = "IGRF-14",
    .version = "igrf14",
    .setup_data = NULL,
    .setup_fun = internal_setup,
    .calc_fun = internal_calc,
    .kind = COR_MODEL_INTERNAL_IGRF,
};

static const Model ts_05 = {
    .name = "Tsyganenko-Sitnov 05",
    .version = "ts05",
    .setup_data = NULL,
    .setup_fun = ts05_setup,
    .calc_fun = ts05_calc,
    .kind = COR_MODEL_EXTERNAL_TS05,
};

static const Model t_96 = {
    .name = "Tsyganenko 96",
    .version = "t96",
    .setup_data = NULL,
    .setup_fun = t96_setup,
    .calc_fun = t96_calc,
    .kind = COR_MODEL_EXTERNAL_T96,
};

/* src/models.c:102-108: no input lines (setup_fun NULL) */
static const Model chaos_7_16 = {
    .name = "CHAOS-7.16",
    .version = "chaos7.16",
    .setup_data = NULL,
    .setup_fun = NULL,
    .calc_fun = chaos716_calc,
    .kind = COR_MODEL_CRUSTAL_SH,
    .sh_nm_max = COR_CHAOS716_NM,
    .sh_g = chaos716_g,
    .sh_h = chaos716_h,
};

static void internal_calc(Vector g, Vector *t, const ModelsParams *p) { calc_one(COR_MODEL_INTERNAL_IGRF, &igrf_14, g, t, p); }
static void ts05_calc(Vector g, Vector *t, const ModelsParams *p) { calc_one(COR_MODEL_EXTERNAL_TS05, &ts_05, g, t, p); }
static void t96_calc(Vector g, Vector *t, const ModelsParams *p) { calc_one(COR_MODEL_EXTERNAL_T96, &t_96, g, t, p); }
static void chaos716_calc(Vector g, Vector *t, const ModelsParams *p) {
    calc_one(COR_MODEL_CRUSTAL_SH, &chaos_7_16, g, t, p);
}

/* src/models.c:110-112 (the IGRF-14 chain is the internal model built here) */
const Model *internal_models[] = {&igrf_14, NULL};
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
        switch (models[i]->kind) {
            case COR_MODEL_INTERNAL_IGRF: set->internal = COR_INTERNAL_IGRF; break;
            case COR_MODEL_EXTERNAL_T96: set->external = COR_EXTERNAL_T96; break;
            case COR_MODEL_EXTERNAL_TS05: set->external = COR_EXTERNAL_TS05; break;
            case COR_MODEL_CRUSTAL_SH:
                set->crustal = COR_CRUSTAL_SH;
                set->sh_nm_max = models[i]->sh_nm_max;
                set->sh_g = models[i]->sh_g;
                set->sh_h = models[i]->sh_h;
                break;
            default: return COR_E_INVALID;
        }
    }
    return COR_OK;
}

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
