/*
 * field.c -- field_simulation of the reference API (src/field.c:6-54): the lat/lon
 * grid is built on the host exactly as the reference walks it (latitudes 0..n_the,
 * longitudes 0..n_phi-1), evaluated in one GPU batch, and printed in the same order
 * and format.
 */
#include <math.h>
#include <stdlib.h>
#include "cor_host.h"

void field_simulation(InfileHeader *header, ModelsParams *params, const Model **models, FILE *outfile,
                      uint32_t parallel) {
    (void)parallel;
    const double del_the = header->the1 - header->the0;
    const double del_phi = header->fi1 - header->fi0;
    int n_the = 0, n_phi = 0;
    if (header->lat_step > 0.) n_the = (int)(del_the / header->lat_step);
    if (header->lon_step > 0.) n_phi = (int)(del_phi / header->lon_step);
    if (n_the < 0 || n_phi <= 0) return;
    const size_t n = (size_t)(n_the + 1) * (size_t)n_phi;
    double *pos = (double *)malloc(sizeof(double) * 3 * n);
    double *fld = (double *)malloc(sizeof(double) * 3 * n);
    cor_model_set set;
    int rc = (pos && fld) ? cor_models_to_set(models, &set) : COR_E_NOMEM;
    if (rc == COR_OK) {
        size_t k = 0;
        for (int j = 0; j <= n_the; j++) {
            const double the0a = header->the0 + (double)(j * header->lat_step);
            for (int i = 0; i < n_phi; i++) {
                const double phi0a = header->fi0 + (double)(i * header->lon_step);
                Vector sph = {{header->r0 * DEF_RE, (90. - the0a) * DEF_TO_RAD, phi0a * DEF_TO_RAD}};
                Vector car, gsm;
                sph_to_car(&sph, &car);
                geo_to_gsm(&car, &gsm, params->A1, params->A2, params->A3);
                pos[3 * k] = gsm.x;
                pos[3 * k + 1] = gsm.y;
                pos[3 * k + 2] = gsm.z;
                k++;
            }
        }
        rc = cor_gpu_field(&set, params, pos, fld, (int64_t)n, COR_PREC_FP64);
    }
    if (rc != COR_OK) {
        fprintf(stderr, "Error: GPU field evaluation failed: %s (%s)\n", cor_gpu_error_string(rc), cor_gpu_last_hip_error());
    } else {
        size_t k = 0;
        for (int j = 0; j <= n_the; j++) {
            const double the0a = header->the0 + (double)(j * header->lat_step);
            for (int i = 0; i < n_phi; i++) {
                const double phi0a = header->fi0 + (double)(i * header->lon_step);
                const double *f = fld + 3 * k++;
                const double b = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
                fprintf(outfile, "%9.2f%9.2f%11.2f%11.2f%11.2f%11.2f\n", the0a, phi0a, b, f[0], f[1], f[2]);
            }
        }
    }
    free(pos);
    free(fld);
}
