/*
 * utils.c -- number parsing and coordinate transforms of the reference API
 * (src/utils.c:8-101).  Host-side: command-line values, per-epoch setup, grid
 * construction and output geometry.
 */
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include "cor_host.h"

/* src/utils.c:8-16: the whole string as a base-10 unsigned number.  The failure test
 * is the reference's (nothing parsed, trailing characters, or an ERANGE result equal to
 * LONG_MIN / LONG_MAX with errno as the caller left it), so values are accepted and
 * refused exactly as the reference's command line does. */
int string_to_uint64(char *str_num, uint64_t *num) {
    char *end = NULL;
    *num = strtoul(str_num, &end, 10);
    const int range = (*num == (uint64_t)LONG_MIN || *num == (uint64_t)LONG_MAX) && errno == ERANGE;
    return (end == str_num || *end != '\0' || range) ? -1 : 0;
}

/* src/utils.c:18-26: the whole string as a double, with the same failure test. */
int string_to_double(char *str_num, double *num) {
    char *end = NULL;
    *num = strtod(str_num, &end);
    const int range = (*num == (double)LONG_MIN || *num == (double)LONG_MAX) && errno == ERANGE;
    return (end == str_num || *end != '\0' || range) ? -1 : 0;
}

/* src/utils.c:28-41 */
void sph_to_car(Vector *sph, Vector *car) {
    const double rs = sph->r * sin(sph->theta);
    const double x = rs * cos(sph->phi);
    const double y = rs * sin(sph->phi);
    const double z = sph->r * cos(sph->theta);
    car->x = x;
    car->y = y;
    car->z = z;
}

/* src/utils.c:43-73: phi in [0, 2pi) with the truncated 2pi; polar axis -> theta 0/pi */
void car_to_sph(Vector *car, Vector *sph) {
    const double x = car->x, y = car->y, z = car->z;
    const double rho2 = x * x + y * y;
    const double r = sqrt(rho2 + z * z);
    if (rho2 != 0.) {
        double phi = atan2(y, x);
        const double theta = atan2(sqrt(rho2), z);
        if (phi < 0.) phi += DEF_2PI;
        sph->r = r;
        sph->theta = theta;
        sph->phi = phi;
        return;
    }
    sph->r = r;
    sph->theta = z < 0. ? DEF_PI : 0.;
    sph->phi = 0.;
}

/* src/utils.c:75-87 */
void geo_to_gsm(Vector *geo, Vector *gsm, const double *A1, const double *A2, const double *A3) {
    const double x = geo->x, y = geo->y, z = geo->z;
    const double nx = A1[0] * x + A1[1] * y + A1[2] * z;
    const double ny = A2[0] * x + A2[1] * y + A2[2] * z;
    const double nz = A3[0] * x + A3[1] * y + A3[2] * z;
    gsm->x = nx;
    gsm->y = ny;
    gsm->z = nz;
}

/* src/utils.c:89-101 */
void gsm_to_geo(Vector *gsm, Vector *geo, const double *A1, const double *A2, const double *A3) {
    const double x = gsm->x, y = gsm->y, z = gsm->z;
    const double nx = A1[0] * x + A2[0] * y + A3[0] * z;
    const double ny = A1[1] * x + A2[1] * y + A3[1] * z;
    const double nz = A1[2] * x + A2[2] * y + A3[2] * z;
    geo->x = nx;
    geo->y = ny;
    geo->z = nz;
}
