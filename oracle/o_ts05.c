/*
 * o_ts05.c -- oracle: Tsyganenko-Sitnov 2005 external field (zero tilt).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Not restated yet: returns NaN so that no test can pass against it by accident.
 */
#include <math.h>
#include "oracle.h"

void o_ts05_gsm(const double *parmod, double x, double y, double z, double b[3]) {
    (void)parmod; (void)x; (void)y; (void)z;
    b[0] = b[1] = b[2] = NAN;
}
