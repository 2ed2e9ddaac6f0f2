/*
 * internal.c -- per-epoch setup of the internal field models on the host
 * (lib/internal/internal.c:12-180, internal_recalc).  Runs once per epoch in FP64; its
 * outputs (recursion-scaled G/H, REC, GEO->GSM rows) are uploaded to the GPU, whose
 * IGRF-form synthesis serves every internal model (IGRF-9..14, the historical model
 * and CALS10k.2: the same degree-13 triangle, different tables).
 */
#include <math.h>
#include <stdint.h>
#include "cor_host.h"

/* lib/internal/internal.c:38-62: Russell (1971) sun position and sidereal time. */
static void sun_position(int iy, int iday, int ih, int min, int is, double *gst, double *slong,
                         double *srasn, double *sdec) {
    const double fday = (ih * 3600 + min * 60 + is) / 86400.;
    const double dj = 365 * (iy - 1900) + (iy - 1901) / 4 + iday - 0.5 + fday;
    const double t = dj / 36525.;
    *gst = fmod(279.690983 + 0.9856473354 * dj + 360. * fday + 180., 360.) / COR_TO_DEG;
    const double vl = fmod(279.696678 + 0.9856473354 * dj, 360.);
    const double g = fmod(358.475845 + 0.985600267 * dj, 360.) / COR_TO_DEG;
    double sl = (vl + (1.91946 - 0.004789 * t) * sin(g) + 0.020094 * sin(2. * g)) / COR_TO_DEG;
    if (sl > 6.2831853) {
        sl = sl - 6.2831853;
    } else if (sl < 0) {
        sl = sl + 6.2831853;
    }
    *slong = sl;
    const double obliq = (23.45229 - 0.0130125 * t) / COR_TO_DEG;
    const double slp = sl - 9.924E-5;
    const double sob = sin(obliq);
    const double sind = sob * sin(slp);
    const double cosd = sqrt(1. - sind * sind);
    const double sc = sind / cosd;
    *sdec = atan(sc);
    *srasn = 3.141592654 - atan2(cos(obliq) / sob * sc, -cos(slp) / cosd);
}

/* lib/internal/internal.c:12-36,64-87: coefficients at (year, day of year): linear
 * interpolation between the table's rows, or main field + secular variation (index
 * <= 45) within one step of the last year where the model extrapolates.  The
 * descriptor's unsigned arithmetic (int year against uint32_t years) is kept, so out
 * of range and negative years behave as in the reference. */
static int coefs_at(const InternalModelCoefs *c, int year, int iday, double *G, double *H) {
    const double step = (double)c->years_step;
    if (c->extrapolate && c->max_year - (uint32_t)year <= c->years_step) {
        const double *g0 = c->coefs[c->coefs_size - 4], *h0 = c->coefs[c->coefs_size - 3];
        const double *gs = c->coefs[c->coefs_size - 2], *hs = c->coefs[c->coefs_size - 1];
        const double dt = year + (iday - 1) / 365.25 - ((int)c->max_year - step);
        for (int k = 0; k < INTERNAL_COEFS_SIZE; k++) {
            G[k] = g0[k];
            H[k] = h0[k];
            if (k <= 45) {
                G[k] += gs[k] * dt;
                H[k] += hs[k] * dt;
            }
        }
        return 0;
    }
    if ((uint32_t)year >= c->max_year) return -1;
    const int back = (uint32_t)year <= c->min_year
                         ? (int)(c->max_year / c->years_step - (c->min_year + 1) / c->years_step)
                         : (int)(c->max_year / c->years_step - (uint32_t)year / c->years_step);
    const int row = (int)(c->coefs_size - 2 * back - 1);
    const int edge = (int)(c->min_year + ((row - 1) / 2) * c->years_step);
    const double *gx = c->coefs[row - 1], *hx = c->coefs[row];
    const double *gy = c->coefs[row + 1], *hy = c->coefs[row + 2];
    const double f2 = (year + (iday - 1) / 365.25 - edge) / step;
    const double f1 = 1. - f2;
    for (int k = 0; k < INTERNAL_COEFS_SIZE; k++) {
        G[k] = gx[k] * f1 + gy[k] * f2;
        H[k] = hx[k] * f1 + hy[k] * f2;
    }
    return 0;
}

/* lib/internal/internal.c:89-180. */
int internal_recalc(const InternalModelCoefs *coefs, const RecalcTime *time, double *G, double *H, double *REC,
                    double *A1, double *A2, double *A3) {
    if ((uint32_t)time->iy > coefs->max_year || (uint32_t)time->iy < coefs->min_year) return -1;
    for (int n = 1; n < 15; n++) {
        double den = 2 * n - 1;
        den = den * (den - 2);
        for (int m = 1; m <= n; m++) REC[n * (n - 1) / 2 + m - 1] = ((n - m) * (n + m - 2)) / den;
    }
    if (coefs_at(coefs, time->iy, time->id, G, H) != 0) return -1;

    /* Schmidt semi-normalisation -> the recursion's scaling */
    double s = 1.;
    for (int n = 2; n <= 14; n++) {
        const int base = n * (n - 1) / 2;
        s = s * (2 * n - 3) / (n - 1);
        G[base] *= s;
        H[base] *= s;
        double f = s;
        for (int m = 2; m <= n; m++) {
            const double two = (m == 2) ? 2. : 1.;
            f = f * sqrt(two * (n - m + 1) / (n + m - 2));
            G[base + m - 1] *= f;
            H[base + m - 1] *= f;
        }
    }

    /* dipole axis and the GEO -> GSM rows */
    const double g10 = -G[1], g11 = G[2], h11 = H[2];
    const double sq = g11 * g11 + h11 * h11;
    const double sqq = sqrt(sq);
    const double sqr = sqrt(g10 * g10 + sq);
    double gst = 0., slong = 0., srasn = 0., sdec = 0.;
    if (time->iy > 1900 && time->iy < 2100) {
        sun_position(time->iy, time->id, time->ih, time->min, time->is, &gst, &slong, &srasn, &sdec);
    }
    (void)slong;
    const double ct0 = g10 / sqr, cl0 = -g11 / sqq, sl0 = -h11 / sqq, st0 = sqq / sqr;
    const double stcl = st0 * cl0, stsl = st0 * sl0;
    const double cgst = cos(gst), sgst = sin(gst);
    const double d1 = stcl * cgst - stsl * sgst;
    const double d2 = stcl * sgst + stsl * cgst;
    const double d3 = ct0;
    const double csdec = cos(sdec);
    const double s1 = cos(srasn) * csdec, s2 = sin(srasn) * csdec, s3 = sin(sdec);
    double y1 = d2 * s3 - d3 * s2;
    double y2 = d3 * s1 - d1 * s3;
    double y3 = d1 * s2 - d2 * s1;
    const double yn = sqrt(y1 * y1 + y2 * y2 + y3 * y3);
    y1 = y1 / yn;
    y2 = y2 / yn;
    y3 = y3 / yn;
    const double z1 = s2 * y3 - s3 * y2;
    const double z2 = s3 * y1 - s1 * y3;
    const double z3 = s1 * y2 - s2 * y1;
    A1[0] = s1 * cgst + s2 * sgst;
    A1[1] = -s1 * sgst + s2 * cgst;
    A1[2] = s3;
    A2[0] = y1 * cgst + y2 * sgst;
    A2[1] = -y1 * sgst + y2 * cgst;
    A2[2] = y3;
    A3[0] = z1 * cgst + z2 * sgst;
    A3[1] = -z1 * sgst + z2 * cgst;
    A3[2] = z3;
    return 0;
}
