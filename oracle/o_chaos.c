/*
 * o_chaos.c -- oracle: generic spherical-harmonic synthesis as lib/chaos/chaos.c
 * does it (non-OLD_CODE branch, lib/chaos/chaos.c:118-227), restated in plain C.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Schmidt semi-normalised P_n^m and
 * dP_n^m/dtheta by the standard three-term recursion, (a/r)^(n+1) powers with
 * a = 6371.2 km, fresh sin/cos(m phi) per term.  G/H are row-major
 * [nm_max][nm_max] tables indexed [n][m] in nT.  Input GEO Cartesian [Re], output
 * GEO Cartesian field [nT].  At an exact pole B_phi divides by sin(theta) = 0 as the
 * reference does.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

#define CH_RE_KM 6371.2                 /* lib/chaos/chaos.h:6 */
#define CH_2PI 6.283185307179586        /* lib/chaos/chaos.h:7 */
#define CH_PI 3.141592653589793         /* lib/chaos/chaos.h:8 */
#define CH_NMAX 200

void o_chaos_geo(double x, double y, double z, int nm_max, const double *g,
                 const double *h, double b[3]) {
    static _Thread_local double pnm[CH_NMAX][CH_NMAX], dpnm[CH_NMAX][CH_NMAX];
    double ar[CH_NMAX], dar[CH_NMAX];
    if (nm_max < 2 || nm_max > CH_NMAX) {
        b[0] = b[1] = b[2] = NAN;
        return;
    }
    const double rho2 = x * x + y * y;
    const double rr = sqrt(rho2 + z * z);
    const double c = z / rr;
    const double r = rr * CH_RE_KM;
    double theta = 0., phi = 0., s = 0., cf = 1., sf = 0.;
    if (rho2 != 0.) {
        const double rho = sqrt(rho2);
        s = rho / rr;
        cf = x / rho;
        sf = y / rho;
        phi = atan2(y, x);
        theta = atan2(rho, z);
        if (phi < 0.) phi += CH_2PI;
    } else if (z < 0.) {
        theta = CH_PI;
    }

    const double q = CH_RE_KM / r;
    ar[0] = q;
    dar[0] = -q / r;
    for (int n = 1; n < nm_max; n++) {
        ar[n] = ar[n - 1] * q;
        dar[n] = -(n + 1) * ar[n] / r;
    }

    for (int n = 0; n < nm_max; n++) {
        memset(pnm[n], 0, (size_t)nm_max * sizeof(double));
        memset(dpnm[n], 0, (size_t)nm_max * sizeof(double));
    }
    const double st = sin(theta), ct = cos(theta);
    pnm[0][0] = 1.;
    pnm[1][0] = ct;
    pnm[1][1] = st;
    dpnm[0][0] = 0.;
    dpnm[1][0] = -st;
    dpnm[1][1] = ct;
    for (int n = 2; n < nm_max; n++) {
        const double n2 = n * 2;
        const double n21 = n2 - 1;
        /* sectoral term from the previous sectoral one */
        const double kd = sqrt(n21 / n2);
        const double kp = kd * pnm[n - 1][n - 1];
        pnm[n][n] = st * kp;
        dpnm[n][n] = ct * kp + st * kd * dpnm[n - 1][n - 1];
        const double nn = (double)n * n;
        for (int m = 0; m < n; m++) {
            const double mm = m * m;
            const double den = nn - mm;
            const double k1 = n21 / sqrt(den);
            const double k2 = sqrt(((n - 1) * (n - 1) - mm) / den);
            const double p1 = pnm[n - 1][m];
            pnm[n][m] = k1 * ct * p1 - k2 * pnm[n - 2][m];
            dpnm[n][m] = k1 * (-st * p1 + ct * dpnm[n - 1][m]) - k2 * dpnm[n - 2][m];
        }
    }

    double br = 0., bt = 0., bp = 0.;
    for (int n = 0; n < nm_max; n++) {
        for (int m = 0; m <= n; m++) {
            const double mphi = m * phi;
            const double smp = sin(mphi), cmp = cos(mphi);
            const double gg = g[n * nm_max + m], hh = h[n * nm_max + m];
            const double t = gg * cmp + hh * smp;
            br += t * pnm[n][m] * dar[n];
            bt += t * dpnm[n][m] * ar[n];
            bp += (-gg * m * smp + hh * m * cmp) * pnm[n][m] * ar[n];
        }
    }
    bt = -bt * CH_RE_KM / r;
    bp = -bp * CH_RE_KM / (r * st);
    br = -br * CH_RE_KM;
    const double he = br * s + bt * c;
    b[0] = he * cf - bp * sf;
    b[1] = he * sf + bp * cf;
    b[2] = br * c - bt * s;
}
