/*
 * o_ts05.c -- oracle: Tsyganenko-Sitnov 2005 external field at zero dipole tilt, as
 * lib/ts05/{ts05,deformed,birk_tot,full_rc}.c compute it, restated in plain C.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The model is a published empirical fit, so this restatement keeps every coefficient
 * and every expression's operation order (compiled with -ffp-contract=off, like the
 * reference): it is checked bit for bit against the reference compiled from its own
 * sources (tests/golden/).  The organisation is this file's own: per-epoch driver
 * terms in one struct, the Cartesian-harmonic sums of the four shield/warping
 * expansions as one helper over (P, R) tables, and only the outputs that reach the
 * total field (the reference also computes the Birkeland mode-2 fields R12/R22, which
 * ts05_gsm discards, lib/ts05/ts05.c:112-114).
 *
 * The helpers have external linkage on purpose: with file-local helpers GCC inlines
 * them differently from the reference's translation units and then merges some
 * sin/cos pairs of one argument into sincos calls, which can round differently in
 * the last bit.  With external linkage the restatement is bit-identical to the
 * reference (checked on 1.3e5 points).
 */
#include <math.h>
#include "oracle.h"

/* lib/ts05/ts05.c:8-17: the 69 model parameters */
static const double TA[69] = {
    1.00000, 5.44118, 0.891995, 9.09684, 0.00000, -7.18972, 12.2700, -4.89408, 0.00000, 0.870536,
    1.36081, 0.00000, 0.688650, 0.602330, 0.00000, 0.316346, 1.22728, -0.363620E-01, -0.405821,
    0.452536, 0.755831, 0.215662, 0.152759, 5.96235, 23.2036, 11.2994, 69.9596, 0.989596,
    -0.132131E-01, 0.985681, 0.344212E-01, 1.02389, 0.207867, 1.51220, 0.682715E-01, 1.84714,
    1.76977, 1.37690, 0.696350, 0.343280, 3.28846, 111.293, 5.82287, 4.39664, 0.383403,
    0.648176, 0.318752E-01, 0.581168, 1.15070, 0.843004, 0.394732, 0.846509, 0.916555, 0.550920,
    0.180725, 0.898772, 0.387365, 2.26596, 1.29123, 0.436819, 1.28211, 1.33199, .405553, 1.6229,
    .699074, 1.26131, 2.42297, .537116, .619441};

/* ------------------------------------------------------------ per-epoch terms */

typedef struct {
    double pdyn, dst, byimf, bzimf, w[6];
    double xappa, xappa3, am, x0, oimfy, oimfz;
    double dxshift1, dxshift2, d, xkappa1, xkappa2, sc_sy, sc_pr;
    double tamp1, tamp2, a_src, a_prc, a_r11, a_r21;
} ts05_drv;

/* lib/ts05/ts05.c:19-34,41-43,56-110: everything that depends only on PARMOD */
void ts05_drivers(const double *pm, ts05_drv *v) {
    v->pdyn = pm[0];
    v->dst = pm[1] * 0.8 - 13. * sqrt(v->pdyn);
    v->byimf = pm[2];
    v->bzimf = pm[3];
    for (int k = 0; k < 6; k++) v->w[k] = pm[4 + k];
    v->xappa = pow((v->pdyn / 2.), TA[22]);
    v->xappa3 = v->xappa * v->xappa * v->xappa;
    v->am = 34.586 / v->xappa;
    v->x0 = 3.4397 / v->xappa;
    v->oimfy = v->byimf * TA[19];
    v->oimfz = v->bzimf * TA[19];

    double dstt = -20.;
    if (v->dst < dstt) dstt = v->dst;
    double znam = pow(fabs(dstt), 0.37);
    v->dxshift1 = TA[23] - TA[24] / znam;
    v->dxshift2 = TA[25] - TA[26] / znam;
    v->d = TA[35] * exp(-v->w[0] / TA[36]) + TA[68];

    znam = fabs(v->dst);
    if (v->dst >= -20.) znam = 20.;
    v->xkappa1 = TA[31] * pow(znam / 20., TA[32]);
    v->xkappa2 = TA[33] * pow(znam / 20., TA[34]);
    v->sc_sy = TA[27] * pow((20. / znam), TA[28]) * v->xappa;
    v->sc_pr = TA[29] * pow((20. / znam), TA[30]) * v->xappa;

    const double dlp1 = pow((v->pdyn / 2.), TA[20]);
    const double dlp2 = pow((v->pdyn / 2.), TA[21]);
    const double *w = v->w;
    v->tamp1 = TA[1] + TA[2] * dlp1 + TA[3] * (TA[38] * w[0]) / sqrt(w[0] * w[0] + TA[38] * TA[38]);
    v->tamp2 = TA[5] + TA[6] * dlp2 + TA[7] * (TA[39] * w[1]) / sqrt(w[1] * w[1] + TA[39] * TA[39]);
    v->a_src = TA[9] + TA[10] * (TA[40] * w[2]) / sqrt(w[2] * w[2] + TA[40] * TA[40]) + TA[11] * v->dst;
    v->a_prc = TA[12] + TA[13] * (TA[41] * w[3]) / sqrt(w[3] * w[3] + TA[41] * TA[41]) + TA[14] * v->dst;
    v->a_r11 = TA[15] + TA[16] * (TA[42] * w[4]) / sqrt(w[4] * w[4] + TA[42] * TA[42]);
    v->a_r21 = TA[17] + TA[18] * (TA[43] * w[5]) / sqrt(w[5] * w[5] + TA[43] * TA[43]);
}

/* ------------------------------------------------------------ dipole, outer shield */

/* lib/ts05/ts05.c:233-246 (its own dipole constant and signs) */
void ts05_dipole(double x, double y, double z, double b[3]) {
    const double p = x * x, u = z * z, t = y * y;
    const double s = sqrt(p + t + u);
    const double q = 30115. / (s * s * s * s * s);
    b[0] = q * 3. * z * x;
    b[1] = q * -3. * z * y;
    b[2] = q * (p + t - 2. * u);
}

/* lib/ts05/ts05.c:136-231: 3 x 3 harmonics; the third z scale has its own term form */
void ts05_shlcar3x3(double x, double y, double z, double b[3]) {
    static const double P[3] = {9.620648151, 6.082014949, 27.75216226};
    static const double R[3] = {12.44199571, 5.122226936, 6.982039615};
    static const double A[9] = {-901.2327248 + (double)895.8011176, 817.6208321 - (double)845.5880889,
                                -83.73539535 + (double)86.58542841, 336.8781402 - (double)329.3619944,
                                -311.2947120 + (double)308.6011161, 31.94469304 - (double)31.30824526,
                                125.8739681 - (double)372.3384278, -235.4720434 + (double)286.7594095,
                                21.86305585 - (double)27.42344605};
    double rinv2[3], czr[3], szr[3], zr[3];
    for (int k = 0; k < 3; k++) {
        zr[k] = z / R[k];
        czr[k] = cos(zr[k]);
        szr[k] = sin(zr[k]);
    }
    rinv2[0] = 1. / (R[0] * R[0]);
    rinv2[1] = 1. / (R[1] * R[1]);
    const double r32 = R[2] * R[2];
    rinv2[2] = 1. / r32;
    double bx = 0., by = 0., bz = 0.;
    for (int i = 0; i < 3; i++) {
        const double pinv2 = 1. / (P[i] * P[i]);
        const double cyp = cos(y / P[i]), syp = sin(y / P[i]);
        for (int k = 0; k < 3; k++) {
            const double sq = sqrt(pinv2 + rinv2[k]);
            const double e = exp(sq * x);
            double hx, hy, hz;
            if (k < 2) {
                hx = -sq * e * cyp * szr[k];
                hy = e / P[i] * syp * szr[k];
                hz = -e * cyp / R[k] * czr[k];
            } else {
                hx = -e * cyp * (sq * z * czr[2] + szr[2] / R[2] * (x + 1. / sq));
                hy = e / P[i] * syp * (z * czr[2] + x / R[2] * szr[2] / sq);
                hz = -e * cyp * (czr[2] * (1. + x / r32 / sq) - zr[2] * szr[2]);
            }
            const double a = A[3 * i + k];
            bx = (i == 0 && k == 0) ? a * hx : bx + a * hx;
            by = (i == 0 && k == 0) ? a * hy : by + a * hy;
            bz = (i == 0 && k == 0) ? a * hz : bz + a * hz;
        }
    }
    b[0] = bx;
    b[1] = by;
    b[2] = bz;
}

/* ------------------------------------------------------------ warped tail */

/* lib/ts05/deformed.c:119-180: five conical current sheets of the tail */
void ts05_taildisk(double d0, double deltadx, double x, double y, double z, double b[3]) {
    static const double F[5] = {-71.09346626, -1014.308601, -1272.939359, -3224.935936, -44546.86232};
    static const double B[5] = {10.90101242, 12.68393898, 13.51791954, 14.86775017, 15.12306404};
    static const double C[5] = {.7954069972, .6716601849, 1.174866319, 2.565249920, 10.01986790};
    const double rho = sqrt(x * x + y * y);
    const double drhodx = x / rho, drhody = y / rho;
    const double dex = exp(x / 7.);
    const double y20 = y / 20.;
    const double d = d0 + 4.7 * y20 * y20 + deltadx * dex;
    const double dddy = 4.7 * y * 0.005;
    const double dddx = deltadx / 7. * dex;
    const double dzeta = sqrt(z * z + d * d);
    const double ddzdx = d * dddx / dzeta, ddzdy = d * dddy / dzeta, ddzdz = z / dzeta;
    double dbx = 0., dby = 0., dbz = 0.;
    for (int i = 0; i < 5; i++) {
        const double bi = B[i], dzc = dzeta + C[i], dzc2 = dzc * dzc;
        const double rpb = rho + bi, rmb = rho - bi;
        const double s1 = sqrt(rpb * rpb + dzc2), s2 = sqrt(rmb * rmb + dzc2);
        const double ds1drho = (rho + bi) / s1, ds2drho = (rho - bi) / s2;
        const double ds1ddz = dzc / s1, ds2ddz = dzc / s2;
        const double ds1dx = ds1drho * drhodx + ds1ddz * ddzdx;
        const double ds1dy = ds1drho * drhody + ds1ddz * ddzdy;
        const double ds1dz = ds1ddz * ddzdz;
        const double ds2dx = ds2drho * drhodx + ds2ddz * ddzdx;
        const double ds2dy = ds2drho * drhody + ds2ddz * ddzdy;
        const double ds2dz = ds2ddz * ddzdz;
        const double s1ts2 = s1 * s2, s1ps2 = s1 + s2, s1ps2sq = s1ps2 * s1ps2;
        const double bi2 = 2. * bi;
        const double fac1 = sqrt(s1ps2sq - bi2 * bi2);
        const double as = fac1 / (s1ts2 * s1ps2sq);
        const double dasds1 = (1. / (fac1 * s2) - as / s1ps2 * (s2 * s2 + s1 * (3. * s1 + 4. * s2))) / (s1 * s1ps2);
        const double dasds2 = (1. / (fac1 * s1) - as / s1ps2 * (s1 * s1 + s2 * (3. * s2 + 4. * s1))) / (s2 * s1ps2);
        const double dasdx = dasds1 * ds1dx + dasds2 * ds2dx;
        const double dasdy = dasds1 * ds1dy + dasds2 * ds2dy;
        const double dasdz = dasds1 * ds1dz + dasds2 * ds2dz;
        dbx = dbx - F[i] * x * dasdz;
        dby = dby - F[i] * y * dasdz;
        dbz = dbz + F[i] * (2. * as + x * dasdx + y * dasdy);
    }
    b[0] = dbx;
    b[1] = dby;
    b[2] = dbz;
}

/* lib/ts05/deformed.c:182-208: 5 x 5 harmonics whose coefficients are linear in the
 * tail shift; table layout as the reference's (50 coefficient pairs, then 1/P, 1/R) */
void ts05_shlcar5x5(const double *a, double x, double y, double z, double dshift, double b[3]) {
    double hx = 0., hy = 0., hz = 0.;
    int l = 0;
    for (int i = 0; i < 5; i++) {
        const double rp = 1. / a[50 + i];
        const double rp2 = rp * rp;
        const double cy = cos(y * rp), sy = sin(y * rp);
        for (int k = 0; k < 5; k++, l += 2) {
            const double rr = 1. / a[55 + k];
            const double sq = sqrt(rp2 + rr * rr);
            const double e = exp(x * sq);
            const double zr = z * rr;
            const double sz = sin(zr), cz = cos(zr);
            const double coef = a[l] + a[l + 1] * dshift;
            hx = hx + coef * (-sq * e * cy * sz);
            hy = hy + coef * (rp * e * sy * sz);
            hz = hz + coef * (-rr * e * cy * cz);
        }
    }
    b[0] = hx;
    b[1] = hy;
    b[2] = hz;
}

/* lib/ts05/deformed.c:47-117 */
static const double WARP_A1[60] = {
    -25.45869857, 57.35899080, 317.5501869, -2.626756717, -93.38053698, -199.6467926, -858.8129729,
    34.09192395, 845.4214929, -29.07463068, 47.10678547, -128.9797943, -781.7512093, 6.165038619,
    167.8905046, 492.0680410, 1654.724031, -46.77337920, -1635.922669, 40.86186772, -.1349775602,
    -.9661991179E-01, -.1662302354, .002810467517, .2487355077, .1025565237, -14.41750229,
    -.8185333989, 11.07693629, .7569503173, -9.655264745, 112.2446542, 777.5948964, -5.745008536,
    -83.03921993, -490.2278695, -1155.004209, 39.08023320, 1172.780574, -39.44349797, -14.07211198,
    -40.41201127, -313.2277343, 2.203920979, 8.232835341, 197.7065115, 391.2733948, -18.57424451,
    -437.2779053, 23.04976898, 11.75673963, 13.60497313, 4.691927060, 18.20923547, 27.59044809,
    6.677425469, 1.398283308, 2.839005878, 31.24817706, 24.53577264};
static const double WARP_A2[60] = {
    -287187.1962, 4970.499233, 410490.1952, -1347.839052, -386370.3240, 3317.983750, -143462.3895,
    5706.513767, 171176.2904, 250.8882750, -506570.8891, 5733.592632, 397975.5842, 9771.762168,
    -941834.2436, 7990.975260, 54313.10318, 447.5388060, 528046.3449, 12751.04453, -21920.98301,
    -21.05075617, 31971.07875, 3012.641612, -301822.9103, -3601.107387, 1797.577552, -6.315855803,
    142578.8406, 13161.93640, 804184.8410, -14168.99698, -851926.6360, -1890.885671, 972475.6869,
    -8571.862853, 26432.49197, -2554.752298, -482308.3431, -4391.473324, 105155.9160, -1134.622050,
    -74353.53091, -5382.670711, 695055.0788, -916.3365144, -12111.06667, 67.20923358, -367200.9285,
    -21414.14421, 14.75567902, 20.75638190, 59.78601609, 16.86431444, 32.58482365, 23.69472951,
    17.24977936, 13.64902647, 68.40989058, 11.67828167};

/* lib/ts05/deformed.c:4-45,47-117: the two tail modes at the unwarped point (at zero
 * tilt the warping is the identity up to the polar round trip, kept for parity) */
void ts05_warped(double x, double y, double z, const ts05_drv *v, double b1[3], double b2[3]) {
    const double rho = sqrt(y * y + z * z);
    double phi = 0., cphi = 1., sphi = 0.;
    if (y != 0. || z != 0.) {
        phi = atan2(z, y);
        cphi = y / rho;
        sphi = z / rho;
    }
    const double cf = cos(phi), sf = sin(phi);
    const double yu = rho * cf, zu = rho * sf;
    double f[3], h[3], u1[3], u2[3];
    ts05_taildisk(v->d * 1.1, 1., (x - 6. - v->dxshift1) * 1.1 + 1.2, yu * 1.1, zu * 1.1, f);
    ts05_shlcar5x5(WARP_A1, x, yu, zu, v->dxshift1, h);
    for (int c = 0; c < 3; c++) u1[c] = f[c] + h[c];
    ts05_taildisk(v->d * 0.25, 0., (x - 4. - v->dxshift2) * 0.25 - 9., yu * 0.25, zu * 0.25, f);
    ts05_shlcar5x5(WARP_A2, x, yu, zu, v->dxshift2, h);
    for (int c = 0; c < 3; c++) u2[c] = f[c] + h[c];
    const double *u[2] = {u1, u2};
    double *o[2] = {b1, b2};
    for (int m = 0; m < 2; m++) {
        const double brho = u[m][1] * cf + u[m][2] * sf;
        const double bphi = -u[m][1] * sf + u[m][2] * cf;
        o[m][0] = u[m][0];
        o[m][1] = brho * cphi - bphi * sphi;
        o[m][2] = brho * sphi + bphi * cphi;
    }
}

/* ------------------------------------------------------------ Birkeland currents */

/* lib/ts05/birk_tot.c:284-299 (r_s) and :301-307 (theta_s): the deformation of the
 * conical current system, a = one of the 31-coefficient sets */
double ts05_rs(const double *a, double r, double th) {
    const double r2 = r * r;
    const double r2aa = r2 + a[15] * a[15];
    return r + a[1] / r + a[2] * r / sqrt(r2 + a[10] * a[10]) + a[3] * r / (r2 + a[11] * a[11]) +
           (a[4] + a[5] / r + a[6] * r / sqrt(r2 + a[12] * a[12]) + a[7] * r / (r2 + a[13] * a[13])) * cos(th) +
           (a[8] * r / sqrt(r2 + a[14] * a[14]) + a[9] * r / (r2aa * r2aa)) * cos(2. * th);
}
double ts05_thetas(const double *a, double r, double th) {
    const double r2 = r * r;
    return th + (a[16] + a[17] / r + a[18] / r2 + a[19] * r / sqrt(r2 + a[26] * a[26])) * sin(th) +
           (a[20] + a[21] * r / sqrt(r2 + a[27] * a[27]) + a[22] * r / (r2 + a[28] * a[28])) * sin(2. * th) +
           (a[23] + a[24] / r + a[25] * r / (r2 + a[29] * a[29])) * sin(3. * th);
}

/* lib/ts05/birk_tot.c:309-377 for mode 1 (the only mode ts05_gsm uses) */
void ts05_fialcos1(double r, double th, double ph, double th0, double dt, double *bth, double *bph) {
    const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    const double ro = r * st;
    const double tg = st / (1. + ct), ctg = st / (1. - ct);
    const double tnp = th0 + dt, tnm = th0 - dt;
    double tgp = 0., tgm = 0., tgm2 = 0., tgp2 = 0.;
    if (th >= tnm) {
        tgp = tan(tnp * 0.5);
        tgm = tan(tnm * 0.5);
        tgm2 = tgm * tgm;
        tgp2 = tgp * tgp;
    }
    const double tm = tg;                  /* tg^1 */
    const double cc = cp - 0. * sp;        /* cos(1 phi) by the reference's recurrence */
    const double ss = 0. * cp + sp;
    double t, dtt;
    if (th < tnm) {
        t = tm;
        dtt = 0.5 * tm * (tg + ctg);
    } else if (th < tnp) {
        const double fc = 1. / (tgp - tgm), fc1 = 1. / 3.;
        const double tgm2m1 = tgm2 * tgm;
        const double tg21 = 1. + tg * tg;
        t = fc * (tm * (tgp - tg) + fc1 * (tm * tg - tgm2m1 / tm));
        dtt = 0.5 * fc * tg21 * (tm / tg * (tgp - tg) - fc1 * (tm - tgm2m1 / (tm * tg)));
    } else {
        const double fc = 1. / (tgp - tgm), fc1 = 1. / 3.;
        t = fc * fc1 * (tgp2 * tgp - tgm2 * tgm) / tm;
        dtt = -t * 0.5 * (tg + ctg);
    }
    *bth = t * cc / ro * 800.;
    *bph = -dtt * ss / r * 800.;
}

/* lib/ts05/birk_tot.c:252-282: one cone, with the r_s/theta_s derivatives by central
 * differences of +-1e-6 */
void ts05_one_cone(const double *a, double x, double y, double z, double dth, double b[3]) {
    const double rho2 = x * x + y * y;
    const double rho = sqrt(rho2);
    const double r = sqrt(rho2 + z * z);
    const double th = atan2(rho, z);
    const double ph = atan2(y, x);
    const double rs = ts05_rs(a, r, th);
    const double ths = ts05_thetas(a, r, th);
    double btast, bfast;
    ts05_fialcos1(rs, ths, ph, a[30], dth, &btast, &bfast);
    const double rsr = rs / r;
    const double stsst = sin(ths) / sin(th);
    const double drsdr = (ts05_rs(a, r + 1E-6, th) - ts05_rs(a, r - 1E-6, th)) / 2E-6;
    const double drsdt = (ts05_rs(a, r, th + 1E-6) - ts05_rs(a, r, th - 1E-6)) / 2E-6;
    const double dtsdr = (ts05_thetas(a, r + 1E-6, th) - ts05_thetas(a, r - 1E-6, th)) / 2E-6;
    const double dtsdt = (ts05_thetas(a, r, th + 1E-6) - ts05_thetas(a, r, th - 1E-6)) / 2E-6;
    const double br = -rsr / r * stsst * btast * drsdt;
    const double bth = rsr * stsst * btast * drsdr;
    const double bph = rsr * bfast * (drsdr * dtsdt - drsdt * dtsdr);
    const double s = rho / r, c = z / r, sf = y / rho, cf = x / rho;
    const double be = br * s + bth * c;
    b[0] = a[0] * (be * cf - bph * sf);
    b[1] = a[0] * (be * sf + bph * cf);
    b[2] = a[0] * (br * c - bth * s);
}

/* lib/ts05/birk_tot.c:117-227 for mode 1: Region 1 (numb 1) or Region 2 (numb 2) */
void ts05_birk_1n2(int numb, double x, double y, double z, double xkappa, double b[3]) {
    static const double A11[31] = {
        .1618068350, -.1797957553, 2.999642482, -.9322708978, -.6811059760, .2099057262, -8.358815746,
        -14.86033550, .3838362986, -16.30945494, 4.537022847, 2.685836007, 27.97833029, 6.330871059,
        1.876532361, 18.95619213, .9651528100, .4217195118, -.08957770020, -1.823555887, .7457045438,
        -.5785916524, -1.010200918, .01112389357, .09572927448, -.3599292276, 8.713700514, .9763932955,
        3.834602998, 2.492118385, .7113544659};
    static const double A21[31] = {
        .1278764024, -.2320034273, 1.805623266, -32.37241440, -.9931490648, .3175085630, -2.492465814,
        -16.21600096, .2695393416, -6.752691265, 3.971794901, 14.54477563, 41.10158386, 7.912889730,
        1.258297372, 9.583547721, 1.014141963, .5104134759, -.1790430468, -1.756358428, .7561986717,
        -.6775248254, -.04014016420, .01446794851, .1200521731, -.2203584559, 4.508963850, .8221623576,
        1.779933730, 1.102649543, .8867880020};
    const double *a = numb == 1 ? A11 : A21;
    const double dphi = numb == 1 ? 0.055 : 0.030;
    const double dth = numb == 1 ? 0.06 : 0.09;
    const double xs = x * xkappa, ys = y * xkappa, zs = z * xkappa;
    const double rho = sqrt(xs * xs + zs * zs);
    const double rho2 = rho * rho;
    double phi = 0.;
    if (xs != 0. || zs != 0.) phi = atan2(-zs, xs);
    const double sphic = sin(phi), cphic = cos(phi);
    const double den = 49. + rho2;
    const double brack = dphi + 0.5 * 49. / 50. * (rho2 - 1.) / den;
    const double phis = phi - brack * sin(phi);
    const double dphisphi = 1. - brack * cos(phi);
    const double dphisrho = -2. * 0.5 * 49. * rho / (den * den) * sin(phi);
    const double sphics = sin(phis), cphics = cos(phis);
    const double xc = rho * cphics, zc = -rho * sphics;
    double bn[3], bs[3];
    ts05_one_cone(a, xc, ys, zc, dth, bn);
    ts05_one_cone(a, xc, -ys, -zc, dth, bs);
    const double bxs = bn[0] - bs[0], byas = bn[1] + bs[1], bzs = bn[2] + bs[2];
    const double brhoas = bxs * cphics - bzs * sphics;
    const double bphias = -bxs * sphics - bzs * cphics;
    const double brho_s = brhoas * dphisphi * xkappa;
    const double bphi_s = (bphias - rho * (brhoas * dphisrho)) * xkappa;
    b[0] = brho_s * cphic - bphi_s * sphic;
    b[1] = byas * dphisphi * xkappa;
    b[2] = -brho_s * sphic - bphi_s * cphic;
}

/* Shield harmonics shared by birk_shl (lib/ts05/birk_tot.c:379-403) and rc_shield
 * (lib/ts05/full_rc.c:697-721): 3 x 3 terms, coefficient (a0 + s a1 + a2 + s a3) of a
 * zero-tilt table, P = a[72..74], R = a[75..77]; fac multiplies each term (rc_shield's
 * (x_sc + 1)^3, written into the term as the reference does, or none). */
void ts05_shield(const double *a, double xsc, double x, double y, double z, int use_fac, double fac,
                        double b[3]) {
    double gx = 0., gy = 0., gz = 0.;
    int l = 0;
    for (int j = 0; j < 3; j++) {
        const double p = a[72 + j];
        const double sy = sin(y / p), cy = cos(y / p);
        const double pinv2 = 1. / (p * p);
        for (int k = 0; k < 3; k++, l += 4) {
            const double r = a[75 + k];
            const double sz = sin(z / r), cz = cos(z / r);
            const double sq = sqrt(pinv2 + 1. / (r * r));
            const double e = exp(x * sq);
            const double coef = a[l] + xsc * a[l + 1] + a[l + 2] + xsc * a[l + 3];
            if (use_fac) {
                gx += (-sq * e * cy * sz * fac) * coef;
                gy += (e * sy * sz / p * fac) * coef;
                gz += (-e * cy * cz / r * fac) * coef;
            } else {
                gx += (-sq * e * cy * sz) * coef;
                gy += (e * sy * sz / p) * coef;
                gz += (-e * cy * cz / r) * coef;
            }
        }
    }
    b[0] = gx;
    b[1] = gy;
    b[2] = gz;
}

static const double SH11[78] = {
    46488.84663, -15541.95244, -23210.09824, -32625.03856, -109894.4551, -71415.32808, 58168.94612,
    55564.87578, -22890.60626, -6056.763968, 5091.368100, 239.7001538, -13899.49253, 4648.016991,
    6971.310672, 9699.351891, 32633.34599, 21028.48811, -17395.96190, -16461.11037, 7447.621471,
    2528.844345, -1934.094784, -588.3108359, -32588.88216, 10894.11453, 16238.25044, 22925.60557,
    77251.11274, 50375.97787, -40763.78048, -39088.60660, 15546.53559, 3559.617561, -3187.730438,
    309.1487975, 88.22153914, -243.0721938, -63.63543051, 191.1109142, 69.94451996, -187.9539415,
    -49.89923833, 104.0902848, -120.2459738, 253.5572433, 89.25456949, -205.6516252, -44.93654156,
    124.7026309, 32.53005523, -98.85321751, -36.51904756, 98.88241690, 24.88493459, -55.04058524,
    61.14493565, -128.4224895, -45.35023460, 105.0548704, -43.66748755, 119.3284161, 31.38442798,
    -92.87946767, -33.52716686, 89.98992001, 25.87341323, -48.86305045, 59.69362881, -126.5353789,
    -44.39474251, 101.5196856, 59.41537992, 41.18892281, 80.86101200, 3.066809418, 7.893523804,
    30.56212082};
static const double SH21[78] = {
    162294.6224, 503885.1125, -27057.67122, -531450.1339, 84747.05678, -237142.1712, 84133.61490,
    259530.0402, 69196.05160, -189093.5264, -19278.55134, 195724.5034, -263082.6367, -818899.6923,
    43061.10073, 863506.6932, -139707.9428, 389984.8850, -135167.5555, -426286.9206, -109504.0387,
    295258.3531, 30415.07087, -305502.9405, 100785.3400, 315010.9567, -15999.50673, -332052.2548,
    54964.34639, -152808.3750, 51024.67566, 166720.0603, 40389.67945, -106257.7272, -11126.14442,
    109876.2047, 2.978695024, 558.6019011, 2.685592939, -338.0004730, -81.99724090, -444.1102659,
    89.44617716, 212.0849592, -32.58562625, -982.7336105, -35.10860935, 567.8931751, -1.917212423,
    -260.2023543, -1.023821735, 157.5533477, 23.00200055, 232.0603673, -36.79100036, -111.9110936,
    18.05429984, 447.0481000, 15.10187415, -258.7297813, -1.032340149, -298.6402478, -1.676201415,
    180.5856487, 64.52313024, 209.0160857, -53.85574010, -98.52164290, 14.35891214, 536.7666279,
    20.09318806, -309.7349530, 58.54144539, 67.45226850, 97.92374406, 4.752449760, 10.46824379,
    32.91856110};

/* ------------------------------------------------------------ ring currents */

/* lib/ts05/full_rc.c:213-327 (ap) and :370-423 (apprc) share the vector-potential of
 * two loops at the dipole-mapped point: */
double ts05_loops_aphi(double rhos, double zs2, double rrc, double dd) {
    const double rr = rrc + rhos;
    const double p = rr * rr + zs2 + dd * dd;
    const double xk2 = 4. * rrc * rhos / p;
    const double xk = sqrt(xk2);
    const double xkrho12 = xk * sqrt(rhos);
    const double xk2s = 1. - xk2;
    const double dl = log(1. / xk2s);
    const double elk = 1.38629436112 + xk2s * (0.09666344259 + xk2s * (0.03590092383 + xk2s * (0.03742563713 + xk2s * 0.01451196212))) +
                       dl * (0.5 + xk2s * (0.12498593597 + xk2s * (0.06880248576 + xk2s * (0.03328355346 + xk2s * 0.00441787012))));
    const double ele = 1. + xk2s * (0.44325141463 + xk2s * (0.0626060122 + xk2s * (0.04757383546 + xk2s * 0.01736506451))) +
                       dl * xk2s * (0.2499836831 + xk2s * (0.09200180037 + xk2s * (0.04069697526 + xk2s * 0.00526449639)));
    return ((1. - xk2 * 0.5) * elk - ele) / xkrho12;
}

/* the dipole-line mapping (alpha_s, gamma_s) -> (rho_s, z_s) common to ap and apprc */
void ts05_map(double alpha_s, double gamma_s, double *rhos, double *zs) {
    const double gammas2 = gamma_s * gamma_s;
    const double alsqh = (alpha_s * alpha_s) / 2.;
    const double f = 64. / 27. * gammas2 + (alsqh * alsqh);
    const double q = pow((sqrt(f) + alsqh), (1. / 3.));
    double c = q - 4. * pow(gammas2, (1. / 3.)) / (3. * q);
    if (c < 0.) c = 0.;
    const double g = sqrt(c * c + 4. * pow(gammas2, (1. / 3.)));
    const double rs = 4. / ((sqrt(2. * g - c) + sqrt(c)) * (g + c));
    const double costs = gamma_s * (rs * rs);
    const double sints = sqrt(1. - costs * costs);
    *rhos = rs * sints;
    *zs = rs * costs;
}

/* lib/ts05/full_rc.c:213-327: symmetric ring current vector potential */
double ts05_ap(double r, double sint, double cost) {
    int prox = 0;
    double sint1 = sint, cost1 = cost;
    if (sint1 < 1.E-2) {
        sint1 = 1.E-2;
        cost1 = .99994999875;
        prox = 1;
    }
    const double u1 = (r - 3.746064740) / 1.508802177, u2 = (r - 4.993638842) / 3.324180497;
    const double u3 = (r - 2.969226745) / 2.243367377;
    const double c1 = cost1 / 0.5873525737, c2 = cost1 / 0.4368407663;
    const double arg1 = -(u1 * u1) - (c1 * c1), arg2 = -(u2 * u2) - (c2 * c2), arg3 = -(u3 * u3);
    const double e1 = arg1 >= -500. ? exp(arg1) : 0.;
    const double e2 = arg2 >= -500. ? exp(arg2) : 0.;
    const double e3 = arg3 >= -500. ? exp(arg3) : 0.;
    const double alpha = (sint1 * sint1) / r;
    const double alpha_s = alpha * (1. + -0.2291904607 * e1 + 0.1556236119 * e2 + 0.1855957207 * e3);
    const double gamma_s = cost1 / (r * r);
    double rhos, zs;
    ts05_map(alpha_s, gamma_s, &rhos, &zs);
    const double zs2 = zs * zs;
    const double a1 = ts05_loops_aphi(rhos, zs2, 4.274684950, 2.439528329);
    const double a2 = ts05_loops_aphi(rhos, zs2, 3.367557287, 3.146382545);
    double ap = -456.5289941 * a1 + 375.9055332 * a2;
    if (prox) ap = ap * sint / sint1;
    return ap;
}

/* lib/ts05/full_rc.c:329-423: partial ring current, symmetric part */
double ts05_apprc(double r, double sint, double cost) {
    int prox = 0;
    double sint1 = sint, cost1 = cost;
    if (sint1 < 1.E-2) {
        sint1 = 1.E-2;
        cost1 = .99994999875;
        prox = 1;
    }
    const double alpha = (sint1 * sint1) / r;
    const double gamma = cost1 / pow(r, 2);
    const double gdg1 = gamma / .04742939676, gdg4 = gamma / .01512963586;
    const double a4 = (alpha - .2647095287) / .07091230197;
    const double arg1 = -(gdg1 * gdg1), arg2 = -(a4 * a4) - (gdg4 * gdg4);
    const double e1 = arg1 >= -500. ? exp(arg1) : 0.;
    const double e2 = arg2 >= -500. ? exp(arg2) : 0.;
    const double alpha_s =
        alpha * (1. + .3058309043 / pow((1. + pow(((alpha - .1817139853) / .1257532909), 2)), 3.422509402) * e1 +
                 -4.800458958 * (alpha - -.02845643596) / pow((1. + pow(((alpha - -.02845643596) / .2188114228), 2)), 2.545944574) /
                     pow((1. + pow((gamma / .00813272793), 2)), .35868244) +
                 103.1601001 * pow((alpha - -.00764731187), 2) /
                     pow((1. + pow(((alpha - -.00764731187) / .1046487459), 2)), 2.958863546) /
                     pow((1. + pow((gamma / .01172314188), 2)), .4382872938));
    const double gamma_s =
        gamma * (1. + .01134908150 + 14.51339943 * (alpha - .2647095287) * e2 +
                 6.861329631 * (alpha - .1677400816) / pow((1. + pow(((alpha - .1677400816) / .04433648846), 2)), .7665599464) /
                     pow((1. + pow((gamma / .05553741389), 2)), .7277854652));
    double rhos, zs;
    ts05_map(alpha_s, gamma_s, &rhos, &zs);
    const double zs2 = zs * zs;
    const double a1 = ts05_loops_aphi(rhos, zs2, 6.560486035, 1.930711037);
    const double a2 = ts05_loops_aphi(rhos, zs2, 3.827208119, .7789990504);
    double ap = -80.11202281 * a1 + 12.58246758 * a2;
    if (prox) ap = ap * sint / sint1;
    return ap;
}

/* lib/ts05/full_rc.c:161-211 and :329-368: B = curl(A_phi) of a vector potential by
 * central differences (kind 0: ap, kind 1: apprc) */
double ts05_aphi(int kind, double r, double s, double c) {
    return kind == 0 ? ts05_ap(r, s, c) : ts05_apprc(r, s, c);
}
void ts05_symm(int kind, double x, double y, double z, double b[3]) {
    const double ds = 1.E-2, dc = 0.99994999875, d = 1.E-4, drd = 5E3;
    const double rho2 = x * x + y * y;
    const double r2 = rho2 + z * z;
    const double r = sqrt(r2);
    const double rp = r + d, rm = r - d;
    const double sint = sqrt(rho2) / r, cost = z / r;
    double fxy, bz;
    if (sint < ds) {
        const double a = ts05_aphi(kind, r, ds, dc) / ds;
        const double dardr = (rp * ts05_aphi(kind, rp, ds, dc) - rm * ts05_aphi(kind, rm, ds, dc)) * drd;
        fxy = z * (2. * a - dardr) / (r * r2);
        bz = (2. * a * (cost * cost) + dardr * (sint * sint)) / r;
    } else {
        const double th = atan2(sint, cost);
        const double tp = th + d, tm = th - d;
        const double stp = sin(tp), stm = sin(tm), ctp = cos(tp), ctm = cos(tm);
        const double br = (stp * ts05_aphi(kind, r, stp, ctp) - stm * ts05_aphi(kind, r, stm, ctm)) / (r * sint) * drd;
        const double bt = (rm * ts05_aphi(kind, rm, sint, cost) - rp * ts05_aphi(kind, rp, sint, cost)) / r * drd;
        fxy = (br + bt * cost / sint) / r;
        bz = br * cost - bt * sint;
    }
    b[0] = fxy * x;
    b[1] = fxy * y;
    b[2] = bz;
}

/* lib/ts05/full_rc.c:689-695 */
void ts05_ffs(double a, double a0, double da, double *f, double *fa, double *fs) {
    const double da2 = da * da;
    const double p = a + a0, m = a - a0;
    const double sq1 = sqrt(p * p + da2), sq2 = sqrt(m * m + da2);
    *fa = 2. / (sq1 + sq2);
    *f = *fa * a;
    *fs = 0.5 * (sq1 + sq2) / (sq1 * sq2) * (1. - *f * *f);
}

/* lib/ts05/full_rc.c:482-593: radial component of the quadrupole partial ring current */
double ts05_br_prc_q(double r, double sint, double cost) {
    static const double A[18] = {-21.2666329, 32.24527521, -6.062894078, 7.515660734, 233.7341288,
                                 -227.1195714, 8.483233889, 16.80642754, -24.63534184, 9.067120578,
                                 -1.052686913, -12.08384538, 18.61969572, -12.71686069, 47017.35679,
                                 -50646.71204, 7746.058231, 1.531069371};
    const double sint2 = sint * sint, cost2 = cost * cost, sc = sint * cost;
    const double alpha = sint2 / r, gamma = cost / (r * r);
    double f, fa, fs, d[18];
    ts05_ffs(alpha, .1417519429, .6388013110E-02, &f, &fa, &fs);
    d[0] = sc * pow(f, 2.318824273) / (pow((r / 5.303934488), 4.213397467) + 1.);
    d[1] = d[0] * cost2;
    ts05_ffs(alpha, .1401142771, .2306094179E-01, &f, &fa, &fs);
    d[2] = sc * pow(fs, .7955534018) / (pow((r / 3.462235072), 2.568743010) + 1.);
    d[3] = d[2] * cost2;
    ts05_ffs(alpha, .1485233485, .2319676273E-01, &f, &fa, &fs);
    d[4] = sc * (pow(alpha, 3.477425908)) * (pow(fs, 1.922155110)) / (pow((r / 7.830223587), 8.492933868) + 1.);
    d[5] = d[4] * cost2;
    const double a4 = (alpha - .1295221828) / .01753008801, g1 = gamma / .01125504083;
    double arga = a4 * a4 + 1., argg = 1. + g1 * g1;
    d[6] = sc / arga / argg;
    d[7] = d[6] / arga;
    d[8] = d[7] / arga;
    d[9] = d[8] / arga;
    const double a5 = (alpha - .1811846095) / .04841237481, g2 = gamma / .01981805097;
    arga = a5 * a5 + 1.;
    argg = 1. + g2 * g2;
    d[10] = sc / arga / argg;
    d[11] = d[10] / arga;
    d[12] = d[11] / arga;
    d[13] = d[12] / arga;
    const double r4 = r * r * r * r;
    const double c1 = 6.557801891, c2 = 6.348576071, c3 = 5.744436687;
    d[14] = sc / (r4 + c1 * c1 * c1 * c1);
    d[15] = sc / (r4 + c2 * c2 * c2 * c2) * cost2;
    d[16] = sc / (r4 + c3 * c3 * c3 * c3) * (cost2 * cost2);
    ts05_ffs(alpha, .2265212965, .1301957209, &f, &fa, &fs);
    const double rd = (r - 1.2) / .5654023158;
    d[17] = sc * fs / (1. + rd * rd);
    double s = A[0] * d[0];
    for (int k = 1; k < 18; k++) s = s + A[k] * d[k];
    return s;
}

/* lib/ts05/full_rc.c:595-687: polar component of the quadrupole partial ring current */
double ts05_bt_prc_q(double r, double sint, double cost) {
    static const double A[17] = {12.74640393, -7.516393516, -5.476233865, 3.212704645, -59.10926169,
                                 46.62198189, -.01644280062, .1234229112, -.08579198697, .01321366966,
                                 .8970494003, 9.136186247, -38.19301215, 21.73775846, -410.0783424,
                                 -69.90832690, -848.8543440};
    const double sint2 = sint * sint, cost2 = cost * cost;
    const double alpha = sint2 / r, gamma = cost / (r * r);
    double f, fa, fs, d[17];
    ts05_ffs(alpha, .2071721360, .05030555417, &f, &fa, &fs);
    d[0] = pow(f, 1.243288286) / (pow((r / 7.471332374), 3.180533613) + 1.);
    d[1] = d[0] * cost2;
    ts05_ffs(alpha, .1568504222, .02092910682, &f, &fa, &fs);
    d[2] = pow(fa, 1.376743507) / pow(r, 1.985148197);
    d[3] = d[2] * cost2;
    ts05_ffs(alpha, .1701395257, .1019870070, &f, &fa, &fs);
    d[4] = pow(fs, .3157139940) * pow(alpha, 1.056309517) / (pow((r / 6.293740981), 5.671824276) + 1.);
    d[5] = d[4] * cost2;
    ts05_ffs(gamma, 0., .01040696080, &f, &fa, &fs);
    const double a4 = (alpha - .1280772299) / .02189060799;
    const double fcc = (1. + a4 * a4);
    d[6] = 1. / fcc * fs;
    d[7] = d[6] / fcc;
    d[8] = d[7] / fcc;
    d[9] = d[8] / fcc;
    const double a5 = (alpha - .1648265607) / .04701592613, g2 = gamma / .01526400086;
    const double arg = 1. + a5 * a5;
    d[10] = 1. / arg / (1. + g2 * g2);
    d[11] = d[10] / arg;
    d[12] = d[11] / arg;
    d[13] = d[12] / arg;
    const double r4 = r * r * r * r;
    const double c1 = 12.88384229, c2 = 3.361775101, c3 = 23.44173897;
    d[14] = 1. / (r4 + c1 * c1);
    d[15] = cost2 / (r4 + c2 * c2);
    d[16] = (cost2 * cost2) / (r4 + c3 * c3);
    double s = A[0] * d[0];
    for (int k = 1; k < 17; k++) s = s + A[k] * d[k];
    return s;
}

/* lib/ts05/full_rc.c:425-480 */
void ts05_prc_quad(double x, double y, double z, double b[3]) {
    const double d = 1.E-4, dd = 2.E-4, ds = 1.E-2, dc = 0.99994999875;
    const double rho2 = x * x + y * y;
    const double r = sqrt(rho2 + z * z);
    const double rho = sqrt(rho2);
    const double sint = rho / r, cost = z / r;
    const double rp = r + d, rm = r - d;
    if (sint > ds) {
        const double cphi = x / rho, sphi = y / rho;
        const double br = ts05_br_prc_q(r, sint, cost);
        const double bt = ts05_bt_prc_q(r, sint, cost);
        const double dbrr = (ts05_br_prc_q(rp, sint, cost) - ts05_br_prc_q(rm, sint, cost)) / dd;
        const double th = atan2(sint, cost);
        const double tp = th + d, tm = th - d;
        const double stp = sin(tp), ctp = cos(tp), stm = sin(tm), ctm = cos(tm);
        const double dbtt = (ts05_bt_prc_q(r, stp, ctp) - ts05_bt_prc_q(r, stm, ctm)) / dd;
        b[0] = sint * (br + (br + r * dbrr + dbtt) * (sphi * sphi)) + cost * bt;
        b[1] = -sint * sphi * cphi * (br + r * dbrr + dbtt);
        b[2] = (br * cost - bt * sint) * cphi;
    } else {
        const double st = ds;
        const double ct = z < 0. ? -dc : dc;
        const double th = atan2(st, ct);
        const double tp = th + d, tm = th - d;
        const double stp = sin(tp), ctp = cos(tp), stm = sin(tm), ctm = cos(tm);
        const double br = ts05_br_prc_q(r, st, ct);
        const double bt = ts05_bt_prc_q(r, st, ct);
        const double dbrr = (ts05_br_prc_q(rp, st, ct) - ts05_br_prc_q(rm, st, ct)) / dd;
        const double dbtt = (ts05_bt_prc_q(r, stp, ctp) - ts05_bt_prc_q(r, stm, ctm)) / dd;
        const double fcxy = r * dbrr + dbtt;
        const double rst = r * st;
        const double rst2 = rst * rst;
        const double y2 = y * y;
        b[0] = (br * (x * x + 2. * y2) + fcxy * y2) / rst2 + bt * cost;
        b[1] = -(br + fcxy) * x * y / rst2;
        b[2] = (br * cost / st - bt) * x / r;
    }
}

static const double C_SY[78] = {
    -957.2534900, -817.5450246, 583.2991249, 758.8568270, 13.17029064, 68.94173502, -15.29764089,
    -53.43151590, 27.34311724, 149.5252826, -11.00696044, -179.7031814, 953.0914774, 817.2340042,
    -581.0791366, -757.5387665, -13.10602697, -68.58155678, 15.22447386, 53.15535633, -27.07982637,
    -149.1413391, 10.91433279, 179.3251739, -6.028703251, 1.303196101, -1.345909343, -1.138296330,
    -0.06642634348, -0.3795246458, .07487833559, .2891156371, -.5506314391, -.4443105812,
    0.2273682152, 0.01086886655, -9.130025352, 1.118684840, 1.110838825, .1219761512, -.06263009645,
    -.1896093743, .03434321042, .01523060688, -.4913171541, -.2264814165, -.04791374574,
    .1981955976, -68.32678140, -48.72036263, 14.03247808, 16.56233733, 2.369921099, 6.200577111,
    -1.415841250, -0.8184867835, -3.401307527, -8.490692287, 3.217860767, -9.037752107, 66.09298105,
    48.23198578, -13.67277141, -16.27028909, -2.309299411, -6.016572391, 1.381468849, 0.7935312553,
    3.436934845, 8.260038635, -3.136213782, 8.833214943, 8.041075485, 8.024818618, 35.54861873,
    12.55415215, 1.738167799, 3.721685353};
static const double C_PR[78] = {
    -64820.58481, -63965.62048, 66267.93413, 135049.7504, -36.56316878, 124.6614669, 56.75637955,
    -87.56841077, 5848.631425, 4981.097722, -6233.712207, -10986.40188, 68716.52057, 65682.69473,
    -69673.32198, -138829.3568, 43.45817708, -117.9565488, -62.14836263, 79.83651604, -6211.451069,
    -5151.633113, 6544.481271, 11353.03491, 23.72352603, -256.4846331, 25.77629189, 145.2377187,
    -4.472639098, -3.554312754, 2.936973114, 2.682302576, 2.728979958, 26.43396781, -9.312348296,
    -29.65427726, -247.5855336, -206.9111326, 74.25277664, 106.4069993, 15.45391072, 16.35943569,
    -5.965177750, -6.079451700, 115.6748385, -35.27377307, -32.28763497, -32.53122151, 93.74409310,
    84.25677504, -29.23010465, -43.79485175, -6.434679514, -6.620247951, 2.443524317, 2.266538956,
    -43.82903825, 6.904117876, 12.24289401, 17.62014361, 152.3078796, 124.5505289, -44.58690290,
    -63.02382410, -8.999368955, -9.693774119, 3.510930306, 3.770949738, -77.96705716, 22.07730961,
    20.46491655, 18.67728847, 9.451290614, 9.313661792, 644.7620970, 418.2515954, 7.183754387,
    35.62128817};

/* lib/ts05/full_rc.c:4-159: symmetric (src) and partial (prc) ring currents with
 * their shields */
void ts05_full_rc(double x, double y, double z, const ts05_drv *v, double bs[3], double bp[3]) {
    const double phi = TA[37];
    double hs[3], ha[3], hq[3], fs[3], fp[3];
    ts05_symm(0, x / v->sc_sy, y / v->sc_sy, z / v->sc_sy, hs);
    const double xta = x / v->sc_pr, yta = y / v->sc_pr, zta = z / v->sc_pr;
    ts05_symm(1, xta, yta, zta, ha);
    const double cp = cos(phi), sp = sin(phi);
    ts05_prc_quad(xta * cp - yta * sp, xta * sp + yta * cp, zta, hq);
    const double bxq = hq[0] * cp + hq[1] * sp;
    const double byq = -hq[0] * sp + hq[1] * cp;
    double xsc = v->sc_sy - 1.;
    double x1 = xsc + 1.;
    ts05_shield(C_SY, xsc, x, y, z, 1, x1 * x1 * x1, fs);
    xsc = v->sc_pr - 1.;
    x1 = xsc + 1.;
    ts05_shield(C_PR, xsc, x, y, z, 1, x1 * x1 * x1, fp);
    bs[0] = hs[0] + fs[0];
    bs[1] = hs[1] + fs[1];
    bs[2] = hs[2] + fs[2];
    bp[0] = (ha[0] + bxq) + fp[0];
    bp[1] = (ha[1] + byq) + fp[1];
    bp[2] = (ha[2] + hq[2]) + fp[2];
}

/* ------------------------------------------------------------ the model */

/* lib/ts05/ts05.c:7-134 */
void o_ts05_gsm(const double *parmod, double x, double y, double z, double b[3]) {
    ts05_drv v;
    ts05_drivers(parmod, &v);
    const double xx = x * v.xappa, yy = y * v.xappa, zz = z * v.xappa;
    double xmxm = v.am + x - v.x0;
    if (xmxm < 0) xmxm = 0.;
    const double axx0 = xmxm * xmxm;
    const double asq = v.am * v.am;
    const double rho2 = y * y + z * z;
    const double aro = (asq + rho2) + axx0;
    const double sigma = sqrt((aro + sqrt(aro * aro - 4. * asq * axx0)) / (2. * asq));
    double q[3];
    if (!(sigma < 1.1960 + 0.005)) {
        ts05_dipole(x, y, z, q);
        b[0] = -q[0];
        b[1] = v.oimfy - q[1];
        b[2] = v.oimfz - q[2];
        return;
    }
    double cf[3], t1[3], t2[3], r11[3], r21[3], src[3], prc[3];
    ts05_shlcar3x3(xx, yy, zz, cf);
    for (int c = 0; c < 3; c++) cf[c] = cf[c] * v.xappa3;
    ts05_warped(xx, yy, zz, &v, t1, t2);
    double f[3], h[3];
    ts05_birk_1n2(1, xx, yy, zz, v.xkappa1, f);
    ts05_shield(SH11, v.xkappa1 - 1.1, xx, yy, zz, 0, 1., h);
    for (int c = 0; c < 3; c++) r11[c] = f[c] + h[c];
    ts05_birk_1n2(2, xx, yy, zz, v.xkappa2, f);
    ts05_shield(SH21, v.xkappa2 - 1.0, xx, yy, zz, 0, 1., h);
    for (int c = 0; c < 3; c++) r21[c] = f[c] + h[c];
    ts05_full_rc(xx, yy, zz, &v, src, prc);
    const double him[3] = {0., v.byimf, v.bzimf};
    double bb[3];
    for (int c = 0; c < 3; c++)
        bb[c] = TA[0] * cf[c] + v.tamp1 * t1[c] + v.tamp2 * t2[c] + v.a_src * src[c] + v.a_prc * prc[c] +
                v.a_r11 * r11[c] + v.a_r21 * r21[c] + TA[19] * him[c];
    if (sigma < 1.1960 - 0.005) {
        b[0] = bb[0];
        b[1] = bb[1];
        b[2] = bb[2];
        return;
    }
    ts05_dipole(x, y, z, q);
    const double fint = 0.5 * (1. - (sigma - 1.1960) / 0.005);
    const double fext = 0.5 * (1. + (sigma - 1.1960) / 0.005);
    b[0] = (bb[0] + q[0]) * fint - q[0];
    b[1] = (bb[1] + q[1]) * fint + v.oimfy * fext - q[1];
    b[2] = (bb[2] + q[2]) * fint + v.oimfz * fext - q[2];
}
