/*
 * o_t96.c -- oracle: Tsyganenko-96 external field at zero dipole tilt, restated
 * from lib/t96/t96.c in plain C (operation order kept for bit parity).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  GSM position in Re -> GSM field in nT.
 * The model is the published T96_01 with the tilt terms dropped, plus the
 * reference's own dipole form (lib/t96/t96.c:1645-1656).
 */
#include <math.h>
#include "oracle.h"

typedef struct { double x, y, z; } v3;

/* lib/t96/t96.c:1645-1656: dipole as the reference writes it; the x component
 * carries the (T+U-2P) term of its own form. */
static v3 t96_dipole(double x, double y, double z) {
    const double p = x * x, u = z * z, v = 3.0 * z * x, t = y * y;
    const double r = sqrt(p + t + u);
    const double q = 30574.0 / (r * r * r * r * r);
    v3 b = {q * ((t + u - 2.0 * p) - v), -3.0 * y * q * z, q * (p + t - 2.0 * u)};
    return b;
}

/* lib/t96/t96.c:212-234: polynomial / asymptotic J0 and J1. */
static double bessel_j0(double x) {
    if (fabs(x) < 3.0) {
        const double q = (x / 3.0) * (x / 3.0);
        return 1.0 - q * (2.2499997 - q * (1.2656208 - q * (0.3163866 - q * (0.0444479 - q * (0.0039444 - q * 0.00021)))));
    }
    const double u = 3.0 / x;
    const double f = 0.79788456 - u * (0.00000077 + u * (0.00552740 + u * (0.00009512 - u * (0.00137237 - u * (0.00072805 - u * 0.00014476)))));
    const double th = x - 0.78539816 - u * (0.04166397 + u * (0.00003954 - u * (0.00262573 - u * (0.00054125 + u * (0.00029333 - u * 0.00013558)))));
    return f / sqrt(x) * cos(th);
}

static double bessel_j1(double x) {
    if (fabs(x) < 3.0) {
        const double q = (x / 3.0) * (x / 3.0);
        const double j1x = 0.5 - q * (0.56249985 - q * (0.21093573 - q * (0.03954289 - q * (0.00443319 - q * (0.00031761 - q * 0.00001109)))));
        return j1x * x;
    }
    const double u = 3.0 / x;
    const double f = 0.79788456 + u * (0.00000156 + u * (0.01659667 + u * (0.00017105 - u * (0.00249511 - u * (0.00113653 - u * 0.00020033)))));
    const double th = x - 2.35619449 + u * (0.12499612 + u * (0.0000565 - u * (0.00637879 - u * (0.00074348 + u * (0.00079824 - u * 0.00029166)))));
    return f / sqrt(x) * cos(th);
}

/* lib/t96/t96.c:113-163: Chapman-Ferraro shielding in cylindrical harmonics. */
static v3 t96_cylharm(double x, double y, double z) {
    static const double a[12] = {0.24777, -27.003, -0.46815, 7.0637, -1.5918, -0.90317E-01,
                                 57.522, 13.757, 2.0100, 10.458, 4.5798, 2.1695};
    double rho = sqrt(y * y + z * z);
    double sinfi, cosfi;
    if (rho < 1.0e-8) {
        sinfi = 1.0;
        cosfi = 0.0;
        rho = 1.0e-8;
    } else {
        sinfi = z / rho;
        cosfi = y / rho;
    }
    const double sinfi2 = sinfi * sinfi;
    const double si2co2 = sinfi2 - cosfi * cosfi;
    v3 b = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        const double dz = rho / a[i + 6];
        const double j0 = bessel_j0(dz), j1 = bessel_j1(dz);
        const double e = exp(x / a[i + 6]);
        b.x -= a[i] * j1 * e * sinfi;
        b.y += a[i] * (2.0 * j1 / dz - j0) * e * sinfi * cosfi;
        b.z += a[i] * (j1 / dz * si2co2 - j0 * sinfi2) * e;
    }
    for (int i = 3; i < 6; i++) {
        const double dz = rho / a[i + 6];
        const double xk = x / a[i + 6];
        const double j0 = bessel_j0(dz), j1 = bessel_j1(dz);
        const double e = exp(xk);
        const double brho = (xk * j0 - (dz * dz + xk - 1.0) * j1 / dz) * e * sinfi;
        const double bphi = (j0 + j1 / dz * (xk - 1.0)) * e * cosfi;
        b.x += a[i] * (dz * j0 + xk * j1) * e * sinfi;
        b.y += a[i] * (brho * cosfi - bphi * sinfi);
        b.z += a[i] * (brho * sinfi + bphi * cosfi);
    }
    return b;
}

/* lib/t96/t96.c:236-271: IMF interconnection field (3x3 Cartesian harmonics). */
static v3 t96_intercon(double x, double y, double z) {
    static const double a[9] = {-8.411078731, 5932254.951, -9073284.93, -11.68794634,
                                6027598.824, -9218378.368, -6.508798398, -11824.42793,
                                18015.66212};
    static const double rp[3] = {0.12503844191332988, 0.071597395204476152, 0.011080976889832743};
    static const double rr[3] = {0.059675527522452816, 0.00098459526668325296, 0.00064371058922777988};
    v3 b = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        const double cy = cos(y * rp[i]), sy = sin(y * rp[i]);
        for (int k = 0; k < 3; k++) {
            const double sz = sin(z * rr[k]), cz = cos(z * rr[k]);
            const double sq = sqrt(rp[i] * rp[i] + rr[k] * rr[k]);
            const double e = exp(x * sq);
            const double c = a[3 * i + k];
            b.x += c * (-sq * e * cy * sz);
            b.y += c * (rp[i] * e * sy * sz);
            b.z += c * (-rr[k] * e * cy * cz);
        }
    }
    return b;
}

/* lib/t96/t96.c:322-381: ring current (two conical-sheet terms). */
static v3 t96_ringcurr(double x, double y, double z) {
    static const double f[2] = {569.895366, -1603.386993};
    static const double beta[2] = {2.722188, 3.766875};
    const double dzs = sqrt(z * z + 4.0);
    const double rhos = sqrt(x * x + y * y);
    const double ddz_dz = z / dzs;
    double drho_dx = 0.0, drho_dy = (y >= 0) ? 1.0 : -1.0;
    if (rhos >= 1.0e-5) {
        drho_dx = x / rhos;
        drho_dy = y / rhos;
    }
    v3 b = {0.0, 0.0, 0.0};
    for (int i = 0; i < 2; i++) {
        const double bi = beta[i];
        const double zb = dzs + bi, rp = rhos + bi, rm = rhos - bi;
        const double zb2 = zb * zb, b2 = 2.0 * bi;
        const double s1 = sqrt(zb2 + rp * rp), s2 = sqrt(zb2 + rm * rm);
        const double ds1_dz = zb / s1, ds2_dz = zb / s2;
        const double ds1_dr = rp / s1, ds2_dr = rm / s2;
        const double s12 = s1 * s2, sp = s1 + s2, sp2 = sp * sp;
        const double fac1 = sqrt(sp2 - b2 * b2);
        const double as = fac1 / (s12 * sp2);
        const double term1 = 1.0 / (s12 * sp * fac1);
        const double fac2 = as / sp2;
        const double das1 = term1 - fac2 / s1 * (s2 * s2 + s1 * (3.0 * s1 + 4.0 * s2));
        const double das2 = term1 - fac2 / s2 * (s1 * s1 + s2 * (3.0 * s2 + 4.0 * s1));
        const double dasx = das1 * (ds1_dr * drho_dx) + das2 * (ds2_dr * drho_dx);
        const double dasy = das1 * (ds1_dr * drho_dy) + das2 * (ds2_dr * drho_dy);
        const double dasz = das1 * (ds1_dz * ddz_dz) + das2 * (ds2_dz * ddz_dz);
        b.x += -f[i] * x * dasz;
        b.y -= f[i] * y * dasz;
        b.z += f[i] * ((2.0 * as + y * dasy) + x * dasx);
    }
    return b;
}

/* lib/t96/t96.c:383-443: tail disk (four conical-sheet terms, warped in y). */
static v3 t96_taildisk(double x, double y, double z) {
    static const double f[4] = {-745796.7338, 1176470.141, -444610.529, -57508.01028};
    static const double beta[4] = {7.9250000, 8.0850000, 8.4712500, 27.89500};
    const double y20 = y / 20.0;
    const double dd = 2.0 + 10.0 * (y20 * y20);
    const double dzs = sqrt(z * z + dd * dd);
    const double ddz_dy = (dd * 10.0 * y * 0.005) / dzs;
    const double ddz_dz = z / dzs;
    const double xs = x - 4.5;
    const double rhos = sqrt(xs * xs + y * y);
    double drho_dx = 0.0, drho_dy = (y >= 0) ? 1.0 : -1.0;
    if (rhos >= 1.0e-5) {
        drho_dx = xs / rhos;
        drho_dy = y / rhos;
    }
    v3 b = {0.0, 0.0, 0.0};
    for (int i = 0; i < 4; i++) {
        const double bi = beta[i], b2 = 2.0 * bi;
        const double zb = dzs + bi, rp = rhos + bi, rm = rhos - bi;
        const double s1 = sqrt(zb * zb + rp * rp), s2 = sqrt(zb * zb + rm * rm);
        const double ds1_dz = zb / s1, ds2_dz = zb / s2;
        const double ds1_dr = rp / s1, ds2_dr = rm / s2;
        const double s12 = s1 * s2, sp = s1 + s2, sp2 = sp * sp;
        const double fac1 = sqrt(sp2 - b2 * b2);
        const double as = fac1 / (s12 * sp2);
        const double term1 = 1.0 / (s12 * sp * fac1);
        const double fac2 = as / sp2;
        const double das1 = term1 - fac2 / s1 * (s2 * s2 + s1 * (3.0 * s1 + 4.0 * s2));
        const double das2 = term1 - fac2 / s2 * (s1 * s1 + s2 * (3.0 * s2 + 4.0 * s1));
        const double dasx = das1 * (ds1_dr * drho_dx) + das2 * (ds2_dr * drho_dx);
        const double dasy = das1 * (ds1_dz * ddz_dy + ds1_dr * drho_dy) + das2 * (ds2_dz * ddz_dy + ds2_dr * drho_dy);
        const double dasz = das1 * (ds1_dz * ddz_dz) + das2 * (ds2_dz * ddz_dz);
        b.x += f[i] * ((4.5 - x) * dasz);
        b.y -= f[i] * y * dasz;
        b.z += f[i] * ((2.0 * as + y * dasy) + xs * dasx);
    }
    return b;
}

/* lib/t96/t96.c:445-499: the T87-style long tail current sheet (no y component). */
static void t96_tail87(double x, double z, double *bx, double *bz) {
    const double zp = z - 40.0, zm = z + 40.0;
    const double xnx = -10.0 - x, xnx2 = xnx * xnx;
    const double xc1 = x + 1.261, xc2 = x + 0.663;
    const double xc22 = xc2 * xc2, xr2 = xc2 * -0.1071, xc12 = xc1 * xc1;
    const double d2 = 3.0 * 3.0;
    const double b20 = z * z + d2, b2p = zp * zp + d2, b2m = zm * zm + d2;
    const double b = sqrt(b20), bp = sqrt(b2p), bm = sqrt(b2m);
    const double xa1 = xc12 + b20, xap1 = xc12 + b2p, xam1 = xc12 + b2m;
    const double xa2 = 1. / (xc22 + b20), xap2 = 1. / (xc22 + b2p), xam2 = 1. / (xc22 + b2m);
    const double xna = xnx2 + b20, xnap = xnx2 + b2p, xnam = xnx2 + b2m;
    const double ff = b20 - xc22, fp = b2p - xc22, fm = b2m - xc22;
    const double ln1 = log(76.37 / xna), lnp1 = log(76.37 / xnap), lnm1 = log(76.37 / xnam);
    const double ln2 = ln1 + 0.13238005, lnp2 = lnp1 + 0.13238005, lnm2 = lnm1 + 0.13238005;
    const double aln = 0.25 * (lnp1 + lnm1 - 2. * ln1);
    const double s0 = (atan(xnx / b) + 1.5707963) / b;
    const double s0p = (atan(xnx / bp) + 1.5707963) / bp;
    const double s0m = (atan(xnx / bm) + 1.5707963) / bm;
    const double s1 = (ln1 * .5 + xc1 * s0) / xa1;
    const double s1p = (lnp1 * .5 + xc1 * s0p) / xap1;
    const double s1m = (lnm1 * .5 + xc1 * s0m) / xam1;
    const double s2 = (xc2 * xa2 * ln2 + 0.1071 - ff * xa2 * s0) * xa2;
    const double s2p = (xc2 * xap2 * lnp2 + 0.1071 - fp * xap2 * s0p) * xap2;
    const double s2m = (xc2 * xam2 * lnm2 + 0.1071 - fm * xam2 * s0m) * xam2;
    const double g1 = (b20 * s0 - 0.5 * xc1 * ln1) / xa1;
    const double g1p = (b2p * s0p - 0.5 * xc1 * lnp1) / xap1;
    const double g1m = (b2m * s0m - 0.5 * xc1 * lnm1) / xam1;
    const double g2 = ((0.5 * ff * ln2 + 2. * s0 * b20 * xc2) * xa2 + xr2) * xa2;
    const double g2p = ((0.5 * fp * lnp2 + 2. * s0p * b2p * xc2) * xap2 + xr2) * xap2;
    const double g2m = ((0.5 * fm * lnm2 + 2. * s0m * b2m * xc2) * xam2 + xr2) * xam2;
    *bx = 0.391734 * (z * s0 - 0.5 * (zp * s0p + zm * s0m)) + 5.89715 * (z * s1 - 0.5 * (zp * s1p + zm * s1m)) + 24.6833 * (z * s2 - 0.5 * (zp * s2p + zm * s2m));
    *bz = 0.391734 * aln + 5.89715 * (g1 - 0.5 * (g1p + g1m)) + 24.6833 * (g2 - 0.5 * (g2p + g2m));
}

/* lib/t96/t96.c:501-617: 3x3 Cartesian-harmonic shielding field; the pair sums
 * a[2j] + a[2j+1] are the zero-tilt combination of the two parity halves.
 * a[18..20] are the y scales, a[21..23] the z scales. */
static v3 t96_shlcar3x3(const double *a, double x, double y, double z) {
    v3 b = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        const double p = a[18 + i];
        const double pp = 1.0 / (p * p);
        const double yp = y / p;
        const double sy = sin(yp), cy = cos(yp);
        for (int k = 0; k < 3; k++) {
            const double r = a[21 + k];
            const double zr = z / r;
            const double sz = sin(zr), cz = cos(zr);
            const double sq = sqrt(pp + 1.0 / (r * r));
            const double e = exp(x * sq);
            const int j = 3 * i + k;
            const double c = a[2 * j] + a[2 * j + 1];
            b.x += (-sq * e * cy * sz) * c;
            b.y += (e / p * sy * sz) * c;
            b.z += (-e / r * cy * cz) * c;
        }
    }
    return b;
}

/* lib/t96/t96.c:273-320: ring current + tail modes 2 and 3 with their shields. */
static void t96_tailrc(double x, double y, double z, v3 *rc, v3 *t2, v3 *t3) {
    static const double arc[24] = {
        -3.087699646, 3.516259114, 18.81380577, -13.95772338, -5.497076303,
        0.1712890838, 2.392629189, -2.728020808, -14.79349936, 11.08738083,
        4.388174084, 0.2492163197E-01, 0.7030375685, -.7966023165, -3.835041334,
        2.642228681, -0.2405352424, -0.7297705678, 12.75434981, 11.37659788,
        636.4346279, 1.752483754, 3.604231143, 12.83078674};
    static const double atail2[24] = {
        0.8747515218, -0.9116821411, 2.209365387, -2.159059518, -7.059828867,
        5.924671028, -1.916935691, 1.996707344, -3.877101873, 3.947666061,
        11.38715899, -8.343210833, 1.194109867, -1.244316975, 3.73895491,
        -4.406522465, -20.66884863, 3.020952989, 5.187666221, 6.802472048,
        39.13543412, 2.784722096, 6.979576616, 25.71716760};
    static const double atail3[24] = {
        -19091.95061, -3011.613928, 20582.16203, 4242.918430, -2377.091102,
        -1504.820043, 19884.04650, 2725.150544, -21389.04845, -3990.475093,
        2401.610097, 1548.171792, -946.5493963, 490.1528941, 986.9156625,
        -489.3265930, -67.99278499, 8.711175710, 19.69877970, 20.30095680,
        86.45407420, 22.50403727, 23.41617329, 48.48140573};
    const v3 h = t96_ringcurr(x, y, z);
    v3 w = t96_shlcar3x3(arc, x, y, z);
    rc->x = w.x + h.x;
    rc->y = w.y + h.y;
    rc->z = w.z + h.z;
    w = t96_shlcar3x3(atail2, x, y, z);
    const v3 d = t96_taildisk(x, y, z);
    t2->x = w.x + d.x;
    t2->y = w.y + d.y;
    t2->z = w.z + d.z;
    w = t96_shlcar3x3(atail3, x, y, z);
    double hx, hz;
    t96_tail87(x, z, &hx, &hz);
    t3->x = w.x + hx;
    t3->y = w.y;
    t3->z = w.z + hz;
}

/* lib/t96/t96.c:854-881: field of a circular loop of radius rl in the z=0 plane
 * (elliptic integrals by Hastings polynomials). */
static v3 t96_circle(double x, double y, double z, double rl) {
    const double rho2 = x * x + y * y;
    const double rho = sqrt(rho2);
    const double rr = rho + rl;
    const double r22 = z * z + rr * rr;
    const double r2 = sqrt(r22);
    const double r12 = r22 - 4.0 * rho * rl;
    const double r32 = 0.5 * (r12 + r22);
    const double xk2 = 1.0 - r12 / r22;
    const double xk2s = 1.0 - xk2;
    const double dl = log(1.0 / xk2s);
    const double kk = 1.38629436112 + xk2s * (0.09666344259 + xk2s * (0.03590092383 +
                      xk2s * (0.03742563713 + xk2s * 0.01451196212))) + dl *
                      (0.5 + xk2s * (0.12498593597 + xk2s * (0.06880248576 +
                      xk2s * (0.03328355346 + xk2s * 0.00441787012))));
    const double ee = 1.0 + xk2s * (0.44325141463 + xk2s * (0.0626060122 + xk2s *
                      (0.04757383546 + xk2s * 0.01736506451))) + dl *
                      xk2s * (0.2499836831 + xk2s * (0.09200180037 + xk2s *
                      (0.04069697526 + xk2s * 0.00526449639)));
    double brho;
    if (rho > 1e-6) {
        brho = z / (rho2 * r2) * (r32 / r12 * ee - kk);
    } else {
        brho = 3.141592654 * rl / r2 * (rl - rho) / r12 * z / (r32 - rho2);
    }
    v3 b = {brho * x, brho * y, (kk - ee * (r32 - 2.0 * rl * rl) / r12) / r2};
    return b;
}

/* lib/t96/t96.c:883-898: two loops tilted by +-al about the x axis, shifted by xc. */
static v3 t96_crosslp(double x, double y, double z, double xc, double rl, double al) {
    const double cal = cos(al), sal = sin(al);
    const double y1 = y * cal - z * sal, z1 = y * sal + z * cal;
    const double y2 = y * cal + z * sal, z2 = -y * sal + z * cal;
    const v3 b1 = t96_circle(x - xc, y1, z1, rl);
    const v3 b2 = t96_circle(x - xc, y2, z2, rl);
    v3 b = {b1.x + b2.x, (b1.y + b2.y) * cal + (b1.z - b2.z) * sal,
            -(b1.y - b2.y) * sal + (b1.z + b2.z) * cal};
    return b;
}

/* lib/t96/t96.c:900-918: gradient tensor of a unit dipole (columns x, y, z moments). */
static void t96_dipxyz(double x, double y, double z, v3 *mx, v3 *my, v3 *mz) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    const double q = 30574.0 / (r2 * r2 * sqrt(r2));
    const double q3 = 3.0 * q;
    mx->x = q * (3.0 * x2 - r2);
    mx->y = q3 * x * y;
    mx->z = q3 * x * z;
    my->x = mx->y;
    my->y = q * (3.0 * y2 - r2);
    my->z = q3 * y * z;
    mz->x = mx->z;
    mz->y = my->z;
    mz->z = q * (3.0 * z2 - r2);
}

/* lib/t96/t96.c:809-852: Region-1 basis near the noon/midnight meridian, 12 dipole
 * pairs plus one crossed-loop pair and one shifted loop -> d[3][14]. */
static void t96_diploop1(double x, double y, double z, double d[3][14]) {
    static const double xx[12] = {-11.0, -7.0, -7.0, -3.0, -3.0, 1.0, 1.0, 1.0, 5.0, 5.0, 9.0, 9.0};
    static const double yy[12] = {2.0, 0.0, 4.0, 2.0, 6.0, 0.0, 4.0, 8.0, 2.0, 6.0, 0.0, 4.0};
    const double z2 = z * z;
    for (int i = 0; i < 12; i++) {
        const double yd = yy[i] * 0.945719;
        const double xd = xx[i] * 1.12541;
        const double dx = x - xd;
        double dy = y - yd;
        double r2 = dx * dx + dy * dy + z2;
        double q = 30574.0 / (r2 * r2 * sqrt(r2));
        double q3 = 3.0 * q;
        const double b1x = q3 * dx * z, b1y = q3 * dy * z, b1z = q * (3.0 * z2 - r2);
        double b2x = 0.0, b2y = 0.0, b2z = 0.0;
        if (fabs(yd) > 1.0e-10) {
            dy = y + yd;
            r2 = dx * dx + dy * dy + z2;
            q = 30574.0 / (r2 * r2 * sqrt(r2));
            q3 = 3.0 * q;
            b2x = q3 * dx * z;
            b2y = q3 * dy * z;
            b2z = q * (3.0 * z2 - r2);
        }
        d[0][i] = b1x + b2x;
        d[1][i] = b1y + b2y;
        d[2][i] = b1z + b2z;
    }
    const v3 oct = t96_crosslp(x, y, z, 2.28397, 1.86106, 1.00891);
    d[0][12] = oct.x;
    d[1][12] = oct.y;
    d[2][12] = oct.z;
    const v3 cl = t96_circle(x + 5.60831, y, z, 7.83281);
    d[0][13] = cl.x;
    d[1][13] = cl.y;
    d[2][13] = cl.z;
}

/* lib/t96/t96.c:920-1028: Region-1 basis away from the meridian: 5 conical
 * harmonics, 9 quadruplets of dipoles (3 moments each), 5 dipole pairs on the axis
 * (2 moments each) -> d[3][42]. */
static void t96_condip1(double x, double y, double z, double d[3][42]) {
    static const double xx[14] = {-10.0, -7.0, -4.0, -4.0, 0.0, 4.0, 4.0, 7.0, 10.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    static const double yy[14] = {3.0, 6.0, 3.0, 9.0, 6.0, 3.0, 9.0, 6.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    static const double zz[14] = {20.0, 20.0, 4.0, 20.0, 4.0, 4.0, 20.0, 20.0, 20.0, 2.0, 3.0, 4.50, 7.0, 10.0};
    const double xsm = x + 0.16;
    const double ro2 = xsm * xsm + y * y;
    const double ro = sqrt(ro2);
    double cf[5], sf[5];
    cf[0] = xsm / ro;
    sf[0] = y / ro;
    cf[1] = cf[0] * cf[0] - sf[0] * sf[0];
    sf[1] = 2.0 * sf[0] * cf[0];
    for (int k = 2; k < 5; k++) {
        cf[k] = cf[k - 1] * cf[0] - sf[k - 1] * sf[0];
        sf[k] = sf[k - 1] * cf[0] + cf[k - 1] * sf[0];
    }
    const double r2 = ro2 + z * z;
    const double r = sqrt(r2);
    const double c = z / r, s = ro / r;
    const double ch = sqrt(0.5 * (1.0 + c)), sh = sqrt(0.5 * (1.0 - c));
    const double tnh = sh / ch, cnh = 1.0 / tnh;
    for (int m = 0; m < 5; m++) {
        const double m1 = m + 1;
        const double bt = m1 * cf[m] / (r * s) * (pow(tnh, m1) + pow(cnh, m1));
        const double bf = -0.5 * m1 * sf[m] / r * (pow(tnh, m) / (ch * ch) - pow(cnh, m) / (sh * sh));
        d[0][m] = bt * c * cf[0] - bf * sf[0];
        d[1][m] = bt * c * sf[0] + bf * cf[0];
        d[2][m] = -bt * s;
    }
    for (int i = 0; i < 9; i++) {
        const double sc = (i == 2 || i == 4 || i == 5) ? 0.08 : 0.4;
        const double xd = xx[i] * sc, yd = yy[i] * sc, zd = zz[i];
        v3 m1x, m1y, m1z, m2x, m2y, m2z, m3x, m3y, m3z, m4x, m4y, m4z;
        t96_dipxyz(x - xd, y - yd, z - zd, &m1x, &m1y, &m1z);
        t96_dipxyz(x - xd, y + yd, z - zd, &m2x, &m2y, &m2z);
        t96_dipxyz(x - xd, y - yd, z + zd, &m3x, &m3y, &m3z);
        t96_dipxyz(x - xd, y + yd, z + zd, &m4x, &m4y, &m4z);
        const int ix = (i + 1) * 3 + 2, iy = ix + 1, iz = iy + 1;
        d[0][ix] = m1x.x + m2x.x - m3x.x - m4x.x;
        d[1][ix] = m1x.y + m2x.y - m3x.y - m4x.y;
        d[2][ix] = m1x.z + m2x.z - m3x.z - m4x.z;
        d[0][iy] = m1y.x - m2y.x - m3y.x + m4y.x;
        d[1][iy] = m1y.y - m2y.y - m3y.y + m4y.y;
        d[2][iy] = m1y.z - m2y.z - m3y.z + m4y.z;
        d[0][iz] = m1z.x + m2z.x + m3z.x + m4z.x;
        d[1][iz] = m1z.y + m2z.y + m3z.y + m4z.y;
        d[2][iz] = m1z.z + m2z.z + m3z.z + m4z.z;
    }
    for (int i = 1; i <= 5; i++) {
        const double zd = zz[i + 8];
        v3 m1x, m1y, m1z, m2x, m2y, m2z;
        t96_dipxyz(x, y, z - zd, &m1x, &m1y, &m1z);
        t96_dipxyz(x, y, z + zd, &m2x, &m2y, &m2z);
        const int ix = 30 + i * 2, iz = ix + 1;
        d[0][ix] = m1x.x - m2x.x;
        d[1][ix] = m1x.y - m2x.y;
        d[2][ix] = m1x.z - m2x.z;
        d[0][iz] = m1z.x + m2z.x;
        d[1][iz] = m1z.y + m2z.y;
        d[2][iz] = m1z.z + m2z.z;
    }
}

/* lib/t96/t96.c:1030-1194: Region-1 shield, 4x4 Cartesian harmonics; the
 * coefficient of each term is the zero-tilt pair sum. */
static v3 t96_birk1shld(double x, double y, double z) {
    static const double py[4] = {5.303648988, 10.40368955, 69.65230348, 466.5099509};
    static const double pz[4] = {1.645049286, 3.825838190, 11.66675599, 558.9781177};
    static const double amp[4][4] = {
        {1.174198045 + (double)-1.463820502, 4.840161537 + (double)-3.674506864,
         82.18368896 + (double)-94.94071588, -4122.331796 + (double)4670.278676},
        {-21.54975037 + (double)26.72661293, -72.81365728 + (double)44.09887902,
         40.08073706 + (double)-51.23563510, 1955.348537 + (double)-1940.971550},
        {794.0496433 + (double)-982.2441344, 1889.837171 + (double)-558.9779727,
         -1260.543238 + (double)1260.063802, -293.5942373 + (double)344.7250789},
        {-773.7002492 + (double)957.0094135, -1824.143669 + (double)520.7994379,
         1192.484774 + (double)-1192.184565, 89.15537624 + (double)-98.52042999}};
    double rk[4], rk2[4], szk[4], czk[4];
    for (int k = 0; k < 4; k++) {
        rk[k] = 1.0 / pz[k];
        const double zr = z * rk[k];
        szk[k] = sin(zr);
        czk[k] = cos(zr);
        rk2[k] = rk[k] * rk[k];
    }
    v3 b = {0.0, 0.0, 0.0};
    for (int i = 0; i < 4; i++) {
        const double rp = 1.0 / py[i];
        const double rp2 = rp * rp;
        const double ry = rp * y;
        const double sy = sin(ry), cy = cos(ry);
        for (int k = 0; k < 4; k++) {
            const double sq = sqrt(rp2 + rk2[k]);
            const double e = exp(x * sq);
            const double c = amp[i][k];
            b.x += (-sq * e * cy * szk[k]) * c;
            b.y += (rp * e * sy * szk[k]) * c;
            b.z += (-rk[k] * e * cy * czk[k]) * c;
        }
    }
    return b;
}

/* lib/t96/t96.c:619-807: total Region-1 field: diploop1 / condip1 expansions chosen
 * by the dipole-mapped colatitude TET0, linear blend across the +-2 degree bands
 * around the noon and midnight boundaries, plus the shield. */
static v3 t96_birk1tot(double x, double y, double z) {
    static const double c1[14] = {
        -0.911582E-03, -0.376654E-02, -0.727423E-02, -0.270084E-02,
        -0.123899E-02, -0.154387E-02, -0.340040E-02, -0.191858E-01,
        -0.518979E-01, 0.635061E-01, 0.440680, -0.396570, -89.5587, 23.2806};
    static const double c2[42] = {
        6.04133, 0.305415, 0.606066E-02, 0.128379E-03, -0.179406E-04,
        1.41714, -27.2586, -4.28833, -1.30675, 35.5607,
        8.95792, 0.961617E-03, -0.801477E-03, -0.782795E-03, -1.65242,
        -16.5242, -5.33798, 0.424878E-03, 0.331787E-03, -0.704305E-03,
        0.844342E-03, 0.953682E-04, 0.886271E-03, 25.1120, 20.9299,
        5.14569, -44.1670, -51.0672, -1.87725, 20.2998,
        48.7505, -2.97415, 0.174615E-03, -0.789777E-06, 0.686047E-03,
        0.460104E-04, -0.345216E-02, 0.221871E-02, 0.110078E-01, -0.661373E-02,
        0.249201E-02, 0.343978E-01};
    const double tnoonn = 12.0 * 0.01745329;
    const double tnoons = 3.141592654 - tnoonn;
    const double r2 = x * x + y * y + z * z;
    const double r = sqrt(r2);
    const double r3 = r * r2;
    double pas = 0.0;
    if (x != 0.0 || y != 0.0) pas = atan2(y, x);
    const double tas = atan2(sqrt(x * x + y * y), z);
    const double stas = sin(tas);
    double tet0 = asin(stas / pow((stas * stas * stas * stas * stas * stas) * (1.0 - r3) + r3, 0.1666666667));
    if (tas > 1.5707963) tet0 = 3.141592654 - tet0;
    const double dtet = (8.0 * 0.01745329) * (sin(pas * 0.5) * sin(pas * 0.5));
    const double tetr1n = tnoonn + dtet, tetr1s = tnoons - dtet;
    const double t01n = tetr1n - 0.034906, t01s = tetr1s - 0.034906;
    const double t02s = tetr1s + 0.034906, t02n = tetr1n + 0.034906;

    double d1[3][14], d2[3][42];
    v3 b = {0.0, 0.0, 0.0};
    if (tet0 < t01n || tet0 > t02s) {
        t96_diploop1(x, y, z, d1);
        for (int i = 0; i < 14; i++) {
            b.x += c1[i] * d1[0][i];
            b.y += c1[i] * d1[1][i];
            b.z += c1[i] * d1[2][i];
        }
    }
    if (tet0 > t02n && tet0 < t01s) {
        t96_condip1(x, y, z, d2);
        for (int i = 0; i < 42; i++) {
            b.x += c2[i] * d2[0][i];
            b.y += c2[i] * d2[1][i];
            b.z += c2[i] * d2[2][i];
        }
    }
    /* the two blend bands: north (loop side first) and south (conical side first) */
    for (int band = 0; band < 2; band++) {
        const double ta = band == 0 ? t01n : t01s;
        const double tb = band == 0 ? t02n : t02s;
        if (!(tet0 >= ta && tet0 <= tb)) continue;
        const double sqr = sqrt(r);
        const double sa = sin(ta);
        const double st1 = sqr / pow(r3 + 1.0 / (sa * sa * sa * sa * sa * sa) - 1.0, 0.1666666667);
        const double sb = sin(tb);
        const double st2 = sqr / pow(r3 + 1.0 / (sb * sb * sb * sb * sb * sb) - 1.0, 0.1666666667);
        double ct1 = sqrt(1.0 - st1 * st1), ct2 = sqrt(1.0 - st2 * st2);
        if (band == 1) {
            ct1 = -ct1;
            ct2 = -ct2;
        }
        const double x1 = r * st1 * cos(pas), y1 = r * st1 * sin(pas), z1 = r * ct1;
        const double x2 = r * st2 * cos(pas), y2 = r * st2 * sin(pas), z2 = r * ct2;
        v3 e1 = {0.0, 0.0, 0.0}, e2 = {0.0, 0.0, 0.0};
        if (band == 0) {
            t96_diploop1(x1, y1, z1, d1);
            for (int i = 0; i < 14; i++) {
                e1.x += c1[i] * d1[0][i];
                e1.y += c1[i] * d1[1][i];
                e1.z += c1[i] * d1[2][i];
            }
            t96_condip1(x2, y2, z2, d2);
            for (int i = 0; i < 42; i++) {
                e2.x += c2[i] * d2[0][i];
                e2.y += c2[i] * d2[1][i];
                e2.z += c2[i] * d2[2][i];
            }
        } else {
            t96_condip1(x1, y1, z1, d2);
            for (int i = 0; i < 42; i++) {
                e1.x += c2[i] * d2[0][i];
                e1.y += c2[i] * d2[1][i];
                e1.z += c2[i] * d2[2][i];
            }
            t96_diploop1(x2, y2, z2, d1);
            for (int i = 0; i < 14; i++) {
                e2.x += c1[i] * d1[0][i];
                e2.y += c1[i] * d1[1][i];
                e2.z += c1[i] * d1[2][i];
            }
        }
        const double ddx = x2 - x1, ddy = y2 - y1, ddz = z2 - z1;
        const double ss = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        const double qx = x - x1, qy = y - y1, qz = z - z1;
        const double ds = sqrt(qx * qx + qy * qy + qz * qz);
        const double frac = ds / ss;
        b.x = e1.x * (1.0 - frac) + e2.x * frac;
        b.y = e1.y * (1.0 - frac) + e2.y * frac;
        b.z = e1.z * (1.0 - frac) + e2.z * frac;
    }
    const v3 sh = t96_birk1shld(x, y, z);
    v3 out = {b.x + sh.x, b.y + sh.y, b.z + sh.z};
    return out;
}

/* lib/t96/t96.c:1196-1263: Region-2 shield (2x2 Cartesian harmonics). */
static v3 t96_birk2shl(double x, double y, double z) {
    static const double ai[2] = {13.85650567, 14.90554500};
    static const double ak[2] = {10.21914434, 10.09021632};
    static const double amp[2][2] = {
        {-111.6371348 + (double)124.5402702, 110.3735178 + (double)-122.0095905},
        {111.9448247 + (double)-129.1957743, -110.7586562 + (double)126.5649012}};
    v3 b = {0.0, 0.0, 0.0};
    for (int i = 0; i < 2; i++) {
        const double yi = y / ai[i];
        const double cy = cos(yi), sy = sin(yi);
        const double aiai = 1.0 / (ai[i] * ai[i]);
        for (int k = 0; k < 2; k++) {
            const double zk = z / ak[k];
            const double sz = sin(zk), cz = cos(zk);
            const double sq = sqrt(aiai + 1.0 / (ak[k] * ak[k]));
            const double e = exp(x * sq);
            const double c = amp[i][k];
            b.x += (-sq * e * cy * sz) * c;
            b.y += (e / ai[i] * sy * sz) * c;
            b.z += (-e / ak[k] * cy * cz) * c;
        }
    }
    return b;
}

/* lib/t96/t96.c:1596-1627: xi coordinate of the Region-2 sheet (distorted
 * dipole-line invariant). */
static double t96_xksi(double x, double y, double z) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    const double r = sqrt(r2);
    const double xr = x / r, yr = y / r, zr = z / r;
    const double dr = 7.50937, dr2 = dr * dr, r0 = 1.21563;
    double pr = 0.0;
    if (r >= r0) pr = sqrt((r - r0) * (r - r0) + dr2) - dr;
    const double f = x + pr * (0.305662 + -0.383593 * xr + 0.2677733 * xr * xr + -0.097656 * yr * yr + -0.636034 * zr * zr);
    const double g = y + pr * (-0.359862 * yr + 0.424706 * xr * yr);
    const double h = z + pr * (-0.126366 * zr + 0.292578 * xr * zr);
    const double g2 = g * g;
    const double fgh = f * f + g2 + h * h;
    const double sq = sqrt(fgh);
    const double fgh32 = sq * sq * sq;
    const double fg2 = (f * f) + g2;
    if (fg2 < 1.0e-5) return -1.0;
    const double ths = sin(0.3665191 + 0.5 * 0.09599309 * (1.0 - f / sqrt(fg2)));
    return (fg2 / fgh32) - (ths * ths);
}

/* lib/t96/t96.c:1629-1643: smooth step from 0 to 1 across xks0 +- dxksi. */
static double t96_tksi(double xksi, double xks0, double dxksi) {
    const double tdz3 = 2.0 * dxksi * dxksi * dxksi;
    if (xksi - xks0 < -dxksi) return 0.0;
    if (xksi - xks0 >= dxksi) return 1.0;
    if (xksi >= xks0 - dxksi && xksi < xks0) {
        const double u = xksi - xks0 + dxksi;
        const double u3 = u * u * u;
        return 1.5 * u3 / (tdz3 + u3);
    }
    const double u = xksi - xks0 - dxksi;
    const double u3 = u * u * u;
    return 1.0 + 1.5 * u3 / (tdz3 - u3);
}

/* lib/t96/t96.c:1338-1366: conical harmonics m = 1..5. */
static void t96_bconic(double x, double y, double z, double cbx[5], double cby[5], double cbz[5]) {
    const double ro2 = x * x + y * y;
    const double ro = sqrt(ro2);
    const double cf = x / ro, sf = y / ro;
    const double r2 = ro2 + z * z;
    const double r = sqrt(r2);
    const double c = z / r, s = ro / r;
    const double ch = sqrt(0.5 * (1.0 + c)), sh = sqrt(0.5 * (1.0 - c));
    double tnh_prev = 1.0, cnh_prev = 1.0;
    const double tnh = sh / ch, cnh = 1.0 / tnh;
    double cfm_prev = 1.0, sfm_prev = 0.0;
    for (int m = 0; m < 5; m++) {
        const double cfm = cfm_prev * cf - sfm_prev * sf;
        const double sfm = cfm_prev * sf + sfm_prev * cf;
        cfm_prev = cfm;
        sfm_prev = sfm;
        const double tnhm = tnh_prev * tnh, cnhm = cnh_prev * cnh;
        const double bt = (m + 1) * cfm / (r * s) * (tnhm + cnhm);
        const double bf = -0.5 * (m + 1) * sfm / r * (tnh_prev / (ch * ch) - cnh_prev / (sh * sh));
        tnh_prev = tnhm;
        cnh_prev = cnhm;
        cbx[m] = bt * c * cf - bf * sf;
        cby[m] = bt * c * sf + bf * cf;
        cbz[m] = -bt * s;
    }
}

/* lib/t96/t96.c:1368-1386: field of a distributed dipole line (mode 0) or its
 * far form (mode 1). */
static v3 t96_dipdistr(double x, double y, double z, int mode) {
    const double x2 = x * x;
    const double rho2 = x2 + y * y;
    v3 b;
    if (mode == 0) {
        const double r2 = rho2 + z * z;
        const double r3 = r2 * sqrt(r2);
        b.x = z / (rho2 * rho2) * (r2 * (y * y - x2) - rho2 * x2) / r3;
        b.y = -x * y * z / (rho2 * rho2) * (2.0 * r2 + rho2) / r3;
        b.z = x / r3;
        return b;
    }
    b.x = z / (rho2 * rho2) * (y * y - x2);
    b.y = -2.0 * x * y * z / (rho2 * rho2);
    b.z = x / rho2;
    return b;
}

/* lib/t96/t96.c:1414-1470: four loops placed symmetrically about the x-z and
 * x-y planes. */
static v3 t96_loops4(double x, double y, double z, double xc, double yc, double zc,
                     double rl, double theta, double phi) {
    const double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
    double xs, yss, zs, xss, zss, bxs;
    v3 bl;

    xs = (x - xc) * cp + (y - yc) * sp;
    yss = (y - yc) * cp - (x - xc) * sp;
    zs = z - zc;
    xss = xs * ct - zs * st;
    zss = zs * ct + xs * st;
    bl = t96_circle(xss, yss, zss, rl);
    bxs = bl.x * ct + bl.z * st;
    const double bz1 = bl.z * ct - bl.x * st;
    const double bx1 = bxs * cp - bl.y * sp;
    const double by1 = bxs * sp + bl.y * cp;

    xs = (x - xc) * cp - (y + yc) * sp;
    yss = (y + yc) * cp + (x - xc) * sp;
    zs = z - zc;
    xss = xs * ct - zs * st;
    zss = zs * ct + xs * st;
    bl = t96_circle(xss, yss, zss, rl);
    bxs = bl.x * ct + bl.z * st;
    const double bz2 = bl.z * ct - bl.x * st;
    const double bx2 = bxs * cp + bl.y * sp;
    const double by2 = -bxs * sp + bl.y * cp;

    xs = -(x - xc) * cp + (y + yc) * sp;
    yss = -(y + yc) * cp - (x - xc) * sp;
    zs = z + zc;
    xss = xs * ct - zs * st;
    zss = zs * ct + xs * st;
    bl = t96_circle(xss, yss, zss, rl);
    bxs = bl.x * ct + bl.z * st;
    const double bz3 = bl.z * ct - bl.x * st;
    const double bx3 = -bxs * cp - bl.y * sp;
    const double by3 = bxs * sp - bl.y * cp;

    xs = -(x - xc) * cp - (y - yc) * sp;
    yss = -(y - yc) * cp + (x - xc) * sp;
    zs = z + zc;
    xss = xs * ct - zs * st;
    zss = zs * ct + xs * st;
    bl = t96_circle(xss, yss, zss, rl);
    bxs = bl.x * ct + bl.z * st;
    v3 b = {bx1 + bx2 + bx3 + (-bxs * cp + bl.y * sp),
            by1 + by2 + by3 + (-bxs * sp - bl.y * cp),
            bz1 + bz2 + bz3 + (bl.z * ct - bl.x * st)};
    return b;
}

/* lib/t96/t96.c:1323-1336 */
static v3 t96_r2inner(double x, double y, double z) {
    double cbx[5], cby[5], cbz[5];
    t96_bconic(x, y, z, cbx, cby, cbz);
    const v3 d8 = t96_loops4(x, y, z, -8.1902, 6.5239, 5.504, 7.7815, .8573, 3.0986);
    const v3 d6 = t96_dipdistr(x - .0774, y, z, 0);
    const v3 d7 = t96_dipdistr(x + .038, y, z, 1);
    v3 b;
    b.x = 154.185 * cbx[0] + -2.12446 * cbx[1] + 0.601735E-01 * cbx[2] + -.153954E-02 * cbx[3] + .355077E-04 * cbx[4] + 29.9996 * d6.x + 262.886 * d7.x + 99.9132 * d8.x;
    b.y = 154.185 * cby[0] + -2.12446 * cby[1] + 0.601735E-01 * cby[2] + -.153954E-02 * cby[3] + .355077E-04 * cby[4] + 29.9996 * d6.y + 262.886 * d7.y + 99.9132 * d8.y;
    b.z = 154.185 * cbz[0] + -2.12446 * cbz[1] + 0.601735E-01 * cbz[2] + -.153954E-02 * cbz[3] + .355077E-04 * cbz[4] + 29.9996 * d6.z + 262.886 * d7.z + 99.9132 * d8.z;
    return b;
}

/* lib/t96/t96.c:1388-1412 */
static v3 t96_r2outer(double x, double y, double z) {
    const v3 d1 = t96_crosslp(x, y, z, 0.55, 0.694, 0.0031);
    const v3 d2 = t96_crosslp(x, y, z, 1.55, 2.8, 0.1375);
    const v3 d3 = t96_crosslp(x, y, z, -0.7, 0.2, 0.9625);
    const v3 d4 = t96_circle(x + 2.994, y, z, 2.925);
    const v3 d5 = t96_loops4(x, y, z, -1.775, 4.3, -0.275, 2.7, 0.4312, 1.55);
    v3 b;
    b.x = -34.105 * d1.x + -2.00019 * d2.x + 628.639 * d3.x + 73.4847 * d4.x + 12.5162 * d5.x;
    b.y = -34.105 * d1.y + -2.00019 * d2.y + 628.639 * d3.y + 73.4847 * d4.y + 12.5162 * d5.y;
    b.z = -34.105 * d1.z + -2.00019 * d2.z + 628.639 * d3.z + 73.4847 * d4.z + 12.5162 * d5.z;
    return b;
}

/* lib/t96/t96.c:1472-1594: Region-2 current sheet.  The three t-factor triples
 * differ per component; the sheet is expanded in exp(a ct^2)-type latitude terms
 * times cos/sin(m phi).  Coefficient tables are laid out [term][phi-harmonic][t]. */
static void t96_tfac(double xks, double dd1, double dd2, double dd3, double t[4]) {
    /* t[0] unused (the constant term); t[1..3] as lib/t96/t96.c:1474-1487 */
    const double xks2 = xks * xks;
    const double p72 = dd2 * dd2;
    const double s7 = sqrt(xks2 + p72);
    t[0] = 1.0;
    t[1] = xks / sqrt(xks2 + dd1 * dd1);
    t[2] = (p72 * dd2) / (s7 * s7 * s7);
    t[3] = xks / pow(sqrt(xks2 + dd3 * dd3), 5) * 3.493856 * pow(dd3, 4);
}

static double t96_quad(const double c[4], const double t[4]) {
    return c[0] + c[1] * t[1] + c[2] * t[2] + c[3] * t[3];
}

static v3 t96_r2sheet(double x, double y, double z) {
    static const double cx[5][4][4] = {
        {{8.07190, -7.39582, -7.62341, 0.684671}, {-13.5672, 11.6681, 13.1154, -0.890217},
         {7.78726, -5.38346, -8.08738, 0.609385}, {-2.70410, 3.53741, 3.15549, -1.11069}},
        {{-8.47555, 0.278122, 2.73514, 4.55625}, {13.1134, 1.15848, -3.52648, -8.24698},
         {-6.85710, -2.81369, 2.03795, 4.64383}, {2.49309, -1.22041, -1.67432, -0.422526}},
        {{-5.39796, 7.10326, 5.53730, -13.1918}, {4.67853, -7.60329, -2.53066, 7.76338},
         {5.60165, 5.34816, -4.56441, 7.05976}, {-2.62723, -0.529078, 1.42019, -2.93919}},
        {{55.6338, -1.55181, 39.8311, -80.6561}, {-46.9655, 32.8925, -6.32296, 19.7841},
         {124.731, 10.4347, -30.7581, 102.680}, {-47.4037, -3.31278, 9.37141, -50.0268}},
        {{-533.319, 110.426, 1000.20, -1051.40}, {1619.48, 589.855, -1462.73, 1087.10},
         {-1994.73, -1654.12, 1263.33, -260.210}, {1424.84, 1255.71, -956.733, 219.946}}};
    static const double cy[5][4][4] = {
        {{-9.08427, 10.6777, 10.3288, -0.969987}, {6.45257, -8.42508, -7.97464, 1.41996},
         {-1.92490, 3.93575, 2.83283, -1.48621}, {0.244033, -0.757941, -0.386557, 0.344566}},
        {{9.56674, -2.5365, -3.32916, -5.86712}, {-6.19625, 1.83879, 2.52772, 4.34417},
         {1.87268, -2.13213, -1.69134, -0.176379}, {-0.261359, 0.566419, 0.3138, -0.134699}},
        {{-3.83086, -8.4154, 4.77005, -9.31479}, {37.5715, 19.3992, -17.9582, 36.4604},
         {-14.9993, -3.1442, 6.17409, -15.5519}, {2.28621, -0.00891549, -0.462912, 2.47314}},
        {{41.7555, 208.614, -45.7861, -77.8687}, {239.357, -67.9226, 66.8743, 238.534},
         {-112.136, 16.2069, -40.4706, -134.328}, {21.56, -0.201725, 2.21, 32.5855}},
        {{-108.217, -1005.98, 585.753, 323.668}, {-817.056, 235.750, -560.965, -576.892},
         {684.193, 85.0275, 168.394, 477.776}, {-289.253, -123.216, 75.6501, -178.605}}};
    static const double cz[5][4][4] = {
        {{1167.61, -917.782, -1253.2, -274.128}, {-1538.75, 1257.62, 1745.07, 113.479},
         {393.326, -426.858, -641.1, 190.833}, {-29.9435, -1.04881, 117.125, -25.7663}},
        {{-1168.16, 910.247, 1239.31, 289.515}, {1540.56, -1248.29, -1727.61, -131.785},
         {-394.577, 426.163, 637.422, -187.965}, {30.0348, 0.221898, -116.68, 26.0291}},
        {{12.6804, 4.84091, 1.18166, -2.75946}, {-17.9822, -6.80357, -1.47134, 3.02266},
         {4.79648, 0.665255, -0.256229, -0.0857282}, {-0.588997, 0.0634812, 0.164303, -0.15285}},
        {{22.2524, -22.4376, -3.85595, 6.07625}, {-105.959, -41.6698, 0.378615, 1.55958},
         {44.3981, 18.8521, 3.19466, 5.89142}, {-8.63227, -2.36418, -1.027, -2.31515}},
        {{1035.38, 2040.66, -131.881, -744.533}, {-3274.93, -4845.61, 482.438, 1567.43},
         {1354.02, 2040.47, -151.653, -845.012}, {-111.723, -265.343, -26.1171, 216.632}}};

    const double xks = t96_xksi(x, y, z);
    double tx[4], ty[4], tz[4];
    t96_tfac(xks, 0.483750e-01, 0.396953e-01, 0.579023e-01, tx);
    t96_tfac(xks, 0.478750e-01, 0.363750e-01, 0.567500e-01, ty);
    t96_tfac(xks, 0.355625e-01, 0.318750e-01, 0.538750e-01, tz);
    const double rho2 = x * x + y * y;
    const double r = sqrt(rho2 + z * z);
    const double rho = sqrt(rho2);
    const double c1p = x / rho, s1p = y / rho;
    const double s2p = 2.0 * s1p * c1p;
    const double c2p = c1p * c1p - s1p * s1p;
    const double s3p = s2p * c1p + c2p * s1p;
    const double c3p = c2p * c1p - s2p * s1p;
    const double s4p = s3p * c1p + c3p * s1p;
    const double ct = z / r;
    const double ct2 = ct * ct;
    const double ct2m1 = ct2 - 1.0;
    const double stmp = -2. * 2.718281828459;

    double sx[5] = {sqrt(stmp * -19.0969) * ct * exp(-19.0969 * ct2),
                    sqrt(stmp * -9.28828) * ct * exp(-9.28828 * ct2),
                    sqrt(stmp * -0.129687) * ct * exp(-0.129687 * ct2),
                    ct * exp(5.58594 * ct2m1), ct * exp(22.5055 * ct2m1)};
    double sy[5] = {sqrt(stmp * -13.6750) * ct * exp(-13.6750 * ct2),
                    sqrt(stmp * -6.70625) * ct * exp(-6.70625 * ct2),
                    ct * exp(2.31875 * ct2m1), ct * exp(11.4062 * ct2m1),
                    ct * exp(20.4562 * ct2m1)};
    double sz[5] = {exp(-16.7125 * ct2), exp(-16.4625 * ct2), exp(-0.1625 * ct2),
                    exp(5.1 * (ct2 - 1.0)), exp(23.7125 * (ct2 - 1.0))};
    const double cphi[4] = {1.0, c1p, c2p, c3p};
    const double sphi[4] = {s1p, s2p, s3p, s4p};

    v3 b = {0.0, 0.0, 0.0};
    for (int k = 0; k < 5; k++) {
        /* x and z: constant term first, then cos(m phi) m = 1..3 */
        const double qx = ((t96_quad(cx[k][0], tx) + c1p * t96_quad(cx[k][1], tx)) +
                           c2p * t96_quad(cx[k][2], tx)) + c3p * t96_quad(cx[k][3], tx);
        const double qy = ((sphi[0] * t96_quad(cy[k][0], ty) + sphi[1] * t96_quad(cy[k][1], ty)) +
                           sphi[2] * t96_quad(cy[k][2], ty)) + sphi[3] * t96_quad(cy[k][3], ty);
        const double qz = ((t96_quad(cz[k][0], tz) + c1p * t96_quad(cz[k][1], tz)) +
                           c2p * t96_quad(cz[k][2], tz)) + c3p * t96_quad(cz[k][3], tz);
        if (k == 0) {
            b.x = sx[k] * qx;
            b.y = sy[k] * qy;
            b.z = sz[k] * qz;
        } else {
            b.x = b.x + sx[k] * qx;
            b.y = b.y + sy[k] * qy;
            b.z = b.z + sz[k] * qz;
        }
    }
    (void)cphi;
    return b;
}

/* lib/t96/t96.c:1265-1321: Region-2 field, switching between the inner, sheet and
 * outer representations on xi with cubic blends of half-width 0.015. */
static v3 t96_r2birk(double x, double y, double z) {
    const double del = 0.03, del1 = 0.015;
    const double xks = t96_xksi(x, y, z);
    v3 b = {0.0, 0.0, 0.0};
    if (xks < -(del + del1)) {
        b = t96_r2outer(x, y, z);
        b.x *= -0.02;
        b.y *= -0.02;
        b.z *= -0.02;
    }
    if (xks >= -(del + del1) && xks < -del + del1) {
        const v3 b1 = t96_r2outer(x, y, z);
        const v3 b2 = t96_r2sheet(x, y, z);
        const double f2 = -0.02 * t96_tksi(xks, -del, del1);
        const double f1 = -0.02 - f2;
        b.x = b1.x * f1 + b2.x * f2;
        b.y = b1.y * f1 + b2.y * f2;
        b.z = b1.z * f1 + b2.z * f2;
    }
    if (xks >= -del + del1 && xks < del - del1) {
        b = t96_r2sheet(x, y, z);
        b.x *= -0.02;
        b.y *= -0.02;
        b.z *= -0.02;
    }
    if (xks >= del - del1 && xks < del + del1) {
        const v3 b1 = t96_r2inner(x, y, z);
        const v3 b2 = t96_r2sheet(x, y, z);
        const double f1 = -0.02 * t96_tksi(xks, del, del1);
        const double f2 = -0.02 - f1;
        b.x = b1.x * f1 + b2.x * f2;
        b.y = b1.y * f1 + b2.y * f2;
        b.z = b1.z * f1 + b2.z * f2;
    }
    if (xks >= del + del1) {
        b = t96_r2inner(x, y, z);
        b.x = -b.x * 0.02;
        b.y = -b.y * 0.02;
        b.z = -b.z * 0.02;
    }
    return b;
}

/* lib/t96/t96.c:4-111: T96 at zero tilt.  PARMOD = {Pdyn, Dst, By, Bz, ...}. */
void o_t96_gsm(const double *parmod, double x, double y, double z, double out[3]) {
    const double pdyn = parmod[0];
    const double depr = 0.8 * parmod[1] - 13.0 * sqrt(pdyn);
    const double byimf = parmod[2], bzimf = parmod[3];
    const double bt = sqrt(byimf * byimf + bzimf * bzimf);
    double theta = 0.0;
    if (byimf != 0.0 || bzimf != 0.0) {
        theta = atan2(byimf, bzimf);
        if (theta <= 0.0) theta += 6.2831853;
    }
    const double ct = cos(theta), st = sin(theta);
    const double eps = 718.5 * sqrt(pdyn) * bt * sin(theta / 2.);
    const double ys = y * ct - z * st, zs = z * ct + y * st;
    const double ysd = ys / 10.0;
    const double factimf = exp(x / 20.0 - ysd * ysd);
    const double oimfy = 0.7850 * byimf * factimf, oimfz = 0.7850 * bzimf * factimf;
    const double xappa = pow(pdyn / 2.0, 0.14);
    const double xappa3 = xappa * xappa * xappa;
    const double am = 70.0 / xappa;
    const double rho2 = y * y + z * z;
    const double asq = am * am;
    double xmxm = am + x - (5.48 / xappa);
    if (xmxm < 0.0) xmxm = 0.0;
    const double axx0 = xmxm * xmxm;
    const double aro = asq + rho2;
    const double aro_axx0 = aro + axx0;
    const double sigma = sqrt((aro + axx0 + sqrt(aro_axx0 * aro_axx0 - 4.0 * asq * axx0)) / (2.0 * asq));

    if (sigma >= 1.08 + 0.005) {
        const v3 q = t96_dipole(x, y, z);
        out[0] = -q.x;
        out[1] = oimfy - q.y;
        out[2] = oimfz - q.z;
        return;
    }

    const double facteps = eps / 3630.7 - 1.0;
    const double factpd = sqrt(pdyn / 2.0) - 1.0;
    const double rcampl = -1.162 * depr;
    const double tampl2 = 22.344 + 18.50 * factpd + 2.602 * facteps;
    const double tampl3 = 6.903 + 5.287 * factpd;
    const double b1ampl = 0.5790 + 0.4462 * facteps;
    const double b2ampl = 20.0 * b1ampl;
    const double xx = x * xappa, yy = y * xappa, zz = z * xappa;

    const v3 cf = t96_cylharm(xx, yy, zz);
    v3 rc, t2, t3;
    t96_tailrc(xx, yy, zz, &rc, &t2, &t3);
    const v3 r1 = t96_birk1tot(xx, yy, zz);
    v3 r2 = t96_birk2shl(xx, yy, zz);
    const v3 r2b = t96_r2birk(xx, yy, zz);
    r2.x += r2b.x;
    r2.y += r2b.y;
    r2.z += r2b.z;
    const v3 im = t96_intercon(xx, ys * xappa, zs * xappa);
    const double rimfy = im.y * ct + im.z * st;
    const double rimfz = im.z * ct - im.y * st;
    const double rimfampl = 0.7850 * bt;

    const double fx = cf.x * xappa3 + rcampl * rc.x + tampl2 * t2.x + tampl3 * t3.x + b1ampl * r1.x + b2ampl * r2.x + rimfampl * im.x;
    const double fy = cf.y * xappa3 + rcampl * rc.y + tampl2 * t2.y + tampl3 * t3.y + b1ampl * r1.y + b2ampl * r2.y + rimfampl * rimfy;
    const double fz = cf.z * xappa3 + rcampl * rc.z + tampl2 * t2.z + tampl3 * t3.z + b1ampl * r1.z + b2ampl * r2.z + rimfampl * rimfz;

    if (sigma < 1.08 - 0.005) {
        out[0] = fx;
        out[1] = fy;
        out[2] = fz;
        return;
    }
    const double fint = 0.5 * (1.0 - (sigma - 1.08) / 0.005);
    const double fext = 0.5 * (1.0 + (sigma - 1.08) / 0.005);
    const v3 q = t96_dipole(x, y, z);
    out[0] = (fx + q.x) * fint - q.x;
    out[1] = (fy + q.y) * fint + oimfy * fext - q.y;
    out[2] = (fz + q.z) * fint + oimfz * fext - q.z;
}
