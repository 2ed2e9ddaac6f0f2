/*
 * o_core.c -- oracle: coordinate transforms, IGRF epoch setup and synthesis, the
 * field aggregate and the cutoff tracer, restated from the reference in plain C.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Arithmetic is written in the same
 * operation order as the reference so that, compiled without FMA contraction on
 * the same libm, results are bit-identical; the code structure is this file's own.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

/* ------------------------------------------------------------------ transforms */

/* src/utils.c:28-41 */
void o_sph_to_car(const double sph[3], double car[3]) {
    const double rs = sph[0] * sin(sph[1]);
    car[0] = rs * cos(sph[2]);
    car[1] = rs * sin(sph[2]);
    car[2] = sph[0] * cos(sph[1]);
}

/* src/utils.c:43-73: phi wrapped into [0, 2pi) with the truncated 2pi constant;
 * on the polar axis theta is 0 or pi and phi is 0. */
void o_car_to_sph(const double car[3], double sph[3]) {
    const double rho2 = car[0] * car[0] + car[1] * car[1];
    sph[0] = sqrt(rho2 + car[2] * car[2]);
    if (rho2 != 0.) {
        double phi = atan2(car[1], car[0]);
        sph[1] = atan2(sqrt(rho2), car[2]);
        if (phi < 0.) phi += O_2PI;
        sph[2] = phi;
        return;
    }
    sph[1] = car[2] < 0. ? O_PI : 0.;
    sph[2] = 0.;
}

/* src/utils.c:75-87: GSM = A . GEO with A's rows A1, A2, A3. */
void o_geo_to_gsm(const double geo[3], double gsm[3], const o_params *p) {
    const double x = geo[0], y = geo[1], z = geo[2];
    gsm[0] = p->A1[0] * x + p->A1[1] * y + p->A1[2] * z;
    gsm[1] = p->A2[0] * x + p->A2[1] * y + p->A2[2] * z;
    gsm[2] = p->A3[0] * x + p->A3[1] * y + p->A3[2] * z;
}

/* src/utils.c:89-101: GEO = A^T . GSM. */
void o_gsm_to_geo(const double gsm[3], double geo[3], const o_params *p) {
    const double x = gsm[0], y = gsm[1], z = gsm[2];
    geo[0] = p->A1[0] * x + p->A2[0] * y + p->A3[0] * z;
    geo[1] = p->A1[1] * x + p->A2[1] * y + p->A3[1] * z;
    geo[2] = p->A1[2] * x + p->A2[2] * y + p->A3[2] * z;
}

/* ------------------------------------------------------- IGRF-14 epoch setup */

#define IGRF14_STEP 5u
#define IGRF14_MIN 1900u
#define IGRF14_MAX 2030u

/* lib/internal/internal.c:38-62 -- Russell (1971) low-precision sun position and
 * Greenwich sidereal time.  out = {GST, SLONG, SRASN, SDEC} in radians. */
void o_sun(int iy, int iday, int ih, int min, int is, double out[4]) {
    const double fday = (ih * 3600 + min * 60 + is) / 86400.;
    const double dj = 365 * (iy - 1900) + (iy - 1901) / 4 + iday - 0.5 + fday;
    const double t = dj / 36525.;
    const double gst = fmod(279.690983 + 0.9856473354 * dj + 360. * fday + 180., 360.) / O_TO_DEG;
    const double vl = fmod(279.696678 + 0.9856473354 * dj, 360.);
    const double g = fmod(358.475845 + 0.985600267 * dj, 360.) / O_TO_DEG;
    double slong = (vl + (1.91946 - 0.004789 * t) * sin(g) + 0.020094 * sin(2. * g)) / O_TO_DEG;
    if (slong > 6.2831853) {
        slong = slong - 6.2831853;
    } else if (slong < 0) {
        slong = slong + 6.2831853;
    }
    const double obliq = (23.45229 - 0.0130125 * t) / O_TO_DEG;
    const double slp = slong - 9.924E-5;
    const double sob = sin(obliq);
    const double sind = sob * sin(slp);
    const double cosd = sqrt(1. - sind * sind);
    const double sc = sind / cosd;
    out[0] = gst;
    out[1] = slong;
    out[2] = 3.141592654 - atan2(cos(obliq) / sob * sc, -cos(slp) / cosd);
    out[3] = atan(sc);
}

/* lib/internal/internal.c:64-87 (+ interpolate :12-22, extrapolate :24-36) for the
 * IGRF-14 descriptor (igrf_14_coefs.h:88-95).  The unsigned arithmetic of the
 * descriptor fields is kept. */
static int igrf14_coefs_at(int year, int iday, double *G, double *H) {
    const uint32_t nrows = O_IGRF14_NROWS;
    const double step = (double)IGRF14_STEP;
    if (IGRF14_MAX - (uint32_t)year <= IGRF14_STEP) {
        /* last main-field row plus secular variation, SV only for index <= 45 */
        const double *g0 = o_igrf14_rows[nrows - 4], *h0 = o_igrf14_rows[nrows - 3];
        const double *gs = o_igrf14_rows[nrows - 2], *hs = o_igrf14_rows[nrows - 1];
        const double dt = year + (iday - 1) / 365.25 - ((int)IGRF14_MAX - step);
        for (int k = 0; k < O_NCOEF; k++) {
            G[k] = g0[k];
            H[k] = h0[k];
            if (k <= 45) {
                G[k] += gs[k] * dt;
                H[k] += hs[k] * dt;
            }
        }
        return 0;
    }
    if ((uint32_t)year >= IGRF14_MAX) return -1;
    uint32_t back;
    if ((uint32_t)year <= IGRF14_MIN) {
        back = IGRF14_MAX / IGRF14_STEP - (IGRF14_MIN + 1) / IGRF14_STEP;
    } else {
        back = IGRF14_MAX / IGRF14_STEP - (uint32_t)year / IGRF14_STEP;
    }
    const int hrow = (int)(nrows - 2 * back - 1);
    const int edge = (int)(IGRF14_MIN + ((hrow - 1) / 2) * IGRF14_STEP);
    const double *gx = o_igrf14_rows[hrow - 1], *hx = o_igrf14_rows[hrow];
    const double *gy = o_igrf14_rows[hrow + 1], *hy = o_igrf14_rows[hrow + 2];
    const double f2 = (year + (iday - 1) / 365.25 - edge) / step;
    const double f1 = 1. - f2;
    for (int k = 0; k < O_NCOEF; k++) {
        G[k] = gx[k] * f1 + gy[k] * f2;
        H[k] = hx[k] * f1 + hy[k] * f2;
    }
    return 0;
}

/* lib/internal/internal.c:89-180: recursion table, coefficients for the epoch,
 * Schmidt -> recursion scaling, dipole axis, sun, and the GEO->GSM rows. */
int o_igrf14_recalc(int iy, int mes, int ide, int id, int ih, int min, int is, o_params *p) {
    (void)mes;
    (void)ide;
    if ((uint32_t)iy > IGRF14_MAX || (uint32_t)iy < IGRF14_MIN) return -1;

    for (int n = 1; n < 15; n++) {
        double den = 2 * n - 1;
        den = den * (den - 2);
        for (int m = 1; m <= n; m++) {
            p->REC[n * (n - 1) / 2 + m - 1] = ((n - m) * (n + m - 2)) / den;
        }
    }
    if (igrf14_coefs_at(iy, id, p->G, p->H) != 0) return -1;

    double s = 1.;
    for (int n = 2; n <= 14; n++) {
        const int base = n * (n - 1) / 2;
        s = s * (2 * n - 3) / (n - 1);
        p->G[base] *= s;
        p->H[base] *= s;
        double f = s;
        for (int m = 2; m <= n; m++) {
            const double two = (m == 2) ? 2. : 1.;
            f = f * sqrt(two * (n - m + 1) / (n + m - 2));
            p->G[base + m - 1] *= f;
            p->H[base + m - 1] *= f;
        }
    }

    const double g10 = -p->G[1], g11 = p->G[2], h11 = p->H[2];
    const double sq = g11 * g11 + h11 * h11;
    const double sqq = sqrt(sq);
    const double sqr = sqrt(g10 * g10 + sq);

    double sun[4] = {0., 0., 0., 0.};
    if (iy > 1900 && iy < 2100) o_sun(iy, id, ih, min, is, sun);
    const double gst = sun[0], srasn = sun[2], sdec = sun[3];

    const double ct0 = g10 / sqr;
    const double cl0 = -g11 / sqq;
    const double sl0 = -h11 / sqq;
    const double st0 = sqq / sqr;
    const double stcl = st0 * cl0, stsl = st0 * sl0;
    const double cgst = cos(gst), sgst = sin(gst);
    const double dip1 = stcl * cgst - stsl * sgst;
    const double dip2 = stcl * sgst + stsl * cgst;
    const double dip3 = ct0;

    const double csdec = cos(sdec);
    const double s1 = cos(srasn) * csdec, s2 = sin(srasn) * csdec, s3 = sin(sdec);
    double y1 = dip2 * s3 - dip3 * s2;
    double y2 = dip3 * s1 - dip1 * s3;
    double y3 = dip1 * s2 - dip2 * s1;
    const double yn = sqrt(y1 * y1 + y2 * y2 + y3 * y3);
    y1 = y1 / yn;
    y2 = y2 / yn;
    y3 = y3 / yn;
    const double z1 = s2 * y3 - s3 * y2;
    const double z2 = s3 * y1 - s1 * y3;
    const double z3 = s1 * y2 - s2 * y1;

    p->A1[0] = s1 * cgst + s2 * sgst;
    p->A1[1] = -s1 * sgst + s2 * cgst;
    p->A1[2] = s3;
    p->A2[0] = y1 * cgst + y2 * sgst;
    p->A2[1] = -y1 * sgst + y2 * cgst;
    p->A2[2] = y3;
    p->A3[0] = z1 * cgst + z2 * sgst;
    p->A3[1] = -z1 * sgst + z2 * cgst;
    p->A3[2] = z3;
    return 0;
}

/* --------------------------------------------------------- IGRF synthesis */

/* lib/internal/internal.c:182-275.  GEOPACK-style recursion: for each order m the
 * associated Legendre values (and theta derivatives) are carried up in degree with
 * the REC table, cos/sin(m phi) are advanced by rotation, and the expansion is cut
 * at degree NM = min(13, 3 + 30/((int)R + 2)).  Input GEO Cartesian [Re], output
 * GEO Cartesian field [nT]. */
void o_igrf_geo(double x, double y, double z, const double *G, const double *H,
                const double *REC, double b[3]) {
    double ra[14], rb[14];
    const double rho2 = x * x + y * y;
    const double r = sqrt(rho2 + z * z);
    const double c = z / r;
    const double rho = sqrt(rho2);
    const double s = rho / r;
    const int polar = s < 1.E-5;
    const double cf = polar ? 1. : x / rho;
    const double sf = polar ? 0. : y / rho;

    const double inv_r = 1. / r;
    double pw = inv_r;
    int nm = 3 + 30 / ((int)r + 2);
    if (nm > 13) nm = 13;
    const int kmax = nm + 1;
    for (int n = 1; n <= kmax; n++) {
        pw = pw * inv_r;
        ra[n - 1] = pw;
        rb[n - 1] = pw * n;
    }

    double pmm = 1., dmm = 0.;          /* sectoral P and dP carried between orders */
    double cm = 0., sm = 1.;            /* rotated so that (cm, sm) = (sin m phi, cos m phi) */
    double br = 0., bt = 0., bf = 0.;
    for (int m = 1; m <= kmax; m++) {
        if (m != 1) {
            const double w = cm;
            cm = w * cf + sm * sf;
            sm = sm * cf - w * sf;
        }
        double q = pmm, dq = dmm, q2 = 0., dq2 = 0., bphi = 0.;
        for (int n = m; n <= kmax; n++) {
            const double an = ra[n - 1];
            const int idx = n * (n - 1) / 2 + m - 1;
            const double g = G[idx], h = H[idx];
            const double w = g * sm + h * cm;
            br = br + rb[n - 1] * w * q;
            bt = bt - an * w * dq;
            if (m != 1) {
                const double qq = polar ? dq : q;
                bphi = bphi + an * (g * cm - h * sm) * qq;
            }
            const double rec = REC[idx];
            const double dnext = c * dq - s * q - rec * dq2;
            const double pnext = c * q - rec * q2;
            dq2 = dq;
            q2 = q;
            dq = dnext;
            q = pnext;
        }
        dmm = s * dmm + c * pmm;
        pmm = s * pmm;
        if (m != 1) {
            bphi = bphi * (m - 1);
            bf = bf + bphi;
        }
    }

    double bphi_tot;
    if (polar) {
        if (c < 0.) bf = -bf;
        bphi_tot = bf;
    } else {
        bphi_tot = bf / s;
    }
    const double he = br * s + bt * c;
    b[0] = he * cf - bphi_tot * sf;
    b[1] = he * sf + bphi_tot * cf;
    b[2] = br * c - bt * s;
}

/* ----------------------------------------------------------- field aggregate */

static void add3(double acc[3], const double v[3]) {
    acc[0] += v[0];
    acc[1] += v[1];
    acc[2] += v[2];
}

/* src/models.c:194-207 with the model adapters of src/models.c:146-192: position
 * scaled to Re, every model adds its GSM field (SH models round-trip through GEO),
 * and the GSM sum is returned in GEO. */
void o_calculate_field(const double gsm_pos_m[3], double geo_field_nT[3],
                       const o_models *m, const o_params *p) {
    const double gsm_re[3] = {gsm_pos_m[0] / O_RE, gsm_pos_m[1] / O_RE, gsm_pos_m[2] / O_RE};
    double total[3] = {0., 0., 0.};
    double geo[3], bgeo[3], bgsm[3];
    if (m->internal == O_INT_IGRF) {
        o_gsm_to_geo(gsm_re, geo, p);
        o_igrf_geo(geo[0], geo[1], geo[2], p->G, p->H, p->REC, bgeo);
        o_geo_to_gsm(bgeo, bgsm, p);
        add3(total, bgsm);
    }
    if (m->external == O_EXT_T96) {
        o_t96_gsm(p->PARMOD, gsm_re[0], gsm_re[1], gsm_re[2], bgsm);
        add3(total, bgsm);
    } else if (m->external == O_EXT_TS05) {
        o_ts05_gsm(p->PARMOD, gsm_re[0], gsm_re[1], gsm_re[2], bgsm);
        add3(total, bgsm);
    }
    if (m->crustal == O_CRU_CHAOS) {
        o_gsm_to_geo(gsm_re, geo, p);
        o_chaos_geo(geo[0], geo[1], geo[2], m->chaos_nm_max, m->chaos_g, m->chaos_h, bgeo);
        o_geo_to_gsm(bgeo, bgsm, p);
        add3(total, bgsm);
    }
    o_gsm_to_geo(total, geo_field_nT, p);
}

void o_field_batch(const o_models *m, const o_params *p, const double *gsm_pos_m,
                   double *geo_field_nT, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        o_calculate_field(gsm_pos_m + 3 * i, geo_field_nT + 3 * i, m, p);
    }
}

/* ------------------------------------------------------------------ tracer */

/* src/trajectory.c:275-356.  Six-stage Kutta-Nystrom step with the field frozen:
 * the velocity stages integrate dv/dt = d (v x B); the position stages are the same
 * tableau applied to f(x) = v + x, i.e. each stage is h * (v + combination of
 * earlier position stages).  Division by the tableau denominators is kept. */
void o_rk_step(double h, double d, const double fb[3], double vel[3], double pos[3]) {
    static const double den[6] = {1., 3., 25., 4., 81., 75.};
    static const double w[6][5] = {
        {0., 0., 0., 0., 0.},
        {1., 0., 0., 0., 0.},
        {4., 6., 0., 0., 0.},
        {1., -12., 15., 0., 0.},
        {6., 90., -50., 8., 0.},
        {6., 36., 10., 8., 0.},
    };
    const double bx = fb[0], by = fb[1], bz = fb[2];
    const double vx = vel[0], vy = vel[1], vz = vel[2];
    double kx[6], ky[6], kz[6];   /* velocity stages */
    double px[6], py[6], pz[6];   /* position stages */

    kx[0] = h * (d * (vy * bz - vz * by));
    ky[0] = h * (d * (vz * bx - vx * bz));
    kz[0] = h * (d * (vx * by - vy * bx));
    for (int s = 1; s < 6; s++) {
        /* combination of the earlier stages, summed in tableau order */
        double cx, cy, cz;
        if (s == 1) {
            cx = kx[0] / 3.;
            cy = ky[0] / 3.;
            cz = kz[0] / 3.;
        } else {
            cx = w[s][0] * kx[0];
            cy = w[s][0] * ky[0];
            cz = w[s][0] * kz[0];
            for (int j = 1; j < s; j++) {
                cx = cx + w[s][j] * kx[j];
                cy = cy + w[s][j] * ky[j];
                cz = cz + w[s][j] * kz[j];
            }
            cx = cx / den[s];
            cy = cy / den[s];
            cz = cz / den[s];
        }
        const double ux = vx + cx, uy = vy + cy, uz = vz + cz;
        kx[s] = h * (d * (uy * bz - uz * by));
        ky[s] = h * (d * (uz * bx - ux * bz));
        kz[s] = h * (d * (ux * by - uy * bx));
    }

    px[0] = h * vx;
    py[0] = h * vy;
    pz[0] = h * vz;
    for (int s = 1; s < 6; s++) {
        double cx, cy, cz;
        if (s == 1) {
            cx = px[0] / 3.;
            cy = py[0] / 3.;
            cz = pz[0] / 3.;
        } else {
            cx = w[s][0] * px[0];
            cy = w[s][0] * py[0];
            cz = w[s][0] * pz[0];
            for (int j = 1; j < s; j++) {
                cx = cx + w[s][j] * px[j];
                cy = cy + w[s][j] * py[j];
                cz = cz + w[s][j] * pz[j];
            }
            cx = cx / den[s];
            cy = cy / den[s];
            cz = cz / den[s];
        }
        px[s] = h * (vx + cx);
        py[s] = h * (vy + cy);
        pz[s] = h * (vz + cz);
    }

    vel[0] += (23. * kx[0] + 125. * kx[2] - 81. * kx[4] + 125. * kx[5]) / 192.;
    vel[1] += (23. * ky[0] + 125. * ky[2] - 81. * ky[4] + 125. * ky[5]) / 192.;
    vel[2] += (23. * kz[0] + 125. * kz[2] - 81. * kz[4] + 125. * kz[5]) / 192.;
    pos[0] += (23. * px[0] + 125. * px[2] - 81. * px[4] + 125. * px[5]) / 192.;
    pos[1] += (23. * py[0] + 125. * py[2] - 81. * py[4] + 125. * py[5]) / 192.;
    pos[2] += (23. * pz[0] + 125. * pz[2] - 81. * pz[4] + 125. * pz[5]) / 192.;
}

/* src/trajectory.c:7 */
int32_t o_scan_count(const o_scan *s) {
    return (int32_t)round((s->rk - s->rig) / s->del) + 1;
}

static double norm3(const double v[3]) {
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* src/trajectory.c:13-153 for one rigidity index i.  Same step control: the field is
 * frozen over a step and its rejected attempts; an attempt is rejected (nk0 doubled,
 * state restored) when the velocity turns by >= tu_angle and nk0 <= 50000, or when
 * it leaves the escape surface outside the 0.002 band (no nk0 cap).  The speed used
 * by the next attempt is the one the last attempt produced, rejected or not. */
void o_trace_one(const o_scan *sc, int32_t i, const o_models *m, const o_params *p,
                 uint64_t step_limit, double tu_angle, double mass, o_traj *out) {
    memset(out, 0, sizeof(*out));
    int nk0 = sc->nk1;
    const double rig = sc->rig + i * sc->del;
    double alength = 0., time = 0.;
    const double a = rig * 1.0e+09 * sc->zn * O_Q / (mass * O_C * O_C);

    const double sph_vel[3] = {O_C * sqrt(1. - 1. / (1. + a * a)), (90. - sc->the1) * O_TO_RAD,
                               sc->fi1 * O_TO_RAD};
    double sph_pos[3] = {sc->r0 * O_RE, (90. - sc->the0) * O_TO_RAD, sc->fi0 * O_TO_RAD};
    double pos[3], vel[3], gsm[3], fld[3];
    o_sph_to_car(sph_pos, pos);
    o_geo_to_gsm(pos, gsm, p);
    o_calculate_field(gsm, fld, m, p);
    double b = 1.E-09 * norm3(fld);
    o_sph_to_car(sph_vel, vel);
    double v = sph_vel[0];
    unsigned int step = 0;
    uint32_t attempts = 0;

    for (;;) {
        step++;
        const double vc = v / O_C;
        if (vc > 1.) {
            out->status = O_FORBIDDEN_REENTRY;
            break;
        }
        const double hm = mass / sqrt(1. - vc * vc);
        const double period = (2. * O_PI * hm) / (fabs(sc->zn) * O_Q * b);
        const double h = period / nk0;
        const double d = (1.E-09 * sc->zn * O_Q) / hm;
        const double pos0[3] = {pos[0], pos[1], pos[2]};
        const double vel0[3] = {vel[0], vel[1], vel[2]};
        const double vprev = v;

        o_rk_step(h, d, fld, vel, pos);
        attempts++;
        v = norm3(vel);

        const double gc = (vel[0] * vel0[0] + vel[1] * vel0[1] + vel[2] * vel0[2]) / (v * vprev);
        if (gc < 1.) {
            const double gu = atan2(sqrt((1. - gc) * (1. + gc)), gc);
            if (gu >= tu_angle && nk0 <= 50000) {
                nk0 *= 2;
                memcpy(pos, pos0, sizeof pos);
                memcpy(vel, vel0, sizeof vel);
                step--;
                continue;
            }
        }

        o_car_to_sph(pos, sph_pos);
        o_geo_to_gsm(pos, gsm, p);
        const double pax = (gsm[0] / O_RE + 25.3) / 36.08;
        const double paz = (gsm[1] * gsm[1] + gsm[2] * gsm[2]) / (O_RE * O_RE);
        const double r = sph_pos[0] / O_RE;
        const double paus = pax * pax + paz / 459.6736;

        if (r >= 25. || paus > 1.) {
            if ((r >= 25. && fabs(r - 25.) < 0.002) || (r <= 25. && paus > 1. && fabs(paus - 1.) < 0.002)) {
                /* src/trajectory.c:91-112: exit point and asymptotic direction */
                const double st = sin(sph_pos[1]), ct = cos(sph_pos[1]);
                const double sf = sin(sph_pos[2]), cf = cos(sph_pos[2]);
                const double be = vel[0] * cf + vel[1] * sf;
                const double vr = be * st + vel[2] * ct;
                const double vth = be * ct - vel[2] * st;
                const double vph = -vel[0] * sf + vel[1] * cf;
                const double va = vth * ct + vr * st;
                const double dm = sqrt(vth * vth + va * va);
                double asf = (sph_pos[2] + atan2(vph, va)) / O_TO_RAD;
                if (asf > 360.) asf -= 360.;
                out->status = O_ALLOWED;
                out->vc = v / O_C;
                out->r = r;
                out->td = 90. - (sph_pos[1] / O_TO_RAD);
                out->fei = sph_pos[2] / O_TO_RAD;
                out->ast = atan2(-vth * st + vr * ct, dm) / O_TO_RAD;
                out->asf = asf;
                out->time = time;
                out->length_m = alength;
                break;
            }
            nk0 *= 2;
            memcpy(pos, pos0, sizeof pos);
            memcpy(vel, vel0, sizeof vel);
            step--;
            continue;
        }
        if (r < 1.) {
            out->status = O_FORBIDDEN_REENTRY;
            break;
        }
        if (step > step_limit || sph_pos[2] > 31.4) {
            out->status = O_FORBIDDEN_LIMIT;
            break;
        }

        o_calculate_field(gsm, fld, m, p);
        b = 1.E-09 * norm3(fld);
        time = time + period / nk0;
        const double dx = pos[0] - pos0[0], dy = pos[1] - pos0[1], dz = pos[2] - pos0[2];
        alength += sqrt(dx * dx + dy * dy + dz * dz);
        nk0 = sc->nk1;
    }
    out->steps = step;
    out->attempts = attempts;
}

void o_trace_scan(const o_scan *s, const o_models *m, const o_params *p,
                  uint64_t step_limit, double tu_angle, double mass, int32_t threads,
                  o_traj *out) {
    const int32_t n = o_scan_count(s);
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int32_t i = 0; i < n; i++) {
        o_trace_one(s, i, m, p, step_limit, tu_angle, mass, out + i);
    }
}

/* src/trajectory.c:49-54,114-115,133-145,156-160: forbidden count, largest forbidden
 * rigidity per kind, smallest allowed rigidity (starting at rk), then
 * cut = {Rl, Ru, Rc} in the order the reference prints them (P(S), P(C), P(M)). */
void o_penumbra(const o_scan *s, const o_traj *t, int32_t n, double cut[3]) {
    int nza = 0;
    double rmx1 = 0., rmx2 = 0., rmi = s->rk;
    for (int32_t i = 0; i < n; i++) {
        const double rig = s->rig + i * s->del;
        if (t[i].status == O_ALLOWED) {
            rmi = fmin(rmi, rig);
        } else if (t[i].status == O_FORBIDDEN_REENTRY) {
            nza += 1;
            rmx1 = fmax(rmx1, rig);
        } else {
            nza += 1;
            rmx2 = fmax(rmx2, rig);
        }
    }
    double rmx = fmax(rmx1, rmx2);
    const double zan = nza - (rmi - s->rig) / s->del;
    const double rms = rmi + zan * s->del;
    rmx = rmx + s->del;
    cut[0] = rmi;
    cut[1] = rmx;
    cut[2] = rms;
}
