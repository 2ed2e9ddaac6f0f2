// ts05_host.hpp -- per-epoch TS05 constants on the host (FP64, libm), in the
// reference's arithmetic so the FP64 device path sees the reference's values:
// lib/ts05/ts05.c:19-110 (drivers and amplitudes), lib/ts05/deformed.c:182-208 (warped
// tail coefficients A[2l] + A[2l+1] * shift and the 1/P, 1/R scales),
// lib/ts05/birk_tot.c:379-403 and lib/ts05/full_rc.c:697-721 (shield coefficient sums
// a0 + s a1 + a2 + s a3 and (s + 1)^3, and the shields' scales), lib/ts05/birk_tot.c:322-328
// (tan of the cone band edges).
#pragma once
#include <math.h>
#include "ts05_dev.hpp"

namespace cor {

inline void ts05_build_params(const double *pm, TS05Params<double> &P) {
    static const double A[69] = {TS05_A};
    static const double W1[60] = {TS05_WARP1}, W2[60] = {TS05_WARP2};
    static const double SH11[78] = {TS05_SH11}, SH21[78] = {TS05_SH21};
    static const double CSY[78] = {TS05_CSY}, CPR[78] = {TS05_CPR};
    static const double A11[31] = {TS05_A11}, A21[31] = {TS05_A21};
    const double pdyn = pm[0];
    const double dst = pm[1] * 0.8 - 13. * sqrt(pdyn);
    const double w[6] = {pm[4], pm[5], pm[6], pm[7], pm[8], pm[9]};
    P.byimf = pm[2];
    P.bzimf = pm[3];
    P.xappa = pow((pdyn / 2.), A[22]);
    P.xappa3 = P.xappa * P.xappa * P.xappa;
    P.am = 34.586 / P.xappa;
    P.x0 = 3.4397 / P.xappa;
    P.oimfy = pm[2] * A[19];
    P.oimfz = pm[3] * A[19];
    double dstt = -20.;
    if (dst < dstt) dstt = dst;
    double znam = pow(fabs(dstt), 0.37);
    P.dxshift1 = A[23] - A[24] / znam;
    P.dxshift2 = A[25] - A[26] / znam;
    P.d0 = A[35] * exp(-w[0] / A[36]) + A[68];
    znam = fabs(dst);
    if (dst >= -20.) znam = 20.;
    P.xkappa1 = A[31] * pow(znam / 20., A[32]);
    P.xkappa2 = A[33] * pow(znam / 20., A[34]);
    P.sc_sy = A[27] * pow((20. / znam), A[28]) * P.xappa;
    P.sc_pr = A[29] * pow((20. / znam), A[30]) * P.xappa;
    const double dlp1 = pow((pdyn / 2.), A[20]);
    const double dlp2 = pow((pdyn / 2.), A[21]);
    P.tamp1 = A[1] + A[2] * dlp1 + A[3] * (A[38] * w[0]) / sqrt(w[0] * w[0] + A[38] * A[38]);
    P.tamp2 = A[5] + A[6] * dlp2 + A[7] * (A[39] * w[1]) / sqrt(w[1] * w[1] + A[39] * A[39]);
    P.a_src = A[9] + A[10] * (A[40] * w[2]) / sqrt(w[2] * w[2] + A[40] * A[40]) + A[11] * dst;
    P.a_prc = A[12] + A[13] * (A[41] * w[3]) / sqrt(w[3] * w[3] + A[41] * A[41]) + A[14] * dst;
    P.a_r11 = A[15] + A[16] * (A[42] * w[4]) / sqrt(w[4] * w[4] + A[42] * A[42]);
    P.a_r21 = A[17] + A[18] * (A[43] * w[5]) / sqrt(w[5] * w[5] + A[43] * A[43]);
    P.cphi = cos(A[37]);
    P.sphi = sin(A[37]);
    // cone band edges: theta0 = A[30] of each cone set, dtheta 0.06 (Region 1), 0.09 (Region 2)
    P.tgp[0] = tan((A11[30] + 0.06) * 0.5);
    P.tgm[0] = tan((A11[30] - 0.06) * 0.5);
    P.tgp[1] = tan((A21[30] + 0.09) * 0.5);
    P.tgm[1] = tan((A21[30] - 0.09) * 0.5);
    const double *wt[2] = {W1, W2};
    const double shift[2] = {P.dxshift1, P.dxshift2};
    double *wc[2] = {P.w1, P.w2};
    for (int m = 0; m < 2; m++) {
        for (int i = 0; i < 5; i++) P.wrp[m][i] = 1. / wt[m][50 + i];
        for (int k = 0; k < 5; k++) P.wrp[m][5 + k] = 1. / wt[m][55 + k];
        for (int i = 0; i < 5; i++)
            for (int k = 0; k < 5; k++) {
                const int l = 5 * i + k;
                wc[m][l] = wt[m][2 * l] + wt[m][2 * l + 1] * shift[m];
                const double rp = P.wrp[m][i], rr = P.wrp[m][5 + k];
                P.wsq[m][l] = sqrt(rp * rp + rr * rr);
            }
    }
    const double *st[4] = {SH11, SH21, CSY, CPR};
    const double xsc[4] = {P.xkappa1 - 1.1, P.xkappa2 - 1.0, P.sc_sy - 1., P.sc_pr - 1.};
    double *sc[4] = {P.r11, P.r21, P.src, P.prc};
    for (int t = 0; t < 4; t++)
        for (int l = 0; l < 9; l++) {
            const double *a = st[t] + 4 * l;
            sc[t][l] = a[0] + xsc[t] * a[1] + a[2] + xsc[t] * a[3];
        }
    // the shields' scales for the mixed path's harmonic sums: 1/P_j, 1/R_k and
    // sqrt(1/P_j^2 + 1/R_k^2) (the last as the FP64 path forms it per term)
    for (int t = 0; t < 4; t++)
        for (int j = 0; j < 3; j++) {
            const double p = st[t][72 + j];
            P.shinv[t][j] = 1. / p;
            P.shinv[t][3 + j] = 1. / st[t][75 + j];
            const double pinv2 = 1. / (p * p);
            for (int k = 0; k < 3; k++) {
                const double r = st[t][75 + k];
                P.shsq[t][3 * j + k] = sqrt(pinv2 + 1. / (r * r));
            }
        }
    const double x1 = xsc[2] + 1., x2 = xsc[3] + 1.;
    P.fac_sy = x1 * x1 * x1;
    P.fac_pr = x2 * x2 * x2;
}

inline void ts05_params_to_float(const TS05Params<double> &d, TS05Params<float> &f) {
    const double *s = reinterpret_cast<const double *>(&d);
    float *o = reinterpret_cast<float *>(&f);
    static_assert(sizeof(TS05Params<double>) == 2 * sizeof(TS05Params<float>), "plain arrays of T");
    for (size_t k = 0; k < sizeof(TS05Params<float>) / sizeof(float); k++) o[k] = (float)s[k];
}

}  // namespace cor
