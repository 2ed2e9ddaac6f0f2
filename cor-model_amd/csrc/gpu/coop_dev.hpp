// coop_dev.hpp -- cooperative mode of the mixed tracer: a group of lanes (8, 16, 32 or
// the whole wavefront) per trajectory.
//
// Cutoff scans have heavy-tailed step counts: a few trapped trajectories (lowest
// rigidities) run 1e5..1e6 steps, hundreds of times the median.  In the per-lane mode a
// wavefront's step takes as long with one live lane as with 64, so once the queue is
// drained the run time is the latency of those few trajectories.  In cooperative mode
// the lanes of a group share one trajectory:
//   * the trajectory state is held by every lane of the group (uniform values,
//     identical arithmetic);
//   * the field model is split into families of terms of identical form -- Cartesian
//     harmonics, dipoles, current loops, conical harmonics, conical sheets, cylindrical
//     harmonics, the tail-87 sheets, the Region-2 sheet terms, IGRF orders -- one term
//     per lane, so a family is one SIMD pass, or a few for small groups;
//   * each lane accumulates its terms; one cross-lane sum per step (DPP, read back from
//     one lane) gives every lane the same bits;
//   * the RK attempts and step control are the per-lane mode's code, run redundantly.
// The region switches of T96 (magnetopause blend, Region-1 bands, Region-2 xi bands) are
// uniform per trajectory, so a lone trajectory's family passes never diverge on them
// (groups of a wavefront at different points may).
//
// Arithmetic is FP32 like the per-lane mixed path.  The interconnection harmonics (a fit
// whose terms cancel by four orders of magnitude) are summed in a form without the
// cancellation (see the lane-table builder).
#pragma once
#include <utility>
#include "dev_types.hpp"
#include "igrf_dev.hpp"
#include "t96_dev.hpp"

namespace cor {

// ---------------------------------------------------------------- lane-table builder

// Per-epoch lane tables from the T96 coefficient tables (lib/t96/t96.c function-level
// citations at each family) and the epoch's amplitudes.  One thread; runs once per plan.
template <class F>
__device__ inline void co_circ(CoopTabT<F> &T, int &l, const double R[9], double cx, double cy, double cz,
                               double rl, double coef, int group) {
    for (int k = 0; k < 9; k++) T.circ[l][k] = (F)R[k];
    T.circ[l][9] = (F)cx;
    T.circ[l][10] = (F)cy;
    T.circ[l][11] = (F)cz;
    T.circ[l][12] = (F)rl;
    T.circ[l][13] = (F)coef;
    T.circ[l][14] = (F)group;
    T.circ[l][15] = F(0);
    l++;
}

// crosslp (lib/t96/t96.c:883-898): two loops tilted by +-alpha about the x axis
template <class F>
__device__ inline void co_crosslp(CoopTabT<F> &T, int &l, double xc, double rl, double cal, double sal,
                                  double coef, int group) {
    const double R1[9] = {1, 0, 0, 0, cal, -sal, 0, sal, cal};
    const double R2[9] = {1, 0, 0, 0, cal, sal, 0, -sal, cal};
    co_circ(T, l, R1, xc, 0, 0, rl, coef, group);
    co_circ(T, l, R2, xc, 0, 0, rl, coef, group);
}

// loops4 (lib/t96/t96.c:1414-1470): four mirrored loops, each rotated about z by the
// azimuth and then tilted about y; local = Tilt * Rz * (p - centre)
template <class F>
__device__ inline void co_loops4(CoopTabT<F> &T, int &l, double xc, double yc, double zc, double rl, double st,
                                 double ct, double sp, double cp, double coef, int group) {
    const double M[4][9] = {{cp, sp, 0, -sp, cp, 0, 0, 0, 1},
                            {cp, -sp, 0, sp, cp, 0, 0, 0, 1},
                            {-cp, sp, 0, -sp, -cp, 0, 0, 0, 1},
                            {-cp, -sp, 0, sp, -cp, 0, 0, 0, 1}};
    const double C[4][3] = {{xc, yc, zc}, {xc, -yc, zc}, {xc, -yc, -zc}, {xc, yc, -zc}};
    const double Tt[9] = {ct, 0, -st, 0, 1, 0, st, 0, ct};
    for (int j = 0; j < 4; j++) {
        double R[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                R[3 * r + c] = Tt[3 * r] * M[j][c] + Tt[3 * r + 1] * M[j][3 + c] + Tt[3 * r + 2] * M[j][6 + c];
        co_circ(T, l, R, C[j][0], C[j][1], C[j][2], rl, coef, group);
    }
}

template <class F>
__device__ inline void build_coop_tab(const T96Params<double> &P, CoopTabT<F> &T) {
    using namespace t96;
    F *z = &T.harm[0][0];
    for (size_t k = 0; k < sizeof(CoopTabT<F>) / sizeof(F); k++) z[k] = F(0.);
    const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

    // Cartesian harmonics: tailrc_96 shield sets (lib/t96/t96.c:501-617), birk1shld
    // (:1030-1194), birk2shl (:1196-1263)
    int l = 0;
    const double set_amp[3] = {P.rcampl, P.tampl2, P.tampl3};
    for (int set = 0; set < 3; set++) {
        const double *tab = k_shl_d + 15 * set;
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++, l++) {
                const double p = tab[9 + i], r = tab[12 + k];
                T.harm[l][0] = (F)(1.0 / p);
                T.harm[l][1] = (F)(1.0 / r);
                T.harm[l][2] = (F)sqrt(1.0 / (p * p) + 1.0 / (r * r));
                T.harm[l][3] = (F)(tab[3 * i + k] * set_amp[set]);
            }
    }
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++, l++) {
            const double rp = 1.0 / k_b1s_py_d[i], rk = 1.0 / k_b1s_pz_d[k];
            T.harm[l][0] = (F)rp;
            T.harm[l][1] = (F)rk;
            T.harm[l][2] = (F)sqrt(rp * rp + rk * rk);
            T.harm[l][3] = (F)(k_b1s_amp_d[4 * i + k] * P.b1ampl);
        }
    {
        const double ai[2] = {13.85650567, 14.90554500}, ak[2] = {10.21914434, 10.09021632};
        const double amp[4] = {-111.6371348 + 124.5402702, 110.3735178 + -122.0095905,
                               111.9448247 + -129.1957743, -110.7586562 + 126.5649012};
        for (int i = 0; i < 2; i++)
            for (int k = 0; k < 2; k++, l++) {
                T.harm[l][0] = (F)(1.0 / ai[i]);
                T.harm[l][1] = (F)(1.0 / ak[k]);
                T.harm[l][2] = (F)sqrt(1.0 / (ai[i] * ai[i]) + 1.0 / (ak[k] * ak[k]));
                T.harm[l][3] = (F)(amp[2 * i + k] * P.b2ampl);
            }
    }
    // interconnection field (lib/t96/t96.c:236-271).  Per y harmonic i, the k = 0 term
    // as it is; the k = 1, 2 pair (coefficients ~1e7 on sin(z r_k), r_k ~ 1e-3, whose sum
    // cancels by four orders of magnitude) as the series
    //   sum_k c_k s_k e^{x s_k} sin(z r_k)
    //     = e^{x s_1} (sum_j (-1)^j z^{2j+1} A_{2j+1} / (2j+1)!  +  c_2 s_2 delta sin(z r_2)),
    //   A_n = c_1 s_1 r_1^n + c_2 s_2 r_2^n,  delta = e^{x (s_2 - s_1)} - 1,
    // and the same with B_n = c_1 r_1^n + c_2 r_2^n for the y component (sine) and the z
    // component (cosine, odd n).  The A_n, B_n are formed here in FP64, so the device sums
    // terms that no longer cancel; |z r_k| < 0.1 makes the n <= 7 series exact in FP32.
    {
        const double a[9] = {-8.411078731, 5932254.951, -9073284.93, -11.68794634, 6027598.824,
                             -9218378.368, -6.508798398, -11824.42793, 18015.66212};
        const double rp[3] = {0.12503844191332988, 0.071597395204476152, 0.011080976889832743};
        const double rr[3] = {0.059675527522452816, 0.00098459526668325296, 0.00064371058922777988};
        for (int i = 0; i < 3; i++) {
            F *t = T.icon[i];
            const double c0 = a[3 * i] * P.rimfampl, c1 = a[3 * i + 1] * P.rimfampl, c2 = a[3 * i + 2] * P.rimfampl;
            const double p2 = rp[i] * rp[i];
            const double s0 = sqrt(p2 + rr[0] * rr[0]), s1 = sqrt(p2 + rr[1] * rr[1]), s2 = sqrt(p2 + rr[2] * rr[2]);
            double A[4], B[4];
            for (int j = 0; j < 4; j++) {
                const double q1 = pow(rr[1], 2 * j + 1), q2 = pow(rr[2], 2 * j + 1);
                A[j] = c1 * s1 * q1 + c2 * s2 * q2;
                B[j] = c1 * q1 + c2 * q2;
            }
            const double v[21] = {rp[i], rr[0], s0, c0, s1, (rr[2] * rr[2] - rr[1] * rr[1]) / (s1 + s2), rr[2],
                                  c2 * s2, c2, c2 * rr[2], A[0], A[1] / 6., A[2] / 120., A[3] / 5040.,
                                  B[0], B[1] / 6., B[2] / 120., B[3] / 5040., B[1] / 2., B[2] / 24., B[3] / 720.};
            for (int k = 0; k < 21; k++) t[k] = (F)v[k];
        }
    }
    // conical sheets: ringcurr_96 (lib/t96/t96.c:322-381), taildisk_t96 (:383-443)
    {
        const double fr[2] = {569.895366, -1603.386993}, br[2] = {2.722188, 3.766875};
        const double fd[4] = {-745796.7338, 1176470.141, -444610.529, -57508.01028};
        const double bd[4] = {7.9250000, 8.0850000, 8.4712500, 27.89500};
        for (int i = 0; i < 2; i++) {
            T.cone[i][0] = (F)(fr[i] * P.rcampl);
            T.cone[i][1] = (F)br[i];
        }
        for (int i = 0; i < 4; i++) {
            T.cone[2 + i][0] = (F)(fd[i] * P.tampl2);
            T.cone[2 + i][1] = (F)bd[i];
            T.cone[2 + i][2] = F(4.5);
            T.cone[2 + i][3] = F(10.);
        }
    }
    // cylindrical harmonics (lib/t96/t96.c:113-163)
    {
        const double a[12] = {0.24777, -27.003, -0.46815, 7.0637, -1.5918, -0.90317E-01,
                              57.522, 13.757, 2.0100, 10.458, 4.5798, 2.1695};
        for (int i = 0; i < 6; i++) {
            T.cyl[i][0] = (F)(a[i] * P.xappa3);
            T.cyl[i][1] = (F)a[i + 6];
            T.cyl[i][2] = i < 3 ? F(0.) : F(1.);
        }
    }
    // tail87 (lib/t96/t96.c:445-499): the sheet at z and its images at z -+ 40
    T.t87[0][0] = F(0.);
    T.t87[0][1] = (F)P.tampl3;
    T.t87[1][0] = -F(40.);
    T.t87[1][1] = (F)(-0.5 * P.tampl3);
    T.t87[2][0] = F(40.);
    T.t87[2][1] = (F)(-0.5 * P.tampl3);
    // dipoles: diploop1 (lib/t96/t96.c:809-852) then condip1 (:920-1028), each a
    // point dipole of moment M at position D: field 30574 (3 (M.r) r - r^2 M) / r^5
    {
        int s = 0;
        auto dip = [&](double dx, double dy, double dz, double mx, double my, double mz) {
            T.dip[s][0] = (F)dx;
            T.dip[s][1] = (F)dy;
            T.dip[s][2] = (F)dz;
            T.dip[s][4] = (F)mx;
            T.dip[s][5] = (F)my;
            T.dip[s][6] = (F)mz;
            s++;
        };
        for (int i = 0; i < 12; i++) {
            const double xd = k_dl_xx_d[i] * 1.12541, yd = k_dl_yy_d[i] * 0.945719;
            const double c = k_dl_c1_d[i] * P.b1ampl;
            dip(xd, yd, 0., 0., 0., c);
            if (k_dl_yy_d[i] != 0.) dip(xd, -yd, 0., 0., 0., c);
        }
        for (int i = 0; i < 9; i++) {
            const double xd = k_cd_xd_d[i], yd = k_cd_yd_d[i], zd = k_cd_zz_d[i];
            const int ix = (i + 1) * 3 + 2;
            const double cx = k_cd_c2_d[ix] * P.b1ampl, cy = k_cd_c2_d[ix + 1] * P.b1ampl,
                         cz = k_cd_c2_d[ix + 2] * P.b1ampl;
            dip(xd, yd, zd, cx, cy, cz);
            dip(xd, -yd, zd, cx, -cy, cz);
            dip(xd, yd, -zd, -cx, -cy, cz);
            dip(xd, -yd, -zd, -cx, cy, cz);
        }
        for (int i = 1; i <= 5; i++) {
            const double zd = k_cd_zz_d[i + 8];
            const int ix = 30 + i * 2;
            const double cx = k_cd_c2_d[ix] * P.b1ampl, cz = k_cd_c2_d[ix + 1] * P.b1ampl;
            dip(0., 0., zd, cx, 0., cz);
            dip(0., 0., -zd, -cx, 0., cz);
        }
    }
    // current loops: diploop1's crossed pair and circle (group 0), r2outer (group 1,
    // lib/t96/t96.c:1388-1412), r2inner's loops4 (group 2, :1323-1366)
    {
        int c = 0;
        co_crosslp(T, c, 2.28397, 1.86106, 0.5327834519503802, 0.8462516134861054, k_dl_c1_d[12] * P.b1ampl, 0);
        co_circ(T, c, I3, -5.60831, 0., 0., 7.83281, k_dl_c1_d[13] * P.b1ampl, 0);
        co_crosslp(T, c, 0.55, 0.694, 0.999995195003848, 0.003099995034835719, -34.105 * P.b2ampl, 1);
        co_crosslp(T, c, 1.55, 2.8, 0.9905617592124779, 0.13706714116038543, -2.00019 * P.b2ampl, 1);
        co_crosslp(T, c, -0.7, 0.2, 0.5714702170359919, 0.8206228068003206, 628.639 * P.b2ampl, 1);
        co_circ(T, c, I3, -2.994, 0., 0., 2.925, 73.4847 * P.b2ampl, 1);
        co_loops4(T, c, -1.775, 4.3, -0.275, 2.7, 0.4179612609200968, 0.9084648503767676, 0.999783764189357,
                  0.020794827803092428, 12.5162 * P.b2ampl, 1);
        co_loops4(T, c, -8.1902, 6.5239, 5.504, 7.7815, 0.7560782215370907, 0.6544812624646408,
                  0.04297941043769308, -0.9990759582126018, 99.9132 * P.b2ampl, 2);
    }
    // conical harmonics: condip1's (group 0, at x + 0.16) and r2inner's bconic (group 1)
    {
        const double cc[5] = {154.185, -2.12446, 0.601735E-01, -.153954E-02, .355077E-04};
        for (int m = 1; m <= 5; m++) {
            T.conic[m - 1][0] = (F)m;
            T.conic[m - 1][1] = (F)(k_cd_c2_d[m - 1] * P.b1ampl);
            T.conic[m - 1][2] = F(0.16);
            T.conic[m - 1][3] = F(0.);
            T.conic[4 + m][0] = (F)m;
            T.conic[4 + m][1] = (F)(cc[m - 1] * P.b2ampl);
            T.conic[4 + m][2] = F(0.);
            T.conic[4 + m][3] = F(1.);
        }
    }
    // Region-2 sheet (lib/t96/t96.c:1472-1594): per component and latitude term k the
    // factor alpha (ct)^[comp != z] exp(beta ct^2 + gamma), the t-factor constants and
    // the 16 coefficients of the four phi harmonics
    {
        const double alpha[3][5] = {{10.189284199579348, 7.106076658978449, 0.839673526422457, 1., 1.},
                                    {8.622355131189718, 6.038125124921339, 1., 1., 1.},
                                    {1., 1., 1., 1., 1.}};
        const double beta[3][5] = {{-19.0969, -9.28828, -0.129687, 5.58594, 22.5055},
                                   {-13.6750, -6.70625, 2.31875, 11.4062, 20.4562},
                                   {-16.7125, -16.4625, -0.1625, 5.1, 23.7125}};
        const double gamma[3][5] = {{0., 0., 0., -5.58594, -22.5055},
                                    {0., 0., -2.31875, -11.4062, -20.4562},
                                    {0., 0., 0., -5.1, -23.7125}};
        const double tf[3][4] = {{0.483750e-01, 0.396953e-01, 0.579023e-01, 1.124043867626711e-05},
                                 {0.478750e-01, 0.363750e-01, 0.567500e-01, 1.0372022816406252e-05},
                                 {0.355625e-01, 0.318750e-01, 0.538750e-01, 8.42459695336914e-06}};
        for (int comp = 0; comp < 3; comp++)
            for (int k = 0; k < 5; k++) {
                F *s = T.sheet[comp * 5 + k];
                s[0] = (F)alpha[comp][k];
                s[1] = (F)beta[comp][k];
                s[2] = (F)gamma[comp][k];
                s[3] = (F)comp;
                for (int j = 0; j < 4; j++) s[4 + j] = (F)tf[comp][j];
                for (int j = 0; j < 16; j++) s[8 + j] = (F)(k_r2s_d[comp * 80 + 16 * k + j] * P.b2ampl);
            }
    }
}

// ---------------------------------------------------------------- wavefront helpers

// Sum of each of N values over the 64 lanes with DPP row operations (pairs, quads,
// half rows and rows by permutation, then rows 1/3 and 2/3 by row broadcasts); the
// total ends in lane 63 and is read from there, so every lane gets the same bits.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
// the same for a double: both halves move with the same lane pattern
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWS, 0xF, false);
    return v + __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float readlane_any(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane_any(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int N, class F>
__device__ __forceinline__ void wave_sum(F *p) {
#pragma unroll
    for (int k = 0; k < N; k++) p[k] = dpp_add<0xB1, 0xF>(p[k]);    // quad_perm [1,0,3,2]
#pragma unroll
    for (int k = 0; k < N; k++) p[k] = dpp_add<0x4E, 0xF>(p[k]);    // quad_perm [2,3,0,1]
#pragma unroll
    for (int k = 0; k < N; k++) p[k] = dpp_add<0x141, 0xF>(p[k]);   // row_half_mirror
#pragma unroll
    for (int k = 0; k < N; k++) p[k] = dpp_add<0x140, 0xF>(p[k]);   // row_mirror
#pragma unroll
    for (int k = 0; k < N; k++) p[k] = dpp_add<0x142, 0xA>(p[k]);   // row_bcast:15 -> rows 1, 3
#pragma unroll
    for (int k = 0; k < N; k++) p[k] = dpp_add<0x143, 0xC>(p[k]);   // row_bcast:31 -> rows 2, 3
#pragma unroll
    for (int k = 0; k < N; k++) p[k] = readlane_any(p[k], 63);
}

// Sum of each of N values over a group of G lanes (G = 4: the two DPP steps within a
// quad; G = 8: three steps within a half row; G = 16: the four row steps; G = 32: the
// row sums plus the partner row's, read by a lane permute; every lane of the group ends
// with the same bits, since each step adds two values that the partner lane adds in the
// other order).  G = 64: wave_sum.
template <int N, int G, class F>
__device__ __forceinline__ void group_sum(F *p) {
    static_assert(G == 4 || G == 8 || G == 16 || G == 32 || G == 64, "groups of 4, 8, 16, 32 or 64 lanes");
    if constexpr (G == 64) {
        wave_sum<N>(p);
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) p[k] = dpp_add<0xB1, 0xF>(p[k]);    // quad_perm [1,0,3,2]
#pragma unroll
        for (int k = 0; k < N; k++) p[k] = dpp_add<0x4E, 0xF>(p[k]);    // quad_perm [2,3,0,1]
        if constexpr (G >= 8) {
#pragma unroll
            for (int k = 0; k < N; k++) p[k] = dpp_add<0x141, 0xF>(p[k]);   // row_half_mirror
        }
        if constexpr (G >= 16) {
#pragma unroll
            for (int k = 0; k < N; k++) p[k] = dpp_add<0x140, 0xF>(p[k]);   // row_mirror
        }
        if constexpr (G == 32) {
#pragma unroll
            for (int k = 0; k < N; k++) p[k] = p[k] + __shfl_xor(p[k], 16);
        }
    }
}

// ---------------------------------------------------------------- canonical order
//
// The field of a trajectory is a sum of terms of several families.  So that a
// trajectory's bits do not depend on how many lanes evaluate it, every sum has one
// canonical order for all group widths: term t of a family lives in slot brev6(t) of
// 64 slots (the 6-bit reversal of t), a slot's value is its terms added in family order,
// and the field is the balanced pairwise sum of the 64 slots in slot order.  A group of G
// lanes gives lane gl the contiguous slots gl*64/G .. (gl+1)*64/G - 1, i.e. the terms
// t = brev(gl) + G*p (p = 0 .. 64/G - 1, slot brev(p) of the lane's block): the lane
// sums its block pairwise and the cross-lane steps (balanced in lane order) finish the
// same tree.  Floating-point addition being commutative, each pairwise step gives the
// same bits in either operand order, so every lane of a group ends with the total.

constexpr int ilog2c(int g) { return g <= 1 ? 0 : 1 + ilog2c(g / 2); }
constexpr int brevc(int v, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}
// brev(gl) over log2(G) bits (the lane's first term)
template <int G>
__device__ __forceinline__ int lane_residue(int gl) {
    constexpr int L = ilog2c(G);
    if constexpr (L == 0) return 0;
    else return (int)(__builtin_bitreverse32((unsigned)gl) >> (32 - L));
}

// Pairwise sum of a lane's slots pushed in slot order J (a template argument, so every
// index below is a compile-time constant and the levels stay in registers): level k
// holds the pending left half of size 2^k.
template <class V, int NP>
struct SlotTree {
    V lv[ilog2c(NP) + 1];
    template <int J, int K = 0>
    __device__ __forceinline__ void push(V v) {
        if constexpr ((J >> K) & 1) {
            push<J, K + 1>(lv[K] + v);
        } else {
            lv[K] = v;
        }
    }
    __device__ __forceinline__ V total() const { return lv[ilog2c(NP)]; }
};

// f(integral_constant<int, J>) for J = 0 .. N-1, in order
template <int... J, class Fn>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, J...>, Fn &&f) {
    (f(std::integral_constant<int, J>{}), ...);
}
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn &&f) {
    static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

template <class F>
struct Acc3 {
    F x, y, z;
    __device__ __forceinline__ Acc3 operator+(const Acc3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
};

// four consecutive table entries (a 16-byte-aligned row start)
template <class F>
struct Row4 {
    F x, y, z, w;
};
__device__ __forceinline__ Row4<float> ld4(const float *p) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    return {v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ Row4<double> ld4(const double *p) {
    const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
    return {a.x, a.y, b.x, b.y};
}
// 1/sqrt: the hardware estimate in FP32 (the conical harmonics' mixed-precision form),
// correctly rounded root and division in FP64
__device__ __forceinline__ float co_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double co_rsq(double x) { return 1.0 / sqrt(x); }

// One term of the harmonic form shared by the shield sets, birk1shld, birk2shl and the
// interconnection field: c e^{x sq} (-sq cos(y rp) sin(z rk), rp sin(y rp) sin(z rk),
// -rk cos(y rp) cos(z rk)).
template <class T>
__device__ __forceinline__ v3<T> co_harm(T x, T y, T z, T rp, T rk, T sq, T c) {
    T sy, cy, sz, cz;
    Dm<T>::sincos(y * rp, &sy, &cy);
    Dm<T>::sincos(z * rk, &sz, &cz);
    const T ce = c * Dm<T>::exp(x * sq);
    return mk3<T>(-sq * ce * cy * sz, rp * ce * sy * sz, -rk * ce * cy * cz);
}

// ---------------------------------------------------------------- T96, cooperative

// This lane's share of the sum FX of lib/t96/t96.c:91-93 plus the uniform terms that
// finish the model (lib/t96/t96.c:4-111).
template <class F>
struct T96Share {
    v3<F> fx;           // this lane's terms of FX
    v3<F> q;            // dipole field (for the magnetopause blend)
    F sigma, oimfy, oimfz;
};

// The point and region state a T96 slot's terms need (uniform per trajectory).
template <class F>
struct T96Geo {
    F xx, yy, zz;                   // the point scaled by xappa
    F X_i, Y_i, Z_i;                // the IMF-rotated point (interconnection field)
    F lx, ly, lz, cxp, cyp, czp;    // Region-1 loop-basis and conical-basis points
    F wl, wc, wo, ws, wi;           // blend weights: Region-1 loop, cone; Region-2 outer, sheet, inner
    F xks;                          // Region-2 xi
    bool need_loop, need_cone;
};

// One Region-1 dipole (lib/t96/t96.c:809-852, 920-1028): 30574 w (3 (M.r) r - r^2 M) / r^5.
template <class F>
__device__ __forceinline__ Acc3<F> coop_dipole(const CoopTabT<F> &T, int s, F w, F qx0, F qy0, F qz0) {
    const Row4<F> dp = ld4(&T.dip[s][0]);
    const Row4<F> mm = ld4(&T.dip[s][4]);
    const F X = qx0 - dp.x, Y = qy0 - dp.y, Z = qz0 - dp.z;
    const F r2 = X * X + Y * Y + Z * Z;
    const F qd = w * F(30574.0) / (r2 * r2 * Dm<F>::sqrt(r2));
    const F md = F(3.0) * (mm.x * X + mm.y * Y + mm.z * Z);
    return Acc3<F>{qd * (md * X - r2 * mm.x), qd * (md * Y - r2 * mm.y), qd * (md * Z - r2 * mm.z)};
}

// The terms of slot t (all families, in family order).  `base` is the smallest t of the
// pass (a compile-time constant), so families whose terms all lie below it drop out.
template <class F>
__device__ __forceinline__ Acc3<F> coop_t96_slot(const T96Params<F> &P, const CoopTabT<F> &T, const T96Geo<F> gm, const int base,
                                            const int t) {
    using namespace t96;
    F px = 0, py = 0, pz = 0;
    // family 1: Cartesian harmonics (47)
    if (base < CO_NHARM && t < CO_NHARM) {
        const Row4<F> h = ld4(T.harm[t]);
        const v3<F> b = co_harm<F>(gm.xx, gm.yy, gm.zz, h.x, h.y, h.z, h.w);
        px += b.x;
        py += b.y;
        pz += b.z;
    }
    // family 2: interconnection harmonics (3, one y harmonic each) at the
    // IMF-rotated point, rotated back (lib/t96/t96.c:66-72)
    if (base < CO_NICON && t < CO_NICON) {
        const F *c = T.icon[t];
        F sy, cy, sz, cz, s2, c2;
        Dm<F>::sincos(gm.Y_i * c[0], &sy, &cy);
        Dm<F>::sincos(gm.Z_i * c[1], &sz, &cz);
        Dm<F>::sincos(gm.Z_i * c[6], &s2, &c2);
        const F e0 = c[3] * Dm<F>::exp(gm.X_i * c[2]);
        const F e1 = Dm<F>::exp(gm.X_i * c[4]);
        const F u = gm.X_i * c[5];
        const F del = u * (F(1.0) + u * (F(0.5) + u * F(1.0 / 6.0)));
        const F z2 = gm.Z_i * gm.Z_i;
        const F sx = gm.Z_i * (c[10] - z2 * (c[11] - z2 * (c[12] - z2 * c[13]))) + c[7] * del * s2;
        const F sy2 = gm.Z_i * (c[14] - z2 * (c[15] - z2 * (c[16] - z2 * c[17]))) + c[8] * del * s2;
        const F sz2 = c[14] - z2 * (c[18] - z2 * (c[19] - z2 * c[20])) + c[9] * del * c2;
        const F bx = -cy * (c[2] * e0 * sz + e1 * sx);
        const F by = c[0] * sy * (e0 * sz + e1 * sy2);
        const F bz = -cy * (c[1] * e0 * cz + e1 * sz2);
        px += bx;
        py += by * P.ct + bz * P.st;
        pz += bz * P.ct - by * P.st;
    }
    // family 3: conical sheets of the ring current and the tail disk (6)
    if (base < CO_NCONE && t < CO_NCONE) {
        const Row4<F> c = ld4(T.cone[t]);
        const F f = c.x, beta = c.y, xs = gm.xx - c.z, d1 = c.w;
        const F y20 = gm.yy / F(20.0);
        const F dd = F(2.0) + d1 * (y20 * y20);
        const F dzs = Dm<F>::sqrt(gm.zz * gm.zz + dd * dd);
        const F ddz_dy = (dd * d1 * gm.yy * F(0.005)) / dzs;
        const F ddz_dz = gm.zz / dzs;
        const F rhos = Dm<F>::sqrt(xs * xs + gm.yy * gm.yy);
        F drx = F(0.0), dry = (gm.yy >= 0) ? F(1.0) : F(-1.0);
        if (rhos >= F(1.0e-5)) {
            drx = xs / rhos;
            dry = gm.yy / rhos;
        }
        F as, da1, da2, s1z, s2z, s1r, s2r;
        cone_terms(dzs + beta, rhos, beta, &as, &da1, &da2, &s1z, &s2z, &s1r, &s2r);
        const F dasx = da1 * (s1r * drx) + da2 * (s2r * drx);
        const F dasy = da1 * (s1z * ddz_dy + s1r * dry) + da2 * (s2z * ddz_dy + s2r * dry);
        const F dasz = da1 * (s1z * ddz_dz) + da2 * (s2z * ddz_dz);
        px += -f * xs * dasz;
        py += -f * gm.yy * dasz;
        pz += f * ((F(2.0) * as + gm.yy * dasy) + xs * dasx);
    }
    // family 4: cylindrical harmonics (6)
    if (base < CO_NCYL && t < CO_NCYL) {
        const Row4<F> c = ld4(T.cyl[t]);
        const F a = c.x, scale = c.y;
        F rho = Dm<F>::sqrt(gm.yy * gm.yy + gm.zz * gm.zz);
        F sinfi, cosfi;
        if (rho < F(1.0e-8)) {
            sinfi = F(1.0);
            cosfi = F(0.0);
            rho = F(1.0e-8);
        } else {
            sinfi = gm.zz / rho;
            cosfi = gm.yy / rho;
        }
        const F dz = rho / scale, xk = gm.xx / scale;
        F j0, j1;
        bessel01(dz, &j0, &j1);
        const F e = Dm<F>::exp(xk);
        if (c.z == F(0.0)) {
            const F sinfi2 = sinfi * sinfi;
            const F si2co2 = sinfi2 - cosfi * cosfi;
            px += -a * j1 * e * sinfi;
            py += a * (F(2.0) * j1 / dz - j0) * e * sinfi * cosfi;
            pz += a * (j1 / dz * si2co2 - j0 * sinfi2) * e;
        } else {
            const F brho = (xk * j0 - (dz * dz + xk - F(1.0)) * j1 / dz) * e * sinfi;
            const F bphi = (j0 + j1 / dz * (xk - F(1.0))) * e * cosfi;
            px += a * (dz * j0 + xk * j1) * e * sinfi;
            py += a * (brho * cosfi - bphi * sinfi);
            pz += a * (brho * sinfi + bphi * cosfi);
        }
    }
    // family 5: tail87 sheet and images (3)
    if (base < CO_NT87 && t < CO_NT87) {
        const Row4<F> c = ld4(T.t87[t]);
        const F zl = gm.zz + c.x, w = c.y;
        const F xnx = F(-10.0) - gm.xx, xnx2 = xnx * xnx;
        const F xc1 = gm.xx + F(1.261), xc2 = gm.xx + F(0.663);
        const F xc22 = xc2 * xc2, xr2 = xc2 * F(-0.1071), xc12 = xc1 * xc1;
        const F b2 = zl * zl + F(9.0);
        const F bb = Dm<F>::sqrt(b2);
        const F xa1 = xc12 + b2, xa2 = F(1.) / (xc22 + b2), xna = xnx2 + b2;
        const F ff = b2 - xc22;
        const F ln1 = Dm<F>::log(F(76.37) / xna);
        const F ln2 = ln1 + F(0.13238005);
        const F s0 = (Dm<F>::atan(xnx / bb) + F(1.5707963)) / bb;
        const F s1 = (ln1 * F(.5) + xc1 * s0) / xa1;
        const F s2 = (xc2 * xa2 * ln2 + F(0.1071) - ff * xa2 * s0) * xa2;
        const F g1 = (b2 * s0 - F(0.5) * xc1 * ln1) / xa1;
        const F g2 = ((F(0.5) * ff * ln2 + F(2.) * s0 * b2 * xc2) * xa2 + xr2) * xa2;
        px += w * zl * (F(0.391734) * s0 + F(5.89715) * s1 + F(24.6833) * s2);
        pz += w * (F(-0.5) * F(0.391734) * ln1 + F(5.89715) * g1 + F(24.6833) * g2);
    }
    // family 6: Region-1 dipoles of the loop basis (21) at the loop point, then of the
    // conical basis (46) at the cone point
    if (base < CO_NDIP_LOOP && gm.need_loop && t < CO_NDIP_LOOP) {
        const Acc3<F> d = coop_dipole(T, t, gm.wl, gm.lx, gm.ly, gm.lz);
        px += d.x;
        py += d.y;
        pz += d.z;
    }
    if (base < CO_NDIP - CO_NDIP_LOOP && gm.need_cone && t < CO_NDIP - CO_NDIP_LOOP) {
        const Acc3<F> d = coop_dipole(T, CO_NDIP_LOOP + t, gm.wc, gm.cxp, gm.cyp, gm.czp);
        px += d.x;
        py += d.y;
        pz += d.z;
    }
    // family 7: current loops (18; groups weighted by wl, wo, wi)
    if (base < CO_NCIRC && t < CO_NCIRC) {
        const F *c = T.circ[t];
        const int g = (int)c[14];
        const F w = g == 0 ? gm.wl : (g == 1 ? gm.wo : gm.wi);
        if (w != F(0.0)) {
            const F ux = (g == 0 ? gm.lx : gm.xx) - c[9], uy = (g == 0 ? gm.ly : gm.yy) - c[10], uz = (g == 0 ? gm.lz : gm.zz) - c[11];
            const F qx = c[0] * ux + c[1] * uy + c[2] * uz;
            const F qy = c[3] * ux + c[4] * uy + c[5] * uz;
            const F qz = c[6] * ux + c[7] * uy + c[8] * uz;
            const v3<F> bl = circle(qx, qy, qz, c[12]);
            const F k = w * c[13];
            px += k * (c[0] * bl.x + c[3] * bl.y + c[6] * bl.z);
            py += k * (c[1] * bl.x + c[4] * bl.y + c[7] * bl.z);
            pz += k * (c[2] * bl.x + c[5] * bl.y + c[8] * bl.z);
        }
    }
    // family 8: conical harmonics (10: Region-1 conical basis, Region-2 inner)
    if (base < CO_NCONIC && t < CO_NCONIC) {
        const Row4<F> c = ld4(T.conic[t]);
        const bool g0 = c.w == F(0.0);
        const F w = g0 ? gm.wc : gm.wi;
        if (w != F(0.0)) {
            const int m = (int)c.x;
            const F X = (g0 ? gm.cxp : gm.xx) + c.z, Y = g0 ? gm.cyp : gm.yy, Z = g0 ? gm.czp : gm.zz;
            // reciprocal square roots in place of the roots and divisions:
            // 1/(r s) = 1/ro, 1/ch^2 = 1/(0.5 (1 + cos)), 1/sh^2 = 1/(0.5 (1 - cos))
            const F ro2 = X * X + Y * Y;
            const F irs = co_rsq(ro2);
            const F ro = ro2 * irs;
            const F cf = X * irs, sf = Y * irs;
            const F rr2 = ro2 + Z * Z;
            const F ir = co_rsq(rr2);
            const F cc = Z * ir, sn = ro * ir;
            const F ch2 = F(0.5) * (F(1.0) + cc), sh2 = F(0.5) * (F(1.0) - cc);
            const F ich = co_rsq(ch2), ish = co_rsq(sh2);
            const F tnh = sh2 * ish * ich, cnh = ch2 * ich * ish;
            // cos/sin(m phi), tnh^m, cnh^m and the (m-1)-th powers
            F cfm = cf, sfm = sf, tm = tnh, cm = cnh, tm1 = F(1.0), cm1 = F(1.0);
#pragma unroll
            for (int j = 2; j <= 5; j++) {
                if (j <= m) {
                    const F cn = cfm * cf - sfm * sf;
                    sfm = sfm * cf + cfm * sf;
                    cfm = cn;
                    tm1 = tm;
                    cm1 = cm;
                    tm = tm * tnh;
                    cm = cm * cnh;
                }
            }
            const F fm = (F)m;
            const F bt = fm * cfm * irs * (tm + cm);
            const F bf = F(-0.5) * fm * sfm * ir * (tm1 * (ich * ich) - cm1 * (ish * ish));
            const F k = w * c.y;
            px += k * (bt * cc * cf - bf * sf);
            py += k * (bt * cc * sf + bf * cf);
            pz += k * (-bt * sn);
        }
    }
    // family 9: Region-2 sheet (15)
    if (base < CO_NSHEET && gm.ws != F(0.0) && t < CO_NSHEET) {
        const F *c = T.sheet[t];
        const int comp = (int)c[3];
        F t1, t2, t3;
        tfac(gm.xks, c[4], c[5], c[6], c[7], &t1, &t2, &t3);
        const F rh2 = gm.xx * gm.xx + gm.yy * gm.yy;
        const F r = Dm<F>::sqrt(rh2 + gm.zz * gm.zz);
        const F rho = Dm<F>::sqrt(rh2);
        const F c1p = gm.xx / rho, s1p = gm.yy / rho;
        const F s2p = F(2.0) * s1p * c1p;
        const F c2p = c1p * c1p - s1p * s1p;
        const F s3p = s2p * c1p + c2p * s1p;
        const F c3p = c2p * c1p - s2p * s1p;
        const F s4p = s3p * c1p + c3p * s1p;
        const F ct = gm.zz / r;
        const F ct2 = ct * ct;
        const F g = c[0] * (comp == 2 ? F(1.0) : ct) * Dm<F>::exp(c[1] * ct2 + c[2]);
        const bool ycomp = comp == 1;
        const F m0 = ycomp ? s1p : F(1.0), m1 = ycomp ? s2p : c1p, m2 = ycomp ? s3p : c2p, m3 = ycomp ? s4p : c3p;
        const F qv = ((m0 * quad(c + 8, t1, t2, t3) + m1 * quad(c + 12, t1, t2, t3)) + m2 * quad(c + 16, t1, t2, t3)) +
                     m3 * quad(c + 20, t1, t2, t3);
        const F v = gm.ws * g * qv;
        if (comp == 0) px += v;
        else if (comp == 1) py += v;
        else pz += v;
    }
    // Region-2 inner dipole distributions (one term, t = 0)
    if (base == 0 && gm.wi != F(0.0) && t == 0) {
        const v3<F> d6 = dipdistr(gm.xx - F(.0774), gm.yy, gm.zz, 0);
        const v3<F> d7 = dipdistr(gm.xx + F(.038), gm.yy, gm.zz, 1);
        const F k = gm.wi * P.b2ampl;
        px += k * (F(29.9996) * d6.x + F(262.886) * d7.x);
        py += k * (F(29.9996) * d6.y + F(262.886) * d7.y);
        pz += k * (F(29.9996) * d6.z + F(262.886) * d7.z);
    }
    return Acc3<F>{px, py, pz};
}

template <class F, int G>
__device__ __forceinline__ T96Share<F> coop_t96_share(const T96Params<F> &P, const CoopTabT<F> &T, F x, F y, F z,
                                                      int gl) {
    using namespace t96;
    T96Share<F> S;
    // ---- uniform preamble (lib/t96/t96.c:22-60)
    const F ys = y * P.ct - z * P.st, zs = z * P.ct + y * P.st;
    const F ysd = ys / F(10.0);
    const F factimf = Dm<F>::exp(x / F(20.0) - ysd * ysd);
    S.oimfy = P.oimfy0 * factimf;
    S.oimfz = P.oimfz0 * factimf;
    const F rho2 = y * y + z * z;
    F xmxm = P.am + x - P.c548;
    if (xmxm < F(0.0)) xmxm = F(0.0);
    const F axx0 = xmxm * xmxm;
    const F aro = P.asq + rho2;
    const F aro_axx0 = aro + axx0;
    S.sigma = Dm<F>::sqrt((aro + axx0 + Dm<F>::sqrt(aro_axx0 * aro_axx0 - P.asq4 * axx0)) / P.asq2);
    S.q = dipole(x, y, z);
    S.fx = mk3<F>(0, 0, 0);
    if (S.sigma >= F(1.08 + 0.005)) return S;
    const F xx = x * P.xappa, yy = y * P.xappa, zz = z * P.xappa;

    // ---- uniform region geometry.  Region 1 (lib/t96/t96.c:619-807): which basis
    // applies, at which point, with which blend weight.
    F lx = xx, ly = yy, lz = zz, cxp = xx, cyp = yy, czp = zz, wl = 0, wc = 0;
    bool need_loop, need_cone;
    {
        const F tnoonn = F(12.0 * 0.01745329), tnoons = F(3.141592654 - 12.0 * 0.01745329);
        const F r2 = xx * xx + yy * yy + zz * zz;
        const F r = Dm<F>::sqrt(r2);
        const F r3 = r * r2;
        // the angles pas, tas by their sines and cosines: sin(tas) = rho/r,
        // cos/sin(pas) = x/rho, y/rho, sin^2(pas/2) = (1 - cos pas)/2, tas > pi/2 for z < 0
        const F rho = Dm<F>::sqrt(xx * xx + yy * yy);
        const F stas = rho / r;
        F cps = F(1.0), sps = F(0.0);
        if (rho > F(0.0)) {
            cps = xx / rho;
            sps = yy / rho;
        }
        F tet0 = Dm<F>::asin(stas / Dm<F>::pow((stas * stas * stas * stas * stas * stas) * (F(1.0) - r3) + r3,
                                                 F(0.1666666667)));
        if (zz < F(0.0)) tet0 = F(3.141592654) - tet0;
        const F dtet = F(8.0 * 0.01745329) * (F(0.5) * (F(1.0) - cps));
        const F tetr1n = tnoonn + dtet, tetr1s = tnoons - dtet;
        const F t01n = tetr1n - F(0.034906), t01s = tetr1s - F(0.034906);
        const F t02s = tetr1s + F(0.034906), t02n = tetr1n + F(0.034906);
        const bool loop_only = (tet0 < t01n || tet0 > t02s);
        const bool cone_only = (tet0 > t02n && tet0 < t01s);
        const bool north = (tet0 >= t01n && tet0 <= t02n);
        const bool south = (tet0 >= t01s && tet0 <= t02s);
        need_loop = loop_only || north || south;
        need_cone = cone_only || north || south;
        if (north || south) {
            const F sqr = Dm<F>::sqrt(r);
            const F ta = north ? t01n : t01s, tb = north ? t02n : t02s;
            const F sa = Dm<F>::sin(ta), sb = Dm<F>::sin(tb);
            const F st1 = sqr / Dm<F>::pow(r3 + F(1.0) / (sa * sa * sa * sa * sa * sa) - F(1.0), F(0.1666666667));
            const F st2 = sqr / Dm<F>::pow(r3 + F(1.0) / (sb * sb * sb * sb * sb * sb) - F(1.0), F(0.1666666667));
            F ct1 = Dm<F>::sqrt(F(1.0) - st1 * st1), ct2 = Dm<F>::sqrt(F(1.0) - st2 * st2);
            if (south) {
                ct1 = -ct1;
                ct2 = -ct2;
            }
            const F px1 = r * st1 * cps, py1 = r * st1 * sps, pz1 = r * ct1;
            const F px2 = r * st2 * cps, py2 = r * st2 * sps, pz2 = r * ct2;
            const F ddx = px2 - px1, ddy = py2 - py1, ddz = pz2 - pz1;
            const F ss = Dm<F>::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            const F qx = xx - px1, qy = yy - py1, qz = zz - pz1;
            const F frac = Dm<F>::sqrt(qx * qx + qy * qy + qz * qz) / ss;
            // loop basis at the first point (north) / second point (south), conical
            // basis at the other one; the first point's basis weighs 1 - frac
            lx = south ? px2 : px1;
            ly = south ? py2 : py1;
            lz = south ? pz2 : pz1;
            cxp = north ? px2 : px1;
            cyp = north ? py2 : py1;
            czp = north ? pz2 : pz1;
            wl = north ? F(1.0) - frac : frac;
            wc = north ? frac : F(1.0) - frac;
        } else {
            wl = loop_only ? F(1.0) : F(0.0);
            wc = cone_only ? F(1.0) : F(0.0);
        }
    }
    // Region 2 (lib/t96/t96.c:1265-1321): weights of the outer, sheet and inner
    // representations, already times the -0.02 scale of r2birk
    F wo = 0, ws = 0, wi = 0;
    const F xks = xksi(xx, yy, zz);
    {
        const F del = F(0.03), del1 = F(0.015);
        if (xks < -(del + del1)) {
            wo = F(-0.02);
        } else if (xks >= -(del + del1) && xks < -del + del1) {
            ws = F(-0.02) * tksi(xks, -del, del1);
            wo = F(-0.02) - ws;
        } else if (xks >= -del + del1 && xks < del - del1) {
            ws = F(-0.02);
        } else if (xks >= del - del1 && xks < del + del1) {
            wi = F(-0.02) * tksi(xks, del, del1);
            ws = F(-0.02) - wi;
        } else if (xks >= del + del1) {
            wi = F(-0.02);
        }
    }

    const T96Geo<F> geo = {xx, yy, zz, xx, ys * P.xappa, zs * P.xappa, lx, ly, lz, cxp, cyp, czp,
                        wl, wc, wo, ws, wi, xks, need_loop, need_cone};
    // ---- this lane's block of slots, pairwise in slot order
    constexpr int NP = 64 / G;
    constexpr int LP = ilog2c(NP);
    const int r = lane_residue<G>(gl);
    SlotTree<Acc3<F>, NP> tree;
    static_for<NP>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        constexpr int base = G * brevc(j, LP);
        tree.template push<j>(coop_t96_slot(P, T, geo, base, r + base));
    });
    const Acc3<F> tot = tree.total();
    S.fx = mk3<F>(tot.x, tot.y, tot.z);
    return S;
}

// Uniform end of lib/t96/t96.c:94-111 once FX is summed.
template <class F>
__device__ __forceinline__ v3<F> t96_finish(const T96Share<F> &S, v3<F> fx) {
    const v3<F> q = S.q;
    if (S.sigma >= F(1.08 + 0.005)) return mk3<F>(-q.x, S.oimfy - q.y, S.oimfz - q.z);
    if (S.sigma < F(1.08 - 0.005)) return fx;
    const F fint = F(0.5) * (F(1.0) - (S.sigma - F(1.08)) / F(0.005));
    const F fext = F(0.5) * (F(1.0) + (S.sigma - F(1.08)) / F(0.005));
    return mk3<F>((fx.x + q.x) * fint - q.x, (fx.y + q.y) * fint + S.oimfy * fext - q.y,
                  (fx.z + q.z) * fint + S.oimfz * fext - q.z);
}

// ---------------------------------------------------------------- IGRF, cooperative

// lib/internal/internal.c:182-275 with one order m per lane (lanes 0..kw-1; groups of 8
// take a second order, m + 8, in a second pass): each lane advances the sectoral terms
// to its order, then runs the degree recursion of its column.  Returns this lane's share
// of (Br, Btheta, Bphi*sin).
template <class F>
struct IgrfShare {
    F br, bt, bf;
    F s, c, cf, sf;
    bool polar;
};

template <class F, int GL>
__device__ __forceinline__ IgrfShare<F> coop_igrf_share(const F *__restrict__ G, const F *__restrict__ H,
                                                        const F *__restrict__ REC, F x, F y, F z,
                                                     int gl) {
    IgrfShare<F> S;
    const F rho2 = x * x + y * y;
    const F r = Dm<F>::sqrt(rho2 + z * z);
    const F c = z / r;
    const F rho = Dm<F>::sqrt(rho2);
    const F s = rho / r;
    S.polar = s < F(1.E-5);
    const F cf = S.polar ? F(1.) : x / rho;
    const F sf = S.polar ? F(0.) : y / rho;
    S.s = s;
    S.c = c;
    S.cf = cf;
    S.sf = sf;
    const F inv_r = F(1.) / r;
    int nm = 3 + 30 / ((int)r + 2);
    if (nm > 13) nm = 13;
    const int kw = nm + 1;
    F am = inv_r * inv_r, pmm = F(1.), dmm = F(0.), cm = F(0.), sm = F(1.);
    // orders m = t + 1 in the canonical slots: this lane's terms t = r + GL*p; the
    // sectoral terms advance monotonically over the passes
    constexpr int NPI = (14 + GL - 1) / GL;   // passes that can hold an order (m <= 14)
    const int r0 = lane_residue<GL>(gl);
    F cbr[NPI], cbt[NPI], cbf[NPI];
#pragma unroll
    for (int pass = 0; pass < NPI; pass++) {
        const int m = r0 + 1 + GL * pass;
        const int m0 = pass == 0 ? 1 : m - GL;    // order the sectoral terms are at
        // the pass's lanes hold orders GL*pass+1 .. GL*(pass+1), so only j in
        // [GL*(pass-1)+1, GL*(pass+1)) can advance one of them
        const int j_lo = pass == 0 ? 1 : GL * (pass - 1) + 1;
        const int j_hi = kw < GL * (pass + 1) ? kw : GL * (pass + 1);
        for (int j = j_lo; j < j_hi; j++) {
            if (j >= m0 && j < m) {
                const F w = cm;
                cm = w * cf + sm * sf;
                sm = sm * cf - w * sf;
                const F d2 = s * dmm + c * pmm;
                pmm = s * pmm;
                dmm = d2;
                am = am * inv_r;
            }
        }
        F br = F(0.), bt = F(0.), bf = F(0.);
        if (m <= kw) {
            F bphi = F(0.);
            F q = pmm, dq = dmm, q2 = F(0.), dq2 = F(0.), an = am;
            int idx = m * (m - 1) / 2 + m - 1;
            // coefficients two degrees ahead are read while this degree is summed (LDS
            // latency; the tables are zero-padded past the last index read)
            F g = G[idx], h = H[idx], rec = REC[idx];
            F gn = G[idx + m], hn = H[idx + m], recn = REC[idx + m];
            for (int n = m; n <= kw; n++) {
                const int i2 = idx + n + (n + 1);
                const F g2 = G[i2], h2 = H[i2], rec2 = REC[i2];
                const F w = g * sm + h * cm;
                br = br + (an * n) * w * q;
                bt = bt - an * w * dq;
                bphi = bphi + an * (g * cm - h * sm) * (S.polar ? dq : q);
                const F dnext = c * dq - s * q - rec * dq2;
                const F pnext = c * q - rec * q2;
                dq2 = dq;
                q2 = q;
                dq = dnext;
                q = pnext;
                an = an * inv_r;
                idx += n;
                g = gn;
                h = hn;
                rec = recn;
                gn = g2;
                hn = h2;
                recn = rec2;
            }
            if (m >= 2) bf = bphi * (F)(m - 1);
        }
        cbr[pass] = br;
        cbt[pass] = bt;
        cbf[pass] = bf;
    }
    constexpr int NP = 64 / GL;
    constexpr int LP = ilog2c(NP);
    SlotTree<Acc3<F>, NP> tree;
    static_for<NP>([&](auto jc) __attribute__((always_inline)) {
        constexpr int pass = brevc(decltype(jc)::value, LP);
        if constexpr (pass < NPI) {
            tree.template push<decltype(jc)::value>(Acc3<F>{cbr[pass], cbt[pass], cbf[pass]});
        } else {
            tree.template push<decltype(jc)::value>(Acc3<F>{F(0.), F(0.), F(0.)});
        }
    });
    const Acc3<F> tot = tree.total();
    S.br = tot.x;
    S.bt = tot.y;
    S.bf = tot.z;
    return S;
}

template <class F>
__device__ __forceinline__ v3<F> igrf_finish(const IgrfShare<F> &S, F br, F bt, F bf) {
    F bp;
    if (S.polar) {
        bp = S.c < F(0.) ? -bf : bf;
    } else {
        bp = bf / S.s;
    }
    const F he = br * S.s + bt * S.c;
    return mk3<F>(he * S.cf - bp * S.sf, he * S.sf + bp * S.cf, br * S.c - bt * S.s);
}

// ---------------------------------------------------------------- total field

// GEO field [nT] of the model set at the GEO position geo_m / GSM position gsm_m [m],
// identical in every lane of a group of GS lanes (gl: the lane's index in the group;
// each group may be at its own position).
// Position in Earth radii: a multiplication in FP32 (the same FP32 value to its last
// bit), the reference's division in FP64.
__device__ __forceinline__ float to_re(double x, float) { return (float)(x * K_INV_RE); }
__device__ __forceinline__ double to_re(double x, double) { return x / K_RE; }

template <class F, int INT, int EXT, int GS>
__device__ __forceinline__ v3<F> coop_field(const F *G, const F *H, const F *REC, const F *Af,
                                            const T96Params<F> &P, const CoopTabT<F> &T, v3<double> geo_m,
                                            v3<double> gsm_m, int gl) {
    static_assert(GS == 4 || GS == 8 || GS == 16 || GS == 32 || GS == 64, "groups of 4, 8, 16, 32 or 64 lanes");
    IgrfShare<F> I;
    T96Share<F> X;
    F p[6] = {0, 0, 0, 0, 0, 0};
    if constexpr (INT == INT_IGRF) {
        I = coop_igrf_share<F, GS>(G, H, REC, to_re(geo_m.x, F(0)), to_re(geo_m.y, F(0)), to_re(geo_m.z, F(0)), gl);
        p[0] = I.br;
        p[1] = I.bt;
        p[2] = I.bf;
    }
    if constexpr (EXT == EXT_T96) {
        X = coop_t96_share<F, GS>(P, T, to_re(gsm_m.x, F(0)), to_re(gsm_m.y, F(0)), to_re(gsm_m.z, F(0)), gl);
        p[3] = X.fx.x;
        p[4] = X.fx.y;
        p[5] = X.fx.z;
    }
    group_sum<6, GS>(p);
    v3<F> tot = mk3<F>(0, 0, 0);
    if constexpr (INT == INT_IGRF) tot = igrf_finish(I, p[0], p[1], p[2]);
    if constexpr (EXT == EXT_T96) {
        const v3<F> e = t96_finish(X, mk3<F>(p[3], p[4], p[5]));
        tot.x += Af[0] * e.x + Af[3] * e.y + Af[6] * e.z;
        tot.y += Af[1] * e.x + Af[4] * e.y + Af[7] * e.z;
        tot.z += Af[2] * e.x + Af[5] * e.y + Af[8] * e.z;
    }
    static_assert(EXT != EXT_TS05, "model not compiled into the cooperative tracer");
    return tot;
}

// Copy the lane tables into the workgroup's LDS (all threads of the block).
template <class F>
__device__ __forceinline__ void stage_coop(const CoopTabT<F> *__restrict__ src, CoopTabT<F> *dst) {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d = reinterpret_cast<uint32_t *>(dst);
    for (int k = threadIdx.x; k < (int)(sizeof(CoopTabT<F>) / 4); k += blockDim.x) d[k] = s[k];
    __syncthreads();
}

}  // namespace cor
