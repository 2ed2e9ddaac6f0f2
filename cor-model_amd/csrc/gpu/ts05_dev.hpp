// ts05_dev.hpp -- Tsyganenko-Sitnov 2005 external field (zero tilt) for one lane,
// templated on the arithmetic type.  GSM position [Re] -> GSM field [nT].
//
// Follows lib/ts05/{ts05,deformed,birk_tot,full_rc}.c (function-level citations below):
//   * FP64 (parity path): the reference's expression order, its central differences
//     (r_s/theta_s with +-1e-6, the ring-current vector potentials with +-1e-4), and
//     per-epoch terms hoisted to the host with the reference's arithmetic (TS05Params):
//     the driver amplitudes, the tail shifts, the shield coefficient sums, the
//     harmonic scales and the two tan() constants of the cone model.
//   * FP32 (mixed path): the same model, but every derivative the reference takes by
//     central differences is taken exactly by forward-mode dual numbers (Dual<float>):
//     in FP32 a 1e-6 step is below the resolution of r, and a 1e-4 step leaves 3 digits.
//     The Cartesian-harmonic shield and warping sums stay FP64 (their terms cancel by
//     4-6 orders of magnitude).
// Only the fields that reach the total are evaluated (ts05_gsm discards the Birkeland
// mode-2 fields R12/R22, lib/ts05/ts05.c:112-114).
#pragma once
#include "devmath.hpp"
#include "t96_dev.hpp"
#include "ts05_tables.h"
#include "fastmath64.hpp"

namespace cor {

// Per-epoch constants (lib/ts05/ts05.c:19-110 and the coefficient sums of the shields).
template <class T>
struct TS05Params {
    T xappa, xappa3, am, x0, oimfy, oimfz, byimf, bzimf;
    T dxshift1, dxshift2, d0, xkappa1, xkappa2, sc_sy, sc_pr;
    T tamp1, tamp2, a_src, a_prc, a_r11, a_r21, fac_sy, fac_pr;
    T tgp[2], tgm[2];            // tan((theta0 +- dtheta)/2) of the Region-1 / Region-2 cones
    T cphi, sphi;                // cos/sin of the partial ring current's rotation A[37]
    T w1[25], w2[25];            // warped-tail harmonic coefficients A[2l] + A[2l+1] * dxshift
    T wrp[2][10];                // their 1/P, 1/R scales
    T wsq[2][25];                // sqrt(1/P^2 + 1/R^2)
    T r11[9], r21[9], src[9], prc[9];   // shield coefficient sums a0 + s a1 + a2 + s a3
    T shinv[4][6];               // the four shields' 1/P_j, 1/R_k (order r11, r21, src, prc)
    T shsq[4][9];                // and their sqrt(1/P_j^2 + 1/R_k^2)
};

// Forward-mode dual number: value and derivative along one direction.
template <class T>
struct Dual {
    T v, d;
};
template <class T>
__device__ __forceinline__ Dual<T> dl(T v, T d) {
    Dual<T> r;
    r.v = v;
    r.d = d;
    return r;
}

// Arithmetic shared by plain and dual scalars: S is T or Dual<T>.
template <class T> __device__ __forceinline__ T val(T x) { return x; }
template <class T> __device__ __forceinline__ T val(Dual<T> x) { return x.v; }
template <class T> __device__ __forceinline__ Dual<T> operator+(Dual<T> a, Dual<T> b) { return dl(a.v + b.v, a.d + b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a, Dual<T> b) { return dl(a.v - b.v, a.d - b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a) { return dl(-a.v, -a.d); }
template <class T> __device__ __forceinline__ Dual<T> operator*(Dual<T> a, Dual<T> b) { return dl(a.v * b.v, a.d * b.v + a.v * b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator/(Dual<T> a, Dual<T> b) {
    const T q = a.v / b.v;
    return dl(q, (a.d - q * b.d) / b.v);
}
template <class T> __device__ __forceinline__ Dual<T> operator+(Dual<T> a, T b) { return dl(a.v + b, a.d); }
template <class T> __device__ __forceinline__ Dual<T> operator+(T a, Dual<T> b) { return dl(a + b.v, b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator-(Dual<T> a, T b) { return dl(a.v - b, a.d); }
template <class T> __device__ __forceinline__ Dual<T> operator-(T a, Dual<T> b) { return dl(a - b.v, -b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator*(Dual<T> a, T b) { return dl(a.v * b, a.d * b); }
template <class T> __device__ __forceinline__ Dual<T> operator*(T a, Dual<T> b) { return dl(a * b.v, a * b.d); }
template <class T> __device__ __forceinline__ Dual<T> operator/(Dual<T> a, T b) { return dl(a.v / b, a.d / b); }
template <class T> __device__ __forceinline__ Dual<T> operator/(T a, Dual<T> b) {
    const T q = a / b.v;
    return dl(q, -q * b.d / b.v);
}
template <class T> __device__ __forceinline__ T fsqrt(T x) { return Dm<T>::sqrt(x); }
template <class T> __device__ __forceinline__ Dual<T> fsqrt(Dual<T> x) {
    const T s = Dm<T>::sqrt(x.v);
    return dl(s, x.d / (T(2) * s));
}
template <class T> __device__ __forceinline__ T fexp(T x) { return Dm<T>::exp(x); }
template <class T> __device__ __forceinline__ Dual<T> fexp(Dual<T> x) {
    const T e = Dm<T>::exp(x.v);
    return dl(e, e * x.d);
}
template <class T> __device__ __forceinline__ T flog(T x) { return Dm<T>::log(x); }
template <class T> __device__ __forceinline__ Dual<T> flog(Dual<T> x) { return dl(Dm<T>::log(x.v), x.d / x.v); }
template <class T> __device__ __forceinline__ T fpow(T x, T c) { return Dm<T>::pow(x, c); }
template <class T> __device__ __forceinline__ Dual<T> fpow(Dual<T> x, T c) {
    const T p = Dm<T>::pow(x.v, c);
    return dl(p, c * p / x.v * x.d);
}
template <class T> __device__ __forceinline__ T fsin(T x) { return Dm<T>::sin(x); }
template <class T> __device__ __forceinline__ Dual<T> fsin(Dual<T> x) { return dl(Dm<T>::sin(x.v), Dm<T>::cos(x.v) * x.d); }
template <class T> __device__ __forceinline__ T fcos(T x) { return Dm<T>::cos(x); }
template <class T> __device__ __forceinline__ Dual<T> fcos(Dual<T> x) { return dl(Dm<T>::cos(x.v), -Dm<T>::sin(x.v) * x.d); }

namespace ts05 {

#define CT2(name) name##_d
using t96::pin;
using t96::tab_sel;

COR_CTAB(k_ts05_a11, 31, TS05_A11)
COR_CTAB(k_ts05_a21, 31, TS05_A21)
// the four shield tables; only entries 72..77 (P and R scales) are read on the device,
// the coefficient sums are per-epoch (TS05Params)
COR_CTAB(k_ts05_sh11, 78, TS05_SH11)
COR_CTAB(k_ts05_sh21, 78, TS05_SH21)
COR_CTAB(k_ts05_csy, 78, TS05_CSY)
COR_CTAB(k_ts05_cpr, 78, TS05_CPR)


// lib/ts05/ts05.c:233-246
template <class T>
__device__ __forceinline__ v3<T> dipole(T x, T y, T z) {
    const T p = x * x, u = z * z, t = y * y;
    const T s = Dm<T>::sqrt(p + t + u);
    const T q = T(30115.) / (s * s * s * s * s);
    return mk3<T>(q * T(3.) * z * x, q * T(-3.) * z * y, q * (p + t - T(2.) * u));
}

// lib/ts05/ts05.c:136-231: 3 x 3 harmonics; the third z scale has its own term form
template <class T>
__device__ __forceinline__ v3<T> shlcar3x3(T x, T y, T z) {
    const T P[3] = {T(9.620648151), T(6.082014949), T(27.75216226)};
    const T R[3] = {T(12.44199571), T(5.122226936), T(6.982039615)};
    const T A[9] = {T(-901.2327248 + (double)895.8011176), T(817.6208321 - (double)845.5880889),
                    T(-83.73539535 + (double)86.58542841), T(336.8781402 - (double)329.3619944),
                    T(-311.2947120 + (double)308.6011161), T(31.94469304 - (double)31.30824526),
                    T(125.8739681 - (double)372.3384278), T(-235.4720434 + (double)286.7594095),
                    T(21.86305585 - (double)27.42344605)};
    T czr[3], szr[3], zr[3], rinv2[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        zr[k] = z / R[k];
        Dm<T>::sincos(zr[k], &szr[k], &czr[k]);
    }
    rinv2[0] = T(1.) / (R[0] * R[0]);
    rinv2[1] = T(1.) / (R[1] * R[1]);
    const T r32 = R[2] * R[2];
    rinv2[2] = T(1.) / r32;
    T bx = 0, by = 0, bz = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T pinv2 = T(1.) / (P[i] * P[i]);
        T syp, cyp;
        Dm<T>::sincos(y / P[i], &syp, &cyp);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const T sq = Dm<T>::sqrt(pinv2 + rinv2[k]);
            const T e = Dm<T>::exp(sq * x);
            T hx, hy, hz;
            if (k < 2) {
                hx = -sq * e * cyp * szr[k];
                hy = e / P[i] * syp * szr[k];
                hz = -e * cyp / R[k] * czr[k];
            } else {
                hx = -e * cyp * (sq * z * czr[2] + szr[2] / R[2] * (x + T(1.) / sq));
                hy = e / P[i] * syp * (z * czr[2] + x / R[2] * szr[2] / sq);
                hz = -e * cyp * (czr[2] * (T(1.) + x / r32 / sq) - zr[2] * szr[2]);
            }
            const T a = A[3 * i + k];
            if (i == 0 && k == 0) {
                bx = a * hx;
                by = a * hy;
                bz = a * hz;
            } else {
                bx = bx + a * hx;
                by = by + a * hy;
                bz = bz + a * hz;
            }
        }
    }
    return mk3<T>(bx, by, bz);
}

// lib/ts05/deformed.c:119-180: five conical current sheets of the tail
template <class T>
__device__ __forceinline__ v3<T> taildisk(T d0, T deltadx, T x, T y, T z) {
    const T F[5] = {T(-71.09346626), T(-1014.308601), T(-1272.939359), T(-3224.935936), T(-44546.86232)};
    const T B[5] = {T(10.90101242), T(12.68393898), T(13.51791954), T(14.86775017), T(15.12306404)};
    const T C[5] = {T(.7954069972), T(.6716601849), T(1.174866319), T(2.565249920), T(10.01986790)};
    const T rho = Dm<T>::sqrt(x * x + y * y);
    const T drhodx = x / rho, drhody = y / rho;
    const T dex = Dm<T>::exp(x / T(7.));
    const T y20 = y / T(20.);
    const T d = d0 + T(4.7) * y20 * y20 + deltadx * dex;
    const T dddy = T(4.7) * y * T(0.005);
    const T dddx = deltadx / T(7.) * dex;
    const T dzeta = Dm<T>::sqrt(z * z + d * d);
    const T ddzdx = d * dddx / dzeta, ddzdy = d * dddy / dzeta, ddzdz = z / dzeta;
    T dbx = 0, dby = 0, dbz = 0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const T bi = B[i], dzc = dzeta + C[i], dzc2 = dzc * dzc;
        const T rpb = rho + bi, rmb = rho - bi;
        const T s1 = Dm<T>::sqrt(rpb * rpb + dzc2), s2 = Dm<T>::sqrt(rmb * rmb + dzc2);
        const T ds1drho = (rho + bi) / s1, ds2drho = (rho - bi) / s2;
        const T ds1ddz = dzc / s1, ds2ddz = dzc / s2;
        const T ds1dx = ds1drho * drhodx + ds1ddz * ddzdx;
        const T ds1dy = ds1drho * drhody + ds1ddz * ddzdy;
        const T ds1dz = ds1ddz * ddzdz;
        const T ds2dx = ds2drho * drhodx + ds2ddz * ddzdx;
        const T ds2dy = ds2drho * drhody + ds2ddz * ddzdy;
        const T ds2dz = ds2ddz * ddzdz;
        const T s1ts2 = s1 * s2, s1ps2 = s1 + s2, s1ps2sq = s1ps2 * s1ps2;
        const T bi2 = T(2.) * bi;
        const T fac1 = Dm<T>::sqrt(s1ps2sq - bi2 * bi2);
        const T as = fac1 / (s1ts2 * s1ps2sq);
        const T dasds1 = (T(1.) / (fac1 * s2) - as / s1ps2 * (s2 * s2 + s1 * (T(3.) * s1 + T(4.) * s2))) / (s1 * s1ps2);
        const T dasds2 = (T(1.) / (fac1 * s1) - as / s1ps2 * (s1 * s1 + s2 * (T(3.) * s2 + T(4.) * s1))) / (s2 * s1ps2);
        const T dasdx = dasds1 * ds1dx + dasds2 * ds2dx;
        const T dasdy = dasds1 * ds1dy + dasds2 * ds2dy;
        const T dasdz = dasds1 * ds1dz + dasds2 * ds2dz;
        dbx = dbx - F[i] * x * dasdz;
        dby = dby - F[i] * y * dasdz;
        dbz = dbz + F[i] * (T(2.) * as + x * dasdx + y * dasdy);
        COR_SEQ();
    }
    return mk3<T>(dbx, dby, dbz);
}

// e^x and sin/cos of the FP64 harmonic sums: in the mixed path (FAST) one shared
// reduction and Taylor series (fastmath64.hpp), within a few ulp; the FP64 path keeps
// the library's correctly rounded routines.
template <class T, bool FAST>
__device__ __forceinline__ T hexp(T v) {
    if constexpr (FAST) return fm64::exp(v);
    else return Dm<T>::exp(v);
}
template <class T, bool FAST>
__device__ __forceinline__ void hsincos(T v, T *s, T *c) {
    if constexpr (FAST) fm64::sincos(v, s, c);
    else Dm<T>::sincos(v, s, c);
}

// lib/ts05/deformed.c:182-208 with the per-epoch coefficients and scales
template <class T, bool FAST = false>
__device__ __forceinline__ v3<T> shlcar5x5(const T *coef, const T *rpr, const T *sqt, T x, T y, T z) {
    static_assert(!FAST || Dm<T>::fp64, "the fast harmonic-sum functions are FP64");
    T hx = 0, hy = 0, hz = 0;
    T szk[5], czk[5];
#pragma unroll
    for (int k = 0; k < 5; k++) hsincos<T, FAST>(z * rpr[5 + k], &szk[k], &czk[k]);
    for (int i = 0; i < 5; i++) {
        const T rp = rpr[i];
        T sy, cy;
        hsincos<T, FAST>(y * rp, &sy, &cy);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const T rr = rpr[5 + k];
            const T sq = sqt[5 * i + k];
            const T e = hexp<T, FAST>(x * sq);
            const T c = coef[5 * i + k];
            hx = hx + c * (-sq * e * cy * szk[k]);
            hy = hy + c * (rp * e * sy * szk[k]);
            hz = hz + c * (-rr * e * cy * czk[k]);
        }
    }
    return mk3<T>(hx, hy, hz);
}

// lib/ts05/deformed.c:4-117: the two tail modes at the unwarped point (identity at
// zero tilt, up to the polar round trip, kept)
template <class T>
__device__ __forceinline__ v3<T> to3(v3<double> v) {
    return mk3<T>((T)v.x, (T)v.y, (T)v.z);
}

template <class T>
__device__ __forceinline__ void warped(const TS05Params<T> &P, const TS05Params<double> &Pd, T x, T y, T z,
                                       v3<T> *o1, v3<T> *o2) {
    const T rho = Dm<T>::sqrt(y * y + z * z);
    T phi = 0, cphi = 1, sphi = 0;
    if (y != T(0.) || z != T(0.)) {
        phi = Dm<T>::atan2(z, y);
        cphi = y / rho;
        sphi = z / rho;
    }
    T sf, cf;
    Dm<T>::sincos(phi, &sf, &cf);
    const T yu = rho * cf, zu = rho * sf;
    v3<T> u[2];
    {
        v3<T> f = taildisk(P.d0 * T(1.1), T(1.), (x - T(6.) - P.dxshift1) * T(1.1) + T(1.2), yu * T(1.1), zu * T(1.1));
        pin(f);
        COR_SEQ();
        const v3<T> h = to3<T>(shlcar5x5<double, !Dm<T>::fp64>(Pd.w1, Pd.wrp[0], Pd.wsq[0], (double)x, (double)yu,
                                                               (double)zu));
        u[0] = mk3<T>(f.x + h.x, f.y + h.y, f.z + h.z);
        pin(u[0]);
    }
    COR_SEQ();
    {
        v3<T> f = taildisk(P.d0 * T(0.25), T(0.), (x - T(4.) - P.dxshift2) * T(0.25) - T(9.), yu * T(0.25), zu * T(0.25));
        pin(f);
        COR_SEQ();
        const v3<T> h = to3<T>(shlcar5x5<double, !Dm<T>::fp64>(Pd.w2, Pd.wrp[1], Pd.wsq[1], (double)x, (double)yu,
                                                               (double)zu));
        u[1] = mk3<T>(f.x + h.x, f.y + h.y, f.z + h.z);
        pin(u[1]);
    }
    v3<T> *o[2] = {o1, o2};
#pragma unroll
    for (int m = 0; m < 2; m++) {
        const T brho = u[m].y * cf + u[m].z * sf;
        const T bphi = -u[m].y * sf + u[m].z * cf;
        *o[m] = mk3<T>(u[m].x, brho * cphi - bphi * sphi, brho * sphi + bphi * cphi);
    }
}

// lib/ts05/birk_tot.c:284-307: deformation of the cone system (S: plain or dual)
template <class T, class S>
__device__ __forceinline__ S rs_fn(const T *a, S r, S th) {
    const S r2 = r * r;
    const S r2aa = r2 + a[15] * a[15];
    return r + a[1] / r + a[2] * r / fsqrt(r2 + a[10] * a[10]) + a[3] * r / (r2 + a[11] * a[11]) +
           (a[4] + a[5] / r + a[6] * r / fsqrt(r2 + a[12] * a[12]) + a[7] * r / (r2 + a[13] * a[13])) * fcos(th) +
           (a[8] * r / fsqrt(r2 + a[14] * a[14]) + a[9] * r / (r2aa * r2aa)) * fcos(T(2.) * th);
}
template <class T, class S>
__device__ __forceinline__ S ths_fn(const T *a, S r, S th) {
    const S r2 = r * r;
    return th + (a[16] + a[17] / r + a[18] / r2 + a[19] * r / fsqrt(r2 + a[26] * a[26])) * fsin(th) +
           (a[20] + a[21] * r / fsqrt(r2 + a[27] * a[27]) + a[22] * r / (r2 + a[28] * a[28])) * fsin(T(2.) * th) +
           (a[23] + a[24] / r + a[25] * r / (r2 + a[29] * a[29])) * fsin(T(3.) * th);
}

// lib/ts05/birk_tot.c:309-377 for mode 1
template <class T>
__device__ __forceinline__ void fialcos1(T r, T th, T ph, T th0, T dt, T tgp, T tgm, T *bth, T *bph) {
    T st, ct, sp, cp;
    Dm<T>::sincos(th, &st, &ct);
    Dm<T>::sincos(ph, &sp, &cp);
    const T ro = r * st;
    T tg, ctg;   // tan(th/2), cot(th/2)
    if constexpr (Dm<T>::fp64) {
        tg = st / (T(1.) + ct);
        ctg = st / (T(1.) - ct);
    } else {
        // the same half-angle tangents in the form without the cancellation of 1 -+ cos th:
        // in FP32, cos th rounds to +-1 within ~3e-4 rad of the cone's axis, where the
        // reference's form divides by zero and the field would turn NaN
        if (ct >= T(0.)) {
            tg = st / (T(1.) + ct);
            ctg = (T(1.) + ct) / st;
        } else {
            tg = (T(1.) - ct) / st;
            ctg = st / (T(1.) - ct);
        }
    }
    const T tnp = th0 + dt, tnm = th0 - dt;
    T tgm2 = 0, tgp2 = 0;
    if (th >= tnm) {
        tgm2 = tgm * tgm;
        tgp2 = tgp * tgp;
    } else {
        tgp = 0;
        tgm = 0;
    }
    const T tm = tg;
    const T cc = cp - T(0.) * sp;
    const T ss = T(0.) * cp + sp;
    T t, dtt;
    if (th < tnm) {
        t = tm;
        dtt = T(0.5) * tm * (tg + ctg);
    } else if (th < tnp) {
        const T fc = T(1.) / (tgp - tgm), fc1 = T(1.) / T(3.);
        const T tgm2m1 = tgm2 * tgm;
        const T tg21 = T(1.) + tg * tg;
        t = fc * (tm * (tgp - tg) + fc1 * (tm * tg - tgm2m1 / tm));
        dtt = T(0.5) * fc * tg21 * (tm / tg * (tgp - tg) - fc1 * (tm - tgm2m1 / (tm * tg)));
    } else {
        const T fc = T(1.) / (tgp - tgm), fc1 = T(1.) / T(3.);
        t = fc * fc1 * (tgp2 * tgp - tgm2 * tgm) / tm;
        dtt = -t * T(0.5) * (tg + ctg);
    }
    *bth = t * cc / ro * T(800.);
    *bph = -dtt * ss / r * T(800.);
}

// lib/ts05/birk_tot.c:252-282: one cone.  FP64: the reference's central differences;
// FP32: exact derivatives by dual numbers.
template <class T>
__device__ __forceinline__ v3<T> one_cone(const T *a, T x, T y, T z, T dth, T tgp, T tgm) {
    const T rho2 = x * x + y * y;
    const T rho = Dm<T>::sqrt(rho2);
    const T r = Dm<T>::sqrt(rho2 + z * z);
    const T th = Dm<T>::atan2(rho, z);
    const T ph = Dm<T>::atan2(y, x);
    T rs, ths, drsdr, drsdt, dtsdr, dtsdt;
    if constexpr (Dm<T>::fp64) {
        rs = rs_fn<T, T>(a, r, th);
        ths = ths_fn<T, T>(a, r, th);
        drsdr = (rs_fn<T, T>(a, r + T(1E-6), th) - rs_fn<T, T>(a, r - T(1E-6), th)) / T(2E-6);
        drsdt = (rs_fn<T, T>(a, r, th + T(1E-6)) - rs_fn<T, T>(a, r, th - T(1E-6))) / T(2E-6);
        dtsdr = (ths_fn<T, T>(a, r + T(1E-6), th) - ths_fn<T, T>(a, r - T(1E-6), th)) / T(2E-6);
        dtsdt = (ths_fn<T, T>(a, r, th + T(1E-6)) - ths_fn<T, T>(a, r, th - T(1E-6))) / T(2E-6);
    } else {
        const Dual<T> rr = rs_fn<T, Dual<T>>(a, dl(r, T(1)), dl(th, T(0)));
        const Dual<T> rt = rs_fn<T, Dual<T>>(a, dl(r, T(0)), dl(th, T(1)));
        const Dual<T> tr = ths_fn<T, Dual<T>>(a, dl(r, T(1)), dl(th, T(0)));
        const Dual<T> tt = ths_fn<T, Dual<T>>(a, dl(r, T(0)), dl(th, T(1)));
        rs = rr.v;
        ths = tr.v;
        drsdr = rr.d;
        drsdt = rt.d;
        dtsdr = tr.d;
        dtsdt = tt.d;
    }
    T btast, bfast;
    fialcos1(rs, ths, ph, a[30], dth, tgp, tgm, &btast, &bfast);
    const T rsr = rs / r;
    const T stsst = Dm<T>::sin(ths) / Dm<T>::sin(th);
    const T br = -rsr / r * stsst * btast * drsdt;
    const T bth = rsr * stsst * btast * drsdr;
    const T bph = rsr * bfast * (drsdr * dtsdt - drsdt * dtsdr);
    const T s = rho / r, c = z / r, sf = y / rho, cf = x / rho;
    const T be = br * s + bth * c;
    return mk3<T>(a[0] * (be * cf - bph * sf), a[0] * (be * sf + bph * cf), a[0] * (br * c - bth * s));
}

// lib/ts05/birk_tot.c:117-250 for mode 1: Region 1 (numb 0) or Region 2 (numb 1)
template <class T>
__device__ __forceinline__ v3<T> birk_1n2(const TS05Params<T> &P, int numb, T x, T y, T z) {
    const T *a = numb == 0 ? CT(k_ts05_a11) : CT(k_ts05_a21);
    const T xkappa = numb == 0 ? P.xkappa1 : P.xkappa2;
    const T dphi = numb == 0 ? T(0.055) : T(0.030);
    const T dth = numb == 0 ? T(0.06) : T(0.09);
    const T xs = x * xkappa, ys = y * xkappa, zs = z * xkappa;
    const T rho = Dm<T>::sqrt(xs * xs + zs * zs);
    const T rho2 = rho * rho;
    T phi = 0;
    if (xs != T(0.) || zs != T(0.)) phi = Dm<T>::atan2(-zs, xs);
    T sphic, cphic;
    Dm<T>::sincos(phi, &sphic, &cphic);
    const T den = T(49.) + rho2;
    const T brack = dphi + T(0.5) * T(49.) / T(50.) * (rho2 - T(1.)) / den;
    const T phis = phi - brack * Dm<T>::sin(phi);
    const T dphisphi = T(1.) - brack * Dm<T>::cos(phi);
    const T dphisrho = T(-2.) * T(0.5) * T(49.) * rho / (den * den) * Dm<T>::sin(phi);
    T sphics, cphics;
    Dm<T>::sincos(phis, &sphics, &cphics);
    const T xc = rho * cphics, zc = -rho * sphics;
    v3<T> bn = one_cone(a, xc, ys, zc, dth, P.tgp[numb], P.tgm[numb]);
    pin(bn);
    COR_SEQ();
    const v3<T> bs = one_cone(a, xc, -ys, -zc, dth, P.tgp[numb], P.tgm[numb]);
    const T bxs = bn.x - bs.x, byas = bn.y + bs.y, bzs = bn.z + bs.z;
    const T brhoas = bxs * cphics - bzs * sphics;
    const T bphias = -bxs * sphics - bzs * cphics;
    const T brho_s = brhoas * dphisphi * xkappa;
    const T bphi_s = (bphias - rho * (brhoas * dphisrho)) * xkappa;
    return mk3<T>(brho_s * cphic - bphi_s * sphic, byas * dphisphi * xkappa, -brho_s * sphic - bphi_s * cphic);
}

// Shield harmonics (lib/ts05/birk_tot.c:379-403, lib/ts05/full_rc.c:697-721) with the
// per-epoch coefficient sums; pr = P[3], R[3]; fac multiplies each term when use_fac.
// FAST (the mixed path): the scales' reciprocals and roots from the per-epoch tables
// (inv = 1/P, 1/R; sqt = sqrt(1/P^2 + 1/R^2)) and the fast exp and sin/cos.
template <class T, bool FAST = false>
__device__ __forceinline__ v3<T> shield(const T *coef, const T *pr, const T *inv, const T *sqt, bool use_fac, T fac,
                                        T x, T y, T z) {
    static_assert(!FAST || Dm<T>::fp64, "the fast harmonic-sum functions are FP64");
    T gx = 0, gy = 0, gz = 0;
    T sz[3], cz[3];
#pragma unroll
    for (int k = 0; k < 3; k++) hsincos<T, FAST>(FAST ? z * inv[3 + k] : z / pr[3 + k], &sz[k], &cz[k]);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const T p = pr[j];
        T sy, cy;
        hsincos<T, FAST>(FAST ? y * inv[j] : y / p, &sy, &cy);
        const T pinv2 = FAST ? T(0.) : T(1.) / (p * p);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const T r = pr[3 + k];
            const T sq = FAST ? sqt[3 * j + k] : Dm<T>::sqrt(pinv2 + T(1.) / (r * r));
            const T e = hexp<T, FAST>(x * sq);
            const T c = coef[3 * j + k];
            if constexpr (FAST) {
                const T ip = inv[j], ir = inv[3 + k];
                if (use_fac) {
                    gx += (-sq * e * cy * sz[k] * fac) * c;
                    gy += (e * sy * sz[k] * ip * fac) * c;
                    gz += (-e * cy * cz[k] * ir * fac) * c;
                } else {
                    gx += (-sq * e * cy * sz[k]) * c;
                    gy += (e * sy * sz[k] * ip) * c;
                    gz += (-e * cy * cz[k] * ir) * c;
                }
            } else if (use_fac) {
                gx += (-sq * e * cy * sz[k] * fac) * c;
                gy += (e * sy * sz[k] / p * fac) * c;
                gz += (-e * cy * cz[k] / r * fac) * c;
            } else {
                gx += (-sq * e * cy * sz[k]) * c;
                gy += (e * sy * sz[k] / p) * c;
                gz += (-e * cy * cz[k] / r) * c;
            }
        }
    }
    return mk3<T>(gx, gy, gz);
}

// ---------------------------------------------------------------- ring currents

// lib/ts05/full_rc.c:297-323 (the two-loop vector potential at the mapped point)
template <class T, class S>
__device__ __forceinline__ S loops_aphi(S rhos, S zs2, T rrc, T dd) {
    const S rr = rrc + rhos;
    const S p = rr * rr + zs2 + dd * dd;
    const S xk2 = T(4.) * rrc * rhos / p;
    const S xk = fsqrt(xk2);
    const S xkrho12 = xk * fsqrt(rhos);
    const S xk2s = T(1.) - xk2;
    const S dlg = flog(T(1.) / xk2s);
    const S elk = T(1.38629436112) + xk2s * (T(0.09666344259) + xk2s * (T(0.03590092383) + xk2s * (T(0.03742563713) + xk2s * T(0.01451196212)))) +
                  dlg * (T(0.5) + xk2s * (T(0.12498593597) + xk2s * (T(0.06880248576) + xk2s * (T(0.03328355346) + xk2s * T(0.00441787012)))));
    const S ele = T(1.) + xk2s * (T(0.44325141463) + xk2s * (T(0.0626060122) + xk2s * (T(0.04757383546) + xk2s * T(0.01736506451)))) +
                  dlg * xk2s * (T(0.2499836831) + xk2s * (T(0.09200180037) + xk2s * (T(0.04069697526) + xk2s * T(0.00526449639))));
    return ((T(1.) - xk2 * T(0.5)) * elk - ele) / xkrho12;
}

// dipole-line mapping (alpha_s, gamma_s) -> (rho_s, z_s^2) (lib/ts05/full_rc.c:270-288)
//
// FP32: near the dipole axis (alpha_s -> 0) the reference's c = q - 4 g3 / (3 q) is a
// difference of two nearly equal cube roots, u - v with u^3 - v^3 = 2 A (A = alpha_s^2/2,
// v = (sqrt f - A)^(1/3) = 4 g3 / (3 u)), so it is formed as 2 A / (u^2 + u v + v^2);
// and sin(theta_s) as sqrt(alpha_s r_s) (a dipole line has sin^2 theta / r = alpha)
// instead of sqrt(1 - cos^2 theta_s).  Both are exact identities; in FP32 they avoid
// losing ~2 digits at sin(theta) ~ 0.04.
template <class T, class S>
__device__ __forceinline__ void dmap(S alpha_s, S gamma_s, S *rhos, S *zs2) {
    const S gammas2 = gamma_s * gamma_s;
    const S alsqh = (alpha_s * alpha_s) / T(2.);
    const S f = T(64.) / T(27.) * gammas2 + (alsqh * alsqh);
    const S q = fpow(fsqrt(f) + alsqh, T(1.) / T(3.));
    const S g3 = fpow(gammas2, T(1.) / T(3.));
    S c;
    if constexpr (Dm<T>::fp64) {
        c = q - T(4.) * g3 / (T(3.) * q);
        if (val(c) < T(0.)) c = c * T(0.);
    } else {
        const S v = T(4.) * g3 / (T(3.) * q);
        c = T(2.) * alsqh / (q * q + q * v + v * v);
    }
    const S g = fsqrt(c * c + T(4.) * g3);
    const S rs = T(4.) / ((fsqrt(T(2.) * g - c) + fsqrt(c)) * (g + c));
    const S costs = gamma_s * (rs * rs);
    S sints;
    if constexpr (Dm<T>::fp64) sints = fsqrt(T(1.) - costs * costs);
    else sints = fsqrt(alpha_s * rs);
    *rhos = rs * sints;
    const S zs = rs * costs;
    *zs2 = zs * zs;
}

// lib/ts05/full_rc.c:213-327
template <class T, class S>
__device__ __forceinline__ S ap_fn(S r, S sint, S cost) {
    bool prox = false;
    S sint1 = sint, cost1 = cost;
    if (val(sint1) < T(1.E-2)) {
        sint1 = sint1 * T(0.) + T(1.E-2);
        cost1 = cost1 * T(0.) + T(.99994999875);
        prox = true;
    }
    const S u1 = (r - T(3.746064740)) / T(1.508802177), u2 = (r - T(4.993638842)) / T(3.324180497);
    const S u3 = (r - T(2.969226745)) / T(2.243367377);
    const S c1 = cost1 / T(0.5873525737), c2 = cost1 / T(0.4368407663);
    const S arg1 = -(u1 * u1) - (c1 * c1), arg2 = -(u2 * u2) - (c2 * c2), arg3 = -(u3 * u3);
    const S e1 = val(arg1) >= T(-500.) ? fexp(arg1) : arg1 * T(0.);
    const S e2 = val(arg2) >= T(-500.) ? fexp(arg2) : arg2 * T(0.);
    const S e3 = val(arg3) >= T(-500.) ? fexp(arg3) : arg3 * T(0.);
    const S alpha = (sint1 * sint1) / r;
    const S alpha_s = alpha * (T(1.) + T(-0.2291904607) * e1 + T(0.1556236119) * e2 + T(0.1855957207) * e3);
    const S gamma_s = cost1 / (r * r);
    S rhos, zs2;
    dmap<T, S>(alpha_s, gamma_s, &rhos, &zs2);
    const S a1 = loops_aphi<T, S>(rhos, zs2, T(4.274684950), T(2.439528329));
    const S a2 = loops_aphi<T, S>(rhos, zs2, T(3.367557287), T(3.146382545));
    S ap = T(-456.5289941) * a1 + T(375.9055332) * a2;
    if (prox) ap = ap * sint / sint1;
    return ap;
}

// lib/ts05/full_rc.c:329-423
template <class T, class S>
__device__ __forceinline__ S apprc_fn(S r, S sint, S cost) {
    bool prox = false;
    S sint1 = sint, cost1 = cost;
    if (val(sint1) < T(1.E-2)) {
        sint1 = sint1 * T(0.) + T(1.E-2);
        cost1 = cost1 * T(0.) + T(.99994999875);
        prox = true;
    }
    const S alpha = (sint1 * sint1) / r;
    S gamma;
    if constexpr (Dm<T>::fp64) gamma = cost1 / Dm<T>::pow(r, 2.);
    else gamma = cost1 / (r * r);
    const S gdg1 = gamma / T(.04742939676), gdg4 = gamma / T(.01512963586);
    const S a4 = (alpha - T(.2647095287)) / T(.07091230197);
    const S arg1 = -(gdg1 * gdg1), arg2 = -(a4 * a4) - (gdg4 * gdg4);
    const S e1 = val(arg1) >= T(-500.) ? fexp(arg1) : arg1 * T(0.);
    const S e2 = val(arg2) >= T(-500.) ? fexp(arg2) : arg2 * T(0.);
    // pow(x, 2) of the reference is x*x except in FP64, where it is kept as pow()
    auto sq2 = [](S v) -> S {
        if constexpr (Dm<T>::fp64) return Dm<T>::pow(v, 2.);
        else return v * v;
    };
    const S alpha_s =
        alpha * (T(1.) + T(.3058309043) / fpow(T(1.) + sq2((alpha - T(.1817139853)) / T(.1257532909)), T(3.422509402)) * e1 +
                 T(-4.800458958) * (alpha - T(-.02845643596)) / fpow(T(1.) + sq2((alpha - T(-.02845643596)) / T(.2188114228)), T(2.545944574)) /
                     fpow(T(1.) + sq2(gamma / T(.00813272793)), T(.35868244)) +
                 T(103.1601001) * sq2(alpha - T(-.00764731187)) /
                     fpow(T(1.) + sq2((alpha - T(-.00764731187)) / T(.1046487459)), T(2.958863546)) /
                     fpow(T(1.) + sq2(gamma / T(.01172314188)), T(.4382872938)));
    const S gamma_s =
        gamma * (T(1.) + T(.01134908150) + T(14.51339943) * (alpha - T(.2647095287)) * e2 +
                 T(6.861329631) * (alpha - T(.1677400816)) / fpow(T(1.) + sq2((alpha - T(.1677400816)) / T(.04433648846)), T(.7665599464)) /
                     fpow(T(1.) + sq2(gamma / T(.05553741389)), T(.7277854652)));
    S rhos, zs2;
    dmap<T, S>(alpha_s, gamma_s, &rhos, &zs2);
    const S a1 = loops_aphi<T, S>(rhos, zs2, T(6.560486035), T(1.930711037));
    const S a2 = loops_aphi<T, S>(rhos, zs2, T(3.827208119), T(.7789990504));
    S ap = T(-80.11202281) * a1 + T(12.58246758) * a2;
    if (prox) ap = ap * sint / sint1;
    return ap;
}

template <int KIND, class T, class S>
__device__ __forceinline__ S aphi(S r, S s, S c) {
    if constexpr (KIND == 0) return ap_fn<T, S>(r, s, c);
    else return apprc_fn<T, S>(r, s, c);
}

// B = curl(A_phi e_phi): lib/ts05/full_rc.c:161-211 (rc_symm) and :329-368 (prc_symm).
// FP64: the reference's central differences; FP32: dual numbers in r and theta.
template <int KIND, class T>
__device__ __forceinline__ v3<T> symm(T x, T y, T z) {
    const T ds = T(1.E-2), dc = T(0.99994999875), d = T(1.E-4), drd = T(5E3);
    const T rho2 = x * x + y * y;
    const T r2 = rho2 + z * z;
    const T r = Dm<T>::sqrt(r2);
    const T rp = r + d, rm = r - d;
    const T sint = Dm<T>::sqrt(rho2) / r, cost = z / r;
    T fxy, bz;
    if (sint < ds) {
        T a, dardr;
        if constexpr (Dm<T>::fp64) {
            a = aphi<KIND, T, T>(r, ds, dc) / ds;
            dardr = (rp * aphi<KIND, T, T>(rp, ds, dc) - rm * aphi<KIND, T, T>(rm, ds, dc)) * drd;
        } else {
            const Dual<T> ra = aphi<KIND, T, Dual<T>>(dl(r, T(1)), dl(ds, T(0)), dl(dc, T(0)));
            a = ra.v / ds;
            dardr = ra.v + r * ra.d;              // d(r A)/dr
        }
        fxy = z * (T(2.) * a - dardr) / (r * r2);
        bz = (T(2.) * a * (cost * cost) + dardr * (sint * sint)) / r;
    } else {
        const T th = Dm<T>::atan2(sint, cost);
        T br, bt;
        if constexpr (Dm<T>::fp64) {
            const T tp = th + d, tm = th - d;
            const T stp = Dm<T>::sin(tp), stm = Dm<T>::sin(tm), ctp = Dm<T>::cos(tp), ctm = Dm<T>::cos(tm);
            br = (stp * aphi<KIND, T, T>(r, stp, ctp) - stm * aphi<KIND, T, T>(r, stm, ctm)) / (r * sint) * drd;
            bt = (rm * aphi<KIND, T, T>(rm, sint, cost) - rp * aphi<KIND, T, T>(rp, sint, cost)) / r * drd;
        } else {
            // d/dtheta of sin(theta) A(r, sin theta, cos theta), d/dr of r A(r, ...)
            const Dual<T> st = dl(sint, cost), ct = dl(cost, -sint);
            const Dual<T> at = aphi<KIND, T, Dual<T>>(dl(r, T(0)), st, ct);
            br = (cost * at.v + sint * at.d) / (r * sint);
            const Dual<T> ar = aphi<KIND, T, Dual<T>>(dl(r, T(1)), dl(sint, T(0)), dl(cost, T(0)));
            bt = -(ar.v + r * ar.d) / r;
        }
        fxy = (br + bt * cost / sint) / r;
        bz = br * cost - bt * sint;
    }
    return mk3<T>(fxy * x, fxy * y, bz);
}

// lib/ts05/full_rc.c:689-695
template <class T, class S>
__device__ __forceinline__ void ffs(S a, T a0, T da, S *f, S *fa, S *fs) {
    const T da2 = da * da;
    const S p = a + a0, m = a - a0;
    const S sq1 = fsqrt(p * p + da2), sq2 = fsqrt(m * m + da2);
    *fa = T(2.) / (sq1 + sq2);
    *f = *fa * a;
    *fs = T(0.5) * (sq1 + sq2) / (sq1 * sq2) * (T(1.) - *f * *f);
}

// lib/ts05/full_rc.c:482-593
template <class T, class S>
__device__ __forceinline__ S br_prc_q(S r, S sint, S cost) {
    const T A[18] = {T(-21.2666329), T(32.24527521), T(-6.062894078), T(7.515660734), T(233.7341288),
                     T(-227.1195714), T(8.483233889), T(16.80642754), T(-24.63534184), T(9.067120578),
                     T(-1.052686913), T(-12.08384538), T(18.61969572), T(-12.71686069), T(47017.35679),
                     T(-50646.71204), T(7746.058231), T(1.531069371)};
    const S sint2 = sint * sint, cost2 = cost * cost, sc = sint * cost;
    const S alpha = sint2 / r, gamma = cost / (r * r);
    S f, fa, fs, d[18];
    ffs<T, S>(alpha, T(.1417519429), T(.6388013110E-02), &f, &fa, &fs);
    d[0] = sc * fpow(f, T(2.318824273)) / (fpow(r / T(5.303934488), T(4.213397467)) + T(1.));
    d[1] = d[0] * cost2;
    ffs<T, S>(alpha, T(.1401142771), T(.2306094179E-01), &f, &fa, &fs);
    d[2] = sc * fpow(fs, T(.7955534018)) / (fpow(r / T(3.462235072), T(2.568743010)) + T(1.));
    d[3] = d[2] * cost2;
    ffs<T, S>(alpha, T(.1485233485), T(.2319676273E-01), &f, &fa, &fs);
    d[4] = sc * (fpow(alpha, T(3.477425908))) * (fpow(fs, T(1.922155110))) / (fpow(r / T(7.830223587), T(8.492933868)) + T(1.));
    d[5] = d[4] * cost2;
    const S a4 = (alpha - T(.1295221828)) / T(.01753008801), g1 = gamma / T(.01125504083);
    S arga = a4 * a4 + T(1.), argg = T(1.) + g1 * g1;
    d[6] = sc / arga / argg;
    d[7] = d[6] / arga;
    d[8] = d[7] / arga;
    d[9] = d[8] / arga;
    const S a5 = (alpha - T(.1811846095)) / T(.04841237481), g2 = gamma / T(.01981805097);
    arga = a5 * a5 + T(1.);
    argg = T(1.) + g2 * g2;
    d[10] = sc / arga / argg;
    d[11] = d[10] / arga;
    d[12] = d[11] / arga;
    d[13] = d[12] / arga;
    const S r4 = r * r * r * r;
    const T c1 = T(6.557801891), c2 = T(6.348576071), c3 = T(5.744436687);
    d[14] = sc / (r4 + c1 * c1 * c1 * c1);
    d[15] = sc / (r4 + c2 * c2 * c2 * c2) * cost2;
    d[16] = sc / (r4 + c3 * c3 * c3 * c3) * (cost2 * cost2);
    ffs<T, S>(alpha, T(.2265212965), T(.1301957209), &f, &fa, &fs);
    const S rd = (r - T(1.2)) / T(.5654023158);
    d[17] = sc * fs / (T(1.) + rd * rd);
    S s = A[0] * d[0];
#pragma unroll
    for (int k = 1; k < 18; k++) s = s + A[k] * d[k];
    return s;
}

// lib/ts05/full_rc.c:595-687
template <class T, class S>
__device__ __forceinline__ S bt_prc_q(S r, S sint, S cost) {
    const T A[17] = {T(12.74640393), T(-7.516393516), T(-5.476233865), T(3.212704645), T(-59.10926169),
                     T(46.62198189), T(-.01644280062), T(.1234229112), T(-.08579198697), T(.01321366966),
                     T(.8970494003), T(9.136186247), T(-38.19301215), T(21.73775846), T(-410.0783424),
                     T(-69.90832690), T(-848.8543440)};
    const S sint2 = sint * sint, cost2 = cost * cost;
    const S alpha = sint2 / r, gamma = cost / (r * r);
    S f, fa, fs, d[17];
    ffs<T, S>(alpha, T(.2071721360), T(.05030555417), &f, &fa, &fs);
    d[0] = fpow(f, T(1.243288286)) / (fpow(r / T(7.471332374), T(3.180533613)) + T(1.));
    d[1] = d[0] * cost2;
    ffs<T, S>(alpha, T(.1568504222), T(.02092910682), &f, &fa, &fs);
    d[2] = fpow(fa, T(1.376743507)) / fpow(r, T(1.985148197));
    d[3] = d[2] * cost2;
    ffs<T, S>(alpha, T(.1701395257), T(.1019870070), &f, &fa, &fs);
    d[4] = fpow(fs, T(.3157139940)) * fpow(alpha, T(1.056309517)) / (fpow(r / T(6.293740981), T(5.671824276)) + T(1.));
    d[5] = d[4] * cost2;
    ffs<T, S>(gamma, T(0.), T(.01040696080), &f, &fa, &fs);
    const S a4 = (alpha - T(.1280772299)) / T(.02189060799);
    const S fcc = (T(1.) + a4 * a4);
    d[6] = T(1.) / fcc * fs;
    d[7] = d[6] / fcc;
    d[8] = d[7] / fcc;
    d[9] = d[8] / fcc;
    const S a5 = (alpha - T(.1648265607)) / T(.04701592613), g2 = gamma / T(.01526400086);
    const S arg = T(1.) + a5 * a5;
    d[10] = T(1.) / arg / (T(1.) + g2 * g2);
    d[11] = d[10] / arg;
    d[12] = d[11] / arg;
    d[13] = d[12] / arg;
    const S r4 = r * r * r * r;
    const T c1 = T(12.88384229), c2 = T(3.361775101), c3 = T(23.44173897);
    d[14] = T(1.) / (r4 + c1 * c1);
    d[15] = cost2 / (r4 + c2 * c2);
    d[16] = (cost2 * cost2) / (r4 + c3 * c3);
    S s = A[0] * d[0];
#pragma unroll
    for (int k = 1; k < 17; k++) s = s + A[k] * d[k];
    return s;
}

// lib/ts05/full_rc.c:425-480.  FP64: central differences; FP32: dual numbers.
template <class T>
__device__ __forceinline__ v3<T> prc_quad(T x, T y, T z) {
    const T d = T(1.E-4), dd = T(2.E-4), ds = T(1.E-2), dc = T(0.99994999875);
    const T rho2 = x * x + y * y;
    const T r = Dm<T>::sqrt(rho2 + z * z);
    const T rho = Dm<T>::sqrt(rho2);
    const T sint = rho / r, cost = z / r;
    const T rp = r + d, rm = r - d;
    const bool off_axis = sint > ds;
    const T st = off_axis ? sint : ds;
    const T ct = off_axis ? cost : (z < T(0.) ? -dc : dc);
    T br, bt, dbrr, dbtt;
    if constexpr (Dm<T>::fp64) {
        const T th = Dm<T>::atan2(st, ct);
        const T tp = th + d, tm = th - d;
        const T stp = Dm<T>::sin(tp), ctp = Dm<T>::cos(tp), stm = Dm<T>::sin(tm), ctm = Dm<T>::cos(tm);
        br = br_prc_q<T, T>(r, st, ct);
        bt = bt_prc_q<T, T>(r, st, ct);
        dbrr = (br_prc_q<T, T>(rp, st, ct) - br_prc_q<T, T>(rm, st, ct)) / dd;
        dbtt = (bt_prc_q<T, T>(r, stp, ctp) - bt_prc_q<T, T>(r, stm, ctm)) / dd;
    } else {
        const Dual<T> brr = br_prc_q<T, Dual<T>>(dl(r, T(1)), dl(st, T(0)), dl(ct, T(0)));
        const Dual<T> btt = bt_prc_q<T, Dual<T>>(dl(r, T(0)), dl(st, ct), dl(ct, -st));
        br = brr.v;
        bt = btt.v;
        dbrr = brr.d;
        dbtt = btt.d;
    }
    if (off_axis) {
        const T cphi = x / rho, sphi = y / rho;
        return mk3<T>(sint * (br + (br + r * dbrr + dbtt) * (sphi * sphi)) + cost * bt,
                      -sint * sphi * cphi * (br + r * dbrr + dbtt), (br * cost - bt * sint) * cphi);
    }
    const T fcxy = r * dbrr + dbtt;
    const T rst = r * st;
    const T rst2 = rst * rst;
    const T y2 = y * y;
    return mk3<T>((br * (x * x + T(2.) * y2) + fcxy * y2) / rst2 + bt * cost, -(br + fcxy) * x * y / rst2,
                  (br * cost / st - bt) * x / r);
}

// lib/ts05/full_rc.c:4-159: symmetric and partial ring currents with their shields,
// already times their amplitudes and summed: a_src * src + a_prc * prc.
template <class T>
__device__ __forceinline__ void full_rc(const TS05Params<T> &P, const TS05Params<double> &Pd, T x, T y, T z,
                                        v3<T> *src, v3<T> *prc) {
    v3<T> hs = symm<0, T>(x / P.sc_sy, y / P.sc_sy, z / P.sc_sy);
    pin(hs);
    COR_SEQ();
    const T xta = x / P.sc_pr, yta = y / P.sc_pr, zta = z / P.sc_pr;
    v3<T> ha = symm<1, T>(xta, yta, zta);
    pin(ha);
    COR_SEQ();
    const T cp = P.cphi, sp = P.sphi;
    v3<T> hq = prc_quad(xta * cp - yta * sp, xta * sp + yta * cp, zta);
    pin(hq);
    COR_SEQ();
    const T bxq = hq.x * cp + hq.y * sp;
    const T byq = -hq.x * sp + hq.y * cp;
    v3<T> fs = to3<T>(shield<double, !Dm<T>::fp64>(Pd.src, CT2(k_ts05_csy) + 72, Pd.shinv[2], Pd.shsq[2], true,
                                                     Pd.fac_sy, (double)x, (double)y, (double)z));
    pin(fs);
    COR_SEQ();
    const v3<T> fp = to3<T>(shield<double, !Dm<T>::fp64>(Pd.prc, CT2(k_ts05_cpr) + 72, Pd.shinv[3], Pd.shsq[3], true,
                                                           Pd.fac_pr, (double)x, (double)y, (double)z));
    *src = mk3<T>(hs.x + fs.x, hs.y + fs.y, hs.z + fs.z);
    *prc = mk3<T>((ha.x + bxq) + fp.x, (ha.y + byq) + fp.y, (ha.z + hq.z) + fp.z);
}

}  // namespace ts05

// lib/ts05/ts05.c:7-134: the whole model at zero tilt.
// Pd: the FP64 constants, used by both paths for the Cartesian-harmonic shield and
// warping sums, whose terms cancel by 4-6 orders of magnitude (coefficients ~1e4-1e5,
// e^{x s} up to ~1e4 on the dayside): those sums stay FP64 in the mixed path too.
template <class T>
__device__ __forceinline__ v3<T> ts05_field(const TS05Params<T> &P, const TS05Params<double> &Pd, T x, T y, T z) {
    using namespace ts05;
    const T xx = x * P.xappa, yy = y * P.xappa, zz = z * P.xappa;
    T xmxm = P.am + x - P.x0;
    if (xmxm < T(0)) xmxm = T(0.);
    const T axx0 = xmxm * xmxm;
    const T asq = P.am * P.am;
    const T rho2 = y * y + z * z;
    const T aro = (asq + rho2) + axx0;
    const T sigma = Dm<T>::sqrt((aro + Dm<T>::sqrt(aro * aro - T(4.) * asq * axx0)) / (T(2.) * asq));
    if (!(sigma < T(1.1960 + 0.005))) {
        const v3<T> q = ts05::dipole(x, y, z);
        return mk3<T>(-q.x, P.oimfy - q.y, P.oimfz - q.z);
    }

    // BB = A0 cf + tamp1 t1 + tamp2 t2 + a_src src + a_prc prc + a_r11 r11 + a_r21 r21 +
    // A19 imf (lib/ts05/ts05.c:106-108), accumulated left to right as each term is made
    v3<T> bb;
    {
        const v3<T> cf = shlcar3x3(xx, yy, zz);
        bb = mk3<T>(T(1.00000) * (cf.x * P.xappa3), T(1.00000) * (cf.y * P.xappa3), T(1.00000) * (cf.z * P.xappa3));
        pin(bb);
    }
    COR_SEQ();
    {
        v3<T> t1, t2;
        warped(P, Pd, xx, yy, zz, &t1, &t2);
        bb = mk3<T>(bb.x + P.tamp1 * t1.x + P.tamp2 * t2.x, bb.y + P.tamp1 * t1.y + P.tamp2 * t2.y,
                    bb.z + P.tamp1 * t1.z + P.tamp2 * t2.z);
        pin(bb);
    }
    COR_SEQ();
    v3<T> r11, r21;
    {
        v3<T> f = birk_1n2(P, 0, xx, yy, zz);
        pin(f);
        COR_SEQ();
        const v3<T> h = to3<T>(shield<double, !Dm<T>::fp64>(Pd.r11, CT2(k_ts05_sh11) + 72, Pd.shinv[0], Pd.shsq[0],
                                                            false, 1., (double)xx, (double)yy, (double)zz));
        r11 = mk3<T>(f.x + h.x, f.y + h.y, f.z + h.z);
        pin(r11);
    }
    COR_SEQ();
    {
        v3<T> f = birk_1n2(P, 1, xx, yy, zz);
        pin(f);
        COR_SEQ();
        const v3<T> h = to3<T>(shield<double, !Dm<T>::fp64>(Pd.r21, CT2(k_ts05_sh21) + 72, Pd.shinv[1], Pd.shsq[1],
                                                            false, 1., (double)xx, (double)yy, (double)zz));
        r21 = mk3<T>(f.x + h.x, f.y + h.y, f.z + h.z);
        pin(r21);
    }
    COR_SEQ();
    {
        v3<T> src, prc;
        full_rc(P, Pd, xx, yy, zz, &src, &prc);
        bb = mk3<T>(bb.x + P.a_src * src.x + P.a_prc * prc.x, bb.y + P.a_src * src.y + P.a_prc * prc.y,
                    bb.z + P.a_src * src.z + P.a_prc * prc.z);
    }
    // A[19] * (0, By, Bz) is the product oimfy/oimfz already hold (A[19] * By == By * A[19])
    bb = mk3<T>(bb.x + P.a_r11 * r11.x + P.a_r21 * r21.x + T(0.),
                bb.y + P.a_r11 * r11.y + P.a_r21 * r21.y + P.oimfy,
                bb.z + P.a_r11 * r11.z + P.a_r21 * r21.z + P.oimfz);
    if (sigma < T(1.1960 - 0.005)) return bb;
    const v3<T> q = ts05::dipole(x, y, z);
    const T fint = T(0.5) * (T(1.) - (sigma - T(1.1960)) / T(0.005));
    const T fext = T(0.5) * (T(1.) + (sigma - T(1.1960)) / T(0.005));
    return mk3<T>((bb.x + q.x) * fint - q.x, (bb.y + q.y) * fint + P.oimfy * fext - q.y,
                  (bb.z + q.z) * fint + P.oimfz * fext - q.z);
}

}  // namespace cor
