// sh_dev.hpp -- generic spherical-harmonic synthesis (the CHAOS-form crustal/core
// model) for one lane: lib/chaos/chaos.c:118-227, GEO position [Re] -> GEO field [nT].
//
// The reference fills P_n^m / dP_n^m row by row (n outer) and sums with a fresh
// sin/cos(m phi) per term.  Here the order m is the outer loop and each lane carries
// one degree column: P_m^m from the sectoral chain, then the three-term recursion in n.
// Every P and dP element is produced by the reference's own expression, so the FP64
// values are identical; only the order of the final sum changes (the m-outer order).
// G/H tables are [nm][nm] (row n, column m) in device memory, read with
// wavefront-uniform indices (scalar or broadcast loads).
//   * FP64: theta/phi by atan2 and sin/cos(m phi) per order, as the reference.
//   * FP32: sin(theta), cos(theta) as rho/r, z/r and cos/sin(m phi) by rotation.
#pragma once
#include "devmath.hpp"

namespace cor {

constexpr double SH_RE_KM = 6371.2;   // lib/chaos/chaos.h:6

template <class T>
__device__ __forceinline__ v3<T> sh_geo(const T *__restrict__ G, const T *__restrict__ H, int nm, T x, T y, T z) {
    const T rho2 = x * x + y * y;
    const T rr = Dm<T>::sqrt(rho2 + z * z);
    const T c = z / rr;
    const T r = rr * T(SH_RE_KM);
    T theta = T(0.), phi = T(0.), s = T(0.), cf = T(1.), sf = T(0.);
    if (rho2 != T(0.)) {
        const T rho = Dm<T>::sqrt(rho2);
        s = rho / rr;
        cf = x / rho;
        sf = y / rho;
        if constexpr (Dm<T>::fp64) {
            phi = Dm<T>::atan2(y, x);
            theta = Dm<T>::atan2(rho, z);
            if (phi < T(0.)) phi += T(6.283185307179586);
        }
    } else if (z < T(0.)) {
        theta = T(3.141592653589793);
    }
    T st, ct;
    if constexpr (Dm<T>::fp64) {
        st = Dm<T>::sin(theta);
        ct = Dm<T>::cos(theta);
    } else {
        st = s;
        ct = rho2 != T(0.) ? c : (z < T(0.) ? T(-1.) : T(1.));
    }
    const T q = T(SH_RE_KM) / r;
    T br = T(0.), bt = T(0.), bp = T(0.);
    T pmm = T(1.), dpmm = T(0.);     // P_m^m, dP_m^m
    T arm = q;                       // (a/r)^(m+1)
    T cm = T(1.), sm = T(0.);        // cos/sin(m phi), FP32 rotation
    for (int m = 0; m < nm; m++) {
        if (m == 1) {
            pmm = st;
            dpmm = ct;
        } else if (m >= 2) {
            // lib/chaos/chaos.c:182-186 (sectoral term from the previous sectoral one)
            const T n2 = (T)(m * 2), n21 = n2 - T(1.);
            const T kd = Dm<T>::sqrt(n21 / n2);
            const T kp = kd * pmm;
            const T dnext = ct * kp + st * kd * dpmm;
            pmm = st * kp;
            dpmm = dnext;
        }
        T smp, cmp;
        if constexpr (Dm<T>::fp64) {
            const T mphi = (T)m * phi;
            smp = Dm<T>::sin(mphi);
            cmp = Dm<T>::cos(mphi);
        } else {
            if (m >= 1) {
                const T cn = cm * cf - sm * sf;
                sm = sm * cf + cm * sf;
                cm = cn;
            }
            smp = sm;
            cmp = cm;
        }
        // degree column n = m .. nm-1 (lib/chaos/chaos.c:187-196 and :198-214)
        T p = pmm, dp = dpmm, p1 = T(0.), dp1 = T(0.);   // P_n^m, and P_(n-1)^m
        T arn = arm;
        const T fm = (T)m;
        for (int n = m; n < nm; n++) {
            if (n > m) {
                T pn, dpn;
                if (n == 1) {                 // m = 0, n = 1 is set explicitly (:172-176)
                    pn = ct;
                    dpn = -st;
                } else {
                    const T nn = (T)n * n;
                    const T mm = (T)(m * m);
                    const T den = nn - mm;
                    const T k1 = (T)(n * 2 - 1) / Dm<T>::sqrt(den);
                    const T k2 = Dm<T>::sqrt(((T)((n - 1) * (n - 1)) - mm) / den);
                    pn = k1 * ct * p - k2 * p1;
                    dpn = k1 * (-st * p + ct * dp) - k2 * dp1;
                }
                p1 = p;
                dp1 = dp;
                p = pn;
                dp = dpn;
                arn = arn * q;
            }
            const T gg = G[n * nm + m], hh = H[n * nm + m];
            const T t = gg * cmp + hh * smp;
            const T darn = (T)(-(n + 1)) * arn / r;
            br += t * p * darn;
            bt += t * dp * arn;
            bp += (-gg * fm * smp + hh * fm * cmp) * p * arn;
        }
        arm = arm * q;
    }
    bt = -bt * T(SH_RE_KM) / r;
    bp = -bp * T(SH_RE_KM) / (r * st);
    br = -br * T(SH_RE_KM);
    const T he = br * s + bt * c;
    return mk3<T>(he * cf - bp * sf, he * sf + bp * cf, br * c - bt * s);
}

}  // namespace cor
