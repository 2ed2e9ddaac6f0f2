// igrf_dev.hpp -- IGRF spherical-harmonic synthesis for one lane (lib/internal/internal.c:182-275).
//
// GEOPACK-style recursion: for each order m the Schmidt-scaled P_n^m and dP_n^m/dtheta
// are carried up in degree with the REC table, cos/sin(m phi) are advanced by rotation,
// and the expansion stops at the radius-dependent degree NM = min(13, 3 + 30/((int)R+2)).
//
// GPU layout: G/H/REC of the epoch are staged in LDS once per workgroup (the kernels
// are persistent, so this is a one-time copy); the (m, n) loops run to the
// wavefront's largest truncation (found with ballots, so only active lanes count)
// with wave-uniform trip counts, so every lane reads the same coefficient in the same
// cycle (an LDS broadcast) and lanes whose own truncation is lower mask their
// accumulation.  The next degree's coefficients are read one iteration ahead so the
// LDS latency overlaps the recursion arithmetic.  The (1/R)^(n+1) powers are
// produced by the same left-to-right product chain as the reference's A[] table,
// without a register array.
#pragma once
#include "devmath.hpp"

namespace cor {

// Coefficient tables as staged in LDS: 105 entries zero-padded so the prefetches of the
// last degrees (one degree ahead here, two in the cooperative evaluator: index at most
// 133) stay inside the array.
constexpr int IGRF_LDS_N = 136;

template <class T>
__device__ __forceinline__ v3<T> igrf_geo(const T *__restrict__ G, const T *__restrict__ H,
                                          const T *__restrict__ REC, T x, T y, T z) {
    const T rho2 = x * x + y * y;
    const T r = Dm<T>::sqrt(rho2 + z * z);
    const T c = z / r;
    const T rho = Dm<T>::sqrt(rho2);
    const T s = rho / r;
    const bool polar = s < T(1.E-5);
    const T cf = polar ? T(1.) : x / rho;
    const T sf = polar ? T(0.) : y / rho;
    const T inv_r = T(1.) / r;

    int nm = 3 + 30 / ((int)r + 2);
    if (nm > 13) nm = 13;
    const int klane = nm + 1;
    // largest truncation among the active lanes (4..14)
    int kw = 14;
    while (kw > 4 && !__any(klane >= kw)) kw--;
    kw = __builtin_amdgcn_readfirstlane(kw);

    T am = inv_r * inv_r;       // (1/R)^(m+1), the first power of each order's column
    T pmm = T(1.), dmm = T(0.); // sectoral P and dP carried between orders
    T cm = T(0.), sm = T(1.);   // (sin m phi, cos m phi) by rotation
    T br = T(0.), bt = T(0.), bf = T(0.);
    for (int m = 1; m <= kw; m++) {
        if (m != 1) {
            const T w = cm;
            cm = w * cf + sm * sf;
            sm = sm * cf - w * sf;
        }
        T q = pmm, dq = dmm, q2 = T(0.), dq2 = T(0.), bphi = T(0.);
        T an = am;
        int idx = m * (m - 1) / 2 + m - 1;
        T g = G[idx], h = H[idx], rec = REC[idx];
        for (int n = m; n <= kw; n++) {
            // prefetch degree n+1 (the LDS tables are padded, so idx + n stays in range)
            const T gn = G[idx + n], hn = H[idx + n], recn = REC[idx + n];
            const bool on = n <= klane;
            const T w = g * sm + h * cm;
            const T bn = an * n;
            if (on) {
                br = br + bn * w * q;
                bt = bt - an * w * dq;
                if (m != 1) {
                    const T qq = polar ? dq : q;
                    bphi = bphi + an * (g * cm - h * sm) * qq;
                }
            }
            const T dnext = c * dq - s * q - rec * dq2;
            const T pnext = c * q - rec * q2;
            dq2 = dq;
            q2 = q;
            dq = dnext;
            q = pnext;
            an = an * inv_r;
            idx += n;   // n(n+1)/2 + m - 1 for the next degree
            g = gn;
            h = hn;
            rec = recn;
        }
        dmm = s * dmm + c * pmm;
        pmm = s * pmm;
        if (m != 1 && m <= klane) {
            bphi = bphi * (m - 1);
            bf = bf + bphi;
        }
        am = am * inv_r;
    }
    T bp;
    if (polar) {
        if (c < T(0.)) bf = -bf;
        bp = bf;
    } else {
        bp = bf / s;
    }
    const T he = br * s + bt * c;
    return mk3<T>(he * cf - bp * sf, he * sf + bp * cf, br * c - bt * s);
}

}  // namespace cor
