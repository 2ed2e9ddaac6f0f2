// tracer_dev.hpp -- the persistent cutoff tracer: one back-traced particle per lane.
//
// Per lane this is src/trajectory.c:12-153 (trajectory_simulation_cutoff's body) with
// calculate_position_and_speed (:275-356) inlined.  Per wavefront it is a loop of
//   1. refill   -- lanes whose trajectory ended take the next ids of a global queue
//                  (one atomic per wavefront per refill, ranks by mbcnt);
//   2. field    -- every live lane evaluates the field at its current point
//                  (converged: all lanes are in the same phase);
//   3. step     -- each lane repeats RK attempts until one is accepted or the
//                  trajectory terminates (divergent, but only the cheap RK + tests).
// A wavefront retires when the queue is drained and all its lanes have ended, so the
// grid drains on its own.
//
// Precision variants:
//   PREC_FP64  -- everything in double, the reference's operation order (compiled
//                 without FMA contraction): the parity path.
//   PREC_MIXED -- field models and RK stage arithmetic in FP32; position, velocity,
//                 time and path length stay FP64 and absorb the FP32 increments;
//                 step-size control (period, h, turn angle), the escape / re-entry
//                 tests and the exit/asymptotic geometry stay FP64.
#pragma once
#include "dev_types.hpp"
#include "igrf_dev.hpp"
#include "t96_dev.hpp"
#include "coop_dev.hpp"
#include "sh_dev.hpp"

namespace cor {

// ---------------------------------------------------------------- epoch constants

// What the field evaluators read per step: the IGRF tables from LDS, the rotation rows
// and the T96 constants from registers.  Built once per workgroup before the
// persistent loop (stage_epoch), so the step does no global-memory reads.
template <class FT>
struct EpochView {
    const FT *G, *H, *REC;
    double A[9];
    float Af[9];
    T96Params<FT> t96;
    const TS05Params<FT> *ts05;          // in HBM (read with scalar loads)
    const TS05Params<double> *ts05d;     // FP64 constants of the shield sums
    const FT *shG, *shH;                 // SH model tables in HBM (uniform indices)
    int sh_nm;
    const float *icon;                   // T96 interconnection series table (mixed path)
};
__device__ __forceinline__ const double *sh_tab(const EpochDev *E, int k, double) { return k ? E->shH : E->shG; }
__device__ __forceinline__ const float *sh_tab(const EpochDev *E, int k, float) { return k ? E->shHf : E->shGf; }
__device__ __forceinline__ const TS05Params<double> *ts05_of(const EpochDev *E, double) { return &E->ts05d; }
__device__ __forceinline__ const TS05Params<float> *ts05_of(const EpochDev *E, float) { return &E->ts05f; }

__device__ __forceinline__ const T96Params<double> &t96_of(const EpochDev *E, double) { return E->t96d; }
__device__ __forceinline__ const T96Params<float> &t96_of(const EpochDev *E, float) { return E->t96f; }
__device__ __forceinline__ const CoopTabT<double> *coop_tab(const EpochDev *E, double) { return &E->coopd; }
__device__ __forceinline__ const CoopTabT<float> *coop_tab(const EpochDev *E, float) { return &E->coop; }
// the GSM -> GEO rotation in the field type (the cooperative evaluator's last step)
__device__ __forceinline__ const double *rot_rows(const EpochView<double> &E) { return E.A; }
__device__ __forceinline__ const float *rot_rows(const EpochView<float> &E) { return E.Af; }
__device__ __forceinline__ const double *igrf_tab(const EpochDev *E, int k, double) {
    return k == 0 ? E->G : (k == 1 ? E->H : E->REC);
}
__device__ __forceinline__ const float *igrf_tab(const EpochDev *E, int k, float) {
    return k == 0 ? E->Gf : (k == 1 ? E->Hf : E->RECf);
}

// Copy the epoch's IGRF tables into the workgroup's LDS (zero padding) and the small
// constants into the view.  Called by every thread of the workgroup.
template <class FT>
__device__ __forceinline__ EpochView<FT> stage_epoch(const EpochDev *__restrict__ E, FT *sg, FT *sh, FT *srec) {
    const FT *g = igrf_tab(E, 0, FT(0)), *h = igrf_tab(E, 1, FT(0)), *rc = igrf_tab(E, 2, FT(0));
    for (int k = threadIdx.x; k < IGRF_LDS_N; k += blockDim.x) {
        sg[k] = k < NCOEF ? g[k] : FT(0);
        sh[k] = k < NCOEF ? h[k] : FT(0);
        srec[k] = k < NCOEF ? rc[k] : FT(0);
    }
    __syncthreads();
    EpochView<FT> V;
    V.G = sg;
    V.H = sh;
    V.REC = srec;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        V.A[k] = E->A[k];
        V.Af[k] = E->Af[k];
    }
    V.t96 = t96_of(E, FT(0));
    V.ts05 = ts05_of(E, FT(0));
    V.ts05d = &E->ts05d;
    V.shG = sh_tab(E, 0, FT(0));
    V.shH = sh_tab(E, 1, FT(0));
    V.sh_nm = E->sh_nm;
    V.icon = &E->coop.icon[0][0];
    return V;
}

// The view of one epoch block read straight from device memory, for the multi-epoch
// tracer whose lanes may each be at another epoch: the IGRF tables and the TS05 and SH
// constants through pointers into the block (vector loads that the caches serve), the
// rotation rows loaded.  Its T96 constants are not set: that tracer has no T96 model set.
template <class FT>
__device__ __forceinline__ EpochView<FT> global_view(const EpochDev *__restrict__ E) {
    EpochView<FT> V;
    V.G = igrf_tab(E, 0, FT(0));
    V.H = igrf_tab(E, 1, FT(0));
    V.REC = igrf_tab(E, 2, FT(0));
#pragma unroll
    for (int k = 0; k < 9; k++) {
        V.A[k] = E->A[k];
        V.Af[k] = E->Af[k];
    }
    V.ts05 = ts05_of(E, FT(0));
    V.ts05d = &E->ts05d;
    V.shG = sh_tab(E, 0, FT(0));
    V.shH = sh_tab(E, 1, FT(0));
    V.sh_nm = E->sh_nm;
    V.icon = &E->coop.icon[0][0];
    return V;
}

// ---------------------------------------------------------------- field aggregates

__device__ __forceinline__ v3<double> rot_geo_to_gsm(const double *A, v3<double> g) {
    return mk3(A[0] * g.x + A[1] * g.y + A[2] * g.z, A[3] * g.x + A[4] * g.y + A[5] * g.z,
               A[6] * g.x + A[7] * g.y + A[8] * g.z);
}
__device__ __forceinline__ v3<double> rot_gsm_to_geo(const double *A, v3<double> s) {
    return mk3(A[0] * s.x + A[3] * s.y + A[6] * s.z, A[1] * s.x + A[4] * s.y + A[7] * s.z,
               A[2] * s.x + A[5] * s.y + A[8] * s.z);
}
__device__ __forceinline__ v3<float> rot_gsm_to_geo_f(const float *A, v3<float> s) {
    return mk3(A[0] * s.x + A[3] * s.y + A[6] * s.z, A[1] * s.x + A[4] * s.y + A[7] * s.z,
               A[2] * s.x + A[5] * s.y + A[8] * s.z);
}

// src/models.c:194-207 with the adapters :146-192, in the reference's order: GSM [m]
// -> Re, each model adds its GSM field (the SH model round-trips through GEO), the
// sum is returned in GEO [nT].
template <int INT, int EXT, int CRU>
__device__ __forceinline__ v3<double> field_fp64(const EpochView<double> &E, v3<double> gsm_m) {
    const v3<double> g = mk3(gsm_m.x / K_RE, gsm_m.y / K_RE, gsm_m.z / K_RE);
    v3<double> tot = mk3(0., 0., 0.);
    if (INT == INT_IGRF) {
        const v3<double> geo = rot_gsm_to_geo(E.A, g);
        const v3<double> bg = igrf_geo<double>(E.G, E.H, E.REC, geo.x, geo.y, geo.z);
        const v3<double> bs = rot_geo_to_gsm(E.A, bg);
        tot.x += bs.x;
        tot.y += bs.y;
        tot.z += bs.z;
    }
    if (EXT == EXT_T96) {
        const v3<double> bs = t96_field<double>(E.t96, g.x, g.y, g.z);
        tot.x += bs.x;
        tot.y += bs.y;
        tot.z += bs.z;
    }
    if (EXT == EXT_TS05) {
        const v3<double> bs = ts05_field<double>(*E.ts05, *E.ts05d, g.x, g.y, g.z);
        tot.x += bs.x;
        tot.y += bs.y;
        tot.z += bs.z;
    }
    if (CRU == CRU_SH) {
        // src/models.c:179-192: GSM -> GEO, synthesis, GEO -> GSM, accumulate
        const v3<double> geo = rot_gsm_to_geo(E.A, g);
        const v3<double> bg = sh_geo<double>(E.shG, E.shH, E.sh_nm, geo.x, geo.y, geo.z);
        const v3<double> bs = rot_geo_to_gsm(E.A, bg);
        tot.x += bs.x;
        tot.y += bs.y;
        tot.z += bs.z;
    }
    return rot_gsm_to_geo(E.A, tot);
}

// Mixed path: the SH model takes the GEO position directly (no GSM round trip), the
// external model's GSM field is rotated to GEO in FP32.  Positions go to Re by a
// multiplication (the FP32 argument is the same to within its last bit).
template <int INT, int EXT, int CRU>
__device__ __forceinline__ v3<float> field_mixed(const EpochView<float> &E, v3<double> geo_m,
                                                 v3<double> gsm_m) {
    v3<float> tot = mk3(0.f, 0.f, 0.f);
    if (INT == INT_IGRF) {
        const float x = (float)(geo_m.x * K_INV_RE), y = (float)(geo_m.y * K_INV_RE), z = (float)(geo_m.z * K_INV_RE);
        tot = igrf_geo<float>(E.G, E.H, E.REC, x, y, z);
    }
    if (EXT == EXT_T96) {
        const float x = (float)(gsm_m.x * K_INV_RE), y = (float)(gsm_m.y * K_INV_RE), z = (float)(gsm_m.z * K_INV_RE);
        const v3<float> bg = rot_gsm_to_geo_f(E.Af, t96_field<float>(E.t96, x, y, z, E.icon));
        tot.x += bg.x;
        tot.y += bg.y;
        tot.z += bg.z;
    }
    if (EXT == EXT_TS05) {
        const float x = (float)(gsm_m.x * K_INV_RE), y = (float)(gsm_m.y * K_INV_RE), z = (float)(gsm_m.z * K_INV_RE);
        const v3<float> bg = rot_gsm_to_geo_f(E.Af, ts05_field<float>(*E.ts05, *E.ts05d, x, y, z));
        tot.x += bg.x;
        tot.y += bg.y;
        tot.z += bg.z;
    }
    if (CRU == CRU_SH) {
        const float x = (float)(geo_m.x * K_INV_RE), y = (float)(geo_m.y * K_INV_RE), z = (float)(geo_m.z * K_INV_RE);
        const v3<float> bg = sh_geo<float>(E.shG, E.shH, E.sh_nm, x, y, z);
        tot.x += bg.x;
        tot.y += bg.y;
        tot.z += bg.z;
    }
    return tot;
}

// ---------------------------------------------------------------- RK step

// src/trajectory.c:275-356 in FP64: velocity stages of dv/dt = d (v x B) with B frozen,
// position stages of the same tableau applied to f(x) = v + x; divisions kept.
__device__ __forceinline__ void rk_step_fp64(double h, double d, v3<double> B, v3<double> &vel,
                                             v3<double> &pos) {
    const double vx = vel.x, vy = vel.y, vz = vel.z;
    double kx[6], ky[6], kz[6];
    kx[0] = h * (d * (vy * B.z - vz * B.y));
    ky[0] = h * (d * (vz * B.x - vx * B.z));
    kz[0] = h * (d * (vx * B.y - vy * B.x));
#define COR_STAGE(S, CX, CY, CZ)                                   \
    {                                                              \
        const double ux = vx + (CX), uy = vy + (CY), uz = vz + (CZ); \
        kx[S] = h * (d * (uy * B.z - uz * B.y));                   \
        ky[S] = h * (d * (uz * B.x - ux * B.z));                   \
        kz[S] = h * (d * (ux * B.y - uy * B.x));                   \
    }
    COR_STAGE(1, kx[0] / 3., ky[0] / 3., kz[0] / 3.)
    COR_STAGE(2, (4. * kx[0] + 6. * kx[1]) / 25., (4. * ky[0] + 6. * ky[1]) / 25.,
              (4. * kz[0] + 6. * kz[1]) / 25.)
    COR_STAGE(3, (kx[0] - 12. * kx[1] + 15. * kx[2]) / 4., (ky[0] - 12. * ky[1] + 15. * ky[2]) / 4.,
              (kz[0] - 12. * kz[1] + 15. * kz[2]) / 4.)
    COR_STAGE(4, (6. * kx[0] + 90. * kx[1] - 50. * kx[2] + 8. * kx[3]) / 81.,
              (6. * ky[0] + 90. * ky[1] - 50. * ky[2] + 8. * ky[3]) / 81.,
              (6. * kz[0] + 90. * kz[1] - 50. * kz[2] + 8. * kz[3]) / 81.)
    COR_STAGE(5, (6. * kx[0] + 36. * kx[1] + 10. * kx[2] + 8. * kx[3]) / 75.,
              (6. * ky[0] + 36. * ky[1] + 10. * ky[2] + 8. * ky[3]) / 75.,
              (6. * kz[0] + 36. * kz[1] + 10. * kz[2] + 8. * kz[3]) / 75.)
#undef COR_STAGE
    // position stages: p_s = h (v + combination of earlier p)
    double p[3][6];
    const double vv[3] = {vx, vy, vz};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double v0 = vv[c];
        double *q = p[c];
        q[0] = h * v0;
        q[1] = h * (v0 + q[0] / 3.);
        q[2] = h * (v0 + (4. * q[0] + 6. * q[1]) / 25.);
        q[3] = h * (v0 + (q[0] - 12. * q[1] + 15. * q[2]) / 4.);
        q[4] = h * (v0 + (6. * q[0] + 90. * q[1] - 50. * q[2] + 8. * q[3]) / 81.);
        q[5] = h * (v0 + (6. * q[0] + 36. * q[1] + 10. * q[2] + 8. * q[3]) / 75.);
    }
    vel.x += (23. * kx[0] + 125. * kx[2] - 81. * kx[4] + 125. * kx[5]) / 192.;
    vel.y += (23. * ky[0] + 125. * ky[2] - 81. * ky[4] + 125. * ky[5]) / 192.;
    vel.z += (23. * kz[0] + 125. * kz[2] - 81. * kz[4] + 125. * kz[5]) / 192.;
    pos.x += (23. * p[0][0] + 125. * p[0][2] - 81. * p[0][4] + 125. * p[0][5]) / 192.;
    pos.y += (23. * p[1][0] + 125. * p[1][2] - 81. * p[1][4] + 125. * p[1][5]) / 192.;
    pos.z += (23. * p[2][0] + 125. * p[2][2] - 81. * p[2][4] + 125. * p[2][5]) / 192.;
}

// The same step in FP32 increments added to the FP64 state, in closed form.  With the
// field frozen over the step the velocity stages are linear in v: k_s = hd (u_s x B),
// u_s = v + sum_j a_sj k_j.  The tableau's stability polynomial is exactly
// 1 + z + z^2/2 + z^3/6 + z^4/24 + z^5/120 (no z^6 term), and for z = hd [. x B] the
// powers reduce to two vectors, c = v x B and e = c x B (z^3 v = -hd^2 |B|^2 z v, ...):
//   dv = hd (1 - s/6 + s^2/120) c + hd^2 (1/2 - s/24) e,   s = hd^2 |B|^2.
// The position stages use v only (the reference's form: q_s = h (v + sum_j a_sj q_j)),
// so every component is v times one polynomial in h, formed once.
// Equal to the staged FP64 step in exact arithmetic; in FP32 to its rounding.
__device__ __forceinline__ void rk_step_mixed(double h_d, double d_d, v3<float> B, v3<double> &vel,
                                              v3<double> &pos) {
    const float h = (float)h_d, hd = (float)(h_d * d_d);
    const float vx = (float)vel.x, vy = (float)vel.y, vz = (float)vel.z;
    const float cx = vy * B.z - vz * B.y, cy = vz * B.x - vx * B.z, cz = vx * B.y - vy * B.x;
    const float ex = cy * B.z - cz * B.y, ey = cz * B.x - cx * B.z, ez = cx * B.y - cy * B.x;
    const float s = hd * hd * (B.x * B.x + B.y * B.y + B.z * B.z);
    const float al = hd * (1.f + s * (-1.f / 6.f + s * (1.f / 120.f)));
    const float be = hd * hd * (0.5f + s * (-1.f / 24.f));
    // position polynomial: q_s / v for the stage recurrence of src/trajectory.c's step
    const float r3 = 1.f / 3.f, r25 = 1.f / 25.f, r4 = 0.25f, r81 = 1.f / 81.f, r75 = 1.f / 75.f;
    const float q0 = h;
    const float q1 = h * (1.f + q0 * r3);
    const float q2 = h * (1.f + (4.f * q0 + 6.f * q1) * r25);
    const float q3 = h * (1.f + (q0 - 12.f * q1 + 15.f * q2) * r4);
    const float q4 = h * (1.f + (6.f * q0 + 90.f * q1 - 50.f * q2 + 8.f * q3) * r81);
    const float q5 = h * (1.f + (6.f * q0 + 36.f * q1 + 10.f * q2 + 8.f * q3) * r75);
    const float qp = (23.f * q0 + 125.f * q2 - 81.f * q4 + 125.f * q5) * (1.f / 192.f);
    vel.x += (double)(al * cx + be * ex);
    vel.y += (double)(al * cy + be * ey);
    vel.z += (double)(al * cz + be * ez);
    pos.x += (double)(vx * qp);
    pos.y += (double)(vy * qp);
    pos.z += (double)(vz * qp);
}

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ double norm3(v3<double> v) {
    return sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
}

// global id -> scan (scans are sorted by their first global id)
__device__ __forceinline__ int find_scan(const ScanDev *__restrict__ scans, int n, int64_t gid) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (scans[mid].first <= gid) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// src/trajectory.c:81,91-112: exit point (car_to_sph) and asymptotic direction of an
// allowed trajectory, written to its result record.
// A trajectory's end: status and counters; the exit record of a forbidden one is
// zeroed, so every run of a plan writes every output (a re-run at another epoch leaves
// nothing of the previous one).
__device__ __forceinline__ void write_end(const TraceArgs &a, int64_t gid, int status, unsigned int step,
                                          uint32_t attempts, uint32_t handover = 0) {
    unsigned char *rec = reinterpret_cast<unsigned char *>(a.out + gid);
    *reinterpret_cast<uint4 *>(rec) = make_uint4((uint32_t)status, step, attempts, handover);
    if (status != ST_ALLOWED) {
        const double2 z = make_double2(0., 0.);
        for (int k = 0; k < 4; k++) reinterpret_cast<double2 *>(rec + 16)[k] = z;
    }
}

// One line of the trajectory mode (src/trajectory.c:232-269): the GSM position of the
// new point (geo_to_gsm, src/utils.c:75-87), its spherical coordinates (car_to_sph,
// src/utils.c:43-73), speed, time and path length after an accepted step that does not
// end the trajectory.
__device__ __forceinline__ void write_track(const TraceArgs &a, int64_t gid, unsigned int step, const double *A,
                                            v3<double> pos, double v, double time, double alength) {
    const int64_t k = a.trk_off[gid] + (int64_t)step - 1;
    if (step < 1u || k >= a.trk_off[gid + 1]) return;
    const double x = pos.x, y = pos.y, z = pos.z;
    const double sq = x * x + y * y;
    const double rr = sqrt(sq + z * z);
    double th = 0., ph = 0.;
    if (sq != 0.) {
        ph = atan2(y, x);
        th = atan2(sqrt(sq), z);
        if (ph < 0.) ph += K_2PI;
    } else if (z < 0.) {
        th = K_PI;
    }
    cor_track_point &t = a.trk[k];
    t.step = step;
    t.pad = 0u;
    t.gsm[0] = A[0] * x + A[1] * y + A[2] * z;
    t.gsm[1] = A[3] * x + A[4] * y + A[5] * z;
    t.gsm[2] = A[6] * x + A[7] * y + A[8] * z;
    t.r = rr / K_RE;
    t.td = 90. - (th / K_TO_RAD);
    t.fei = ph / K_TO_RAD;
    t.v = v;
    t.time = time;
    t.length_m = alength;
}

__device__ __forceinline__ void write_exit(const TraceArgs &a, int64_t gid, v3<double> pos, v3<double> vel,
                                           double v, double r, double time, double alength) {
    const double rho2 = pos.x * pos.x + pos.y * pos.y;
    double th = 0., ph = 0.;
    if (rho2 != 0.) {
        ph = atan2(pos.y, pos.x);
        th = atan2(sqrt(rho2), pos.z);
        if (ph < 0.) ph += K_2PI;
    } else if (pos.z < 0.) {
        th = K_PI;
    }
    const double st = sin(th), ct = cos(th), sf = sin(ph), cf = cos(ph);
    const double be = vel.x * cf + vel.y * sf;
    const double vr = be * st + vel.z * ct;
    const double vth = be * ct - vel.z * st;
    const double vph = -vel.x * sf + vel.y * cf;
    const double va = vth * ct + vr * st;
    const double dm = sqrt(vth * vth + va * va);
    double asf = (ph + atan2(vph, va)) / K_TO_RAD;
    if (asf > 360.) asf -= 360.;
    double2 *d = reinterpret_cast<double2 *>(reinterpret_cast<unsigned char *>(a.out + gid) + 16);
    d[0] = make_double2(v / K_C, r);
    d[1] = make_double2(90. - (th / K_TO_RAD), ph / K_TO_RAD);
    d[2] = make_double2(atan2(-vth * st + vr * ct, dm) / K_TO_RAD, asf);
    d[3] = make_double2(time, alength);
}

// src/trajectory.c:46-152 for one lane, FP64 and in the reference's operation order:
// RK attempts with the frozen field are repeated until one is accepted (returns -1) or
// the trajectory ends (returns its status; an allowed trajectory also writes its exit
// record).  The speed carried into the next attempt is the last attempt's, rejected or
// not, as in the reference.
template <class EV>
__device__ __forceinline__ int attempts_fp64(const TraceArgs &a, const EV &E, v3<double> fld,
                                             v3<double> &pos, v3<double> &vel, double &v, double b,
                                             double zn, int &nk0, int nk1, unsigned int &step,
                                             uint32_t &attempts, double &time, double &alength, int64_t gid,
                                             bool writer = true) {
    for (;;) {
        step++;
        const double vc = v / K_C;
        if (vc > 1.) return ST_FORBIDDEN_REENTRY;
        const double hm = a.mass / sqrt(1. - vc * vc);
        const double period = (2. * K_PI * hm) / (fabs(zn) * K_Q * b);
        const double h = period / nk0;
        const double d = (1.E-09 * zn * K_Q) / hm;
        const v3<double> pos0 = pos, vel0 = vel;
        const double vprev = v;
        rk_step_fp64(h, d, fld, vel, pos);
        attempts++;
        v = norm3(vel);
        if (__builtin_isnan(v)) {
            // a NaN state meets no termination test: the reference runs it to the step
            // limit and classifies it as forbidden there
            step = (unsigned int)(a.step_limit + 1);
            return ST_FORBIDDEN_LIMIT;
        }
        const double gc = (vel.x * vel0.x + vel.y * vel0.y + vel.z * vel0.z) / (v * vprev);
        bool reject = false;
        if (gc < 1.) {
            const double gu = atan2(sqrt((1. - gc) * (1. + gc)), gc);
            reject = gu >= a.tu_angle;
        }
        if (reject && nk0 <= 50000) {
            nk0 = (int)((unsigned)nk0 * 2u);   // wraps like the reference's int
            pos = pos0;
            vel = vel0;
            step--;
            continue;
        }
        const v3<double> gsm = rot_geo_to_gsm(E.A, pos);
        const double pax = (gsm.x / K_RE + 25.3) / 36.08;
        const double paz = (gsm.y * gsm.y + gsm.z * gsm.z) / (K_RE * K_RE);
        const double r = norm3(pos) / K_RE;
        const double paus = pax * pax + paz / 459.6736;
        if (r >= 25. || paus > 1.) {
            if ((r >= 25. && fabs(r - 25.) < 0.002) || (r <= 25. && paus > 1. && fabs(paus - 1.) < 0.002)) {
                if (writer) write_exit(a, gid, pos, vel, v, r, time, alength);
                return ST_ALLOWED;
            }
            nk0 = (int)((unsigned)nk0 * 2u);
            pos = pos0;
            vel = vel0;
            step--;
            continue;
        }
        if (r < 1.) return ST_FORBIDDEN_REENTRY;
        if ((uint64_t)step > a.step_limit) return ST_FORBIDDEN_LIMIT;
        // accepted, continuing: src/trajectory.c:150-152
        time = time + period / nk0;
        const double dx = pos.x - pos0.x, dy = pos.y - pos0.y, dz = pos.z - pos0.z;
        alength += sqrt(dx * dx + dy * dy + dz * dz);
        nk0 = nk1;
        return -1;
    }
}

// The same step control for the mixed path: FP32 RK increments on the FP64 state, all
// control arithmetic in FP64 but with reciprocal multiplications instead of the
// reference's divisions, the turn test as gc <= cos(tu_angle) (acos is decreasing),
// and the sphere tests on |x|^2 with a square root only near the escape radius.
// 1/sqrt(q) in FP64 for the mixed step: the hardware estimate and one Newton step
// (full precision for its uses: speeds, lengths, the Lorentz factor, 1/|B|).
__device__ __forceinline__ double rsq_fp64(double q) {
    const double r = __builtin_amdgcn_rsq(q);
    return r * (1.5 - 0.5 * q * r * r);
}

// 1/b of the mixed step (b = 1e-9 |B| in T).
__device__ __forceinline__ double inv_b_mixed(v3<float> f) {
    const double fx = f.x, fy = f.y, fz = f.z;
    return 1.E+09 * rsq_fp64(fx * fx + fy * fy + fz * fz);
}

template <class EV>
__device__ __forceinline__ int attempts_mixed(const TraceArgs &a, const EV &E, v3<float> fld,
                                              v3<double> &pos, v3<double> &vel, double &v, double inv_b,
                                              double zn, int &nk0, int nk1, unsigned int &step,
                                              uint32_t &attempts, double &time, double &alength, int64_t gid,
                                              bool writer = true) {
    constexpr double inv_c = 1.0 / K_C, inv_re = 1.0 / K_RE, inv_re2 = 1.0 / (K_RE * K_RE);
    constexpr double re2 = K_RE * K_RE, r25sq = (25. * K_RE) * (25. * K_RE);
    const double inv_mass = a.inv_mass;
    const double azn = fabs(zn);                         // period = kt * hm / b
    const double kt = azn == 1. ? (2. * K_PI) / K_Q : (2. * K_PI) / (azn * K_Q);
    const double kd = 1.E-09 * zn * K_Q;                 // d = kd / hm
    double inv_nk0 = 1.0 / nk0;                          // halved exactly when nk0 doubles
    for (;;) {
        step++;
        const double vc = v * inv_c;
        if (vc > 1.) return ST_FORBIDDEN_REENTRY;
        // 1/sqrt(1 - vc^2) by the hardware estimate and one Newton step (full FP64
        // precision), in place of a square root and a division
        const double g2 = 1. - vc * vc;
        const double rg = rsq_fp64(g2);
        const double inv_hm = g2 * rg * inv_mass;
        const double hm = a.mass * rg;
        const double period = kt * hm * inv_b;
        const double h = period * inv_nk0;
        const double d = kd * inv_hm;
        const v3<double> pos0 = pos, vel0 = vel;
        const double vprev = v;
        rk_step_mixed(h, d, fld, vel, pos);
        attempts++;
        const double v2 = vel.x * vel.x + vel.y * vel.y + vel.z * vel.z;
        v = v2 * rsq_fp64(v2);   // NaN stays NaN (and is caught below)
        if (__builtin_isnan(v)) {
            step = (unsigned int)(a.step_limit + 1);
            return ST_FORBIDDEN_LIMIT;
        }
        const double dot = vel.x * vel0.x + vel.y * vel0.y + vel.z * vel0.z;
        if (dot <= a.cos_tu * (v * vprev) && nk0 <= 50000) {   // acos(gc) >= tu_angle
            nk0 = (int)((unsigned)nk0 * 2u);
            inv_nk0 *= 0.5;
            pos = pos0;
            vel = vel0;
            step--;
            continue;
        }
        const v3<double> gsm = rot_geo_to_gsm(E.A, pos);
        const double pax = (gsm.x * inv_re + 25.3) * (1.0 / 36.08);
        const double paz = (gsm.y * gsm.y + gsm.z * gsm.z) * inv_re2;
        const double paus = pax * pax + paz * (1.0 / 459.6736);
        const double rr2 = pos.x * pos.x + pos.y * pos.y + pos.z * pos.z;
        if (rr2 >= r25sq || paus > 1.) {
            const double r = sqrt(rr2) * inv_re;
            if ((r >= 25. && fabs(r - 25.) < 0.002) || (r <= 25. && paus > 1. && fabs(paus - 1.) < 0.002)) {
                if (writer) write_exit(a, gid, pos, vel, v, r, time, alength);
                return ST_ALLOWED;
            }
            nk0 = (int)((unsigned)nk0 * 2u);
            inv_nk0 *= 0.5;
            pos = pos0;
            vel = vel0;
            step--;
            continue;
        }
        if (rr2 < re2) return ST_FORBIDDEN_REENTRY;
        if ((uint64_t)step > a.step_limit) return ST_FORBIDDEN_LIMIT;
        time = time + h;
        const double dx = pos.x - pos0.x, dy = pos.y - pos0.y, dz = pos.z - pos0.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        alength += d2 > 0. ? d2 * rsq_fp64(d2) : 0.;
        nk0 = nk1;
        return -1;
    }
}

template <int PREC>
struct FieldT;
template <>
struct FieldT<PREC_FP64> { typedef double T; };
template <>
struct FieldT<PREC_MIXED> { typedef float T; };

// ---------------------------------------------------------------- cooperative tail

__device__ __forceinline__ unsigned long long load_acquire(unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int load_acquire(int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
// relaxed load of a shared counter, made wavefront-uniform
// A lane's double read by every lane, the lane index being wavefront-uniform (two
// readlanes through scalar registers instead of an LDS permute).
__device__ __forceinline__ double read_lane_d(double v, int src) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, src);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ unsigned long long load_relaxed_uniform(unsigned long long *p) {
    const unsigned long long x = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)x);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(x >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Hand this wavefront's live trajectories to the cooperative workers (called with the
// whole wavefront active, at the top of a step): lane 0 reserves one park slot per live
// lane, each live lane writes its state and publishes the slot with the run id.
__device__ __forceinline__ void park_lanes(const TraceArgs &a, uint32_t run_id, int lane, bool active, int64_t gid,
                                           v3<double> pos, v3<double> vel, double v, double time,
                                           double alength, double zn, unsigned int step, uint32_t attempts,
                                           int nk1) {
    const uint64_t live = __ballot(active);
    if (!live) return;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&a.coop->res, (unsigned long long)__popcll(live));
    base = __shfl(base, 0);
    if (active) {
        ParkDev *P = a.park + base + __popcll(live & ((1ull << lane) - 1ull));
        P->pos[0] = pos.x;
        P->pos[1] = pos.y;
        P->pos[2] = pos.z;
        P->vel[0] = vel.x;
        P->vel[1] = vel.y;
        P->vel[2] = vel.z;
        P->v = v;
        P->time = time;
        P->alength = alength;
        P->zn = zn;
        P->gid = gid;
        P->step = step;
        P->attempts = attempts;
        P->nk1 = nk1;
        P->handover = 0u;
        __hip_atomic_store(&P->ready, run_id, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The field of a wavefront's few trajectories with FG lanes each (fewer, wider groups
// than the bookkeeping groups of coop_group<PREC>() lanes: a wavefront step costs about the same
// whatever the number of trajectories, and wider groups finish it sooner).  Field group
// j evaluates the j-th live trajectory (leads: the first lanes of the bookkeeping groups
// that hold one); each bookkeeping group then reads its trajectory's field from that
// field group, so all its lanes hold the same bits.
template <class FT, int INT, int EXT, int FG>
__device__ __forceinline__ v3<FT> coop_field_spread(const EpochView<FT> &E, const CoopTabT<FT> &T, uint64_t leads,
                                                    v3<double> pos, int lane, int lead) {
    const int fg = lane / FG;
    uint64_t m = leads;
#pragma unroll
    for (int j = 0; j < 64 / FG - 1; j++)
        if (j < fg) m &= m - 1ull;
    const int src = (m ? __ffsll((unsigned long long)m) : __ffsll((unsigned long long)leads)) - 1;
    const v3<double> p1 = mk3(__shfl(pos.x, src), __shfl(pos.y, src), __shfl(pos.z, src));
    const v3<double> gsm = rot_geo_to_gsm(E.A, p1);
    const v3<FT> fo = coop_field<FT, INT, EXT, FG>(E.G, E.H, E.REC, rot_rows(E), E.t96, T, p1, gsm, lane & (FG - 1));
    const int r = __popcll(leads & ((1ull << lead) - 1ull));
    const int from = (r < 64 / FG ? r : 0) * FG;
    return mk3(__shfl(fo.x, from), __shfl(fo.y, from), __shfl(fo.z, from));
}

// Cooperative worker; returns true when the run is over, false when the wavefront should
// go back to the per-lane mode (it is idle, no parked trajectory waits and the queue still
// has new ones).  A wavefront runs up to 64 / GS parked trajectories (eight in either
// precision), one per group of GS = coop_group<PREC>() lanes, while parked
// work is plentiful; with eight or fewer the field is evaluated with wider groups (a lone
// trajectory with all 64 lanes: its step then takes about two thirds of the time).  Groups of a wavefront hold the same state in all
// their lanes and compute the same bits.
//   * An idle wavefront takes one ticket for the next park slot (tickets are served in
//     slot order) and waits until the slot is reserved and published.
//   * A busy wavefront takes tickets for its free groups only for slots already
//     reserved and not yet claimed (compare-and-swap on the ticket counter).
//   * A wavefront running two or more trajectories while another waits hands back the
//     one with the most steps (re-park, at most one park slot per lane of the grid in
//     all), so the trajectories spread one per wavefront as the work thins out, the
//     longest-running first.
// An idle wavefront keeps its ticket until every trajectory of the run has finished
// (a counter both modes add to), so a trajectory handed back is always taken; then it
// retires.  Every slot is served by the holder of its ticket, so the count is reached.
// The only other exit is a guard against a lost record: with no wavefront left in the
// per-lane mode and no trajectory finished for idle_guard ticks -- set longer than one
// trajectory can run to the step limit at the slowest step of any mode, so it cannot
// fire while a trajectory is still running -- an idle wavefront retires and the host
// reports the shortfall.  Workers never wait for one another and parkers never wait.
template <int PREC, int INT, int EXT>
__device__ bool coop_worker(const TraceArgs &a, uint32_t run_id, const EpochView<typename FieldT<PREC>::T> &E,
                            const CoopTabT<typename FieldT<PREC>::T> &T, int lane) {
    typedef typename FieldT<PREC>::T FT;
    constexpr int GS = coop_group<PREC>();
    const int gl = lane & (GS - 1);
    const int lead = lane & ~(GS - 1);            // first lane of this lane's group
    const bool is_lead = gl == 0;
    bool has = false, holding = false;
    unsigned long long ticket = 0;
    v3<double> pos = mk3(2.0 * K_RE, 1.0 * K_RE, 0.5 * K_RE), vel = mk3(0., 0., 0.);
    double v = 0., time = 0., alength = 0., zn = 0.;
    int64_t gid = 0;
    unsigned int step = 0;
    uint32_t attempts = 0, handover = 0;
    int nk1 = 1, nk0 = 1;
    unsigned long long evals = 0;
    bool repark_ok = true;
    unsigned iter = 0;
    unsigned long long done_seen = ~0ull, idle_since = 0;
    for (;;) {
        iter++;
        uint64_t has_leads = __ballot(has && is_lead);
        // ---- 1. tickets and the start of parked trajectories (counters read relaxed:
        // an acquire load would invalidate the caches of the model constants; acquire
        // is used where it orders data: a published slot, and running == 0 before res
        // is taken as final)
        // a busy wavefront looks at the shared counters every 16th step only
        const bool look = !has_leads || (iter & 15u) == 0;
        if (look && __popcll(has_leads) < 64 / GS) {
            const bool free_g = !has && !holding;
            const uint64_t free_leads = __ballot(free_g && is_lead);
            const unsigned long long head0 = load_relaxed_uniform(&a.coop->head);
            const unsigned long long res0 = load_relaxed_uniform(&a.coop->res);
            int want = 0;
            bool wait_ok = false;
            if (!has_leads && !__any(holding)) {
                // idle: with no parked trajectory waiting and the queue not yet drained,
                // this wavefront goes back to taking new trajectories in the per-lane mode
                if (head0 >= res0 && (load_relaxed_uniform(a.head) < (unsigned long long)a.n_items ||
                                      load_relaxed_uniform(&a.coop->lres) > load_relaxed_uniform(&a.coop->lhead)))
                    return false;
                want = 1;
                wait_ok = true;
            } else if (free_leads && head0 < res0) {
                want = (int)min((unsigned long long)__popcll(free_leads), res0 - head0);
            }
            if (want > 0) {
                const int first = __ffsll((unsigned long long)free_leads) - 1;
                unsigned long long base = ~0ull;
                if (lane == first) {
                    if (wait_ok) {
                        base = atomicAdd(&a.coop->head, (unsigned long long)want);
                    } else if (atomicCAS(&a.coop->head, head0, head0 + (unsigned long long)want) == head0) {
                        base = head0;
                    }
                }
                base = __shfl(base, first);
                if (base != ~0ull && free_g) {
                    const int rank = __popcll(free_leads & ((1ull << lead) - 1ull));
                    if (rank < want) {
                        ticket = base + (unsigned long long)rank;
                        holding = true;
                    }
                }
            }
            int running = __hip_atomic_load(&a.coop->running, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned long long res;
            if (running == 0) {
                running = load_acquire(&a.coop->running);
                res = load_acquire(&a.coop->res);
            } else {
                res = __hip_atomic_load(&a.coop->res, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (holding && ticket < res) {
                ParkDev *P = a.park + ticket;
                if (__hip_atomic_load(&P->ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == run_id) {
                    pos = mk3(P->pos[0], P->pos[1], P->pos[2]);
                    vel = mk3(P->vel[0], P->vel[1], P->vel[2]);
                    v = P->v;
                    time = P->time;
                    alength = P->alength;
                    zn = P->zn;
                    gid = P->gid;
                    step = P->step;
                    attempts = P->attempts;
                    nk1 = P->nk1;
                    nk0 = nk1;
                    handover = P->handover ? P->handover : step + 1u;
                    evals = 0;
                    has = true;
                    holding = false;
                }
            }
            has_leads = __ballot(has && is_lead);
            if (!has_leads) {
                const unsigned long long done = load_relaxed_uniform(&a.coop->done);
                if (done >= (unsigned long long)a.n_items) return true;
                const unsigned long long now = wall_clock64();
                if (done != done_seen || running != 0) {
                    done_seen = done;
                    idle_since = now;
                } else if (now - idle_since > a.idle_guard) {
                    return true;
                }
                __builtin_amdgcn_s_sleep(127);
                continue;
            }
        }
        // ---- 2. spread: with two or more trajectories here and a wavefront waiting for
        // work, hand back the one with the most steps so far (most likely the longest to
        // go), which then runs alone
        if (look && __popcll(has_leads) >= 2 && repark_ok) {
            const unsigned long long head1 = load_relaxed_uniform(&a.coop->head);
            const unsigned long long res1 = load_relaxed_uniform(&a.coop->res);
            int src = -1;
            if (head1 > res1) {
                unsigned best = 0;
                for (uint64_t m = has_leads; m; m &= m - 1) {
                    const int l = __ffsll((unsigned long long)m) - 1;
                    const unsigned s_l = (unsigned)__shfl((int)step, l);
                    if (src < 0 || s_l > best) {
                        best = s_l;
                        src = l;
                    }
                }
            }
            if (src >= 0) {
                unsigned long long r = 0;
                if (lane == src) r = atomicAdd(&a.coop->reparks, 1ull);
                r = __shfl(r, src);
                if (r < (unsigned long long)a.repark_limit) {
                    if (lead == src) {
                        if (is_lead) {
                            const unsigned long long slot = atomicAdd(&a.coop->res, 1ull);
                            ParkDev *P = a.park + slot;
                            P->pos[0] = pos.x;
                            P->pos[1] = pos.y;
                            P->pos[2] = pos.z;
                            P->vel[0] = vel.x;
                            P->vel[1] = vel.y;
                            P->vel[2] = vel.z;
                            P->v = v;
                            P->time = time;
                            P->alength = alength;
                            P->zn = zn;
                            P->gid = gid;
                            P->step = step;
                            P->attempts = attempts;
                            P->nk1 = nk1;
                            P->handover = handover;
                            if (evals) atomicAdd(&a.stats->coop_steps, evals);
                            __hip_atomic_store(&P->ready, run_id, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        has = false;
                    }
                    has_leads = __ballot(has && is_lead);
                } else {
                    repark_ok = false;
                }
            }
        }
        // ---- 3. one step of each trajectory.  The field: a lone trajectory with all 64
        // lanes, two with 32 lanes each, three or four with 16, five to eight with 8, more
        // with their own group of GS lanes (groups without one evaluate at their last
        // point).
        v3<FT> f;
        const int k_live = __popcll(has_leads);
        if (k_live == 1) {
            const int src = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)has_leads) - 1);
            const v3<double> p1 = mk3(read_lane_d(pos.x, src), read_lane_d(pos.y, src), read_lane_d(pos.z, src));
            const v3<double> gsm = rot_geo_to_gsm(E.A, p1);
            f = coop_field<FT, INT, EXT, 64>(E.G, E.H, E.REC, rot_rows(E), E.t96, T, p1, gsm, lane);
        } else if (k_live == 2) {
            f = coop_field_spread<FT, INT, EXT, 32>(E, T, has_leads, pos, lane, lead);
        } else if (k_live <= 4) {
            f = coop_field_spread<FT, INT, EXT, 16>(E, T, has_leads, pos, lane, lead);
        } else if (GS < 8 && k_live <= 8) {
            f = coop_field_spread<FT, INT, EXT, (GS < 8 ? 8 : GS)>(E, T, has_leads, pos, lane, lead);
        } else {
            const v3<double> gsm = rot_geo_to_gsm(E.A, pos);
            f = coop_field<FT, INT, EXT, GS>(E.G, E.H, E.REC, rot_rows(E), E.t96, T, pos, gsm, gl);
        }
        if (has) {
            evals++;
            int status;
            if constexpr (PREC == PREC_FP64) {
                const double b = 1.E-09 * sqrt(f.x * f.x + f.y * f.y + f.z * f.z);
                status = attempts_fp64(a, E, f, pos, vel, v, b, zn, nk0, nk1, step, attempts, time, alength, gid, is_lead);
            } else {
                status = attempts_mixed(a, E, f, pos, vel, v, inv_b_mixed(f), zn, nk0, nk1, step, attempts, time, alength,
                                        gid, is_lead);
            }
            if (status >= 0) {
                if (is_lead) {
                    write_end(a, gid, status, step, attempts, handover);
                    atomicAdd(&a.stats->coop_steps, evals);
                    atomicAdd(&a.coop->done, 1ull);
                    __hip_atomic_fetch_add(&a.stats->finished, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                has = false;
            }
        }
    }
}

// ---------------------------------------------------------------- per-lane compaction

// After the queue is drained, a per-lane wavefront with at most POOL_DONATE_LANES live
// trajectories puts them into the per-lane pool and turns worker; a wavefront with idle
// lanes takes pooled trajectories into them (idle workers come back to the per-lane mode
// for them).  A pooled trajectory continues in the per-lane mode, with the same
// arithmetic, so compaction changes no result.  A wavefront that took from the pool
// runs POOL_HOLD_ITERS steps before it may give to it again.
#ifndef COR_POOL_DONATE_LANES
#define COR_POOL_DONATE_LANES 16
#endif
#ifndef COR_POOL_HOLD_ITERS
#define COR_POOL_HOLD_ITERS 256
#endif
constexpr int POOL_DONATE_LANES = COR_POOL_DONATE_LANES;
constexpr unsigned POOL_HOLD_ITERS = COR_POOL_HOLD_ITERS;

__device__ __forceinline__ void pool_write(ParkDev *P, uint32_t run_id, int64_t gid, v3<double> pos, v3<double> vel,
                                           double v, double time, double alength, double zn, unsigned int step,
                                           uint32_t attempts, int nk1, uint32_t ho_at) {
    P->pos[0] = pos.x;
    P->pos[1] = pos.y;
    P->pos[2] = pos.z;
    P->vel[0] = vel.x;
    P->vel[1] = vel.y;
    P->vel[2] = vel.z;
    P->v = v;
    P->time = time;
    P->alength = alength;
    P->zn = zn;
    P->gid = gid;
    P->step = step;
    P->attempts = attempts;
    P->nk1 = nk1;
    P->handover = ho_at;          // the trajectory's handover step count travels with it
    __hip_atomic_store(&P->ready, run_id, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- the kernel

// Register budget: COR_WAVES_PER_EU wavefronts per SIMD (set per translation unit; the
// FP64 T96 tracer needs ~250 VGPRs, so 2; the FP32 field allows more).
#ifndef COR_WAVES_PER_EU
#define COR_WAVES_PER_EU 2
#endif

// COOP: the kernel carries the cooperative worker (mixed precision always; FP64 in a
// kernel of its own, so the FP64 per-lane kernel keeps its registers free of it).
// MULTI: the multi-epoch tracer -- a.epoch is an array of epoch blocks and each scan
// names its block, so one launch drains trajectories of many epochs; each lane reads
// its epoch's constants from device memory (per-lane mode only, no T96 model set).
template <int PREC, int INT, int EXT, int CRU, bool COOP = PREC == PREC_MIXED, bool MULTI = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(COR_WAVES_PER_EU)))
trace_kernel(TraceArgs a) {
    typedef typename FieldT<PREC>::T FT;
    static_assert(!MULTI || (!COOP && EXT != EXT_T96), "the multi-epoch tracer is per-lane, without T96");
    EpochView<FT> E;
    if constexpr (!MULTI) {
        __shared__ FT s_g[IGRF_LDS_N], s_h[IGRF_LDS_N], s_rec[IGRF_LDS_N];
        E = stage_epoch<FT>(a.epoch, s_g, s_h, s_rec);
    }
    const EpochDev *ep = a.epoch;   // this lane's epoch block (multi-epoch tracer)
    const int lane = __lane_id();
    const CoopTabT<FT> *ctab = nullptr;
    constexpr bool kCoop = COOP && EXT != EXT_TS05 && CRU == CRU_NONE;   // the cooperative evaluator's models
    // this run's id (advanced by the reset kernel that precedes every launch)
    uint32_t run_id = 0u;
    if constexpr (kCoop) {
        __shared__ CoopTabT<FT> s_coop;
        if (a.coop_on) {
            run_id = __hip_atomic_load(&a.coop->run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            stage_coop(coop_tab(a.epoch, FT(0)), &s_coop);
            if (lane == 0) atomicAdd(&a.coop->running, 1);
        }
        ctab = &s_coop;
    }

    bool active = false;
    // this wavefront refills ended lanes from the queue (until the queue is drained, or
    // it turns into a cooperative worker)
    bool feeding, converting;
    int64_t gid = 0;
    int scan = 0;
    v3<double> pos = mk3(0., 0., 0.), vel = mk3(0., 0., 0.);
    v3<FT> fld = mk3<FT>(0, 0, 0);
    double v = 0., b = 0., time = 0., alength = 0., zn = 0.;
    int nk0 = 0, nk1 = 0;
    unsigned int step = 0;
    uint32_t attempts = 0;
    uint32_t ho_at = 0xffffffffu;   // this trajectory's handover step (cooperative mode)
    unsigned iter = 0;
    unsigned no_donate_until = 0;   // per-lane pool: iteration before which this wavefront keeps its lanes
    // fill of this wavefront's per-lane loop
    unsigned long long wave_iters = 0, lane_iters = 0, wave_iters_open = 0, lane_iters_open = 0;
    unsigned long long ended_here = 0;
    if (lane == 0) atomicMin(&a.stats->t_start, (unsigned long long)wall_clock64());
    // per-lane phase, then (mixed, cooperative model sets) the cooperative worker, which
    // may hand the wavefront back to the per-lane mode while the queue has trajectories
    for (int phase = 0;; phase++) {
        feeding = true;
        converting = false;
        ended_here = 0;
        if constexpr (kCoop) {
            if (a.coop_on && lane == 0) {
                if (phase > 0) atomicAdd(&a.coop->running, 1);
                atomicAdd(&a.coop->feeders, 1);
            }
        }

        for (;;) {
            iter++;
            // ---- 1. refill ended lanes from the queue
            if (feeding) {
                const uint64_t idle = __ballot(!active);
                if (idle) {
                    const int leader = __ffsll((unsigned long long)idle) - 1;
                    unsigned long long base = 0;
                    if (lane == leader) base = atomicAdd(a.head, (unsigned long long)__popcll(idle));
                    base = __shfl(base, leader);
                    if (base + __popcll(idle) >= (unsigned long long)a.n_items) {
                        feeding = false;
                        if constexpr (kCoop) {
                            if (a.coop_on && lane == leader) atomicSub(&a.coop->feeders, 1);
                        }
                        // this wavefront took the queue's last slot
                        if (lane == leader && base < (unsigned long long)a.n_items)
                            a.stats->t_drain = (unsigned long long)wall_clock64();
                    }
                    if (!active) {
                        const unsigned long long slot = base + __popcll(idle & ((1ull << lane) - 1ull));
                        if (slot < (unsigned long long)a.n_items) {
                            gid = a.order ? (int64_t)a.order[slot] + a.id_base : (int64_t)slot + a.id_base;
                            scan = find_scan(a.scans, a.n_scans, gid);
                            const ScanDev &S = a.scans[scan];
                            const int i = (int)(gid - S.first);
                            if constexpr (MULTI) ep = a.epoch + S.epoch;
                            // src/trajectory.c:13-43
                            zn = S.zn;
                            nk1 = S.nk1;
                            nk0 = nk1;
                            ho_at = i < a.handover_low_count ? a.handover_low_steps : a.handover_steps;
                            const double rig = S.rig + i * S.del;
                            const double aa = rig * 1.0e+09 * zn * K_Q / (a.mass * K_C * K_C);
                            v = K_C * sqrt(1. - 1. / (1. + aa * aa));
                            const double rs = v * S.st1;
                            vel = mk3(rs * S.cp1, rs * S.sp1, v * S.ct1);
                            pos = mk3(S.pos0[0], S.pos0[1], S.pos0[2]);
                            time = 0.;
                            alength = 0.;
                            step = 0;
                            attempts = 0;
                            active = true;
                        }
                    }
                }
            }
            if constexpr (kCoop) {
                if (a.coop_on) {
                    // ---- 2. handover: a trajectory that has run its handover step count
                    // (handover_low_steps for the lowest rigidities of a scan, where trapped
                    // trajectories concentrate, handover_steps for the others) continues in
                    // the cooperative mode.  The trajectory's own step count and rigidity
                    // index alone decide, so its arithmetic (per-lane up to that step,
                    // cooperative after it, whose sums have one order for every group width)
                    // is the same in every run, whatever the scheduling.
                    const bool ho = active && step >= ho_at;
                    if (__ballot(ho)) {
                        park_lanes(a, run_id, lane, ho, gid, pos, vel, v, time, alength, zn, step, attempts, nk1);
                        if (ho) active = false;
                    }
                    // ---- 3. a feeding wavefront stops taking new trajectories when parked ones
                    // wait for workers beyond what converting wavefronts will take (one per group),
                    // as long as min_feeders others still feed; it runs its lanes down (each at
                    // most handover_steps steps) and turns worker
                    if (feeding && (iter & 15u) == 0) {
                        const long long backlog =
                            (long long)(load_relaxed_uniform(&a.coop->res) - load_relaxed_uniform(&a.coop->head));
                        if (backlog > (long long)a.convert_backlog) {
                            int c = 0;
                            if (lane == 0) {
                                // leave the feeders only above their floor (decrement first, so
                                // concurrent conversions cannot take the count below it)
                                const int f = atomicSub(&a.coop->feeders, 1);
                                c = atomicAdd(&a.coop->converting, 1);
                                if (f <= a.min_feeders ||
                                    backlog <= (long long)(64 / coop_group<PREC>()) * c + a.convert_backlog) {
                                    atomicSub(&a.coop->converting, 1);
                                    atomicAdd(&a.coop->feeders, 1);
                                    c = -1;
                                }
                            }
                            c = __shfl(c, 0);
                            if (c >= 0) {
                                feeding = false;
                                converting = true;
                            }
                        }
                    }
                }
            }
            if constexpr (kCoop) {
                // ---- 3b. per-lane compaction once this wavefront takes no new trajectories
                const uint64_t live = __ballot(active);
                if (a.coop_on && a.pool_cap > 0 && !feeding && ((iter & 15u) == 0 || !live)) {
                    const int nl = __popcll(live);
                    const unsigned long long ph = load_relaxed_uniform(&a.coop->lhead);
                    const unsigned long long pr = load_relaxed_uniform(&a.coop->lres);
                    if (pr > ph && nl < 64) {
                        // take waiting records into the idle lanes (claimed by compare-and-swap;
                        // each is published by its donor right after the reservation)
                        const int want = (int)min((unsigned long long)(64 - nl), pr - ph);
                        unsigned long long base = ~0ull;
                        if (lane == 0 && atomicCAS(&a.coop->lhead, ph, ph + (unsigned long long)want) == ph) base = ph;
                        base = __shfl(base, 0);
                        if (base != ~0ull) {
                            const int rank = __popcll(~live & ((1ull << lane) - 1ull));
                            if (!active && rank < want) {
                                ParkDev *P = a.pool + base + rank;
                                while (__hip_atomic_load(&P->ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != run_id)
                                    __builtin_amdgcn_s_sleep(1);
                                pos = mk3(P->pos[0], P->pos[1], P->pos[2]);
                                vel = mk3(P->vel[0], P->vel[1], P->vel[2]);
                                v = P->v;
                                time = P->time;
                                alength = P->alength;
                                zn = P->zn;
                                gid = P->gid;
                                step = P->step;
                                attempts = P->attempts;
                                nk1 = P->nk1;
                                nk0 = nk1;          // a step's top: nk0 is back at nk1
                                ho_at = P->handover;
                                active = true;
                            }
                            no_donate_until = iter + POOL_HOLD_ITERS;
                        }
                    } else if (nl > 0 && nl <= POOL_DONATE_LANES && iter >= no_donate_until) {
                        // give the live trajectories to the pool (while it has room) and turn worker
                        unsigned long long base = ~0ull;
                        if (lane == 0 && pr + (unsigned long long)nl <= a.pool_cap &&
                            atomicCAS(&a.coop->lres, pr, pr + (unsigned long long)nl) == pr)
                            base = pr;
                        base = __shfl(base, 0);
                        if (base != ~0ull) {
                            if (active)
                                pool_write(a.pool + base + __popcll(live & ((1ull << lane) - 1ull)), run_id, gid, pos,
                                           vel, v, time, alength, zn, step, attempts, nk1, ho_at);
                            active = false;
                        }
                    }
                }
            }
            const uint64_t live_mask = __ballot(active);
            if (!live_mask) {
                if (feeding) continue;   // every lane was just handed over: refill
                break;
            }
            wave_iters++;
            lane_iters += __popcll(live_mask);
            if (feeding) {
                wave_iters_open++;
                lane_iters_open += __popcll(live_mask);
            }

            bool ended_now = false;
            if (active) {
                // the epoch of this lane: the workgroup's staged view, or (multi-epoch)
                // the lane's own block read from device memory
                EpochView<FT> EL;
                if constexpr (MULTI) EL = global_view<FT>(ep);
                const EpochView<FT> &EV = MULTI ? EL : E;
                // ---- 4. field at the current point (src/trajectory.c:37-38,147-148); the
                // GSM position is recomputed from the GEO state (same arithmetic as the
                // reference's geo_to_gsm after an accepted step)
                const v3<double> gsm = rot_geo_to_gsm(EV.A, pos);
                if constexpr (PREC == PREC_FP64) {
                    const v3<double> f = field_fp64<INT, EXT, CRU>(EV, gsm);
                    fld = mk3<FT>((FT)f.x, (FT)f.y, (FT)f.z);
                    b = 1.E-09 * sqrt(f.x * f.x + f.y * f.y + f.z * f.z);
                } else {
                    const v3<float> f = field_mixed<INT, EXT, CRU>(EV, pos, gsm);
                    fld = mk3<FT>((FT)f.x, (FT)f.y, (FT)f.z);
                    b = inv_b_mixed(f);   // the mixed step takes 1/b
                }

                // ---- 5. attempts until one is accepted or the trajectory ends
                int status;
                if constexpr (PREC == PREC_FP64) {
                    status = attempts_fp64(a, EV, fld, pos, vel, v, b, zn, nk0, nk1, step, attempts, time, alength, gid);
                } else {
                    status = attempts_mixed(a, EV, fld, pos, vel, v, b, zn, nk0, nk1, step, attempts, time, alength, gid);
                }
                if (status >= 0) {
                    write_end(a, gid, status, step, attempts);
                    active = false;
                    ended_now = true;
                } else if constexpr (PREC == PREC_FP64) {
                    if (a.trk) write_track(a, gid, step, EV.A, pos, v, time, alength);
                }
            }
            // trajectories ended here: the progress counter now, the run's count (which
            // the cooperative workers' exit test reads) when the wavefront leaves the
            // per-lane mode
            const uint64_t ended_mask = __ballot(ended_now);
            if (ended_mask) {
                ended_here += __popcll(ended_mask);
                if (lane == 0)
                    __hip_atomic_fetch_add(&a.stats->finished, (unsigned long long)__popcll(ended_mask),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0) {
            atomicAdd(&a.stats->wave_iters, wave_iters);
            atomicAdd(&a.stats->lane_iters, lane_iters);
            atomicAdd(&a.stats->wave_iters_open, wave_iters_open);
            atomicAdd(&a.stats->lane_iters_open, lane_iters_open);
            atomicMax(&a.stats->t_lane_end, (unsigned long long)wall_clock64());
        }
        wave_iters = lane_iters = wave_iters_open = lane_iters_open = 0;
        if constexpr (kCoop) {
            if (a.coop_on) {
                if (lane == 0) {
                    if (ended_here) atomicAdd(&a.coop->done, ended_here);
                    if (converting) atomicSub(&a.coop->converting, 1);
                    __hip_atomic_fetch_add(&a.coop->running, -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (coop_worker<PREC, INT, EXT>(a, run_id, E, *ctab, lane)) break;
                continue;
            }
        }
        break;
    }
}

}  // namespace cor
