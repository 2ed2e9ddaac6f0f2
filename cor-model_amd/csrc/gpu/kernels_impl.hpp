// kernels_impl.hpp -- instantiation of the tracer and field-batch kernels for one
// precision; included once by kernels_fp64.hip and once by kernels_mixed.hip.
#pragma once
#include "kernels.hpp"
#include "tracer_dev.hpp"

namespace cor {

// calculate_field over a batch of points (src/models.c:194-207), one point per lane.
template <int PREC, int INT, int EXT, int CRU>
__global__ void __launch_bounds__(256) field_batch_kernel(FieldBatchArgs a) {
    typedef typename FieldT<PREC>::T FT;
    __shared__ FT s_g[IGRF_LDS_N], s_h[IGRF_LDS_N], s_rec[IGRF_LDS_N];
    const EpochView<FT> E = stage_epoch<FT>(a.epoch, s_g, s_h, s_rec);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const v3<double> gsm = mk3(a.gsm_m[3 * i], a.gsm_m[3 * i + 1], a.gsm_m[3 * i + 2]);
    v3<double> out;
    if constexpr (PREC == PREC_FP64) {
        out = field_fp64<INT, EXT, CRU>(E, gsm);
    } else {
        const v3<double> geo = rot_gsm_to_geo(E.A, gsm);
        const v3<float> f = field_mixed<INT, EXT, CRU>(E, geo, gsm);
        out = mk3((double)f.x, (double)f.y, (double)f.z);
    }
    a.geo_nT[3 * i] = out.x;
    a.geo_nT[3 * i + 1] = out.y;
    a.geo_nT[3 * i + 2] = out.z;
}

// The same batch with the cooperative evaluator as the tracer's cooperative mode runs
// it: one group of GS lanes (8, 16, 32 or 64) per point, points strided over the grid's
// groups.  Every lane of a wavefront takes part in each pass (the group sums need all
// lanes); groups past the end evaluate the last point again and write nothing.
template <int PREC, int INT, int EXT, int CRU, int GS>
__global__ void __launch_bounds__(256) field_coop_kernel(FieldBatchArgs a) {
    typedef typename FieldT<PREC>::T FT;
    __shared__ FT s_g[IGRF_LDS_N], s_h[IGRF_LDS_N], s_rec[IGRF_LDS_N];
    __shared__ CoopTabT<FT> s_coop;
    const EpochView<FT> E = stage_epoch<FT>(a.epoch, s_g, s_h, s_rec);
    stage_coop(coop_tab(a.epoch, FT(0)), &s_coop);
    const int gl = __lane_id() & (GS - 1);
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x / GS);
    const int64_t first = (int64_t)blockIdx.x * (blockDim.x / GS) + threadIdx.x / GS;
    // the wavefront's groups loop together: as many rounds as its first group needs
    const int64_t wave_first = first - (int64_t)((threadIdx.x & 63) / GS);
    for (int64_t base = wave_first; base < a.n; base += groups) {
        const int64_t i = base + (first - wave_first);
        const int64_t ic = i < a.n ? i : a.n - 1;
        const v3<double> gsm = mk3(a.gsm_m[3 * ic], a.gsm_m[3 * ic + 1], a.gsm_m[3 * ic + 2]);
        const v3<double> geo = rot_gsm_to_geo(E.A, gsm);
        const v3<FT> f = coop_field<FT, INT, EXT, GS>(E.G, E.H, E.REC, rot_rows(E), E.t96, s_coop, geo, gsm, gl);
        if (gl == 0 && i < a.n) {
            a.geo_nT[3 * i] = f.x;
            a.geo_nT[3 * i + 1] = f.y;
            a.geo_nT[3 * i + 2] = f.z;
        }
    }
}

#define COR_COOP_MODEL_SETS(X)             \
    X(INT_IGRF, EXT_NONE, CRU_NONE)        \
    X(INT_IGRF, EXT_T96, CRU_NONE)         \
    X(INT_NONE, EXT_T96, CRU_NONE)
#define COR_MODEL_SETS(X)                  \
    COR_COOP_MODEL_SETS(X)                 \
    X(INT_IGRF, EXT_TS05, CRU_NONE)        \
    X(INT_NONE, EXT_TS05, CRU_NONE)        \
    X(INT_NONE, EXT_NONE, CRU_SH)          \
    X(INT_NONE, EXT_TS05, CRU_SH)          \
    X(INT_IGRF, EXT_NONE, CRU_SH)          \
    X(INT_IGRF, EXT_T96, CRU_SH)           \
    X(INT_IGRF, EXT_TS05, CRU_SH)

template <int PREC, int GS>
const void *field_coop_table(int intk, int ext, int cru);

template <int PREC>
const void *field_coop_table_lanes(int intk, int ext, int cru, int lanes) {
    switch (lanes) {
    case 64: return field_coop_table<PREC, 64>(intk, ext, cru);
    case 32: return field_coop_table<PREC, 32>(intk, ext, cru);
    case 16: return field_coop_table<PREC, 16>(intk, ext, cru);
    case 4:
        if constexpr (PREC == PREC_MIXED) return field_coop_table<PREC, 4>(intk, ext, cru);
        return nullptr;
    default: return field_coop_table<PREC, 8>(intk, ext, cru);
    }
}

// the FP64 tracer with the cooperative tail (the cooperative model sets only)
template <int PREC>
const void *trace_kernel_coop_table(int intk, int ext, int cru) {
#define COR_PICK(I, E, C) \
    if (intk == I && ext == E && cru == C) return (const void *)&trace_kernel<PREC, I, E, C, true>;
    COR_COOP_MODEL_SETS(COR_PICK)
#undef COR_PICK
    return nullptr;
}

template <int PREC>
const void *trace_kernel_table(int intk, int ext, int cru) {
#define COR_PICK(I, E, C) \
    if (intk == I && ext == E && cru == C) return (const void *)&trace_kernel<PREC, I, E, C>;
    COR_MODEL_SETS(COR_PICK)
#undef COR_PICK
    return nullptr;
}

// the multi-epoch tracer: per-lane mode, model sets without T96 (whose constants the
// per-lane evaluator keeps in registers)
#define COR_EPOCHS_MODEL_SETS(X)           \
    X(INT_IGRF, EXT_TS05, CRU_NONE)        \
    X(INT_NONE, EXT_TS05, CRU_NONE)        \
    X(INT_IGRF, EXT_NONE, CRU_NONE)

template <int PREC>
const void *trace_kernel_epochs_table(int intk, int ext, int cru) {
#define COR_PICK(I, E, C) \
    if (intk == I && ext == E && cru == C) return (const void *)&trace_kernel<PREC, I, E, C, false, true>;
    COR_EPOCHS_MODEL_SETS(COR_PICK)
#undef COR_PICK
    return nullptr;
}

template <int PREC>
const void *field_kernel_table(int intk, int ext, int cru) {
#define COR_PICK(I, E, C) \
    if (intk == I && ext == E && cru == C) return (const void *)&field_batch_kernel<PREC, I, E, C>;
    COR_MODEL_SETS(COR_PICK)
#undef COR_PICK
    return nullptr;
}

template <int PREC, int GS>
const void *field_coop_table(int intk, int ext, int cru) {
#define COR_PICK(I, E, C) \
    if (intk == I && ext == E && cru == C) return (const void *)&field_coop_kernel<PREC, I, E, C, GS>;
    COR_COOP_MODEL_SETS(COR_PICK)
#undef COR_PICK
    return nullptr;
}

}  // namespace cor
