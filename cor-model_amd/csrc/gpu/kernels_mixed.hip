// kernels_mixed.hip -- mixed-precision instantiations (FP32 field and RK increments,
// FP64 state); built with FMA contraction and FP32 denormal flushing.
#include "kernels_impl.hpp"

namespace cor {
const void *trace_kernel_mixed(int intk, int ext, int cru) {
    return trace_kernel_table<PREC_MIXED>(intk, ext, cru);
}
const void *field_kernel_mixed(int intk, int ext, int cru) {
    return field_kernel_table<PREC_MIXED>(intk, ext, cru);
}
const void *field_kernel_coop(int intk, int ext, int cru, int lanes) {
    return field_coop_table_lanes<PREC_MIXED>(intk, ext, cru, lanes);
}

__global__ void build_coop_kernel(EpochDev *E) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        build_coop_tab(E->t96d, E->coop);
        build_coop_tab(E->t96d, E->coopd);
    }
}
const void *coop_build_kernel() { return (const void *)&build_coop_kernel; }
}  // namespace cor
