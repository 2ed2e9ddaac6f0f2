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
    switch (lanes) {
    case 64: return field_coop_table<64>(intk, ext, cru);
    case 32: return field_coop_table<32>(intk, ext, cru);
    case 16: return field_coop_table<16>(intk, ext, cru);
    default: return field_coop_table<COOP_GROUP>(intk, ext, cru);
    }
}

__global__ void build_coop_kernel(EpochDev *E) {
    if (blockIdx.x == 0 && threadIdx.x == 0) build_coop_tab(E->t96d, E->coop);
}
const void *coop_build_kernel() { return (const void *)&build_coop_kernel; }
}  // namespace cor
