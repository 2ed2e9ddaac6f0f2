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
}  // namespace cor
