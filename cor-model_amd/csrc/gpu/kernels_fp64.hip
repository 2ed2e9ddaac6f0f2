// kernels_fp64.hip -- FP64 (parity) instantiations; built with -ffp-contract=off.
#include "kernels_impl.hpp"

namespace cor {
const void *trace_kernel_fp64(int intk, int ext, int cru) {
    return trace_kernel_table<PREC_FP64>(intk, ext, cru);
}
const void *field_kernel_fp64(int intk, int ext, int cru) {
    return field_kernel_table<PREC_FP64>(intk, ext, cru);
}
}  // namespace cor
