// kernels_epochs_mixed.hip -- the mixed-precision multi-epoch tracer instantiations (a
// translation unit of their own, built in parallel with the others).
#include "kernels_impl.hpp"

namespace cor {
const void *trace_kernel_epochs_mixed(int intk, int ext, int cru) {
    return trace_kernel_epochs_table<PREC_MIXED>(intk, ext, cru);
}
}  // namespace cor
