// kernels_epochs_fp64.hip -- the FP64 multi-epoch tracer instantiations (a translation
// unit of their own, built in parallel with the others); -ffp-contract=off.
#include "kernels_impl.hpp"

namespace cor {
const void *trace_kernel_epochs_fp64(int intk, int ext, int cru) {
    return trace_kernel_epochs_table<PREC_FP64>(intk, ext, cru);
}
}  // namespace cor
