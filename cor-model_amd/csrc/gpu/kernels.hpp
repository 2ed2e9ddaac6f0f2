// kernels.hpp -- kernel tables exported by the per-precision translation units.
//
// kernels_fp64.hip is compiled without FMA contraction (the reference's arithmetic);
// kernels_mixed.hip with contraction and FP32 denormal flushing.  Each exports the
// tracer instantiations for the supported model sets and a field-batch kernel.
#pragma once
#include "dev_types.hpp"

namespace cor {

struct FieldBatchArgs {
    const EpochDev *epoch;
    const double *gsm_m;    // n x 3
    double *geo_nT;         // n x 3
    int64_t n;
};

struct RkStepArgs {
    const double *h, *d;    // n step lengths [s] and 1e-9 zn q / hm
    const double *field;    // n x 3 GEO field [nT]
    double *vel, *pos;      // n x 3, advanced in place
    int64_t n;
};

// Returns the kernel for (internal, external, crustal) or nullptr when that model set
// is not compiled in.
const void *trace_kernel_fp64(int intk, int ext, int cru);
const void *trace_kernel_mixed(int intk, int ext, int cru);
// the FP64 tracer with the cooperative tail (IGRF / T96 model sets; nullptr otherwise)
const void *trace_kernel_fp64_coop(int intk, int ext, int cru);
// the multi-epoch tracer (kernels_epochs_*.hip; model sets without T96, else nullptr)
const void *trace_kernel_epochs_fp64(int intk, int ext, int cru);
const void *trace_kernel_epochs_mixed(int intk, int ext, int cru);
const void *field_kernel_fp64(int intk, int ext, int cru);
const void *field_kernel_mixed(int intk, int ext, int cru);
// Mixed-precision field batch evaluated like the tracer's cooperative mode (one
// wavefront per point), and the one-thread kernel that builds the cooperative lane
// tables of an epoch in place (EpochDev::coop).
const void *field_kernel_coop(int intk, int ext, int cru, int lanes);
const void *field_kernel_coop_fp64(int intk, int ext, int cru, int lanes);
const void *coop_build_kernel();
const void *rk_step_kernel_fp64();

}  // namespace cor
