// kernels_fp64.hip -- FP64 (parity) instantiations; built with -ffp-contract=off.
#include "kernels_impl.hpp"

namespace cor {
const void *trace_kernel_fp64(int intk, int ext, int cru) {
    return trace_kernel_table<PREC_FP64>(intk, ext, cru);
}
const void *trace_kernel_fp64_coop(int intk, int ext, int cru) {
    return trace_kernel_coop_table<PREC_FP64>(intk, ext, cru);
}
const void *field_kernel_fp64(int intk, int ext, int cru) {
    return field_kernel_table<PREC_FP64>(intk, ext, cru);
}
const void *field_kernel_coop_fp64(int intk, int ext, int cru, int lanes) {
    return field_coop_table_lanes<PREC_FP64>(intk, ext, cru, lanes);
}

// One RK step per state with the tracer's FP64 step (src/trajectory.c:275-356).
__global__ void __launch_bounds__(256) rk_step_kernel(RkStepArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    v3<double> vel = mk3(a.vel[3 * i], a.vel[3 * i + 1], a.vel[3 * i + 2]);
    v3<double> pos = mk3(a.pos[3 * i], a.pos[3 * i + 1], a.pos[3 * i + 2]);
    rk_step_fp64(a.h[i], a.d[i], mk3(a.field[3 * i], a.field[3 * i + 1], a.field[3 * i + 2]), vel, pos);
    a.vel[3 * i] = vel.x;
    a.vel[3 * i + 1] = vel.y;
    a.vel[3 * i + 2] = vel.z;
    a.pos[3 * i] = pos.x;
    a.pos[3 * i + 1] = pos.y;
    a.pos[3 * i + 2] = pos.z;
}
const void *rk_step_kernel_fp64() { return (const void *)&rk_step_kernel; }
}  // namespace cor
