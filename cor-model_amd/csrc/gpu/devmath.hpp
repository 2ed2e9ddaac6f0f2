// devmath.hpp -- precision traits for the field evaluators and the tracer (gfx950).
//
// Every evaluator is written once as a template over the arithmetic type T and uses
// only the functions below for anything beyond + - * /.
//   * Dm<double>: the OCML double routines (IEEE sqrt and division), used by the FP64
//     path, which keeps the reference's expression order and is compiled without FMA
//     contraction.
//   * Dm<float>: the CDNA transcendental unit directly -- v_exp_f32, v_log_f32,
//     v_sin_f32/v_cos_f32 (argument in revolutions), v_sqrt_f32, v_rcp_f32 -- and
//     integer powers by multiplication instead of pow().  Used by the mixed path.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace cor {

template <class T>
struct v3 {
    T x, y, z;
};

template <class T>
__device__ __forceinline__ v3<T> mk3(T x, T y, T z) {
    v3<T> r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}

template <class T>
struct Dm;

template <>
struct Dm<double> {
    typedef double T;
    static constexpr bool fp64 = true;
    __device__ static __forceinline__ T sqrt(T x) { return ::sqrt(x); }
    __device__ static __forceinline__ T exp(T x) { return ::exp(x); }
    __device__ static __forceinline__ T log(T x) { return ::log(x); }
    __device__ static __forceinline__ T sin(T x) { return ::sin(x); }
    __device__ static __forceinline__ T cos(T x) { return ::cos(x); }
    __device__ static __forceinline__ void sincos(T x, T *s, T *c) {
        *s = ::sin(x);
        *c = ::cos(x);
    }
    __device__ static __forceinline__ T atan(T x) { return ::atan(x); }
    __device__ static __forceinline__ T atan2(T y, T x) { return ::atan2(y, x); }
    __device__ static __forceinline__ T asin(T x) { return ::asin(x); }
    __device__ static __forceinline__ T pow(T x, T y) { return ::pow(x, y); }
    // x^n for a small integer n, as pow(x, (double)n) like the reference.
    __device__ static __forceinline__ T powi(T x, int n) { return ::pow(x, (double)n); }
    __device__ static __forceinline__ T fabs(T x) { return ::fabs(x); }
    __device__ static __forceinline__ bool isnan(T x) { return __builtin_isnan(x); }
};

template <>
struct Dm<float> {
    typedef float T;
    static constexpr bool fp64 = false;
    __device__ static __forceinline__ T sqrt(T x) { return __builtin_amdgcn_sqrtf(x); }
    // exp(x) = 2^n * 2^r with n = rint(x log2 e) and r = x log2 e - n formed with a
    // two-part log2 e (Cody-Waite), so v_exp_f32 only sees |r| <= 1/2: ~1 ulp instead
    // of the |x log2 e| * 2^-24 relative error of exp2(x * log2e).  The shielding
    // fields sum exponentials with coefficients 1e4-1e5 times their result, where that
    // difference is visible.
    __device__ static __forceinline__ T exp(T x) {
        const float l2e_hi = 1.44269502162933349609375f;   // float(log2 e)
        const float l2e_lo = 1.925963033500011079e-08f;    // log2 e - l2e_hi
        const float n = __builtin_rintf(x * l2e_hi);
        float r = __builtin_fmaf(x, l2e_hi, -n);
        r = __builtin_fmaf(x, l2e_lo, r);
        return __builtin_ldexpf(__builtin_amdgcn_exp2f(r), (int)n);
    }
    __device__ static __forceinline__ T log(T x) {
        return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
    }
    // v_sin/v_cos take revolutions; every argument in the models is O(10) rad.
    __device__ static __forceinline__ T sin(T x) {
        return __builtin_amdgcn_sinf(x * 0.159154943091895336f);
    }
    __device__ static __forceinline__ T cos(T x) {
        return __builtin_amdgcn_cosf(x * 0.159154943091895336f);
    }
    __device__ static __forceinline__ void sincos(T x, T *s, T *c) {
        const float rev = x * 0.159154943091895336f;
        *s = __builtin_amdgcn_sinf(rev);
        *c = __builtin_amdgcn_cosf(rev);
    }
    __device__ static __forceinline__ T atan(T x) { return ::atanf(x); }
    __device__ static __forceinline__ T atan2(T y, T x) { return ::atan2f(y, x); }
    __device__ static __forceinline__ T asin(T x) { return ::asinf(x); }
    // x^y for x > 0 through the hardware log2/exp2
    __device__ static __forceinline__ T pow(T x, T y) {
        return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
    }
    __device__ static __forceinline__ T powi(T x, int n) {
        T r = 1.0f;
        for (int i = 0; i < n; i++) r *= x;
        return r;
    }
    __device__ static __forceinline__ T fabs(T x) { return ::fabsf(x); }
    __device__ static __forceinline__ bool isnan(T x) { return __builtin_isnan(x); }
};

}  // namespace cor
