// Device helpers shared by the field kernels. Internal to libfarrsight_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fs {
namespace dev {

// 1/sqrt(a), a > 0: hardware estimate (v_rsq_f64, relative error about
// 2^-23) plus one third-order correction y0 (1 + e/2 + 3e^2/8), e = 1 - a y0^2,
// which leaves a truncation error of order e^3 below FP64 rounding.
__device__ __forceinline__ double rsq_f64(double a) {
    const double y0 = __builtin_amdgcn_rsq(a);
    const double t = a * y0;
    const double e = __builtin_fma(-t, y0, 1.0);
    const double r = y0 * e;
    const double p = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(r, p, y0);
}

// 1/sqrt(a): v_rsq_f64 and one Newton step (relative error ~1e-13).
__device__ __forceinline__ double rsq_newton(double a) {
    const double y0 = __builtin_amdgcn_rsq(a);
    const double t = a * y0;
    const double e = __builtin_fma(-t, y0, 1.0);
    const double h = 0.5 * y0;
    return __builtin_fma(h, e, y0);
}

// s + lo += b, error-free (Knuth TwoSum).
__device__ __forceinline__ void two_sum_acc(double& s, double& lo, double b) {
    const double t = s + b;
    const double bp = t - s;
    const double ap = t - bp;
    lo += (s - ap) + (b - bp);
    s = t;
}

// Does the cyclic interval [a, a + len] (mod 2^64) meet [0, width)?
__device__ __forceinline__ bool cyc_meets(uint64_t a, uint64_t len, uint64_t width) {
    return a < width || a + len < a;
}

}  // namespace dev
}  // namespace fs
