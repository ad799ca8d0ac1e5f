// fsqrt.h -- correctly rounded FP64 square root without the rescaling of tiny arguments (device
// code only).
#pragma once
#include <hip/hip_runtime.h>

namespace titan {

// Correctly rounded FP64 square root for arguments of at least 2^-767 (and +-0, +inf): the
// instruction sequence the compiler emits for sqrt -- a v_rsq_f64 seed, one coupled Newton step
// on (g, h) ~ (sqrt x, 1 / (2 sqrt x)) and two remainder corrections -- without its rescaling of
// tiny arguments (a shift by 0 at and above 2^-767).  Every square root on the step path takes a
// squared length, density, pressure or wave-speed term, many orders of magnitude above that.
__device__ __forceinline__ double sqrt_n(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return __builtin_amdgcn_class(x, 0x260) ? x : g;  // +-0 and +inf are their own square roots
}

template <typename T>
__device__ __forceinline__ T sqrt_t(T x) {
    if constexpr (sizeof(T) == sizeof(double)) return sqrt_n(x);
    else return sqrt(x);
}

}  // namespace titan
