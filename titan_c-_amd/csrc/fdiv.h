// fdiv.h -- correctly rounded FP64 division by a denominator whose correctly rounded reciprocal
// is already known: a constant, or a value several quotients share.
//
// The IEEE division sequence on gfx950 costs about ten VALU instructions (two v_div_scale, a
// reciprocal seed, Newton steps, v_div_fmas, v_div_fixup).  With y = RN(1/b) known, Markstein's
// correction gives the same correctly rounded quotient in three: q = RN(a y) lies within one ulp
// of a/b, the remainder r = a - b q is exact when formed with one FMA, and RN(q + r y) = RN(a/b)
// (Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point Arithmetic,
// 2nd ed., sec. 4.7).  A zero remainder returns q unchanged, which also gives a zero quotient the
// sign IEEE division gives it.  The theorem needs a, b and a/b in the normal range and finite;
// callers use it only where the denominator is a positive constant or a quantity the reference's
// own logic keeps finite and away from zero (volumes, densities, sums of square roots).
#pragma once
#include <hip/hip_runtime.h>

namespace titan {

__host__ __device__ __forceinline__ double div_by(double a, double b, double y) {
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return r == 0.0 ? q : __builtin_fma(r, y, q);
}

// a / b through the reciprocal for FP64 (y = RN(1/b)); plain division for other types, whose
// units are compiled with fast division already.
template <typename T>
__host__ __device__ __forceinline__ T div_by_t(T a, T b, T y) {
    if constexpr (sizeof(T) == sizeof(double)) return div_by(a, b, y);
    else { (void)y; return a / b; }
}


}  // namespace titan
