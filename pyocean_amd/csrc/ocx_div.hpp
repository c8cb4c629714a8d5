// Correctly rounded FP64 divisions from the reciprocal instruction and FMA (gfx950), bit-identical to the IEEE
// quotient in their stated ranges.
#pragma once
#include <math.h>

namespace ocx {

// Correctly rounded x / b for a constant divisor b with rb = RN(1/b): q = RN(x rb) is within an ulp of x/b, the
// remainder x - b q is exact under an FMA, and RN(q + (x - b q) rb) = RN(x/b) (Markstein's correction; needs rb
// within half an ulp of 1/b and no underflow, so tiny |x| takes the IEEE division).
__device__ __forceinline__ double div_const(double x, double b, double rb) {
    if (fabs(x) < 0x1p-960 && x != 0.0) return x / b;
    const double q = x * rb;
    const double r = __fma_rn(-b, q, x);
    return __fma_rn(r, rb, q);
}

// x / y for y > 0 by the Newton-Raphson sequence that the compiler emits for an IEEE FP64 division (reciprocal,
// two refinements, quotient, one residual correction), without the range scaling and special-value fix-up steps.
// Where those steps would not change anything (neither operand near the overflow or denormal range) the result is
// the IEEE quotient bit for bit; callers guarantee that range.
__device__ __forceinline__ double div_nr(double x, double y) {
    double r = __builtin_amdgcn_rcp(y);
    double e = __fma_rn(-y, r, 1.0);
    r = __fma_rn(r, e, r);
    e = __fma_rn(-y, r, 1.0);
    r = __fma_rn(r, e, r);
    const double q = x * r;
    const double t = __fma_rn(-y, q, x);
    return __fma_rn(t, r, q);
}

// x / y for y > 0, correctly rounded: div_nr where both operands lie within 2^-500 .. 2^500 in magnitude (x may be
// 0), so that the quotient, the reciprocal and the residual stay far from overflow and underflow; the IEEE
// division elsewhere
__device__ __forceinline__ double qdiv(double x, double y) {
    const double ax = fabs(x);
    if (y >= 0x1p-500 && y <= 0x1p500 && ax <= 0x1p500 && (ax >= 0x1p-500 || x == 0.0)) return div_nr(x, y);
    return x / y;
}

}  // namespace ocx
