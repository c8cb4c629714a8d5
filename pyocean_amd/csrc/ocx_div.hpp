// Correctly rounded FP64 divisions from the reciprocal instruction and FMA (gfx950), bit-identical to the IEEE
// quotient in their stated ranges.
#pragma once
#include <float.h>
#include <math.h>

namespace ocx {

// Correctly rounded x / b for a constant divisor b with rb = RN(1/b) and a finite x: q = RN(x rb) is within an ulp
// of x/b, the remainder b q - x is exact under an FMA, and RN(q - (b q - x) rb) = RN(x/b) (Markstein's correction;
// needs rb within half an ulp of 1/b and no underflow, so tiny nonzero |x| takes the IEEE division).  The remainder
// is formed as b q - x and subtracted, so a zero x keeps its sign: RN(-0 + -0) = -0, while x - b q would give +0
// for x = -0.  A NaN x gives a NaN; an infinite x gives a NaN where the IEEE quotient is infinite (the callers
// divide finite sums of finite fluxes).
__device__ __forceinline__ double div_const(double x, double b, double rb) {
    if (fabs(x) < 0x1p-960 && x != 0.0) return x / b;
    const double q = x * rb;
    const double r = __fma_rn(b, q, -x);
    return __fma_rn(-r, rb, q);
}

// x / y for y > 0 by the Newton-Raphson sequence that the compiler emits for an IEEE FP64 division (reciprocal,
// two refinements, quotient, one residual correction), without the range scaling and special-value fix-up steps.
// Where those steps would not change anything (neither operand near the overflow or denormal range, x finite) the
// result is the IEEE quotient bit for bit; callers guarantee that range.  A zero x gives the zero of its own sign
// (the residual is formed as y q - x and subtracted, as in div_const).
__device__ __forceinline__ double div_nr(double x, double y) {
    double r = __builtin_amdgcn_rcp(y);
    double e = __fma_rn(-y, r, 1.0);
    r = __fma_rn(r, e, r);
    e = __fma_rn(-y, r, 1.0);
    r = __fma_rn(r, e, r);
    const double q = x * r;
    const double t = __fma_rn(y, q, -x);
    return __fma_rn(-t, r, q);
}

// x / y for y > 0, correctly rounded: div_nr where both operands lie within 2^-500 .. 2^500 in magnitude (x may be
// +-0), so that the quotient, the reciprocal and the residual stay far from overflow and underflow; the IEEE
// division elsewhere (also for infinite or NaN operands)
__device__ __forceinline__ double qdiv(double x, double y) {
    const double ax = fabs(x);
    if (y >= 0x1p-500 && y <= 0x1p500 && ax <= 0x1p500 && (ax >= 0x1p-500 || x == 0.0)) return div_nr(x, y);
    return x / y;
}

// qdiv with the divisor's reciprocal refined ahead of the dividend: recip_of(y) holds div_nr's refined reciprocal
// of y and whether y lies in qdiv's range, and qdiv_r(x, R) is then qdiv(x, y) bit for bit (the same operations on
// the same operands).  Where y is known before x (a pivot loaded ahead of a recurrence), only the quotient and its
// residual correction wait on x: three dependent operations instead of the IEEE division's sequence.
struct Recip {
    double y, r;
    bool ok;
};
__device__ __forceinline__ Recip recip_of(double y) {
    double r = __builtin_amdgcn_rcp(y);
    double e = __fma_rn(-y, r, 1.0);
    r = __fma_rn(r, e, r);
    e = __fma_rn(-y, r, 1.0);
    r = __fma_rn(r, e, r);
    return Recip{y, r, y >= 0x1p-500 && y <= 0x1p500};
}
__device__ __forceinline__ double qdiv_r(double x, const Recip& R) {
    const double ax = fabs(x);
    if (R.ok && ax <= 0x1p500 && (ax >= 0x1p-500 || x == 0.0)) {
        const double q = x * R.r;
        const double t = __fma_rn(R.y, q, -x);
        return __fma_rn(-t, R.r, q);
    }
    return x / R.y;
}

}  // namespace ocx
