// Device-side face reconstruction of the tracer advection (gfx950).
//
// FP64 arithmetic with the operand order of core/dynamics/sigma_baroclinic.py:204-233 (_wval): every product and
// sum rounded on its own (the library is compiled with -ffp-contract=off) and IEEE-correct divisions, so a face
// value equals the CPU reference's bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ocx_div.hpp"

namespace ocx {

constexpr unsigned OPEN_W = 1u, OPEN_E = 2u, OPEN_S = 4u, OPEN_N = 8u;
constexpr unsigned WENO_JS = 16u;      // WENO5-JS instead of WENO5-Z
constexpr unsigned UNLIMITED = 32u;    // FCT coefficients = 1 (the RK3 result)
constexpr unsigned UP3 = 64u;          // third-order upwind faces (tracer_scheme == "upwind3"), RK3 step only
constexpr unsigned UP1 = 128u;         // first-order upwind faces (tracer_scheme == "upwind1"), RK3 step only
constexpr unsigned CENTRED = 256u;     // centred 6th/4th/2nd-order faces (rhs(centred=True), the 'pc' predictor)

// Face schemes of the tile kernels: WENO5-JS, WENO5-Z, UP3, UP1, centred.  From SCH_UP3 on, the faces are those of
// the reference's vectorised rhs_numpy: every face is Q times its value (no Q == 0 test), a closed edge Q * 0.
constexpr int SCH_JS = 0, SCH_Z = 1, SCH_UP3 = 2, SCH_UP1 = 3, SCH_C6 = 4;

__device__ __forceinline__ double sq(double x) { return x * x; }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2): physical block b runs on XCD b % 8.
// This maps b to a logical index so that every XCD works through one contiguous range of logical blocks; blocks
// that are neighbours in logical order then share an L2 (a bijection for any block count).
__device__ __forceinline__ long long xcd_logical(long long b, long long nb) {
    const long long q = nb / 8, r = nb % 8, x = b % 8, s = b / 8;
    return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

// Index of the calling wavefront in its workgroup, as a wave-uniform (scalar) value: everything derived from it
// (tracer, plane base pointers) stays in scalar registers instead of being recomputed per lane.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// x / b for the WENO constant divisors, without div_const's tiny-x test: for x that is 0 or at least 2^-960 in
// magnitude (a sum of products of stencil values that are each 0 or at least 2^-400 in magnitude is a multiple of
// 2^-452, so it is either 0 or beyond that bound).
__device__ __forceinline__ double div_const_n(double x, double b, double rb) {
    const double q = x * rb;
    const double r = __fma_rn(b, q, -x);
    return __fma_rn(-r, rb, q);
}

// Division modes of the face value: IEEE divisions; the fast divisions (div_const with its tiny-x test, div_nr);
// the fast divisions when every stencil value is known to be 0 or of magnitude in [2^-400, 2^100] (no test at all).
constexpr int DIV_IEEE = 0, DIV_FAST = 1, DIV_FAST_KNOWN = 2;

// A stencil value that keeps every division of its faces on the fast path: 0, or of magnitude in [2^-400, 2^100].
__device__ __forceinline__ bool fast_value(double v) {
    const double a = fabs(v);
    return a <= 0x1p100 && (a >= 0x1p-400 || a == 0.0);
}

// The constants of the WENO5 value that are not exact 32-bit literals of an FP64 instruction.  By default they are
// compile-time constants; a kernel short of scalar registers can hold them in vector registers instead (weno_k_v).
struct WenoK {
    double r6 = 1.0 / 6.0, c1312 = 13.0 / 12.0, d0 = 0.1, d1 = 0.6, d2 = 0.3, eps_z = 1e-40, eps_js = 1e-6;
};
// the same values; those selected by SEL (bit 0: 1/6, 1: 13/12, 2: the linear weights d, 3: the two epsilons) each
// pass through an empty asm statement that places them in a vector register pair
template <unsigned SEL = 15u>
__device__ __forceinline__ WenoK weno_k_v() {
    WenoK K;
    if constexpr ((SEL & 1u) != 0) asm volatile("" : "+v"(K.r6));
    if constexpr ((SEL & 2u) != 0) asm volatile("" : "+v"(K.c1312));
    if constexpr ((SEL & 4u) != 0) {
        asm volatile("" : "+v"(K.d0));
        asm volatile("" : "+v"(K.d1));
        asm volatile("" : "+v"(K.d2));
    }
    if constexpr ((SEL & 8u) != 0) {
        asm volatile("" : "+v"(K.eps_z));
        asm volatile("" : "+v"(K.eps_js));
    }
    return K;
}

// WENO5 value from the upwind-ordered stencil v1..v5 (v3 = upwind cell).  Z: WENO5-Z (Borges et al. 2008,
// tau5 = |b2 - b0|, eps = 1e-40); otherwise WENO5-JS (Jiang & Shu 1996, eps = 1e-6).  d = (0.1, 0.6, 0.3).
// MODE: the divisions (DIV_IEEE, DIV_FAST, DIV_FAST_KNOWN); all three give identical results in their ranges.
template <bool Z, int MODE>
__device__ __forceinline__ double weno5(double v1, double v2, double v3, double v4, double v5,
                                       const WenoK& K = WenoK{}) {
    constexpr bool FAST = MODE != DIV_IEEE;
    const double r6 = K.r6;
    const double n0 = 2.0 * v1 - 7.0 * v2 + 11.0 * v3;
    const double n1 = -v2 + 5.0 * v3 + 2.0 * v4;
    const double n2 = 2.0 * v3 + 5.0 * v4 - v5;
    const double q0 = MODE == DIV_FAST_KNOWN ? div_const_n(n0, 6.0, r6) : FAST ? div_const(n0, 6.0, r6) : n0 / 6.0;
    const double q1 = MODE == DIV_FAST_KNOWN ? div_const_n(n1, 6.0, r6) : FAST ? div_const(n1, 6.0, r6) : n1 / 6.0;
    const double q2 = MODE == DIV_FAST_KNOWN ? div_const_n(n2, 6.0, r6) : FAST ? div_const(n2, 6.0, r6) : n2 / 6.0;
    const double c1312 = K.c1312;
    const double b0 = c1312 * sq(v1 - 2.0 * v2 + v3) + 0.25 * sq(v1 - 4.0 * v2 + 3.0 * v3);
    const double b1 = c1312 * sq(v2 - 2.0 * v3 + v4) + 0.25 * sq(v2 - v4);
    const double b2 = c1312 * sq(v3 - 2.0 * v4 + v5) + 0.25 * sq(3.0 * v3 - 4.0 * v4 + v5);
    double a0, a1, a2;
    if (Z) {
        const double tau = fabs(b2 - b0);
        const double e = K.eps_z;
        a0 = K.d0 * (1.0 + (FAST ? div_nr(tau, e + b0) : tau / (e + b0)));
        a1 = K.d1 * (1.0 + (FAST ? div_nr(tau, e + b1) : tau / (e + b1)));
        a2 = K.d2 * (1.0 + (FAST ? div_nr(tau, e + b2) : tau / (e + b2)));
    } else {
        const double e = K.eps_js;
        a0 = FAST ? div_nr(K.d0, sq(e + b0)) : K.d0 / sq(e + b0);
        a1 = FAST ? div_nr(K.d1, sq(e + b1)) : K.d1 / sq(e + b1);
        a2 = FAST ? div_nr(K.d2, sq(e + b2)) : K.d2 / sq(e + b2);
    }
    const double num = a0 * q0 + a1 * q1 + a2 * q2, den = a0 + a1 + a2;
    return FAST ? div_nr(num, den) : num / den;
}

// WENO3 from v2..v4 (candidates (-v2 + 3 v3)/2 and (v3 + v4)/2, d = (1/3, 2/3), eps = 1e-6).
template <bool FAST>
__device__ __forceinline__ double weno3(double v2, double v3, double v4) {
    const double c0 = (-v2 + 3.0 * v3) / 2.0, c1 = (v3 + v4) / 2.0;
    const double y0 = sq(1e-6 + sq(v3 - v2)), y1 = sq(1e-6 + sq(v4 - v3));
    const double e0 = FAST ? div_nr(1.0 / 3.0, y0) : (1.0 / 3.0) / y0;
    const double e1 = FAST ? div_nr(2.0 / 3.0, y1) : (2.0 / 3.0) / y1;
    const double num = e0 * c0 + e1 * c1, den = e0 + e1;
    return FAST ? div_nr(num, den) : num / den;
}

// Face value with the validity rule of the reference: WENO5 when all five cells are valid, WENO3 when v2..v4
// are, else the upwind cell.  Cells outside the domain carry value 0 and validity 0.  The fast divisions are used
// when the stencil's largest magnitude m lies in [2^-400, 2^100] (or m = 0): then no operand of a division is near
// the overflow or denormal range (the smoothness indicators stay below 2^210, the weight sums above 2^-430, a
// nonzero numerator above 2^-520), which is where the plain sequence and the IEEE division can differ.  known: the
// caller has checked that every value of the stencil is 0 or in that range (fast_value), which implies that test.
template <bool Z>
__device__ __forceinline__ double face_value(double v1, double v2, double v3, double v4, double v5,
                                             bool k1, bool k2, bool k3, bool k4, bool k5, bool known = false,
                                             const WenoK& K = WenoK{}) {
    if (known) {
        if (k1 && k2 && k3 && k4 && k5) return weno5<Z, DIV_FAST_KNOWN>(v1, v2, v3, v4, v5, K);
        if (k2 && k3 && k4) return weno3<true>(v2, v3, v4);
        return v3;
    }
    const double m = fmax(fmax(fmax(fabs(v1), fabs(v2)), fmax(fabs(v3), fabs(v4))), fabs(v5));
    const bool fast = m <= 0x1p100 && (m >= 0x1p-400 || m == 0.0);
    if (k1 && k2 && k3 && k4 && k5)
        return fast ? weno5<Z, DIV_FAST>(v1, v2, v3, v4, v5, K) : weno5<Z, DIV_IEEE>(v1, v2, v3, v4, v5, K);
    if (k2 && k3 && k4) return fast ? weno3<true>(v2, v3, v4) : weno3<false>(v2, v3, v4);
    return v3;
}

// Face value between cells c-1 and c of a line held in LDS (values s, validity o, element stride st), upwind
// side from the sign of q (q >= 0: low side).  The line must hold valid indices c-3 .. c+2.  known: as for
// face_value, every value of the line's stencil is 0 or of magnitude in [2^-400, 2^100].
template <bool Z>
__device__ __forceinline__ double lds_face(const double* s, const uint8_t* o, int st, int c, double q,
                                           bool known = false) {
    const bool pos = q >= 0.0;
    const int i1 = pos ? c - 3 : c + 2, i2 = pos ? c - 2 : c + 1, i3 = pos ? c - 1 : c;
    const int i4 = pos ? c : c - 1, i5 = pos ? c + 1 : c - 2;
    return face_value<Z>(s[i1 * st], s[i2 * st], s[i3 * st], s[i4 * st], s[i5 * st],
                         o[i1 * st] != 0, o[i2 * st] != 0, o[i3 * st] != 0, o[i4 * st] != 0, o[i5 * st] != 0,
                         known);
}

// UP3 / UP1 face value between cells c-1 and c of a line (values s, validity o, stride st), _up3_faces
// (sigma_baroclinic.py:164-187): q >= 0 takes (-c[i-2] + 5 c[i-1] + 2 c[i]) / 6 where cells i-2 .. i are ok, else
// c[i-1]; q < 0 (or NaN) takes (2 c[i-1] + 5 c[i] - c[i+1]) / 6 where cells i-1 .. i+1 are ok, else c[i].  UP1 is
// the donor value.  The line must hold valid indices c-2 .. c+1 (validity 0 outside the domain).
template <bool O3>
__device__ __forceinline__ double up_face(const double* s, const uint8_t* o, int st, int c, double q) {
    const double lo = s[(c - 1) * st], hi = s[c * st];
    if (!O3) return q >= 0.0 ? lo : hi;
    if (q >= 0.0)
        return (o[(c - 2) * st] && o[(c - 1) * st] && o[c * st]) ? (-s[(c - 2) * st] + 5.0 * lo + 2.0 * hi) / 6.0 : lo;
    return (o[(c - 1) * st] && o[c * st] && o[(c + 1) * st]) ? (2.0 * lo + 5.0 * hi - s[(c + 1) * st]) / 6.0 : hi;
}

// Centred face value between cells c-1 and c of a line (values s, validity o, stride st), _centred_faces
// (sigma_baroclinic.py:632-652): (37 (c[i-1] + c[i]) - 8 (c[i-2] + c[i+1]) + (c[i-3] + c[i+2])) / 60 where cells
// i-3 .. i+2 are ok, else (7 (c[i-1] + c[i]) - (c[i-2] + c[i+1])) / 12 where cells i-2 .. i+1 are ok, else
// (c[i-1] + c[i]) / 2, whatever the flux.  The line must hold valid indices c-3 .. c+2 (validity 0 outside).
__device__ __forceinline__ double centred_face(const double* s, const uint8_t* o, int st, int c) {
    const double a = s[(c - 1) * st] + s[c * st];
    const bool m4 = o[(c - 2) * st] && o[(c - 1) * st] && o[c * st] && o[(c + 1) * st];
    if (m4 && o[(c - 3) * st] && o[(c + 2) * st])
        return (37.0 * a - 8.0 * (s[(c - 2) * st] + s[(c + 1) * st]) + (s[(c - 3) * st] + s[(c + 2) * st])) / 60.0;
    if (m4) return (7.0 * a - (s[(c - 2) * st] + s[(c + 1) * st])) / 12.0;
    return 0.5 * a;
}

}  // namespace ocx
