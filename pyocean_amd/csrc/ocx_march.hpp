// Wave-marching kernels of the FCT step (gfx950).
//
// One 64-lane wavefront owns a strip of 63 cells in x of one (tracer, layer) plane and marches down a chunk of
// rows.  Lane l sits on column i0 + l: it reconstructs the x-face i0 + l (64 faces for the strip's 63 cells, so
// every lane is busy and no face is computed twice inside a strip) and the y-face north of its cell.  The y stencil
// (rows j-3 .. j+3 of the lane's column) lives in registers and slides one row per step, so each input row is read
// from HBM once; the x stencil of the current row goes through a 69-entry LDS row private to the wave (no block
// barrier).  The east face of a cell comes from lane + 1 (cross-lane shuffle), the south face from the previous
// step.  Faces are written to global memory only by their owner (x-faces i0 .. i0+62 and the domain's last face;
// y-faces jb .. je-1 of the chunk and the domain's last face), so a read-modify-write of a face stack never races.
//
// Stages (core/components/tracers.py:208-254 with the kernels of core/dynamics/sigma_baroclinic.py:265-530):
//   1: F0 = faces(C) -> fx/fy;  c1 = (q0 - dt div F0)/Va (safe)
//   2: F1 = faces(c1);  S = F0 + F1 -> s_x/s_y (F0 stays read-only: a wave reads F0 on faces its neighbours own);
//      q1 recomputed from F0 (bit-identical to stage 1's), c2 = (0.75 q0 + 0.25 (q1 - dt div F1))/Vb (safe)
//   3: F2 = faces(c2);  F_H = (F0 + F1)/6 + 2 F2/3;  donor F_L;  A = F_H - F_L -> ax/ay;  c_L = (q0 - dt div F_L)/V1
//      (safe) -> cout.  UNLIM: out = (q0 - dt div F_H)/V1 (safe) -> cout and the face column of F_H.
#pragma once
#include <type_traits>

#include "ocx_weno.hpp"

namespace ocx {

constexpr int MS = 63;     // own cells per strip
constexpr int MRB = 72;    // LDS row: columns i0-3 .. i0+65 (69 used)

struct MarchArgs {
    const void* cin;       // the stage's input values: C (double), or c1, c2 (the precision's stage type)
    const double* C;       // the step's initial values
    const double* ext;     // inflow values, or nullptr (stage's own values; C for the donor flux)
    const double* V0;
    const double* V1;
    const double* Qx;
    const double* Qy;
    const uint8_t* ok;
    void* fx;              // F0 (stage 1 out, stage 2 in; read-only in stage 2)      [flux type]
    void* fy;
    void* s_x;             // S = F0 + F1 (stage 2 out, stage 3 in)                     [flux type]
    void* s_y;
    void* ax;              // A (stage 3; may be the F0 buffers, dead by then)         [flux type]
    void* ay;
    void* cout;            // c1, c2 [stage type]; c_L or the unlimited result [double]
    double* fxi;           // unlimited: face column of F_H
    int flux_face_i;
    int nt, n, ny, nx;
    int rch;               // rows per chunk
    int nsx, ncy;          // strips in x, chunks in y
    double dt;
    unsigned flags;
};

// Neighbour-lane moves by DPP wavefront shifts (a VALU move, no LDS round trip): lane l receives lane l+1
// (shfl_down1) or lane l-1 (shfl_up1); the lane without a source keeps its own value.  Every lane of the wave must
// execute them (they sit outside divergent branches).
__device__ __forceinline__ int dpp_down1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_up1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ double shfl_down1(double v) {
    return __hiloint2double(dpp_down1(__double2hiint(v)), dpp_down1(__double2loint(v)));
}
__device__ __forceinline__ double shfl_up1(double v) {
    return __hiloint2double(dpp_up1(__double2hiint(v)), dpp_up1(__double2loint(v)));
}

// The stage value c = q / V (V > 0) as the precision stores it: FP64 keeps the IEEE quotient; the split form, which
// stores c in FP32, multiplies by the reciprocal refined once by Newton-Raphson (relative error near 2^-50 before
// the rounding to FP32, so the stored value is the FP32 rounding of q / V up to one ulp in rare ties).
template <class PR>
__device__ __forceinline__ typename PR::st stage_value(double q, double V, double c0) {
    if (!(V > 0)) return (typename PR::st)c0;
    if constexpr (PR::split) {
        double r = __builtin_amdgcn_rcp(V);
        r = __fma_rn(r, __fma_rn(-V, r, 1.0), r);
        return (float)(q * r);
    } else {
        return q / V;
    }
}

// A load or store at a byte offset from a wave-uniform base: the address is the scalar base plus a 32-bit lane
// offset, so no 64-bit address arithmetic runs on the vector ALUs.  Every offset passed here lies inside the array
// (rows and columns outside the plane are clamped to row 0 / column 0 and their values are never used).
template <class T>
__device__ __forceinline__ T ldo(const T* base, uint32_t off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <class T>
__device__ __forceinline__ void sto(T* base, uint32_t off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}

// Precision of a step.  F64: everything in FP64, bit-identical to the reference.  Split: stage values c1, c2 and
// the face fluxes F0, F0 + F1, A stored in FP32 and the WENO reconstruction (smoothness indicators, weights, face
// values, fluxes) computed in FP32; the h*T accumulation q = a q0 + b (q' - dt div F), c_L, the FCT ratios, the
// limited update and the content sums stay FP64.
struct PrecF64 {
    using st = double;     // stored stage values
    using fl = double;     // stored face fluxes
    using w = double;      // reconstruction arithmetic
    static constexpr bool split = false;
};
struct PrecSplit {
    using st = float;
    using fl = float;
    using w = float;
    static constexpr bool split = true;
};

__device__ __forceinline__ float sqf(float x) { return x * x; }

// FP32 WENO5 value from the upwind-ordered stencil, written on the first differences e1..e4 of v1..v5 (the same
// scheme up to rounding, with less cancellation and fewer operations):
//   candidates q_k = v3 + p_k / 6, p0 = 5 e2 - 2 e1, p1 = e2 + 2 e3, p2 = 4 e3 - e4;
//   4 beta_k = (13/3) s_k^2 + r_k^2 with second differences s_k and r0 = s0 + 2 e2, r1 = e2 + e3, r2 = s2 - 2 e3;
//   value = v3 + (sum alpha_k p_k) / (6 sum alpha_k).
// The weights depend on beta only through ratios, so 4 beta, 4 tau and 4 eps (Z) or (4e-6 + 4 beta)^2 (JS) give the
// reference's weights.  WENO-Z's eps = 1e-40 is below the FP32 normal range, so eps = 1e-30 (1 + m^2) with m the
// stencil's largest magnitude: it only keeps a vanishing indicator from dividing by zero, and tau/(eps + beta) stays
// below about 1e31 whatever the data.  Hardware reciprocals, contraction allowed.
struct Diff5 {
    float e1, e2, e3, e4;
};
__device__ __forceinline__ Diff5 diff5(float v1, float v2, float v3, float v4, float v5) {
    return Diff5{v2 - v1, v3 - v2, v4 - v3, v5 - v4};
}

template <bool Z>
__device__ __forceinline__ float weno5f(float v1, float v2, float v3, float v4, float v5, const Diff5& d) {
#pragma clang fp contract(fast)
    const float s0 = d.e2 - d.e1, s1 = d.e3 - d.e2, s2 = d.e4 - d.e3;
    const float r0 = s0 + 2.0f * d.e2, r1 = d.e2 + d.e3, r2 = s2 - 2.0f * d.e3;
    constexpr float c133 = 13.0f / 3.0f;
    const float b0 = c133 * s0 * s0 + r0 * r0;
    const float b1 = c133 * s1 * s1 + r1 * r1;
    const float b2 = c133 * s2 * s2 + r2 * r2;
    const float p0 = 5.0f * d.e2 - 2.0f * d.e1, p1 = d.e2 + 2.0f * d.e3, p2 = 4.0f * d.e3 - d.e4;
    float a0, a1, a2;
    if (Z) {
        const float m = fmaxf(fmaxf(fmaxf(fabsf(v1), fabsf(v2)), fmaxf(fabsf(v3), fabsf(v4))), fabsf(v5));
        const float eps = 4e-30f * (1.0f + m * m);
        const float tau = fabsf(b2 - b0);
        a0 = 0.1f + 0.1f * tau * __builtin_amdgcn_rcpf(eps + b0);
        a1 = 0.6f + 0.6f * tau * __builtin_amdgcn_rcpf(eps + b1);
        a2 = 0.3f + 0.3f * tau * __builtin_amdgcn_rcpf(eps + b2);
    } else {
        a0 = 0.1f * __builtin_amdgcn_rcpf(sqf(4e-6f + b0));
        a1 = 0.6f * __builtin_amdgcn_rcpf(sqf(4e-6f + b1));
        a2 = 0.3f * __builtin_amdgcn_rcpf(sqf(4e-6f + b2));
    }
    return v3 + (a0 * p0 + a1 * p1 + a2 * p2) * (__builtin_amdgcn_rcpf(a0 + a1 + a2) * (1.0f / 6.0f));
}

// WENO3 from v2..v4: candidates (-v2 + 3 v3)/2 = v3 + e2/2 and (v3 + v4)/2 = v3 + e3/2, d = (1/3, 2/3),
// eps = 1e-6; value = v3 + (w0 e2 + w1 e3) / 2.
__device__ __forceinline__ float weno3f(float v3, const Diff5& d) {
#pragma clang fp contract(fast)
    const float g0 = (1.0f / 3.0f) * __builtin_amdgcn_rcpf(sqf(1e-6f + d.e2 * d.e2));
    const float g1 = (2.0f / 3.0f) * __builtin_amdgcn_rcpf(sqf(1e-6f + d.e3 * d.e3));
    return v3 + (g0 * d.e2 + g1 * d.e3) * (0.5f * __builtin_amdgcn_rcpf(g0 + g1));
}

template <bool Z>
__device__ __forceinline__ float face_value(float v1, float v2, float v3, float v4, float v5, bool k1, bool k2,
                                            bool k3, bool k4, bool k5, bool = false, const WenoK& = WenoK{}) {
    const Diff5 d = diff5(v1, v2, v3, v4, v5);
    if (k1 && k2 && k3 && k4 && k5) return weno5f<Z>(v1, v2, v3, v4, v5, d);
    if (k2 && k3 && k4) return weno3f(v3, d);
    return v3;
}

// WENO/edge face flux of a face whose upwind-ordered candidates are u[0..5] = cells f-3 .. f+2 along the axis
// (validity o[0..5]); lo = cell f-1, hi = cell f; rules of _advect_faces_h.  W: the reconstruction type (the flux
// q * value is formed in it too).
template <bool Z, class W>
__device__ __forceinline__ W line_face(double qd, int f, int nf, const W* u, const bool* o, bool open_lo,
                                       bool open_hi, W ext_lo, W ext_hi, bool known = false,
                                       const WenoK& K = WenoK{}) {
    const W q = (W)qd;
    if (f > 0 && f < nf) {
        if (qd == 0.0) return (W)0;
        const bool pos = qd >= 0.0;
        const W v1 = pos ? u[0] : u[5], v2 = pos ? u[1] : u[4], v3 = pos ? u[2] : u[3];
        const W v4 = pos ? u[3] : u[2], v5 = pos ? u[4] : u[1];
        const bool k1 = pos ? o[0] : o[5], k2 = pos ? o[1] : o[4], k3 = pos ? o[2] : o[3];
        const bool k4 = pos ? o[3] : o[2], k5 = pos ? o[4] : o[1];
        return q * face_value<Z>(v1, v2, v3, v4, v5, k1, k2, k3, k4, k5, known, K);
    }
    if (f == 0) return open_lo ? q * (qd > 0.0 ? ext_lo : u[3]) : (W)0;
    return open_hi ? q * (qd < 0.0 ? ext_hi : u[2]) : (W)0;
}

// donor-cell flux of a face between cells lo and hi (interior), or of an edge face (rules of _fct_fluxes); the
// exterior values are read (ext_lo[0] / ext_hi[0]) only on an open edge face
__device__ __forceinline__ double donor(double q, int f, int nf, double lo, double hi, bool open_lo, bool open_hi,
                                        const double* ext_lo, const double* ext_hi) {
    if (f > 0 && f < nf) return q * (q >= 0 ? lo : hi);
    if (f == 0) return open_lo ? q * (q > 0 ? ext_lo[0] : hi) : 0.0;
    return open_hi ? q * (q < 0 ? ext_hi[0] : lo) : 0.0;
}

// Work decomposition: a workgroup holds the waves of up to MAXW tracers of one (layer, strip, chunk), so the
// per-layer fields (Qx, Qy, V0, V1, ok) they all read come from the compute unit's caches after the first wave;
// logical blocks run strip-fastest, each XCD through a contiguous range.
#ifndef OCX_MAXW
#define OCX_MAXW 4
#endif
constexpr int MAXW = OCX_MAXW;

// Logical blocks of one marching launch over a domain: tracer groups x strips x chunks x layers.
__host__ __device__ inline long long march_blocks(const MarchArgs& a, int wpb) {
    return (long long)((a.nt + wpb - 1) / wpb) * a.nsx * a.ncy * a.n;
}

// One logical block L of a marching stage (wave wv of wpb).  rbuf/obuf: this wave's LDS row.
template <int STAGE, bool Z, bool UNLIM, class PR>
__device__ __forceinline__ void march_block(const MarchArgs& a, long long L, int wv, int wpb, typename PR::w* rbw,
                                            uint8_t* obw) {
    using W = typename PR::w;
    using ST = typename PR::st;
    using FL = typename PR::fl;
    using IN = typename std::conditional<STAGE == 1, double, ST>::type;   // type of the stage's input values
    const int lane = threadIdx.x & 63;
    const int ntb = (a.nt + wpb - 1) / wpb;
    const int tg = (int)(L % ntb);
    L /= ntb;
    const int sx = (int)(L % a.nsx);
    L /= a.nsx;
    const int cy = (int)(L % a.ncy);
    const int k = (int)(L / a.ncy);
    const int t = tg * wpb + wv;
    if (t >= a.nt || k >= a.n) return;            // whole wave leaves; no block barrier is used below
    const int p = t * a.n + k;
    const int ny = a.ny, nx = a.nx;
    const int i0 = sx * MS, i = i0 + lane;
    const int jb = cy * a.rch, je = min(jb + a.rch, ny);
    const size_t plane = (size_t)ny * nx;
    const IN* cin = static_cast<const IN*>(a.cin) + (size_t)p * plane;
    const double* Cp = a.C + (size_t)p * plane;
    const double* extp = a.ext ? a.ext + (size_t)p * plane : nullptr;   // stage faces: nullptr = cin
    const double* ext0 = a.ext ? a.ext + (size_t)p * plane : Cp;        // donor flux: inflow or C
    const uint8_t* okp = a.ok + (size_t)k * plane;
    const double* Qxk = a.Qx + (size_t)k * ny * (nx + 1);
    const double* Qyk = a.Qy + (size_t)k * (ny + 1) * nx;
    const double* V0k = a.V0 + (size_t)k * plane;
    const double* V1k = a.V1 + (size_t)k * plane;
    FL* fxp = static_cast<FL*>(a.fx) + (size_t)p * ny * (nx + 1);
    FL* fyp = static_cast<FL*>(a.fy) + (size_t)p * (ny + 1) * nx;
    FL* sxp = static_cast<FL*>(a.s_x) + (size_t)p * ny * (nx + 1);
    FL* syp = static_cast<FL*>(a.s_y) + (size_t)p * (ny + 1) * nx;
    FL* axp = static_cast<FL*>(a.ax) + (size_t)p * ny * (nx + 1);
    FL* ayp = static_cast<FL*>(a.ay) + (size_t)p * (ny + 1) * nx;
    ST* coutp = static_cast<ST*>(a.cout) + (size_t)p * plane;     // stages 1 and 2
    const bool ow = a.flags & OPEN_W, oe = a.flags & OPEN_E, os = a.flags & OPEN_S, on = a.flags & OPEN_N;
    const bool col_in = i < nx;                     // lane's column is a cell
    const bool own = lane < MS && col_in;           // lane's cell belongs to this strip
    const bool xface = i <= nx;                     // lane's x-face exists
    const bool xown = xface && (lane < MS || i == nx);
    // halo column of lanes 0..4 in the LDS row: i0-3, i0-2, i0-1, i0+64, i0+65
    const int hcol = lane < 3 ? i0 - 3 + lane : i0 + 61 + lane;
    const int hidx = lane < 3 ? lane : 64 + lane;
    const bool hin = lane < 5 && hcol >= 0 && hcol < nx;
    // byte offsets of the lane's column inside a row (column 0 for lanes outside it); rows are clamped on the scalar
    // unit, so each load costs one 32-bit add on the vector ALUs
    const uint32_t ic = col_in ? (uint32_t)i : 0u, ix = xface ? (uint32_t)i : 0u, ih = hin ? (uint32_t)hcol : 0u;
    constexpr uint32_t sIN = sizeof(IN), sFL = sizeof(FL);
    const uint32_t nxu = (uint32_t)nx, nx1 = (uint32_t)nx + 1u;
    const uint32_t c8 = ic * 8u, cin_b = ic * sIN, x8 = ix * 8u, xfl = ix * sFL, cfl = ic * sFL, h_in = ih * sIN;
    // lane 0 of a stage-3 wave also reads C west of the strip (the donor flux of the strip's first x-face)
    const uint32_t cw8 = (STAGE == 3 && lane == 0 && i0 > 0) ? (uint32_t)(i0 - 1) * 8u : c8;
    auto rowc = [&](int r) -> uint32_t { return (r >= 0 && r < ny) ? (uint32_t)r : 0u; };    // cell / x-face rows
    auto rowy = [&](int r) -> uint32_t { return (r >= 0 && r <= ny) ? (uint32_t)r : 0u; };   // y-face rows

    // Loads are split into an issue half (unconditional, from an offset that is always inside the array) and a
    // use half (mask and convert); the use half of a prefetched row runs one loop step later, so the wave never
    // waits on a load right after issuing it.
    struct Raw {
        IN v;
        uint8_t o;
        bool in;
    };
    auto ld = [&](int j) {
        const bool in = col_in && j >= 0 && j < ny;
        const uint32_t r = rowc(j);
        return Raw{ldo(cin, cin_b + r * nxu * sIN), ldo(okp, ic + r * nxu), in};
    };
    auto ldh = [&](int j) {
        const bool in = hin && j >= 0 && j < ny;
        const uint32_t r = rowc(j);
        return Raw{ldo(cin, h_in + r * nxu * sIN), ldo(okp, ih + r * nxu), in};
    };
    auto use = [&](const Raw& r, W& v, bool& o) {
        v = r.in ? (W)r.v : (W)0;
        o = r.in && r.o != 0;
    };
    // exterior value of an edge face: the inflow, or (nullptr) the stage's own edge cell
    auto ex = [&](size_t x, W own_val) -> W { return extp ? (W)extp[x] : own_val; };

    // inputs of one row j that the loop consumes, fetched one row ahead (raw: masked when used)
    struct RowIn {
        double qx, qy;           // Qx[j, i], Qy[j+1, i]
        double c, v0, v1;        // C[j, i], V0, V1 (own cell)
        FL fw, fn;               // stage 2: F0, stage 3: S on x-face (j, i) and y-face (j+1, i)
        double cl;               // stage 3: C[j, i0-1] (lane 0)
        bool xf, ci;             // the x-face exists, the cell exists (and j is a row of the chunk)
    };
    auto ld_row = [&](int j) {
        RowIn r;
        const bool row = j < je;
        r.xf = row && xface;
        r.ci = row && col_in;
        const uint32_t rc = rowc(j), ry = rowy(j + 1);
        const uint32_t oc = c8 + rc * nxu * 8u;
        r.qx = ldo(Qxk, x8 + rc * nx1 * 8u);
        r.qy = ldo(Qyk, c8 + ry * nxu * 8u);
        r.c = ldo(Cp, oc);
        r.v0 = ldo(V0k, oc);
        r.v1 = ldo(V1k, oc);
        if (STAGE >= 2) {
            const FL* gx = STAGE == 2 ? fxp : sxp;
            const FL* gy = STAGE == 2 ? fyp : syp;
            r.fw = ldo(gx, xfl + rc * nx1 * sFL);
            r.fn = ldo(gy, cfl + ry * nxu * sFL);
        } else {
            r.fw = r.fn = (FL)0;
        }
        r.cl = STAGE == 3 ? ldo(Cp, cw8 + rc * nxu * 8u) : 0.0;
        return r;
    };
    auto mask_row = [&](RowIn r) {
        if (!(STAGE == 3 && lane == 0 && i0 > 0 && r.ci)) r.cl = 0.0;
        if (!r.xf) {
            r.qx = 0.0;
            r.fw = (FL)0;
        }
        if (!r.ci) {
            r.qy = r.c = r.v0 = r.v1 = 0.0;
            r.fn = (FL)0;
        }
        return r;
    };

    // FP64: WENO constants held in vector registers, since the stage kernels are short of scalar registers: all of
    // them in stage 1 (scalar spills 74 -> 10), 1/6 and 13/12 in stages 2 and 3 (their VGPR budget of 128 has no
    // room for more); stage 1 7.59 -> 7.50 ms, stage 2 8.70 -> 8.66 at 1024 x 1024 x 50 x 10
#ifndef OCX_WENOK_S1
#define OCX_WENOK_S1 15u
#endif
#ifndef OCX_WENOK_S23
#define OCX_WENOK_S23 3u
#endif
    constexpr unsigned KSEL = PR::split ? 0u : (STAGE == 1 ? OCX_WENOK_S1 : OCX_WENOK_S23);
    const WenoK K = weno_k_v<KSEL>();
    W w[7];
    bool kw[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) use(ld(jb - 3 + m), w[m], kw[m]);
    W hv;
    bool hk;
    use(ldh(jb), hv, hk);
    RowIn cur = mask_row(ld_row(jb));
    // FP64: the last row of the window (or its x-stencil halo) holding a value outside fast_value's range in any
    // lane; the faces of row j take the untested fast divisions when that row lies below j - 3 (wave-uniform)
    auto row_bad = [&](W v) -> bool {
        if constexpr (PR::split) return false;
        else return __builtin_amdgcn_ballot_w64(!fast_value(v)) != 0;
    };
    int bad_row = jb - 4;
#pragma unroll
    for (int m = 0; m < 7; ++m)
        if (row_bad(w[m])) bad_row = jb - 3 + m;
    if (row_bad(hv)) bad_row = max(bad_row, jb);

    // y-face jb (the chunk's first south face)
    double fys;
    {
        const double q = (col_in) ? Qyk[(size_t)jb * nx + i] : 0.0;
        const W elo = col_in ? ex(i, w[3]) : (W)0;
        const W ehi = col_in ? ex((size_t)(ny - 1) * nx + i, w[2]) : (W)0;
        fys = col_in ? (double)line_face<Z>(q, jb, ny, w, kw, os, on, elo, ehi) : 0.0;
    }
    // stage 1: F0 on the owned face jb; stage 2: F0 there and S = F0 + F1; stage 3: S there
    double f0s = 0.0;
    if (STAGE == 1 && col_in) fyp[(size_t)jb * nx + i] = (FL)fys;
    if (STAGE >= 2 && col_in) f0s = (double)(STAGE == 2 ? fyp : syp)[(size_t)jb * nx + i];
    if (STAGE == 2 && col_in) syp[(size_t)jb * nx + i] = (FL)(f0s + fys);
    // stage 3 south-face pieces: F_H and F_L of face jb
    double fhs = 0.0, fls = 0.0;
    if (STAGE == 3) {
        if (col_in) {
            fhs = PR::split ? (double)((float)f0s * (1.0f / 6.0f) + 2.0f * (float)fys * (1.0f / 3.0f))
                            : div_const(f0s, 6.0, 1.0 / 6.0) + div_const(2.0 * fys, 3.0, 1.0 / 3.0);
            const double q = Qyk[(size_t)jb * nx + i];
            const double clo = jb > 0 ? Cp[(size_t)(jb - 1) * nx + i] : 0.0;
            fls = donor(q, jb, ny, clo, cur.c, os, on, ext0 + i, ext0 + (size_t)(ny - 1) * nx + i);
            if (!UNLIM) ayp[(size_t)jb * nx + i] = (FL)(fhs - fls);
        }
    }

    for (int j = jb; j < je; ++j) {
        // prefetch the next row's inputs, the next row of the window and the next halo
        const RowIn nxt_raw = ld_row(j + 1);
        const Raw wn_raw = ld(j + 4), hn_raw = ldh(j + 1);
        // x-face i of row j
        rbw[lane + 3] = w[3];
        obw[lane + 3] = kw[3];
        if (lane < 5) {
            rbw[hidx] = hv;
            obw[hidx] = hk;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bool known = !PR::split && bad_row < j - 3;
        double fxw = 0.0;
        if (xface) {
            W u[6];
            bool o[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                u[m] = rbw[lane + m];
                o[m] = obw[lane + m] != 0;
            }
            fxw = (double)line_face<Z>(cur.qx, i, nx, u, o, ow, oe, ex((size_t)j * nx, u[3]),
                                       ex((size_t)j * nx + nx - 1, u[2]), known, K);
        }
        // y-face j+1 (north of row j): rows j-2 .. j+3 = w[1..6]
        double fyn = 0.0;
        const double qyn = cur.qy;
        if (col_in) fyn = (double)line_face<Z>(qyn, j + 1, ny, w + 1, kw + 1, os, on, ex(i, w[4]),
                                               ex((size_t)(ny - 1) * nx + i, w[3]), known, K);
        const double fxe = shfl_down1(fxw);
        const bool yown = (j + 1 < je) || (j + 1 == ny);

        if (STAGE == 1) {
            if (xown) sto(fxp, xfl + (uint32_t)j * nx1 * sFL, (FL)fxw);
            if (col_in && yown) sto(fyp, cfl + (uint32_t)(j + 1) * nxu * sFL, (FL)fyn);
            if (own) {
                const double d = fxe - fxw + fyn - fys;
                const double c0 = cur.c, v0 = cur.v0, v1 = cur.v1;
                const double q0 = v0 * c0, dV = v1 - v0;
                const double qn = 0.0 * q0 + 1.0 * (q0 - a.dt * d);
                const double V = v0 + dV;
                sto(coutp, ic * (uint32_t)sizeof(ST) + (uint32_t)j * nxu * (uint32_t)sizeof(ST), stage_value<PR>(qn, V, c0));
            }
        } else if (STAGE == 2) {
            const double f0w = (double)cur.fw, f0n = (double)cur.fn;
            const double f0e = shfl_down1(f0w);
            if (xown) sto(sxp, xfl + (uint32_t)j * nx1 * sFL, (FL)(f0w + fxw));
            if (col_in && yown) sto(syp, cfl + (uint32_t)(j + 1) * nxu * sFL, (FL)(f0n + fyn));
            if (own) {
                const double c0 = cur.c, v0 = cur.v0, v1 = cur.v1;
                const double q0 = v0 * c0, dV = v1 - v0;
                const double d0 = f0e - f0w + f0n - f0s;
                const double q1 = 0.0 * q0 + 1.0 * (q0 - a.dt * d0);
                const double d1 = fxe - fxw + fyn - fys;
                const double qn = 0.75 * q0 + 0.25 * (q1 - a.dt * d1);
                const double V = v0 + 0.5 * dV;
                sto(coutp, ic * (uint32_t)sizeof(ST) + (uint32_t)j * nxu * (uint32_t)sizeof(ST), stage_value<PR>(qn, V, c0));
            }
            f0s = f0n;
        } else {
            // F_H on the lane's x-face and on the north face
            double fhw, fhn;
            if (PR::split) {
                fhw = (double)((float)cur.fw * (1.0f / 6.0f) + 2.0f * (float)fxw * (1.0f / 3.0f));
                fhn = (double)((float)cur.fn * (1.0f / 6.0f) + 2.0f * (float)fyn * (1.0f / 3.0f));
            } else {
                // (F0 + F1)/6 + 2 F2/3 with correctly rounded divisions by the constants (div_const)
                fhw = div_const((double)cur.fw, 6.0, 1.0 / 6.0) + div_const(2.0 * fxw, 3.0, 1.0 / 3.0);
                fhn = div_const((double)cur.fn, 6.0, 1.0 / 6.0) + div_const(2.0 * fyn, 3.0, 1.0 / 3.0);
            }
            const double fhe = shfl_down1(fhw);
            // donor fluxes: x-face i between C[j, i-1] and C[j, i]; y-face j+1 between C[j, i] and C[j+1, i]
            const double cC = cur.c;
            double cCl = shfl_up1(cC);
            if (lane == 0) cCl = cur.cl;
            const double cCn = (j + 1 < je) ? (nxt_raw.ci ? nxt_raw.c : 0.0)
                                             : ((col_in && j + 1 < ny) ? ldo(Cp, c8 + rowc(j + 1) * nxu * 8u) : 0.0);
            double flw = 0.0, fln = 0.0;
            if (xface) {
                const double chi = col_in ? cC : 0.0;
                flw = donor(cur.qx, i, nx, cCl, chi, ow, oe, ext0 + (size_t)j * nx, ext0 + (size_t)j * nx + nx - 1);
            }
            if (col_in) fln = donor(qyn, j + 1, ny, cC, cCn, os, on, ext0 + i, ext0 + (size_t)(ny - 1) * nx + i);
            const double fle = shfl_down1(flw);
            if (UNLIM) {
                if (a.fxi && xown && i == a.flux_face_i) a.fxi[(size_t)p * ny + j] = fhw;
            } else {
                if (xown) sto(axp, xfl + (uint32_t)j * nx1 * sFL, (FL)(fhw - flw));
                if (col_in && yown) sto(ayp, cfl + (uint32_t)(j + 1) * nxu * sFL, (FL)(fhn - fln));
            }
            if (own) {
                const double c0 = cC, v0 = cur.v0, v1 = cur.v1;
                const double q0 = v0 * c0;
                double qn;
                if (UNLIM) {
                    const double d = fhe - fhw + fhn - fhs;
                    qn = q0 - a.dt * d;
                } else {
                    const double d = fle - flw + fln - fls;
                    qn = 0.0 * q0 + 1.0 * (q0 - a.dt * d);
                }
                sto(static_cast<double*>(a.cout) + (size_t)p * plane, c8 + (uint32_t)j * nxu * 8u, v1 > 0 ? qn / v1 : c0);
            }
            fhs = fhn;
            fls = fln;
        }
        fys = fyn;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            w[m] = w[m + 1];
            kw[m] = kw[m + 1];
        }
        use(wn_raw, w[6], kw[6]);
        use(hn_raw, hv, hk);
        if (row_bad(w[6])) bad_row = j + 4;
        if (row_bad(hv)) bad_row = max(bad_row, j + 1);
        cur = mask_row(nxt_raw);
    }
}

// waves per SIMD the FP64 stage kernels are compiled for (their register budget: 512 / waves VGPRs)
#ifndef OCX_MARCH_WAVES
#define OCX_MARCH_WAVES 4
#endif

template <int STAGE, bool Z, bool UNLIM, class PR>
__global__ __launch_bounds__(64 * MAXW) __attribute__((amdgpu_waves_per_eu(OCX_MARCH_WAVES))) void k_march(
    MarchArgs a) {
    __shared__ typename PR::w rbuf[MAXW][MRB];
    __shared__ uint8_t obuf[MAXW][MRB];
    const int wv = wave_id();
    march_block<STAGE, Z, UNLIM, PR>(a, xcd_logical(blockIdx.x, gridDim.x), wv, blockDim.x >> 6, rbuf[wv], obuf[wv]);
}

}  // namespace ocx
