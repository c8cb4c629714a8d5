// ALE layer-transport kernels (gfx950): the producer of the advection step's layer fluxes, all FP64 in the
// reference's per-element arithmetic (no contraction, IEEE divisions), so results are bit-identical.
//
//   transport velocity   u^θ and the free-surface flux Q_bt per face column   (core/vertical/ale.py:789-808)
//   transport thickness  q_e / q_w, the donor layer cut at the shallower bed  (core/vertical/ale.py:168-199)
//   PPM edges            limited slopes + positive-definite parabola per cell (core/vertical/ale.py:578-623)
//   PPM flux             swept-volume mean thickness, a Newton correction du per face column with the global stop
//                        of the reference, then the residual share             (core/vertical/ale.py:626-740)
//   outflow limiter      outflow > content check, cut, give-back passes       (core/vertical/ale.py:299-437)
//   Lagrangian step      h† = h - Δt ∇·Q / A, emptied-layer check, V0, V1, ok (core/vertical/ale.py:278-295)
//
// Faces of one direction are addressed in their natural layout: AX 0 the u-faces (n, ny, nx+1), AX 1 the v-faces
// (n, ny+1, nx).  A face column (a, f) is face f along the flux axis in row (AX 0) or column (AX 1) a; threads
// take face columns with the unit-stride index fastest, so every layer's loads are coalesced.  Each thread loops
// over the layers from the bed up, which is the order of every sum over k in the reference.
#pragma once
#include <float.h>

#include "ocx_weno.hpp"

namespace ocx {
namespace ale {

constexpr double H_PPM_MIN = 2e-10;   // positive-definite floor of the PPM thickness [m] (core/vertical/ale.py:298)

// Python's min(a, b) / max(a, b) on floats: the first argument unless the second compares smaller / larger
__device__ __forceinline__ double amin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double amax(double a, double b) { return b > a ? b : a; }

// numpy.nan_to_num: NaN -> 0, +-inf -> +-largest finite
__device__ __forceinline__ double nan0(double x) {
    if (x != x) return 0.0;
    if (x == __builtin_inf()) return DBL_MAX;
    if (x == -__builtin_inf()) return -DBL_MAX;
    return x;
}

// numpy.minimum: the smaller, NaN if either is NaN
__device__ __forceinline__ double npmin(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return b < a ? b : a;
}

// numpy.maximum: the larger, NaN if either is NaN
__device__ __forceinline__ double npmax(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return b > a ? b : a;
}

// numpy.sign: -1, 0 (of the zero's sign) or 1; NaN stays NaN
__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x * 0.0); }

template <int AX>
struct Geo {
    int n, ny, nx;
    __host__ __device__ __forceinline__ int ncell() const { return AX == 0 ? nx : ny; }           // cells along the flux axis
    __host__ __device__ __forceinline__ long long cols() const {                                    // face columns
        return AX == 0 ? (long long)ny * (nx + 1) : (long long)(ny + 1) * nx;
    }
    __host__ __device__ __forceinline__ size_t pc() const { return (size_t)ny * nx; }               // cell plane
    __host__ __device__ __forceinline__ size_t pf() const { return (size_t)cols(); }                // face plane
    __host__ __device__ __forceinline__ void col(long long t, int& a, int& f) const {
        if (AX == 0) {
            a = (int)(t / (nx + 1));
            f = (int)(t - (long long)a * (nx + 1));
        } else {
            f = (int)(t / nx);
            a = (int)(t - (long long)f * nx);
        }
    }
    __host__ __device__ __forceinline__ size_t cell(int a, int c) const {
        return AX == 0 ? (size_t)a * nx + c : (size_t)c * nx + a;
    }
    __host__ __device__ __forceinline__ size_t face(int a, int f) const {
        return AX == 0 ? (size_t)a * (nx + 1) + f : (size_t)f * nx + a;
    }
};

__device__ __forceinline__ long long gtid() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// one lane per wave with `bad` set ORs 1 into *flag (a vector atomic)
__device__ __forceinline__ void wave_flag(bool bad, unsigned* flag) {
    const unsigned long long b = __ballot(bad);
    if (b && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)b) - 1)) atomicOr(flag, 1u);
}

// ------------------------------------------------------------------------------------------ transport velocity
// u (n, M) columns; last_wet (M) bytes; w (n, M); L, D (M).  ut (n, M), Qb (M).
__global__ void k_transport_velocity(const double* __restrict__ u_new, const double* __restrict__ u_old,
                                     const uint8_t* __restrict__ last_wet, const double* __restrict__ w,
                                     const double* __restrict__ L, const double* __restrict__ D, double theta,
                                     double theta_bc, int n, long long M, double* __restrict__ ut,
                                     double* __restrict__ Qb) {
    const long long p = gtid();
    if (p >= M) return;
    const bool wet = last_wet[p] != 0;
    double m1 = 0.0, m0 = 0.0;
    for (int k = 0; k < n; ++k) {
        const size_t x = (size_t)k * M + p;
        const double un = wet ? u_old[x] : 0.0;
        m1 += w[x] * u_new[x];
        m0 += w[x] * un;
    }
    const double d = (theta - theta_bc) * (m1 - m0);
    const double omt = 1.0 - theta_bc;
    double s = 0.0;
    for (int k = 0; k < n; ++k) {
        const size_t x = (size_t)k * M + p;
        const double un = wet ? u_old[x] : 0.0;
        const double v = theta_bc * u_new[x] + omt * un + d;
        ut[x] = v;
        s += w[x] * v;
    }
    Qb[p] = L[p] * D[p] * s;
}

// ------------------------------------------------------------------------------------------ transport thickness
// q_e: from the low-side cell (flow towards +), q_w: from the high-side cell; the domain-edge faces use the inner
// cell, closed faces are 0.  clip: only the part of the donor layer above the shallower bed passes.
template <int AX>
__global__ void k_face_thickness(const double* __restrict__ h, const double* __restrict__ H,
                                 const uint8_t* __restrict__ act, int clip, Geo<AX> g, double* __restrict__ qe,
                                 double* __restrict__ qw) {
    const long long t = gtid();
    if (t >= g.cols()) return;
    int a, f;
    g.col(t, a, f);
    const int nc = g.ncell();
    const size_t cw = g.cell(a, f > 0 ? f - 1 : 0), ce = g.cell(a, f < nc ? f : nc - 1);
    const size_t fo = g.face(a, f), PC = g.pc(), PF = g.pf();
    if (!(act[cw] && act[ce])) {
        for (int k = 0; k < g.n; ++k) {
            qe[k * PF + fo] = 0.0;
            qw[k * PF + fo] = 0.0;
        }
        return;
    }
    if (clip) {
        double Hw = H[cw], He = H[ce];
        Hw = Hw == Hw ? Hw : 0.0;
        He = He == He ? He : 0.0;
        const double z_bed = amax(-Hw, -He);
        double sw = 0.0, se = 0.0;
        for (int k = 0; k < g.n; ++k) {
            double hw = h[k * PC + cw], he = h[k * PC + ce];
            hw = hw == hw ? hw : 0.0;
            he = he == he ? he : 0.0;
            sw += hw;
            se += he;
            qe[k * PF + fo] = amin(hw, amax(0.0, (-Hw + sw) - z_bed));
            qw[k * PF + fo] = amin(he, amax(0.0, (-He + se) - z_bed));
        }
    } else {
        for (int k = 0; k < g.n; ++k) {
            const double hw = h[k * PC + cw], he = h[k * PC + ce];
            qe[k * PF + fo] = hw == hw ? hw : 0.0;
            qw[k * PF + fo] = he == he ? he : 0.0;
        }
    }
}

// ------------------------------------------------------------------------------------------ PPM edges
// limited slope of a cell from its neighbours' values (already the cell itself where the neighbour is land or
// outside) and its masks
__device__ __forceinline__ double ppm_slope(double hm1, double hc, double hp1, double mw, double me, double m) {
    double s = 0.5 * (hp1 - hm1);
    const double dmx = amax(amax(hp1, hm1), hc) - hc;
    const double dmn = hc - amin(amin(hp1, hm1), hc);
    const double a = amin(fabs(s), 2.0 * amin(dmx, dmn));
    s = s > 0.0 ? a : (s < 0.0 ? -a : 0.0);
    return mw * me * m > 0.0 ? s : 0.0;
}

// one thread per cell (natural index j * nx + i), looping over the layers.  The thickness is read as the flux
// passes it: inactive cells 0, NaN 0.  Edges: lo = the low-index (west / south) edge, hi = the high one.
template <int AX>
__global__ void k_ppm_edges(const double* __restrict__ h, const uint8_t* __restrict__ act, double h_min, Geo<AX> g,
                            double* __restrict__ e_lo, double* __restrict__ e_hi) {
    const long long t = gtid();
    const size_t PC = g.pc();
    if (t >= (long long)PC) return;
    const int nx = g.nx;
    const int a = AX == 0 ? (int)(t / nx) : (int)(t % nx);
    const int c = AX == 0 ? (int)(t % nx) : (int)(t / nx);
    const int nc = g.ncell();
    // the five cells c-2 .. c+2 along the axis (clamped indices; existence flags decide)
    size_t o[5];
    bool ex[5];
    double m[5];
    for (int d = 0; d < 5; ++d) {
        const int cc = c + d - 2;
        ex[d] = cc >= 0 && cc < nc;
        o[d] = g.cell(a, ex[d] ? cc : c);
        m[d] = ex[d] && act[o[d]] ? 1.0 : 0.0;
    }
    const double r6 = 1.0 / 6.0;
    for (int k = 0; k < g.n; ++k) {
        double v[5];
        for (int d = 0; d < 5; ++d) v[d] = m[d] != 0.0 ? nan0(h[k * PC + o[d]]) : 0.0;
        // hm1 / hp1 / slope of the cell at offset d (1..3) with its neighbours d-1, d+1
        double hm1[5], hp1[5], slp[5];
        for (int d = 1; d <= 3; ++d) {
            const bool lo_ex = ex[d - 1], hi_ex = ex[d + 1];      // cell d has a low / high neighbour
            const double mw = lo_ex ? m[d - 1] : 0.0, me = hi_ex ? m[d + 1] : 0.0;
            hm1[d] = lo_ex ? mw * v[d - 1] + (1.0 - mw) * v[d] : v[d];
            hp1[d] = hi_ex ? me * v[d + 1] + (1.0 - me) * v[d] : v[d];
            slp[d] = ex[d] ? ppm_slope(hm1[d], v[d], hp1[d], mw, me, m[d]) : 0.0;
        }
        const double hc = v[2];
        const double sw = ex[1] ? slp[1] : 0.0, se = ex[3] ? slp[3] : 0.0;
        const double hw = 0.5 * (hm1[2] + hc) + div_const(sw - slp[2], 6.0, r6);
        const double he = 0.5 * (hp1[2] + hc) + div_const(slp[2] - se, 6.0, r6);
        const double curv = 3.0 * ((hw + he) - 2.0 * hc);
        const double dh = he - hw;
        const bool inside = curv > 0.0 && fabs(dh) < curv;
        double ow, oe;
        if (inside && hc <= h_min) {
            ow = hc;
            oe = hc;
        } else {
            double sc = 1.0;
            if (inside) {
                const double num = 12.0 * curv * (hc - h_min);
                const double den = curv * curv + 3.0 * dh * dh;
                if (num < den) sc = num / den;
            }
            ow = hc + sc * (hw - hc);
            oe = hc + sc * (he - hc);
        }
        e_lo[k * PC + t] = ow;
        e_hi[k * PC + t] = oe;
    }
}

// ------------------------------------------------------------------------------------------ PPM flux
template <int AX>
struct PpmArgs {
    const double *h, *e_lo, *e_hi, *ut, *L, *Qb, *qe, *qw, *dxt;
    const uint8_t* act;
    double dt;
    Geo<AX> g;
    double *du, *lo, *hi, *err, *dsum;   // state per face column
    unsigned* notconv;                   // notconv[m]: some column unconverged after iteration m
    double* out;
};

// the face column's geometry: donor cells on both sides and their widths
struct Column {
    size_t cw, ce, fo;
    double dxw, dxe, Lf, Qb;
    bool open;
};

template <int AX>
__device__ __forceinline__ Column column(const PpmArgs<AX>& p, int a, int f) {
    Column c;
    const int nc = p.g.ncell();
    c.cw = p.g.cell(a, f > 0 ? f - 1 : 0);
    c.ce = p.g.cell(a, f < nc ? f : nc - 1);
    c.fo = p.g.face(a, f);
    c.dxw = p.dxt[c.cw];
    c.dxe = p.dxt[c.ce];
    c.Lf = p.L[c.fo];
    c.Qb = p.Qb[c.fo];
    c.open = p.act[c.cw] && p.act[c.ce];
    return c;
}

// (F, dF/du) of layer k at velocity u: the swept mean thickness of the donor parabola over the Courant fraction,
// capped by the transport thickness where that is below the donor layer
template <int AX>
__device__ __forceinline__ void face_flux(const PpmArgs<AX>& p, const Column& c, int k, double u, double& F,
                                          double& dF) {
    const size_t PC = p.g.pc(), PF = p.g.pf();
    const bool pos = u >= 0.0;
    const size_t ci = k * PC + (pos ? c.cw : c.ce);
    const double cr = amin(fabs(u) * p.dt / (pos ? c.dxw : c.dxe), 1.0);
    const double hu = nan0(p.h[ci]);
    const double ww = p.e_lo[ci], ee = p.e_hi[ci];
    const double curv3 = (ww + ee) - 2.0 * hu;
    const double dh = pos ? ww - ee : ee - ww;
    const double edge = pos ? ee : ww;
    const double q = pos ? p.qe[k * PF + c.fo] : p.qw[k * PF + c.fo];
    double mean = edge + cr * (0.5 * dh + curv3 * (cr - 1.5));
    double marg = edge + cr * (dh + 3.0 * curv3 * (cr - 1.0));
    if (q < hu) {
        mean = amin(mean, q);
        marg = amin(marg, q);
    }
    F = c.Lf * u * amax(mean, 0.0);
    dF = c.Lf * amax(marg, 0.0);
}

template <int AX>
__device__ __forceinline__ void column_sums(const PpmArgs<AX>& p, const Column& c, double du, double& s,
                                            double& sd) {
    s = 0.0;
    sd = 0.0;
    if (!c.open) return;
    const size_t PF = p.g.pf();
    for (int k = 0; k < p.g.n; ++k) {
        double F, dF;
        face_flux(p, c, k, p.ut[k * PF + c.fo] + du, F, dF);
        s += F;
        sd += dF;
    }
}

// the CFL bracket [lo, hi] of du and the uncorrected column sums (du = 0 adds nothing: the reference evaluates
// at u^θ itself)
template <int AX>
__global__ void k_ppm_init(PpmArgs<AX> p) {
    const long long t = gtid();
    if (t >= p.g.cols()) return;
    int a, f;
    p.g.col(t, a, f);
    const Column c = column(p, a, f);
    const size_t PF = p.g.pf();
    const double limw = 0.5 * c.dxw / p.dt, lime = 0.5 * c.dxe / p.dt;
    double mnw = __builtin_inf(), mxe = -__builtin_inf();
    double s = 0.0, sd = 0.0;
    for (int k = 0; k < p.g.n; ++k) {
        const double u = p.ut[k * PF + c.fo];
        mnw = amin(mnw, limw - u);
        mxe = amax(mxe, -(lime + u));
        if (c.open) {
            double F, dF;
            face_flux(p, c, k, u, F, dF);
            s += F;
            sd += dF;
        }
    }
    p.hi[t] = amax(amin(limw, mnw), 0.0);
    p.lo[t] = amin(amax(-lime, mxe), 0.0);
    p.du[t] = 0.0;
    p.err[t] = s - c.Qb;
    p.dsum[t] = sd;
}

// Newton iteration m (0-based) with bracketing bisection; runs only while the previous iteration left some
// column unconverged (every column iterates until all have converged, as the reference)
template <int AX>
__global__ void k_ppm_iter(PpmArgs<AX> p, int m) {
    if (m > 0 && p.notconv[m - 1] == 0u) return;
    const long long t = gtid();
    bool bad = false;
    if (t < p.g.cols()) {
        int a, f;
        p.g.col(t, a, f);
        const Column c = column(p, a, f);
        const double e = p.err[t], d = p.du[t], ds = p.dsum[t];
        double hi = p.hi[t], lo = p.lo[t];
        if (e > 0.0) hi = d;
        if (e < 0.0) lo = d;
        const double step = ds > 0.0 ? -e / ds : 0.0;
        double nw = d + step;
        if (step > 0.0 && nw >= hi) nw = 0.5 * (d + hi);
        if (step < 0.0 && nw <= lo) nw = 0.5 * (d + lo);
        double s, sd;
        column_sums(p, c, nw, s, sd);
        const double er = s - c.Qb;
        p.du[t] = nw;
        p.hi[t] = hi;
        p.lo[t] = lo;
        p.err[t] = er;
        p.dsum[t] = sd;
        bad = fabs(er) > 1e-12 * amax(fabs(c.Qb), 1.0);
    }
    wave_flag(bad, p.notconv + m);
}

// the layer fluxes at u^θ + du plus the residual R = Q_bt - Σ F shared in proportion to the transport thickness
// on R's side (the other side's where that column is empty); closed faces 0
template <int AX>
__global__ void k_ppm_final(PpmArgs<AX> p) {
    const long long t = gtid();
    if (t >= p.g.cols()) return;
    int a, f;
    p.g.col(t, a, f);
    const Column c = column(p, a, f);
    const size_t PF = p.g.pf();
    if (!c.open) {
        for (int k = 0; k < p.g.n; ++k) p.out[k * PF + c.fo] = 0.0;
        return;
    }
    const double du = p.du[t];
    double s = 0.0, te = 0.0, tw = 0.0;
    for (int k = 0; k < p.g.n; ++k) {
        double F, dF;
        face_flux(p, c, k, p.ut[k * PF + c.fo] + du, F, dF);
        p.out[k * PF + c.fo] = F;
        s += F;
        te += p.qe[k * PF + c.fo];
        tw += p.qw[k * PF + c.fo];
    }
    const double R = c.Qb - s;
    bool use_e = R >= 0.0;
    double tot = use_e ? te : tw;
    if (tot <= 0.0) {
        use_e = !use_e;
        tot = use_e ? te : tw;
    }
    const double den = tot > 0.0 ? tot : 1.0;
    const double* q = use_e ? p.qe : p.qw;
    for (int k = 0; k < p.g.n; ++k) {
        const size_t x = k * PF + c.fo;
        p.out[x] = p.out[x] + R * q[x] / den;
    }
}

// ------------------------------------------------------------------------------------------ upwind thickness flux
// upwind_transport_x (core/vertical/ale.py:233-251), one thread per face column: F_k = L u_k q_k with the transport
// thickness of the side the water comes from, R = Q_bt - sum_k F_k, and R shared in proportion to the transport
// thicknesses on R's side (the other side where that column's sum is not positive).  The sums run from layer 0
// up and start from the first term, as NumPy's sum over the first axis does.  Face arrays are indexed by the face
// column t directly (t is the face's index in its layer plane for both axes).
template <int AX>
__global__ void k_upwind_transport(const double* __restrict__ ut, const double* __restrict__ L,
                                   const double* __restrict__ Qb, const double* __restrict__ qe,
                                   const double* __restrict__ qw, Geo<AX> g, double* __restrict__ out) {
    const long long t = gtid();
    if (t >= g.cols()) return;
    const size_t PF = g.pf();
    const double Lf = L[t];
    double s = 0.0, se = 0.0, sw = 0.0;
    for (int k = 0; k < g.n; ++k) {
        const size_t x = k * PF + t;
        const double u = ut[x], e = qe[x], w = qw[x];
        const double F = Lf * u * (u >= 0.0 ? e : w);
        out[x] = F;
        s = k == 0 ? F : s + F;
        se = k == 0 ? e : se + e;
        sw = k == 0 ? w : sw + w;
    }
    const double R = Qb[t] - s;
    bool use_e = R >= 0.0;
    if (!((use_e ? se : sw) > 0.0)) use_e = !use_e;
    const double tot = use_e ? se : sw;
    const double den = tot > 0.0 ? tot : 1.0;
    const double* q = use_e ? qe : qw;
    for (int k = 0; k < g.n; ++k) {
        const size_t x = k * PF + t;
        out[x] = out[x] + R * q[x] / den;
    }
}

// ------------------------------------------------------------------------------------------ open-edge faces
// open_face_transport (core/vertical/ale.py:841-861) on the m faces of one open edge, one thread per face: the
// layer fluxes (h_k / sum h) Q_bt + L h_k (tb u'_new + (1 - tb) u'_old) of the interior cell's thicknesses,
// written where the face depth D > 0 (open_boundary.py:350-351).  Every array is addressed by an offset, a layer
// stride and a face stride, so the four edges of the u- and v-face arrays share the kernel.
struct EdgeArgs {
    double* Q;                       // face fluxes
    const double *h, *un, *uo;       // interior cells; new and old velocities at the faces
    const double *D, *Qo, *L;        // one layer of faces
    long long q0, qk, qf;            // Q, un, uo: offset, layer stride, face stride
    long long h0, hk, hf;            // h
    long long f0, ff;                // D, Qo, L
    int n, m;
    double tb;
};

__global__ void k_open_face_transport(EdgeArgs e) {
    const long long f = gtid();
    if (f >= e.m) return;
    const long long fq = e.q0 + f * e.qf, fh = e.h0 + f * e.hf, fo = e.f0 + f * e.ff;
    if (!(e.D[fo] > 0.0)) return;
    double tot = 0.0;
    for (int k = 0; k < e.n; ++k) {
        const double hk = e.h[fh + k * e.hk];
        tot = k == 0 ? hk : tot + hk;
    }
    const bool wet = tot > 0.0;
    const double den = wet ? tot : 1.0;
    double sn = 0.0, so = 0.0;
    for (int k = 0; k < e.n; ++k) {
        const double w = wet ? e.h[fh + k * e.hk] / den : 0.0;
        const double an = w * e.un[fq + k * e.qk], ao = w * e.uo[fq + k * e.qk];
        sn = k == 0 ? an : sn + an;
        so = k == 0 ? ao : so + ao;
    }
    const double Qb = e.Qo[fo], Lf = e.L[fo], tb = e.tb;
    for (int k = 0; k < e.n; ++k) {
        const double hk = e.h[fh + k * e.hk];
        const double w = wet ? hk / den : 0.0;
        const double pn = e.un[fq + k * e.qk] - sn, po = e.uo[fq + k * e.qk] - so;
        const double q = w * Qb + Lf * hk * (tb * pn + (1.0 - tb) * po);
        e.Q[fq + k * e.qk] = wet ? q : 0.0;
    }
}

// ------------------------------------------------------------------------------------------ freshwater
// One lane per wavefront with `hit` set adds the wavefront's count of set lanes to *count (a vector atomic).
__device__ __forceinline__ void wave_count(bool hit, unsigned long long* count) {
    const unsigned long long b = __ballot(hit);
    if (b && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)b) - 1))
        atomicAdd(count, (unsigned long long)__popcll(b));
}

// freshwater_volumes (core/components/freshwater.py:31-58), one thread per column of M: the volume scale dV joins
// the top layer; a removal that would take it below the floor is taken from the original layers top-down, first
// keeping each at the floor, then down to zero.  Columns where that happens are counted.  With area, the volumes
// out A are written to Vout as well (layer_volumes of the thicknesses).
__global__ void k_freshwater_volumes(const double* __restrict__ V, const double* __restrict__ dV, double scale,
                                     const double* __restrict__ floor_c, double floor_s,
                                     const double* __restrict__ area, int n, long long M, double* __restrict__ out,
                                     double* __restrict__ Vout, unsigned long long* count) {
    const long long c = gtid();
    bool spill = false;
    if (c < M) {
        const double d = scale * dV[c];
        const double fl = floor_c ? floor_c[c] : floor_s;
        for (int k = 0; k < n - 1; ++k) out[(size_t)k * M + c] = V[(size_t)k * M + c];
        const size_t top = (size_t)(n - 1) * M + c;
        const double vt = V[top] + d;
        out[top] = vt;
        spill = d < 0.0 && vt < fl;
        if (spill) {
            double deficit = -d;
            out[top] = V[top];
            for (int pass = 0; pass < 2; ++pass) {
                const double keep = pass == 0 ? fl : 0.0;
                for (int k = n - 1; k >= 0; --k) {
                    const size_t x = (size_t)k * M + c;
                    const double col = out[x];
                    const double take = npmin(deficit, npmax(col - keep, 0.0));
                    out[x] = col - take;
                    deficit = deficit - take;
                }
            }
        }
        if (Vout) {
            const double A = area[c];
            for (int k = 0; k < n; ++k) Vout[(size_t)k * M + c] = out[(size_t)k * M + c] * A;
        }
    }
    wave_count(spill, count);
}

// apply_freshwater (core/components/freshwater.py:61-122), one thread per column, in place on the stack (nt, n, M):
// every member but T (t_index) and T^2 (t2_index) takes c V_adv / V_fw where the column has freshwater; in a
// column whose volume changed below the top layer the touched layers (from the lowest changed one up) share
// sum V_adv c / sum V_fw.  var_surf (n, M), for member s_index: (V_fw S_new^2 - V_adv S_old^2) / dt where the
// column has freshwater, else 0.  The sums run from layer 0 up and start from the first term (NumPy's sum over
// the first axis).  Columns with a change below the top layer are counted.
struct FreshArgs {
    double* stack;
    const double *Va, *Vf, *fwa;
    double* var_surf;
    int nt, n, t_index, t2_index, s_index;
    long long M;
    double dt;
    unsigned long long* count;
};

__global__ void k_apply_freshwater(FreshArgs a) {
    const long long c = gtid();
    bool spill = false;
    if (c < a.M) {
        const int n = a.n;
        const long long M = a.M;
        const bool cols = a.fwa[c] != 0.0;
        int kmin = -1;
        for (int k = 0; k < n; ++k) {
            const bool ch = a.Va[(size_t)k * M + c] != a.Vf[(size_t)k * M + c];
            if (ch && kmin < 0) kmin = k;
            if (ch && k < n - 1) spill = true;
        }
        if (kmin < 0) kmin = 0;
        double vsum = 0.0;
        if (spill)
            for (int k = 0; k < n; ++k) {
                const double v = k >= kmin ? a.Vf[(size_t)k * M + c] : 0.0;
                vsum = k == 0 ? v : vsum + v;
            }
        const bool has_water = vsum > 0.0;
        for (int i = 0; i < a.nt; ++i) {
            if (i == a.t_index || i == a.t2_index) continue;
            double* st = a.stack + (size_t)i * n * M + c;
            double cbar = 0.0;
            if (spill) {
                double num = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double v = k >= kmin ? a.Va[(size_t)k * M + c] * st[(size_t)k * M] : 0.0;
                    num = k == 0 ? v : num + v;
                }
                cbar = has_water ? num / vsum : 0.0;
            }
            for (int k = 0; k < n; ++k) {
                const size_t x = (size_t)k * M;
                const double va = a.Va[x + c], vf = a.Vf[x + c], s = st[x];
                const double ratio = (cols && vf > 0.0) ? va / vf : 1.0;
                double cv = cols ? s * ratio : s;
                if (spill && k >= kmin && has_water) cv = cbar;
                if (i == a.s_index && a.var_surf) a.var_surf[x + c] = cols ? (vf * (cv * cv) - va * (s * s)) / a.dt : 0.0;
                st[x] = cv;
            }
        }
    }
    wave_count(spill, a.count);
}

// ------------------------------------------------------------------------------------------ outflow limiter
struct OutArgs {
    const double *h, *area;
    const uint8_t* act;
    double *Qx, *Qy;
    int n, ny, nx;
    double dt, h_floor;
    double *r, *room;                    // (n, ny, nx)
    double* delta;                       // per face column of the current axis
    unsigned long long* count;           // cell-layers over the bound
    unsigned* any;                       // some cell-layer over the bound
    unsigned long long* bad_face;        // first face whose cut cannot be given back (linear index), else ~0
};

// what a cell-layer can give in Δt, (h - h_floor)+ A / Δt with A = 0 on land
__device__ __forceinline__ double room0(const OutArgs& o, size_t x, size_t c) {
    const double A = o.act[c] ? o.area[c] : 0.0;
    return amax(o.h[x] - o.h_floor, 0.0) * A / o.dt;
}

// outgoing flux of cell-layer (k, j, i) over its four faces
__device__ __forceinline__ double outflow(const OutArgs& o, int k, int j, int i) {
    const size_t fx = ((size_t)k * o.ny + j) * (o.nx + 1) + i;
    const size_t fy = ((size_t)k * (o.ny + 1) + j) * o.nx + i;
    return amax(o.Qx[fx + 1], 0.0) - amin(o.Qx[fx], 0.0) + amax(o.Qy[fy + o.nx], 0.0) - amin(o.Qy[fy], 0.0);
}

// over = active & O > room0 x factor; r = room0 / O where over, else 1; counts and flags the cells over
__global__ void k_out_over(OutArgs o, double factor, int first) {
    const long long t = gtid();
    const size_t PC = (size_t)o.ny * o.nx;
    bool over = false;
    if (t < (long long)(PC * o.n)) {
        const int k = (int)(t / PC);
        const size_t c = (size_t)t - (size_t)k * PC;
        const int j = (int)(c / o.nx), i = (int)(c - (size_t)j * o.nx);
        const double O = outflow(o, k, j, i);
        const double r0 = room0(o, t, c);
        over = o.act[c] && (first ? O > r0 : O > r0 * factor);
        o.r[t] = over ? r0 / (O > 0.0 ? O : 1.0) : 1.0;
    }
    const unsigned long long b = __ballot(over);
    if (b && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)b) - 1)) {
        atomicAdd(o.count, (unsigned long long)__popcll(b));
        atomicOr(o.any, 1u);
    }
}

// room = room0 - outflow with the current Qx, Qy
__global__ void k_out_room(OutArgs o) {
    const long long t = gtid();
    const size_t PC = (size_t)o.ny * o.nx;
    if (t >= (long long)(PC * o.n)) return;
    const int k = (int)(t / PC);
    const size_t c = (size_t)t - (size_t)k * PC;
    const int j = (int)(c / o.nx), i = (int)(c - (size_t)j * o.nx);
    o.room[t] = room0(o, t, c) - outflow(o, k, j, i);
}

// the face arrays of one axis and the donor cells of face column (a, f): lo = the low-index side (the donor of
// flow towards +), hi = the high-index side; the edge faces take the cell across the wrap on a periodic axis
template <int AX>
struct OutFace {
    size_t fo, PF, clo, chi;
    bool has_lo, has_hi;
};

template <int AX>
__device__ __forceinline__ OutFace<AX> out_face(const OutArgs& o, long long t, int wrap) {
    Geo<AX> g{o.n, o.ny, o.nx};
    int a, f;
    g.col(t, a, f);
    const int nc = g.ncell();
    OutFace<AX> r;
    r.fo = g.face(a, f);
    r.PF = g.pf();
    r.has_lo = f > 0 || wrap;
    r.has_hi = f < nc || wrap;
    r.clo = g.cell(a, f > 0 ? f - 1 : nc - 1);
    r.chi = g.cell(a, f < nc ? f : 0);
    return r;
}

// cut every layer's flux by the ratio of its donor cell; delta = Σ_k (Q - Q_cut), the net cut of the face
template <int AX>
__global__ void k_out_cut(OutArgs o, int wrap) {
    const long long t = gtid();
    Geo<AX> g{o.n, o.ny, o.nx};
    if (t >= g.cols()) return;
    const OutFace<AX> fc = out_face<AX>(o, t, wrap);
    double* Q = AX == 0 ? o.Qx : o.Qy;
    const size_t PC = g.pc();
    double d = 0.0;
    for (int k = 0; k < o.n; ++k) {
        const size_t x = k * fc.PF + fc.fo;
        const double q = Q[x];
        const double rlo = fc.has_lo ? o.r[k * PC + fc.clo] : 1.0;
        const double rhi = fc.has_hi ? o.r[k * PC + fc.chi] : 1.0;
        const double nw = q > 0.0 ? q * rlo : (q < 0.0 ? q * rhi : q);
        d += q - nw;
        Q[x] = nw;
    }
    o.delta[t] = d;
}

// give the net cut back to the face's layers: same-way layers by donor room, then counter-flowing layers by |Q|
// (never past 0), then the donor's layers by content; a face that needs the last step out of an empty donor column
// is recorded (the reference raises there)
template <int AX>
__global__ void k_out_give(OutArgs o, int wrap) {
    const long long t = gtid();
    Geo<AX> g{o.n, o.ny, o.nx};
    if (t >= g.cols()) return;
    const OutFace<AX> fc = out_face<AX>(o, t, wrap);
    double* Q = AX == 0 ? o.Qx : o.Qy;
    const size_t PC = g.pc();
    const double delta = o.delta[t];
    const double s = sgn(delta);
    auto room_of = [&](int k, bool lo_side) -> double {
        const bool has = lo_side ? fc.has_lo : fc.has_hi;
        return amax(has ? o.room[k * PC + (lo_side ? fc.clo : fc.chi)] : 0.0, 0.0);
    };
    auto rm_of = [&](int k, double q) -> double {
        const bool same = (sgn(q) == s) && (q != 0.0) && (s != 0.0);
        return same ? (s > 0.0 ? room_of(k, true) : room_of(k, false)) : 0.0;
    };
    double tot = 0.0;
    for (int k = 0; k < o.n; ++k) tot += rm_of(k, Q[k * fc.PF + fc.fo]);
    const double need_ = fabs(delta);
    const double give = npmin(need_, tot);
    const double gq = tot > 0.0 ? give / tot : 0.0;
    double cap = 0.0;
    for (int k = 0; k < o.n; ++k) {
        const size_t x = k * fc.PF + fc.fo;
        const double q = Q[x];
        const double nq = q + s * rm_of(k, q) * gq;
        Q[x] = nq;
        const bool opp = (sgn(nq) == -s) && (s != 0.0);
        cap += opp ? fabs(nq) : 0.0;
    }
    double rest = need_ - give;
    const double take = npmin(rest, cap);
    const double frac = cap > 0.0 ? take / cap : 0.0;
    rest = rest - take;
    const double on = rest > 0.0 ? 1.0 : 0.0;
    double tot2 = 0.0;
    for (int k = 0; k < o.n; ++k) {
        const size_t x = k * fc.PF + fc.fo;
        const double q = Q[x];
        const bool opp = (sgn(q) == -s) && (s != 0.0);
        Q[x] = q - (opp ? q * frac : 0.0);
        const bool lo_side = s > 0.0;
        const bool has = lo_side ? fc.has_lo : fc.has_hi;
        const double full = has ? room0(o, k * PC + (lo_side ? fc.clo : fc.chi), lo_side ? fc.clo : fc.chi) : 0.0;
        tot2 += full * on;
    }
    if (rest > 0.0 && tot2 <= 0.0) atomicMin(o.bad_face, (unsigned long long)t);
    const double gr = tot2 > 0.0 ? rest / tot2 : 0.0;
    for (int k = 0; k < o.n; ++k) {
        const size_t x = k * fc.PF + fc.fo;
        const bool lo_side = s > 0.0;
        const bool has = lo_side ? fc.has_lo : fc.has_hi;
        const double full = has ? room0(o, k * PC + (lo_side ? fc.clo : fc.chi), lo_side ? fc.clo : fc.chi) : 0.0;
        Q[x] = Q[x] + s * (full * on) * gr;
    }
}

// ------------------------------------------------------------------------------------------ Lagrangian step
// h† = h - Δt div / A (A = 1 where area <= 0), div = ∇·Q from divq or from Qx, Qy; the first cell-layer below
// -tol is recorded (linear index, atomicMin); h† = max(h†, 0).  With V0: V0 = h A, V1 = h† A, ok = active & V1 > 0.
__global__ void k_lagrangian(const double* __restrict__ h, const double* __restrict__ divq,
                             const double* __restrict__ Qx, const double* __restrict__ Qy,
                             const double* __restrict__ area, const uint8_t* __restrict__ act, int n, int ny,
                             int nx, double dt, double tol, double* __restrict__ out, double* __restrict__ V0,
                             double* __restrict__ V1, uint8_t* __restrict__ ok, unsigned long long* bad) {
    const long long t = gtid();
    const size_t PC = (size_t)ny * nx;
    if (t >= (long long)(PC * n)) return;
    const int k = (int)(t / PC);
    const size_t c = (size_t)t - (size_t)k * PC;
    double dv;
    if (divq) {
        dv = divq[t];
    } else {
        const int j = (int)(c / nx), i = (int)(c - (size_t)j * nx);
        const size_t fx = ((size_t)k * ny + j) * (nx + 1) + i;
        const size_t fy = ((size_t)k * (ny + 1) + j) * nx + i;
        dv = Qx[fx + 1] - Qx[fx] + Qy[fy + nx] - Qy[fy];
    }
    const double ar = area[c];
    const double hk = h[t];
    double v = hk - dt * dv / (ar > 0.0 ? ar : 1.0);
    if (v < -tol) atomicMin(bad, (unsigned long long)t);
    v = v < 0.0 ? 0.0 : v;
    out[t] = v;
    if (V0) {
        const double a1 = hk * ar, a2 = v * ar;
        V0[t] = a1;
        V1[t] = a2;
        ok[t] = (act[c] && a2 > 0.0) ? 1 : 0;
    }
}

// h_dag = active ? h - Δt div / A : 0 (A = area on active cells, 1 elsewhere); min over all cell-layers as an
// order-preserving u64 key (atomicMin), then the first index holding that minimum
__device__ __forceinline__ unsigned long long order_key(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double h_dag(const OutArgs& o, long long t) {
    const size_t PC = (size_t)o.ny * o.nx;
    const int k = (int)(t / PC);
    const size_t c = (size_t)t - (size_t)k * PC;
    if (!o.act[c]) return 0.0;
    const int j = (int)(c / o.nx), i = (int)(c - (size_t)j * o.nx);
    const size_t fx = ((size_t)k * o.ny + j) * (o.nx + 1) + i;
    const size_t fy = ((size_t)k * (o.ny + 1) + j) * o.nx + i;
    const double div = o.Qx[fx + 1] - o.Qx[fx] + o.Qy[fy + o.nx] - o.Qy[fy];
    const double A = o.area[c];
    return o.h[t] - o.dt * div / (A > 0.0 ? A : 1.0);
}

__global__ void k_hdag_min(OutArgs o, unsigned long long* key) {
    const long long t = gtid();
    const long long N = (long long)o.n * o.ny * o.nx;
    if (t >= N) return;
    atomicMin(key, order_key(h_dag(o, t)));
}

__global__ void k_hdag_where(OutArgs o, const unsigned long long* key, unsigned long long* idx) {
    const long long t = gtid();
    const long long N = (long long)o.n * o.ny * o.nx;
    if (t >= N) return;
    if (order_key(h_dag(o, t)) == *key) atomicMin(idx, (unsigned long long)t);
}

}  // namespace ale
}  // namespace ocx
