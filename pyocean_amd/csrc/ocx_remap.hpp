// Vertical regrid (z*, sigma) and conservative remap (PCM, PLM, PPM_H4) of the ALE coordinate (gfx950), FP64 in the
// reference's per-element arithmetic (no contraction, IEEE divisions), so results are bit-identical.
//
//   regrid   zstar_column / sigma_column / enforce_min_thickness (core/vertical/regrid.py:41-95) under
//            ALECoordinate.regrid (core/components/vertical_coordinate.py:69-109)
//   remap    remap_column and its reconstruction (core/vertical/remap.py:42-383): edges, limiters and the
//            overlap integration with the thickest-piece residual
//
// Columns are independent: one thread per (column, field), columns contiguous across lanes, so every level of a
// column is one coalesced load for the wavefront; the waves of a workgroup are fields of the same 64 columns and
// share the thickness loads through the caches.  The PPM_H4 path is one streaming pass up the column: a 5-cell
// window of (h, u) in registers yields each interface value once, its bounded and continuity-checked edges, the
// cell's limited parabola, and at once that source cell's share of the target cells, which are written as soon as
// no later source cell can reach them.  Nothing of the column is stored.
#pragma once
#include <float.h>
#include <math.h>
#include <stdint.h>

#include "ocx_div.hpp"

namespace ocx {
namespace rmp {

constexpr int PCM = 0, PLM = 1, PPM_H4 = 2;
constexpr double H_NEGLECT = 1e-30;        // negligible thickness of the cell reconstructions
constexpr double H_NEGLECT_EDGE = 1e-10;   // of the edge values
constexpr double H_MIN_FRAC = 1e-5;
constexpr double EPS = DBL_EPSILON;

// Python's min / max of two floats: the first argument unless the second compares smaller / larger
__device__ __forceinline__ double pmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double pmax(double a, double b) { return b > a ? b : a; }
__device__ __forceinline__ double clampu(double e, double a, double b) {   // max(min(e, max(a, b)), min(a, b))
    return pmax(pmin(e, pmax(a, b)), pmin(a, b));
}

// strided view of one column: level k at p[k * M]
struct Col {
    const double* p;
    long long M;
    __device__ __forceinline__ double operator[](int k) const { return p[(size_t)k * M]; }
};

// ------------------------------------------------------------------------------------------ PPM_H4 pieces
// the end cubic's weights, which depend on the 4 end thicknesses only (dz from the end, > 0)
struct EndW {
    double w11, w21, w31, w12, w22, w32, w13, w23, w33, w14, w24, w34, d;
};
__device__ EndW end_weights(const double dz[4]) {
    const double min_frac = 1.0e-6;
    const double h1 = dz[0], h2 = dz[1];
    double h3 = dz[2], h4 = dz[3];
    if ((h2 + h3) < min_frac * h1) h3 = min_frac * h1 - h2;
    if ((h3 + h4) < min_frac * h1) h4 = min_frac * h1 - h3;
    const double h12 = h1 + h2, h23 = h2 + h3, h34 = h3 + h4;
    const double h123 = h12 + h3, h234 = h2 + h34, h1234 = h12 + h34;
    const double I_denB3 = qdiv(1.0, h123 * h12 * h23);
    const double I_h12 = (h123 * h23) * I_denB3;
    const double I_h23 = (h12 * h123) * I_denB3;
    const double I_h123 = (h12 * h23) * I_denB3;
    const double I_denom = qdiv(1.0, h1234 * (h234 * h34));
    const double I_h234 = (h1234 * h34) * I_denom;
    const double I_h1234 = (h234 * h34) * I_denom;
    EndW W;
    W.w11 = -h1 * (I_h1234 + I_h123 + I_h12);
    W.w21 = h1 * h12 * (I_h234 * I_h1234 + I_h23 * (I_h234 + I_h123));
    W.w31 = -h1 * h12 * h123 * I_denom;
    W.w12 = 2.0 * (I_h12 * (1.0 + (h1 + h12) * (I_h1234 + I_h123)) + h1 * I_h1234 * I_h123);
    W.w22 = -2.0 * ((h1 * h12 * I_h1234) * (I_h23 * (I_h234 + I_h123)) +
                    (h1 + h12) * (I_h1234 * I_h234 + I_h23 * (I_h234 + I_h123)));
    W.w32 = 2.0 * ((h1 + h12) * h123 + h1 * h12) * I_denom;
    W.w13 = -3.0 * I_h12 * I_h123 * (1.0 + I_h1234 * ((h1 + h12) + h123));
    W.w23 = 3.0 * I_h23 * (I_h123 + I_h1234 * ((h1 + h12) + h123) * (I_h123 + I_h234));
    W.w33 = -3.0 * ((h1 + h12) + h123) * I_denom;
    W.w14 = 4.0 * I_h1234 * I_h123 * I_h12;
    W.w24 = -4.0 * I_h1234 * (I_h23 * (I_h123 + I_h234));
    W.w34 = 4.0 * I_denom;
    W.d = dz[0];                                  // evaluated at the unmodified end thickness
    return W;
}

// (outer, inner) edge values at one end: the cubic whose means over the 4 end cells are u
__device__ __forceinline__ void end_apply(const EndW& W, double u0, double u1, double u2, double u3, double& e_out,
                                          double& e_in) {
    const double d1 = u1 - u0, d2 = u2 - u1, d3 = u3 - u2;
    const double c0 = ((u0 + (W.w11 * d1)) + (W.w21 * d2)) + (W.w31 * d3);
    const double c1 = ((W.w12 * d1) + (W.w22 * d2)) + (W.w32 * d3);
    const double c2 = ((W.w13 * d1) + (W.w23 * d2)) + (W.w33 * d3);
    const double c3 = ((W.w14 * d1) + (W.w24 * d2)) + (W.w34 * d3);
    e_out = c0;
    e_in = c0 + W.d * (c1 + W.d * (c2 + W.d * c3));
}

__device__ void end_values_h4(const double dz[4], const double u[4], double& e_out, double& e_in) {
    end_apply(end_weights(dz), u[0], u[1], u[2], u[3], e_out, e_in);
}

// the explicit 4th-order interface value between cells 1 and 2 of the 4-cell stencil: its thickness factors ...
struct EdgeW {
    double A1, B1, C2, C3, Dsum;
};
__device__ EdgeW interior_weights(double h0, double h1, double h2, double h3) {
    if (h0 + h1 == 0.0 || h1 + h2 == 0.0 || h2 + h3 == 0.0) {
        const double h_min = H_MIN_FRAC * pmax(H_NEGLECT_EDGE, (h0 + h1) + (h2 + h3));
        h0 = pmax(h_min, h0);
        h1 = pmax(h_min, h1);
        h2 = pmax(h_min, h2);
        h3 = pmax(h_min, h3);
    }
    const double I_h12 = qdiv(1.0, h1 + h2);
    const double I_den_et2 = qdiv(1.0, ((h0 + h1) + h2) * (h0 + h1));
    const double I_h012 = (h0 + h1) * I_den_et2;
    const double I_den_et3 = qdiv(1.0, (h1 + (h2 + h3)) * (h2 + h3));
    const double I_h123 = (h2 + h3) * I_den_et3;
    EdgeW W;
    W.A1 = (1.0 + (h1 * I_h012 + (h0 + h1) * I_h123)) * I_h12 * (h2 * (h2 + h3));
    W.B1 = (1.0 + (h2 * I_h123 + (h2 + h3) * I_h012)) * I_h12 * (h1 * (h0 + h1));
    W.C2 = h1 * (h2 * (h2 + h3)) * I_den_et2;
    W.C3 = h2 * (h1 * (h0 + h1)) * I_den_et3;
    W.Dsum = (h0 + h1) + (h2 + h3);
    return W;
}
// ... applied to the cell means u0..u3
__device__ __forceinline__ double interior_apply(const EdgeW& W, double u0, double u1, double u2, double u3) {
    const double et1 = W.A1 * u1 + W.B1 * u2;
    const double et2 = W.C2 * (u1 - u0);
    const double et3 = W.C3 * (u2 - u3);
    return qdiv(et1 + (et2 + et3), W.Dsum);
}

__device__ double interior_edge(double h0, double h1, double h2, double h3, double u0, double u1, double u2,
                                double u3) {
    return interior_apply(interior_weights(h0, h1, h2, h3), u0, u1, u2, u3);
}

// the thickness ratio h_k / (h_km1 + h_kp1 + 2 h_k) of the edge-bounding slope, and whether it is defined
__device__ __forceinline__ double slope_ratio(double hm, double hc, double hp, bool& ok) {
    const double den = (hm + hp) + 2.0 * hc;
    ok = den > 0.0;
    return ok ? qdiv(hc, den) : 0.0;
}
// the cell's limited slope times h used to bound its edges, from its ratio (neighbours clamped at the column ends)
__device__ __forceinline__ double bound_slope_r(bool ok, double ratio, double um, double uc, double up) {
    double slope_x_h = 0.0;
    if (ok) {
        const double sl = uc - um;
        const double sc = (up - um) * ratio;
        const double sr = up - uc;
        if (sl * sr > 0.0) {
            const double m = pmin(pmin(fabs(sl), fabs(sc)), fabs(sr));
            slope_x_h = sc >= 0.0 ? m : -m;
        }
    }
    return slope_x_h;
}
__device__ __forceinline__ double bound_slope(double hm, double hc, double hp, double um, double uc, double up) {
    bool ok;
    const double ratio = slope_ratio(hm, hc, hp, ok);
    return bound_slope_r(ok, ratio, um, uc, up);
}
// an edge value bounded by its cell and the neighbour on its side (unb = u of that neighbour, or the cell itself at
// the column end); right: the upper edge
__device__ __forceinline__ double bound_edge(double e, double slope_x_h, double unb, double uc, bool right) {
    const double sgn = slope_x_h >= 0.0 ? 1.0 : -1.0;
    if ((unb - e) * (e - uc) < 0.0) {
        const double m = pmin(fabs(slope_x_h), fabs(e - uc));
        e = right ? uc + sgn * m : uc - sgn * m;
    }
    return clampu(e, unb, uc);
}

// the Colella-Woodward limited parabola of an interior cell from its final edges
__device__ __forceinline__ void cw_coef(double u_l, double u_c, double u_r, double el, double er, double& c0,
                                        double& c1, double& c2) {
    if ((u_r - u_c) * (u_c - u_l) <= 0.0) {
        el = u_c;
        er = u_c;
    } else {
        const double expr1 = 3.0 * (er - el) * ((u_c - el) + (u_c - er));
        const double expr2 = (er - el) * (er - el);
        if (expr1 > expr2) {
            el = u_c + 2.0 * (u_c - er);
            el = clampu(el, u_l, u_c);
        } else if (expr1 < -expr2) {
            er = u_c + 2.0 * (u_c - el);
            er = clampu(er, u_r, u_c);
        }
    }
    if (fabs(er - el) < pmax(1e-60, EPS * fabs(u_c))) {
        el = u_c;
        er = u_c;
    }
    c0 = el;
    c1 = 4.0 * (u_c - el) + 2.0 * (u_c - er);
    c2 = 3.0 * ((er - u_c) + (el - u_c));
}

// ------------------------------------------------------------------------------------------ PLM (short columns)
__device__ double plm_slope_wa(double h_l, double h_c, double h_r, double u_l, double u_c, double u_r) {
    const double sigma_r = u_r - u_c;
    const double sigma_l = u_c - u_l;
    const double sigma_c = 2.0 * (u_r - u_l) * qdiv(h_c, h_l + 2.0 * h_c + h_r + H_NEGLECT);
    const double u_min = pmin(pmin(u_l, u_c), u_r);
    const double u_max = pmax(pmax(u_l, u_c), u_r);
    double slope;
    if (sigma_l * sigma_r > 0.0) {
        const double s = pmin(fabs(sigma_c), 2.0 * pmin(u_c - u_min, u_max - u_c));
        slope = sigma_c >= 0.0 ? s : -s;
    } else {
        slope = 0.0;
    }
    if (u_c - 0.5 * fabs(slope) < u_min || u_c + 0.5 * fabs(slope) > u_max) slope = slope * (1.0 - EPS);
    if (fabs(slope) < 1e-140) slope = 0.0;
    return slope;
}

__device__ double plm_mono(double u_l, double u_c, double u_r, double s_l, double s_c, double s_r) {
    const double almost_two = 2.0 * (1.0 - EPS);
    const double e_r = u_l + 0.5 * s_l;
    const double e_l = u_r - 0.5 * s_r;
    double slp = fabs(s_c);
    double edge = u_c - 0.5 * s_c;
    if ((edge - e_r) * (u_c - edge) < 0.0) {
        edge = 0.5 * (edge + e_r);
        slp = pmin(slp, fabs(edge - u_c) * almost_two);
    }
    edge = u_c + 0.5 * s_c;
    if ((edge - u_c) * (e_l - edge) < 0.0) {
        edge = 0.5 * (edge + e_l);
        slp = pmin(slp, fabs(edge - u_c) * almost_two);
    }
    return s_c >= 0.0 ? slp : -slp;
}

// slope of cell k (0 at the column ends)
__device__ __forceinline__ double plm_slp(const Col& h, const Col& u, int n, int k) {
    if (k <= 0 || k >= n - 1) return 0.0;
    return plm_slope_wa(h[k - 1], h[k], h[k + 1], u[k - 1], u[k], u[k + 1]);
}

__device__ __forceinline__ double plm_mslp(const Col& h, const Col& u, int n, int k) {
    if (k <= 0 || k >= n - 1) return 0.0;
    return plm_mono(u[k - 1], u[k], u[k + 1], plm_slp(h, u, n, k - 1), plm_slp(h, u, n, k), plm_slp(h, u, n, k + 1));
}

// PLM coefficients of cell k (n >= 3), recomputed from the column (short columns only reach this)
__device__ void plm_coef(const Col& h, const Col& u, int n, int k, double& c0, double& c1, double& c2) {
    c2 = 0.0;
    if (k <= 0 || k >= n - 1) {
        c0 = u[k];
        c1 = 0.0;
        return;
    }
    const double uk = u[k];
    const double m = plm_mslp(h, u, n, k);
    const double u_l = uk - 0.5 * m;
    const double u_r = uk + 0.5 * m;
    c0 = u_l;
    c1 = u_r - u_l;
    const double edge = c1 + c0;
    const double m1 = plm_mslp(h, u, n, k + 1);
    const double s1 = plm_slp(h, u, n, k + 1);
    const double e_r = u[k + 1] - 0.5 * (s1 >= 0.0 ? fabs(m1) : -fabs(m1));
    if ((edge - uk) * (e_r - edge) < 0.0) c1 = c1 * (1.0 - EPS);
}

// ------------------------------------------------------------------------------------------ overlap integration
__device__ __forceinline__ double average(double c0, double c1, double c2, double xa, double xb) {
    if (xb - xa <= 1e-14) {
        const double x = 0.5 * (xa + xb);
        return c0 + x * (c1 + x * c2);
    }
    return c0 + 0.5 * c1 * (xa + xb) + div_const(c2 * (xa * xa + xa * xb + xb * xb), 3.0, 1.0 / 3.0);
}

// the target side of one column: the cells the source has reached, their running content and the vanished cells
// still waiting for a local value
struct Targets {
    Col ht;
    double* out;        // level k at out[k * M]
    long long M;
    int n1;
    int j;              // open target cell of the overlap walk
    double tj;          // its lower interface t_edge[j]
    double content;     // its content so far
    int jv;             // next target whose lower interface has not been located in a source cell
    double tv;          // t_edge[jv]

    __device__ __forceinline__ void put(int k, double v) { out[(size_t)k * M] = v; }
    __device__ __forceinline__ double top(int jj, double t_lo, double b) const {   // upper end of target jj
        return jj < n1 - 1 ? t_lo + ht[jj] : pmax(t_lo + ht[jj], b);
    }
    // a target the walk has passed: its mean (vanished targets are set by their local value instead)
    __device__ __forceinline__ void finish(int jj, double c) {
        const double h = ht[jj];
        if (h > 0.0) put(jj, qdiv(c, h));
    }
};

// one active source cell [a, a + hi) with mean us and polynomial (c0, c1, c2): its pieces over the targets, the
// largest taking the residual u h - Σ others so the cell's content is exact
__device__ void merge_cell(Targets& T, double a, double hi, double us, double c0, double c1, double c2) {
    const double b = a + hi;
    // pass 1: the geometry of the pieces, which one is thickest
    int npc = 0, ithick = 0, jthick = T.j;
    double thick = -1.0;
    {
        int j = T.j;
        double t = T.tj;
        while (true) {
            const double lo = pmax(a, t);
            const double hi_t = T.top(j, t, b);
            const double up = pmin(b, hi_t);
            const double L = up - lo;
            if (L > 0.0) {
                if (L > thick) {
                    thick = L;
                    ithick = npc;
                    jthick = j;
                }
                ++npc;
            }
            if (hi_t <= b && j < T.n1 - 1) {
                t = t + T.ht[j];
                ++j;
            } else {
                break;
            }
        }
    }
    // pass 2: the pieces in order; every target passed here except the thickest piece's is final
    double s = 0.0, c_thick = 0.0;
    int p = 0;
    int j = T.j;
    double t = T.tj, c = T.content;
    while (true) {
        const double lo = pmax(a, t);
        const double hi_t = T.top(j, t, b);
        const double up = pmin(b, hi_t);
        const double L = up - lo;
        if (L > 0.0) {
            if (p != ithick) {
                const double piece = L * average(c0, c1, c2, qdiv(lo - a, hi), qdiv(up - a, hi));
                s += piece;
                c = c + piece;
            } else {
                c_thick = c;                 // the content before this cell's piece
            }
            ++p;
        }
        if (hi_t <= b && j < T.n1 - 1) {
            if (j != jthick || npc == 0) T.finish(j, c);
            t = t + T.ht[j];
            ++j;
            c = 0.0;
        } else {
            break;
        }
    }
    if (npc > 0) {
        const double tp = us * hi - s;
        if (jthick == j) {
            c = c_thick + tp;                // the open target holds the thickest piece
        } else {
            T.finish(jthick, c_thick + tp);
        }
    }
    T.j = j;
    T.tj = t;
    T.content = c;
}

// vanished targets whose lower interface lies in the source cell ending at b get the cell's polynomial there
__device__ __forceinline__ void locate_vanished(Targets& T, double a, double hi, double b, double c0, double c1,
                                                double c2) {
    while (T.jv < T.n1 && T.tv <= b) {
        const double h = T.ht[T.jv];
        if (!(h > 0.0)) {
            const double xi = qdiv(T.tv - a, hi);
            T.put(T.jv, c0 + xi * (c1 + xi * c2));
        }
        T.tv = T.tv + h;
        ++T.jv;
    }
}

// after the last active source cell: the open target and those beyond are final; vanished ones above the source
// take the last active cell's mean
__device__ void finish_targets(Targets& T, double u_last) {
    T.finish(T.j, T.content);
    for (int jj = T.j + 1; jj < T.n1; ++jj) T.finish(jj, 0.0);
    while (T.jv < T.n1) {
        const double h = T.ht[T.jv];
        if (!(h > 0.0)) T.put(T.jv, u_last);
        ++T.jv;
    }
}


// ------------------------------------------------------------------------------------------ grouped PPM_H4 remap
// G fields of one column in one thread: the thickness work - the interface stencils' reciprocals and products, the
// slope ratios, the end cubics' weights and the whole overlap geometry - is done once for the G fields; per field
// remain the sums with its values, one division per interface and one per target cell.  Every value is computed
// by the same operations as the single-field path.
template <int G>
struct GTargets {
    Col ht;
    double* out[G];
    long long M;
    int n1, ng;
    int j;
    double tj;
    double content[G];
    int jv;
    double tv;

    __device__ __forceinline__ double top(int jj, double t_lo, double b) const {
        return jj < n1 - 1 ? t_lo + ht[jj] : pmax(t_lo + ht[jj], b);
    }
    __device__ __forceinline__ void put(int g, int k, double v) {
        if (g < ng) out[g][(size_t)k * M] = v;
    }
    __device__ __forceinline__ void finish(int jj, const double* c) {
        const double h = ht[jj];
        if (h > 0.0) {
#pragma unroll
            for (int g = 0; g < G; ++g) put(g, jj, qdiv(c[g], h));
        }
    }
};

template <int G>
__device__ void merge_cell_g(GTargets<G>& T, double a, double hi, const double* us, const double* c0,
                             const double* c1, const double* c2) {
    const double b = a + hi;
    int npc = 0, ithick = 0, jthick = T.j;
    double thick = -1.0;
    {
        int j = T.j;
        double t = T.tj;
        while (true) {
            const double lo = pmax(a, t);
            const double hi_t = T.top(j, t, b);
            const double up = pmin(b, hi_t);
            const double L = up - lo;
            if (L > 0.0) {
                if (L > thick) {
                    thick = L;
                    ithick = npc;
                    jthick = j;
                }
                ++npc;
            }
            if (hi_t <= b && j < T.n1 - 1) {
                t = t + T.ht[j];
                ++j;
            } else {
                break;
            }
        }
    }
    double s[G], c[G], c_thick[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        s[g] = 0.0;
        c[g] = T.content[g];
        c_thick[g] = 0.0;
    }
    int p = 0;
    int j = T.j;
    double t = T.tj;
    while (true) {
        const double lo = pmax(a, t);
        const double hi_t = T.top(j, t, b);
        const double up = pmin(b, hi_t);
        const double L = up - lo;
        if (L > 0.0) {
            if (p != ithick) {
                const double xa = qdiv(lo - a, hi), xb = qdiv(up - a, hi);
                if (xb - xa <= 1e-14) {
                    const double x = 0.5 * (xa + xb);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const double piece = L * (c0[g] + x * (c1[g] + x * c2[g]));
                        s[g] += piece;
                        c[g] = c[g] + piece;
                    }
                } else {
                    const double sx = xa + xb;
                    const double s2 = xa * xa + xa * xb + xb * xb;
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const double piece = L * (c0[g] + 0.5 * c1[g] * sx + div_const(c2[g] * s2, 3.0, 1.0 / 3.0));
                        s[g] += piece;
                        c[g] = c[g] + piece;
                    }
                }
            } else {
#pragma unroll
                for (int g = 0; g < G; ++g) c_thick[g] = c[g];
            }
            ++p;
        }
        if (hi_t <= b && j < T.n1 - 1) {
            if (j != jthick || npc == 0) T.finish(j, c);
            t = t + T.ht[j];
            ++j;
#pragma unroll
            for (int g = 0; g < G; ++g) c[g] = 0.0;
        } else {
            break;
        }
    }
    if (npc > 0) {
        double tp[G];
#pragma unroll
        for (int g = 0; g < G; ++g) tp[g] = c_thick[g] + (us[g] * hi - s[g]);
        if (jthick == j) {
#pragma unroll
            for (int g = 0; g < G; ++g) c[g] = tp[g];
        } else {
            T.finish(jthick, tp);
        }
    }
    T.j = j;
    T.tj = t;
#pragma unroll
    for (int g = 0; g < G; ++g) T.content[g] = c[g];
}

template <int G>
__device__ __forceinline__ void locate_vanished_g(GTargets<G>& T, double a, double hi, double b, const double* c0,
                                                  const double* c1, const double* c2) {
    while (T.jv < T.n1 && T.tv <= b) {
        const double h = T.ht[T.jv];
        if (!(h > 0.0)) {
            const double xi = qdiv(T.tv - a, hi);
#pragma unroll
            for (int g = 0; g < G; ++g) T.put(g, T.jv, c0[g] + xi * (c1[g] + xi * c2[g]));
        }
        T.tv = T.tv + h;
        ++T.jv;
    }
}

// ------------------------------------------------------------------------------------------ the remap kernel
struct RemapArgs {
    const double *h_src, *f, *h_tgt;
    const uint8_t* wet;
    double* out;
    int n0, n1, nf, scheme;
    long long M;
};

// column col of field q; ht: the column's target thicknesses (in HBM, or staged in LDS)
__device__ void remap_one(const RemapArgs& r, long long col, int q, Col ht) {
    const long long M = r.M;
    const Col hs{r.h_src + col, M};
    const Col u{r.f + (size_t)q * r.n0 * M + col, M};
    double* out = r.out + (size_t)q * r.n1 * M + col;
    const int n0 = r.n0, n1 = r.n1;
    if (!r.wet[col]) {
        for (int k = 0; k < n1; ++k) out[(size_t)k * M] = n0 == n1 ? u[k] : __builtin_nan("");
        return;
    }
    int n_active = 0, last = -1;
    for (int k = 0; k < n0; ++k) {
        if (hs[k] > 0.0) {
            ++n_active;
            last = k;
        }
    }
    if (last < 0) {                                   // no source thickness: keep the values
        double mean = 0.0;
        for (int k = 0; k < n0; ++k) mean += u[k] / n0;
        for (int k = 0; k < n1; ++k) out[(size_t)k * M] = n0 == n1 ? u[k] : mean;
        return;
    }
    int scheme = r.scheme;
    if (n_active <= 1) scheme = PCM;
    else if (n_active <= 3 && scheme > PLM) scheme = PLM;
    if ((scheme == PLM && n0 < 3) || (scheme == PPM_H4 && n0 < 4)) scheme = PCM;

    Targets T;
    T.ht = ht;
    T.out = out;
    T.M = M;
    T.n1 = n1;
    T.j = 0;
    T.tj = 0.0;
    T.content = 0.0;
    T.jv = 0;
    T.tv = 0.0;

    if (scheme != PPM_H4) {
        double a = 0.0;
        for (int i = 0; i < n0; ++i) {
            const double hi = hs[i];
            if (!(hi > 0.0)) continue;
            double c0, c1, c2;
            if (scheme == PLM) {
                plm_coef(hs, u, n0, i, c0, c1, c2);
            } else {
                c0 = u[i];
                c1 = 0.0;
                c2 = 0.0;
            }
            const double b = a + hi;
            merge_cell(T, a, hi, u[i], c0, c1, c2);
            locate_vanished(T, a, hi, b, c0, c1, c2);
            a = b;
        }
        finish_targets(T, u[last]);
        return;
    }

    // PPM_H4: the window w[0..4] holds cells i-2 .. i+2 (entries outside the column are never used)
    double wh[5], wu[5];
    for (int d = 0; d < 5; ++d) {
        const int k = d - 2;
        wh[d] = (k >= 0 && k < n0) ? hs[k] : 0.0;
        wu[d] = (k >= 0 && k < n0) ? u[k] : 0.0;
    }
    // end values: bottom (interfaces 0 and 1) and top (interfaces n0 and n0 - 1)
    double e_bot_out, e_bot_in, e_top_out, e_top_in;
    {
        double dz[4], uu[4];
        for (int d = 0; d < 4; ++d) {
            dz[d] = pmax(H_NEGLECT_EDGE, hs[d]);
            uu[d] = u[d];
        }
        end_values_h4(dz, uu, e_bot_out, e_bot_in);
        for (int d = 0; d < 4; ++d) {
            dz[d] = pmax(H_NEGLECT_EDGE, hs[n0 - 1 - d]);
            uu[d] = u[n0 - 1 - d];
        }
        end_values_h4(dz, uu, e_top_out, e_top_in);
    }
    // raw value of interface m (0 .. n0) from the window, which must hold cells m-2 .. m+1 at positions
    // (m - i) + 0 .. (m - i) + 3 when called at step i with m = i + 1
    auto interface_raw = [&](int m) -> double {
        if (m == 0) return e_bot_out;
        if (m == 1) return e_bot_in;
        if (m == n0) return e_top_out;
        if (m == n0 - 1) return e_top_in;
        return interior_edge(wh[1], wh[2], wh[3], wh[4], wu[1], wu[2], wu[3], wu[4]);
    };
    // bound slope of the cell at window position d (cell i - 2 + d), neighbours clamped at the column ends
    auto slope_at = [&](int d, int k) -> double {
        const int dm = k > 0 ? d - 1 : d, dp = k < n0 - 1 ? d + 1 : d;
        return bound_slope(wh[dm], wh[d], wh[dp], wu[dm], wu[d], wu[dp]);
    };
    // cell 0: its lower edge (no continuity check below it)
    double s_cur = slope_at(2, 0);                                  // slope of cell i (window position 2)
    double fl = bound_edge(e_bot_out, s_cur, wu[2], wu[2], false);  // final lower edge of cell i
    double a = 0.0;
    for (int i = 0; i < n0; ++i) {
        // interface i + 1: raw value, the upper edge of cell i and the lower edge of cell i + 1, their check
        double fr, fl_next = 0.0, s_next = 0.0;
        if (i + 1 < n0) {
            const double raw = interface_raw(i + 1);
            s_next = slope_at(3, i + 1);
            const double br = bound_edge(raw, s_cur, wu[3], wu[2], true);
            const double bl = bound_edge(raw, s_next, wu[2], wu[3], false);
            fr = br;
            fl_next = bl;
            if ((bl - br) * (wu[3] - wu[2]) < 0.0) {
                const double av = clampu(0.5 * (br + bl), wu[2], wu[3]);
                fr = av;
                fl_next = av;
            }
        } else {
            fr = bound_edge(e_top_out, s_cur, wu[2], wu[2], true);
        }
        // the cell's polynomial: PCM at both ends, the limited parabola inside
        double c0 = wu[2], c1 = 0.0, c2 = 0.0;
        if (i > 0 && i < n0 - 1) cw_coef(wu[1], wu[2], wu[3], fl, fr, c0, c1, c2);
        const double hi = wh[2];
        if (hi > 0.0) {
            const double b = a + hi;
            merge_cell(T, a, hi, wu[2], c0, c1, c2);
            locate_vanished(T, a, hi, b, c0, c1, c2);
            a = b;
        }
        // shift the window up one cell
        for (int d = 0; d < 4; ++d) {
            wh[d] = wh[d + 1];
            wu[d] = wu[d + 1];
        }
        const int k_in = i + 3;
        wh[4] = k_in < n0 ? hs[k_in] : 0.0;
        wu[4] = k_in < n0 ? u[k_in] : 0.0;
        fl = fl_next;
        s_cur = s_next;
    }
    finish_targets(T, u[last]);
}


// fields q0 .. q0 + G - 1 of column col (those beyond nf are computed on field q0 and not stored): the PPM_H4 path
// shares the thickness work; the other schemes and dry or vanished columns take the single-field path per field
template <int G>
__device__ void remap_group(const RemapArgs& r, long long col, int q0, Col ht) {
    const long long M = r.M;
    const int n0 = r.n0, n1 = r.n1;
    const int ng = r.nf - q0 < G ? r.nf - q0 : G;
    const Col hs{r.h_src + col, M};
    bool ppm = false;
    int last = -1;
    if (r.wet[col] && r.scheme == PPM_H4 && n0 >= 4) {
        int n_active = 0;
        for (int k = 0; k < n0; ++k) {
            if (hs[k] > 0.0) {
                ++n_active;
                last = k;
            }
        }
        ppm = n_active > 3;
    }
    if (!ppm) {
        for (int g = 0; g < ng; ++g) remap_one(r, col, q0 + g, ht);
        return;
    }
    Col u[G];
    GTargets<G> T;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int q = q0 + (g < ng ? g : 0);
        u[g] = Col{r.f + (size_t)q * n0 * M + col, M};
        T.out[g] = r.out + (size_t)q * n1 * M + col;
        T.content[g] = 0.0;
    }
    T.ht = ht;
    T.M = M;
    T.n1 = n1;
    T.ng = ng;
    T.j = 0;
    T.tj = 0.0;
    T.jv = 0;
    T.tv = 0.0;

    double wh[5], wu[G][5];
#pragma unroll
    for (int d = 0; d < 5; ++d) {
        const int k = d - 2;
        const bool in = k >= 0 && k < n0;
        wh[d] = in ? hs[k] : 0.0;
#pragma unroll
        for (int g = 0; g < G; ++g) wu[g][d] = in ? u[g][k] : 0.0;
    }
    double e_bot_in[G], e_top_out[G], e_top_in[G], fl[G], s_cur[G];
    {
        double dz[4];
        for (int d = 0; d < 4; ++d) dz[d] = pmax(H_NEGLECT_EDGE, hs[d]);
        const EndW Wb = end_weights(dz);
        bool ok;
        const double ratio = slope_ratio(wh[2], wh[2], wh[3], ok);      // cell 0: km1 = 0
#pragma unroll
        for (int g = 0; g < G; ++g) {
            double e_out;
            end_apply(Wb, u[g][0], u[g][1], u[g][2], u[g][3], e_out, e_bot_in[g]);
            s_cur[g] = bound_slope_r(ok, ratio, wu[g][2], wu[g][2], wu[g][3]);
            fl[g] = bound_edge(e_out, s_cur[g], wu[g][2], wu[g][2], false);
            e_top_out[g] = 0.0;
            e_top_in[g] = 0.0;
        }
    }
    double a = 0.0;
    for (int i = 0; i < n0; ++i) {
        if (i == n0 - 2) {
            // the top end cubic, from cells n0-1 .. n0-4 (window positions 3 - d at step n0 - 2)
            double dz[4];
            for (int d = 0; d < 4; ++d) dz[d] = pmax(H_NEGLECT_EDGE, hs[n0 - 1 - d]);
            const EndW Wt = end_weights(dz);
#pragma unroll
            for (int g = 0; g < G; ++g)
                end_apply(Wt, u[g][n0 - 1], u[g][n0 - 2], u[g][n0 - 3], u[g][n0 - 4], e_top_out[g], e_top_in[g]);
        }
        double fr[G], fl_next[G], s_next[G];
        if (i + 1 < n0) {
            const int m = i + 1;
            const bool interior = m >= 2 && m <= n0 - 2;
            EdgeW W;
            if (interior) W = interior_weights(wh[1], wh[2], wh[3], wh[4]);
            bool ok;
            const double ratio = i + 1 < n0 - 1 ? slope_ratio(wh[2], wh[3], wh[4], ok) : slope_ratio(wh[2], wh[3], wh[3], ok);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const double raw = m == 1 ? e_bot_in[g] : (m == n0 - 1 ? e_top_in[g]
                                                            : interior_apply(W, wu[g][1], wu[g][2], wu[g][3], wu[g][4]));
                const double up = i + 1 < n0 - 1 ? wu[g][4] : wu[g][3];
                s_next[g] = bound_slope_r(ok, ratio, wu[g][2], wu[g][3], up);
                const double br = bound_edge(raw, s_cur[g], wu[g][3], wu[g][2], true);
                const double bl = bound_edge(raw, s_next[g], wu[g][2], wu[g][3], false);
                fr[g] = br;
                fl_next[g] = bl;
                if ((bl - br) * (wu[g][3] - wu[g][2]) < 0.0) {
                    const double av = clampu(0.5 * (br + bl), wu[g][2], wu[g][3]);
                    fr[g] = av;
                    fl_next[g] = av;
                }
            }
        } else {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                fr[g] = bound_edge(e_top_out[g], s_cur[g], wu[g][2], wu[g][2], true);
                fl_next[g] = 0.0;
                s_next[g] = 0.0;
            }
        }
        const double hi = wh[2];
        if (hi > 0.0) {
            double c0[G], c1[G], c2[G], us[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                c0[g] = wu[g][2];
                c1[g] = 0.0;
                c2[g] = 0.0;
                us[g] = wu[g][2];
                if (i > 0 && i < n0 - 1) cw_coef(wu[g][1], wu[g][2], wu[g][3], fl[g], fr[g], c0[g], c1[g], c2[g]);
            }
            const double b = a + hi;
            merge_cell_g<G>(T, a, hi, us, c0, c1, c2);
            locate_vanished_g<G>(T, a, hi, b, c0, c1, c2);
            a = b;
        }
        for (int d = 0; d < 4; ++d) {
            wh[d] = wh[d + 1];
#pragma unroll
            for (int g = 0; g < G; ++g) wu[g][d] = wu[g][d + 1];
        }
        const int k_in = i + 3;
        wh[4] = k_in < n0 ? hs[k_in] : 0.0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            wu[g][4] = k_in < n0 ? u[g][k_in] : 0.0;
            fl[g] = fl_next[g];
            s_cur[g] = s_next[g];
        }
    }
    // the open target and those beyond; vanished targets above the source take the last active cell's mean
    T.finish(T.j, T.content);
    double zero[G];
#pragma unroll
    for (int g = 0; g < G; ++g) zero[g] = 0.0;
    for (int jj = T.j + 1; jj < n1; ++jj) T.finish(jj, zero);
    while (T.jv < n1) {
        const double h = T.ht[T.jv];
        if (!(h > 0.0)) {
#pragma unroll
            for (int g = 0; g < G; ++g) T.put(g, T.jv, u[g][last]);
        }
        ++T.jv;
    }
}

// The target thicknesses of the workgroup's 64 columns (blockIdx.x), staged in LDS as [layer][lane]: the overlap
// walks read them at layers that differ from lane to lane, which from HBM costs a cache line per lane and an L2
// round trip per step of the walk; in LDS a lane's column sits at stride 64, so every lane reads its own bank.
__device__ __forceinline__ Col stage_targets(double* sht, const RemapArgs& r) {
    const long long col0 = (long long)blockIdx.x * 64;
    const int nthr = blockDim.x * blockDim.y, tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int idx = tid; idx < r.n1 * 64; idx += nthr) {
        const int k = idx >> 6, l = idx & 63;
        sht[idx] = col0 + l < r.M ? r.h_tgt[(size_t)k * r.M + col0 + l] : 0.0;
    }
    __syncthreads();
    return Col{sht + threadIdx.x, 64};
}

// block (64, FW): x over 64 columns, y over groups of G fields from group goff on; grid (ceil(M / 64), rows).  STAGED:
// the target thicknesses come from LDS (n1 * 512 bytes of dynamic shared memory), shared by the FW waves.
#ifndef REMAP_FW_MAX
#define REMAP_FW_MAX 4
#endif

template <int G, bool STAGED>
__global__ __launch_bounds__(64 * REMAP_FW_MAX) void k_remap_g(RemapArgs r, int goff) {
    extern __shared__ double sht[];
    const long long col = (long long)blockIdx.x * 64 + threadIdx.x;
    const int q0 = (goff + blockIdx.y * blockDim.y + threadIdx.y) * G;
    Col ht{r.h_tgt + col, r.M};
    if constexpr (STAGED) ht = stage_targets(sht, r);
    if (col >= r.M || q0 >= r.nf) return;
    remap_group<G>(r, col, q0, ht);
}

// block (64, FW): x over columns, y over fields; grid (ceil(M / 64), ceil(nf / FW))
template <bool STAGED>
__global__ __launch_bounds__(64 * REMAP_FW_MAX) void k_remap(RemapArgs r) {
    extern __shared__ double sht[];
    const long long col = (long long)blockIdx.x * 64 + threadIdx.x;
    const int q = blockIdx.y * blockDim.y + threadIdx.y;
    Col ht{r.h_tgt + col, r.M};
    if constexpr (STAGED) ht = stage_targets(sht, r);
    if (col >= r.M || q >= r.nf) return;
    remap_one(r, col, q, ht);
}

// ------------------------------------------------------------------------------------------ regrid
// one thread per column: wet = active & H + eta > 0 & Σ h > 0 (H with NaN read as 0); h_new = the rule's layers where
// wet, h elsewhere.  rule 0: z* (dz = nominal thicknesses, surface first); 1: sigma (dz = fractions, bed first).
// A wet column the rule cannot build (non-finite H or D, D <= 0, z* with H <= 0) is recorded in *bad (atomicMin of
// its index).
__global__ void k_regrid(const double* __restrict__ H, const double* __restrict__ eta,
                         const uint8_t* __restrict__ active, const double* __restrict__ h, int n, long long M,
                         int rule, const double* __restrict__ dz, double h_min, double* __restrict__ h_new,
                         uint8_t* __restrict__ wet_out, unsigned long long* bad) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M) return;
    double Hc = H[c];
    Hc = Hc != Hc ? 0.0 : (Hc == __builtin_inf() ? DBL_MAX : (Hc == -__builtin_inf() ? -DBL_MAX : Hc));
    const double et = eta[c];
    double tot = 0.0;
    for (int k = 0; k < n; ++k) tot += h[(size_t)k * M + c];
    const bool wet = active[c] && (Hc + et > 0.0) && (tot > 0.0);
    if (wet_out) wet_out[c] = wet ? 1 : 0;
    if (!wet) {
        for (int k = 0; k < n; ++k) h_new[(size_t)k * M + c] = h[(size_t)k * M + c];
        return;
    }
    const double D = Hc + et;
    if (!(fabs(Hc) <= DBL_MAX) || !(fabs(D) <= DBL_MAX) || (rule == 0 && !(Hc > 0.0))) {
        atomicMin(bad, (unsigned long long)c);
        for (int k = 0; k < n; ++k) h_new[(size_t)k * M + c] = h[(size_t)k * M + c];
        return;
    }
    // interfaces are produced top-down (z*) or bottom-up (sigma) and written as thicknesses in a second sweep
    if (rule == 0) {
        // z*: e[0] = eta, e[k+1] = e[k] - s dz[k], e[n] = -H, raised from the bed by hm; h_k = e[n-1-k] - e[n-k]
        const double hm = pmin(h_min, D / n);
        const double s = D / Hc;
        // nominal positions from the top, then the floor from the bed up: walk the interfaces bottom-up, each
        // one's nominal value recomputed by the same sequential sum (kept in h_new as scratch)
        double e = et;
        for (int k = 0; k < n; ++k) {                 // h_new[n-1-k] temporarily holds the nominal e[k+1]
            e = e - s * dz[k];
            h_new[(size_t)(n - 1 - k) * M + c] = e;
        }
        double below = -Hc;                           // e[n]
        for (int k = n - 1; k >= 0; --k) {
            double ek = k == 0 ? et : h_new[(size_t)(n - k) * M + c];     // nominal e[k]
            if (ek < below + hm) ek = below + hm;
            // layer n-1-k (bed first) spans e[k+1] .. e[k]: thickness e[k] - e[k+1]
            h_new[(size_t)(n - 1 - k) * M + c] = ek - below;
            below = ek;
        }
    } else {
        // sigma: zw[0] = -H, zw[j+1] = zw[j] + D dsigma[j], zw[n] = eta, then the minimum thickness from the top
        // and from the bed; interfaces kept in h_new (zw[j+1] at level j) until the thicknesses are formed
        double z = -Hc;
        for (int j = 0; j < n; ++j) {
            z = z + D * dz[j];
            h_new[(size_t)j * M + c] = j == n - 1 ? et : z;              // zw[j + 1]
        }
        auto zw = [&](int j) -> double { return j == 0 ? -Hc : h_new[(size_t)(j - 1) * M + c]; };
        const double hm = pmin(h_min, (zw(n) - zw(0)) / n);
        for (int j = n - 1; j > 0; --j) {
            const double up = zw(j + 1);
            if (zw(j) > up - hm) h_new[(size_t)(j - 1) * M + c] = up - hm;
        }
        for (int j = 1; j < n; ++j) {
            const double dn = zw(j - 1);
            if (zw(j) < dn + hm) h_new[(size_t)(j - 1) * M + c] = dn + hm;
        }
        double prev = -Hc;
        for (int j = 0; j < n; ++j) {
            const double zj = h_new[(size_t)j * M + c];
            h_new[(size_t)j * M + c] = zj - prev;
            prev = zj;
        }
    }
}

}  // namespace rmp
}  // namespace ocx

namespace ocx {
namespace rmp {

// ------------------------------------------------------------------------------------------ hybrid (HYCOM1) rule
// hycom1_column / _p1m_h2_positions / _keep_consistent / window_shallow_sigma (core/vertical/regrid.py:98-303) per
// column, one thread each; the column's working arrays live in a caller-provided scratch laid out level-major
// (level k of column c at k * M + c), so the lanes of a wavefront stay coalesced.

struct WCol {                     // writable strided column
    double* p;
    long long M;
    __device__ __forceinline__ double& operator[](int k) const { return p[(size_t)k * M]; }
};

__device__ void enforce_min_thickness_w(const WCol& zw, int n, double h_min) {
    const double hm = pmin(h_min, (zw[n] - zw[0]) / n);
    for (int j = n - 1; j > 0; --j)
        if (zw[j] > zw[j + 1] - hm) zw[j] = zw[j + 1] - hm;
    for (int j = 1; j < n; ++j)
        if (zw[j] < zw[j - 1] + hm) zw[j] = zw[j - 1] + hm;
}

__device__ void sigma_column_w(double H, double eta, const double* dsigma, int n, double h_min, const WCol& zw) {
    const double D = H + eta;
    zw[0] = -H;
    for (int j = 0; j < n; ++j) zw[j + 1] = zw[j] + D * dsigma[j];
    zw[n] = eta;
    enforce_min_thickness_w(zw, n, h_min);
}

// z* interfaces bed first: zw[n - k] is the surface-first e[k]
__device__ void zstar_column_w(double H, double eta, const double* dz_top, int n, double h_min, const WCol& zw) {
    const double D = H + eta;
    const double hm = pmin(h_min, D / n);
    const double s = D / H;
    zw[n] = eta;
    for (int k = 0; k < n; ++k) zw[n - k - 1] = zw[n - k] - s * dz_top[k];
    zw[0] = -H;
    for (int j = 1; j <= n; ++j)
        if (zw[j] < zw[j - 1] + hm) zw[j] = zw[j - 1] + hm;
}

__device__ __forceinline__ double transition_weight(double D, double d_sigma, double d_hybrid) {
    if (d_hybrid <= d_sigma) return D > d_sigma ? 1.0 : 0.0;
    return pmin(pmax((D - d_sigma) / (d_hybrid - d_sigma), 0.0), 1.0);
}

struct HybridArgs {
    const double *H, *eta, *h, *rho, *dz_top, *dsigma, *rho_target, *max_h_top;
    const uint8_t* active;
    double d_sigma, d_hybrid, h_min;
    int keep, n;
    long long M;
    double* h_new;
    uint8_t* wet;
    double* scratch;              // (6 n + 3) levels x M
    unsigned long long* bad;
};

__global__ void k_regrid_hybrid(HybridArgs r) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= r.M) return;
    const long long M = r.M;
    const int n = r.n, m = n;
    double Hc = r.H[c];
    Hc = Hc != Hc ? 0.0 : (Hc == __builtin_inf() ? DBL_MAX : (Hc == -__builtin_inf() ? -DBL_MAX : Hc));
    const double et = r.eta[c];
    const Col hb{r.h + c, M}, rb{r.rho + c, M};
    double tot = 0.0;
    for (int k = 0; k < n; ++k) tot += hb[k];
    const bool wet = r.active[c] && (Hc + et > 0.0) && (tot > 0.0);
    if (r.wet) r.wet[c] = wet ? 1 : 0;
    auto keep_h = [&]() {
        for (int k = 0; k < n; ++k) r.h_new[(size_t)k * M + c] = hb[k];
    };
    if (!wet) {
        keep_h();
        return;
    }
    const double D = Hc + et;
    double q = r.d_hybrid > 0.0 ? transition_weight(D, r.d_sigma, r.d_hybrid) : 1.0;
    if (r.d_hybrid > 0.0 && Hc <= 0.0) q = 0.0;
    bool rho_ok = true;                                       // every wet column's densities must be finite
    for (int k = 0; k < n; ++k) rho_ok = rho_ok && fabs(rb[k]) <= DBL_MAX;
    if (!(fabs(D) <= DBL_MAX) || (r.d_hybrid <= 0.0 && !(Hc > 0.0)) || !rho_ok) {
        atomicMin(r.bad, (unsigned long long)c);
        keep_h();
        return;
    }
    double* base = r.scratch + c;
    const WCol rho{base, M};                                  // n: the densities the rule follows (surface first)
    const WCol e{base + (size_t)n * M, M};                    // 2n: P1M edges (left, right) per layer
    const WCol zw{base + (size_t)3 * n * M, M};               // n + 1: interfaces (first the positions x)
    const WCol sig{base + (size_t)(4 * n + 1) * M, M};        // n + 1
    const WCol zs{base + (size_t)(5 * n + 2) * M, M};         // n + 1
    if (q <= 0.0) {
        sigma_column_w(Hc, et, r.dsigma, n, r.h_min, zw);
    } else {
        // layer densities, surface first: massless layers take the density of the nearest layer with mass above
        // (below for a top run without mass), then the running minimum from the bed up
        const double h_mass = 2.0 * r.h_min;
        auto h_top = [&](int k) -> double { return hb[n - 1 - k]; };
        auto rho_own = [&](int k) -> double { return rb[n - 1 - k]; };
        double last = -1.0;
        for (int k = 0; k < n; ++k) {
            const double rk = rho_own(k);
            rho[k] = rk;
            if (h_top(k) >= h_mass)
                last = rk;
            else if (last >= 0.0)
                rho[k] = last;
        }
        int k0 = 0;
        for (k0 = 0; k0 < n; ++k0)
            if (h_top(k0) >= h_mass) break;
        if (k0 == n) k0 = n - 1;
        if (k0 > 0 && h_top(k0) >= h_mass)
            for (int kk = 0; kk < k0; ++kk) rho[kk] = rho[k0];
        for (int kk = n - 2; kk >= 0; --kk) rho[kk] = pmin(rho[kk], rho[kk + 1]);
        // P1M with H2 edges: edges from the thickness-weighted neighbours, bounded, averaged where discontinuous
        e[0] = rho[0];
        e[2 * (n - 1) + 1] = rho[n - 1];
        for (int k = 1; k < n; ++k) {
            const double ha = h_top(k - 1), hc = h_top(k);
            double v;
            if (ha + hc == 0.0)
                v = 0.5 * (rho[k - 1] + rho[k]);
            else
                v = (rho[k - 1] * hc + rho[k] * ha) / (ha + hc);
            e[2 * k] = v;
            e[2 * (k - 1) + 1] = v;
        }
        for (int k = 0; k < n; ++k) {
            const int km1 = k - 1 > 0 ? k - 1 : 0, kp1 = k + 1 < n - 1 ? k + 1 : n - 1;
            const double s = bound_slope(h_top(km1), h_top(k), h_top(kp1), rho[km1], rho[k], rho[kp1]);
            e[2 * k] = bound_edge(e[2 * k], s, rho[km1], rho[k], false);
            e[2 * k + 1] = bound_edge(e[2 * k + 1], s, rho[kp1], rho[k], true);
        }
        for (int k = 0; k < n - 1; ++k) {
            if (e[2 * k + 1] != e[2 * (k + 1)]) {
                const double a = 0.5 * (e[2 * k + 1] + e[2 * (k + 1)]);
                e[2 * k + 1] = a;
                e[2 * (k + 1)] = a;
            }
        }
        // positions x of the target densities, depths from the surface (x[k] in zw[k])
        double xgn = 0.0;
        for (int k = 0; k < n; ++k) xgn += h_top(k);
        {
            double xg = 0.0;
            for (int k = 0; k <= m; ++k) {
                zw[k] = xg;                                   // m == n: the input grid
                if (k < n) xg += h_top(k);
            }
        }
        // The edges E_0 .. E_n (E_k = e[2k], E_n = e[2n-1]; the averaging made e[2k-1] = e[2k]) of a column are
        // monotone after the running minimum (each averaged edge lies between its two layers' densities).  Then
        // the first match of each of the reference's scans is found by bisection: t equal to an inner edge, or
        // strictly inside one layer.  A column whose edges are not monotone takes the scans themselves.
        auto E = [&](int k) -> double { return k < n ? e[2 * k] : e[2 * n - 1]; };
        bool mono = true;
        for (int k = 0; k < n; ++k) mono = mono && E(k) <= E(k + 1);
        if (mono) {
            const WCol xgv = sig;                             // the input grid's positions (sig is rebuilt later)
            {
                double xg = 0.0;
                for (int k = 0; k <= n; ++k) {
                    xgv[k] = xg;
                    if (k < n) xg += h_top(k);
                }
            }
            const double E0 = E(0), En = E(n);
            for (int kk = 1; kk < m; ++kk) {
                const double t = r.rho_target[kk];
                if (!(t == t)) continue;                      // NaN matches no case: the input grid stays
                if (t <= E0) {
                    zw[kk] = 0.0;
                    continue;
                }
                int lo = 1, hi = n;                           // the first inner edge >= t (n: none)
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (E(mid) < t)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                if (lo < n && E(lo) == t) {
                    zw[kk] = xgv[lo];
                    continue;
                }
                if (t >= En) {
                    zw[kk] = xgn;
                    continue;
                }
                const int k = lo - 1;                         // E_k < t < E_k+1
                const double el = E(k), er = E(k + 1);
                const double xi = (t - el) / (er - el);
                zw[kk] = xgv[k] + pmin(pmax(xi, 0.0), 1.0) * h_top(k);
            }
        } else {
            for (int kk = 1; kk < m; ++kk) {
                const double t = r.rho_target[kk];
                if (t <= e[0]) {
                    zw[kk] = 0.0;
                    continue;
                }
                bool found = false;
                double xg = h_top(0);
                for (int k = 1; k < n; ++k) {
                    if (t >= e[2 * (k - 1) + 1] && t <= e[2 * k]) {
                        zw[kk] = xg;
                        found = true;
                        break;
                    }
                    xg += h_top(k);
                }
                if (found) continue;
                if (t >= e[2 * (n - 1) + 1]) {
                    zw[kk] = xgn;
                    continue;
                }
                xg = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double el = e[2 * k], er = e[2 * k + 1];
                    if (el < t && t < er) {
                        const double xi = (t - el) / (er - el);
                        zw[kk] = xg + pmin(pmax(xi, 0.0), 1.0) * h_top(k);
                        break;
                    }
                    xg += h_top(k);
                }
            }
        }
        if (r.keep) {
            // isopycnal preservation: interfaces between consistent layers, or straddling their own target, stay
            auto ok = [&](int j) -> bool {
                const double rj = rho_own(j);
                return h_top(j) < h_mass || (r.rho_target[j] <= rj && rj <= r.rho_target[j + 1]);
            };
            auto kept = [&](int j) -> bool {
                if (j == 0 || j == n) return true;
                const bool straddle = h_top(j - 1) >= h_mass && h_top(j) >= h_mass &&
                                      rho_own(j - 1) <= r.rho_target[j] && r.rho_target[j] <= rho_own(j);
                return (ok(j - 1) && ok(j)) || straddle;
            };
            // the kept interfaces once, one bit each, for columns of up to 63 layers (each layer's test reads its
            // thickness, density and targets, and the three passes below ask for every interface)
            const bool bits = n <= 63;
            unsigned long long kb = 0ull;
            if (bits) {
                bool ok_prev = ok(0);
                for (int j = 1; j < n; ++j) {
                    const bool ok_j = ok(j);
                    const bool straddle = h_top(j - 1) >= h_mass && h_top(j) >= h_mass &&
                                          rho_own(j - 1) <= r.rho_target[j] && r.rho_target[j] <= rho_own(j);
                    if ((ok_prev && ok_j) || straddle) kb |= 1ull << j;
                    ok_prev = ok_j;
                }
            }
            auto K = [&](int j) -> bool { return bits ? ((kb >> j) & 1ull) != 0ull : kept(j); };
            double xg = 0.0;
            for (int j = 1; j < n; ++j) {
                xg += h_top(j - 1);
                if (K(j)) zw[j] = xg;
            }
            double lo = 0.0;
            for (int j = 1; j < n; ++j) {
                if (K(j))
                    lo = zw[j];
                else if (zw[j] < lo)
                    zw[j] = lo;
            }
            double hi = xgn;
            for (int j = n - 1; j > 0; --j) {
                if (K(j))
                    hi = zw[j];
                else if (zw[j] > hi)
                    zw[j] = hi;
            }
        }
        // z* floor, maximum thickness, order, then the interfaces bed first
        const double s = D / Hc;
        double nominal = 0.0;
        for (int kk = 1; kk <= m; ++kk) {
            nominal += r.dz_top[kk - 1] * s;
            zw[kk] = pmin(pmax(zw[kk], nominal), D);
        }
        for (int kk = 1; kk < m; ++kk) zw[kk] = pmin(zw[kk], zw[kk - 1] + r.max_h_top[kk - 1]);
        for (int kk = 1; kk <= m; ++kk) zw[kk] = pmax(zw[kk], zw[kk - 1]);
        zw[m] = D;
        // zw[j] = eta - x[m - j]: reverse in place
        for (int j = 0; j < (m + 1) / 2; ++j) {
            const double a = zw[j], b = zw[m - j];
            zw[j] = et - b;
            zw[m - j] = et - a;
        }
        if ((m + 1) % 2 == 1) zw[m / 2] = et - zw[m / 2];
        zw[0] = -Hc;
        zw[m] = et;
        enforce_min_thickness_w(zw, m, r.h_min);
        if (q < 1.0) {
            sigma_column_w(Hc, et, r.dsigma, n, r.h_min, sig);
            zstar_column_w(Hc, et, r.dz_top, n, r.h_min, zs);
            for (int j = 0; j <= n; ++j) {
                const double lo = q * (-Hc) + (1.0 - q) * sig[j];
                const double hi = q * zs[j] + (1.0 - q) * sig[j];
                zw[j] = pmin(pmax(zw[j], lo), hi);
            }
            zw[0] = -Hc;
            zw[n] = sig[n];
            enforce_min_thickness_w(zw, n, r.h_min);
        }
    }
    for (int k = 0; k < n; ++k) r.h_new[(size_t)k * M + c] = zw[k + 1] - zw[k];
}

// ---------------------------------------------------------------------------------------- layer density
// The densities the hybrid rule follows (ALECoordinate.regrid_density, core/components/vertical_coordinate.py:80-100):
// the linear EOS rho0 - t_coef (T - t_ref) [+ s_coef (S - s_ref)] (core/components/eos.py:62-69), or Wright's (1997)
// full-range density at sea pressure p (physics/eos/wright_full.py:80-89; the MOM6 WRIGHT_FULL parenthesization and
// the coefficients of Wright's Table 1, second column).  Elementwise over n cells, every operation rounded on its own.
constexpr int EOS_LINEAR = 0, EOS_WRIGHT_FULL = 1;

struct DensityArgs {
    const double* T;
    const double* S;     // null: the linear EOS without salinity
    const double* p;     // null: the scalar pressure ps for every cell
    double ps;
    double lin[5];       // rho0, t_coef, t_ref, s_coef, s_ref
    int kind;
    long long n;
    double* rho;
};

__device__ __forceinline__ double wright_density(double T, double S, double p) {
    constexpr double a0 = 7.133718e-4, a1 = 2.724670e-7, a2 = -1.646582e-7;
    constexpr double b0 = 5.613770e8, b1 = 3.600337e6, b2 = -3.727194e4, b3 = 1.660557e2, b4 = 6.844158e5,
                     b5 = -8.389457e3;
    constexpr double c0 = 1.609893e5, c1 = 8.427815e2, c2 = -6.931554, c3 = 3.869318e-2, c4 = -1.664201e2,
                     c5 = -2.765195;
    const double al0 = a0 + (a1 * T + a2 * S);
    const double p0 = b0 + (b4 * S + T * (b1 + (T * (b2 + b3 * T) + b5 * S)));
    const double lam = c0 + (c4 * S + T * (c1 + (T * (c2 + c3 * T) + c5 * S)));
    return (p + p0) / (lam + al0 * (p + p0));
}

__global__ void k_regrid_density(DensityArgs a) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < a.n; c += stride) {
        const double T = a.T[c];
        double r;
        if (a.kind == EOS_WRIGHT_FULL) {
            r = wright_density(T, a.S[c], a.p ? a.p[c] : a.ps);
        } else {
            r = a.lin[0] - a.lin[1] * (T - a.lin[2]);
            if (a.S) r = r + a.lin[3] * (a.S[c] - a.lin[4]);
        }
        a.rho[c] = r;
    }
}

}  // namespace rmp
}  // namespace ocx
