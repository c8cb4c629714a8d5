/*
 * C loop restatement of pyOCEAN's vertical regrid (z*, sigma) and conservative remap (PCM, PLM, PPM_H4) of the ALE
 * coordinate (test infrastructure only; see oracle/__init__.py).  Built with -ffp-contract=off, every product and
 * sum rounded on its own as Numba (no fastmath) rounds them, and with the operand order of the reference:
 *
 *   remap       core/vertical/remap.py:42-427 (_effective_scheme, _plm_slope_wa, _plm_monotonized_slope,
 *               _end_values_h4, _edges_h4, _bound_edges, _check_discontinuous, reconstruct, _average,
 *               remap_column, _remap_columns)
 *   regrid      core/vertical/regrid.py:41-95 (zstar_column, sigma_column, enforce_min_thickness), the z* / sigma
 *               branches of _regrid_columns (:325-357) and ALECoordinate.regrid / regrid_remap
 *               (core/components/vertical_coordinate.py:69-109)
 *
 * Python's min / max of two floats return the first argument unless the second compares smaller / larger; the
 * helpers below keep that rule (only NaN could tell it apart).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef long long i64;

enum { PCM = 0, PLM = 1, PPM_H4 = 2 };
#define H_NEGLECT 1e-30
#define H_NEGLECT_EDGE 1e-10
#define H_MIN_FRAC 1e-5
#define EPS DBL_EPSILON

static inline double pmin(double a, double b) { return b < a ? b : a; }
static inline double pmax(double a, double b) { return b > a ? b : a; }
static inline double pmin3(double a, double b, double c) { return pmin(pmin(a, b), c); }
static inline double pmax3(double a, double b, double c) { return pmax(pmax(a, b), c); }

static int effective_scheme(int n_active, int scheme) {
    if (n_active <= 1) return PCM;
    if (n_active <= 3 && scheme > PLM) return PLM;
    return scheme;
}

static double plm_slope_wa(double h_l, double h_c, double h_r, double h_neglect, double u_l, double u_c, double u_r) {
    const double sigma_r = u_r - u_c;
    const double sigma_l = u_c - u_l;
    const double sigma_c = 2.0 * (u_r - u_l) * (h_c / (h_l + 2.0 * h_c + h_r + h_neglect));
    const double u_min = pmin3(u_l, u_c, u_r);
    const double u_max = pmax3(u_l, u_c, u_r);
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

static double plm_monotonized_slope(double u_l, double u_c, double u_r, double s_l, double s_c, double s_r) {
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

/* (C(0), C(dz[0])) of the cubic through the means of the 4 end cells, closed form */
static void end_values_h4(const double* dz, const double* u, double* e_out, double* e_in) {
    const double min_frac = 1.0e-6;
    const double h1 = dz[0], h2 = dz[1];
    double h3 = dz[2], h4 = dz[3];
    if ((h2 + h3) < min_frac * h1) h3 = min_frac * h1 - h2;
    if ((h3 + h4) < min_frac * h1) h4 = min_frac * h1 - h3;
    const double h12 = h1 + h2, h23 = h2 + h3, h34 = h3 + h4;
    const double h123 = h12 + h3, h234 = h2 + h34, h1234 = h12 + h34;
    const double I_denB3 = 1.0 / (h123 * h12 * h23);
    const double I_h12 = (h123 * h23) * I_denB3;
    const double I_h23 = (h12 * h123) * I_denB3;
    const double I_h123 = (h12 * h23) * I_denB3;
    const double I_denom = 1.0 / (h1234 * (h234 * h34));
    const double I_h234 = (h1234 * h34) * I_denom;
    const double I_h1234 = (h234 * h34) * I_denom;
    const double w11 = -h1 * (I_h1234 + I_h123 + I_h12);
    const double w21 = h1 * h12 * (I_h234 * I_h1234 + I_h23 * (I_h234 + I_h123));
    const double w31 = -h1 * h12 * h123 * I_denom;
    const double w12 = 2.0 * (I_h12 * (1.0 + (h1 + h12) * (I_h1234 + I_h123)) + h1 * I_h1234 * I_h123);
    const double w22 = -2.0 * ((h1 * h12 * I_h1234) * (I_h23 * (I_h234 + I_h123)) +
                               (h1 + h12) * (I_h1234 * I_h234 + I_h23 * (I_h234 + I_h123)));
    const double w32 = 2.0 * ((h1 + h12) * h123 + h1 * h12) * I_denom;
    const double w13 = -3.0 * I_h12 * I_h123 * (1.0 + I_h1234 * ((h1 + h12) + h123));
    const double w23 = 3.0 * I_h23 * (I_h123 + I_h1234 * ((h1 + h12) + h123) * (I_h123 + I_h234));
    const double w33 = -3.0 * ((h1 + h12) + h123) * I_denom;
    const double w14 = 4.0 * I_h1234 * I_h123 * I_h12;
    const double w24 = -4.0 * I_h1234 * (I_h23 * (I_h123 + I_h234));
    const double w34 = 4.0 * I_denom;
    const double d1 = u[1] - u[0], d2 = u[2] - u[1], d3 = u[3] - u[2];
    const double c0 = ((u[0] + (w11 * d1)) + (w21 * d2)) + (w31 * d3);
    const double c1 = ((w12 * d1) + (w22 * d2)) + (w32 * d3);
    const double c2 = ((w13 * d1) + (w23 * d2)) + (w33 * d3);
    const double c3 = ((w14 * d1) + (w24 * d2)) + (w34 * d3);
    const double d = dz[0];
    *e_out = c0;
    *e_in = c0 + d * (c1 + d * (c2 + d * c3));
}

/* e[2k] = left edge of cell k, e[2k+1] = right edge */
static void edges_h4(const double* h, const double* u, int n, double h_neglect, double* e) {
    for (int i = 2; i < n - 1; ++i) {
        double h0 = h[i - 2], h1 = h[i - 1], h2 = h[i], h3 = h[i + 1];
        if (h0 + h1 == 0.0 || h1 + h2 == 0.0 || h2 + h3 == 0.0) {
            const double h_min = H_MIN_FRAC * pmax(h_neglect, (h0 + h1) + (h2 + h3));
            h0 = pmax(h_min, h0);
            h1 = pmax(h_min, h1);
            h2 = pmax(h_min, h2);
            h3 = pmax(h_min, h3);
        }
        const double I_h12 = 1.0 / (h1 + h2);
        const double I_den_et2 = 1.0 / (((h0 + h1) + h2) * (h0 + h1));
        const double I_h012 = (h0 + h1) * I_den_et2;
        const double I_den_et3 = 1.0 / ((h1 + (h2 + h3)) * (h2 + h3));
        const double I_h123 = (h2 + h3) * I_den_et3;
        const double et1 = (1.0 + (h1 * I_h012 + (h0 + h1) * I_h123)) * I_h12 * (h2 * (h2 + h3)) * u[i - 1] +
                           (1.0 + (h2 * I_h123 + (h2 + h3) * I_h012)) * I_h12 * (h1 * (h0 + h1)) * u[i];
        const double et2 = (h1 * (h2 * (h2 + h3)) * I_den_et2) * (u[i - 1] - u[i - 2]);
        const double et3 = (h2 * (h1 * (h0 + h1)) * I_den_et3) * (u[i] - u[i + 1]);
        e[2 * i] = (et1 + (et2 + et3)) / ((h0 + h1) + (h2 + h3));
        e[2 * (i - 1) + 1] = e[2 * i];
    }
    double dz[4], ut[4];
    for (int i = 0; i < 4; ++i) {
        dz[i] = pmax(h_neglect, h[i]);
        ut[i] = u[i];
    }
    end_values_h4(dz, ut, &e[0], &e[1]);
    e[2] = e[1];
    for (int i = 0; i < 4; ++i) {
        dz[i] = pmax(h_neglect, h[n - 1 - i]);
        ut[i] = u[n - 1 - i];
    }
    end_values_h4(dz, ut, &e[2 * (n - 1) + 1], &e[2 * (n - 1)]);
    e[2 * (n - 2) + 1] = e[2 * (n - 1)];
}

static void bound_edges(const double* h, const double* u, int n, double* e) {
    for (int k = 0; k < n; ++k) {
        const int km1 = k - 1 > 0 ? k - 1 : 0;
        const int kp1 = k + 1 < n - 1 ? k + 1 : n - 1;
        double slope_x_h = 0.0;
        if ((h[km1] + h[kp1]) + 2.0 * h[k] > 0.0) {
            const double sl = u[k] - u[km1];
            const double sc = (u[kp1] - u[km1]) * (h[k] / ((h[km1] + h[kp1]) + 2.0 * h[k]));
            const double sr = u[kp1] - u[k];
            if (sl * sr > 0.0) {
                const double m = pmin3(fabs(sl), fabs(sc), fabs(sr));
                slope_x_h = sc >= 0.0 ? m : -m;
            }
        }
        const double sgn = slope_x_h >= 0.0 ? 1.0 : -1.0;
        double* el = &e[2 * k];
        double* er = &e[2 * k + 1];
        if ((u[km1] - *el) * (*el - u[k]) < 0.0) *el = u[k] - sgn * pmin(fabs(slope_x_h), fabs(*el - u[k]));
        if ((u[kp1] - *er) * (*er - u[k]) < 0.0) *er = u[k] + sgn * pmin(fabs(slope_x_h), fabs(*er - u[k]));
        *el = pmax(pmin(*el, pmax(u[km1], u[k])), pmin(u[km1], u[k]));
        *er = pmax(pmin(*er, pmax(u[kp1], u[k])), pmin(u[kp1], u[k]));
    }
}

static void check_discontinuous(const double* u, int n, double* e) {
    for (int k = 0; k < n - 1; ++k) {
        if ((e[2 * (k + 1)] - e[2 * k + 1]) * (u[k + 1] - u[k]) < 0.0) {
            double a = 0.5 * (e[2 * k + 1] + e[2 * (k + 1)]);
            a = pmax(pmin(a, pmax(u[k], u[k + 1])), pmin(u[k], u[k + 1]));
            e[2 * k + 1] = a;
            e[2 * (k + 1)] = a;
        }
    }
}

/* coef[3k .. 3k+2] = (c0, c1, c2) of p(xi) = c0 + c1 xi + c2 xi^2 in cell k */
void oc_reconstruct(const double* h, const double* u, int n, int scheme, double* coef) {
    for (int k = 0; k < n; ++k) {
        coef[3 * k] = u[k];
        coef[3 * k + 1] = 0.0;
        coef[3 * k + 2] = 0.0;
    }
    if (scheme == PCM || n < 3) return;
    if (scheme == PLM) {
        double* slp = calloc((size_t)n, sizeof(double));
        double* mslp = calloc((size_t)n, sizeof(double));
        for (int k = 1; k < n - 1; ++k)
            slp[k] = plm_slope_wa(h[k - 1], h[k], h[k + 1], H_NEGLECT, u[k - 1], u[k], u[k + 1]);
        for (int k = 1; k < n - 1; ++k)
            mslp[k] = plm_monotonized_slope(u[k - 1], u[k], u[k + 1], slp[k - 1], slp[k], slp[k + 1]);
        const double almost_one = 1.0 - EPS;
        for (int k = 1; k < n - 1; ++k) {
            const double u_l = u[k] - 0.5 * mslp[k];
            const double u_r = u[k] + 0.5 * mslp[k];
            coef[3 * k] = u_l;
            coef[3 * k + 1] = u_r - u_l;
            const double edge = coef[3 * k + 1] + coef[3 * k];
            const double e_r = u[k + 1] - 0.5 * (slp[k + 1] >= 0.0 ? fabs(mslp[k + 1]) : -fabs(mslp[k + 1]));
            if ((edge - u[k]) * (e_r - edge) < 0.0) coef[3 * k + 1] = coef[3 * k + 1] * almost_one;
        }
        free(slp);
        free(mslp);
        return;
    }
    if (scheme == PPM_H4 && n >= 4) {
        double* e = malloc(sizeof(double) * 2 * (size_t)n);
        edges_h4(h, u, n, H_NEGLECT_EDGE, e);
        bound_edges(h, u, n, e);
        check_discontinuous(u, n, e);
        for (int k = 1; k < n - 1; ++k) {
            const double u_l = u[k - 1], u_c = u[k], u_r = u[k + 1];
            double el = e[2 * k], er = e[2 * k + 1];
            if ((u_r - u_c) * (u_c - u_l) <= 0.0) {
                el = u_c;
                er = u_c;
            } else {
                const double expr1 = 3.0 * (er - el) * ((u_c - el) + (u_c - er));
                const double expr2 = (er - el) * (er - el);
                if (expr1 > expr2) {
                    el = u_c + 2.0 * (u_c - er);
                    el = pmax(pmin(el, pmax(u_l, u_c)), pmin(u_l, u_c));
                } else if (expr1 < -expr2) {
                    er = u_c + 2.0 * (u_c - el);
                    er = pmax(pmin(er, pmax(u_r, u_c)), pmin(u_r, u_c));
                }
            }
            if (fabs(er - el) < pmax(1e-60, EPS * fabs(u_c))) {
                el = u_c;
                er = u_c;
            }
            coef[3 * k] = el;
            coef[3 * k + 1] = 4.0 * (u_c - el) + 2.0 * (u_c - er);
            coef[3 * k + 2] = 3.0 * ((er - u_c) + (el - u_c));
        }
        free(e);
    }
}

static double average(double c0, double c1, double c2, double xa, double xb) {
    if (xb - xa <= 1e-14) {
        const double x = 0.5 * (xa + xb);
        return c0 + x * (c1 + x * c2);
    }
    return c0 + 0.5 * c1 * (xa + xb) + c2 * (xa * xa + xa * xb + xb * xb) / 3.0;
}

void oc_remap_column(const double* h_src, const double* u_src, int n0, const double* h_tgt, int n1, int scheme,
                     double* u_tgt) {
    int n_active = 0;
    for (int k = 0; k < n0; ++k)
        if (h_src[k] > 0.0) ++n_active;
    double* coef = malloc(sizeof(double) * 3 * (size_t)n0);
    oc_reconstruct(h_src, u_src, n0, effective_scheme(n_active, scheme), coef);
    double* t_edge = malloc(sizeof(double) * (size_t)(n1 + 1));
    t_edge[0] = 0.0;
    for (int j = 0; j < n1; ++j) t_edge[j + 1] = t_edge[j] + h_tgt[j];
    double* content = calloc((size_t)(n1 > 0 ? n1 : 1), sizeof(double));
    double* piece = malloc(sizeof(double) * (size_t)(n1 + 1));
    int* piece_j = malloc(sizeof(int) * (size_t)(n1 + 1));
    double a = 0.0;
    int j = 0, last = -1;
    for (int i = 0; i < n0; ++i) {
        const double hi = h_src[i];
        if (hi <= 0.0) continue;
        last = i;
        const double b = a + hi;
        int npc = 0, ithick = 0;
        double thick = -1.0;
        while (j < n1) {
            const double lo = pmax(a, t_edge[j]);
            const double hi_t = j < n1 - 1 ? t_edge[j + 1] : pmax(t_edge[n1], b);
            const double up = pmin(b, hi_t);
            const double L = up - lo;
            if (L > 0.0) {
                piece[npc] = L * average(coef[3 * i], coef[3 * i + 1], coef[3 * i + 2], (lo - a) / hi, (up - a) / hi);
                piece_j[npc] = j;
                if (L > thick) {
                    thick = L;
                    ithick = npc;
                }
                ++npc;
            }
            if (hi_t <= b && j < n1 - 1)
                ++j;
            else
                break;
        }
        if (npc > 0) {
            double s = 0.0;
            for (int p = 0; p < npc; ++p)
                if (p != ithick) s += piece[p];
            piece[ithick] = u_src[i] * hi - s;
            for (int p = 0; p < npc; ++p) content[piece_j[p]] += piece[p];
        }
        a = b;
    }
    if (last < 0) {
        double mean = 0.0;
        for (int i = 0; i < n0; ++i) mean += u_src[i] / n0;
        for (int jj = 0; jj < n1; ++jj) u_tgt[jj] = n0 == n1 ? u_src[jj] : mean;
    } else {
        for (int jj = 0; jj < n1; ++jj) {
            if (h_tgt[jj] > 0.0) {
                u_tgt[jj] = content[jj] / h_tgt[jj];
            } else {
                const double x = t_edge[jj];
                double aa = 0.0, val = u_src[last];
                for (int i = 0; i < n0; ++i) {
                    if (h_src[i] <= 0.0) continue;
                    if (x <= aa + h_src[i]) {
                        const double xi = (x - aa) / h_src[i];
                        val = coef[3 * i] + xi * (coef[3 * i + 1] + xi * coef[3 * i + 2]);
                        break;
                    }
                    aa += h_src[i];
                }
                u_tgt[jj] = val;
            }
        }
    }
    free(coef);
    free(t_edge);
    free(content);
    free(piece);
    free(piece_j);
}

/* every wet column of the fields f (nf, n0, M) from h_src (n0, M) to h_tgt (n1, M); dry columns: f copied when
 * n0 == n1, else NaN */
void oc_remap_columns(const double* h_src, const double* f, int n0, const double* h_tgt, int n1, i64 M, int nf,
                      int scheme, const uint8_t* wet, double* out) {
#pragma omp parallel for schedule(dynamic, 256)
    for (i64 c = 0; c < M * nf; ++c) {
        const i64 col = c % M;
        const int q = (int)(c / M);
        const double* fq = f + (size_t)q * n0 * M;
        double* oq = out + (size_t)q * n1 * M;
        if (!wet[col]) {
            for (int k = 0; k < n1; ++k) oq[(size_t)k * M + col] = n0 == n1 ? fq[(size_t)k * M + col] : NAN;
            continue;
        }
        double hs[n0], us[n0], ht[n1], ut[n1];
        for (int k = 0; k < n0; ++k) {
            hs[k] = h_src[(size_t)k * M + col];
            us[k] = fq[(size_t)k * M + col];
        }
        for (int k = 0; k < n1; ++k) ht[k] = h_tgt[(size_t)k * M + col];
        oc_remap_column(hs, us, n0, ht, n1, scheme, ut);
        for (int k = 0; k < n1; ++k) oq[(size_t)k * M + col] = ut[k];
    }
}

/* ------------------------------------------------------------------------------------------------ regrid */
void oc_zstar_column(double H, double eta, const double* dz_top, int n, double h_min, double* zw) {
    const double D = H + eta;
    const double hm = pmin(h_min, D / n);
    const double s = D / H;
    double e[n + 1];
    e[0] = eta;
    for (int k = 0; k < n; ++k) e[k + 1] = e[k] - s * dz_top[k];
    e[n] = -H;
    for (int k = n - 1; k >= 0; --k)
        if (e[k] < e[k + 1] + hm) e[k] = e[k + 1] + hm;
    for (int j = 0; j <= n; ++j) zw[j] = e[n - j];
}

static void enforce_min_thickness(double* zw, int n, double h_min) {
    const double hm = pmin(h_min, (zw[n] - zw[0]) / n);
    for (int j = n - 1; j > 0; --j)
        if (zw[j] > zw[j + 1] - hm) zw[j] = zw[j + 1] - hm;
    for (int j = 1; j < n; ++j)
        if (zw[j] < zw[j - 1] + hm) zw[j] = zw[j - 1] + hm;
}

void oc_sigma_column(double H, double eta, const double* dsigma, int n, double h_min, double* zw) {
    const double D = H + eta;
    zw[0] = -H;
    for (int j = 0; j < n; ++j) zw[j + 1] = zw[j] + D * dsigma[j];
    zw[n] = eta;
    enforce_min_thickness(zw, n, h_min);
}

/* ALECoordinate.regrid for the z* (rule 0) and sigma (rule 1) rules: wet = active & H + eta > 0 & sum h > 0
 * (H with NaN read as 0); h_new = diff of the rule's interfaces where wet, else h.  wet_out (M) optional. */
void oc_regrid(const double* H, const double* eta, const uint8_t* active, const double* h, int n, i64 M, int rule,
               const double* dz, double h_min, double* h_new, uint8_t* wet_out) {
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < M; ++c) {
        const double Hc = H[c] == H[c] ? (H[c] == INFINITY ? DBL_MAX : (H[c] == -INFINITY ? -DBL_MAX : H[c])) : 0.0;
        double tot = 0.0;
        for (int k = 0; k < n; ++k) tot += h[(size_t)k * M + c];
        const int wet = active[c] && (Hc + eta[c] > 0.0) && (tot > 0.0);
        if (wet_out) wet_out[c] = (uint8_t)wet;
        if (!wet) {
            for (int k = 0; k < n; ++k) h_new[(size_t)k * M + c] = h[(size_t)k * M + c];
            continue;
        }
        double zw[n + 1];
        if (rule == 0)
            oc_zstar_column(Hc, eta[c], dz, n, h_min, zw);
        else
            oc_sigma_column(Hc, eta[c], dz, n, h_min, zw);
        for (int k = 0; k < n; ++k) h_new[(size_t)k * M + c] = zw[k + 1] - zw[k];
    }
}

/* ------------------------------------------------------------------------------------------------ hybrid rule */
/* P1M_H2 positions (depths from the surface, 0 .. D) of the target interface densities (surface first, m+1) in the
 * source column h_top, rho_top (surface first, n layers): _p1m_h2_positions (core/vertical/regrid.py:98-151) */
static void p1m_h2_positions(const double* h_top, const double* rho_top, int n, const double* targets, int m,
                             double* x_new) {
    double* e = malloc(sizeof(double) * 2 * (size_t)n);
    double* xg = malloc(sizeof(double) * (size_t)(n + 1));
    e[0] = rho_top[0];
    e[2 * (n - 1) + 1] = rho_top[n - 1];
    for (int k = 1; k < n; ++k) {
        if (h_top[k - 1] + h_top[k] == 0.0)
            e[2 * k] = 0.5 * (rho_top[k - 1] + rho_top[k]);
        else
            e[2 * k] = (rho_top[k - 1] * h_top[k] + rho_top[k] * h_top[k - 1]) / (h_top[k - 1] + h_top[k]);
        e[2 * (k - 1) + 1] = e[2 * k];
    }
    bound_edges(h_top, rho_top, n, e);
    for (int k = 0; k < n - 1; ++k) {
        if (e[2 * k + 1] != e[2 * (k + 1)]) {
            const double a = 0.5 * (e[2 * k + 1] + e[2 * (k + 1)]);
            e[2 * k + 1] = a;
            e[2 * (k + 1)] = a;
        }
    }
    xg[0] = 0.0;
    for (int k = 0; k < n; ++k) xg[k + 1] = xg[k] + h_top[k];
    for (int k = 0; k <= m; ++k) x_new[k] = m == n ? xg[k] : xg[n] * k / m;
    for (int kk = 1; kk < m; ++kk) {
        const double t = targets[kk];
        if (t <= e[0]) {
            x_new[kk] = xg[0];
            continue;
        }
        int found = 0;
        for (int k = 1; k < n; ++k) {
            if (t >= e[2 * (k - 1) + 1] && t <= e[2 * k]) {
                x_new[kk] = xg[k];
                found = 1;
                break;
            }
        }
        if (found) continue;
        if (t >= e[2 * (n - 1) + 1]) {
            x_new[kk] = xg[n];
            continue;
        }
        for (int k = 0; k < n; ++k) {
            if (e[2 * k] < t && t < e[2 * k + 1]) {
                const double xi = (t - e[2 * k]) / (e[2 * k + 1] - e[2 * k]);
                x_new[kk] = xg[k] + pmin(pmax(xi, 0.0), 1.0) * h_top[k];
                break;
            }
        }
    }
    free(e);
    free(xg);
}

/* isopycnal layer preservation: _keep_consistent (core/vertical/regrid.py:157-203) */
static void keep_consistent(const double* h_top, const double* rho_top, const double* targets, int n, double h_mass,
                            double* x) {
    uint8_t* ok = malloc((size_t)n);
    uint8_t* kept = calloc((size_t)(n + 1), 1);
    double* xg = malloc(sizeof(double) * (size_t)(n + 1));
    for (int j = 0; j < n; ++j)
        ok[j] = h_top[j] < h_mass || (targets[j] <= rho_top[j] && rho_top[j] <= targets[j + 1]);
    xg[0] = 0.0;
    for (int j = 0; j < n; ++j) xg[j + 1] = xg[j] + h_top[j];
    kept[0] = 1;
    kept[n] = 1;
    for (int j = 1; j < n; ++j) {
        const int straddle = h_top[j - 1] >= h_mass && h_top[j] >= h_mass && rho_top[j - 1] <= targets[j] &&
                             targets[j] <= rho_top[j];
        if ((ok[j - 1] && ok[j]) || straddle) {
            x[j] = xg[j];
            kept[j] = 1;
        }
    }
    double lo = xg[0];
    for (int j = 1; j < n; ++j) {
        if (kept[j])
            lo = x[j];
        else if (x[j] < lo)
            x[j] = lo;
    }
    double hi = xg[n];
    for (int j = n - 1; j > 0; --j) {
        if (kept[j])
            hi = x[j];
        else if (x[j] > hi)
            x[j] = hi;
    }
    free(ok);
    free(kept);
    free(xg);
}

/* hybrid z*-isopycnal interfaces (bed first) of one column with n source layers onto n target layers:
 * hycom1_column (core/vertical/regrid.py:206-259) */
void oc_hycom1_column(double H, double eta, const double* h_bed, const double* rho_bed, int n,
                      const double* rho_target, const double* dz_top, const double* max_h_top, double h_min, int keep,
                      double* zw) {
    const int m = n;
    const double D = H + eta;
    double* h_top = malloc(sizeof(double) * (size_t)n);
    double* rho_own = malloc(sizeof(double) * (size_t)n);
    double* rho = malloc(sizeof(double) * (size_t)n);
    double* x = malloc(sizeof(double) * (size_t)(m + 1));
    for (int k = 0; k < n; ++k) {
        h_top[k] = h_bed[n - 1 - k];
        rho_own[k] = rho_bed[n - 1 - k];
        rho[k] = rho_own[k];
    }
    const double h_mass = 2.0 * h_min;
    double last = -1.0;
    for (int k = 0; k < n; ++k) {
        if (h_top[k] >= h_mass)
            last = rho[k];
        else if (last >= 0.0)
            rho[k] = last;
    }
    int k = 0;
    for (k = 0; k < n; ++k)
        if (h_top[k] >= h_mass) break;
    if (k == n) k = n - 1;                           /* a Python for loop leaves k at its last value */
    if (k > 0 && h_top[k] >= h_mass)
        for (int kk = 0; kk < k; ++kk) rho[kk] = rho[k];
    for (int kk = n - 2; kk >= 0; --kk) rho[kk] = pmin(rho[kk], rho[kk + 1]);
    p1m_h2_positions(h_top, rho, n, rho_target, m, x);
    if (keep) keep_consistent(h_top, rho_own, rho_target, n, 2.0 * h_min, x);
    const double s = D / H;
    double nominal = 0.0;
    for (int kk = 1; kk <= m; ++kk) {
        nominal += dz_top[kk - 1] * s;
        x[kk] = pmin(pmax(x[kk], nominal), D);
    }
    for (int kk = 1; kk < m; ++kk) x[kk] = pmin(x[kk], x[kk - 1] + max_h_top[kk - 1]);
    for (int kk = 1; kk <= m; ++kk) x[kk] = pmax(x[kk], x[kk - 1]);
    x[m] = D;
    for (int j = 0; j <= m; ++j) zw[j] = eta - x[m - j];
    zw[0] = -H;
    zw[m] = eta;
    enforce_min_thickness(zw, m, h_min);
    free(h_top);
    free(rho_own);
    free(rho);
    free(x);
}

static double transition_weight(double D, double d_sigma, double d_hybrid) {
    if (d_hybrid <= d_sigma) return D > d_sigma ? 1.0 : 0.0;
    return pmin(pmax((D - d_sigma) / (d_hybrid - d_sigma), 0.0), 1.0);
}

void oc_window_shallow_sigma(const double* zw_hyb, const double* zw_sig, const double* zw_zs, double H, double D,
                                 double d_sigma, double d_hybrid, double h_min, int n, double* zw) {
    const double q = transition_weight(D, d_sigma, d_hybrid);
    if (q >= 1.0) {
        for (int j = 0; j <= n; ++j) zw[j] = zw_hyb[j];
        return;
    }
    for (int j = 0; j <= n; ++j) {
        const double lo = q * (-H) + (1.0 - q) * zw_sig[j];
        const double hi = q * zw_zs[j] + (1.0 - q) * zw_sig[j];
        zw[j] = pmin(pmax(zw_hyb[j], lo), hi);
    }
    zw[0] = -H;
    zw[n] = zw_sig[n];
    enforce_min_thickness(zw, n, h_min);
}

/* ALECoordinate.regrid for the hybrid rule (_regrid_columns code 2, core/vertical/regrid.py:325-357): rho (n, M)
 * the layer densities; dz_top (z* floor, surface first), dsigma (bed first), rho_target (n+1, surface first),
 * max_h_top (n); d_sigma, d_hybrid the shallow window (d_hybrid 0: none); keep the isopycnal preservation. */
void oc_regrid_hybrid(const double* H, const double* eta, const uint8_t* active, const double* h, const double* rho,
                      int n, i64 M, const double* dz_top, const double* dsigma, const double* rho_target,
                      const double* max_h_top, double d_sigma, double d_hybrid, double h_min, int keep, double* h_new,
                      uint8_t* wet_out) {
#pragma omp parallel for schedule(dynamic, 64)
    for (i64 c = 0; c < M; ++c) {
        const double Hc = H[c] == H[c] ? (H[c] == INFINITY ? DBL_MAX : (H[c] == -INFINITY ? -DBL_MAX : H[c])) : 0.0;
        double tot = 0.0;
        for (int k = 0; k < n; ++k) tot += h[(size_t)k * M + c];
        const int wet = active[c] && (Hc + eta[c] > 0.0) && (tot > 0.0);
        if (wet_out) wet_out[c] = (uint8_t)wet;
        if (!wet) {
            for (int k = 0; k < n; ++k) h_new[(size_t)k * M + c] = h[(size_t)k * M + c];
            continue;
        }
        double zw[n + 1], hb[n], rb[n];
        double q = d_hybrid > 0.0 ? transition_weight(Hc + eta[c], d_sigma, d_hybrid) : 1.0;
        if (d_hybrid > 0.0 && Hc <= 0.0) q = 0.0;
        if (q <= 0.0) {
            oc_sigma_column(Hc, eta[c], dsigma, n, h_min, zw);
        } else {
            for (int k = 0; k < n; ++k) {
                hb[k] = h[(size_t)k * M + c];
                rb[k] = rho[(size_t)k * M + c];
            }
            oc_hycom1_column(Hc, eta[c], hb, rb, n, rho_target, dz_top, max_h_top, h_min, keep, zw);
            if (q < 1.0) {
                double sig[n + 1], zs[n + 1], hyb[n + 1];
                oc_sigma_column(Hc, eta[c], dsigma, n, h_min, sig);
                oc_zstar_column(Hc, eta[c], dz_top, n, h_min, zs);
                for (int j = 0; j <= n; ++j) hyb[j] = zw[j];
                oc_window_shallow_sigma(hyb, sig, zs, Hc, Hc + eta[c], d_sigma, d_hybrid, h_min, n, zw);
            }
        }
        for (int k = 0; k < n; ++k) h_new[(size_t)k * M + c] = zw[k + 1] - zw[k];
    }
}
