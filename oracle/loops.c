/*
 * C loop restatement of pyOCEAN's compiled tracer-advection kernels.  TEST INFRASTRUCTURE ONLY: the checker of
 * the HIP path and the timed CPU baseline; nothing in the product links it.
 *
 * Each function follows one Numba kernel of core/dynamics/sigma_baroclinic.py (file:line in its comment) with
 * the same operand order.  Build with -ffp-contract=off (no fused multiply-add) so that every product and sum is
 * rounded on its own, as Numba without fastmath does.  Parallel over (tracer, layer) planes with OpenMP, like
 * the reference's prange; planes are independent, so results do not depend on the thread count.
 *
 * Layout: stacks (nt, n, ny, nx), cells (n, ny, nx), Qx (n, ny, nx+1), Qy (n, ny+1, nx), Fz (n+1, ny, nx),
 * C-contiguous, k = 0 at the bed.  ok is one byte per cell (0 or 1).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* OpenMP team size of the loops below (the CPU baseline times them on all threads and on one) */
void oc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

#define OPEN_W 1
#define OPEN_E 2
#define OPEN_S 4
#define OPEN_N 8

typedef long long i64;

static inline double sq(double x) { return x * x; }
static inline double dmax(double a, double b) { return b > a ? b : a; }
static inline double dmin(double a, double b) { return b < a ? b : a; }

/* _wval, sigma_baroclinic.py:204-233 */
static inline double wval(double v1, double v2, double v3, double v4, double v5,
                          int k1, int k2, int k3, int k4, int k5, int z) {
    if (k1 && k2 && k3 && k4 && k5) {
        double q0 = (2.0 * v1 - 7.0 * v2 + 11.0 * v3) / 6.0;
        double q1 = (-v2 + 5.0 * v3 + 2.0 * v4) / 6.0;
        double q2 = (2.0 * v3 + 5.0 * v4 - v5) / 6.0;
        const double c1312 = 13.0 / 12.0;
        double b0 = c1312 * sq(v1 - 2.0 * v2 + v3) + 0.25 * sq(v1 - 4.0 * v2 + 3.0 * v3);
        double b1 = c1312 * sq(v2 - 2.0 * v3 + v4) + 0.25 * sq(v2 - v4);
        double b2 = c1312 * sq(v3 - 2.0 * v4 + v5) + 0.25 * sq(3.0 * v3 - 4.0 * v4 + v5);
        double a0, a1, a2;
        if (z) {
            double tau = fabs(b2 - b0);
            a0 = 0.1 * (1.0 + tau / (1e-40 + b0));
            a1 = 0.6 * (1.0 + tau / (1e-40 + b1));
            a2 = 0.3 * (1.0 + tau / (1e-40 + b2));
        } else {
            a0 = 0.1 / sq(1e-6 + b0);
            a1 = 0.6 / sq(1e-6 + b1);
            a2 = 0.3 / sq(1e-6 + b2);
        }
        return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
    }
    if (k2 && k3 && k4) {
        double c0 = (-v2 + 3.0 * v3) / 2.0, c1 = (v3 + v4) / 2.0;
        double e0 = (1.0 / 3.0) / sq(1e-6 + sq(v3 - v2));
        double e1 = (2.0 / 3.0) / sq(1e-6 + sq(v4 - v3));
        return (e0 * c0 + e1 * c1) / (e0 + e1);
    }
    return v3;
}

/* WENO face value along a strided line: cells base[m*stride], m in [0, n), validity okb[m*ostride];
 * face f between cells f-1 and f, upwind side from the sign of q (>= 0: low side). */
static inline double line_face(const double* base, const uint8_t* okb, i64 stride, i64 ostride, int n, int f,
                               double q, int z) {
    int i1, i2, i3, i4, i5;
    if (q >= 0.0) { i1 = f - 3; i2 = f - 2; i3 = f - 1; i4 = f; i5 = f + 1; }
    else          { i1 = f + 2; i2 = f + 1; i3 = f; i4 = f - 1; i5 = f - 2; }
    int b1 = 0 <= i1 && i1 < n, b2 = 0 <= i2 && i2 < n, b3 = 0 <= i3 && i3 < n, b4 = 0 <= i4 && i4 < n,
        b5 = 0 <= i5 && i5 < n;
    return wval(b1 ? base[i1 * stride] : 0.0, b2 ? base[i2 * stride] : 0.0, b3 ? base[i3 * stride] : 0.0,
                b4 ? base[i4 * stride] : 0.0, b5 ? base[i5 * stride] : 0.0,
                b1 && okb[i1 * ostride], b2 && okb[i2 * ostride], b3 && okb[i3 * ostride],
                b4 && okb[i4 * ostride], b5 && okb[i5 * ostride], z);
}

/* _advect_faces_h, sigma_baroclinic.py:265-321.  fx, fy must be zeroed by the caller (np.zeros). */
void oc_faces_h(const double* c, const double* Qx, const double* Qy, const uint8_t* ok, const double* ext,
                int nt, int n, int ny, int nx, int open_mask, int z, double* fx, double* fy) {
    #pragma omp parallel for schedule(static)
    for (i64 p = 0; p < (i64)nt * n; ++p) {
        i64 t = p / n, k = p % n;
        const double* cp = c + (t * n + k) * ny * nx;
        const double* ep = ext + (t * n + k) * ny * nx;
        const uint8_t* op = ok + k * ny * nx;
        double* fxp = fx + (t * n + k) * ny * (nx + 1);
        double* fyp = fy + (t * n + k) * (ny + 1) * nx;
        for (int j = 0; j < ny; ++j) {
            for (int i = 1; i < nx; ++i) {
                double q = Qx[(k * ny + j) * (nx + 1) + i];
                if (q == 0.0) continue;
                fxp[j * (nx + 1) + i] = q * line_face(cp + j * nx, op + j * nx, 1, 1, nx, i, q, z);
            }
            double q = Qx[(k * ny + j) * (nx + 1) + 0];
            if (open_mask & OPEN_W) fxp[j * (nx + 1)] = q * (q > 0.0 ? ep[j * nx] : cp[j * nx]);
            q = Qx[(k * ny + j) * (nx + 1) + nx];
            if (open_mask & OPEN_E)
                fxp[j * (nx + 1) + nx] = q * (q < 0.0 ? ep[j * nx + nx - 1] : cp[j * nx + nx - 1]);
        }
        for (int j = 1; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double q = Qy[(k * (ny + 1) + j) * nx + i];
                if (q == 0.0) continue;
                fyp[j * nx + i] = q * line_face(cp + i, op + i, nx, nx, ny, j, q, z);
            }
        for (int i = 0; i < nx; ++i) {
            double q = Qy[(k * (ny + 1) + 0) * nx + i];
            if (open_mask & OPEN_S) fyp[i] = q * (q > 0.0 ? ep[i] : cp[i]);
            q = Qy[(k * (ny + 1) + ny) * nx + i];
            if (open_mask & OPEN_N)
                fyp[ny * nx + i] = q * (q < 0.0 ? ep[(ny - 1) * nx + i] : cp[(ny - 1) * nx + i]);
        }
    }
}

/* _advect_div, sigma_baroclinic.py:324-364: out = -div(Q c) with the vertical pass when vert. */
void oc_advect_div(const double* c, const double* Qx, const double* Qy, const double* Fz, const uint8_t* ok,
                   const double* ext, int nt, int n, int ny, int nx, int open_mask, int z, int vert, double* out) {
    i64 plane = (i64)ny * nx;
    double* fx = calloc((size_t)nt * n * ny * (nx + 1), sizeof(double));
    double* fy = calloc((size_t)nt * n * (ny + 1) * nx, sizeof(double));
    double* fz = vert ? calloc((size_t)nt * (n + 1) * plane, sizeof(double)) : NULL;
    oc_faces_h(c, Qx, Qy, ok, ext, nt, n, ny, nx, open_mask, z, fx, fy);
    if (vert) {
        #pragma omp parallel for schedule(static)
        for (i64 p = 0; p < (i64)nt * ny; ++p) {
            i64 t = p / ny, j = p % ny;
            for (int k = 1; k < n; ++k)
                for (int i = 0; i < nx; ++i) {
                    double q = Fz[(k * ny + j) * nx + i];
                    if (q == 0.0) continue;
                    fz[(t * (n + 1) + k) * plane + j * nx + i] =
                        q * line_face(c + t * n * plane + j * nx + i, ok + j * nx + i, plane, plane, n, k, q, z);
                }
        }
    }
    #pragma omp parallel for schedule(static)
    for (i64 p = 0; p < (i64)nt * n; ++p) {
        i64 t = p / n, k = p % n;
        const double* fxp = fx + p * ny * (nx + 1);
        const double* fyp = fy + p * (ny + 1) * nx;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double d = fxp[j * (nx + 1) + i + 1] - fxp[j * (nx + 1) + i] + fyp[(j + 1) * nx + i] - fyp[j * nx + i];
                if (vert)
                    d = d + fz[(t * (n + 1) + k + 1) * plane + j * nx + i] - fz[(t * (n + 1) + k) * plane + j * nx + i];
                out[p * plane + j * nx + i] = -d;
            }
    }
    free(fx); free(fy); free(fz);
}

/* _fct_stage, sigma_baroclinic.py:452-468 */
void oc_fct_stage(const double* q0, const double* qp, const double* fx, const double* fy, double a, double b,
                  double dt, const double* V, const double* C, int nt, int n, int ny, int nx,
                  double* q_out, double* c_out) {
    i64 plane = (i64)ny * nx;
    #pragma omp parallel for schedule(static)
    for (i64 p = 0; p < (i64)nt * n; ++p) {
        i64 k = p % n;
        const double* fxp = fx + p * ny * (nx + 1);
        const double* fyp = fy + p * (ny + 1) * nx;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                i64 x = p * plane + j * nx + i;
                double d = fxp[j * (nx + 1) + i + 1] - fxp[j * (nx + 1) + i] + fyp[(j + 1) * nx + i] - fyp[j * nx + i];
                double q = a * q0[x] + b * (qp[x] - dt * d);
                q_out[x] = q;
                double v = V[k * plane + j * nx + i];
                c_out[x] = v > 0 ? q / v : C[x];
            }
    }
}

/* _fct_fluxes, sigma_baroclinic.py:471-511.  FLx, FLy zeroed by the caller. */
void oc_fct_fluxes(const double* c, const double* Qx, const double* Qy, const double* ext, int nt, int n, int ny,
                   int nx, int open_mask, const double* fx0, const double* fy0, const double* fx1,
                   const double* fy1, const double* fx2, const double* fy2, double* FLx, double* FLy, double* Ax,
                   double* Ay) {
    #pragma omp parallel for schedule(static)
    for (i64 p = 0; p < (i64)nt * n; ++p) {
        i64 k = p % n;
        const double* cp = c + p * ny * nx;
        const double* ep = ext + p * ny * nx;
        i64 ox = p * ny * (nx + 1), oy = p * (ny + 1) * nx;
        for (int j = 0; j < ny; ++j) {
            const double* qx = Qx + (k * ny + j) * (nx + 1);
            for (int f = 1; f < nx; ++f) {
                double q = qx[f];
                FLx[ox + j * (nx + 1) + f] = q * (q >= 0 ? cp[j * nx + f - 1] : cp[j * nx + f]);
            }
            double q = qx[0];
            if (open_mask & OPEN_W) FLx[ox + j * (nx + 1)] = q * (q > 0 ? ep[j * nx] : cp[j * nx]);
            q = qx[nx];
            if (open_mask & OPEN_E)
                FLx[ox + j * (nx + 1) + nx] = q * (q < 0 ? ep[j * nx + nx - 1] : cp[j * nx + nx - 1]);
            for (int f = 0; f <= nx; ++f) {
                i64 x = ox + j * (nx + 1) + f;
                double fh = (fx0[x] + fx1[x]) / 6.0 + 2.0 * fx2[x] / 3.0;
                Ax[x] = fh - FLx[x];
            }
        }
        for (int i = 0; i < nx; ++i) {
            for (int f = 1; f < ny; ++f) {
                double q = Qy[(k * (ny + 1) + f) * nx + i];
                FLy[oy + f * nx + i] = q * (q >= 0 ? cp[(f - 1) * nx + i] : cp[f * nx + i]);
            }
            double q = Qy[(k * (ny + 1)) * nx + i];
            if (open_mask & OPEN_S) FLy[oy + i] = q * (q > 0 ? ep[i] : cp[i]);
            q = Qy[(k * (ny + 1) + ny) * nx + i];
            if (open_mask & OPEN_N)
                FLy[oy + ny * nx + i] = q * (q < 0 ? ep[(ny - 1) * nx + i] : cp[(ny - 1) * nx + i]);
            for (int f = 0; f <= ny; ++f) {
                i64 x = oy + f * nx + i;
                double fh = (fy0[x] + fy1[x]) / 6.0 + 2.0 * fy2[x] / 3.0;
                Ay[x] = fh - FLy[x];
            }
        }
    }
}

/* _fct_coeff_kernel, sigma_baroclinic.py:407-449.  Cx, Cy zeroed by the caller. */
void oc_fct_coeff(const double* c0, const double* cL, const double* V1, const double* Ax, const double* Ay,
                  const uint8_t* ok, double dt, int nt, int n, int ny, int nx, double* Cx, double* Cy) {
    i64 plane = (i64)ny * nx;
    double* r_in = calloc((size_t)nt * n * plane, sizeof(double));
    double* r_out = calloc((size_t)nt * n * plane, sizeof(double));
    #pragma omp parallel for schedule(static)
    for (i64 p = 0; p < (i64)nt * n; ++p) {
        i64 t = p / n, k = p % n;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (!ok[k * plane + j * nx + i]) continue;
                double cmax = -INFINITY, cmin = INFINITY;
                const int nb[7][3] = {{0, 0, 0}, {0, 0, -1}, {0, 0, 1}, {0, -1, 0}, {0, 1, 0}, {-1, 0, 0}, {1, 0, 0}};
                for (int m = 0; m < 7; ++m) {
                    i64 kk = k + nb[m][0], jj = j + nb[m][1], ii = i + nb[m][2];
                    if (0 <= kk && kk < n && 0 <= jj && jj < ny && 0 <= ii && ii < nx && ok[kk * plane + jj * nx + ii]) {
                        i64 x = (t * n + kk) * plane + jj * nx + ii;
                        cmax = dmax(dmax(cmax, c0[x]), cL[x]);
                        cmin = dmin(dmin(cmin, c0[x]), cL[x]);
                    }
                }
                i64 ox = p * ny * (nx + 1) + j * (nx + 1) + i, oy = p * (ny + 1) * nx + j * nx + i;
                double p_in = dt * (dmax(Ax[ox], 0.0) - dmin(Ax[ox + 1], 0.0) + dmax(Ay[oy], 0.0) - dmin(Ay[oy + nx], 0.0));
                double p_out = dt * (dmax(Ax[ox + 1], 0.0) - dmin(Ax[ox], 0.0) + dmax(Ay[oy + nx], 0.0) - dmin(Ay[oy], 0.0));
                i64 x = p * plane + j * nx + i;
                double q_in = dmax(V1[k * plane + j * nx + i] * (cmax - cL[x]), 0.0);
                double q_out = dmax(V1[k * plane + j * nx + i] * (cL[x] - cmin), 0.0);
                r_in[x] = p_in > 0.0 ? dmin(1.0, q_in / p_in) : 0.0;
                r_out[x] = p_out > 0.0 ? dmin(1.0, q_out / p_out) : 0.0;
            }
    }
    #pragma omp parallel for schedule(static)
    for (i64 p = 0; p < (i64)nt * n; ++p) {
        const double* ri = r_in + p * plane;
        const double* ro = r_out + p * plane;
        for (int j = 0; j < ny; ++j)
            for (int f = 1; f < nx; ++f) {
                i64 x = p * ny * (nx + 1) + j * (nx + 1) + f;
                Cx[x] = Ax[x] >= 0.0 ? dmin(ri[j * nx + f], ro[j * nx + f - 1]) : dmin(ri[j * nx + f - 1], ro[j * nx + f]);
            }
        for (int f = 1; f < ny; ++f)
            for (int i = 0; i < nx; ++i) {
                i64 x = p * (ny + 1) * nx + f * nx + i;
                Cy[x] = Ay[x] >= 0.0 ? dmin(ri[f * nx + i], ro[(f - 1) * nx + i]) : dmin(ri[(f - 1) * nx + i], ro[f * nx + i]);
            }
    }
    free(r_in); free(r_out);
}

/* _fct_final, sigma_baroclinic.py:514-530 */
void oc_fct_final(const double* q0, const double* FLx, const double* FLy, const double* Ax, const double* Ay,
                  const double* Cx, const double* Cy, double dt, const double* V1, const double* C, int nt, int n,
                  int ny, int nx, double* out) {
    i64 plane = (i64)ny * nx;
    #pragma omp parallel for schedule(static)
    for (i64 p = 0; p < (i64)nt * n; ++p) {
        i64 k = p % n;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                i64 e = p * ny * (nx + 1) + j * (nx + 1) + i + 1, w = e - 1;
                i64 s = p * (ny + 1) * nx + j * nx + i, nn = s + nx;
                double d = FLx[e] + Cx[e] * Ax[e] - (FLx[w] + Cx[w] * Ax[w]) + (FLy[nn] + Cy[nn] * Ay[nn])
                           - (FLy[s] + Cy[s] * Ay[s]);
                i64 x = p * plane + j * nx + i;
                double q = q0[x] - dt * d;
                double v = V1[k * plane + j * nx + i];
                out[x] = v > 0 ? q / v : C[x];
            }
    }
}

/* _rk_stage, sigma_baroclinic.py:856-875 (stage 1, 2 or 3; V broadcast over the tracers) */
void oc_rk_stage(const double* qa, const double* qb, const double* r, double dt, const double* V, const double* C,
                 int stage, i64 size, i64 nv, double* q_out, double* c_out) {
    #pragma omp parallel for schedule(static)
    for (i64 p = 0; p < size; ++p) {
        double q;
        if (stage == 1) q = qa[p] + dt * r[p];
        else if (stage == 2) q = 0.75 * qa[p] + 0.25 * (qb[p] + dt * r[p]);
        else q = qa[p] / 3.0 + 2.0 / 3.0 * (qb[p] + dt * r[p]);
        q_out[p] = q;
        double v = V[p % nv];
        c_out[p] = v > 0 ? q / v : C[p];
    }
}

static double* zalloc(size_t count) { return calloc(count, sizeof(double)); }

/* One step of TracerAdvection.advect_fct (core/components/tracers.py:208-254), the reference's buffer sequence.
 * ext: the inflow stack or NULL (then stage 1 and the donor flux use C, stages 2 and 3 their own values).
 * unlimited: C = 1 (the RK3 result).  fx_i (nt, n, ny) optional (face column flux_face_i); open_flux (nt)
 * optional: per-face edge fluxes summed in the caller. */
int oc_advect_fct(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                  const uint8_t* ok, const double* ext, int nt, int n, int ny, int nx, double dt, int open_mask,
                  int z, int unlimited, int flux_face_i, double* out, double* fx_i, double* Gx_out, double* Gy_out) {
    size_t S = (size_t)nt * n * ny * nx, P = (size_t)n * ny * nx;
    size_t SX = (size_t)nt * n * ny * (nx + 1), SY = (size_t)nt * n * (ny + 1) * nx;
    double *Va = malloc(P * 8), *Vb = malloc(P * 8), *q0 = malloc(S * 8);
    double *q1 = malloc(S * 8), *c1 = malloc(S * 8), *q2 = malloc(S * 8), *c2 = malloc(S * 8);
    double *fx0 = zalloc(SX), *fy0 = zalloc(SY), *fx1 = zalloc(SX), *fy1 = zalloc(SY), *fx2 = zalloc(SX),
           *fy2 = zalloc(SY), *FLx = zalloc(SX), *FLy = zalloc(SY), *Ax = malloc(SX * 8), *Ay = malloc(SY * 8),
           *Cx = zalloc(SX), *Cy = zalloc(SY);
    if (!Va || !Vb || !q0 || !q1 || !c1 || !q2 || !c2 || !fx0 || !fy0 || !fx1 || !fy1 || !fx2 || !fy2 || !FLx ||
        !FLy || !Ax || !Ay || !Cx || !Cy) return -1;
    for (size_t x = 0; x < P; ++x) {
        double dV = V1[x] - V0[x];
        Va[x] = V0[x] + dV;
        Vb[x] = V0[x] + 0.5 * dV;
    }
    #pragma omp parallel for schedule(static)
    for (i64 x = 0; x < (i64)S; ++x) q0[x] = V0[x % P] * C[x];
    const double* e0 = ext ? ext : C;
    oc_faces_h(C, Qx, Qy, ok, e0, nt, n, ny, nx, open_mask, z, fx0, fy0);
    oc_fct_stage(q0, q0, fx0, fy0, 0.0, 1.0, dt, Va, C, nt, n, ny, nx, q1, c1);
    oc_faces_h(c1, Qx, Qy, ok, ext ? ext : c1, nt, n, ny, nx, open_mask, z, fx1, fy1);
    oc_fct_stage(q0, q1, fx1, fy1, 0.75, 0.25, dt, Vb, C, nt, n, ny, nx, q2, c2);
    oc_faces_h(c2, Qx, Qy, ok, ext ? ext : c2, nt, n, ny, nx, open_mask, z, fx2, fy2);
    oc_fct_fluxes(C, Qx, Qy, e0, nt, n, ny, nx, open_mask, fx0, fy0, fx1, fy1, fx2, fy2, FLx, FLy, Ax, Ay);
    if (unlimited) {
        /* tracers.py:287-289: V1 c = V0 c0 - dt div F_H, F_H = (F0 + F1)/6 + 2 F2/3 formed again into fx0/fy0 */
        for (size_t x = 0; x < SX; ++x) fx0[x] = (fx0[x] + fx1[x]) / 6.0 + 2.0 * fx2[x] / 3.0;
        for (size_t x = 0; x < SY; ++x) fy0[x] = (fy0[x] + fy1[x]) / 6.0 + 2.0 * fy2[x] / 3.0;
        oc_fct_stage(q0, q0, fx0, fy0, 0.0, 1.0, dt, V1, C, nt, n, ny, nx, q1, out);
    } else {
        oc_fct_stage(q0, q0, FLx, FLy, 0.0, 1.0, dt, V1, C, nt, n, ny, nx, q1, c1);     /* c_L in c1 */
        oc_fct_coeff(C, c1, V1, Ax, Ay, ok, dt, nt, n, ny, nx, Cx, Cy);
        oc_fct_final(q0, FLx, FLy, Ax, Ay, Cx, Cy, dt, V1, C, nt, n, ny, nx, out);
    }
    if (fx_i || Gx_out) {
        for (size_t p = 0; p < (size_t)nt * n; ++p)
            for (int j = 0; j < ny; ++j)
                for (int f = 0; f <= nx; ++f) {
                    size_t x = p * ny * (nx + 1) + (size_t)j * (nx + 1) + f;
                    double g = unlimited ? fx0[x] : FLx[x] + Cx[x] * Ax[x];
                    if (Gx_out) Gx_out[x] = g;
                    if (fx_i && f == flux_face_i) fx_i[p * ny + j] = g;
                }
    }
    if (Gy_out)
        for (size_t x = 0; x < SY; ++x) Gy_out[x] = unlimited ? fy0[x] : FLy[x] + Cy[x] * Ay[x];
    free(Va); free(Vb); free(q0); free(q1); free(c1); free(q2); free(c2); free(fx0); free(fy0); free(fx1);
    free(fy1); free(fx2); free(fy2); free(FLx); free(FLy); free(Ax); free(Ay); free(Cx); free(Cy);
    return 0;
}

/* One step of TracerAdvection.advect (core/components/tracers.py:342-375): SSP-RK3 of _advect_div RHS with
 * _rk_stage.  ext: inflow stack or NULL (each stage its own values).  Fz (n+1, ny, nx); vert as any(Fz != 0). */
int oc_advect_rk3(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                  const double* Fz, const uint8_t* ok, const double* ext, int nt, int n, int ny, int nx, double dt,
                  int open_mask, int z, double* out) {
    size_t S = (size_t)nt * n * ny * nx, P = (size_t)n * ny * nx;
    int vert = 0;
    for (size_t x = 0; x < (size_t)(n + 1) * ny * nx; ++x) if (Fz[x] != 0.0) { vert = 1; break; }
    double *Va = malloc(P * 8), *Vb = malloc(P * 8), *q0 = malloc(S * 8), *r = malloc(S * 8);
    double *q1 = malloc(S * 8), *c1 = malloc(S * 8), *q2 = malloc(S * 8), *c2 = malloc(S * 8), *q3 = malloc(S * 8);
    if (!Va || !Vb || !q0 || !r || !q1 || !c1 || !q2 || !c2 || !q3) return -1;
    for (size_t x = 0; x < P; ++x) {
        double dV = V1[x] - V0[x];
        Va[x] = V0[x] + dV;
        Vb[x] = V0[x] + 0.5 * dV;
    }
    #pragma omp parallel for schedule(static)
    for (i64 x = 0; x < (i64)S; ++x) q0[x] = V0[x % P] * C[x];
    oc_advect_div(C, Qx, Qy, Fz, ok, ext ? ext : C, nt, n, ny, nx, open_mask, z, vert, r);
    oc_rk_stage(q0, q0, r, dt, Va, C, 1, (i64)S, (i64)P, q1, c1);
    oc_advect_div(c1, Qx, Qy, Fz, ok, ext ? ext : c1, nt, n, ny, nx, open_mask, z, vert, r);
    oc_rk_stage(q0, q1, r, dt, Vb, C, 2, (i64)S, (i64)P, q2, c2);
    oc_advect_div(c2, Qx, Qy, Fz, ok, ext ? ext : c2, nt, n, ny, nx, open_mask, z, vert, r);
    oc_rk_stage(q0, q2, r, dt, V1, C, 3, (i64)S, (i64)P, q3, out);
    free(Va); free(Vb); free(q0); free(r); free(q1); free(c1); free(q2); free(c2); free(q3);
    return 0;
}
