// Tracer diffusion of the ALE step (gfx950): the explicit Laplacian along the layers and the implicit vertical
// diffusion, both in the reference's per-element arithmetic (FP64, no contraction), so results are bit-identical.
//
// Horizontal (_diffuse_ale_kernel, core/dynamics/sigma_baroclinic.py:533-561): on each interior face between
// active cells with okx / oky set, F = ((-K_h c) h_f) (Tw_hi - Tw_lo) with c = L/dx of the face and h_f the thinner
// adjacent layer; out = T - dt div F / V where V > 0.  One thread per cell of a layer, looping over the tracers;
// rows are read coalesced and the four neighbours come from the caches.
//
// Vertical (_thomas_columns, core/dynamics/sigma_wd.py:57-88): backward Euler with interface diffusivity nu and the
// layer thicknesses dz, a tridiagonal system per column solved by the Thomas algorithm.  The forward-sweep
// coefficients cp_k depend on dz, nu and the rates only, so one pass stores them (k_thomas_cp) and a second solves
// the fields of each column with them (k_thomas_solve); columns are contiguous across threads.
#pragma once
#include "ocx_weno.hpp"

namespace ocx {

__device__ __forceinline__ double pymin(double a, double b) { return b < a ? b : a; }   // Python's min(a, b)

// grid (ceil(nx / 64), ceil(ny / 4), n), block (64, 4): one thread per cell of layer k for every tracer of the
// stack (the per-layer fields are read once and serve all tracers); rows are coalesced across lanes
__global__ __launch_bounds__(256) void k_diffuse_h(const double* __restrict__ T, const double* __restrict__ Tw,
                                                   const double* __restrict__ h, const uint8_t* __restrict__ act,
                                                   const double* __restrict__ cx, const double* __restrict__ cy,
                                                   const uint8_t* __restrict__ okx, const uint8_t* __restrict__ oky,
                                                   const double* __restrict__ V, int nt, int n, int ny, int nx,
                                                   double kh, double dt, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const size_t plane = (size_t)ny * nx;
    const int k = blockIdx.z;
    const double* hk = h + (size_t)k * plane;
    const double* Vk = V + (size_t)k * plane;
    const size_t c = (size_t)j * nx + i;
    const size_t fx = (size_t)j * (nx + 1) + i;
    // every load is issued up front from an in-bounds index (a missing neighbour reads the cell itself) and the
    // face conditions select afterwards, so no load waits behind another's branch; a selected flux is computed
    // exactly as the reference computes it
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    const size_t cw = hw ? c - 1 : c, ce = he ? c + 1 : c, cs = hs ? c - nx : c, cn = hn ? c + nx : c;
    const double h0 = hk[c], hw_ = hk[cw], he_ = hk[ce], hs_ = hk[cs], hn_ = hk[cn];
    const bool a0 = act[c], aw = act[cw], ae = act[ce], as = act[cs], an = act[cn];
    const bool fwk = hw && okx[fx] != 0, fek = he && okx[fx + 1] != 0, fsk = hs && oky[c] != 0;
    const bool fnk = hn && oky[c + nx] != 0;
    const double cxw = cx[fx], cxe = cx[fx + 1], cys = cy[c], cyn = cy[c + nx];
    const double v0 = Vk[c];
    const double huw = (aw && a0) ? pymin(hw_, h0) : 0.0, hue = (a0 && ae) ? pymin(h0, he_) : 0.0;
    const double hvs = (as && a0) ? pymin(hs_, h0) : 0.0, hvn = (a0 && an) ? pymin(h0, hn_) : 0.0;
    for (int t = 0; t < nt; ++t) {
        const size_t pl = ((size_t)t * n + k) * plane;
        const double* Wp = Tw + pl;
        const double w0 = Wp[c], ww = Wp[cw], we = Wp[ce], ws = Wp[cs], wn = Wp[cn];
        const double t0 = T[pl + c];
        const double fw = fwk ? -kh * cxw * huw * (w0 - ww) : 0.0;
        const double fe = fek ? -kh * cxe * hue * (we - w0) : 0.0;
        const double fs = fsk ? -kh * cys * hvs * (w0 - ws) : 0.0;
        const double fn = fnk ? -kh * cyn * hvn * (wn - w0) : 0.0;
        const double div = fe - fw + fn - fs;
        out[pl + c] = v0 > 0.0 ? t0 - dt * div / v0 : t0;
    }
}

// coefficients of layer k of column col: lower, the diagonal and upper of the tridiagonal row
struct TriRow {
    double lower, diag, upper;
};
__device__ __forceinline__ TriRow tri_row(const double* __restrict__ dz, const double* __restrict__ nu,
                                          const double* __restrict__ r, int r_layers, int n, long long M, int k,
                                          long long col, double dt) {
    TriRow t{0.0, dz[(size_t)k * M + col], 0.0};
    if (k > 0) {
        t.lower = -dt * nu[(size_t)k * M + col] / (0.5 * (dz[(size_t)k * M + col] + dz[(size_t)(k - 1) * M + col]));
        t.diag -= t.lower;
    }
    if (k < n - 1) {
        t.upper = -dt * nu[(size_t)(k + 1) * M + col] /
                  (0.5 * (dz[(size_t)k * M + col] + dz[(size_t)(k + 1) * M + col]));
        t.diag -= t.upper;
    }
    // rates: (M,) acts on layer 0 only (zero elsewhere, which still adds dt * 0), (n, M) on every layer
    const double rk = r ? (r_layers == 1 ? (k == 0 ? r[col] : 0.0) : r[(size_t)k * M + col]) : 0.0;
    t.diag += dt * rk;
    return t;
}

// cp (n, M): forward-sweep coefficients, shared by every tracer
__global__ void k_thomas_cp(const double* __restrict__ dz, const double* __restrict__ nu, const double* __restrict__ r,
                            int r_layers, int n, long long M, double dt, double* __restrict__ cp) {
    for (long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x; col < M;
         col += (long long)gridDim.x * blockDim.x) {
        double prev = 0.0;
        for (int k = 0; k < n; ++k) {
            const TriRow t = tri_row(dz, nu, r, r_layers, n, M, k, col, dt);
            const double c = k == 0 ? t.upper / t.diag : t.upper / (t.diag - t.lower * prev);
            cp[(size_t)k * M + col] = c;
            prev = c;
        }
    }
}

// out (nt, n, M) = the solution for u (nt, n, M); out may be u itself (each thread reads u[k] before writing out[k]).
// One thread per column solves TC fields at a time, so each level's dz, nu and cp are read once for all of them.
constexpr int TC = 4;
__global__ void k_thomas_solve(const double* u, const double* __restrict__ dz, const double* __restrict__ nu,
                               const double* __restrict__ r, int r_layers, const double* __restrict__ tau_top, int nt,
                               int n, long long M, double dt, const double* __restrict__ cp, double* out) {
    for (long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x; col < M;
         col += (long long)gridDim.x * blockDim.x) {
        const double tau = tau_top ? tau_top[col] : 0.0;     // the reference adds a zero stress too
        for (int t0 = 0; t0 < nt; t0 += TC) {
            const int tn = nt - t0 < TC ? nt - t0 : TC;
            double dprev[TC];
            for (int k = 0; k < n; ++k) {
                const TriRow tr = tri_row(dz, nu, r, r_layers, n, M, k, col, dt);
                const double dzk = dz[(size_t)k * M + col];
                const double m = k == 0 ? tr.diag : tr.diag - tr.lower * cp[(size_t)(k - 1) * M + col];
#pragma unroll
                for (int q = 0; q < TC; ++q) {
                    if (q < tn) {
                        const size_t x = ((size_t)(t0 + q) * n + k) * M + col;
                        double rhs = dzk * u[x];
                        if (k == n - 1) rhs += dt * tau;
                        const double d = k == 0 ? rhs / m : (rhs - tr.lower * dprev[q]) / m;
                        out[x] = d;
                        dprev[q] = d;
                    }
                }
            }
            for (int k = n - 2; k >= 0; --k) {
                const double ck = cp[(size_t)k * M + col];
#pragma unroll
                for (int q = 0; q < TC; ++q) {
                    if (q < tn) {
                        const size_t x = ((size_t)(t0 + q) * n + k) * M + col;
                        const double v = out[x] - ck * dprev[q];
                        out[x] = v;
                        dprev[q] = v;
                    }
                }
            }
        }
    }
}

}  // namespace ocx
