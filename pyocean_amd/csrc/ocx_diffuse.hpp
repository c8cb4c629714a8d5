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
// coefficients cp_k depend on dz, nu and the rates only, so one pass stores them (k_thomas_cp) with each level's
// lower coefficient and pivot, and a second solves the fields of each column with them: k_thomas_solve_r for columns
// of up to 64 levels (the forward values in registers), k_thomas_solve beyond; columns are contiguous across
// threads.
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

// cp (n, M): forward-sweep coefficients, shared by every tracer.  With lw and mm (not null) it also stores each
// level's lower coefficient and pivot m_k = diag_k - lower_k cp_{k-1} (m_0 = diag_0), the values the solve of
// every field needs, so that k_thomas_solve_r reads them instead of recomputing the row
__global__ void k_thomas_cp(const double* __restrict__ dz, const double* __restrict__ nu, const double* __restrict__ r,
                            int r_layers, int n, long long M, double dt, double* __restrict__ cp,
                            double* __restrict__ lw, double* __restrict__ mm) {
    for (long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x; col < M;
         col += (long long)gridDim.x * blockDim.x) {
        double prev = 0.0;
        for (int k = 0; k < n; ++k) {
            const TriRow t = tri_row(dz, nu, r, r_layers, n, M, k, col, dt);
            const double m = k == 0 ? t.diag : t.diag - t.lower * prev;
            const double c = t.upper / m;
            cp[(size_t)k * M + col] = c;
            if (lw) {
                lw[(size_t)k * M + col] = t.lower;
                mm[(size_t)k * M + col] = m;
            }
            prev = c;
        }
    }
}

// The same solve for columns of n <= NB levels, one thread per (column, field): the forward sweep's values stay
// in registers (d[NB], every index static after unrolling), so each field is read once and written once; the
// per-level lower, pivot and cp come from k_thomas_cp.  Block (64, fy): 64 columns of fy fields.  A 1-D grid of
// column blocks x field groups, remapped XCD-contiguously with the field groups of a column block adjacent, so
// that the groups sharing a column's coefficients run together on one XCD and find them in its L2.
template <int NB>
__global__ __launch_bounds__(256) void k_thomas_solve_r(const double* u, const double* __restrict__ dz,
                                                        const double* __restrict__ tau_top, int nt, int n, long long M,
                                                        double dt, const double* __restrict__ cp,
                                                        const double* __restrict__ lw, const double* __restrict__ mm,
                                                        double* out, int groups) {
    const long long b = xcd_logical(blockIdx.x, gridDim.x);
    const long long col = (b / groups) * 64 + threadIdx.x;
    const int t = (int)(b % groups) * blockDim.y + threadIdx.y;
    // cp of the block's 64 columns, staged in LDS as [level][lane] for the back substitution, whose chain of
    // dependent steps would otherwise wait on a global load at every level (vertical solve 4.04 -> 3.30 ms at
    // 1024 x 1024 x 50 x 10)
    __shared__ double scp[NB * 64];
    {
        const long long c0 = (b / groups) * 64;
        const int tid = threadIdx.y * 64 + threadIdx.x, nthr = 64 * blockDim.y;
        for (int idx = tid; idx < n * 64; idx += nthr) {
            const int k = idx >> 6, l = idx & 63;
            scp[idx] = c0 + l < M ? cp[(size_t)k * M + c0 + l] : 0.0;
        }
        __syncthreads();
    }
    if (col >= M || t >= nt) return;
    const double tau = tau_top ? tau_top[col] : 0.0;     // the reference adds a zero stress too
    const double* uc = u + (size_t)t * n * M + col;
    double* oc = out + (size_t)t * n * M + col;
    const double* dzc = dz + col;
    const double* mmc = mm + col;
    const double* lwc = lw + col;
    double d[NB];
    double dprev = 0.0;
    // the inputs of level k are loaded PF levels ahead of its division (a ring of PF slots); levels k >= n repeat
    // level n-1 (in-bounds reads whose results are discarded), so the loop has no branch to hold the loads back
    constexpr int PF = 4;
    double ru[PF], rz[PF], rm[PF], rl[PF];
#pragma unroll
    for (int q = 0; q < PF; ++q) {
        const size_t x = (size_t)(q < n ? q : n - 1) * M;
        ru[q] = uc[x];
        rz[q] = dzc[x];
        rm[q] = mmc[x];
        rl[q] = lwc[x];
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int q = k % PF;
        const double uu = ru[q], zz = rz[q], m = rm[q], l = rl[q];
        if (k + PF < NB) {
            const size_t x = (size_t)(k + PF < n ? k + PF : n - 1) * M;
            ru[q] = uc[x];
            rz[q] = dzc[x];
            rm[q] = mmc[x];
            rl[q] = lwc[x];
        }
        double rhs = zz * uu;
        rhs = k == n - 1 ? rhs + dt * tau : rhs;
        // the pivot's reciprocal is refined off the recurrence; qdiv_r gives the IEEE quotient (rhs - l dprev) / m
        const Recip R = recip_of(m);
        const double v = qdiv_r(k == 0 ? rhs : rhs - l * dprev, R);
        d[k] = v;
        dprev = k < n ? v : dprev;
    }
#pragma unroll
    for (int k = NB - 1; k >= 0; --k) {
        if (k < n) {
            const size_t x = (size_t)k * M;
            if (k < n - 1) dprev = d[k] - scp[k * 64 + threadIdx.x] * dprev;
            oc[x] = dprev;
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
                const Recip R = recip_of(m);            // one refined reciprocal of the pivot for the TC fields
#pragma unroll
                for (int q = 0; q < TC; ++q) {
                    if (q < tn) {
                        const size_t x = ((size_t)(t0 + q) * n + k) * M + col;
                        double rhs = dzk * u[x];
                        if (k == n - 1) rhs += dt * tau;
                        const double d = qdiv_r(k == 0 ? rhs : rhs - tr.lower * dprev[q], R);
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
