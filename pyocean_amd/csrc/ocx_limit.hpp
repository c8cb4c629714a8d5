// Zalesak limiter and final update of the FCT step as one wave-marching kernel (gfx950).
//
// One wavefront owns a strip of 60 cells of one (tracer, layer) plane and marches down a chunk of rows.  Lane l sits
// on column i0 - 2 + l: lanes 2..61 are the strip's cells, lanes 1 and 62 the halo columns whose ratios R+- the
// edge cells need, lanes 0 and 63 the columns those ratios take their bounds from.  At row j the wave computes R+- of row j+1 (bounds over the cell, i +- 1, j +- 1, k +- 1 of C
// and c_L, ok cells only; P+- from the antidiffusive fluxes on the cell's faces; _fct_coeff_kernel,
// core/dynamics/sigma_baroclinic.py:407-431) and finishes row j: face coefficients from the ratios of the two
// adjacent cells (:436-447), G = F_L + C A, q = q0 - dt div G, out = q/V1 (safe) (_fct_final :514-530).  R+- of
// the west/east neighbours come from the adjacent lanes, those of rows j-1, j+1 from a register ring, so R+- never
// goes through HBM.  The waves of a workgroup are the tracers of one layer; the workgroups of layers k-1, k, k+1
// side in one workgroup, so the k +- 1 reads of C and c_L hit the caches of that compute unit.
#pragma once
#include "ocx_march.hpp"

namespace ocx {

constexpr int LS = 60;     // own cells per strip of the limiter kernel

struct LimitArgs {
    const double* C;
    const double* cL;
    const double* ext;     // inflow values or nullptr (then C)
    const double* V0;
    const double* V1;
    const double* Qx;
    const double* Qy;
    const uint8_t* ok;
    const void* ax;        // antidiffusive fluxes A (the precision's flux type)
    const void* ay;
    double* out;
    double* fxi;           // optional face column flux_face_i of G
    int flux_face_i;
    double* gedge;         // optional open-edge face fluxes
    int nt, n, ny, nx;
    int rch, nsx, ncy;
    double dt;
    unsigned flags;
};

// loads of one row r that the ratios of that row need (fetched one row ahead)
struct RatioIn {
    double aw, an, as;       // A on the lane's west, north and south faces
    double cd, ld, cu, lu;   // C and c_L in the layers below and above
    double v1;
    bool od, ou;             // validity below / above
};

// loads of one row j that the final update needs (fetched one row ahead)
struct FinalIn {
    double qx, qys, qyn, v0, v1;
};

struct RowR {
    double rin, rout;        // ratios of the lane's cell in this row
    double aw, an;           // A on the lane's west face and on its north face
};

// Work decomposition: a workgroup holds the waves of up to MAXW tracers of one (layer, strip, chunk) (shared
// per-layer fields from the CU's caches); logical blocks run layer-fastest, so the workgroups of layers k-1, k, k+1
// of a strip run side by side on one XCD and the k +- 1 reads of C and c_L hit its L2.
// Logical blocks of one limiter launch over a domain: tracer groups x layers x strips x chunks.
__host__ __device__ inline long long limit_blocks(const LimitArgs& a, int wpb) {
    return (long long)((a.nt + wpb - 1) / wpb) * a.n * a.nsx * a.ncy;
}

// One logical block L of the limiter (wave wv of wpb).
template <class PR>
__device__ __forceinline__ void limit_block(const LimitArgs& a, long long L, int wv, int wpb) {
    using FL = typename PR::fl;
    const int lane = threadIdx.x & 63;
    const int ntb = (a.nt + wpb - 1) / wpb;
    const int tg = (int)(L % ntb);
    L /= ntb;
    const int k = (int)(L % a.n);
    L /= a.n;
    const int sx = (int)(L % a.nsx);
    const int cy = (int)(L / a.nsx);
    const int t = tg * wpb + wv;
    if (t >= a.nt || cy >= a.ncy) return;         // whole wave leaves; no block barrier is used
    const int p = t * a.n + k;
    const int ny = a.ny, nx = a.nx, n = a.n;
    const int i0 = sx * LS, i = i0 - 2 + lane;
    const int jb = cy * a.rch, je = min(jb + a.rch, ny);
    const size_t plane = (size_t)ny * nx;
    const bool col_in = i >= 0 && i < nx;
    const bool own = lane >= 2 && lane < 2 + LS && col_in;
    const bool rlane = lane >= 1 && lane <= 62;    // lanes whose ratios are used
    const double* Cp = a.C + (size_t)p * plane;
    const double* cLp = a.cL + (size_t)p * plane;
    const double* ext0 = a.ext ? a.ext + (size_t)p * plane : Cp;
    const uint8_t* okk = a.ok + (size_t)k * plane;
    const double* V0k = a.V0 + (size_t)k * plane;
    const double* V1k = a.V1 + (size_t)k * plane;
    const double* Qxk = a.Qx + (size_t)k * ny * (nx + 1);
    const double* Qyk = a.Qy + (size_t)k * (ny + 1) * nx;
    const FL* axp = static_cast<const FL*>(a.ax) + (size_t)p * ny * (nx + 1);
    const FL* ayp = static_cast<const FL*>(a.ay) + (size_t)p * (ny + 1) * nx;
    const bool has_dn = k > 0, has_up = k + 1 < n;
    const double dt = a.dt;
    const unsigned flags = a.flags;

    // Loads are issued unconditionally from addresses that are always inside the arrays and masked one loop step
    // later, when the row is used, so a wave never waits on a load right after issuing it.
    struct RawC {
        double c, l;
        uint8_t o;
        bool in;
    };
    auto ldc = [&](int r) {
        const bool in = col_in && r >= 0 && r < ny;
        const size_t x = in ? (size_t)r * nx + i : 0;
        return RawC{Cp[x], cLp[x], okk[x], in};
    };
    auto usec = [&](const RawC& q, double& c, double& l, bool& o) {
        c = q.in ? q.c : 0.0;
        l = q.in ? q.l : 0.0;
        o = q.in && q.o != 0;
    };
    struct RawR {
        RatioIn q;
        uint8_t od, ou;
        bool row, aw_in, cell;
    };
    auto ld_ratio = [&](int r) {
        RawR w;
        w.row = r >= 0 && r < ny;
        w.aw_in = w.row && i >= 0 && i <= nx;
        w.cell = w.row && col_in && rlane;
        const size_t x = w.cell ? (size_t)r * nx + i : 0;
        w.q.aw = (double)axp[w.aw_in ? (size_t)r * (nx + 1) + i : 0];
        w.q.an = (double)ayp[w.cell ? x + nx : 0];
        w.q.as = (double)ayp[x];
        w.q.v1 = V1k[x];
        const size_t xd = (w.cell && has_dn) ? x - plane : 0, xu = (w.cell && has_up) ? x + plane : 0;
        w.od = okk[xd];
        w.q.cd = Cp[xd];
        w.q.ld = cLp[xd];
        w.ou = okk[xu];
        w.q.cu = Cp[xu];
        w.q.lu = cLp[xu];
        return w;
    };
    auto use_ratio = [&](const RawR& w) {
        RatioIn q = w.q;
        if (!w.aw_in) q.aw = 0.0;
        if (!w.cell) q.an = q.as = q.v1 = 0.0;
        q.od = w.cell && has_dn && w.od != 0;
        q.ou = w.cell && has_up && w.ou != 0;
        if (!q.od) q.cd = q.ld = 0.0;
        if (!q.ou) q.cu = q.lu = 0.0;
        return q;
    };
    struct RawF {
        FinalIn f;
        bool xf, ow_;
    };
    auto ld_final = [&](int j) {
        RawF w;
        const bool row = j < je;
        w.xf = row && i >= 0 && i <= nx;
        w.ow_ = row && own;
        const size_t x = w.ow_ ? (size_t)j * nx + i : 0;
        w.f.qx = Qxk[w.xf ? (size_t)j * (nx + 1) + i : 0];
        w.f.qys = Qyk[x];
        w.f.qyn = Qyk[w.ow_ ? x + nx : 0];
        w.f.v0 = V0k[x];
        w.f.v1 = V1k[x];
        return w;
    };
    auto use_final = [&](const RawF& w) {
        FinalIn f = w.f;
        if (!w.xf) f.qx = 0.0;
        if (!w.ow_) f.qys = f.qyn = f.v0 = f.v1 = 0.0;
        return f;
    };
    // R+- of row r from the lane's column values of rows r-1, r, r+1 and the row's prefetched loads
    auto ratios = [&](int r, const RatioIn& q, double cs, double ls, bool os, double cc, double lc, bool oc,
                      double cn, double ln, bool on) -> RowR {
        RowR res{0.0, 0.0, q.aw, q.an};
        const double cw = shfl_up1(cc), lw = shfl_up1(lc);
        const double ce = shfl_down1(cc), le = shfl_down1(lc);
        const int owi = dpp_up1((int)oc), oei = dpp_down1((int)oc);
        const double ae = shfl_down1(q.aw);
        if (!(rlane && oc && r >= 0 && r < ny)) return res;      // oc implies the cell is in the domain
        double cmax = -INFINITY, cmin = INFINITY;
        auto take = [&](bool v, double c0, double l0) {
            if (v) {
                cmax = dmax(dmax(cmax, c0), l0);
                cmin = dmin(dmin(cmin, c0), l0);
            }
        };
        take(true, cc, lc);
        take(owi != 0, cw, lw);
        take(oei != 0, ce, le);
        take(os, cs, ls);
        take(on, cn, ln);
        take(q.od, q.cd, q.ld);
        take(q.ou, q.cu, q.lu);
        const double p_in = dt * (dmax(q.aw, 0.0) - dmin(ae, 0.0) + dmax(q.as, 0.0) - dmin(q.an, 0.0));
        const double p_out = dt * (dmax(ae, 0.0) - dmin(q.aw, 0.0) + dmax(q.an, 0.0) - dmin(q.as, 0.0));
        const double q_in = dmax(q.v1 * (cmax - lc), 0.0);
        const double q_out = dmax(q.v1 * (lc - cmin), 0.0);
        res.rin = p_in > 0.0 ? dmin(1.0, q_in / p_in) : 0.0;
        res.rout = p_out > 0.0 ? dmin(1.0, q_out / p_out) : 0.0;
        return res;
    };

    // ring of the lane's column: C, c_L, validity of rows j-1 .. j+3 (row j+3 is one row ahead of its use)
    double cr[5], lr[5];
    bool orr[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) usec(ldc(jb - 1 + m), cr[m], lr[m], orr[m]);
    double cm2, lm2;
    bool om2;
    usec(ldc(jb - 2), cm2, lm2, om2);
    RowR Rm = ratios(jb - 1, use_ratio(ld_ratio(jb - 1)), cm2, lm2, om2, cr[0], lr[0], orr[0], cr[1], lr[1], orr[1]);
    RowR Rc = ratios(jb, use_ratio(ld_ratio(jb)), cr[0], lr[0], orr[0], cr[1], lr[1], orr[1], cr[2], lr[2], orr[2]);
    RatioIn qn1 = use_ratio(ld_ratio(jb + 1));      // inputs of R(j+1)
    FinalIn fc = use_final(ld_final(jb));           // inputs of row j's update
    const bool ow = flags & OPEN_W, oe = flags & OPEN_E, osb = flags & OPEN_S, onb = flags & OPEN_N;

    for (int j = jb; j < je; ++j) {
        // one row ahead: the ring row j+4, the ratio inputs of row j+2, the update inputs of row j+1
        const RawC c4 = ldc(j + 4);
        const RawR qn2 = ld_ratio(j + 2);
        const RawF fn = ld_final(j + 1);
        // ratios of row j+1 (rows j .. j+2 = cr[1..3])
        const RowR Rn = ratios(j + 1, qn1, cr[1], lr[1], orr[1], cr[2], lr[2], orr[2], cr[3], lr[3], orr[3]);
        // finish row j
        const double rin_w = shfl_up1(Rc.rin), rout_w = shfl_up1(Rc.rout);
        const double rin_e = shfl_down1(Rc.rin), rout_e = shfl_down1(Rc.rout);
        const double ae = shfl_down1(Rc.aw);
        const double cw = shfl_up1(cr[1]), ce = shfl_down1(cr[1]);
        const double qxe = shfl_down1(fc.qx);
        if (own) {
            const size_t xc = (size_t)j * nx + i;
            const double c0 = cr[1];
            const double aw = Rc.aw, as_ = Rm.an, an = Rc.an;
            const double cfw = i > 0 ? coeff(aw, Rc.rin, rout_w, rin_w, Rc.rout) : 0.0;
            const double cfe = i + 1 < nx ? coeff(ae, rin_e, Rc.rout, Rc.rin, rout_e) : 0.0;
            const double cfs = j > 0 ? coeff(as_, Rc.rin, Rm.rout, Rm.rin, Rc.rout) : 0.0;
            const double cfn = j + 1 < ny ? coeff(an, Rn.rin, Rc.rout, Rc.rin, Rn.rout) : 0.0;
            const double* erow = ext0 + (size_t)j * nx;
            const double* ecol_s = ext0 + i;
            const double* ecol_n = ext0 + (size_t)(ny - 1) * nx + i;
            const double Gw = donor(fc.qx, i, nx, cw, c0, ow, oe, erow, erow + nx - 1) + cfw * aw;
            const double Ge = donor(qxe, i + 1, nx, c0, ce, ow, oe, erow, erow + nx - 1) + cfe * ae;
            const double Gs = donor(fc.qys, j, ny, cr[0], c0, osb, onb, ecol_s, ecol_n) + cfs * as_;
            const double Gn = donor(fc.qyn, j + 1, ny, c0, cr[2], osb, onb, ecol_s, ecol_n) + cfn * an;
            const double d = Ge - Gw + Gn - Gs;
            const double q = fc.v0 * c0 - dt * d;
            const double v1 = fc.v1;
            a.out[(size_t)p * plane + xc] = v1 > 0 ? q / v1 : c0;
            if (a.fxi) {
                if (a.flux_face_i == i) a.fxi[(size_t)p * ny + j] = Gw;
                if (a.flux_face_i == nx && i == nx - 1) a.fxi[(size_t)p * ny + j] = Ge;
            }
            if (a.gedge) {
                const size_t nyl = (size_t)a.nt * n * ny, nxl = (size_t)a.nt * n * nx;
                if (i == 0) a.gedge[(size_t)p * ny + j] = Gw;
                if (i == nx - 1) a.gedge[nyl + (size_t)p * ny + j] = Ge;
                if (j == 0) a.gedge[2 * nyl + (size_t)p * nx + i] = Gs;
                if (j == ny - 1) a.gedge[2 * nyl + nxl + (size_t)p * nx + i] = Gn;
            }
        }
        // advance
        Rm = Rc;
        Rc = Rn;
        qn1 = use_ratio(qn2);
        fc = use_final(fn);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            cr[m] = cr[m + 1];
            lr[m] = lr[m + 1];
            orr[m] = orr[m + 1];
        }
        usec(c4, cr[4], lr[4], orr[4]);
    }
}

template <class PR>
__global__ __launch_bounds__(64 * MAXW) void k_fct_limit(LimitArgs a) {
    limit_block<PR>(a, xcd_logical(blockIdx.x, gridDim.x), threadIdx.x >> 6, blockDim.x >> 6);
}

}  // namespace ocx
