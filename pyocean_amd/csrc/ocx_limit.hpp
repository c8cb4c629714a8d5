// Zalesak limiter and final update of the FCT step as one wave-marching kernel (gfx950).
//
// One wavefront owns a strip of 60 cells of one (tracer, layer) plane and marches down a chunk of rows.  Lane l sits
// on column i0 - 2 + l: lanes 2..61 are the strip's cells, lanes 1 and 62 the halo columns whose ratios R+- the
// edge cells need, lanes 0 and 63 the columns those ratios take their bounds from.  At row j the wave computes R+-
// of row j+1 (_fct_coeff_kernel, core/dynamics/sigma_baroclinic.py:407-431: bounds over the cell, i +- 1, j +- 1 and
// k +- 1 of C and c_L, ok cells only; P+- from the antidiffusive fluxes on the cell's faces) and finishes row j
// (_fct_final :514-530): face coefficients from the ratios of the two adjacent cells (:436-447), G = F_L + C A,
// q = q0 - dt div G, out = q/V1 (safe).
//
// Each cell's bound contribution (max and min of C and c_L, or -inf/+inf when the cell is not valid) is formed once
// when its row enters the register ring and reaches the neighbouring lanes by DPP shifts, so the 7-point bounds are
// twelve max/min instructions.  R+- of the west/east neighbours come from the adjacent lanes and those of rows
// j-1, j+1 from registers, so R+- never goes through HBM.  Each face flux G is formed once: the east face of a cell
// is the west face of the next lane, the south face of row j is the north face of row j-1 of the same lane.
// Loads address a wave-uniform plane base with 32-bit lane offsets (rows and columns outside the plane clamped to an
// in-bounds element whose value the validity masks discard).
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

// Work decomposition: a workgroup holds LG consecutive layers x (up to LMAXW) tracers of one (strip, chunk); its
// waves march the rows in step (one workgroup barrier per row).  Each wave publishes the bound contributions of its
// newest row in LDS, so a wave reads those of layers k +- 1 from the waves above and below it and only the group's
// bottom and top layers read their outer neighbour layer from memory; the waves of one layer (its tracers) read the
// shared per-layer fields together, through the compute unit's cache.
#ifndef OCX_LG
#define OCX_LG 4
#endif
constexpr int LG = OCX_LG;  // layers per limiter workgroup
constexpr int LMAXW = 4;   // tracers per limiter workgroup at most (LMAXW x LG waves <= 1024 threads)

// Logical blocks of one limiter launch over a domain: tracer groups x strips x layer groups x chunks.
__host__ __device__ inline long long limit_blocks(const LimitArgs& a, int wpb) {
    return (long long)((a.nt + wpb - 1) / wpb) * ((a.n + LG - 1) / LG) * a.nsx * a.ncy;
}

// LDS of one limiter workgroup: per wave, the bound contributions of three rows (a ring) for its 64 lanes.
__host__ __device__ constexpr size_t limit_lds_bytes(int wpb) { return (size_t)wpb * LG * 3 * 64 * 2 * sizeof(double); }

// Upper and lower bound contributions of one cell to the 7-point bounds of its neighbours: max and min of C and
// c_L where the cell is valid, -inf / +inf (no contribution) elsewhere.
struct Bnd {
    double hi, lo;
};
__device__ __forceinline__ Bnd bnd(bool v, double c, double l) {
    return v ? Bnd{fmax(c, l), fmin(c, l)} : Bnd{-INFINITY, INFINITY};
}

// One logical block L of the limiter: wave wv of the workgroup's wpb x LG waves (tracer wv % wpb of the group,
// layer wv / wpb of the layer group).  Every wave runs the whole row loop, also a wave whose tracer or layer lies
// past the stack (it computes on clamped in-bounds data and stores nothing), so every wave meets every barrier.
// lds: the workgroup's limit_lds_bytes(wpb) bytes of LDS.
template <class PR>
__device__ __forceinline__ void limit_block(const LimitArgs& a, long long L, int wv, int wpb, Bnd* lds) {
    using FL = typename PR::fl;
    const int lane = threadIdx.x & 63;
    const int ntb = (a.nt + wpb - 1) / wpb;
    const int nkg = (a.n + LG - 1) / LG;
    // tracer groups fastest, then strips: the strips of a row band run together, so the cache lines a strip's
    // halo lanes share with its neighbours are read from L2 (7.9 -> 7.4 ms at 1024 x 1024 x 50 x 10 against
    // layer groups before strips)
    const int tg = (int)(L % ntb);
    L /= ntb;
    const int sx = (int)(L % a.nsx);
    L /= a.nsx;
    const int kg = (int)(L % nkg);
    const int cy = (int)(L / nkg);
    const int tw = wv % wpb, kw = wv / wpb;
    const int t_raw = tg * wpb + tw, k_raw = kg * LG + kw;
    const bool active = t_raw < a.nt && k_raw < a.n;
    const int t = min(t_raw, a.nt - 1), k = min(k_raw, a.n - 1);    // in-bounds planes for an idle wave
    const int p = t * a.n + k;
    const int ny = a.ny, nx = a.nx, n = a.n;
    const int i0 = sx * LS, i = i0 - 2 + lane;
    const int jb = cy * a.rch, je = min(jb + a.rch, ny);
    const size_t plane = (size_t)ny * nx;
    const bool col_in = i >= 0 && i < nx;
    const bool xcol = i >= 0 && i <= nx;
    const bool own = active && lane >= 2 && lane < 2 + LS && col_in;
    const double dt = a.dt;
    const unsigned flags = a.flags;
    const bool ow = flags & OPEN_W, oe = flags & OPEN_E, osb = flags & OPEN_S, onb = flags & OPEN_N;
    const bool has_dn = k > 0, has_up = k + 1 < n;
    // layers k -+ 1 from the neighbouring waves of the workgroup, or (group edge) from memory
    const bool dn_lds = kw > 0, up_lds = kw + 1 < LG && k_raw + 1 < n;
    const bool dn_mem = has_dn && !dn_lds, up_mem = has_up && !up_lds;
    Bnd* my = lds + (size_t)wv * 3 * 64;
    const Bnd* below = lds + (size_t)(wv - wpb) * 3 * 64;           // used only when dn_lds
    const Bnd* above = lds + (size_t)(wv + wpb) * 3 * 64;           // used only when up_lds
    auto slot = [](int r) { return ((r % 3) + 3) % 3; };
    // wave-uniform plane bases
    const double* Cp = a.C + (size_t)p * plane;
    const double* cLp = a.cL + (size_t)p * plane;
    const double* Cd = has_dn ? Cp - plane : Cp;
    const double* Ld = has_dn ? cLp - plane : cLp;
    const double* Cu = has_up ? Cp + plane : Cp;
    const double* Lu = has_up ? cLp + plane : cLp;
    const uint8_t* okk = a.ok + (size_t)k * plane;
    const uint8_t* okd = has_dn ? okk - plane : okk;
    const uint8_t* oku = has_up ? okk + plane : okk;
    const double* V0k = a.V0 + (size_t)k * plane;
    const double* V1k = a.V1 + (size_t)k * plane;
    const double* Qxk = a.Qx + (size_t)k * ny * (nx + 1);
    const double* Qyk = a.Qy + (size_t)k * (ny + 1) * nx;
    const FL* axp = static_cast<const FL*>(a.ax) + (size_t)p * ny * (nx + 1);
    const FL* ayp = static_cast<const FL*>(a.ay) + (size_t)p * (ny + 1) * nx;
    const double* ext0 = a.ext ? a.ext + (size_t)p * plane : Cp;
    double* outp = a.out + (size_t)p * plane;
    // lane byte offsets inside a row (column 0 for lanes outside the row)
    const uint32_t o8 = col_in ? (uint32_t)i * 8u : 0u, o1 = col_in ? (uint32_t)i : 0u;
    const uint32_t ox8 = xcol ? (uint32_t)i * 8u : 0u;
    const uint32_t of = col_in ? (uint32_t)i * (uint32_t)sizeof(FL) : 0u;
    const uint32_t oxf = xcol ? (uint32_t)i * (uint32_t)sizeof(FL) : 0u;
    const uint32_t rc8 = (uint32_t)nx * 8u, rx8 = (uint32_t)(nx + 1) * 8u;
    const uint32_t rcf = (uint32_t)nx * (uint32_t)sizeof(FL), rxf = (uint32_t)(nx + 1) * (uint32_t)sizeof(FL);
    auto rowc = [&](int r) -> uint32_t { return (r >= 0 && r < ny) ? (uint32_t)r : 0u; };     // cell / x-face rows
    auto rowy = [&](int r) -> uint32_t { return (r >= 0 && r <= ny) ? (uint32_t)r : 0u; };    // y-face rows
    auto row_in = [&](int r) { return r >= 0 && r < ny; };

    // Loads are issued one loop step before their use, so a wave never waits on a load right after issuing it.
    struct RingRaw {
        double c, l;
        uint8_t o;
    };
    auto ld_ring = [&](int r) {
        const uint32_t rr = rowc(r);
        return RingRaw{ldo(Cp, o8 + rr * rc8), ldo(cLp, o8 + rr * rc8), ldo(okk, o1 + rr * (uint32_t)nx)};
    };
    // ratio inputs of row r: A on the west face (r, i) and north face (r+1, i), V1, and (group edge) the layer
    // below / above from memory
    struct RatRaw {
        FL aw, an;
        double v1, cv, lv;
        uint8_t ov;
    };
    auto ld_rat = [&](int r) {
        const uint32_t rr = rowc(r), ry = rowy(r + 1);
        const uint32_t xc = o8 + rr * rc8, xb = o1 + rr * (uint32_t)nx;
        RatRaw w{ldo(axp, oxf + rr * rxf), ldo(ayp, of + ry * rcf), ldo(V1k, xc), 0.0, 0.0, 0};
        if (dn_mem) {
            w.cv = ldo(Cd, xc);
            w.lv = ldo(Ld, xc);
            w.ov = ldo(okd, xb);
        } else if (up_mem) {
            w.cv = ldo(Cu, xc);
            w.lv = ldo(Lu, xc);
            w.ov = ldo(oku, xb);
        }
        return w;
    };
    // a group-edge wave with both outer neighbours (LG = 1 would need it) cannot occur: LG > 1
    static_assert(LG > 1, "a layer group has a bottom and a top wave");
    struct Rat {
        double aw, an, v1;
        Bnd mem;                 // the outer neighbour layer's contribution (group-edge waves)
    };
    auto use_rat = [&](const RatRaw& w, int r) {
        const bool in = col_in && row_in(r);
        return Rat{(double)w.aw, (double)w.an, w.v1, bnd(in && (dn_mem || up_mem) && w.ov != 0, w.cv, w.lv)};
    };
    // the layers below and above at row r: from LDS (published two or more barriers ago) or from memory
    auto vert = [&](int r, const Rat& q, Bnd& d, Bnd& u) {
        const int sl = slot(r);
        d = dn_lds ? below[sl * 64 + lane] : (dn_mem ? q.mem : Bnd{-INFINITY, INFINITY});
        u = up_lds ? above[sl * 64 + lane] : (up_mem ? q.mem : Bnd{-INFINITY, INFINITY});
    };
    // update inputs of row j: V0, Qx on the west face, Qy on the north face
    struct FinRaw {
        double v0, qx, qyn;
    };
    auto ld_fin = [&](int j) {
        const uint32_t rr = rowc(j);
        return FinRaw{ldo(V0k, o8 + rr * rc8), ldo(Qxk, ox8 + rr * rx8), ldo(Qyk, o8 + rowy(j + 1) * rc8)};
    };

    // R+- of row r (_fct_coeff_kernel, core/dynamics/sigma_baroclinic.py:407-431): bounds over the cell, i +- 1
    // (neighbour lanes), r +- 1 (bS, bN) and k +- 1 (d, u), valid cells only; as = A on the south face
    struct RR {
        double rin, rout;
    };
    auto ratios = [&](const Bnd& b, const Bnd& bS, const Bnd& bN, const Bnd& d, const Bnd& u, const Rat& q,
                      double as, double lc, bool oc) -> RR {
        const double hw = shfl_up1(b.hi), he = shfl_down1(b.hi);
        const double lw = shfl_up1(b.lo), le = shfl_down1(b.lo);
        const double ae = shfl_down1(q.aw);
        const double cmax = fmax(fmax(fmax(b.hi, hw), fmax(he, bS.hi)), fmax(fmax(bN.hi, d.hi), u.hi));
        const double cmin = fmin(fmin(fmin(b.lo, lw), fmin(le, bS.lo)), fmin(fmin(bN.lo, d.lo), u.lo));
        const double p_in = dt * (dmax(q.aw, 0.0) - dmin(ae, 0.0) + dmax(as, 0.0) - dmin(q.an, 0.0));
        const double p_out = dt * (dmax(ae, 0.0) - dmin(q.aw, 0.0) + dmax(q.an, 0.0) - dmin(as, 0.0));
        const double q_in = dmax(q.v1 * (cmax - lc), 0.0);
        const double q_out = dmax(q.v1 * (lc - cmin), 0.0);
        RR res;
        res.rin = (oc && p_in > 0.0) ? dmin(1.0, q_in / p_in) : 0.0;
        res.rout = (oc && p_out > 0.0) ? dmin(1.0, q_out / p_out) : 0.0;
        return res;
    };

    // ring of the lane's column, rows j .. j+2: C, c_L, validity and bound contributions
    double cr[3], lr[3];
    bool orr[3];
    Bnd br[3];
    auto ring_in = [&](const RingRaw& w, int r, int m) {
        cr[m] = w.c;
        lr[m] = w.l;
        orr[m] = col_in && row_in(r) && w.o != 0;
        br[m] = bnd(orr[m], w.c, w.l);
    };
    auto publish = [&](int r, const Bnd& b) { my[slot(r) * 64 + lane] = b; };
    // rows jb-2 .. jb+1; rows jb-1, jb, jb+1 published for the layers above and below
    const RingRaw rm2 = ld_ring(jb - 2), rm1 = ld_ring(jb - 1), r0 = ld_ring(jb), r1 = ld_ring(jb + 1);
    const Rat qm1 = use_rat(ld_rat(jb - 1), jb - 1);
    const Rat q0 = use_rat(ld_rat(jb), jb);
    Rat q1 = use_rat(ld_rat(jb + 1), jb + 1);
    const double as_m1 = (double)ldo(ayp, of + rowy(jb - 1) * rcf);     // A on the south face of row jb-1
    ring_in(rm2, jb - 2, 0);
    ring_in(rm1, jb - 1, 1);
    ring_in(r0, jb, 2);
    publish(jb - 1, br[1]);
    publish(jb, br[2]);
    const Bnd b_p1 = bnd(col_in && row_in(jb + 1) && r1.o != 0, r1.c, r1.l);
    publish(jb + 1, b_p1);
    __syncthreads();
    Bnd vd, vu;
    vert(jb - 1, qm1, vd, vu);
    const RR Rm = ratios(br[1], br[0], br[2], vd, vu, qm1, as_m1, lr[1], orr[1]);
    const double cm1 = cr[1];                                           // C(jb-1)
    ring_in(r1, jb + 1, 0);                                             // rows jb .. jb+1 in slots 2, 0
    vert(jb, q0, vd, vu);
    RR Rc = ratios(br[2], br[1], br[0], vd, vu, q0, qm1.an, lr[2], orr[2]);
    // reorder the ring to rows jb, jb+1 in slots 0, 1 and load row jb+2 into slot 2
    {
        const double c0 = cr[2], l0 = lr[2];
        const bool o0 = orr[2];
        const Bnd b0 = br[2];
        cr[1] = cr[0];
        lr[1] = lr[0];
        orr[1] = orr[0];
        br[1] = br[0];
        cr[0] = c0;
        lr[0] = l0;
        orr[0] = o0;
        br[0] = b0;
        ring_in(ld_ring(jb + 2), jb + 2, 2);
    }
    __syncthreads();                                                    // row jb-1's slot is read by everyone
    publish(jb + 2, br[2]);
    __syncthreads();
    FinRaw fc = ld_fin(jb);
    double aw0 = q0.aw, an0 = q0.an, v1_0 = q0.v1;                     // row jb: A west / north, V1
    // G on the south face of row jb (_fct_final :514-530 with the face coefficient of :436-447)
    double Gs = 0.0;
    if (own) {
        const double* ecol_s = ext0 + i;
        const double* ecol_n = ext0 + (size_t)(ny - 1) * nx + i;
        const double qys = ldo(Qyk, o8 + rowy(jb) * rc8);
        const double cfs = jb > 0 ? coeff(qm1.an, Rc.rin, Rm.rout, Rm.rin, Rc.rout) : 0.0;
        Gs = donor(qys, jb, ny, cm1, cr[0], osb, onb, ecol_s, ecol_n) + cfs * qm1.an;
    }

    for (int j = jb; j < je; ++j) {
        // one row ahead: ring row j+3, ratio inputs of row j+2, update inputs of row j+1
        const RingRaw rn = ld_ring(j + 3);
        const RatRaw qn = ld_rat(j + 2);
        const FinRaw fn = ld_fin(j + 1);
        // ratios of row j+1 (ring rows j .. j+2; layers k -+ 1 of row j+1)
        vert(j + 1, q1, vd, vu);
        const RR Rn = ratios(br[1], br[0], br[2], vd, vu, q1, an0, lr[1], orr[1]);
        // finish row j
        const double rin_w = shfl_up1(Rc.rin), rout_w = shfl_up1(Rc.rout);
        const double c0 = cr[0];
        const double cw = shfl_up1(c0);
        // G on the lane's west face (lanes whose column is 0 .. nx): donor flux + coefficient x A
        const double* erow = ext0 + (size_t)j * nx;
        const double cfw = (i > 0 && i < nx) ? coeff(aw0, Rc.rin, rout_w, rin_w, Rc.rout) : 0.0;
        const double Gw = donor(fc.qx, i, nx, cw, c0, ow, oe, erow, erow + nx - 1) + cfw * aw0;
        const double Ge = shfl_down1(Gw);
        double Gn = 0.0;
        if (own) {
            const double* ecol_s = ext0 + i;
            const double* ecol_n = ext0 + (size_t)(ny - 1) * nx + i;
            const double cfn = j + 1 < ny ? coeff(an0, Rn.rin, Rc.rout, Rc.rin, Rn.rout) : 0.0;
            Gn = donor(fc.qyn, j + 1, ny, c0, cr[1], osb, onb, ecol_s, ecol_n) + cfn * an0;
            const double d = Ge - Gw + Gn - Gs;
            const double q = fc.v0 * c0 - dt * d;
            const double v1 = v1_0;
            sto(outp, o8 + (uint32_t)j * rc8, v1 > 0 ? q / v1 : c0);
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
        Gs = Gn;
        Rc = Rn;
        aw0 = q1.aw;
        an0 = q1.an;
        v1_0 = q1.v1;
        q1 = use_rat(qn, j + 2);
        fc = fn;
        cr[0] = cr[1];
        lr[0] = lr[1];
        orr[0] = orr[1];
        br[0] = br[1];
        cr[1] = cr[2];
        lr[1] = lr[2];
        orr[1] = orr[2];
        br[1] = br[2];
        ring_in(rn, j + 3, 2);
        publish(j + 3, br[2]);                 // row j's slot: every wave read it in the previous step
        __syncthreads();
    }
}

// blockDim.x = 64 * wpb * LG; dynamic LDS limit_lds_bytes(wpb)
template <class PR>
__global__ __launch_bounds__(64 * LMAXW * LG) void k_fct_limit(LimitArgs a) {
    extern __shared__ double limit_lds[];
    limit_block<PR>(a, xcd_logical(blockIdx.x, gridDim.x), wave_id(), (blockDim.x >> 6) / LG,
                    reinterpret_cast<Bnd*>(limit_lds));
}

}  // namespace ocx
