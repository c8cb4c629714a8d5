// C ABI of the FP64 tracer-advection library (declarations and contracts: include/ocx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ocx.h"
#include "ocx_kernels_f64.hpp"
#include "ocx_march.hpp"
#include "ocx_limit.hpp"
#include "ocx_batch.hpp"
#include "ocx_diffuse.hpp"
#include "ocx_host.hpp"

namespace ocx_host {
thread_local std::string g_err;
std::mutex g_prof_mu;
std::vector<ProfRec> g_prof;
bool g_prof_on = false;
}  // namespace ocx_host

namespace {

using namespace ocx_host;

constexpr int TX = 64, TY = 4;
// Tiles of the RK3 / pc / RHS layer-marching kernel: 256 cells in 4 wavefronts, RK_TX columns x 256 / RK_TX rows
// with the WENO faces, RK_TX_UP columns with the UP3 / UP1 / centred faces.  16 x 16 tiles form 3.1 WENO faces per
// cell (their 16 south and 16 east faces one evaluation) from a 22 x 22 halo tile, against 3.3 from 70 x 10 for
// 64 x 4: the sigma RK3 step took 36.3 ms against 38.6.  The cheaper faces keep 64 x 4, where the kernel needs
// fewer registers (the pc step 13.3 ms against 15.5 with 16 x 16 tiles).
#ifndef OCX_RK_TX
#define OCX_RK_TX 16
#endif
#ifndef OCX_RK_TX_UP
#define OCX_RK_TX_UP 64
#endif
constexpr int RK_TX = OCX_RK_TX, RK_TX_UP = OCX_RK_TX_UP;
template <int SCH>
constexpr int rk_tx() {
    return SCH <= ocx::SCH_Z ? RK_TX : RK_TX_UP;
}
constexpr bool rk_tile_ok(int tx) { return 256 % tx == 0 && (tx == 64 || tx + 256 / tx <= 64); }
static_assert(rk_tile_ok(RK_TX) && rk_tile_ok(RK_TX_UP), "RK tile shape");
long long rk_blocks(int nt, int ny, int nx, int tx) {
    const int ty = 256 / tx;
    return (long long)((nx + tx - 1) / tx) * ((ny + ty - 1) / ty) * nt;
}

int check_shape(int nt, int nz, int ny, int nx) {
    if (nt < 1 || nz < 1 || ny < 1 || nx < 1)
        return fail(OCX_E_ARG, "shape (nt=%d, nz=%d, ny=%d, nx=%d) must be positive", nt, nz, ny, nx);
    if ((long long)nt * nz > (1LL << 30) || ny > (1 << 24) || nx > (1 << 24))
        return fail(OCX_E_ARG, "shape (nt=%d, nz=%d, ny=%d, nx=%d) too large", nt, nz, ny, nx);
    // the kernels address one (tracer, layer) plane with 32-bit byte offsets
    if ((long long)(ny + 1) * (nx + 1) * 8 >= (1LL << 31))
        return fail(OCX_E_ARG, "plane ny=%d x nx=%d too large (at most 2^28 cells per layer)", ny, nx);
    return OCX_OK;
}

struct Sizes {
    size_t S, SX, SY, P;
    Sizes(int nt, int nz, int ny, int nx) {
        S = (size_t)nt * nz * ny * nx;
        SX = (size_t)nt * nz * ny * (nx + 1);
        SY = (size_t)nt * nz * (ny + 1) * nx;
        P = (size_t)nt * nz;
    }
};

// FCT workspace.  FP64: c1 (later c_L), c2, F0 (later A), S, open-edge fluxes.  Split: c1, c2 and the face
// stacks in FP32, c_L in its own FP64 stack.
struct FctWs {
    void *c1, *c2, *fx, *fy, *sx, *sy;
    double *cl, *gedge;
};

template <class PR>
size_t fct_layout(void* base, int nt, int nz, int ny, int nx, FctWs* w) {
    using ST = typename PR::st;
    using FL = typename PR::fl;
    Sizes s(nt, nz, ny, nx);
    Carver cv(base);
    FctWs t;
    t.c1 = cv.take<ST>(s.S);
    t.c2 = cv.take<ST>(s.S);
    t.fx = cv.take<FL>(s.SX);
    t.fy = cv.take<FL>(s.SY);
    t.sx = cv.take<FL>(s.SX);
    t.sy = cv.take<FL>(s.SY);
    t.cl = PR::split ? cv.take<double>(s.S) : static_cast<double*>(t.c1);
    t.gedge = cv.take<double>(2 * s.P * ny + 2 * s.P * nx);
    if (w) *w = t;
    return cv.used;
}

// launch shape of the single-domain calls (ocx_set_tuning): rows per chunk, waves per workgroup (0: default)
std::atomic<int> g_rch{0}, g_wpb{0};
constexpr int DEFAULT_RCH = 64;
constexpr int DEFAULT_WPB = 1;    // one tracer of a layer per workgroup (measured best at 1024 x 1024 x 50 x 10)
constexpr int DEFAULT_BATCH_WPB = 2;    // the persistent batch's default (measured best on the 64-domain batch)

int rows_per_chunk() {
    const int r = g_rch.load();
    return r > 0 ? r : DEFAULT_RCH;
}

// waves per workgroup = tracers per group: the tracers split into groups of at most `cap` (<= MAXW), as even as
// possible
int tracer_group(int nt, int cap = 0) {
    if (cap <= 0) cap = g_wpb.load();
    if (cap <= 0) cap = DEFAULT_WPB;
    if (cap > ocx::MAXW) cap = ocx::MAXW;
    const int groups = (nt + cap - 1) / cap;
    return (nt + groups - 1) / groups;
}

// tracers per limiter workgroup: the same grouping, capped at LMAXW (its workgroups also span LG layers)
int limit_group(int nt, int cap = 0) {
    if (cap <= 0) cap = g_wpb.load();
    if (cap <= 0) cap = DEFAULT_WPB;
    return tracer_group(nt, cap < ocx::LMAXW ? cap : ocx::LMAXW);
}

template <int STAGE, bool UNLIM, class PR>
void launch_march(const ocx::MarchArgs& a, bool js, hipStream_t s) {
    const int wpb = tracer_group(a.nt);
    const long long ntb = (a.nt + wpb - 1) / wpb;
    const unsigned blocks = (unsigned)(ntb * a.nsx * a.ncy * a.n);
    if (js)
        ocx::k_march<STAGE, false, UNLIM, PR><<<blocks, 64 * wpb, 0, s>>>(a);
    else
        ocx::k_march<STAGE, true, UNLIM, PR><<<blocks, 64 * wpb, 0, s>>>(a);
}

struct Rk3Ws {
    double *q, *c1, *c2, *fxs;
};

size_t rk3_layout(void* base, int nt, int nz, int ny, int nx, Rk3Ws* w) {
    Sizes s(nt, nz, ny, nx);
    Carver cv(base);
    Rk3Ws t;
    t.q = cv.take<double>(s.S);
    t.c1 = cv.take<double>(s.S);
    t.c2 = cv.take<double>(s.S);
    t.fxs = cv.take<double>(3 * s.P * ny);
    if (w) *w = t;
    return cv.used;
}

dim3 tile_grid(int nt, int nz, int ny, int nx) {
    const long long planes = (long long)nt * nz;
    return dim3((unsigned)((nx + TX - 1) / TX), (unsigned)((ny + TY - 1) / TY), (unsigned)std::min(planes, 65535LL));
}

// Face scheme of the flags.  The caller has rejected flags with more than one of CENTRED, UP3, UP1, WENO_JS.
int scheme_of(unsigned flags) {
    using namespace ocx;
    return (flags & CENTRED) ? SCH_C6
           : (flags & UP1)   ? SCH_UP1
           : (flags & UP3)   ? SCH_UP3
           : (flags & WENO_JS) ? SCH_JS
                               : SCH_Z;
}

// Arguments of one k_rk_stage launch.
struct RkArgs {
    const double *cin, *C, *ext, *V0, *V1, *Qx, *Qy, *Fz;
    const uint8_t* ok;
    double *q, *cout, *fxs;
    int fi, nt, nz, ny, nx;
    double dt;
    unsigned flags;
};

// k_rk_stage runs one block per (tile, tracer) on a 1-D grid; the host checks the block count fits it.
template <int STAGE, int SCH, bool VERT>
void rk_launch(const RkArgs& a, dim3, hipStream_t s) {
    constexpr int tx = rk_tx<SCH>(), ty = 256 / tx;
    const long long nb = rk_blocks(a.nt, a.ny, a.nx, tx);
    ocx::k_rk_stage<STAGE, SCH, VERT, tx, ty><<<dim3((unsigned)nb), dim3(64, 4), 0, s>>>(
        a.cin, a.C, a.ext, a.V0, a.V1, a.Qx, a.Qy, a.Fz, a.ok, a.q, a.cout, a.fxs, a.fi, a.nt, a.nz, a.ny, a.nx, a.dt,
        a.flags);
}

// One RK3 or 'pc' stage (STAGE 0: the RHS only) with the scheme's faces; UP3 / UP1 / centred always form the z
// faces.
template <int STAGE>
void rk_stage(const RkArgs& a, int sch, bool vert, dim3 grid, hipStream_t s) {
    using namespace ocx;
    switch (sch) {
        case SCH_C6: rk_launch<STAGE, SCH_C6, true>(a, grid, s); break;
        case SCH_UP3: rk_launch<STAGE, SCH_UP3, true>(a, grid, s); break;
        case SCH_UP1: rk_launch<STAGE, SCH_UP1, true>(a, grid, s); break;
        case SCH_JS: vert ? rk_launch<STAGE, SCH_JS, true>(a, grid, s) : rk_launch<STAGE, SCH_JS, false>(a, grid, s); break;
        default: vert ? rk_launch<STAGE, SCH_Z, true>(a, grid, s) : rk_launch<STAGE, SCH_Z, false>(a, grid, s);
    }
}

// The RK3 / pc / RHS kernel's 1-D grid of (tile, tracer) blocks must fit a launch.
int check_rk_grid(int nt, int ny, int nx) {
    const long long nb = std::max(rk_blocks(nt, ny, nx, RK_TX), rk_blocks(nt, ny, nx, RK_TX_UP));
    if (nb > 0x7fffffffLL) return fail(OCX_E_ARG, "%lld tiles x tracers exceed one launch", nb);
    return OCX_OK;
}

// Flags of the RK3 step, the RHS and faces_h: open edges plus at most one of WENO_JS, UP3, UP1 and (when
// `centred` allows it) CENTRED.
int check_rk_flags(unsigned flags, const char* what, bool centred = false) {
    using namespace ocx;
    if (flags & ~(31u | UP3 | UP1 | (centred ? CENTRED : 0u)))
        return fail(OCX_E_ARG, "flags 0x%x not valid for %s", flags, what);
    const int schemes = !!(flags & WENO_JS) + !!(flags & UP3) + !!(flags & UP1) + !!(flags & CENTRED);
    if (schemes > 1) return fail(OCX_E_ARG, "flags 0x%x name more than one face scheme", flags);
    return OCX_OK;
}

}  // namespace

extern "C" {

const char* ocx_last_error(void) { return g_err.c_str(); }

int ocx_version(void) { return 1; }

size_t ocx_fct_workspace_bytes(int nt, int nz, int ny, int nx) {
    if (check_shape(nt, nz, ny, nx) != OCX_OK) return 0;
    return fct_layout<ocx::PrecF64>(nullptr, nt, nz, ny, nx, nullptr);
}

size_t ocx_fct_split_workspace_bytes(int nt, int nz, int ny, int nx) {
    if (check_shape(nt, nz, ny, nx) != OCX_OK) return 0;
    return fct_layout<ocx::PrecSplit>(nullptr, nt, nz, ny, nx, nullptr);
}

size_t ocx_rk3_workspace_bytes(int nt, int nz, int ny, int nx) {
    if (check_shape(nt, nz, ny, nx) != OCX_OK) return 0;
    return rk3_layout(nullptr, nt, nz, ny, nx, nullptr);
}

}  // extern "C"

namespace {

template <class PR>
int advect_fct_impl(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                    const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx, double dt,
                    unsigned flags, int flux_face_i, double* C_out, double* fx_i, double* open_flux, void* workspace,
                    size_t workspace_bytes, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(C, "C"));
    OCX_TRY(need(V0, "V0"));
    OCX_TRY(need(V1, "V1"));
    OCX_TRY(need(Qx, "Qx"));
    OCX_TRY(need(Qy, "Qy"));
    OCX_TRY(need(ok, "ok"));
    OCX_TRY(need(C_out, "C_out"));
    if (C_out == C) return fail(OCX_E_ARG, "C_out must not alias C");
    if (flags & ~63u) return fail(OCX_E_ARG, "unknown flag bits 0x%x", flags & ~63u);
    if (flux_face_i != -1 && (flux_face_i < 0 || flux_face_i > nx))
        return fail(OCX_E_ARG, "flux_face_i %d outside 0..%d", flux_face_i, nx);
    if (flux_face_i >= 0 && !fx_i) return fail(OCX_E_ARG, "fx_i is NULL with flux_face_i = %d", flux_face_i);
    if (PR::split && (flags & UNLIMITED))
        return fail(OCX_E_ARG, "the unlimited step has no split-precision form");
    const size_t wsb = fct_layout<PR>(nullptr, nt, nz, ny, nx, nullptr);
    if (!workspace || workspace_bytes < wsb)
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed, %zu given", wsb, workspace ? workspace_bytes : 0);
    FctWs w;
    fct_layout<PR>(workspace, nt, nz, ny, nx, &w);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool js = flags & WENO_JS, unlim = flags & UNLIMITED;
    const int fi = fx_i ? flux_face_i : -1;
    const int rch = rows_per_chunk();
    if ((long long)((nx + LS - 1) / LS) * ((ny + rch - 1) / rch) * nt * nz >= (1LL << 31))
        return fail(OCX_E_ARG, "shape (nt=%d, nz=%d, ny=%d, nx=%d) too large for one launch", nt, nz, ny, nx);
    MarchArgs a{};
    a.C = C;
    a.ext = inflow;
    a.V0 = V0;
    a.V1 = V1;
    a.Qx = Qx;
    a.Qy = Qy;
    a.ok = ok;
    a.fx = w.fx;
    a.fy = w.fy;
    a.s_x = w.sx;
    a.s_y = w.sy;
    a.ax = w.fx;                     // A overwrites F0, dead after stage 2
    a.ay = w.fy;
    a.fxi = nullptr;
    a.flux_face_i = -1;
    a.nt = nt;
    a.n = nz;
    a.ny = ny;
    a.nx = nx;
    a.rch = rch;
    a.nsx = (nx + MS - 1) / MS;
    a.ncy = (ny + rch - 1) / rch;
    a.dt = dt;
    a.flags = flags;
    // stage 1: F0, c1
    {
        ProfScope _ps("fct_stage1", s);
        a.cin = C;
        a.cout = w.c1;
        launch_march<1, false, PR>(a, js, s);
    }
    OCX_TRY(launch_check("fct stage 1"));
    // stage 2: S = F0 + F1, c2
    {
        ProfScope _ps("fct_stage2", s);
        a.cin = w.c1;
        a.cout = w.c2;
        launch_march<2, false, PR>(a, js, s);
    }
    OCX_TRY(launch_check("fct stage 2"));
    if (unlim) {
        {
            ProfScope _ps("fct_unlimited", s);
            a.cin = w.c2;
            a.cout = C_out;
            a.fxi = fx_i;
            a.flux_face_i = fi;
            launch_march<3, true, PR>(a, js, s);
        }
        OCX_TRY(launch_check("fct unlimited"));
        if (open_flux) {
            // the unlimited step keeps no open-flux record (advect_fct_numpy returns before it, tracers.py:287-289)
            k_fill<<<1, 256, 0, s>>>(open_flux, 0.0, (size_t)nt);
            OCX_TRY(launch_check("open_flux clear"));
        }
        return OCX_OK;
    }
    // stage 3: F2, A = F_H - F_L (into the F0 buffers), c_L (into c1)
    {
        ProfScope _ps("fct_stage3_split", s);
        a.cin = w.c2;
        a.cout = w.cl;
        launch_march<3, false, PR>(a, js, s);
    }
    OCX_TRY(launch_check("fct high/low split"));
    // limiter and final update (c_L in c1, A in the F0 buffers)
    {
        ProfScope _ps("fct_limit_final", s);
        LimitArgs l{};
        l.C = C;
        l.cL = w.cl;
        l.ext = inflow;
        l.V0 = V0;
        l.V1 = V1;
        l.Qx = Qx;
        l.Qy = Qy;
        l.ok = ok;
        l.ax = w.fx;
        l.ay = w.fy;
        l.out = C_out;
        l.fxi = fx_i;
        l.flux_face_i = fi;
        l.gedge = open_flux ? w.gedge : nullptr;
        l.nt = nt;
        l.n = nz;
        l.ny = ny;
        l.nx = nx;
        l.rch = rch;
        l.nsx = (nx + LS - 1) / LS;
        l.ncy = (ny + rch - 1) / rch;
        l.dt = dt;
        l.flags = flags;
        const int wpb = limit_group(nt);
        const unsigned blocks = (unsigned)limit_blocks(l, wpb);
        k_fct_limit<PR><<<blocks, 64 * wpb * LG, limit_lds_bytes(wpb), s>>>(l);
    }
    OCX_TRY(launch_check("fct limiter"));
    if (open_flux) {
        k_open_flux<<<nt, 256, 0, s>>>(w.gedge, open_flux, nt, nz, ny, nx, flags);
        OCX_TRY(launch_check("open flux"));
    }
    return OCX_OK;
}

}  // namespace

extern "C" {

int ocx_advect_fct_f64(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                       const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx, double dt,
                       unsigned flags, int flux_face_i, double* C_out, double* fx_i, double* open_flux,
                       void* workspace, size_t workspace_bytes, void* stream) {
    return advect_fct_impl<ocx::PrecF64>(C, V0, V1, Qx, Qy, ok, inflow, nt, nz, ny, nx, dt, flags, flux_face_i,
                                         C_out, fx_i, open_flux, workspace, workspace_bytes, stream);
}

int ocx_advect_fct_split(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                         const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx, double dt,
                         unsigned flags, int flux_face_i, double* C_out, double* fx_i, double* open_flux,
                         void* workspace, size_t workspace_bytes, void* stream) {
    return advect_fct_impl<ocx::PrecSplit>(C, V0, V1, Qx, Qy, ok, inflow, nt, nz, ny, nx, dt, flags, flux_face_i,
                                           C_out, fx_i, open_flux, workspace, workspace_bytes, stream);
}

}  // extern "C"

namespace {

// Scratch of the allocation-free FCT entry: the step's workspace (stage values, face stacks, open-edge fluxes) in a
// file-scope device array.  Each GPU the library runs on holds its own copy, at its own address, until the process
// ends.  In host code `fct_scratch` names only a host-side stand-in: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is ever passed to a kernel.
constexpr size_t NOALLOC_SCRATCH_BYTES = OCX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char fct_scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's address of fct_scratch (looked up on the first call made on that device).
int scratch_address(void** out) {
    int dev = 0;
    OCX_TRY(hip_check(hipGetDevice(&dev), "hipGetDevice"));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(OCX_E_ARG, "device %d beyond %d", dev, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        OCX_TRY(hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(fct_scratch)), "hipGetSymbolAddress(fct_scratch)"));
        if (!p) return fail(OCX_E_HIP, "hipGetSymbolAddress(fct_scratch) gave a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return OCX_OK;
}

}  // namespace

extern "C" {

int ocx_advect_fct_f64_noalloc(const double* C, const double* V0, const double* V1, const double* Qx,
                               const double* Qy, const uint8_t* ok, const double* inflow, int nt, int nz, int ny,
                               int nx, double dt, unsigned flags, int flux_face_i, double* C_out, double* fx_i,
                               double* open_flux, void* stream) {
    OCX_TRY(check_shape(nt, nz, ny, nx));
    const size_t need_b = fct_layout<ocx::PrecF64>(nullptr, nt, nz, ny, nx, nullptr);
    if (need_b > NOALLOC_SCRATCH_BYTES)
        return fail(OCX_E_WORKSPACE,
                    "shape (nt=%d, nz=%d, ny=%d, nx=%d) needs %zu bytes of workspace; the allocation-free entry holds "
                    "%zu (use ocx_advect_fct_f64 with a workspace)", nt, nz, ny, nx, need_b, NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    OCX_TRY(scratch_address(&ws));
    return advect_fct_impl<ocx::PrecF64>(C, V0, V1, Qx, Qy, ok, inflow, nt, nz, ny, nx, dt, flags, flux_face_i,
                                         C_out, fx_i, open_flux, ws, NOALLOC_SCRATCH_BYTES, stream);
}

int ocx_advect_rk3_f64(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                       const double* Fz, const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx,
                       double dt, unsigned flags, int vert, int flux_face_i, double* C_out, double* fx_i,
                       void* workspace, size_t workspace_bytes, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(C, "C"));
    OCX_TRY(need(V0, "V0"));
    OCX_TRY(need(V1, "V1"));
    OCX_TRY(need(Qx, "Qx"));
    OCX_TRY(need(Qy, "Qy"));
    OCX_TRY(need(ok, "ok"));
    OCX_TRY(need(C_out, "C_out"));
    const bool up = flags & (UP3 | UP1);
    if (vert || up) OCX_TRY(need(Fz, "Fz"));
    if (C_out == C) return fail(OCX_E_ARG, "C_out must not alias C");
    OCX_TRY(check_rk_flags(flags, "the RK3 step"));
    OCX_TRY(check_rk_grid(nt, ny, nx));
    if (flux_face_i != -1 && (flux_face_i < 0 || flux_face_i > nx))
        return fail(OCX_E_ARG, "flux_face_i %d outside 0..%d", flux_face_i, nx);
    if (flux_face_i >= 0 && !fx_i) return fail(OCX_E_ARG, "fx_i is NULL with flux_face_i = %d", flux_face_i);
    const size_t wsb = rk3_layout(nullptr, nt, nz, ny, nx, nullptr);
    if (!workspace || workspace_bytes < wsb)
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed, %zu given", wsb, workspace ? workspace_bytes : 0);
    Rk3Ws w;
    rk3_layout(workspace, nt, nz, ny, nx, &w);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid = tile_grid(nt, nz, ny, nx);
    const int sch = scheme_of(flags);
    const int fi = fx_i ? flux_face_i : -1;
    const size_t col = (size_t)nt * nz * ny;
    double* f0 = fx_i ? w.fxs : nullptr;
    double* f1 = fx_i ? w.fxs + col : nullptr;
    double* f2 = fx_i ? w.fxs + 2 * col : nullptr;
    RkArgs a{nullptr, C, inflow, V0, V1, Qx, Qy, Fz, ok, w.q, nullptr, nullptr, fi, nt, nz, ny, nx, dt, flags};
    {
        ProfScope _ps("rk3_stage1", s);
        a.cin = C, a.cout = w.c1, a.fxs = f0;
        rk_stage<1>(a, sch, vert, grid, s);
        OCX_TRY(launch_check("rk3 stage 1"));
    }
    {
        ProfScope _ps("rk3_stage2", s);
        a.cin = w.c1, a.cout = w.c2, a.fxs = f1;
        rk_stage<2>(a, sch, vert, grid, s);
        OCX_TRY(launch_check("rk3 stage 2"));
    }
    {
        ProfScope _ps("rk3_stage3", s);
        a.cin = w.c2, a.cout = C_out, a.fxs = f2;
        rk_stage<3>(a, sch, vert, grid, s);
        OCX_TRY(launch_check("rk3 stage 3"));
    }
    if (fx_i) {
        k_rk_face_column<<<(unsigned)((col + 255) / 256), 256, 0, s>>>(f0, f1, f2, fx_i, col);
        OCX_TRY(launch_check("rk3 face column"));
    }
    return OCX_OK;
}

int ocx_advect_pc_f64(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                      const double* Fz, const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx,
                      double dt, unsigned flags, int vert, int flux_face_i, double* C_out, double* fx_i,
                      void* workspace, size_t workspace_bytes, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(C, "C"));
    OCX_TRY(need(V0, "V0"));
    OCX_TRY(need(V1, "V1"));
    OCX_TRY(need(Qx, "Qx"));
    OCX_TRY(need(Qy, "Qy"));
    OCX_TRY(need(ok, "ok"));
    OCX_TRY(need(C_out, "C_out"));
    OCX_TRY(need(Fz, "Fz"));
    if (C_out == C) return fail(OCX_E_ARG, "C_out must not alias C");
    OCX_TRY(check_rk_flags(flags, "the pc step"));
    OCX_TRY(check_rk_grid(nt, ny, nx));
    if (flux_face_i != -1 && (flux_face_i < 0 || flux_face_i > nx))
        return fail(OCX_E_ARG, "flux_face_i %d outside 0..%d", flux_face_i, nx);
    if (flux_face_i >= 0 && !fx_i) return fail(OCX_E_ARG, "fx_i is NULL with flux_face_i = %d", flux_face_i);
    const size_t wsb = rk3_layout(nullptr, nt, nz, ny, nx, nullptr);
    if (!workspace || workspace_bytes < wsb)
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed, %zu given", wsb, workspace ? workspace_bytes : 0);
    Rk3Ws w;
    rk3_layout(workspace, nt, nz, ny, nx, &w);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid = tile_grid(nt, nz, ny, nx);
    const unsigned edges = flags & 15u;
    RkArgs a{C, C, inflow, V0, V1, Qx, Qy, Fz, ok, nullptr, w.c1, nullptr, -1, nt, nz, ny, nx, dt, edges | CENTRED};
    {
        ProfScope _ps("pc_predictor", s);
        rk_stage<4>(a, SCH_C6, true, grid, s);
        OCX_TRY(launch_check("pc predictor"));
    }
    {
        ProfScope _ps("pc_corrector", s);
        a.cin = w.c1, a.cout = C_out, a.flags = flags;
        a.fxs = fx_i ? fx_i : nullptr;
        a.fi = fx_i ? flux_face_i : -1;
        rk_stage<5>(a, scheme_of(flags), vert, grid, s);
        OCX_TRY(launch_check("pc corrector"));
    }
    return OCX_OK;
}

int ocx_faces_h_f64(const double* c, const double* Qx, const double* Qy, const uint8_t* ok, const double* ext,
                    int nt, int nz, int ny, int nx, unsigned flags, double* fx, double* fy, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(c, "c"));
    OCX_TRY(need(Qx, "Qx"));
    OCX_TRY(need(Qy, "Qy"));
    OCX_TRY(need(ok, "ok"));
    OCX_TRY(need(fx, "fx"));
    OCX_TRY(need(fy, "fy"));
    OCX_TRY(check_rk_flags(flags, "faces_h", true));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid = tile_grid(nt, nz, ny, nx), block(TX, TY);
    switch (scheme_of(flags)) {
        case SCH_C6: k_faces<SCH_C6, TX, TY><<<grid, block, 0, s>>>(c, ext, Qx, Qy, ok, fx, fy, nt, nz, ny, nx, flags); break;
        case SCH_UP3: k_faces<SCH_UP3, TX, TY><<<grid, block, 0, s>>>(c, ext, Qx, Qy, ok, fx, fy, nt, nz, ny, nx, flags); break;
        case SCH_UP1: k_faces<SCH_UP1, TX, TY><<<grid, block, 0, s>>>(c, ext, Qx, Qy, ok, fx, fy, nt, nz, ny, nx, flags); break;
        case SCH_JS: k_faces<SCH_JS, TX, TY><<<grid, block, 0, s>>>(c, ext, Qx, Qy, ok, fx, fy, nt, nz, ny, nx, flags); break;
        default: k_faces<SCH_Z, TX, TY><<<grid, block, 0, s>>>(c, ext, Qx, Qy, ok, fx, fy, nt, nz, ny, nx, flags);
    }
    return launch_check("faces_h");
}

int ocx_faces_f64(const double* c, const double* flux, const uint8_t* ok, long long outer, int n, int nf,
                  long long inner, unsigned flags, double* out, void* stream) {
    using namespace ocx;
    OCX_TRY(need(c, "c"));
    OCX_TRY(need(ok, "ok"));
    OCX_TRY(need(out, "out"));
    if (flags & ~(WENO_JS | UP3 | UP1 | CENTRED))
        return fail(OCX_E_ARG, "flags 0x%x: faces take only OCX_WENO_JS, OCX_UP3, OCX_UP1 or OCX_CENTRED", flags);
    OCX_TRY(check_rk_flags(flags, "faces", true));
    if (!(flags & CENTRED)) OCX_TRY(need(flux, "flux"));
    if (outer < 1 || inner < 1 || n < 1 || nf < 1)
        return fail(OCX_E_ARG, "shape (outer=%lld, n=%d, nf=%d, inner=%lld) must be positive", outer, n, nf, inner);
    if (outer * nf > (1LL << 40) / inner || outer * n > (1LL << 40) / inner)
        return fail(OCX_E_ARG, "shape (outer=%lld, n=%d, nf=%d, inner=%lld) too large", outer, n, nf, inner);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long total = outer * nf * inner;
    const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 1LL << 20);
    switch (scheme_of(flags)) {
        case SCH_C6: k_line_faces<SCH_C6><<<blocks, 256, 0, s>>>(c, flux, ok, outer, n, nf, inner, out); break;
        case SCH_UP3: k_line_faces<SCH_UP3><<<blocks, 256, 0, s>>>(c, flux, ok, outer, n, nf, inner, out); break;
        case SCH_UP1: k_line_faces<SCH_UP1><<<blocks, 256, 0, s>>>(c, flux, ok, outer, n, nf, inner, out); break;
        case SCH_JS: k_line_faces<SCH_JS><<<blocks, 256, 0, s>>>(c, flux, ok, outer, n, nf, inner, out); break;
        default: k_line_faces<SCH_Z><<<blocks, 256, 0, s>>>(c, flux, ok, outer, n, nf, inner, out);
    }
    return launch_check("faces");
}

int ocx_open_faces_f64(double* face, const double* c, const double* Q, const double* ext, int nt, int nz, int ny,
                       int nx, int axis, unsigned open_mask, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(face, "face"));
    OCX_TRY(need(c, "c"));
    OCX_TRY(need(Q, "Q"));
    if (axis != 0 && axis != 1) return fail(OCX_E_ARG, "axis %d: 0 (x faces) or 1 (y faces)", axis);
    if (open_mask & ~15u) return fail(OCX_E_ARG, "open mask 0x%x", open_mask);
    const int lo = axis == 0 ? !!(open_mask & OPEN_W) : !!(open_mask & OPEN_S);
    const int hi = axis == 0 ? !!(open_mask & OPEN_E) : !!(open_mask & OPEN_N);
    if (!lo && !hi) return OCX_OK;
    const long long total = (long long)nt * nz * (axis == 0 ? ny : nx) * 2;
    const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 1LL << 20);
    k_open_faces<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(face, c, Q, ext ? ext : c, nt, nz, ny, nx,
                                                                        axis, lo, hi);
    return launch_check("open_faces");
}

int ocx_upwind_faces_h_f64(const double* c, const double* Qx, const double* Qy, const double* ext, int nt, int nz,
                           int ny, int nx, unsigned flags, double* fx, double* fy, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(c, "c"));
    OCX_TRY(need(Qx, "Qx"));
    OCX_TRY(need(Qy, "Qy"));
    OCX_TRY(need(fx, "fx"));
    OCX_TRY(need(fy, "fy"));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t total = (size_t)nt * nz * ny * (nx + 1) + (size_t)nt * nz * (ny + 1) * nx;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 1u << 20);
    k_upwind_faces<<<blocks, 256, 0, s>>>(c, ext, Qx, Qy, fx, fy, nt, nz, ny, nx, flags);
    return launch_check("upwind_faces_h");
}

int ocx_rhs_f64(const double* c, const double* Qx, const double* Qy, const double* Fz, const uint8_t* ok,
                const double* ext, int nt, int nz, int ny, int nx, unsigned flags, int vert, double* out,
                void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(c, "c"));
    OCX_TRY(need(Qx, "Qx"));
    OCX_TRY(need(Qy, "Qy"));
    OCX_TRY(need(ok, "ok"));
    OCX_TRY(need(out, "out"));
    const bool up = flags & (UP3 | UP1 | CENTRED);
    if (vert || up) OCX_TRY(need(Fz, "Fz"));
    OCX_TRY(check_rk_flags(flags, "the RHS", true));
    OCX_TRY(check_rk_grid(nt, ny, nx));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const RkArgs a{c, c, ext, nullptr, nullptr, Qx, Qy, Fz, ok, nullptr, out, nullptr, -1, nt, nz, ny, nx, 0.0, flags};
    rk_stage<0>(a, scheme_of(flags), vert, tile_grid(nt, nz, ny, nx), s);
    return launch_check("rhs");
}

int ocx_fct_coefficients_f64(const double* c0, const double* cL, const double* V1, const double* Ax,
                             const double* Ay, const uint8_t* ok, double dt, int nt, int nz, int ny, int nx,
                             double* Cx, double* Cy, void* workspace, size_t workspace_bytes, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(c0, "c0"));
    OCX_TRY(need(cL, "cL"));
    OCX_TRY(need(V1, "V1"));
    OCX_TRY(need(Ax, "Ax"));
    OCX_TRY(need(Ay, "Ay"));
    OCX_TRY(need(ok, "ok"));
    OCX_TRY(need(Cx, "Cx"));
    OCX_TRY(need(Cy, "Cy"));
    const size_t S = (size_t)nt * nz * ny * nx;
    if (!workspace || workspace_bytes < 2 * S * sizeof(double))
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed", 2 * S * sizeof(double));
    double* r_in = static_cast<double*>(workspace);
    double* r_out = r_in + S;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid = tile_grid(nt, nz, ny, nx), block(TX, TY);
    k_fct_ratio<TX, TY><<<grid, block, 0, s>>>(c0, cL, V1, Ax, Ay, ok, r_in, r_out, nt, nz, ny, nx, dt);
    OCX_TRY(launch_check("fct ratios"));
    const size_t total = (size_t)nt * nz * ny * (nx + 1) + (size_t)nt * nz * (ny + 1) * nx;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 1u << 20);
    k_fct_face_coeff<<<blocks, 256, 0, s>>>(Ax, Ay, r_in, r_out, Cx, Cy, nt, nz, ny, nx);
    return launch_check("fct face coefficients");
}

int ocx_periodic_pad(const void* src, void* dst, size_t lead, int sy, int sx, int kind, int pad_y, int pad_x,
                     int elem, int crop, void* stream) {
    using namespace ocx;
    OCX_TRY(need(src, "src"));
    OCX_TRY(need(dst, "dst"));
    if (sy < 1 || sx < 1 || pad_x < 0 || pad_y < 0 || kind < 0 || kind > 2 || (elem != 8 && elem != 1))
        return fail(OCX_E_ARG, "bad pad arguments (sy=%d sx=%d kind=%d pad=%d,%d elem=%d)", sy, sx, kind, pad_y, pad_x,
                    elem);
    const int fx = kind == 1, fy = kind == 2;
    int dy, dx, ny_u = sy - fy, nx_u = sx - fx;
    if (crop) {
        dy = sy - 2 * pad_y;
        dx = sx - 2 * pad_x;
        if (dy < 1 || dx < 1) return fail(OCX_E_ARG, "crop of %dx%d by %d,%d is empty", sy, sx, pad_y, pad_x);
    } else {
        dy = sy + 2 * pad_y;
        dx = sx + 2 * pad_x;
        if ((pad_y && ny_u < 1) || (pad_x && nx_u < 1)) return fail(OCX_E_ARG, "nothing to wrap");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t total = lead * (size_t)dy * dx;
    if (total == 0) return OCX_OK;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 1u << 20);
    if (elem == 8)
        k_pad<double><<<blocks, 256, 0, s>>>(static_cast<const double*>(src), static_cast<double*>(dst), lead, sy, sx, dy,
                                             dx, pad_y, pad_x, ny_u, nx_u, crop);
    else
        k_pad<uint8_t><<<blocks, 256, 0, s>>>(static_cast<const uint8_t*>(src), static_cast<uint8_t*>(dst), lead, sy,
                                              sx, dy, dx, pad_y, pad_x, ny_u, nx_u, crop);
    return launch_check("periodic pad");
}

int ocx_profile_enable(int on) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof_on = on != 0;
    return OCX_OK;
}

int ocx_profile_collect(int max_kernels, const char** names, double* ms_total, long long* launches) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    std::vector<const char*> keys;
    std::vector<double> ms;
    std::vector<long long> cnt;
    for (auto& r : g_prof) {
        if (hipEventSynchronize(r.b) != hipSuccess) return fail(OCX_E_HIP, "profile event wait failed");
        float e = 0.f;
        if (hipEventElapsedTime(&e, r.a, r.b) != hipSuccess) return fail(OCX_E_HIP, "profile event time failed");
        size_t m = 0;
        while (m < keys.size() && std::strcmp(keys[m], r.name) != 0) ++m;
        if (m == keys.size()) {
            keys.push_back(r.name);
            ms.push_back(0.0);
            cnt.push_back(0);
        }
        ms[m] += e;
        cnt[m] += 1;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof.clear();
    const int n = (int)keys.size();
    for (int m = 0; m < n && m < max_kernels; ++m) {
        if (names) names[m] = keys[m];
        if (ms_total) ms_total[m] = ms[m];
        if (launches) launches[m] = cnt[m];
    }
    return n;
}

}  // extern "C"

struct ocx_batch {
    int ndom = 0, wpb = 0, wpl = 0, rch = 0, ncu = 0;   // wpl: tracers per limiter workgroup
    int pm = 0;            // member precisions: 1 FP64 only, 2 split only, 3 both
    int gpc[4] = {};       // resident workgroups per compute unit of stage 1, 2, 3 and the limiter
    ocx::BatchDomain* d_dom = nullptr;
    long long* d_pm = nullptr;
    long long* d_pl = nullptr;
    unsigned long long* d_ctr = nullptr;   // work-queue counters: NQ per kernel of the step
    long long tm = 0, tl = 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
};

namespace {

template <int PM>
void batch_kernels_pm(const void** k) {
    k[0] = reinterpret_cast<const void*>(&ocx::k_march_batch<1, PM>);
    k[1] = reinterpret_cast<const void*>(&ocx::k_march_batch<2, PM>);
    k[2] = reinterpret_cast<const void*>(&ocx::k_march_batch<3, PM>);
    k[3] = reinterpret_cast<const void*>(&ocx::k_limit_batch<PM>);
}

void batch_kernels(int pm, const void** k) {
    if (pm == 1) batch_kernels_pm<1>(k);
    else if (pm == 2) batch_kernels_pm<2>(k);
    else batch_kernels_pm<3>(k);
}

void batch_free(ocx_batch* b) {
    if (!b) return;
    for (int q = 0; q < 2; ++q) {
        if (b->exec[q]) (void)hipGraphExecDestroy(b->exec[q]);
        if (b->graph[q]) (void)hipGraphDestroy(b->graph[q]);
    }
    if (b->d_dom) (void)hipFree(b->d_dom);
    if (b->d_pm) (void)hipFree(b->d_pm);
    if (b->d_pl) (void)hipFree(b->d_pl);
    if (b->d_ctr) (void)hipFree(b->d_ctr);
    delete b;
}

}  // namespace

extern "C" {

int ocx_set_tuning(int rows_per_chunk, int waves_per_group) {
    if (rows_per_chunk < 0 || rows_per_chunk > (1 << 20) || waves_per_group < 0 || waves_per_group > ocx::MAXW)
        return fail(OCX_E_ARG, "tuning (rows_per_chunk=%d, waves_per_group=%d) out of range (waves <= %d)",
                    rows_per_chunk, waves_per_group, ocx::MAXW);
    g_rch.store(rows_per_chunk);
    g_wpb.store(waves_per_group);
    return OCX_OK;
}

int ocx_batch_create(const ocx_domain* domains, int ndomains, int rows_per_chunk, int waves_per_group,
                     int groups_per_cu, ocx_batch** out) {
    using namespace ocx;
    OCX_TRY(need(out, "out"));
    *out = nullptr;
    if (ndomains < 1 || !domains) return fail(OCX_E_ARG, "no domains");
    if (rows_per_chunk < 0 || waves_per_group < 0 || waves_per_group > MAXW || groups_per_cu < 0)
        return fail(OCX_E_ARG, "bad tuning (rows_per_chunk=%d, waves_per_group=%d, groups_per_cu=%d)",
                    rows_per_chunk, waves_per_group, groups_per_cu);
    const int rch = rows_per_chunk > 0 ? rows_per_chunk : DEFAULT_RCH;
    int maxnt = 1;
    for (int d = 0; d < ndomains; ++d) maxnt = std::max(maxnt, domains[d].nt);
    // two tracers per workgroup unless asked otherwise (the 64-domain batch measured 149.5 ms per step with two,
    // 150.7 with one)
    const int wg = waves_per_group > 0 ? waves_per_group : DEFAULT_BATCH_WPB;
    const int wpb = tracer_group(maxnt, wg), wpl = limit_group(maxnt, wg);
    std::vector<BatchDomain> h(ndomains);
    std::vector<long long> pm(ndomains + 1, 0), pl(ndomains + 1, 0);
    for (int d = 0; d < ndomains; ++d) {
        const ocx_domain& D = domains[d];
        OCX_TRY(check_shape(D.nt, D.nz, D.ny, D.nx));
        if (!D.C || !D.C_out || !D.V0 || !D.V1 || !D.Qx || !D.Qy || !D.ok || !D.workspace)
            return fail(OCX_E_ARG, "domain %d: a required pointer is NULL", d);
        if (D.C == D.C_out) return fail(OCX_E_ARG, "domain %d: C and C_out alias", d);
        if (D.flags & ~31u) return fail(OCX_E_ARG, "domain %d: flags 0x%x not valid in a batch", d, D.flags);
        if (D.precision != 0 && D.precision != 1) return fail(OCX_E_ARG, "domain %d: precision %d", d, D.precision);
        const bool split = D.precision == 1;
        const size_t need_b = split ? fct_layout<PrecSplit>(nullptr, D.nt, D.nz, D.ny, D.nx, nullptr)
                                    : fct_layout<PrecF64>(nullptr, D.nt, D.nz, D.ny, D.nx, nullptr);
        if (D.workspace_bytes < need_b)
            return fail(OCX_E_WORKSPACE, "domain %d: workspace of %zu bytes needed, %zu given", d, need_b,
                        D.workspace_bytes);
        FctWs w;
        if (split) fct_layout<PrecSplit>(D.workspace, D.nt, D.nz, D.ny, D.nx, &w);
        else fct_layout<PrecF64>(D.workspace, D.nt, D.nz, D.ny, D.nx, &w);
        BatchDomain& B = h[d];
        B = BatchDomain{};
        MarchArgs& a = B.m;
        a.ext = D.inflow;
        a.V0 = D.V0;
        a.V1 = D.V1;
        a.Qx = D.Qx;
        a.Qy = D.Qy;
        a.ok = D.ok;
        a.fx = w.fx;
        a.fy = w.fy;
        a.s_x = w.sx;
        a.s_y = w.sy;
        a.ax = w.fx;
        a.ay = w.fy;
        a.fxi = nullptr;
        a.flux_face_i = -1;
        a.nt = D.nt;
        a.n = D.nz;
        a.ny = D.ny;
        a.nx = D.nx;
        a.rch = rch;
        a.nsx = (D.nx + MS - 1) / MS;
        a.ncy = (D.ny + rch - 1) / rch;
        a.dt = D.dt;
        a.flags = D.flags;
        LimitArgs& l = B.l;
        l.cL = w.cl;
        l.ext = D.inflow;
        l.V0 = D.V0;
        l.V1 = D.V1;
        l.Qx = D.Qx;
        l.Qy = D.Qy;
        l.ok = D.ok;
        l.ax = w.fx;
        l.ay = w.fy;
        l.fxi = nullptr;
        l.flux_face_i = -1;
        l.gedge = nullptr;
        l.nt = D.nt;
        l.n = D.nz;
        l.ny = D.ny;
        l.nx = D.nx;
        l.rch = rch;
        l.nsx = (D.nx + LS - 1) / LS;
        l.ncy = (D.ny + rch - 1) / rch;
        l.dt = D.dt;
        l.flags = D.flags;
        B.C = D.C;
        B.C_out = D.C_out;
        B.c1 = w.c1;
        B.c2 = w.c2;
        B.cl = w.cl;
        B.split = split;
        B.js = (D.flags & WENO_JS) != 0;
        pm[d + 1] = pm[d] + march_blocks(a, wpb);
        pl[d + 1] = pl[d] + limit_blocks(l, wpl);
    }
    ocx_batch* b = new ocx_batch();
    b->ndom = ndomains;
    b->wpb = wpb;
    b->wpl = wpl;
    b->rch = rch;
    b->tm = pm[ndomains];
    b->tl = pl[ndomains];
    for (int d = 0; d < ndomains; ++d) b->pm |= domains[d].precision == 1 ? 2 : 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || b->ncu < 1) {
        batch_free(b);
        return fail(OCX_E_HIP, "cannot query the compute-unit count");
    }
    // the persistent grid: groups_per_cu workgroups per compute unit, or (0) as many as the kernel's register and
    // LDS use let stay resident
    const void* kern[4];
    batch_kernels(b->pm, kern);
    for (int q = 0; q < 4; ++q) {
        int n = groups_per_cu;
        const int threads = q < 3 ? 64 * wpb : 64 * wpl * LG;
        const size_t lds = q < 3 ? 0 : limit_lds_bytes(wpl);
        if (n <= 0 && hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern[q], threads, lds) != hipSuccess) {
            batch_free(b);
            return fail(OCX_E_HIP, "occupancy query failed");
        }
        b->gpc[q] = n > 0 ? n : 1;
    }
    if (hipMalloc(&b->d_dom, sizeof(BatchDomain) * ndomains) != hipSuccess ||
        hipMalloc(&b->d_pm, sizeof(long long) * (ndomains + 1)) != hipSuccess ||
        hipMalloc(&b->d_pl, sizeof(long long) * (ndomains + 1)) != hipSuccess ||
        hipMalloc(&b->d_ctr, sizeof(unsigned long long) * 4 * NQ) != hipSuccess) {
        batch_free(b);
        return fail(OCX_E_HIP, "batch table allocation failed");
    }
    if (hipMemcpy(b->d_dom, h.data(), sizeof(BatchDomain) * ndomains, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_pm, pm.data(), sizeof(long long) * (ndomains + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_pl, pl.data(), sizeof(long long) * (ndomains + 1), hipMemcpyHostToDevice) != hipSuccess) {
        batch_free(b);
        return fail(OCX_E_HIP, "batch table copy failed");
    }
    *out = b;
    return OCX_OK;
}

int ocx_batch_step(ocx_batch* b, int parity, void* stream) {
    using namespace ocx;
    OCX_TRY(need(b, "batch"));
    if (parity != 0 && parity != 1) return fail(OCX_E_ARG, "parity %d", parity);
    hipStream_t s = static_cast<hipStream_t>(stream);
    static const char* const names[4] = {"batch_stage1", "batch_stage2", "batch_stage3_split", "batch_limit_final"};
    const void* kern[4];
    batch_kernels(b->pm, kern);
    const dim3 block(64u * b->wpb);
    // the work-queue counters start at zero (a kernel, not a memset, so that the step can be graph-captured)
    k_zero_u64<<<1, 64, 0, s>>>(b->d_ctr, 4 * NQ);
    OCX_TRY(launch_check("batch queue reset"));
    for (int q = 0; q < 4; ++q) {
        // a resident grid taking items from the work queues
        const long long total = q < 3 ? b->tm : b->tl;
        const long long slots = (long long)b->ncu * b->gpc[q];
        const dim3 grid((unsigned)(total < slots ? total : slots));
        const long long* prefix = q < 3 ? b->d_pm : b->d_pl;
        unsigned long long* ctr = b->d_ctr + (size_t)q * NQ;
        void* args[] = {&b->d_dom, const_cast<long long**>(&prefix), &b->ndom, const_cast<long long*>(&total),
                        &parity, &ctr};
        ProfScope _ps(names[q], s);
        const dim3 blk = q < 3 ? block : dim3(64u * b->wpl * LG);
        const size_t lds = q < 3 ? 0 : limit_lds_bytes(b->wpl);
        OCX_TRY(hip_check(hipLaunchKernel(kern[q], grid, blk, args, lds, s), names[q]));
    }
    return launch_check("batch step");
}

int ocx_batch_graph_step(ocx_batch* b, int parity, void* stream) {
    OCX_TRY(need(b, "batch"));
    OCX_TRY(need(stream, "stream (graph capture needs a created stream)"));
    if (parity != 0 && parity != 1) return fail(OCX_E_ARG, "parity %d", parity);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!b->exec[parity]) {
        OCX_TRY(hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture"));
        const int rc = ocx_batch_step(b, parity, stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != OCX_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        OCX_TRY(hip_check(e, "end capture"));
        b->graph[parity] = g;
        OCX_TRY(hip_check(hipGraphInstantiate(&b->exec[parity], g, nullptr, nullptr, 0), "graph instantiate"));
    }
    return hip_check(hipGraphLaunch(b->exec[parity], s), "graph launch");
}

int ocx_batch_destroy(ocx_batch* b) {
    batch_free(b);
    return OCX_OK;
}

int ocx_diffuse_h_f64(const double* T, const double* Tw, const double* h, const uint8_t* act, const double* cx,
                      const double* cy, const uint8_t* okx, const uint8_t* oky, const double* V, int nt, int nz, int ny,
                      int nx, double kh, double dt, double* out, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    const void* need_p[] = {T, Tw, h, act, cx, cy, okx, oky, V, out};
    const char* need_n[] = {"T", "Tw", "h", "act", "cx", "cy", "okx", "oky", "V", "out"};
    for (int q = 0; q < 10; ++q) OCX_TRY(need(need_p[q], need_n[q]));
    if (out == T || out == Tw) return fail(OCX_E_ARG, "out must not alias T or Tw (neighbours are read)");
    if (nz > 65535) return fail(OCX_E_ARG, "nz = %d layers: at most 65535", nz);
    const dim3 grid((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nz), block(64, 4);
    hipStream_t s = static_cast<hipStream_t>(stream);
    ProfScope _ps("diffuse_h", s);
    k_diffuse_h<<<grid, block, 0, s>>>(T, Tw, h, act, cx, cy, okx, oky, V, nt, nz, ny, nx, kh, dt, out);
    return launch_check("diffuse_h");
}

// cp, then (columns of at most 64 levels) the lower coefficients and pivots of the register-resident solve
size_t ocx_vertical_implicit_workspace_bytes(int n, long long columns) {
    if (n < 1 || columns < 1) return 0;
    const size_t level = align256((size_t)n * (size_t)columns * sizeof(double));
    return n <= 64 ? 3 * level : level;
}

int ocx_vertical_implicit_f64(const double* u, const double* dz, const double* nu, const double* r_bed, int r_layers,
                              const double* tau_top, int nt, int n, long long columns, double dt, double* out,
                              void* workspace, size_t workspace_bytes, void* stream) {
    using namespace ocx;
    if (nt < 1 || n < 1 || columns < 1) return fail(OCX_E_ARG, "shape (nt=%d, n=%d, columns=%lld) must be positive",
                                                     nt, n, columns);
    OCX_TRY(need(u, "u"));
    OCX_TRY(need(dz, "dz"));
    OCX_TRY(need(nu, "nu"));
    OCX_TRY(need(out, "out"));
    if (r_bed && r_layers != 1 && r_layers != n)
        return fail(OCX_E_ARG, "r_layers = %d: 1 (a bed rate on layer 0) or n = %d (a rate per layer)", r_layers, n);
    const size_t wsb = ocx_vertical_implicit_workspace_bytes(n, columns);
    if (!workspace || workspace_bytes < wsb)
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed, %zu given", wsb, workspace ? workspace_bytes : 0);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t level = align256((size_t)n * (size_t)columns * sizeof(double));
    double* cp = static_cast<double*>(workspace);
    // the workspace holds lw and mm whenever n <= 64; the register solve's 1-D grid needs fewer than 2^32 blocks
    const bool reg = n <= 64 && ((columns + 63) / 64) * ((nt + 3) / 4) < (1ll << 32);
    double* lw = reg ? reinterpret_cast<double*>(static_cast<char*>(workspace) + level) : nullptr;
    double* mm = reg ? reinterpret_cast<double*>(static_cast<char*>(workspace) + 2 * level) : nullptr;
    const unsigned gc = (unsigned)std::min<long long>((columns + 255) / 256, 1 << 16);
    {
        ProfScope _ps("vertical_coefficients", s);
        k_thomas_cp<<<gc, 256, 0, s>>>(dz, nu, r_bed, r_layers, n, columns, dt, cp, lw, mm);
    }
    OCX_TRY(launch_check("vertical_implicit coefficients"));
    ProfScope _ps("vertical_solve", s);
    if (reg) {
        const int fy = std::min(nt, 4), groups = (nt + fy - 1) / fy;
        const unsigned grid = (unsigned)(((columns + 63) / 64) * groups);
        const dim3 block(64, fy);
#define OCX_SOLVE_R(NB) \
    k_thomas_solve_r<NB><<<grid, block, 0, s>>>(u, dz, tau_top, nt, n, columns, dt, cp, lw, mm, out, groups)
        switch ((n + 7) / 8) {     // the smallest multiple of 8 levels that holds the column
            case 1: OCX_SOLVE_R(8); break;
            case 2: OCX_SOLVE_R(16); break;
            case 3: OCX_SOLVE_R(24); break;
            case 4: OCX_SOLVE_R(32); break;
            case 5: OCX_SOLVE_R(40); break;
            case 6: OCX_SOLVE_R(48); break;
            case 7: OCX_SOLVE_R(56); break;
            default: OCX_SOLVE_R(64); break;
        }
#undef OCX_SOLVE_R
    } else {
        const unsigned gs = (unsigned)std::min<long long>((columns + 255) / 256, 1 << 16);
        k_thomas_solve<<<gs, 256, 0, s>>>(u, dz, nu, r_bed, r_layers, tau_top, nt, n, columns, dt, cp, out);
    }
    return launch_check("vertical_implicit");
}

int ocx_selftest_div_f64(const double* x, const double* y, long long n, double* out, void* stream) {
    OCX_TRY(need(x, "x"));
    OCX_TRY(need(y, "y"));
    OCX_TRY(need(out, "out"));
    if (n < 1 || n > (1LL << 30)) return fail(OCX_E_ARG, "n = %lld outside 1 .. 2^30", n);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    ocx::k_selftest_div<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(x, y, (size_t)n, out);
    return launch_check("selftest div");
}

int ocx_mix_num_f64(const double* c, const double* c2, const double* V1, int nz, int ny, int nx, double dt,
                    int accumulate, double* out, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(1, nz, ny, nx));
    OCX_TRY(need(c, "c"));
    OCX_TRY(need(c2, "c2"));
    OCX_TRY(need(V1, "V1"));
    OCX_TRY(need(out, "out"));
    if (!(dt > 0)) return fail(OCX_E_ARG, "dt = %g must be positive", dt);
    const size_t cells = (size_t)nz * ny * nx;
    const unsigned blocks = (unsigned)std::min<size_t>((cells + 255) / 256, 65536);
    k_mix_num<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(c, c2, V1, dt, accumulate != 0, out, cells);
    return launch_check("mix_num");
}

size_t ocx_content_workspace_bytes(int nt) {
    return nt > 0 ? (size_t)nt * ocx::CONTENT_PARTS * 2 * sizeof(double) : 0;
}

int ocx_content_f64(const double* c, const double* V, int nt, int nz, int ny, int nx, double* out, void* workspace,
                    size_t workspace_bytes, void* stream) {
    using namespace ocx;
    OCX_TRY(check_shape(nt, nz, ny, nx));
    OCX_TRY(need(c, "c"));
    OCX_TRY(need(V, "V"));
    OCX_TRY(need(out, "out"));
    const size_t wsb = ocx_content_workspace_bytes(nt);
    if (!workspace || workspace_bytes < wsb)
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed, %zu given", wsb, workspace ? workspace_bytes : 0);
    if (nt > 65535) return fail(OCX_E_ARG, "nt=%d: at most 65535 tracers per content call", nt);
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* part = static_cast<double*>(workspace);
    k_content_part<<<dim3(CONTENT_PARTS, nt), 256, 0, s>>>(c, V, part, (size_t)nz * ny * nx);
    OCX_TRY(launch_check("content partials"));
    k_content_fold<<<nt, CONTENT_PARTS, 0, s>>>(part, out);
    return launch_check("content");
}

}  // extern "C"
