// C ABI of the ALE layer-transport kernels (declarations and contracts: include/ocx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/ocx.h"
#include "ocx_ale.hpp"
#include "ocx_host.hpp"

namespace {

using namespace ocx_host;
using namespace ocx::ale;

constexpr int BS = 256;

unsigned blocks_for(long long n) { return (unsigned)((n + BS - 1) / BS); }

int check_ale_shape(int n, int ny, int nx) {
    if (n < 1 || ny < 1 || nx < 1) return fail(OCX_E_ARG, "shape (n=%d, ny=%d, nx=%d) must be positive", n, ny, nx);
    if ((long long)(ny + 1) * (nx + 1) >= (1LL << 31))
        return fail(OCX_E_ARG, "plane ny=%d x nx=%d too large (face columns must fit 31 bits)", ny, nx);
    if ((long long)n * (ny + 1) * (nx + 1) >= (1LL << 40))
        return fail(OCX_E_ARG, "shape (n=%d, ny=%d, nx=%d) too large", n, ny, nx);
    return OCX_OK;
}

int check_axis(int axis) {
    if (axis != 0 && axis != 1) return fail(OCX_E_ARG, "axis = %d: 0 (u-faces) or 1 (v-faces)", axis);
    return OCX_OK;
}

long long face_cols(int axis, int ny, int nx) {
    return axis == 0 ? (long long)ny * (nx + 1) : (long long)(ny + 1) * nx;
}

// PPM workspace: the two edge stacks, the Newton state per face column, the convergence flags
struct PpmWs {
    double *e_lo, *e_hi, *du, *lo, *hi, *err, *dsum;
    unsigned* notconv;
};

constexpr int PPM_MAX_ITER = 1000;

size_t ppm_layout(void* base, int n, int ny, int nx, int axis, int max_iter, PpmWs* w) {
    Carver cv(base);
    const size_t cells = (size_t)n * ny * nx;
    const size_t cols = (size_t)face_cols(axis, ny, nx);
    PpmWs t;
    t.e_lo = cv.take<double>(cells);
    t.e_hi = cv.take<double>(cells);
    t.du = cv.take<double>(cols);
    t.lo = cv.take<double>(cols);
    t.hi = cv.take<double>(cols);
    t.err = cv.take<double>(cols);
    t.dsum = cv.take<double>(cols);
    t.notconv = cv.take<unsigned>((size_t)std::max(max_iter, 1));
    if (w) *w = t;
    return cv.used;
}

template <int AX>
void launch_edges(const double* h, const uint8_t* act, double h_min, int n, int ny, int nx, double* lo, double* hi,
                  hipStream_t s) {
    Geo<AX> g{n, ny, nx};
    k_ppm_edges<AX><<<blocks_for((long long)ny * nx), BS, 0, s>>>(h, act, h_min, g, lo, hi);
}

template <int AX>
int ppm_impl(const double* h, const double* ut, const double* L, const double* Qb, const double* qe, const double* qw,
             const uint8_t* act, const double* dxt, int n, int ny, int nx, double dt, int max_iter, double* out,
             const PpmWs& w, hipStream_t s) {
    PpmArgs<AX> p;
    p.h = h;
    p.e_lo = w.e_lo;
    p.e_hi = w.e_hi;
    p.ut = ut;
    p.L = L;
    p.Qb = Qb;
    p.qe = qe;
    p.qw = qw;
    p.dxt = dxt;
    p.act = act;
    p.dt = dt;
    p.g = Geo<AX>{n, ny, nx};
    p.du = w.du;
    p.lo = w.lo;
    p.hi = w.hi;
    p.err = w.err;
    p.dsum = w.dsum;
    p.notconv = w.notconv;
    p.out = out;
    const unsigned gb = blocks_for(face_cols(AX, ny, nx));
    {
        ProfScope _ps("ppm_edges", s);
        launch_edges<AX>(h, act, H_PPM_MIN, n, ny, nx, w.e_lo, w.e_hi, s);
    }
    OCX_TRY(launch_check("ppm edges"));
    k_fill_u32<<<blocks_for(std::max(max_iter, 1)), BS, 0, s>>>(w.notconv, std::max(max_iter, 1), 0u);
    OCX_TRY(launch_check("ppm flags"));
    {
        ProfScope _ps("ppm_init", s);
        k_ppm_init<AX><<<gb, BS, 0, s>>>(p);
    }
    OCX_TRY(launch_check("ppm init"));
    for (int m = 0; m < max_iter; ++m) {
        ProfScope _ps("ppm_newton", s);
        k_ppm_iter<AX><<<gb, BS, 0, s>>>(p, m);
    }
    OCX_TRY(launch_check("ppm newton"));
    {
        ProfScope _ps("ppm_final", s);
        k_ppm_final<AX><<<gb, BS, 0, s>>>(p);
    }
    return launch_check("ppm final");
}

// limiter workspace: r, room (cell stacks), the cut per face column (the larger axis), counters and flags
struct OutWs {
    double *r, *room, *delta;
    unsigned long long *count, *bad_face, *key, *idx;
    unsigned* any;
};

size_t out_layout(void* base, int n, int ny, int nx, OutWs* w) {
    Carver cv(base);
    const size_t cells = (size_t)n * ny * nx;
    OutWs t;
    t.r = cv.take<double>(cells);
    t.room = cv.take<double>(cells);
    t.delta = cv.take<double>((size_t)std::max(face_cols(0, ny, nx), face_cols(1, ny, nx)));
    t.count = cv.take<unsigned long long>(1);
    t.bad_face = cv.take<unsigned long long>(1);
    t.key = cv.take<unsigned long long>(1);
    t.idx = cv.take<unsigned long long>(1);
    t.any = cv.take<unsigned>(1);
    if (w) *w = t;
    return cv.used;
}

__global__ void k_out_reset(unsigned long long* count, unsigned* any, unsigned long long* bad_face) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *count = 0ull;
        *any = 0u;
        *bad_face = ~0ull;
    }
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

struct OverResult {
    unsigned long long count;
    unsigned any;
    unsigned long long bad_face;
};

// one over-the-bound test (and r), then the counters to the host (synchronous)
int over_pass(OutArgs& o, const OutWs& w, double factor, int first, hipStream_t s, OverResult* res) {
    k_out_reset<<<1, 64, 0, s>>>(w.count, w.any, w.bad_face);
    OCX_TRY(launch_check("outflow reset"));
    {
        ProfScope _ps("outflow_over", s);
        k_out_over<<<blocks_for((long long)o.n * o.ny * o.nx), BS, 0, s>>>(o, factor, first);
    }
    OCX_TRY(launch_check("outflow over"));
    OCX_TRY(hip_check(hipMemcpyAsync(&res->count, w.count, sizeof(unsigned long long), hipMemcpyDeviceToHost, s),
                      "outflow count"));
    OCX_TRY(hip_check(hipMemcpyAsync(&res->any, w.any, sizeof(unsigned), hipMemcpyDeviceToHost, s), "outflow flag"));
    return hip_check(hipStreamSynchronize(s), "outflow sync");
}

template <int AX>
int cut_and_give(OutArgs& o, int wrap, hipStream_t s) {
    Geo<AX> g{o.n, o.ny, o.nx};
    const unsigned gb = blocks_for(g.cols());
    {
        ProfScope _ps("outflow_cut", s);
        k_out_cut<AX><<<gb, BS, 0, s>>>(o, wrap);
    }
    OCX_TRY(launch_check("outflow cut"));
    {
        ProfScope _ps("outflow_room", s);
        k_out_room<<<blocks_for((long long)o.n * o.ny * o.nx), BS, 0, s>>>(o);
    }
    OCX_TRY(launch_check("outflow room"));
    {
        ProfScope _ps("outflow_give", s);
        k_out_give<AX><<<gb, BS, 0, s>>>(o, wrap);
    }
    return launch_check("outflow give");
}

int check_lag(const double* h, const double* area, int n, int ny, int nx, const double* out) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    OCX_TRY(need(h, "h"));
    OCX_TRY(need(area, "area"));
    OCX_TRY(need(out, "out"));
    return OCX_OK;
}

}  // namespace

extern "C" {

int ocx_transport_velocity_f64(const double* u_new, const double* u_old, const uint8_t* last_wet, const double* w,
                               const double* L, const double* D, int n, long long columns, double theta,
                               double theta_bc, double* ut, double* Q_bt, void* stream) {
    if (n < 1 || columns < 1) return fail(OCX_E_ARG, "shape (n=%d, columns=%lld) must be positive", n, columns);
    const void* p[] = {u_new, u_old, last_wet, w, L, D, ut, Q_bt};
    const char* nm[] = {"u_new", "u_old", "last_wet", "w", "L", "D", "ut", "Q_bt"};
    for (int q = 0; q < 8; ++q) OCX_TRY(need(p[q], nm[q]));
    hipStream_t s = static_cast<hipStream_t>(stream);
    ProfScope _ps("transport_velocity", s);
    k_transport_velocity<<<blocks_for(columns), BS, 0, s>>>(u_new, u_old, last_wet, w, L, D, theta, theta_bc, n,
                                                            columns, ut, Q_bt);
    return launch_check("transport_velocity");
}

int ocx_face_transport_thickness_f64(const double* h, const double* H, const uint8_t* active, int n, int ny, int nx,
                                     int axis, int clip, double* q_pos, double* q_neg, void* stream) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    OCX_TRY(check_axis(axis));
    OCX_TRY(need(h, "h"));
    OCX_TRY(need(active, "active"));
    OCX_TRY(need(q_pos, "q_pos"));
    OCX_TRY(need(q_neg, "q_neg"));
    if (clip) OCX_TRY(need(H, "H"));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned gb = blocks_for(face_cols(axis, ny, nx));
    ProfScope _ps("face_thickness", s);
    if (axis == 0)
        k_face_thickness<0><<<gb, BS, 0, s>>>(h, H, active, clip, Geo<0>{n, ny, nx}, q_pos, q_neg);
    else
        k_face_thickness<1><<<gb, BS, 0, s>>>(h, H, active, clip, Geo<1>{n, ny, nx}, q_pos, q_neg);
    return launch_check("face_transport_thickness");
}

int ocx_ppm_edges_f64(const double* h, const uint8_t* active, int n, int ny, int nx, int axis, double h_min,
                      double* e_lo, double* e_hi, void* stream) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    OCX_TRY(check_axis(axis));
    OCX_TRY(need(h, "h"));
    OCX_TRY(need(active, "active"));
    OCX_TRY(need(e_lo, "e_lo"));
    OCX_TRY(need(e_hi, "e_hi"));
    if (e_lo == h || e_hi == h || e_lo == e_hi) return fail(OCX_E_ARG, "e_lo, e_hi and h must be distinct");
    hipStream_t s = static_cast<hipStream_t>(stream);
    ProfScope _ps("ppm_edges", s);
    if (axis == 0)
        launch_edges<0>(h, active, h_min, n, ny, nx, e_lo, e_hi, s);
    else
        launch_edges<1>(h, active, h_min, n, ny, nx, e_lo, e_hi, s);
    return launch_check("ppm_edges");
}

size_t ocx_ppm_transport_workspace_bytes(int n, int ny, int nx, int axis, int max_iter) {
    if (check_ale_shape(n, ny, nx) != OCX_OK || check_axis(axis) != OCX_OK) return 0;
    if (max_iter < 0 || max_iter > PPM_MAX_ITER) return 0;
    return ppm_layout(nullptr, n, ny, nx, axis, max_iter, nullptr);
}

int ocx_ppm_transport_f64(const double* h, const double* ut, const double* L, const double* Q_bt, const double* q_pos,
                          const double* q_neg, const uint8_t* active, const double* d_t, int n, int ny, int nx,
                          int axis, double dt, int max_iter, double* out, void* workspace, size_t workspace_bytes,
                          void* stream) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    OCX_TRY(check_axis(axis));
    const void* p[] = {h, ut, L, Q_bt, q_pos, q_neg, active, d_t, out};
    const char* nm[] = {"h", "ut", "L", "Q_bt", "q_pos", "q_neg", "active", "d_t", "out"};
    for (int q = 0; q < 9; ++q) OCX_TRY(need(p[q], nm[q]));
    if (max_iter < 0 || max_iter > PPM_MAX_ITER)
        return fail(OCX_E_ARG, "max_iter = %d outside 0..%d", max_iter, PPM_MAX_ITER);
    if (!(dt > 0)) return fail(OCX_E_ARG, "dt = %g must be positive", dt);
    if (out == ut || out == q_pos || out == q_neg) return fail(OCX_E_ARG, "out must not alias ut, q_pos or q_neg");
    const size_t wsb = ppm_layout(nullptr, n, ny, nx, axis, max_iter, nullptr);
    if (!workspace || workspace_bytes < wsb)
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed, %zu given", wsb, workspace ? workspace_bytes : 0);
    PpmWs w;
    ppm_layout(workspace, n, ny, nx, axis, max_iter, &w);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (axis == 0) return ppm_impl<0>(h, ut, L, Q_bt, q_pos, q_neg, active, d_t, n, ny, nx, dt, max_iter, out, w, s);
    return ppm_impl<1>(h, ut, L, Q_bt, q_pos, q_neg, active, d_t, n, ny, nx, dt, max_iter, out, w, s);
}

int ocx_upwind_transport_f64(const double* ut, const double* L, const double* Q_bt, const double* q_pos,
                             const double* q_neg, int n, int ny, int nx, int axis, double* out, void* stream) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    OCX_TRY(check_axis(axis));
    const void* p[] = {ut, L, Q_bt, q_pos, q_neg, out};
    const char* nm[] = {"ut", "L", "Q_bt", "q_pos", "q_neg", "out"};
    for (int q = 0; q < 6; ++q) OCX_TRY(need(p[q], nm[q]));
    if (out == ut || out == q_pos || out == q_neg) return fail(OCX_E_ARG, "out must not alias ut, q_pos or q_neg");
    hipStream_t s = static_cast<hipStream_t>(stream);
    ProfScope _ps("upwind_transport", s);
    const unsigned gb = blocks_for(face_cols(axis, ny, nx));
    if (axis == 0)
        k_upwind_transport<0><<<gb, BS, 0, s>>>(ut, L, Q_bt, q_pos, q_neg, Geo<0>{n, ny, nx}, out);
    else
        k_upwind_transport<1><<<gb, BS, 0, s>>>(ut, L, Q_bt, q_pos, q_neg, Geo<1>{n, ny, nx}, out);
    return launch_check("upwind_transport");
}

int ocx_ppm_transport_iterations(const void* workspace, int n, int ny, int nx, int axis, int max_iter,
                                 void* stream) {
    if (check_ale_shape(n, ny, nx) != OCX_OK || check_axis(axis) != OCX_OK) return OCX_E_ARG;
    if (max_iter < 0 || max_iter > PPM_MAX_ITER)
        return fail(OCX_E_ARG, "max_iter = %d outside 0..%d", max_iter, PPM_MAX_ITER);
    if (!workspace) return fail(OCX_E_ARG, "workspace is NULL");
    if (max_iter == 0) return 0;
    PpmWs w;
    ppm_layout(const_cast<void*>(workspace), n, ny, nx, axis, max_iter, &w);
    static thread_local unsigned flags[PPM_MAX_ITER];
    hipStream_t s = static_cast<hipStream_t>(stream);
    OCX_TRY(hip_check(hipMemcpyAsync(flags, w.notconv, sizeof(unsigned) * max_iter, hipMemcpyDeviceToHost, s),
                      "ppm flags"));
    OCX_TRY(hip_check(hipStreamSynchronize(s), "ppm flags sync"));
    int it = 1;
    while (it < max_iter && flags[it - 1] != 0u) ++it;
    return it;
}

size_t ocx_limit_outflow_workspace_bytes(int n, int ny, int nx) {
    if (check_ale_shape(n, ny, nx) != OCX_OK) return 0;
    return out_layout(nullptr, n, ny, nx, nullptr);
}

int ocx_limit_outflow_f64(const double* h, double* Qx, double* Qy, const double* area, const uint8_t* active, int n,
                          int ny, int nx, double dt, double h_floor, int max_pass, int periodic_x, int periodic_y,
                          int check_only, void* workspace, size_t workspace_bytes, long long* n_limited, int* status,
                          long long* where, double* value, void* stream) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    const void* p[] = {h, Qx, Qy, area, active, n_limited, status};
    const char* nm[] = {"h", "Qx", "Qy", "area", "active", "n_limited", "status"};
    for (int q = 0; q < 7; ++q) OCX_TRY(need(p[q], nm[q]));
    if (!(dt > 0)) return fail(OCX_E_ARG, "dt = %g must be positive", dt);
    if (max_pass < 0) return fail(OCX_E_ARG, "max_pass = %d must be >= 0", max_pass);
    const size_t wsb = out_layout(nullptr, n, ny, nx, nullptr);
    if (!workspace || workspace_bytes < wsb)
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed, %zu given", wsb, workspace ? workspace_bytes : 0);
    OutWs w;
    out_layout(workspace, n, ny, nx, &w);
    hipStream_t s = static_cast<hipStream_t>(stream);
    OutArgs o;
    o.h = h;
    o.area = area;
    o.act = active;
    o.Qx = Qx;
    o.Qy = Qy;
    o.n = n;
    o.ny = ny;
    o.nx = nx;
    o.dt = dt;
    o.h_floor = h_floor;
    o.r = w.r;
    o.room = w.room;
    o.delta = w.delta;
    o.count = w.count;
    o.any = w.any;
    o.bad_face = w.bad_face;
    *n_limited = 0;
    if (where) where[0] = where[1] = where[2] = -1;
    OverResult res{};
    OCX_TRY(over_pass(o, w, 1.0, 1, s, &res));
    if (!res.any) {
        *status = OCX_OUTFLOW_NOOP;
        return OCX_OK;
    }
    if (check_only) {
        *status = OCX_OUTFLOW_EXCEEDS;
        return OCX_OK;
    }
    const double factor = 1.0 + 1e-12;
    long long nl = 0;
    for (int pass = 0; pass < max_pass; ++pass) {
        nl += (long long)res.count;
        for (int ax = 0; ax < 2; ++ax) {
            OCX_TRY(ax == 0 ? cut_and_give<0>(o, periodic_x != 0, s) : cut_and_give<1>(o, periodic_y != 0, s));
            unsigned long long bad = 0;
            OCX_TRY(hip_check(hipMemcpyAsync(&bad, w.bad_face, sizeof bad, hipMemcpyDeviceToHost, s), "outflow bad"));
            OCX_TRY(hip_check(hipStreamSynchronize(s), "outflow sync"));
            if (bad != ~0ull) {
                *n_limited = nl;
                *status = OCX_OUTFLOW_EMPTY_DONOR;
                const long long nf = ax == 0 ? nx + 1 : nx;
                if (where) {
                    where[0] = ax;
                    where[1] = (long long)bad / nf;
                    where[2] = (long long)bad % nf;
                }
                return fail(OCX_E_RANGE, "limit_outflow: %s-face (%lld, %lld) must carry its cut out of a donor column "
                            "that holds no water — barotropic CFL", ax == 0 ? "u" : "v", (long long)bad / nf,
                            (long long)bad % nf);
            }
        }
        OCX_TRY(over_pass(o, w, factor, 0, s, &res));
        if (!res.any) {
            *n_limited = nl;
            *status = OCX_OUTFLOW_OK;
            return OCX_OK;
        }
    }
    *n_limited = nl;
    // still over after max_pass: accepted when no layer would fall below -h_floor, else the reference raises
    const long long cells = (long long)n * ny * nx;
    k_set_u64<<<1, 64, 0, s>>>(w.key, ~0ull);
    k_set_u64<<<1, 64, 0, s>>>(w.idx, ~0ull);
    k_hdag_min<<<blocks_for(cells), BS, 0, s>>>(o, w.key);
    k_hdag_where<<<blocks_for(cells), BS, 0, s>>>(o, w.key, w.idx);
    OCX_TRY(launch_check("outflow h_dag"));
    unsigned long long key = 0, idx = 0;
    OCX_TRY(hip_check(hipMemcpyAsync(&key, w.key, sizeof key, hipMemcpyDeviceToHost, s), "h_dag min"));
    OCX_TRY(hip_check(hipMemcpyAsync(&idx, w.idx, sizeof idx, hipMemcpyDeviceToHost, s), "h_dag where"));
    OCX_TRY(hip_check(hipStreamSynchronize(s), "h_dag sync"));
    const unsigned long long bits = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
    double mn;
    std::memcpy(&mn, &bits, sizeof mn);
    if (value) *value = mn;
    if (where) {
        where[0] = (long long)(idx / ((unsigned long long)ny * nx));
        where[1] = (long long)(idx / nx % ny);
        where[2] = (long long)(idx % nx);
    }
    if (mn >= -h_floor) {
        *status = OCX_OUTFLOW_ACCEPTED;
        return OCX_OK;
    }
    *status = OCX_OUTFLOW_NEGATIVE;
    return fail(OCX_E_RANGE, "limit_outflow: layer %lld of cell (j=%lld, i=%lld) would reach h = %.3g m after %d "
                "passes (dt = %g s): the column cannot supply the free-surface flux — barotropic CFL",
                (long long)(idx / ((unsigned long long)ny * nx)), (long long)(idx / nx % ny), (long long)(idx % nx),
                mn, max_pass, dt);
}

size_t ocx_lagrangian_workspace_bytes(void) { return 256; }

int ocx_lagrangian_thickness_f64(const double* h, const double* divq, const double* Qx, const double* Qy,
                                 const double* area, const uint8_t* active, int n, int ny, int nx, double dt,
                                 double tol, double* out, double* V0, double* V1, uint8_t* ok, void* workspace,
                                 size_t workspace_bytes, long long* bad_index, void* stream) {
    OCX_TRY(check_lag(h, area, n, ny, nx, out));
    if (!divq && !(Qx && Qy)) return fail(OCX_E_ARG, "divq or both Qx and Qy must be given");
    if (divq && (Qx || Qy)) return fail(OCX_E_ARG, "give divq or Qx and Qy, not both");
    if (V0 || V1 || ok) {
        if (!(V0 && V1 && ok && active)) return fail(OCX_E_ARG, "V0, V1, ok and active go together");
        if (V0 == out || V1 == out || V0 == h || V1 == h) return fail(OCX_E_ARG, "V0 / V1 must not alias h or out");
    }
    if (!workspace || workspace_bytes < ocx_lagrangian_workspace_bytes())
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed", ocx_lagrangian_workspace_bytes());
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* bad = static_cast<unsigned long long*>(workspace);
    k_set_u64<<<1, 64, 0, s>>>(bad, ~0ull);
    OCX_TRY(launch_check("lagrangian reset"));
    {
        ProfScope _ps("lagrangian", s);
        k_lagrangian<<<blocks_for((long long)n * ny * nx), BS, 0, s>>>(h, divq, Qx, Qy, area, active, n, ny, nx, dt,
                                                                       tol, out, V0, V1, ok, bad);
    }
    OCX_TRY(launch_check("lagrangian"));
    if (!bad_index) return OCX_OK;
    unsigned long long b = 0;
    OCX_TRY(hip_check(hipMemcpyAsync(&b, bad, sizeof b, hipMemcpyDeviceToHost, s), "lagrangian flag"));
    OCX_TRY(hip_check(hipStreamSynchronize(s), "lagrangian sync"));
    *bad_index = b == ~0ull ? -1 : (long long)b;
    if (b == ~0ull) return OCX_OK;
    const long long k = (long long)(b / ((unsigned long long)ny * nx)), j = (long long)(b / nx % ny),
                    i = (long long)(b % nx);
    return fail(OCX_E_RANGE, "layer %lld of cell (j=%lld, i=%lld) emptied in the Lagrangian step (reduce dt)", k, j, i);
}

int ocx_open_face_transports_f64(double* Qx, double* Qy, const double* h, const double* u_old, const double* v_old,
                                 const double* u_new, const double* v_new, const double* open_Du,
                                 const double* open_Qu, const double* open_Dv, const double* open_Qv, const double* Lu,
                                 const double* Lv, int n, int ny, int nx, unsigned open_mask, double theta_b,
                                 void* stream) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    const void* p[] = {Qx, Qy, h, u_old, v_old, u_new, v_new, open_Du, open_Qu, open_Dv, open_Qv, Lu, Lv};
    const char* nm[] = {"Qx", "Qy", "h", "u_old", "v_old", "u_new", "v_new", "open_Du", "open_Qu", "open_Dv",
                        "open_Qv", "Lu", "Lv"};
    for (int q = 0; q < 13; ++q) OCX_TRY(need(p[q], nm[q]));
    if (open_mask & ~15u) return fail(OCX_E_ARG, "open_mask %u: bits 1 (west), 2 (east), 4 (south), 8 (north)", open_mask);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long NX1 = nx + 1, NY1 = ny + 1;
    const long long pu = (long long)ny * NX1, pv = NY1 * nx, pc = (long long)ny * nx;
    ProfScope _ps("open_face_transport", s);
    for (int edge = 0; edge < 4; ++edge) {
        if (!(open_mask & (1u << edge))) continue;
        EdgeArgs e{};
        e.n = n;
        e.tb = theta_b;
        if (edge < 2) {                                  // west / east: the u-faces of column 0 / nx
            const long long fc = edge == 0 ? 0 : nx, cc = edge == 0 ? 0 : nx - 1;
            e.Q = Qx;
            e.un = u_new;
            e.uo = u_old;
            e.q0 = fc; e.qk = pu; e.qf = NX1;
            e.h = h;
            e.h0 = cc; e.hk = pc; e.hf = nx;
            e.D = open_Du; e.Qo = open_Qu; e.L = Lu;
            e.f0 = fc; e.ff = NX1;
            e.m = ny;
        } else {                                         // south / north: the v-faces of row 0 / ny
            const long long fr = edge == 2 ? 0 : ny, cr = edge == 2 ? 0 : ny - 1;
            e.Q = Qy;
            e.un = v_new;
            e.uo = v_old;
            e.q0 = fr * nx; e.qk = pv; e.qf = 1;
            e.h = h;
            e.h0 = cr * nx; e.hk = pc; e.hf = 1;
            e.D = open_Dv; e.Qo = open_Qv; e.L = Lv;
            e.f0 = fr * nx; e.ff = 1;
            e.m = nx;
        }
        k_open_face_transport<<<blocks_for(e.m), BS, 0, s>>>(e);
        OCX_TRY(launch_check("open_face_transport"));
    }
    return OCX_OK;
}

size_t ocx_freshwater_workspace_bytes(void) { return 256; }

int ocx_freshwater_volumes_f64(const double* V, const double* dV, double dv_scale, const double* floor_cols,
                               double floor, const double* area, int n, int ny, int nx, double* out, double* V_out,
                               void* workspace, size_t workspace_bytes, long long* n_limited, void* stream) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    OCX_TRY(need(V, "V"));
    OCX_TRY(need(dV, "dV"));
    OCX_TRY(need(out, "out"));
    if (out == V) return fail(OCX_E_ARG, "out must not alias V");
    if ((area == nullptr) != (V_out == nullptr)) return fail(OCX_E_ARG, "area and V_out go together");
    if (V_out && (V_out == V || V_out == out)) return fail(OCX_E_ARG, "V_out must not alias V or out");
    if (!workspace || workspace_bytes < ocx_freshwater_workspace_bytes())
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed", ocx_freshwater_workspace_bytes());
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* cnt = static_cast<unsigned long long*>(workspace);
    k_set_u64<<<1, 64, 0, s>>>(cnt, 0ull);
    OCX_TRY(launch_check("freshwater reset"));
    const long long M = (long long)ny * nx;
    {
        ProfScope _ps("freshwater_volumes", s);
        k_freshwater_volumes<<<blocks_for(M), BS, 0, s>>>(V, dV, dv_scale, floor_cols, floor, area, n, M, out, V_out,
                                                          cnt);
    }
    OCX_TRY(launch_check("freshwater_volumes"));
    if (!n_limited) return OCX_OK;
    unsigned long long b = 0;
    OCX_TRY(hip_check(hipMemcpyAsync(&b, cnt, sizeof b, hipMemcpyDeviceToHost, s), "freshwater count"));
    OCX_TRY(hip_check(hipStreamSynchronize(s), "freshwater sync"));
    *n_limited = (long long)b;
    return OCX_OK;
}

int ocx_apply_freshwater_f64(double* stack, int nt, int n, int ny, int nx, int t_index, int t2_index, int s_index,
                             const double* V_adv, const double* V_fw, const double* fw_area_dt, double dt,
                             double* var_surf, void* workspace, size_t workspace_bytes, long long* n_limited,
                             void* stream) {
    OCX_TRY(check_ale_shape(n, ny, nx));
    if (nt < 1) return fail(OCX_E_ARG, "nt = %d must be positive", nt);
    const void* p[] = {stack, V_adv, V_fw, fw_area_dt};
    const char* nm[] = {"stack", "V_adv", "V_fw", "fw_area_dt"};
    for (int q = 0; q < 4; ++q) OCX_TRY(need(p[q], nm[q]));
    if (t_index >= nt || t2_index >= nt || s_index >= nt)
        return fail(OCX_E_ARG, "member indices (T %d, T2 %d, S %d) must lie below nt = %d", t_index, t2_index,
                    s_index, nt);
    if (s_index >= 0 && (s_index == t_index || s_index == t2_index))
        return fail(OCX_E_ARG, "S (%d) cannot be T or T2", s_index);
    if (var_surf && s_index < 0) return fail(OCX_E_ARG, "var_surf needs the index of S");
    if (!(dt > 0)) return fail(OCX_E_ARG, "dt = %g must be positive", dt);
    if (!workspace || workspace_bytes < ocx_freshwater_workspace_bytes())
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed", ocx_freshwater_workspace_bytes());
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* cnt = static_cast<unsigned long long*>(workspace);
    k_set_u64<<<1, 64, 0, s>>>(cnt, 0ull);
    OCX_TRY(launch_check("freshwater reset"));
    const long long M = (long long)ny * nx;
    FreshArgs a{stack, V_adv, V_fw, fw_area_dt, var_surf, nt, n, t_index, t2_index, s_index, M, dt, cnt};
    {
        ProfScope _ps("apply_freshwater", s);
        k_apply_freshwater<<<blocks_for(M), BS, 0, s>>>(a);
    }
    OCX_TRY(launch_check("apply_freshwater"));
    if (!n_limited) return OCX_OK;
    unsigned long long b = 0;
    OCX_TRY(hip_check(hipMemcpyAsync(&b, cnt, sizeof b, hipMemcpyDeviceToHost, s), "freshwater count"));
    OCX_TRY(hip_check(hipStreamSynchronize(s), "freshwater sync"));
    *n_limited = (long long)b;
    return OCX_OK;
}

}  // extern "C"
