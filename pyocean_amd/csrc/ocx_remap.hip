// C ABI of the vertical regrid and remap kernels (declarations and contracts: include/ocx.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/ocx.h"
#include "ocx_host.hpp"
#include "ocx_remap.hpp"

namespace {

using namespace ocx_host;
using namespace ocx::rmp;

#ifndef REMAP_G
#define REMAP_G 2
#endif
#ifndef REMAP_STAGE
#define REMAP_STAGE 1
#endif
#ifndef REMAP_SPLIT_ROWS
#define REMAP_SPLIT_ROWS 1
#endif
constexpr int REMAP_LDS_LAYERS = 128;   // 64 KB of target thicknesses per workgroup

__global__ void k_set_bad(unsigned long long* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = ~0ull;
}

}  // namespace

extern "C" {

int ocx_remap_columns_f64(const double* h_src, const double* f, const double* h_tgt, const uint8_t* wet, int n0,
                          int n1, long long columns, int nf, int scheme, double* out, void* stream) {
    if (n0 < 1 || n1 < 1 || columns < 1 || nf < 1)
        return fail(OCX_E_ARG, "shape (n0=%d, n1=%d, columns=%lld, nf=%d) must be positive", n0, n1, columns, nf);
    if (nf > 65535 * 4) return fail(OCX_E_ARG, "nf = %d fields: at most %d per call", nf, 65535 * 4);
    if (scheme < PCM || scheme > PPM_H4) return fail(OCX_E_ARG, "scheme = %d: 0 (PCM), 1 (PLM) or 2 (PPM_H4)", scheme);
    const void* p[] = {h_src, f, h_tgt, wet, out};
    const char* nm[] = {"h_src", "f", "h_tgt", "wet", "out"};
    for (int q = 0; q < 5; ++q) OCX_TRY(need(p[q], nm[q]));
    if (out == f || out == h_src || out == h_tgt) return fail(OCX_E_ARG, "out must not alias the inputs");
    RemapArgs r{h_src, f, h_tgt, wet, out, n0, n1, nf, scheme, columns};
    hipStream_t s = static_cast<hipStream_t>(stream);
    ProfScope _ps("remap", s);
    // one field per thread for a single field; otherwise groups of REMAP_G fields per thread, up to REMAP_FW_MAX
    // groups (waves) of the same 64 columns per workgroup, balanced over the workgroups of a column block; with at
    // most REMAP_LDS_LAYERS target layers the workgroup stages the 64 columns' target thicknesses in LDS once for
    // its waves.  Workgroups of at most 4 waves (launch bound 256) ran the 10-field remap at 1024^2 x 50 in 8.7 ms
    // (8.5 with the rows split so that no wave idles); 5 waves under a launch bound of 512 took 16.4 ms.
    const int g = nf == 1 ? 1 : REMAP_G;
    const int groups = (nf + g - 1) / g;
    const int nby = (groups + REMAP_FW_MAX - 1) / REMAP_FW_MAX;
    const int fw = (groups + nby - 1) / nby;
    const unsigned ncb = (unsigned)((columns + 63) / 64);
    const bool staged = REMAP_STAGE && n1 <= REMAP_LDS_LAYERS;
    const size_t lds = staged ? (size_t)n1 * 64 * sizeof(double) : 0;
    if (g == 1) {
        const dim3 grid(ncb, (unsigned)nby), block(64, fw);
        if (staged) k_remap<true><<<grid, block, lds, s>>>(r);
        else k_remap<false><<<grid, block, 0, s>>>(r);
    } else {
        // the groups as rows of workgroups of fw and fw - 1 waves, so that no wave idles (5 groups: a row of 3 and
        // a row of 2, each row of a column block one workgroup), each kind its own launch
        const int full = REMAP_SPLIT_ROWS ? groups - nby * (fw - 1) : nby;    // rows of fw waves
        const int rest = nby - full;                                           // rows of fw - 1 waves
        auto launch = [&](int rows, int width, int goff) {
            if (rows <= 0) return;
            const dim3 grid(ncb, (unsigned)rows), block(64, width);
            if (staged) k_remap_g<REMAP_G, true><<<grid, block, lds, s>>>(r, goff);
            else k_remap_g<REMAP_G, false><<<grid, block, 0, s>>>(r, goff);
        };
        launch(full, fw, 0);
        launch(rest, fw - 1, full * fw);
    }
    return launch_check("remap_columns");
}

size_t ocx_regrid_workspace_bytes(void) { return 256; }

int ocx_regrid_f64(const double* H, const double* eta, const uint8_t* active, const double* h, int n,
                   long long columns, int rule, const double* dz, double h_min, double* h_new, uint8_t* wet,
                   void* workspace, size_t workspace_bytes, long long* bad_column, void* stream) {
    if (n < 1 || columns < 1) return fail(OCX_E_ARG, "shape (n=%d, columns=%lld) must be positive", n, columns);
    if (rule != 0 && rule != 1) return fail(OCX_E_ARG, "rule = %d: 0 (z*) or 1 (sigma)", rule);
    const void* p[] = {H, eta, active, h, dz, h_new};
    const char* nm[] = {"H", "eta", "active", "h", "dz", "h_new"};
    for (int q = 0; q < 6; ++q) OCX_TRY(need(p[q], nm[q]));
    if (h_new == h) return fail(OCX_E_ARG, "h_new must not alias h");
    if (!workspace || workspace_bytes < ocx_regrid_workspace_bytes())
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed", ocx_regrid_workspace_bytes());
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long* bad = static_cast<unsigned long long*>(workspace);
    k_set_bad<<<1, 64, 0, s>>>(bad);
    OCX_TRY(launch_check("regrid reset"));
    {
        ProfScope _ps("regrid", s);
        k_regrid<<<(unsigned)((columns + 255) / 256), 256, 0, s>>>(H, eta, active, h, n, columns, rule, dz, h_min,
                                                                   h_new, wet, bad);
    }
    OCX_TRY(launch_check("regrid"));
    if (!bad_column) return OCX_OK;
    unsigned long long b = 0;
    OCX_TRY(hip_check(hipMemcpyAsync(&b, bad, sizeof b, hipMemcpyDeviceToHost, s), "regrid flag"));
    OCX_TRY(hip_check(hipStreamSynchronize(s), "regrid sync"));
    *bad_column = b == ~0ull ? -1 : (long long)b;
    if (b == ~0ull) return OCX_OK;
    return fail(OCX_E_RANGE, "wet column %llu: %s", b,
                rule == 0 ? "z* needs finite H > 0 and D = H + eta > 0" : "sigma needs finite H and D = H + eta > 0");
}

size_t ocx_regrid_hybrid_workspace_bytes(int n, long long columns) {
    if (n < 1 || columns < 1) return 0;
    return align256((size_t)(6 * (long long)n + 3) * (size_t)columns * sizeof(double)) + 256;
}

int ocx_regrid_hybrid_f64(const double* H, const double* eta, const uint8_t* active, const double* h, const double* rho,
                          int n, long long columns, const double* dz_top, const double* dsigma,
                          const double* rho_target, const double* max_h_top, double d_sigma, double d_hybrid,
                          double h_min, int keep, double* h_new, uint8_t* wet, void* workspace,
                          size_t workspace_bytes, long long* bad_column, void* stream) {
    if (n < 1 || columns < 1) return fail(OCX_E_ARG, "shape (n=%d, columns=%lld) must be positive", n, columns);
    const void* p[] = {H, eta, active, h, rho, dz_top, dsigma, rho_target, max_h_top, h_new};
    const char* nm[] = {"H", "eta", "active", "h", "rho", "dz_top", "dsigma", "rho_target", "max_h_top", "h_new"};
    for (int q = 0; q < 10; ++q) OCX_TRY(need(p[q], nm[q]));
    if (h_new == h) return fail(OCX_E_ARG, "h_new must not alias h");
    const size_t wsb = ocx_regrid_hybrid_workspace_bytes(n, columns);
    if (!workspace || workspace_bytes < wsb)
        return fail(OCX_E_WORKSPACE, "workspace of %zu bytes needed, %zu given", wsb, workspace ? workspace_bytes : 0);
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* base = static_cast<char*>(workspace);
    unsigned long long* bad = reinterpret_cast<unsigned long long*>(base);
    HybridArgs r{H, eta, h, rho, dz_top, dsigma, rho_target, max_h_top, active, d_sigma, d_hybrid, h_min,
                 keep != 0, n, columns, h_new, wet, reinterpret_cast<double*>(base + 256), bad};
    k_set_bad<<<1, 64, 0, s>>>(bad);
    OCX_TRY(launch_check("regrid reset"));
    {
        ProfScope _ps("regrid_hybrid", s);
        k_regrid_hybrid<<<(unsigned)((columns + 127) / 128), 128, 0, s>>>(r);
    }
    OCX_TRY(launch_check("regrid_hybrid"));
    if (!bad_column) return OCX_OK;
    unsigned long long b = 0;
    OCX_TRY(hip_check(hipMemcpyAsync(&b, bad, sizeof b, hipMemcpyDeviceToHost, s), "regrid flag"));
    OCX_TRY(hip_check(hipStreamSynchronize(s), "regrid sync"));
    *bad_column = b == ~0ull ? -1 : (long long)b;
    if (b == ~0ull) return OCX_OK;
    return fail(OCX_E_RANGE, "wet column %llu: the hybrid rule needs finite densities, D = H + eta > 0, and H > 0 "
                "unless it has a shallow sigma zone (d_hybrid > 0)", b);
}

int ocx_regrid_density_f64(int eos, const double* T, const double* S, const double* p, double p_scalar,
                           const double* linear, long long n, double* rho, void* stream) {
    if (n < 1) return fail(OCX_E_ARG, "n = %lld cells must be positive", n);
    if (eos != EOS_LINEAR && eos != EOS_WRIGHT_FULL) return fail(OCX_E_ARG, "eos = %d: 0 (linear) or 1 (wright_full)", eos);
    OCX_TRY(need(T, "T"));
    OCX_TRY(need(rho, "rho"));
    if (eos == EOS_WRIGHT_FULL && !S) return fail(OCX_E_ARG, "wright_full EOS needs salinity");
    if (eos == EOS_LINEAR && !linear) return fail(OCX_E_ARG, "the linear EOS needs its 5 coefficients");
    DensityArgs a{T, S, eos == EOS_WRIGHT_FULL ? p : nullptr, p_scalar, {0, 0, 0, 0, 0}, eos, n, rho};
    if (eos == EOS_LINEAR)
        for (int q = 0; q < 5; ++q) a.lin[q] = linear[q];
    hipStream_t s = static_cast<hipStream_t>(stream);
    ProfScope _ps("regrid_density", s);
    const long long blocks = std::min<long long>((n + 255) / 256, 65536);
    k_regrid_density<<<(unsigned)blocks, 256, 0, s>>>(a);
    return launch_check("regrid_density");
}

}  // extern "C"
