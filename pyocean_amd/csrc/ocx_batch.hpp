// Batched FCT step: many independent domains (ensemble members or regional domains of uneven nx x ny x nz, FP64
// or split precision) advanced by one persistent launch per stage.
//
// Every domain contributes the logical blocks its own launch would have (march_blocks / limit_blocks); the batch
// concatenates them in domain order (prefix sums) and a fixed grid of G workgroups, as many as stay resident, walks
// them grid-stride: workgroup g takes items g, g + G, g + 2G, ... with g numbered XCD-contiguously, so the items in
// flight at any moment are neighbours in item order (the tracer groups and layers of one strip of one domain) and
// share the per-layer fields and the limiter's k +- 1 planes through the XCD's L2, as a one-domain launch does.
// Each item runs exactly the per-domain code (march_block / limit_block) with the domain's own sizes and
// precision, so a domain's result does not depend on what else is in the batch.
#pragma once
#include "ocx_limit.hpp"

namespace ocx {

struct BatchDomain {
    MarchArgs m;           // shared fields of the three stages (cin/cout/fxi set per stage)
    LimitArgs l;           // limiter (out set per parity)
    const double* C;       // the two state buffers: step parity 0 reads C and writes C_out, parity 1 the reverse
    double* C_out;
    void* c1;              // stage buffers from the domain's workspace
    void* c2;
    double* cl;
    int split;             // 1: split precision
    int js;                // 1: WENO5-JS
};

__device__ __forceinline__ int find_domain(const long long* prefix, int ndom, long long item) {
    int lo = 0, hi = ndom;                     // prefix[0] = 0 <= item < prefix[ndom]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= item) lo = mid; else hi = mid;
    }
    return lo;
}

// PM: which member precisions the batch holds (1: FP64 only, 2: split only, 3: both); a batch of one precision
// compiles only that path, so its kernels need fewer registers and keep more waves resident
template <int STAGE, int PM>
__global__ __launch_bounds__(64 * MAXW) void k_march_batch(const BatchDomain* __restrict__ dom,
                                                           const long long* __restrict__ prefix, int ndom,
                                                           long long total, int parity) {
    __shared__ double rbuf[MAXW][MRB];
    __shared__ uint8_t obuf[MAXW][MRB];
    const int wv = wave_id(), wpb = blockDim.x >> 6;
    for (long long it = xcd_logical(blockIdx.x, gridDim.x); it < total; it += gridDim.x) {
        const int d = find_domain(prefix, ndom, it);
        const BatchDomain& D = dom[d];
        MarchArgs a = D.m;
        const double* Cin = parity ? D.C_out : D.C;
        a.C = Cin;
        if (STAGE == 1) {
            a.cin = Cin;
            a.cout = D.c1;
        } else if (STAGE == 2) {
            a.cin = D.c1;
            a.cout = D.c2;
        } else {
            a.cin = D.c2;
            a.cout = D.cl;
        }
        const long long L = it - prefix[d];
        if (PM == 2 || (PM == 3 && D.split)) {
            if (D.js)
                march_block<STAGE, false, false, PrecSplit>(a, L, wv, wpb, reinterpret_cast<float*>(rbuf[wv]), obuf[wv]);
            else
                march_block<STAGE, true, false, PrecSplit>(a, L, wv, wpb, reinterpret_cast<float*>(rbuf[wv]), obuf[wv]);
        } else {
            if (D.js)
                march_block<STAGE, false, false, PrecF64>(a, L, wv, wpb, rbuf[wv], obuf[wv]);
            else
                march_block<STAGE, true, false, PrecF64>(a, L, wv, wpb, rbuf[wv], obuf[wv]);
        }
    }
}

template <int PM>
__global__ __launch_bounds__(64 * MAXW) void k_limit_batch(const BatchDomain* __restrict__ dom,
                                                           const long long* __restrict__ prefix, int ndom,
                                                           long long total, int parity) {
    const int wv = wave_id(), wpb = blockDim.x >> 6;
    for (long long it = xcd_logical(blockIdx.x, gridDim.x); it < total; it += gridDim.x) {
        const int d = find_domain(prefix, ndom, it);
        const BatchDomain& D = dom[d];
        LimitArgs l = D.l;
        l.C = parity ? D.C_out : D.C;
        l.out = parity ? const_cast<double*>(D.C) : D.C_out;
        const long long L = it - prefix[d];
        if (PM == 2 || (PM == 3 && D.split))
            limit_block<PrecSplit>(l, L, wv, wpb);
        else
            limit_block<PrecF64>(l, L, wv, wpb);
    }
}

}  // namespace ocx
