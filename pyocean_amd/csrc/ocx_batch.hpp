// Batched FCT step: many independent domains (ensemble members or regional domains of uneven nx x ny x nz, FP64
// or split precision) advanced by one persistent launch per stage.
//
// Every domain contributes the logical blocks its own launch would have (march_blocks / limit_blocks); the batch
// concatenates them in domain order (prefix sums) into one item space.  A resident grid of workgroups takes items
// from eight work queues, one per XCD, each a contiguous eighth of the item space: a workgroup draws from the queue
// of the XCD it runs on (workgroups are dispatched to the XCDs round-robin) and, once that queue is empty, from the
// others.  Items therefore run in item order on each XCD, as the hardware dispatches a one-domain launch, so the
// items in flight on an XCD are neighbours (the tracer groups and layers of one strip of one domain) and share the
// per-layer fields and the limiter's k +- 1 planes through that XCD's L2; the queues balance the load at the end.
// Each item runs exactly the per-domain code (march_block / limit_block) with the domain's own sizes and
// precision, so a domain's result does not depend on what else is in the batch.
//
// The queue counters are vector atomics issued by one lane; the next item is claimed while the current one runs.
// Every workgroup leaves once all eight queues are exhausted.
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

constexpr int NQ = 8;       // work queues (one per XCD)

// Work queue state of one workgroup, held by its first lane: the queue it draws from and the claim in flight.
struct WorkQueue {
    unsigned long long* ctr;   // NQ counters, zero at launch
    long long total;
    int q;                     // current queue
    unsigned long long raw;    // counter value returned by the last claim on queue q

    __device__ long long lo(int k) const { return total * k / NQ; }
    __device__ long long hi(int k) const { return total * (k + 1) / NQ; }
    // issue a claim on the current queue (its result is looked at by take())
    __device__ void claim() { raw = atomicAdd(&ctr[q], 1ull); }
    // the item of the last claim, or the next item of another queue, or total when every queue is empty
    __device__ long long take() {
        long long v = lo(q) + (long long)raw;
        int tried = 0;
        while (v >= hi(q) && ++tried < NQ) {
            q = (q + 1) % NQ;
            v = lo(q) + (long long)atomicAdd(&ctr[q], 1ull);
        }
        return v < hi(q) ? v : total;
    }
};

// Runs body(item) for the items this workgroup takes; every wave of the workgroup runs every iteration.
template <class Body>
__device__ __forceinline__ void drain(unsigned long long* ctr, long long total, Body&& body) {
    __shared__ long long s_item;
    WorkQueue wq{ctr, total, (int)(blockIdx.x % NQ), 0ull};
    const bool leader = threadIdx.x == 0;
    if (leader) {
        wq.claim();
        s_item = wq.take();
    }
    __syncthreads();
    long long it = s_item;
    while (it < total) {
        if (leader) wq.claim();                // the next item, claimed while this one runs
        body(it);
        __syncthreads();                       // every wave has read s_item and finished the item
        if (leader) s_item = wq.take();
        __syncthreads();
        it = s_item;
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// PM: which member precisions the batch holds (1: FP64 only, 2: split only, 3: both); a batch of one precision
// compiles only that path, so its kernels need fewer registers and keep more waves resident
template <int STAGE, int PM>
__global__ __launch_bounds__(64 * MAXW) void k_march_batch(const BatchDomain* __restrict__ dom,
                                                           const long long* __restrict__ prefix, int ndom,
                                                           long long total, int parity,
                                                           unsigned long long* __restrict__ ctr) {
    __shared__ double rbuf[MAXW][MRB];
    __shared__ uint8_t obuf[MAXW][MRB];
    const int wv = wave_id(), wpb = blockDim.x >> 6;
    drain(ctr, total, [&](long long it) {
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
    });
}

// blockDim.x = 64 * wpb * LG; dynamic LDS limit_lds_bytes(wpb)
template <int PM>
__global__ __launch_bounds__(64 * LMAXW * LG) void k_limit_batch(const BatchDomain* __restrict__ dom,
                                                                const long long* __restrict__ prefix, int ndom,
                                                                long long total, int parity,
                                                                unsigned long long* __restrict__ ctr) {
    extern __shared__ double limit_lds[];
    Bnd* lds = reinterpret_cast<Bnd*>(limit_lds);
    const int wv = wave_id(), wpb = (blockDim.x >> 6) / LG;
    drain(ctr, total, [&](long long it) {
        const int d = find_domain(prefix, ndom, it);
        const BatchDomain& D = dom[d];
        LimitArgs l = D.l;
        l.C = parity ? D.C_out : D.C;
        l.out = parity ? const_cast<double*>(D.C) : D.C_out;
        const long long L = it - prefix[d];
        if (PM == 2 || (PM == 3 && D.split))
            limit_block<PrecSplit>(l, L, wv, wpb, lds);
        else
            limit_block<PrecF64>(l, L, wv, wpb, lds);
    });
}

}  // namespace ocx
