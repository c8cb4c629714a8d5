// Host-side helpers shared by the translation units of the C ABI: the per-thread error text behind
// ocx_last_error(), argument checks, workspace carving and the optional per-kernel event timing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ocx.h"

namespace ocx_host {

extern thread_local std::string g_err;

inline int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

inline int hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) return fail(OCX_E_HIP, "%s: %s", what, hipGetErrorString(e));
    return OCX_OK;
}

#define OCX_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != OCX_OK) return _rc; \
    } while (0)

inline int need(const void* p, const char* name) {
    if (!p) return fail(OCX_E_ARG, "%s is NULL", name);
    return OCX_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Carves a workspace into consecutive 256-byte-aligned pieces (a null base only measures).
struct Carver {
    char* base;
    size_t used = 0;
    explicit Carver(void* b) : base(static_cast<char*>(b)) {}
    template <typename T>
    T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
        used += align256(count * sizeof(T));
        return p;
    }
};

inline int launch_check(const char* what) { return hip_check(hipGetLastError(), what); }

// Optional per-kernel timing: a hipEvent pair on the launch stream around every kernel launch while enabled.
struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};
extern std::mutex g_prof_mu;
extern std::vector<ProfRec> g_prof;
extern bool g_prof_on;

struct ProfScope {
    hipStream_t s;
    bool on;
    ProfRec rec{};
    ProfScope(const char* name, hipStream_t st) : s(st), on(g_prof_on) {
        if (!on) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (st && hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
            on = false;                // no event nodes inside a captured graph
            return;
        }
        rec.name = name;
        if (hipEventCreate(&rec.a) != hipSuccess || hipEventCreate(&rec.b) != hipSuccess) {
            on = false;
            return;
        }
        (void)hipEventRecord(rec.a, s);
    }
    ~ProfScope() {
        if (!on) return;
        (void)hipEventRecord(rec.b, s);
        std::lock_guard<std::mutex> g(g_prof_mu);
        g_prof.push_back(rec);
    }
};

}  // namespace ocx_host
