// common.hpp -- shared definitions of the gfx950 Laplace unwrapping library.
//
// Every HIP source is compiled with -ffp-contract=off: the only fused
// multiply-adds are the explicit fmaf/fma calls below, so each kernel's
// floating-point operation order is fixed and can be restated on the CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cuphu.h"

namespace cuphu {

// ── errors: thrown inside the library, converted to status codes at the ABI ──
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define CUPHU_HIP_CHECK(expr)                                                         \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw ::cuphu::Error(CUPHU_ERR_HIP, std::string(hipGetErrorString(e_)) +  \
                                                    " (" #expr ") at " __FILE__ ":" + \
                                                    std::to_string(__LINE__));        \
    } while (0)

#define CUPHU_REQUIRE(cond, msg)                                                      \
    do {                                                                              \
        if (!(cond)) throw ::cuphu::Error(CUPHU_ERR_INVALID, (msg));                  \
    } while (0)

// Runs f() and maps exceptions to status codes + the thread-local message.
template <class F> int guarded(F &&f) {
    try {
        return f();
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return CUPHU_ERR_NOMEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return CUPHU_ERR_INTERNAL;
    } catch (...) {
        set_last_error("unknown error");
        return CUPHU_ERR_INTERNAL;
    }
}

// ── device memory RAII ──────────────────────────────────────────────────────
template <class T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) CUPHU_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T *get() const { return p; }
};

// Pinned host buffer (allocated pinned; never registers existing memory).
template <class T> struct PinBuf {
    T *p = nullptr;
    size_t n = 0;
    PinBuf() = default;
    explicit PinBuf(size_t count) { alloc(count); }
    ~PinBuf() { if (p) (void)hipHostFree(p); }
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    void alloc(size_t count) {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = count;
        if (count) CUPHU_HIP_CHECK(hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault));
    }
};

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        CUPHU_HIP_CHECK(hipGetDevice(&prev));
        if (dev != prev) CUPHU_HIP_CHECK(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

// A named host range for `rocprofv3 --marker-trace` (roctx; costs a call
// into the roctx library when no profiler is attached).
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }
inline int pitch_of(int ncol) { return (ncol + 63) / 64 * 64; }

// ── constants (cuphu_laplace.cu:59-62, cuphu_util.cu:23) ──────────────────
constexpr float kTwoPiF = 6.28318530717958647692f;
constexpr float kInvTwoPiF = 0.15915494309189534561f;
constexpr float kEpsF = 1e-12f;
constexpr double kTwoPiD = 6.28318530717958647692;
constexpr double kEpsD = 1e-12;
constexpr float kPiF = 3.14159265358979323846f;
constexpr int kLargeShort = 32000;
constexpr uint32_t kInvalidLabel = 0xFFFFFFFFu;

// ── device helpers ──────────────────────────────────────────────────────────
__device__ __forceinline__ float wrap_pi(float x) {
    return __builtin_fmaf(-kTwoPiF, __builtin_rintf(x * kInvTwoPiF), x);
}
__device__ __forceinline__ double wrap_pi_d(double x) {
    return __builtin_fma(-kTwoPiD, __builtin_rint(x / kTwoPiD), x);
}
__device__ __forceinline__ float wrap_cycles(float d) { return d - __builtin_floorf(d + 0.5f); }
__device__ __forceinline__ float wrap_to_2pi(float p) {
    float t = __builtin_floorf(p / kTwoPiF);
    return p - kTwoPiF * t;
}
// Wrapped phase of one complex sample: atan2 in FP64 rounded once, +2pi if
// negative (cuphu_solver.cpp:1707-1712), then p - 2pi*floor(p/2pi)
// (cuphu_util.cu:33-46).
__device__ __forceinline__ float phase_of(float re, float im) {
    float a = (float)atan2((double)im, (double)re);
    if (a < 0.0f) a += 6.28318530717958648f;
    return wrap_to_2pi(a);
}
__device__ __forceinline__ float arc_weight_f(short s) {
    if (s <= 0 || s >= kLargeShort) return 0.0f;
    return 1.0f / (float)s;
}
__device__ __forceinline__ double arc_weight_d(short s) {
    if (s <= 0 || s >= kLargeShort) return 0.0;
    return 1.0 / (double)s;
}

// Precision traits: FP32 mode keeps vectors/weights in float, FP64 mode in
// double.  fmaT / wrapT / weightT dispatch to the right precision.
template <class T> struct Prec;
template <> struct Prec<float> {
    using T2 = float2;
    using T4 = float4;
    static __device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    static __device__ __forceinline__ float wrap(float x) { return wrap_pi(x); }
    static __device__ __forceinline__ float weight(short s) { return arc_weight_f(s); }
    static __device__ __forceinline__ float sqrtT(float x) { return __builtin_sqrtf(x); }
    static constexpr float eps = kEpsF;
};
template <> struct Prec<double> {
    using T2 = double2;
    static __device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
    static __device__ __forceinline__ double wrap(double x) { return wrap_pi_d(x); }
    static __device__ __forceinline__ double weight(short s) { return arc_weight_d(s); }
    static __device__ __forceinline__ double sqrtT(double x) { return __builtin_sqrt(x); }
    static constexpr double eps = kEpsD;
};

// Statistical-cost constants for one instance (cuphu_cost.cu:41-59), passed
// by value / per instance instead of process-global __constant__ symbols.
struct CostConsts {
    double rho0, thresh, rhopow, sigsqrhoconst, sigsqcorr, nsc, nsc2, costscale, defomax;
    int sqmin;
    int pad_;
};
CostConsts make_cost_consts(const CuPhuParams &p, double ncorrlooks);

}  // namespace cuphu
