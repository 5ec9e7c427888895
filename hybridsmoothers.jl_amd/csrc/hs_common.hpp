// hs_common.hpp -- error handling, device buffers and small device helpers
// shared by every translation unit of libhsmooth.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/hsmooth.h"

namespace hs {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(int code, const std::string &msg, int64_t aux = 0);
int last_code();

struct Error {
    int code;
};

#define HS_CHECK_HIP(expr)                                                             \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            int _c = (_e == hipErrorOutOfMemory) ? HS_ENOMEM : HS_EDEVICE;            \
            ::hs::set_error(_c, std::string(#expr) + ": " + hipGetErrorString(_e));   \
            throw ::hs::Error{_c};                                                     \
        }                                                                              \
    } while (0)

#define HS_REQUIRE(cond, code, msg)                                                    \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::hs::set_error((code), (msg));                                            \
            throw ::hs::Error{(code)};                                                 \
        }                                                                              \
    } while (0)

// Runs `body` and converts exceptions into status codes for the C ABI.
template <class F>
int guard(F &&body) {
    try {
        body();
        return HS_OK;
    } catch (const Error &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error(HS_ENOMEM, "host allocation failed");
        return HS_ENOMEM;
    } catch (...) {
        set_error(HS_EDEVICE, "unexpected exception");
        return HS_EDEVICE;
    }
}

// ---------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------
extern std::atomic<int64_t> g_device_bytes;

// Owning device buffer (hipMalloc / hipFree), movable, zero-size safe.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) {
            HS_CHECK_HIP(hipMalloc(&p, count * sizeof(T)));
            g_device_bytes += (int64_t)(count * sizeof(T));
        }
    }
    void release() {
        if (p) {
            (void)hipFree(p);
            g_device_bytes -= (int64_t)(n * sizeof(T));
        }
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
    T *get() const { return p; }
};

// Process-unique id of a handle (cached graphs are keyed by ids, never by
// addresses that a later allocation could reuse).
uint64_t next_serial();

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Resolves the current device's address of the allocation-free smoother's
// scratch array (hs_pcg.hip); called by hs_init, once per device.
void noalloc_scratch_init();

inline void finish(hipStream_t s, int flags) {
    if (!(flags & HS_FLAG_ASYNC)) HS_CHECK_HIP(hipStreamSynchronize(s));
}

inline void check_launch() { HS_CHECK_HIP(hipGetLastError()); }

inline unsigned grid_for(int64_t work, int block) {
    int64_t g = (work + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------
// scalar types of the smoother storage
// ---------------------------------------------------------------------------
template <class T>
struct Scalar;
template <>
struct Scalar<double> {
    __host__ __device__ static double to_f64(double v) { return v; }
    __host__ __device__ static double from_f64(double v) { return v; }
};
template <>
struct Scalar<float> {
    __host__ __device__ static double to_f64(float v) { return (double)v; }
    __host__ __device__ static float from_f64(double v) { return (float)v; }
};
template <>
struct Scalar<__hip_bfloat16> {
    __host__ __device__ static double to_f64(__hip_bfloat16 v) { return (double)__bfloat162float(v); }
    __host__ __device__ static __hip_bfloat16 from_f64(double v) { return __float2bfloat16((float)v); }
};

// compute-type conversion of a stored value
template <class TC, class TS>
__device__ __forceinline__ TC load_as(TS v);
template <>
__device__ __forceinline__ double load_as<double, double>(double v) { return v; }
template <>
__device__ __forceinline__ float load_as<float, float>(float v) { return v; }
template <>
__device__ __forceinline__ float load_as<float, __hip_bfloat16>(__hip_bfloat16 v) {
    return __bfloat162float(v);
}

// ---------------------------------------------------------------------------
// host staging for calls that pass host vectors
// ---------------------------------------------------------------------------
struct Staging {
    DBuf<double> a, b;
    void ensure(size_t n) {
        if (a.n < n) a.alloc(n);
        if (b.n < n) b.alloc(n);
    }
};

}  // namespace hs
