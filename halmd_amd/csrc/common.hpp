// Shared definitions of the MI355X (gfx950) MD backend: status codes, device
// buffers, launch helpers and double-word ("float-float") arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/halmd_hip.h"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace halmd {

// ---------------------------------------------------------------------------
// Errors inside the C++ module layer are C++ exceptions carrying one of the
// HALMD_ERR_* codes of the C-ABI; the C-ABI converts them to return values
// and keeps the message for halmd_last_error().
// ---------------------------------------------------------------------------
struct error : std::runtime_error {
    int code;
    error(int c, std::string const& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, char const* what) {
    if (e != hipSuccess) throw error(HALMD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIP_CHECK(x) ::halmd::hip_check((x), #x)

inline void require(bool cond, std::string const& msg) {
    if (!cond) throw error(HALMD_ERR_INVALID_ARGUMENT, msg);
}

// ---------------------------------------------------------------------------
// Device buffer (owning, non-copyable).  Contents are uninitialised after
// allocation; kernels never allocate.
// ---------------------------------------------------------------------------
template <typename T>
class dbuf {
public:
    dbuf() = default;
    explicit dbuf(size_t n) { resize(n); }
    ~dbuf() { release(); }
    dbuf(dbuf const&) = delete;
    dbuf& operator=(dbuf const&) = delete;
    dbuf(dbuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    dbuf& operator=(dbuf&& o) noexcept { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; return *this; }

    void resize(size_t n) {
        if (n == n_) return;
        release();
        if (n) HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
        n_ = n;
    }
    // grow only (keeps the allocation when large enough)
    void reserve(size_t n) { if (n > n_) resize(n); }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr; n_ = 0;
    }
    T* get() const { return p_; }
    operator T*() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

    void upload(T const* h, size_t n, hipStream_t s) const {
        HIP_CHECK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t n, hipStream_t s) const {
        HIP_CHECK(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Page-locked host buffer for small device-to-host reads (flags, partials).
template <typename T>
class hbuf {
public:
    hbuf() = default;
    explicit hbuf(size_t n) { resize(n); }
    ~hbuf() { if (p_) (void)hipHostFree(p_); }
    hbuf(hbuf const&) = delete;
    hbuf& operator=(hbuf const&) = delete;
    void resize(size_t n) {
        if (n == n_) return;
        if (p_) (void)hipHostFree(p_);
        p_ = nullptr;
        if (n) HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault));
        n_ = n;
    }
    T* get() const { return p_; }
    T& operator[](size_t i) const { return p_[i]; }
    size_t size() const { return n_; }
private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

constexpr unsigned WAVE = 64;                 // CDNA wavefront width
constexpr unsigned PAD = 256;                 // particle arrays are padded to this multiple

inline unsigned div_up(size_t a, size_t b) { return unsigned((a + b - 1) / b); }
inline size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// exact split of a double into an unevaluated float sum hi + lo
inline void split_ff(double x, float& hi, float& lo) {
    hi = float(x);
    lo = float(x - double(hi));
}

// ---------------------------------------------------------------------------
// Double-word arithmetic on (hi, lo) float pairs.
// Algorithms: Knuth's TwoSum, Dekker's Fast2Sum, FMA-based 2Prod, and the
// double-word operations DWPlusFP (Alg. 4), AccurateDWPlusDW (Alg. 6) and
// DWTimesFP3 (Alg. 9) of Joldes, Muller & Popescu, "Tight and rigorous error
// bounds for basic building blocks of double-word arithmetic", ACM TOMS 44
// (2017).  Contraction is disabled inside each routine so that the error-free
// transformations stay exact.
// ---------------------------------------------------------------------------
struct ff {
    float hi, lo;
};

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
#pragma clang fp contract(off)
    s = a + b;
    float ap = s - b;
    float bp = s - ap;
    e = (a - ap) + (b - bp);
}

__device__ __forceinline__ void fast_two_sum(float a, float b, float& s, float& e) {
#pragma clang fp contract(off)
    s = a + b;
    e = b - (s - a);
}

// (xh, xl) + y
__device__ __forceinline__ ff ff_add_f(ff x, float y) {
#pragma clang fp contract(off)
    float sh, sl;
    two_sum(x.hi, y, sh, sl);
    float v = x.lo + sl;
    ff z;
    fast_two_sum(sh, v, z.hi, z.lo);
    return z;
}

// (xh, xl) + (yh, yl), accurate variant
__device__ __forceinline__ ff ff_add(ff x, ff y) {
#pragma clang fp contract(off)
    float sh, sl, th, tl, vh, vl;
    two_sum(x.hi, y.hi, sh, sl);
    two_sum(x.lo, y.lo, th, tl);
    float c = sl + th;
    fast_two_sum(sh, c, vh, vl);
    float w = tl + vl;
    ff z;
    fast_two_sum(vh, w, z.hi, z.lo);
    return z;
}

// (xh, xl) * y
__device__ __forceinline__ ff ff_mul_f(ff x, float y) {
#pragma clang fp contract(off)
    float ch = x.hi * y;
    float cl1 = __builtin_fmaf(x.hi, y, -ch);
    float cl3 = __builtin_fmaf(x.lo, y, cl1);
    ff z;
    fast_two_sum(ch, cl3, z.hi, z.lo);
    return z;
}

__device__ __forceinline__ ff ff_neg(ff x) { return ff{-x.hi, -x.lo}; }

// (xh, xl) * (yh, yl) (DWTimesDW1 of the same paper)
__device__ __forceinline__ ff ff_mul(ff x, ff y) {
#pragma clang fp contract(off)
    float const ch = x.hi * y.hi;
    float const cl1 = __builtin_fmaf(x.hi, y.hi, -ch);
    float const tl = x.hi * y.lo;
    float const cl2 = __builtin_fmaf(x.lo, y.hi, tl);
    float const cl3 = cl1 + cl2;
    ff z;
    fast_two_sum(ch, cl3, z.hi, z.lo);
    return z;
}

// square root of a double word: one Newton correction of the FP32 root
__device__ __forceinline__ ff ff_sqrt(ff x) {
#pragma clang fp contract(off)
    float const s = __builtin_sqrtf(x.hi);
    float const e = __builtin_fmaf(-s, s, x.hi);
    float const c = (e + x.lo) / (2.0f * s);
    ff z;
    fast_two_sum(s, c, z.hi, z.lo);
    return z;
}

// exact double word of a 32-bit integer
__device__ __forceinline__ ff ff_from_int(int32_t d) {
    float const hi = float(d);
    return ff{hi, float(d - int32_t(hi))};
}

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
    for (int off = WAVE / 2; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, WAVE));
    return v;
}

} // namespace halmd
