// Shared definitions of the wb97mv_amd library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>
#include "wb97mv_amd.h"

namespace wb {

constexpr int WAVE = 64;      // CDNA wavefront
constexpr int SUB = 64;       // grid points per sub-block = one wave, one point per lane
constexpr int LMAX = 3;
constexpr int NCLASS = (LMAX + 1) * (LMAX + 2) / 2;   // (la >= lb) classes

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define WB_HIP(call)                                                             \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            int code_ = (e_ == hipErrorOutOfMemory) ? WB_OOM : WB_DEVICE_ERROR;         \
            throw wb::Error(code_, std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                        \
    } while (0)

inline int class_id(int la, int lb) { return la * (la + 1) / 2 + lb; }   // la >= lb
inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Device buffer owned by the library, or a view of memory owned elsewhere
// (borrow(): never freed, never grown).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = true;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = true;
    }
    void borrow(void* q, size_t n) {
        release();
        p = q;
        bytes = n;
        owned = false;
    }
    // grow-only (re)allocation; contents are not preserved
    void* ensure(size_t n) {
        if (n > bytes) {
            if (!owned) throw Error(WB_INVALID, "borrowed device buffer too small");
            release();
            if (n) WB_HIP(hipMalloc(&p, n));
            bytes = n;
        }
        return p;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// ---- wave-level helpers (64 lanes) ------------------------------------------
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace wb
