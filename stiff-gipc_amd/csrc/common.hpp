// Shared host/device utilities for the sgipc HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace sgipc {

constexpr int kWave = 64;

struct hip_error : std::runtime_error {
    explicit hip_error(const std::string& s) : std::runtime_error(s) {}
};
struct arg_error : std::runtime_error {
    explicit arg_error(const std::string& s) : std::runtime_error(s) {}
};
struct capacity_error : std::runtime_error {
    explicit capacity_error(const std::string& s) : std::runtime_error(s) {}
};
struct state_error : std::runtime_error {
    explicit state_error(const std::string& s) : std::runtime_error(s) {}
};

#define SGIPC_HIP(call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::sgipc::hip_error(std::string(#call) + " failed: " + hipGetErrorString(e_) + \
                                     " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

#define SGIPC_LAUNCH_CHECK() SGIPC_HIP(hipGetLastError())

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// Device buffer that only grows; contents are not preserved across growth.
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    T* reserve(size_t n) {
        if (n <= cap && p) return p;
        release();
        size_t bytes = (n ? n : 1) * sizeof(T);
        SGIPC_HIP(hipMalloc((void**)&p, bytes));
        cap = n ? n : 1;
        return p;
    }
    T* get() const { return p; }
};

// ---- device helpers -------------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
inline void fill(T* p, T v, long long n, hipStream_t s) {
    if (n <= 0) return;
    int blocks = div_up(n, 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, s, p, v, n);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
