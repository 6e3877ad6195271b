// Shared host/device infrastructure of the MI355X Newton-with-barrier path:
// HIP error handling, growable device buffers, deterministic reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/idp_amd.h"

namespace idp {

struct HipError : std::runtime_error {
    int code;
    HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define IDP_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::idp::HipError(IDP_E_HIP, std::string(#call " failed: ") + hipGetErrorString(e_) +   \
                                                  " at " __FILE__ ":" + std::to_string(__LINE__));      \
    } while (0)

#define IDP_LAUNCH(kernel, grid, block, shmem, stream, ...)                                             \
    do {                                                                                                \
        if ((grid) > 0) {                                                                               \
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, stream, __VA_ARGS__);           \
            IDP_HIP(hipGetLastError());                                                                 \
        }                                                                                               \
    } while (0)

inline int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

// Growable device buffer (capacity only grows; contents not preserved on grow
// unless keep=true).
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    T* ensure(size_t n, hipStream_t s = 0, bool keep = false)
    {
        if (n <= cap && p) return p;
        size_t nc = std::max<size_t>(n, 16);
        nc = std::max(nc, cap + cap / 2);
        T* q = nullptr;
        IDP_HIP(hipMalloc(&q, nc * sizeof(T)));
        if (keep && p && cap) IDP_HIP(hipMemcpyAsync(q, p, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (p) {
            if (keep) IDP_HIP(hipStreamSynchronize(s));
            IDP_HIP(hipFree(p));
        }
        p = q;
        cap = nc;
        return p;
    }
    void upload(const T* h, size_t n, hipStream_t s)
    {
        ensure(n, s);
        if (n) IDP_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t n, hipStream_t s) const
    {
        if (n) IDP_HIP(hipMemcpyAsync(h, p, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// ---------------------------------------------------------------------------
// Deterministic reductions.  Every sum is taken in a fixed order for a fixed
// problem size: block b reduces a fixed index range through a fixed LDS tree,
// and a single block then reduces the per-block partials the same way.  No
// floating-point atomics are used anywhere in the library.
// ---------------------------------------------------------------------------
constexpr int RB = 256;  // reduction block size

template <int B>
__device__ inline double block_sum(double v, double* sh)
{
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = B / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// Sum of n partials in a fixed order, computed redundantly by every block that
// needs it (so a dependent kernel needs no extra launch).
__device__ inline double sum_partials_block(const double* part, int n, double* sh)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
    return block_sum<RB>(acc, sh);
}

}  // namespace idp
