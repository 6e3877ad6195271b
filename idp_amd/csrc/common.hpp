// Shared host/device infrastructure of the MI355X Newton-with-barrier path:
// HIP error handling, growable device buffers, deterministic reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/idp_amd.h"

namespace idp {

// A roctx range named like the reference's TIMER_FLAG scope it stands for
// (Library/Utils/PROFILER.h:66-105), so a rocprofv3 --marker-trace shows the
// step in the reference's own terms; a no-op when no tool is attached.
struct TimerFlag {
    explicit TimerFlag(const char* name) { roctxRangePushA(name); }
    ~TimerFlag() { roctxRangePop(); }
    TimerFlag(const TimerFlag&) = delete;
    TimerFlag& operator=(const TimerFlag&) = delete;
};
#define IDP_TIMER_CAT2(a, b) a##b
#define IDP_TIMER_CAT(a, b) IDP_TIMER_CAT2(a, b)
#define IDP_TIMER_FLAG(name) ::idp::TimerFlag IDP_TIMER_CAT(idp_timer_flag_, __LINE__)(name)

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

// launch with a start/stop event pair recorded by the dispatch itself (kernel
// time only, as a profiler measures it); null events: a plain launch
#define IDP_LAUNCH_TIMED(kernel, grid, block, shmem, stream, ev0, ev1, ...)                             \
    do {                                                                                                \
        if ((grid) > 0) {                                                                               \
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, stream, ev0, ev1, 0, __VA_ARGS__); \
            IDP_HIP(hipGetLastError());                                                                 \
        }                                                                                               \
    } while (0)

inline int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

// Device bytes held by the library's buffers, process-wide (every context),
// and the most ever held at once: what a serving loop budgets against.
struct DevMem {
    std::atomic<long long> in_use{0}, peak{0};
    void add(long long b)
    {
        const long long now = in_use.fetch_add(b) + b;
        long long pk = peak.load();
        while (now > pk && !peak.compare_exchange_weak(pk, now)) {
        }
    }
};
inline DevMem& dev_mem()
{
    static DevMem m;
    return m;
}

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
        if (p) {
            (void)hipFree(p);
            dev_mem().add(-(long long)(cap * sizeof(T)));
        }
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
        dev_mem().add((long long)(nc * sizeof(T)));
        if (keep && p && cap) IDP_HIP(hipMemcpyAsync(q, p, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (p) {
            if (keep) IDP_HIP(hipStreamSynchronize(s));
            IDP_HIP(hipFree(p));
            dev_mem().add(-(long long)(cap * sizeof(T)));
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
    T* upload_get(const T* h, size_t n, hipStream_t s)
    {
        upload(h, n, s);
        return p;
    }
    void download(T* h, size_t n, hipStream_t s) const
    {
        if (n) IDP_HIP(hipMemcpyAsync(h, p, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// ---------------------------------------------------------------------------
// Optional per-family device timing with HIP events on one stream.  Records
// stay open until collect(); a caller may discard records it tagged (e.g. the
// kernels a chunked solver launched past its stopping point).
// ---------------------------------------------------------------------------
struct Profiler {
    struct Rec {
        int fam, a, b, tag;
        double work;
    };
    // 0 off; 1 every family (an event pair around each launch); 2 only the
    // sampled PCG kernel pairs, so a timed run keeps its launch gaps
    int level = 0;
    bool on = false;  // level != 0
    bool full() const { return level == 1; }
    hipStream_t s = nullptr;
    std::vector<hipEvent_t> pool;
    std::vector<Rec> open;
    size_t next = 0;
    double ms[IDP_K_COUNT] = {0};
    long launches[IDP_K_COUNT] = {0};
    double work[IDP_K_COUNT] = {0};
    double ccd_steps = 0, ccd_max_steps = 0;  // ACCD iterations: total, longest pair
    ~Profiler()
    {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
    int ev()
    {
        if (next == pool.size()) {
            hipEvent_t e;
            IDP_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
        return (int)next++;
    }
    void reset()
    {
        collect();
        for (int i = 0; i < IDP_K_COUNT; ++i) ms[i] = work[i] = 0, launches[i] = 0;
        ccd_steps = ccd_max_steps = 0;
    }
    // returns a token for end()
    int begin(int fam)
    {
        if (!full()) return -1;
        int a = ev();
        IDP_HIP(hipEventRecord(pool[a], s));
        return a;
    }
    void end(int fam, int tok, double w = 0.0, int tag = 0)
    {
        if (!full() || tok < 0) return;
        int b = ev();
        IDP_HIP(hipEventRecord(pool[b], s));
        open.push_back({fam, tok, b, tag, w});
    }
    // an event pair for a launch that records them itself (IDP_LAUNCH_TIMED):
    // returns a token for end_pair(), or -1 (and null events) when off
    int pair(hipEvent_t& e0, hipEvent_t& e1)
    {
        e0 = e1 = nullptr;
        if (!on) return -1;
        const int a = ev(), b = ev();
        e0 = pool[a];
        e1 = pool[b];
        return a;
    }
    void end_pair(int fam, int tok, double w = 0.0, int tag = 0)
    {
        if (!on || tok < 0) return;
        open.push_back({fam, tok, tok + 1, tag, w});
    }
    void discard_from(int fam, int tag)
    {
        size_t k = 0;
        for (size_t i = 0; i < open.size(); ++i)
            if (!(open[i].fam == fam && open[i].tag >= tag)) open[k++] = open[i];
        open.resize(k);
    }
    void add_work(int fam, double w)
    {
        if (full()) work[fam] += w;
    }
    // safe point: collect when many records are open
    void checkpoint()
    {
        if (open.size() > 16384) collect();
    }
    void collect()
    {
        if (open.empty()) {
            next = 0;
            return;
        }
        IDP_HIP(hipStreamSynchronize(s));
        for (auto& o : open) {
            float t = 0.f;
            IDP_HIP(hipEventElapsedTime(&t, pool[o.a], pool[o.b]));
            ms[o.fam] += t;
            launches[o.fam] += 1;
            work[o.fam] += o.work;
        }
        open.clear();
        next = 0;
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
// needs it (so a dependent kernel needs no extra launch).  B = block size.
template <int B>
__device__ inline double sum_partials_block_n(const double* part, int n, double* sh)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += B) acc += part[i];
    return block_sum<B>(acc, sh);
}

__device__ inline double sum_partials_block(const double* part, int n, double* sh)
{
    return sum_partials_block_n<RB>(part, n, sh);
}

}  // namespace idp
