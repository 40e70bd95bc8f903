// Shared host/device plumbing: status codes, HIP error checks, grow-only
// device buffers, wave-level helpers for wave64 (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/ipcg.h"

namespace ipcg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define IPCG_HIP(expr)                                                                  \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            throw ::ipcg::Error(IPCG_ERR_HIP, std::string(#expr) + ": "                 \
                                                  + hipGetErrorString(_e));             \
    } while (0)

#define IPCG_CHECK_LAUNCH() IPCG_HIP(hipGetLastError())

constexpr int WAVE = 64;

// Workgroups of `block` threads for n items.  By default the grid covers
// every item (one thread each); kernels that loop over their input with a
// grid stride pass their own cap.  A launch beyond the 2^32 work-items of
// one grid dimension is an error, never a silently truncated grid.
inline unsigned grid_for(int64_t n, int block, int64_t max_blocks = 0)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (max_blocks > 0 && g > max_blocks) g = max_blocks;
    if (g * (int64_t)block > (int64_t)UINT32_MAX)
        throw Error(IPCG_ERR_ARG, "a launch over " + std::to_string(n)
                                      + " items exceeds the 2^32 work-items of a grid");
    return (unsigned)g;
}

// Grow-only device allocation owned by a context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void reserve(size_t n)
    {
        if (n <= bytes) return;
        if (p) IPCG_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
        size_t want = n + n / 4 + 256;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error(IPCG_ERR_OOM, "device allocation of " + std::to_string(want)
                                          + " bytes failed");
        }
        bytes = want;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8) and every XCD has its own L2.  xcd_block() maps the workgroup to a
// logical index so that each XCD runs one contiguous range of logical
// workgroups: spatially ordered work (Morton-sorted queries, sorted
// collisions) then shares one L2.  A bijection of [0, gridDim.x).
constexpr unsigned NUM_XCD = 8;
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned b = blockIdx.x, nb = gridDim.x;
    const unsigned q = nb / NUM_XCD, r = nb % NUM_XCD, x = b % NUM_XCD;
    return x * q + (x < r ? x : r) + b / NUM_XCD;
}

// ---- wave64 helpers -------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// exclusive prefix count of `pred` among lower lanes, and the wave total
__device__ __forceinline__ unsigned wave_prefix(bool pred, unsigned& total)
{
    const unsigned long long m = __ballot(pred);
    total = (unsigned)__popcll(m);
    return (unsigned)__popcll(m & ((1ull << lane_id()) - 1ull));
}

// wave-aggregated append of `cnt` items; returns this lane's first slot
__device__ __forceinline__ unsigned long long wave_append(unsigned long long* counter,
                                                          unsigned cnt)
{
    // inclusive scan of cnt over active lanes via shuffles
    unsigned x = cnt;
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const unsigned y = __shfl_up(x, o, WAVE);
        if (l >= o) x += y;
    }
    const unsigned total = __shfl(x, WAVE - 1, WAVE);
    unsigned long long base = 0;
    if (l == WAVE - 1 && total) base = atomicAdd(counter, (unsigned long long)total);
    base = __shfl(base, WAVE - 1, WAVE);
    return base + (x - cnt);
}

// Appends one item for each lane with `a` set and one for each lane with
// `b` set (all lanes of the wave must call it together): ballots give the
// lane's offset (items of lower lanes, a-items then b-items per lane) and
// one atomic per wave reserves the slots.  Returns the wave's base slot;
// `pre` receives the lane's offset from it.
__device__ __forceinline__ unsigned long long wave_append_bits(unsigned long long* counter, bool a,
                                                               bool b, unsigned& pre)
{
    const unsigned long long ma = __ballot(a), mb = __ballot(b);
    const unsigned total = (unsigned)(__popcll(ma) + __popcll(mb));
    pre = __builtin_amdgcn_mbcnt_hi((unsigned)(ma >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)ma, 0u))
        + __builtin_amdgcn_mbcnt_hi((unsigned)(mb >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mb, 0u));
    unsigned long long base = 0;
    if (total) {
        if (lane_id() == 0) base = atomicAdd(counter, (unsigned long long)total);
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)base);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(base >> 32));
        base = ((unsigned long long)hi << 32) | lo;
    }
    return base;
}

} // namespace ipcg
