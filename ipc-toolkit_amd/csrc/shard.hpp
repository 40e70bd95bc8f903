// Sharded append lists.  One device-wide counter serialises every wave
// that appends to it (about 12 ns per atomic on MI355X, whatever the
// launch width), so producers that emit many small batches spread their
// appends over SHARDS counters, each on its own 128-byte line and owning a
// region of `scap` slots of a staging array.  The wave picks its shard from
// its global wave index.  After the producers, the host reads the shard
// counts (one small copy), checks that no region overflowed and packs the
// regions into a contiguous array with shard_compact.
#pragma once

#include "common.hpp"

#include <algorithm>

namespace ipcg {

constexpr int SHARDS = 256;
constexpr int SHARD_PAD = 16; // u64 slots between counters (128 B)

struct Shards {
    unsigned long long* cnt; // SHARDS * SHARD_PAD
    int64_t scap;            // slots per shard region
};

__device__ __forceinline__ int wave_shard()
{
    const unsigned wpb = blockDim.x / WAVE;
    return (int)((blockIdx.x * wpb + threadIdx.x / WAVE) & (SHARDS - 1));
}

// Reserves `cnt` slots for this lane in shard `sh`; returns the first slot's
// position inside the shard region (may be >= scap: the region overflowed,
// the caller must not write and the host retries with a larger region).
__device__ __forceinline__ unsigned long long shard_append(const Shards& S, int sh, unsigned cnt)
{
    return wave_append(S.cnt + (int64_t)sh * SHARD_PAD, cnt);
}

// Ballot form for lanes appending at most two items (a, then b).
__device__ __forceinline__ unsigned long long shard_append_bits(const Shards& S, int sh, bool a,
                                                                bool b, unsigned& pre)
{
    return wave_append_bits(S.cnt + (int64_t)sh * SHARD_PAD, a, b, pre);
}

struct ShardCounts {
    int64_t total = 0, max = 0;
};

// Counts of the last producer launch over nsets consecutive counter sets
// (set k at cnt_dev + k * SHARDS * SHARD_PAD): each set's total and largest
// shard are reduced on the device into sum_dev (2 * nsets u64) and read
// back in one small copy with one synchronisation of the stream, together
// with *extra_dev when it is given.
ShardCounts shard_counts(const unsigned long long* cnt_dev, unsigned long long* sum_dev,
                         void* pinned, hipStream_t s);
void shard_counts_n(const unsigned long long* cnt_dev, int nsets, unsigned long long* sum_dev,
                    void* pinned, hipStream_t s, ShardCounts* out, const int* extra_dev = nullptr,
                    int* extra = nullptr);

// exclusive offsets of the shard counts, each clipped to scap (SHARDS + 1
// int64 on the device)
void shard_offsets(const unsigned long long* cnt_dev, int64_t scap, int64_t* off_dev,
                   hipStream_t s);

template <class T>
__global__ __launch_bounds__(256) void k_shard_copy(const T* in, int64_t scap,
                                                    const unsigned long long* cnt,
                                                    const int64_t* off, T* out)
{
    const int sh = blockIdx.y;
    const int64_t n = (int64_t)cnt[(int64_t)sh * SHARD_PAD];
    const int64_t lim = n < scap ? n : scap;
    const int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int64_t j = base + 256 * k;
        if (j < lim) out[off[sh] + j] = in[(int64_t)sh * scap + j];
    }
}

// Packs the shard regions of `in` into `out` (off_dev: SHARDS + 1 int64).
template <class T>
void shard_compact(const T* in, int64_t scap, const unsigned long long* cnt_dev, int64_t max_count,
                   T* out, int64_t* off_dev, hipStream_t s)
{
    if (max_count <= 0) return;
    shard_offsets(cnt_dev, scap, off_dev, s);
    const int64_t m = std::min<int64_t>(max_count, scap);
    hipLaunchKernelGGL(k_shard_copy<T>, dim3((unsigned)((m + 1023) / 1024), SHARDS), dim3(256), 0,
                       s, in, scap, cnt_dev, off_dev, out);
    IPCG_CHECK_LAUNCH();
}

constexpr size_t SHARD_CNT_BYTES = sizeof(unsigned long long) * SHARDS * SHARD_PAD;

} // namespace ipcg
