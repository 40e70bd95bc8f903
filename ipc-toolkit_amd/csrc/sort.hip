// Device-wide scan, stable LSD radix sort and run-length encoding.
//
// Radix sort pass (8-bit digit, 256-thread blocks, 2048-key tiles):
//   1. per-tile digit histogram (LDS atomics) -> counts[digit][tile]
//   2. exclusive scan of counts in digit-major order -> tile base offsets
//   3. per-tile stable ranking: each wave owns 512 consecutive keys, ranks
//      a 64-key chunk at a time by matching digits with 8 ballots, keeps a
//      running per-digit count in LDS; wave counts are prefix-summed across
//      the block, then every key is written to base + wave prefix + rank.
// 64-bit or 32-bit keys.  Passes whose digit is constant over all keys are
// skipped where the caller asks for it (a read-back of the varying bits).
#include "shard.hpp"
#include "sort.hpp"

#include <utility>

namespace ipcg {

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <class T> __device__ T block_exclusive_scan(T v, T& total)
{
    __shared__ T wsum[SCAN_BLOCK / WAVE];
    const int l = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    T x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const T y = __shfl_up(x, o, WAVE);
        if (l >= o) x += y;
    }
    if (l == WAVE - 1) wsum[w] = x;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SCAN_BLOCK / WAVE; i++) {
        const T s = wsum[i];
        if (i < w) off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

template <class T>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_partials(const T* in, int64_t n, T* partial)
{
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    T s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        const int64_t i = base + (int64_t)j * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += in[i];
    }
    T tot;
    block_exclusive_scan<T>(s, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// single block scans the partials in place (any count, looping)
template <class T>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_single(T* partial, int64_t m, T* total)
{
    T carry = 0;
    for (int64_t base = 0; base < m; base += SCAN_TILE) {
        T v[SCAN_ITEMS];
        T s = 0;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; j++) {
            const int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + j;
            v[j] = i < m ? partial[i] : 0;
            s += v[j];
        }
        T tot;
        T off = block_exclusive_scan<T>(s, tot) + carry;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; j++) {
            const int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + j;
            if (i < m) partial[i] = off;
            off += v[j];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <class T>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_final(const T* in, T* out, int64_t n,
                                                           const T* partial)
{
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        const int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + j;
        v[j] = i < n ? in[i] : 0;
        s += v[j];
    }
    T tot;
    T off = block_exclusive_scan<T>(s, tot) + partial[blockIdx.x];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        const int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + j;
        if (i < n) out[i] = off;
        off += v[j];
    }
}

// n <= SCAN_TILE: the whole scan in one workgroup (one launch instead of
// three)
template <class T>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_one(const T* in, T* out, int64_t n,
                                                         T* total)
{
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        const int64_t i = (int64_t)threadIdx.x * SCAN_ITEMS + j;
        v[j] = i < n ? in[i] : 0;
        s += v[j];
    }
    T tot;
    T off = block_exclusive_scan<T>(s, tot);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        const int64_t i = (int64_t)threadIdx.x * SCAN_ITEMS + j;
        if (i < n) out[i] = off;
        off += v[j];
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

template <class T>
void exclusive_scan(const T* in, T* out, int64_t n, T* total, void* tmp, hipStream_t s)
{
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    T* partial = static_cast<T*>(tmp);
    if (n <= 0) {
        if (total) IPCG_HIP(hipMemsetAsync(total, 0, sizeof(T), s));
        return;
    }
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_one<T>, dim3(1), dim3(SCAN_BLOCK), 0, s, in, out, n, total);
        IPCG_CHECK_LAUNCH();
        return;
    }
    hipLaunchKernelGGL(k_scan_partials<T>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n,
                       partial);
    hipLaunchKernelGGL(k_scan_single<T>, dim3(1), dim3(SCAN_BLOCK), 0, s, partial, nb, total);
    hipLaunchKernelGGL(k_scan_final<T>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, out, n,
                       partial);
    IPCG_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
constexpr int RS_BLOCK = 256;
constexpr int RS_WAVES = RS_BLOCK / WAVE;
constexpr int RS_CHUNKS = 8; // 64-key chunks per wave
constexpr int RS_TILE = RS_BLOCK * RS_CHUNKS;

template <class K>
__global__ __launch_bounds__(RS_BLOCK) void k_radix_hist(const K* keys, int64_t n, int shift,
                                                         uint32_t* counts, int64_t ntiles)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_CHUNKS; j++) {
        const int64_t i = base + (int64_t)j * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one 8-bit digit.  Each wave ranks its keys per digit
// with ballots; the block turns the per-wave counts into offsets, places
// the tile's keys in LDS in digit order, and writes each digit's run to its
// global range so that neighbouring lanes write neighbouring addresses.
template <class K>
__global__ __launch_bounds__(RS_BLOCK) void k_radix_scatter(
    const K* keys, const uint32_t* vals, K* okeys, uint32_t* ovals, int64_t n, int shift,
    const uint32_t* offsets, int64_t ntiles)
{
    __shared__ uint32_t wcnt[RS_WAVES][256];
    __shared__ uint32_t tile_off[256];
    __shared__ uint32_t dstart[256]; // first slot of each digit in the tile
    __shared__ uint32_t wtot[RS_WAVES];
    __shared__ K sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_BLOCK) (&wcnt[0][0])[i] = 0;
    tile_off[threadIdx.x] = offsets[(int64_t)threadIdx.x * ntiles + blockIdx.x];
    __syncthreads();
    const int64_t tile0 = (int64_t)blockIdx.x * RS_TILE;
    const int64_t base = tile0 + (int64_t)w * (RS_CHUNKS * WAVE);
    const unsigned long long lt = (1ull << lane) - 1ull;
    K k[RS_CHUNKS];
    uint32_t v[RS_CHUNKS];
    uint32_t rank[RS_CHUNKS];
#pragma unroll
    for (int c = 0; c < RS_CHUNKS; c++) {
        const int64_t i = base + c * WAVE + lane;
        const bool valid = i < n;
        k[c] = valid ? keys[i] : 0;
        v[c] = valid ? vals[i] : 0;
        const uint32_t d = (uint32_t)(k[c] >> shift) & 0xFF;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const unsigned long long m = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const uint32_t r = (uint32_t)__popcll(peers & lt);
        const uint32_t before = valid ? wcnt[w][d] : 0;
        __builtin_amdgcn_wave_barrier();
        if (valid && r == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[c] = before + r;
    }
    __syncthreads();
    // per digit: wave counts -> offsets within the digit, digit total
    uint32_t total;
    {
        const int d = threadIdx.x; // RS_BLOCK == 256 digits
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ww++) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += c;
        }
        total = run;
    }
    // exclusive scan of the digit totals: within each wave, then across
    uint32_t x = total;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) wtot[w] = x;
    __syncthreads();
    uint32_t wpre = 0;
#pragma unroll
    for (int ww = 0; ww < RS_WAVES; ww++)
        if (ww < w) wpre += wtot[ww];
    dstart[threadIdx.x] = wpre + x - total;
    __syncthreads();
    // keys in digit order in LDS
#pragma unroll
    for (int c = 0; c < RS_CHUNKS; c++) {
        const int64_t i = base + c * WAVE + lane;
        if (i < n) {
            const uint32_t d = (uint32_t)(k[c] >> shift) & 0xFF;
            const uint32_t lp = dstart[d] + wcnt[w][d] + rank[c];
            sk[lp] = k[c];
            sv[lp] = v[c];
        }
    }
    __syncthreads();
    const int tn = (int)(n - tile0 < RS_TILE ? n - tile0 : RS_TILE);
    for (int j = threadIdx.x; j < tn; j += RS_BLOCK) {
        const K key = sk[j];
        const uint32_t d = (uint32_t)(key >> shift) & 0xFF;
        const int64_t pos = (int64_t)tile_off[d] + (j - (int)dstart[d]);
        okeys[pos] = key;
        ovals[pos] = sv[j];
    }
}

// xor-fold of all keys against the first one: bits that vary
template <class K>
__global__ __launch_bounds__(256) void k_vary_bits(const K* keys, int64_t n,
                                                   unsigned long long* out)
{
    const uint64_t k0 = keys[0];
    uint64_t acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        acc |= keys[i] ^ k0;
    // wave reduce
    for (int o = 32; o > 0; o >>= 1) acc |= __shfl_xor(acc, o, WAVE);
    if (threadIdx.x % WAVE == 0 && acc) atomicOr(out, (unsigned long long)acc);
}

__global__ void k_rle_flags(const uint64_t* keys, int64_t n, uint32_t* flags)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_rle_scatter(const uint64_t* keys, int64_t n, const uint32_t* flags,
                              const uint32_t* pos, uint64_t* ukeys, uint32_t* start,
                              const uint32_t* nu)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) {
        ukeys[pos[i]] = keys[i];
        start[pos[i]] = (uint32_t)i;
    }
    if (i == 0) start[*nu] = (uint32_t)n;
}

} // namespace

size_t scan_tmp_bytes(int64_t n)
{
    return sizeof(uint64_t) * (size_t)((n + SCAN_TILE - 1) / SCAN_TILE + 1) + 256;
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev,
                        void* tmp, hipStream_t s)
{
    exclusive_scan<uint32_t>(in, out, n, total_dev, tmp, s);
}

void exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, uint64_t* total_dev,
                        void* tmp, hipStream_t s)
{
    exclusive_scan<uint64_t>(in, out, n, total_dev, tmp, s);
}

size_t radix_tmp_bytes(int64_t n)
{
    const int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    const int64_t nc = 256 * ntiles;
    return 2 * sizeof(uint32_t) * (size_t)nc + scan_tmp_bytes(nc) + 1024;
}

// LSD passes over the low key_bits bits.  skip_constant: one pass over the
// keys and a read-back find the digits that vary, and only those passes run;
// otherwise every pass runs without any host round trip (a constant digit
// costs a stable identity pass).  Returns true when the sorted pairs ended
// in keys_alt / vals_alt (copy_back false and an odd number of passes).
template <class K>
static bool radix_sort(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n,
                       int key_bits, void* tmp, hipStream_t s, bool skip_constant,
                       bool copy_back, unsigned long long bit_mask = ~0ull)
{
    if (n <= 1) return false;
    const int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    const int64_t nc = 256 * ntiles;
    char* t = static_cast<char*>(tmp);
    unsigned long long* vary = reinterpret_cast<unsigned long long*>(t);
    uint32_t* counts = reinterpret_cast<uint32_t*>(t + 256);
    uint32_t* offsets = counts + nc;
    void* stmp = offsets + nc;
    unsigned long long hv = bit_mask; // digits without a bit here are constant
    if (skip_constant) { // which bits vary at all (one small D2H read per sort)
        IPCG_HIP(hipMemsetAsync(vary, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_vary_bits<K>, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, keys,
                           n, vary);
        IPCG_CHECK_LAUNCH();
        IPCG_HIP(hipMemcpyAsync(&hv, vary, sizeof(hv), hipMemcpyDeviceToHost, s));
        IPCG_HIP(hipStreamSynchronize(s));
    }
    K* ki = keys;
    uint32_t* vi = vals;
    K* ko = keys_alt;
    uint32_t* vo = vals_alt;
    for (int shift = 0; shift < key_bits; shift += 8) {
        if (((hv >> shift) & 0xFFull) == 0) continue;
        hipLaunchKernelGGL(k_radix_hist<K>, dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, s, ki, n,
                           shift, counts, ntiles);
        exclusive_scan_u32(counts, offsets, nc, nullptr, stmp, s);
        hipLaunchKernelGGL(k_radix_scatter<K>, dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, s, ki,
                           vi, ko, vo, n, shift, offsets, ntiles);
        IPCG_CHECK_LAUNCH();
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (ki == keys) return false;
    if (!copy_back) return true;
    IPCG_HIP(hipMemcpyAsync(keys, ki, sizeof(K) * n, hipMemcpyDeviceToDevice, s));
    IPCG_HIP(hipMemcpyAsync(vals, vi, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    return false;
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                      int64_t n, int key_bits, void* tmp, hipStream_t s)
{
    radix_sort(keys, vals, keys_alt, vals_alt, n, key_bits, tmp, s, true, true);
}

bool radix_sort_pairs_all(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                          uint32_t* vals_alt, int64_t n, int key_bits, void* tmp, hipStream_t s)
{
    return radix_sort(keys, vals, keys_alt, vals_alt, n, key_bits, tmp, s, false, false);
}

bool radix_sort_pairs_masked(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                             uint32_t* vals_alt, int64_t n, unsigned long long bit_mask,
                             void* tmp, hipStream_t s)
{
    return radix_sort(keys, vals, keys_alt, vals_alt, n, 64, tmp, s, false, false, bit_mask);
}

bool radix_sort_pairs_all(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                          uint32_t* vals_alt, int64_t n, int key_bits, void* tmp, hipStream_t s)
{
    if (key_bits > 32) throw Error(IPCG_ERR_ARG, "32-bit keys hold at most 32 key bits");
    return radix_sort(keys, vals, keys_alt, vals_alt, n, key_bits, tmp, s, false, false);
}

size_t rle_tmp_bytes(int64_t n)
{
    return 2 * sizeof(uint32_t) * (size_t)(n + 1) + scan_tmp_bytes(n) + 256;
}

void run_length_encode(const uint64_t* keys, int64_t n, uint64_t* unique_keys,
                       uint32_t* run_start, uint32_t* n_unique_dev, void* tmp, hipStream_t s)
{
    if (n <= 0) {
        IPCG_HIP(hipMemsetAsync(n_unique_dev, 0, sizeof(uint32_t), s));
        IPCG_HIP(hipMemsetAsync(run_start, 0, sizeof(uint32_t), s));
        return;
    }
    uint32_t* flags = static_cast<uint32_t*>(tmp);
    uint32_t* pos = flags + (n + 1);
    void* stmp = pos + (n + 1);
    // one thread per key: the grid must cover all n keys
    hipLaunchKernelGGL(k_rle_flags, dim3(grid_for(n, 256, INT32_MAX)), dim3(256), 0, s, keys, n,
                       flags);
    exclusive_scan_u32(flags, pos, n, n_unique_dev, stmp, s);
    hipLaunchKernelGGL(k_rle_scatter, dim3(grid_for(n, 256, INT32_MAX)), dim3(256), 0, s, keys, n, flags,
                       pos, unique_keys, run_start, n_unique_dev);
    IPCG_CHECK_LAUNCH();
}

__global__ void k_shard_offsets(const unsigned long long* cnt, int64_t scap, int64_t* off)
{
    __shared__ int64_t v[SHARDS];
    const int t = threadIdx.x;
    const int64_t c = (int64_t)cnt[(int64_t)t * SHARD_PAD];
    v[t] = c < scap ? c : scap;
    __syncthreads();
    if (t == 0) {
        int64_t acc = 0;
        for (int i = 0; i < SHARDS; i++) {
            off[i] = acc;
            acc += v[i];
        }
        off[SHARDS] = acc;
    }
}

// one workgroup per counter set: its total and its largest shard count
__global__ __launch_bounds__(SHARDS) void k_shard_summary(const unsigned long long* cnt,
                                                          unsigned long long* out)
{
    __shared__ unsigned long long wsum[SHARDS / WAVE], wmax[SHARDS / WAVE];
    const unsigned long long v =
        cnt[(int64_t)blockIdx.x * SHARDS * SHARD_PAD + (int64_t)threadIdx.x * SHARD_PAD];
    unsigned long long sm = v, mx = v;
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        sm += __shfl_xor(sm, o, WAVE);
        const unsigned long long y = __shfl_xor(mx, o, WAVE);
        mx = y > mx ? y : mx;
    }
    if (threadIdx.x % WAVE == 0) {
        wsum[threadIdx.x / WAVE] = sm;
        wmax[threadIdx.x / WAVE] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0, m = 0;
        for (int w = 0; w < SHARDS / WAVE; w++) {
            t += wsum[w];
            m = wmax[w] > m ? wmax[w] : m;
        }
        out[2 * blockIdx.x] = t;
        out[2 * blockIdx.x + 1] = m;
    }
}

void shard_counts_n(const unsigned long long* cnt_dev, int nsets, unsigned long long* sum_dev,
                    void* pinned, hipStream_t s, ShardCounts* out, const int* extra_dev,
                    int* extra)
{
    hipLaunchKernelGGL(k_shard_summary, dim3(nsets), dim3(SHARDS), 0, s, cnt_dev, sum_dev);
    IPCG_CHECK_LAUNCH();
    unsigned long long* h = static_cast<unsigned long long*>(pinned);
    IPCG_HIP(hipMemcpyAsync(h, sum_dev, sizeof(unsigned long long) * 2 * nsets,
                            hipMemcpyDeviceToHost, s));
    if (extra_dev)
        IPCG_HIP(hipMemcpyAsync(h + 2 * nsets, extra_dev, sizeof(int), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < nsets; k++) {
        out[k].total = (int64_t)h[2 * k];
        out[k].max = (int64_t)h[2 * k + 1];
    }
    if (extra_dev && extra) *extra = *reinterpret_cast<const int*>(h + 2 * nsets);
}

ShardCounts shard_counts(const unsigned long long* cnt_dev, unsigned long long* sum_dev,
                         void* pinned, hipStream_t s)
{
    ShardCounts r;
    shard_counts_n(cnt_dev, 1, sum_dev, pinned, s, &r);
    return r;
}

void shard_offsets(const unsigned long long* cnt_dev, int64_t scap, int64_t* off_dev,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_shard_offsets, dim3(1), dim3(SHARDS), 0, s, cnt_dev, scap, off_dev);
    IPCG_CHECK_LAUNCH();
}

} // namespace ipcg
