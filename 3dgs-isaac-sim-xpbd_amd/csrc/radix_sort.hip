// radix_sort.hip — stable LSD radix sort of (uint32 key, int32 value) pairs.
//
// 8-bit digits, one three-kernel pass per digit:
//   hist    one block per tile of kTileKeys keys: digit histogram in LDS,
//           written digit-major hist[digit * tiles + tile];
//   scan    one block per digit: exclusive scan of its tile counts plus the
//           digit total (digit, then tile, then position: stable);
//   scatter one block per tile: keys are ranked in rounds of 256 (one per
//           thread, in index order); within a wave64 the rank among equal
//           digits comes from 8 ballots (match mask) and a popcount, across
//           the four waves from LDS counts, across rounds from running
//           per-digit counters.  Every key keeps its input order within its
//           digit, so the sort is stable like an LSD sort must be.
// Values start as the identity when `iota` is set, else as given in v0.
// Only the low `bits` bits of the keys are sorted (keys must not have higher
// bits set).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace xpbd {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kTileKeys = kSortThreads * kSortRounds;   // 2048 keys per tile

__global__ void __launch_bounds__(kSortThreads) radix_hist_kernel(int n, const unsigned int* __restrict__ keys, int shift,
                                                                 int tiles, unsigned int* __restrict__ hist,
                                                                 const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTileKeys;
    for (int r = 0; r < kSortRounds; ++r) {
        const int i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Per digit d (one block each): exclusive scan of its `tiles` counts
// hist[d * tiles + tile] in place, and the digit's total in dsum[d].  The
// scatter adds the digits' exclusive offsets, so the global order is digit,
// then tile, then position (stable).
__global__ void __launch_bounds__(kSortThreads) radix_scan_kernel(int tiles, unsigned int* __restrict__ hist,
                                                                 unsigned int* __restrict__ dsum,
                                                                 const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ unsigned int wsum[kSortThreads / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    unsigned int* h = hist + (size_t)blockIdx.x * tiles;
    unsigned int carry = 0;
    for (int base = 0; base < tiles; base += kSortThreads) {
        const int i = base + t;
        const unsigned int v = i < tiles ? h[i] : 0u;
        unsigned int x = v;   // inclusive scan inside the wave
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const unsigned int o = __shfl_up(x, m, 64);
            if (lane >= m) x += o;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        unsigned int before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (i < tiles) h[i] = before + x - v;
        unsigned int chunk = 0;
        for (int w = 0; w < kSortThreads / 64; ++w) chunk += wsum[w];
        carry += chunk;
        __syncthreads();
    }
    if (t == 0) dsum[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(kSortThreads) radix_scatter_kernel(
    int n, const unsigned int* __restrict__ keysIn, const int* __restrict__ valsIn, unsigned int* __restrict__ keysOut,
    int* __restrict__ valsOut, int shift, int tiles, const unsigned int* __restrict__ offsets,
    const unsigned int* __restrict__ dsum, int iota, const int* __restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ unsigned int run[256];          // keys of each digit already placed by this tile
    __shared__ unsigned int wcnt[4][256];      // per-wave digit counts of the current round
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // exclusive offsets of the digits (scan of the 256 digit totals)
    run[t] = dsum[t];
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const unsigned int v = (t >= off) ? run[t - off] : 0u;
        __syncthreads();
        run[t] += v;
        __syncthreads();
    }
    const unsigned int digitBase = run[t] - dsum[t];
    __syncthreads();
    run[t] = digitBase + offsets[t * tiles + blockIdx.x];
    const int base = blockIdx.x * kTileKeys;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kSortRounds; ++r) {
        const int i = base + r * kSortThreads + t;
        const bool valid = i < n;
        unsigned int key = 0, digit = 0;
        int val = 0;
        if (valid) {
            key = keysIn[i];
            digit = (key >> shift) & 255u;
            val = iota ? i : valsIn[i];
        }
        wcnt[0][t] = 0; wcnt[1][t] = 0; wcnt[2][t] = 0; wcnt[3][t] = 0;
        __syncthreads();
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bal = __ballot(valid && ((digit >> b) & 1u));
            peers &= ((digit >> b) & 1u) ? bal : ~bal;
        }
        const unsigned int rankInWave = (unsigned int)__popcll(peers & lt);
        if (valid && rankInWave == 0) wcnt[wave][digit] = (unsigned int)__popcll(peers);
        __syncthreads();
        unsigned int before = 0;
        if (valid) {
            for (int w = 0; w < wave; ++w) before += wcnt[w][digit];
            const unsigned int pos = run[digit] + before + rankInWave;
            keysOut[pos] = key;
            valsOut[pos] = val;
        }
        __syncthreads();
        run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();
    }
}

int radix_sort_tiles(int n) { return (n + kTileKeys - 1) / kTileKeys; }
size_t radix_sort_hist_size(int n) { return 256 * (size_t)radix_sort_tiles(n) + 256; }

// keys/vals ping-pong between (k0, v0) and (k1, v1); the sorted result ends
// in (k0, v0) when the number of passes is even, else in (k1, v1).  Returns
// true when the result is in (k1, v1).
bool launch_radix_sort(hipStream_t s, int n, int bits, unsigned int* k0, int* v0, unsigned int* k1, int* v1,
                       unsigned int* hist, bool iota, const int* gate)
{
    if (n <= 0) return false;
    const int tiles = radix_sort_tiles(n);
    const int passes = (bits + 7) / 8;
    unsigned int* kin = k0;
    int* vin = v0;
    unsigned int* kout = k1;
    int* vout = v1;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        radix_hist_kernel<<<tiles, kSortThreads, 0, s>>>(n, kin, shift, tiles, hist, gate);
        radix_scan_kernel<<<256, kSortThreads, 0, s>>>(tiles, hist, hist + 256 * (size_t)tiles, gate);
        radix_scatter_kernel<<<tiles, kSortThreads, 0, s>>>(n, kin, vin, kout, vout, shift, tiles, hist,
                                                            hist + 256 * (size_t)tiles, (iota && p == 0) ? 1 : 0, gate);
        unsigned int* tk = kin; kin = kout; kout = tk;
        int* tv = vin; vin = vout; vout = tv;
    }
    return kin == k1;
}

// ---------------------------------------------------------------------------
// Exclusive prefix sum of int counts (reduce, scan of the block sums, apply).
// Each block takes kScanItems consecutive values, kScanPer per thread.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanItems = kScanThreads * kScanPer;

size_t exclusive_scan_tmp_size(int n) { return (size_t)(n + kScanItems - 1) / kScanItems + 1; }

// Block sum of 256 thread values (every thread gets it).
__device__ __forceinline__ int scan_block_total(int v, int* wsum)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    const int t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return t;
}

__global__ void __launch_bounds__(kScanThreads) scan_reduce_kernel(int n, const int* __restrict__ in,
                                                                  int* __restrict__ bsum)
{
    __shared__ int wsum[kScanThreads / 64];
    const int base = blockIdx.x * kScanItems;
    int v = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const int i = base + k * kScanThreads + (int)threadIdx.x;   // coalesced
        if (i < n) v += in[i];
    }
    const int t = scan_block_total(v, wsum);
    if (threadIdx.x == 0) bsum[blockIdx.x] = t;
}

// One block: exclusive scan of the nb block sums in place, total in bsum[nb].
__global__ void __launch_bounds__(kScanThreads) scan_top_kernel(int nb, int* __restrict__ bsum)
{
    __shared__ int wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int carry = 0;
    for (int base = 0; base < nb; base += kScanThreads) {
        const int i = base + (int)threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        int x = v;
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const int o = __shfl_up(x, m, 64);
            if (lane >= m) x += o;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        int before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        const int chunk = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
        if (i < nb) bsum[i] = before + x - v;
        carry += chunk;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

// Thread k of a block owns values base + 16 k .. base + 16 k + 15 (staged
// through LDS so that the global reads and writes stay coalesced).
__global__ void __launch_bounds__(kScanThreads) scan_apply_kernel(int n, const int* __restrict__ in,
                                                                 int* __restrict__ out, const int* __restrict__ bsum,
                                                                 int nb)
{
    __shared__ int vals[kScanItems + kScanItems / 32];   // padded: thread rows start on distinct banks
    __shared__ int wsum[kScanThreads / 64];
    const int base = blockIdx.x * kScanItems;
    auto slot = [](int k) { return k + (k >> 5); };
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const int j = k * kScanThreads + (int)threadIdx.x;
        vals[slot(j)] = (base + j < n) ? in[base + j] : 0;
    }
    __syncthreads();
    int own[kScanPer];
    int run = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        own[k] = run;
        run += vals[slot((int)threadIdx.x * kScanPer + k)];
    }
    // exclusive scan of the thread totals over the block
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = run;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const int o = __shfl_up(x, m, 64);
        if (lane >= m) x += o;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int before = bsum[blockIdx.x] + x - run;
    for (int w = 0; w < wave; ++w) before += wsum[w];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) vals[slot((int)threadIdx.x * kScanPer + k)] = before + own[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const int j = k * kScanThreads + (int)threadIdx.x;
        if (base + j < n) out[base + j] = vals[slot(j)];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = bsum[nb];
}

void launch_exclusive_scan(hipStream_t s, int n, const int* in, int* out, int* tmp)
{
    if (n <= 0) return;
    const int nb = (n + kScanItems - 1) / kScanItems;
    scan_reduce_kernel<<<nb, kScanThreads, 0, s>>>(n, in, tmp);
    scan_top_kernel<<<1, kScanThreads, 0, s>>>(nb, tmp);
    scan_apply_kernel<<<nb, kScanThreads, 0, s>>>(n, in, out, tmp, nb);
}

}  // namespace xpbd
