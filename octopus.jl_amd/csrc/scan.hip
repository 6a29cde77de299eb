// Device-wide exclusive prefix sum of uint32 values: tile sums, then tile
// scans with carry-in (two launches up to 1024 tiles, three beyond).  Used for the radix
// histograms, the BFS node numbering and the per-query edge offsets.
//
// Tile sums, carries and the total are 64-bit, so the total is exact for any
// input; the uint32 outputs are exact whenever the total is below 2^32 (the
// callers check the total before using the offsets).
#include "internal.hpp"

namespace oct {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 4096

__device__ inline uint64_t wave_incl_scan_u64(uint64_t v) {
    const int lane = lane_id();
    for (int o = 1; o < OCT_WAVE; o <<= 1) {
        const uint64_t w = __shfl_up(v, o, OCT_WAVE);
        if (lane >= o) v += w;
    }
    return v;
}

// block-wide exclusive scan of one value per thread; returns the block total
__device__ inline uint64_t block_excl_scan(uint64_t v, uint64_t* lds_waves, uint64_t& total) {
    const int lane = lane_id();
    const int w = threadIdx.x / OCT_WAVE;
    const int nw = blockDim.x / OCT_WAVE;
    const uint64_t incl = wave_incl_scan_u64(v);
    if (lane == OCT_WAVE - 1) lds_waves[w] = incl;
    __syncthreads();
    uint64_t wpre = 0, tot = 0;
    for (int i = 0; i < nw; ++i) {
        const uint64_t s = lds_waves[i];
        if (i < w) wpre += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wpre + incl - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tile_sums(const uint32_t* __restrict__ in,
                                                                 int64_t n, uint64_t* __restrict__ sums) {
    __shared__ uint64_t lw[SCAN_THREADS / OCT_WAVE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int64_t k = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
        if (k < n) s += in[k];
    }
    uint64_t tot;
    block_excl_scan(s, lw, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// one block scans the m tile sums in place and writes the 64-bit total
__global__ void __launch_bounds__(1024) k_scan_sums(uint64_t* __restrict__ sums, int64_t m,
                                                    uint64_t* __restrict__ total_out) {
    __shared__ uint64_t lw[1024 / OCT_WAVE];
    const int per = (int)((m + 1023) / 1024);
    const int64_t lo = (int64_t)threadIdx.x * per;
    uint64_t s = 0;
    for (int i = 0; i < per; ++i)
        if (lo + i < m) s += sums[lo + i];
    uint64_t tot;
    uint64_t pre = block_excl_scan(s, lw, tot);
    for (int i = 0; i < per; ++i) {
        if (lo + i < m) {
            const uint64_t v = sums[lo + i];
            sums[lo + i] = pre;
            pre += v;
        }
    }
    if (threadIdx.x == 0 && total_out) *total_out = tot;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out, int64_t n,
                                                             const uint64_t* __restrict__ sums) {
    __shared__ uint64_t lw[SCAN_THREADS / OCT_WAVE];
    __shared__ uint32_t tile[SCAN_TILE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int64_t k = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
        tile[i * SCAN_THREADS + threadIdx.x] = k < n ? in[k] : 0u;
    }
    __syncthreads();
    // each thread owns SCAN_ITEMS consecutive entries
    uint32_t v[SCAN_ITEMS];
    uint64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = tile[threadIdx.x * SCAN_ITEMS + i];
        s += v[i];
    }
    uint64_t tot;
    uint64_t pre = block_excl_scan(s, lw, tot) + sums[blockIdx.x];
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        tile[threadIdx.x * SCAN_ITEMS + i] = (uint32_t)pre;
        pre += v[i];
    }
    __syncthreads();
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int64_t k = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
        if (k < n) out[k] = tile[i * SCAN_THREADS + threadIdx.x];
    }
}

// Up to SCAN_CARRY_MAX tiles: every tile block sums the totals of the tiles
// before it itself, so the scan takes two launches.
constexpr int64_t SCAN_CARRY_MAX = 1024;
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles_carry(const uint32_t* __restrict__ in,
                                                                   uint32_t* __restrict__ out, int64_t n,
                                                                   const uint64_t* __restrict__ sums,
                                                                   uint64_t* __restrict__ total_out) {
    __shared__ uint64_t lw[SCAN_THREADS / OCT_WAVE];
    __shared__ uint32_t tile[SCAN_TILE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    uint64_t c = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += SCAN_THREADS) c += sums[t];
    uint64_t carry;
    block_excl_scan(c, lw, carry);
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int64_t k = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
        tile[i * SCAN_THREADS + threadIdx.x] = k < n ? in[k] : 0u;
    }
    __syncthreads();
    uint32_t v[SCAN_ITEMS];
    uint64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = tile[threadIdx.x * SCAN_ITEMS + i];
        s += v[i];
    }
    uint64_t tot;
    uint64_t pre = block_excl_scan(s, lw, tot) + carry;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        tile[threadIdx.x * SCAN_ITEMS + i] = (uint32_t)pre;
        pre += v[i];
    }
    __syncthreads();
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int64_t k = base + (int64_t)i * SCAN_THREADS + threadIdx.x;
        if (k < n) out[k] = tile[i * SCAN_THREADS + threadIdx.x];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0 && total_out) *total_out = carry + tot;
}

// words of uint32 scratch exclusive_scan needs for n inputs (64-bit tile sums)
int64_t scan_tmp_words(int64_t n) { return 2 * ((n + SCAN_TILE - 1) / SCAN_TILE + 1); }

hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                          uint64_t* total_out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles <= SCAN_CARRY_MAX) {
        uint64_t* sums = (uint64_t*)tmp;
        hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, in, n, sums);
        hipLaunchKernelGGL(k_scan_tiles_carry, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, in, out, n, sums,
                           total_out);
        return hipGetLastError();
    }
    if (tiles > (int64_t)1024 * 64) return hipErrorInvalidValue;
    uint64_t* sums = (uint64_t*)tmp;
    hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, tiles, total_out);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, in, out, n, sums);
    return hipGetLastError();
}

}  // namespace oct
