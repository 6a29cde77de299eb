// run! on the GPU: origin/extent reduction, Morton keys, stable LSD radix
// sort, octree topology, bounds refit.
//
// Reference behaviour (src/api.jl:194-250, src/cpu/*.jl): origin = per-axis
// min - Float32(1e-6); key = 21-bit-per-axis Morton code of
// clamp(floor((p - origin) / r)); perm sorts the keys; BFS octree where a node
// splits on the 3 key bits of its level, is a leaf iff count <= leaf size or
// level == 21, children in octant order and node ids in BFS order; node bounds
// are the exact min/max of the node's points.
//
// The topology here is not built level by level.  For sorted keys sk, let
// m_k = number of levels sk[k-1] and sk[k] share (m_0 = -1).  Position S
// starts the level-L prefix range iff m_S < L.  The node at level L starting
// at S exists iff its parent range has more than T points, and the prefix
// ranges shrink with L, so the nodes starting at S are exactly levels
// m_S < L <= lambda(S) (when S = 0 or the parent range is big), where
//   lambda(S) = min(21, max(m_S, min{m_j : S < j <= S + T}) + 1)
// is the level at which the range starting at S first holds <= T points.
// Counting those (level, S) pairs per block, one exclusive scan over
// (level, block) gives every node its BFS id directly, since BFS order is
// level-major and Morton-ordered within a level.  Four launches build the
// whole topology regardless of depth.
#include "internal.hpp"

namespace oct {

// =============================================================== bounds

// Per-block partial min/max of one set's coordinates, copied to `copy` on
// the way when asked.  Each thread takes four consecutive points per pass:
// three 16-byte loads (and stores) when the coordinates are 16-byte aligned
// (VEC), else twelve 4-byte ones.  (A last-block fold in the same launch was
// measured slower: the device-scope fences it needs write back the L2 of
// each XCD, 72 us at 1M points against 3.6 us for the separate fold.)
template <bool VEC>
__global__ void __launch_bounds__(BND_THREADS)
k_bounds_partial(const float* __restrict__ coords, int64_t n, float* __restrict__ copy,
                 uint32_t* __restrict__ part) {
    {
        uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
        uint32_t mx[3] = {0u, 0u, 0u};
        for (int64_t p0 = 4 * ((int64_t)blockIdx.x * BND_THREADS + threadIdx.x); p0 < n;
             p0 += 4 * (int64_t)gridDim.x * BND_THREADS) {
        float v[12];
        int np = 0;
        if (p0 + 4 <= n) {
            np = 4;
            if (VEC) {
                const float4* src = reinterpret_cast<const float4*>(coords + 3 * p0);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const float4 t = src[q];
                    v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
                }
                if (copy) {
                    float4* dst = reinterpret_cast<float4*>(copy + 3 * p0);
#pragma unroll
                    for (int q = 0; q < 3; ++q) dst[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 12; ++q) v[q] = coords[3 * p0 + q];
                if (copy)
#pragma unroll
                    for (int q = 0; q < 12; ++q) copy[3 * p0 + q] = v[q];
            }
        } else if (p0 < n) {  // the last, partial group
            np = (int)(n - p0);
            for (int q = 0; q < 3 * np; ++q) {
                v[q] = coords[3 * p0 + q];
                if (copy) copy[3 * p0 + q] = v[q];
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= np) break;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const uint32_t o = f2o(v[3 * k + d]);
                mn[d] = o < mn[d] ? o : mn[d];
                mx[d] = o > mx[d] ? o : mx[d];
            }
        }
        }
        __shared__ uint32_t red[6][BND_THREADS / OCT_WAVE];
        const int w = threadIdx.x / OCT_WAVE;
        for (int d = 0; d < 3; ++d) {
            const uint32_t a = wave_min_u(mn[d]);
            const uint32_t b = wave_max_u(mx[d]);
            if (lane_id() == 0) { red[d][w] = a; red[3 + d][w] = b; }
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            const int d = threadIdx.x;
            uint32_t r = red[d][0];
            for (int k = 1; k < BND_THREADS / OCT_WAVE; ++k)
                r = d < 3 ? (red[d][k] < r ? red[d][k] : r) : (red[d][k] > r ? red[d][k] : r);
            part[6 * blockIdx.x + d] = r;
        }
    }
}

// (also zeroes `zero[0 .. nz)`: the tile sums k_refit_keys adds up next; and
// with csrc, stores the coordinates' address there for the refit kernels,
// which a replayed graph runs on whatever array the caller passed this time)
__global__ void __launch_bounds__(BND_THREADS)
k_bounds_final(const uint32_t* __restrict__ partials, int nblocks, float radius, BuildParams* __restrict__ prm,
               uint64_t* __restrict__ zero, int64_t nz, const float* coords, const float** __restrict__ csrc) {
    for (int64_t i = threadIdx.x; i < nz; i += BND_THREADS) zero[i] = 0ull;
    if (csrc && threadIdx.x == 0) *csrc = coords;
    bounds_fold(partials, 0, nblocks, radius, prm);
}

// ================================================================ keys

__global__ void __launch_bounds__(256)
k_keys(const float* __restrict__ coords, int64_t n, const BuildParams* __restrict__ prm,
       uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float ox = prm->origin[0], oy = prm->origin[1], oz = prm->origin[2];
    const float inv = prm->inv_cs;
    const float px = coords[3 * i], py = coords[3 * i + 1], pz = coords[3 * i + 2];
    const uint32_t ix = floor_clamp((px - ox) * inv, 0x1fffffu);
    const uint32_t iy = floor_clamp((py - oy) * inv, 0x1fffffu);
    const uint32_t iz = floor_clamp((pz - oz) * inv, 0x1fffffu);
    keys[i] = morton3(ix, iy, iz);
    vals[i] = (int32_t)i;
}

// ========================================================== radix sort

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 4096 keys per tile
constexpr int RS_WAVES = RS_THREADS / OCT_WAVE;
constexpr int RS_BITS = 9;
constexpr int RS_BUCKETS = 1 << RS_BITS;

// Tile digit histograms, stored digit-major (hist[d * ntiles + tile]) so one
// exclusive scan yields every (digit, tile) output offset.
__global__ void __launch_bounds__(RS_THREADS)
k_rs_hist(const uint64_t* __restrict__ keys, int64_t n, int shift, uint32_t* __restrict__ hist,
          int ntiles) {
    __shared__ uint32_t h[RS_BUCKETS];
    for (int d = threadIdx.x; d < RS_BUCKETS; d += RS_THREADS) h[d] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int i = 0; i < RS_ITEMS; ++i) {
        const int64_t k = base + (int64_t)i * RS_THREADS + threadIdx.x;
        if (k < n) atomicAdd(&h[(uint32_t)(keys[k] >> shift) & (RS_BUCKETS - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RS_BUCKETS; d += RS_THREADS) hist[(int64_t)d * ntiles + blockIdx.x] = h[d];
}

// Stable scatter of one tile.  Items are ranked in index order with wave
// ballots (one "same digit" peer mask per item), staged in LDS in digit order,
// then written out so consecutive lanes write consecutive addresses of one
// digit run.  The last pass also writes perm (+ index base) and the
// Morton-sorted coordinates with the original index in w.
__global__ void __launch_bounds__(RS_THREADS)
k_rs_scatter(const uint64_t* __restrict__ kin, const int32_t* __restrict__ vin,
             uint64_t* __restrict__ kout, int32_t* __restrict__ vout, int64_t n, int shift,
             const uint32_t* __restrict__ hist_off, int ntiles, int last_pass,
             int32_t* __restrict__ perm_out, float4* __restrict__ scoord,
             const float* __restrict__ coords, int index_base) {
    __shared__ uint64_t skey[RS_TILE];
    __shared__ int32_t sval[RS_TILE];
    __shared__ uint32_t wcount[RS_WAVES][RS_BUCKETS];
    __shared__ uint32_t dstart[RS_BUCKETS];
    __shared__ uint32_t gstart[RS_BUCKETS];

    const int lane = lane_id();
    const int w = threadIdx.x / OCT_WAVE;
    for (int d = threadIdx.x; d < RS_BUCKETS; d += RS_THREADS) {
        for (int k = 0; k < RS_WAVES; ++k) wcount[k][d] = 0;
        gstart[d] = hist_off[(int64_t)d * ntiles + blockIdx.x];
    }
    __syncthreads();

    const int64_t tile_base = (int64_t)blockIdx.x * RS_TILE;
    const int64_t wave_base = tile_base + (int64_t)w * (RS_ITEMS * OCT_WAVE);
    uint64_t key[RS_ITEMS];
    int32_t val[RS_ITEMS];
    uint32_t lpos[RS_ITEMS];  // position within (wave, digit)
    for (int t = 0; t < RS_ITEMS; ++t) {
        const int64_t k = wave_base + (int64_t)t * OCT_WAVE + lane;
        const bool valid = k < n;
        key[t] = valid ? kin[k] : 0ull;
        val[t] = valid ? vin[k] : 0;
        const uint32_t d = (uint32_t)(key[t] >> shift) & (RS_BUCKETS - 1);
        uint64_t peers = ballot(valid);
        for (int b = 0; b < RS_BITS; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)mbcnt(peers);
        const uint32_t old = wcount[w][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) wcount[w][d] = old + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        lpos[t] = old + rank;
    }
    __syncthreads();
    // digit starts inside the tile and per-wave offsets inside each digit
    for (int d = threadIdx.x; d < RS_BUCKETS; d += RS_THREADS) {
        uint32_t run = 0;
        for (int k = 0; k < RS_WAVES; ++k) {
            const uint32_t c = wcount[k][d];
            wcount[k][d] = run;
            run += c;
        }
        dstart[d] = run;  // temporarily the digit total
    }
    __syncthreads();
    if (threadIdx.x < OCT_WAVE) {
        // exclusive scan of the digit totals by one wave (RS_PER_LANE per lane)
        constexpr int RS_PER_LANE = RS_BUCKETS / OCT_WAVE;
        uint32_t tv[RS_PER_LANE], s = 0;
        for (int j = 0; j < RS_PER_LANE; ++j) { tv[j] = dstart[RS_PER_LANE * lane + j]; s += tv[j]; }
        const uint32_t incl = (uint32_t)wave_incl_scan((int)s);
        uint32_t pre = incl - s;
        for (int j = 0; j < RS_PER_LANE; ++j) { dstart[RS_PER_LANE * lane + j] = pre; pre += tv[j]; }
    }
    __syncthreads();
    for (int t = 0; t < RS_ITEMS; ++t) {
        const int64_t k = wave_base + (int64_t)t * OCT_WAVE + lane;
        if (k < n) {
            const uint32_t d = (uint32_t)(key[t] >> shift) & (RS_BUCKETS - 1);
            const uint32_t p = dstart[d] + wcount[w][d] + lpos[t];
            skey[p] = key[t];
            sval[p] = val[t];
        }
    }
    __syncthreads();
    const int64_t tile_n = n - tile_base < RS_TILE ? n - tile_base : RS_TILE;
    for (int p = threadIdx.x; p < tile_n; p += RS_THREADS) {
        const uint64_t kk = skey[p];
        const int32_t vv = sval[p];
        const uint32_t d = (uint32_t)(kk >> shift) & (RS_BUCKETS - 1);
        const int64_t g = (int64_t)gstart[d] + (p - (int64_t)dstart[d]);
        kout[g] = kk;
        if (!last_pass) {
            vout[g] = vv;
        } else {
            perm_out[g] = vv + index_base;
            const float* c = coords + 3 * (int64_t)vv;
            scoord[g] = make_float4(c[0], c[1], c[2], __int_as_float(vv));
        }
    }
}

// ================================================== almost-sorted refit
//
// Time stepping moves few points across a cell boundary, so the previous
// sorted order (keys PK, point indices PI, sorted by (key, index)) is almost
// the new one.  Points whose key did not change keep their relative order;
// only the changed points C are placed: each one's rank within C (by key,
// then index) plus the number of unchanged points before it, found by binary
// search in the previous order.  The result is exactly the stable sort's
// order, so the tree is the one the full sort would build.

constexpr int RF_THREADS = 256;
constexpr int RF_TILE = 1024;  // C entries staged in LDS per round of the rank kernel

__device__ inline bool key_idx_less(uint64_t ka, int32_t ia, uint64_t kb, int32_t ib) {
    return ka < kb || (ka == kb && ia < ib);
}

// K1: new keys in the previous order (pi = PI, and per group of 64
// positions the mask of those whose key changed, wmask) and the changed
// count of each tile of RF_TILE_POS positions (tile sums, zeroed by the
// bounds fold).
constexpr int RF_TILE_POS = 4096;
constexpr int RF_TILE_GROUPS = RF_TILE_POS / OCT_WAVE;  // 64: one lane each in K3
static_assert(RF_TILE_GROUPS == OCT_WAVE, "K3 gives each group of a tile one lane");
__global__ void __launch_bounds__(RF_THREADS)
k_refit_keys(const float* const* __restrict__ csrc, int64_t n, const BuildParams* __restrict__ prm,
             const uint64_t* __restrict__ pk, const int32_t* __restrict__ prev_perm, int index_base,
             int32_t* __restrict__ pi, RefitLook lk) {
    const float* __restrict__ coords = *csrc;  // this step's coordinates
    const int64_t i = (int64_t)blockIdx.x * RF_THREADS + threadIdx.x;
    bool changed = false;
    if (i < n) {
        const int32_t idx = prev_perm[i] - index_base;
        const float ox = prm->origin[0], oy = prm->origin[1], oz = prm->origin[2];
        const float inv = prm->inv_cs;
        const float px = coords[3 * (int64_t)idx], py = coords[3 * (int64_t)idx + 1],
                    pz = coords[3 * (int64_t)idx + 2];
        const uint64_t k = morton3(floor_clamp((px - ox) * inv, 0x1fffffu),
                                   floor_clamp((py - oy) * inv, 0x1fffffu), floor_clamp((pz - oz) * inv, 0x1fffffu));
        pi[i] = idx;
        changed = k != pk[i];
    }
    // this block's changed count into its tile's sum
    __shared__ uint32_t bc;
    if (threadIdx.x == 0) bc = 0;
    __syncthreads();
    const uint64_t m = ballot(changed);
    const int64_t g = (int64_t)blockIdx.x * (RF_THREADS / OCT_WAVE) + (int)threadIdx.x / OCT_WAVE;
    if (lane_id() == 0) {
        if (g * OCT_WAVE < n) lk.wmask[g] = m;
        if (m) atomicAdd(&bc, (uint32_t)__popcll(m));
    }
    __syncthreads();
    if (threadIdx.x == 0 && bc) {
        const int64_t tile = ((int64_t)blockIdx.x * RF_THREADS) / RF_TILE_POS;
        atomicAdd(reinterpret_cast<unsigned long long*>(lk.tsum + tile), (unsigned long long)bc);
    }
}

// K3: one wave per tile: the changed count before the tile (its tile sums
// below it), before each of its groups (wpre), and the changed points C
// compacted in the previous order (ck, ci; their new keys computed again, as
// K1 did).  Tile 0 records the total and whether the refit can run (within
// `limit`, key width unchanged) for the kernels after it and the host.
__global__ void __launch_bounds__(OCT_WAVE)
k_refit_compact(const float* const* __restrict__ csrc, int64_t n, BuildParams* __restrict__ prm,
                const int32_t* __restrict__ pi, RefitLook lk, uint64_t* __restrict__ ck, int32_t* __restrict__ ci,
                int64_t limit, int key_bits, uint64_t* __restrict__ total) {
    const int64_t ntiles = (n + RF_TILE_POS - 1) / RF_TILE_POS;
    const int64_t t = blockIdx.x;
    const int lane = lane_id();
    uint64_t below = 0, all = 0;
    for (int64_t u = lane; u < ntiles; u += OCT_WAVE) {
        const uint64_t v = lk.tsum[u];
        all += v;
        below += u < t ? v : 0ull;
    }
    below = wave_sum_u64(below);
    all = wave_sum_u64(all);
    const int64_t g = t * RF_TILE_GROUPS + lane;
    const uint64_t m = g * OCT_WAVE < n ? lk.wmask[g] : 0ull;
    const int c = __popcll(m);
    const uint64_t gpre = below + (uint64_t)(wave_incl_scan(c) - c);
    if (g * OCT_WAVE < n) lk.wpre[g] = (uint32_t)gpre;
    if (t == 0 && lane == 0) {
        *total = all;
        prm->changed = (uint32_t)(all < 0xffffffffull ? all : 0xffffffffull);
        prm->refit_fail = (all > (uint64_t)limit || prm->key_bits != key_bits) ? 1u : 0u;
    }
    if (!m || all > (uint64_t)limit) return;
    const float* __restrict__ coords = *csrc;
    const float ox = prm->origin[0], oy = prm->origin[1], oz = prm->origin[2];
    const float inv = prm->inv_cs;
    uint64_t rest = m;
    for (uint64_t pos = gpre; rest; ++pos) {
        const int b = __builtin_ctzll(rest);
        rest &= rest - 1ull;
        const int64_t i = g * OCT_WAVE + b;
        const int32_t idx = pi[i];
        const float* p = coords + 3 * (int64_t)idx;
        ck[pos] = morton3(floor_clamp((p[0] - ox) * inv, 0x1fffffu), floor_clamp((p[1] - oy) * inv, 0x1fffffu),
                          floor_clamp((p[2] - oz) * inv, 0x1fffffu));
        ci[pos] = idx;
    }
}

// changed positions of the previous order before position i (i <= n)
__device__ inline uint32_t changed_before(const RefitLook& lk, int64_t i, int64_t n, uint32_t nc) {
    if (i >= n) return nc;
    const uint64_t m = lk.wmask[i / OCT_WAVE];
    const int l = (int)(i % OCT_WAVE);
    return lk.wpre[i / OCT_WAVE] + (uint32_t)__popcll(m & ((1ull << l) - 1ull));
}

// K4: rank of each changed point within C (all-pairs count, C staged in LDS
// tiles) and its number of unchanged predecessors (a search of the previous
// order); writes C sorted (csk, csi) with final positions cspos.  While the
// grid holds RF_LANES threads per changed point (the usual case: a time step
// changes few keys), the RF_LANES consecutive lanes of a point split its
// count and search the previous order RF_LANES-ary: each round probes
// RF_LANES positions at once and keeps the gap between the last probe below
// the point and the first one not below, so 1M positions take 5 rounds of
// dependent loads instead of 20.  Otherwise one thread per point.
constexpr int RF_LANES = 16;
__global__ void __launch_bounds__(RF_THREADS)
k_refit_rank(int64_t n, const uint64_t* __restrict__ total, const uint64_t* __restrict__ ck,
             const int32_t* __restrict__ ci, const uint64_t* __restrict__ pk, const int32_t* __restrict__ pi,
             RefitLook lk, uint64_t* __restrict__ csk, int32_t* __restrict__ csi,
             uint32_t* __restrict__ cspos, const BuildParams* __restrict__ prm) {
    __shared__ uint64_t tk[RF_TILE];
    __shared__ int32_t ti[RF_TILE];
    const int64_t nc = (int64_t)*total;
    const bool wide = nc * RF_LANES <= (int64_t)gridDim.x * RF_THREADS;
    const int S = wide ? RF_LANES : 1;
    const int64_t first = (int64_t)blockIdx.x * (RF_THREADS / S);  // this block's first point
    // block-uniform exits (before any barrier): a failed refit, or a block past the changed points
    if (prm->refit_fail || first >= nc) return;
    const int sub = (int)threadIdx.x % S;
    const int64_t c = first + (int64_t)threadIdx.x / S;
    const bool mine = c < nc;  // the same for the S lanes of a point
    const uint64_t key = mine ? ck[c] : 0ull;
    const int32_t idx = mine ? ci[c] : 0;
    uint32_t rank = 0;
    for (int64_t t0 = 0; t0 < nc; t0 += RF_TILE) {
        const int tn = (int)(nc - t0 < RF_TILE ? nc - t0 : RF_TILE);
        __syncthreads();
        for (int j = threadIdx.x; j < tn; j += RF_THREADS) {
            tk[j] = ck[t0 + j];
            ti[j] = ci[t0 + j];
        }
        __syncthreads();
        if (mine)
            for (int j = sub; j < tn; j += S) rank += key_idx_less(tk[j], ti[j], key, idx) ? 1u : 0u;
    }
    int64_t lo = 0, hi = n;  // first position of the previous order not below (key, idx)
    if (wide) {              // block-uniform
#pragma unroll
        for (int o = 1; o < RF_LANES; o <<= 1) rank += (uint32_t)__shfl_xor((int)rank, o, RF_LANES);
        const int g0 = (lane_id() / RF_LANES) * RF_LANES;  // the point's first lane in the wave
        while (mine && lo < hi) {  // the same for the point's lanes
            const int64_t m = hi - lo;
            const int64_t p = lo + ((int64_t)(sub + 1) * m) / (RF_LANES + 1);  // in [lo, hi), ascending
            const bool below = key_idx_less(pk[p], pi[p], key, idx);
            const int t = __popc((uint32_t)((ballot(below) >> g0) & ((1ull << RF_LANES) - 1ull)));
            const int64_t nlo = t > 0 ? lo + ((int64_t)t * m) / (RF_LANES + 1) + 1 : lo;
            const int64_t nhi = t < RF_LANES ? lo + ((int64_t)(t + 1) * m) / (RF_LANES + 1) : hi;
            lo = nlo;
            hi = nhi;
        }
        if (sub != 0) return;
    } else {
        while (mine && lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (key_idx_less(pk[mid], pi[mid], key, idx)) lo = mid + 1;
            else hi = mid;
        }
    }
    if (!mine) return;
    const uint32_t before_changed = changed_before(lk, lo, n, (uint32_t)nc);
    const uint32_t unchanged_before = (uint32_t)lo - before_changed;
    csk[rank] = key;
    csi[rank] = idx;
    cspos[rank] = unchanged_before + rank;
}

__device__ inline void refit_write(int64_t pos, uint64_t key, int32_t idx, const float* __restrict__ coords,
                                   uint64_t* __restrict__ kout, int32_t* __restrict__ perm, float4* __restrict__ sc,
                                   int index_base) {
    kout[pos] = key;
    perm[pos] = idx + index_base;
    const float* p = coords + 3 * (int64_t)idx;
    sc[pos] = make_float4(p[0], p[1], p[2], __int_as_float(idx));
}

// K5: unchanged points to their new positions; K6 (c < |C|): changed points.
// An unchanged point moves by the changed points sorted below it, found by
// binary search of C sorted; while C has at most RF_PLACE_LDS points (a time
// step changes few) each block stages it in LDS first, so the search's ten
// or so dependent steps are LDS reads rather than global loads.
constexpr int RF_PLACE_LDS = 1024;
__global__ void __launch_bounds__(RF_THREADS)
k_refit_place(const float* const* __restrict__ csrc, int64_t n, const uint64_t* __restrict__ total,
              const uint64_t* __restrict__ pk, const int32_t* __restrict__ pi, RefitLook lk,
              const uint64_t* __restrict__ csk, const int32_t* __restrict__ csi, const uint32_t* __restrict__ cspos,
              uint64_t* __restrict__ kout, int32_t* __restrict__ perm, float4* __restrict__ sc, int index_base,
              const BuildParams* __restrict__ prm) {
    __shared__ uint64_t s_ck[RF_PLACE_LDS];
    __shared__ int32_t s_ci[RF_PLACE_LDS];
    const float* __restrict__ coords = *csrc;  // this step's coordinates
    const int64_t nc = (int64_t)*total;
    const int64_t i = (int64_t)blockIdx.x * RF_THREADS + threadIdx.x;
    if (prm->refit_fail) {  // keep the previous order (perm and scoord stay as they are)
        if (i < n) kout[i] = pk[i];
        return;
    }
    // loads that do not depend on C first
    uint64_t wm = 0, key = 0;
    int32_t idx = 0;
    if (i < n) {
        wm = lk.wmask[i / OCT_WAVE];
        key = pk[i];
        idx = pi[i];
    }
    const bool staged = nc <= RF_PLACE_LDS;  // block-uniform
    if (staged) {
        for (int j = threadIdx.x; j < (int)nc; j += RF_THREADS) {
            s_ck[j] = csk[j];
            s_ci[j] = csi[j];
        }
        __syncthreads();
    }
    if (i < nc) refit_write(cspos[i], csk[i], csi[i], coords, kout, perm, sc, index_base);
    if (i >= n) return;
    const int l = (int)(i % OCT_WAVE);
    if ((wm >> l) & 1ull) return;  // changed: placed above
    const uint32_t c = lk.wpre[i / OCT_WAVE] + (uint32_t)__popcll(wm & ((1ull << l) - 1ull));
    int64_t lo = 0, hi = nc;  // changed points below (key, idx)
    if (staged) {
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (key_idx_less(s_ck[mid], s_ci[mid], key, idx)) lo = mid + 1;
            else hi = mid;
        }
    } else {
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (key_idx_less(csk[mid], csi[mid], key, idx)) lo = mid + 1;
            else hi = mid;
        }
    }
    refit_write((int64_t)(i - c) + lo, key, idx, coords, kout, perm, sc, index_base);
}

// ============================================================ topology

constexpr int TOPO_THREADS = 256;
constexpr int TOPO_MAX_T = 256;  // largest leaf size the window supports

// Per sorted position k: m_k (levels shared with the previous key), lambda
// and whether nodes start at k, plus the length of the leaf starting at k;
// per-block counts of nodes per level (rows 0..21) and of leaves (row 22).
// Every window test is one comparison of two keys: for sorted keys the levels
// keys i and j share (i < j) are the minimum of m over (i, j], so the minimum
// of m over (k, k+T] is key_shared(k, k+T), and the level-P node holding k
// spans the positions whose keys share at least P levels with key k, found by
// binary searches over the block's keys in LDS.
constexpr uint32_t TOPO_LEN_SHIFT = 11;  // topo bits 11..19: leaf length (0: more than T)
constexpr uint32_t EMIT_LIST = 512;      // k_topo_emit's nodes per round
__global__ void __launch_bounds__(TOPO_THREADS)
k_topo_count(const uint64_t* __restrict__ sk, int64_t n, int T, int batched, uint32_t* __restrict__ topo,
             uint32_t* __restrict__ hist, int nblk) {
    __shared__ uint64_t kk[TOPO_THREADS + 2 * TOPO_MAX_T + 1];  // keys of positions lo .. lo + span - 1
    // the block's node starts per level as differences: a position starting
    // nodes at levels mk+1 .. lam adds 1 at mk+1 and takes 1 at lam+1; [23]
    // counts the leaves (one per starting position)
    __shared__ int32_t dl[OCT_NLEVELS + 2];
    if (threadIdx.x < OCT_NLEVELS + 2) dl[threadIdx.x] = 0;
    const int64_t base = (int64_t)blockIdx.x * TOPO_THREADS;
    const int64_t lo = base - T - 1;
    const int span = TOPO_THREADS + 2 * T + 1;
    for (int i = threadIdx.x; i < span; i += TOPO_THREADS) {
        const int64_t j = lo + i;
        kk[i] = (j >= 0 && j < n) ? sk[j] : 0ull;  // outside [0, n): never compared
    }
    __syncthreads();
    const int64_t k = base + threadIdx.x;
    const int li = threadIdx.x + T + 1;  // kk[li] is key k
    int mk = -1, lam = -1, len = 0;
    bool emit = false;
    if (k < n) {
        const uint64_t key = kk[li];
        mk = k > 0 ? key_shared(kk[li - 1], key, batched) : -1;  // -1 where a set starts
        // lambda: first level whose range starting at k holds <= T points.
        // A repeat of the previous key (mk == 21) starts no node.
        int mw;
        if (mk >= OCT_MAX_LEVEL) mw = OCT_MAX_LEVEL;
        else if (k + T >= n) mw = -1;
        else mw = key_shared(key, kk[li + T], batched);
        lam = min(OCT_MAX_LEVEL, max(mk, mw) + 1);
        // keys sharing at least P levels with key k lie in [key & mask, key | ~mask]
        // (a forest compares its set bits from level 0 and its Morton bits,
        // which start at level 7, from there on)
        auto mask = [&](int P) {
            const int sh = batched ? min(63 - 3 * P, 42) : 63 - 3 * P;
            return ~0ull << sh;
        };
        const int e_n = (int)min<int64_t>(n - lo, span);  // index of position n, if staged
        if (mk < 0) {  // position 0 or the first point of a set: its root chain starts here
            emit = true;
        } else if (lam > mk) {
            // nodes start here iff the level-mk node [s, e) holding k has
            // more than T points: decided by two keys when s < k - T or
            // e > k + T, else s and e are searched (both within T of k)
            const uint64_t mp = mask(mk), klo = key & mp, khi = key | ~mp;
            emit = (k + T < n && mw >= mk) || (k > T && kk[li - T - 1] >= klo);
            if (!emit) {
                int a = li - (int)min<int64_t>(k, T), b = li;  // s: first index with key >= klo
                while (a < b) {
                    const int mid = (a + b) >> 1;
                    if (kk[mid] >= klo) b = mid;
                    else a = mid + 1;
                }
                const int s_i = a;
                a = li + 1;  // e: first index past li that is n or holds a key > khi
                b = min(li + T, e_n);
                while (a < b) {
                    const int mid = (a + b) >> 1;
                    if (kk[mid] > khi) b = mid;
                    else a = mid + 1;
                }
                emit = a - s_i > T;
            }
        }
        if (emit) {
            // the leaf starting at k (level lam) ends at the first position
            // past k that is n or holds a key above the leaf's range; within
            // T unless lam == 21 holds more than T equal keys
            const uint64_t khi = key | ~mask(lam);
            int a = li + 1, b = min(li + T, e_n);
            if (b == e_n || kk[b] > khi) {
                while (a < b) {
                    const int mid = (a + b) >> 1;
                    if (kk[mid] > khi) b = mid;
                    else a = mid + 1;
                }
                len = a - li;
            }
        }
    }
    // m is stored +1 so k == 0 (m = -1) fits the low 5 bits
    if (k < n)
        topo[k] = (uint32_t)((mk + 1) | (lam << 5) | (emit ? (1 << 10) : 0)) | ((uint32_t)len << TOPO_LEN_SHIFT);
    if (emit) {  // mk + 1 >= 0 and lam + 1 <= OCT_NLEVELS
        atomicAdd(&dl[mk + 1], 1);
        atomicSub(&dl[lam + 1], 1);
        atomicAdd(&dl[OCT_NLEVELS + 1], 1);
    }
    __syncthreads();
    if (threadIdx.x <= OCT_NLEVELS) {
        int32_t s = 0;
        if (threadIdx.x < OCT_NLEVELS)
            for (int l = 0; l <= (int)threadIdx.x; ++l) s += dl[l];
        else
            s = dl[OCT_NLEVELS + 1];
        hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = (uint32_t)s;
    }
}

// The topology table holds one row per level (0..21) and one for the leaves
// (22), each with one count per k_topo_count block.  One block per row scans
// it in place (exclusive, in chunks of TR_THREADS * TR_ITEMS with the carry
// kept) and writes the row's total after the table (row 23): k_topo_emit
// turns the totals of the rows before a level into the level's first id.
constexpr int TR_THREADS = 1024;
constexpr int TR_ITEMS = 4;
__global__ void __launch_bounds__(TR_THREADS) k_topo_rows(uint32_t* __restrict__ hist, int nblk) {
    __shared__ uint32_t s_w[TR_THREADS / OCT_WAVE];
    uint32_t* row = hist + (int64_t)blockIdx.x * nblk;
    const int lane = lane_id(), w = (int)threadIdx.x / OCT_WAVE;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < nblk; c0 += TR_THREADS * TR_ITEMS) {  // block-uniform
        uint32_t v[TR_ITEMS];
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < TR_ITEMS; ++q) {
            const int i = c0 + (int)threadIdx.x * TR_ITEMS + q;
            v[q] = i < nblk ? row[i] : 0u;
            sum += v[q];
        }
        const uint32_t incl = (uint32_t)wave_incl_scan((int)sum);
        if (lane == OCT_WAVE - 1) s_w[w] = incl;
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
        for (int i = 0; i < TR_THREADS / OCT_WAVE; ++i) {
            const uint32_t x = s_w[i];
            wpre += i < w ? x : 0u;
            tot += x;
        }
        uint32_t pre = carry + wpre + incl - sum;
#pragma unroll
        for (int q = 0; q < TR_ITEMS; ++q) {
            const int i = c0 + (int)threadIdx.x * TR_ITEMS + q;
            if (i < nblk) row[i] = pre;
            pre += v[q];
        }
        carry += tot;
        __syncthreads();  // s_w is rewritten by the next chunk
    }
    if (threadIdx.x == 0) hist[(int64_t)(OCT_NLEVELS + 1) * nblk + blockIdx.x] = carry;
}

// first j in [lo, hi) with pred(j) true, pred monotone false..true; hi if none
template <typename P>
__device__ inline int64_t lower_bound_pred(int64_t lo, int64_t hi, P pred) {
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (pred(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Emits every node starting at each sorted position, with no search: the
// BFS id of the node at (level L, position k) is the scanned (L, block)
// offset plus its rank among the block's level-L starts.
// The block lists its nodes in LDS, level-major and in position order (in
// rounds of EMIT_LIST entries), and one thread per node writes it, so
// consecutive threads write consecutive ids.  Per node: node_first, level, the packed
// record, its octant in its parent and the id its first child would have (a
// node's children are the ids from its own to the next node's, so
// k_link_refit writes the child slots and counts from these two arrays);
// leaves also get their range end, their bounds (the exact min / max of their
// points) and their ancestor at every refit level (ranc, level-major, leaf
// order), and the nodes covering all of a set's points (its chain down to the
// first split) get the set's bounds.  Internal bounds and range ends come
// from k_link_refit and k_refit_merge.
__global__ void __launch_bounds__(TOPO_THREADS)
k_topo_emit(const uint64_t* __restrict__ sk, int64_t n, int T,
            const uint32_t* __restrict__ topo, const uint32_t* __restrict__ hoff, int nblk,
            BuildParams* __restrict__ prm, Forest forest, int index_base, const float4* __restrict__ scoord, NodeRec* __restrict__ nodes, int32_t* __restrict__ nfirst,
            int32_t* __restrict__ nlast, float* __restrict__ bmin, float* __restrict__ bmax,
            uint8_t* __restrict__ nlevel, int32_t* __restrict__ nfc,
            uint8_t* __restrict__ noct, int32_t* __restrict__ leaf_ids, NodeRec* __restrict__ leaves,
            int32_t* __restrict__ ranc, int64_t astride) {
    constexpr int NW = TOPO_THREADS / OCT_WAVE;
    __shared__ uint32_t wc[NW][OCT_NLEVELS + 1];
    __shared__ uint64_t wmask[NW][OCT_NLEVELS + 1];  // per wave and level: lanes starting a node
    __shared__ uint32_t woff[NW][OCT_NLEVELS + 1];   // ids before the wave's first such lane
    __shared__ uint32_t s_loff[OCT_NLEVELS + 1];     // the block's nodes at levels below L
    __shared__ int32_t s_M, s_nl;                    // nodes and leaves in the tree
    __shared__ uint32_t s_topo[TOPO_THREADS];
    __shared__ uint64_t s_key[TOPO_THREADS];
    __shared__ uint16_t s_list[EMIT_LIST];           // position | level << 8, one round's nodes
    const int64_t base = (int64_t)blockIdx.x * TOPO_THREADS;
    // loads that do not depend on the ranks are issued first, so their
    // latencies overlap the ballots and barriers below
    // row L of the table: this block's offset within the row, and the row's total
    const uint32_t hrel = threadIdx.x <= OCT_NLEVELS ? hoff[(int64_t)threadIdx.x * nblk + blockIdx.x] : 0u;
    const uint32_t rtot = threadIdx.x <= OCT_NLEVELS ? hoff[(int64_t)(OCT_NLEVELS + 1) * nblk + threadIdx.x] : 0u;
    const int64_t k0 = base + threadIdx.x;
    const uint64_t key0 = k0 < n ? sk[k0] : 0ull;
    int64_t s_lo1 = 0, s_hi1 = 0;  // one set: its range and chain depth
    int L01 = 0;
    if (!forest.batched && n > 0) {
        s_lo1 = forest.set_off[0];
        s_hi1 = forest.set_off[1];
        L01 = key_shared(sk[s_lo1], sk[s_hi1 - 1], 0);
    }
    uint32_t t0 = 0;
    if (k0 < n) t0 = topo[k0];
    s_topo[threadIdx.x] = t0;
    s_key[threadIdx.x] = key0;
    const int mk0 = (int)(t0 & 31) - 1, lam0 = (t0 >> 5) & 31;
    const bool emit0 = k0 < n && ((t0 >> 10) & 1);
    const int w0 = threadIdx.x / OCT_WAVE;
    const int lane = lane_id();
    for (int L = 0; L <= OCT_NLEVELS; ++L) {
        const bool ind = L < OCT_NLEVELS ? (emit0 && mk0 < L && L <= lam0) : emit0;
        const uint64_t b = ballot(ind);
        if (lane == 0) {
            wmask[w0][L] = b;
            wc[w0][L] = (uint32_t)__popcll(b);
        }
    }
    __syncthreads();
    if (threadIdx.x < OCT_WAVE) {  // the block's offset per level, then each wave's
        const int L = threadIdx.x;
        // ids before level L: the totals of the rows before it (row 22, the
        // leaves, then starts at the node count)
        const int rb = wave_incl_scan((int)rtot) - (int)rtot;
        if (L == OCT_NLEVELS) {
            s_M = rb;
            s_nl = (int)rtot;
        }
        const uint32_t hrow = hrel + (uint32_t)rb;
        uint32_t s = hrow;
        if (L <= OCT_NLEVELS)
            for (int q = 0; q < NW; ++q) {
                woff[q][L] = s;
                s += wc[q][L];
            }
        // the block's nodes at levels below L (row 22, the leaves, is no
        // level); the scan runs with the whole wave active
        const int c = L < OCT_NLEVELS ? (int)(s - hrow) : 0;
        const int incl = wave_incl_scan(c);
        if (L <= OCT_NLEVELS) s_loff[L] = (uint32_t)(incl - c);
    }
    __syncthreads();
    const int32_t M = s_M;
    if (k0 == 0) {
        prm->n_nodes = M;
        prm->n_leaves = s_nl;
        nfc[M] = M;  // the first child of a node past the last
    }
    // count of level-L starts before position base + p, plus the level's offset
    auto before = [&](int L, int p) -> int32_t {
        const int w = p / OCT_WAVE, l = p % OCT_WAVE;
        return (int32_t)(woff[w][L] + (uint32_t)__popcll(wmask[w][L] & ((1ull << l) - 1ull)));
    };
    const uint32_t nb = s_loff[OCT_NLEVELS];  // the block's nodes (block-uniform)
    for (uint32_t r0 = 0; r0 < nb; r0 += EMIT_LIST) {
        if (emit0)
            for (int L = mk0 + 1; L <= lam0; ++L) {
                const uint32_t slot = s_loff[L] + (uint32_t)(before(L, threadIdx.x) - (int32_t)woff[0][L]);
                if (slot >= r0 && slot < r0 + EMIT_LIST) s_list[slot - r0] = (uint16_t)(threadIdx.x | (L << 8));
            }
        __syncthreads();
        const uint32_t nr = min(nb - r0, (uint32_t)EMIT_LIST);
        for (uint32_t j = threadIdx.x; j < nr; j += TOPO_THREADS) {
            const int p = s_list[j] & 255, L = s_list[j] >> 8;
            const int32_t id = before(L, p);
            const uint32_t t = s_topo[p];
            const int mk = (int)(t & 31) - 1, lam = (t >> 5) & 31, len = (int)(t >> TOPO_LEN_SHIFT);
            const int64_t k = base + p;
            const uint64_t key = s_key[p];
            const int Lfirst = mk + 1;  // 0 where a set starts
            // the set of this position, its first point and (at its first
            // point) its end and chain depth: levels 0..L0 hold all its points
            const int set = key_set(key, forest.batched);
            int64_t s_lo = s_lo1, s_hi = s_hi1;
            int L0 = L01;
            if (forest.batched) {
                s_lo = forest.set_off[set];
                if (k == s_lo) {
                    s_hi = forest.set_off[set + 1];
                    L0 = key_shared(sk[s_lo], sk[s_hi - 1], 1);
                }
            }
            BuildParams* sp = forest.sprm + set;
            if (k == s_lo && L == 0) sp->chain_level = L0;
            const bool chain = k == s_lo && L <= L0;
            nlevel[id] = (uint8_t)L;
            nfirst[id] = (int32_t)k + index_base;
            noct[id] = L == 0 ? (uint8_t)0 : (uint8_t)((key_morton(key, forest.batched) >> level_shift(L - 1)) & 7ull);
            const int32_t fc = L < OCT_MAX_LEVEL ? before(L + 1, p) : M;
            nfc[id] = fc;
            NodeRec rec;
            if (L == lam) {
                // leaf: its length from k_topo_count, else (0) a level-21 leaf
                // holding more than T equal keys
                const int64_t e = len > 0 ? k + len
                                          : lower_bound_pred(k + T + 1, n, [&](int64_t j) { return sk[j] != key; });
                const int32_t cnt_pts = (int32_t)(e - k);
                nlast[id] = (int32_t)(e - 1) + index_base;
                // the exact min / max of its points, four loads in flight (a
                // short last group re-reads the first point; eight took the
                // kernel to 77 VGPRs and six waves per SIMD: 21.0 -> 20.4 us)
                uint32_t lo3[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
                uint32_t hi3[3] = {0u, 0u, 0u};
                for (int32_t j0 = 0; j0 < cnt_pts; j0 += 4) {
                    float4 pp[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) pp[q] = scoord[k + (j0 + q < cnt_pts ? j0 + q : 0)];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float c3[3] = {pp[q].x, pp[q].y, pp[q].z};
                        for (int d = 0; d < 3; ++d) {
                            const uint32_t o = f2o(c3[d]);
                            lo3[d] = o < lo3[d] ? o : lo3[d];
                            hi3[d] = o > hi3[d] ? o : hi3[d];
                        }
                    }
                }
                rec.lo = make_float4(o2f(lo3[0]), o2f(lo3[1]), o2f(lo3[2]), __int_as_float((int)k));
                rec.hi = make_float4(o2f(hi3[0]), o2f(hi3[1]), o2f(hi3[2]), __int_as_float(cnt_pts));
                const int32_t lr = before(OCT_NLEVELS, p) - M;
                // the leaf's ancestor at each level the refit visits (-1 from its
                // own level on): the last node of that level starting at or before k
                for (int La = forest.batched ? 0 : L0 + 1; La < OCT_MAX_LEVEL; ++La)
                    ranc[(int64_t)La * astride + lr] = La >= lam ? -1 : before(La, p) - (La >= Lfirst ? 0 : 1);
                leaf_ids[lr] = id;
                leaves[lr] = rec;
                if (chain) sp->branch_root = id;  // the whole set sits in one leaf
            } else {
                rec.lo = make_float4(0.f, 0.f, 0.f, __int_as_float(fc));  // first child
                rec.hi = make_float4(0.f, 0.f, 0.f, __int_as_float(0));   // -children, from k_link_refit
                if (chain) {
                    rec.lo.x = o2f(sp->omin[0]); rec.lo.y = o2f(sp->omin[1]); rec.lo.z = o2f(sp->omin[2]);
                    rec.hi.x = o2f(sp->omax[0]); rec.hi.y = o2f(sp->omax[1]); rec.hi.z = o2f(sp->omax[2]);
                    nlast[id] = (int32_t)(s_hi - 1) + index_base;
                    if (L == L0) sp->branch_root = id;
                }
            }
            nodes[id] = rec;
            if (L == lam || chain) {
                bmin[3 * (int64_t)id + 0] = rec.lo.x; bmin[3 * (int64_t)id + 1] = rec.lo.y; bmin[3 * (int64_t)id + 2] = rec.lo.z;
                bmax[3 * (int64_t)id + 0] = rec.hi.x; bmax[3 * (int64_t)id + 1] = rec.hi.y; bmax[3 * (int64_t)id + 2] = rec.hi.z;
            }
        }
        __syncthreads();  // s_list is rewritten by the next round
    }
}

// Internal bounds and range ends travel as seven maxima: ~f2o of the lower
// bounds, f2o of the upper ones, the range end.  A node's results, stored.
__device__ inline void store_refit(int32_t id, const uint32_t* v, NodeRec* __restrict__ nodes,
                                   int32_t* __restrict__ nlast, float* __restrict__ bmin, float* __restrict__ bmax) {
    float* nr = reinterpret_cast<float*>(nodes + id);  // x, y, z of lo and hi: the counts in w stay
    for (int d = 0; d < 3; ++d) {
        const float lo = o2f(~v[d]), hi = o2f(v[3 + d]);
        nr[d] = lo;
        nr[4 + d] = hi;
        bmin[3 * (int64_t)id + d] = lo;
        bmax[3 * (int64_t)id + d] = hi;
    }
    nlast[id] = (int32_t)v[6];
}

// An unfinished run: the node (-1: none) and its seven maxima.
constexpr int REFIT_SLOT_WORDS = 8;
constexpr int LR_THREADS = 512;  // k_link_refit's blocks; the refit takes LR_THREADS leaves per block
constexpr int LR_WAVES = LR_THREADS / OCT_WAVE;
int64_t refit_blocks(int64_t n) { return (n + LR_THREADS - 1) / LR_THREADS; }
int64_t refit_slot_words(int64_t n) { return (int64_t)OCT_MAX_LEVEL * 2 * refit_blocks(n) * REFIT_SLOT_WORDS; }
__device__ inline void put_slot(uint32_t* p, int32_t id, const uint32_t* v) {
    reinterpret_cast<uint4*>(p)[0] = make_uint4((uint32_t)id, v[0], v[1], v[2]);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(v[3], v[4], v[5], v[6]);
}

// Two jobs in one launch, both on k_topo_emit's output and independent of
// each other.
// The last link_blocks blocks, one thread per node: a node's children are
// the ids from its first child to the next node's first child (BFS order
// keeps siblings together, and a node without children has the same first
// child as the node after it), so the thread writes the node's eight child
// slots from its children's octants and, for an internal node, its record's
// b = -children.  No slot is written twice and no atomics are needed.
// The blocks before them: internal bounds and range ends from the leaves'
// bounds, one (block of LR_THREADS consecutive leaves, level) pair per
// block, all levels at once.  The lanes under one node of the level form a
// run of consecutive lanes; a segmented wave scan leaves the maxima of each
// wave's part of a run in its last lane, and the parts that cross into later
// waves are carried over in LDS.  A run that holds all of its node's leaves
// (the leaves either side of the block tell) stores the results; a run cut
// by an edge of the block goes to the block's slot for that level and edge
// (slots[L][block][edge]; a run over the whole block fills both), and
// k_refit_merge finishes those nodes.  The chain of a single set keeps its
// set's bounds.
__global__ void __launch_bounds__(LR_THREADS)
k_link_refit(const BuildParams* __restrict__ prm, int batched, int index_base, int link_blocks, int refit_chunks,
             const int32_t* __restrict__ nfc, const uint8_t* __restrict__ noct, const int32_t* __restrict__ ranc,
             int64_t astride, const NodeRec* __restrict__ leaves, int32_t* __restrict__ nchild,
             NodeRec* __restrict__ nodes, int32_t* __restrict__ nlast, float* __restrict__ bmin,
             float* __restrict__ bmax, uint32_t* __restrict__ slots, int64_t bstride) {
    __shared__ int32_t s_edge[2];                                  // ancestors of the leaves before / after the block
    __shared__ int32_t s_fid[LR_WAVES], s_lid[LR_WAVES], s_one[LR_WAVES];  // per wave: ids at lanes 0 / 63, one run
    __shared__ uint32_t s_lv[LR_WAVES][7];                         // the maxima at lane 63
    __shared__ int32_t s_cut[2];                                   // the edge slots were written
    const int nrb = (int)gridDim.x - link_blocks;  // the refit's blocks come first
    if ((int)blockIdx.x >= nrb) {
        const int32_t M = prm->n_nodes;
        for (int32_t id = (int32_t)((blockIdx.x - nrb) * blockDim.x + threadIdx.x); id < M;
             id += (int32_t)(link_blocks * blockDim.x)) {
            const int32_t f0 = nfc[id], f1 = nfc[id + 1];
            int32_t row[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
            for (int32_t c = f0; c < f1; ++c) {
                const int o = noct[c];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (q == o) row[q] = c + index_base;
            }
            int4* dst = reinterpret_cast<int4*>(nchild + 8 * (int64_t)id);
            dst[0] = make_int4(row[0], row[1], row[2], row[3]);
            dst[1] = make_int4(row[4], row[5], row[6], row[7]);
            if (f1 > f0) reinterpret_cast<int32_t*>(nodes + id)[7] = f0 - f1;
        }
        return;
    }
    const int L = (int)blockIdx.x / refit_chunks;  // levels 0..20; level-21 nodes are leaves
    const int32_t nl = prm->n_leaves;
    const int L0 = batched ? -1 : prm->chain_level;  // levels 0..L0: the chain, not refit
    if (L <= L0) return;
    const int lane = lane_id();
    const int wv = threadIdx.x / OCT_WAVE;
    const uint64_t upto = lane == OCT_WAVE - 1 ? ~0ull : ((2ull << lane) - 1ull);
    const int32_t* anc = ranc + (int64_t)L * astride;
    uint32_t* lslots = slots + (int64_t)L * bstride * 2 * REFIT_SLOT_WORDS;
    for (int64_t blk = (int)blockIdx.x % refit_chunks; blk * LR_THREADS < nl; blk += refit_chunks) {  // block-uniform
        const int32_t bbase = (int32_t)(blk * LR_THREADS);
        const int32_t lr = bbase + (int32_t)threadIdx.x;
        const bool valid = lr < nl;
        // the first thread also reads the leaf before the block's, the last
        // one the leaf after
        const bool edge = (threadIdx.x == 0 && bbase > 0) || (threadIdx.x == LR_THREADS - 1 && lr + 1 < nl);
        const int64_t ej = threadIdx.x == 0 ? (int64_t)bbase - 1 : (int64_t)lr + 1;
        const int32_t cur = valid ? anc[lr] : -1;
        const int32_t x = edge ? anc[ej] : -1;
        uint32_t v[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
        if (valid) {
            const NodeRec rec = leaves[lr];  // bounds from k_topo_emit
            v[0] = ~f2o(rec.lo.x); v[1] = ~f2o(rec.lo.y); v[2] = ~f2o(rec.lo.z);
            v[3] = f2o(rec.hi.x); v[4] = f2o(rec.hi.y); v[5] = f2o(rec.hi.z);
            v[6] = (uint32_t)(__float_as_int(rec.lo.w) + __float_as_int(rec.hi.w) - 1 + index_base);
        }
        uint32_t* sl = lslots + blk * 2 * REFIT_SLOT_WORDS;
        __syncthreads();  // the last block's LDS reads are done
        if (threadIdx.x == 0) {
            s_edge[0] = x;
            s_cut[0] = 0;
            s_cut[1] = 0;
        }
        if (threadIdx.x == LR_THREADS - 1) s_edge[1] = x;
        const bool active = cur >= 0;
        const uint64_t am = ballot(active);
        const int32_t prev = __shfl_up(cur, 1, OCT_WAVE), next = __shfl_down(cur, 1, OCT_WAVE);
        const bool head = active && (lane == 0 || !((am >> (lane - 1)) & 1ull) || prev != cur);
        const bool tail = active && (lane == OCT_WAVE - 1 || !((am >> (lane + 1)) & 1ull) || next != cur);
        const uint64_t hm = ballot(head) & upto;
        // an inactive lane keeps its values: its segment starts past itself
        const int seg = active ? 63 - __clzll((long long)hm) : lane + 1;
#pragma unroll
        for (int q = 0; q < 7; ++q) v[q] = seg_max_scan(v[q], lane, seg);
        if (lane == 0) s_fid[wv] = cur;
        if (lane == OCT_WAVE - 1) {
            s_lid[wv] = cur;
            s_one[wv] = active && seg == 0;
            for (int q = 0; q < 7; ++q) s_lv[wv][q] = v[q];
        }
        __syncthreads();
        // the run open into this wave from the block's earlier ones, and
        // whether it reaches back to the block's first lane
        int32_t cid = -1;
        bool creach = false;
        uint32_t cv[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
        for (int y = 0; y < wv; ++y) {
            if (s_one[y] && s_lid[y] == cid) {
                for (int q = 0; q < 7; ++q) cv[q] = s_lv[y][q] > cv[q] ? s_lv[y][q] : cv[q];
            } else {
                cid = s_lid[y];
                creach = y == 0 && s_one[y];
                for (int q = 0; q < 7; ++q) cv[q] = s_lv[y][q];
            }
        }
        bool reach0 = false;  // this lane's run reaches the block's first lane
        if (active && seg == 0) {
            if (cid == cur) {
                for (int q = 0; q < 7; ++q) v[q] = cv[q] > v[q] ? cv[q] : v[q];
                reach0 = creach;
            } else {
                reach0 = wv == 0;
            }
        }
        const bool cont = tail && lane == OCT_WAVE - 1 && wv < LR_WAVES - 1 && s_fid[wv + 1] == cur;
        const bool reachE = tail && lane == OCT_WAVE - 1 && wv == LR_WAVES - 1;  // to the block's last lane
        // all of the node's leaves lie in this run: the leaves either side
        // of the block are not the node's where the run reaches that edge
        const bool whole = (!reach0 || s_edge[0] != cur) && (!reachE || s_edge[1] != cur);
        if (tail && !cont) {
            if (whole) {
                store_refit(cur, v, nodes, nlast, bmin, bmax);
            } else {
                if (reach0) {
                    put_slot(sl, cur, v);
                    s_cut[0] = 1;
                }
                if (reachE) {
                    put_slot(sl + REFIT_SLOT_WORDS, cur, v);
                    s_cut[1] = 1;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (!s_cut[0]) sl[0] = ~0u;
            if (!s_cut[1]) sl[REFIT_SLOT_WORDS] = ~0u;
        }
    }
}

// The runs k_link_refit left unfinished, one block per level.  A level's
// slots (two per k_link_refit block) lie in leaf order and the slots of one node are consecutive (its
// leaves are), so a segmented max over them, in chunks of 1024 slots with
// the open segment carried from chunk to chunk, finishes every such node.
constexpr int RM_THREADS = 1024;
__global__ void __launch_bounds__(RM_THREADS)
k_refit_merge(const BuildParams* __restrict__ prm, int batched, const uint32_t* __restrict__ slots,
              int64_t bstride, NodeRec* __restrict__ nodes, int32_t* __restrict__ nlast, float* __restrict__ bmin,
              float* __restrict__ bmax) {
    constexpr int NW = RM_THREADS / OCT_WAVE;
    __shared__ int32_t s_last[NW], s_one[NW];  // per wave: the id at lane 63, one segment over the wave
    __shared__ uint32_t s_lastv[NW][7];        // the maxima at lane 63
    __shared__ int32_t s_cid[NW + 1];          // the open segment into each wave (NW: into the next chunk)
    __shared__ uint32_t s_cv[NW + 1][7];
    const int L = blockIdx.x;
    const int L0 = batched ? -1 : prm->chain_level;
    if (L <= L0) return;
    const int64_t nslots = 2 * (((int64_t)prm->n_leaves + LR_THREADS - 1) / LR_THREADS);
    const uint32_t* ls = slots + (int64_t)L * bstride * 2 * REFIT_SLOT_WORDS;
    const int w = threadIdx.x / OCT_WAVE, lane = lane_id();
    if (threadIdx.x == 0) {
        s_cid[NW] = -1;
        for (int q = 0; q < 7; ++q) s_cv[NW][q] = 0u;
    }
    auto load = [&](int64_t e, int32_t& id, uint32_t* v) {
        id = -1;
        for (int q = 0; q < 7; ++q) v[q] = 0u;
        if (e < nslots) {
            const uint4 a = reinterpret_cast<const uint4*>(ls + e * REFIT_SLOT_WORDS)[0];
            const uint4 b = reinterpret_cast<const uint4*>(ls + e * REFIT_SLOT_WORDS)[1];
            id = (int32_t)a.x;
            v[0] = a.y; v[1] = a.z; v[2] = a.w; v[3] = b.x; v[4] = b.y; v[5] = b.z; v[6] = b.w;
        }
    };
    int32_t idn;
    uint32_t vn[7];
    load(threadIdx.x, idn, vn);
    const uint64_t upto = lane == OCT_WAVE - 1 ? ~0ull : ((2ull << lane) - 1ull);
    for (int64_t c0 = 0; c0 < nslots; c0 += RM_THREADS) {  // block-uniform
        const int64_t e = c0 + threadIdx.x;
        int32_t id = idn;
        uint32_t v[7];
        for (int q = 0; q < 7; ++q) v[q] = vn[q];
        load(e + RM_THREADS, idn, vn);  // the next chunk, in flight meanwhile
        int32_t nid = -1;
        if (lane == OCT_WAVE - 1 && e + 1 < nslots) nid = (int32_t)ls[(e + 1) * REFIT_SLOT_WORDS];
        const int32_t pid = __shfl_up(id, 1, OCT_WAVE), dn = __shfl_down(id, 1, OCT_WAVE);
        if (lane < OCT_WAVE - 1) nid = dn;
        const bool act = id >= 0;
        const bool head = act && (lane == 0 || pid != id);
        const uint64_t hm = ballot(head) & upto;
        const int seg = act ? 63 - __clzll((long long)hm) : lane + 1;
#pragma unroll
        for (int q = 0; q < 7; ++q) v[q] = seg_max_scan(v[q], lane, seg);
        __syncthreads();  // the last chunk's reads of s_cid / s_cv are done
        if (lane == OCT_WAVE - 1) {
            s_last[w] = id;
            s_one[w] = act && seg == 0;
            for (int q = 0; q < 7; ++q) s_lastv[w][q] = v[q];
        }
        __syncthreads();
        if (w == 0) {
            // lane 0: the segment open into the chunk; lane j: the one open
            // after wave j - 1, which continues lane j - 1's when wave j - 1 is
            // one segment of that node
            int32_t cid = -1;
            uint32_t cv[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
            bool one = false;
            if (lane == 0) {
                cid = s_cid[NW];
                for (int q = 0; q < 7; ++q) cv[q] = s_cv[NW][q];
            } else if (lane <= NW) {
                cid = s_last[lane - 1];
                one = s_one[lane - 1] != 0;
                for (int q = 0; q < 7; ++q) cv[q] = s_lastv[lane - 1][q];
            }
            const int32_t cprev = __shfl_up(cid, 1, OCT_WAVE);
            const bool cact = lane <= NW;
            const bool chead = cact && (lane == 0 || !(one && cid == cprev));
            const uint64_t chm = ballot(chead) & upto;
            const int cseg = cact ? 63 - __clzll((long long)chm) : lane + 1;
#pragma unroll
            for (int q = 0; q < 7; ++q) cv[q] = seg_max_scan(cv[q], lane, cseg);
            if (cact) {
                s_cid[lane] = cid;
                for (int q = 0; q < 7; ++q) s_cv[lane][q] = cv[q];
            }
        }
        __syncthreads();
        if (act && seg == 0 && s_cid[w] == id)
            for (int q = 0; q < 7; ++q) v[q] = s_cv[w][q] > v[q] ? s_cv[w][q] : v[q];
        if (act && nid != id) store_refit(id, v, nodes, nlast, bmin, bmax);
    }
}

// ============================================================== driver

int bounds_blocks(int64_t n) {  // four points per thread and pass (k_bounds), at most 4096 blocks
    const int64_t b = (n + BND_THREADS * 4 - 1) / (BND_THREADS * 4);
    return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}
int64_t sort_tiles(int64_t n) { return (n + RS_TILE - 1) / RS_TILE; }
int64_t topo_blocks(int64_t n) { return (n + TOPO_THREADS - 1) / TOPO_THREADS; }
int64_t hist_words(int64_t n) {
    int64_t a = (int64_t)RS_BUCKETS * sort_tiles(n);
    int64_t b = (int64_t)(OCT_NLEVELS + 1) * topo_blocks(n) + OCT_NLEVELS + 1;  // and the row totals
    return a > b ? a : b;
}
// Nodes: at most n leaves, and at most n / (T + 1) internal nodes per level.
int64_t max_nodes(int64_t n, int T) { return n + 21 * (n / (T + 1)) + 22; }

// Stage 1: bounds reduction (the caller reads prm back to learn key_bits).
hipError_t launch_bounds_stage(const float* coords, int64_t n, float radius, float* copy, const BuildBufs& b,
                               hipStream_t s, const float** csrc) {
    const int nb = bounds_blocks(n);
    const bool vec = ((uintptr_t)coords & 15u) == 0 && (!copy || ((uintptr_t)copy & 15u) == 0);
    if (vec)
        hipLaunchKernelGGL(k_bounds_partial<true>, dim3(nb), dim3(BND_THREADS), 0, s, coords, n, copy, b.partials);
    else
        hipLaunchKernelGGL(k_bounds_partial<false>, dim3(nb), dim3(BND_THREADS), 0, s, coords, n, copy, b.partials);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(BND_THREADS), 0, s, b.partials, nb, radius, b.prm, b.rlook,
                       b.rlook ? refit_look_zero_words(n) : 0, coords, csrc);
    return hipGetLastError();
}

// Radix passes over keys[0] / vals[0] at the given digit shifts (9 bits
// each); *cur = the keys buffer holding the sorted keys.  The last pass
// writes perm (+ index base) and the sorted float4 coordinates.
hipError_t launch_sort_passes(const float* coords, int64_t n, const int* shifts, int npasses, int index_base,
                              const BuildBufs& b, int* cur_out, hipStream_t s) {
    const int ntiles = (int)sort_tiles(n);
    int cur = 0;
    for (int p = 0; p < npasses; ++p) {
        const int shift = shifts[p];
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, s, b.keys[cur], n, shift, b.hist, ntiles);
        hipError_t e = exclusive_scan(b.hist, b.hist, (int64_t)RS_BUCKETS * ntiles, b.scan_tmp, nullptr, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, s, b.keys[cur], b.vals[cur],
                           b.keys[cur ^ 1], b.vals[cur ^ 1], n, shift, b.hist, ntiles, (int)(p == npasses - 1),
                           b.perm, b.scoord, coords, index_base);
        cur ^= 1;
    }
    *cur_out = cur;
    return hipGetLastError();
}

// Stage 2a: keys and the full stable sort over the live bits; *cur = the
// keys buffer that ends up holding the sorted keys.
hipError_t launch_sort_stage(const float* coords, int64_t n, int key_bits, int index_base, const BuildBufs& b,
                             int* cur_out, hipStream_t s) {
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_keys, dim3(g), dim3(256), 0, s, coords, n, b.prm, b.keys[0], b.vals[0]);
    int passes = (key_bits + RS_BITS - 1) / RS_BITS;
    if (passes < 1) passes = 1;
    int shifts[8];
    for (int p = 0; p < passes; ++p) shifts[p] = p * RS_BITS;
    return launch_sort_passes(coords, n, shifts, passes, index_base, b, cur_out, s);
}

// Stage 2a' (almost-sorted path), part 1: new keys in the previous order
// (keys[cur] and perm hold the previous sort), the changed points compacted,
// their count in *b.total and, in prm, whether the refit can run.
RefitLook refit_look(uint64_t* base, int64_t n) {
    const int64_t ntiles = (n + RF_TILE_POS - 1) / RF_TILE_POS, ngroups = (n + OCT_WAVE - 1) / OCT_WAVE;
    RefitLook lk;
    lk.tsum = base;
    lk.wmask = base + ntiles;
    lk.wpre = reinterpret_cast<uint32_t*>(base + ntiles + ngroups);
    return lk;
}
int64_t refit_look_zero_words(int64_t n) { return (n + RF_TILE_POS - 1) / RF_TILE_POS; }
int64_t refit_look_words(int64_t n) {
    const int64_t ngroups = (n + OCT_WAVE - 1) / OCT_WAVE;
    return refit_look_zero_words(n) + ngroups + (ngroups + 1) / 2;
}

hipError_t launch_refit_prepare(const float* const* csrc, int64_t n, int64_t limit, int key_bits, int cur,
                                int index_base, const BuildBufs& b, const RefitBufs& r, hipStream_t s) {
    if (limit > REFIT_MAX_CHANGED) return hipErrorInvalidValue;  // the changed-point buffers hold that many
    const RefitLook lk = refit_look(b.rlook, n);
    // the tile sums were zeroed by the bounds fold just before
    hipLaunchKernelGGL(k_refit_keys, dim3((unsigned)((n + RF_THREADS - 1) / RF_THREADS)), dim3(RF_THREADS), 0, s,
                       csrc, n, b.prm, b.keys[cur], b.perm, index_base, b.vals[0], lk);
    hipLaunchKernelGGL(k_refit_compact, dim3((unsigned)((n + RF_TILE_POS - 1) / RF_TILE_POS)), dim3(OCT_WAVE), 0, s,
                       csrc, n, b.prm, b.vals[0], lk, r.ck, r.ci, limit, key_bits, b.total);
    return hipGetLastError();
}

// Part 2: the sorted keys end in keys[cur ^ 1].  Grids cover `limit` changed
// points; the kernels read the real count (and refit_fail) on the device.
hipError_t launch_refit_finish(const float* const* csrc, int64_t n, int64_t limit, int cur, int index_base,
                               const BuildBufs& b, const RefitBufs& r, hipStream_t s) {
    const unsigned g = (unsigned)((n + RF_THREADS - 1) / RF_THREADS);
    const RefitLook lk = refit_look(b.rlook, n);
    if (limit > 0) {
        const unsigned gc = (unsigned)((limit + RF_THREADS - 1) / RF_THREADS);
        hipLaunchKernelGGL(k_refit_rank, dim3(gc), dim3(RF_THREADS), 0, s, n, b.total, r.ck, r.ci, b.keys[cur],
                           b.vals[0], lk, r.csk, r.csi, r.cspos, b.prm);
    }
    hipLaunchKernelGGL(k_refit_place, dim3(g), dim3(RF_THREADS), 0, s, csrc, n, b.total, b.keys[cur], b.vals[0],
                       lk, r.csk, r.csi, r.cspos, b.keys[cur ^ 1], b.perm, b.scoord, index_base, b.prm);
    return hipGetLastError();
}

// Stage 2b: topology and bounds from the sorted keys in keys[cur].  A forest
// (batched sets) refits every level; a single set skips its top chain, whose
// nodes already hold the set's bounds.
hipError_t launch_topology_stage(int64_t n, int key_bits, int T, int index_base, int cur, const BuildBufs& b,
                                 const Forest& f, hipStream_t s) {
    const uint64_t* sk = b.keys[cur];
    const int nblk = (int)topo_blocks(n);
    hipLaunchKernelGGL(k_topo_count, dim3(nblk), dim3(TOPO_THREADS), 0, s, sk, n, T, f.batched, b.topo, b.hist,
                       nblk);
    hipLaunchKernelGGL(k_topo_rows, dim3(OCT_NLEVELS + 1), dim3(TR_THREADS), 0, s, b.hist, nblk);
    hipLaunchKernelGGL(k_topo_emit, dim3(nblk), dim3(TOPO_THREADS), 0, s, sk, n, T, b.topo, b.hist,
                       nblk, b.prm, f, index_base, b.scoord, b.nodes, b.nfirst, b.nlast, b.bmin, b.bmax,
                       b.nlevel, b.nfc, b.noct, b.leaf_ids, b.leaves, b.ranc, n);
    // child links over the nodes, and the refit over (block of leaves, level)
    // pairs (at most max_nodes nodes and n leaves; both loops stride over the
    // real counts)
    const int gl = (int)std::min<int64_t>((max_nodes(n, T) + LR_THREADS - 1) / LR_THREADS, 1024);
    const int rc = (int)std::min<int64_t>(refit_blocks(n), 256);
    const int64_t bstride = refit_blocks(n);  // refit blocks of leaves, at most
    hipLaunchKernelGGL(k_link_refit, dim3(rc * OCT_MAX_LEVEL + gl), dim3(LR_THREADS), 0, s, b.prm, f.batched,
                       index_base, gl, rc, b.nfc, b.noct, b.ranc, n, b.leaves, b.nchild, b.nodes, b.nlast, b.bmin,
                       b.bmax, b.rslots, bstride);
    hipLaunchKernelGGL(k_refit_merge, dim3(OCT_MAX_LEVEL), dim3(RM_THREADS), 0, s, b.prm, f.batched, b.rslots,
                       bstride, b.nodes, b.nlast, b.bmin, b.bmax);
    return hipGetLastError();
}

__global__ void k_single_set_offsets(int64_t* off, int64_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        off[0] = 0;
        off[1] = n;
    }
}

hipError_t launch_single_set_offsets(int64_t* off, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_single_set_offsets, dim3(1), dim3(64), 0, s, off, n);
    return hipGetLastError();
}

}  // namespace oct
