// Shared device helpers for the octohip HIP library (gfx950, wave64).
//
// FP contract: every translation unit is compiled with -ffp-contract=off and
// without fast-math, so each float expression below rounds once per operation
// in the written order.  That is what keeps the edge predicate bit-identical
// to the reference CPU path (src/cpu/iterator_cpu.jl:96-98).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OCT_WAVE 64
#define OCT_MAX_LEVEL 21  // 21 bits per axis, 63-bit keys (src/morton.jl:8)
#define OCT_NLEVELS 22    // levels 0..21

namespace oct {

// ------------------------------------------------------------ float order

// Monotone map float -> uint32 (total order on non-NaN floats, -0 < +0).
__host__ __device__ inline uint32_t f2o(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float o2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}

// Julia's Base.min / Base.max on non-NaN floats: among equal values min
// prefers -0.0 and max prefers +0.0.  The ordered-uint comparison gives
// exactly that (-0.0 orders below +0.0).
__device__ inline float jmin(float a, float b) { return f2o(b) < f2o(a) ? b : a; }
__device__ inline float jmax(float a, float b) { return f2o(b) > f2o(a) ? b : a; }

// One rel_displacement component as the CPU path writes it
// (src/cpu/edges_cpu.jl:111-113): the traversal's difference is target minus
// query (src/cpu/iterator_cpu.jl:96) and the fill negates it.  Equal to
// (q - t) * inv_r in every bit except when t == q, where it is -0.0.
__device__ inline float rel_disp(float q, float t, float inv_r) { return -(t - q) * inv_r; }

// ------------------------------------------------------------------ Morton

// Magic-bits spread of the low 21 bits (src/morton.jl:13-23).
__host__ __device__ inline uint64_t split_by_3(uint64_t x) {
    x &= 0x00000000001fffffULL;
    x = (x | (x << 32)) & 0x001f00000000ffffULL;
    x = (x | (x << 16)) & 0x001f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

// x -> bit 0, y -> bit 1, z -> bit 2 (src/morton.jl:35-40).
__host__ __device__ inline uint64_t morton3(uint32_t ix, uint32_t iy, uint32_t iz) {
    return (split_by_3(iz) << 2) | (split_by_3(iy) << 1) | split_by_3(ix);
}

// floor(Int32, v) then clamp(., 0, hi) (src/morton.jl:83-88,
// src/cpu/bin_cpu.jl:64-68), clamped in float so out-of-range values stay
// defined.
__host__ __device__ inline uint32_t floor_clamp(float v, uint32_t hi) {
    float f = floorf(v);
    if (!(f >= 0.0f)) return 0u;
    if (f >= (float)hi) return hi;
    return (uint32_t)f;
}

// Number of leading levels two keys share: the largest L with
// key_a >> (63-3L) == key_b >> (63-3L), 21 when equal.
__device__ inline int shared_levels(uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    if (x == 0) return OCT_MAX_LEVEL;
    int h = 63 - __clzll((long long)x);  // highest differing bit, <= 62
    return (62 - h) / 3;
}

// Bit position of the lowest of the 3 bits a level-L node splits on
// (src/cpu/build_cpu.jl:46-49): children of a level-L node differ there.
__host__ __device__ inline int level_shift(int L) { return 63 - 3 * (L + 1); }

// Set id and Morton part of a (possibly batched) key; levels two keys share,
// -1 across a set boundary (a new tree starts there).
__device__ inline int key_set(uint64_t k, int batched) { return batched ? (int)(k >> 42) : 0; }
__device__ inline uint64_t key_morton(uint64_t k, int batched) { return batched ? (k & ((1ull << 42) - 1ull)) : k; }
__device__ inline int key_shared(uint64_t a, uint64_t b, int batched) {
    if (batched && ((a ^ b) >> 42)) return -1;
    return shared_levels(key_morton(a, batched), key_morton(b, batched));
}

// ------------------------------------------------------------ XCD mapping

// MI355X dispatches workgroups round-robin over its 8 XCDs, each with its own
// L2.  For work counted on the device, with a grid sized by an upper bound:
// the XCD of this block (of the first min(grid, 8)) takes an equal contiguous
// share [lo, hi) of [0, n), and this block is number k of the nb blocks on
// that XCD.  Morton-adjacent work (which shares tree nodes and candidate
// points) then stays in one L2, and every XCD gets its share whatever the
// grid's excess.
#define OCT_NXCD 8u
struct XcdShare {
    int64_t lo, hi;
    uint32_t k, nb;
};
__device__ inline XcdShare xcd_share(int64_t n) {
    const uint32_t nx = gridDim.x < OCT_NXCD ? gridDim.x : OCT_NXCD;
    const uint32_t x = blockIdx.x % nx;
    XcdShare s;
    s.k = blockIdx.x / nx;
    s.nb = (gridDim.x - x + nx - 1) / nx;
    s.lo = n * (int64_t)x / (int64_t)nx;
    s.hi = n * (int64_t)(x + 1) / (int64_t)nx;
    return s;
}

// ------------------------------------------------------------- wave utils

__device__ inline int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// popcount of mask bits below this lane
__device__ inline int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

template <typename T>
__device__ inline T wave_min_u(T v) {
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, OCT_WAVE);
        v = w < v ? w : v;
    }
    return v;
}
template <typename T>
__device__ inline T wave_max_u(T v) {
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, OCT_WAVE);
        v = w > v ? w : v;
    }
    return v;
}
// inclusive prefix sum across the wave in DPP moves (no LDS round trips):
// row_shr 1, 2, 4, 8 scan each row of 16 lanes (lanes shifted in from
// outside the row read 0), then row_bcast:15 adds row 0's total to row 1
// and row 2's to row 3, and row_bcast:31 adds the total of rows 0-1 to
// rows 2-3.  Call with every lane of the wave active.
__device__ inline int wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
    return v;
}
// sum across the wave (the same in every lane)
__device__ inline int wave_sum(int v) { return __builtin_amdgcn_readlane(wave_incl_scan(v), OCT_WAVE - 1); }

// Inclusive max-scan inside segments of consecutive lanes: lane i gets the
// max over lanes seg .. i, where seg <= i is the first lane of its segment.
// The steps of wave_incl_scan, each taken only where its source lane lies in
// the segment (row_shr:s reads lane i - s, row_bcast:15 the lane before the
// row, row_bcast:31 lane 31).  0 is the identity.  Every lane active.
__device__ inline uint32_t seg_max_scan(uint32_t v, int lane, int seg) {
    auto step = [&](uint32_t t, int src) { v = (src >= seg && t > v) ? t : v; };
    step((uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true), lane - 1);
    step((uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true), lane - 2);
    step((uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true), lane - 4);
    step((uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true), lane - 8);
    step((uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false), (lane & ~15) - 1);
    step((uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false), 31);
    return v;
}
__device__ inline uint64_t wave_sum_u64(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, OCT_WAVE);
    return v;
}

// Device-coherent loads that bypass the non-coherent per-CU cache, for data
// another workgroup wrote during this same launch.
__device__ inline float load_coherent(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int load_coherent(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Correctly rounded FP32 square root of x >= SQRT_DIRECT_MIN (finite) and of
// +0: the hardware estimate (within 1 ulp), then the residual of each
// neighbour, x - s_down * s and x - s_up * s, in one fused multiply-add:
// x <= s_down * s puts sqrt(x) below the midpoint of s_down and s, x > s_up *
// s above that of s and s_up (no float x lies between such a product and
// the midpoint's square).  These are the steps sqrtf itself takes on this
// target for such inputs; below SQRT_DIRECT_MIN it first scales x by 2^32 so
// the residuals stay normal, which callers get by calling sqrtf there.
constexpr float SQRT_DIRECT_MIN = 0x1p-96f;
__device__ inline float sqrt_rn_direct(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sdn = __int_as_float(__float_as_int(s) - 1);
    const float sup = __int_as_float(__float_as_int(s) + 1);
    const float rdn = fmaf(-sdn, s, x);
    const float rup = fmaf(-sup, s, x);
    const float y = rdn <= 0.0f ? sdn : s;
    return rup > 0.0f ? sup : y;
}

// FP32 squared distance from a point to a box, in the reference's order
// (src/cpu/iterator_cpu.jl:31-40).
__device__ inline float point_box_d2(float px, float py, float pz, float4 lo, float4 hi) {
    float dx = fmaxf(lo.x - px, 0.0f); dx = fmaxf(dx, px - hi.x);
    float dy = fmaxf(lo.y - py, 0.0f); dy = fmaxf(dy, py - hi.y);
    float dz = fmaxf(lo.z - pz, 0.0f); dz = fmaxf(dz, pz - hi.z);
    return dx * dx + dy * dy + dz * dz;
}

// FP32 squared gap between two boxes.  For any point q in A and t in B,
// fl(box_box_d2(A, B)) <= fl(d2(q, t)) because each step is a monotone IEEE
// operation on a smaller-or-equal magnitude, so pruning on it never drops an
// edge.
__device__ inline float box_box_d2(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    float dx = fmaxf(fmaxf(blo.x - ahi.x, alo.x - bhi.x), 0.0f);
    float dy = fmaxf(fmaxf(blo.y - ahi.y, alo.y - bhi.y), 0.0f);
    float dz = fmaxf(fmaxf(blo.z - ahi.z, alo.z - bhi.z), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

}  // namespace oct
