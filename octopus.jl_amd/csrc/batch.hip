// Many point sets in one forest: the kernels that differ from the one-set
// build (per-set bounds, keys carrying the set id, per-set edge offsets, the
// check a replayed graph needs).  The reference loops over the sets one at a
// time (src/api.jl:196); here every stage is one launch for all sets, so a
// batch of small sets costs a few dozen launches, and the whole step can be
// captured once and replayed as a HIP graph.
//
// Layout: the caller's coordinates are the sets concatenated (set s holds
// points off[s] .. off[s + 1]); keys are set_id << 42 | the set's own 63-bit
// Morton key (which, with at most 2^14 cells per axis, fits in 42 bits and
// has the same tree as the set built alone); one stable sort then keeps every
// set contiguous and in order, and a change of set id starts a new tree.
#include "internal.hpp"

namespace oct {

constexpr int BB_THREADS = 256;

// set s with prefix[s] <= x < prefix[s + 1] (prefix nondecreasing, S + 1 entries)
template <typename T>
__device__ inline int set_of(const T* __restrict__ prefix, int S, T x) {
    int lo = 0, hi = S;  // invariant: prefix[lo] <= x < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Partial min/max: block b works for the set whose block range holds b.
__global__ void __launch_bounds__(BB_THREADS)
k_bounds_batch_partial(const float* __restrict__ coords, const int64_t* __restrict__ off,
                       const int32_t* __restrict__ bprefix, int S, uint32_t* __restrict__ partials) {
    const int b = (int)blockIdx.x;
    const int s = set_of<int32_t>(bprefix, S, b);
    bounds_part(coords, off[s], off[s + 1], b - bprefix[s], bprefix[s + 1] - bprefix[s], partials + 6 * b);
}

// One block per set folds its partials into the set's parameters.
__global__ void __launch_bounds__(BB_THREADS)
k_bounds_batch_final(const uint32_t* __restrict__ partials, const int32_t* __restrict__ bprefix,
                     const float* __restrict__ radii, BuildParams* __restrict__ sprm) {
    const int s = (int)blockIdx.x;
    bounds_fold(partials, bprefix[s], bprefix[s + 1], radii[s], sprm + s);
}

// Keys of all points: the set id above bit BATCH_SHIFT, the set's Morton key
// (its own origin and cell size) below; vals = point index in the batch.
__global__ void __launch_bounds__(BB_THREADS)
k_keys_batch(const float* __restrict__ coords, int64_t n, const int64_t* __restrict__ off, int S,
             const BuildParams* __restrict__ sprm, uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * BB_THREADS + threadIdx.x;
    if (i >= n) return;
    const int s = set_of<int64_t>(off, S, i);
    const BuildParams& p = sprm[s];
    const float px = coords[3 * i], py = coords[3 * i + 1], pz = coords[3 * i + 2];
    const uint32_t ix = floor_clamp((px - p.origin[0]) * p.inv_cs, 0x1fffffu);
    const uint32_t iy = floor_clamp((py - p.origin[1]) * p.inv_cs, 0x1fffffu);
    const uint32_t iz = floor_clamp((pz - p.origin[2]) * p.inv_cs, 0x1fffffu);
    keys[i] = morton3(ix, iy, iz) | ((uint64_t)s << BATCH_SHIFT);
    vals[i] = (int32_t)i;
}

// flags[0] |= 1 when a set needs more key bits than the sort was set up for
// (max_bits), or more than the batched layout holds.
__global__ void k_batch_check_bits(const BuildParams* __restrict__ sprm, const int64_t* __restrict__ off, int S,
                                   int max_bits, uint32_t* __restrict__ flags) {
    for (int s = (int)(blockIdx.x * blockDim.x + threadIdx.x); s < S; s += (int)(gridDim.x * blockDim.x)) {
        if (off[s + 1] == off[s]) continue;  // an empty set has no keys
        if (sprm[s].key_bits > max_bits || sprm[s].key_bits > BATCH_SHIFT) atomicOr(&flags[0], 1u);
    }
}

// Per-set edge offsets from the exclusive scan of the per-leaf hit counts:
// the first leaf of set s is the first leaf whose first point is at or after
// the set's first position.
__global__ void k_batch_edge_offsets(const uint32_t* __restrict__ loff, const uint64_t* __restrict__ total,
                                     const NodeRec* __restrict__ leaves, const BuildParams* __restrict__ gprm,
                                     const int64_t* __restrict__ off, int S, int64_t* __restrict__ eoff) {
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s > S) return;
    if (s == S) {
        eoff[S] = (int64_t)*total;
        return;
    }
    const int32_t nl = gprm->n_leaves;
    int32_t lo = 0, hi = nl;  // first leaf with first position >= off[s]
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if ((int64_t)__float_as_int(leaves[mid].lo.w) < off[s]) lo = mid + 1;
        else hi = mid;
    }
    eoff[s] = lo < nl ? (int64_t)loff[lo] : (int64_t)*total;
}

hipError_t launch_bounds_batch(const float* coords, const int64_t* off, const int32_t* bprefix, int S,
                               int nblocks, const float* radii, uint32_t* partials, BuildParams* sprm,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_bounds_batch_partial, dim3(nblocks), dim3(BB_THREADS), 0, s, coords, off, bprefix, S,
                       partials);
    hipLaunchKernelGGL(k_bounds_batch_final, dim3(S), dim3(BB_THREADS), 0, s, partials, bprefix, radii, sprm);
    return hipGetLastError();
}

hipError_t launch_keys_batch(const float* coords, int64_t n, const int64_t* off, int S, const BuildParams* sprm,
                             uint64_t* keys, int32_t* vals, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys_batch, dim3((unsigned)((n + BB_THREADS - 1) / BB_THREADS)), dim3(BB_THREADS), 0, s,
                       coords, n, off, S, sprm, keys, vals);
    return hipGetLastError();
}

hipError_t launch_batch_check_bits(const BuildParams* sprm, const int64_t* off, int S, int max_bits,
                                   uint32_t* flags, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_check_bits, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s, sprm, off, S,
                       max_bits, flags);
    return hipGetLastError();
}

hipError_t launch_batch_edge_offsets(const uint32_t* loff, const uint64_t* total, const NodeRec* leaves,
                                     const BuildParams* gprm, const int64_t* off, int S, int64_t* eoff,
                                     hipStream_t s) {
    hipLaunchKernelGGL(k_batch_edge_offsets, dim3((unsigned)((S + 1 + 255) / 256)), dim3(256), 0, s, loff, total,
                       leaves, gprm, off, S, eoff);
    return hipGetLastError();
}

}  // namespace oct
