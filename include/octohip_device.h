/*
 * Device-side neighbour iteration for user HIP kernels, the counterpart of
 * for_each_neighbor_device / @for_each_neighbor_device_inline
 * (ext/OctopusCUDAExt.jl:90-147, src/macros.jl:60-115).
 *
 *   __global__ void my_kernel(oct_view v, int* counts) {
 *       int64_t i = blockIdx.x * blockDim.x + threadIdx.x;
 *       if (i >= v.n_q) return;
 *       int c = 0;
 *       oct_for_each_neighbor_device(v, i, [&](int32_t j) { ++c; });
 *       counts[i] = c;
 *   }
 *
 * i is a 0-based query row; the callback gets j with the view's index base.
 * As in the reference, the self pair is NOT filtered.  The predicate is the
 * reference's FP32 test d2 = (dx*dx + dy*dy) + dz*dz <= r*r, so the visited
 * set equals the CPU path's when the kernel is built with -ffp-contract=off.
 * The traversal state (cursor, stack) is kept in registers / scratch of the
 * calling thread; nodes are tested before they are pushed.
 */
#ifndef OCTOHIP_DEVICE_H
#define OCTOHIP_DEVICE_H

#include <hip/hip_runtime.h>

#include "octohip.h"

/* Packed node record (32 B): lo = (min x, min y, min z, a), hi = (max x, max y,
 * max z, b).  Leaf: a = first sorted position, b = point count (> 0).
 * Internal: a = first child id (children are consecutive), b = -child count. */
typedef struct oct_node_rec {
    float4 lo;
    float4 hi;
} oct_node_rec;

#define OCT_DEVICE_STACK (1 + 7 * 21)

__device__ inline float oct_point_box_d2(float px, float py, float pz, float4 lo, float4 hi) {
    float dx = fmaxf(lo.x - px, 0.0f); dx = fmaxf(dx, px - hi.x);
    float dy = fmaxf(lo.y - py, 0.0f); dy = fmaxf(dy, py - hi.y);
    float dz = fmaxf(lo.z - pz, 0.0f); dz = fmaxf(dz, pz - hi.z);
    return dx * dx + dy * dy + dz * dz;
}

template <typename F>
__device__ inline void oct_for_each_neighbor_device(const oct_view& v, int64_t i, F&& f) {
    if (v.n_nodes == 0) return;
    const oct_node_rec* nodes = (const oct_node_rec*)v.nodes;
    const float4* st = (const float4*)v.sorted_t;
    const float px = v.coords_q[3 * i], py = v.coords_q[3 * i + 1], pz = v.coords_q[3 * i + 2];
    const float r_sq = v.radius_sq;
    int32_t stack[OCT_DEVICE_STACK];
    int sp = 0;
    {
        const oct_node_rec rr = nodes[v.root];
        if (oct_point_box_d2(px, py, pz, rr.lo, rr.hi) <= r_sq) stack[sp++] = v.root;
    }
    while (sp > 0) {
        const oct_node_rec nd = nodes[stack[--sp]];
        const int a = __float_as_int(nd.lo.w), b = __float_as_int(nd.hi.w);
        if (b > 0) {
            for (int k = a; k < a + b; ++k) {
                const float4 q = st[k];
                const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (d2 <= r_sq) f((int32_t)(__float_as_int(q.w) + v.index_base));
            }
        } else {
            for (int c = 0; c < -b; ++c) {
                const oct_node_rec ch = nodes[a + c];
                if (oct_point_box_d2(px, py, pz, ch.lo, ch.hi) <= r_sq) stack[sp++] = a + c;
            }
        }
    }
}

#endif
