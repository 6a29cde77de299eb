// Adjoint of build_edges_diff (reference rrule:
// ext/OctopusChainRulesCoreExt.jl:137-192, GPU kernel
// ext/OctopusCUDAChainRulesCoreExt.jl:56-99).
//
// For edge k with receiver i and sender j:
//   v_d = d_disp[d,k] / r  +  (rdist_k > 1e-8 ? rdisp[d,k] * (d_dist[k] / (rdist_k * r)) : 0)
//   grad[d,i] += v_d,  grad[d,j] -= v_d
// Edges are stored receiver-grouped (each receiver's edges are one run), so
// the receiver side is a segmented reduction in registers: a wave walks 64
// consecutive edges, sums the runs with shuffles and issues one atomic per
// run boundary instead of one per edge.  The sender side is a scatter with
// FP32 atomics (order-dependent rounding, as in the reference GPU path).
#include "internal.hpp"

namespace oct {

__global__ void __launch_bounds__(256)
k_pullback(int64_t n_edges, const int32_t* __restrict__ senders, const int32_t* __restrict__ receivers,
           const float* __restrict__ rdisp, const float* __restrict__ rdist,
           const float* __restrict__ d_disp, const float* __restrict__ d_dist, float r, int index_base,
           float* __restrict__ grad) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = k < n_edges;
    float v[3] = {0.f, 0.f, 0.f};
    int32_t i = -1, j = -1;
    if (valid) {
        i = receivers[k] - index_base;
        j = senders[k] - index_base;
        if (d_disp) {
            for (int d = 0; d < 3; ++d) v[d] = d_disp[3 * k + d] / r;
        }
        if (d_dist) {
            const float dn = rdist[k];
            if (dn > 1e-8f) {
                const float inv_dnr = d_dist[k] / (dn * r);
                for (int d = 0; d < 3; ++d) v[d] = v[d] + rdisp[3 * k + d] * inv_dnr;
            }
        }
        for (int d = 0; d < 3; ++d) atomicAdd(&grad[3 * (int64_t)j + d], -v[d]);
    }
    // receiver side: segmented sum over runs of equal receivers within the wave
    const int lane = lane_id();
    const int prev_i = __shfl_up(i, 1, OCT_WAVE);
    const bool head = valid && (lane == 0 || prev_i != i);
    const uint64_t heads = ballot(head);
    for (int d = 0; d < 3; ++d) {
        float s = v[d];
        // inclusive segmented scan: add from lane-o when no head lies in (lane-o, lane]
        for (int o = 1; o < OCT_WAVE; o <<= 1) {
            const float w = __shfl_up(s, o, OCT_WAVE);
            const uint64_t span = (lane >= o) ? (((2ull << lane) - 1) & ~((2ull << (lane - o)) - 1)) : 0ull;
            if (lane >= o && (heads & span) == 0) s = s + w;
        }
        v[d] = s;
    }
    const int next_i = __shfl_down(i, 1, OCT_WAVE);
    const bool next_valid = __shfl_down((int)valid, 1, OCT_WAVE);
    const bool tail = valid && (lane == OCT_WAVE - 1 || !next_valid || next_i != i);
    if (tail)
        for (int d = 0; d < 3; ++d) atomicAdd(&grad[3 * (int64_t)i + d], v[d]);
}

__global__ void k_zero_f32(float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

hipError_t launch_pullback(int64_t n_points, int64_t n_edges, const int32_t* senders,
                           const int32_t* receivers, const float* rdisp, const float* rdist,
                           const float* d_disp, const float* d_dist, float r, int index_base,
                           float* grad, hipStream_t s) {
    if (n_points > 0)
        hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)((3 * n_points + 255) / 256)), dim3(256), 0, s, grad,
                           3 * n_points);
    if (n_edges > 0 && (d_disp || d_dist))
        hipLaunchKernelGGL(k_pullback, dim3((unsigned)((n_edges + 255) / 256)), dim3(256), 0, s, n_edges,
                           senders, receivers, rdisp, rdist, d_disp, d_dist, r, index_base, grad);
    return hipGetLastError();
}

}  // namespace oct
