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
// FP32 atomics (order-dependent rounding, as in the reference GPU path) in
// the push form, or pulled from the reverse edges in the pull form below.
#include "internal.hpp"

namespace oct {

template <int D, typename T>
__global__ void __launch_bounds__(256)
k_pullback(int64_t n_edges, const int32_t* __restrict__ senders, const int32_t* __restrict__ receivers,
           const T* __restrict__ rdisp, const T* __restrict__ rdist,
           const T* __restrict__ d_disp, const T* __restrict__ d_dist, T r, int index_base,
           T* __restrict__ grad) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = k < n_edges;
    T v[D] = {};
    int32_t i = -1, j = -1;
    if (valid) {
        i = receivers[k] - index_base;
        j = senders[k] - index_base;
        if (d_disp) {
            for (int d = 0; d < D; ++d) v[d] = d_disp[D * k + d] / r;
        }
        if (d_dist) {
            const T dn = rdist[k];
            if (dn > T(1e-8)) {
                const T inv_dnr = d_dist[k] / (dn * r);
                for (int d = 0; d < D; ++d) v[d] = v[d] + rdisp[D * k + d] * inv_dnr;
            }
        }
        for (int d = 0; d < D; ++d) atomicAdd(&grad[D * (int64_t)j + d], -v[d]);
    }
    // receiver side: segmented sum over runs of equal receivers within the wave
    const int lane = lane_id();
    const int prev_i = __shfl_up(i, 1, OCT_WAVE);
    const bool head = valid && (lane == 0 || prev_i != i);
    const uint64_t heads = ballot(head);
    for (int d = 0; d < D; ++d) {
        T s = v[d];
        // inclusive segmented scan: add from lane-o when no head lies in (lane-o, lane]
        for (int o = 1; o < OCT_WAVE; o <<= 1) {
            const T w = __shfl_up(s, o, OCT_WAVE);
            const uint64_t span = (lane >= o) ? (((2ull << lane) - 1) & ~((2ull << (lane - o)) - 1)) : 0ull;
            if (lane >= o && (heads & span) == 0) s = s + w;
        }
        v[d] = s;
    }
    const int next_i = __shfl_down(i, 1, OCT_WAVE);
    const bool next_valid = __shfl_down((int)valid, 1, OCT_WAVE);
    const bool tail = valid && (lane == OCT_WAVE - 1 || !next_valid || next_i != i);
    if (tail)
        for (int d = 0; d < D; ++d) atomicAdd(&grad[D * (int64_t)i + d], v[d]);
}

// ---- pull form: no per-edge atomics on the sender side --------------------
//
// With every receiver's edges in one contiguous run (what build_edges
// writes, in either order), the sender-side term of edge k = (i -> j) can be
// pulled by j instead of pushed to it: for the self pair every edge has its
// reverse (j -> i) in j's run (the predicate is symmetric bit for bit), and
//   grad_i = sum over k in run(i) of (v_k - v_rev(k)).
// Pairing is by occurrence: the t-th edge (i -> j) of run(i) pairs with the
// t-th edge (j -> i) of run(j), which is symmetric, so each sender term is
// taken exactly once; an edge with no partner (a cross pair, an edge list
// from elsewhere) pushes its sender term with an atomic as before, and so do
// edges of runs longer than PULL_MAX_RUN (the partner search is linear).  A
// receiver with two runs sets a flag and the whole call takes the push form.
constexpr int PULL_MAX_RUN = 96;
constexpr int PULL_BATCH = 8;

// start/end of each receiver's run, and the flag for a receiver seen in two
// runs (run_count, start, end zeroed before).
__global__ void __launch_bounds__(256)
k_pull_runs(int64_t n_edges, const int32_t* __restrict__ receivers, int index_base, int32_t* __restrict__ run_start,
            int32_t* __restrict__ run_end, int32_t* __restrict__ run_count, uint32_t* __restrict__ flag) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_edges) return;
    const int32_t i = receivers[k];
    if (k == 0 || receivers[k - 1] != i) {
        run_start[i - index_base] = (int32_t)k;
        if (atomicAdd(&run_count[i - index_base], 1) != 0) atomicOr(flag, 1u);
    }
    if (k == n_edges - 1 || receivers[k + 1] != i) run_end[i - index_base] = (int32_t)(k + 1);
}

template <int D, typename T>
__device__ inline void pull_value(int64_t k, const T* __restrict__ rdisp, const T* __restrict__ rdist,
                                  const T* __restrict__ d_disp, const T* __restrict__ d_dist, T r, T (&v)[D]) {
    for (int d = 0; d < D; ++d) v[d] = T(0);
    if (d_disp)
        for (int d = 0; d < D; ++d) v[d] = d_disp[D * k + d] / r;
    if (d_dist) {
        const T dn = rdist[k];
        if (dn > T(1e-8)) {  // T(1e-8), as the reference's rrule
            const T inv_dnr = d_dist[k] / (dn * r);
            for (int d = 0; d < D; ++d) v[d] = v[d] + rdisp[D * k + d] * inv_dnr;
        }
    }
}

template <int D, typename T>
__global__ void __launch_bounds__(256)
k_pullback_pull(int64_t n_edges, const int32_t* __restrict__ senders, const int32_t* __restrict__ receivers,
                const T* __restrict__ rdisp, const T* __restrict__ rdist, const T* __restrict__ d_disp,
                const T* __restrict__ d_dist, T r, int index_base, const int32_t* __restrict__ run_start,
                const int32_t* __restrict__ run_end, const uint32_t* __restrict__ flag, T* __restrict__ grad) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = k < n_edges;
    const bool runs_ok = *flag == 0u;
    T v[D] = {};
    int32_t i = -1;
    if (valid) {
        i = receivers[k] - index_base;
        const int32_t j = senders[k] - index_base;
        T w[D];
        pull_value<D, T>(k, rdisp, rdist, d_disp, d_dist, r, w);
        int64_t kr = -1;
        if (runs_ok) {
            const int32_t si = run_start[i], sj = run_start[j], ej = run_end[j];
            if (run_end[i] - si <= PULL_MAX_RUN && ej - sj <= PULL_MAX_RUN) {
                // both scans load PULL_BATCH senders before comparing any, so
                // the loads of a batch are in flight together
                const int32_t js = senders[k];
                int t = 0;  // earlier edges (i -> j) in run(i)
                for (int32_t m0 = si; m0 < (int32_t)k; m0 += PULL_BATCH) {
                    int32_t sv[PULL_BATCH];
#pragma unroll
                    for (int u = 0; u < PULL_BATCH; ++u) sv[u] = m0 + u < (int32_t)k ? senders[m0 + u] : -1;
#pragma unroll
                    for (int u = 0; u < PULL_BATCH; ++u) t += sv[u] == js;
                }
                const int32_t is = i + index_base;
                for (int32_t m0 = sj; m0 < ej && kr < 0; m0 += PULL_BATCH) {
                    int32_t sv[PULL_BATCH];
#pragma unroll
                    for (int u = 0; u < PULL_BATCH; ++u) sv[u] = m0 + u < ej ? senders[m0 + u] : -1;
#pragma unroll
                    for (int u = 0; u < PULL_BATCH; ++u) {
                        if (kr < 0 && sv[u] == is) {
                            if (t == 0) kr = m0 + u;
                            else --t;
                        }
                    }
                }
            }
        }
        if (kr >= 0) {  // the partner's sender term belongs to i: pull it
            T u[D];
            pull_value<D, T>(kr, rdisp, rdist, d_disp, d_dist, r, u);
            for (int d = 0; d < D; ++d) v[d] = w[d] - u[d];
        } else {  // no partner: push this edge's sender term to j
            for (int d = 0; d < D; ++d) {
                v[d] = w[d];
                atomicAdd(&grad[D * (int64_t)j + d], -w[d]);
            }
        }
    }
    // receiver side: segmented sum over runs of equal receivers within the wave
    const int lane = lane_id();
    const int prev_i = __shfl_up(i, 1, OCT_WAVE);
    const bool head = valid && (lane == 0 || prev_i != i);
    const uint64_t heads = ballot(head);
    for (int d = 0; d < D; ++d) {
        T s = v[d];
        for (int o = 1; o < OCT_WAVE; o <<= 1) {
            const T w = __shfl_up(s, o, OCT_WAVE);
            const uint64_t span = (lane >= o) ? (((2ull << lane) - 1) & ~((2ull << (lane - o)) - 1)) : 0ull;
            if (lane >= o && (heads & span) == 0) s = s + w;
        }
        v[d] = s;
    }
    const int next_i = __shfl_down(i, 1, OCT_WAVE);
    const bool next_valid = __shfl_down((int)valid, 1, OCT_WAVE);
    const bool tail = valid && (lane == OCT_WAVE - 1 || !next_valid || next_i != i);
    if (tail)
        for (int d = 0; d < D; ++d) atomicAdd(&grad[D * (int64_t)i + d], v[d]);
}

template <typename T>
__global__ void k_zero(T* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

int64_t pullback_scratch_words(int64_t n_points) { return 3 * n_points + 64; }

template <int D, typename T>
static hipError_t launch_pullback_d(int64_t n_points, int64_t n_edges, const int32_t* senders,
                                    const int32_t* receivers, const T* rdisp, const T* rdist, const T* d_disp,
                                    const T* d_dist, T r, int index_base, T* grad, uint32_t* scratch,
                                    hipStream_t s) {
    if (n_points > 0)
        hipLaunchKernelGGL(k_zero<T>, dim3((unsigned)((D * n_points + 255) / 256)), dim3(256), 0, s, grad,
                           D * n_points);
    if (n_edges <= 0 || !(d_disp || d_dist)) return hipGetLastError();
    const unsigned g = (unsigned)((n_edges + 255) / 256);
    if (!scratch) {
        hipLaunchKernelGGL((k_pullback<D, T>), dim3(g), dim3(256), 0, s, n_edges, senders, receivers, rdisp, rdist,
                           d_disp, d_dist, r, index_base, grad);
        return hipGetLastError();
    }
    int32_t* run_start = (int32_t*)scratch;
    int32_t* run_end = run_start + n_points;
    int32_t* run_count = run_end + n_points;
    uint32_t* flag = (uint32_t*)(run_count + n_points);
    const int64_t words = 3 * n_points + 1;
    hipLaunchKernelGGL(k_zero<uint32_t>, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, scratch, words);
    hipLaunchKernelGGL(k_pull_runs, dim3(g), dim3(256), 0, s, n_edges, receivers, index_base, run_start, run_end,
                       run_count, flag);
    hipLaunchKernelGGL((k_pullback_pull<D, T>), dim3(g), dim3(256), 0, s, n_edges, senders, receivers, rdisp, rdist,
                       d_disp, d_dist, r, index_base, run_start, run_end, flag, grad);
    return hipGetLastError();
}

hipError_t launch_pullback(int64_t n_points, int64_t n_edges, const int32_t* senders,
                           const int32_t* receivers, const float* rdisp, const float* rdist,
                           const float* d_disp, const float* d_dist, float r, int index_base,
                           float* grad, uint32_t* scratch, int dims, hipStream_t s) {
    if (dims == 2)
        return launch_pullback_d<2, float>(n_points, n_edges, senders, receivers, rdisp, rdist, d_disp, d_dist, r,
                                           index_base, grad, scratch, s);
    if (dims == 3)
        return launch_pullback_d<3, float>(n_points, n_edges, senders, receivers, rdisp, rdist, d_disp, d_dist, r,
                                           index_base, grad, scratch, s);
    return hipErrorInvalidValue;
}

hipError_t launch_pullback64(int64_t n_points, int64_t n_edges, const int32_t* senders, const int32_t* receivers,
                             const double* rdisp, const double* rdist, const double* d_disp, const double* d_dist,
                             double r, int index_base, double* grad, uint32_t* scratch, int dims, hipStream_t s) {
    if (dims == 2)
        return launch_pullback_d<2, double>(n_points, n_edges, senders, receivers, rdisp, rdist, d_disp, d_dist, r,
                                            index_base, grad, scratch, s);
    if (dims == 3)
        return launch_pullback_d<3, double>(n_points, n_edges, senders, receivers, rdisp, rdist, d_disp, d_dist, r,
                                            index_base, grad, scratch, s);
    return hipErrorInvalidValue;
}

}  // namespace oct
