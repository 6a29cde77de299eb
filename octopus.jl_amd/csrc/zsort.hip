// apply_zsort and its adjoint: columns of a user array gathered into the
// z-order of a point set, and the transpose (scatter) for the rrule.
// Reference: apply_zsort_cpu! (src/cpu/zsort_cpu.jl:11-32), the CUDA gather
// user_array[:, perm] (ext/OctopusCUDAExt.jl:422-432) and the rrule's
// grad[:, perm] = dy (ext/OctopusChainRulesCoreExt.jl:65-94).
//
// An array is n columns of `width` elements (column-major (width, n), or a
// vector with width 1).  One thread per output element: consecutive threads
// write consecutive elements, the reads are whole columns (width contiguous
// elements) at the permuted position.  Every column is written once (perm is
// a permutation), so the scatter needs no zero fill.  HBM-bound: 2 x n x
// width x elem_bytes + 4 n bytes.
#include "internal.hpp"

namespace oct {

template <typename E, bool SCATTER>
__global__ void __launch_bounds__(256)
k_zsort(const E* __restrict__ src, E* __restrict__ dst, const int32_t* __restrict__ perm, int64_t n, int64_t width,
        int index_base) {
    const int64_t total = n * width;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t k = e / width;
        const int64_t w = e - k * width;
        const int64_t p = (int64_t)(perm[k] - index_base);
        if (SCATTER)
            dst[p * width + w] = src[e];  // grad[:, perm[k]] = dy[:, k]
        else
            dst[e] = src[p * width + w];  // out[:, k] = in[:, perm[k]]
    }
}

template <typename E>
static hipError_t launch_zsort_t(const void* src, void* dst, const int32_t* perm, int64_t n, int64_t width,
                                 int index_base, bool scatter, hipStream_t s) {
    const int64_t total = n * width;
    const int64_t want = (total + 255) / 256;
    const unsigned g = (unsigned)(want < 1 ? 1 : (want > 65536 ? 65536 : want));
    if (scatter)
        hipLaunchKernelGGL((k_zsort<E, true>), dim3(g), dim3(256), 0, s, (const E*)src, (E*)dst, perm, n, width,
                           index_base);
    else
        hipLaunchKernelGGL((k_zsort<E, false>), dim3(g), dim3(256), 0, s, (const E*)src, (E*)dst, perm, n, width,
                           index_base);
    return hipGetLastError();
}

hipError_t launch_zsort(const void* src, void* dst, const int32_t* perm, int64_t n, int64_t width, int elem_bytes,
                        int index_base, bool scatter, hipStream_t s) {
    if (n <= 0 || width <= 0) return hipSuccess;
    switch (elem_bytes) {
        case 1: return launch_zsort_t<uint8_t>(src, dst, perm, n, width, index_base, scatter, s);
        case 2: return launch_zsort_t<uint16_t>(src, dst, perm, n, width, index_base, scatter, s);
        case 4: return launch_zsort_t<uint32_t>(src, dst, perm, n, width, index_base, scatter, s);
        case 8: return launch_zsort_t<uint64_t>(src, dst, perm, n, width, index_base, scatter, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace oct
