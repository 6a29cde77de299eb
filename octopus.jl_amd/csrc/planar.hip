// 2D point sets (quadtrees, src/morton.jl:50-80 and the NDIMS = 2 branches
// of src/cpu/*.jl) on the 3D engine.
//
// A 2D set is built as the 3D set (x, y, PLANE_Z) with PLANE_Z = 2^20: the
// origin's z is min z - Float32(1e-6), which rounds back to exactly 2^20, so
// every point's z cell is 0 and its octant is its quadrant (x bit | y bit << 1).
// The 3D distance (dx*dx + dy*dy) + 0*0 is the 2D one bit for bit, and the
// octree of those points is the reference's quadtree without its top ten
// levels: with cells below 2^21 per axis the quadtree's levels 0-9 split on
// zero bits (one child each, all points), and its levels 10-31 are the
// octree's levels 0-21 node for node.  These kernels pad the caller's (2, N)
// coordinates, drop the z column of rel_displacement, and write the
// reference-layout quadtree arrays from the octree's.
#include "internal.hpp"

namespace oct {

__global__ void __launch_bounds__(256)
k_pad_planar(const float* __restrict__ xy, int64_t n, float* __restrict__ xyz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 p = reinterpret_cast<const float2*>(xy)[i];
    xyz[3 * i + 0] = p.x;
    xyz[3 * i + 1] = p.y;
    xyz[3 * i + 2] = PLANE_Z;
}

// rel_displacement (3, E) -> (2, E)
__global__ void __launch_bounds__(256)
k_disp_planar(const float* __restrict__ d3, int64_t n_edges, float* __restrict__ d2) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_edges) return;
    reinterpret_cast<float2*>(d2)[e] = make_float2(d3[3 * e], d3[3 * e + 1]);
}

// Quadtree arrays (reference layout, index base applied like the octree's)
// from the octree's: `chain` one-child nodes on top, then octree node k as
// node k + chain with its first four children (the last four are empty).
template <typename T>
__global__ void __launch_bounds__(256)
k_quadtree(const int32_t* __restrict__ ofirst, const int32_t* __restrict__ olast, const int32_t* __restrict__ okids,
           const T* __restrict__ obmin, const T* __restrict__ obmax, int32_t m3, int32_t chain,
           int64_t n, int index_base, int32_t* __restrict__ qfirst, int32_t* __restrict__ qlast,
           int32_t* __restrict__ qkids, T* __restrict__ qbmin, T* __restrict__ qbmax) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)m3 + chain) return;
    const int64_t k = id < chain ? 0 : id - chain;
    if (id < chain) {
        qfirst[id] = index_base;
        qlast[id] = (int32_t)(n - 1) + index_base;
        qkids[4 * id] = (int32_t)id + 1 + index_base;
        for (int c = 1; c < 4; ++c) qkids[4 * id + c] = -1;
    } else {
        qfirst[id] = ofirst[k];
        qlast[id] = olast[k];
        for (int c = 0; c < 4; ++c) {
            const int32_t ch = okids[8 * k + c];
            qkids[4 * id + c] = ch == -1 ? -1 : ch + chain;
        }
    }
    qbmin[2 * id] = obmin[3 * k];
    qbmin[2 * id + 1] = obmin[3 * k + 1];
    qbmax[2 * id] = obmax[3 * k];
    qbmax[2 * id + 1] = obmax[3 * k + 1];
}

hipError_t launch_pad_planar(const float* xy, int64_t n, float* xyz, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pad_planar, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xy, n, xyz);
    return hipGetLastError();
}

hipError_t launch_disp_planar(const float* d3, int64_t n_edges, float* d2, hipStream_t s) {
    if (n_edges <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_disp_planar, dim3((unsigned)((n_edges + 255) / 256)), dim3(256), 0, s, d3, n_edges, d2);
    return hipGetLastError();
}

hipError_t launch_quadtree(const int32_t* ofirst, const int32_t* olast, const int32_t* okids, const float* obmin,
                           const float* obmax, int32_t m3, int32_t chain, int64_t n, int index_base, int32_t* qfirst,
                           int32_t* qlast, int32_t* qkids, float* qbmin, float* qbmax, hipStream_t s) {
    const int64_t m2 = (int64_t)m3 + chain;
    if (m2 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_quadtree<float>, dim3((unsigned)((m2 + 255) / 256)), dim3(256), 0, s, ofirst, olast, okids,
                       obmin, obmax, m3, chain, n, index_base, qfirst, qlast, qkids, qbmin, qbmax);
    return hipGetLastError();
}

hipError_t launch_quadtree64(const int32_t* ofirst, const int32_t* olast, const int32_t* okids, const double* obmin,
                             const double* obmax, int32_t m3, int32_t chain, int64_t n, int index_base,
                             int32_t* qfirst, int32_t* qlast, int32_t* qkids, double* qbmin, double* qbmax,
                             hipStream_t s) {
    const int64_t m2 = (int64_t)m3 + chain;
    if (m2 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_quadtree<double>, dim3((unsigned)((m2 + 255) / 256)), dim3(256), 0, s, ofirst, olast,
                       okids, obmin, obmax, m3, chain, n, index_base, qfirst, qlast, qkids, qbmin, qbmax);
    return hipGetLastError();
}

}  // namespace oct
