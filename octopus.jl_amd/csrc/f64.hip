// Float64 point sets (TNS(Float64), test/float64.jl): every value the
// reference computes in T is computed in double.
//
//   bounds, origin = min - Float64(1e-6), inv_cs = 1 / r: k_bounds64_*
//   cells floor((p - o) * inv_cs) clamped to 21 bits, Morton keys: k_keys64
//   the stable radix sort and the search-free topology: shared with Float32
//     (they only see keys; a rounded Float32 copy of the coordinates feeds
//     the sort's sorted-coordinate output, which this path does not use)
//   node bounds, exact min/max in double, levels bottom-up: k_refit64
//   traversal and edges: one thread per query, the reference's LIFO walk
//     over the reference-layout tree (src/cpu/iterator_cpu.jl:54-120, the
//     same order as ext/OctopusCUDAExt.jl:262-317), d2 = (dx*dx + dy*dy) +
//     dz*dz and the features (src/cpu/edges_cpu.jl:106-115) in double, each
//     query's edges at its scanned offset (receiver order): k_edges64
// This path favours plainness over speed: the reference calls Float64
// "works but unbenchmarked" (test/float64.jl:11-13).
#include "internal.hpp"

namespace oct {

// order-preserving uint64 image of a double (-0.0 orders below +0.0)
__device__ inline uint64_t d2o(double d) {
    const uint64_t u = (uint64_t)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double o2d(uint64_t o) {
    return __longlong_as_double((long long)((o >> 63) ? (o & 0x7fffffffffffffffull) : ~o));
}
// Julia's Base.min / Base.max on non-NaN doubles
__device__ inline double jmin64(double a, double b) { return d2o(b) < d2o(a) ? b : a; }
__device__ inline double jmax64(double a, double b) { return d2o(b) > d2o(a) ? b : a; }

// floor(Int32, v) clamped to [0, hi] (src/cpu/bin_cpu.jl:64-68)
__device__ inline uint32_t floor_clamp64(double v, uint32_t hi) {
    const double f = floor(v);
    if (!(f >= 0.0)) return 0u;
    if (f >= (double)hi) return hi;
    return (uint32_t)f;
}

constexpr int B64_THREADS = 256;

__global__ void __launch_bounds__(B64_THREADS)
k_bounds64_partial(const double* __restrict__ c, int64_t n, uint64_t* __restrict__ partials) {
    uint64_t mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0ull, 0ull, 0ull};
    for (int64_t i = (int64_t)blockIdx.x * B64_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * B64_THREADS)
        for (int d = 0; d < 3; ++d) {
            const uint64_t o = d2o(c[3 * i + d]);
            mn[d] = o < mn[d] ? o : mn[d];
            mx[d] = o > mx[d] ? o : mx[d];
        }
    __shared__ uint64_t red[6][B64_THREADS];
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = mn[d];
        red[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int o = B64_THREADS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int d = 0; d < 3; ++d) {
                const uint64_t a = red[d][threadIdx.x + o], b = red[3 + d][threadIdx.x + o];
                red[d][threadIdx.x] = a < red[d][threadIdx.x] ? a : red[d][threadIdx.x];
                red[3 + d][threadIdx.x] = b > red[3 + d][threadIdx.x] ? b : red[3 + d][threadIdx.x];
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partials[6 * (int64_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One thread folds the partials: origin, 1/r, the largest cell and the live
// key bits, in double (Params64) and as the Float32 build parameters the
// shared topology stage reads (key width, radius).
__global__ void k_bounds64_fold(const uint64_t* __restrict__ partials, int nb, double radius, Params64* __restrict__ p,
                                BuildParams* __restrict__ prm) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0ull, 0ull, 0ull};
    for (int b = 0; b < nb; ++b)
        for (int d = 0; d < 3; ++d) {
            mn[d] = partials[6 * b + d] < mn[d] ? partials[6 * b + d] : mn[d];
            mx[d] = partials[6 * b + 3 + d] > mx[d] ? partials[6 * b + 3 + d] : mx[d];
        }
    const double inv_cs = 1.0 / radius;  // one(T) / cell_size
    uint32_t cmax = 0;
    for (int d = 0; d < 3; ++d) {
        const double o = o2d(mn[d]) - 1e-6;  // ox - T(1e-6)
        p->origin[d] = o;
        const uint32_t c = floor_clamp64((o2d(mx[d]) - o) * inv_cs, 0x1fffffu);
        cmax = c > cmax ? c : cmax;
        prm->origin[d] = (float)o;
    }
    p->inv_cs = inv_cs;
    p->radius = radius;
    prm->inv_cs = (float)inv_cs;
    prm->radius = (float)radius;
    prm->changed = 0;
    prm->refit_fail = 0;
    prm->cmax = cmax;
    prm->key_bits = cmax == 0 ? 0 : 3 * (32 - __clz((int)cmax));
}

// Keys (src/morton.jl:35-40, 83-88 in double), the index payload and the
// rounded Float32 copy the sort's last pass reads.
__global__ void __launch_bounds__(256)
k_keys64(const double* __restrict__ c, int64_t n, const Params64* __restrict__ p, uint64_t* __restrict__ keys,
         int32_t* __restrict__ vals, float* __restrict__ c32) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double x = c[3 * i], y = c[3 * i + 1], z = c[3 * i + 2];
    const double inv = p->inv_cs;
    keys[i] = morton3(floor_clamp64((x - p->origin[0]) * inv, 0x1fffffu),
                      floor_clamp64((y - p->origin[1]) * inv, 0x1fffffu),
                      floor_clamp64((z - p->origin[2]) * inv, 0x1fffffu));
    vals[i] = (int32_t)i;
    c32[3 * i] = (float)x;
    c32[3 * i + 1] = (float)y;
    c32[3 * i + 2] = (float)z;
}

// Node bounds of one level (src/cpu/refit_cpu.jl:21-70 in double): a leaf
// folds its points (through perm), an internal node its children.  Launched
// from the deepest level up; node arrays carry the index base.
__global__ void __launch_bounds__(256)
k_refit64(const double* __restrict__ c, const BuildParams* __restrict__ prm, int64_t max_nodes, int level,
          const uint8_t* __restrict__ nlevel, const int32_t* __restrict__ nfirst, const int32_t* __restrict__ nlast,
          const int32_t* __restrict__ nchild, const int32_t* __restrict__ perm, int base, double* __restrict__ bmin,
          double* __restrict__ bmax) {
    const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t m = prm->n_nodes;
    if (id >= m || id >= max_nodes || (int)nlevel[id] != level) return;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const int32_t* kids = nchild + 8 * id;
    bool leaf = true;
    for (int k = 0; k < 8; ++k) leaf = leaf && kids[k] == -1;
    if (leaf) {
        for (int32_t s = nfirst[id] - base; s <= nlast[id] - base; ++s) {
            const int64_t j = perm[s] - base;
            for (int d = 0; d < 3; ++d) {
                lo[d] = jmin64(lo[d], c[3 * j + d]);
                hi[d] = jmax64(hi[d], c[3 * j + d]);
            }
        }
    } else {
        for (int k = 0; k < 8; ++k) {
            if (kids[k] == -1) continue;
            const int64_t ch = kids[k] - base;
            for (int d = 0; d < 3; ++d) {
                lo[d] = jmin64(lo[d], bmin[3 * ch + d]);
                hi[d] = jmax64(hi[d], bmax[3 * ch + d]);
            }
        }
    }
    for (int d = 0; d < 3; ++d) {
        bmin[3 * id + d] = lo[d];
        bmax[3 * id + d] = hi[d];
    }
}

// src/cpu/iterator_cpu.jl:31-40 (_aabb_dist_sq) in double
__device__ inline double box_d2_64(double px, double py, double pz, const double* lo, const double* hi) {
    double dx = fmax(lo[0] - px, 0.0);
    dx = fmax(dx, px - hi[0]);
    double dy = fmax(lo[1] - py, 0.0);
    dy = fmax(dy, py - hi[1]);
    double dz = fmax(lo[2] - pz, 0.0);
    dz = fmax(dz, pz - hi[2]);
    return dx * dx + dy * dy + dz * dz;
}

constexpr int STACK64 = 1 + 7 * OCT_MAX_LEVEL;  // the LIFO bound (src/cpu/build_cpu.jl:25-43)

// One query per thread: the LIFO walk of the reference (a node whose first
// child slot is empty scans its range like a leaf, as the reference's
// kids[1] == -1 test does; the edge set is the same either way).
//   MODE EDGE_COUNT: qcount[i] = neighbours (j == i dropped for the self pair)
//   MODE EDGE_FILL:  edges at qoff[i], features in double
//   MODE EDGE_INDEX: targets only (CSR)
template <int MODE>
__global__ void __launch_bounds__(128)
k_edges64(const double* __restrict__ cq, int64_t nq, const double* __restrict__ ct, const double* __restrict__ bmin,
          const double* __restrict__ bmax, const int32_t* __restrict__ nfirst, const int32_t* __restrict__ nlast,
          const int32_t* __restrict__ nchild, const int32_t* __restrict__ perm, int base, double r_sq, double inv_r,
          int self_pair, uint32_t* __restrict__ qcount, const uint32_t* __restrict__ qoff,
          int32_t* __restrict__ senders, int32_t* __restrict__ receivers, double* __restrict__ rdisp,
          double* __restrict__ rdist) {
    const int64_t i = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (i >= nq) return;
    const double px = cq[3 * i], py = cq[3 * i + 1], pz = cq[3 * i + 2];
    uint32_t cnt = 0;
    uint32_t e = MODE == EDGE_COUNT ? 0u : qoff[i];
    int32_t stack[STACK64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int32_t nid = stack[--sp];
        if (box_d2_64(px, py, pz, bmin + 3 * (int64_t)nid, bmax + 3 * (int64_t)nid) > r_sq) continue;
        const int32_t* kids = nchild + 8 * (int64_t)nid;
        if (kids[0] == -1) {
            for (int32_t s = nfirst[nid] - base; s <= nlast[nid] - base; ++s) {
                const int32_t j = perm[s] - base;
                const double dx = ct[3 * (int64_t)j] - px, dy = ct[3 * (int64_t)j + 1] - py,
                             dz = ct[3 * (int64_t)j + 2] - pz;
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > r_sq || (self_pair && (int64_t)j == i)) continue;
                if (MODE == EDGE_COUNT) {
                    ++cnt;
                } else {
                    senders[e] = j + base;
                    if (MODE == EDGE_FILL) {
                        receivers[e] = (int32_t)i + base;
                        rdisp[3 * (int64_t)e] = -dx * inv_r;
                        rdisp[3 * (int64_t)e + 1] = -dy * inv_r;
                        rdisp[3 * (int64_t)e + 2] = -dz * inv_r;
                        rdist[e] = sqrt(d2) * inv_r;
                    }
                    ++e;
                }
            }
        } else {
            for (int k = 0; k < 8; ++k) {
                const int32_t cid = kids[k];
                if (cid == -1) continue;
                const int32_t ch = cid - base;
                if (box_d2_64(px, py, pz, bmin + 3 * (int64_t)ch, bmax + 3 * (int64_t)ch) <= r_sq && sp < STACK64)
                    stack[sp++] = ch;
            }
        }
    }
    if (MODE == EDGE_COUNT) qcount[i] = cnt;
}

__global__ void __launch_bounds__(256)
k_pad_planar64(const double* __restrict__ xy, int64_t n, double* __restrict__ xyz) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    xyz[3 * i] = xy[2 * i];
    xyz[3 * i + 1] = xy[2 * i + 1];
    xyz[3 * i + 2] = PLANE_Z64;
}

__global__ void __launch_bounds__(256)
k_disp_planar64(const double* __restrict__ d3, int64_t n_edges, double* __restrict__ d2) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n_edges) return;
    d2[2 * e] = d3[3 * e];
    d2[2 * e + 1] = d3[3 * e + 1];
}

// ------------------------------------------------------------ launchers

int bounds64_blocks(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>((n + 4095) / 4096, 1), 1024); }

hipError_t launch_build64(const double* coords, int64_t n, double radius, uint64_t* partials, Params64* p,
                          BuildParams* prm, hipStream_t s) {
    const int nb = bounds64_blocks(n);
    hipLaunchKernelGGL(k_bounds64_partial, dim3(nb), dim3(B64_THREADS), 0, s, coords, n, partials);
    hipLaunchKernelGGL(k_bounds64_fold, dim3(1), dim3(64), 0, s, partials, nb, radius, p, prm);
    return hipGetLastError();
}

hipError_t launch_keys64(const double* coords, int64_t n, const Params64* p, uint64_t* keys, int32_t* vals,
                         float* c32, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, coords, n, p, keys, vals, c32);
    return hipGetLastError();
}

hipError_t launch_refit64(const double* coords, const BuildParams* prm, int64_t max_nodes, const uint8_t* nlevel,
                          const int32_t* nfirst, const int32_t* nlast, const int32_t* nchild, const int32_t* perm,
                          int base, double* bmin, double* bmax, hipStream_t s) {
    const unsigned g = (unsigned)((max_nodes + 255) / 256);
    for (int level = OCT_MAX_LEVEL; level >= 0; --level)
        hipLaunchKernelGGL(k_refit64, dim3(g), dim3(256), 0, s, coords, prm, max_nodes, level, nlevel, nfirst, nlast,
                           nchild, perm, base, bmin, bmax);
    return hipGetLastError();
}

hipError_t launch_edges64(int mode, const double* cq, int64_t nq, const double* ct, const double* bmin,
                          const double* bmax, const int32_t* nfirst, const int32_t* nlast, const int32_t* nchild,
                          const int32_t* perm, int base, double r_sq, double inv_r, int self_pair, uint32_t* qcount,
                          const uint32_t* qoff, int32_t* senders, int32_t* receivers, double* rdisp, double* rdist,
                          hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    const unsigned g = (unsigned)((nq + 127) / 128);
#define OCT_E64(M)                                                                                             \
    hipLaunchKernelGGL(k_edges64<M>, dim3(g), dim3(128), 0, s, cq, nq, ct, bmin, bmax, nfirst, nlast, nchild, perm, \
                       base, r_sq, inv_r, self_pair, qcount, qoff, senders, receivers, rdisp, rdist)
    if (mode == EDGE_COUNT) OCT_E64(EDGE_COUNT);
    else if (mode == EDGE_FILL) OCT_E64(EDGE_FILL);
    else OCT_E64(EDGE_INDEX);
#undef OCT_E64
    return hipGetLastError();
}

hipError_t launch_pad_planar64(const double* xy, int64_t n, double* xyz, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pad_planar64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xy, n, xyz);
    return hipGetLastError();
}

hipError_t launch_disp_planar64(const double* d3, int64_t n_edges, double* d2, hipStream_t s) {
    if (n_edges <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_disp_planar64, dim3((unsigned)((n_edges + 255) / 256)), dim3(256), 0, s, d3, n_edges, d2);
    return hipGetLastError();
}

}  // namespace oct
