// Internal structures shared by the kernels and the C-ABI host code.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace oct {

// Device-resident build parameters of one point set, written by the kernels
// of run! and read by later kernels (no host round trip needed between them).
struct BuildParams {
    uint32_t omin[3];      // ordered-uint min per axis (f2o)
    uint32_t omax[3];      // ordered-uint max per axis
    float origin[3];       // min - Float32(1e-6)  (src/cpu/bin_cpu.jl:12-23)
    float inv_cs;          // 1 / cell_size, cell_size = radius (src/api.jl:217)
    uint32_t cmax;         // largest cell coordinate over the three axes
    int32_t key_bits;      // live bits of the Morton keys (3 * bit length of cmax)
    uint32_t done;         // last-block counter of the bounds reduction (self-resetting)
    int32_t n_nodes;       // M
    int32_t n_leaves;
    int32_t branch_root;   // first node with count == n that is a leaf or has > 1 child
    int32_t chain_level;   // L0: levels 0..L0 hold all n points (one node each)
    int32_t pad;
};

// Packed traversal record of one node (32 B, two 16-B loads):
//   lo = (min x, min y, min z, a), hi = (max x, max y, max z, b)
// leaf:     a = first sorted position, b = point count (> 0)
// internal: a = first child id (children are consecutive BFS ids), b = -child count
struct NodeRec {
    float4 lo;
    float4 hi;
};

__device__ inline int rec_a(const NodeRec& r) { return __float_as_int(r.lo.w); }
__device__ inline int rec_b(const NodeRec& r) { return __float_as_int(r.hi.w); }

// Everything build_edges needs from a built point set.
struct TreeView {
    const NodeRec* nodes;
    const int32_t* leaf_ids;   // leaf node ids in Morton order
    const NodeRec* leaves;     // leaf records in Morton order (a = first position, b = count)
    const float4* scoord;      // Morton-sorted coords, w = original index (int bits)
    const BuildParams* params;
    int32_t n_leaves;
    int32_t n_nodes;
    int32_t branch_root;
    int64_t n;
};

// Device buffers of one point set's build (see build.hip).
struct BuildBufs {
    BuildParams* prm;
    uint32_t* partials;  // 6 per bounds block
    uint64_t* keys[2];   // radix sort ping-pong
    int32_t* vals[2];
    uint32_t* hist;      // radix histograms, reused for the topology table
    uint32_t* scan_tmp;
    uint64_t* total;     // scan total of the topology table
    uint16_t* topo;      // per sorted position: m + 1, lambda, emit flag
    int32_t* perm;       // exported, index base applied
    float4* scoord;      // Morton-sorted coords, w = original index bits
    NodeRec* nodes;
    int32_t *nfirst, *nlast, *nchild, *parent, *leaf_ids;
    NodeRec* leaves;     // leaf records in Morton order
    uint8_t* nlevel;
    float *bmin, *bmax;
};

int bounds_blocks(int64_t n);
int64_t sort_tiles(int64_t n);
int64_t topo_blocks(int64_t n);
int64_t hist_words(int64_t n);
int64_t max_nodes(int64_t n, int T);
hipError_t launch_bounds_stage(const float* coords, int64_t n, float radius, const BuildBufs& b,
                               hipStream_t s);
hipError_t launch_build_stage(const float* coords, int64_t n, int key_bits, int T, int index_base,
                              const BuildBufs& b, hipStream_t s);

// Edge outputs (see edges.hip).
struct EdgeOut {
    int32_t* senders;
    int32_t* receivers;
    float* rdisp;
    float* rdist;
};

enum { EDGE_COUNT = 0, EDGE_FILL = 1 };
// flags[0]: largest candidate count of an overflowing query leaf; flags[1]:
// a walk ran out of stack (rerun with safe = 1)
hipError_t launch_leaf_lists(const TreeView& q, const TreeView& t, float r_sq, int cap, int safe, int2* lists,
                             int32_t* lcount, uint32_t* flags, hipStream_t s);
hipError_t launch_edges(int mode, const TreeView& q, const TreeView& t, const int2* lists,
                        const int32_t* lcount, int cap, float r_sq, float inv_r, int self_pair,
                        int by_receiver, uint32_t* qcount, const uint32_t* qoff, int64_t n_edges, EdgeOut out,
                        int index_base, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
int64_t single_tiles(int64_t n_leaves);
// Single pass: flags[0] capacity exceeded, flags[1] look-back gave up,
// flags[2..3] 64-bit edge total (written by the last tile).
hipError_t launch_edges_single(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                               int cap, float r_sq, float inv_r, int self_pair, uint32_t* qcount, uint32_t* qoff,
                               uint64_t* tile_state, uint32_t* ticket, uint32_t* flags, uint64_t cap_edges,
                               EdgeOut out, int index_base, hipStream_t s);

// Adjoint (see backward.hip).
hipError_t launch_pullback(int64_t n_points, int64_t n_edges, const int32_t* senders,
                           const int32_t* receivers, const float* rdisp, const float* rdist,
                           const float* d_disp, const float* d_dist, float r, int index_base,
                           float* grad, hipStream_t s);

int64_t scan_tmp_words(int64_t n);
// Exclusive scan of n uint32 counts; the optional 64-bit total is exact even
// when the uint32 offsets would wrap (callers check it first).
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                          uint64_t* total_out, hipStream_t s);

}  // namespace oct
