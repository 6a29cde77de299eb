// Internal structures shared by the kernels and the C-ABI host code.
#pragma once

#include <algorithm>

#include "common.hpp"
#include "octohip.h"

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
    const uint8_t* nlevel;     // tree level per node
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
hipError_t launch_sort_stage(const float* coords, int64_t n, int key_bits, int index_base, const BuildBufs& b,
                             int* cur_out, hipStream_t s);
hipError_t launch_topology_stage(int64_t n, int key_bits, int T, int index_base, int cur, const BuildBufs& b,
                                 hipStream_t s);

// Almost-sorted refit (time stepping): the previous sort is keys[cur] and
// scoord.  prepare: new keys in that order and the changed count in *b.total;
// finish (changed count <= REFIT_MAX_CHANGED): sorted keys in keys[cur ^ 1],
// perm and scoord rewritten, identical to the full sort's.
constexpr int64_t REFIT_MAX_CHANGED = 32768;
struct RefitBufs {
    uint64_t* ck;    // changed keys (compacted)
    int32_t* ci;     // changed point indices
    uint64_t* csk;   // changed, sorted by (key, index)
    int32_t* csi;
    uint32_t* cspos; // their final positions
};
hipError_t launch_refit_prepare(const float* coords, int64_t n, int cur, const BuildBufs& b, hipStream_t s);
hipError_t launch_refit_finish(const float* coords, int64_t n, int64_t n_changed, int cur, int index_base,
                               const BuildBufs& b, const RefitBufs& r, hipStream_t s);

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
// ldir (optional): one direction code per list entry, for the predictor.
hipError_t launch_leaf_lists(const TreeView& q, const TreeView& t, float r_sq, int cap, int safe, int2* lists,
                             uint8_t* ldir, int32_t* lcount, uint32_t* flags, hipStream_t s);
hipError_t launch_edges(int mode, const TreeView& q, const TreeView& t, const int2* lists,
                        const int32_t* lcount, int cap, float r_sq, float inv_r, int self_pair,
                        int by_receiver, uint32_t* qcount, const uint32_t* qoff, uint32_t* lhits, int64_t n_edges,
                        EdgeOut out, int index_base, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
int64_t single_tiles(int64_t n_leaves);
// Single pass: flags[0] capacity exceeded, flags[1] look-back gave up (1) or
// malformed plan (2), flags[2..3] 64-bit edge total (written by the last
// tile); stats[0] recounted leaves, stats[1] recorded leaves.  wave_first /
// plan_waves: optional plan of leaf runs per wave (null: fixed runs).
hipError_t launch_edges_single(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                               int cap, float r_sq, float inv_r, int self_pair, uint32_t* qcount, uint32_t* qoff,
                               const int32_t* wave_first, const int32_t* plan_waves, uint32_t* lhits,
                               uint64_t* tile_state, uint32_t* ticket, uint32_t* flags, uint32_t* stats,
                               uint64_t cap_edges, EdgeOut out, int index_base, hipStream_t s);

// Per-leaf count predictor (see predictor.hip).  Parameter block layout in
// floats: feature mean and std, then W1 (29 x 64, row = input), b1, W2
// (64 x 64, row = input), b2, W3 (64), b3.
constexpr int MLP_IN = 29;
constexpr int MLP_ADJ = 26;
constexpr int MLP_H = 64;
constexpr int MLP_MU = 0;
constexpr int MLP_SD = MLP_MU + MLP_IN;
constexpr int MLP_W1 = MLP_SD + MLP_IN;
constexpr int MLP_B1 = MLP_W1 + MLP_IN * MLP_H;
constexpr int MLP_W2 = MLP_B1 + MLP_H;
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H;
constexpr int MLP_B3 = MLP_W3 + MLP_H;
constexpr int MLP_PARAMS = MLP_B3 + 1;
constexpr int MLP_TRAINED = MLP_PARAMS - MLP_W1;  // everything but the feature statistics
constexpr int MLP_ROWS_PER_CHUNK = 256;           // rows per partial gradient sum
static_assert(MLP_PARAMS == OCT_MLP_PARAMS, "parameter block size is part of the C-ABI");

hipError_t launch_leaf_features(const TreeView& q, const int2* lists, const uint8_t* ldir, const int32_t* lcount,
                                int cap, float r, float* feat, hipStream_t s);
hipError_t launch_mlp_infer(const float* X, int64_t n, const float* P, float* y, float* hits, hipStream_t s);
int64_t mlp_scratch_floats(int batch);
// One training step on `batch` rows (seed 0: rows 0..batch-1, else hashed
// rows); lr 0 computes the gradient (grad_out) without updating.
hipError_t launch_mlp_step(const float* X, const float* T, int64_t n_data, int batch, uint64_t seed, float* P,
                           float* m, float* v, int step, float lr, float* scratch, float* grad_out, float* loss_out,
                           hipStream_t s);
hipError_t launch_feature_stats(const float* X, int64_t n, float* P, hipStream_t s);
hipError_t launch_log1p_u32(const uint32_t* in, int64_t n, float* out, hipStream_t s);
// Predictions and emission plan from per-leaf features: pred = predicted
// hits; runs of consecutive leaves packed greedily within chunks of 64 leaves
// so that a run's predicted hits fit `budget` hit-list entries and it holds at
// most max_leaves leaves; wave_first[n_runs + 1]; pred_total = sum of the
// rounded-up predictions.  work: 4 * plan_chunks(n) words; scan_tmp:
// scan_tmp_words(plan_chunks(n)) words.
int64_t plan_chunks(int64_t n_leaves);
hipError_t launch_predict_plan(const float* feat, int64_t n, const float* P, float* pred, uint32_t budget,
                               int max_leaves, uint32_t* work, uint32_t* scan_tmp, int32_t* wave_first,
                               int32_t* n_runs, uint64_t* pred_total, hipStream_t s);
constexpr uint32_t SP_PLAN_BUDGET = 1024;  // hit-list entries of one wave (edges.hip)
constexpr int SP_PLAN_MAX_LEAVES = 64;     // leaves one wave may own (edges.hip)

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
