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
    float radius;          // search radius of the set (batched builds: per set)
    uint32_t changed;      // refit: points whose key changed
    uint32_t refit_fail;   // refit: 1 when it could not run (too many changed points, or a new key width)
};

// Batched builds (many sets in one forest): keys carry the set id above bit
// BATCH_SHIFT and the set's own Morton key below it (at most 14 bits per
// axis), so one sort orders all sets and a change of set id is a tree
// boundary.  A single set has batched = 0 and plain 63-bit keys.
constexpr int BATCH_SHIFT = 42;
constexpr uint64_t BATCH_MORTON = (1ull << BATCH_SHIFT) - 1ull;
struct Forest {
    int batched;
    const int64_t* set_off;  // S + 1 offsets of the sets in sorted order (positions)
    BuildParams* sprm;       // per-set parameters (the single set: the set's own)
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

// ----- bounds reduction pieces (build.hip, batch.hip)

constexpr int BND_THREADS = 256;

// Ordered-uint min/max of the coordinates of points i0 .. i1, the share of
// block lb of nb (grid-stride).
__device__ inline void bounds_part(const float* __restrict__ coords, int64_t i0, int64_t i1, int lb, int nb,
                                   uint32_t* __restrict__ out) {
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
    uint32_t mx[3] = {0u, 0u, 0u};
    const int64_t stride = (int64_t)nb * blockDim.x;
    for (int64_t i = i0 + (int64_t)lb * blockDim.x + threadIdx.x; i < i1; i += stride) {
        for (int d = 0; d < 3; ++d) {
            const uint32_t o = f2o(coords[3 * i + d]);
            mn[d] = o < mn[d] ? o : mn[d];
            mx[d] = o > mx[d] ? o : mx[d];
        }
    }
    __shared__ uint32_t red[6][BND_THREADS / OCT_WAVE];
    const int w = threadIdx.x / OCT_WAVE;
    for (int d = 0; d < 3; ++d) {
        const uint32_t a = wave_min_u(mn[d]);
        const uint32_t b = wave_max_u(mx[d]);
        if (lane_id() == 0) { red[d][w] = a; red[3 + d][w] = b; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        uint32_t v = red[d][0];
        for (int k = 1; k < BND_THREADS / OCT_WAVE; ++k)
            v = d < 3 ? (red[d][k] < v ? red[d][k] : v) : (red[d][k] > v ? red[d][k] : v);
        out[d] = v;
    }
}

// One block folds the partials b0 .. b1 and derives origin, 1/r and the live
// key bits of one set.
__device__ inline void bounds_fold(const uint32_t* __restrict__ partials, int b0, int b1, float radius,
                                   BuildParams* __restrict__ prm) {
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
    uint32_t mx[3] = {0u, 0u, 0u};
    for (int b = b0 + (int)threadIdx.x; b < b1; b += BND_THREADS)
        for (int d = 0; d < 3; ++d) {
            const uint32_t a = partials[6 * b + d], c = partials[6 * b + 3 + d];
            mn[d] = a < mn[d] ? a : mn[d];
            mx[d] = c > mx[d] ? c : mx[d];
        }
    __shared__ uint32_t red[6][BND_THREADS / OCT_WAVE];
    const int w = threadIdx.x / OCT_WAVE;
    for (int d = 0; d < 3; ++d) {
        const uint32_t a = wave_min_u(mn[d]);
        const uint32_t b = wave_max_u(mx[d]);
        if (lane_id() == 0) { red[d][w] = a; red[3 + d][w] = b; }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint32_t v[6];
    for (int d = 0; d < 6; ++d) v[d] = red[d][0];
    for (int k = 1; k < BND_THREADS / OCT_WAVE; ++k)
        for (int d = 0; d < 3; ++d) {
            v[d] = red[d][k] < v[d] ? red[d][k] : v[d];
            v[3 + d] = red[3 + d][k] > v[3 + d] ? red[3 + d][k] : v[3 + d];
        }
    const float inv_cs = 1.0f / radius;  // one(T) / cell_size
    uint32_t cmax = 0;
    for (int d = 0; d < 3; ++d) {
        prm->omin[d] = v[d];
        prm->omax[d] = v[3 + d];
        const float o = o2f(v[d]) - 1e-6f;  // ox - T(1e-6)
        prm->origin[d] = o;
        const uint32_t c = floor_clamp((o2f(v[3 + d]) - o) * inv_cs, 0x1fffffu);
        cmax = c > cmax ? c : cmax;
    }
    prm->inv_cs = inv_cs;
    prm->radius = radius;
    prm->changed = 0;
    prm->refit_fail = 0;
    prm->cmax = cmax;
    prm->key_bits = cmax == 0 ? 0 : 3 * (32 - __clz((int)cmax));
}

// Device buffers of one point set's build (see build.hip).
struct BuildBufs {
    BuildParams* prm;
    uint32_t* partials;  // 6 per bounds block
    uint64_t* keys[2];   // radix sort ping-pong
    int32_t* vals[2];
    uint32_t* hist;      // radix histograms, reused for the topology table
    uint32_t* scan_tmp;
    uint64_t* total;     // scan total of the topology table
    uint32_t* topo;      // per sorted position: m + 1, lambda, emit flag, leaf length
    int32_t* perm;       // exported, index base applied
    float4* scoord;      // Morton-sorted coords, w = original index bits
    NodeRec* nodes;
    int32_t *nfirst, *nlast, *nchild, *leaf_ids;
    NodeRec* leaves;     // leaf records in Morton order
    uint8_t* nlevel;
    float *bmin, *bmax;
    uint32_t* rslots;    // refit: the runs cut by a block's edges, refit_slot_words(n)
    int32_t* ranc;       // refit: per level (OCT_MAX_LEVEL rows of n) each leaf's ancestor, -1 none
    int32_t* nfc;        // per node (and one past the last): the id its first child has or would have
    uint8_t* noct;       // per node: its octant in its parent
    uint64_t* rlook;     // refit: tile sums, group masks and prefixes (refit_look_words(n))
};

int bounds_blocks(int64_t n);
int64_t sort_tiles(int64_t n);
int64_t topo_blocks(int64_t n);
int64_t hist_words(int64_t n);
int64_t max_nodes(int64_t n, int T);
int64_t refit_slot_words(int64_t n);
// Bounds, origin and key width of one set (per-block partials, then a fold);
// with `copy`, the coordinates are copied there on the way; with `csrc`, their
// address is stored there (device memory) for the refit kernels.
hipError_t launch_bounds_stage(const float* coords, int64_t n, float radius, float* copy, const BuildBufs& b,
                               hipStream_t s, const float** csrc = nullptr);
hipError_t launch_sort_stage(const float* coords, int64_t n, int key_bits, int index_base, const BuildBufs& b,
                             int* cur_out, hipStream_t s);
constexpr int RADIX_BITS = 9;
hipError_t launch_sort_passes(const float* coords, int64_t n, const int* shifts, int npasses, int index_base,
                              const BuildBufs& b, int* cur_out, hipStream_t s);
hipError_t launch_topology_stage(int64_t n, int key_bits, int T, int index_base, int cur, const BuildBufs& b,
                                 const Forest& f, hipStream_t s);
hipError_t launch_single_set_offsets(int64_t* off, int64_t n, hipStream_t s);

// Almost-sorted refit (time stepping): the previous sort is keys[cur] and
// scoord.  prepare: new keys in that order and the changed count in *b.total;
// finish: prm->changed, and prm->refit_fail = 1 when that count passes
// `limit` (at most REFIT_MAX_CHANGED) or the key width is no longer key_bits;
// sorted keys in keys[cur ^ 1], perm and scoord rewritten, identical to the
// then the full sort's order, or (refit_fail) the previous sort copied to keys[cur ^ 1] with
// perm and scoord left as they were, a valid tree of the previous order that
// the host replaces by a full rebuild.  No host wait in between.
constexpr int64_t REFIT_MAX_CHANGED = 32768;
// The refit's changed-point bookkeeping in BuildBufs::rlook: the changed
// count per tile of 4096 positions (zeroed by the bounds fold,
// refit_look_zero_words), and per group of 64 positions the changed mask and
// the changed count before the group.
struct RefitLook {
    uint64_t* tsum;
    uint64_t* wmask;
    uint32_t* wpre;
};
int64_t refit_look_words(int64_t n);
int64_t refit_look_zero_words(int64_t n);
struct RefitBufs {
    uint64_t* ck;    // changed keys (compacted)
    int32_t* ci;     // changed point indices
    uint64_t* csk;   // changed, sorted by (key, index)
    int32_t* csi;
    uint32_t* cspos; // their final positions
};
// The refit reads the coordinates through *csrc (device memory, written by
// the bounds stage), so a captured refit follows the caller's array.
hipError_t launch_refit_prepare(const float* const* csrc, int64_t n, int64_t limit, int key_bits, int cur,
                                int index_base, const BuildBufs& b, const RefitBufs& r, hipStream_t s);
hipError_t launch_refit_finish(const float* const* csrc, int64_t n, int64_t limit, int cur, int index_base,
                               const BuildBufs& b, const RefitBufs& r, hipStream_t s);

// Edge outputs (see edges.hip).
struct EdgeOut {
    int32_t* senders;
    int32_t* receivers;
    float* rdisp;
    float* rdist;
};

// EDGE_INDEX: the fill pass writing only the target index of each hit
// (CSR neighbour lists, materialize_all_neighbors!), into out.senders.
enum { EDGE_COUNT = 0, EDGE_FILL = 1, EDGE_INDEX = 2 };
// flags[0]: largest candidate count of an overflowing query leaf; flags[1]:
// a walk ran out of stack (rerun with safe = 1)
// Per-leaf set parameters of a forest (batched sets, self pairs): the set of
// a query leaf comes from the key of its first point; batched = 0: one set,
// the kernel arguments apply.
struct SetLookup {
    int batched;
    const uint64_t* qsk;      // sorted keys of the query forest
    const int64_t* set_off;   // S + 1 positions
    const BuildParams* sprm;  // per-set parameters
};
// ldir (optional): one direction code per list entry, for the predictor.
// lheavy (optional): per query leaf, 1 when the single pass cannot record it
// in one sweep ("heavy": more than 64 queries or candidate leaves, or more
// than 1024 candidate points); heavy_list[0 .. *n_heavy) lists those leaves
// in no particular order (*n_heavy must be zero before the call).
hipError_t launch_leaf_lists(const TreeView& q, const TreeView& t, float r_sq, int cap, int safe, int2* lists,
                             uint8_t* ldir, int32_t* lcount, uint32_t* flags, const SetLookup& sets,
                             uint8_t* lheavy, int32_t* heavy_list, uint32_t* n_heavy, hipStream_t s);
// *dst = *src with n_leaves clamped to rows (*over = 1 if it was, else 0):
// the edge kernels of a tree whose leaf count is not on the host yet read the
// clamped copy, so they never index past per-leaf arrays of `rows` entries.
// raw[0..1] receive *src and *tsrc unchanged (read back with the flags).
// The same launch zeroes z0[0 .. n0) and z1[0 .. n1) (the call's flag words
// and the single pass's run states), which saves two fill launches.
hipError_t launch_clamp_leaves(const BuildParams* src, const BuildParams* tsrc, BuildParams* dst, int64_t rows,
                               uint32_t* over, BuildParams* raw, uint32_t* z0, int64_t n0, uint32_t* z1, int64_t n1,
                               hipStream_t s);
// Edges of the listed heavy leaves after the single pass (which leaves each
// one's output position in qoff at its first query); flags = the single
// pass's (nothing runs when it raised flags[0] or flags[1]).
hipError_t launch_edges_heavy_fill(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                                   int cap, float r_sq, float inv_r, int self_pair, const int32_t* heavy_list,
                                   const uint32_t* n_heavy, uint32_t* qcount, uint32_t* qoff, const uint32_t* flags,
                                   EdgeOut out, int index_base, const SetLookup& sets, hipStream_t s);
// The heavy leaves' edges after the single pass, one block per leaf, query by
// query through the leaf's candidates staged in LDS (k_edges_dense_fill):
// the same edges in the same order as launch_edges_heavy_fill.
hipError_t launch_edges_dense_fill(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                                   int cap, float r_sq, float inv_r, int self_pair, const int32_t* heavy_list,
                                   const uint32_t* n_heavy, uint32_t* qcount, const uint32_t* qoff,
                                   const uint32_t* flags, EdgeOut out, int index_base, const SetLookup& sets,
                                   hipStream_t s);
// The same staging and sweep counting instead: per-query counts into qcount,
// each heavy leaf's total into lhits (in place of launch_edges_heavy).
hipError_t launch_edges_dense_count(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                                    int cap, float r_sq, float inv_r, int self_pair, const int32_t* heavy_list,
                                    const uint32_t* n_heavy, uint32_t* qcount, uint32_t* lhits,
                                    const SetLookup& sets, hipStream_t s);
// Exact counts of the listed heavy leaves ahead of the single pass: per-query
// counts into qcount, leaf totals into lhits.
hipError_t launch_edges_heavy(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                              int cap, float r_sq, float inv_r, int self_pair, const int32_t* heavy_list,
                              const uint32_t* n_heavy, uint32_t* qcount, uint32_t* lhits, const SetLookup& sets,
                              hipStream_t s);
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
// lheavy (optional, needs lhits): leaves launch_edges_heavy counted already.
// run_chunk: runs a block takes per ticket (1..SP_MAX_CHUNK).
// sig (optional, fixed runs only): host-visible memory; the last run copies
// sig_words words of sig_src there and then writes 1 + the edge total as the
// 64-bit word at sig + SP_SIG_MARK, while the other runs' phase B goes on.
hipError_t launch_edges_single(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                               int cap, float r_sq, float inv_r, int self_pair, uint32_t* qcount, uint32_t* qoff,
                               const int32_t* wave_first, const int32_t* plan_waves, uint32_t* lhits,
                               uint64_t* tile_state, uint32_t* ticket, uint32_t* flags, uint32_t* stats,
                               uint64_t cap_edges, EdgeOut out, int index_base, int leaves_per_run,
                               int run_chunk, const SetLookup& sets, const uint8_t* lheavy, uint32_t* sig,
                               const uint32_t* sig_src, int sig_words, hipStream_t s);
constexpr int SP_SIG_MARK = 128;      // 32-bit word of the early-total marker in the signal block
constexpr int SP_DEFAULT_RUN = 6;    // leaves per run of the single pass without a plan
constexpr int SP_DEFAULT_CHUNK = 4;  // runs per block ticket of the single pass
constexpr int SP_MAX_CHUNK = 4;      // at most one run per wave of a block

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
// The same as a compile-time constant: one row record (normalised inputs,
// both layers' activations and deltas, the output delta and the loss) per
// row, then the trained parameters' partial sums per chunk of rows.
constexpr int64_t MLP_ROW_FLOATS = MLP_IN + 4 * MLP_H + 2;
constexpr int64_t mlp_scratch_floats_c(int64_t batch) {
    return batch * MLP_ROW_FLOATS + ((batch + MLP_ROWS_PER_CHUNK - 1) / MLP_ROWS_PER_CHUNK) * MLP_TRAINED;
}
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
// most max_leaves leaves (leaves flagged in lheavy, when given, cost nothing:
// the single pass does not record them); wave_first[n_runs + 1]; pred_total =
// sum of the rounded-up predictions, heavy leaves included.  work: 4 *
// plan_chunks(n) words; scan_tmp: scan_tmp_words(plan_chunks(n)) words.
int64_t plan_chunks(int64_t n_leaves);
hipError_t launch_predict_plan(const float* feat, int64_t n, const float* P, float* pred, uint32_t budget,
                               int max_leaves, uint32_t* work, uint32_t* scan_tmp, int32_t* wave_first,
                               int32_t* n_runs, uint64_t* pred_total, const uint8_t* lheavy, hipStream_t s);
constexpr uint32_t SP_PLAN_BUDGET = 1536;  // hit-list entries of one wave (edges.hip)
constexpr int SP_PLAN_MAX_LEAVES = 64;     // leaves one wave may own (edges.hip)

// Batched forests (see batch.hip).
hipError_t launch_bounds_batch(const float* coords, const int64_t* off, const int32_t* bprefix, int S,
                               int nblocks, const float* radii, uint32_t* partials, BuildParams* sprm,
                               hipStream_t s);
hipError_t launch_keys_batch(const float* coords, int64_t n, const int64_t* off, int S, const BuildParams* sprm,
                             uint64_t* keys, int32_t* vals, hipStream_t s);
hipError_t launch_batch_check_bits(const BuildParams* sprm, const int64_t* off, int S, int max_bits,
                                   uint32_t* flags, hipStream_t s);
hipError_t launch_batch_edge_offsets(const uint32_t* loff, const uint64_t* total, const NodeRec* leaves,
                                     const BuildParams* gprm, const int64_t* off, int S, int64_t* eoff,
                                     hipStream_t s);

// 2D sets on the 3D engine (see planar.hip): z = PLANE_Z for every point.
constexpr float PLANE_Z = 1048576.0f;  // 2^20
constexpr int PLANAR_CHAIN = 10;       // quadtree levels above the octree's root (31 - 21)
hipError_t launch_pad_planar(const float* xy, int64_t n, float* xyz, hipStream_t s);
hipError_t launch_disp_planar(const float* d3, int64_t n_edges, float* d2, hipStream_t s);
hipError_t launch_quadtree(const int32_t* ofirst, const int32_t* olast, const int32_t* okids, const float* obmin,
                           const float* obmax, int32_t m3, int32_t chain, int64_t n, int index_base, int32_t* qfirst,
                           int32_t* qlast, int32_t* qkids, float* qbmin, float* qbmax, hipStream_t s);

// Float64 point sets (see f64.hip).
struct Params64 {
    double origin[3];  // min - Float64(1e-6)
    double inv_cs;     // 1 / r
    double radius;
};
constexpr double PLANE_Z64 = 17179869184.0;  // 2^34: 2^34 - 1e-6 rounds back to 2^34 in double
int bounds64_blocks(int64_t n);
hipError_t launch_build64(const double* coords, int64_t n, double radius, uint64_t* partials, Params64* p,
                          BuildParams* prm, hipStream_t s);
hipError_t launch_keys64(const double* coords, int64_t n, const Params64* p, uint64_t* keys, int32_t* vals,
                         float* c32, hipStream_t s);
hipError_t launch_refit64(const double* coords, const BuildParams* prm, int64_t max_nodes, const uint8_t* nlevel,
                          const int32_t* nfirst, const int32_t* nlast, const int32_t* nchild, const int32_t* perm,
                          int base, double* bmin, double* bmax, hipStream_t s);
hipError_t launch_edges64(int mode, const double* cq, int64_t nq, const double* ct, const double* bmin,
                          const double* bmax, const int32_t* nfirst, const int32_t* nlast, const int32_t* nchild,
                          const int32_t* perm, int base, double r_sq, double inv_r, int self_pair, uint32_t* qcount,
                          const uint32_t* qoff, int32_t* senders, int32_t* receivers, double* rdisp, double* rdist,
                          hipStream_t s);
hipError_t launch_pad_planar64(const double* xy, int64_t n, double* xyz, hipStream_t s);
hipError_t launch_disp_planar64(const double* d3, int64_t n_edges, double* d2, hipStream_t s);
hipError_t launch_quadtree64(const int32_t* ofirst, const int32_t* olast, const int32_t* okids, const double* obmin,
                             const double* obmax, int32_t m3, int32_t chain, int64_t n, int index_base,
                             int32_t* qfirst, int32_t* qlast, int32_t* qkids, double* qbmin, double* qbmax,
                             hipStream_t s);

// apply_zsort gather (scatter = its adjoint) of n columns of `width`
// elements of elem_bytes (1, 2, 4 or 8) bytes (see zsort.hip).
hipError_t launch_zsort(const void* src, void* dst, const int32_t* perm, int64_t n, int64_t width, int elem_bytes,
                        int index_base, bool scatter, hipStream_t s);

// Adjoint (see backward.hip).
// scratch (pullback_scratch_words(n_points) words, or null for the push
// form): the pull form's run table; the result is the same up to FP32
// summation order.
int64_t pullback_scratch_words(int64_t n_points);
hipError_t launch_pullback(int64_t n_points, int64_t n_edges, const int32_t* senders,
                           const int32_t* receivers, const float* rdisp, const float* rdist,
                           const float* d_disp, const float* d_dist, float r, int index_base,
                           float* grad, uint32_t* scratch, int dims, hipStream_t s);
// The same in double (Float64 sets): every pointer holds doubles.
hipError_t launch_pullback64(int64_t n_points, int64_t n_edges, const int32_t* senders, const int32_t* receivers,
                             const double* rdisp, const double* rdist, const double* d_disp, const double* d_dist,
                             double r, int index_base, double* grad, uint32_t* scratch, int dims, hipStream_t s);

int64_t scan_tmp_words(int64_t n);
// Exclusive scan of n uint32 counts; the optional 64-bit total is exact even
// when the uint32 offsets would wrap (callers check it first).
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp,
                          uint64_t* total_out, hipStream_t s);

}  // namespace oct
