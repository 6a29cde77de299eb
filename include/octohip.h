/*
 * octohip: C ABI of the MI355X fixed-radius neighbour search engine.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses on
 * the context's device.  Every function returns OCT_OK (0) or a negative
 * status; oct_last_error() gives the message.  A context is not reentrant:
 * all work of one context is ordered on its stream, and the host waits only
 * where a host value is returned.
 *
 * Each entry point names the reference interface it replaces
 * (una-auxme/Octopus.jl, file:line).  Index conventions follow index_base
 * (1 for a Julia caller, 0 for C / Python): set ids, point indices, perm
 * entries, node ids in node_first/node_last/node_children and
 * senders/receivers all carry it; -1 stays the empty-child sentinel.
 */
#ifndef OCTOHIP_H
#define OCTOHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the front-ends map them to the reference's exception types */
#define OCT_OK 0
#define OCT_ERR_ARGUMENT (-1)     /* ArgumentError   (src/api.jl:25,184,190) */
#define OCT_ERR_DIMENSION (-2)    /* DimensionMismatch (src/api.jl:97,137-138) */
#define OCT_ERR_STATE (-3)        /* ErrorException  (src/api.jl:306,452) */
#define OCT_ERR_DEVICE (-4)       /* HIP runtime failure */
#define OCT_ERR_CAPACITY (-5)     /* size beyond a documented limit (E >= 2^31) */

/* build_edges modes */
#define OCT_EDGES_EXACT 0         /* count pass, scan, fill pass (two traversals) */
#define OCT_EDGES_LEARNED 1       /* single pass sized by the per-leaf count predictor */

/* edge order */
#define OCT_ORDER_LEAF 0          /* receiver runs in Morton (tree leaf) order */
#define OCT_ORDER_RECEIVER 1      /* receivers ascending, as the reference writes them */

typedef struct oct_ctx oct_ctx;

/* Flat-COO edges of one active pair (src/types.jl:84-100 EdgeBuffer).
 * Device arrays owned by the context, valid until the next build on the pair
 * (the reference aliases its EdgeBuffer the same way, test/contracts.jl:223-240). */
typedef struct oct_edges {
    int32_t* senders;          /* E, target index j */
    int32_t* receivers;        /* E, query index i */
    float* rel_displacement;   /* D*E (D = the context's ndims), (q_i - t_j) / r, D components per edge */
    float* rel_dist_norm;      /* E, |q_i - t_j| / r */
    int64_t n_edges;
} oct_edges;

/* Tree of one point set, reference layout (src/types.jl:23-42), device arrays
 * owned by the context, valid until the next oct_run.  In a 2D context
 * (quadtrees) the bounds hold 2 floats and node_children 4 ids per node, and
 * origin[2] is 0. */
typedef struct oct_tree {
    float* node_bounds_min;    /* D*M */
    float* node_bounds_max;    /* D*M */
    int32_t* node_first;       /* M, inclusive first position into perm */
    int32_t* node_last;        /* M, inclusive last position */
    int32_t* node_children;    /* 2^D*M, octant (quadrant) order, -1 empty; a leaf has all empty */
    int32_t* permutation;      /* N, sorted position -> point index */
    int32_t n_nodes;
    int32_t n_leaves;
    int64_t n_points;
    float origin[3];
} oct_tree;

/* Borrowed device handle for user kernels (ext/OctopusCUDAExt.jl:23-35
 * DeviceView); pass it to oct_for_each_neighbor_device (octohip_device.h).
 * Invalidated by the next oct_run. */
typedef struct oct_view {
    const float* coords_q;     /* query coords, 3*Nq (2D: the (x, y, 2^20) copy) */
    const float* coords_t;     /* target coords, 3*Nt (likewise) */
    const float* bounds_min;   /* reference-layout tree, as in oct_tree */
    const float* bounds_max;
    const int32_t* node_first;
    const int32_t* node_last;
    const int32_t* node_children;
    const int32_t* perm;
    const void* nodes;         /* packed traversal records of the target tree */
    const float* sorted_t;     /* 4*Nt Morton-sorted target coords, w = point index bits */
    int32_t n_nodes;
    int32_t root;              /* first node with more than one child (or the lone leaf) */
    float radius;
    float radius_sq;
    int64_t n_q;
    int32_t index_base;
    int32_t pad;
} oct_view;

/* TNS(T; ndims) (src/types.jl:167-183).  ndims 3 (octrees) or 2
 * (quadtrees: point sets are (2, n) arrays, xy interleaved; the library
 * builds a 2D set as the octree of (x, y, 2^20), whose edges are the 2D
 * ones bit for bit and whose nodes are the quadtree's below its top ten
 * one-child levels, and writes the reference's quadtree arrays from it;
 * batches are 3D only).  device is a HIP ordinal.  The context creates its
 * own stream unless oct_set_stream is used. */
int oct_create(oct_ctx** out, int device, int ndims, int index_base);
/* TNS(Float64; ndims) (test/float64.jl): elem_bytes 8 makes every
 * coordinate, radius, bound and feature a double (elem_bytes 4 is
 * oct_create).  A Float64 context builds with double bounds, origin and
 * cells, walks the tree in double one query per thread, and writes edges in
 * receiver order; float* fields of the result structs (oct_edges features,
 * oct_tree bounds) then point at doubles.  Its radius is set with
 * oct_set_search_radius64, its adjoint is oct_build_edges_backward64; the
 * learned single pass, the refit mode, batches, the device view and the
 * predictor are Float32 only (OCT_ERR_ARGUMENT / OCT_ERR_STATE). */
int oct_create_typed(oct_ctx** out, int device, int ndims, int index_base, int elem_bytes);
int oct_destroy(oct_ctx* ctx);
const char* oct_last_error(const oct_ctx* ctx);
/* Run the context's work on this stream from now on (work already queued on
 * the previous stream stays ahead of it; no host wait unless the previous
 * stream was the context's own).  Setting the current stream again is free,
 * so a caller may pass its current stream before every call. */
int oct_set_stream(oct_ctx* ctx, void* hip_stream);
/* No counterpart in the reference (Julia tasks order work on one stream).
 * direction 0: the context's stream waits for the work queued so far on
 * hip_stream; 1: hip_stream waits for the work queued so far on the
 * context's stream.  One reused event per direction, no host wait. */
int oct_stream_join(oct_ctx* ctx, void* hip_stream, int direction);

/* set_search_radius! (src/api.jl:24-29): r > 0, marks every set dirty. */
int oct_set_search_radius(oct_ctx* ctx, float r);
/* set_search_radius! with a double radius: T(r) of the context's type. */
int oct_set_search_radius64(oct_ctx* ctx, double r);
/* set_refit_mode! (src/api.jl:43-47): keep keys and perm between builds and
 * try the almost-sorted path first; marks every set dirty. */
int oct_set_refit_mode(oct_ctx* ctx, int on);
/* target_leaf_size (src/types.jl:177, fixed to 32 there); 1..256. */
int oct_set_leaf_size(oct_ctx* ctx, int leaf_size);

/* add_point_set! (src/api.jl:79-109): coords is a caller-owned device array of
 * n points, xyz (2D: xy) interleaved (the bytes of a column-major (D, n)
 * matrix); it is referenced, never copied (a 2D set's build reads a padded
 * copy made at each run).  Writes the set id. */
int oct_add_point_set(oct_ctx* ctx, const float* coords, int64_t n, int* set_id);
/* update_point_set! (src/api.jl:131-140): same n required. */
int oct_update_point_set(oct_ctx* ctx, int set_id, const float* coords, int64_t n);
/* resize_point_set! (src/api.jl:111-129): n may change. */
int oct_resize_point_set(oct_ctx* ctx, int set_id, const float* coords, int64_t n);
/* set_active_search! (src/api.jl:142-165): directed pair, idempotent. */
int oct_set_active_search(oct_ctx* ctx, int q, int t);

/* run! (src/api.jl:183-250, ext/OctopusCUDAExt.jl:214-248): rebuild every
 * dirty set; invalidates all edge buffers. */
int oct_run(oct_ctx* ctx);
/* The built tree of a set (tns.trees[id], tns.permutation[id]). */
int oct_get_tree(oct_ctx* ctx, int set_id, oct_tree* out);

/* How the last oct_run built a set.  In refit mode a set rebuilt with the
 * same point count first tries the almost-sorted path (no counterpart in the
 * reference, which only stores the flag, src/api.jl:43-47,234-238): the
 * previous sorted order is kept and only the points whose key changed are
 * placed, which gives exactly the full sort's order; with more than
 * min(32768, n / 8) changed points the full sort runs instead. */
typedef struct oct_build_info {
    int64_t changed_points;    /* points whose key changed (-1: no previous sort tried) */
    int32_t refit_used;        /* 1: the almost-sorted path built the order */
    int32_t pad;
} oct_build_info;
int oct_build_info_get(oct_ctx* ctx, int set_id, oct_build_info* out);

/* build_edges / build_edges! (src/api.jl:304-337,
 * ext/OctopusCUDAExt.jl:345-415): flat-COO edges of the active pair (q, t);
 * the self pair (q == t) excludes j == i.  mode: OCT_EDGES_*; order:
 * OCT_ORDER_* (OCT_ORDER_RECEIVER needs OCT_EDGES_EXACT).  A set never run
 * has an empty tree and gives no edges (test/backend_errors_cpu.jl:40-57);
 * a set changed since its last run is OCT_ERR_STATE. */
int oct_build_edges(oct_ctx* ctx, int q, int t, int mode, int order, oct_edges* out);

/* CSR neighbour lists of one active pair (tns.neighbor_buffers[pair],
 * src/types.jl:44-48).  Query i's neighbours are flat[offsets[i] ..
 * offsets[i+1]) (offsets are 0-based counts, offsets[0] = 0, as the
 * reference's exclusive scan gives them; flat carries index_base), each
 * slice in traversal order.  Device arrays owned by the context, valid until
 * the next materialize on the pair. */
typedef struct oct_neighbors {
    int32_t* offsets;          /* n_queries + 1 */
    int32_t* flat;             /* total */
    int64_t n_queries;
    int64_t total;
    int32_t materialized;      /* 0 before the first materialize and after each oct_run */
    int32_t pad;
} oct_neighbors;
/* materialize_all_neighbors! for the active pair (q, t) (src/api.jl:348-362,
 * materialize_cpu! src/cpu/query_cpu.jl:139-192, which the reference runs on
 * the CPU only, src/api.jl:350): count pass, exclusive scan, and a fill pass
 * that writes only the target indices, in query order.  The self pair
 * excludes j == i.  Int32 limit on the total as build_edges. */
int oct_materialize_neighbors(oct_ctx* ctx, int q, int t, oct_neighbors* out);
/* The pair's lists as last materialized (materialized = 0: none valid). */
int oct_get_neighbors(oct_ctx* ctx, int q, int t, oct_neighbors* out);

/* apply_zsort (src/api.jl:405-410; apply_zsort_cpu! src/cpu/zsort_cpu.jl:11-32,
 * _apply_zsort_cuda ext/OctopusCUDAExt.jl:422-432): out[:, k] = in[:, perm[k]]
 * with perm the set's z-order permutation.  in and out are distinct device
 * arrays of n (the set's size) columns of `width` elements of elem_bytes
 * (1, 2, 4 or 8) bytes each: a column-major (width, n) matrix, or a vector
 * with width 1.  The permutation must be current (run! after the last
 * change of the set, as prepare_zsort checks, src/api.jl:381-388):
 * OCT_ERR_STATE otherwise. */
int oct_apply_zsort(oct_ctx* ctx, int set_id, const void* in, void* out, int64_t width, int elem_bytes);
/* Its adjoint, the rrule's scatter (ext/OctopusChainRulesCoreExt.jl:65-94):
 * grad[:, perm[k]] = d_out[:, k]; every column of grad is written. */
int oct_apply_zsort_adjoint(oct_ctx* ctx, int set_id, const void* d_out, void* grad, int64_t width,
                            int elem_bytes);

/* device_view (src/api.jl:451-454, ext/OctopusCUDAExt.jl:37-56). */
int oct_device_view(oct_ctx* ctx, int q, int t, oct_view* out);

/* Per-query neighbour counts through the device iterator of octohip_device.h,
 * self not filtered (the reference's test kernel, test/query_cuda.jl:40-56).
 * counts is a device array of view->n_q int32. */
int oct_count_neighbors_device(oct_ctx* ctx, const oct_view* view, int32_t* counts);

/* Pullback of build_edges_diff (ext/OctopusChainRulesCoreExt.jl:137-192,
 * ext/OctopusCUDAChainRulesCoreExt.jl:56-167): grad (D*n_points, device) is
 * overwritten with dL/dcoords.  d_disp (D*E) or d_dist (E) may be NULL for a
 * zero tangent.  Edge indices carry the context's index_base. */
int oct_build_edges_backward(oct_ctx* ctx, int64_t n_points, const oct_edges* edges,
                             const float* d_disp, const float* d_dist, float radius, float* grad);
/* The same for a Float64 context: every array holds doubles. */
int oct_build_edges_backward64(oct_ctx* ctx, int64_t n_points, const oct_edges* edges, const double* d_disp,
                               const double* d_dist, double radius, double* grad);
/* No counterpart in the reference (whose GPU adjoint pushes every term with
 * atomics, ext/OctopusCUDAChainRulesCoreExt.jl:56-99).  pull != 0 (the
 * default): oct_build_edges_backward pulls each edge's sender term from its
 * reverse edge when every receiver's edges form one run (build_edges
 * output, either order) and pushes the rest; 0: every sender term is pushed
 * with an atomic.  Both give dL/dcoords up to FP32 summation order. */
int oct_set_backward_form(oct_ctx* ctx, int pull);

/* Kernel timing by phase with HIP events on the context's stream (off by
 * default).  oct_profile_read waits for the stream, then writes the summed
 * milliseconds and launch counts of the OCT_PROF_NPHASES phases below. */
#define OCT_PROF_BOUNDS 0      /* origin / extent reduction */
#define OCT_PROF_BUILD 1       /* keys, radix sort, topology, refit (a refit step: with its bounds) */
#define OCT_PROF_LEAF_LISTS 2  /* candidate target leaves per query leaf */
#define OCT_PROF_COUNT 3       /* per-query edge counts */
#define OCT_PROF_SCAN 4        /* exclusive scan of the counts */
#define OCT_PROF_FILL 5        /* edge emission */
#define OCT_PROF_BACKWARD 6    /* build_edges_diff adjoint */
#define OCT_PROF_SINGLE 7      /* learned single pass */
#define OCT_PROF_PREDICT 8     /* predictor features, inference and plan */
#define OCT_PROF_NPHASES 9
int oct_profile(oct_ctx* ctx, int enable);
/* The same for the phases whose bits (1 << OCT_PROF_*) are set: timing one
 * phase costs two events per launch of it. */
int oct_profile_phases(oct_ctx* ctx, uint32_t mask);
int oct_profile_read(oct_ctx* ctx, int reset, double* ms, int64_t* calls);

/* Block until all work queued on the context's stream is done. */
int oct_synchronize(oct_ctx* ctx);

/* Buffer lifetime.  Arrays the library hands out (oct_edges, oct_tree,
 * oct_view, oct_edge_stats, oct_batch_edges) hold the last call's results
 * until the next call that rewrites them, as the reference's EdgeBuffer and
 * tree arrays do (test/contracts.jl:223-240).  When a later call needs a
 * larger buffer, the old allocation is not freed but kept, with its stale
 * contents, on the context's retire list, so a view of it stays readable
 * memory (a replaced Julia array stays valid likewise); buffers grow by at
 * least 1.5x, so the list holds at most about twice the live size.
 * oct_trim frees the list (after waiting for the context's stream): views
 * taken before a buffer was replaced must not be read after it.
 * oct_destroy frees everything.  No counterpart in the reference, whose
 * garbage collector frees unreachable arrays. */
int oct_trim(oct_ctx* ctx);
/* Allocations on the retire list. */
int oct_retired_count(const oct_ctx* ctx, int64_t* count);

/* ---- Learned per-leaf hit-count predictor ------------------------------
 * No counterpart in the reference (whose GPU build_edges always runs the
 * exact count pass, ext/OctopusCUDAExt.jl:373-387): a 29 -> 64 -> 64 -> 1
 * ReLU MLP over per-leaf features predicts log1p(hits) of every query leaf.
 * OCT_EDGES_LEARNED uses the predictions to size the edge buffers on a
 * pair's first call and to plan which leaves each wave emits in the single
 * pass (see oct_set_predictor_policy); the edges never depend on it (every leaf is counted exactly before
 * anything is written, a misjudged leaf is recounted).  Parameter block
 * (OCT_MLP_PARAMS floats): feature mean[29], feature std[29], W1[29*64]
 * (row = input), b1[64], W2[64*64] (row = input), b2[64], W3[64], b3[1]. */
#define OCT_MLP_FEATURES 29
#define OCT_MLP_HIDDEN 64
#define OCT_MLP_PARAMS 6203
#define OCT_MLP_TRAINED (OCT_MLP_PARAMS - 2 * OCT_MLP_FEATURES)

/* When OCT_EDGES_LEARNED runs the predictor: by default on a pair's first
 * call and again after a call that needed a second attempt, fell back to the
 * exact path or recounted more than 1/64 of its leaves (otherwise the call
 * reuses the buffers and fixed runs of oct_set_run_leaves leaves per wave,
 * 6 by default); always != 0: on every call. */
int oct_set_predictor_policy(oct_ctx* ctx, int always);
/* Install a parameter block from host memory (NULL removes the predictor). */
int oct_predictor_load(oct_ctx* ctx, const float* host_params);
/* Copy the installed parameter block to host memory (OCT_ERR_STATE if none). */
int oct_predictor_save(oct_ctx* ctx, float* host_params);
/* Count the exact per-leaf hits of the active pair (q, t) (the count pass
 * only, no edges are written) and append its rows (features, log1p of the
 * hits) to the training set; writes the number of rows added. */
int oct_predictor_collect(oct_ctx* ctx, int q, int t, int64_t* rows);
/* Drop the training set. */
int oct_predictor_clear_data(oct_ctx* ctx);
/* Adam on the training set, `steps` steps of `batch` hashed rows (seed
 * picks them).  With init != 0, or no predictor installed, the network
 * starts from seeded He-normal weights and the feature statistics of the
 * training set.  Writes the mean squared error of the last batch. */
int oct_predictor_train(oct_ctx* ctx, int steps, int batch, float lr, uint64_t seed, int init, float* loss);

/* What the last build of the pair did, with device arrays (valid until the
 * next build on the pair) of per-leaf features (n_leaves * 29, row-major),
 * predicted hits and exact hits; predicted is NULL when the predictor was not
 * used, features NULL when they were not computed. */
typedef struct oct_edge_stats {
    const float* features;
    const float* predicted;
    const uint32_t* exact;
    int64_t n_leaves;
    int64_t n_edges;
    int64_t recounted_leaves;  /* single pass: leaves swept twice */
    int64_t recorded_leaves;   /* single pass: leaves swept once */
    double predicted_total;    /* sum of the per-leaf predictions */
    int32_t single_pass;       /* the edges came from the single pass */
    int32_t used_predictor;
    int32_t attempts;          /* passes the call needed (1 when nothing overflowed) */
    int32_t heavy_leaves;      /* single pass: leaves counted ahead and filled after it (too large to
                                  record in one sweep, or beside a dense region) */
} oct_edge_stats;
/* A learned call that returned on the single pass's early total knows its
 * recount counters (recounted / recorded leaves) only when its kernels end:
 * oct_edge_stats_get waits for them; oct_edge_stats_peek never waits and
 * leaves those two counters at the previous call's values until they are in
 * (everything else is the last call's).  For callers that read statistics
 * every step without holding up the stream. */
int oct_edge_stats_get(oct_ctx* ctx, int q, int t, oct_edge_stats* out);
int oct_edge_stats_peek(oct_ctx* ctx, int q, int t, oct_edge_stats* out);

/* ---- Batched sets ---------------------------------------------------------
 * No counterpart in the reference, which runs the sets of a TNS one after
 * the other (src/api.jl:196).  A batch holds many point sets, each searched
 * against itself with its own radius: coords is one device array with the
 * sets concatenated (set s = points offsets[s] .. offsets[s+1], host array of
 * n_sets + 1), radii a host array of n_sets.  Every stage of run! and the
 * single-pass build_edges is one launch for all sets (the sets form one
 * forest); the first oct_batch_run of a configuration runs eagerly and
 * captures the step as a HIP graph, later calls replay it (after updating the
 * coordinates in place; a replay that would overflow a buffer falls back to
 * an eager step).  Each set needs at most 2^14 cells per axis (extent / r).
 * Edge indices are local to their set (plus index_base). */
typedef struct oct_batch_edges {
    int32_t* senders;
    int32_t* receivers;
    float* rel_displacement;      /* 3 * n_edges */
    float* rel_dist_norm;
    int64_t n_edges;
    const int64_t* set_offsets;   /* device, n_sets + 1: set s owns edges [off[s], off[s+1]) */
    int32_t n_sets;
    int32_t replayed;             /* 1: this call replayed the captured graph */
    int64_t recounted_leaves;     /* single pass: leaves swept twice */
    int64_t recorded_leaves;      /* single pass: leaves swept once */
    int64_t heavy_leaves;         /* single pass: leaves counted ahead and filled after it (too large to
                                     record in one sweep, or beside a dense region) */
} oct_batch_edges;
int oct_batch_create(oct_ctx* ctx, int n_sets, const float* coords, const int64_t* offsets, const float* radii,
                     int* batch_id);
int oct_batch_destroy(oct_ctx* ctx, int batch_id);
/* Leaf size (1..256), leaves per single-pass run (1..64), single-pass runs
 * per block ticket (1..4) and graph replay on or off; a change drops the
 * captured graph. */
int oct_batch_config(oct_ctx* ctx, int batch_id, int leaf_size, int run_leaves, int run_chunk, int use_graph);
int oct_batch_run(oct_ctx* ctx, int batch_id, oct_batch_edges* out);

/* ---- Autotuner -------------------------------------------------------------
 * No counterpart in the reference (leaf size 32 and the launch shapes are
 * fixed there, src/types.jl:177).  Each candidate (leaf size, leaves per
 * single-pass run, runs per block ticket) is timed with events over `reps`
 * steps after one untimed step; the fastest is kept and its index
 * ((leaf index * n_runs + run index) * n_chunks + chunk index) written to
 * *best, the times (ms per step) to ms_out[n_leaf_sizes * n_runs * n_chunks]
 * when not NULL.  Results never depend on the choice. */
int oct_batch_autotune(oct_ctx* ctx, int batch_id, const int* leaf_sizes, int n_leaf_sizes, const int* runs,
                       int n_runs, const int* chunks, int n_chunks, int reps, double* ms_out, int* best);
/* The same for the active pair (q, t) of a context: a step is run! of every
 * set plus build_edges(q, t, mode); sets the context's leaf size, run length
 * and runs per ticket. */
int oct_autotune(oct_ctx* ctx, int q, int t, int mode, const int* leaf_sizes, int n_leaf_sizes, const int* runs,
                 int n_runs, const int* chunks, int n_chunks, int reps, double* ms_out, int* best);
/* Leaves per single-pass run without a plan (1..64, default 6). */
int oct_set_run_leaves(oct_ctx* ctx, int run_leaves);
/* Single-pass runs a block takes per ticket (1..4, default 4): larger shares
 * candidate points in a CU's caches, smaller starts every run at once (better
 * for uneven runs). */
int oct_set_run_chunk(oct_ctx* ctx, int run_chunk);

/* The network on n rows of features (device arrays): y = log1p(hits). */
int oct_mlp_forward(oct_ctx* ctx, const float* params, const float* features, int64_t n, float* y);
/* Gradient of mean((f(x) - target)^2) over n rows with respect to the
 * OCT_MLP_TRAINED trained parameters (the block without the feature
 * statistics, same order); loss (device, 1 float) may be NULL. */
int oct_mlp_gradient(oct_ctx* ctx, const float* params, const float* features, const float* targets, int n,
                     float* grad, float* loss);

/* No counterpart in the reference.  oct_mlp_gradient for callers that cannot
 * allocate (code under HIP graph capture, tight training loops, harnesses
 * that forbid hipMalloc): it calls no allocating routine and no host wait;
 * its row records and partial sums live in a fixed device array of the
 * library (~10 MB on each GPU it runs on, held until the process ends), so
 * n is at most OCT_MLP_STATIC_ROWS (more: OCT_ERR_CAPACITY).  grad and loss
 * (loss may be NULL) are written in order on the context's stream.  Every
 * call on one GPU uses that GPU's one array: calls that overlap in time on
 * one GPU (other streams or contexts) race. */
#define OCT_MLP_STATIC_ROWS 8192
int oct_mlp_gradient_noalloc(oct_ctx* ctx, const float* params, const float* features, const float* targets, int n,
                             float* grad, float* loss);

#ifdef __cplusplus
}
#endif
#endif
