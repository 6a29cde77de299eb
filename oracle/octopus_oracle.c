/*
 * CPU oracle: a plain-C restatement of the reference's CPU path for
 * run! (origin, bin, Morton, sort, BFS build, refit) and build_edges
 * (count pass, serial scan, fill pass), plus the build_edges_diff pullback.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against, and the timed CPU baseline.  Nothing in the product
 * (octopus.jl_amd/) links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Indices here are
 * 0-based (the reference is 1-based Julia): node ids, perm entries, point
 * indices and node_first/node_last all carry the reference value minus one;
 * the empty-child sentinel stays -1.
 *
 * Build with -ffp-contract=off and no fast-math: every FP32 operation below
 * is a single IEEE-rounded operation in the reference's evaluation order,
 * which is what makes the edge set a closed form (see the header of
 * octopus_oracle.h).
 */
#include "octopus_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Julia's min/max on IEEE floats (Base.min/max): NaN-free inputs here; among
 * equal values min prefers -0.0 and max prefers +0.0, which fminf/fmaxf
 * leave unspecified. */
static inline float jmin(float x, float y) {
    if (x < y) return x;
    if (y < x) return y;
    return signbit(x) ? x : y;
}
static inline float jmax(float x, float y) {
    if (x > y) return x;
    if (y > x) return y;
    return signbit(x) ? y : x;
}

/* ------------------------------------------------------------------ Morton */

/* src/morton.jl:13-23 (_split_by_3): spread the low 21 bits, one every 3rd. */
uint64_t orc_split_by_3(uint64_t x) {
    x &= 0x00000000001fffffULL;
    x = (x | (x << 32)) & 0x001f00000000ffffULL;
    x = (x | (x << 16)) & 0x001f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

/* src/morton.jl:25-33 (_compact_by_3). */
uint64_t orc_compact_by_3(uint64_t x) {
    x &= 0x1249249249249249ULL;
    x = (x | (x >> 2)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x >> 4)) & 0x100f00f00f00f00fULL;
    x = (x | (x >> 8)) & 0x001f0000ff0000ffULL;
    x = (x | (x >> 16)) & 0x001f00000000ffffULL;
    x = (x | (x >> 32)) & 0x00000000001fffffULL;
    return x;
}

/* src/morton.jl:35-40: x -> bit 0, y -> bit 1, z -> bit 2. */
uint64_t orc_morton_encode3(uint32_t ix, uint32_t iy, uint32_t iz) {
    return (orc_split_by_3(iz) << 2) | (orc_split_by_3(iy) << 1) | orc_split_by_3(ix);
}

/* src/morton.jl:42-47. */
void orc_morton_decode3(uint64_t code, uint32_t out[3]) {
    out[0] = (uint32_t)orc_compact_by_3(code);
    out[1] = (uint32_t)orc_compact_by_3(code >> 1);
    out[2] = (uint32_t)orc_compact_by_3(code >> 2);
}

/* floor(Int32, v) followed by clamp(., 0, hi) (src/morton.jl:83-88,
 * src/cpu/bin_cpu.jl:64-68).  Julia's floor(Int32, v) throws when v is out of
 * Int32 range; clamping in float first gives the same value whenever the
 * reference does not throw, and a defined value when it would. */
static int32_t floor_clamp(float v, int32_t hi) {
    float f = floorf(v);
    if (!(f >= 0.0f)) return 0; /* negatives (and NaN) clamp to 0 */
    if (f >= (float)hi) return hi;
    return (int32_t)f;
}

/* src/morton.jl:83-88 (bin_point, 3D). */
void orc_bin_point3(float px, float py, float pz, float ox, float oy, float oz,
                    float inv_cs, int32_t out[3]) {
    /* No clamp here: bin_point itself does floor only.  Values outside Int32
     * are the reference's InexactError; callers in tests stay in range. */
    out[0] = (int32_t)floorf((px - ox) * inv_cs);
    out[1] = (int32_t)floorf((py - oy) * inv_cs);
    out[2] = (int32_t)floorf((pz - oz) * inv_cs);
}

/* -------------------------------------------------------------- run! pieces */

/* src/cpu/bin_cpu.jl:12-23 (_point_origin_3d): per-axis min from point 1,
 * then subtract T(1e-6) in FP32. */
void orc_point_origin3(const float* coords, int64_t n, float origin[3]) {
    float ox = coords[0], oy = coords[1], oz = coords[2];
    for (int64_t i = 1; i < n; ++i) {
        const float* p = coords + 3 * i;
        ox = jmin(ox, p[0]);
        oy = jmin(oy, p[1]);
        oz = jmin(oz, p[2]);
    }
    const float eps = 1e-6f; /* Float32(1e-6) */
    origin[0] = ox - eps;
    origin[1] = oy - eps;
    origin[2] = oz - eps;
}

/* src/cpu/bin_cpu.jl:52-73 (bin_cpu!, 3D): parallel over points in chunks of
 * at least 4096 (Polyester @batch minbatch=4096). */
void orc_bin3(const float* coords, int64_t n, const float origin[3], float cell_size,
              uint64_t* keys) {
    const float inv_cs = 1.0f / cell_size; /* one(T) / cell_size */
    const float ox = origin[0], oy = origin[1], oz = origin[2];
#pragma omp parallel for schedule(static, 4096) if (n > 4096)
    for (int64_t i = 0; i < n; ++i) {
        const float* p = coords + 3 * i;
        int32_t ix = floor_clamp((p[0] - ox) * inv_cs, 0x1fffff);
        int32_t iy = floor_clamp((p[1] - oy) * inv_cs, 0x1fffff);
        int32_t iz = floor_clamp((p[2] - oz) * inv_cs, 0x1fffff);
        keys[i] = orc_morton_encode3((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
    }
}

/* src/sort.jl:15-24 (sort_by_key_cpu!): perm with keys[perm] nondecreasing,
 * serial.  The reference uses Julia's QuickSort, whose order among equal keys
 * is unspecified and not observable in any edge set; this restatement breaks
 * ties by point index (a stable merge sort), which the HIP radix sort also
 * does, so perm is comparable entry by entry. */
static void merge_pass(const int32_t* src, int32_t* dst, const uint64_t* keys, int64_t n,
                       int64_t width) {
    for (int64_t lo = 0; lo < n; lo += 2 * width) {
        int64_t mid = lo + width < n ? lo + width : n;
        int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
        int64_t a = lo, b = mid, o = lo;
        while (a < mid && b < hi) {
            if (keys[src[b]] < keys[src[a]]) dst[o++] = src[b++];
            else dst[o++] = src[a++];
        }
        while (a < mid) dst[o++] = src[a++];
        while (b < hi) dst[o++] = src[b++];
    }
}

void orc_sort_by_key(const uint64_t* keys, int64_t n, int32_t* perm) {
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    if (n < 2) return;
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* src = perm;
    int32_t* dst = tmp;
    for (int64_t w = 1; w < n; w *= 2) {
        merge_pass(src, dst, keys, n, w);
        int32_t* t = src; src = dst; dst = t;
    }
    if (src != perm) memcpy(perm, src, sizeof(int32_t) * (size_t)n);
    free(tmp);
}

/* src/cpu/build_cpu.jl:46-49: lowest bit of the 3 Morton bits at level L. */
static int level_shift(int L) { return 3 * 21 - 3 * (L + 1); }

/* src/cpu/build_cpu.jl:55-82 (_partition_octants!): linear scan of the sorted
 * range [f, l]; empty octants keep first = -1. */
static void partition_octants(int32_t oc_first[8], int32_t oc_last[8], const uint64_t* keys,
                              const int32_t* perm, int32_t f, int32_t l, int L) {
    for (int o = 0; o < 8; ++o) { oc_first[o] = -1; oc_last[o] = -2; }
    const int shift = level_shift(L);
    int prev = -1;
    for (int32_t i = f; i <= l; ++i) {
        int o = (int)((keys[perm[i]] >> shift) & 7ULL);
        if (o != prev) {
            if (prev != -1) oc_last[prev] = i - 1;
            oc_first[o] = i;
            prev = o;
        }
    }
    if (prev != -1) oc_last[prev] = l;
}

typedef struct {
    int32_t* first; int32_t* last; int32_t* level; int32_t* children; int64_t cap;
} node_arrays;

static void grow_nodes(node_arrays* a, int64_t needed) {
    if (needed <= a->cap) return;
    int64_t newcap = needed + 16 > (a->cap * 3) / 2 ? needed + 16 : (a->cap * 3) / 2;
    a->first = (int32_t*)realloc(a->first, sizeof(int32_t) * (size_t)newcap);
    a->last = (int32_t*)realloc(a->last, sizeof(int32_t) * (size_t)newcap);
    a->level = (int32_t*)realloc(a->level, sizeof(int32_t) * (size_t)newcap);
    a->children = (int32_t*)realloc(a->children, sizeof(int32_t) * 8 * (size_t)newcap);
    a->cap = newcap;
}

/* src/cpu/build_cpu.jl:100-194 (build_cpu!): BFS top-down; a node is a leaf
 * iff count <= target_leaf_size or level >= 21; children are emitted in
 * octant order 1..8, so node ids are BFS order. */
static void build_tree(orc_tree* t, int32_t leaf_size) {
    const int64_t n = t->n;
    node_arrays a = {0};
    int64_t est = 8 * ((n + leaf_size - 1) / (leaf_size > 1 ? leaf_size : 1)) + 16;
    if (est < 16) est = 16;
    grow_nodes(&a, est);
    a.first[0] = 0;
    a.last[0] = (int32_t)(n - 1);
    a.level[0] = 0;
    int64_t n_nodes = 1, next_open = 0;
    int32_t oc_first[8], oc_last[8];
    while (next_open < n_nodes) {
        const int64_t frontier_end = n_nodes;
        for (int64_t nid = next_open; nid < frontier_end; ++nid) {
            const int32_t f = a.first[nid], l = a.last[nid], lvl = a.level[nid];
            const int32_t count = l - f + 1;
            if (count <= leaf_size || lvl >= 21) {
                for (int o = 0; o < 8; ++o) a.children[8 * nid + o] = -1;
                continue;
            }
            partition_octants(oc_first, oc_last, t->keys, t->perm, f, l, lvl);
            for (int o = 0; o < 8; ++o) {
                if (oc_first[o] == -1) {
                    a.children[8 * nid + o] = -1;
                } else {
                    grow_nodes(&a, n_nodes + 1);
                    const int64_t cid = n_nodes;
                    a.first[cid] = oc_first[o];
                    a.last[cid] = oc_last[o];
                    a.level[cid] = lvl + 1;
                    a.children[8 * nid + o] = (int32_t)cid;
                    n_nodes = cid + 1;
                }
            }
        }
        next_open = frontier_end;
    }
    t->n_nodes = (int32_t)n_nodes;
    t->node_first = a.first;
    t->node_last = a.last;
    t->node_level = a.level;
    t->node_children = a.children;
    t->bmin = (float*)malloc(sizeof(float) * 3 * (size_t)n_nodes);
    t->bmax = (float*)malloc(sizeof(float) * 3 * (size_t)n_nodes);
}

/* src/cpu/refit_cpu.jl:21-70 (refit_bounds_cpu!): reverse node-id order;
 * leaves fold over their points, internal nodes over non-empty children. */
static void refit_tree(orc_tree* t, const float* coords) {
    for (int64_t nid = (int64_t)t->n_nodes - 1; nid >= 0; --nid) {
        const int32_t* kids = t->node_children + 8 * nid;
        for (int d = 0; d < 3; ++d) {
            float lo = INFINITY, hi = -INFINITY;
            if (kids[0] == -1) {
                for (int32_t k = t->node_first[nid]; k <= t->node_last[nid]; ++k) {
                    const float x = coords[3 * (int64_t)t->perm[k] + d];
                    lo = jmin(lo, x);
                    hi = jmax(hi, x);
                }
            } else {
                for (int c = 0; c < 8; ++c) {
                    const int32_t cid = kids[c];
                    if (cid == -1) continue;
                    lo = jmin(lo, t->bmin[3 * (int64_t)cid + d]);
                    hi = jmax(hi, t->bmax[3 * (int64_t)cid + d]);
                }
            }
            t->bmin[3 * nid + d] = lo;
            t->bmax[3 * nid + d] = hi;
        }
    }
}

/* src/api.jl:194-250 (_run_cpu!) for one dirty, non-empty 3D set:
 * origin -> bin -> sort -> build -> refit.  cell_size is the search radius
 * (src/api.jl:217). */
orc_tree* orc_run(const float* coords, int64_t n, float radius, int32_t leaf_size) {
    orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
    t->n = n;
    t->radius = radius;
    t->leaf_size = leaf_size;
    if (n == 0) return t; /* src/api.jl:205-212: empty set keeps n_nodes = 0 */
    t->keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    t->perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    orc_point_origin3(coords, n, t->origin);
    orc_bin3(coords, n, t->origin, radius, t->keys);
    orc_sort_by_key(t->keys, n, t->perm);
    build_tree(t, leaf_size);
    refit_tree(t, coords);
    return t;
}

void orc_tree_free(orc_tree* t) {
    if (!t) return;
    free(t->keys); free(t->perm);
    free(t->node_first); free(t->node_last); free(t->node_level); free(t->node_children);
    free(t->bmin); free(t->bmax);
    free(t);
}

/* ---------------------------------------------------------------- traversal */

/* src/cpu/iterator_cpu.jl:31-40 (_aabb_dist_sq, 3D). */
static inline float aabb_dist_sq(float px, float py, float pz, const float* lo,
                                 const float* hi) {
    float dx = jmax(lo[0] - px, 0.0f); dx = jmax(dx, px - hi[0]);
    float dy = jmax(lo[1] - py, 0.0f); dy = jmax(dy, py - hi[1]);
    float dz = jmax(lo[2] - pz, 0.0f); dz = jmax(dz, pz - hi[2]);
    return dx * dx + dy * dy + dz * dz;
}

/* Stack bound 1 + 7*21 = 148 (src/cpu/build_cpu.jl:25-43). */
#define ORC_STACK_DEPTH 148

typedef void (*orc_visit_fn)(void* ctx, int32_t j, float dx, float dy, float dz, float d2);

/* src/cpu/iterator_cpu.jl:54-120 (_traverse_with_diff!, 3D): LIFO stack, the
 * popped node is re-tested, a leaf loops its points in perm order with
 * dx = target - query, and children are tested before they are pushed. */
static inline void traverse(const orc_tree* t, const float* coords_t, float px, float py,
                            float pz, float r_sq, orc_visit_fn f, void* ctx) {
    if (t->n_nodes == 0) return;
    int32_t stack[ORC_STACK_DEPTH];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int32_t nid = stack[--sp];
        const float dsq = aabb_dist_sq(px, py, pz, t->bmin + 3 * (int64_t)nid,
                                       t->bmax + 3 * (int64_t)nid);
        if (dsq > r_sq) continue;
        const int32_t* kids = t->node_children + 8 * (int64_t)nid;
        if (kids[0] == -1) {
            for (int32_t k = t->node_first[nid]; k <= t->node_last[nid]; ++k) {
                const int32_t j = t->perm[k];
                const float* q = coords_t + 3 * (int64_t)j;
                const float dx = q[0] - px, dy = q[1] - py, dz = q[2] - pz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (d2 <= r_sq) f(ctx, j, dx, dy, dz, d2);
            }
        } else {
            for (int c = 0; c < 8; ++c) {
                const int32_t cid = kids[c];
                if (cid == -1) continue;
                const float cdsq = aabb_dist_sq(px, py, pz, t->bmin + 3 * (int64_t)cid,
                                                t->bmax + 3 * (int64_t)cid);
                if (cdsq <= r_sq) stack[sp++] = cid;
            }
        }
    }
}

typedef struct { int32_t* out; int64_t cap; int64_t n; } list_ctx;
static void visit_list(void* c, int32_t j, float dx, float dy, float dz, float d2) {
    (void)dx; (void)dy; (void)dz; (void)d2;
    list_ctx* l = (list_ctx*)c;
    if (l->n < l->cap) l->out[l->n] = j;
    l->n++;
}

/* src/cpu/iterator_cpu.jl:245-300 (for_each_neighbor, host): every target j
 * within r of query point i, in traversal order, self NOT filtered.  Writes up
 * to cap indices and returns the total count. */
int64_t orc_for_each_neighbor(const orc_tree* t, const float* coords_q, const float* coords_t,
                              int64_t i, int32_t* out, int64_t cap) {
    list_ctx l = {out, cap, 0};
    const float r = t->radius, r_sq = r * r;
    const float* p = coords_q + 3 * i;
    traverse(t, coords_t, p[0], p[1], p[2], r_sq, visit_list, &l);
    return l.n;
}

/* ------------------------------------------------------------- build_edges */

typedef struct { int32_t i; int exclude_self; int32_t c; } count_ctx;
static void visit_count(void* c, int32_t j, float dx, float dy, float dz, float d2) {
    (void)dx; (void)dy; (void)dz; (void)d2;
    count_ctx* k = (count_ctx*)c;
    if (!k->exclude_self || j != k->i) k->c++;
}

typedef struct {
    int32_t i; int exclude_self; int64_t cur; float inv_r;
    int32_t* senders; int32_t* receivers; float* rdisp; float* rdist;
} fill_ctx;
static void visit_fill(void* c, int32_t j, float dx, float dy, float dz, float d2) {
    fill_ctx* k = (fill_ctx*)c;
    if (k->exclude_self && j == k->i) return;
    const int64_t e = k->cur++;
    k->senders[e] = j;
    k->receivers[e] = k->i;
    k->rdisp[3 * e + 0] = -dx * k->inv_r;
    k->rdisp[3 * e + 1] = -dy * k->inv_r;
    k->rdisp[3 * e + 2] = -dz * k->inv_r;
    k->rdist[e] = sqrtf(d2) * k->inv_r;
}

/* src/util.jl:38-50 (chunk_bounds / n_query_chunks): contiguous query chunks
 * of at least 1024, at most one per thread. */
static int n_query_chunks(int64_t n, int nthreads) {
    if (n <= 1024) return 1;
    int64_t c = (n + 1023) / 1024;
    if (c > nthreads) c = nthreads;
    return c < 1 ? 1 : (int)c;
}
static void chunk_bounds(int64_t n, int nchunks, int c, int64_t* lo, int64_t* hi) {
    const int64_t base = n / nchunks, rem = n % nchunks;
    *lo = c * base + (c < rem ? c : rem);
    *hi = *lo + base + (c < rem ? 1 : 0); /* exclusive */
}

/* src/cpu/edges_cpu.jl:172-247 (build_edges_cpu!): pass A counts per query
 * (src/cpu/edges_cpu.jl:36-61), a serial exclusive scan (:209-212), pass B
 * fills senders = j, receivers = i, rdisp = -dx*inv_r, rdist =
 * sqrt(d2)*inv_r (:92-130).  Receivers come out ascending; within a receiver
 * the order is the DFS visit order. */
orc_edges* orc_build_edges(const orc_tree* t, const float* coords_q, int64_t nq,
                           const float* coords_t, int exclude_self, int nthreads) {
    orc_edges* e = (orc_edges*)calloc(1, sizeof(orc_edges));
    e->nq = nq;
    e->offsets = (int64_t*)calloc((size_t)nq + 1, sizeof(int64_t));
    const float r = t->radius;
    const float r_sq = r * r;
    const float inv_r = 1.0f / r;
    if (nthreads < 1) nthreads = 1;
    const int nchunks = n_query_chunks(nq, nthreads);
    int32_t* counts = (int32_t*)calloc((size_t)nq + 1, sizeof(int32_t));
#pragma omp parallel for schedule(static, 1) num_threads(nthreads) if (nchunks > 1)
    for (int c = 0; c < nchunks; ++c) {
        int64_t lo, hi;
        chunk_bounds(nq, nchunks, c, &lo, &hi);
        for (int64_t i = lo; i < hi; ++i) {
            count_ctx k = {(int32_t)i, exclude_self, 0};
            const float* p = coords_q + 3 * i;
            traverse(t, coords_t, p[0], p[1], p[2], r_sq, visit_count, &k);
            counts[i] = k.c;
        }
    }
    for (int64_t i = 0; i < nq; ++i) e->offsets[i + 1] = e->offsets[i] + counts[i];
    free(counts);
    const int64_t ne = e->offsets[nq];
    e->n_edges = ne;
    e->senders = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    e->receivers = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    e->rdisp = (float*)malloc(sizeof(float) * 3 * (size_t)(ne > 0 ? ne : 1));
    e->rdist = (float*)malloc(sizeof(float) * (size_t)(ne > 0 ? ne : 1));
#pragma omp parallel for schedule(static, 1) num_threads(nthreads) if (nchunks > 1)
    for (int c = 0; c < nchunks; ++c) {
        int64_t lo, hi;
        chunk_bounds(nq, nchunks, c, &lo, &hi);
        for (int64_t i = lo; i < hi; ++i) {
            fill_ctx k = {(int32_t)i, exclude_self, e->offsets[i], inv_r,
                          e->senders, e->receivers, e->rdisp, e->rdist};
            const float* p = coords_q + 3 * i;
            traverse(t, coords_t, p[0], p[1], p[2], r_sq, visit_fill, &k);
        }
    }
    return e;
}

void orc_edges_free(orc_edges* e) {
    if (!e) return;
    free(e->offsets); free(e->senders); free(e->receivers); free(e->rdisp); free(e->rdist);
    free(e);
}

/* Pass A of build_edges alone (src/cpu/edges_cpu.jl:36-61,
 * _count_edges_range!): counts[i] for every query i and the total.  For
 * edge sets too large to hold on the host (the 4M-point clustered workload
 * has ~6e8 edges); dynamic chunks of 1024 queries, since per-query work is
 * uneven there (the counts do not depend on the schedule). */
int64_t orc_count_edges(const orc_tree* t, const float* coords_q, int64_t nq, const float* coords_t,
                        int exclude_self, int nthreads, int32_t* counts) {
    const float r = t->radius;
    const float r_sq = r * r;
    if (nthreads < 1) nthreads = 1;
    const int64_t nblk = (nq + 1023) / 1024;
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : total) if (nblk > 1)
    for (int64_t b = 0; b < nblk; ++b) {
        const int64_t hi = (b + 1) * 1024 < nq ? (b + 1) * 1024 : nq;
        for (int64_t i = b * 1024; i < hi; ++i) {
            count_ctx k = {(int32_t)i, exclude_self, 0};
            const float* p = coords_q + 3 * i;
            traverse(t, coords_t, p[0], p[1], p[2], r_sq, visit_count, &k);
            counts[i] = k.c;
            total += k.c;
        }
    }
    return total;
}

/* Pass B of build_edges (src/cpu/edges_cpu.jl:92-130, _fill_edges_range!)
 * restricted to the queries rows[0 .. nrows): their edges in that order,
 * each query's edges in DFS visit order, receivers = the query's own index
 * (so the self pair still excludes j == rows[k]).  A sample of receivers of
 * an edge set too large to fill whole. */
orc_edges* orc_build_edges_rows(const orc_tree* t, const float* coords_q, const int64_t* rows,
                                int64_t nrows, const float* coords_t, int exclude_self, int nthreads) {
    orc_edges* e = (orc_edges*)calloc(1, sizeof(orc_edges));
    e->nq = nrows;
    e->offsets = (int64_t*)calloc((size_t)nrows + 1, sizeof(int64_t));
    const float r = t->radius;
    const float r_sq = r * r;
    const float inv_r = 1.0f / r;
    if (nthreads < 1) nthreads = 1;
    int32_t* counts = (int32_t*)calloc((size_t)nrows + 1, sizeof(int32_t));
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t k = 0; k < nrows; ++k) {
        count_ctx c = {(int32_t)rows[k], exclude_self, 0};
        const float* p = coords_q + 3 * rows[k];
        traverse(t, coords_t, p[0], p[1], p[2], r_sq, visit_count, &c);
        counts[k] = c.c;
    }
    for (int64_t k = 0; k < nrows; ++k) e->offsets[k + 1] = e->offsets[k] + counts[k];
    free(counts);
    const int64_t ne = e->offsets[nrows];
    e->n_edges = ne;
    e->senders = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    e->receivers = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    e->rdisp = (float*)malloc(sizeof(float) * 3 * (size_t)(ne > 0 ? ne : 1));
    e->rdist = (float*)malloc(sizeof(float) * (size_t)(ne > 0 ? ne : 1));
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t k = 0; k < nrows; ++k) {
        fill_ctx f = {(int32_t)rows[k], exclude_self, e->offsets[k], inv_r,
                      e->senders, e->receivers, e->rdisp, e->rdist};
        const float* p = coords_q + 3 * rows[k];
        traverse(t, coords_t, p[0], p[1], p[2], r_sq, visit_fill, &f);
    }
    return e;
}

/* --------------------------------------------------- build_edges_diff rrule */

/* ext/OctopusChainRulesCoreExt.jl:137-192 (CPU pullback): a serial loop over
 * edges in storage order, FP32 accumulation.  For edge k with receiver i and
 * sender j: disp part val = d_disp[d,k] / r (a division, as the reference),
 * grad[d,i] += val, grad[d,j] -= val; dist part when rdist > 1e-8:
 * inv_dnr = d_dist[k] / (rdist * r), g = rdisp[d,k] * inv_dnr, grad[d,i] += g,
 * grad[d,j] -= g.  d_disp or d_dist may be NULL (a zero tangent). */
void orc_pullback(int64_t n_points, int64_t n_edges, const int32_t* senders,
                  const int32_t* receivers, const float* rdisp, const float* rdist,
                  const float* d_disp, const float* d_dist, float r, float* grad) {
    memset(grad, 0, sizeof(float) * 3 * (size_t)n_points);
    if (!d_disp && !d_dist) return;
    for (int64_t k = 0; k < n_edges; ++k) {
        const int64_t i = receivers[k], j = senders[k];
        if (d_disp) {
            for (int d = 0; d < 3; ++d) {
                const float val = d_disp[3 * k + d] / r;
                grad[3 * i + d] += val;
                grad[3 * j + d] -= val;
            }
        }
        if (d_dist) {
            const float dn = rdist[k];
            if (dn > 1e-8f) {
                const float inv_dnr = d_dist[k] / (dn * r);
                for (int d = 0; d < 3; ++d) {
                    const float g = rdisp[3 * k + d] * inv_dnr;
                    grad[3 * i + d] += g;
                    grad[3 * j + d] -= g;
                }
            }
        }
    }
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
