/*
 * CPU oracle of the 2D path (quadtrees) of run! + build_edges and the
 * build_edges_diff pullback.  TEST INFRASTRUCTURE ONLY: the checker of the
 * engine's 2D point sets, never part of the product.
 *
 * Restates the reference's 2D branches: 31-bit Morton keys
 * (src/morton.jl:50-80), the 2D origin and binning (src/cpu/bin_cpu.jl:25-35,
 * 75-95), the same stable sort as 3D (orc_sort_by_key), the BFS build with
 * four children and 31 levels (src/cpu/build_cpu.jl:20-52,100-194 with
 * NDIMS = 2), the bounds refit (src/cpu/refit_cpu.jl:21-70), the 2D
 * traversal (src/cpu/iterator_cpu.jl:122-185) and the 2D fill
 * (src/cpu/edges_cpu.jl:132-170).  As in 3D the edge set is a closed form:
 * {(i, j) : fl(fl(dx*dx) + fl(dy*dy)) <= fl(r*r)} (self pair: j != i).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "octopus_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* Julia's Base.min / Base.max on floats (-0.0 below +0.0). */
static inline float jmin2(float x, float y) { return (y < x || (y == x && signbit(y))) ? y : x; }
static inline float jmax2(float x, float y) { return (y > x || (y == x && !signbit(y))) ? y : x; }

/* src/morton.jl:50-59 (_split_by_2): the low 31 bits to every second bit. */
uint64_t orc_split_by_2(uint64_t x) {
    x &= 0x000000007fffffffULL;
    x = (x | (x << 16)) & 0x0000ffff0000ffffULL;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffULL;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

/* src/morton.jl:61-69 (_compact_by_2). */
uint64_t orc_compact_by_2(uint64_t x) {
    x &= 0x5555555555555555ULL;
    x = (x | (x >> 1)) & 0x3333333333333333ULL;
    x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0fULL;
    x = (x | (x >> 4)) & 0x00ff00ff00ff00ffULL;
    x = (x | (x >> 8)) & 0x0000ffff0000ffffULL;
    x = (x | (x >> 16)) & 0x000000007fffffffULL;
    return x;
}

/* src/morton.jl:71-75 (morton_encode2): x in bit 0, y in bit 1. */
uint64_t orc_morton_encode2(uint32_t ix, uint32_t iy) {
    return (orc_split_by_2(iy) << 1) | orc_split_by_2(ix);
}

void orc_morton_decode2(uint64_t code, uint32_t out[2]) {
    out[0] = (uint32_t)orc_compact_by_2(code);
    out[1] = (uint32_t)orc_compact_by_2(code >> 1);
}

/* src/morton.jl:91-95 (bin_point, 2D): floor only. */
void orc_bin_point2(float px, float py, float ox, float oy, float inv_cs, int32_t out[2]) {
    out[0] = (int32_t)floorf((px - ox) * inv_cs);
    out[1] = (int32_t)floorf((py - oy) * inv_cs);
}

static int32_t floor_clamp31(float v) {
    const float f = floorf(v);
    if (!(f >= 0.0f)) return 0;
    if (f >= 2147483648.0f) return 0x7fffffff; /* clamp(ix, 0, 0x7fffffff) */
    return (int32_t)f;
}

/* src/cpu/bin_cpu.jl:25-35 (_point_origin_2d). */
void orc_point_origin2(const float* coords, int64_t n, float origin[2]) {
    float ox = coords[0], oy = coords[1];
    for (int64_t i = 1; i < n; ++i) {
        ox = jmin2(ox, coords[2 * i]);
        oy = jmin2(oy, coords[2 * i + 1]);
    }
    origin[0] = ox - 1e-6f;
    origin[1] = oy - 1e-6f;
}

/* src/cpu/bin_cpu.jl:75-95 (bin_cpu!, 2D): 31-bit clamp. */
void orc_bin2(const float* coords, int64_t n, const float origin[2], float cell_size, uint64_t* keys) {
    const float inv_cs = 1.0f / cell_size;
    for (int64_t i = 0; i < n; ++i) {
        const int32_t ix = floor_clamp31((coords[2 * i] - origin[0]) * inv_cs);
        const int32_t iy = floor_clamp31((coords[2 * i + 1] - origin[1]) * inv_cs);
        keys[i] = orc_morton_encode2((uint32_t)ix, (uint32_t)iy);
    }
}

/* src/cpu/build_cpu.jl:46-49 with NDIMS = 2: 2 * 31 - 2 * (L + 1). */
static int level_shift2(int L) { return 2 * 31 - 2 * (L + 1); }

/* BFS quadtree (src/cpu/build_cpu.jl:100-194, NDIMS = 2): a node is a leaf
 * iff count <= leaf_size or level >= 31; children in quadrant order. */
static void build_tree2(orc_tree2* t, int32_t leaf_size) {
    const int64_t n = t->n;
    int64_t cap = 4 * ((n + leaf_size - 1) / (leaf_size > 1 ? leaf_size : 1)) + 64;
    int32_t* first = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    int32_t* last = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    int32_t* level = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    int32_t* kids = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)cap);
    first[0] = 0;
    last[0] = (int32_t)(n - 1);
    level[0] = 0;
    int64_t n_nodes = 1, next_open = 0;
    while (next_open < n_nodes) {
        const int64_t frontier_end = n_nodes;
        for (int64_t nid = next_open; nid < frontier_end; ++nid) {
            const int32_t f = first[nid], l = last[nid], lvl = level[nid];
            for (int c = 0; c < 4; ++c) kids[4 * nid + c] = -1;
            if (l - f + 1 <= leaf_size || lvl >= 31) continue;
            const int shift = level_shift2(lvl);
            int32_t qf[4] = {-1, -1, -1, -1}, ql[4] = {-2, -2, -2, -2};
            int prev = -1;
            for (int32_t i = f; i <= l; ++i) { /* _partition_octants! */
                const int q = (int)((t->keys[t->perm[i]] >> shift) & 3ULL);
                if (q != prev) {
                    if (prev != -1) ql[prev] = i - 1;
                    qf[q] = i;
                    prev = q;
                }
            }
            if (prev != -1) ql[prev] = l;
            for (int c = 0; c < 4; ++c) {
                if (qf[c] == -1) continue;
                if (n_nodes + 1 > cap) {
                    cap = cap * 3 / 2 + 16;
                    first = (int32_t*)realloc(first, sizeof(int32_t) * (size_t)cap);
                    last = (int32_t*)realloc(last, sizeof(int32_t) * (size_t)cap);
                    level = (int32_t*)realloc(level, sizeof(int32_t) * (size_t)cap);
                    kids = (int32_t*)realloc(kids, sizeof(int32_t) * 4 * (size_t)cap);
                }
                first[n_nodes] = qf[c];
                last[n_nodes] = ql[c];
                level[n_nodes] = lvl + 1;
                kids[4 * nid + c] = (int32_t)n_nodes;
                ++n_nodes;
            }
        }
        next_open = frontier_end;
    }
    t->n_nodes = (int32_t)n_nodes;
    t->node_first = first;
    t->node_last = last;
    t->node_level = level;
    t->node_children = kids;
    t->bmin = (float*)malloc(sizeof(float) * 2 * (size_t)n_nodes);
    t->bmax = (float*)malloc(sizeof(float) * 2 * (size_t)n_nodes);
}

/* src/cpu/refit_cpu.jl:21-70 in 2D: reverse node order. */
static void refit_tree2(orc_tree2* t, const float* coords) {
    for (int64_t nid = (int64_t)t->n_nodes - 1; nid >= 0; --nid) {
        const int32_t* kids = t->node_children + 4 * nid;
        for (int d = 0; d < 2; ++d) {
            float lo = INFINITY, hi = -INFINITY;
            if (kids[0] == -1) {
                for (int32_t k = t->node_first[nid]; k <= t->node_last[nid]; ++k) {
                    const float x = coords[2 * (int64_t)t->perm[k] + d];
                    lo = jmin2(lo, x);
                    hi = jmax2(hi, x);
                }
            } else {
                for (int c = 0; c < 4; ++c) {
                    const int32_t cid = kids[c];
                    if (cid == -1) continue;
                    lo = jmin2(lo, t->bmin[2 * (int64_t)cid + d]);
                    hi = jmax2(hi, t->bmax[2 * (int64_t)cid + d]);
                }
            }
            t->bmin[2 * nid + d] = lo;
            t->bmax[2 * nid + d] = hi;
        }
    }
}

orc_tree2* orc_run2(const float* coords, int64_t n, float radius, int32_t leaf_size) {
    orc_tree2* t = (orc_tree2*)calloc(1, sizeof(orc_tree2));
    t->n = n;
    t->radius = radius;
    t->leaf_size = leaf_size;
    if (n == 0) return t;
    t->keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    t->perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    orc_point_origin2(coords, n, t->origin);
    orc_bin2(coords, n, t->origin, radius, t->keys);
    orc_sort_by_key(t->keys, n, t->perm);
    build_tree2(t, leaf_size);
    refit_tree2(t, coords);
    return t;
}

void orc_tree2_free(orc_tree2* t) {
    if (!t) return;
    free(t->keys); free(t->perm);
    free(t->node_first); free(t->node_last); free(t->node_level); free(t->node_children);
    free(t->bmin); free(t->bmax);
    free(t);
}

/* src/cpu/iterator_cpu.jl:43-51 (_aabb_dist_sq, 2D). */
static inline float aabb_dist_sq2(float px, float py, const float* lo, const float* hi) {
    float dx = jmax2(lo[0] - px, 0.0f); dx = jmax2(dx, px - hi[0]);
    float dy = jmax2(lo[1] - py, 0.0f); dy = jmax2(dy, py - hi[1]);
    return dx * dx + dy * dy;
}

#define ORC_STACK_DEPTH2 94 /* 1 + 3 * 31 (src/cpu/build_cpu.jl:31,43) */

/* src/cpu/iterator_cpu.jl:122-185 (_traverse_with_diff!, 2D): dx = target -
 * query; visit(j, dx, dy, d2) in DFS order; stop = 1 to count only. */
typedef struct { int32_t i; int exclude_self; int64_t cur; int64_t count; float inv_r;
                 int32_t* senders; int32_t* receivers; float* rdisp; float* rdist; } visit2_ctx;

static void traverse2(const orc_tree2* t, const float* coords_t, float px, float py, float r_sq, visit2_ctx* v,
                      int fill) {
    if (t->n_nodes == 0) return;
    int32_t stack[ORC_STACK_DEPTH2];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int32_t nid = stack[--sp];
        if (aabb_dist_sq2(px, py, t->bmin + 2 * (int64_t)nid, t->bmax + 2 * (int64_t)nid) > r_sq) continue;
        const int32_t* kids = t->node_children + 4 * (int64_t)nid;
        if (kids[0] == -1) {
            for (int32_t k = t->node_first[nid]; k <= t->node_last[nid]; ++k) {
                const int32_t j = t->perm[k];
                const float dx = coords_t[2 * (int64_t)j] - px, dy = coords_t[2 * (int64_t)j + 1] - py;
                const float d2 = dx * dx + dy * dy;
                if (d2 > r_sq || (v->exclude_self && j == v->i)) continue;
                if (fill) { /* src/cpu/edges_cpu.jl:150-156 */
                    const int64_t e = v->cur++;
                    v->senders[e] = j;
                    v->receivers[e] = v->i;
                    v->rdisp[2 * e + 0] = -dx * v->inv_r;
                    v->rdisp[2 * e + 1] = -dy * v->inv_r;
                    v->rdist[e] = sqrtf(d2) * v->inv_r;
                } else {
                    v->count++;
                }
            }
        } else {
            for (int c = 0; c < 4; ++c) {
                const int32_t cid = kids[c];
                if (cid == -1) continue;
                if (aabb_dist_sq2(px, py, t->bmin + 2 * (int64_t)cid, t->bmax + 2 * (int64_t)cid) <= r_sq)
                    stack[sp++] = cid;
            }
        }
    }
}

/* src/cpu/edges_cpu.jl:172-247 in 2D: count, serial scan, fill (rdisp 2 per
 * edge); receivers ascending, DFS order within a receiver. */
orc_edges* orc_build_edges2(const orc_tree2* t, const float* coords_q, int64_t nq, const float* coords_t,
                            int exclude_self, int nthreads) {
    orc_edges* e = (orc_edges*)calloc(1, sizeof(orc_edges));
    e->nq = nq;
    e->offsets = (int64_t*)calloc((size_t)nq + 1, sizeof(int64_t));
    const float r = t->radius, r_sq = r * r, inv_r = 1.0f / r;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
    for (int64_t i = 0; i < nq; ++i) {
        visit2_ctx v = {(int32_t)i, exclude_self, 0, 0, inv_r, 0, 0, 0, 0};
        traverse2(t, coords_t, coords_q[2 * i], coords_q[2 * i + 1], r_sq, &v, 0);
        e->offsets[i + 1] = v.count;
    }
    for (int64_t i = 0; i < nq; ++i) e->offsets[i + 1] += e->offsets[i];
    const int64_t ne = e->offsets[nq];
    e->n_edges = ne;
    e->senders = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    e->receivers = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    e->rdisp = (float*)malloc(sizeof(float) * 2 * (size_t)(ne > 0 ? ne : 1));
    e->rdist = (float*)malloc(sizeof(float) * (size_t)(ne > 0 ? ne : 1));
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
    for (int64_t i = 0; i < nq; ++i) {
        visit2_ctx v = {(int32_t)i, exclude_self, e->offsets[i], 0, inv_r, e->senders, e->receivers, e->rdisp,
                        e->rdist};
        traverse2(t, coords_t, coords_q[2 * i], coords_q[2 * i + 1], r_sq, &v, 1);
    }
    return e;
}

/* ext/OctopusChainRulesCoreExt.jl:137-192 with D = 2 (the reference's loop
 * runs over d in 1:D). */
void orc_pullback2(int64_t n_points, int64_t n_edges, const int32_t* senders, const int32_t* receivers,
                   const float* rdisp, const float* rdist, const float* d_disp, const float* d_dist, float r,
                   float* grad) {
    memset(grad, 0, sizeof(float) * 2 * (size_t)n_points);
    for (int64_t k = 0; k < n_edges; ++k) {
        const int64_t i = receivers[k], j = senders[k];
        if (d_disp) {
            for (int d = 0; d < 2; ++d) {
                const float val = d_disp[2 * k + d] / r;
                grad[2 * i + d] += val;
                grad[2 * j + d] -= val;
            }
        }
        if (d_dist) {
            const float dn = rdist[k];
            if (dn > 1e-8f) {
                const float inv_dnr = d_dist[k] / (dn * r);
                for (int d = 0; d < 2; ++d) {
                    const float g = rdisp[2 * k + d] * inv_dnr;
                    grad[2 * i + d] += g;
                    grad[2 * j + d] -= g;
                }
            }
        }
    }
}
