/*
 * CPU oracle for the run! + build_edges path (3D, FP32).  TEST
 * INFRASTRUCTURE ONLY: the checker and the timed CPU baseline, never part of
 * the product.
 *
 * Why the edge set is a closed form: the traversal prunes a node when the
 * FP32 point-to-box distance exceeds r*r (src/cpu/iterator_cpu.jl:31-40), and
 * that distance is built from the same monotone IEEE operations as the point
 * test (src/cpu/iterator_cpu.jl:96-98).  Rounding is monotone, so for a point
 * inside a box fl(box_d2) <= fl(point_d2); pruning never drops an edge, and
 * the emitted set is exactly
 *   {(i, j) : j != i for the self pair,
 *             fl(fl(fl(dx*dx) + fl(dy*dy)) + fl(dz*dz)) <= fl(r*r)}
 * whatever the tree looks like.
 */
#ifndef OCTOPUS_ORACLE_H
#define OCTOPUS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tree {
    int64_t n;            /* points */
    float radius;         /* search radius == cell size */
    int32_t leaf_size;    /* target_leaf_size (32 in the reference) */
    int32_t n_nodes;
    float origin[3];
    uint64_t* keys;       /* Morton key per point (input order) */
    int32_t* perm;        /* sorted position -> point index */
    int32_t* node_first;  /* inclusive first sorted position */
    int32_t* node_last;   /* inclusive last sorted position */
    int32_t* node_level;  /* BFS level of each node */
    int32_t* node_children; /* 8 per node, octant order, -1 = empty; [0] == -1 marks a leaf */
    float* bmin;          /* 3 per node */
    float* bmax;          /* 3 per node */
} orc_tree;

typedef struct orc_edges {
    int64_t nq;
    int64_t n_edges;
    int64_t* offsets;     /* nq + 1 exclusive prefix of per-query counts */
    int32_t* senders;     /* target index j */
    int32_t* receivers;   /* query index i */
    float* rdisp;         /* 3 per edge: (q_i - t_j) / r */
    float* rdist;         /* 1 per edge: |q_i - t_j| / r */
} orc_edges;

uint64_t orc_split_by_3(uint64_t x);
uint64_t orc_compact_by_3(uint64_t x);
uint64_t orc_morton_encode3(uint32_t ix, uint32_t iy, uint32_t iz);
void orc_morton_decode3(uint64_t code, uint32_t out[3]);
void orc_bin_point3(float px, float py, float pz, float ox, float oy, float oz, float inv_cs,
                    int32_t out[3]);
void orc_point_origin3(const float* coords, int64_t n, float origin[3]);
void orc_bin3(const float* coords, int64_t n, const float origin[3], float cell_size,
              uint64_t* keys);
void orc_sort_by_key(const uint64_t* keys, int64_t n, int32_t* perm);

orc_tree* orc_run(const float* coords, int64_t n, float radius, int32_t leaf_size);
void orc_tree_free(orc_tree* t);

int64_t orc_for_each_neighbor(const orc_tree* t, const float* coords_q, const float* coords_t,
                              int64_t i, int32_t* out, int64_t cap);
orc_edges* orc_build_edges(const orc_tree* t, const float* coords_q, int64_t nq,
                           const float* coords_t, int exclude_self, int nthreads);
void orc_edges_free(orc_edges* e);
int64_t orc_count_edges(const orc_tree* t, const float* coords_q, int64_t nq, const float* coords_t,
                        int exclude_self, int nthreads, int32_t* counts);
orc_edges* orc_build_edges_rows(const orc_tree* t, const float* coords_q, const int64_t* rows,
                                int64_t nrows, const float* coords_t, int exclude_self, int nthreads);

void orc_pullback(int64_t n_points, int64_t n_edges, const int32_t* senders,
                  const int32_t* receivers, const float* rdisp, const float* rdist,
                  const float* d_disp, const float* d_dist, float r, float* grad);

int orc_max_threads(void);

/* ---- 2D (quadtrees): octopus_oracle2d.c -------------------------------- */
typedef struct orc_tree2 {
    int64_t n;
    float radius;
    int32_t leaf_size;
    int32_t n_nodes;
    float origin[2];
    uint64_t* keys;       /* 62-bit 2D Morton key per point (input order) */
    int32_t* perm;
    int32_t* node_first;
    int32_t* node_last;
    int32_t* node_level;
    int32_t* node_children; /* 4 per node, quadrant order, -1 = empty */
    float* bmin;          /* 2 per node */
    float* bmax;
} orc_tree2;

uint64_t orc_split_by_2(uint64_t x);
uint64_t orc_compact_by_2(uint64_t x);
uint64_t orc_morton_encode2(uint32_t ix, uint32_t iy);
void orc_morton_decode2(uint64_t code, uint32_t out[2]);
void orc_bin_point2(float px, float py, float ox, float oy, float inv_cs, int32_t out[2]);
void orc_point_origin2(const float* coords, int64_t n, float origin[2]);
void orc_bin2(const float* coords, int64_t n, const float origin[2], float cell_size, uint64_t* keys);
orc_tree2* orc_run2(const float* coords, int64_t n, float radius, int32_t leaf_size);
void orc_tree2_free(orc_tree2* t);
/* edges with rdisp 2 per edge */
orc_edges* orc_build_edges2(const orc_tree2* t, const float* coords_q, int64_t nq, const float* coords_t,
                            int exclude_self, int nthreads);
void orc_pullback2(int64_t n_points, int64_t n_edges, const int32_t* senders, const int32_t* receivers,
                   const float* rdisp, const float* rdist, const float* d_disp, const float* d_dist, float r,
                   float* grad);

#ifdef __cplusplus
}
#endif
#endif
