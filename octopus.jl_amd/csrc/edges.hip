// build_edges on the GPU: count, scan and emit of flat-COO edges
// (reference: src/cpu/edges_cpu.jl:172-247, ext/OctopusCUDAExt.jl:262-415).
//
// Work is blocked by query leaf.  Stage 1 gives every query leaf the list of
// target leaves whose box lies within r of its box (FP32 box-box gap, a
// monotone lower bound of every point distance between the two boxes, so no
// edge is lost).  Stage 2 runs one wavefront per query leaf: lanes take
// candidate target points (coalesced from the Morton-sorted copy), the leaf's
// queries sit one per lane and are broadcast with readlane, and each
// (query, 64 candidates) step is one ballot.  The predicate is the
// reference's FP32 test
//   d2 = (dx*dx + dy*dy) + dz*dz <= r*r,
// so the edge set is bit-identical to the CPU path.  The emit pass stages a
// leaf's edges in LDS and writes them out with whole-wave contiguous stores.
#include <type_traits>

#include "internal.hpp"

namespace oct {

// ---------------------------------------------------------------- stage 1

static_assert(1 + 7 * OCT_MAX_LEVEL <= 160, "LIFO walk bound (src/cpu/build_cpu.jl:25-43) must fit the stack");
constexpr int LL_SAFE_STACK = 160;  // entries of the safe walk (scratch)
constexpr int LL_FAST_STACK = 32;   // stack entries per query leaf of the cooperative walk (LDS, 4 bytes each)
constexpr int LL_THREADS = 256;
constexpr int EW_WAVES = 4;     // waves per block, each owns one query leaf at a time
constexpr int EW_TILE = 4;      // candidate chunks of 64 held in registers per tile
// A query leaf the single pass records in one sweep has at most 64 queries,
// 64 candidate leaves and REC_MAX_CAND candidate points, of which at most
// one register tile (256) passes the point-box filter.
constexpr int REC_MAX_CAND = 1024;
// Heavy leaf: one that will not be recorded in one sweep (more than 64
// queries or candidate leaves, an estimated filtered candidate count above
// one tile), whose estimated query x candidate pairs exceed HEAVY_PAIRS
// (about four hit lists of hits at the usual hit rate of a quarter), or whose
// candidate leaves hold more than HEAVY_RAW points (a small leaf beside a
// dense region: the single pass, one query per lane, would load and filter
// them all for a few queries, where the heavy-leaf kernels spread the
// candidates over the lanes).  The single pass takes their counts from a
// parallel count pass run before it, so no run's aggregate waits on them.
constexpr float HEAVY_PAIRS = 4096.f;
constexpr int HEAVY_RAW = 512;

// Direction of box c seen from box q, one of 27 codes sx + 3 sy + 9 sz with
// s = 0 (below), 1 (level), 2 (above) per axis: the centres must differ by
// more than a quarter of the two extents summed to count as off-level.  Code
// 13 is the same position.  Feeds the per-leaf count predictor's
// adjacent-leaf features.
__device__ inline uint8_t dir_code(const float4& qlo, const float4& qhi, const float4& clo, const float4& chi) {
    const float d[3] = {0.5f * (clo.x + chi.x) - 0.5f * (qlo.x + qhi.x), 0.5f * (clo.y + chi.y) - 0.5f * (qlo.y + qhi.y),
                        0.5f * (clo.z + chi.z) - 0.5f * (qlo.z + qhi.z)};
    const float tau[3] = {0.25f * ((qhi.x - qlo.x) + (chi.x - clo.x)), 0.25f * ((qhi.y - qlo.y) + (chi.y - clo.y)),
                          0.25f * ((qhi.z - qlo.z) + (chi.z - clo.z))};
    int code = 0, mul = 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int s = d[k] > tau[k] ? 2 : (d[k] < -tau[k] ? 0 : 1);
        code += s * mul;
        mul *= 3;
    }
    return (uint8_t)code;
}

// Appends a passing child: a leaf goes to the list (and its direction code to
// dout when asked), an internal node's child range to the stack.  Returns
// false when the stack has no room.
// Fraction of [lo, hi] inside [a, b] (1 for a degenerate extent).
__device__ inline float span_frac(float lo, float hi, float a, float b) {
    const float ext = hi - lo;
    const float ov = fminf(hi, b) - fmaxf(lo, a);
    return ext > 0.f ? fminf(fmaxf(__fdividef(ov, ext), 0.f), 1.f) : 1.f;
}

// Estimated candidates of a leaf that pass the point-box filter: the part of
// its box inside the query box grown by r, times its point count.
__device__ inline float filtered_estimate(const float4& qlo, const float4& qhi, const NodeRec& ch, float r) {
    return (float)rec_b(ch) * span_frac(ch.lo.x, ch.hi.x, qlo.x - r, qhi.x + r) *
           span_frac(ch.lo.y, ch.hi.y, qlo.y - r, qhi.y + r) * span_frac(ch.lo.z, ch.hi.z, qlo.z - r, qhi.z + r);
}

template <bool DIRS, bool HEAVY, typename Push>
__device__ inline bool ll_visit(const float4& qlo, const float4& qhi, const NodeRec& ch, float r_sq, float r,
                                int2* out, uint8_t* dout, int cap, int& cnt, float& est, int& raw, Push push) {
    if (box_box_d2(qlo, qhi, ch.lo, ch.hi) > r_sq) return true;
    if (rec_b(ch) > 0) {
        if (HEAVY) {
            est += filtered_estimate(qlo, qhi, ch, r);
            raw += rec_b(ch);
        }
        if (cnt < cap) {
            out[cnt] = make_int2(rec_a(ch), rec_b(ch));
            if (DIRS) dout[cnt] = dir_code(qlo, qhi, ch.lo, ch.hi);
        }
        ++cnt;
        return true;
    }
    return push(rec_a(ch), -rec_b(ch));
}

// Safe walk: one thread per query leaf, LIFO depth-first walk of the target
// tree from its branch root.  A stack entry is the child range of a node that
// passed the test; the stack holds 1 + 7 * 21 entries, the bound of a LIFO
// walk over a 21-level tree (src/cpu/build_cpu.jl:25-43), so it never runs
// out.  The host runs it when the cooperative walk below (the usual one)
// outgrew its LDS stack.  A list longer than cap is counted in full and
// flagged through flags[0] (the host regrows cap to that count and reruns).
template <bool DIRS, bool HEAVY>
__global__ void __launch_bounds__(LL_THREADS)
k_leaf_lists(const NodeRec* __restrict__ qleaves,
             const BuildParams* __restrict__ qprm, const NodeRec* __restrict__ tnodes,
             const BuildParams* __restrict__ tprm, float r_sq0, int cap, int2* __restrict__ lists,
             uint8_t* __restrict__ ldir, int32_t* __restrict__ lcount, uint32_t* __restrict__ flags,
             SetLookup sets, uint8_t* __restrict__ lheavy, int32_t* __restrict__ heavy_list,
             uint32_t* __restrict__ n_heavy) {
    const int32_t nleaves = qprm->n_leaves;
    const int32_t root0 = tprm->branch_root;
    const XcdShare sh = xcd_share(nleaves);
    const int tid = threadIdx.x;
    for (int32_t r = (int32_t)(sh.lo + (int64_t)sh.k * LL_THREADS) + tid; r < (int32_t)sh.hi;
         r += (int32_t)(sh.nb * LL_THREADS)) {
        const NodeRec qa = qleaves[r];
        int32_t root = root0;
        float r_sq = r_sq0;
        if (sets.batched) {  // a forest: the leaf's own set's tree and radius
            const BuildParams& sp = sets.sprm[key_set(sets.qsk[rec_a(qa)], 1)];
            root = sp.branch_root;
            r_sq = sp.radius * sp.radius;
        }
        int2* out = lists + (int64_t)r * cap;
        uint8_t* dout = DIRS ? ldir + (int64_t)r * cap : nullptr;
        int cnt = 0;
        float est = 0.f;  // HEAVY: estimated filtered candidate points
        int raw = 0;      // HEAVY: points of the candidate leaves
        const float rad = HEAVY ? sqrtf(r_sq) : 0.f;
        bool full = false;
        const NodeRec rr = tnodes[root];
        if (box_box_d2(qa.lo, qa.hi, rr.lo, rr.hi) <= r_sq) {
            if (rec_b(rr) > 0) {
                if (HEAVY) {
                    est += filtered_estimate(qa.lo, qa.hi, rr, rad);
                    raw += rec_b(rr);
                }
                if (cnt < cap) {
                    out[cnt] = make_int2(rec_a(rr), rec_b(rr));
                    if (DIRS) dout[cnt] = dir_code(qa.lo, qa.hi, rr.lo, rr.hi);
                }
                ++cnt;
            } else {
                int2 stack[LL_SAFE_STACK];
                int sp = 0;
                auto push = [&](int32_t first, int32_t nch) {
                    if (sp >= LL_SAFE_STACK) return false;
                    stack[sp++] = make_int2(first, nch);
                    return true;
                };
                push(rec_a(rr), -rec_b(rr));
                while (sp > 0 && !full) {
                    const int2 ent = stack[--sp];
                    NodeRec ch[8];
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        if (c < ent.y) ch[c] = tnodes[ent.x + c];
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        if (c < ent.y && !ll_visit<DIRS, HEAVY>(qa.lo, qa.hi, ch[c], r_sq, rad, out, dout, cap, cnt, est, raw, push)) full = true;
                }
            }
        }
        lcount[r] = cnt;
        if (cnt > cap) atomicMax(&flags[0], (uint32_t)cnt);  // the host regrows cap to this
        if (full) atomicOr(&flags[1], 1u);                   // the host reruns in safe mode
        if (HEAVY) {  // heavy leaves: flagged, and listed compactly (one atomic per wave)
            const int qc = rec_b(qa);
            const bool heavy = qc > OCT_WAVE || cnt > OCT_WAVE || est > (float)(EW_TILE * OCT_WAVE) ||
                               (float)qc * est > HEAVY_PAIRS || raw > HEAVY_RAW;
            lheavy[r] = heavy ? 1 : 0;
            const uint64_t hm = ballot(heavy);
            if (hm) {
                const int lead = __ffsll((long long)hm) - 1;
                uint32_t base = 0;
                if (lane_id() == lead) base = atomicAdd(n_heavy, (uint32_t)__popcll(hm));
                base = (uint32_t)__shfl((int)base, lead, OCT_WAVE);
                const uint32_t idx = base + (uint32_t)mbcnt(hm);
                if (heavy && idx < (uint32_t)nleaves) heavy_list[idx] = r;
            }
        }
    }
}

// Cooperative form of the fast walk: eight lanes per query leaf, lane c
// testing child c of the popped node.  The one-thread walk loaded a node's
// children with sixteen 16-byte loads per lane, every lane at its own node,
// so each wave-wide load touched 64 cache lines and the L1 request rate
// bounded the kernel; here a load instruction covers eight nodes' contiguous
// children.  The visit order is the one-thread walk's: the passing leaves of
// a node are appended in child order (ballot ranks) and its passing internal
// children pushed in child order, so the last one is popped first.  The
// stack (LL_FAST_STACK entries per leaf, in LDS) overflowing flags the leaf
// for the safe walk, as before.
constexpr int LL_COOP = 8;                              // lanes per query leaf
constexpr int LL_COOP_GROUPS = LL_THREADS / LL_COOP;    // query leaves per block

template <bool DIRS, bool HEAVY>
__global__ void __launch_bounds__(LL_THREADS)
k_leaf_lists_coop(const NodeRec* __restrict__ qleaves, const BuildParams* __restrict__ qprm,
                  const NodeRec* __restrict__ tnodes, const BuildParams* __restrict__ tprm, float r_sq0, int cap,
                  int2* __restrict__ lists, uint8_t* __restrict__ ldir, int32_t* __restrict__ lcount,
                  uint32_t* __restrict__ flags, SetLookup sets, uint8_t* __restrict__ lheavy,
                  int32_t* __restrict__ heavy_list, uint32_t* __restrict__ n_heavy) {
    __shared__ uint32_t stk[LL_COOP_GROUPS][LL_FAST_STACK];
    const int32_t nleaves = qprm->n_leaves;
    const int32_t root0 = tprm->branch_root;
    const XcdShare sh = xcd_share(nleaves);
    const int lane = lane_id();
    const int c = lane & (LL_COOP - 1);        // this lane's child slot
    const int gsh = lane & ~(LL_COOP - 1);     // the group's first lane in the wave
    const int g = threadIdx.x / LL_COOP;       // the group in the block
    const uint32_t below = (1u << c) - 1u;
    for (int32_t r = (int32_t)(sh.lo + (int64_t)sh.k * LL_COOP_GROUPS) + g; r < (int32_t)sh.hi;
         r += (int32_t)(sh.nb * LL_COOP_GROUPS)) {  // uniform in the group
        const NodeRec qa = qleaves[r];
        int32_t root = root0;
        float r_sq = r_sq0;
        if (sets.batched) {  // a forest: the leaf's own set's tree and radius
            const BuildParams& sp = sets.sprm[key_set(sets.qsk[rec_a(qa)], 1)];
            root = sp.branch_root;
            r_sq = sp.radius * sp.radius;
        }
        int2* out = lists + (int64_t)r * cap;
        uint8_t* dout = DIRS ? ldir + (int64_t)r * cap : nullptr;
        int cnt = 0;
        float est = 0.f;  // HEAVY: estimated filtered candidate points
        int raw = 0;      // HEAVY: points of the candidate leaves
        const float rad = HEAVY ? sqrtf(r_sq) : 0.f;
        bool full = false;
        const NodeRec rr = tnodes[root];
        if (box_box_d2(qa.lo, qa.hi, rr.lo, rr.hi) <= r_sq) {
            if (rec_b(rr) > 0) {
                if (HEAVY) {
                    est += filtered_estimate(qa.lo, qa.hi, rr, rad);
                    raw += rec_b(rr);
                }
                if (c == 0 && cap > 0) {
                    out[0] = make_int2(rec_a(rr), rec_b(rr));
                    if (DIRS) dout[0] = dir_code(qa.lo, qa.hi, rr.lo, rr.hi);
                }
                cnt = 1;
            } else {
                if (c == 0) stk[g][0] = ((uint32_t)rec_a(rr) << 3) | (uint32_t)(-rec_b(rr) - 1);
                int sp = 1;
                while (sp > 0 && !full) {  // uniform in the group
                    const uint32_t ent = stk[g][--sp];
                    const int32_t first = (int32_t)(ent >> 3);
                    const int32_t nch = (int32_t)(ent & 7u) + 1;
                    NodeRec ch{};
                    bool pass = false;
                    if (c < nch) {
                        ch = tnodes[first + c];
                        pass = box_box_d2(qa.lo, qa.hi, ch.lo, ch.hi) <= r_sq;
                    }
                    const bool leaf = pass && rec_b(ch) > 0;
                    const bool inner = pass && rec_b(ch) <= 0;
                    const uint32_t lm = (uint32_t)(ballot(leaf) >> gsh) & 0xffu;
                    const uint32_t im = (uint32_t)(ballot(inner) >> gsh) & 0xffu;
                    if (leaf) {
                        const int pos = cnt + __popc(lm & below);
                        if (pos < cap) {
                            out[pos] = make_int2(rec_a(ch), rec_b(ch));
                            if (DIRS) dout[pos] = dir_code(qa.lo, qa.hi, ch.lo, ch.hi);
                        }
                    }
                    if (HEAVY && lm) {  // the group's sum (a fixed order: the estimate is a heuristic)
                        float e = leaf ? filtered_estimate(qa.lo, qa.hi, ch, rad) : 0.f;
                        int w = leaf ? rec_b(ch) : 0;
                        e += __shfl_xor(e, 1, OCT_WAVE);
                        e += __shfl_xor(e, 2, OCT_WAVE);
                        e += __shfl_xor(e, 4, OCT_WAVE);
                        w += __shfl_xor(w, 1, OCT_WAVE);
                        w += __shfl_xor(w, 2, OCT_WAVE);
                        w += __shfl_xor(w, 4, OCT_WAVE);
                        est += e;
                        raw += w;
                    }
                    cnt += __popc(lm);
                    const int np = __popc(im);
                    if (sp + np > LL_FAST_STACK) {
                        full = true;
                    } else {
                        if (inner) stk[g][sp + __popc(im & below)] = ((uint32_t)rec_a(ch) << 3) | (uint32_t)(-rec_b(ch) - 1);
                        sp += np;
                    }
                }
            }
        }
        if (c == 0) {
            lcount[r] = cnt;
            if (cnt > cap) atomicMax(&flags[0], (uint32_t)cnt);  // the host regrows cap to this
            if (full) atomicOr(&flags[1], 1u);                   // the host reruns in safe mode
        }
        if (HEAVY && c == 0) {  // heavy leaves: flagged here, listed below
            const int qc = rec_b(qa);
            const bool heavy = qc > OCT_WAVE || cnt > OCT_WAVE || est > (float)(EW_TILE * OCT_WAVE) ||
                               (float)qc * est > HEAVY_PAIRS || raw > HEAVY_RAW;
            lheavy[r] = heavy ? 1 : 0;
        }
    }
    if (HEAVY) {
        // The block's heavy leaves go to heavy_list with one atomic per block
        // (one per wave, with eight leaves a wave, made the shared counter the
        // kernel's bottleneck): count, reserve, then write, re-reading the
        // flags this block's threads wrote.
        __shared__ uint32_t hcount, hbase, hnext;
        if (threadIdx.x == 0) {
            hcount = 0;
            hnext = 0;
        }
        __syncthreads();
        uint32_t mine = 0;
        if (c == 0)
            for (int32_t r = (int32_t)(sh.lo + (int64_t)sh.k * LL_COOP_GROUPS) + g; r < (int32_t)sh.hi;
                 r += (int32_t)(sh.nb * LL_COOP_GROUPS))
                mine += lheavy[r];
        if (mine) atomicAdd(&hcount, mine);
        __syncthreads();
        if (threadIdx.x == 0) hbase = hcount ? atomicAdd(n_heavy, hcount) : 0u;
        __syncthreads();
        if (mine) {
            uint32_t idx = hbase + atomicAdd(&hnext, mine);
            for (int32_t r = (int32_t)(sh.lo + (int64_t)sh.k * LL_COOP_GROUPS) + g; r < (int32_t)sh.hi;
                 r += (int32_t)(sh.nb * LL_COOP_GROUPS))
                if (lheavy[r]) {
                    if (idx < (uint32_t)nleaves) heavy_list[idx] = r;
                    ++idx;
                }
        }
    }
}

// ---------------------------------------------------------------- stage 2


// Per-wave LDS of the edge passes.
struct WaveLds {
    int32_t* start;  // 65: prefix of candidate-leaf sizes
    int32_t* first;  // 64: first sorted position of each candidate leaf
    int32_t* cand;   // EW_TILE * 64: sorted positions of the candidates that pass the query-box filter
    float4* tile;    // FILL: EW_TILE * 64, the candidate tile being swept
    int2* rec;       // FILL: the tile's edges as (output position, query slot << 8 | tile slot)
    uint32_t nrec;   // FILL: records rec holds
};

// Writes the n edges recorded in L.rec (output position, query slot << 8 |
// tile slot) against the tile in L.tile: one lane per edge, the query's
// coordinates taken from the lane that holds it (myq of lane `slot`) with a
// permute, the target's from LDS.  The FP32 operations are the sweep's, in
// the same order, and the square root is correctly rounded (as sqrtf), so
// the edges are the ones a per-hit store from the sweep would write.
__device__ inline void write_tile_edges(const WaveLds& L, uint32_t n, float4 myq, float inv_r, EdgeOut out,
                                        int index_base, int lane) {
    __builtin_amdgcn_wave_barrier();
    for (uint32_t t0 = 0; t0 < n; t0 += OCT_WAVE) {  // wave-uniform
        const uint32_t t = t0 + (uint32_t)lane;
        const bool valid = t < n;
        int2 h = make_int2(0, 0);
        if (valid) h = L.rec[t];
        const int qs = (int)((uint32_t)h.y >> 8) << 2;  // the query's lane, as a byte address
        const float4 p = L.tile[h.y & 255];
        const float qx = __int_as_float(__builtin_amdgcn_ds_bpermute(qs, __float_as_int(myq.x)));
        const float qy = __int_as_float(__builtin_amdgcn_ds_bpermute(qs, __float_as_int(myq.y)));
        const float qz = __int_as_float(__builtin_amdgcn_ds_bpermute(qs, __float_as_int(myq.z)));
        const int32_t qi = __builtin_amdgcn_ds_bpermute(qs, __float_as_int(myq.w));
        const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float sd = ballot(valid && d2 < SQRT_DIRECT_MIN) ? sqrtf(d2) : sqrt_rn_direct(d2);
        if (valid) {
            const int64_t e = (int64_t)(uint32_t)h.x;
            out.senders[e] = __float_as_int(p.w) + index_base;
            out.receivers[e] = qi + index_base;
            *reinterpret_cast<float3*>(out.rdisp + 3 * e) =
                make_float3(rel_disp(qx, p.x, inv_r), rel_disp(qy, p.y, inv_r), rel_disp(qz, p.z, inv_r));
            out.rdist[e] = sd * inv_r;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// Loads candidate points t0 .. t0 + 255 of the sequence that L.start /
// L.first describe into tp / tpos (w = point index; -1 marks an empty slot).
// With `filter`, only points whose FP32 distance to the query box is within r
// are kept, compacted in sequence order: that distance is a lower bound of
// the distance to every query in the box (each step is a monotone IEEE
// operation), so no edge is lost.  Returns the number of points loaded.
__device__ inline int load_tile(int t0, int total, const float4* __restrict__ tsc, const WaveLds& L, int lane,
                                bool filter, float4 qlo, float4 qhi, float r_sq, float4 (&tp)[EW_TILE],
                                int32_t (&tpos)[EW_TILE]) {
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        const int ci = t0 + u * OCT_WAVE + lane;
        tp[u] = make_float4(INFINITY, INFINITY, INFINITY, __int_as_float(-1));  // never within r
        tpos[u] = 0;
        if (ci < total) {
            int lo = 0, hi = OCT_WAVE;  // owner: last o with start[o] <= ci
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (L.start[mid] <= ci) lo = mid;
                else hi = mid;
            }
            tpos[u] = L.first[lo] + (ci - L.start[lo]);
            tp[u] = tsc[tpos[u]];
        }
    }
    const int n = min(total - t0, EW_TILE * OCT_WAVE);
    if (!filter) return n;
    int kept = 0;
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        const bool keep = t0 + u * OCT_WAVE + lane < total &&
                          point_box_d2(tp[u].x, tp[u].y, tp[u].z, qlo, qhi) <= r_sq;
        const uint64_t m = ballot(keep);
        if (keep) L.cand[kept + mbcnt(m)] = tpos[u];
        kept += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        const int k = u * OCT_WAVE + lane;
        tp[u] = make_float4(INFINITY, INFINITY, INFINITY, __int_as_float(-1));  // never within r
        if (k < kept) {
            tpos[u] = L.cand[k];
            tp[u] = tsc[tpos[u]];
        }
    }
    __builtin_amdgcn_wave_barrier();
    return kept;
}

// One query leaf through one wave.  Lanes take candidate target points (a
// tile of 4 x 64 in registers); the leaf's queries sit one per lane and are
// broadcast with readlane, query-outer, so each (query, 64 candidates) step
// is one ballot.
//   COUNT: writes the per-query counts to qcount[slot]; returns the leaf total.
//   FILL:  reads qoff[slot] and writes the edges: each tile's hits are
//          recorded in LDS (output position, query slot, tile slot) and
//          written by write_tile_edges, one lane per edge.
// slot = sorted position (leaf order) or original index (receiver order).
template <int MODE>
__device__ inline uint32_t leaf_pass(int32_t qf, int32_t qc, float4 qlo, float4 qhi, int32_t ncand, int2 lf0,
                                     const int2* __restrict__ lst, const float4* __restrict__ qsc,
                                     const float4* __restrict__ tsc, float r_sq, float inv_r, int self_pair,
                                     int by_receiver, uint32_t* __restrict__ qcount,
                                     const uint32_t* __restrict__ qoff, EdgeOut out, int index_base,
                                     const WaveLds& L, int lane) {
    uint32_t leaf_total = 0;
    const bool filter = qc <= OCT_WAVE;  // the leaf box bounds the one query block
    for (int qb = 0; qb < qc; qb += OCT_WAVE) {
        const int nqb = min(OCT_WAVE, qc - qb);
        // query qb + lane lives in this lane
        float4 myq = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane < nqb) myq = qsc[qf + qb + lane];
        const int32_t slot = by_receiver ? __float_as_int(myq.w) : qf + qb + lane;
        uint32_t my = 0;  // COUNT: count; FILL: next output position of this lane's query
        if (MODE != EDGE_COUNT && lane < nqb) my = qoff[slot];
        // FILL: the sweep records each hit as (output position, query slot,
        // tile slot) and write_tile_edges computes and stores the edges with
        // one lane per edge, from the tile in LDS and the queries in
        // registers: the features are computed by whole waves, not by the
        // few lanes of a 64-candidate step that hit, and the stores of one
        // instruction are consecutive edges.
        uint32_t nrec = 0;
        auto flush = [&]() {
            write_tile_edges(L, nrec, myq, inv_r, out, index_base, lane);
            nrec = 0;
        };
        for (int g0 = 0; g0 < ncand; g0 += OCT_WAVE) {
            int2 lf = make_int2(0, 0);
            if (g0 + lane < ncand) lf = (g0 == 0) ? lf0 : lst[g0 + lane];
            const int incl = wave_incl_scan(lf.y);
            const int total = __builtin_amdgcn_readlane(incl, OCT_WAVE - 1);
            L.start[lane + 1] = incl;
            if (lane == 0) L.start[0] = 0;
            L.first[lane] = lf.x;
            __builtin_amdgcn_wave_barrier();
            for (int t0 = 0; t0 < total; t0 += EW_TILE * OCT_WAVE) {
                float4 tp[EW_TILE];
                int32_t tpos[EW_TILE];
                const int got = load_tile(t0, total, tsc, L, lane, filter, qlo, qhi, r_sq, tp, tpos);
                const int nu = (got + OCT_WAVE - 1) / OCT_WAVE;
                if (MODE == EDGE_FILL) {
#pragma unroll
                    for (int u = 0; u < EW_TILE; ++u) L.tile[u * OCT_WAVE + lane] = tp[u];
                    __builtin_amdgcn_wave_barrier();
                }
                for (int qi = 0; qi < nqb; ++qi) {
                    const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.x), qi));
                    const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.y), qi));
                    const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.z), qi));
                    const int32_t i = __builtin_amdgcn_readlane(__float_as_int(myq.w), qi);
                    const int32_t ix = self_pair ? i : -2;  // -2 matches no point: only the self pair drops j == i
                    uint32_t run = MODE != EDGE_COUNT ? (uint32_t)__builtin_amdgcn_readlane((int)my, qi) : 0u;
                    const uint32_t run0 = run;
#pragma unroll
                    for (int u = 0; u < EW_TILE; ++u) {
                        if (u >= nu) break;
                        if (MODE == EDGE_FILL && nrec + (uint32_t)OCT_WAVE > L.nrec) flush();  // wave-uniform
                        const int32_t j = __float_as_int(tp[u].w);
                        const float dx = qx - tp[u].x;
                        const float dy = qy - tp[u].y;
                        const float dz = qz - tp[u].z;
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        const bool in_r = d2 <= r_sq;
                        const bool other = j != ix;
                        // compare masks ANDed, no branch; COUNT counts the self pair
                        // here and takes it off once per query below
                        const uint64_t m = MODE == EDGE_COUNT ? ballot(in_r) : ballot(in_r) & ballot(other);
                        if (MODE == EDGE_INDEX && (in_r & other)) out.senders[run + (uint32_t)mbcnt(m)] = j + index_base;
                        if (MODE == EDGE_FILL && (in_r & other))
                            L.rec[nrec + (uint32_t)mbcnt(m)] =
                                make_int2((int)(run + (uint32_t)mbcnt(m)), (qi << 8) | (u * OCT_WAVE + lane));
                        run += (uint32_t)__popcll(m);
                        if (MODE == EDGE_FILL) nrec += (uint32_t)__popcll(m);
                    }
                    if (lane == qi) my += run - run0;
                }
                if (MODE == EDGE_FILL && nrec) flush();  // before the tile is replaced
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (MODE == EDGE_COUNT) {
            // The self pair was counted: a query's own point is among its
            // candidates (its leaf is in its list at box distance 0, and the
            // point-box filter keeps it) and within r, unless a coordinate is
            // not finite (q - q is then NaN, as the sweep computed it).
            if (self_pair && lane < nqb) {
                const float sx = myq.x - myq.x, sy = myq.y - myq.y, sz = myq.z - myq.z;
                if (sx * sx + sy * sy + sz * sz <= r_sq) my -= 1u;
            }
            if (lane < nqb) qcount[slot] = my;
            leaf_total += (uint32_t)wave_sum(lane < nqb ? (int)my : 0);
        }
    }
    return leaf_total;
}

// Hit-list entry: compact = query slot in the leaf (6 bits) << 26 | target
// sorted position (needs fewer than 2^26 target points); wide = (query
// sorted position, target sorted position).
constexpr uint32_t HIT_POS_BITS = 26;
constexpr uint32_t HIT_POS_MASK = (1u << HIT_POS_BITS) - 1u;

// Filtered candidates one wave stages in LDS (x, y, z, sorted position); the
// same space holds the slot -> sorted position table before that.
constexpr int SP_STAGE_PTS = 128;

// The query-outer sweep of one register tile (nu chunks of 64 candidates,
// w = sorted position) for the leaf's queries: hits are appended at
// hits[run ..] while they fit `room`; returns the leaf's exact hit count.
// The self pair is the candidate at the query's own sorted position.  With
// COUNTS the per-query counts go to qcount instead (the list overflowed).
template <bool WIDE, bool COUNTS>
__device__ inline uint32_t sweep_tile(int32_t qf, int nqb, float4 myq, const float4 (&tp)[EW_TILE],
                                      const int32_t (&tpos)[EW_TILE], int nu, float r_sq, int self_pair,
                                      uint32_t* hits, uint32_t room, uint32_t* __restrict__ qcount, int lane) {
    uint32_t run = 0;
    uint32_t mine = 0;
    for (int qi = 0; qi < nqb; ++qi) {
        const uint32_t run0 = run;
        const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.x), qi));
        const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.y), qi));
        const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.z), qi));
        const int32_t sp = self_pair ? qf + qi : -1;  // -1 matches no sorted position
        const uint32_t tag = (uint32_t)qi << HIT_POS_BITS;
        // one query against the tile; CHECK tests each hit against the room
        // (only for a query whose hits might not all fit)
        auto query = [&](auto check) {
            constexpr bool CHECK = decltype(check)::value;
#pragma unroll
            for (int u = 0; u < EW_TILE; ++u) {
                if (u >= nu) break;
                const float dx = qx - tp[u].x;
                const float dy = qy - tp[u].y;
                const float dz = qz - tp[u].z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                const bool in_r = d2 <= r_sq;
                const bool other = tpos[u] != sp;
                const uint64_t m = ballot(in_r) & ballot(other);  // compare masks ANDed, no branch
                if (!COUNTS && (in_r & other)) {
                    const uint32_t k = (uint32_t)mbcnt(m);
                    if (!CHECK || run + k < room) {
                        if (WIDE) (reinterpret_cast<int2*>(hits) + run)[k] = make_int2(qf + qi, tpos[u]);
                        else (hits + run)[k] = tag | (uint32_t)tpos[u];
                    }
                }
                run += (uint32_t)__popcll(m);
            }
        };
        if (COUNTS || run + (uint32_t)nu * OCT_WAVE <= room) query(std::false_type{});
        else query(std::true_type{});
        if (COUNTS && lane == qi) mine = run - run0;
    }
    if (COUNTS && lane < nqb) qcount[qf + lane] = mine;
    return run;
}

// A leaf whose candidates fit one register tile (at most 64 queries, 64
// candidate leaves and 4 x 64 candidate points after the point-box filter)
// is handled in one sweep: the loop is query-outer, so the hits come out
// grouped by query, queries in sorted order and each query's hits in
// candidate order, which is exactly the output order.  Hits are appended to
// the wave's LDS hit list as (query slot, target sorted position) while they
// fit; the return value is the exact hit count either way (a count above
// `room` means the list could not hold them all, and the per-query counts are
// then in qcount).
constexpr uint32_t REC_NOT_TILE = 0xffffffffu;        // the filtered candidates do not fit one register tile

__device__ inline bool recordable_leaf(int32_t qc, int32_t ncand, int total) {
    return qc <= OCT_WAVE && ncand <= OCT_WAVE && total <= REC_MAX_CAND;
}

// All candidates of the sequence L.start / L.first (at most REC_MAX_CAND),
// filtered by their distance to the query box and compacted into one register
// tile.  Returns the number kept, or -1 when more than one tile survives.
__device__ inline int load_filtered_tile(int total, const float4* __restrict__ tsc, const WaveLds& L, int lane,
                                         float4 qlo, float4 qhi, float r_sq, float4 (&tp)[EW_TILE],
                                         int32_t (&tpos)[EW_TILE]) {
    int kept = 0;
    for (int t0 = 0; t0 < total; t0 += EW_TILE * OCT_WAVE) {
        bool keep[EW_TILE];
#pragma unroll
        for (int u = 0; u < EW_TILE; ++u) {
            const int ci = t0 + u * OCT_WAVE + lane;
            keep[u] = false;
            tpos[u] = 0;
            if (ci < total) {
                int lo = 0, hi = OCT_WAVE;  // owner: last o with start[o] <= ci
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (L.start[mid] <= ci) lo = mid;
                    else hi = mid;
                }
                tpos[u] = L.first[lo] + (ci - L.start[lo]);
                const float4 p = tsc[tpos[u]];
                keep[u] = point_box_d2(p.x, p.y, p.z, qlo, qhi) <= r_sq;
            }
        }
        int add = 0;
#pragma unroll
        for (int u = 0; u < EW_TILE; ++u) add += __popcll(ballot(keep[u]));
        if (kept + add > EW_TILE * OCT_WAVE) return -1;  // wave-uniform
#pragma unroll
        for (int u = 0; u < EW_TILE; ++u) {
            const uint64_t m = ballot(keep[u]);
            if (keep[u]) L.cand[kept + mbcnt(m)] = tpos[u];
            kept += __popcll(m);
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        const int k = u * OCT_WAVE + lane;
        tp[u] = make_float4(INFINITY, INFINITY, INFINITY, __int_as_float(-1));  // never within r
        tpos[u] = 0;
        if (k < kept) {
            tpos[u] = L.cand[k];
            tp[u] = tsc[tpos[u]];
        }
    }
    __builtin_amdgcn_wave_barrier();
    return kept;
}

// Loads the leaf's candidates (the sequence of its candidate leaves, at most
// one register tile of them) for leaf_sweep_record: a slot -> sorted
// position table in LDS (each candidate leaf writes its own slots, so no
// search), the gather, then the point-box filter (its FP32 distance to the
// query box is a lower bound of the distance to every query in the box, so
// no edge is lost).  When the filter saves at least one chunk of 64 and the
// kept points fit the stage, they are compacted through LDS in sequence order
// (no second gather).  Returns the number of candidates in tp / tpos.
__device__ inline int load_record_tile(int2 lf, int incl, int total, const float4* __restrict__ tsc, float4* stage,
                                       int lane, float4 qlo, float4 qhi, float r_sq, float4 (&tp)[EW_TILE],
                                       int32_t (&tpos)[EW_TILE]) {
    int32_t* tbl = reinterpret_cast<int32_t*>(stage);  // total <= 256 entries: 1 KB of the stage
    const int ex = incl - lf.y;
    for (int k = 0; ballot(k < lf.y) != 0; ++k)
        if (k < lf.y) tbl[ex + k] = lf.x + k;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        const int ci = u * OCT_WAVE + lane;
        tpos[u] = ci < total ? tbl[ci] : 0;
    }
    __builtin_amdgcn_wave_barrier();
    bool keep[EW_TILE];
    uint64_t km[EW_TILE];
    int kept = 0;
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        const int ci = u * OCT_WAVE + lane;
        tp[u] = make_float4(INFINITY, INFINITY, INFINITY, 0.f);  // never within r
        if (ci < total) tp[u] = tsc[tpos[u]];
        keep[u] = ci < total && point_box_d2(tp[u].x, tp[u].y, tp[u].z, qlo, qhi) <= r_sq;
        km[u] = ballot(keep[u]);
        kept += __popcll(km[u]);
    }
    const int nu_all = (total + OCT_WAVE - 1) / OCT_WAVE;
    if (kept > SP_STAGE_PTS || (kept + OCT_WAVE - 1) / OCT_WAVE >= nu_all) return total;  // wave-uniform
    int base = 0;
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        if (keep[u])
            stage[base + mbcnt(km[u])] = make_float4(tp[u].x, tp[u].y, tp[u].z, __int_as_float(tpos[u]));
        base += __popcll(km[u]);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        const int k = u * OCT_WAVE + lane;
        tp[u] = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
        tpos[u] = 0;
        if (u * OCT_WAVE < SP_STAGE_PTS && k < kept) {
            tp[u] = stage[k];
            tpos[u] = __float_as_int(tp[u].w);
        }
    }
    __builtin_amdgcn_wave_barrier();
    return kept;
}

template <bool WIDE>
__device__ inline uint32_t leaf_sweep_record(int32_t qf, int32_t qc, float4 qlo, float4 qhi, int32_t ncand, int2 lf0,
                                             const float4* __restrict__ qsc, const float4* __restrict__ tsc,
                                             float r_sq, int self_pair, const WaveLds& L, float4* stage, int lane,
                                             uint32_t* hits, uint32_t room, uint32_t* __restrict__ qcount) {
    const int nqb = qc;
    float4 myq = make_float4(0.f, 0.f, 0.f, 0.f);
    if (lane < nqb) myq = qsc[qf + lane];
    int2 lf = make_int2(0, 0);
    if (lane < ncand) lf = lf0;
    const int incl = wave_incl_scan(lf.y);
    const int total = __builtin_amdgcn_readlane(incl, OCT_WAVE - 1);
    float4 tp[EW_TILE];
    int32_t tpos[EW_TILE];
    int got;
    if (total <= EW_TILE * OCT_WAVE) {
        got = load_record_tile(lf, incl, total, tsc, stage, lane, qlo, qhi, r_sq, tp, tpos);
    } else {
        L.start[lane + 1] = incl;
        if (lane == 0) L.start[0] = 0;
        L.first[lane] = lf.x;
        __builtin_amdgcn_wave_barrier();
        got = load_filtered_tile(total, tsc, L, lane, qlo, qhi, r_sq, tp, tpos);
        if (got < 0) return REC_NOT_TILE;
    }
    const int nu = (got + OCT_WAVE - 1) / OCT_WAVE;
    const uint32_t run = sweep_tile<WIDE, false>(qf, nqb, myq, tp, tpos, nu, r_sq, self_pair, hits, room, qcount,
                                                 lane);
    // a list that could not hold every hit: the per-query counts let the
    // caller write this leaf with the FILL body without another count sweep
    if (run > room) sweep_tile<WIDE, true>(qf, nqb, myq, tp, tpos, nu, r_sq, self_pair, hits, room, qcount, lane);
    return run;
}

// Writes n recorded hits of the leaf whose first query is qf at out[base ..):
// every lane takes consecutive hits, so the stores of one instruction are
// contiguous (the displacement as one 12-byte store per edge).  EMIT_BATCH
// hits per lane are loaded before any is written, so their gathers overlap.
// The leaf's output range is addressed from scalar bases with 32-bit
// per-lane byte offsets.
constexpr int EMIT_BATCH = 4;

template <bool WIDE>
__device__ inline void emit_hits(const uint32_t* hits, uint32_t n, uint32_t base, int32_t qf,
                                 const float4* __restrict__ qsc, const float4* __restrict__ tsc, float inv_r,
                                 EdgeOut out, int index_base, int lane) {
    char* const sen = reinterpret_cast<char*>(out.senders + base);
    char* const rcv = reinterpret_cast<char*>(out.receivers + base);
    char* const dsp = reinterpret_cast<char*>(out.rdisp + 3 * (int64_t)base);
    char* const dst = reinterpret_cast<char*>(out.rdist + base);
    const float4* const qb = qsc + qf;
    for (uint32_t t0 = 0; t0 < n; t0 += EMIT_BATCH * OCT_WAVE) {
        float4 q[EMIT_BATCH], p[EMIT_BATCH];
#pragma unroll
        for (int k = 0; k < EMIT_BATCH; ++k) {
            const uint32_t t = t0 + (uint32_t)(k * OCT_WAVE + lane);
            if (WIDE) {
                int2 h = make_int2(qf, 0);
                if (t < n) h = reinterpret_cast<const int2*>(hits)[t];
                q[k] = qsc[h.x];
                p[k] = tsc[h.y];
            } else {
                uint32_t h = 0;
                if (t < n) h = hits[t];
                // the query slot (< 64) and the target position (< 2^26) as byte offsets that fit 32 bits
                q[k] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(qb) + ((h >> HIT_POS_BITS) << 4));
                p[k] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(tsc) + ((h & HIT_POS_MASK) << 4));
            }
        }
#pragma unroll
        for (int k = 0; k < EMIT_BATCH; ++k) {
            const uint32_t t = t0 + (uint32_t)(k * OCT_WAVE + lane);
            if (t >= n) break;
            const float dx = q[k].x - p[k].x, dy = q[k].y - p[k].y, dz = q[k].z - p[k].z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            const float sd = ballot(d2 < SQRT_DIRECT_MIN) ? sqrtf(d2) : sqrt_rn_direct(d2);
            *reinterpret_cast<int32_t*>(sen + (t << 2)) = __float_as_int(p[k].w) + index_base;
            *reinterpret_cast<int32_t*>(rcv + (t << 2)) = __float_as_int(q[k].w) + index_base;
            *reinterpret_cast<float3*>(dsp + t * 12u) =
                make_float3(rel_disp(q[k].x, p[k].x, inv_r), rel_disp(q[k].y, p[k].y, inv_r),
                            rel_disp(q[k].z, p[k].z, inv_r));
            *reinterpret_cast<float*>(dst + (t << 2)) = sd * inv_r;
        }
    }
}

template <int MODE>
struct EdgeLds {
    static constexpr int FT = MODE == EDGE_FILL ? EW_TILE * OCT_WAVE : 1;  // tile slots and records
    int32_t start[EW_WAVES][OCT_WAVE + 1];
    int32_t first[EW_WAVES][OCT_WAVE];
    int32_t cand[EW_WAVES][EW_TILE * OCT_WAVE];
    float4 tile[EW_WAVES][FT];
    int2 rec[EW_WAVES][FT];
    __device__ WaveLds wave(int w) { return WaveLds{start[w], first[w], cand[w], tile[w], rec[w], (uint32_t)FT}; }
};

// Two-pass driver: one wave per query leaf over a contiguous, XCD-local range
// of leaves, the next leaf's metadata loaded while the current one runs.
template <int MODE>
__global__ void __launch_bounds__(EW_WAVES * OCT_WAVE)
k_edges(const NodeRec* __restrict__ qleaves, const float4* __restrict__ qsc, const BuildParams* __restrict__ qprm,
        const float4* __restrict__ tsc, const int2* __restrict__ lists, const int32_t* __restrict__ lcount, int cap,
        float r_sq, float inv_r, int self_pair, int by_receiver, uint32_t* __restrict__ qcount,
        const uint32_t* __restrict__ qoff, uint32_t* __restrict__ lhits, EdgeOut out, int index_base) {
    __shared__ EdgeLds<MODE> lds;
    const int wv = threadIdx.x / OCT_WAVE;
    const int lane = lane_id();
    const WaveLds L = lds.wave(wv);
    const int32_t nleaves = qprm->n_leaves;
    // this XCD's share of the leaves, split into contiguous runs per wave
    const XcdShare sh = xcd_share(nleaves);
    const int32_t nwaves = (int32_t)(sh.nb * EW_WAVES);
    const int32_t w = (int32_t)sh.k * EW_WAVES + wv;
    const int32_t len = (int32_t)(sh.hi - sh.lo);
    const int32_t per = (len + nwaves - 1) / nwaves;
    const int32_t r_lo = __builtin_amdgcn_readfirstlane((int32_t)sh.lo + min(w * per, len));
    const int32_t r_hi = min((int32_t)sh.hi, r_lo + per);
    NodeRec nx_rec{};
    int2 nx_lf = make_int2(0, 0);
    int32_t nx_cnt = 0;
    if (r_lo < r_hi) {
        nx_rec = qleaves[r_lo];
        nx_cnt = lcount[r_lo];
        if (lane < cap) nx_lf = lists[(int64_t)r_lo * cap + lane];
    }
    for (int32_t r = r_lo; r < r_hi; ++r) {
        const NodeRec qa = nx_rec;
        const int32_t qf = __builtin_amdgcn_readfirstlane(rec_a(qa));
        const int32_t qc = __builtin_amdgcn_readfirstlane(rec_b(qa));
        const int32_t ncand = __builtin_amdgcn_readfirstlane(min(nx_cnt, cap));
        const int2 lf0 = nx_lf;
        if (r + 1 < r_hi) {
            nx_rec = qleaves[r + 1];
            nx_cnt = lcount[r + 1];
            if (lane < cap) nx_lf = lists[(int64_t)(r + 1) * cap + lane];
        }
        const uint32_t lt = leaf_pass<MODE>(qf, qc, qa.lo, qa.hi, ncand, lf0, lists + (int64_t)r * cap, qsc, tsc,
                                            r_sq, inv_r, self_pair, by_receiver, qcount, qoff, out, index_base, L,
                                            lane);
        if (MODE == EDGE_COUNT && lhits && lane == 0) lhits[r] = lt;
    }
}

// Radius and output index base of the query leaf whose first sorted position
// is qf: the call's own, or in a forest those of the leaf's set (indices stay
// local to each set).
__device__ inline void leaf_set_params(const SetLookup& sets, int32_t qf, float r_sq, float inv_r, int index_base0,
                                       float& lr_sq, float& linv_r, int& libase) {
    lr_sq = r_sq;
    linv_r = inv_r;
    libase = index_base0;
    if (sets.batched) {
        const int s = key_set(sets.qsk[qf], 1);
        const float rr = sets.sprm[s].radius;
        lr_sq = rr * rr;
        linv_r = 1.0f / rr;
        libase = index_base0 - (int)sets.set_off[s];
    }
}

// ------------------------------------------------------------ heavy leaves

// Exact counts of the heavy query leaves that the candidate walk listed, one wave
// per leaf, before the single pass: per-query counts into qcount, the leaf's
// total into lhits[r].  Same sweep as the two-pass COUNT kernel.
__global__ void __launch_bounds__(EW_WAVES * OCT_WAVE)
k_edges_heavy(const NodeRec* __restrict__ qleaves, const float4* __restrict__ qsc,
              const BuildParams* __restrict__ qprm, const float4* __restrict__ tsc, const int2* __restrict__ lists,
              const int32_t* __restrict__ lcount, int cap, float r_sq, float inv_r, int self_pair,
              const int32_t* __restrict__ heavy_list, const uint32_t* __restrict__ n_heavy,
              uint32_t* __restrict__ qcount, uint32_t* __restrict__ lhits, SetLookup sets) {
    __shared__ EdgeLds<EDGE_COUNT> lds;
    const int wv = threadIdx.x / OCT_WAVE;
    const int lane = lane_id();
    const WaveLds L = lds.wave(wv);
    const int32_t nleaves = qprm->n_leaves;
    const int32_t nh = (int32_t)min(*n_heavy, (uint32_t)max(nleaves, 0));
    const int32_t nwaves = (int32_t)(gridDim.x * EW_WAVES);
    for (int32_t i = (int32_t)blockIdx.x * EW_WAVES + wv; i < nh; i += nwaves) {
        const int32_t r = __builtin_amdgcn_readfirstlane(heavy_list[i]);
        if (r < 0 || r >= nleaves) continue;  // wave-uniform; the candidate walk lists only leaves
        const NodeRec qa = qleaves[r];
        const int32_t qf = __builtin_amdgcn_readfirstlane(rec_a(qa));
        const int32_t qc = __builtin_amdgcn_readfirstlane(rec_b(qa));
        const int32_t ncand = __builtin_amdgcn_readfirstlane(min(lcount[r], cap));
        const int2* lst = lists + (int64_t)r * cap;
        const int2 lf0 = lst[lane < ncand ? lane : 0];
        float lr_sq, linv_r;
        int libase;
        leaf_set_params(sets, qf, r_sq, inv_r, 0, lr_sq, linv_r, libase);
        const uint32_t c = leaf_pass<EDGE_COUNT>(qf, qc, qa.lo, qa.hi, ncand, lf0, lst, qsc, tsc, lr_sq, linv_r,
                                                 self_pair, 0, qcount, nullptr, EdgeOut{}, libase, L, lane);
        if (lane == 0) lhits[r] = c;
    }
}

// Writes the edges of the heavy leaves after the single pass, one wave per
// leaf (the batched sets, whose heavy leaves are seldom dense; a single pair
// takes k_edges_dense_fill): the single pass left each one's output position
// in qoff at its first query, and its per-query counts in qcount came from
// k_edges_heavy.  Nothing runs when the single pass raised flags[0]
// (capacity) or flags[1] (look-back or plan): the host then redoes the call.
__global__ void __launch_bounds__(EW_WAVES * OCT_WAVE)
k_edges_heavy_fill(const NodeRec* __restrict__ qleaves, const float4* __restrict__ qsc,
                   const BuildParams* __restrict__ qprm, const float4* __restrict__ tsc,
                   const int2* __restrict__ lists, const int32_t* __restrict__ lcount, int cap, float r_sq,
                   float inv_r, int self_pair, const int32_t* __restrict__ heavy_list,
                   const uint32_t* __restrict__ n_heavy, uint32_t* __restrict__ qcount, uint32_t* __restrict__ qoff,
                   const uint32_t* __restrict__ flags, EdgeOut out, int index_base0, SetLookup sets) {
    __shared__ EdgeLds<EDGE_FILL> lds;
    if (flags[0] != 0 || flags[1] != 0) return;
    const int wv = threadIdx.x / OCT_WAVE;
    const int lane = lane_id();
    const WaveLds L = lds.wave(wv);
    const int32_t nleaves = qprm->n_leaves;
    const int32_t nh = (int32_t)min(*n_heavy, (uint32_t)max(nleaves, 0));
    const int32_t nwaves = (int32_t)(gridDim.x * EW_WAVES);
    for (int32_t i = (int32_t)blockIdx.x * EW_WAVES + wv; i < nh; i += nwaves) {
        const int32_t r = __builtin_amdgcn_readfirstlane(heavy_list[i]);
        if (r < 0 || r >= nleaves) continue;  // wave-uniform; the candidate walk lists only leaves
        const NodeRec qa = qleaves[r];
        const int32_t qf = __builtin_amdgcn_readfirstlane(rec_a(qa));
        const int32_t qc = __builtin_amdgcn_readfirstlane(rec_b(qa));
        uint32_t run = (uint32_t)__builtin_amdgcn_readfirstlane((int)qoff[qf]);  // before the scan rewrites it
        __builtin_amdgcn_wave_barrier();
        for (int qb = 0; qb < qc; qb += OCT_WAVE) {
            const int nqb = min(OCT_WAVE, qc - qb);
            const uint32_t c = lane < nqb ? qcount[qf + qb + lane] : 0u;
            const uint32_t inc = (uint32_t)wave_incl_scan((int)c);
            if (lane < nqb) qoff[qf + qb + lane] = run + inc - c;
            run += (uint32_t)__builtin_amdgcn_readlane((int)inc, OCT_WAVE - 1);
        }
        const int32_t ncand = __builtin_amdgcn_readfirstlane(min(lcount[r], cap));
        const int2* lst = lists + (int64_t)r * cap;
        const int2 lf0 = lst[lane < ncand ? lane : 0];
        float lr_sq, linv_r;
        int libase;
        leaf_set_params(sets, qf, r_sq, inv_r, index_base0, lr_sq, linv_r, libase);
        leaf_pass<EDGE_FILL>(qf, qc, qa.lo, qa.hi, ncand, lf0, lst, qsc, tsc, lr_sq, linv_r, self_pair, 0, qcount,
                             qoff, out, libase, L, lane);
    }
}

// Writes the edges of the heavy leaves after the single pass, one block per
// leaf, query by query.  Heavy leaves are dense: each query has hundreds of
// hits among thousands of candidates.  Swept tile by tile with every query of
// a 64-query block (leaf_pass), each query's output is written a piece per
// tile, and the pieces of 64 rows land in HBM as partial lines; here the block
// stages the leaf's candidates (those within r of the leaf box, in sequence
// order, across candidate leaf groups, a stage at a time) in LDS, and each
// wave takes one query at a time through all of them, recording its hits
// (stage slots) and writing them out as consecutive edges of that query, whole
// waves at a time.  The edges and their order (candidate sequence order per
// query) are leaf_pass's.  Queries go 256 at a time: their cursors (the
// leaf's output position, left in qoff at its first query by the single pass,
// plus the counts of the COUNT launch in qcount) live in LDS, each wave takes
// a contiguous share of them, and lane j holds the j-th query's position in
// registers, so a query's turn reads no global memory.  Nothing runs when the
// single pass raised flags[0] (capacity) or flags[1] (look-back or plan).
constexpr int DF_THREADS = 256;
constexpr int DF_WAVES = DF_THREADS / OCT_WAVE;
// Candidates staged per round: FILL keeps rounds long (each round costs every
// query a flush), COUNT keeps blocks small (it has no records; 960 against
// 1,984 took its launch 1.48 -> 1.24 ms at the clustered workload, where the
// fill went 3.99 -> 5.02 ms).
constexpr int DF_STAGE_FILL = 1984;
constexpr int DF_STAGE_COUNT = 960;
constexpr int DF_UNROLL = 4;    // candidate chunks of 64 per step of the query sweep
constexpr int DF_RECS = 512;    // hit records per wave (stage slots)
constexpr int DF_LOADS_MAX = 4;

// Per-block LDS of k_edges_dense.  The stage array: 64 slots past the last
// staged candidate hold points at infinity, so that the sweep's last chunk
// reads them and finds no hit (32 KB for FILL).
template <int STAGE, bool RECS>
struct DenseLds {
    static constexpr int stage = STAGE;
    float4 pts[STAGE + OCT_WAVE];
    uint16_t rec[DF_WAVES][RECS ? DF_RECS : 1];
    uint32_t cur[DF_THREADS];
    int32_t start[OCT_WAVE + 1];
    int32_t first[OCT_WAVE];
    uint32_t wsum[DF_LOADS_MAX][DF_WAVES];
    uint32_t carry;
};

// One staging step: candidates [t, t + SU * DF_THREADS) of the current group
// (clipped to total), those within r of the box appended to L.pts in sequence
// order.  Block-uniform; returns the new stage size.
template <int SU, typename LDS>
__device__ inline int df_stage_step(LDS& L, int t, int total, int kept, const float4* __restrict__ tsc,
                                    float4 lo, float4 hi, float lr_sq, int wv) {
    static_assert(SU <= DF_LOADS_MAX, "wsum holds DF_LOADS_MAX steps");
    float4 p[SU];
    bool keep[SU];
#pragma unroll
    for (int k = 0; k < SU; ++k) {
        const int ci = t + k * DF_THREADS + (int)threadIdx.x;
        keep[k] = false;
        p[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ci < total) {
            int o = 0, hi2 = OCT_WAVE;  // owner: last o with start[o] <= ci
            while (hi2 - o > 1) {
                const int mid = (o + hi2) >> 1;
                if (L.start[mid] <= ci) o = mid;
                else hi2 = mid;
            }
            p[k] = tsc[L.first[o] + (ci - L.start[o])];
        }
    }
    uint64_t m[SU];
#pragma unroll
    for (int k = 0; k < SU; ++k) {
        const int ci = t + k * DF_THREADS + (int)threadIdx.x;
        keep[k] = ci < total && point_box_d2(p[k].x, p[k].y, p[k].z, lo, hi) <= lr_sq;
        m[k] = ballot(keep[k]);
        if (lane_id() == 0) L.wsum[k][wv] = (uint32_t)__popcll(m[k]);
    }
    __syncthreads();
    int base = kept;
#pragma unroll
    for (int k = 0; k < SU; ++k) {
        int wb = base;
        for (int w = 0; w < DF_WAVES; ++w) {
            const int v = (int)L.wsum[k][w];
            wb += w < wv ? v : 0;
            base += v;
        }
        if (keep[k]) L.pts[wb + mbcnt(m[k])] = p[k];
    }
    if (threadIdx.x < OCT_WAVE)  // the padding after the new end (base <= stage)
        L.pts[base + (int)threadIdx.x] = make_float4(INFINITY, INFINITY, INFINITY, __int_as_float(-1));
    __syncthreads();  // the stage is complete and wsum free again
    return base;
}

// MODE EDGE_COUNT: the same sweep counts each query's hits (qcount) and the
// leaf's total (lhits[r]) before the single pass, in place of k_edges_heavy's
// one wave per leaf (whose lanes hold queries: a heavy leaf of 20 queries
// leaves 44 lanes idle, where here the lanes hold candidates).
template <int MODE>
__global__ void __launch_bounds__(DF_THREADS)
k_edges_dense(const NodeRec* __restrict__ qleaves, const float4* __restrict__ qsc,
              const BuildParams* __restrict__ qprm, const float4* __restrict__ tsc, const int2* __restrict__ lists,
              const int32_t* __restrict__ lcount, int cap, float r_sq, float inv_r, int self_pair,
              const int32_t* __restrict__ heavy_list, const uint32_t* __restrict__ n_heavy,
              uint32_t* __restrict__ qcount, const uint32_t* __restrict__ qoff, uint32_t* __restrict__ lhits,
              const uint32_t* __restrict__ flags, EdgeOut out, int index_base0, SetLookup sets) {
    static_assert(MODE == EDGE_COUNT || MODE == EDGE_FILL, "count or fill");
    constexpr bool FILL = MODE == EDGE_FILL;
    constexpr int STAGE = FILL ? DF_STAGE_FILL : DF_STAGE_COUNT;
    constexpr int LOADS = FILL ? 4 : 2;  // a first step of LOADS * 256 fits the empty stage
    static_assert(LOADS * DF_THREADS <= STAGE, "a first staging step fits");
    __shared__ DenseLds<STAGE, FILL> L;
    if (FILL && (flags[0] != 0 || flags[1] != 0)) return;  // block-uniform
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / OCT_WAVE));
    const int lane = lane_id();
    const int32_t nleaves = qprm->n_leaves;
    const int32_t nh = (int32_t)min(*n_heavy, (uint32_t)max(nleaves, 0));
    uint16_t* rec = L.rec[wv];
    for (int32_t hi = (int32_t)blockIdx.x; hi < nh; hi += (int32_t)gridDim.x) {  // block-uniform
        const int32_t r = heavy_list[hi];
        if (r < 0 || r >= nleaves) continue;  // the candidate walk lists only leaves
        const NodeRec qa = qleaves[r];
        const int32_t qf = rec_a(qa), qc = rec_b(qa);
        const int32_t ncand = min(lcount[r], cap);
        const int2* lst = lists + (int64_t)r * cap;
        float lr_sq, linv_r;
        int libase;
        leaf_set_params(sets, qf, r_sq, inv_r, index_base0, lr_sq, linv_r, libase);
        if (FILL && threadIdx.x == 0) L.carry = qoff[qf];
        uint32_t mytotal = 0;  // COUNT: hits of this thread's queries
        for (int32_t b0 = 0; b0 < qc; b0 += DF_THREADS) {  // block-uniform: a batch of 256 queries
            const int32_t nb = min(qc - b0, DF_THREADS);
            if (FILL) {  // cursors of the batch, in query order
                const int32_t qi = (int32_t)threadIdx.x;
                const uint32_t c = qi < nb ? qcount[qf + b0 + qi] : 0u;
                const uint32_t inc = (uint32_t)wave_incl_scan((int)c);
                if (lane == OCT_WAVE - 1) L.wsum[0][wv] = inc;
                __syncthreads();  // wsum and carry written
                uint32_t before = L.carry;
                for (int k = 0; k < wv; ++k) before += L.wsum[0][k];
                L.cur[qi] = before + inc - c;
                __syncthreads();  // every thread has read carry
                if (threadIdx.x == DF_THREADS - 1) L.carry = before + inc;
            } else {
                L.cur[threadIdx.x] = 0u;  // counts; read by others only after the syncs below
            }
            // wave w takes a contiguous share of the batch's queries, lane j its
            // j-th, so that each wave's output is one run (partial lines only at
            // the ends of the shares)
            const int32_t per = (nb + DF_WAVES - 1) / DF_WAVES;  // <= 64
            const int32_t qw0 = wv * per;
            const int nmine = max(0, min(nb, qw0 + per) - qw0);  // wave-uniform
            const float4 myq = lane < nmine ? qsc[qf + b0 + qw0 + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            // the candidate sequence, streamed through the stage
            int32_t g0 = -OCT_WAVE;  // current group of candidate leaves
            int total = 0, t = 0;
            bool more = ncand > 0;
            while (more) {  // block-uniform
                int kept = 0;
                while (true) {  // fill the stage
                    if (t >= total) {  // next group of 64 candidate leaves
                        g0 += OCT_WAVE;
                        if (g0 >= ncand) {
                            more = false;
                            break;
                        }
                        __syncthreads();  // every thread is done with start
                        if (wv == 0) {
                            int2 lf = make_int2(0, 0);
                            if (g0 + lane < ncand) lf = lst[g0 + lane];
                            const int incl = wave_incl_scan(lf.y);
                            L.start[lane + 1] = incl;
                            if (lane == 0) L.start[0] = 0;
                            L.first[lane] = lf.x;
                        }
                        __syncthreads();
                        total = L.start[OCT_WAVE];
                        t = 0;
                        continue;
                    }
                    if (kept + LOADS * DF_THREADS <= STAGE) {
                        kept = df_stage_step<LOADS>(L, t, total, kept, tsc, qa.lo, qa.hi, lr_sq, wv);
                        t += LOADS * DF_THREADS;
                    } else if (kept + DF_THREADS <= STAGE) {
                        kept = df_stage_step<1>(L, t, total, kept, tsc, qa.lo, qa.hi, lr_sq, wv);
                        t += DF_THREADS;
                    } else {
                        break;
                    }
                }
                if (kept == 0) continue;  // block-uniform; the loop ends when more is false
                // each wave: its queries through the staged candidates
                for (int j = 0; j < nmine; ++j) {
                    const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.x), j));
                    const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.y), j));
                    const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.z), j));
                    const int32_t i = __builtin_amdgcn_readlane(__float_as_int(myq.w), j);
                    const int32_t ix = self_pair ? i : -2;  // -2 matches no point: only the self pair drops j == i
                    const int qslot = qw0 + j;
                    uint32_t cur = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.cur[qslot]);
                    uint32_t nrec = 0;
                    auto flush = [&]() {
                        __builtin_amdgcn_wave_barrier();
                        // the query's next nrec (< 2^16) edges, from wave-uniform bases
                        // with 32-bit per-lane byte offsets
                        char* const os = reinterpret_cast<char*>(out.senders + cur);
                        char* const orc = reinterpret_cast<char*>(out.receivers + cur);
                        char* const od = reinterpret_cast<char*>(out.rdisp + 3 * (int64_t)cur);
                        char* const on = reinterpret_cast<char*>(out.rdist + cur);
                        for (uint32_t u0 = 0; u0 < nrec; u0 += OCT_WAVE) {
                            const uint32_t u = u0 + (uint32_t)lane;
                            if (u < nrec) {
                                const float4 pp = L.pts[rec[u]];
                                const float dx = qx - pp.x, dy = qy - pp.y, dz = qz - pp.z;
                                const float d2 = dx * dx + dy * dy + dz * dz;
                                const float sd = ballot(d2 < SQRT_DIRECT_MIN) ? sqrtf(d2) : sqrt_rn_direct(d2);
                                *reinterpret_cast<int32_t*>(os + 4u * u) = __float_as_int(pp.w) + libase;
                                *reinterpret_cast<int32_t*>(orc + 4u * u) = i + libase;
                                *reinterpret_cast<float3*>(od + 12u * u) =
                                    make_float3(rel_disp(qx, pp.x, linv_r), rel_disp(qy, pp.y, linv_r),
                                                rel_disp(qz, pp.z, linv_r));
                                *reinterpret_cast<float*>(on + 4u * u) = sd * linv_r;
                            }
                        }
                        __builtin_amdgcn_wave_barrier();
                        cur += nrec;
                        nrec = 0;
                    };
                    // one chunk of 64 staged candidates: hits recorded (FILL) and counted;
                    // the padding past kept is at infinity, so no slot past kept hits
                    auto chunk = [&](const float4& p, int sl) {
                        const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        const bool in_r = d2 <= lr_sq, other = __float_as_int(p.w) != ix;
                        // the mask from the two compares' own masks (a ballot of
                        // their conjunction would rebuild it through a select)
                        const uint64_t m = ballot(in_r) & ballot(other);
                        if (FILL && (in_r & other)) rec[nrec + (uint32_t)mbcnt(m)] = (uint16_t)sl;
                        nrec += (uint32_t)__popcll(m);
                    };
                    // DF_UNROLL chunks at a time while they lie inside the stage: their
                    // LDS reads are issued together, so one latency serves them all
                    int s0 = 0;
                    for (; s0 + DF_UNROLL * OCT_WAVE <= kept; s0 += DF_UNROLL * OCT_WAVE) {  // wave-uniform
                        if (FILL && nrec + (uint32_t)(DF_UNROLL * OCT_WAVE) > (uint32_t)DF_RECS) flush();
                        float4 pp[DF_UNROLL];
#pragma unroll
                        for (int k = 0; k < DF_UNROLL; ++k) pp[k] = L.pts[s0 + k * OCT_WAVE + lane];
#pragma unroll
                        for (int k = 0; k < DF_UNROLL; ++k) chunk(pp[k], s0 + k * OCT_WAVE + lane);
                    }
                    for (; s0 < kept; s0 += OCT_WAVE) {  // wave-uniform; reads end inside the padding
                        if (FILL && nrec + (uint32_t)OCT_WAVE > (uint32_t)DF_RECS) flush();
                        chunk(L.pts[s0 + lane], s0 + lane);
                    }
                    if (FILL) {
                        if (nrec) flush();
                        if (lane == 0) L.cur[qslot] = cur;  // this wave's query: the next round continues here
                    } else {
                        if (lane == 0) L.cur[qslot] = cur + nrec;
                    }
                }
                __syncthreads();  // the stage is rewritten next round
            }
            __syncthreads();  // cur and start are rewritten for the next batch or leaf
            if (!FILL && (int32_t)threadIdx.x < nb) {
                const uint32_t c = L.cur[threadIdx.x];
                qcount[qf + b0 + (int32_t)threadIdx.x] = c;
                mytotal += c;
            }
        }
        if (!FILL) {  // the leaf's total
            const uint32_t ws = (uint32_t)wave_sum((int)mytotal);
            if (lane == 0) L.wsum[0][wv] = ws;
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t tot = 0;
                for (int k = 0; k < DF_WAVES; ++k) tot += L.wsum[0][k];
                lhits[r] = tot;
            }
        }
    }
}

// ------------------------------------------------------------ single pass

constexpr int SP_MAX_LEAVES = SP_PLAN_MAX_LEAVES;  // leaves one wave may own under a plan
constexpr uint64_t SP_FLAG_AGG = 1ull << 62;   // tile aggregate published
constexpr uint64_t SP_FLAG_PRE = 2ull << 62;   // tile inclusive prefix published
constexpr uint64_t SP_VALUE = (1ull << 62) - 1;
constexpr uint32_t SP_SPIN_LIMIT = 1u << 22;   // bounded wait: a stuck look-back reports an error

constexpr int SP_HIT_WORDS = 1536;                 // hit list of one wave: 6 KB
static_assert(SP_SIG_MARK % 2 == 0, "the marker is a 64-bit word");
static_assert(SP_HIT_WORDS * 4 >= EW_TILE * OCT_WAVE * (16 + 8), "phase B's FILL tile and records fit the hit list");
static_assert(SP_PLAN_BUDGET == SP_HIT_WORDS, "the plan packs one wave's compact hit list");
static_assert(SP_STAGE_PTS * 4 >= EW_TILE * OCT_WAVE, "the stage holds the slot table and the filtered positions");

// Run hand-out of the single pass.  A block takes `chunk` (1..EW_WAVES)
// consecutive runs per global ticket (neighbouring runs share candidate
// points in the CU's caches, and one atomic serves several runs); its waves
// take those runs one at a time as each becomes free, with no block barrier,
// so a slow run holds back no other wave.  A larger chunk gives more cache
// reuse; a smaller one starts every run as soon as it is handed out (a run
// that waits for a wave of its block delays the look-back of all later runs),
// which suits uneven runs.  The LDS word q holds (ticket << SP_Q_SHIFT) | runs
// handed out of it.  The wave that finds the count at `chunk` takes the next
// ticket, keeps its first run and installs it with count 1; a wave that finds
// a larger count (at most EW_WAVES - 1 of them, each adding once) waits for
// the ticket to change.  Returns the run (ticket * chunk + k); runs at or
// past the end mean that no work is left.  chunk 1 takes the global ticket
// directly.
constexpr int SP_Q_SHIFT = 8;
constexpr unsigned long long SP_Q_NONE = 0xffffffffull;  // no ticket yet: never a real one

__device__ inline int64_t take_run(unsigned long long* q, uint32_t* ticket, int lane, int chunk) {
    if (chunk == 1) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(ticket, 1u);
        return (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)t);
    }
    for (;;) {
        unsigned long long old = 0;
        if (lane == 0) old = atomicAdd(q, 1ull);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)old);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(old >> 32));
        old = ((unsigned long long)hi << 32) | lo;
        const unsigned long long tk = old >> SP_Q_SHIFT;
        const uint32_t k = (uint32_t)(old & ((1u << SP_Q_SHIFT) - 1u));
        if (k < (uint32_t)chunk) return (int64_t)tk * chunk + k;
        if (k == (uint32_t)chunk) {  // used up: this wave fetches the next ticket
            uint32_t t = 0;
            if (lane == 0) {
                t = atomicAdd(ticket, 1u);
                __hip_atomic_store(q, ((unsigned long long)t << SP_Q_SHIFT) | 1ull, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
            return (int64_t)t * chunk;
        }
        // another wave is fetching the next ticket
        for (;;) {
            unsigned long long cur = 0;
            if (lane == 0) cur = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const uint32_t chi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(cur >> 32));
            const uint32_t clo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)cur);
            if (((((unsigned long long)chi << 32) | clo) >> SP_Q_SHIFT) != tk) break;
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// Decoupled look-back by one whole wave: lanes read the states of the 64
// tiles before `hi` at once, the nearest tile with a published inclusive
// prefix ends the walk, and every tile between must have published at least
// its aggregate (else the wave waits and reads the window again).  Tiles
// before 0 count as an inclusive prefix of 0.  Returns the exclusive prefix of
// tile t; *ok = 0 when the bounded wait ran out.
__device__ inline uint64_t wave_look_back(const uint64_t* tile_state, int32_t t, int lane, uint32_t* ok) {
    uint64_t excl = 0;
    uint32_t spins = 0;
    int32_t hi = t - 1;
    while (hi >= 0) {
        const int32_t k = hi - lane;  // lane 0 is the nearest tile
        const uint64_t v = k >= 0 ? __hip_atomic_load(&tile_state[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : SP_FLAG_PRE;
        const uint64_t flag = v & ~SP_VALUE;
        const uint64_t pre = ballot(flag == SP_FLAG_PRE);
        const uint64_t none = ballot(flag == 0);
        // lanes up to and including the nearest inclusive prefix (or all 64)
        const uint64_t need = pre ? (pre ^ (pre - 1)) : ~0ull;
        if (none & need) {
            if (++spins > SP_SPIN_LIMIT) {
                *ok = 0;
                return excl;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += wave_sum_u64(((need >> lane) & 1ull) ? (v & SP_VALUE) : 0ull);
        if (pre) break;
        hi -= OCT_WAVE;
    }
    return excl;
}

// Single traversal in leaf order.  Every wave owns a run of consecutive query
// leaves, which is also its output tile: leaves_per_run leaves per run, or
// the runs of a plan (wave_first[w] .. wave_first[w + 1], at most
// SP_MAX_LEAVES, *plan_waves runs) that packs the predicted hits of a run into
// the wave's hit list.  Waves take runs through take_run (run_chunk runs per
// block ticket) until none is left.  Phase A sweeps each
// leaf once: a leaf that fits one register tile records its hits in the
// wave's LDS hit list (query-grouped, already in output order) and counts
// them; a leaf whose hits no longer fit the list keeps its per-query counts
// from that sweep, any other leaf is counted with the COUNT body, and both
// are swept again in phase B (stats[0] counts those, stats[1] the recorded
// ones).  lhits (optional) receives every leaf's exact
// hit count.  The run's dense output offset comes from a decoupled look-back
// over the earlier runs' states.  Phase B writes the recorded hits with
// contiguous stores, then re-sweeps the other leaves (the exact recount, with
// the FILL body, whose LDS stage lies over the hit list once it is written
// out).  No count kernel, no scan and no host round trip precede the emit.
// A run whose range would pass cap_edges writes nothing and raises flags[0];
// the host then reruns with room for the total the last run reports in
// flags[2..3].  flags[1] reports a look-back that gave up (1) or a run longer
// than SP_MAX_LEAVES (2); the host then takes the exact path.  A look-back
// only waits on lower runs.  Runs are handed out in increasing order, and a
// run a block holds but has not started is above every run its waves hold, so
// the lowest run without an aggregate is either in a phase A (which waits on
// nothing) or next in line at a block whose waves only wait on lower, already
// published runs: every wait ends (and is bounded anyway).  With lheavy, the
// leaves it flags were counted ahead by k_edges_heavy (totals in lhits,
// per-query counts in qcount): phase A takes their totals as they are, so a
// run publishes its aggregate without waiting on its heaviest leaves, and
// phase B only leaves their output positions for k_edges_heavy_fill: the
// two-pass path, for exactly the leaves the single pass cannot record.
// WIDE: 8-byte hits (any point count; half the list).
template <bool WIDE>
__global__ void __launch_bounds__(EW_WAVES * OCT_WAVE)
k_edges_single(const NodeRec* __restrict__ qleaves, const float4* __restrict__ qsc,
               const BuildParams* __restrict__ qprm, const float4* __restrict__ tsc, const int2* __restrict__ lists,
               const int32_t* __restrict__ lcount, int cap, float r_sq, float inv_r, int self_pair,
               uint32_t* __restrict__ qcount, uint32_t* __restrict__ qoff, const int32_t* __restrict__ wave_first,
               const int32_t* __restrict__ plan_waves, uint32_t* __restrict__ lhits,
               uint64_t* __restrict__ tile_state, uint32_t* __restrict__ ticket, uint32_t* __restrict__ flags,
               uint32_t* __restrict__ stats, uint64_t cap_edges, EdgeOut out, int index_base0,
               int leaves_per_run, int run_chunk, SetLookup sets, const uint8_t* __restrict__ lheavy,
               uint32_t* __restrict__ sig, const uint32_t* __restrict__ sig_src, int sig_words) {
    constexpr uint32_t HITS = WIDE ? SP_HIT_WORDS / 2 : SP_HIT_WORDS;
    __shared__ int32_t s_start[EW_WAVES][OCT_WAVE + 1];
    __shared__ int32_t s_first[EW_WAVES][OCT_WAVE];
    __shared__ alignas(16) uint32_t s_hits[EW_WAVES][SP_HIT_WORDS];
    __shared__ uint32_t s_cnt[EW_WAVES][SP_MAX_LEAVES];
    __shared__ int32_t s_qf[EW_WAVES][SP_MAX_LEAVES];
    __shared__ float4 s_stage[EW_WAVES][SP_STAGE_PTS];  // record tiles; positions of the filtered candidates
    __shared__ unsigned long long s_q;  // run hand-out: ticket << SP_Q_SHIFT | runs taken from it
    __shared__ uint32_t s_stat[2];
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / OCT_WAVE));  // scalar LDS bases
    const int lane = lane_id();
    uint32_t* hits = s_hits[wv];
    uint32_t* cntw = s_cnt[wv];
    int32_t* qfw = s_qf[wv];
    // phase B's FILL sweep lays its tile (4 KB) and records (2 KB) over the hit list
    const WaveLds L{s_start[wv], s_first[wv], reinterpret_cast<int32_t*>(s_stage[wv]), reinterpret_cast<float4*>(hits),
                    reinterpret_cast<int2*>(hits + 4 * EW_TILE * OCT_WAVE), (uint32_t)(EW_TILE * OCT_WAVE)};
    const int32_t nleaves = qprm->n_leaves;
    const int32_t nruns = wave_first ? *plan_waves : (nleaves + leaves_per_run - 1) / leaves_per_run;
    if (threadIdx.x == 0) s_q = (SP_Q_NONE << SP_Q_SHIFT) | (unsigned long long)run_chunk;
    __syncthreads();
    uint32_t n_recount = 0, n_record = 0;  // this wave's, over all its runs
    for (;;) {
        const int64_t run = take_run(&s_q, ticket, lane, run_chunk);
        if (run >= nruns) break;
        const int32_t w = (int32_t)run;  // this wave's run and output tile
        int32_t r_lo, r_hi;
        if (wave_first) {
            r_lo = wave_first[w];
            r_hi = wave_first[w + 1];
        } else {
            r_lo = w * leaves_per_run;
            r_hi = min(nleaves, r_lo + leaves_per_run);
        }
        r_lo = __builtin_amdgcn_readfirstlane(r_lo);
        r_hi = __builtin_amdgcn_readfirstlane(r_hi);
        bool bad = false;
        if (r_lo < 0 || r_hi > nleaves || r_hi < r_lo || r_hi - r_lo > SP_MAX_LEAVES) {
            bad = true;  // a malformed plan: nothing is swept, the host takes the exact path
            r_hi = r_lo = 0;
        }
        // phase A: one sweep per leaf, exact counts
        uint64_t recorded = 0;  // bit k: leaf r_lo + k sits in the hit list
        uint64_t counted = 0;   // bit k: leaf r_lo + k was counted by the heavy-leaf pass
        uint32_t hfill = 0, wtot = 0, nrec = 0;
        for (int32_t r = r_lo; r < r_hi; ++r) {
            const int k = r - r_lo;
            const bool pre = lheavy && __builtin_amdgcn_readfirstlane(lheavy[r]) != 0;
            uint32_t c;
            if (pre) {
                c = __builtin_amdgcn_readfirstlane(lhits[r]);  // counted by the heavy-leaf pass
                counted |= 1ull << k;
            } else {
                const NodeRec qa = qleaves[r];
                const int32_t qf = __builtin_amdgcn_readfirstlane(rec_a(qa));
                const int32_t qc = __builtin_amdgcn_readfirstlane(rec_b(qa));
                const int32_t ncand = __builtin_amdgcn_readfirstlane(min(lcount[r], cap));
                const int2* lst = lists + (int64_t)r * cap;
                const int2 lf0 = lst[lane < ncand ? lane : 0];
                const int total = __builtin_amdgcn_readfirstlane(wave_sum(lane < ncand ? lf0.y : 0));
                float lr_sq, linv_r;
                int libase;
                leaf_set_params(sets, qf, r_sq, inv_r, index_base0, lr_sq, linv_r, libase);
                c = REC_NOT_TILE;
                if (recordable_leaf(qc, ncand, total))
                    c = leaf_sweep_record<WIDE>(qf, qc, qa.lo, qa.hi, ncand, lf0, qsc, tsc, lr_sq, self_pair, L,
                                                s_stage[wv], lane, WIDE ? hits + 2 * hfill : hits + hfill, HITS - hfill, qcount);
                if (c != REC_NOT_TILE) {
                    if (c <= HITS - hfill) {  // else the list is full: qcount holds the per-query counts
                        recorded |= 1ull << k;
                        hfill += c;
                        ++nrec;
                    }
                } else {
                    c = leaf_pass<EDGE_COUNT>(qf, qc, qa.lo, qa.hi, ncand, lf0, lst, qsc, tsc, lr_sq, linv_r,
                                              self_pair, 0, qcount, nullptr, out, libase, L, lane);
                }
                if (lane == 0) {
                    qfw[k] = qf;
                    if (lhits) lhits[r] = c;
                }
            }
            if (lane == 0) cntw[k] = c;
            wtot += c;
        }
        n_recount += (uint32_t)(r_hi - r_lo) - nrec;
        n_record += nrec;
        // this run's output offset
        if (lane == 0) {
            if (bad) atomicOr(&flags[1], 2u);
            __hip_atomic_store(&tile_state[w], SP_FLAG_AGG | (uint64_t)wtot, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        uint32_t ok = 1;
        const uint64_t excl = wave_look_back(tile_state, w, lane, &ok);
        const uint64_t incl = excl + wtot;
        if (lane == 0) {
            __hip_atomic_store(&tile_state[w], SP_FLAG_PRE | incl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!ok) atomicOr(&flags[1], 1u);
            if (incl > cap_edges) atomicOr(&flags[0], 1u);
            if (w == nruns - 1)
                __hip_atomic_store((uint64_t*)(flags + 2), incl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (sig && w == nruns - 1) {  // wave-uniform
            // The total is known: hand it to the host now, with the call's flag
            // words (sig_src), so it can return while phase B still runs.  By
            // now every run has published its aggregate (this look-back read
            // them back to an inclusive prefix, which needed the same of the
            // runs below it), so no later look-back can wait or give up.
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
            for (int i = lane; i < sig_words; i += OCT_WAVE)
                sig[i] = __hip_atomic_load(&sig_src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            if (lane == 0)
                __hip_atomic_store(reinterpret_cast<unsigned long long*>(sig + SP_SIG_MARK), 1ull + incl,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        if (!ok || incl > cap_edges) continue;  // wave-uniform
        __builtin_amdgcn_wave_barrier();
        // phase B: recorded leaves straight from the hit list ...
        // (scalar offsets, so the emit addresses from scalar bases)
        uint32_t pos = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)excl);
        uint32_t hread = 0;
        for (int32_t r = r_lo; r < r_hi; ++r) {
            const int k = r - r_lo;
            const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)cntw[k]);
            if (recorded & (1ull << k)) {
                const int32_t qf = __builtin_amdgcn_readfirstlane(qfw[k]);
                float lr_sq, linv_r;
                int libase;
                leaf_set_params(sets, qf, r_sq, inv_r, index_base0, lr_sq, linv_r, libase);
                emit_hits<WIDE>(WIDE ? hits + 2 * hread : hits + hread, c, pos, qf, qsc, tsc, linv_r, out,
                                libase, lane);
                hread += c;
            }
            pos += c;
        }
        __builtin_amdgcn_wave_barrier();
        // ... then the others re-swept; their stage lies over the hit list.
        // A leaf the heavy-leaf pass counted only leaves its output position
        // (in qoff at its first query) for k_edges_heavy_fill, so no run
        // holds its wave for a long fill.
        pos = (uint32_t)excl;
        for (int32_t r = r_lo; r < r_hi; ++r) {
            const int k = r - r_lo;
            if (counted & (1ull << k)) {
                if (lane == 0) qoff[rec_a(qleaves[r])] = pos;
            } else if (!(recorded & (1ull << k))) {
                const NodeRec qa = qleaves[r];
                const int32_t qf = __builtin_amdgcn_readfirstlane(rec_a(qa));
                const int32_t qc = __builtin_amdgcn_readfirstlane(rec_b(qa));
                uint32_t run = pos;
                for (int qb = 0; qb < qc; qb += OCT_WAVE) {
                    const int nqb = min(OCT_WAVE, qc - qb);
                    const uint32_t c = lane < nqb ? qcount[qf + qb + lane] : 0u;  // this lane's, phase A
                    const uint32_t inc = (uint32_t)wave_incl_scan((int)c);
                    if (lane < nqb) qoff[qf + qb + lane] = run + inc - c;
                    run += (uint32_t)__builtin_amdgcn_readlane((int)inc, OCT_WAVE - 1);
                }
                const int32_t ncand = __builtin_amdgcn_readfirstlane(min(lcount[r], cap));
                const int2* lst = lists + (int64_t)r * cap;
                const int2 lf0 = lst[lane < ncand ? lane : 0];
                float lr_sq, linv_r;
                int libase;
                leaf_set_params(sets, qf, r_sq, inv_r, index_base0, lr_sq, linv_r, libase);
                leaf_pass<EDGE_FILL>(qf, qc, qa.lo, qa.hi, ncand, lf0, lst, qsc, tsc, lr_sq, linv_r, self_pair, 0,
                                     qcount, qoff, out, libase, L, lane);
            }
            pos += cntw[k];
        }
        __builtin_amdgcn_wave_barrier();  // the next run reuses the wave's LDS
    }
    // one pair of global atomics per block
    if (threadIdx.x == 0) s_stat[0] = s_stat[1] = 0;
    __syncthreads();
    if (lane == 0) {
        atomicAdd(&s_stat[0], n_recount);
        atomicAdd(&s_stat[1], n_record);
    }
    __syncthreads();
    if (threadIdx.x == 0 && (s_stat[0] | s_stat[1])) {
        atomicAdd(&stats[0], s_stat[0]);
        atomicAdd(&stats[1], s_stat[1]);
    }
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------- launchers

hipError_t launch_leaf_lists(const TreeView& q, const TreeView& t, float r_sq, int cap, int safe, int2* lists,
                             uint8_t* ldir, int32_t* lcount, uint32_t* flags, const SetLookup& sets,
                             uint8_t* lheavy, int32_t* heavy_list, uint32_t* n_heavy, hipStream_t s) {
    if ((lheavy != nullptr) != (heavy_list != nullptr) || (lheavy && !n_heavy)) return hipErrorInvalidValue;
    // the cooperative walk (eight lanes per query leaf), or the one-thread
    // safe walk when asked or when the packed stack entries cannot hold the
    // node ids (2^29 nodes)
    const bool coop = !safe && t.n_nodes < (1 << 29);
    const int64_t per = coop ? LL_COOP_GROUPS : LL_THREADS;  // query leaves per block
    const int64_t want = ((int64_t)q.n_leaves + per - 1) / per;
    const unsigned g = (unsigned)(want < 1 ? 1 : (want > 65535 ? 65535 : want));
#define OCT_LL_LAUNCH(K, D, H)                                                                                   \
    hipLaunchKernelGGL((K<D, H>), dim3(g), dim3(LL_THREADS), 0, s, q.leaves, q.params, t.nodes, t.params, r_sq, cap, \
                       lists, ldir, lcount, flags, sets, lheavy, heavy_list, n_heavy)
#define OCT_LL_PICK(K)                                    \
    if (ldir && lheavy) OCT_LL_LAUNCH(K, true, true);     \
    else if (ldir) OCT_LL_LAUNCH(K, true, false);         \
    else if (lheavy) OCT_LL_LAUNCH(K, false, true);       \
    else OCT_LL_LAUNCH(K, false, false)
    if (coop) OCT_LL_PICK(k_leaf_lists_coop);
    else OCT_LL_PICK(k_leaf_lists);
#undef OCT_LL_PICK
#undef OCT_LL_LAUNCH
    return hipGetLastError();
}

hipError_t launch_edges(int mode, const TreeView& q, const TreeView& t, const int2* lists,
                        const int32_t* lcount, int cap, float r_sq, float inv_r, int self_pair,
                        int by_receiver, uint32_t* qcount, const uint32_t* qoff, uint32_t* lhits, int64_t n_edges,
                        EdgeOut out, int index_base, hipStream_t s) {
    (void)n_edges;
    // about four leaves per wave, XCD-local contiguous ranges
    const int64_t want = ((int64_t)q.n_leaves + 4 * EW_WAVES - 1) / (4 * EW_WAVES);
    const unsigned g = (unsigned)(want < 1 ? 1 : (want > 16384 ? 16384 : want));
    if (mode == EDGE_COUNT)
        hipLaunchKernelGGL(k_edges<EDGE_COUNT>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord,
                           q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, by_receiver, qcount, qoff,
                           lhits, out, index_base);
    else if (mode == EDGE_FILL)
        hipLaunchKernelGGL(k_edges<EDGE_FILL>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord,
                           q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, by_receiver, qcount, qoff,
                           lhits, out, index_base);
    else
        hipLaunchKernelGGL(k_edges<EDGE_INDEX>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord,
                           q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, by_receiver, qcount, qoff,
                           lhits, out, index_base);
    return hipGetLastError();
}

// Upper bound of the runs (output tiles) of any plan: every run holds a leaf.
int64_t single_tiles(int64_t n_leaves) { return n_leaves + EW_WAVES; }

// tile_state (single_tiles words of 8 bytes) and ticket must be zero.
hipError_t launch_edges_single(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                               int cap, float r_sq, float inv_r, int self_pair, uint32_t* qcount, uint32_t* qoff,
                               const int32_t* wave_first, const int32_t* plan_waves, uint32_t* lhits,
                               uint64_t* tile_state, uint32_t* ticket, uint32_t* flags, uint32_t* stats,
                               uint64_t cap_edges, EdgeOut out, int index_base, int leaves_per_run,
                               int run_chunk, const SetLookup& sets, const uint8_t* lheavy, uint32_t* sig,
                               const uint32_t* sig_src, int sig_words, hipStream_t s) {
    if (sig && (wave_first || sig_words < 0 || sig_words > SP_SIG_MARK)) return hipErrorInvalidValue;
    if (leaves_per_run < 1 || leaves_per_run > SP_MAX_LEAVES) return hipErrorInvalidValue;
    if (run_chunk < 1 || run_chunk > EW_WAVES) return hipErrorInvalidValue;
    if (lheavy && !lhits) return hipErrorInvalidValue;  // the heavy totals live in lhits
    const int64_t runs = wave_first ? q.n_leaves : (q.n_leaves + leaves_per_run - 1) / leaves_per_run;
    const int64_t tickets = (runs + EW_WAVES - 1) / EW_WAVES;
    if (tickets <= 0) return hipSuccess;
    const unsigned g = (unsigned)std::min<int64_t>(tickets, 4096);  // blocks loop over tickets
    if (t.n < ((int64_t)1 << HIT_POS_BITS))
        hipLaunchKernelGGL(k_edges_single<false>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord,
                           q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, qcount, qoff, wave_first,
                           plan_waves, lhits, tile_state, ticket, flags, stats, cap_edges, out, index_base,
                           leaves_per_run, run_chunk, sets, lheavy, sig, sig_src, sig_words);
    else
        hipLaunchKernelGGL(k_edges_single<true>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord,
                           q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, qcount, qoff, wave_first,
                           plan_waves, lhits, tile_state, ticket, flags, stats, cap_edges, out, index_base,
                           leaves_per_run, run_chunk, sets, lheavy, sig, sig_src, sig_words);
    return hipGetLastError();
}

hipError_t launch_edges_heavy(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                              int cap, float r_sq, float inv_r, int self_pair, const int32_t* heavy_list,
                              const uint32_t* n_heavy, uint32_t* qcount, uint32_t* lhits, const SetLookup& sets,
                              hipStream_t s) {
    // a fixed grid (the count is on the device); idle waves exit at once
    const int64_t want = ((int64_t)q.n_leaves + EW_WAVES - 1) / EW_WAVES;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, 2048));
    hipLaunchKernelGGL(k_edges_heavy, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord, q.params,
                       t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, heavy_list, n_heavy, qcount, lhits,
                       sets);
    return hipGetLastError();
}

// dst = *src with n_leaves clamped to rows; *over = 1 when it was (else 0);
// raw[0] = *src and raw[1] = *tsrc, for the host to read back with the flags.
__global__ void __launch_bounds__(256)
k_clamp_leaves(const BuildParams* __restrict__ src, const BuildParams* __restrict__ tsrc,
               BuildParams* __restrict__ dst, int64_t rows, uint32_t* __restrict__ over,
               BuildParams* __restrict__ raw, uint32_t* __restrict__ z0, int64_t n0, uint32_t* __restrict__ z1,
               int64_t n1) {
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = i0; i < n0; i += step) z0[i] = 0u;
    for (int64_t i = i0; i < n1; i += step) z1[i] = 0u;
    if (i0 != 0) return;
    BuildParams p = *src;
    raw[0] = p;
    raw[1] = *tsrc;
    const bool big = (int64_t)p.n_leaves > rows;
    if (big) p.n_leaves = (int32_t)rows;
    *dst = p;
    *over = big ? 1u : 0u;
}

hipError_t launch_clamp_leaves(const BuildParams* src, const BuildParams* tsrc, BuildParams* dst, int64_t rows,
                               uint32_t* over, BuildParams* raw, uint32_t* z0, int64_t n0, uint32_t* z1, int64_t n1,
                               hipStream_t s) {
    const int64_t n = std::max<int64_t>(std::max<int64_t>(n0, n1), 1);
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 512);
    hipLaunchKernelGGL(k_clamp_leaves, dim3(g), dim3(256), 0, s, src, tsrc, dst, rows, over, raw, z0, n0, z1, n1);
    return hipGetLastError();
}

hipError_t launch_edges_heavy_fill(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                                   int cap, float r_sq, float inv_r, int self_pair, const int32_t* heavy_list,
                                   const uint32_t* n_heavy, uint32_t* qcount, uint32_t* qoff, const uint32_t* flags,
                                   EdgeOut out, int index_base, const SetLookup& sets, hipStream_t s) {
    const int64_t want = ((int64_t)q.n_leaves + EW_WAVES - 1) / EW_WAVES;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, 2048));
    hipLaunchKernelGGL(k_edges_heavy_fill, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord, q.params,
                       t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, heavy_list, n_heavy, qcount, qoff,
                       flags, out, index_base, sets);
    return hipGetLastError();
}

hipError_t launch_edges_dense_fill(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                                   int cap, float r_sq, float inv_r, int self_pair, const int32_t* heavy_list,
                                   const uint32_t* n_heavy, uint32_t* qcount, const uint32_t* qoff,
                                   const uint32_t* flags, EdgeOut out, int index_base, const SetLookup& sets,
                                   hipStream_t s) {
    // a block per leaf row (the heavy count is on the device; blocks past it
    // exit at once): the dispatcher then hands each free slot the next heavy
    // leaf, which balances their uneven sizes (a grid of 2,048 blocks looping
    // over the leaves was 20% slower at the clustered workload)
    const unsigned g = (unsigned)std::max<int64_t>(1, q.n_leaves);
    hipLaunchKernelGGL(k_edges_dense<EDGE_FILL>, dim3(g), dim3(DF_THREADS), 0, s, q.leaves, q.scoord, q.params,
                       t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, heavy_list, n_heavy, qcount, qoff,
                       nullptr, flags, out, index_base, sets);
    return hipGetLastError();
}

hipError_t launch_edges_dense_count(const TreeView& q, const TreeView& t, const int2* lists, const int32_t* lcount,
                                    int cap, float r_sq, float inv_r, int self_pair, const int32_t* heavy_list,
                                    const uint32_t* n_heavy, uint32_t* qcount, uint32_t* lhits,
                                    const SetLookup& sets, hipStream_t s) {
    const unsigned g = (unsigned)std::max<int64_t>(1, q.n_leaves);  // as the fill
    hipLaunchKernelGGL(k_edges_dense<EDGE_COUNT>, dim3(g), dim3(DF_THREADS), 0, s, q.leaves, q.scoord, q.params,
                       t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, heavy_list, n_heavy, qcount, nullptr,
                       lhits, nullptr, EdgeOut{}, 0, sets);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace oct
