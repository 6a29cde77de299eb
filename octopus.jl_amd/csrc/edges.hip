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
#include "internal.hpp"

namespace oct {

// ---------------------------------------------------------------- stage 1

static_assert(1 + 7 * OCT_MAX_LEVEL <= 192, "LIFO walk bound (src/cpu/build_cpu.jl:25-43) must fit the stack");
constexpr int LL_STACK = 192;                              // room for multi-pops
constexpr int LL_POP = 2;

// One thread per query leaf: depth-first walk of the target tree from its
// branch root.  A stack entry is the child range of a node that passed the
// test, so each round costs one dependent load; up to LL_POP entries are
// expanded per round to overlap their loads.  Popping one entry at a time is
// plain LIFO depth-first order, whose stack never exceeds 1 + 7 * 21 entries
// (the reference's bound); multi-pop rounds are only taken while the stack
// has room for all their pushes.  Every push is bounds-checked: a walk that
// would still run out of room stops, and the leaf is flagged so the host
// reruns the lists with safe = 1 (single pops, inside the proven bound).
__global__ void __launch_bounds__(256)
k_leaf_lists(const NodeRec* __restrict__ qleaves,
             const BuildParams* __restrict__ qprm, const NodeRec* __restrict__ tnodes,
             const BuildParams* __restrict__ tprm, float r_sq, int cap, int safe, int2* __restrict__ lists,
             int32_t* __restrict__ lcount, uint32_t* __restrict__ flags) {
    const int32_t nleaves = qprm->n_leaves;
    const int32_t root = tprm->branch_root;
    const int32_t lb = (int32_t)xcd_block(blockIdx.x, gridDim.x);
    for (int32_t r = lb * (int32_t)blockDim.x + (int32_t)threadIdx.x; r < nleaves;
         r += (int32_t)(gridDim.x * blockDim.x)) {
        const NodeRec qa = qleaves[r];
        int2* out = lists + (int64_t)r * cap;
        int cnt = 0;
        bool stack_full = false;
        const NodeRec rr = tnodes[root];
        if (box_box_d2(qa.lo, qa.hi, rr.lo, rr.hi) <= r_sq) {
            if (rec_b(rr) > 0) {
                if (cnt < cap) out[cnt] = make_int2(rec_a(rr), rec_b(rr));
                ++cnt;
            } else {
                int2 stack[LL_STACK];
                int sp = 0;
                stack[sp++] = make_int2(rec_a(rr), -rec_b(rr));
                while (sp > 0 && !stack_full) {
                    int k = 1;
                    if (!safe) {
                        k = sp < LL_POP ? sp : LL_POP;
                        while (k > 1 && sp - k + 8 * k > LL_STACK) --k;  // pushes must fit
                    }
                    int2 ent[LL_POP];
#pragma unroll
                    for (int i = 0; i < LL_POP; ++i) ent[i] = i < k ? stack[sp - 1 - i] : make_int2(0, 0);
                    sp -= k;
                    NodeRec ch[LL_POP][8];
#pragma unroll
                    for (int i = 0; i < LL_POP; ++i)
#pragma unroll
                        for (int c = 0; c < 8; ++c)
                            if (c < ent[i].y) ch[i][c] = tnodes[ent[i].x + c];
#pragma unroll
                    for (int i = 0; i < LL_POP; ++i)
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            if (c >= ent[i].y) continue;
                            if (box_box_d2(qa.lo, qa.hi, ch[i][c].lo, ch[i][c].hi) > r_sq) continue;
                            if (rec_b(ch[i][c]) > 0) {
                                if (cnt < cap) out[cnt] = make_int2(rec_a(ch[i][c]), rec_b(ch[i][c]));
                                ++cnt;
                            } else if (sp < LL_STACK) {
                                stack[sp++] = make_int2(rec_a(ch[i][c]), -rec_b(ch[i][c]));
                            } else {
                                stack_full = true;
                            }
                        }
                }
            }
        }
        lcount[r] = cnt;
        if (cnt > cap) atomicMax(&flags[0], (uint32_t)cnt);  // the host regrows cap to this
        if (stack_full) atomicOr(&flags[1], 1u);             // the host reruns with safe = 1
    }
}

// ---------------------------------------------------------------- stage 2

constexpr int EW_WAVES = 4;     // waves per block, each owns one query leaf at a time
constexpr int EW_HITS = 1024;   // hit list entries per wave (8 KB of LDS)
constexpr int EW_TILE = 4;      // candidate chunks of 64 held in registers per tile

// Candidate enumeration shared by the passes: the candidate leaves of a query
// leaf, 64 at a time, prefix-summed into LDS; points are then taken 4 x 64 at
// a time into registers (coords + sorted position).
struct CandTile {
    float4 p[EW_TILE];
    int32_t pos[EW_TILE];
};

__device__ inline void load_tile(CandTile& t, int t0, int total, const int32_t* s_start, const int32_t* s_first,
                                 const float4* __restrict__ tsc, int lane) {
#pragma unroll
    for (int u = 0; u < EW_TILE; ++u) {
        const int ci = t0 + u * OCT_WAVE + lane;
        t.p[u] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
        t.pos[u] = -1;
        if (ci < total) {
            int lo = 0, hi = OCT_WAVE;  // owner leaf: last o with s_start[o] <= ci
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_start[mid] <= ci) lo = mid;
                else hi = mid;
            }
            t.pos[u] = s_first[lo] + (ci - s_start[lo]);
            t.p[u] = tsc[t.pos[u]];
        }
    }
}

// The reference predicate for query (qx, qy, qz, i) against candidate p.
__device__ inline bool edge_test(float qx, float qy, float qz, int32_t i, float4 p, float r_sq, int self_pair) {
    const int32_t j = __float_as_int(p.w);
    const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    return j >= 0 && d2 <= r_sq && !(self_pair && j == i);
}

// Writes edge e of query q against target point p (features exactly as the
// reference: rdisp = (q - t) * inv_r, rdist = sqrt(d2) * inv_r).
__device__ inline void emit_edge(EdgeOut out, uint32_t e, float4 q, float4 p, float inv_r, int index_base) {
    const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    out.senders[e] = __float_as_int(p.w) + index_base;
    out.receivers[e] = __float_as_int(q.w) + index_base;
    out.rdisp[3 * (int64_t)e + 0] = dx * inv_r;
    out.rdisp[3 * (int64_t)e + 1] = dy * inv_r;
    out.rdisp[3 * (int64_t)e + 2] = dz * inv_r;
    out.rdist[e] = sqrtf(d2) * inv_r;
}

// One wave per query leaf (contiguous, XCD-local ranges of leaves).
//   COUNT: per-query neighbour counts (qcount).
//   FILL:  with the exact counts and offsets, queries are split into groups
//          whose edges fit the hit list; phase 1 records each hit (query,
//          target position) at its final place in the list, phase 2 has all
//          lanes compute the features of consecutive hits and write them with
//          contiguous stores.  A lone query with more hits than the list holds
//          writes directly.
template <int MODE>
__global__ void __launch_bounds__(EW_WAVES * OCT_WAVE)
k_edges(const NodeRec* __restrict__ qleaves, const float4* __restrict__ qsc, const BuildParams* __restrict__ qprm,
        const float4* __restrict__ tsc, const int2* __restrict__ lists, const int32_t* __restrict__ lcount, int cap,
        float r_sq, float inv_r, int self_pair, int by_receiver, uint32_t* __restrict__ qcount,
        const uint32_t* __restrict__ qoff, EdgeOut out, int index_base) {
    __shared__ int32_t s_start[EW_WAVES][OCT_WAVE + 1];
    __shared__ int32_t s_first[EW_WAVES][OCT_WAVE];
    constexpr int HITS = MODE == EDGE_FILL ? EW_HITS : 1;
    constexpr int QS = MODE == EDGE_FILL ? OCT_WAVE : 1;
    __shared__ int2 s_hit[EW_WAVES][HITS];
    __shared__ float4 s_q[EW_WAVES][QS];
    __shared__ uint32_t s_qoff[EW_WAVES][QS];
    const int wv = threadIdx.x / OCT_WAVE;
    const int lane = lane_id();
    const int32_t nleaves = qprm->n_leaves;
    const int32_t nwaves = (int32_t)(gridDim.x * EW_WAVES);
    const int32_t w = (int32_t)xcd_block(blockIdx.x, gridDim.x) * EW_WAVES + wv;
    const int32_t per = (nleaves + nwaves - 1) / nwaves;
    const int32_t r_lo = __builtin_amdgcn_readfirstlane(w * per);
    const int32_t r_hi = min(nleaves, r_lo + per);
    // leaf metadata of the next leaf is loaded while the current one runs
    int2 nx_meta = make_int2(0, 0), nx_lf = make_int2(0, 0);
    int32_t nx_cnt = 0;
    if (r_lo < r_hi) {
        const NodeRec qa = qleaves[r_lo];
        nx_meta = make_int2(rec_a(qa), rec_b(qa));
        nx_cnt = lcount[r_lo];
        if (lane < cap) nx_lf = lists[(int64_t)r_lo * cap + lane];
    }
    for (int32_t r = r_lo; r < r_hi; ++r) {
        const int32_t qf = __builtin_amdgcn_readfirstlane(nx_meta.x);
        const int32_t qc = __builtin_amdgcn_readfirstlane(nx_meta.y);
        const int32_t ncand = __builtin_amdgcn_readfirstlane(min(nx_cnt, cap));
        const int2 lf0 = nx_lf;
        if (r + 1 < r_hi) {
            const NodeRec qa = qleaves[r + 1];
            nx_meta = make_int2(rec_a(qa), rec_b(qa));
            nx_cnt = lcount[r + 1];
            if (lane < cap) nx_lf = lists[(int64_t)(r + 1) * cap + lane];
        }
        const int2* lst = lists + (int64_t)r * cap;
        for (int qb = 0; qb < qc; qb += OCT_WAVE) {
            const int nqb = min(OCT_WAVE, qc - qb);
            float4 myq = make_float4(0.f, 0.f, 0.f, 0.f);  // query qb + lane
            if (lane < nqb) myq = qsc[qf + qb + lane];
            const int32_t slot = by_receiver ? __float_as_int(myq.w) : qf + qb + lane;
            // groups of consecutive queries (COUNT: all nqb at once)
            uint32_t c_lane = 0, o_lane = 0;
            if (MODE == EDGE_FILL && lane < nqb) {
                c_lane = qcount[slot];
                o_lane = qoff[slot];
                s_q[wv][lane] = myq;
            }
            const uint32_t incl = (uint32_t)wave_incl_scan((int)c_lane);
            int g = 0;
            while (g < nqb) {
                int gend = nqb;
                uint32_t gtot = 0;
                bool direct = false;
                uint32_t run_lane = 0;  // COUNT: count; FILL: next hit slot of this lane's query
                if (MODE == EDGE_FILL) {
                    const uint32_t base_incl = g > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)incl, g - 1) : 0u;
                    const uint64_t fit = ballot(lane >= g && lane < nqb && incl - base_incl <= (uint32_t)EW_HITS);
                    gend = fit ? 64 - __clzll((long long)fit) : g + 1;
                    gtot = (uint32_t)__builtin_amdgcn_readlane((int)incl, gend - 1) - base_incl;
                    direct = gtot > (uint32_t)EW_HITS;
                    const uint32_t lstart = incl - c_lane - base_incl;
                    run_lane = lstart;
                    if (lane >= g && lane < gend) s_qoff[wv][lane] = o_lane - lstart;
                }
                for (int g0 = 0; g0 < ncand; g0 += OCT_WAVE) {
                    int2 lf = make_int2(0, 0);
                    if (g0 + lane < ncand) lf = (g0 == 0) ? lf0 : lst[g0 + lane];
                    const int cincl = wave_incl_scan(lf.y);
                    const int total = __builtin_amdgcn_readlane(cincl, OCT_WAVE - 1);
                    s_start[wv][lane + 1] = cincl;
                    if (lane == 0) s_start[wv][0] = 0;
                    s_first[wv][lane] = lf.x;
                    __builtin_amdgcn_wave_barrier();
                    for (int t0 = 0; t0 < total; t0 += EW_TILE * OCT_WAVE) {
                        CandTile ct;
                        load_tile(ct, t0, total, s_start[wv], s_first[wv], tsc, lane);
                        const int nu = min(EW_TILE, (total - t0 + OCT_WAVE - 1) / OCT_WAVE);
                        for (int qi = g; qi < gend; ++qi) {
                            const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.x), qi));
                            const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.y), qi));
                            const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myq.z), qi));
                            const int32_t i = __builtin_amdgcn_readlane(__float_as_int(myq.w), qi);
                            uint32_t run = (uint32_t)__builtin_amdgcn_readlane((int)run_lane, qi);
                            const uint32_t run0 = run;
#pragma unroll
                            for (int u = 0; u < EW_TILE; ++u) {
                                if (u >= nu) break;
                                const bool pass = edge_test(qx, qy, qz, i, ct.p[u], r_sq, self_pair);
                                const uint64_t m = ballot(pass);
                                if (MODE == EDGE_FILL && pass) {
                                    const uint32_t hpos = run + (uint32_t)mbcnt(m);
                                    if (!direct) {
                                        s_hit[wv][hpos] = make_int2(qi, ct.pos[u]);
                                    } else {
                                        const uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)o_lane, qi);
                                        const float4 q4 = make_float4(qx, qy, qz, __int_as_float(i));
                                        emit_edge(out, o + hpos, q4, ct.p[u], inv_r, index_base);
                                    }
                                }
                                run += (uint32_t)__popcll(m);
                            }
                            if (lane == qi) run_lane += run - run0;
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                if (MODE == EDGE_COUNT) {
                    if (lane < nqb) qcount[slot] = run_lane;
                } else if (!direct) {
                    // phase 2: dense, position-ordered emission of the group's hits
                    __builtin_amdgcn_wave_barrier();
                    for (uint32_t t = lane; t < gtot; t += OCT_WAVE) {
                        const int2 h = s_hit[wv][t];
                        emit_edge(out, s_qoff[wv][h.x] + t, s_q[wv][h.x], tsc[h.y], inv_r, index_base);
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                g = gend;
            }
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------- launchers

hipError_t launch_leaf_lists(const TreeView& q, const TreeView& t, float r_sq, int cap, int safe, int2* lists,
                             int32_t* lcount, uint32_t* flags, hipStream_t s) {
    const int64_t want = ((int64_t)q.n_leaves + 255) / 256;
    const unsigned g = (unsigned)(want < 1 ? 1 : (want > 65535 ? 65535 : want));
    hipLaunchKernelGGL(k_leaf_lists, dim3(g), dim3(256), 0, s, q.leaves, q.params, t.nodes,
                       t.params, r_sq, cap, safe, lists, lcount, flags);
    return hipGetLastError();
}

hipError_t launch_edges(int mode, const TreeView& q, const TreeView& t, const int2* lists,
                        const int32_t* lcount, int cap, float r_sq, float inv_r, int self_pair,
                        int by_receiver, uint32_t* qcount, const uint32_t* qoff, int64_t n_edges, EdgeOut out,
                        int index_base, hipStream_t s) {
    (void)n_edges;
    // about four leaves per wave, XCD-local contiguous ranges
    const int64_t want = ((int64_t)q.n_leaves + 4 * EW_WAVES - 1) / (4 * EW_WAVES);
    const unsigned g = (unsigned)(want < 1 ? 1 : (want > 16384 ? 16384 : want));
    if (mode == EDGE_COUNT)
        hipLaunchKernelGGL(k_edges<EDGE_COUNT>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord,
                           q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, by_receiver, qcount, qoff,
                           out, index_base);
    else
        hipLaunchKernelGGL(k_edges<EDGE_FILL>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.leaves, q.scoord,
                           q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair, by_receiver, qcount, qoff,
                           out, index_base);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace oct
