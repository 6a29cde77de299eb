// build_edges on the GPU: count, scan and emit of flat-COO edges
// (reference: src/cpu/edges_cpu.jl:172-247, ext/OctopusCUDAExt.jl:262-415).
//
// Work is blocked by query leaf.  Stage 1 gives every query leaf the list of
// target leaves whose box lies within r of its box (FP32 box-box gap, a
// monotone lower bound of every point distance between the two boxes, so no
// edge is lost).  Stage 2 runs one wavefront per query leaf: lanes take
// candidate target points (coalesced from the Morton-sorted copy), the wave
// loops over the leaf's queries, and each (query, 64 candidates) step is one
// ballot.  The predicate is the reference's FP32 test
//   d2 = (dx*dx + dy*dy) + dz*dz <= r*r,
// so the edge set is bit-identical to the CPU path.
#include "internal.hpp"

namespace oct {

constexpr int STACK_DEPTH = 1 + 7 * OCT_MAX_LEVEL;  // 148 (src/cpu/build_cpu.jl:25-43)

// ---------------------------------------------------------------- stage 1

// One thread per query leaf: depth-first walk of the target tree from its
// branch root, testing children before pushing them; leaves that pass are
// recorded as (first sorted position, count).
__global__ void __launch_bounds__(256)
k_leaf_lists(const NodeRec* __restrict__ qnodes, const int32_t* __restrict__ qleaf_ids,
             const BuildParams* __restrict__ qprm, const NodeRec* __restrict__ tnodes,
             const BuildParams* __restrict__ tprm, float r_sq, int cap, int2* __restrict__ lists,
             int32_t* __restrict__ lcount, uint32_t* __restrict__ overflow) {
    const int32_t r = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= qprm->n_leaves) return;
    const NodeRec qa = qnodes[qleaf_ids[r]];
    int2* out = lists + (int64_t)r * cap;
    int cnt = 0;
    const int32_t root = tprm->branch_root;
    const NodeRec rr = tnodes[root];
    if (box_box_d2(qa.lo, qa.hi, rr.lo, rr.hi) <= r_sq) {
        if (rec_b(rr) > 0) {
            if (cnt < cap) out[cnt] = make_int2(rec_a(rr), rec_b(rr));
            ++cnt;
        } else {
            int32_t stack[STACK_DEPTH];
            int sp = 0;
            stack[sp++] = root;
            while (sp > 0) {
                const NodeRec nd = tnodes[stack[--sp]];
                const int c0 = rec_a(nd), nc = -rec_b(nd);
                for (int c = 0; c < nc; ++c) {
                    const NodeRec ch = tnodes[c0 + c];
                    if (box_box_d2(qa.lo, qa.hi, ch.lo, ch.hi) > r_sq) continue;
                    if (rec_b(ch) > 0) {
                        if (cnt < cap) out[cnt] = make_int2(rec_a(ch), rec_b(ch));
                        ++cnt;
                    } else {
                        stack[sp++] = c0 + c;
                    }
                }
            }
        }
    }
    lcount[r] = cnt;
    if (cnt > cap) atomicMax(overflow, (uint32_t)cnt);  // the host regrows cap to this
}

// ---------------------------------------------------------------- stage 2

enum { EDGE_COUNT = 0, EDGE_FILL = 1 };



constexpr int EW_WAVES = 4;  // waves per block, one query leaf each

template <int MODE>
__global__ void __launch_bounds__(EW_WAVES * OCT_WAVE)
k_edges(const NodeRec* __restrict__ qnodes, const int32_t* __restrict__ qleaf_ids,
        const float4* __restrict__ qsc, const BuildParams* __restrict__ qprm,
        const float4* __restrict__ tsc, const int2* __restrict__ lists,
        const int32_t* __restrict__ lcount, int cap, float r_sq, float inv_r, int self_pair,
        int by_receiver, uint32_t* __restrict__ qcount, const uint32_t* __restrict__ qoff,
        EdgeOut out, int index_base) {
    __shared__ int32_t s_start[EW_WAVES][OCT_WAVE + 1];
    __shared__ int32_t s_first[EW_WAVES][OCT_WAVE];
    const int wv = threadIdx.x / OCT_WAVE;
    const int lane = lane_id();
    const int32_t r = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * EW_WAVES + wv));
    if (r >= qprm->n_leaves) return;
    const NodeRec qa = qnodes[qleaf_ids[r]];
    const int32_t qf = __builtin_amdgcn_readfirstlane(rec_a(qa));
    const int32_t qc = __builtin_amdgcn_readfirstlane(rec_b(qa));
    const int32_t ncand = __builtin_amdgcn_readfirstlane(min(lcount[r], cap));
    const int2* lst = lists + (int64_t)r * cap;

    for (int qb = 0; qb < qc; qb += OCT_WAVE) {
        const int nqb = min(OCT_WAVE, qc - qb);
        // per-query state lives in lane (qi - qb)
        uint32_t my = 0;
        if (MODE == EDGE_FILL && lane < nqb) {
            const float4 p = qsc[qf + qb + lane];
            const int32_t slot = by_receiver ? __float_as_int(p.w) : qf + qb + lane;
            my = qoff[slot];
        }
        for (int g0 = 0; g0 < ncand; g0 += OCT_WAVE) {
            int2 lf = make_int2(0, 0);
            if (g0 + lane < ncand) lf = lst[g0 + lane];
            const int incl = wave_incl_scan(lf.y);
            const int total = __builtin_amdgcn_readlane(incl, OCT_WAVE - 1);
            s_start[wv][lane + 1] = incl;
            if (lane == 0) s_start[wv][0] = 0;
            s_first[wv][lane] = lf.x;
            __builtin_amdgcn_wave_barrier();
            for (int c0 = 0; c0 < total; c0 += OCT_WAVE) {
                const int ci = c0 + lane;
                const bool valid = ci < total;
                float4 tp = make_float4(0.f, 0.f, 0.f, 0.f);
                if (valid) {
                    // owner leaf: last o with s_start[o] <= ci
                    int lo = 0, hi = OCT_WAVE;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (s_start[wv][mid] <= ci) lo = mid;
                        else hi = mid;
                    }
                    tp = tsc[s_first[wv][lo] + (ci - s_start[wv][lo])];
                }
                const int32_t j = __float_as_int(tp.w);
                for (int qi = 0; qi < nqb; ++qi) {
                    const float4 qp = qsc[qf + qb + qi];
                    const int32_t i = __float_as_int(qp.w);
                    const float dx = qp.x - tp.x;
                    const float dy = qp.y - tp.y;
                    const float dz = qp.z - tp.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    const bool pass = valid && d2 <= r_sq && !(self_pair && j == i);
                    const uint64_t m = ballot(pass);
                    const uint32_t pc = (uint32_t)__popcll(m);
                    if (MODE == EDGE_COUNT) {
                        if (lane == qi) my += pc;
                    } else {
                        const uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)my, qi);
                        if (pass) {
                            const int64_t e = (int64_t)base + (uint32_t)mbcnt(m);
                            out.senders[e] = j + index_base;
                            out.receivers[e] = i + index_base;
                            out.rdisp[3 * e + 0] = dx * inv_r;
                            out.rdisp[3 * e + 1] = dy * inv_r;
                            out.rdisp[3 * e + 2] = dz * inv_r;
                            out.rdist[e] = sqrtf(d2) * inv_r;
                        }
                        if (lane == qi) my += pc;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (MODE == EDGE_COUNT && lane < nqb) {
            const float4 p = qsc[qf + qb + lane];
            const int32_t slot = by_receiver ? __float_as_int(p.w) : qf + qb + lane;
            qcount[slot] = my;
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------- launchers

hipError_t launch_leaf_lists(const TreeView& q, const TreeView& t, float r_sq, int cap, int2* lists,
                             int32_t* lcount, uint32_t* overflow, hipStream_t s) {
    if (q.n_leaves <= 0) return hipSuccess;
    const unsigned g = (unsigned)((q.n_leaves + 255) / 256);
    hipLaunchKernelGGL(k_leaf_lists, dim3(g), dim3(256), 0, s, q.nodes, q.leaf_ids, q.params, t.nodes,
                       t.params, r_sq, cap, lists, lcount, overflow);
    return hipGetLastError();
}

hipError_t launch_edges(int mode, const TreeView& q, const TreeView& t, const int2* lists,
                        const int32_t* lcount, int cap, float r_sq, float inv_r, int self_pair,
                        int by_receiver, uint32_t* qcount, const uint32_t* qoff, EdgeOut out,
                        int index_base, hipStream_t s) {
    if (q.n_leaves <= 0) return hipSuccess;
    const unsigned g = (unsigned)((q.n_leaves + EW_WAVES - 1) / EW_WAVES);
    if (mode == EDGE_COUNT)
        hipLaunchKernelGGL(k_edges<EDGE_COUNT>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.nodes, q.leaf_ids,
                           q.scoord, q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair,
                           by_receiver, qcount, qoff, out, index_base);
    else
        hipLaunchKernelGGL(k_edges<EDGE_FILL>, dim3(g), dim3(EW_WAVES * OCT_WAVE), 0, s, q.nodes, q.leaf_ids,
                           q.scoord, q.params, t.scoord, lists, lcount, cap, r_sq, inv_r, self_pair,
                           by_receiver, qcount, qoff, out, index_base);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace oct
