// Learned per-leaf hit-count predictor: features, MLP inference, training
// (forward, backward, Adam) and the emission plan built from the predictions.
// The reference has no counterpart; it sizes and schedules the single-pass
// emit (edges.hip), whose results never depend on it: every leaf is counted
// exactly before anything is written, and a leaf the plan misjudges is
// recounted.
//
// Per query leaf, 29 features (all FP32):
//   0      log1p(point count)
//   1      tree level / 21
//   2      log2(largest box side / r), floored at -20
//   3..28  log1p(points of the adjacent candidate leaves) in each of the 26
//          directions (code 13, the same position, is left out)
// Network: 29 -> 64 -> 64 -> 1, ReLU, output y = log1p(predicted hits).
// Inputs are standardised by mean / std stored with the weights.
#include "internal.hpp"

namespace oct {

// ------------------------------------------------------------ features

constexpr int FT_THREADS = 128;

// One thread per query leaf (many leaves per wave keep the chain of
// dependent loads overlapped); the 26 direction bins live in LDS.
__global__ void __launch_bounds__(FT_THREADS)
k_leaf_features(const NodeRec* __restrict__ qleaves, const int32_t* __restrict__ leaf_ids,
                const uint8_t* __restrict__ nlevel, int64_t n_leaves, const int2* __restrict__ lists,
                const uint8_t* __restrict__ ldir, const int32_t* __restrict__ lcount, int cap, float r,
                float* __restrict__ feat) {
    __shared__ uint32_t acc[MLP_ADJ][FT_THREADS];
    const int tid = threadIdx.x;
    const int64_t l = (int64_t)blockIdx.x * FT_THREADS + tid;
    if (l >= n_leaves) return;
#pragma unroll
    for (int b = 0; b < MLP_ADJ; ++b) acc[b][tid] = 0u;
    const int n = min(lcount[l], cap);
    const int2* lst = lists + l * cap;
    const uint8_t* dir = ldir + l * cap;
    for (int e = 0; e < n; ++e) {
        const int d = dir[e];
        if (d == 13 || d > 26) continue;
        acc[d < 13 ? d : d - 1][tid] += (uint32_t)lst[e].y;
    }
    const NodeRec qa = qleaves[l];
    const float ext = fmaxf(fmaxf(qa.hi.x - qa.lo.x, qa.hi.y - qa.lo.y), qa.hi.z - qa.lo.z);
    float* f = feat + l * MLP_IN;
    f[0] = log1pf((float)rec_b(qa));
    f[1] = (float)nlevel[leaf_ids[l]] / 21.f;
    f[2] = fmaxf(log2f(fmaxf(ext / r, 1e-30f)), -20.f);
#pragma unroll
    for (int b = 0; b < MLP_ADJ; ++b) f[3 + b] = log1pf((float)acc[b][tid]);
}

// ------------------------------------------------------------ inference

constexpr int MLP_THREADS = 128;

__device__ inline void load_params(float* sp, const float* __restrict__ P) {
    for (int i = threadIdx.x; i < MLP_PARAMS; i += blockDim.x) sp[i] = P[i];
    __syncthreads();
}

// first layer on one standardised row; h1 after ReLU
__device__ inline void mlp_layer1(const float* sp, const float* x, float* xn, float* h1) {
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) h1[j] = sp[MLP_B1 + j];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) {
        xn[i] = (x[i] - sp[MLP_MU + i]) / sp[MLP_SD + i];
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) h1[j] = fmaf(xn[i], sp[MLP_W1 + i * MLP_H + j], h1[j]);
    }
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) h1[j] = fmaxf(h1[j], 0.f);
}

// second layer pre-activation
__device__ inline void mlp_layer2(const float* sp, const float* h1, float* h2) {
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) h2[j] = sp[MLP_B2 + j];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i)
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) h2[j] = fmaf(h1[i], sp[MLP_W2 + i * MLP_H + j], h2[j]);
}

constexpr int PLAN_CHUNK = OCT_WAVE;  // leaves the plan packs per chunk: one wave of the inference

__device__ inline uint32_t plan_cost(float p, uint32_t budget) {
    if (!(p >= 0.f)) return budget;
    const float h = ceilf(p * 1.0625f + 2.f);
    return h >= (float)budget ? budget : (uint32_t)h;
}

// y[l] = network output (log1p of the predicted hits); hits[l] (optional) =
// expm1(y) clamped at 0.  With runs != null each wave (64 consecutive leaves,
// one plan chunk) also counts the runs of its greedy packing into runs[chunk]
// and the rounded-up sum of its predictions into psum[chunk] (see k_plan_write;
// lheavy as there).
__global__ void __launch_bounds__(MLP_THREADS)
k_mlp_infer(const float* __restrict__ X, int64_t n, const float* __restrict__ P, float* __restrict__ y,
            float* __restrict__ hits, uint32_t* __restrict__ runs, uint32_t* __restrict__ psum, uint32_t budget,
            int max_leaves, const uint8_t* __restrict__ lheavy) {
    __shared__ float sp[MLP_PARAMS];
    load_params(sp, P);
    const int64_t l = (int64_t)blockIdx.x * MLP_THREADS + threadIdx.x;
    const bool valid = l < n;
    float x[MLP_IN], xn[MLP_IN], h1[MLP_H], h2[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) x[i] = valid ? X[l * MLP_IN + i] : 0.f;
    mlp_layer1(sp, x, xn, h1);
    mlp_layer2(sp, h1, h2);
    float out = sp[MLP_B3];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) out = fmaf(fmaxf(h2[j], 0.f), sp[MLP_W3 + j], out);
    const float hv = fmaxf(expm1f(fminf(out, 40.f)), 0.f);
    if (valid && y) y[l] = out;
    if (valid && hits) hits[l] = hv;
    if (!runs) return;
    const int64_t chunk = l / PLAN_CHUNK;
    const int64_t l0 = chunk * PLAN_CHUNK;
    if (l0 >= n) return;  // wave-uniform
    const int cnt = (int)(n - l0 < PLAN_CHUNK ? n - l0 : PLAN_CHUNK);
    const uint32_t c = valid && !(lheavy && lheavy[l]) ? plan_cost(hv, budget) : 0u;
    uint32_t nr = 0, fill = budget + 1;
    int len = 0;
    for (int i = 0; i < cnt; ++i) {
        const uint32_t ci = (uint32_t)__builtin_amdgcn_readlane((int)c, i);
        if (fill + ci > budget || len == max_leaves) {
            ++nr;
            fill = 0;
            len = 0;
        }
        fill += ci;
        ++len;
    }
    const uint64_t whole = !valid ? 0ull : (!(hv >= 0.f) ? (uint64_t)budget : (hv >= 33554432.f ? 33554432ull
                                                                                                  : (uint64_t)ceilf(hv)));
    const uint64_t s = wave_sum_u64(whole);
    if (lane_id() == 0) {
        runs[chunk] = nr;
        psum[chunk] = s > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)s;
    }
}

// ------------------------------------------------------------ training

// Row record of one training sample in the step scratch.
constexpr int TR_XN = 0;
constexpr int TR_H1 = TR_XN + MLP_IN;
constexpr int TR_D1 = TR_H1 + MLP_H;
constexpr int TR_H2 = TR_D1 + MLP_H;
constexpr int TR_D2 = TR_H2 + MLP_H;
constexpr int TR_G = TR_D2 + MLP_H;
constexpr int TR_LOSS = TR_G + 1;
constexpr int TR_ROW = TR_LOSS + 1;
static_assert(TR_ROW == MLP_ROW_FLOATS, "row record size is shared with the fixed scratch of capi.hip");

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One sample per thread: forward, loss (y - t)^2 averaged over the batch,
// and the backward through the network down to the first layer's
// pre-activations.  Sample b is row b when seed == 0 (the batch is the data),
// else a hashed row of the data.  Values indexed by a loop counter go through
// the sample's row record rather than registers.
__global__ void __launch_bounds__(MLP_THREADS)
k_mlp_rows(const float* __restrict__ X, const float* __restrict__ T, int64_t n_data, int batch, uint64_t seed,
           const float* __restrict__ P, float* __restrict__ S) {
    __shared__ float sp[MLP_PARAMS];
    load_params(sp, P);
    const int b = blockIdx.x * MLP_THREADS + threadIdx.x;
    if (b >= batch) return;
    const int64_t row = seed == 0 ? (int64_t)b : (int64_t)(mix64(seed ^ ((uint64_t)b << 20)) % (uint64_t)n_data);
    float* s = S + (int64_t)b * TR_ROW;
    float h1[MLP_H];
    {
        float x[MLP_IN], xn[MLP_IN];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) x[i] = X[row * MLP_IN + i];
        mlp_layer1(sp, x, xn, h1);
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) s[TR_XN + i] = xn[i];
    }
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) s[TR_H1 + j] = h1[j];
    float out = sp[MLP_B3];
#pragma unroll 1
    for (int j = 0; j < MLP_H; ++j) {
        float a = sp[MLP_B2 + j];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) a = fmaf(h1[i], sp[MLP_W2 + i * MLP_H + j], a);
        s[TR_H2 + j] = a;  // pre-activation for now
        out += fmaxf(a, 0.f) * sp[MLP_W3 + j];
    }
    const float diff = out - T[row];
    const float g = 2.f * diff / (float)batch;
    float d2[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        const float a = s[TR_H2 + j];
        d2[j] = a > 0.f ? g * sp[MLP_W3 + j] : 0.f;
        s[TR_H2 + j] = fmaxf(a, 0.f);
        s[TR_D2 + j] = d2[j];
    }
    s[TR_G] = g;
    s[TR_LOSS] = diff * diff;
#pragma unroll 1
    for (int i = 0; i < MLP_H; ++i) {
        float d = 0.f;
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) d = fmaf(sp[MLP_W2 + i * MLP_H + j], d2[j], d);
        s[TR_D1 + i] = s[TR_H1 + i] > 0.f ? d : 0.f;
    }
}

constexpr int WG_THREADS = 256;

// Gradient of the trained parameters summed over one chunk of rows, for all
// parameters at once: thread p owns trained parameter p (W1, b1, W2, b2, W3,
// b3 in storage order).  Sums run in row order, so results are reproducible.
__global__ void __launch_bounds__(WG_THREADS)
k_mlp_wgrad(const float* __restrict__ S, int batch, int rows_per_chunk, float* __restrict__ partial) {
    const int p = blockIdx.x * WG_THREADS + threadIdx.x;
    if (p >= MLP_TRAINED) return;
    const int r0 = blockIdx.y * rows_per_chunk;
    const int r1 = min(batch, r0 + rows_per_chunk);
    int ca, cb;  // the two row fields multiplied (cb < 0: the field alone)
    const int q = p + MLP_W1;
    if (q < MLP_B1) {
        ca = TR_XN + (q - MLP_W1) / MLP_H;
        cb = TR_D1 + (q - MLP_W1) % MLP_H;
    } else if (q < MLP_W2) {
        ca = TR_D1 + (q - MLP_B1);
        cb = -1;
    } else if (q < MLP_B2) {
        ca = TR_H1 + (q - MLP_W2) / MLP_H;
        cb = TR_D2 + (q - MLP_W2) % MLP_H;
    } else if (q < MLP_W3) {
        ca = TR_D2 + (q - MLP_B2);
        cb = -1;
    } else if (q < MLP_B3) {
        ca = TR_G;
        cb = TR_H2 + (q - MLP_W3);
    } else {
        ca = TR_G;
        cb = -1;
    }
    float acc = 0.f;
    for (int r = r0; r < r1; ++r) {
        const float* s = S + (int64_t)r * TR_ROW;
        acc += cb >= 0 ? s[ca] * s[cb] : s[ca];
    }
    partial[(int64_t)blockIdx.y * MLP_TRAINED + p] = acc;
}

// Sums the chunks (in order), optionally exports the gradient and the batch
// loss, and applies one Adam step when lr > 0.
__global__ void __launch_bounds__(WG_THREADS)
k_mlp_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ partial,
           int chunks, int step, float lr, float* __restrict__ grad_out, const float* __restrict__ S, int batch,
           float* __restrict__ loss_out) {
    const int p = blockIdx.x * WG_THREADS + threadIdx.x;
    if (p == 0 && loss_out) {
        float l = 0.f;
        for (int r = 0; r < batch; ++r) l += S[(int64_t)r * TR_ROW + TR_LOSS];
        *loss_out = l / (float)batch;
    }
    if (p >= MLP_TRAINED) return;
    float g = 0.f;
    for (int c = 0; c < chunks; ++c) g += partial[(int64_t)c * MLP_TRAINED + p];
    if (grad_out) grad_out[p] = g;
    if (lr > 0.f) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float mm = b1 * m[p] + (1.f - b1) * g;
        const float vv = b2 * v[p] + (1.f - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        const float mh = mm / (1.f - powf(b1, (float)step));
        const float vh = vv / (1.f - powf(b2, (float)step));
        P[MLP_W1 + p] -= lr * mh / (sqrtf(vh) + eps);
    }
}

// Mean and standard deviation of each feature over the data (one block per
// feature, FP64 sums), written into the parameter block; a feature that does
// not vary gets std 1.
__global__ void __launch_bounds__(256) k_feature_stats(const float* __restrict__ X, int64_t n, float* __restrict__ P) {
    __shared__ double s1[256], s2[256];
    const int f = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += 256) {
        const double x = X[r * MLP_IN + f];
        a += x;
        b += x * x;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double mean = n > 0 ? s1[0] / (double)n : 0.0;
        const double var = n > 0 ? s2[0] / (double)n - mean * mean : 0.0;
        const double sd = var > 1e-12 ? sqrt(var) : 1.0;
        P[MLP_MU + f] = (float)mean;
        P[MLP_SD + f] = (float)sd;
    }
}

// Targets log1p(hits) from exact per-leaf hit counts.
__global__ void k_log1p_u32(const uint32_t* __restrict__ in, int64_t n, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = log1pf((float)in[i]);
}

// ------------------------------------------------------------ plan

// Runs of the plan: greedy packing of consecutive leaves, a run closing
// before the leaf that would take its predicted cost past the budget or its
// length past max_leaves.  A leaf's cost is its prediction with 1/16 headroom
// plus 2 entries (a prediction that is not a number costs the whole budget).
// A heavy leaf (lheavy[l] != 0, when given) costs nothing: the single pass
// takes its count from the heavy-leaf pass and never records its hits, and
// a run per heavy leaf would only multiply the runs.  The predicted total
// (psum) still counts it: it sizes the output.  Runs never cross a chunk of
// PLAN_CHUNK leaves, so the packing is the same on every call.  One wave per
// chunk: the run starts at run_base[chunk]; the last chunk closes the table.
__global__ void __launch_bounds__(256)
k_plan_write(const float* __restrict__ pred, int64_t n, uint32_t budget, int max_leaves,
             const uint32_t* __restrict__ run_base, int32_t* __restrict__ wave_first, int32_t* __restrict__ n_runs,
             const uint8_t* __restrict__ lheavy) {
    const int64_t chunk = ((int64_t)blockIdx.x * 256 + threadIdx.x) / PLAN_CHUNK;
    const int lane = lane_id();
    const int64_t l0 = chunk * PLAN_CHUNK;
    if (l0 >= n) return;  // wave-uniform
    const int64_t l = l0 + lane;
    const int cnt = (int)(n - l0 < PLAN_CHUNK ? n - l0 : PLAN_CHUNK);
    const uint32_t c = l < n && !(lheavy && lheavy[l]) ? plan_cost(pred[l], budget) : 0u;
    uint64_t starts = 0;
    uint32_t fill = budget + 1;
    int len = 0;
    for (int i = 0; i < cnt; ++i) {
        const uint32_t ci = (uint32_t)__builtin_amdgcn_readlane((int)c, i);
        if (fill + ci > budget || len == max_leaves) {
            starts |= 1ull << i;
            fill = 0;
            len = 0;
        }
        fill += ci;
        ++len;
    }
    const uint32_t base = run_base[chunk];
    if ((starts >> lane) & 1ull) wave_first[base + (uint32_t)__popcll(starts & ((1ull << lane) - 1ull))] = (int32_t)l;
    if (l0 + cnt == n && lane == 0) {
        const uint32_t total = base + (uint32_t)__popcll(starts);
        wave_first[total] = (int32_t)n;
        *n_runs = (int32_t)total;
    }
}

// ------------------------------------------------------------ launchers

hipError_t launch_leaf_features(const TreeView& q, const int2* lists, const uint8_t* ldir, const int32_t* lcount,
                                int cap, float r, float* feat, hipStream_t s) {
    if (q.n_leaves <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_leaf_features, dim3((unsigned)((q.n_leaves + FT_THREADS - 1) / FT_THREADS)),
                       dim3(FT_THREADS), 0, s, q.leaves, q.leaf_ids, q.nlevel, (int64_t)q.n_leaves, lists, ldir,
                       lcount, cap, r, feat);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const float* X, int64_t n, const float* P, float* y, float* hits, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + MLP_THREADS - 1) / MLP_THREADS)), dim3(MLP_THREADS), 0, s,
                       X, n, P, y, hits, nullptr, nullptr, 0u, 0, nullptr);
    return hipGetLastError();
}

int64_t mlp_scratch_floats(int batch) { return mlp_scratch_floats_c(batch); }

hipError_t launch_mlp_step(const float* X, const float* T, int64_t n_data, int batch, uint64_t seed, float* P,
                           float* m, float* v, int step, float lr, float* scratch, float* grad_out, float* loss_out,
                           hipStream_t s) {
    if (batch <= 0) return hipErrorInvalidValue;
    float* S = scratch;
    float* partial = scratch + (int64_t)batch * TR_ROW;
    const int chunks = (batch + MLP_ROWS_PER_CHUNK - 1) / MLP_ROWS_PER_CHUNK;
    hipLaunchKernelGGL(k_mlp_rows, dim3((unsigned)((batch + MLP_THREADS - 1) / MLP_THREADS)), dim3(MLP_THREADS), 0,
                       s, X, T, n_data, batch, seed, P, S);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((unsigned)((MLP_TRAINED + WG_THREADS - 1) / WG_THREADS), (unsigned)chunks),
                       dim3(WG_THREADS), 0, s, S, batch, MLP_ROWS_PER_CHUNK, partial);
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((MLP_TRAINED + WG_THREADS - 1) / WG_THREADS)), dim3(WG_THREADS),
                       0, s, P, m, v, partial, chunks, step, lr, grad_out, S, batch, loss_out);
    return hipGetLastError();
}

hipError_t launch_feature_stats(const float* X, int64_t n, float* P, hipStream_t s) {
    hipLaunchKernelGGL(k_feature_stats, dim3(MLP_IN), dim3(256), 0, s, X, n, P);
    return hipGetLastError();
}

hipError_t launch_log1p_u32(const uint32_t* in, int64_t n, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_log1p_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, n, out);
    return hipGetLastError();
}

int64_t plan_chunks(int64_t n_leaves) { return (n_leaves + PLAN_CHUNK - 1) / PLAN_CHUNK; }

hipError_t launch_predict_plan(const float* feat, int64_t n, const float* P, float* pred, uint32_t budget,
                               int max_leaves, uint32_t* work, uint32_t* scan_tmp, int32_t* wave_first,
                               int32_t* n_runs, uint64_t* pred_total, const uint8_t* lheavy, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t nc = plan_chunks(n);
    uint32_t* runs = work;
    uint32_t* psum = work + nc;
    uint32_t* run_base = work + 2 * nc;
    uint32_t* psum_pre = work + 3 * nc;
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + MLP_THREADS - 1) / MLP_THREADS)), dim3(MLP_THREADS), 0, s,
                       feat, n, P, nullptr, pred, runs, psum, budget, max_leaves, lheavy);
    hipError_t e = exclusive_scan(runs, run_base, nc, scan_tmp, nullptr, s);
    if (e != hipSuccess) return e;
    e = exclusive_scan(psum, psum_pre, nc, scan_tmp, pred_total, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_plan_write, dim3((unsigned)((nc * PLAN_CHUNK + 255) / 256)), dim3(256), 0, s, pred, n, budget,
                       max_leaves, run_base, wave_first, n_runs, lheavy);
    return hipGetLastError();
}

}  // namespace oct
