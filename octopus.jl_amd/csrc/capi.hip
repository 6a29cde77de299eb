// C-ABI of the octohip library: context, point sets, active pairs, run!,
// build_edges, device view and the build_edges_diff pullback.  Host code that
// sizes buffers, launches the kernels of build.hip / edges.hip / backward.hip
// and reads back only what a host caller needs.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"
#include "octohip.h"
#include "octohip_device.h"

namespace oct {

// Grow-only device buffer (reused across calls, like ensure_capacity!,
// src/util.jl:12-17).  A buffer the C-ABI hands out (edges, tree arrays,
// device views, statistics) has `retire` set: when it grows, the old
// allocation goes to the context's retire list instead of being freed, so a
// view of it taken earlier stays readable memory (stale contents, as a
// replaced Julia array stays valid) until oct_trim or oct_destroy.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    std::vector<void*>* retire = nullptr;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        // grow by at least 1.5x (and 1/8 headroom, 64 KiB granules) so sizes
        // that drift up step to step do not reallocate every call
        size_t want = std::max(bytes + bytes / 8, cap + cap / 2);
        want = (want + 65535) & ~(size_t)65535;
        if (p) {
            if (retire) {
                retire->push_back(p);
            } else {
                hipError_t e = hipFree(p);
                if (e != hipSuccess) return e;
            }
            p = nullptr;
            cap = 0;
        }
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) { p = nullptr; return e; }
        cap = want;
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T> T* as() const { return (T*)p; }
};

struct SetState {
    const float* coords = nullptr;
    int64_t n = 0;
    bool dirty = true;
    bool built = false;
    int key_bits = 0;
    BuildParams hp{};       // host copy of the build parameters
    bool hp_fresh = false;  // hp holds the last build's tree counts (read at a later wait, not by run!)
    int64_t leaves_known = -1;  // last leaf count read back (sizes the per-leaf arrays of the next build)
    int built_leaf = 32;        // leaf size of the last build
    bool refit_pending = false;  // the last build refitted without a wait: its outcome comes with hp
    // the previous sort (keys[cur], scoord) for the almost-sorted refit
    bool sorted_ok = false;
    int64_t sorted_n = 0;
    int cur = 0;
    oct_build_info info{};
    DevBuf rk, ri, rsk, rsi, rpos;  // refit: changed points
    DevBuf pad;                     // 2D sets: the (x, y, PLANE_Z) copy the build reads (planar.hip)
    DevBuf qfirst, qlast, qkids, qbmin, qbmax;  // 2D sets: the quadtree in the reference layout
    // Float64 sets (f64.hip): double parameters and bounds, the rounded copy the sort reads
    DevBuf p64, partials64, c32, bmin64, bmax64;
    DevBuf soff;                    // {0, n}: the one-set forest descriptor
    int64_t soff_n = -1;            // the n soff holds
    DevBuf prm, partials, keys0, keys1, vals0, vals1, hist, scan_tmp, total, topo, perm, scoord, nodes,
        nfirst, nlast, nchild, nlevel, leaf_ids, leaves, bmin, bmax, rslots, ranc, nfc, noct, rlook;
    // the buffers oct_get_tree and oct_device_view hand out
    void export_to(std::vector<void*>* r) {
        DevBuf* out[] = {&perm, &nfirst, &nlast, &nchild, &bmin, &bmax, &nodes, &scoord, &pad,
                         &qfirst, &qlast, &qkids, &qbmin, &qbmax, &bmin64, &bmax64};
        for (DevBuf* d : out) d->retire = r;
    }
    BuildBufs bufs() const {
        BuildBufs b;
        b.prm = prm.as<BuildParams>();
        b.partials = partials.as<uint32_t>();
        b.keys[0] = keys0.as<uint64_t>();
        b.keys[1] = keys1.as<uint64_t>();
        b.vals[0] = vals0.as<int32_t>();
        b.vals[1] = vals1.as<int32_t>();
        b.hist = hist.as<uint32_t>();
        b.scan_tmp = scan_tmp.as<uint32_t>();
        b.total = total.as<uint64_t>();
        b.topo = topo.as<uint32_t>();
        b.perm = perm.as<int32_t>();
        b.scoord = scoord.as<float4>();
        b.nodes = nodes.as<NodeRec>();
        b.nfirst = nfirst.as<int32_t>();
        b.nlast = nlast.as<int32_t>();
        b.nchild = nchild.as<int32_t>();
        b.nlevel = nlevel.as<uint8_t>();
        b.leaf_ids = leaf_ids.as<int32_t>();
        b.leaves = leaves.as<NodeRec>();
        b.bmin = bmin.as<float>();
        b.bmax = bmax.as<float>();
        b.rslots = rslots.as<uint32_t>();
        b.ranc = ranc.as<int32_t>();
        b.nfc = nfc.as<int32_t>();
        b.noct = noct.as<uint8_t>();
        b.rlook = rlook.as<uint64_t>();
        return b;
    }
    // The refit step (bounds, refit, topology: some 25 launches) as a HIP
    // graph per parity of cur, replayed while its key (sizes, buffers,
    // radius) holds.  It reads the caller's coordinates through csrc, a device
    // word the bounds stage sets every step, so the caller may pass a new
    // array every step.
    DevBuf csrc;
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        uint64_t key = 0;
    };
    StepGraph rg[2];
    void drop_graphs() {
        for (auto& g : rg) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            g.key = 0;
        }
    }
    void release_all() {
        drop_graphs();
        DevBuf* all[] = {&prm, &partials, &keys0, &keys1, &vals0, &vals1, &hist, &scan_tmp, &total, &topo, &perm,
                         &scoord, &nodes, &nfirst, &nlast, &nchild, &nlevel, &leaf_ids, &leaves, &bmin, &bmax,
                         &rslots, &ranc, &nfc, &noct, &rlook, &rk, &ri, &rsk, &rsi, &rpos, &soff, &csrc, &pad, &qfirst, &qlast,
                         &qkids, &qbmin, &qbmax, &p64, &partials64, &c32, &bmin64, &bmax64};
        for (DevBuf* d : all) d->release();
        soff_n = -1;
    }
};

struct PairState {
    int q = 0, t = 0;
    int cap = 64;            // candidate leaves per query leaf
    int64_t cap_edges = 0;   // edges the output buffers hold (single pass sizing)
    bool replan = true;      // run the predictor on the next learned call
    int64_t n_edges = 0;
    bool materialized = false;
    DevBuf senders, receivers, rdisp, rdist, lists, lcount, qcount, qoff, scan_tmp, flags, tiles;
    // predictor inputs and plan: direction codes of the list entries,
    // features, predicted and exact per-leaf hits, plan work / scan / runs
    DevBuf ldir, feat, pred, lhits, pwork, pscan, wave_first;
    // heavy query leaves, counted ahead of the single pass: their list
    // (int32 per leaf) followed by a flag byte per leaf
    DevBuf heavy;
    DevBuf eprm;  // the query tree's build parameters with the leaf count clamped to the rows
    // CSR neighbour lists (materialize_all_neighbors!): offsets (n_q + 1) and targets
    DevBuf csr_off, csr_flat;
    DevBuf rdisp2;  // 2D contexts: rel_displacement (2, E)
    bool csr_ok = false;
    int64_t csr_nq = 0, csr_total = 0;
    bool heavy_on = true;  // count heavy leaves ahead: first call, and while many leaves are not recorded
    // a call that returned on the single pass's early total: its recount
    // statistics are final only when the kernel ends (read later)
    bool stats_pending = false;
    int64_t pending_heavy = 0, pending_leaves = 0;
    int32_t* heavy_list() { return heavy.as<int32_t>(); }
    uint8_t* heavy_flags(int64_t n_leaves) { return (uint8_t*)(heavy.as<int32_t>() + n_leaves); }
    oct_edge_stats st{};
    // the buffers oct_build_edges, oct_batch_run and oct_edge_stats_get hand out
    void export_to(std::vector<void*>* r) {
        DevBuf* out[] = {&senders, &receivers, &rdisp, &rdist, &feat, &pred, &lhits, &csr_off, &csr_flat, &rdisp2};
        for (DevBuf* d : out) d->retire = r;
    }
    void release_all() {
        DevBuf* all[] = {&senders, &receivers, &rdisp, &rdist, &lists, &lcount, &qcount, &qoff, &scan_tmp, &flags,
                         &tiles, &ldir, &feat, &pred, &lhits, &pwork, &pscan, &wave_first, &heavy, &eprm,
                         &csr_off, &csr_flat, &rdisp2};
        for (DevBuf* d : all) d->release();
    }
};

// Predictor parameters, Adam state and training set of a context.
struct Predictor {
    bool loaded = false;
    int adam_step = 0;
    int64_t rows = 0;
    DevBuf params, m, v, scratch, loss, X, T;
    void release_all() {
        DevBuf* all[] = {&params, &m, &v, &scratch, &loss, &X, &T};
        for (DevBuf* d : all) d->release();
    }
};

// A batch: many point sets searched against themselves as one forest (see
// batch.hip).  Every step is enqueued with host parameters fixed after the
// first eager call (sort shifts, capacities, grids sized by the point count),
// so it can be captured once and replayed as a HIP graph.
struct BatchState {
    bool live = false;
    int S = 0;
    int64_t N = 0;
    const float* coords = nullptr;
    std::vector<int64_t> off_h;
    std::vector<float> radii_h;
    int leaf_size = 32;
    int run_leaves = SP_DEFAULT_RUN;
    int run_chunk = SP_DEFAULT_CHUNK;
    bool use_graph = true;
    int nblk_bounds = 0;
    int max_bits = -1;         // Morton bits the captured sort covers (-1: not set up)
    std::vector<int> shifts;   // radix digit shifts
    int cap = 64;              // candidate leaves per query leaf
    int64_t cap_edges = 0;     // edge buffer capacity (edges)
    int64_t n_edges = 0;
    int replayed = 0;
    int steps = 0;
    int64_t recounted = 0, recorded = 0;  // leaves the last step swept twice / once
    int64_t heavy = 0;                    // leaves the last step counted ahead
    SetState set;              // forest build buffers; set.prm holds the forest's totals
    PairState pair;            // edge buffers
    DevBuf off, radii, bprefix, sprm, eoff, loff, bflags, bscan;
    hipGraphExec_t exec = nullptr;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    void release_all() {
        drop_graph();
        set.release_all();
        pair.release_all();
        DevBuf* all[] = {&off, &radii, &bprefix, &sprm, &eoff, &loff, &bflags, &bscan};
        for (DevBuf* d : all) d->release();
        live = false;
    }
};

}  // namespace oct

using namespace oct;

// Fixed device scratch of oct_mlp_gradient_noalloc (the entry point that
// allocates nothing): the row records and partial sums of one gradient over
// at most OCT_MLP_STATIC_ROWS rows, ~10 MB on each GPU the library runs on,
// held until the process ends.  Host code never passes `mlp_noalloc_scratch`
// itself to a kernel (on the host it names a stand-in, not device memory):
// oct_create takes each device's own address of it with hipGetSymbolAddress,
// with that device current, and keeps it in noalloc_scratch_ptr[device].
constexpr int64_t MLP_NOALLOC_FLOATS = mlp_scratch_floats_c(OCT_MLP_STATIC_ROWS);
static __device__ __align__(256) float mlp_noalloc_scratch[MLP_NOALLOC_FLOATS];
constexpr int OCT_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static float* noalloc_scratch_ptr[OCT_MAX_DEVICES] = {nullptr};

// The current device's address of the fixed scratch (device is current).
static hipError_t noalloc_scratch_setup(int device) {
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(mlp_noalloc_scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    noalloc_scratch_ptr[device] = (float*)p;
    return hipSuccess;
}

struct oct_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t join[2] = {nullptr, nullptr};  // oct_stream_join: caller -> context, context -> caller
    hipStream_t cap_stream = nullptr;          // graph captures (the caller's stream may be the legacy default)
    int index_base = 0;
    int ndims = 3;  // 2: quadtree sets, built as planar octrees (planar.hip)
    bool f64 = false;     // Float64 sets (f64.hip)
    double radius64 = 0;  // the radius in double (Float64 contexts)
    float radius = 0.f;
    int leaf_size = 32;
    bool refit_mode = false;
    bool predict_always = false;
    int run_leaves = SP_DEFAULT_RUN;  // leaves per single-pass run without a plan
    int run_chunk = SP_DEFAULT_CHUNK;  // single-pass runs per block ticket
    std::vector<SetState> sets;
    std::vector<PairState> pairs;
    std::vector<BatchState> batches;
    std::vector<void*> retired;  // grown-out allocations of handed-out buffers (DevBuf::retire)
    bool pull_backward = true;   // build_edges_diff adjoint in the pull form (run table), else the push form
    DevBuf bw_scratch;           // its run table
    Predictor mlp;
    std::string err;
    void* pinned = nullptr;  // small host staging area for read-backs
    uint32_t* sig = nullptr;      // host-coherent block the single pass hands its total through (SP_SIG_MARK)
    uint32_t* sig_dev = nullptr;  // its device address
    // optional per-phase kernel timing with HIP events on the context stream
    uint32_t prof_mask = 0;  // phases timed with events (bit per OCT_PROF_*)
    bool capturing = false;  // a graph capture is open: no phase events
    struct Mark { int phase; hipEvent_t a, b; };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> spare;
    double prof_ms[OCT_PROF_NPHASES] = {0};
    int64_t prof_calls[OCT_PROF_NPHASES] = {0};
};

namespace {

// Host wall-clock of a code section, printed to stderr when the environment
// sets OCTOHIP_HOST_TIMING (a diagnostic of launch and wait overheads).
struct HostTimer {
    const char* name;
    std::chrono::steady_clock::time_point t0;
    static bool on() {
        static int v = getenv("OCTOHIP_HOST_TIMING") ? 1 : 0;
        return v != 0;
    }
    explicit HostTimer(const char* n) : name(n), t0(std::chrono::steady_clock::now()) {}
    ~HostTimer() {
        if (on()) {
            const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
            fprintf(stderr, "[octohip host] %s %.1f us\n", name, us);
        }
    }
};

hipEvent_t take_event(oct_ctx* c) {
    if (!c->spare.empty()) {
        hipEvent_t e = c->spare.back();
        c->spare.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Brackets the launches between begin/end with events when profiling is on.
struct Phase {
    oct_ctx* c;
    int ph;
    hipEvent_t a = nullptr;
    Phase(oct_ctx* ctx, int phase) : c(ctx), ph(phase) {
        if (((c->prof_mask >> ph) & 1u) && !c->capturing) {
            a = take_event(c);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~Phase() {
        if (a) {
            hipEvent_t b = take_event(c);
            (void)hipEventRecord(b, c->stream);
            c->marks.push_back({ph, a, b});
        }
    }
};

int fail(oct_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, call)                                                              \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail((ctx), OCT_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define HIP_TRY_RC(call)              \
    do {                              \
        const int rc_ = (call);       \
        if (rc_ != OCT_OK) return rc_; \
    } while (0)

// Captures what `enqueue` launches on c->stream into an executable graph.
// The capture runs on the context's own stream: the caller's stream may be
// the legacy default stream, which cannot capture; the graph then launches
// on any stream.
template <typename F>
int capture_graph(oct_ctx* c, F enqueue, hipGraphExec_t* exec) {
    *exec = nullptr;
    if (!c->cap_stream) HIP_TRY(c, hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
    const hipStream_t saved = c->stream;
    c->stream = c->cap_stream;
    const hipError_t be = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (be != hipSuccess) {
        c->stream = saved;
        HIP_TRY(c, be);
    }
    c->capturing = true;
    const int erc = enqueue();
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(c->stream, &g);
    c->stream = saved;
    if (erc != OCT_OK || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (erc != OCT_OK) return erc;
        HIP_TRY(c, ee);
    }
    const hipError_t ie = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(c, ie);
    return OCT_OK;
}

int set_index(oct_ctx* c, int id) { return id - c->index_base; }

bool valid_set(oct_ctx* c, int id) {
    int k = set_index(c, id);
    return k >= 0 && k < (int)c->sets.size();
}

// Device view of a built set.  Launch shapes use rows (an upper bound of the
// leaves when the counts are not back on the host yet; kernels read the real
// counts from params) and the node bound of the build.
TreeView view_of(const SetState& s, int64_t rows, const BuildParams* params = nullptr) {
    TreeView v;
    v.nodes = s.nodes.as<NodeRec>();
    v.leaf_ids = s.leaf_ids.as<int32_t>();
    v.leaves = s.leaves.as<NodeRec>();
    v.scoord = s.scoord.as<float4>();
    v.params = params ? params : s.prm.as<BuildParams>();
    v.nlevel = s.nlevel.as<uint8_t>();
    v.n_leaves = (int32_t)rows;
    v.n_nodes = s.hp_fresh ? s.hp.n_nodes : (int32_t)std::min<int64_t>(max_nodes(s.n, s.built_leaf), INT32_MAX);
    v.branch_root = s.hp_fresh ? s.hp.branch_root : 0;
    v.n = s.n;
    return v;
}

int mark_all_dirty(oct_ctx* c) {
    for (auto& s : c->sets) s.dirty = true;
    return OCT_OK;
}

// Waits for the stream and copies the build parameters (and, when asked, the
// 64-bit scan total) to the host.
// Waits for the context's stream by polling it: a read-back in the middle of
// a step costs the wake-up latency of a blocking wait every time, which
// polling avoids (the host thread spins for the duration instead).
hipError_t wait_stream(hipStream_t s) {
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e != hipErrorNotReady) return e;
        __builtin_ia32_pause();
    }
}

int read_params(oct_ctx* c, SetState& s, uint64_t* total = nullptr) {
    HIP_TRY(c, hipMemcpyAsync(c->pinned, s.prm.p, sizeof(BuildParams), hipMemcpyDeviceToHost, c->stream));
    if (total)
        HIP_TRY(c, hipMemcpyAsync((char*)c->pinned + 1024, s.total.p, 8, hipMemcpyDeviceToHost, c->stream));
    HostTimer ht("run: wait for build parameters");
    HIP_TRY(c, wait_stream(c->stream));
    memcpy(&s.hp, c->pinned, sizeof(BuildParams));
    if (total) memcpy(total, (char*)c->pinned + 1024, 8);
    return OCT_OK;
}

// Takes a copy of a set's build parameters read back at a later wait.
void take_hp(SetState& s, const void* src) {
    memcpy(&s.hp, src, sizeof(BuildParams));
    s.hp_fresh = true;
    s.leaves_known = s.hp.n_leaves;
}

int build_set(oct_ctx* c, SetState& s, bool force_full = false);

// With fresh build parameters: records how a refit without a wait went, and
// when it could not run (too many changed points, or a new key width)
// rebuilds the set with the full sort.  *rebuilt tells the caller that the
// tree (and hp) changed.
int settle_refit(oct_ctx* c, SetState& s, bool* rebuilt) {
    *rebuilt = false;
    if (!s.refit_pending || !s.hp_fresh) return OCT_OK;
    s.refit_pending = false;
    const int64_t changed = s.hp.changed;
    if (!s.hp.refit_fail) {
        s.info.refit_used = 1;
        s.info.changed_points = changed;
        return OCT_OK;
    }
    HIP_TRY_RC(build_set(c, s, true));
    s.info.refit_used = 0;
    s.info.changed_points = changed;
    *rebuilt = true;
    return OCT_OK;
}

// The tree counts of the last build on the host (waits for the stream when
// they are not back yet; settles a pending refit).
int ensure_hp(oct_ctx* c, SetState& s) {
    for (int k = 0; k < 3 && !(s.hp_fresh && !s.refit_pending) && s.n > 0; ++k) {
        if (!s.hp_fresh) {
            HIP_TRY_RC(read_params(c, s));
            s.hp_fresh = true;
            s.leaves_known = s.hp.n_leaves;
        }
        bool rebuilt = false;
        HIP_TRY_RC(settle_refit(c, s, &rebuilt));
    }
    return OCT_OK;
}

// run! of a Float64 set (f64.hip): double bounds, origin and keys, the
// shared sort and topology, double node bounds.  Always the full sort (the
// almost-sorted refit reads Float32 coordinates).
int build_set_f64(oct_ctx* c, SetState& s);

// The device buffers of one set's build (n points, at most M nodes).
int ensure_build_bufs(oct_ctx* c, SetState& s, int64_t n, int64_t M) {
    const bool fresh_prm = s.prm.p == nullptr;
    HIP_TRY(c, s.prm.ensure(sizeof(BuildParams)));
    if (fresh_prm)
        HIP_TRY(c, launch_fill_u32(s.prm.as<uint32_t>(), sizeof(BuildParams) / 4, 0u, c->stream));
    HIP_TRY(c, s.partials.ensure(6 * sizeof(uint32_t) * (size_t)bounds_blocks(n)));
    HIP_TRY(c, s.keys0.ensure(8 * (size_t)n));
    HIP_TRY(c, s.keys1.ensure(8 * (size_t)n));
    HIP_TRY(c, s.vals0.ensure(4 * (size_t)n));
    HIP_TRY(c, s.vals1.ensure(4 * (size_t)n));
    HIP_TRY(c, s.hist.ensure(4 * (size_t)hist_words(n)));
    // the scan of the radix histograms
    HIP_TRY(c, s.scan_tmp.ensure(4 * (size_t)scan_tmp_words(hist_words(n))));
    HIP_TRY(c, s.total.ensure(16));  // uint64 total
    HIP_TRY(c, s.topo.ensure(4 * (size_t)n));
    HIP_TRY(c, s.perm.ensure(4 * (size_t)n));
    HIP_TRY(c, s.scoord.ensure(16 * (size_t)n));
    HIP_TRY(c, s.nodes.ensure(sizeof(NodeRec) * (size_t)M));
    HIP_TRY(c, s.nfirst.ensure(4 * (size_t)M));
    HIP_TRY(c, s.nlast.ensure(4 * (size_t)M));
    HIP_TRY(c, s.nchild.ensure(32 * (size_t)M));
    HIP_TRY(c, s.nlevel.ensure((size_t)M));
    HIP_TRY(c, s.leaf_ids.ensure(4 * (size_t)n));
    HIP_TRY(c, s.leaves.ensure(sizeof(NodeRec) * (size_t)n));
    HIP_TRY(c, s.bmin.ensure(12 * (size_t)M));
    HIP_TRY(c, s.bmax.ensure(12 * (size_t)M));
    HIP_TRY(c, s.rslots.ensure(4 * (size_t)refit_slot_words(n)));
    HIP_TRY(c, s.ranc.ensure(4 * (size_t)OCT_MAX_LEVEL * (size_t)n));
    HIP_TRY(c, s.nfc.ensure(4 * ((size_t)M + 1)));
    HIP_TRY(c, s.noct.ensure((size_t)M));
    HIP_TRY(c, s.rlook.ensure(8 * (size_t)refit_look_words(n)));
    return OCT_OK;
}

int build_set(oct_ctx* c, SetState& s, bool force_full) {
    const int64_t n = s.n;
    if (s.refit_pending) HIP_TRY_RC(ensure_hp(c, s));  // the previous build's outcome first
    s.built = false;
    if (n == 0) {
        memset(&s.hp, 0, sizeof(s.hp));
        s.hp_fresh = true;
        s.built = true;
        s.dirty = false;
        return OCT_OK;
    }
    if (n >= (int64_t)INT32_MAX) return fail(c, OCT_ERR_CAPACITY, "point sets are limited to 2^31 - 1 points");
    if (c->f64) return build_set_f64(c, s);
    const int T = c->leaf_size;
    const int64_t M = max_nodes(n, T);
    HIP_TRY_RC(ensure_build_bufs(c, s, n, M));
    const BuildBufs b = s.bufs();
    // a 2D set is built from its (x, y, PLANE_Z) copy
    const float* src = s.coords;
    if (c->ndims == 2) {
        HIP_TRY(c, s.pad.ensure(12 * (size_t)n));
        HIP_TRY(c, launch_pad_planar(s.coords, n, s.pad.as<float>(), c->stream));
        src = s.pad.as<float>();
    }
    // almost-sorted refit: the previous sort of the same n points is reused.
    // It runs without a host wait: the device decides whether it can
    // (prm->refit_fail), and whoever reads the build parameters back settles
    // the outcome (settle_refit), rebuilding with the full sort if it could not.
    const bool try_refit = c->refit_mode && s.sorted_ok && s.sorted_n == n && !force_full;
    s.sorted_ok = false;
    memset(&s.info, 0, sizeof(s.info));
    s.info.changed_points = -1;
    HIP_TRY(c, s.soff.ensure(16));
    if (s.soff_n != n) {  // {0, n}: the one-set forest descriptor, written when n changes
        HIP_TRY(c, launch_single_set_offsets(s.soff.as<int64_t>(), n, c->stream));
        s.soff_n = n;
    }
    const Forest f{0, s.soff.as<int64_t>(), s.prm.as<BuildParams>()};
    int cur = 0;
    if (try_refit) {
        const int64_t limit = std::min<int64_t>(REFIT_MAX_CHANGED, n / 8);
        HIP_TRY(c, s.rk.ensure(8 * (size_t)REFIT_MAX_CHANGED));
        HIP_TRY(c, s.ri.ensure(4 * (size_t)REFIT_MAX_CHANGED));
        HIP_TRY(c, s.rsk.ensure(8 * (size_t)REFIT_MAX_CHANGED));
        HIP_TRY(c, s.rsi.ensure(4 * (size_t)REFIT_MAX_CHANGED));
        HIP_TRY(c, s.rpos.ensure(4 * (size_t)REFIT_MAX_CHANGED));
        const RefitBufs r{s.rk.as<uint64_t>(), s.ri.as<int32_t>(), s.rsk.as<uint64_t>(), s.rsi.as<int32_t>(),
                          s.rpos.as<uint32_t>()};
        // the key width must stay s.key_bits (the levels the topology refits depend on it)
        // the graph starts after the bounds, which run eagerly on the caller's
        // coordinates and leave their address in csrc (the graph reads it there)
        HIP_TRY(c, s.csrc.ensure(sizeof(const float*)));
        const float* const* csrc = s.csrc.as<const float*>();
        auto enqueue = [&]() -> int {
            Phase ph(c, OCT_PROF_BUILD);
            HIP_TRY(c, launch_refit_prepare(csrc, n, limit, s.key_bits, s.cur, c->index_base, b, r, c->stream));
            HIP_TRY(c, launch_refit_finish(csrc, n, limit, s.cur, c->index_base, b, r, c->stream));
            HIP_TRY(c, launch_topology_stage(n, s.key_bits, T, c->index_base, s.cur ^ 1, b, f, c->stream));
            return OCT_OK;
        };
        {
            {
                Phase ph(c, OCT_PROF_BOUNDS);
                HIP_TRY(c, launch_bounds_stage(src, n, c->radius, nullptr, b, c->stream,
                                               s.csrc.as<const float*>()));
            }
            // the graph's key: everything its launches bake in
            uint64_t key = 1469598103934665603ull;
            auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
            const DevBuf* bufs[] = {&s.prm, &s.partials, &s.keys0, &s.keys1, &s.vals0, &s.vals1, &s.hist, &s.scan_tmp,
                                    &s.total, &s.topo, &s.perm, &s.scoord, &s.nodes, &s.nfirst, &s.nlast, &s.nchild,
                                    &s.nlevel, &s.leaf_ids, &s.leaves, &s.bmin, &s.bmax, &s.rslots,
                                    &s.ranc, &s.nfc, &s.noct, &s.rlook, &s.rk, &s.ri,
                                    &s.rsk, &s.rsi, &s.rpos, &s.soff, &s.csrc};
            for (const DevBuf* d : bufs) mix((uint64_t)(uintptr_t)d->p);
            uint32_t rbits = 0;
            memcpy(&rbits, &c->radius, 4);
            mix((uint64_t)n);
            mix((uint64_t)T);
            mix((uint64_t)c->index_base);
            mix((uint64_t)s.key_bits);
            mix((uint64_t)rbits);
            SetState::StepGraph& g = s.rg[s.cur & 1];
            if (!g.exec || g.key != key) {
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
                g.exec = nullptr;
                g.key = 0;
                HIP_TRY_RC(capture_graph(c, [&]() { return enqueue(); }, &g.exec));
                g.key = key;
            }
            Phase ph(c, OCT_PROF_BUILD);  // refit and topology
            HIP_TRY(c, hipGraphLaunch(g.exec, c->stream));
        }
        cur = s.cur ^ 1;
        s.refit_pending = true;
    } else {
        {
            Phase ph(c, OCT_PROF_BOUNDS);
            HIP_TRY(c, launch_bounds_stage(src, n, c->radius, nullptr, b, c->stream));
        }
        Phase ph(c, OCT_PROF_BUILD);
        HIP_TRY_RC(read_params(c, s));  // the key width sizes the radix passes
        s.key_bits = s.hp.key_bits;
        HIP_TRY(c, launch_sort_stage(src, n, s.key_bits, c->index_base, b, &cur, c->stream));
        HIP_TRY(c, launch_topology_stage(n, s.key_bits, T, c->index_base, cur, b, f, c->stream));
    }
    // no wait here: the tree counts come back with the next wait that needs them
    s.hp_fresh = false;
    s.built_leaf = T;
    s.cur = cur;
    s.sorted_ok = true;
    s.sorted_n = n;
    s.built = true;
    s.dirty = false;
    return OCT_OK;
}

int build_set_f64(oct_ctx* c, SetState& s) {
    const int64_t n = s.n;
    const int T = c->leaf_size;
    const int64_t M = max_nodes(n, T);
    HIP_TRY_RC(ensure_build_bufs(c, s, n, M));
    HIP_TRY(c, s.p64.ensure(sizeof(Params64)));
    HIP_TRY(c, s.partials64.ensure(6 * sizeof(uint64_t) * (size_t)bounds64_blocks(n)));
    HIP_TRY(c, s.c32.ensure(12 * (size_t)n));
    HIP_TRY(c, s.bmin64.ensure(24 * (size_t)M));
    HIP_TRY(c, s.bmax64.ensure(24 * (size_t)M));
    HIP_TRY(c, s.soff.ensure(16));
    if (s.soff_n != n) {
        HIP_TRY(c, launch_single_set_offsets(s.soff.as<int64_t>(), n, c->stream));
        s.soff_n = n;
    }
    const BuildBufs b = s.bufs();
    const double* src = (const double*)s.coords;
    if (c->ndims == 2) {  // the (x, y, 2^34) copy: its z cell is 0 in double
        HIP_TRY(c, s.pad.ensure(24 * (size_t)n));
        HIP_TRY(c, launch_pad_planar64((const double*)s.coords, n, s.pad.as<double>(), c->stream));
        src = s.pad.as<double>();
    }
    memset(&s.info, 0, sizeof(s.info));
    s.info.changed_points = -1;
    s.refit_pending = false;
    {
        Phase ph(c, OCT_PROF_BOUNDS);
        HIP_TRY(c, launch_build64(src, n, c->radius64, s.partials64.as<uint64_t>(), s.p64.as<Params64>(),
                                  b.prm, c->stream));
    }
    Phase ph(c, OCT_PROF_BUILD);
    HIP_TRY_RC(read_params(c, s));  // the key width sizes the radix passes
    s.key_bits = s.hp.key_bits;
    HIP_TRY(c, launch_keys64(src, n, s.p64.as<Params64>(), b.keys[0], b.vals[0], s.c32.as<float>(), c->stream));
    int passes = std::max(1, (s.key_bits + RADIX_BITS - 1) / RADIX_BITS);
    int shifts[8];
    for (int k = 0; k < passes; ++k) shifts[k] = k * RADIX_BITS;
    int cur = 0;
    HIP_TRY(c, launch_sort_passes(s.c32.as<float>(), n, shifts, passes, c->index_base, b, &cur, c->stream));
    const Forest f{0, s.soff.as<int64_t>(), s.prm.as<BuildParams>()};
    HIP_TRY(c, launch_topology_stage(n, s.key_bits, T, c->index_base, cur, b, f, c->stream));
    HIP_TRY(c, launch_refit64(src, b.prm, M, b.nlevel, b.nfirst, b.nlast, b.nchild, b.perm, c->index_base,
                              s.bmin64.as<double>(), s.bmax64.as<double>(), c->stream));
    s.hp_fresh = false;
    s.built_leaf = T;
    s.cur = cur;
    s.sorted_ok = false;
    s.built = true;
    s.dirty = false;
    return OCT_OK;
}

PairState* find_pair(oct_ctx* c, int q, int t) {
    for (auto& p : c->pairs)
        if (p.q == q && p.t == t) return &p;
    return nullptr;
}

}  // namespace

// ================================================================== C ABI

extern "C" {

int oct_create(oct_ctx** out, int device, int ndims, int index_base) {
    return oct_create_typed(out, device, ndims, index_base, 4);
}

int oct_create_typed(oct_ctx** out, int device, int ndims, int index_base, int elem_bytes) {
    if (!out) return OCT_ERR_ARGUMENT;
    if (elem_bytes != 4 && elem_bytes != 8) return OCT_ERR_ARGUMENT;
    *out = nullptr;
    if (ndims != 3 && ndims != 2) return OCT_ERR_ARGUMENT;
    if (index_base != 0 && index_base != 1) return OCT_ERR_ARGUMENT;
    if (device < 0 || device >= OCT_MAX_DEVICES) return OCT_ERR_ARGUMENT;
    oct_ctx* c = new oct_ctx();
    c->device = device;
    c->index_base = index_base;
    c->ndims = ndims;
    c->f64 = elem_bytes == 8;
    if (hipSetDevice(device) != hipSuccess) { delete c; return OCT_ERR_DEVICE; }
    if (noalloc_scratch_setup(device) != hipSuccess) { delete c; return OCT_ERR_DEVICE; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return OCT_ERR_DEVICE; }
    c->own_stream = true;
    if (hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return OCT_ERR_DEVICE;
    }
    if (hipHostMalloc((void**)&c->sig, 4096, hipHostMallocCoherent) != hipSuccess) {
        (void)hipHostFree(c->pinned);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return OCT_ERR_DEVICE;
    }
    memset(c->sig, 0, 4096);
    if (hipHostGetDevicePointer((void**)&c->sig_dev, c->sig, 0) != hipSuccess || !c->sig_dev) {
        (void)hipHostFree(c->sig);
        (void)hipHostFree(c->pinned);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return OCT_ERR_DEVICE;
    }
    *out = c;
    return OCT_OK;
}

int oct_destroy(oct_ctx* c) {
    if (!c) return OCT_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& s : c->sets) s.release_all();
    for (auto& p : c->pairs) p.release_all();
    for (auto& b : c->batches) b.release_all();
    c->mlp.release_all();
    for (void* p : c->retired) (void)hipFree(p);
    c->bw_scratch.release();
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->sig) (void)hipHostFree(c->sig);
    for (auto& m : c->marks) { (void)hipEventDestroy(m.a); (void)hipEventDestroy(m.b); }
    for (auto e : c->spare) (void)hipEventDestroy(e);
    for (auto e : c->join)
        if (e) (void)hipEventDestroy(e);
    if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OCT_OK;
}

const char* oct_last_error(const oct_ctx* c) { return c ? c->err.c_str() : "null context"; }

int oct_stream_join(oct_ctx* c, void* other, int direction) {
    if (!c || (direction != 0 && direction != 1)) return OCT_ERR_ARGUMENT;
    hipStream_t o = (hipStream_t)other;
    if (o == c->stream) return OCT_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    hipEvent_t& e = c->join[direction];
    if (!e) HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // the waiting stream takes the event's state now, so the event is reusable at once
    HIP_TRY(c, hipEventRecord(e, direction == 0 ? o : c->stream));
    HIP_TRY(c, hipStreamWaitEvent(direction == 0 ? c->stream : o, e, 0));
    return OCT_OK;
}

int oct_set_stream(oct_ctx* c, void* stream) {
    if (!c) return OCT_ERR_ARGUMENT;
    const hipStream_t s = (hipStream_t)stream;
    if (s == c->stream) return OCT_OK;  // the common case of a caller that follows its current stream
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->own_stream) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipStreamDestroy(c->stream));
        c->own_stream = false;
    } else if (c->stream) {  // work queued on the previous stream stays ahead of the new one's
        hipEvent_t& e = c->join[0];
        if (!e) HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIP_TRY(c, hipEventRecord(e, c->stream));
        HIP_TRY(c, hipStreamWaitEvent(s, e, 0));
    }
    c->stream = s;
    return OCT_OK;
}

int oct_set_search_radius(oct_ctx* c, float r) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (!(r > 0.f)) return fail(c, OCT_ERR_ARGUMENT, "search radius must be positive");
    // the kernels rely on a finite r * r (empty candidate slots sit at infinity)
    if (!std::isfinite(r * r)) return fail(c, OCT_ERR_ARGUMENT, "search radius squared must be finite in Float32");
    c->radius = r;
    c->radius64 = (double)r;
    return mark_all_dirty(c);
}

int oct_set_search_radius64(oct_ctx* c, double r) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (!(r > 0.0)) return fail(c, OCT_ERR_ARGUMENT, "search radius must be positive");
    if (!std::isfinite(r * r)) return fail(c, OCT_ERR_ARGUMENT, "search radius squared must be finite");
    if (!c->f64 && !std::isfinite((float)r * (float)r))
        return fail(c, OCT_ERR_ARGUMENT, "search radius squared must be finite in Float32");
    c->radius64 = r;
    c->radius = (float)r;  // T(r) of a Float32 context
    return mark_all_dirty(c);
}

int oct_set_refit_mode(oct_ctx* c, int on) {
    if (!c) return OCT_ERR_ARGUMENT;
    c->refit_mode = on != 0;
    return mark_all_dirty(c);
}

int oct_set_leaf_size(oct_ctx* c, int leaf_size) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (leaf_size < 1 || leaf_size > 256) return fail(c, OCT_ERR_ARGUMENT, "leaf size must be in 1..256");
    c->leaf_size = leaf_size;
    return mark_all_dirty(c);
}

int oct_add_point_set(oct_ctx* c, const float* coords, int64_t n, int* set_id) {
    if (!c || !set_id) return OCT_ERR_ARGUMENT;
    if (n < 0) return fail(c, OCT_ERR_DIMENSION, "point count must be >= 0");
    if (n > 0 && !coords) return fail(c, OCT_ERR_ARGUMENT, "null coordinates");
    c->sets.emplace_back();
    c->sets.back().export_to(&c->retired);
    c->sets.back().coords = coords;
    c->sets.back().n = n;
    *set_id = (int)c->sets.size() - 1 + c->index_base;
    return OCT_OK;
}

int oct_update_point_set(oct_ctx* c, int id, const float* coords, int64_t n) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (!valid_set(c, id)) return fail(c, OCT_ERR_ARGUMENT, "unknown point set");
    SetState& s = c->sets[set_index(c, id)];
    if (n != s.n)
        return fail(c, OCT_ERR_DIMENSION, "update_point_set! expects same size; use resize_point_set! otherwise");
    s.coords = coords;
    s.dirty = true;
    return OCT_OK;
}

int oct_resize_point_set(oct_ctx* c, int id, const float* coords, int64_t n) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (!valid_set(c, id)) return fail(c, OCT_ERR_ARGUMENT, "unknown point set");
    if (n < 0) return fail(c, OCT_ERR_DIMENSION, "point count must be >= 0");
    SetState& s = c->sets[set_index(c, id)];
    s.coords = coords;
    s.n = n;
    s.dirty = true;
    return OCT_OK;
}

int oct_set_active_search(oct_ctx* c, int q, int t) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (!valid_set(c, q) || !valid_set(c, t)) return fail(c, OCT_ERR_ARGUMENT, "unknown point set");
    if (find_pair(c, q, t)) return OCT_OK;  // idempotent (src/api.jl:152)
    c->pairs.emplace_back();
    c->pairs.back().export_to(&c->retired);
    c->pairs.back().q = q;
    c->pairs.back().t = t;
    return OCT_OK;
}

int oct_run(oct_ctx* c) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (!(c->radius > 0.f)) return fail(c, OCT_ERR_ARGUMENT, "set_search_radius! before run!");
    if (c->sets.empty()) return fail(c, OCT_ERR_ARGUMENT, "add at least one point set before run!");
    HIP_TRY(c, hipSetDevice(c->device));
    for (auto& s : c->sets) {
        if (!s.dirty) continue;
        int rc = build_set(c, s);
        if (rc != OCT_OK) return rc;
    }
    for (auto& p : c->pairs) {
        p.materialized = false;
        p.csr_ok = false;  // run! invalidates materialized lists (test/contracts.jl:252-270)
    }
    return OCT_OK;
}

int oct_get_tree(oct_ctx* c, int id, oct_tree* out) {
    if (!c || !out) return OCT_ERR_ARGUMENT;
    if (!valid_set(c, id)) return fail(c, OCT_ERR_ARGUMENT, "unknown point set");
    SetState& s = c->sets[set_index(c, id)];
    if (!s.built || s.dirty) return fail(c, OCT_ERR_STATE, "run! before reading the tree");
    memset(out, 0, sizeof(*out));
    out->n_points = s.n;
    if (s.n == 0) return OCT_OK;
    int rc = ensure_hp(c, s);
    if (rc != OCT_OK) return rc;
    out->node_bounds_min = s.bmin.as<float>();
    out->node_bounds_max = s.bmax.as<float>();
    out->node_first = s.nfirst.as<int32_t>();
    out->node_last = s.nlast.as<int32_t>();
    out->node_children = s.nchild.as<int32_t>();
    out->permutation = s.perm.as<int32_t>();
    out->n_nodes = s.hp.n_nodes;
    out->n_leaves = s.hp.n_leaves;
    for (int d = 0; d < 3; ++d) out->origin[d] = s.hp.origin[d];
    if (c->f64) {  // doubles
        out->node_bounds_min = (float*)s.bmin64.p;
        out->node_bounds_max = (float*)s.bmax64.p;
    }
    if (c->ndims == 2) {  // the reference's quadtree: a one-child chain over the planar octree
        if (s.hp.cmax > 0x1fffffu)
            return fail(c, OCT_ERR_CAPACITY, "2D trees are exported for extents below 2^21 cells (radii) per axis");
        const int32_t chain = s.n > (int64_t)s.built_leaf ? PLANAR_CHAIN : 0;
        const int64_t m2 = (int64_t)s.hp.n_nodes + chain;
        HIP_TRY(c, hipSetDevice(c->device));
        HIP_TRY(c, s.qfirst.ensure(4 * (size_t)m2));
        HIP_TRY(c, s.qlast.ensure(4 * (size_t)m2));
        HIP_TRY(c, s.qkids.ensure(16 * (size_t)m2));
        HIP_TRY(c, s.qbmin.ensure((c->f64 ? 16 : 8) * (size_t)m2));
        HIP_TRY(c, s.qbmax.ensure((c->f64 ? 16 : 8) * (size_t)m2));
        if (c->f64)
            HIP_TRY(c, launch_quadtree64(s.nfirst.as<int32_t>(), s.nlast.as<int32_t>(), s.nchild.as<int32_t>(),
                                         s.bmin64.as<double>(), s.bmax64.as<double>(), s.hp.n_nodes, chain, s.n,
                                         c->index_base, s.qfirst.as<int32_t>(), s.qlast.as<int32_t>(),
                                         s.qkids.as<int32_t>(), s.qbmin.as<double>(), s.qbmax.as<double>(), c->stream));
        else
            HIP_TRY(c, launch_quadtree(s.nfirst.as<int32_t>(), s.nlast.as<int32_t>(), s.nchild.as<int32_t>(),
                                       s.bmin.as<float>(), s.bmax.as<float>(), s.hp.n_nodes, chain, s.n,
                                       c->index_base, s.qfirst.as<int32_t>(), s.qlast.as<int32_t>(),
                                       s.qkids.as<int32_t>(), s.qbmin.as<float>(), s.qbmax.as<float>(), c->stream));
        out->node_bounds_min = s.qbmin.as<float>();
        out->node_bounds_max = s.qbmax.as<float>();
        out->node_first = s.qfirst.as<int32_t>();
        out->node_last = s.qlast.as<int32_t>();
        out->node_children = s.qkids.as<int32_t>();
        out->n_nodes = (int32_t)m2;
        out->origin[2] = 0.f;
    }
    return OCT_OK;
}

constexpr size_t HF_BYTES = 128;  // flag words of a pair's edge call; the build parameters follow
constexpr int SIG_STATS = SP_SIG_MARK + 4;  // deferred recount statistics in the signal block

// A single-pass call's recount statistics set the policy of the calls after
// it: a new plan helps leaves that were recounted because a run's list filled
// up, not heavy leaves (no plan records those); uneven leaves (many not
// recorded in one sweep) keep the heavy-leaf count on, even ones switch it off.
void recount_policy(PairState& p, uint32_t recounted, int64_t heavy_n, int64_t n_leaves) {
    if (((int64_t)recounted - heavy_n) * 64 > n_leaves) p.replan = true;
    p.heavy_on = (int64_t)recounted * 64 > n_leaves;
}

// build_edges / materialize of a Float64 pair (f64.hip): one thread per
// query walks the tree in double; count, scan, then the fill at each query's
// offset (receiver order, the reference's layout).  The learned single pass
// computes in Float32, so a Float64 context takes the exact path only.
static int edges_f64(oct_ctx* c, PairState& p, SetState& qs, SetState& ts, bool self_pair, bool csr,
                     oct_edges* out) {
    if (!csr) {
        memset(&p.st, 0, sizeof(p.st));
        p.n_edges = 0;
    }
    const int64_t nq = qs.n;
    if (nq == 0 || ts.n == 0) {
        if (csr) {
            HIP_TRY(c, p.csr_off.ensure(4 * (size_t)(nq + 1)));
            HIP_TRY(c, launch_fill_u32(p.csr_off.as<uint32_t>(), nq + 1, 0u, c->stream));
            p.csr_ok = true;
            p.csr_nq = nq;
            p.csr_total = 0;
        } else {
            p.materialized = true;
        }
        return OCT_OK;
    }
    const double r = c->radius64, r_sq = r * r, inv_r = 1.0 / r;
    const double* cq = c->ndims == 2 ? qs.pad.as<double>() : (const double*)qs.coords;
    const double* ct = c->ndims == 2 ? ts.pad.as<double>() : (const double*)ts.coords;
    HIP_TRY(c, p.flags.ensure(HF_BYTES + 2 * sizeof(BuildParams)));
    HIP_TRY(c, p.qcount.ensure(4 * (size_t)nq));
    HIP_TRY(c, p.qoff.ensure(4 * (size_t)nq));
    HIP_TRY(c, p.scan_tmp.ensure(4 * (size_t)scan_tmp_words(nq)));
    const int32_t* tf = ts.nfirst.as<int32_t>();
    const int32_t* tl = ts.nlast.as<int32_t>();
    const int32_t* tk = ts.nchild.as<int32_t>();
    const int32_t* tp = ts.perm.as<int32_t>();
    const double* bl = ts.bmin64.as<double>();
    const double* bh = ts.bmax64.as<double>();
    uint64_t* total_dev = (uint64_t*)(p.flags.as<uint32_t>() + 2);
    {
        Phase ph(c, OCT_PROF_COUNT);
        HIP_TRY(c, launch_edges64(EDGE_COUNT, cq, nq, ct, bl, bh, tf, tl, tk, tp, c->index_base, r_sq, inv_r,
                                  self_pair, p.qcount.as<uint32_t>(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                  c->stream));
    }
    {
        Phase ph(c, OCT_PROF_SCAN);
        HIP_TRY(c, exclusive_scan(p.qcount.as<uint32_t>(), p.qoff.as<uint32_t>(), nq, p.scan_tmp.as<uint32_t>(),
                                  total_dev, c->stream));
    }
    HIP_TRY(c, hipMemcpyAsync(c->pinned, total_dev, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, wait_stream(c->stream));
    uint64_t total = 0;
    memcpy(&total, c->pinned, 8);
    if (total >= (uint64_t)INT32_MAX)
        return fail(c, OCT_ERR_CAPACITY, "edge count exceeds Int32 (the reference's limit)");
    const int64_t E = (int64_t)total;
    if (csr) {
        HIP_TRY(c, p.csr_off.ensure(4 * (size_t)(nq + 1)));
        HIP_TRY(c, p.csr_flat.ensure(4 * (size_t)std::max<int64_t>(E, 1)));
        HIP_TRY(c, launch_edges64(EDGE_INDEX, cq, nq, ct, bl, bh, tf, tl, tk, tp, c->index_base, r_sq, inv_r,
                                  self_pair, nullptr, p.qoff.as<uint32_t>(), p.csr_flat.as<int32_t>(), nullptr, nullptr,
                                  nullptr, c->stream));
        HIP_TRY(c, hipMemcpyAsync(p.csr_off.p, p.qoff.p, 4 * (size_t)nq, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, launch_fill_u32(p.csr_off.as<uint32_t>() + nq, 1, (uint32_t)E, c->stream));
        p.csr_ok = true;
        p.csr_nq = nq;
        p.csr_total = E;
        out->n_edges = E;
        return OCT_OK;
    }
    HIP_TRY(c, p.senders.ensure(4 * (size_t)std::max<int64_t>(E, 1)));
    HIP_TRY(c, p.receivers.ensure(4 * (size_t)std::max<int64_t>(E, 1)));
    HIP_TRY(c, p.rdisp.ensure(24 * (size_t)std::max<int64_t>(E, 1)));
    HIP_TRY(c, p.rdist.ensure(8 * (size_t)std::max<int64_t>(E, 1)));
    {
        Phase ph(c, OCT_PROF_FILL);
        HIP_TRY(c, launch_edges64(EDGE_FILL, cq, nq, ct, bl, bh, tf, tl, tk, tp, c->index_base, r_sq, inv_r,
                                  self_pair, nullptr, p.qoff.as<uint32_t>(), p.senders.as<int32_t>(),
                                  p.receivers.as<int32_t>(), p.rdisp.as<double>(), p.rdist.as<double>(), c->stream));
    }
    p.n_edges = E;
    p.materialized = true;
    p.st.n_edges = E;
    p.st.attempts = 1;
    out->senders = p.senders.as<int32_t>();
    out->receivers = p.receivers.as<int32_t>();
    out->rel_displacement = (float*)p.rdisp.p;  // doubles in a Float64 context
    out->rel_dist_norm = (float*)p.rdist.p;
    out->n_edges = E;
    if (c->ndims == 2) {
        HIP_TRY(c, p.rdisp2.ensure(16 * (size_t)std::max<int64_t>(E, 1)));
        HIP_TRY(c, launch_disp_planar64(p.rdisp.as<double>(), E, p.rdisp2.as<double>(), c->stream));
        out->rel_displacement = (float*)p.rdisp2.p;
    }
    return OCT_OK;
}

// What an edges_impl call produces: the flat-COO edges; only the counts with
// the predictor's features and the exact per-leaf hits (training rows); or
// CSR neighbour lists (targets only, in query order).
enum Emit { EMIT_EDGES = 0, EMIT_COLLECT = 1, EMIT_CSR = 2 };

// Flat-COO edges (or counts, or CSR lists) of pair (q, t).
static int edges_impl(oct_ctx* c, int q, int t, int mode, int order, int emit, oct_edges* out) {
    const bool collect = emit == EMIT_COLLECT;
    const bool csr = emit == EMIT_CSR;
    if (!c || !out) return OCT_ERR_ARGUMENT;
    PairState* pp = find_pair(c, q, t);
    if (!pp) return fail(c, OCT_ERR_STATE, "build_edges!: pair not registered; call set_active_search! first");
    if (mode != OCT_EDGES_EXACT && mode != OCT_EDGES_LEARNED) return fail(c, OCT_ERR_ARGUMENT, "unknown edge mode");
    if (order != OCT_ORDER_LEAF && order != OCT_ORDER_RECEIVER) return fail(c, OCT_ERR_ARGUMENT, "unknown edge order");
    if (mode == OCT_EDGES_LEARNED && order != OCT_ORDER_LEAF)
        return fail(c, OCT_ERR_ARGUMENT, "the single pass writes leaf order; receiver order needs the exact mode");
    PairState& p = *pp;
    SetState& qs = c->sets[set_index(c, q)];
    SetState& ts = c->sets[set_index(c, t)];
    if (emit == EMIT_EDGES && (!qs.built || !ts.built)) {
        // never run: the tree is empty and the traversal finds no edges, as
        // the reference's does (test/backend_errors_cpu.jl:40-57)
        HIP_TRY(c, hipSetDevice(c->device));
        const size_t w = c->f64 ? 8 : 4;
        HIP_TRY(c, p.senders.ensure(4));
        HIP_TRY(c, p.receivers.ensure(4));
        HIP_TRY(c, p.rdisp.ensure(3 * w));
        HIP_TRY(c, p.rdist.ensure(w));
        memset(out, 0, sizeof(*out));
        out->senders = p.senders.as<int32_t>();
        out->receivers = p.receivers.as<int32_t>();
        out->rel_displacement = (float*)p.rdisp.p;
        out->rel_dist_norm = (float*)p.rdist.p;
        p.n_edges = 0;
        p.materialized = true;
        return OCT_OK;
    }
    if (!qs.built || qs.dirty || !ts.built || ts.dirty) return fail(c, OCT_ERR_STATE, "run! before build_edges");
    HIP_TRY(c, hipSetDevice(c->device));
    memset(out, 0, sizeof(*out));
    if (c->f64) {
        if (mode != OCT_EDGES_EXACT)
            return fail(c, OCT_ERR_ARGUMENT, "the learned single pass computes in Float32; Float64 sets use the exact mode");
        if (collect) return fail(c, OCT_ERR_STATE, "the predictor's training rows are Float32 only");
        return edges_f64(c, p, qs, ts, q == t, csr, out);
    }
    // the previous call's recount statistics (final once its kernels ended)
    // come back with this call's read-back, queued ahead of everything else
    // (a CSR call leaves the edge buffers' state and statistics alone)
    const bool prev_pending = !csr && p.stats_pending;
    if (prev_pending)
        HIP_TRY(c, hipMemcpyAsync(c->sig + SIG_STATS, p.flags.as<uint32_t>() + 9, 8, hipMemcpyDeviceToHost,
                                  c->stream));
    if (!csr) {
        p.stats_pending = false;
        memset(&p.st, 0, sizeof(p.st));
        p.n_edges = 0;
    }
    if (qs.n == 0 || ts.n == 0) {
        if (csr) {  // every query has no neighbour: offsets all zero
            HIP_TRY(c, p.csr_off.ensure(4 * (size_t)(qs.n + 1)));
            HIP_TRY(c, launch_fill_u32(p.csr_off.as<uint32_t>(), qs.n + 1, 0u, c->stream));
            p.csr_ok = true;
            p.csr_nq = qs.n;
            p.csr_total = 0;
            return OCT_OK;
        }
        p.materialized = true;
        return OCT_OK;
    }
    const int self_pair = q == t;
    const float r = c->radius;
    const float r_sq = r * r;
    const float inv_r = 1.0f / r;
    // the predictor runs on the pair's first call and after a call that
    // overflowed or recounted many leaves (or on every call, if asked)
    const bool predict = mode == OCT_EDGES_LEARNED && emit == EMIT_EDGES && c->mlp.loaded &&
                         (c->predict_always || p.cap_edges <= 0 || p.replan);
    const bool features = predict || collect;
    // Rows of the per-leaf arrays: the exact leaf count when it is on the
    // host, else the last one read back plus a margin.  The kernels see the
    // query tree's parameters through a copy whose leaf count is clamped to
    // the rows (flags[18] = 1 when it was), so a tree that grew past the
    // margin is redone with the count read back at this call's wait.
    // Features and the plan need the exact count: they wait for it.
    if (features || qs.leaves_known < 0) HIP_TRY_RC(ensure_hp(c, qs));
    auto rows_for = [&]() -> int64_t {
        if (qs.hp_fresh) return std::max<int64_t>(qs.hp.n_leaves, 1);
        return std::min<int64_t>(qs.n, qs.leaves_known + qs.leaves_known / 4 + 1024);
    };
    int64_t nql = rows_for();
    HIP_TRY(c, p.flags.ensure(HF_BYTES + 2 * sizeof(BuildParams)));
    HIP_TRY(c, p.eprm.ensure(sizeof(BuildParams)));
    HIP_TRY(c, p.qcount.ensure(4 * (size_t)qs.n));
    HIP_TRY(c, p.qoff.ensure(4 * (size_t)qs.n));
    HIP_TRY(c, p.scan_tmp.ensure(4 * (size_t)scan_tmp_words(qs.n)));
    const int by_receiver = order == OCT_ORDER_RECEIVER || csr;
    // flags (uint32 words): [0] largest candidate count of an overflowing
    // query leaf, [1] a list walk ran out of stack, [2..3] exact-path edge
    // total, [4] single pass over capacity, [5] single-pass look-back gave up
    // or malformed plan, [6..7] single-pass edge total, [8] single-pass tile
    // ticket, [9] recounted leaves, [10] recorded leaves, [12] plan runs,
    // [14..15] predicted total; [16] heavy leaves (kept while the lists are)
    // [18] more query leaves than rows (launch_clamp_leaves); the query and
    // target build parameters follow at byte HF_BYTES
    uint32_t* flags = p.flags.as<uint32_t>();
    constexpr int HF_WORDS = 19;
    uint32_t hf[HF_WORDS] = {0};
    BuildParams* raw_prm = (BuildParams*)((char*)p.flags.p + HF_BYTES);
    int safe = 0;
    bool single = mode == OCT_EDGES_LEARNED && emit == EMIT_EDGES && (p.cap_edges > 0 || predict);
    const bool heavy = single && p.heavy_on;
    bool lists_ok = false;
    uint64_t total = 0;
    int attempts = 0;
    EdgeOut eo{};
    // the single pass hands its total over early with fixed runs (a plan's
    // runs are checked by each run, so a plan waits for the kernels to end)
    const bool early = !predict && c->sig != nullptr;
    bool early_done = false;  // the last read-back came from the single pass, not after the kernels
    TreeView qv{}, tv{};
    for (int attempt = 0; attempt < 10; ++attempt) {
        ++attempts;
        const int64_t ntiles = single_tiles(nql);
        HIP_TRY(c, p.lcount.ensure(4 * (size_t)nql));
        HIP_TRY(c, p.tiles.ensure(8 * (size_t)(ntiles + 1)));
        HIP_TRY(c, p.lhits.ensure(4 * (size_t)nql));
        if (heavy) HIP_TRY(c, p.heavy.ensure(5 * (size_t)nql));
        if (features) HIP_TRY(c, p.feat.ensure(4 * (size_t)nql * MLP_IN));
        if (predict) {
            HIP_TRY(c, p.pred.ensure(4 * (size_t)nql));
            HIP_TRY(c, p.pwork.ensure(16 * (size_t)plan_chunks(nql)));
            HIP_TRY(c, p.pscan.ensure(4 * (size_t)scan_tmp_words(plan_chunks(nql))));
            HIP_TRY(c, p.wave_first.ensure(4 * (size_t)(nql + 2)));
        }
        // the same launch clears the flag words (the heavy-leaf count at [16]
        // too when the lists are rebuilt) and the single pass's run states
        HIP_TRY(c, launch_clamp_leaves(qs.prm.as<BuildParams>(), ts.prm.as<BuildParams>(), p.eprm.as<BuildParams>(),
                                       nql, flags + 18, raw_prm, flags, lists_ok ? 16 : 17,
                                       single ? (uint32_t*)p.tiles.p : nullptr, single ? 2 * (ntiles + 1) : 0,
                                       c->stream));
        qv = view_of(qs, nql, p.eprm.as<BuildParams>());
        tv = view_of(ts, ts.hp_fresh ? ts.hp.n_leaves : ts.n);
        HIP_TRY(c, p.lists.ensure(sizeof(int2) * (size_t)nql * (size_t)p.cap));
        if (features) HIP_TRY(c, p.ldir.ensure((size_t)nql * (size_t)p.cap));
        if (!lists_ok) {
            Phase ph(c, OCT_PROF_LEAF_LISTS);
            HIP_TRY(c, launch_leaf_lists(qv, tv, r_sq, p.cap, safe, p.lists.as<int2>(),
                                         features ? p.ldir.as<uint8_t>() : nullptr, p.lcount.as<int32_t>(), flags,
                                         SetLookup{}, heavy ? p.heavy_flags(nql) : nullptr,
                                         heavy ? p.heavy_list() : nullptr, flags + 16, c->stream));
        }
        if (features) {
            Phase ph(c, OCT_PROF_PREDICT);
            HIP_TRY(c, launch_leaf_features(qv, p.lists.as<int2>(), p.ldir.as<uint8_t>(), p.lcount.as<int32_t>(),
                                            p.cap, r, p.feat.as<float>(), c->stream));
            if (predict && single)
                HIP_TRY(c, launch_predict_plan(p.feat.as<float>(), nql, c->mlp.params.as<float>(),
                                               p.pred.as<float>(), SP_PLAN_BUDGET, SP_PLAN_MAX_LEAVES,
                                               p.pwork.as<uint32_t>(), p.pscan.as<uint32_t>(),
                                               p.wave_first.as<int32_t>(), (int32_t*)(flags + 12),
                                               (uint64_t*)(flags + 14), heavy ? p.heavy_flags(nql) : nullptr,
                                               c->stream));
        }
        if (single && p.cap_edges <= 0) {
            // first call on the pair: size the outputs from the predicted total
            HIP_TRY(c, hipMemcpyAsync(c->pinned, flags, 64, hipMemcpyDeviceToHost, c->stream));
            {
                HostTimer ht("edges: wait for predicted total");
                HIP_TRY(c, hipStreamSynchronize(c->stream));
            }
            memcpy(hf, c->pinned, 64);  // (features: the leaf count was exact already)
            uint64_t pt = 0;
            memcpy(&pt, hf + 14, 8);
            p.cap_edges = (int64_t)std::min<uint64_t>(pt + pt / 16 + 4096, (uint64_t)INT32_MAX - 1);
        }
        if (single) {
            HIP_TRY(c, p.senders.ensure(4 * (size_t)p.cap_edges));
            HIP_TRY(c, p.receivers.ensure(4 * (size_t)p.cap_edges));
            HIP_TRY(c, p.rdisp.ensure(12 * (size_t)p.cap_edges));
            HIP_TRY(c, p.rdist.ensure(4 * (size_t)p.cap_edges));
            eo = EdgeOut{p.senders.as<int32_t>(), p.receivers.as<int32_t>(), p.rdisp.as<float>(), p.rdist.as<float>()};
            Phase ph(c, OCT_PROF_SINGLE);
            if (heavy)
                HIP_TRY(c, launch_edges_dense_count(qv, tv, p.lists.as<int2>(), p.lcount.as<int32_t>(), p.cap,
                                                    r_sq, inv_r, self_pair, p.heavy_list(), flags + 16,
                                                    p.qcount.as<uint32_t>(), p.lhits.as<uint32_t>(), SetLookup{},
                                                    c->stream));
            if (early) {  // cleared before the launch; the single pass's last run sets it
                __atomic_store_n((uint64_t*)(c->sig + SP_SIG_MARK), 0ull, __ATOMIC_RELEASE);
            }
            HIP_TRY(c, launch_edges_single(qv, tv, p.lists.as<int2>(), p.lcount.as<int32_t>(), p.cap, r_sq,
                                           inv_r, self_pair, p.qcount.as<uint32_t>(), p.qoff.as<uint32_t>(),
                                           predict ? p.wave_first.as<int32_t>() : nullptr,
                                           predict ? (const int32_t*)(flags + 12) : nullptr,
                                           p.lhits.as<uint32_t>(), p.tiles.as<uint64_t>(), flags + 8, flags + 4,
                                           flags + 9, (uint64_t)p.cap_edges, eo, c->index_base, c->run_leaves,
                                           c->run_chunk, SetLookup{}, heavy ? p.heavy_flags(nql) : nullptr,
                                           early ? c->sig_dev : nullptr, flags,
                                           (int)((HF_BYTES + 2 * sizeof(BuildParams)) / 4), c->stream));
            if (heavy) {  // one block per heavy leaf
                HIP_TRY(c, launch_edges_dense_fill(qv, tv, p.lists.as<int2>(), p.lcount.as<int32_t>(), p.cap, r_sq,
                                                   inv_r, self_pair, p.heavy_list(), flags + 16,
                                                   p.qcount.as<uint32_t>(), p.qoff.as<uint32_t>(), flags + 4, eo,
                                                   c->index_base, SetLookup{}, c->stream));
            }
        } else {
            {
                Phase ph(c, OCT_PROF_COUNT);
                HIP_TRY(c, launch_edges(EDGE_COUNT, qv, tv, p.lists.as<int2>(), p.lcount.as<int32_t>(), p.cap, r_sq,
                                        inv_r, self_pair, by_receiver, p.qcount.as<uint32_t>(), nullptr,
                                        p.lhits.as<uint32_t>(), 0, EdgeOut{}, c->index_base, c->stream));
            }
            Phase ph(c, OCT_PROF_SCAN);
            HIP_TRY(c, exclusive_scan(p.qcount.as<uint32_t>(), p.qoff.as<uint32_t>(), qs.n,
                                      p.scan_tmp.as<uint32_t>(), (uint64_t*)(flags + 2), c->stream));
        }
        // one read-back: flags and both sets' build parameters, either handed
        // over by the single pass as soon as its total is known (the call then
        // returns while the rest of the pass runs) or copied after the kernels
        const uint8_t* rb = (const uint8_t*)c->pinned;
        bool got = false;
        if (single && early) {
            HostTimer ht("edges: wait for the single pass's total");
            const uint64_t* mark = (const uint64_t*)(c->sig + SP_SIG_MARK);
            for (;;) {
                if (__atomic_load_n(mark, __ATOMIC_ACQUIRE) != 0) {
                    got = true;
                    break;
                }
                const hipError_t qe = hipStreamQuery(c->stream);
                if (qe == hipSuccess) {  // the kernels ended: the marker is there or was never written
                    got = __atomic_load_n(mark, __ATOMIC_ACQUIRE) != 0;
                    break;
                }
                if (qe != hipErrorNotReady) HIP_TRY(c, qe);
                __builtin_ia32_pause();
            }
            if (got) rb = (const uint8_t*)c->sig;
        }
        if (!got) {
            HIP_TRY(c, hipMemcpyAsync(c->pinned, flags, HF_BYTES + 2 * sizeof(BuildParams), hipMemcpyDeviceToHost,
                                      c->stream));
            HostTimer ht(single ? "edges: wait for single pass" : "edges: wait for count total");
            HIP_TRY(c, wait_stream(c->stream));
        }
        early_done = got;
        memcpy(hf, rb, 4 * HF_WORDS);
        take_hp(qs, rb + HF_BYTES);
        if (&ts != &qs) take_hp(ts, rb + HF_BYTES + sizeof(BuildParams));
        bool rq = false, rt = false;
        HIP_TRY_RC(settle_refit(c, qs, &rq));
        if (&ts != &qs) HIP_TRY_RC(settle_refit(c, ts, &rt));
        if (rq || rt) {  // a refit could not run: the sets were rebuilt, redo the edges
            nql = rows_for();
            lists_ok = false;
            continue;
        }
        if (hf[18] != 0) {  // more leaves than rows: redo everything with the count now known
            nql = rows_for();
            lists_ok = false;
            continue;
        }
        if (hf[0] != 0 || hf[1] != 0) {  // candidate lists incomplete: regrow / safe walk, redo everything
            if (hf[0] != 0) p.cap = (int)std::min<uint32_t>(hf[0] + 16u, (uint32_t)INT32_MAX / 8u);
            if (hf[1] != 0) safe = 1;
            lists_ok = false;
            continue;
        }
        lists_ok = true;
        if (!single) {
            memcpy(&total, hf + 2, 8);
            break;
        }
        memcpy(&total, hf + 6, 8);
        if (hf[5] != 0) {  // look-back did not finish or the plan was malformed: take the exact path
            single = false;
            continue;
        }
        if (hf[4] != 0) {  // room exceeded: grow to the reported total and rerun
            if (total >= (uint64_t)INT32_MAX) break;
            p.cap_edges = (int64_t)(total + total / 8 + 1024);
            continue;
        }
        break;
    }
    if (!lists_ok) return fail(c, OCT_ERR_DEVICE, "candidate lists did not settle");
    // Int32 indices and counts, as in the reference (src/types.jl:89)
    if (total >= (uint64_t)INT32_MAX)
        return fail(c, OCT_ERR_CAPACITY, "edge count exceeds Int32 (the reference's limit)");
    const int64_t E = (int64_t)total;
    if (csr) {  // targets only, each query's slice at its scanned offset
        HIP_TRY(c, p.csr_off.ensure(4 * (size_t)(qs.n + 1)));
        HIP_TRY(c, p.csr_flat.ensure(4 * (size_t)std::max<int64_t>(E, 1)));
        if (E > 0)
            HIP_TRY(c, launch_edges(EDGE_INDEX, qv, tv, p.lists.as<int2>(), p.lcount.as<int32_t>(), p.cap, r_sq, inv_r,
                                    self_pair, 1, p.qcount.as<uint32_t>(), p.qoff.as<uint32_t>(), nullptr, E,
                                    EdgeOut{p.csr_flat.as<int32_t>(), nullptr, nullptr, nullptr}, c->index_base,
                                    c->stream));
        HIP_TRY(c, hipMemcpyAsync(p.csr_off.p, p.qoff.p, 4 * (size_t)qs.n, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, launch_fill_u32(p.csr_off.as<uint32_t>() + qs.n, 1, (uint32_t)E, c->stream));
        p.csr_ok = true;
        p.csr_nq = qs.n;
        p.csr_total = E;
        out->n_edges = E;
        return OCT_OK;
    }
    if (collect) {  // counts only: the training rows need no edges
        p.n_edges = 0;
        p.materialized = false;
        p.st.features = p.feat.as<float>();
        p.st.exact = p.lhits.as<uint32_t>();
        p.st.n_leaves = qs.hp.n_leaves;
        p.st.n_edges = E;
        p.st.attempts = attempts;
        out->n_edges = E;
        return OCT_OK;
    }
    if (!single) {
        HostTimer ht_fill("edges: allocate + launch fill");
        HIP_TRY(c, p.senders.ensure(4 * (size_t)E));
        HIP_TRY(c, p.receivers.ensure(4 * (size_t)E));
        HIP_TRY(c, p.rdisp.ensure(12 * (size_t)E));
        HIP_TRY(c, p.rdist.ensure(4 * (size_t)E));
        eo = EdgeOut{p.senders.as<int32_t>(), p.receivers.as<int32_t>(), p.rdisp.as<float>(), p.rdist.as<float>()};
        if (E > 0) {
            Phase ph(c, OCT_PROF_FILL);
            HIP_TRY(c, launch_edges(EDGE_FILL, qv, tv, p.lists.as<int2>(), p.lcount.as<int32_t>(), p.cap, r_sq,
                                    inv_r, self_pair, by_receiver, p.qcount.as<uint32_t>(), p.qoff.as<uint32_t>(),
                                    nullptr, E, eo, c->index_base, c->stream));
        }
    }
    // room for the single pass next time: the senders buffer's capacity
    p.cap_edges = (int64_t)std::min<size_t>(p.senders.cap / 4, (size_t)INT32_MAX - 1);
    p.n_edges = E;
    p.materialized = true;
    uint64_t pt = 0;
    memcpy(&pt, hf + 14, 8);
    p.st.features = features ? p.feat.as<float>() : nullptr;
    p.st.predicted = predict && single ? p.pred.as<float>() : nullptr;
    p.st.exact = p.lhits.as<uint32_t>();
    const int64_t n_leaves = qs.hp.n_leaves;  // read back at the last wait
    p.st.n_leaves = n_leaves;
    p.st.n_edges = E;
    // an early total leaves the recount statistics for later (edge_stats or
    // the next call), when the kernels have ended
    p.st.recounted_leaves = single && !early_done ? hf[9] : 0;
    p.st.recorded_leaves = single && !early_done ? hf[10] : 0;
    p.st.heavy_leaves = heavy && single ? (int32_t)hf[16] : 0;
    p.st.predicted_total = predict && single ? (double)pt : 0.0;
    p.st.single_pass = single ? 1 : 0;
    p.st.used_predictor = predict && single ? 1 : 0;
    p.st.attempts = attempts;
    if (mode == OCT_EDGES_LEARNED) p.replan = !single || attempts > 1;
    // the previous call's statistics came back with this read-back (older,
    // so applied first), then this call's when they are final
    if (prev_pending) recount_policy(p, c->sig[SIG_STATS], p.pending_heavy, p.pending_leaves);
    if (mode == OCT_EDGES_LEARNED && single) {
        const int64_t heavy_n = heavy ? (int64_t)hf[16] : 0;
        if (!early_done) {
            recount_policy(p, hf[9], heavy_n, n_leaves);
        } else {
            p.stats_pending = true;
            p.pending_heavy = heavy_n;
            p.pending_leaves = n_leaves;
        }
    }
    out->senders = eo.senders;
    out->receivers = eo.receivers;
    out->rel_displacement = eo.rdisp;
    out->rel_dist_norm = eo.rdist;
    out->n_edges = E;
    if (c->ndims == 2) {  // rel_displacement (2, E): the z column of the planar build dropped
        HIP_TRY(c, p.rdisp2.ensure(8 * (size_t)std::max<int64_t>(E, 1)));
        HIP_TRY(c, launch_disp_planar(eo.rdisp, E, p.rdisp2.as<float>(), c->stream));
        out->rel_displacement = p.rdisp2.as<float>();
    }
    return OCT_OK;
}

int oct_build_edges(oct_ctx* c, int q, int t, int mode, int order, oct_edges* out) {
    return edges_impl(c, q, t, mode, order, EMIT_EDGES, out);
}

static void neighbors_out(const PairState& p, oct_neighbors* out) {
    memset(out, 0, sizeof(*out));
    if (!p.csr_ok) return;
    out->offsets = p.csr_off.as<int32_t>();
    out->flat = p.csr_flat.as<int32_t>();
    out->n_queries = p.csr_nq;
    out->total = p.csr_total;
    out->materialized = 1;
}

int oct_materialize_neighbors(oct_ctx* c, int q, int t, oct_neighbors* out) {
    if (!c || !out) return OCT_ERR_ARGUMENT;
    oct_edges e{};
    const int rc = edges_impl(c, q, t, OCT_EDGES_EXACT, OCT_ORDER_RECEIVER, EMIT_CSR, &e);
    if (rc != OCT_OK) return rc;
    neighbors_out(*find_pair(c, q, t), out);
    return OCT_OK;
}

int oct_get_neighbors(oct_ctx* c, int q, int t, oct_neighbors* out) {
    if (!c || !out) return OCT_ERR_ARGUMENT;
    PairState* p = find_pair(c, q, t);
    if (!p) return fail(c, OCT_ERR_STATE, "neighbours: pair not registered; call set_active_search! first");
    neighbors_out(*p, out);
    return OCT_OK;
}

int oct_device_view(oct_ctx* c, int q, int t, oct_view* out) {
    if (!c || !out) return OCT_ERR_ARGUMENT;
    if (!find_pair(c, q, t)) return fail(c, OCT_ERR_STATE, "device_view: pair not registered");
    SetState& qs = c->sets[set_index(c, q)];
    SetState& ts = c->sets[set_index(c, t)];
    if (!qs.built || qs.dirty || !ts.built || ts.dirty) return fail(c, OCT_ERR_STATE, "run! before device_view");
    if (c->f64) return fail(c, OCT_ERR_STATE, "device_view: the device iterator walks Float32 trees");
    int rc = ensure_hp(c, ts);
    if (rc != OCT_OK) return rc;
    memset(out, 0, sizeof(*out));
    // 2D: the padded (x, y, PLANE_Z) copies, the layout the device iterator reads
    out->coords_q = c->ndims == 2 ? qs.pad.as<float>() : qs.coords;
    out->coords_t = c->ndims == 2 ? ts.pad.as<float>() : ts.coords;
    out->bounds_min = ts.bmin.as<float>();
    out->bounds_max = ts.bmax.as<float>();
    out->node_first = ts.nfirst.as<int32_t>();
    out->node_last = ts.nlast.as<int32_t>();
    out->node_children = ts.nchild.as<int32_t>();
    out->perm = ts.perm.as<int32_t>();
    out->nodes = ts.nodes.p;
    out->sorted_t = ts.scoord.as<float>();
    out->n_nodes = ts.n == 0 ? 0 : ts.hp.n_nodes;
    out->root = ts.hp.branch_root;
    out->radius = c->radius;
    out->radius_sq = c->radius * c->radius;
    out->n_q = qs.n;
    out->index_base = c->index_base;
    return OCT_OK;
}

}  // extern "C"

__global__ void k_count_neighbors_device(oct_view v, int32_t* counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n_q) return;
    int32_t c = 0;
    oct_for_each_neighbor_device(v, i, [&](int32_t) { ++c; });
    counts[i] = c;
}

extern "C" {

static int zsort_impl(oct_ctx* c, int id, const void* src, void* dst, int64_t width, int elem_bytes,
                      bool scatter) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (!valid_set(c, id)) return fail(c, OCT_ERR_ARGUMENT, "unknown point set");
    if (width < 1) return fail(c, OCT_ERR_ARGUMENT, "apply_zsort: width must be >= 1");
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        return fail(c, OCT_ERR_ARGUMENT, "apply_zsort: element size must be 1, 2, 4 or 8 bytes");
    SetState& s = c->sets[set_index(c, id)];
    // prepare_zsort (src/api.jl:381-388): the permutation must be current
    if (!s.built || s.dirty) return fail(c, OCT_ERR_STATE, "apply_zsort: call run! first so the permutation is available");
    if (s.n == 0) return OCT_OK;
    if (!src || !dst) return fail(c, OCT_ERR_ARGUMENT, "apply_zsort: null array");
    if (src == dst) return fail(c, OCT_ERR_ARGUMENT, "apply_zsort: the result must be a new array");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, launch_zsort(src, dst, s.perm.as<int32_t>(), s.n, width, elem_bytes, c->index_base, scatter, c->stream));
    return OCT_OK;
}

int oct_apply_zsort(oct_ctx* c, int set_id, const void* in, void* out, int64_t width, int elem_bytes) {
    return zsort_impl(c, set_id, in, out, width, elem_bytes, false);
}

int oct_apply_zsort_adjoint(oct_ctx* c, int set_id, const void* d_out, void* grad, int64_t width, int elem_bytes) {
    return zsort_impl(c, set_id, d_out, grad, width, elem_bytes, true);
}

int oct_count_neighbors_device(oct_ctx* c, const oct_view* v, int32_t* counts) {
    if (!c || !v) return OCT_ERR_ARGUMENT;
    if (v->n_q == 0) return OCT_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    hipLaunchKernelGGL(k_count_neighbors_device, dim3((unsigned)((v->n_q + 127) / 128)), dim3(128), 0, c->stream,
                       *v, counts);
    HIP_TRY(c, hipGetLastError());
    return OCT_OK;
}

static int backward_impl(oct_ctx* c, int64_t n_points, const oct_edges* e, const void* d_disp, const void* d_dist,
                         double radius, void* grad) {
    if (!c || !e) return OCT_ERR_ARGUMENT;
    if (n_points < 0 || e->n_edges < 0) return fail(c, OCT_ERR_ARGUMENT, "negative size");
    if (n_points > 0 && !grad) return fail(c, OCT_ERR_ARGUMENT, "null gradient buffer");
    HIP_TRY(c, hipSetDevice(c->device));
    // the pull form's run table (edge indices must fit int32, as the reference's)
    uint32_t* scratch = nullptr;
    if (n_points > 0 && e->n_edges > 0 && e->n_edges < (int64_t)INT32_MAX && c->pull_backward) {
        HIP_TRY(c, c->bw_scratch.ensure(4 * (size_t)pullback_scratch_words(n_points)));
        scratch = c->bw_scratch.as<uint32_t>();
    }
    Phase ph(c, OCT_PROF_BACKWARD);
    if (c->f64)  // every array holds doubles
        HIP_TRY(c, launch_pullback64(n_points, e->n_edges, e->senders, e->receivers,
                                     (const double*)e->rel_displacement, (const double*)e->rel_dist_norm,
                                     (const double*)d_disp, (const double*)d_dist, radius, c->index_base,
                                     (double*)grad, scratch, c->ndims, c->stream));
    else
        HIP_TRY(c, launch_pullback(n_points, e->n_edges, e->senders, e->receivers, e->rel_displacement,
                                   e->rel_dist_norm, (const float*)d_disp, (const float*)d_dist, (float)radius,
                                   c->index_base, (float*)grad, scratch, c->ndims, c->stream));
    return OCT_OK;
}

int oct_build_edges_backward(oct_ctx* c, int64_t n_points, const oct_edges* e, const float* d_disp,
                             const float* d_dist, float radius, float* grad) {
    if (c && c->f64)
        return fail(c, OCT_ERR_ARGUMENT, "a Float64 context takes oct_build_edges_backward64 (double radius)");
    return backward_impl(c, n_points, e, d_disp, d_dist, radius, grad);
}

int oct_build_edges_backward64(oct_ctx* c, int64_t n_points, const oct_edges* e, const double* d_disp,
                               const double* d_dist, double radius, double* grad) {
    if (c && !c->f64) return fail(c, OCT_ERR_ARGUMENT, "oct_build_edges_backward64 needs a Float64 context");
    return backward_impl(c, n_points, e, d_disp, d_dist, radius, grad);
}

int oct_set_backward_form(oct_ctx* c, int pull) {
    if (!c) return OCT_ERR_ARGUMENT;
    c->pull_backward = pull != 0;
    return OCT_OK;
}


int oct_profile_phases(oct_ctx* c, uint32_t mask) {
    if (!c) return OCT_ERR_ARGUMENT;
    c->prof_mask = mask & ((1u << OCT_PROF_NPHASES) - 1u);
    // events are created up front so recording them costs no allocation
    while (c->prof_mask && c->spare.size() < 256) {
        hipEvent_t e = nullptr;
        HIP_TRY(c, hipEventCreate(&e));
        c->spare.push_back(e);
    }
    return OCT_OK;
}

int oct_profile(oct_ctx* c, int enable) { return oct_profile_phases(c, enable ? ~0u : 0u); }

int oct_profile_read(oct_ctx* c, int reset, double* ms, int64_t* calls) {
    if (!c || !ms || !calls) return OCT_ERR_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto& m : c->marks) {
        float t = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&t, m.a, m.b));
        c->prof_ms[m.phase] += t;
        c->prof_calls[m.phase] += 1;
        c->spare.push_back(m.a);
        c->spare.push_back(m.b);
    }
    c->marks.clear();
    for (int k = 0; k < OCT_PROF_NPHASES; ++k) {
        ms[k] = c->prof_ms[k];
        calls[k] = c->prof_calls[k];
        if (reset) { c->prof_ms[k] = 0; c->prof_calls[k] = 0; }
    }
    return OCT_OK;
}

int oct_synchronize(oct_ctx* c) {
    if (!c) return OCT_ERR_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return OCT_OK;
}

int oct_trim(oct_ctx* c) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (c->retired.empty()) return OCT_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));  // no queued kernel still reads them
    for (void* p : c->retired) HIP_TRY(c, hipFree(p));
    c->retired.clear();
    return OCT_OK;
}

int oct_retired_count(const oct_ctx* c, int64_t* count) {
    if (!c || !count) return OCT_ERR_ARGUMENT;
    *count = (int64_t)c->retired.size();
    return OCT_OK;
}

// ------------------------------------------------------------ predictor

namespace {

// Grows a device buffer keeping its first `keep` bytes.
int grow_keep(oct_ctx* c, DevBuf& b, size_t bytes, size_t keep) {
    if (bytes <= b.cap && b.p) return OCT_OK;
    DevBuf nb;
    HIP_TRY(c, nb.ensure(bytes));
    if (keep > 0) {
        HIP_TRY(c, hipMemcpyAsync(nb.p, b.p, keep, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    b.release();
    b = nb;
    return OCT_OK;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// He-normal weights (Box-Muller over splitmix64), zero biases, identity
// feature statistics (the caller replaces them).
void he_init(std::vector<float>& P, uint64_t seed) {
    P.assign(MLP_PARAMS, 0.f);
    uint64_t s = seed ^ 0x6F63746F70757321ull;
    auto normal = [&]() {
        const double u1 = ((splitmix64(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        const double u2 = (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < MLP_IN; ++i) P[MLP_SD + i] = 1.f;
    for (int i = 0; i < MLP_IN * MLP_H; ++i) P[MLP_W1 + i] = (float)(normal() * sqrt(2.0 / MLP_IN));
    for (int i = 0; i < MLP_H * MLP_H; ++i) P[MLP_W2 + i] = (float)(normal() * sqrt(2.0 / MLP_H));
    for (int i = 0; i < MLP_H; ++i) P[MLP_W3 + i] = (float)(normal() * sqrt(1.0 / MLP_H));
}

}  // namespace

int oct_predictor_load(oct_ctx* c, const float* host_params) {
    if (!c) return OCT_ERR_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    if (!host_params) {
        c->mlp.loaded = false;
        return OCT_OK;
    }
    for (int i = 0; i < MLP_PARAMS; ++i)
        if (!std::isfinite(host_params[i])) return fail(c, OCT_ERR_ARGUMENT, "predictor parameters must be finite");
    for (int i = 0; i < MLP_IN; ++i)
        if (!(host_params[MLP_SD + i] > 0.f)) return fail(c, OCT_ERR_ARGUMENT, "feature std must be positive");
    HIP_TRY(c, c->mlp.params.ensure(4 * MLP_PARAMS));
    HIP_TRY(c, hipMemcpyAsync(c->mlp.params.p, host_params, 4 * MLP_PARAMS, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->mlp.loaded = true;
    c->mlp.adam_step = 0;
    return OCT_OK;
}

int oct_predictor_save(oct_ctx* c, float* host_params) {
    if (!c || !host_params) return OCT_ERR_ARGUMENT;
    if (!c->mlp.loaded) return fail(c, OCT_ERR_STATE, "no predictor installed");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(host_params, c->mlp.params.p, 4 * MLP_PARAMS, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return OCT_OK;
}

int oct_predictor_collect(oct_ctx* c, int q, int t, int64_t* rows) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (rows) *rows = 0;
    oct_edges e{};
    int rc = edges_impl(c, q, t, OCT_EDGES_EXACT, OCT_ORDER_LEAF, EMIT_COLLECT, &e);
    if (rc != OCT_OK) return rc;
    PairState& p = *find_pair(c, q, t);
    const int64_t n = p.st.features ? p.st.n_leaves : 0;
    if (n == 0) return OCT_OK;
    Predictor& m = c->mlp;
    rc = grow_keep(c, m.X, 4 * (size_t)(m.rows + n) * MLP_IN, 4 * (size_t)m.rows * MLP_IN);
    if (rc != OCT_OK) return rc;
    rc = grow_keep(c, m.T, 4 * (size_t)(m.rows + n), 4 * (size_t)m.rows);
    if (rc != OCT_OK) return rc;
    HIP_TRY(c, hipMemcpyAsync(m.X.as<float>() + m.rows * MLP_IN, p.st.features, 4 * (size_t)n * MLP_IN,
                              hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, launch_log1p_u32(p.st.exact, n, m.T.as<float>() + m.rows, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    m.rows += n;
    if (rows) *rows = n;
    return OCT_OK;
}

int oct_predictor_clear_data(oct_ctx* c) {
    if (!c) return OCT_ERR_ARGUMENT;
    c->mlp.rows = 0;
    return OCT_OK;
}

int oct_predictor_train(oct_ctx* c, int steps, int batch, float lr, uint64_t seed, int init, float* loss) {
    if (!c) return OCT_ERR_ARGUMENT;
    Predictor& m = c->mlp;
    if (m.rows <= 0) return fail(c, OCT_ERR_STATE, "no training rows; call oct_predictor_collect first");
    if (steps < 0 || batch < 1 || batch > (1 << 20) || !(lr > 0.f))
        return fail(c, OCT_ERR_ARGUMENT, "train: steps >= 0, batch in 1..2^20, lr > 0");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, m.params.ensure(4 * MLP_PARAMS));
    HIP_TRY(c, m.m.ensure(4 * MLP_TRAINED));
    HIP_TRY(c, m.v.ensure(4 * MLP_TRAINED));
    HIP_TRY(c, m.loss.ensure(16));
    HIP_TRY(c, m.scratch.ensure(4 * (size_t)mlp_scratch_floats(batch)));
    if (init || !m.loaded) {
        std::vector<float> P;
        he_init(P, seed);
        HIP_TRY(c, hipMemcpyAsync(m.params.p, P.data(), 4 * MLP_PARAMS, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, launch_feature_stats(m.X.as<float>(), m.rows, m.params.as<float>(), c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        m.adam_step = 0;
    }
    if (m.adam_step == 0) {
        HIP_TRY(c, launch_fill_u32(m.m.as<uint32_t>(), MLP_TRAINED, 0u, c->stream));
        HIP_TRY(c, launch_fill_u32(m.v.as<uint32_t>(), MLP_TRAINED, 0u, c->stream));
    }
    for (int s = 0; s < steps; ++s) {
        ++m.adam_step;
        const uint64_t sseed = (seed * 0x9E3779B97F4A7C15ull + (uint64_t)m.adam_step) | 1ull;
        HIP_TRY(c, launch_mlp_step(m.X.as<float>(), m.T.as<float>(), m.rows, batch, sseed, m.params.as<float>(),
                                   m.m.as<float>(), m.v.as<float>(), m.adam_step, lr, m.scratch.as<float>(), nullptr,
                                   s == steps - 1 ? m.loss.as<float>() : nullptr, c->stream));
    }
    float hl = 0.f;
    if (steps > 0) HIP_TRY(c, hipMemcpyAsync(&hl, m.loss.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = hl;
    m.loaded = true;
    return OCT_OK;
}

int oct_build_info_get(oct_ctx* c, int id, oct_build_info* out) {
    if (!c || !out) return OCT_ERR_ARGUMENT;
    if (!valid_set(c, id)) return fail(c, OCT_ERR_ARGUMENT, "unknown point set");
    SetState& s = c->sets[set_index(c, id)];
    if (s.refit_pending) HIP_TRY_RC(ensure_hp(c, s));  // the refit's outcome comes with the build parameters
    *out = s.info;
    return OCT_OK;
}

int oct_set_predictor_policy(oct_ctx* c, int always) {
    if (!c) return OCT_ERR_ARGUMENT;
    c->predict_always = always != 0;
    return OCT_OK;
}

int oct_edge_stats_get(oct_ctx* c, int q, int t, oct_edge_stats* out) {
    if (!c || !out) return OCT_ERR_ARGUMENT;
    PairState* p = find_pair(c, q, t);
    if (!p) return fail(c, OCT_ERR_STATE, "edge stats: pair not registered");
    if (p->stats_pending) {  // the call returned on the early total: its statistics are final at the kernels' end
        HIP_TRY(c, hipSetDevice(c->device));
        HIP_TRY(c, wait_stream(c->stream));
        uint32_t st2[2] = {0, 0};
        HIP_TRY(c, hipMemcpy(st2, p->flags.as<uint32_t>() + 9, 8, hipMemcpyDeviceToHost));
        p->st.recounted_leaves = st2[0];
        p->st.recorded_leaves = st2[1];
        recount_policy(*p, st2[0], p->pending_heavy, p->pending_leaves);
        p->stats_pending = false;
    }
    *out = p->st;
    return OCT_OK;
}

int oct_edge_stats_peek(oct_ctx* c, int q, int t, oct_edge_stats* out) {
    if (!c || !out) return OCT_ERR_ARGUMENT;
    PairState* p = find_pair(c, q, t);
    if (!p) return fail(c, OCT_ERR_STATE, "edge stats: pair not registered");
    *out = p->st;
    return OCT_OK;
}

int oct_mlp_forward(oct_ctx* c, const float* params, const float* X, int64_t n, float* y) {
    if (!c || !params || !X || !y || n < 0) return OCT_ERR_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, launch_mlp_infer(X, n, params, y, nullptr, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return OCT_OK;
}

// ------------------------------------------------------------ batches

namespace {

BatchState* batch_of(oct_ctx* c, int id) {
    const int k = id - c->index_base;
    if (k < 0 || k >= (int)c->batches.size() || !c->batches[k].live) return nullptr;
    return &c->batches[k];
}

// flags of a batch step (uint32 words): [0] a set needs more key bits than
// the sort covers, [1] largest overflowing candidate count, [2] a list walk
// ran out of stack, [4] single pass over capacity, [5] look-back gave up,
// [6..7] edge total, [8] ticket, [9] recounted leaves, [10] recorded leaves,
// [11] heavy leaves
constexpr int BF_WORDS = 16;

int batch_ensure(oct_ctx* c, BatchState& B) {
    const int64_t N = B.N;
    SetState& s = B.set;
    const int64_t M = max_nodes(N, B.leaf_size) + 22 * (int64_t)B.S;
    HIP_TRY(c, s.prm.ensure(sizeof(BuildParams)));
    HIP_TRY(c, s.partials.ensure(6 * sizeof(uint32_t) * (size_t)std::max(B.nblk_bounds, 1)));
    HIP_TRY(c, s.keys0.ensure(8 * (size_t)N));
    HIP_TRY(c, s.keys1.ensure(8 * (size_t)N));
    HIP_TRY(c, s.vals0.ensure(4 * (size_t)N));
    HIP_TRY(c, s.vals1.ensure(4 * (size_t)N));
    HIP_TRY(c, s.hist.ensure(4 * (size_t)hist_words(N)));
    HIP_TRY(c, s.scan_tmp.ensure(4 * (size_t)scan_tmp_words(hist_words(N))));
    HIP_TRY(c, s.total.ensure(16));
    HIP_TRY(c, s.topo.ensure(4 * (size_t)N));
    HIP_TRY(c, s.perm.ensure(4 * (size_t)N));
    HIP_TRY(c, s.scoord.ensure(16 * (size_t)N));
    HIP_TRY(c, s.nodes.ensure(sizeof(NodeRec) * (size_t)M));
    HIP_TRY(c, s.nfirst.ensure(4 * (size_t)M));
    HIP_TRY(c, s.nlast.ensure(4 * (size_t)M));
    HIP_TRY(c, s.nchild.ensure(32 * (size_t)M));
    HIP_TRY(c, s.nlevel.ensure((size_t)M));
    HIP_TRY(c, s.leaf_ids.ensure(4 * (size_t)N));
    HIP_TRY(c, s.leaves.ensure(sizeof(NodeRec) * (size_t)N));
    HIP_TRY(c, s.bmin.ensure(12 * (size_t)M));
    HIP_TRY(c, s.bmax.ensure(12 * (size_t)M));
    HIP_TRY(c, s.rslots.ensure(4 * (size_t)refit_slot_words(N)));
    HIP_TRY(c, s.ranc.ensure(4 * (size_t)OCT_MAX_LEVEL * (size_t)N));
    HIP_TRY(c, s.nfc.ensure(4 * ((size_t)M + 1)));
    HIP_TRY(c, s.noct.ensure((size_t)M));
    PairState& p = B.pair;
    HIP_TRY(c, p.flags.ensure(4 * BF_WORDS));
    HIP_TRY(c, p.lcount.ensure(4 * (size_t)N));
    HIP_TRY(c, p.lists.ensure(sizeof(int2) * (size_t)N * (size_t)B.cap));
    HIP_TRY(c, p.qcount.ensure(4 * (size_t)N));
    HIP_TRY(c, p.qoff.ensure(4 * (size_t)N));
    HIP_TRY(c, p.lhits.ensure(4 * (size_t)N));
    HIP_TRY(c, p.heavy.ensure(5 * (size_t)N));
    HIP_TRY(c, p.tiles.ensure(8 * (size_t)(single_tiles(N) + 1)));
    HIP_TRY(c, B.loff.ensure(4 * (size_t)N));
    HIP_TRY(c, B.bscan.ensure(4 * (size_t)scan_tmp_words(N)));
    HIP_TRY(c, B.eoff.ensure(8 * (size_t)(B.S + 1)));
    const int64_t ce = std::max<int64_t>(B.cap_edges, 1);
    HIP_TRY(c, p.senders.ensure(4 * (size_t)ce));
    HIP_TRY(c, p.receivers.ensure(4 * (size_t)ce));
    HIP_TRY(c, p.rdisp.ensure(12 * (size_t)ce));
    HIP_TRY(c, p.rdist.ensure(4 * (size_t)ce));
    return OCT_OK;
}

// Enqueues one whole batch step with the fixed host parameters (no host
// wait, no allocation: capturable).
int batch_enqueue(oct_ctx* c, BatchState& B) {
    const hipStream_t st = c->stream;
    const int64_t N = B.N;
    SetState& s = B.set;
    PairState& p = B.pair;
    const BuildBufs b = s.bufs();
    uint32_t* flags = p.flags.as<uint32_t>();
    BuildParams* sprm = B.sprm.as<BuildParams>();
    const int64_t* off = B.off.as<int64_t>();
    HIP_TRY(c, launch_fill_u32(flags, BF_WORDS, 0u, st));
    HIP_TRY(c, launch_bounds_batch(B.coords, off, B.bprefix.as<int32_t>(), B.S, B.nblk_bounds, B.radii.as<float>(),
                                   b.partials, sprm, st));
    HIP_TRY(c, launch_batch_check_bits(sprm, off, B.S, B.max_bits, flags, st));
    HIP_TRY(c, launch_keys_batch(B.coords, N, off, B.S, sprm, b.keys[0], b.vals[0], st));
    int cur = 0;
    HIP_TRY(c, launch_sort_passes(B.coords, N, B.shifts.data(), (int)B.shifts.size(), 0, b, &cur, st));
    const Forest f{1, off, sprm};
    HIP_TRY(c, launch_topology_stage(N, 0, B.leaf_size, 0, cur, b, f, st));
    // edges of every set against itself; grids are sized by N (an upper
    // bound of the leaves) and the kernels read the real counts on the device
    TreeView v;
    v.nodes = s.nodes.as<NodeRec>();
    v.leaf_ids = s.leaf_ids.as<int32_t>();
    v.leaves = s.leaves.as<NodeRec>();
    v.scoord = s.scoord.as<float4>();
    v.params = s.prm.as<BuildParams>();
    v.nlevel = s.nlevel.as<uint8_t>();
    v.n_leaves = (int32_t)N;
    // the fast walk packs node ids in 29 bits: a larger forest takes the safe walk
    v.n_nodes = (int32_t)std::min<int64_t>(max_nodes(N, B.leaf_size) + 22 * (int64_t)B.S, INT32_MAX);
    v.branch_root = 0;
    v.n = N;
    const SetLookup sets{1, b.keys[cur], off, sprm};
    HIP_TRY(c, launch_leaf_lists(v, v, 0.f, B.cap, 0, p.lists.as<int2>(), nullptr, p.lcount.as<int32_t>(), flags + 1,
                                 sets, p.heavy_flags(N), p.heavy_list(), flags + 11, st));
    HIP_TRY(c, launch_fill_u32(p.lhits.as<uint32_t>(), N, 0u, st));
    HIP_TRY(c, launch_fill_u32((uint32_t*)p.tiles.p, 2 * (single_tiles(N) + 1), 0u, st));
    const EdgeOut eo{p.senders.as<int32_t>(), p.receivers.as<int32_t>(), p.rdisp.as<float>(), p.rdist.as<float>()};
    HIP_TRY(c, launch_edges_heavy(v, v, p.lists.as<int2>(), p.lcount.as<int32_t>(), B.cap, 0.f, 0.f, 1,
                                  p.heavy_list(), flags + 11, p.qcount.as<uint32_t>(), p.lhits.as<uint32_t>(), sets,
                                  st));
    HIP_TRY(c, launch_edges_single(v, v, p.lists.as<int2>(), p.lcount.as<int32_t>(), B.cap, 0.f, 0.f, 1,
                                   p.qcount.as<uint32_t>(), p.qoff.as<uint32_t>(), nullptr, nullptr,
                                   p.lhits.as<uint32_t>(), p.tiles.as<uint64_t>(), flags + 8, flags + 4, flags + 9,
                                   (uint64_t)B.cap_edges, eo, c->index_base, B.run_leaves, B.run_chunk, sets,
                                   p.heavy_flags(N), nullptr, nullptr, 0, st));
    HIP_TRY(c, launch_edges_heavy_fill(v, v, p.lists.as<int2>(), p.lcount.as<int32_t>(), B.cap, 0.f, 0.f, 1,
                                       p.heavy_list(), flags + 11, p.qcount.as<uint32_t>(), p.qoff.as<uint32_t>(),
                                       flags + 4, eo, c->index_base, sets, st));
    HIP_TRY(c, exclusive_scan(p.lhits.as<uint32_t>(), B.loff.as<uint32_t>(), N, B.bscan.as<uint32_t>(),
                              (uint64_t*)(flags + 6), st));
    HIP_TRY(c, launch_batch_edge_offsets(B.loff.as<uint32_t>(), (const uint64_t*)(flags + 6), v.leaves, v.params,
                                         off, B.S, B.eoff.as<int64_t>(), st));
    return OCT_OK;
}

// The sort's digit shifts for the sets' Morton bits and the set id bits.
void batch_shifts(BatchState& B) {
    B.shifts.clear();
    const int mb = std::max(B.max_bits, 1);
    for (int sh = 0; sh < mb; sh += RADIX_BITS) B.shifts.push_back(sh);
    int set_bits = 1;
    while ((1 << set_bits) < B.S) ++set_bits;
    if (B.S > 1)
        for (int sh = BATCH_SHIFT; sh < BATCH_SHIFT + set_bits; sh += RADIX_BITS) B.shifts.push_back(sh);
}

int batch_read_flags(oct_ctx* c, BatchState& B, uint32_t* hf) {
    HIP_TRY(c, hipMemcpyAsync(c->pinned, B.pair.flags.p, 4 * BF_WORDS, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, wait_stream(c->stream));
    memcpy(hf, c->pinned, 4 * BF_WORDS);
    return OCT_OK;
}

// Eager step: learns the sort bits, grows what overflowed, reruns until the
// step completes; then captures the step for replays.
int batch_eager(oct_ctx* c, BatchState& B) {
    B.drop_graph();
    HIP_TRY(c, B.sprm.ensure(sizeof(BuildParams) * (size_t)B.S));
    HIP_TRY(c, B.set.partials.ensure(6 * sizeof(uint32_t) * (size_t)B.nblk_bounds));
    HIP_TRY(c, launch_bounds_batch(B.coords, B.off.as<int64_t>(), B.bprefix.as<int32_t>(), B.S, B.nblk_bounds,
                                   B.radii.as<float>(), B.set.partials.as<uint32_t>(), B.sprm.as<BuildParams>(),
                                   c->stream));
    std::vector<BuildParams> hp((size_t)B.S);
    HIP_TRY(c, hipMemcpyAsync(hp.data(), B.sprm.p, sizeof(BuildParams) * (size_t)B.S, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    int mb = 0;
    for (int s = 0; s < B.S; ++s)
        if (B.off_h[s + 1] > B.off_h[s]) mb = std::max(mb, hp[s].key_bits);
    if (mb > BATCH_SHIFT)
        return fail(c, OCT_ERR_CAPACITY, "batched sets are limited to 2^14 cells per axis (extent / radius)");
    B.max_bits = mb;
    batch_shifts(B);
    uint32_t hf[BF_WORDS] = {0};
    for (int attempt = 0; attempt < 8; ++attempt) {
        int rc = batch_ensure(c, B);
        if (rc != OCT_OK) return rc;
        rc = batch_enqueue(c, B);
        if (rc != OCT_OK) return rc;
        rc = batch_read_flags(c, B, hf);
        if (rc != OCT_OK) return rc;
        uint64_t total = 0;
        memcpy(&total, hf + 6, 8);
        if (hf[0] != 0) return fail(c, OCT_ERR_DEVICE, "batch: key bits changed between bounds and keys");
        if (hf[2] != 0) return fail(c, OCT_ERR_CAPACITY, "batch: a candidate walk ran out of stack");
        if (hf[1] != 0) {
            B.cap = (int)std::min<uint32_t>(hf[1] + 16u, 4096u);
            continue;
        }
        if (hf[5] != 0) return fail(c, OCT_ERR_DEVICE, "batch: single-pass look-back did not finish");
        if (hf[4] != 0) {
            if (total >= (uint64_t)INT32_MAX) return fail(c, OCT_ERR_CAPACITY, "edge count exceeds Int32");
            B.cap_edges = (int64_t)(total + total / 8 + 4096);
            continue;
        }
        B.n_edges = (int64_t)total;
        B.recounted = hf[9];
        B.recorded = hf[10];
        B.heavy = hf[11];
        B.replayed = 0;
        if (B.use_graph) {
            // capture the same step for the next calls (a graph replays kernels only)
            HIP_TRY_RC(capture_graph(c, [&]() { return batch_enqueue(c, B); }, &B.exec));
        }
        return OCT_OK;
    }
    return fail(c, OCT_ERR_DEVICE, "batch: buffers did not settle");
}

}  // namespace

int oct_batch_create(oct_ctx* c, int n_sets, const float* coords, const int64_t* offsets, const float* radii,
                     int* batch_id) {
    if (!c || !batch_id || !offsets || !radii) return OCT_ERR_ARGUMENT;
    if (c->ndims != 3 || c->f64) return fail(c, OCT_ERR_ARGUMENT, "batches hold 3D Float32 point sets");
    if (n_sets < 1 || n_sets > (1 << 20)) return fail(c, OCT_ERR_ARGUMENT, "batch: n_sets must be in 1..2^20");
    if (offsets[0] != 0) return fail(c, OCT_ERR_ARGUMENT, "batch: offsets[0] must be 0");
    for (int s = 0; s < n_sets; ++s) {
        if (offsets[s + 1] < offsets[s]) return fail(c, OCT_ERR_ARGUMENT, "batch: offsets must be nondecreasing");
        if (!(radii[s] > 0.f) || !std::isfinite(radii[s] * radii[s]))
            return fail(c, OCT_ERR_ARGUMENT, "batch: radii must be positive with a finite square");
    }
    const int64_t N = offsets[n_sets];
    if (N <= 0) return fail(c, OCT_ERR_ARGUMENT, "batch: no points");
    if (N >= (int64_t)INT32_MAX) return fail(c, OCT_ERR_CAPACITY, "batch: at most 2^31 - 1 points");
    if (!coords) return fail(c, OCT_ERR_ARGUMENT, "batch: coords is NULL");
    HIP_TRY(c, hipSetDevice(c->device));
    c->batches.emplace_back();
    BatchState& B = c->batches.back();
    B.pair.export_to(&c->retired);
    B.eoff.retire = &c->retired;
    B.live = true;
    B.S = n_sets;
    B.N = N;
    B.coords = coords;
    B.off_h.assign(offsets, offsets + n_sets + 1);
    B.radii_h.assign(radii, radii + n_sets);
    B.leaf_size = c->leaf_size;
    B.run_leaves = c->run_leaves;
    B.run_chunk = c->run_chunk;
    std::vector<int32_t> bp((size_t)n_sets + 1, 0);
    for (int s = 0; s < n_sets; ++s) bp[s + 1] = bp[s] + bounds_blocks(std::max<int64_t>(offsets[s + 1] - offsets[s], 1));
    B.nblk_bounds = bp[n_sets];
    HIP_TRY(c, B.off.ensure(8 * (size_t)(n_sets + 1)));
    HIP_TRY(c, B.radii.ensure(4 * (size_t)n_sets));
    HIP_TRY(c, B.bprefix.ensure(4 * (size_t)(n_sets + 1)));
    HIP_TRY(c, hipMemcpy(B.off.p, offsets, 8 * (size_t)(n_sets + 1), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(B.radii.p, radii, 4 * (size_t)n_sets, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(B.bprefix.p, bp.data(), 4 * (size_t)(n_sets + 1), hipMemcpyHostToDevice));
    *batch_id = (int)c->batches.size() - 1 + c->index_base;
    return OCT_OK;
}

int oct_batch_destroy(oct_ctx* c, int batch_id) {
    if (!c) return OCT_ERR_ARGUMENT;
    BatchState* B = batch_of(c, batch_id);
    if (!B) return fail(c, OCT_ERR_ARGUMENT, "unknown batch");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    B->release_all();
    return OCT_OK;
}

int oct_batch_config(oct_ctx* c, int batch_id, int leaf_size, int run_leaves, int run_chunk, int use_graph) {
    if (!c) return OCT_ERR_ARGUMENT;
    BatchState* B = batch_of(c, batch_id);
    if (!B) return fail(c, OCT_ERR_ARGUMENT, "unknown batch");
    if (leaf_size < 1 || leaf_size > 256) return fail(c, OCT_ERR_ARGUMENT, "leaf size must be in 1..256");
    if (run_leaves < 1 || run_leaves > SP_PLAN_MAX_LEAVES)
        return fail(c, OCT_ERR_ARGUMENT, "leaves per run must be in 1..64");
    if (run_chunk < 1 || run_chunk > SP_MAX_CHUNK) return fail(c, OCT_ERR_ARGUMENT, "runs per ticket must be in 1..4");
    if (leaf_size != B->leaf_size || run_leaves != B->run_leaves || run_chunk != B->run_chunk ||
        (use_graph != 0) != B->use_graph) {
        B->leaf_size = leaf_size;
        B->run_leaves = run_leaves;
        B->run_chunk = run_chunk;
        B->use_graph = use_graph != 0;
        B->drop_graph();
        B->max_bits = -1;
    }
    return OCT_OK;
}

int oct_batch_run(oct_ctx* c, int batch_id, oct_batch_edges* out) {
    if (!c || !out) return OCT_ERR_ARGUMENT;
    BatchState* Bp = batch_of(c, batch_id);
    if (!Bp) return fail(c, OCT_ERR_ARGUMENT, "unknown batch");
    BatchState& B = *Bp;
    HIP_TRY(c, hipSetDevice(c->device));
    memset(out, 0, sizeof(*out));
    bool done = false;
    if (B.exec && B.max_bits >= 0) {
        HIP_TRY(c, hipGraphLaunch(B.exec, c->stream));
        uint32_t hf[BF_WORDS] = {0};
        int rc = batch_read_flags(c, B, hf);
        if (rc != OCT_OK) return rc;
        uint64_t total = 0;
        memcpy(&total, hf + 6, 8);
        if (hf[0] == 0 && hf[1] == 0 && hf[2] == 0 && hf[4] == 0 && hf[5] == 0) {
            B.n_edges = (int64_t)total;
            B.recounted = hf[9];
            B.recorded = hf[10];
            B.heavy = hf[11];
            B.replayed = 1;
            done = true;
        }
    }
    if (!done) {
        const int rc = batch_eager(c, B);
        if (rc != OCT_OK) return rc;
    }
    ++B.steps;
    PairState& p = B.pair;
    out->senders = p.senders.as<int32_t>();
    out->receivers = p.receivers.as<int32_t>();
    out->rel_displacement = p.rdisp.as<float>();
    out->rel_dist_norm = p.rdist.as<float>();
    out->n_edges = B.n_edges;
    out->set_offsets = B.eoff.as<int64_t>();
    out->n_sets = B.S;
    out->replayed = B.replayed;
    out->recounted_leaves = B.recounted;
    out->recorded_leaves = B.recorded;
    out->heavy_leaves = B.heavy;
    return OCT_OK;
}

// ------------------------------------------------------------ autotuner

extern "C++" {
namespace {

// Milliseconds of device time of `reps` calls of fn, measured with events on
// the context stream after one untimed call.
template <typename F>
int time_device(oct_ctx* c, int reps, F fn, double* ms) {
    int rc = fn();
    if (rc != OCT_OK) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(c, hipEventCreate(&a));
    HIP_TRY(c, hipEventCreate(&b));
    HIP_TRY(c, hipEventRecord(a, c->stream));
    for (int k = 0; k < reps && rc == OCT_OK; ++k) rc = fn();
    HIP_TRY(c, hipEventRecord(b, c->stream));
    HIP_TRY(c, hipEventSynchronize(b));
    float t = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = (double)t / std::max(reps, 1);
    return rc;
}

}  // namespace
}  // extern "C++"

int oct_batch_autotune(oct_ctx* c, int batch_id, const int* leaf_sizes, int n_leaf_sizes, const int* runs,
                       int n_runs, const int* chunks, int n_chunks, int reps, double* ms_out, int* best) {
    if (!c || !leaf_sizes || !runs || !chunks || n_leaf_sizes < 1 || n_runs < 1 || n_chunks < 1 || reps < 1)
        return OCT_ERR_ARGUMENT;
    BatchState* B = batch_of(c, batch_id);
    if (!B) return fail(c, OCT_ERR_ARGUMENT, "unknown batch");
    const int graph = B->use_graph ? 1 : 0;
    double best_ms = 1e300;
    int best_k = -1;
    for (int i = 0; i < n_leaf_sizes; ++i)
        for (int j = 0; j < n_runs; ++j)
            for (int m = 0; m < n_chunks; ++m) {
                const int k = (i * n_runs + j) * n_chunks + m;
                int rc = oct_batch_config(c, batch_id, leaf_sizes[i], runs[j], chunks[m], graph);
                if (rc != OCT_OK) return rc;
                oct_batch_edges e;
                double ms = 0.0;
                rc = time_device(c, reps, [&]() { return oct_batch_run(c, batch_id, &e); }, &ms);
                if (rc != OCT_OK) return rc;
                if (ms_out) ms_out[k] = ms;
                if (ms < best_ms) { best_ms = ms; best_k = k; }
            }
    const int bi = best_k / (n_runs * n_chunks), bj = (best_k / n_chunks) % n_runs, bm = best_k % n_chunks;
    int rc = oct_batch_config(c, batch_id, leaf_sizes[bi], runs[bj], chunks[bm], graph);
    if (rc != OCT_OK) return rc;
    if (best) *best = best_k;
    return OCT_OK;
}

int oct_set_run_leaves(oct_ctx* c, int run_leaves) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (run_leaves < 1 || run_leaves > SP_PLAN_MAX_LEAVES)
        return fail(c, OCT_ERR_ARGUMENT, "leaves per run must be in 1..64");
    c->run_leaves = run_leaves;
    return OCT_OK;
}

int oct_set_run_chunk(oct_ctx* c, int run_chunk) {
    if (!c) return OCT_ERR_ARGUMENT;
    if (run_chunk < 1 || run_chunk > SP_MAX_CHUNK) return fail(c, OCT_ERR_ARGUMENT, "runs per ticket must be in 1..4");
    c->run_chunk = run_chunk;
    return OCT_OK;
}

int oct_autotune(oct_ctx* c, int q, int t, int mode, const int* leaf_sizes, int n_leaf_sizes, const int* runs,
                 int n_runs, const int* chunks, int n_chunks, int reps, double* ms_out, int* best) {
    if (!c || !leaf_sizes || !runs || !chunks || n_leaf_sizes < 1 || n_runs < 1 || n_chunks < 1 || reps < 1)
        return OCT_ERR_ARGUMENT;
    if (!find_pair(c, q, t)) return fail(c, OCT_ERR_STATE, "autotune: pair not registered");
    double best_ms = 1e300;
    int best_k = -1;
    for (int i = 0; i < n_leaf_sizes; ++i)
        for (int j = 0; j < n_runs; ++j)
            for (int m = 0; m < n_chunks; ++m) {
                const int k = (i * n_runs + j) * n_chunks + m;
                int rc = oct_set_leaf_size(c, leaf_sizes[i]);
                if (rc == OCT_OK) rc = oct_set_run_leaves(c, runs[j]);
                if (rc == OCT_OK) rc = oct_set_run_chunk(c, chunks[m]);
                if (rc != OCT_OK) return rc;
                oct_edges e;
                double ms = 0.0;
                // one step: rebuild the dirty sets and the edges of the pair
                rc = time_device(c, reps, [&]() {
                    for (auto& s : c->sets) s.dirty = true;
                    int r = oct_run(c);
                    return r == OCT_OK ? oct_build_edges(c, q, t, mode, OCT_ORDER_LEAF, &e) : r;
                }, &ms);
                if (rc != OCT_OK) return rc;
                if (ms_out) ms_out[k] = ms;
                if (ms < best_ms) { best_ms = ms; best_k = k; }
            }
    const int bi = best_k / (n_runs * n_chunks), bj = (best_k / n_chunks) % n_runs, bm = best_k % n_chunks;
    int rc = oct_set_leaf_size(c, leaf_sizes[bi]);
    if (rc == OCT_OK) rc = oct_set_run_leaves(c, runs[bj]);
    if (rc == OCT_OK) rc = oct_set_run_chunk(c, chunks[bm]);
    if (rc != OCT_OK) return rc;
    if (best) *best = best_k;
    return OCT_OK;
}

int oct_mlp_gradient_noalloc(oct_ctx* c, const float* params, const float* X, const float* T, int n, float* grad,
                             float* loss) {
    if (!c || !params || !X || !T || !grad || n < 1) return OCT_ERR_ARGUMENT;
    if (n > OCT_MLP_STATIC_ROWS)
        return fail(c, OCT_ERR_CAPACITY, "oct_mlp_gradient_noalloc: at most OCT_MLP_STATIC_ROWS (8192) rows");
    HIP_TRY(c, hipSetDevice(c->device));
    float* scratch = noalloc_scratch_ptr[c->device];  // set by oct_create
    if (!scratch) return fail(c, OCT_ERR_STATE, "fixed scratch not set up for this device");
    // every word the kernels read is written first in the same call (row
    // records for rows < n, partial sums for chunks < ceil(n / 256)), so
    // nothing needs clearing; stream-ordered, no host wait
    HIP_TRY(c, launch_mlp_step(X, T, n, n, 0, const_cast<float*>(params), nullptr, nullptr, 1, 0.f, scratch, grad,
                               loss, c->stream));
    return OCT_OK;
}

int oct_mlp_gradient(oct_ctx* c, const float* params, const float* X, const float* T, int n, float* grad,
                     float* loss) {
    if (!c || !params || !X || !T || !grad || n < 1 || n > (1 << 20)) return OCT_ERR_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    Predictor& m = c->mlp;
    HIP_TRY(c, m.scratch.ensure(4 * (size_t)mlp_scratch_floats(n)));
    HIP_TRY(c, launch_mlp_step(X, T, n, n, 0, const_cast<float*>(params), nullptr, nullptr, 1, 0.f,
                               m.scratch.as<float>(), grad, loss, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return OCT_OK;
}

}  // extern "C"
