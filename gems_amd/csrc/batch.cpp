// Batched unlinked partitions: many contexts (one per partition, like the per-partition
// PhyloTrees of PhyloSuperTree, tree/phylosupertree.cpp:662-734) evaluated with one P(t)
// launch and one persistent traversal launch, replayed as a HIP graph while the launch
// shape is unchanged, with the persistent grid chosen by a measuring autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "engine_common.h"
#include "plk_batch.h"

namespace plk {

#define BHIP_TRY(expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(e_ == hipErrorOutOfMemory ? PLK_ERR_NOMEM : PLK_ERR_HIP,                \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                     \
    } while (0)

struct Batch {
    std::vector<plk_ctx *> parts;
    int parts_cap = 0, items_cap = 0;  // allocated d_params/d_results/h_results and d_items
    int prec = PLK_FP64;
    hipStream_t stream = nullptr;
    KParams *d_params = nullptr;
    BatchItem *d_items = nullptr;
    DevOp *d_ops = nullptr;
    PJob *d_jobs = nullptr;
    double *d_results = nullptr;
    int nitems = 0, ops_cap = 0, jobs_cap = 0;
    void *h_stage = nullptr;  // params | ops | jobs  (pinned)
    size_t stage_cap = 0;
    double *h_results = nullptr;
    int grid = 0;             // 0 = one workgroup per item
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int g_ops = -1, g_jobs = -1, g_grid = -1;  // launch shape captured in the graph
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double t_total = 0.0;
    int64_t t_count = 0, updates = 0;
};

static void drop_graph(Batch &b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    b.g_ops = b.g_jobs = b.g_grid = -1;
}

// Enqueue the whole batch evaluation on b.stream (captured into the graph when enabled).
static int enqueue(Batch &b, int nparts, int tot_ops, int njobs) {
    const size_t pb = sizeof(KParams) * nparts, ob = sizeof(DevOp) * tot_ops, jb = sizeof(PJob) * njobs;
    BHIP_TRY(hipMemcpyAsync(b.d_params, b.h_stage, pb, hipMemcpyHostToDevice, b.stream));
    if (ob) BHIP_TRY(hipMemcpyAsync(b.d_ops, (char *)b.h_stage + pb, ob, hipMemcpyHostToDevice, b.stream));
    if (jb) BHIP_TRY(hipMemcpyAsync(b.d_jobs, (char *)b.h_stage + pb + ob, jb, hipMemcpyHostToDevice, b.stream));
    BHIP_TRY(batch_pmatrix(b.prec, b.d_params, b.d_jobs, njobs, b.stream));
    BHIP_TRY(batch_traverse(b.prec, b.d_params, b.d_items, b.nitems, b.grid, b.stream));
    BHIP_TRY(hipMemcpyAsync(b.h_results, b.d_results, sizeof(double) * 2 * nparts, hipMemcpyDeviceToHost, b.stream));
    return PLK_OK;
}

}  // namespace plk

struct plk_batch : public plk::Batch {};

using namespace plk;

static int check_parts(plk_ctx *const *parts, int nparts, const char *who) {
    const int prec = parts[0] ? parts[0]->prec : 0;
    for (int i = 0; i < nparts; i++) {
        plk_ctx *c = parts[i];
        if (!c) return fail(PLK_ERR_ARG, std::string(who) + ": null context");
        if (c->prec != prec) return fail(PLK_ERR_ARG, std::string(who) + ": partitions must share one precision form");
        if (c->n != 4 && c->n != 20) return fail(PLK_ERR_ARG, std::string(who) + ": batched partitions are DNA or 20-state");
        if (c->ncat > 4) return fail(PLK_ERR_ARG, std::string(who) + ": batched partitions need ncat <= 4");
        for (int j = 0; j < i; j++)
            if (parts[j] == c) return fail(PLK_ERR_ARG, std::string(who) + ": a context appears twice");
    }
    return PLK_OK;
}

// all tiles of all partitions, most expensive partitions first
static std::vector<BatchItem> make_items(plk_ctx *const *parts, int nparts) {
    std::vector<int> order(nparts);
    std::iota(order.begin(), order.end(), 0);
    auto cost = [&](int i) { return (double)parts[i]->n * parts[i]->n * parts[i]->ntiles; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
    std::vector<BatchItem> items;
    for (int i : order)
        for (int t = 0; t < parts[i]->ntiles; t++) items.push_back(BatchItem{i, t});
    return items;
}

extern "C" int plk_batch_rebind(plk_batch *b, plk_ctx *const *parts, int nparts) {
    clear_error();
    if (!b || !parts || nparts < 1) return fail(PLK_ERR_ARG, "plk_batch_rebind: bad arguments");
    int rc = check_parts(parts, nparts, "plk_batch_rebind");
    if (rc) return rc;
    if (parts[0]->prec != b->prec) return fail(PLK_ERR_ARG, "plk_batch_rebind: the batch's precision form differs");
    BHIP_TRY(hipStreamSynchronize(b->stream));
    drop_graph(*b);
    const std::vector<BatchItem> items = make_items(parts, nparts);
    if (nparts > b->parts_cap) {
        const int cap = std::max(nparts, 2 * b->parts_cap);
        // (plain hipFree/hipMalloc like plk_batch_create; growth doubles, so this is rare)
        if (b->d_params) (void)hipFree(b->d_params);
        if (b->d_results) (void)hipFree(b->d_results);
        if (b->h_results) (void)hipHostFree(b->h_results);
        b->d_params = nullptr;
        b->d_results = b->h_results = nullptr;
        b->parts_cap = 0;
        BHIP_TRY(hipMalloc((void **)&b->d_params, sizeof(KParams) * cap));
        BHIP_TRY(hipMalloc((void **)&b->d_results, sizeof(double) * 2 * cap));
        BHIP_TRY(hipHostMalloc((void **)&b->h_results, sizeof(double) * 2 * cap, hipHostMallocDefault));
        b->parts_cap = cap;
    }
    if ((int)items.size() > b->items_cap) {
        const int cap = std::max((int)items.size(), 2 * b->items_cap);
        if (b->d_items) (void)hipFree(b->d_items);
        b->d_items = nullptr;
        b->items_cap = 0;
        BHIP_TRY(hipMalloc((void **)&b->d_items, sizeof(BatchItem) * cap));
        b->items_cap = cap;
    }
    BHIP_TRY(hipMemcpyAsync(b->d_items, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    b->parts.assign(parts, parts + nparts);
    b->nitems = (int)items.size();
    return PLK_OK;
}

extern "C" int plk_batch_create(plk_batch **out, plk_ctx *const *parts, int nparts) {
    clear_error();
    if (!out || !parts || nparts < 1) return fail(PLK_ERR_ARG, "plk_batch_create: bad arguments");
    *out = nullptr;
    const int prec = parts[0] ? parts[0]->prec : 0;
    {
        int rc = check_parts(parts, nparts, "plk_batch_create");
        if (rc) return rc;
    }
    auto *b = new plk_batch();
    b->parts.assign(parts, parts + nparts);
    b->prec = prec;
    auto cleanup = [&](int rc) {
        plk_batch_destroy(b);
        return rc;
    };
    const std::vector<BatchItem> items = make_items(parts, nparts);
    b->nitems = (int)items.size();
    b->parts_cap = nparts;
    b->items_cap = b->nitems;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&b->d_params, sizeof(KParams) * nparts)) != hipSuccess ||
        (e = hipMalloc(&b->d_items, sizeof(BatchItem) * items.size())) != hipSuccess ||
        (e = hipMalloc(&b->d_results, sizeof(double) * 2 * nparts)) != hipSuccess ||
        (e = hipHostMalloc((void **)&b->h_results, sizeof(double) * 2 * nparts, hipHostMallocDefault)) != hipSuccess ||
        (e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess ||
        (e = hipMemcpy(b->d_items, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(fail(PLK_ERR_HIP, std::string("plk_batch_create: ") + hipGetErrorString(e)));
    *out = b;
    return PLK_OK;
}

extern "C" int plk_batch_destroy(plk_batch *b) {
    if (!b) return PLK_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(*b);
    void *bufs[] = {b->d_params, b->d_items, b->d_ops, b->d_jobs, b->d_results};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->h_results) (void)hipHostFree(b->h_results);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PLK_OK;
}

extern "C" int plk_batch_set_launch(plk_batch *b, int grid, int use_graph) {
    if (!b || grid < 0) return fail(PLK_ERR_ARG, "plk_batch_set_launch: bad arguments");
    b->grid = grid;
    b->use_graph = use_graph != 0;
    drop_graph(*b);
    return PLK_OK;
}

extern "C" int plk_batch_loglik(plk_batch *b, const int *a, const int *bref, const double *t, double *logl,
                                double *total) {
    clear_error();
    if (!b || !a || !bref || !t) return fail(PLK_ERR_ARG, "plk_batch_loglik: null argument");
    const int nparts = (int)b->parts.size();
    // --- per-partition validation and op/job gathering (the queued ops of each context)
    int tot_ops = 0, njobs = 0;
    for (int i = 0; i < nparts; i++) {
        plk_ctx *c = b->parts[i];
        if (!c->model_set || !c->patterns_set) return fail(PLK_ERR_STATE, "plk_batch_loglik: partition not initialised");
        int ra = a[i], rb = bref[i];
        if (!valid_ref(*c, ra) || !valid_ref(*c, rb) || (ra < 0 && rb < 0) || ra == rb || !std::isfinite(t[i]) || t[i] < 0)
            return fail(PLK_ERR_ARG, "plk_batch_loglik: bad branch for partition " + std::to_string(i));
        tot_ops += (int)c->pending.size();
        njobs += (int)c->jobs.size() + (b->prec == PLK_FP32 ? 1 : 0);
    }
    const size_t pb = sizeof(KParams) * nparts, ob = sizeof(DevOp) * tot_ops, jb = sizeof(PJob) * njobs;
    if (pb + ob + jb > b->stage_cap) {
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        if (b->h_stage) (void)hipHostFree(b->h_stage);
        b->h_stage = nullptr;
        b->stage_cap = 0;
        BHIP_TRY(hipHostMalloc(&b->h_stage, (pb + ob + jb) * 2, hipHostMallocDefault));
        b->stage_cap = (pb + ob + jb) * 2;
        drop_graph(*b);
    }
    if (tot_ops > b->ops_cap) {
        if (b->d_ops) (void)hipFree(b->d_ops);
        b->d_ops = nullptr;
        BHIP_TRY(hipMalloc(&b->d_ops, sizeof(DevOp) * std::max(tot_ops * 2, 64)));
        b->ops_cap = std::max(tot_ops * 2, 64);
        drop_graph(*b);
    }
    if (njobs > b->jobs_cap) {
        if (b->d_jobs) (void)hipFree(b->d_jobs);
        b->d_jobs = nullptr;
        BHIP_TRY(hipMalloc(&b->d_jobs, sizeof(PJob) * std::max(njobs * 2, 64)));
        b->jobs_cap = std::max(njobs * 2, 64);
        drop_graph(*b);
    }
    KParams *hp = (KParams *)b->h_stage;
    DevOp *hop = (DevOp *)((char *)b->h_stage + pb);
    PJob *hj = (PJob *)((char *)b->h_stage + pb + ob);
    int op_off = 0, job_off = 0;
    for (int i = 0; i < nparts; i++) {
        plk_ctx *c = b->parts[i];
        const int nops = (int)c->pending.size();
        int rc = ensure_op_capacity(*c, nops + 1);
        if (rc) return rc;
        int ra = a[i], rb = bref[i];
        if (rb < 0) std::swap(ra, rb);
        KParams p = make_params(*c);
        p.ops = b->d_ops + op_off;
        p.nops = nops;
        p.mats = c->d_mats;
        p.br_a = ra;
        p.br_b = rb;
        p.br_len = t[i];
        p.result = b->d_results + 2 * i;
        hp[i] = p;
        if (nops) std::memcpy(hop + op_off, c->pending.data(), sizeof(DevOp) * nops);
        for (const PJob &j : c->jobs) {
            PJob q = j;
            q.pad = i;
            hj[job_off++] = q;
        }
        if (b->prec == PLK_FP32) hj[job_off++] = PJob{nops, 0, 0, i, t[i]};
        op_off += nops;
        c->theta_valid = false;
    }
    // --- launch: graph replay when the captured launch shape still applies
    const bool replay = b->use_graph && b->exec && b->g_ops == tot_ops && b->g_jobs == njobs && b->g_grid == b->grid;
    BHIP_TRY(hipEventRecord(b->ev0, b->stream));
    if (replay) {
        BHIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else if (b->use_graph) {
        drop_graph(*b);
        BHIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(*b, nparts, tot_ops, njobs);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess) return fail(PLK_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
        b->graph = g;
        BHIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->g_ops = tot_ops;
        b->g_jobs = njobs;
        b->g_grid = b->grid;
        BHIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
        int rc = enqueue(*b, nparts, tot_ops, njobs);
        if (rc) return rc;
    }
    BHIP_TRY(hipEventRecord(b->ev1, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) {
        b->t_total += ms;
        b->t_count += 1;
    }
    double sum = 0.0;
    for (int i = 0; i < nparts; i++) {
        plk_ctx *c = b->parts[i];
        double v = b->h_results[2 * i];
        if (!std::isfinite(v)) {
            // per-pattern fallback of tree/phylokernelnew.h:3167-3184 for this partition
            std::vector<double> pl(c->nptn);
            BHIP_TRY(hipMemcpy(pl.data(), c->d_ptn_lh, sizeof(double) * c->nptn, hipMemcpyDeviceToHost));
            v = 0.0;
            for (int q = 0; q < c->nptn; q++) {
                if (!std::isfinite(pl[q])) pl[q] = kLogScalingThreshold * 4;
                v += pl[q] * c->host_freq[q];
            }
        }
        if (logl) logl[i] = v;
        c->ptn_lh_valid = true;
        sum += v;
        c->updates += (int64_t)c->pending.size() * c->nptn;
        b->updates += (int64_t)c->pending.size() * c->nptn;
        c->pending.clear();
        c->jobs.clear();
    }
    if (total) *total = sum;
    return PLK_OK;
}

extern "C" int plk_batch_autotune(plk_batch *b, const int *a, const int *bref, const double *t, int reps,
                                  int *best_grid, double *best_ms) {
    clear_error();
    if (!b || reps < 1) return fail(PLK_ERR_ARG, "plk_batch_autotune: bad arguments");
    const int nparts = (int)b->parts.size();
    // keep the queued traversal so every candidate runs the same work
    std::vector<std::vector<DevOp>> ops(nparts);
    std::vector<std::vector<PJob>> jobs(nparts);
    for (int i = 0; i < nparts; i++) {
        ops[i] = b->parts[i]->pending;
        jobs[i] = b->parts[i]->jobs;
    }
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    std::vector<int> cands = {0};
    for (int k : {1, 2, 4, 8, 16})
        if (ncu * k < b->nitems) cands.push_back(ncu * k);
    const int saved_grid = b->grid;
    const bool saved_graph = b->use_graph;
    int best = 0;
    double best_t = 1e300;
    for (int g : cands) {
        b->grid = g;
        b->use_graph = saved_graph;
        drop_graph(*b);
        double tmin = 1e300;
        for (int r = 0; r < reps + 1; r++) {
            for (int i = 0; i < nparts; i++) {
                b->parts[i]->pending = ops[i];
                b->parts[i]->jobs = jobs[i];
            }
            const double before = b->t_total;
            int rc = plk_batch_loglik(b, a, bref, t, nullptr, nullptr);
            if (rc) {
                b->grid = saved_grid;
                return rc;
            }
            if (r > 0) tmin = std::min(tmin, b->t_total - before);  // first run captures the graph
        }
        if (tmin < best_t) {
            best_t = tmin;
            best = g;
        }
    }
    for (int i = 0; i < nparts; i++) {
        b->parts[i]->pending = ops[i];
        b->parts[i]->jobs = jobs[i];
    }
    b->grid = best;
    drop_graph(*b);
    if (best_grid) *best_grid = best;
    if (best_ms) *best_ms = best_t;
    return PLK_OK;
}

extern "C" int plk_batch_timing(plk_batch *b, double *total_ms, int64_t *launches, int64_t *updates) {
    if (!b) return fail(PLK_ERR_ARG, "null batch");
    if (total_ms) *total_ms = b->t_total;
    if (launches) *launches = b->t_count;
    if (updates) *updates = b->updates;
    return PLK_OK;
}

extern "C" int plk_batch_items(plk_batch *b) { return b ? b->nitems : 0; }
