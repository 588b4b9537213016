// Serving loop with continuous batching.
//
// Admission: FIFO in arrival order, whenever a request slot is free and the request's HBM
// footprint (plk_memory_required, after PhyloTree::getMemoryRequired,
// tree/phylotree.cpp:992-1052) fits in the remaining budget; a request that can never
// fit fails at once instead of blocking the queue.
//
// Execution: every admitted request is a small state machine over PhyloTree jobs --
// a traversal + logL (computeLikelihood, tree/phylotree.cpp:1258-1299), a Newton solve
// (optimizeOneBranch :2620-2670 or optimizeAllBranches :2724-2801), then the final logL.
// One scheduler thread gathers, at every step, the traversals of ALL requests in flight
// (new admissions and those whose Newton phase just ended) and runs them as one persistent
// batched launch over their (request, tile) work items (plk_batch, re-formed per step).
// Newton phases, which are short chains of launches on the request's own stream, run on
// three worker threads, so at most four streams (the batch's and three Newton streams)
// are busy at once.  Requests that arrive while others are on the GPU join the next
// step.  Trees and their device buffers come from a pool keyed by (states, categories,
// precision, scaling) and are reshaped for each request instead of being allocated anew.
//
// Results do not depend on what else is in flight: a batched launch gives each request
// exactly the value of evaluating it alone (fixed-order reductions per request), and the
// Newton phases run the same kernels on the request's own context.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "plk.h"
#include "plk_batch.h"
#include "plk_internal.h"
#include "plk_serve.h"
#include "plk_tree.h"

namespace plk {

enum : int { ST_FIRST = 0, ST_NEWTON = 1, ST_FINAL = 2, ST_DONE = 3 };
constexpr int kNewtonWorkers = 3;

struct PoolTree {
    plk_tree *t = nullptr;
    int nstates = 0, ncat = 0, prec = 0, scaling = 0;
    bool busy = false;
};

struct Job {
    plk_request *r = nullptr;
    int64_t fp = 0;
    int stage = ST_FIRST;
    bool busy = false;   // in a step or a Newton phase
    int pool = -1;       // index into the tree pool
    bool setup = false;  // model, patterns and topology loaded
    int rc = PLK_OK;
};

struct Server {
    int max_active = 1;
    int64_t budget = 0, in_use = 0, peak_use = 0;
    int peak_active = 0;
    std::deque<plk_request *> queue;
    std::vector<Job *> jobs;
    std::deque<Job *> newton_q;
    std::vector<PoolTree> pool;
    std::mutex mu;
    std::condition_variable cv_sched, cv_newton, cv_done;
    std::thread sched;
    std::vector<std::thread> newton;
    bool stop = false;
    int64_t submitted = 0, completed = 0, failed = 0;
    int64_t launches = 0, batched = 0;  // shared traversal launches, request traversals in them
    std::vector<double> latency;
    double first_submit = -1.0, last_done = 0.0;
    plk_batch *batch[2] = {nullptr, nullptr};  // per precision form
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static int64_t footprint(const plk_request *r) {
    return plk_memory_required(r->nstates, r->ncat, r->ntaxa, r->nptn, r->ntaxa - 2, r->precision, r->scaling);
}

static bool batchable(const plk_request *r) { return (r->nstates == 4 || r->nstates == 20) && r->ncat <= 4; }

// (under the lock) request done: release its tree and budget, record the latency
static void finish(Server &s, Job *j) {
    plk_request *r = j->r;
    r->status = j->rc;
    r->t_done = s.now();
    if (j->pool >= 0) s.pool[j->pool].busy = false;
    s.in_use -= j->fp;
    s.completed++;
    if (j->rc) s.failed++;
    else s.latency.push_back(r->t_done - r->t_submit);
    s.last_done = std::max(s.last_done, r->t_done);
    s.jobs.erase(std::find(s.jobs.begin(), s.jobs.end(), j));
    delete j;
    s.cv_done.notify_all();
}

// (under the lock) an idle pooled tree of the request's class, or -1
static int pool_take(Server &s, const plk_request *r) {
    for (size_t i = 0; i < s.pool.size(); i++) {
        PoolTree &p = s.pool[i];
        if (!p.busy && p.nstates == r->nstates && p.ncat == r->ncat && p.prec == r->precision &&
            p.scaling == r->scaling) {
            p.busy = true;
            return (int)i;
        }
    }
    return -1;
}

// Load a request into its tree (reshaped pooled tree or a fresh one).
static int setup_job(Server &s, Job *j, plk_tree **fresh) {
    plk_request *r = j->r;
    plk_tree *t = nullptr;
    int rc;
    if (j->pool >= 0) {
        t = s.pool[j->pool].t;  // the pool vector only grows under the lock; entries stay
        rc = plk_tree_reshape(t, r->ntaxa, r->nptn);
    } else {
        rc = plk_tree_create(&t, r->nstates, r->ncat, r->ntaxa, r->nptn, r->precision, r->scaling);
        if (!rc) *fresh = t;
    }
    if (!rc) rc = plk_tree_set_model(t, r->eval, r->evec, r->inv_evec, r->rates, r->props, r->state_unknown, r->tip_lh);
    if (!rc) rc = plk_tree_set_patterns(t, r->states, r->freq, r->invar);
    if (!rc) rc = plk_tree_set_topology(t, r->nei, r->len);
    return rc;
}

// One step: the traversal + logL of every job in `ready` (FIRST or FINAL stage); batchable
// requests of one precision form share one launch.
static void run_step(Server &s, std::vector<Job *> &ready, std::vector<plk_tree *> &trees) {
    std::vector<Job *> group[2];
    std::vector<plk_tree *> gtree[2];
    for (size_t k = 0; k < ready.size(); k++) {
        Job *j = ready[k];
        if (j->rc) continue;
        plk_request *r = j->r;
        if (batchable(r)) {
            group[r->precision == PLK_FP32 ? 1 : 0].push_back(j);
            gtree[r->precision == PLK_FP32 ? 1 : 0].push_back(trees[k]);
            continue;
        }
        // 61 states or more than four categories: the single-context kernels
        plk_tree *t = trees[k];
        double v = 0.0;
        if (j->stage == ST_FIRST) j->rc = plk_tree_compute_likelihood(t, &v);
        else j->rc = plk_tree_loglik(t, r->opt_a, r->opt_b, &v, nullptr);
        if (j->stage == ST_FIRST) r->logl_before = v;
        r->logl_after = v;
    }
    for (int g = 0; g < 2; g++) {
        const size_t n = group[g].size();
        if (!n) continue;
        std::vector<plk_ctx *> ctx(n);
        std::vector<int> a(n), b(n);
        std::vector<double> len(n), out(n, 0.0);
        int rc = PLK_OK;
        for (size_t i = 0; i < n && !rc; i++) {
            Job *j = group[g][i];
            plk_tree *t = gtree[g][i];
            ctx[i] = plk_tree_context(t);
            plk_request *r = j->r;
            if (j->stage == ST_FIRST) {
                // computeLikelihood: the branch of the first leaf, every partial recomputed
                rc = plk_tree_queue_branch(t, 0, r->nei[0], 1, &a[i], &b[i], &len[i]);
            } else {
                rc = plk_tree_queue_branch(t, r->opt_a, r->opt_b, 0, &a[i], &b[i], &len[i]);
            }
        }
        if (!rc) {
            if (!s.batch[g]) {
                rc = plk_batch_create(&s.batch[g], ctx.data(), (int)n);
                if (!rc) rc = plk_batch_set_launch(s.batch[g], 0, 0);
            } else {
                rc = plk_batch_rebind(s.batch[g], ctx.data(), (int)n);
            }
        }
        if (!rc) rc = plk_batch_loglik(s.batch[g], a.data(), b.data(), len.data(), out.data(), nullptr);
        if (!rc) {
            std::lock_guard<std::mutex> lk(s.mu);
            s.launches++;
            s.batched += (int64_t)n;
        }
        for (size_t i = 0; i < n; i++) {
            Job *j = group[g][i];
            if (rc) {
                j->rc = rc;
                continue;
            }
            if (j->stage == ST_FIRST) j->r->logl_before = out[i];
            j->r->logl_after = out[i];
        }
    }
}

static void scheduler(Server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // admission, in arrival order (no overtaking, so large requests are not starved)
        while (!s->queue.empty()) {
            plk_request *r = s->queue.front();
            const int64_t fp = r->footprint;
            if (fp > s->budget) {
                s->queue.pop_front();
                auto *j = new Job();
                j->r = r;
                j->rc = PLK_ERR_NOMEM;
                r->t_admit = s->now();
                s->jobs.push_back(j);
                finish(*s, j);
                continue;
            }
            if ((int)s->jobs.size() >= s->max_active || s->in_use + fp > s->budget) break;
            s->queue.pop_front();
            auto *j = new Job();
            j->r = r;
            j->fp = fp;
            s->in_use += fp;
            s->peak_use = std::max(s->peak_use, s->in_use);
            r->t_admit = s->now();
            s->jobs.push_back(j);
            s->peak_active = std::max(s->peak_active, (int)s->jobs.size());
        }
        std::vector<Job *> ready;
        for (Job *j : s->jobs)
            if (!j->busy && (j->stage == ST_FIRST || j->stage == ST_FINAL)) ready.push_back(j);
        if (ready.empty()) {
            if (s->stop && s->queue.empty() && s->jobs.empty()) return;
            s->cv_sched.wait(lk);
            continue;
        }
        for (Job *j : ready) {
            j->busy = true;
            if (!j->setup && j->pool < 0) j->pool = pool_take(*s, j->r);
        }
        lk.unlock();
        // trees: pooled ones are reshaped, new ones created and pooled below
        std::vector<plk_tree *> trees(ready.size(), nullptr), fresh(ready.size(), nullptr);
        for (size_t k = 0; k < ready.size(); k++) {
            Job *j = ready[k];
            if (!j->setup) {
                j->rc = setup_job(*s, j, &fresh[k]);
                j->setup = true;
            }
        }
        lk.lock();
        for (size_t k = 0; k < ready.size(); k++) {
            Job *j = ready[k];
            if (fresh[k]) {
                PoolTree p;
                p.t = fresh[k];
                p.nstates = j->r->nstates;
                p.ncat = j->r->ncat;
                p.prec = j->r->precision;
                p.scaling = j->r->scaling;
                p.busy = true;
                s->pool.push_back(p);
                j->pool = (int)s->pool.size() - 1;
            }
            trees[k] = j->pool >= 0 ? s->pool[j->pool].t : nullptr;
            if (!trees[k] && !j->rc) j->rc = PLK_ERR_STATE;
        }
        lk.unlock();
        run_step(*s, ready, trees);
        lk.lock();
        for (Job *j : ready) {
            j->busy = false;
            if (j->rc || j->stage == ST_FINAL || j->r->kind == PLK_REQ_LOGL) {
                finish(*s, j);
            } else {
                j->stage = ST_NEWTON;
                j->busy = true;
                s->newton_q.push_back(j);
                s->cv_newton.notify_one();
            }
        }
    }
}

static void newton_worker(Server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (s->newton_q.empty()) {
            if (s->stop && s->queue.empty() && s->jobs.empty()) return;
            s->cv_newton.wait(lk);
            continue;
        }
        Job *j = s->newton_q.front();
        s->newton_q.pop_front();
        plk_tree *t = s->pool[j->pool].t;
        plk_request *r = j->r;
        lk.unlock();
        int rc;
        bool done = false;
        if (r->kind == PLK_REQ_OPT_BRANCH) {
            rc = plk_tree_optimize_one_branch(t, r->opt_a, r->opt_b, 1e-6, 10.0, 100, &r->opt_len);
        } else {
            rc = plk_tree_optimize_all_branches(t, std::max(r->iterations, 1), r->tol > 0 ? r->tol : 1e-3, 1e-6, 10.0,
                                                100, &r->logl_after);
            done = true;
        }
        lk.lock();
        j->rc = rc;
        j->busy = false;
        if (rc || done) finish(*s, j);
        else j->stage = ST_FINAL;
        s->cv_sched.notify_one();
    }
}

}  // namespace plk

struct plk_server : public plk::Server {};

using plk::fail;

extern "C" int plk_server_create(plk_server **out, int nworkers, int64_t hbm_budget) {
    if (!out || nworkers < 1 || nworkers > 64 || hbm_budget <= 0) return fail(PLK_ERR_ARG, "plk_server_create: bad arguments");
    auto *s = new plk_server();
    s->max_active = nworkers;
    s->budget = hbm_budget;
    s->pool.reserve(4 * 64);  // entries are never moved while trees are in use
    s->sched = std::thread(plk::scheduler, s);
    for (int i = 0; i < plk::kNewtonWorkers; i++) s->newton.emplace_back(plk::newton_worker, s);
    *out = s;
    return PLK_OK;
}

extern "C" int plk_server_submit(plk_server *s, plk_request *r) {
    if (!s || !r) return fail(PLK_ERR_ARG, "plk_server_submit: null");
    if (r->ntaxa < 3 || r->nptn < 1) return fail(PLK_ERR_ARG, "plk_server_submit: bad problem size");
    r->footprint = plk::footprint(r);
    r->status = 1;  // pending
    std::lock_guard<std::mutex> lk(s->mu);
    r->t_submit = s->now();
    if (s->first_submit < 0) s->first_submit = r->t_submit;
    s->queue.push_back(r);
    s->submitted++;
    s->cv_sched.notify_one();
    return PLK_OK;
}

extern "C" int plk_server_drain(plk_server *s) {
    if (!s) return fail(PLK_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed == s->submitted; });
    return PLK_OK;
}

extern "C" double plk_server_now(plk_server *s) { return s ? s->now() : 0.0; }

extern "C" int plk_server_stats(plk_server *s, double *st) {
    if (!s || !st) return fail(PLK_ERR_ARG, "null");
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> l = s->latency;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        double pos = q * (l.size() - 1);
        size_t i = (size_t)std::floor(pos);
        double f = pos - i;
        return i + 1 < l.size() ? l[i] * (1 - f) + l[i + 1] * f : l[i];
    };
    double mean = 0.0;
    for (double v : l) mean += v;
    if (!l.empty()) mean /= l.size();
    st[0] = (double)s->completed;
    st[1] = (double)s->failed;
    st[2] = pct(0.5);
    st[3] = pct(0.99);
    st[4] = mean;
    st[5] = s->first_submit >= 0 ? s->last_done - s->first_submit : 0.0;
    st[6] = (double)s->peak_use;
    st[7] = (double)s->peak_active;
    return PLK_OK;
}

extern "C" int plk_server_counters(plk_server *s, int64_t *out) {
    if (!s || !out) return fail(PLK_ERR_ARG, "null");
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = s->launches;
    out[1] = s->batched;
    out[2] = (int64_t)s->pool.size();
    return PLK_OK;
}

extern "C" int plk_server_destroy(plk_server *s) {
    if (!s) return PLK_OK;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->completed == s->submitted; });
        s->stop = true;
        s->cv_sched.notify_all();
        s->cv_newton.notify_all();
    }
    s->sched.join();
    for (auto &t : s->newton) t.join();
    for (plk_batch *b : s->batch)
        if (b) plk_batch_destroy(b);
    for (auto &p : s->pool) plk_tree_destroy(p.t);
    delete s;
    return PLK_OK;
}
