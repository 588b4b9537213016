// Serving loop: FIFO admission by HBM footprint onto a pool of worker threads, one
// context (own HIP stream, stream-ordered allocations) per admitted request.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "plk.h"
#include "plk_internal.h"
#include "plk_serve.h"
#include "plk_tree.h"

namespace plk {

struct Server {
    int nworkers = 1;
    int64_t budget = 0, in_use = 0, peak_use = 0;
    int active = 0, peak_active = 0;
    std::deque<plk_request *> queue;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::thread> threads;
    bool stop = false;
    int64_t submitted = 0, completed = 0, failed = 0;
    std::vector<double> latency;
    double first_submit = -1.0, last_done = 0.0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static int64_t footprint(const plk_request *r) {
    return plk_memory_required(r->nstates, r->ncat, r->ntaxa, r->nptn, r->ntaxa - 2, r->precision, r->scaling);
}

// One request, start to finish, on a fresh context.
static int run_request(plk_request *r) {
    plk_tree *t = nullptr;
    int rc = plk_tree_create(&t, r->nstates, r->ncat, r->ntaxa, r->nptn, r->precision, r->scaling);
    if (rc) return rc;
    rc = plk_tree_set_model(t, r->eval, r->evec, r->inv_evec, r->rates, r->props, r->state_unknown, r->tip_lh);
    if (!rc) rc = plk_tree_set_patterns(t, r->states, r->freq, r->invar);
    if (!rc) rc = plk_tree_set_topology(t, r->nei, r->len);
    if (!rc) rc = plk_tree_compute_likelihood(t, &r->logl_before);
    r->logl_after = r->logl_before;
    if (!rc && r->kind == PLK_REQ_OPT_BRANCH) {
        rc = plk_tree_optimize_one_branch(t, r->opt_a, r->opt_b, 1e-6, 10.0, 100, &r->opt_len);
        if (!rc) rc = plk_tree_loglik(t, r->opt_a, r->opt_b, &r->logl_after, nullptr);
    } else if (!rc && r->kind == PLK_REQ_OPT_ALL) {
        rc = plk_tree_optimize_all_branches(t, std::max(r->iterations, 1), r->tol > 0 ? r->tol : 1e-3, 1e-6, 10.0, 100,
                                            &r->logl_after);
    }
    plk_tree_destroy(t);
    return rc;
}

static void worker(Server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (s->stop && s->queue.empty()) return;
        if (!s->queue.empty()) {
            plk_request *r = s->queue.front();
            const int64_t fp = r->footprint;
            if (fp > s->budget) {
                // can never fit: fail it now rather than block the queue
                s->queue.pop_front();
                r->status = PLK_ERR_NOMEM;
                r->t_admit = r->t_done = s->now();
                s->completed++;
                s->failed++;
                s->cv_done.notify_all();
                continue;
            }
            if (s->in_use + fp <= s->budget) {
                // admit in arrival order (no overtaking, so large requests are not starved)
                s->queue.pop_front();
                s->in_use += fp;
                s->active++;
                s->peak_use = std::max(s->peak_use, s->in_use);
                s->peak_active = std::max(s->peak_active, s->active);
                r->t_admit = s->now();
                lk.unlock();
                int rc = run_request(r);
                lk.lock();
                r->status = rc;
                r->t_done = s->now();
                s->in_use -= fp;
                s->active--;
                s->completed++;
                if (rc) s->failed++;
                else s->latency.push_back(r->t_done - r->t_submit);
                s->last_done = std::max(s->last_done, r->t_done);
                s->cv_work.notify_all();
                s->cv_done.notify_all();
                continue;
            }
        }
        s->cv_work.wait(lk);
    }
}

}  // namespace plk

struct plk_server : public plk::Server {};

using plk::fail;

extern "C" int plk_server_create(plk_server **out, int nworkers, int64_t hbm_budget) {
    if (!out || nworkers < 1 || nworkers > 64 || hbm_budget <= 0) return fail(PLK_ERR_ARG, "plk_server_create: bad arguments");
    auto *s = new plk_server();
    s->nworkers = nworkers;
    s->budget = hbm_budget;
    for (int i = 0; i < nworkers; i++) s->threads.emplace_back(plk::worker, s);
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
    s->cv_work.notify_all();
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

extern "C" int plk_server_destroy(plk_server *s) {
    if (!s) return PLK_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto &t : s->threads) t.join();
    delete s;
    return PLK_OK;
}
