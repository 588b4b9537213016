// PhyloTreeGPU: the host half of IQ-TREE's PhyloTree likelihood path, driving the
// GPU engine through the plk.h C ABI.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bfgs.h"
#include "brent.h"
#include "plk.h"
#include "plk_internal.h"
#include "plk_model.h"
#include "plk_tree.h"

namespace plk {

class PhyloTreeGPU {
  public:
    plk_ctx *ctx = nullptr;
    int ntaxa = 0, nnodes = 0;
    std::vector<int> nei;       // nnodes x 3
    std::vector<double> len;    // nnodes x 3
    // LM_PER_NODE (default): internal node v owns slot v - ntaxa, oriented toward
    // orient[v].  all_dirs: every direction (dad -> internal node) has its own slot
    // 3*(node - ntaxa) + k, k = index of dad among node's neighbours (3(T-2) slots; what
    // the many-branch derivative pass needs, and HBM is plentiful).
    bool all_dirs = false;
    std::vector<int> orient;
    std::vector<char> computed;  // per node (LM_PER_NODE) or per slot (all_dirs)
    std::vector<plk_op> ops;
    int64_t n_derivs = 0, n_ops = 0;
    // Newton branch solves run device-resident (plk_newton_branch) unless switched to
    // the host loop below, which issues one derivative call per iteration.
    bool device_newton = true;
    // batched NNI: scratch slots start at nni_base, 3 per candidate, nni_cands candidates
    int nni_base = 0, nni_cands = 0;
    // The model as last set, kept so that a new Gamma shape re-uploads it with new rates.
    // gamma_shape is 0 while the rates are the caller's own (not derived from a shape).
    int nstates = 0, ncat = 0, su = 0;
    std::vector<double> m_eval, m_evec, m_inv, m_rates, m_props, m_tip;
    bool model_stored = false;
    double gamma_shape = 0.0;
    bool gamma_median = false;  // category rates at the medians (RateGamma's cut_median)
    // Reversible substitution model given by exchangeabilities (upper triangle, row-major)
    // and fixed frequencies: the eigen system and tip vectors are derived from them.
    bool gtr_set = false;
    int seq_type = 0;
    std::vector<double> m_exch, m_pi;

    bool is_leaf(int v) const { return v < ntaxa; }

    int dir(int a, int b) const {
        for (int k = 0; k < 3; k++)
            if (nei[a * 3 + k] == b) return a * 3 + k;
        return -1;
    }
    double length(int a, int b) const { return len[dir(a, b)]; }
    // slot of the partial of node's subtree seen from dad (node internal)
    int slot_of(int dad, int node) const { return all_dirs ? 3 * (node - ntaxa) + (dir(node, dad) - node * 3) : node - ntaxa; }
    // node reference of the partial of `to`'s subtree seen from `from`
    int ref_dir(int from, int to) const { return is_leaf(to) ? -(to + 1) : slot_of(from, to); }
    bool is_computed(int dad, int node) const {
        return all_dirs ? computed[slot_of(dad, node)] != 0 : (orient[node] == dad && computed[node]);
    }
    void set_computed(int dad, int node, bool v) {
        if (all_dirs) computed[slot_of(dad, node)] = v;
        else {
            orient[node] = dad;
            computed[node] = v;
        }
    }

    // subtree sizes by direction (dad's neighbour slot of node), -1 while unknown; the
    // topology changes only in plk_tree_set_topology, which clears them
    mutable std::vector<int> ssize;
    int subtree_size(int node, int dad) const {
        if (is_leaf(node)) return 1;
        const int d = dad >= 0 ? dir(dad, node) : -1;
        if (d >= 0 && ssize[d] >= 0) return ssize[d];
        int s = 0;
        for (int k = 0; k < 3; k++) {
            int x = nei[node * 3 + k];
            if (x >= 0 && x != dad) s += subtree_size(x, node);
        }
        if (d >= 0) ssize[d] = s;
        return s;
    }

    // computeTraversalInfo(PhyloNeighbor *dad_branch, PhyloNode *dad)
    // (tree/phylotree.cpp:6031-6139): post-order over dirty directions, children in
    // decreasing subtree size; reorientPartialLh (6011-6025) moves the node's slot.
    void collect(int dad, int node) {
        if (is_leaf(node)) return;
        if (is_computed(dad, node)) return;
        set_computed(dad, node, false);
        int kids[2], nk = 0;
        for (int k = 0; k < 3; k++) {
            int x = nei[node * 3 + k];
            if (x >= 0 && x != dad) kids[nk++] = x;
        }
        if (subtree_size(kids[0], node) >= subtree_size(kids[1], node)) {
            collect(node, kids[0]);
            collect(node, kids[1]);
        } else {
            collect(node, kids[1]);
            collect(node, kids[0]);
        }
        set_computed(dad, node, true);
        plk_op op{};
        op.out_slot = slot_of(dad, node);
        op.child[0] = ref_dir(node, kids[0]);
        op.child[1] = ref_dir(node, kids[1]);
        op.len[0] = length(node, kids[0]);
        op.len[1] = length(node, kids[1]);
        ops.push_back(op);
    }

    // both directions of branch (dad, node), then queue the ops
    int plan(int dad, int node) {
        ops.clear();
        collect(node, dad);
        collect(dad, node);
        n_ops += (int64_t)ops.size();
        if (ops.empty()) return PLK_OK;
        return plk_update_partials(ctx, ops.data(), (int)ops.size());
    }

    // PhyloNode::clearReversePartialLh (tree/phylonode.cpp:22-33)
    void clear_reverse(int node, int dad) {
        for (int k = 0; k < 3; k++) {
            int x = nei[node * 3 + k];
            if (x < 0 || x == dad) continue;
            if (!is_leaf(node)) {
                if (all_dirs) computed[slot_of(x, node)] = 0;
                else if (orient[node] == x) computed[node] = 0;
            }
            clear_reverse(x, node);
        }
    }

    void clear_all() { std::fill(computed.begin(), computed.end(), 0); }

    int set_length(int a, int b, double v) {
        int da = dir(a, b), db = dir(b, a);
        if (da < 0 || db < 0) return fail(PLK_ERR_ARG, "no such branch");
        len[da] = len[db] = v;
        clear_reverse(a, b);
        clear_reverse(b, a);
        return PLK_OK;
    }

    int loglik(int dad, int node, double *logl, double *ptn_lh) {
        if (dir(dad, node) < 0) return fail(PLK_ERR_ARG, "no such branch");
        int rc = plan(dad, node);
        if (rc) return rc;
        return plk_branch_loglik(ctx, ref_dir(node, dad), ref_dir(dad, node), length(dad, node), logl, ptn_lh);
    }

    int derivs(int dad, int node, int reuse, double *df, double *ddf) {
        if (dir(dad, node) < 0) return fail(PLK_ERR_ARG, "no such branch");
        int rc = plan(dad, node);
        if (rc) return rc;
        n_derivs++;
        return plk_branch_derivs(ctx, ref_dir(node, dad), ref_dir(dad, node), length(dad, node), reuse, df, ddf);
    }

    // df/ddf of every branch at its current length (IQ-TREE fills gradient_vector and
    // hessian_diagonal one computeLikelihoodDerv call at a time).  With all_dirs every
    // missing partial is computed in one traversal launch and all branches in one
    // derivative launch; LM_PER_NODE evaluates the branches one after another.
    int all_branch_derivs(std::vector<int> &ba, std::vector<int> &bb, std::vector<double> &df, std::vector<double> &ddf) {
        ba.clear();
        bb.clear();
        for (int u = 0; u < nnodes; u++)
            for (int k = 0; k < 3; k++) {
                int v = nei[u * 3 + k];
                if (v > u) {
                    ba.push_back(u);
                    bb.push_back(v);
                }
            }
        const int nb = (int)ba.size();
        df.assign(nb, 0.0);
        ddf.assign(nb, 0.0);
        if (!all_dirs) {
            for (int i = 0; i < nb; i++) {
                int rc = derivs(ba[i], bb[i], 0, &df[i], &ddf[i]);
                if (rc) return rc;
            }
            return PLK_OK;
        }
        ops.clear();
        std::vector<int> ra(nb), rb(nb);
        std::vector<double> t(nb);
        for (int i = 0; i < nb; i++) {
            collect(bb[i], ba[i]);
            collect(ba[i], bb[i]);
            ra[i] = ref_dir(bb[i], ba[i]);
            rb[i] = ref_dir(ba[i], bb[i]);
            t[i] = length(ba[i], bb[i]);
        }
        n_ops += (int64_t)ops.size();
        if (!ops.empty()) {
            int rc = plk_update_partials(ctx, ops.data(), (int)ops.size());
            if (rc) return rc;
        }
        n_derivs += nb;
        return plk_branch_derivs_multi(ctx, nb, ra.data(), rb.data(), t.data(), df.data(), ddf.data());
    }

    // computeFuncDerv (tree/phylotree.cpp:2555-2562): derivatives of -logL
    int func_derv(int a, int b, double value, bool &first, double &f, double &df) {
        int da = dir(a, b), db = dir(b, a);
        len[da] = len[db] = value;
        double d1 = 0, d2 = 0;
        int rc = derivs(a, b, first ? 0 : 1, &d1, &d2);
        first = false;
        f = -d1;
        df = -d2;
        return rc;
    }

    // Optimization::minimizeNewton (utils/optimization.cpp:422-499)
    int minimize_newton(int a, int b, double x1, double xguess, double x2, double xacc, int max_nr, double &out) {
        double df, dx, f, temp, xh, xl, rts, rts_old;
        bool first = true;
        rts = std::min(std::max(xguess, x1), x2);
        int rc = func_derv(a, b, rts, first, f, df);
        if (rc) return rc;
        if (!std::isfinite(f) || !std::isfinite(df)) return fail(PLK_ERR_NUMERIC, "Wrong computeFuncDerv");
        if (df >= 0.0 && std::fabs(f) < xacc) {
            out = rts;
            return PLK_OK;
        }
        if (f < 0.0) {
            xl = rts;
            xh = x2;
        } else {
            xh = rts;
            xl = x1;
        }
        dx = std::fabs(xh - xl);
        for (int j = 1; j <= max_nr; j++) {
            rts_old = rts;
            if ((df <= 0.0) || (((rts - xh) * df - f) * ((rts - xl) * df - f) >= 0.0)) {
                dx = 0.5 * (xh - xl);
                rts = xl + dx;
                if (xl == rts) {
                    out = rts;
                    return PLK_OK;
                }
            } else {
                dx = f / df;
                temp = rts;
                rts -= dx;
                if (temp == rts) {
                    out = rts;
                    return PLK_OK;
                }
            }
            if (std::fabs(dx) < xacc || j == max_nr) {
                out = rts_old;
                return PLK_OK;
            }
            rc = func_derv(a, b, rts, first, f, df);
            if (rc) return rc;
            if (!std::isfinite(f) || !std::isfinite(df)) return fail(PLK_ERR_NUMERIC, "Wrong computeFuncDerv");
            if (df > 0.0 && std::fabs(f) < xacc) {
                out = rts;
                return PLK_OK;
            }
            if (f < 0.0) xl = rts;
            else if (f > 0.0) xh = rts;
        }
        return fail(PLK_ERR_NUMERIC, "Maximum number of iterations exceeded in minimizeNewton");
    }

    // PhyloTree::optimizeOneBranch (tree/phylotree.cpp:2620-2670), Newton branch
    int optimize_one(int n1, int n2, double min_bl, double max_bl, int max_nr, double &optx) {
        if (dir(n1, n2) < 0) return fail(PLK_ERR_ARG, "no such branch");
        double current = length(n1, n2);
        int rc;
        if (device_newton) {
            rc = plan(n1, n2);
            if (rc) return rc;
            int evals = 0;
            rc = plk_newton_branch(ctx, ref_dir(n2, n1), ref_dir(n1, n2), min_bl, current, max_bl, min_bl, max_nr,
                                   &optx, &evals);
            n_derivs += evals;
        } else {
            rc = minimize_newton(n1, n2, min_bl, current, max_bl, min_bl, max_nr, optx);
        }
        if (rc) return rc;
        int da = dir(n1, n2), db = dir(n2, n1);
        if (optx > max_bl * 0.95) {
            double opt_lh = 0, orig_lh = 0;
            rc = plk_loglik_from_theta(ctx, optx, &opt_lh);
            if (rc) return rc;
            rc = plk_loglik_from_theta(ctx, current, &orig_lh);
            if (rc) return rc;
            if (orig_lh > opt_lh) optx = current;
        }
        len[da] = len[db] = optx;
        if (current != optx) {
            clear_reverse(n1, n2);
            clear_reverse(n2, n1);
        }
        return PLK_OK;
    }

    // getBestNNIForBran (tree/phylotree.cpp:4176-4429) with NNI5, for many internal
    // branches at once: both moves of every branch are evaluated in lockstep.  A move
    // swaps node1's first other neighbour A with node2's move-th other neighbour X (each
    // keeps its branch length), then optimizes node1's two other branches, the central
    // branch and node2's two other branches (NNI_MAX_NR_STEP Newton steps each); the
    // score is the logL at the last branch.  Each step is one traversal launch over all
    // candidates' new partials (scratch slots) plus one many-branch Newton solve; the
    // subtrees around a move are the tree's own directional partials.  The tree is not
    // changed.
    int nni_evaluate(int nbr, const int *n1s, const int *n2s, double min_bl, double max_bl, int max_nr,
                     double *score, double *lens) {
        if (!all_dirs || nni_cands == 0) return fail(PLK_ERR_STATE, "batched NNI needs a tree created with all_directions = 2");
        const int ncand = 2 * nbr;
        if (ncand > nni_cands) return fail(PLK_ERR_ARG, "plk_tree_nni_evaluate: too many branches");
        struct Cand {
            int n1, n2;
            int y[2], z[2];        // node1's / node2's other neighbours after the swap, in slot order
            int yref[2], zref[2];  // their subtrees' node references (tips or directional slots)
            double ylen[2], zlen[2], clen;
            int s1, s2, st;        // scratch slots: node1 side seen from node2, node2 side seen from node1, one more
        };
        std::vector<Cand> cs(ncand);
        for (int b = 0; b < nbr; b++) {
            const int n1 = n1s[b], n2 = n2s[b];
            if (n1 < ntaxa || n2 < ntaxa || n1 >= nnodes || n2 >= nnodes || dir(n1, n2) < 0)
                return fail(PLK_ERR_ARG, "plk_tree_nni_evaluate: not an internal branch");
            // every subtree partial around the branch, oriented toward it
            ops.clear();
            for (int k = 0; k < 3; k++) {
                const int u = nei[n1 * 3 + k], v = nei[n2 * 3 + k];
                if (u != n2) collect(n1, u);
                if (v != n1) collect(n2, v);
            }
            n_ops += (int64_t)ops.size();
            if (!ops.empty()) {
                int rc = plk_update_partials(ctx, ops.data(), (int)ops.size());
                if (rc) return rc;
            }
            int a_slot = -1, xs[2], m = 0;
            for (int k = 0; k < 3; k++)
                if (nei[n1 * 3 + k] != n2 && a_slot < 0) a_slot = k;
            for (int k = 0; k < 3; k++)
                if (nei[n2 * 3 + k] != n1) xs[m++] = k;
            for (int mv = 0; mv < 2; mv++) {
                Cand &c = cs[2 * b + mv];
                c.n1 = n1;
                c.n2 = n2;
                const int x_slot = xs[mv];
                const int A = nei[n1 * 3 + a_slot], X = nei[n2 * 3 + x_slot];
                int qi = 0, ri = 0;
                for (int k = 0; k < 3; k++) {
                    const int u = nei[n1 * 3 + k];
                    if (u == n2) continue;
                    const int yy = (k == a_slot) ? X : u;
                    c.y[qi] = yy;
                    c.yref[qi] = (k == a_slot) ? ref_dir(n2, X) : ref_dir(n1, u);
                    c.ylen[qi] = (k == a_slot) ? len[n2 * 3 + x_slot] : len[n1 * 3 + k];
                    qi++;
                }
                for (int k = 0; k < 3; k++) {
                    const int v = nei[n2 * 3 + k];
                    if (v == n1) continue;
                    const int zz = (k == x_slot) ? A : v;
                    c.z[ri] = zz;
                    c.zref[ri] = (k == x_slot) ? ref_dir(n1, A) : ref_dir(n2, v);
                    c.zlen[ri] = (k == x_slot) ? len[n1 * 3 + a_slot] : len[n2 * 3 + k];
                    ri++;
                }
                c.clen = length(n1, n2);
                const int base = nni_base + 3 * (2 * b + mv);
                c.s1 = base;
                c.s2 = base + 1;
                c.st = base + 2;
            }
        }
        // partials of the directions the moves read must be current before the scratch ops
        int rc = plk_flush(ctx);
        if (rc) return rc;
        std::vector<int> ra(ncand), rb(ncand), st(ncand);
        std::vector<double> guess(ncand), out(ncand), logl(ncand);
        std::vector<plk_op> nops;
        auto mkop = [](int out_slot, int c0, double l0, int c1, double l1) {
            plk_op o{};
            o.out_slot = out_slot;
            o.child[0] = c0;
            o.child[1] = c1;
            o.len[0] = l0;
            o.len[1] = l1;
            return o;
        };
        // one step: queue ops, then Newton on (a, b) from the current length; returns the
        // lengths after optimizeOneBranch's divergence check
        auto step = [&](auto make_ops, auto branch, auto cur) -> int {
            nops.clear();
            for (int i = 0; i < ncand; i++) make_ops(cs[i]);
            if (!nops.empty()) {
                int r = plk_update_partials(ctx, nops.data(), (int)nops.size());
                if (r) return r;
            }
            for (int i = 0; i < ncand; i++) {
                branch(cs[i], ra[i], rb[i]);
                guess[i] = cur(cs[i]);
            }
            int r = plk_newton_multi(ctx, ncand, ra.data(), rb.data(), min_bl, guess.data(), max_bl, min_bl, max_nr,
                                     out.data(), logl.data(), st.data());
            if (r) return r;
            for (int i = 0; i < ncand; i++)
                if (st[i] != PLK_OK) return fail(st[i], "batched NNI: a branch solve failed (non-finite derivative)");
            n_derivs += ncand;
            // optimizeOneBranch's divergence check (tree/phylotree.cpp:2642-2651)
            std::vector<int> div;
            for (int i = 0; i < ncand; i++)
                if (out[i] > max_bl * 0.95) div.push_back(i);
            if (!div.empty()) {
                std::vector<int> da, db;
                std::vector<double> dg, dout(div.size()), dl(div.size());
                std::vector<int> dst(div.size());
                for (int i : div) {
                    da.push_back(ra[i]);
                    db.push_back(rb[i]);
                    dg.push_back(guess[i]);
                }
                // logL at the starting length
                r = plk_newton_multi(ctx, (int)div.size(), da.data(), db.data(), min_bl, dg.data(), max_bl, min_bl, -1,
                                     dout.data(), dl.data(), dst.data());
                if (r) return r;
                for (size_t q = 0; q < div.size(); q++)
                    if (dl[q] > logl[div[q]]) {
                        out[div[q]] = guess[div[q]];
                        logl[div[q]] = dl[q];
                    }
            }
            return PLK_OK;
        };
        // 1. node1's first branch: node2 side seen from node1, node1 side seen from y0
        rc = step([&](Cand &c) {
                      nops.push_back(mkop(c.s2, c.zref[0], c.zlen[0], c.zref[1], c.zlen[1]));
                      nops.push_back(mkop(c.st, c.yref[1], c.ylen[1], c.s2, c.clen));
                  },
                  [&](Cand &c, int &a, int &b) { a = c.yref[0]; b = c.st; }, [&](Cand &c) { return c.ylen[0]; });
        if (rc) return rc;
        for (int i = 0; i < ncand; i++) cs[i].ylen[0] = out[i];
        // 2. node1's second branch
        rc = step([&](Cand &c) { nops.push_back(mkop(c.st, c.yref[0], c.ylen[0], c.s2, c.clen)); },
                  [&](Cand &c, int &a, int &b) { a = c.yref[1]; b = c.st; }, [&](Cand &c) { return c.ylen[1]; });
        if (rc) return rc;
        for (int i = 0; i < ncand; i++) cs[i].ylen[1] = out[i];
        // 3. the central branch
        rc = step([&](Cand &c) { nops.push_back(mkop(c.s1, c.yref[0], c.ylen[0], c.yref[1], c.ylen[1])); },
                  [&](Cand &c, int &a, int &b) { a = c.s1; b = c.s2; }, [&](Cand &c) { return c.clen; });
        if (rc) return rc;
        for (int i = 0; i < ncand; i++) cs[i].clen = out[i];
        // 4. node2's first branch
        rc = step([&](Cand &c) { nops.push_back(mkop(c.st, c.zref[1], c.zlen[1], c.s1, c.clen)); },
                  [&](Cand &c, int &a, int &b) { a = c.zref[0]; b = c.st; }, [&](Cand &c) { return c.zlen[0]; });
        if (rc) return rc;
        for (int i = 0; i < ncand; i++) cs[i].zlen[0] = out[i];
        // 5. node2's second branch; the score is the logL there
        rc = step([&](Cand &c) { nops.push_back(mkop(c.st, c.zref[0], c.zlen[0], c.s1, c.clen)); },
                  [&](Cand &c, int &a, int &b) { a = c.zref[1]; b = c.st; }, [&](Cand &c) { return c.zlen[1]; });
        if (rc) return rc;
        for (int i = 0; i < ncand; i++) {
            cs[i].zlen[1] = out[i];
            if (score) score[i] = logl[i];
            if (lens) {
                double *l = lens + 5 * i;
                l[0] = cs[i].clen;
                l[1] = cs[i].ylen[0];
                l[2] = cs[i].ylen[1];
                l[3] = cs[i].zlen[0];
                l[4] = cs[i].zlen[1];
            }
        }
        return PLK_OK;
    }

    // MTree::findFarthestLeaf (tree/mtree.cpp:3040-3060)
    int farthest_leaf(int node, int dad, std::vector<int> &h) const {
        if (dad >= 0 && is_leaf(node)) {
            h[node] = 0;
            return node;
        }
        int res = -1;
        h[node] = 0;
        for (int k = 0; k < 3; k++) {
            int x = nei[node * 3 + k];
            if (x < 0 || x == dad) continue;
            int leaf = farthest_leaf(x, node, h);
            if (h[node] < h[x] + 1) {
                h[node] = h[x] + 1;
                res = leaf;
            }
        }
        return res;
    }

    // MTree::getPreOrderBranches (tree/mtree.cpp:3062-3082)
    void preorder(int node, int dad, const std::vector<int> &h, std::vector<int> &a, std::vector<int> &b) const {
        if (dad >= 0) {
            a.push_back(node);
            b.push_back(dad);
        }
        int nv[3], m = 0;
        for (int k = 0; k < 3; k++)
            if (nei[node * 3 + k] >= 0) nv[m++] = nei[node * 3 + k];
        for (int i = 0; i < m; i++)
            for (int j = i + 1; j < m; j++)
                if (h[nv[i]] > h[nv[j]]) std::swap(nv[i], nv[j]);
        for (int i = 0; i < m; i++)
            if (nv[i] != dad) preorder(nv[i], node, h, a, b);
    }

    // PhyloTree::optimizeAllBranches(iterations, tol, maxNRStep) (tree/phylotree.cpp:2724-2801)
    int optimize_all(int iterations, double tol, double min_bl, double max_bl, int max_nr, double &out) {
        std::vector<int> h(nnodes, 0), na, nb;
        int far = farthest_leaf(0, -1, h);
        farthest_leaf(far, -1, h);
        preorder(far, -1, h, na, nb);
        double tree_lh = 0.0;
        int rc = loglik(na[0], nb[0], &tree_lh, nullptr);
        if (rc) return rc;
        std::vector<double> saved;
        for (int it = 0; it < iterations; it++) {
            saved = len;
            for (size_t j = 0; j < na.size(); j++) {
                double optx;
                rc = optimize_one(na[j], nb[j], min_bl, max_bl, max_nr, optx);
                if (rc) return rc;
            }
            double new_lh = 0.0;
            rc = plk_loglik_from_theta(ctx, length(na.back(), nb.back()), &new_lh);
            if (rc) return rc;
            if (new_lh < tree_lh - tol * 0.1) {
                len = saved;
                clear_all();
                rc = loglik(na[0], nb[0], &tree_lh, nullptr);
                if (rc) return rc;
                break;
            }
            if (tree_lh <= new_lh && new_lh <= tree_lh + tol) {
                tree_lh = new_lh;
                break;
            }
            tree_lh = new_lh;
        }
        out = tree_lh;
        return PLK_OK;
    }

    int upload_model() {
        clear_all();
        return plk_set_model(ctx, m_eval.data(), m_evec.data(), m_inv.data(), m_rates.data(), m_props.data(), su,
                             m_tip.data());
    }

    // RateGamma::setGammaShape -> computeRates (model/rategamma.cpp:173-176, 98-171), then
    // clearAllPartialLH: every partial depends on the category rates.
    int set_gamma(double alpha) {
        if (!model_stored) return fail(PLK_ERR_STATE, "Gamma shape needs a model (plk_tree_set_model) first");
        std::vector<double> r(ncat);
        int rc = gamma_median ? plk_gamma_rates_median(alpha, ncat, r.data()) : plk_gamma_rates(alpha, ncat, r.data());
        if (rc) return rc;
        m_rates = r;
        gamma_shape = alpha;
        return upload_model();
    }

    // RateGamma::computeFunction (model/rategamma.cpp:178-185): -logL at shape; the
    // partials are recomputed only when the shape changes.
    int gamma_function(double shape, double &f) {
        int rc;
        if (shape != gamma_shape && (rc = set_gamma(shape))) return rc;
        double lnl = 0.0;
        rc = loglik(0, nei[0], &lnl, nullptr);
        f = -lnl;
        return rc;
    }

    // ModelMarkov::decomposeRateMatrix (model/modelmarkov.cpp:1389-1635) for the stored
    // exchangeabilities, then the tip vectors; the model is re-uploaded (partials dirty).
    int decompose_gtr() {
        const int n = nstates;
        std::vector<double> Q((size_t)n * n), ev(n), U((size_t)n * n), Ui((size_t)n * n);
        int rc = plk_rate_matrix(n, m_exch.data(), m_pi.data(), Q.data());
        if (!rc) rc = plk_eigen(n, Q.data(), m_pi.data(), ev.data(), U.data(), Ui.data());
        if (rc) return rc;
        std::vector<double> tip((size_t)(su + 1) * n);
        if ((rc = plk_tip_lh(n, seq_type, Ui.data(), tip.data()))) return rc;
        m_eval = ev;
        m_evec = U;
        m_inv = Ui;
        m_tip = tip;
        return upload_model();
    }

    // ModelMarkov::targetFunk (model/modelmarkov.cpp:1091-1113): the free exchangeabilities
    // (all but the last, which stays 1) from x; a change re-decomposes Q; -logL.
    int gtr_function(const std::vector<double> &x, double &f) {
        bool changed = false;
        for (size_t i = 0; i < x.size(); i++) changed |= m_exch[i] != x[i];
        int rc;
        if (changed) {
            std::copy(x.begin(), x.end(), m_exch.begin());
            if ((rc = decompose_gtr())) return rc;
        }
        double lnl = 0.0;
        rc = loglik(0, nei[0], &lnl, nullptr);
        f = -lnl;
        return rc;
    }

    // ModelMarkov::optimizeParameters (model/modelmarkov.cpp:1166-1243) with fixed
    // frequencies: BFGS (Optimization::minimizeMultiDimen, utils/optimization.cpp:750-906)
    // over the free exchangeabilities in [MIN_RATE, MAX_RATE] = [1e-4, 100] with gradient
    // tolerance max(gradient_epsilon, TOL_RATE = 1e-4) (model/modelmarkov.h:30-32); no
    // bound is checked, so the search is not restarted.  The model is left at the result.
    int optimize_gtr(double gradient_epsilon, double &logl) {
        if (!gtr_set) return fail(PLK_ERR_STATE, "plk_tree_optimize_model: no exchangeabilities set");
        const int ndim = (int)m_exch.size() - 1;
        if (ndim < 1) return fail(PLK_ERR_STATE, "plk_tree_optimize_model: nothing to optimise");
        std::vector<double> x(m_exch.begin(), m_exch.begin() + ndim), lo(ndim, 1e-4), hi(ndim, 100.0);
        auto f = [this](const std::vector<double> &v, double &fv) { return gtr_function(v, fv); };
        double fret = 0.0;
        int rc = bfgs_minimize(f, x, lo, hi, std::max(gradient_epsilon, 1e-4), fret);
        if (rc) return rc;
        double fx = 0.0;
        rc = gtr_function(x, fx);  // getVariables: the model at the returned point
        logl = -fx;
        return rc;
    }

    // RateGamma::optimizeParameters(gradient_epsilon, min_gamma, max_gamma)
    // (model/rategamma.cpp:215-232): minimizeOneDimen over the shape with tolerance
    // max(gradient_epsilon, TOL_GAMMA_SHAPE = 0.001) (utils/tools.h:607); the model is left
    // at the returned shape and its logL is returned.
    int optimize_gamma(double gradient_epsilon, double min_gamma, double max_gamma, double &logl) {
        if (!(gamma_shape > 0.0)) return fail(PLK_ERR_STATE, "plk_tree_optimize_gamma: no Gamma shape set");
        auto f = [this](double x, double &fx) { return gamma_function(x, fx); };
        double xopt = 0.0, fopt = 0.0, f2 = 0.0;
        int rc = minimize_one_dimen(f, min_gamma, gamma_shape, max_gamma, std::max(gradient_epsilon, 0.001), xopt, fopt,
                                    f2);
        if (rc) return rc;
        double fx = 0.0;
        rc = gamma_function(xopt, fx);
        logl = -fx;
        return rc;
    }

    // ModelFactory::optimizeParameters (model/modelfactory.cpp:1558-1720) with branch
    // lengths optimised (BRLEN_OPTIMIZE), the exchangeabilities (when set with
    // plk_tree_set_gtr) and a Gamma rate model (when a shape is set): alternate
    // all-branch Newton passes (min(i, 3) rounds) with the model parameters until an
    // iteration gains no more than logl_epsilon, then a final optimizeAllBranches(100).
    // RateGamma::rescaleRates is the base class's 1.0 (model/rateheterogeneity.h:288), so
    // no length rescaling follows.
    int optimize_parameters(int num_iterations, double logl_epsilon, double gradient_epsilon, double min_gamma,
                            double max_gamma, double min_bl, double max_bl, int max_nr, double &out) {
        const bool free_model = gtr_set && m_exch.size() > 1, free_rates = gamma_shape > 0.0;
        if (!free_model && !free_rates)
            return fail(PLK_ERR_STATE, "plk_tree_optimize_parameters: neither exchangeabilities nor a Gamma shape set");
        double cur = 0.0;
        int rc = loglik(0, nei[0], &cur, nullptr);
        if (rc) return rc;
        for (int i = 2; i < num_iterations; i++) {
            double new_lh = 0.0;
            if ((rc = optimize_all(std::min(i, 3), logl_epsilon, min_bl, max_bl, max_nr, new_lh))) return rc;
            // optimizeParametersOnly (1263-1315): the substitution model, then the rates;
            // the logL is the rate model's when it has one
            if (free_model && (rc = optimize_gtr(gradient_epsilon, new_lh))) return rc;
            if (free_rates && (rc = optimize_gamma(gradient_epsilon, min_gamma, max_gamma, new_lh))) return rc;
            if (new_lh > cur + logl_epsilon) {
                cur = new_lh;
            } else {
                if ((rc = optimize_all(100, logl_epsilon, min_bl, max_bl, max_nr, cur))) return rc;
                break;
            }
        }
        out = cur;
        return PLK_OK;
    }
};

}  // namespace plk

struct plk_tree : public plk::PhyloTreeGPU {};

using plk::fail;

extern "C" int plk_tree_create_ex(plk_tree **out, int nstates, int ncat, int ntaxa, int nptn, int precision,
                                  int scaling, int all_directions) {
    if (!out) return fail(PLK_ERR_ARG, "plk_tree_create: null out");
    *out = nullptr;
    if (ntaxa < 3) return fail(PLK_ERR_ARG, "plk_tree_create: need at least 3 taxa");
    auto *t = new plk_tree();
    t->ntaxa = ntaxa;
    t->nnodes = 2 * ntaxa - 2;
    t->all_dirs = all_directions != 0;
    t->nni_base = 3 * (ntaxa - 2);
    t->nni_cands = all_directions == 2 ? 2 * (ntaxa - 3) : 0;
    t->nstates = nstates;
    t->ncat = ncat;
    t->nei.assign((size_t)t->nnodes * 3, -1);
    t->ssize.assign((size_t)t->nnodes * 3, -1);
    t->len.assign((size_t)t->nnodes * 3, 0.0);
    t->orient.assign(t->nnodes, -1);
    t->computed.assign(t->all_dirs ? (size_t)3 * (ntaxa - 2) : (size_t)t->nnodes, 0);
    // all_directions == 2 also reserves 3 scratch slots per NNI candidate (2 per internal branch)
    const int nslots = t->all_dirs ? 3 * (ntaxa - 2) + 3 * t->nni_cands : ntaxa - 2;
    int rc = plk_create(&t->ctx, nstates, ncat, ntaxa, nptn, nslots, precision, scaling);
    if (rc) {
        delete t;
        return rc;
    }
    *out = t;
    return PLK_OK;
}

extern "C" int plk_tree_create(plk_tree **out, int nstates, int ncat, int ntaxa, int nptn, int precision, int scaling) {
    return plk_tree_create_ex(out, nstates, ncat, ntaxa, nptn, precision, scaling, 0);
}

extern "C" int plk_tree_reshape(plk_tree *t, int ntaxa, int nptn) {
    if (!t) return fail(PLK_ERR_ARG, "plk_tree_reshape: null tree");
    if (t->all_dirs) return fail(PLK_ERR_STATE, "plk_tree_reshape: trees with one slot per direction are not reshaped");
    if (ntaxa < 3) return fail(PLK_ERR_ARG, "plk_tree_reshape: need at least 3 taxa");
    int rc = plk_reshape(t->ctx, ntaxa, nptn, ntaxa - 2);
    if (rc) return rc;
    t->ntaxa = ntaxa;
    t->nnodes = 2 * ntaxa - 2;
    t->nni_base = 3 * (ntaxa - 2);
    t->nei.assign((size_t)t->nnodes * 3, -1);
    t->ssize.assign((size_t)t->nnodes * 3, -1);
    t->len.assign((size_t)t->nnodes * 3, 0.0);
    t->orient.assign(t->nnodes, -1);
    t->computed.assign((size_t)t->nnodes, 0);
    t->ops.clear();
    t->model_stored = false;
    t->gamma_shape = 0.0;
    t->gtr_set = false;
    return PLK_OK;
}

extern "C" int plk_tree_destroy(plk_tree *t) {
    if (!t) return PLK_OK;
    plk_destroy(t->ctx);
    delete t;
    return PLK_OK;
}

extern "C" plk_ctx *plk_tree_context(plk_tree *t) { return t ? t->ctx : nullptr; }

extern "C" int plk_tree_set_model(plk_tree *t, const double *eval, const double *evec, const double *inv_evec,
                                  const double *rates, const double *props, int su, const double *tip_lh) {
    if (!t) return fail(PLK_ERR_ARG, "null tree");
    t->clear_all();
    int rc = plk_set_model(t->ctx, eval, evec, inv_evec, rates, props, su, tip_lh);
    if (rc) return rc;
    const size_t n = (size_t)t->nstates, nc = (size_t)t->ncat;
    t->su = su;
    t->m_eval.assign(eval, eval + n);
    t->m_evec.assign(evec, evec + n * n);
    t->m_inv.assign(inv_evec, inv_evec + n * n);
    t->m_rates.assign(rates, rates + nc);
    t->m_props.assign(props, props + nc);
    t->m_tip.assign(tip_lh, tip_lh + (size_t)(su + 1) * n);
    t->model_stored = true;
    t->gamma_shape = 0.0;
    t->gtr_set = false;
    return PLK_OK;
}

extern "C" int plk_tree_set_gamma_shape(plk_tree *t, double alpha) {
    if (!t || !(alpha > 0.0) || !std::isfinite(alpha)) return fail(PLK_ERR_ARG, "plk_tree_set_gamma_shape: bad arguments");
    return t->set_gamma(alpha);
}

extern "C" int plk_tree_set_gtr(plk_tree *t, int seq_type, const double *exch, const double *pi) {
    if (!t || !exch || !pi) return fail(PLK_ERR_ARG, "plk_tree_set_gtr: bad arguments");
    if (!t->model_stored) return fail(PLK_ERR_STATE, "plk_tree_set_gtr: set a model (rates, proportions) first");
    const int n = t->nstates, ne = n * (n - 1) / 2;
    for (int i = 0; i < ne; i++)
        if (!(exch[i] > 0.0) || !std::isfinite(exch[i])) return fail(PLK_ERR_ARG, "plk_tree_set_gtr: bad exchangeability");
    for (int i = 0; i < n; i++)
        if (!(pi[i] > 0.0)) return fail(PLK_ERR_ARG, "plk_tree_set_gtr: frequencies must be positive");
    t->seq_type = seq_type;
    t->m_exch.assign(exch, exch + ne);
    t->m_pi.assign(pi, pi + n);
    t->gtr_set = true;
    return t->decompose_gtr();
}

extern "C" int plk_tree_get_gtr(plk_tree *t, double *exch) {
    if (!t || !exch || !t->gtr_set) return fail(PLK_ERR_ARG, "plk_tree_get_gtr: bad arguments");
    std::copy(t->m_exch.begin(), t->m_exch.end(), exch);
    return PLK_OK;
}

extern "C" int plk_tree_optimize_model(plk_tree *t, double gradient_epsilon, double *logl) {
    if (!t || !logl || !(gradient_epsilon >= 0.0)) return fail(PLK_ERR_ARG, "plk_tree_optimize_model: bad arguments");
    return t->optimize_gtr(gradient_epsilon, *logl);
}

extern "C" int plk_tree_set_gamma_median(plk_tree *t, int median) {
    if (!t) return fail(PLK_ERR_ARG, "plk_tree_set_gamma_median: null tree");
    t->gamma_median = median != 0;
    return PLK_OK;
}

extern "C" int plk_tree_get_gamma_shape(plk_tree *t, double *alpha) {
    if (!t || !alpha) return fail(PLK_ERR_ARG, "plk_tree_get_gamma_shape: bad arguments");
    *alpha = t->gamma_shape;
    return PLK_OK;
}

static bool gamma_bounds_ok(double min_gamma, double max_gamma) {
    return min_gamma > 0.0 && max_gamma > min_gamma && std::isfinite(max_gamma);
}

extern "C" int plk_tree_optimize_gamma(plk_tree *t, double gradient_epsilon, double min_gamma, double max_gamma,
                                       double *alpha, double *logl) {
    if (!t || !logl || !(gradient_epsilon >= 0.0) || !gamma_bounds_ok(min_gamma, max_gamma))
        return fail(PLK_ERR_ARG, "plk_tree_optimize_gamma: bad arguments");
    int rc = t->optimize_gamma(gradient_epsilon, min_gamma, max_gamma, *logl);
    if (alpha) *alpha = t->gamma_shape;
    return rc;
}

extern "C" int plk_tree_optimize_parameters(plk_tree *t, int num_param_iterations, double logl_epsilon,
                                            double gradient_epsilon, double min_gamma, double max_gamma, double min_bl,
                                            double max_bl, int max_nr_step, double *alpha, double *logl) {
    if (!t || !logl || num_param_iterations < 0 || !(logl_epsilon >= 0.0) || !(gradient_epsilon >= 0.0) ||
        !gamma_bounds_ok(min_gamma, max_gamma) || !(min_bl >= 0.0) || !(max_bl > min_bl) || max_nr_step < 0)
        return fail(PLK_ERR_ARG, "plk_tree_optimize_parameters: bad arguments");
    int rc = t->optimize_parameters(num_param_iterations, logl_epsilon, gradient_epsilon, min_gamma, max_gamma, min_bl,
                                    max_bl, max_nr_step, *logl);
    if (alpha) *alpha = t->gamma_shape;
    return rc;
}

extern "C" int plk_tree_set_patterns(plk_tree *t, const uint8_t *states, const double *freq, const double *invar) {
    if (!t) return fail(PLK_ERR_ARG, "null tree");
    t->clear_all();
    return plk_set_patterns(t->ctx, states, freq, invar);
}

extern "C" int plk_tree_set_topology(plk_tree *t, const int32_t *nei, const double *len) {
    if (!t || !nei || !len) return fail(PLK_ERR_ARG, "plk_tree_set_topology: null argument");
    // validate: symmetric adjacency, leaves degree 1 (slot 0), internal degree 3
    for (int v = 0; v < t->nnodes; v++) {
        int deg = 0;
        for (int k = 0; k < 3; k++) {
            int x = nei[v * 3 + k];
            if (x < -1 || x >= t->nnodes || x == v) return fail(PLK_ERR_ARG, "plk_tree_set_topology: bad neighbour");
            if (x < 0) continue;
            deg++;
            bool back = false;
            for (int q = 0; q < 3; q++)
                if (nei[x * 3 + q] == v) back = std::fabs(len[x * 3 + q] - len[v * 3 + k]) == 0.0;
            if (!back) return fail(PLK_ERR_ARG, "plk_tree_set_topology: adjacency/lengths not symmetric");
            if (!(len[v * 3 + k] >= 0.0) || !std::isfinite(len[v * 3 + k]))
                return fail(PLK_ERR_ARG, "plk_tree_set_topology: bad branch length");
        }
        if (v < t->ntaxa ? (deg != 1 || nei[v * 3] < 0) : deg != 3)
            return fail(PLK_ERR_ARG, "plk_tree_set_topology: tree must be binary with leaves 0..ntaxa-1");
    }
    t->nei.assign(nei, nei + (size_t)t->nnodes * 3);
    std::fill(t->ssize.begin(), t->ssize.end(), -1);
    t->len.assign(len, len + (size_t)t->nnodes * 3);
    std::fill(t->orient.begin(), t->orient.end(), -1);
    t->clear_all();
    return PLK_OK;
}

extern "C" int plk_tree_set_length(plk_tree *t, int a, int b, double len) {
    if (!t || a < 0 || b < 0 || a >= t->nnodes || b >= t->nnodes || !(len >= 0.0))
        return fail(PLK_ERR_ARG, "plk_tree_set_length: bad arguments");
    return t->set_length(a, b, len);
}

extern "C" int plk_tree_get_length(plk_tree *t, int a, int b, double *len) {
    if (!t || !len || a < 0 || b < 0 || a >= t->nnodes || b >= t->nnodes || t->dir(a, b) < 0)
        return fail(PLK_ERR_ARG, "plk_tree_get_length: bad arguments");
    *len = t->length(a, b);
    return PLK_OK;
}

extern "C" int plk_tree_clear_all(plk_tree *t) {
    if (!t) return fail(PLK_ERR_ARG, "null tree");
    t->clear_all();
    return PLK_OK;
}

static bool node_ok(plk_tree *t, int a, int b) {
    return t && a >= 0 && b >= 0 && a < t->nnodes && b < t->nnodes && t->dir(a, b) >= 0;
}

extern "C" int plk_tree_loglik(plk_tree *t, int dad, int node, double *logl, double *ptn_lh) {
    if (!node_ok(t, dad, node) || !logl) return fail(PLK_ERR_ARG, "plk_tree_loglik: bad arguments");
    return t->loglik(dad, node, logl, ptn_lh);
}

extern "C" int plk_tree_compute_likelihood(plk_tree *t, double *logl) {
    if (!t || !logl) return fail(PLK_ERR_ARG, "plk_tree_compute_likelihood: bad arguments");
    // root is leaf 0; current_it = its branch (tree/phylotree.cpp:1258-1299)
    return t->loglik(0, t->nei[0], logl, nullptr);
}

extern "C" int plk_tree_derivs(plk_tree *t, int dad, int node, int reuse, double *df, double *ddf) {
    if (!node_ok(t, dad, node) || !df || !ddf) return fail(PLK_ERR_ARG, "plk_tree_derivs: bad arguments");
    return t->derivs(dad, node, reuse, df, ddf);
}

extern "C" int plk_tree_loglik_from_theta(plk_tree *t, int dad, int node, double *logl) {
    if (!node_ok(t, dad, node) || !logl) return fail(PLK_ERR_ARG, "plk_tree_loglik_from_theta: bad arguments");
    return plk_loglik_from_theta(t->ctx, t->length(dad, node), logl);
}

extern "C" int plk_tree_get_partial(plk_tree *t, int dad, int node, double *partial, uint16_t *scale) {
    if (!node_ok(t, dad, node) || node < t->ntaxa) return fail(PLK_ERR_ARG, "plk_tree_get_partial: bad arguments");
    t->ops.clear();
    t->collect(dad, node);
    if (!t->ops.empty()) {
        int rc = plk_update_partials(t->ctx, t->ops.data(), (int)t->ops.size());
        if (rc) return rc;
    }
    return plk_get_partial(t->ctx, t->slot_of(dad, node), partial, scale);
}

extern "C" int plk_tree_nni_evaluate(plk_tree *t, int nbr, const int32_t *node1, const int32_t *node2, double min_bl,
                                     double max_bl, int max_nr_step, double *score, double *lens) {
    if (!t || nbr < 1 || !node1 || !node2 || !(min_bl >= 0) || !(max_bl > min_bl) || max_nr_step < 0)
        return fail(PLK_ERR_ARG, "plk_tree_nni_evaluate: bad arguments");
    return t->nni_evaluate(nbr, node1, node2, min_bl, max_bl, max_nr_step, score, lens);
}

extern "C" int plk_tree_set_newton(plk_tree *t, int device_resident) {
    if (!t) return fail(PLK_ERR_ARG, "plk_tree_set_newton: null tree");
    t->device_newton = device_resident != 0;
    return PLK_OK;
}

extern "C" int plk_tree_optimize_one_branch(plk_tree *t, int n1, int n2, double min_bl, double max_bl, int max_nr,
                                            double *optx) {
    if (!node_ok(t, n1, n2) || !optx) return fail(PLK_ERR_ARG, "plk_tree_optimize_one_branch: bad arguments");
    return t->optimize_one(n1, n2, min_bl, max_bl, max_nr, *optx);
}

extern "C" int plk_tree_optimize_all_branches(plk_tree *t, int iterations, double tol, double min_bl, double max_bl,
                                              int max_nr, double *logl) {
    if (!t || !logl) return fail(PLK_ERR_ARG, "plk_tree_optimize_all_branches: bad arguments");
    return t->optimize_all(iterations, tol, min_bl, max_bl, max_nr, *logl);
}

extern "C" int plk_tree_full_traversal(plk_tree *t, double *logl) {
    if (!t || !logl) return fail(PLK_ERR_ARG, "plk_tree_full_traversal: bad arguments");
    t->clear_all();
    return t->loglik(0, t->nei[0], logl, nullptr);
}

extern "C" int plk_tree_counters(plk_tree *t, int64_t *out3) {
    if (!t || !out3) return fail(PLK_ERR_ARG, "plk_tree_counters: bad arguments");
    out3[0] = plk_updates(t->ctx);
    out3[1] = t->n_derivs;
    out3[2] = t->n_ops;
    return PLK_OK;
}

extern "C" int plk_tree_queue_branch(plk_tree *t, int dad, int node, int clear, int *ref_a, int *ref_b, double *len) {
    if (!node_ok(t, dad, node) || !ref_a || !ref_b || !len) return fail(PLK_ERR_ARG, "plk_tree_queue_branch: bad arguments");
    if (clear) t->clear_all();
    int rc = t->plan(dad, node);
    if (rc) return rc;
    *ref_a = t->ref_dir(node, dad);
    *ref_b = t->ref_dir(dad, node);
    *len = t->length(dad, node);
    return PLK_OK;
}

// Tile-shape autotuner: time full traversals for each FP32 matrix-core tile shape
// (waves per category) with HIP events and keep the fastest.  FP64 contexts have one
// shape and return it.
extern "C" int plk_tree_autotune(plk_tree *t, int reps, int *best_wpc, double *best_ms) {
    if (!t || reps < 1) return fail(PLK_ERR_ARG, "plk_tree_autotune: bad arguments");
    const int cands[4] = {0, 1, 2, 4};
    int best = 1;
    double best_t = 1e300;
    for (int wpc : cands) {
        if (plk_set_tile_shape(t->ctx, wpc) != PLK_OK) continue;
        double lnl = 0.0;
        int rc = plk_tree_full_traversal(t, &lnl);  // warm-up
        if (rc) return rc;
        plk_set_timing(t->ctx, 1);
        for (int r = 0; r < reps; r++) {
            rc = plk_tree_full_traversal(t, &lnl);
            if (rc) return rc;
        }
        double ms = 0.0;
        int64_t n = 0;
        plk_get_timing(t->ctx, 0, &ms, &n);
        plk_set_timing(t->ctx, 0);
        const double per = n ? ms / n : 1e300;
        if (per < best_t) {
            best_t = per;
            best = wpc;
        }
    }
    plk_set_tile_shape(t->ctx, best);
    if (best_wpc) *best_wpc = best;
    if (best_ms) *best_ms = best_t;
    return PLK_OK;
}

extern "C" int plk_tree_all_branch_derivs(plk_tree *t, int32_t *branch_a, int32_t *branch_b, double *df, double *ddf) {
    if (!t || !df || !ddf) return fail(PLK_ERR_ARG, "plk_tree_all_branch_derivs: bad arguments");
    std::vector<int> a, b;
    std::vector<double> d1, d2;
    int rc = t->all_branch_derivs(a, b, d1, d2);
    for (size_t i = 0; i < a.size(); i++) {
        if (branch_a) branch_a[i] = a[i];
        if (branch_b) branch_b[i] = b[i];
        df[i] = d1[i];
        ddf[i] = d2[i];
    }
    return rc;
}
