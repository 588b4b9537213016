/*
 * plk_oracle.c -- CPU FP64 restatement of IQ-TREE's likelihood kernels.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  See plk_oracle.h
 * for the list of reference functions restated here and what pins each one.
 */
#define _POSIX_C_SOURCE 200809L
#include "plk_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_V
#define ORC_V 4
#endif
#define V ORC_V

/* tree/phylotree.h:74-80 */
#define SCALING_THRESHOLD_EXP 256
static const double SCALING_THRESHOLD = 8.636168555094444625386e-78; /* 2^-256 */
static const double LOG_SCALING_THRESHOLD = -177.4456782233459932741;

typedef struct {
    int dad, node;      /* output = partial of node's subtree seen from dad */
    int left, right;    /* children of node (tree/phylokernelnew.h:1347-1354) */
    double *eleft, *eright;     /* ncat*n*n each */
    double *leaf_left, *leaf_right; /* (su+1)*block each, or NULL */
} orc_op;

struct orc_tree {
    int n, ncat, block, ntaxa, nnodes, nptn, nptn_pad, safe, nthreads, npackets, su;
    double *eval, *evec, *inv_evec, *rates, *props, *tip_lh;
    uint8_t *states;           /* ntaxa * nptn_pad */
    double *freq, *invar;      /* nptn_pad */
    int *nei;                  /* nnodes*3 */
    double *len;               /* nnodes*3 */
    double **part;             /* per direction d = dad*3+k: partial of nei[d]'s subtree */
    uint16_t **scale;
    uint8_t *computed;
    double *theta, *buf_scale;
    int theta_valid, theta_dad, theta_node;
    double *ptn_lh;
    size_t *limits;
    orc_op *ops;
    int nops;
    double *mats;              /* scratch for echildren/leaf tables of the current traversal */
    size_t mats_cap;
    long long updates;
    double gamma_shape;        /* 0: rates are the caller's own */
    int gamma_median;          /* category rates at the medians (cut_median) */
};

static void *xalloc(size_t bytes) {
    void *p = NULL;
    if (bytes == 0) bytes = 64;
    if (posix_memalign(&p, 64, (bytes + 63) / 64 * 64) != 0) abort();
    memset(p, 0, bytes);
    return p;
}

int orc_lanes(void) { return V; }
long long orc_updates(const orc_tree *t) { return t->updates; }

/* utils/tools.h roundUpToMultiple + computeBounds, tree/phylokernelnew.h:1116-1147 */
static void compute_bounds(int threads, int packets, size_t elements, size_t *limits) {
    size_t block_start = 0;
    int k = 0;
    elements = (elements + V - 1) / V * V;
    for (int wave = packets / threads; wave >= 1; --wave) {
        size_t this_wave = elements - block_start;
        if (wave > 1) this_wave = this_wave * 3 / 4;
        this_wave = (this_wave + V - 1) / V * V;
        size_t stop = block_start + this_wave;
        for (int to_go = threads; to_go >= 1; --to_go) {
            limits[k++] = block_start;
            size_t bs = (stop - block_start) / to_go;
            bs = (bs + V - 1) / V * V;
            block_start += bs;
        }
    }
    limits[k] = elements;
    /* the reference aborts when a wave overruns; clamp so slices stay in range */
    for (int i = 0; i <= k; i++)
        if (limits[i] > elements) limits[i] = elements;
}

orc_tree *orc_create(int nstates, int ncat, int ntaxa, int nptn, int safe, int nthreads) {
    orc_tree *t = (orc_tree *)xalloc(sizeof(orc_tree));
    t->n = nstates;
    t->ncat = ncat;
    t->block = nstates * ncat;
    t->ntaxa = ntaxa;
    t->nnodes = 2 * ntaxa - 2;
    t->nptn = nptn;
    t->nptn_pad = (nptn + V - 1) / V * V;
    t->safe = safe;
    t->nthreads = nthreads < 1 ? 1 : nthreads;
    /* tree/phylotreesse.cpp:43-52 */
    t->npackets = (t->nthreads == 1) ? 1 : 2 * t->nthreads;
    t->su = -1;
    int n = nstates;
    t->eval = (double *)xalloc(sizeof(double) * n);
    t->evec = (double *)xalloc(sizeof(double) * n * n);
    t->inv_evec = (double *)xalloc(sizeof(double) * n * n);
    t->rates = (double *)xalloc(sizeof(double) * ncat);
    t->props = (double *)xalloc(sizeof(double) * ncat);
    t->states = (uint8_t *)xalloc((size_t)ntaxa * t->nptn_pad);
    t->freq = (double *)xalloc(sizeof(double) * t->nptn_pad);
    t->invar = (double *)xalloc(sizeof(double) * t->nptn_pad);
    t->nei = (int *)xalloc(sizeof(int) * 3 * t->nnodes);
    t->len = (double *)xalloc(sizeof(double) * 3 * t->nnodes);
    t->part = (double **)xalloc(sizeof(double *) * 3 * t->nnodes);
    t->scale = (uint16_t **)xalloc(sizeof(uint16_t *) * 3 * t->nnodes);
    t->computed = (uint8_t *)xalloc(3 * t->nnodes);
    t->theta = (double *)xalloc(sizeof(double) * (size_t)t->nptn_pad * t->block);
    t->buf_scale = (double *)xalloc(sizeof(double) * t->nptn_pad);
    t->ptn_lh = (double *)xalloc(sizeof(double) * t->nptn_pad);
    t->limits = (size_t *)xalloc(sizeof(size_t) * (t->npackets + 2));
    t->ops = (orc_op *)xalloc(sizeof(orc_op) * t->nnodes);
    compute_bounds(t->nthreads, t->npackets, t->nptn_pad, t->limits);
    return t;
}

void orc_destroy(orc_tree *t) {
    if (!t) return;
    for (int d = 0; d < 3 * t->nnodes; d++) {
        free(t->part[d]);
        free(t->scale[d]);
    }
    free(t->eval); free(t->evec); free(t->inv_evec); free(t->rates); free(t->props);
    free(t->tip_lh); free(t->states); free(t->freq); free(t->invar); free(t->nei);
    free(t->len); free(t->part); free(t->scale); free(t->computed); free(t->theta);
    free(t->buf_scale); free(t->ptn_lh); free(t->limits); free(t->ops); free(t->mats);
    free(t);
}

void orc_set_model(orc_tree *t, const double *eval, const double *evec, const double *inv_evec,
                   const double *rates, const double *props) {
    int n = t->n;
    memcpy(t->eval, eval, sizeof(double) * n);
    memcpy(t->evec, evec, sizeof(double) * n * n);
    memcpy(t->inv_evec, inv_evec, sizeof(double) * n * n);
    memcpy(t->rates, rates, sizeof(double) * t->ncat);
    memcpy(t->props, props, sizeof(double) * t->ncat);
    t->gamma_shape = 0.0;
    orc_clear_all(t);
}

/* a new eigen system (decomposeRateMatrix of a new Q); rates, proportions and the Gamma
 * shape stay */
void orc_set_eigen(orc_tree *t, const double *eval, const double *evec, const double *inv_evec) {
    int n = t->n;
    memcpy(t->eval, eval, sizeof(double) * n);
    memcpy(t->evec, evec, sizeof(double) * n * n);
    memcpy(t->inv_evec, inv_evec, sizeof(double) * n * n);
    orc_clear_all(t);
}

void orc_set_tip_lh(orc_tree *t, int state_unknown, const double *tip_lh) {
    free(t->tip_lh);
    t->su = state_unknown;
    t->tip_lh = (double *)xalloc(sizeof(double) * (state_unknown + 1) * t->n);
    memcpy(t->tip_lh, tip_lh, sizeof(double) * (state_unknown + 1) * t->n);
    orc_clear_all(t);
}

void orc_set_patterns(orc_tree *t, const uint8_t *states, const double *freq, const double *invar) {
    for (int s = 0; s < t->ntaxa; s++) {
        uint8_t *row = t->states + (size_t)s * t->nptn_pad;
        memcpy(row, states + (size_t)s * t->nptn, t->nptn);
        /* padded patterns are STATE_UNKNOWN (tree/phylokernelnew.h:1706-1708) */
        for (int p = t->nptn; p < t->nptn_pad; p++) row[p] = (uint8_t)t->su;
    }
    memset(t->freq, 0, sizeof(double) * t->nptn_pad);
    memset(t->invar, 0, sizeof(double) * t->nptn_pad);
    memcpy(t->freq, freq, sizeof(double) * t->nptn);
    if (invar) memcpy(t->invar, invar, sizeof(double) * t->nptn);
    orc_clear_all(t);
}

void orc_set_tree(orc_tree *t, const int *nei, const double *len) {
    memcpy(t->nei, nei, sizeof(int) * 3 * t->nnodes);
    memcpy(t->len, len, sizeof(double) * 3 * t->nnodes);
    orc_clear_all(t);
}

void orc_clear_all(orc_tree *t) {
    memset(t->computed, 0, 3 * t->nnodes);
    t->theta_valid = 0;
}

static int dir_of(const orc_tree *t, int dad, int node) {
    for (int k = 0; k < 3; k++)
        if (t->nei[dad * 3 + k] == node) return dad * 3 + k;
    abort();
}

static int is_leaf(const orc_tree *t, int node) { return node < t->ntaxa; }

/* PhyloNode::clearReversePartialLh, tree/phylonode.cpp:22-33 */
static void clear_reverse(orc_tree *t, int node, int dad) {
    for (int k = 0; k < 3; k++) {
        int x = t->nei[node * 3 + k];
        if (x < 0 || x == dad) continue;
        t->computed[dir_of(t, x, node)] = 0;
        clear_reverse(t, x, node);
    }
}

void orc_set_length(orc_tree *t, int a, int b, double len) {
    t->len[dir_of(t, a, b)] = len;
    t->len[dir_of(t, b, a)] = len;
    clear_reverse(t, a, b);
    clear_reverse(t, b, a);
}

double orc_get_length(const orc_tree *t, int a, int b) { return t->len[dir_of(t, a, b)]; }

static int subtree_size(const orc_tree *t, int node, int dad) {
    if (is_leaf(t, node)) return 1;
    int s = 0;
    for (int k = 0; k < 3; k++) {
        int x = t->nei[node * 3 + k];
        if (x >= 0 && x != dad) s += subtree_size(t, x, node);
    }
    return s;
}

/* PhyloTree::computeTraversalInfo(PhyloNeighbor*, ...), tree/phylotree.cpp:6031-6139:
 * post-order over dirty directions, children visited in decreasing subtree size. */
static void collect(orc_tree *t, int dad, int node) {
    int d = dir_of(t, dad, node);
    if (is_leaf(t, node) || t->computed[d]) return;
    int kids[2], sizes[2], nk = 0;
    for (int k = 0; k < 3; k++) {
        int x = t->nei[node * 3 + k];
        if (x >= 0 && x != dad) {
            kids[nk] = x;
            sizes[nk] = subtree_size(t, x, node);
            nk++;
        }
    }
    if (nk != 2) abort(); /* binary trees only; multifurcation is out of scope */
    if (sizes[0] >= sizes[1]) {
        collect(t, node, kids[0]);
        collect(t, node, kids[1]);
    } else {
        collect(t, node, kids[1]);
        collect(t, node, kids[0]);
    }
    t->computed[d] = 1;
    orc_op *op = &t->ops[t->nops++];
    op->dad = dad;
    op->node = node;
    op->left = kids[0];   /* FOR_NEIGHBOR_IT order of node->neighbors */
    op->right = kids[1];
    if (!t->part[d]) {
        t->part[d] = (double *)xalloc(sizeof(double) * (size_t)t->nptn_pad * t->block);
        t->scale[d] = (uint16_t *)xalloc(sizeof(uint16_t) * (size_t)t->nptn_pad * t->ncat);
    }
}

/* computePartialInfo, reversible vectorised branch (tree/phylokernelnew.h:1013-1063) */
static void partial_info(orc_tree *t, orc_op *op, double *buf) {
    int n = t->n, ncat = t->ncat, block = t->block, su = t->su;
    int child[2] = {op->left, op->right};
    double *e[2], *leaf[2] = {NULL, NULL};
    for (int ci = 0; ci < 2; ci++) {
        double len = t->len[dir_of(t, op->node, child[ci])];
        e[ci] = buf;
        buf += (size_t)ncat * n * n;
        for (int c = 0; c < ncat; c++) {
            double lc = t->rates[c] * len;
            for (int x = 0; x < n; x++)
                for (int i = 0; i < n; i++)
                    e[ci][(c * n + x) * n + i] = t->evec[x * n + i] * exp(t->eval[i] * lc);
        }
        if (is_leaf(t, child[ci])) {
            leaf[ci] = buf;
            buf += (size_t)(su + 1) * block;
            for (int s = 0; s <= su; s++)
                for (int c = 0; c < ncat; c++)
                    for (int x = 0; x < n; x++) {
                        double v = 0.0;
                        for (int i = 0; i < n; i++)
                            v = fma(e[ci][(c * n + x) * n + i], t->tip_lh[s * n + i], v);
                        leaf[ci][s * block + c * n + x] = v;
                    }
            /* unknown state row is exactly 1 (tree/phylokernelnew.h:1056-1059) */
            for (int x = 0; x < block; x++) leaf[ci][su * block + x] = 1.0;
        }
    }
    op->eleft = e[0];
    op->eright = e[1];
    op->leaf_left = leaf[0];
    op->leaf_right = leaf[1];
    /* a leaf child is always "left" (tree/phylokernelnew.h:1399-1409) */
    if (!is_leaf(t, op->left) && is_leaf(t, op->right)) {
        int ti = op->left; op->left = op->right; op->right = ti;
        double *tp = op->eleft; op->eleft = op->eright; op->eright = tp;
        tp = op->leaf_left; op->leaf_left = op->leaf_right; op->leaf_right = tp;
    }
}

static void prepare_ops(orc_tree *t) {
    int n = t->n, ncat = t->ncat;
    size_t per_op = 2 * ((size_t)ncat * n * n + (size_t)(t->su + 1) * t->block);
    size_t need = per_op * (t->nops > 0 ? t->nops : 1);
    if (need > t->mats_cap) {
        free(t->mats);
        t->mats = (double *)xalloc(sizeof(double) * need);
        t->mats_cap = need;
    }
    for (int i = 0; i < t->nops; i++) partial_info(t, &t->ops[i], t->mats + per_op * i);
    t->updates += t->nops;
}

/* Felsenstein pruning in eigen space for one op over patterns [lo, hi), step V.
 * tip-tip 1636-1750, tip-internal 1752-1909, internal-internal 1911-2036. */
static void partial_slice(orc_tree *t, const orc_op *op, size_t lo, size_t hi) {
    const int n = t->n, ncat = t->ncat, block = t->block, safe = t->safe;
    const double *uinv = t->inv_evec;
    int dout = dir_of(t, op->dad, op->node);
    double *out_all = t->part[dout];
    uint16_t *sout = t->scale[dout];
    int lleaf = is_leaf(t, op->left), rleaf = is_leaf(t, op->right);
    const double *pl = NULL, *pr = NULL;
    const uint16_t *sl = NULL, *sr = NULL;
    if (!lleaf) {
        int d = dir_of(t, op->node, op->left);
        pl = t->part[d];
        sl = t->scale[d];
    }
    if (!rleaf) {
        int d = dir_of(t, op->node, op->right);
        pr = t->part[d];
        sr = t->scale[d];
    }
    const uint8_t *stl = lleaf ? t->states + (size_t)op->left * t->nptn_pad : NULL;
    const uint8_t *str = rleaf ? t->states + (size_t)op->right * t->nptn_pad : NULL;
    double tmp[64 * V]; /* n <= 64 */

    for (size_t ptn = lo; ptn < hi; ptn += V) {
        double *out = out_all + ptn * block;
        double lh_max[V];
        for (int v = 0; v < V; v++) lh_max[v] = 0.0;
        /* scale bookkeeping of the children */
        if (!safe) {
            for (int v = 0; v < V; v++) {
                unsigned s = 0;
                if (lleaf && rleaf) s = 0;                       /* 1643-1644 */
                else if (lleaf) s = sr[ptn + v];                 /* 1757-1760 */
                else s = (unsigned)sl[ptn + v] + sr[ptn + v];    /* 1929-1936 */
                sout[ptn + v] = (uint16_t)s;
            }
        } else {
            for (int v = 0; v < V; v++)
                for (int c = 0; c < ncat; c++) {
                    size_t a = (ptn + v) * ncat + c;
                    unsigned s = 0;
                    if (lleaf && rleaf) s = 0;
                    else if (lleaf) s = sr[a];
                    else s = (unsigned)sl[a] + sr[a];
                    sout[a] = (uint16_t)s;
                }
        }
        for (int c = 0; c < ncat; c++) {
#ifdef ORC_PLANT_CATEGORY_MAX
            /* planted mistake, built only by tests/test_oracle.py: NORM keeps a
             * per-category maximum, so its check sees the last category alone */
            if (1)
#else
            if (safe)
#endif
                for (int v = 0; v < V; v++) lh_max[v] = 0.0;
            const double *el = op->eleft + (size_t)c * n * n;
            const double *er = op->eright + (size_t)c * n * n;
            for (int x = 0; x < n; x++) {
                double a[V], b[V];
                if (lleaf) {
                    for (int v = 0; v < V; v++)
                        a[v] = op->leaf_left[stl[ptn + v] * block + c * n + x];
                } else {
                    const double *L = pl + ptn * block + (size_t)c * n * V;
                    for (int v = 0; v < V; v++) a[v] = 0.0;
                    for (int i = 0; i < n; i++)
                        for (int v = 0; v < V; v++) a[v] = fma(el[x * n + i], L[i * V + v], a[v]);
                }
                if (rleaf) {
                    for (int v = 0; v < V; v++)
                        b[v] = op->leaf_right[str[ptn + v] * block + c * n + x];
                } else {
                    const double *R = pr + ptn * block + (size_t)c * n * V;
                    for (int v = 0; v < V; v++) b[v] = 0.0;
                    for (int i = 0; i < n; i++)
                        for (int v = 0; v < V; v++) b[v] = fma(er[x * n + i], R[i * V + v], b[v]);
                }
                for (int v = 0; v < V; v++) tmp[x * V + v] = a[v] * b[v];
            }
            /* productVecMat with inv_evec (tree/phylokernelnew.h:449-567) */
            double *o = out + (size_t)c * n * V;
            for (int j = 0; j < n; j++) {
                double acc[V];
                for (int v = 0; v < V; v++) acc[v] = 0.0;
                for (int x = 0; x < n; x++)
                    for (int v = 0; v < V; v++) acc[v] = fma(tmp[x * V + v], uinv[j * n + x], acc[v]);
                for (int v = 0; v < V; v++) {
                    o[j * V + v] = acc[v];
                    double aa = fabs(acc[v]);
                    if (aa > lh_max[v]) lh_max[v] = aa;
                }
            }
            if (safe && !(lleaf && rleaf)) {
                /* per-category scaling (1997-2013, 1871-1886) */
                for (int v = 0; v < V; v++) {
                    if (lh_max[v] < SCALING_THRESHOLD && t->invar[ptn + v] == 0.0) {
                        for (int j = 0; j < n; j++) o[j * V + v] = ldexp(o[j * V + v], SCALING_THRESHOLD_EXP);
                        sout[(ptn + v) * ncat + c] += 1;
                    }
                }
            }
        }
        if (!safe && !(lleaf && rleaf)) {
            /* one check over all categories (2019-2034, 1893-1907) */
            for (int v = 0; v < V; v++) {
                if (lh_max[v] < SCALING_THRESHOLD && t->invar[ptn + v] == 0.0) {
                    for (int i = 0; i < block; i++) out[i * V + v] = ldexp(out[i * V + v], SCALING_THRESHOLD_EXP);
                    sout[ptn + v] += 1;
                }
            }
        }
    }
}

static void run_ops_slice(orc_tree *t, size_t lo, size_t hi) {
    for (int i = 0; i < t->nops; i++) partial_slice(t, &t->ops[i], lo, hi);
}

/* computeTraversalInfo(node, dad, false): both directions of the branch. */
static void plan_branch(orc_tree *t, int dad, int node) {
    t->nops = 0;
    collect(t, node, dad); /* dad's subtree seen from node (skipped when dad is a leaf) */
    collect(t, dad, node); /* node's subtree seen from dad */
    prepare_ops(t);
}

/* resolve the per-pattern scale of a branch: NORM sum or SAFE min rule (2918-2951, 3060-3094).
 * cat_factor[c] gets 1, 2^-256 or 0 per category. */
static double branch_scale(const orc_tree *t, const uint16_t *sa, const uint16_t *sb, size_t p,
                           double *cat_factor) {
    int ncat = t->ncat;
    if (!t->safe) {
        unsigned s = sa[p] + (sb ? sb[p] : 0);
        for (int c = 0; c < ncat; c++) cat_factor[c] = 1.0;
        return (double)s;
    }
    unsigned sum[64], mn = 0xffffffffu;
    for (int c = 0; c < ncat; c++) {
        sum[c] = sa[p * ncat + c] + (sb ? sb[p * ncat + c] : 0);
        if (sum[c] < mn) mn = sum[c];
    }
    for (int c = 0; c < ncat; c++) {
        if (sum[c] == mn) cat_factor[c] = 1.0;
        else if (sum[c] == mn + 1) cat_factor[c] = SCALING_THRESHOLD;
        else cat_factor[c] = 0.0;
    }
    return (double)mn;
}

/* computeLikelihoodBranchSIMD (tree/phylokernelnew.h:2658-3184) */
double orc_branch_loglik(orc_tree *t, int dad, int node, double *ptn_lh_out) {
    if (is_leaf(t, node)) { int tmp = dad; dad = node; node = tmp; }
    plan_branch(t, dad, node);
    const int n = t->n, ncat = t->ncat, block = t->block;
    double tlen = t->len[dir_of(t, dad, node)];
    double *val = (double *)xalloc(sizeof(double) * block);
    for (int c = 0; c < ncat; c++)
        for (int i = 0; i < n; i++)
            val[c * n + i] = exp(t->eval[i] * t->rates[c] * tlen) * t->props[c];
    int dleaf = is_leaf(t, dad);
    double *node_tab = NULL;
    if (dleaf) {
        node_tab = (double *)xalloc(sizeof(double) * (t->su + 1) * block);
        for (int s = 0; s <= t->su; s++)
            for (int c = 0; c < ncat; c++)
                for (int i = 0; i < n; i++)
                    node_tab[s * block + c * n + i] = val[c * n + i] * t->tip_lh[s * n + i];
    }
    int ddn = dir_of(t, dad, node);
    double *all_lh = (double *)xalloc(sizeof(double) * t->npackets);
#pragma omp parallel for schedule(dynamic, 1) num_threads(t->nthreads)
    for (int pk = 0; pk < t->npackets; pk++) {
        size_t lo = t->limits[pk], hi = t->limits[pk + 1];
        run_ops_slice(t, lo, hi);
        const double *pd = t->part[ddn];
        const uint16_t *sd = t->scale[ddn];
        const double *pn = NULL;
        const uint16_t *sn = NULL;
        if (!dleaf) {
            int dnd = dir_of(t, node, dad);
            pn = t->part[dnd];
            sn = t->scale[dnd];
        }
        double vc_tree_lh[V];
        for (int v = 0; v < V; v++) vc_tree_lh[v] = 0.0;
        for (size_t ptn = lo; ptn < hi; ptn += V) {
            for (int v = 0; v < V; v++) {
                size_t p = ptn + v;
                double lh_cat[64];
                for (int c = 0; c < ncat; c++) {
                    double acc = 0.0;
                    const double *D = pd + ptn * block + (size_t)c * n * V + v;
                    if (dleaf) {
                        const double *tab = node_tab + t->states[(size_t)dad * t->nptn_pad + p] * block + c * n;
                        for (int i = 0; i < n; i++) acc = fma(tab[i], D[i * V], acc);
                    } else {
                        const double *N = pn + ptn * block + (size_t)c * n * V + v;
                        for (int i = 0; i < n; i++) acc = fma(val[c * n + i] * N[i * V], D[i * V], acc);
                    }
                    lh_cat[c] = acc;
                }
                double fac[64];
                double sc = branch_scale(t, sd, sn, p, fac);
                double lh = 0.0;
                for (int c = 0; c < ncat; c++) lh += lh_cat[c] * fac[c];
                lh = fabs(lh) + t->invar[p];
                double l = log(lh) + sc * LOG_SCALING_THRESHOLD;
                if (p < (size_t)t->nptn) {
                    t->ptn_lh[p] = l;
                    vc_tree_lh[v] = fma(l, t->freq[p], vc_tree_lh[v]);
                }
            }
        }
        double s = 0.0;
        for (int v = 0; v < V; v++) s += vc_tree_lh[v];
        all_lh[pk] = s;
    }
    /* fixed packet order (3146-3148) */
    double tree_lh = 0.0;
    for (int k = 0; k < t->npackets; k++) tree_lh += all_lh[k];
    if (!isfinite(tree_lh)) {
        /* 3167-3184 */
        tree_lh = 0.0;
        for (int p = 0; p < t->nptn; p++) {
            if (!isfinite(t->ptn_lh[p])) t->ptn_lh[p] = LOG_SCALING_THRESHOLD * 4;
            tree_lh += t->ptn_lh[p] * t->freq[p];
        }
    }
    if (ptn_lh_out) memcpy(ptn_lh_out, t->ptn_lh, sizeof(double) * t->nptn);
    free(val);
    free(node_tab);
    free(all_lh);
    return tree_lh;
}

/* computeLikelihoodBufferSIMD over one slice (2052-2235): theta and buffer_scale_all */
static void theta_slice(orc_tree *t, int dad, int node, size_t lo, size_t hi) {
    const int n = t->n, ncat = t->ncat, block = t->block;
    int ddn = dir_of(t, dad, node);
    const double *pd = t->part[ddn];
    const uint16_t *sd = t->scale[ddn];
    int dleaf = is_leaf(t, dad);
    const double *pn = NULL;
    const uint16_t *sn = NULL;
    if (!dleaf) {
        int dnd = dir_of(t, node, dad);
        pn = t->part[dnd];
        sn = t->scale[dnd];
    }
    for (size_t ptn = lo; ptn < hi; ptn += V) {
        double *th = t->theta + ptn * block;
        for (int v = 0; v < V; v++) {
            size_t p = ptn + v;
            const double *tip = dleaf ? t->tip_lh + t->states[(size_t)dad * t->nptn_pad + p] * n : NULL;
            for (int c = 0; c < ncat; c++)
                for (int i = 0; i < n; i++) {
                    size_t o = ((size_t)c * n + i) * V + v;
                    double a = dleaf ? tip[i] : pn[ptn * block + o];
                    th[o] = a * pd[ptn * block + o];
                }
            double fac[64];
            double sc = branch_scale(t, sd, sn, p, fac);
            if (t->safe)
                for (int c = 0; c < ncat; c++)
                    if (fac[c] != 1.0)
                        for (int i = 0; i < n; i++) {
                            size_t o = ((size_t)c * n + i) * V + v;
                            th[o] = (fac[c] == 0.0) ? 0.0 : th[o] * SCALING_THRESHOLD;
                        }
            t->buf_scale[p] = sc * LOG_SCALING_THRESHOLD;
        }
    }
}

/* computeLikelihoodDervSIMD (2237-2650) */
int orc_branch_derivs(orc_tree *t, int dad, int node, int reuse_theta, double *df, double *ddf) {
    if (is_leaf(t, node)) { int tmp = dad; dad = node; node = tmp; }
    const int n = t->n, ncat = t->ncat, block = t->block;
    int recompute = !(reuse_theta && t->theta_valid && t->theta_dad == dad && t->theta_node == node);
    if (recompute) plan_branch(t, dad, node);
    double tlen = t->len[dir_of(t, dad, node)];
    double *v0 = (double *)xalloc(sizeof(double) * block);
    double *v1 = (double *)xalloc(sizeof(double) * block);
    double *v2 = (double *)xalloc(sizeof(double) * block);
    for (int c = 0; c < ncat; c++)
        for (int i = 0; i < n; i++) {
            double cof = t->eval[i] * t->rates[c];
            double val = exp(cof * tlen) * t->props[c];
            double val1 = cof * val;
            v0[c * n + i] = val;
            v1[c * n + i] = val1;
            v2[c * n + i] = cof * val1;
        }
    double *pdf = (double *)xalloc(sizeof(double) * t->npackets);
    double *pddf = (double *)xalloc(sizeof(double) * t->npackets);
#pragma omp parallel for schedule(dynamic, 1) num_threads(t->nthreads)
    for (int pk = 0; pk < t->npackets; pk++) {
        size_t lo = t->limits[pk], hi = t->limits[pk + 1];
        if (recompute) {
            run_ops_slice(t, lo, hi);
            theta_slice(t, dad, node, lo, hi);
        }
        double my_df[V], my_ddf[V];
        for (int v = 0; v < V; v++) my_df[v] = my_ddf[v] = 0.0;
        for (size_t ptn = lo; ptn < hi; ptn += V) {
            const double *th = t->theta + ptn * block;
            double lh[V], d1[V], d2[V];
            for (int v = 0; v < V; v++) lh[v] = d1[v] = d2[v] = 0.0;
            for (int i = 0; i < block; i++)
                for (int v = 0; v < V; v++) {
                    double x = th[i * V + v];
                    lh[v] = fma(v0[i], x, lh[v]);
                    d1[v] = fma(v1[i], x, d1[v]);
                    d2[v] = fma(v2[i], x, d2[v]);
                }
            for (int v = 0; v < V; v++) {
                size_t p = ptn + v;
                if (p >= (size_t)t->nptn) continue;
                double l = fabs(lh[v]) + t->invar[p];
                double inv = 1.0 / l;
                double df_frac = d1[v] * inv, ddf_frac = d2[v] * inv;
                double f = t->freq[p];
                double tmp1 = df_frac * f, tmp2 = ddf_frac * f;
                my_df[v] += tmp1;
                my_ddf[v] += tmp2 - tmp1 * df_frac;
            }
        }
        double a = 0.0, b = 0.0;
        for (int v = 0; v < V; v++) { a += my_df[v]; b += my_ddf[v]; }
        pdf[pk] = a;
        pddf[pk] = b;
    }
    double sdf = 0.0, sddf = 0.0;
    for (int k = 0; k < t->npackets; k++) { sdf += pdf[k]; sddf += pddf[k]; }
    t->theta_valid = 1;
    t->theta_dad = dad;
    t->theta_node = node;
    free(v0); free(v1); free(v2); free(pdf); free(pddf);
    int status = 0;
    if (!isfinite(sdf)) {
        if (!t->safe) status = 1;          /* outError in NORM mode (2594-2596) */
        sdf = sddf = 0.0;                  /* SAFE: warn and zero (2646-2649) */
    }
    *df = sdf;
    *ddf = sddf;
    return status;
}

/* computeLikelihoodFromBufferSIMD (3284-3450) */
double orc_loglik_from_theta(orc_tree *t, int dad, int node) {
    if (is_leaf(t, node)) { int tmp = dad; dad = node; node = tmp; }
    const int n = t->n, ncat = t->ncat, block = t->block;
    double tlen = t->len[dir_of(t, dad, node)];
    double *v0 = (double *)xalloc(sizeof(double) * block);
    for (int c = 0; c < ncat; c++)
        for (int i = 0; i < n; i++)
            v0[c * n + i] = exp(t->eval[i] * t->rates[c] * tlen) * t->props[c];
    size_t ngroups = t->nptn_pad / V;
    double *all_lh = (double *)xalloc(sizeof(double) * (ngroups + 1));
#pragma omp parallel for num_threads(t->nthreads)
    for (size_t g = 0; g < ngroups; g++) {
        size_t ptn = g * V;
        const double *th = t->theta + ptn * block;
        double lh[V], acc = 0.0;
        for (int v = 0; v < V; v++) lh[v] = 0.0;
        for (int i = 0; i < block; i++)
            for (int v = 0; v < V; v++) lh[v] = fma(v0[i], th[i * V + v], lh[v]);
        for (int v = 0; v < V; v++) {
            size_t p = ptn + v;
            double l = fabs(lh[v]) + t->invar[p];
            if (p < (size_t)t->nptn) {
                double ll = log(fabs(l)) + t->buf_scale[p];
                t->ptn_lh[p] = ll;
                acc = fma(ll, t->freq[p], acc);
            }
        }
        all_lh[g] = acc;
    }
    double tree_lh = 0.0;
    for (size_t g = 0; g < ngroups; g++) tree_lh += all_lh[g];
    if (!isfinite(tree_lh)) {
        tree_lh = 0.0;
        for (int p = 0; p < t->nptn; p++) {
            if (!isfinite(t->ptn_lh[p])) t->ptn_lh[p] = LOG_SCALING_THRESHOLD * 4;
            tree_lh += t->ptn_lh[p] * t->freq[p];
        }
    }
    free(v0);
    free(all_lh);
    return tree_lh;
}

int orc_get_partial(orc_tree *t, int dad, int node, double *out, uint16_t *scale) {
    if (is_leaf(t, node)) return -1;
    t->nops = 0;
    collect(t, dad, node);
    prepare_ops(t);
#pragma omp parallel for schedule(dynamic, 1) num_threads(t->nthreads)
    for (int pk = 0; pk < t->npackets; pk++) run_ops_slice(t, t->limits[pk], t->limits[pk + 1]);
    int d = dir_of(t, dad, node);
    const int block = t->block, ncat = t->ncat;
    for (int p = 0; p < t->nptn; p++)
        for (int i = 0; i < block; i++)
            out[(size_t)p * block + i] = t->part[d][((size_t)(p / V) * block + i) * V + p % V];
    if (scale) {
        if (t->safe)
            memcpy(scale, t->scale[d], sizeof(uint16_t) * (size_t)t->nptn * ncat);
        else
            memcpy(scale, t->scale[d], sizeof(uint16_t) * (size_t)t->nptn);
    }
    return 0;
}

/* Optimization::minimizeNewton (utils/optimization.cpp:422-499) on f(t) = -logL along
 * branch (dad, node); computeFuncDerv negates the derivatives (tree/phylotree.cpp:2555-2562). */
static void func_derv(orc_tree *t, int a, int b, double value, int *first, double *f, double *df) {
    double d1, d2;
    t->len[dir_of(t, a, b)] = value;
    t->len[dir_of(t, b, a)] = value;
    orc_branch_derivs(t, a, b, *first ? 0 : 1, &d1, &d2);
    *first = 0;
    *f = -d1;
    *df = -d2;
}

static double minimize_newton(orc_tree *t, int a, int b, double x1, double xguess, double x2,
                              double xacc, int max_nr_step) {
    double df, dx, f, temp, xh, xl, rts, rts_old;
    int first = 1;
    rts = xguess;
    if (rts < x1) rts = x1;
    if (rts > x2) rts = x2;
    func_derv(t, a, b, rts, &first, &f, &df);
    if (!isfinite(f) || !isfinite(df)) return NAN;
    if (df >= 0.0 && fabs(f) < xacc) return rts;
    if (f < 0.0) { xl = rts; xh = x2; }
    else { xh = rts; xl = x1; }
    dx = fabs(xh - xl);
    for (int j = 1; j <= max_nr_step; j++) {
        rts_old = rts;
        if ((df <= 0.0) || (((rts - xh) * df - f) * ((rts - xl) * df - f) >= 0.0)) {
            dx = 0.5 * (xh - xl);
            rts = xl + dx;
            if (xl == rts) return rts;
        } else {
            dx = f / df;
            temp = rts;
            rts -= dx;
            if (temp == rts) return rts;
        }
        if (fabs(dx) < xacc || (j == max_nr_step)) return rts_old;
        func_derv(t, a, b, rts, &first, &f, &df);
        if (!isfinite(f) || !isfinite(df)) return NAN;
        if (df > 0.0 && fabs(f) < xacc) return rts;
        if (f < 0.0) xl = rts;
        else if (f > 0.0) xh = rts;
    }
    return 0.0;
}

/* PhyloTree::optimizeOneBranch (tree/phylotree.cpp:2620-2670), Newton branch */
double orc_optimize_one_branch(orc_tree *t, int node1, int node2, double min_bl, double max_bl,
                               int max_nr_step) {
    double current_len = orc_get_length(t, node1, node2);
    t->theta_valid = 0;
    double optx = minimize_newton(t, node1, node2, min_bl, current_len, max_bl, min_bl, max_nr_step);
    if (optx > max_bl * 0.95) {
        double opt_lh = orc_loglik_from_theta(t, node1, node2);
        t->len[dir_of(t, node1, node2)] = current_len;
        t->len[dir_of(t, node2, node1)] = current_len;
        double orig_lh = orc_loglik_from_theta(t, node1, node2);
        if (orig_lh > opt_lh) optx = current_len;
    }
    t->len[dir_of(t, node1, node2)] = optx;
    t->len[dir_of(t, node2, node1)] = optx;
    if (current_len != optx) {
        int keep = t->theta_valid;
        clear_reverse(t, node1, node2);
        clear_reverse(t, node2, node1);
        t->theta_valid = keep;
    }
    return optx;
}

/* MTree::findFarthestLeaf (tree/mtree.cpp:3040-3060); heights in edges */
static int farthest_leaf(const orc_tree *t, int node, int dad, int *height) {
    if (dad >= 0 && is_leaf(t, node)) { height[node] = 0; return node; }
    int res = -1;
    height[node] = 0;
    for (int k = 0; k < 3; k++) {
        int x = t->nei[node * 3 + k];
        if (x < 0 || x == dad) continue;
        int leaf = farthest_leaf(t, x, node, height);
        if (height[node] < height[x] + 1) {
            height[node] = height[x] + 1;
            res = leaf;
        }
    }
    return res;
}

/* MTree::getPreOrderBranches (tree/mtree.cpp:3062-3082) */
static void preorder_branches(const orc_tree *t, int node, int dad, const int *height, int *nodes,
                              int *nodes2, int *cnt) {
    if (dad >= 0) { nodes[*cnt] = node; nodes2[*cnt] = dad; (*cnt)++; }
    int nv[3], m = 0;
    for (int k = 0; k < 3; k++)
        if (t->nei[node * 3 + k] >= 0) nv[m++] = t->nei[node * 3 + k];
    for (int i = 0; i < m; i++)
        for (int j = i + 1; j < m; j++)
            if (height[nv[i]] > height[nv[j]]) { int x = nv[i]; nv[i] = nv[j]; nv[j] = x; }
    for (int i = 0; i < m; i++)
        if (nv[i] != dad) preorder_branches(t, nv[i], node, height, nodes, nodes2, cnt);
}

/* PhyloTree::optimizeAllBranches(iterations, tol, maxNRStep) (tree/phylotree.cpp:2724-2801);
 * the tree root is leaf 0. */
double orc_optimize_all_branches(orc_tree *t, int iterations, double tol, double min_bl,
                                 double max_bl, int max_nr_step) {
    int nn = t->nnodes;
    int *height = (int *)xalloc(sizeof(int) * nn);
    int *nodes = (int *)xalloc(sizeof(int) * nn);
    int *nodes2 = (int *)xalloc(sizeof(int) * nn);
    double *saved = (double *)xalloc(sizeof(double) * 3 * nn);
    int far = farthest_leaf(t, 0, -1, height);
    farthest_leaf(t, far, -1, height);
    int cnt = 0;
    preorder_branches(t, far, -1, height, nodes, nodes2, &cnt);
    double tree_lh = orc_branch_loglik(t, nodes[0], nodes2[0], NULL);
    for (int it = 0; it < iterations; it++) {
        memcpy(saved, t->len, sizeof(double) * 3 * nn);
        for (int j = 0; j < cnt; j++)
            orc_optimize_one_branch(t, nodes[j], nodes2[j], min_bl, max_bl, max_nr_step);
        double new_lh = orc_loglik_from_theta(t, nodes[cnt - 1], nodes2[cnt - 1]);
        if (new_lh < tree_lh - tol * 0.1) {
            memcpy(t->len, saved, sizeof(double) * 3 * nn);
            orc_clear_all(t);
            tree_lh = orc_branch_loglik(t, nodes[0], nodes2[0], NULL);
            break;
        }
        if (tree_lh <= new_lh && new_lh <= tree_lh + tol) { tree_lh = new_lh; break; }
        tree_lh = new_lh;
    }
    free(height); free(nodes); free(nodes2); free(saved);
    return tree_lh;
}

/* STATE_UNKNOWN (alignment/alignment.cpp:1116-1128) */
int orc_state_unknown(int nstates, int seq_type) {
    if (seq_type == ORC_SEQ_DNA) return 18;
    if (seq_type == ORC_SEQ_PROTEIN) return 23;
    return nstates;
}

/* tip_partial_lh[s][i] = sum_j Uinv[i][j] * tipLik_s[j]  (tree/phylotreesse.cpp:354-375,
 * model/modelsubst.cpp:183-208; DNA ambiguity model/modeldna.cpp:462-476; protein B/Z/J
 * model/modelprotein.cpp:1240-1256) */
void orc_tip_lh(int n, int seq_type, const double *inv_evec, double *out) {
    int su = orc_state_unknown(n, seq_type);
    double lik[64];
    for (int s = 0; s <= su; s++) {
        for (int j = 0; j < n; j++) lik[j] = 0.0;
        if (s < n) {
            lik[s] = 1.0;
        } else if (seq_type == ORC_SEQ_DNA && s < 18) {
            int cstate = s - n + 1;
            for (int j = 0; j < n; j++)
                if (cstate & (1 << j)) lik[j] = 1.0;
        } else if (seq_type == ORC_SEQ_PROTEIN && s < n + 3) {
            static const int ambi[6] = {2, 3, 5, 6, 9, 10};
            int cs = s - n;
            lik[ambi[cs * 2]] = 1.0;
            lik[ambi[cs * 2 + 1]] = 1.0;
        } else {
            for (int j = 0; j < n; j++) lik[j] = 1.0;
        }
        for (int i = 0; i < n; i++) {
            double v = 0.0;
            for (int j = 0; j < n; j++) v += inv_evec[i * n + j] * lik[j];
            out[s * n + i] = v;
        }
    }
}

/* RELL loop of IQTree::saveCurrentTree (tree/iqtree.cpp:3852-3931): pattern_lh is the
 * per-pattern logL rounded to float (3867-3870; a non-finite value first gets the branch
 * logL fallback, tree/phylokernelnew.h:3167-3184), rell[s] = sum_p pattern_lh[p] *
 * boot_samples[s][p] (dotProduct, 3912-3915), summed here sequentially in FP64.  With
 * uniform != NULL the replicate state is updated by the rule of 3917-3931, the caller's
 * uniform[s] standing in for random_double(rstream). */
void orc_rell_save_tree(int nsamples, int nptn, const float *counts, const double *ptn_lh, double cur_logl,
                        double eps, const double *uniform, double *rell, double *boot_logl, int32_t *boot_counts,
                        double *boot_orig, int32_t *improved) {
    float *lh = (float *)malloc(sizeof(float) * (size_t)(nptn > 0 ? nptn : 1));
    for (int p = 0; p < nptn; p++) {
        double l = ptn_lh[p];
        if (!isfinite(l)) l = LOG_SCALING_THRESHOLD * 4;
        lh[p] = (float)l;
    }
    for (int s = 0; s < nsamples; s++) {
        const float *w = counts + (size_t)s * nptn;
        double r = 0.0;
        for (int p = 0; p < nptn; p++) r += (double)lh[p] * (double)w[p];
        if (rell) rell[s] = r;
        if (!uniform) continue;
        int better = r > boot_logl[s] + eps;
        if (!better && r > boot_logl[s] - eps) better = uniform[s] <= 1.0 / (boot_counts[s] + 1);
        if (better) {
            if (r <= boot_logl[s] + eps) boot_counts[s]++;
            else boot_counts[s] = 1;
            boot_logl[s] = boot_logl[s] > r ? boot_logl[s] : r;
            boot_orig[s] = cur_logl;
        }
        if (improved) improved[s] = better;
    }
    free(lh);
}

/* One NNI move on internal branch (node1, node2) scored as getBestNNIForBran does with
 * NNI5 (tree/phylotree.cpp:4176-4429, params.nni5 default true, nni5_num_eval 1):
 * node1's first neighbour other than node2 (A) is swapped with node2's move-th other
 * neighbour (X), each keeping its branch length (updateNeighbor keeps the Neighbor
 * object); then optimizeOneBranch is run on node1's two other branches in neighbour
 * order, the central branch, and node2's two other branches, with max_nr Newton steps;
 * the score is the logL from the theta buffer of the last branch (4356-4357).  The tree
 * is restored afterwards.  lens[5] = new lengths (central, node1's two, node2's two, in
 * the order of newLen[0..4], 4323-4350). */
double orc_nni_eval(orc_tree *t, int node1, int node2, int move, double min_bl, double max_bl, int max_nr,
                    double *lens) {
    const int nn = t->nnodes;
    int *nei0 = (int *)malloc(sizeof(int) * 3 * nn);
    double *len0 = (double *)malloc(sizeof(double) * 3 * nn);
    memcpy(nei0, t->nei, sizeof(int) * 3 * nn);
    memcpy(len0, t->len, sizeof(double) * 3 * nn);
    int a_slot = -1, x_slot = -1, cnt = 0;
    for (int k = 0; k < 3; k++)
        if (t->nei[node1 * 3 + k] != node2 && a_slot < 0) a_slot = k;
    for (int k = 0; k < 3; k++)
        if (t->nei[node2 * 3 + k] != node1) {
            if (cnt == move) x_slot = k;
            cnt++;
        }
    const int A = t->nei[node1 * 3 + a_slot], X = t->nei[node2 * 3 + x_slot];
    const double lenA = t->len[node1 * 3 + a_slot], lenX = t->len[node2 * 3 + x_slot];
    const int dA = dir_of(t, A, node1), dX = dir_of(t, X, node2);
    t->nei[node1 * 3 + a_slot] = X;
    t->len[node1 * 3 + a_slot] = lenX;
    t->nei[dX] = node1;
    t->nei[node2 * 3 + x_slot] = A;
    t->len[node2 * 3 + x_slot] = lenA;
    t->nei[dA] = node2;
    orc_clear_all(t);
    int li = 1;
    for (int k = 0; k < 3; k++) {
        int y = t->nei[node1 * 3 + k];
        if (y == node2) continue;
        lens[li++] = orc_optimize_one_branch(t, node1, y, min_bl, max_bl, max_nr);
    }
    lens[0] = orc_optimize_one_branch(t, node1, node2, min_bl, max_bl, max_nr);
    int last = -1;
    for (int k = 0; k < 3; k++) {
        int z = t->nei[node2 * 3 + k];
        if (z == node1) continue;
        lens[li++] = orc_optimize_one_branch(t, node2, z, min_bl, max_bl, max_nr);
        last = z;
    }
    const double score = orc_loglik_from_theta(t, node2, last);
    memcpy(t->nei, nei0, sizeof(int) * 3 * nn);
    memcpy(t->len, len0, sizeof(double) * 3 * nn);
    orc_clear_all(t);
    free(nei0);
    free(len0);
    return score;
}

/* ---- model-parameter loop ---------------------------------------------------------- */

/* Discrete-Gamma mean rates (RateGamma::computeRatesMean, model/rategamma.cpp:98-171).
 * The reference takes the category boundaries from published algorithms; restated from
 * their papers: log Gamma by the Stirling series after shifting the argument to >= 7
 * (Pike & Hill 1966), the incomplete Gamma ratio by series or continued fraction
 * (Bhattacharjee 1970, AS 32), normal quantiles by a rational approximation (Odeh & Evans
 * 1974), chi-square quantiles by a starting approximation refined with a seventh-order
 * Taylor series (Best & Roberts 1975, AS 91).  Their stopping rules set the accuracy
 * (~1e-6 relative); the rates must carry the same rounding as the reference's for a
 * shape search to take the same steps. */
static double orc_lgamma(double x) {
    double shift = 0.0;
    if (x < 7.0) {
        /* Gamma(x) = Gamma(x + m) / (x (x+1) ... (x+m-1)), x + m the first value >= 7 */
        double prod = 1.0;
        for (; x < 7.0; x += 1.0) prod *= x;
        shift = -log(prod);
    }
    const double z = 1.0 / (x * x);
    const double series = (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z +
                           0.083333333333333) / x;
    return shift + (x - 0.5) * log(x) - x + 0.918938533204673 + series;
}

static double orc_incgamma(double x, double p, double lgp) {
    const double acc = 1e-8, big = 1e30;
    if (x == 0.0) return 0.0;
    if (x < 0.0 || p <= 0.0) return -1.0;
    const double lead = exp(p * log(x) - x - lgp);
    if (x <= 1.0 || x < p) {
        double sum = 1.0, term = 1.0, den = p;
        do {
            den += 1.0;
            term *= x / den;
            sum += term;
        } while (term > acc);
        return sum * lead / p;
    }
    double a = 1.0 - p, b = a + x + 1.0, k = 0.0;
    double A0 = 1.0, B0 = x, A1 = x + 1.0, B1 = x * b;
    double cf = A1 / B1;
    for (;;) {
        a += 1.0;
        b += 2.0;
        k += 1.0;
        const double an = a * k;
        const double A2 = b * A1 - an * A0, B2 = b * B1 - an * B0;
        if (B2 != 0.0) {
            const double r = A2 / B2, d = fabs(cf - r);
            if (d <= acc && d <= acc * r) return 1.0 - lead * cf;
            cf = r;
        }
        A0 = A1; B0 = B1; A1 = A2; B1 = B2;
        if (fabs(A1) >= big) { A0 /= big; B0 /= big; A1 /= big; B1 /= big; }
    }
}

static double orc_norm_quantile(double prob) {
    const double p1 = prob < 0.5 ? prob : 1.0 - prob;
    if (p1 < 1e-20) return -9999.0;
    const double y = sqrt(log(1.0 / (p1 * p1)));
    const double num = (((y * -0.453642210148e-4 - 0.0204231210245) * y - 0.342242088547) * y - 1.0) * y -
                       0.322232431088;
    const double den = (((y * 0.0038560700634 + 0.103537752850) * y + 0.531103462366) * y + 0.588581570495) * y +
                       0.0993484626060;
    const double z = y + num / den;
    return prob < 0.5 ? -z : z;
}

static double orc_chi2_quantile(double prob, double v) {
    const double e = 0.5e-6, ln2 = 0.6931471805;
    if (prob < 0.000002 || prob > 0.999998 || v <= 0.0) return -1.0;
    const double g = orc_lgamma(v / 2.0), xx = v / 2.0, c = xx - 1.0;
    double ch;
    if (v < -1.24 * log(prob)) {
        ch = pow(prob * xx * exp(g + xx * ln2), 1.0 / xx);
        if (ch - e < 0.0) return ch;
    } else if (v <= 0.32) {
        const double la = log(1.0 - prob);
        double prev;
        ch = 0.4;
        do {
            prev = ch;
            const double q1 = 1.0 + ch * (4.67 + ch);
            const double q2 = ch * (6.73 + ch * (6.66 + ch));
            const double tt = -0.5 + (4.67 + 2.0 * ch) / q1 - (6.73 + ch * (13.32 + 3.0 * ch)) / q2;
            ch -= (1.0 - exp(la + g + 0.5 * ch + c * ln2) * q2 / q1) / tt;
        } while (fabs(prev / ch - 1.0) - 0.01 > 0.0);
    } else {
        const double zq = orc_norm_quantile(prob), w = 0.222222 / v;
        ch = v * pow(zq * sqrt(w) + 1.0 - w, 3.0);
        if (ch > 2.2 * v + 6.0) ch = -2.0 * (log(1.0 - prob) - c * log(0.5 * ch) + g);
    }
    double prev;
    do {
        prev = ch;
        const double half = 0.5 * ch;
        double tt = orc_incgamma(half, xx, g);
        if (tt < 0.0) return -1.0;
        tt = (prob - tt) * exp(xx * ln2 + g + half - c * log(ch));
        const double b = tt / ch, a = 0.5 * tt - b * c;
        const double s1 = (210 + a * (140 + a * (105 + a * (84 + a * (70 + 60 * a))))) / 420;
        const double s2 = (420 + a * (735 + a * (966 + a * (1141 + 1278 * a)))) / 2520;
        const double s3 = (210 + a * (462 + a * (707 + 932 * a))) / 2520;
        const double s4 = (252 + a * (672 + 1182 * a) + c * (294 + a * (889 + 1740 * a))) / 5040;
        const double s5 = (84 + 264 * a + c * (175 + 606 * a)) / 2520;
        const double s6 = (120 + c * (346 + 127 * c)) / 5040;
        ch += tt * (1 + 0.5 * tt * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
    } while (fabs(prev / ch - 1.0) > e);
    return ch;
}

void orc_gamma_rates(double alpha, int ncat, double *rates) {
    if (ncat == 1) { rates[0] = 1.0; return; }
    /* cut points of Gamma(alpha, rate alpha) = chi2(2 alpha) / (2 alpha); the mean rate of a
     * category is ncat x (mass of Gamma(alpha + 1) between its cuts) */
    const double lg1 = orc_lgamma(alpha + 1.0);
    double prev = 0.0, sum = 0.0;
    for (int k = 0; k < ncat; k++) {
        double cum = 1.0;
        if (k < ncat - 1) {
            const double cut = orc_chi2_quantile((k + 1.0) / ncat, 2.0 * alpha) / (2.0 * alpha);
            cum = orc_incgamma(cut * alpha, alpha + 1.0, lg1);
        }
        rates[k] = (cum - prev) * ncat;
        prev = cum;
    }
    for (int k = 0; k < ncat; k++) sum += rates[k];
    if (sum != (double)ncat)
        for (int k = 0; k < ncat; k++) rates[k] *= ncat / sum;
}

/* the cut_median branch of RateGamma::computeRates (model/rategamma.cpp:111-126) */
void orc_gamma_rates_median(double alpha, int ncat, double *rates) {
    if (ncat == 1) { rates[0] = 1.0; return; }
    double sum = 0.0;
    for (int k = 0; k < ncat; k++) {
        rates[k] = fabs(orc_chi2_quantile((2.0 * k + 1.0) / (2.0 * ncat), 2.0 * alpha) / (2.0 * alpha));
        sum += rates[k];
    }
    for (int k = 0; k < ncat; k++) rates[k] = rates[k] * ncat / sum;
}

/* Optimization::brent_opt (utils/optimization.cpp:174-310): 100 iterations, golden
 * fraction 0.3819660, tolerance floor 1e-10, curvature of the last three points. */
static double orc_brent(orc_fn1 f, void *ctx, double ax, double bx, double cx, double tol, double fa, double fb,
                        double fc, double *fopt, double *f2opt) {
    double lo = ax < cx ? ax : cx, hi = ax > cx ? ax : cx;
    double x = bx, fx = fb, w, fw, v, fv;
    if (fa < fc) { w = ax; fw = fa; v = cx; fv = fc; }
    else { w = cx; fw = fc; v = ax; fv = fa; }
    double d = 0.0, e = 0.0;
    for (int it = 1; it <= 100; it++) {
        const double xm = 0.5 * (lo + hi), tol1 = tol * fabs(x) + 1.0e-10, tol2 = 2.0 * tol1;
        if (fabs(x - xm) <= tol2 - 0.5 * (hi - lo)) break;
        int parabolic = 0;
        if (fabs(e) > tol1) {
            const double r = (x - w) * (fx - fv);
            double q = (x - v) * (fx - fw), p = (x - v) * q - (x - w) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = fabs(q);
            const double etemp = e;
            e = d;
            if (!(fabs(p) >= fabs(0.5 * q * etemp) || p <= q * (lo - x) || p >= q * (hi - x))) {
                parabolic = 1;
                d = p / q;
                const double u = x + d;
                if (u - lo < tol2 || hi - u < tol2) d = (xm - x >= 0.0) ? fabs(tol1) : -fabs(tol1);
            }
        }
        if (!parabolic) {
            e = (x >= xm) ? lo - x : hi - x;
            d = 0.3819660 * e;
        }
        const double u = fabs(d) >= tol1 ? x + d : x + (d >= 0.0 ? fabs(tol1) : -fabs(tol1));
        const double fu = f(ctx, u);
        if (fu <= fx) {
            if (u >= x) lo = x; else hi = x;
            v = w; fv = fw; w = x; fw = fx; x = u; fx = fu;
        } else {
            if (u < x) lo = u; else hi = u;
            if (fu <= fw || w == x) { v = w; fv = fw; w = u; fw = fu; }
            else if (fu <= fv || v == x || v == w) { v = u; fv = fu; }
        }
    }
    *fopt = fx;
    const double xw = x - w, wv = w - v, vx = v - x;
    *f2opt = 2.0 * (fv * xw + fx * wv + fw * vx) / (v * v * xw + x * x * wv + w * w * vx);
    return x;
}

/* Optimization::minimizeOneDimen (utils/optimization.cpp:322-372) */
double orc_minimize_one_dimen(orc_fn1 f, void *ctx, double xmin, double xguess, double xmax, double tol, double *fx,
                              double *f2x) {
    if (xguess < xmin) xguess = xmin;
    if (xguess > xmax) xguess = xmax;
    const double eps = xguess * tol * 50.0;
    double ax = xguess - eps, bx = xguess, cx = xguess + eps;
    if (ax < xmin) ax = xmin;
    if (cx > xmax) cx = xmax;
    const double fb = f(ctx, bx);
    double fa = ax < bx ? f(ctx, ax) : fb;
    double fc = bx < cx ? f(ctx, cx) : fb;
    if (fa < fb || fc < fb) {
        if (ax != xmin) fa = f(ctx, xmin);
        if (cx != xmax) fc = f(ctx, xmax);
        ax = xmin;
        cx = xmax;
    }
    const double x = orc_brent(f, ctx, ax, bx, cx, tol, fa, fb, fc, fx, f2x);
    if (*fx > fb) {
        *fx = f(ctx, bx);
        return bx;
    }
    return x;
}

/* RateGamma::setGammaShape (model/rategamma.cpp:173-176) + clearAllPartialLH */
void orc_set_gamma_shape(orc_tree *t, double alpha) {
    t->gamma_shape = alpha;
    if (t->gamma_median) orc_gamma_rates_median(alpha, t->ncat, t->rates);
    else orc_gamma_rates(alpha, t->ncat, t->rates);
    orc_clear_all(t);
}

void orc_set_gamma_median(orc_tree *t, int median) { t->gamma_median = median != 0; }

double orc_get_gamma_shape(const orc_tree *t) { return t->gamma_shape; }

/* RateGamma::computeFunction (model/rategamma.cpp:178-185): -computeLikelihood() at the
 * shape, the first leaf's branch */
static double gamma_objective(void *ctx, double shape) {
    orc_tree *t = (orc_tree *)ctx;
    if (t->gamma_shape != shape) orc_set_gamma_shape(t, shape);
    return -orc_branch_loglik(t, 0, t->nei[0], NULL);
}

/* RateGamma::optimizeParameters(gradient_epsilon, min_gamma, max_gamma)
 * (model/rategamma.cpp:215-232), TOL_GAMMA_SHAPE = 0.001 (utils/tools.h:607) */
double orc_optimize_gamma(orc_tree *t, double gradient_epsilon, double min_gamma, double max_gamma) {
    double fx, f2x;
    const double tol = gradient_epsilon > 0.001 ? gradient_epsilon : 0.001;
    const double x = orc_minimize_one_dimen(gamma_objective, t, min_gamma, t->gamma_shape, max_gamma, tol, &fx, &f2x);
    return -gamma_objective(t, x);
}

/* ModelFactory::optimizeParameters (model/modelfactory.cpp:1558-1720), BRLEN_OPTIMIZE,
 * substitution model without free parameters (optimizeParametersOnly returns the rate
 * model's logL, 1263-1315), Gamma rates (rescaleRates = 1, model/rateheterogeneity.h:288) */
double orc_optimize_parameters(orc_tree *t, int num_iterations, double logl_epsilon, double gradient_epsilon,
                               double min_gamma, double max_gamma, double min_bl, double max_bl, int max_nr) {
    double cur = orc_branch_loglik(t, 0, t->nei[0], NULL);
    for (int i = 2; i < num_iterations; i++) {
        orc_optimize_all_branches(t, i < 3 ? i : 3, logl_epsilon, min_bl, max_bl, max_nr);
        const double new_lh = orc_optimize_gamma(t, gradient_epsilon, min_gamma, max_gamma);
        if (new_lh > cur + logl_epsilon) {
            cur = new_lh;
        } else {
            cur = orc_optimize_all_branches(t, 100, logl_epsilon, min_bl, max_bl, max_nr);
            break;
        }
    }
    return cur;
}
