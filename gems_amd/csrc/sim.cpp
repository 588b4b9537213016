// Seeded synthetic problems: Yule-Harding trees, random model parameters, sequence
// simulation down the tree and pattern compression.  Host-only.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "plk.h"
#include "plk_internal.h"
#include "plk_model.h"

namespace {

// xoshiro256** seeded through splitmix64 (public-domain generators by Blackman & Vigna)
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            s[i] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * 0x1.0p-53; }            // [0,1)
    double uniform_open() { return ((next() >> 11) + 0.5) * 0x1.0p-53; }  // (0,1)
    int integer(int n) { return (int)(uniform() * n); }
    double exponential(double mean) { return -mean * std::log(uniform_open()); }
    double normal() {
        double u1 = uniform_open(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    // Marsaglia & Tsang (2000) gamma variate, shape k, scale 1
    double gamma(double k) {
        if (k < 1.0) return gamma(k + 1.0) * std::pow(uniform_open(), 1.0 / k);
        double d = k - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do {
                x = normal();
                v = 1.0 + c * x;
            } while (v <= 0.0);
            v = v * v * v;
            double u = uniform_open();
            if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
            if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
        }
    }
};

}  // namespace

struct plk_problem {
    int seq_type = 0, n = 0, ntaxa = 0, nsites = 0, nptn = 0, ncat = 0, su = 0;
    std::vector<double> pi, Q, eval, evec, inv_evec, rates, props, tip_lh, freq, blen, params;
    std::vector<int32_t> nei;
    std::vector<uint8_t> states;
};

static int sim_create(plk_problem **out, int seq_type, int ntaxa, const int32_t *given_nei, int nsites,
                      double mean_len, double alpha, int ncat, uint64_t seed) {
    if (!out || ntaxa < 3 || ntaxa > 100000 || nsites < 1 || !(mean_len > 0) || ncat < 1 || ncat > PLK_MAX_CATS)
        return plk::fail(PLK_ERR_ARG, "plk_sim_create: bad arguments");
    if (seq_type < PLK_SEQ_DNA || seq_type > PLK_SEQ_CODON) return plk::fail(PLK_ERR_ARG, "plk_sim_create: bad seq_type");
    Rng rng(seed);
    auto *p = new plk_problem();
    p->seq_type = seq_type;
    p->n = seq_type == PLK_SEQ_DNA ? 4 : seq_type == PLK_SEQ_PROTEIN ? 20 : 61;
    p->ntaxa = ntaxa;
    p->nsites = nsites;
    p->ncat = ncat;
    p->su = plk_state_unknown(p->n, seq_type);
    const int n = p->n;

    // --- model parameters
    std::vector<double> exch((size_t)n * (n - 1) / 2);
    double kappa = 0.0, omega = 0.0;
    if (seq_type == PLK_SEQ_DNA) {
        for (auto &e : exch) e = 0.5 + 1.5 * rng.uniform();
    } else if (seq_type == PLK_SEQ_PROTEIN) {
        for (auto &e : exch) e = rng.gamma(1.0);
    } else {
        kappa = 1.0 + 3.0 * rng.uniform();
        omega = 0.1 + 0.9 * rng.uniform();
        plk_codon_exch(kappa, omega, exch.data());
    }
    p->pi.resize(n);
    double sum = 0.0;
    for (auto &v : p->pi) sum += (v = rng.gamma(1.0));
    for (auto &v : p->pi) v /= sum;
    p->params = {alpha, kappa, omega};
    p->Q.resize((size_t)n * n);
    p->eval.resize(n);
    p->evec.resize((size_t)n * n);
    p->inv_evec.resize((size_t)n * n);
    p->rates.resize(ncat);
    p->props.assign(ncat, 1.0 / ncat);
    p->tip_lh.resize((size_t)(p->su + 1) * n);
    int rc = plk_rate_matrix(n, exch.data(), p->pi.data(), p->Q.data());
    if (rc == PLK_OK) rc = plk_eigen(n, p->Q.data(), p->pi.data(), p->eval.data(), p->evec.data(), p->inv_evec.data());
    if (rc == PLK_OK) rc = plk_gamma_rates(alpha, ncat, p->rates.data());
    if (rc == PLK_OK) rc = plk_tip_lh(n, seq_type, p->inv_evec.data(), p->tip_lh.data());
    if (rc != PLK_OK) {
        delete p;
        return rc;
    }

    // --- Yule-Harding tree (tree/mexttree.cpp:432-490): start from a 3-leaf star and
    // repeatedly turn a uniformly chosen leaf into a cherry.
    struct Edge { int a, b; double len; };
    std::vector<Edge> edges;
    std::vector<int> leaves;
    int next_id = 1;  // node 0 = initial centre
    for (int i = 0; i < 3; i++) {
        int leaf = next_id++;
        edges.push_back({0, leaf, rng.exponential(mean_len)});
        leaves.push_back(leaf);
    }
    for (int i = 3; i < ntaxa; i++) {
        int idx = rng.integer(i);
        int node = leaves[idx];
        int l1 = next_id++, l2 = next_id++;
        edges.push_back({node, l1, rng.exponential(mean_len)});
        edges.push_back({node, l2, rng.exponential(mean_len)});
        leaves[idx] = l1;
        leaves.push_back(l2);
    }
    // renumber: leaves 0..T-1 in leaf order, internal nodes T..2T-3
    int nn = 2 * ntaxa - 2;
    std::vector<int> id(next_id, -1);
    for (int i = 0; i < ntaxa; i++) id[leaves[i]] = i;
    int nxt = ntaxa;
    for (int v = 0; v < next_id; v++)
        if (id[v] < 0) id[v] = nxt++;
    p->nei.assign((size_t)nn * 3, -1);
    p->blen.assign((size_t)nn * 3, 0.0);
    auto link = [&](int a, int b, double len) {
        int ka = 0, kb = 0;
        while (p->nei[a * 3 + ka] >= 0) ka++;
        while (p->nei[b * 3 + kb] >= 0) kb++;
        p->nei[a * 3 + ka] = b;
        p->nei[b * 3 + kb] = a;
        p->blen[a * 3 + ka] = p->blen[b * 3 + kb] = len;
    };
    for (auto &e : edges) link(id[e.a], id[e.b], std::max(e.len, 1e-6));
    if (given_nei) {
        // caller's topology, fresh lengths (the random Yule draw above keeps the
        // generator stream identical between the two entry points)
        p->nei.assign(given_nei, given_nei + (size_t)nn * 3);
        p->blen.assign((size_t)nn * 3, 0.0);
        for (int u = 0; u < nn; u++)
            for (int k = 0; k < 3; k++) {
                int v = p->nei[u * 3 + k];
                if (v < 0) continue;
                if (v < 0 || v >= nn) {
                    delete p;
                    return plk::fail(PLK_ERR_ARG, "plk_sim_create_on_tree: bad topology");
                }
                if (v > u) {
                    double len = std::max(rng.exponential(mean_len), 1e-6);
                    p->blen[u * 3 + k] = len;
                    int kb = 0;
                    while (kb < 3 && p->nei[v * 3 + kb] != u) kb++;
                    if (kb == 3) {
                        delete p;
                        return plk::fail(PLK_ERR_ARG, "plk_sim_create_on_tree: topology not symmetric");
                    }
                    p->blen[v * 3 + kb] = len;
                }
            }
    }

    // --- simulate sites down the tree from the first internal node
    std::vector<std::vector<uint8_t>> seq(nn, std::vector<uint8_t>(nsites));
    std::vector<int> site_cat(nsites);
    for (int s = 0; s < nsites; s++) site_cat[s] = rng.integer(ncat);
    std::vector<double> cum_pi(n);
    double acc = 0.0;
    for (int i = 0; i < n; i++) cum_pi[i] = (acc += p->pi[i]);
    auto draw = [&](const double *cum) {
        double u = rng.uniform() * cum[n - 1];
        int k = 0;
        while (k < n - 1 && u >= cum[k]) k++;
        return (uint8_t)k;
    };
    int root = ntaxa;
    for (int s = 0; s < nsites; s++) seq[root][s] = draw(cum_pi.data());
    std::vector<std::pair<int, int>> stack = {{root, -1}};
    std::vector<double> P((size_t)ncat * n * n);
    while (!stack.empty()) {
        auto [u, par] = stack.back();
        stack.pop_back();
        for (int k = 0; k < 3; k++) {
            int v = p->nei[u * 3 + k];
            if (v < 0 || v == par) continue;
            double t = p->blen[u * 3 + k];
            // cumulative rows of P(r_c t) = U exp(L r_c t) U^-1
            std::vector<double> ex((size_t)ncat * n);
            for (int c = 0; c < ncat; c++)
                for (int i = 0; i < n; i++) ex[c * n + i] = std::exp(p->eval[i] * p->rates[c] * t);
            for (int c = 0; c < ncat; c++)
                for (int x = 0; x < n; x++) {
                    double run = 0.0;
                    for (int y = 0; y < n; y++) {
                        double v2 = 0.0;
                        for (int i = 0; i < n; i++)
                            v2 += p->evec[x * n + i] * ex[c * n + i] * p->inv_evec[i * n + y];
                        run += std::max(v2, 0.0);
                        P[((size_t)c * n + x) * n + y] = run;
                    }
                }
            for (int s = 0; s < nsites; s++)
                seq[v][s] = draw(&P[((size_t)site_cat[s] * n + seq[u][s]) * n]);
            if (v >= ntaxa) stack.push_back({v, u});
        }
    }
    // --- pattern compression in first-occurrence order (alignment/alignment.cpp:1511-1560)
    std::unordered_map<std::string, int> index;
    std::string key(ntaxa, '\0');
    std::vector<int> col_of;
    for (int s = 0; s < nsites; s++) {
        for (int t = 0; t < ntaxa; t++) key[t] = (char)seq[t][s];
        auto it = index.find(key);
        if (it == index.end()) {
            index.emplace(key, (int)col_of.size());
            col_of.push_back(s);
            p->freq.push_back(1.0);
        } else {
            p->freq[it->second] += 1.0;
        }
    }
    p->nptn = (int)col_of.size();
    p->states.resize((size_t)ntaxa * p->nptn);
    for (int t = 0; t < ntaxa; t++)
        for (int k = 0; k < p->nptn; k++) p->states[(size_t)t * p->nptn + k] = seq[t][col_of[k]];
    *out = p;
    return PLK_OK;
}

extern "C" int plk_sim_create(plk_problem **out, int seq_type, int ntaxa, int nsites, double mean_len,
                              double alpha, int ncat, uint64_t seed) {
    return sim_create(out, seq_type, ntaxa, nullptr, nsites, mean_len, alpha, ncat, seed);
}

extern "C" int plk_sim_create_on_tree(plk_problem **out, int seq_type, int ntaxa, const int32_t *nei, int nsites,
                                      double mean_len, double alpha, int ncat, uint64_t seed) {
    if (!nei) return plk::fail(PLK_ERR_ARG, "plk_sim_create_on_tree: null topology");
    return sim_create(out, seq_type, ntaxa, nei, nsites, mean_len, alpha, ncat, seed);
}

extern "C" int plk_sim_destroy(plk_problem *p) {
    delete p;
    return PLK_OK;
}

extern "C" int plk_sim_dims(const plk_problem *p, int *dims) {
    if (!p || !dims) return plk::fail(PLK_ERR_ARG, "plk_sim_dims: null");
    dims[0] = p->n; dims[1] = p->ntaxa; dims[2] = p->nsites; dims[3] = p->nptn;
    dims[4] = p->ncat; dims[5] = p->su; dims[6] = p->seq_type;
    return PLK_OK;
}

extern "C" const void *plk_sim_array(const plk_problem *p, const char *name, int64_t *count) {
    if (!p || !name) return nullptr;
    std::string s(name);
    auto ret = [&](const auto &v) -> const void * {
        if (count) *count = (int64_t)v.size();
        return v.data();
    };
    if (s == "pi") return ret(p->pi);
    if (s == "Q") return ret(p->Q);
    if (s == "eval") return ret(p->eval);
    if (s == "evec") return ret(p->evec);
    if (s == "inv_evec") return ret(p->inv_evec);
    if (s == "rates") return ret(p->rates);
    if (s == "props") return ret(p->props);
    if (s == "tip_lh") return ret(p->tip_lh);
    if (s == "freq") return ret(p->freq);
    if (s == "blen") return ret(p->blen);
    if (s == "params") return ret(p->params);
    if (s == "nei") return ret(p->nei);
    if (s == "states") return ret(p->states);
    plk::fail(PLK_ERR_ARG, "plk_sim_array: unknown array " + s);
    return nullptr;
}
