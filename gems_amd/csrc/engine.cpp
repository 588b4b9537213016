// Host runtime of the C ABI (include/plk.h): context, HBM buffers, op validation and
// kernel launches on the context's HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <queue>
#include <vector>

#include "engine_common.h"

namespace plk {

static thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}
void clear_error() { g_last_error.clear(); }

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(e_ == hipErrorOutOfMemory ? PLK_ERR_NOMEM : PLK_ERR_HIP,                \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                     \
    } while (0)

static int64_t round_up(int64_t v, int64_t g) { return (v + g - 1) / g * g; }

static bool supported_states(int n) { return n == 4 || n == 20 || n == 61; }

// per-op matrix block (in reals): transition data for both children + leaf tables for
// both children.  FP64: echild[c][x][i]; FP32: matrix-core fragments of P_c.
static int64_t mat_stride_of(int n, int ncat, int su, int prec) {
    if (prec == PLK_FP32) {
        int64_t fr = (int64_t)f32_xblocks(n) * f32_ksteps(n) * 64;
        return round_up(2LL * ncat * fr + 2LL * (su + 1) * f32_trow(n) * ncat, 64);
    }
    return round_up(2LL * ncat * n * n + 2LL * (su + 1) * n * ncat, 32);
}

int alloc_stage(Ctx &c, size_t bytes) {
    if (bytes <= c.h_stage_cap) return PLK_OK;
    if (c.h_stage) (void)hipHostFree(c.h_stage);
    c.h_stage = nullptr;
    c.h_stage_cap = 0;
    HIP_TRY(hipHostMalloc(&c.h_stage, bytes, hipHostMallocDefault));
    c.h_stage_cap = bytes;
    return PLK_OK;
}

// Stream-ordered allocations: hipFree would synchronise every stream of the device and
// stall other contexts' work (the serving loop runs many contexts concurrently).
int ensure_op_capacity(Ctx &c, int nops) {
    if (nops <= c.ops_cap && nops <= c.mats_cap) return PLK_OK;
    int cap = std::max(nops, std::max(16, c.ops_cap * 2));
    if (c.d_ops) (void)hipFreeAsync(c.d_ops, c.stream);
    if (c.d_jobs) (void)hipFreeAsync(c.d_jobs, c.stream);
    if (c.d_mats) (void)hipFreeAsync(c.d_mats, c.stream);
    c.d_ops = nullptr;
    c.d_jobs = nullptr;
    c.d_mats = nullptr;
    c.ops_cap = c.mats_cap = 0;
    // ops, then (level-parallel kernel) nops + 1 level bounds and up to ntaxa + 1 tip ids
    HIP_TRY(hipMallocAsync((void **)&c.d_ops, sizeof(DevOp) * cap + sizeof(int32_t) * (cap + c.ntaxa + 4), c.stream));
    HIP_TRY(hipMallocAsync((void **)&c.d_jobs, sizeof(PJob) * 2 * cap, c.stream));
    HIP_TRY(hipMallocAsync(&c.d_mats, c.real_size() * c.mat_stride * cap, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));  // usable from any stream (batched launches)
    c.ops_cap = c.mats_cap = cap;
    return PLK_OK;
}

KParams make_params(Ctx &c) {
    KParams p{};
    p.n = c.n;
    p.ncat = c.ncat;
    p.block = c.block;
    p.nptn = c.nptn;
    p.nptn_pad = c.nptn_pad;
    p.ntiles = c.ntiles;
    p.su = c.su;
    p.safe = c.safe;
    p.ncw = std::min(c.ncat, 4);
    p.wpc = c.wpc;
    p.part = c.d_part;
    p.slot_stride = (int64_t)c.nptn_pad * c.block;
    p.scale = c.d_scale;
    p.scale_stride = (int64_t)c.nptn_pad * ((c.safe && c.prec == PLK_FP64) ? c.ncat : 1);
    p.states = c.d_states;
    p.ntaxa = c.ntaxa;
    p.freq = c.d_freq;
    p.invar = c.d_invar;
    p.mats = c.d_mats;
    p.mat_stride = c.mat_stride;
    p.inv_evec = c.d_inv_evec;
    p.eval = c.d_eval;
    p.rates = c.d_rates;
    p.props = c.d_props;
    p.tip_lh = c.d_tip;
    p.evec = c.d_evec;
    p.pi = c.d_pi;
    p.piU = c.d_piU;
    p.tipvec = c.d_tipvec;
    p.uinv_frag = c.d_uinv_frag;
    // measured per traversal: 61 states 3.06 -> 0.64 ms on the matrix cores; 20 states
    // 2.22 ms vector ALU vs 2.61 ms matrix cores (rows padded 20 -> 32)
    const int form = c.f64_form >= 0 ? c.f64_form : (c.n == 61 ? 1 : 0);
    p.f64_form = (c.d_uinv_frag && c.ncat <= 4) ? form : 0;
    p.ops = c.d_ops;
    p.nops = 0;
    p.ptn_lh = c.d_ptn_lh;
    p.ptn_lh_cat = c.d_side;
    p.ptn_scaling = c.d_side + (size_t)c.nptn_pad * c.ncat;
    p.g_frac = c.d_side + (size_t)c.nptn_pad * (c.ncat + 1);
    p.theta = c.d_theta;
    p.buf_scale = c.d_buf_scale;
    p.tile_sums = c.d_tile_sums;
    p.counter = c.d_counter;
    p.result = c.d_result;
    p.sink = c.d_sink;
    p.ncu = c.ncu;
    p.split_full = 0;
    return p;
}

static void timing_begin(Ctx &c, int kind) {
    if (c.timing) (void)hipEventRecord(c.ev[2 * kind], c.stream);
}
static void timing_end(Ctx &c, int kind) {
    if (c.timing) (void)hipEventRecord(c.ev[2 * kind + 1], c.stream);
}
// called after the stream is synchronised
static void timing_collect(Ctx &c, const bool used[3]) {
    if (!c.timing) return;
    for (int k = 0; k < 3; k++) {
        if (!used[k]) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c.ev[2 * k], c.ev[2 * k + 1]) == hipSuccess) {
            c.t_total[k] += ms;
            c.t_count[k] += 1;
        }
    }
}

// Level-parallel 4-state kernel plan: sort the queued ops into dependency levels
// (read-after-write, write-after-read and write-after-write on slots, so any op list
// keeps its meaning), give partials consumed later in the launch an LDS buffer (by
// liveness, at most max_buf), and stage the tips' states.  Reorders c.pending and
// remaps the P(t) jobs.
struct SmallPlan {
    std::vector<int32_t> lvl, tips;
    int nbuf = 0, br_lds_a = -1, br_lds_b = -1;
    bool stage = false;
    size_t lds_bytes = 0;
};

static void plan_small(Ctx &c, bool branch, int br_a, int br_b, SmallPlan &sp) {
    std::vector<DevOp> &ops = c.pending;
    const int n = (int)ops.size();
    auto internal_children = [](const DevOp &o, int *ch) {
        int m = 0;
        if (o.kind == KIND_TI) ch[m++] = o.right;
        else if (o.kind == KIND_II) {
            ch[m++] = o.left;
            ch[m++] = o.right;
        }
        return m;
    };
    // 1. levels
    std::vector<int> level(n, 0), writer(c.nslots, -1), maxread(c.nslots, -1);
    int nlev = 0;
    for (int k = 0; k < n; k++) {
        int ch[2], m = internal_children(ops[k], ch), lv = 0;
        for (int j = 0; j < m; j++)
            if (writer[ch[j]] >= 0) lv = std::max(lv, level[writer[ch[j]]] + 1);
        const int o = ops[k].out_slot;
        if (writer[o] >= 0) lv = std::max(lv, level[writer[o]] + 1);
        lv = std::max(lv, maxread[o] + 1);
        level[k] = lv;
        for (int j = 0; j < m; j++) maxread[ch[j]] = std::max(maxread[ch[j]], lv);
        writer[o] = k;
        maxread[o] = -1;
        nlev = std::max(nlev, lv + 1);
    }
    // 2. stable sort by level, remap the P(t) jobs
    std::vector<int> perm(n), newidx(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return level[a] < level[b]; });
    std::vector<DevOp> sorted(n);
    for (int i = 0; i < n; i++) {
        sorted[i] = ops[perm[i]];
        sorted[i].level = level[perm[i]];
        newidx[perm[i]] = i;
    }
    ops.swap(sorted);
    for (PJob &j : c.jobs)
        if (j.op < n) j.op = newidx[j.op];
    sp.lvl.assign(nlev + 1, 0);
    for (int i = 0; i < n; i++) sp.lvl[ops[i].level + 1] = i + 1;
    for (int l = 1; l <= nlev; l++) sp.lvl[l] = std::max(sp.lvl[l], sp.lvl[l - 1]);
    // 3. producers of every read and the last level that reads each value
    std::vector<int> current(c.nslots, -1), last_use(n, -1), prod_l(n, -1), prod_r(n, -1);
    for (int i = 0; i < n; i++) {
        DevOp &o = ops[i];
        if (o.kind == KIND_II) {
            prod_l[i] = current[o.left];
            if (prod_l[i] >= 0) last_use[prod_l[i]] = std::max(last_use[prod_l[i]], o.level);
        }
        if (o.kind != KIND_TT) {
            prod_r[i] = current[o.right];
            if (prod_r[i] >= 0) last_use[prod_r[i]] = std::max(last_use[prod_r[i]], o.level);
        }
        current[o.out_slot] = i;
    }
    int pa = -1, pb = -1;
    if (branch) {
        if (br_a >= 0) pa = current[br_a];
        if (br_b >= 0) pb = current[br_b];
        if (pa >= 0) last_use[pa] = nlev;
        if (pb >= 0) last_use[pb] = nlev;
    }
    // 4. staged tips (up to 256)
    std::vector<int> tip_index(c.ntaxa, -1);
    auto stage_tip = [&](int taxon) {
        if (tip_index[taxon] < 0 && (int)sp.tips.size() < 256) {
            tip_index[taxon] = (int)sp.tips.size();
            sp.tips.push_back(taxon);
        }
        return tip_index[taxon];
    };
    for (DevOp &o : ops) {
        o.lds_left = o.lds_right = o.lds_out = -1;
        if (o.kind <= KIND_TI) o.lds_left = stage_tip(o.left);
        if (o.kind == KIND_TT) o.lds_right = stage_tip(o.right);
    }
    if (branch && br_a < 0) sp.br_lds_a = stage_tip(-br_a - 1);
    // 5. LDS buffers by liveness within the budget
    const int form = c.prec == PLK_FP64 ? 0 : 1;
    // dynamic LDS for partial buffers: up to 112 KB when every tile has a compute unit to
    // itself (the kernel's own arrays take 24 KB, the staged op list up to 16 KB, of the
    // 160 KB), else 32 KB so that two workgroups share one
    const size_t budget = (c.ntiles <= 256 ? 112 : 32) * 1024;
    const size_t fixed = small_lds_bytes(form, c.ncat, c.safe, c.su, 0, (int)sp.tips.size());
    const size_t per_buf = small_lds_bytes(form, c.ncat, c.safe, c.su, 1, 0) - small_lds_bytes(form, c.ncat, c.safe, c.su, 0, 0);
    const int max_buf = budget > fixed ? (int)std::min<size_t>(64, (budget - fixed) / per_buf) : 0;
    std::vector<int> buf_of(n, -1), holder;  // holder[b] = value in buffer b
    for (int l = 0, i = 0; l < nlev; l++) {
        for (size_t b = 0; b < holder.size(); b++)
            if (holder[b] >= 0 && last_use[holder[b]] < l) holder[b] = -1;
        for (; i < n && ops[i].level == l; i++) {
            if (last_use[i] < 0) continue;
            int b = -1;
            for (size_t q = 0; q < holder.size(); q++)
                if (holder[q] < 0) {
                    b = (int)q;
                    break;
                }
            if (b < 0 && (int)holder.size() < max_buf) {
                b = (int)holder.size();
                holder.push_back(-1);
            }
            if (b < 0) continue;
            holder[b] = i;
            buf_of[i] = b;
            ops[i].lds_out = b;
        }
    }
    for (int i = 0; i < n; i++) {
        if (prod_l[i] >= 0) ops[i].lds_left = buf_of[prod_l[i]];
        if (prod_r[i] >= 0) ops[i].lds_right = buf_of[prod_r[i]];
    }
    if (pa >= 0) sp.br_lds_a = buf_of[pa];
    if (pb >= 0) sp.br_lds_b = buf_of[pb];
    sp.nbuf = (int)holder.size();
    sp.lds_bytes = small_lds_bytes(form, c.ncat, c.safe, c.su, sp.nbuf, (int)sp.tips.size());
    // small op lists are staged in LDS with their P(t) exponentials (computed once per
    // tile by all threads instead of once per lane of every op's wavefronts)
    const size_t stage = small_stage_bytes(c.ncat, n, nlev);
    sp.stage = n > 0 && stage <= 16 * 1024;
    if (sp.stage) sp.lds_bytes = ((sp.lds_bytes + 15) & ~(size_t)15) + stage;
}

// Upload the queued ops, run the P(t) kernel, then one traversal launch in `mode`.
// The FP32 form also needs P(t) of the branch itself (an extra job after the ops) for
// the logL stage, and computes theta in a separate FP64 kernel after the traversal.
// A second valid order of an op list for half of the tiles (see KParams::perm): Kahn's
// algorithm over the read-after-write dependencies, always taking the ready op with the
// largest index (the mirror image of a post-order).  Only for lists in which no slot is
// written twice or written after an earlier op read it; otherwise (and for short lists)
// returns false and every tile keeps the list's order.  Each op computes the same values in
// either order, so results do not depend on it.
static bool alt_order(const Ctx &c, const std::vector<DevOp> &ops, std::vector<int32_t> &perm) {
    const int n = (int)ops.size();
    if (n < 8) return false;
    std::vector<int> writer(c.nslots, -1), reader_last(c.nslots, -1);
    std::vector<std::vector<int>> succ(n);
    std::vector<int> indeg(n, 0);
    auto reads = [&](const DevOp &o, int &a, int &b) {
        a = o.kind == KIND_II ? o.left : -1;
        b = o.kind != KIND_TT ? o.right : -1;
    };
    for (int j = 0; j < n; j++) {
        int a, b;
        reads(ops[j], a, b);
        for (int s : {a, b}) {
            if (s < 0) continue;
            if (writer[s] >= 0) {
                succ[writer[s]].push_back(j);
                indeg[j]++;
            }
            reader_last[s] = j;
        }
        const int w = ops[j].out_slot;
        if (writer[w] >= 0 || reader_last[w] >= 0) return false;  // WAW or WAR
        writer[w] = j;
    }
    std::priority_queue<int> ready;
    for (int j = 0; j < n; j++)
        if (indeg[j] == 0) ready.push(j);
    perm.clear();
    while (!ready.empty()) {
        const int j = ready.top();
        ready.pop();
        perm.push_back(j);
        for (int k : succ[j])
            if (--indeg[k] == 0) ready.push(k);
    }
    return (int)perm.size() == n;
}

static int run_traversal(Ctx &c, KParams &p, int mode, bool used[3]) {
    const int nops = (int)c.pending.size();
    const bool f32 = c.prec == PLK_FP32;
    if (f32 && mode == MODE_LOGL) c.jobs.push_back(PJob{nops, 0, 0, 0, p.br_len});
    int rc = ensure_op_capacity(c, nops + 1);
    if (rc) return rc;
    p.ops = c.d_ops;
    p.mats = c.d_mats;
    p.nops = nops;
    // 4-state models: level-parallel kernel (partials, and the logL stage when asked)
    // auto: the level kernel forms P(t) itself, which saves the P(t) launch (what dominates
    // small alignments) but repeats the exponentials in every tile (which costs on large ones)
    const int slots = c.small_waves >= 0 ? c.small_waves : (c.ntiles <= 320 ? 4 : 0);
    const bool small = c.n == 4 && slots > 0 && (nops > 0 || mode == MODE_LOGL);
    SmallPlan sp;
    if (small) plan_small(c, mode != MODE_PARTIALS, p.br_a, p.br_b, sp);
    if (small) c.jobs.clear();  // the level-parallel kernel forms P(t) itself
    std::vector<int32_t> perm;
    // (not for the 20-state FP32 traversal, which carries each op's output into the next
    // op of the list's own order)
    const bool two_orders = !small && !(f32 && c.n == 20) && alt_order(c, c.pending, perm);
    bool inline_ops = false;
    SmallInline inl;
    p.perm = nullptr;
    const int njobs = (int)c.jobs.size();
    if (njobs > 0 || small || two_orders) {
        size_t ob = sizeof(DevOp) * nops, jb = sizeof(PJob) * njobs;
        size_t xb = small ? sizeof(int32_t) * (sp.lvl.size() + sp.tips.size()) : two_orders ? sizeof(int32_t) * nops : 0;
        rc = alloc_stage(c, ob + xb + jb);
        if (rc) return rc;
        char *h = (char *)c.h_stage;
        if (nops > 0) std::memcpy(h, c.pending.data(), ob);
        if (small) {
            std::memcpy(h + ob, sp.lvl.data(), sizeof(int32_t) * sp.lvl.size());
            if (!sp.tips.empty()) std::memcpy(h + ob + sizeof(int32_t) * sp.lvl.size(), sp.tips.data(), sizeof(int32_t) * sp.tips.size());
            p.lvl = (const int32_t *)((char *)c.d_ops + ob);
            p.tips = p.lvl + sp.lvl.size();
            p.nlevels = (int)sp.lvl.size() - 1;
            p.ntip = (int)sp.tips.size();
            p.nbuf = sp.nbuf;
            p.br_lds_a = sp.br_lds_a;
            p.br_lds_b = sp.br_lds_b;
            p.small_stage = sp.stage ? 1 : 0;
        }
        if (two_orders) {
            std::memcpy(h + ob, perm.data(), sizeof(int32_t) * nops);
            p.perm = (const int32_t *)((char *)c.d_ops + ob);
        }
        if (njobs > 0) std::memcpy(h + ob + xb, c.jobs.data(), jb);
        // a short staged list of the level-parallel kernel travels in its kernel arguments
        inline_ops = small && sp.stage && njobs == 0 && (ob + xb) / 4 <= (size_t)kSmallInlineWords;
        if (inline_ops) {
            inl.nwords = (int32_t)((ob + xb) / 4);
            inl.pad = 0;
            std::memcpy(inl.data, h, ob + xb);
        } else if (ob + xb > 0) {
            HIP_TRY(hipMemcpyAsync(c.d_ops, h, ob + xb, hipMemcpyHostToDevice, c.stream));
        }
        if (njobs > 0) {
            HIP_TRY(hipMemcpyAsync(c.d_jobs, h + ob + xb, jb, hipMemcpyHostToDevice, c.stream));
            timing_begin(c, 1);
            HIP_TRY(f32 ? f32_pmatrix(p, c.d_jobs, njobs, c.stream) : f64_pmatrix(p, c.d_jobs, njobs, c.stream));
            timing_end(c, 1);
            used[1] = true;
        }
    }
    if (small) {
        const int smode = mode == MODE_LOGL ? MODE_LOGL : MODE_PARTIALS;
        timing_begin(c, 0);
        HIP_TRY(small_traverse(p, f32 ? 1 : 0, smode, slots, sp.lds_bytes, inline_ops ? &inl : nullptr, c.stream));
        timing_end(c, 0);
        used[0] = true;
        if (mode == MODE_THETA) {
            // theta and derivatives at the branch from the partials just computed
            KParams q = p;
            q.nops = 0;
            timing_begin(c, 2);
            HIP_TRY(f32 ? f32_theta(q, c.stream) : f64_traverse(q, MODE_THETA, c.stream));
            timing_end(c, 2);
            used[2] = true;
        }
    } else if (f32) {
        const int tmode = mode == MODE_THETA ? MODE_PARTIALS : mode;
        if (nops > 0 || tmode != MODE_PARTIALS) {
            timing_begin(c, 0);
            HIP_TRY(f32_traverse(p, tmode, c.stream));
            timing_end(c, 0);
            used[0] = true;
        }
        if (mode == MODE_THETA) {
            timing_begin(c, 2);
            HIP_TRY(f32_theta(p, c.stream));
            timing_end(c, 2);
            used[2] = true;
        }
    } else if (nops > 0 || mode != MODE_PARTIALS) {
        timing_begin(c, 0);
        HIP_TRY(f64_traverse(p, mode, c.stream));
        timing_end(c, 0);
        used[0] = true;
    }
    c.updates += (int64_t)nops * c.nptn;
    c.pending.clear();
    c.jobs.clear();
    return PLK_OK;
}

static int check_ctx(plk_ctx *ctx) {
    if (!ctx) return fail(PLK_ERR_ARG, "null context");
    if (!ctx->model_set || !ctx->patterns_set) return fail(PLK_ERR_STATE, "model and patterns must be set first");
    return PLK_OK;
}

bool valid_ref(const Ctx &c, int r) { return r >= 0 ? r < c.nslots : (-r - 1) < c.ntaxa; }

}  // namespace plk

using namespace plk;

extern "C" const char *plk_last_error(void) { return plk::g_last_error.c_str(); }
extern "C" int plk_version(void) { return 1; }

extern "C" int plk_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int plk_device_units(void) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return n;
}

extern "C" int64_t plk_memory_required(int nstates, int ncat, int ntaxa, int nptn, int nslots, int precision,
                                       int scaling) {
    int64_t pad = round_up(nptn, 64), block = (int64_t)nstates * ncat, rs = precision == PLK_FP64 ? 8 : 4;
    int64_t su = nstates == 4 ? 18 : nstates == 20 ? 23 : nstates;
    int64_t part = (int64_t)nslots * pad * block * rs;
    int64_t scale = (int64_t)nslots * pad * (precision == PLK_FP64 ? (scaling ? ncat : 1) * 2 : 4);
    int64_t theta = pad * block * 8;
    int64_t states = (int64_t)ntaxa * pad;
    int64_t mats = (int64_t)(ntaxa + 1) * mat_stride_of(nstates, ncat, (int)su, precision) * rs;
    int64_t misc = pad * 8 * 4 + (pad / 64 + 1) * 16;
    return part + scale + theta + states + mats + misc;
}

extern "C" int plk_create(plk_ctx **out, int nstates, int ncat, int ntaxa, int nptn, int nslots, int precision,
                          int scaling) {
    clear_error();
    if (!out) return fail(PLK_ERR_ARG, "plk_create: null out");
    *out = nullptr;
    if (!supported_states(nstates)) return fail(PLK_ERR_ARG, "plk_create: supported state counts are 4, 20, 61");
    if (ncat < 1 || ncat > PLK_GPU_MAX_CATS) return fail(PLK_ERR_ARG, "plk_create: ncat must be in [1, 8]");
    if (ntaxa < 3 || nptn < 1 || nslots < 1) return fail(PLK_ERR_ARG, "plk_create: need ntaxa >= 3, nptn >= 1, nslots >= 1");
    if (precision != PLK_FP64 && precision != PLK_FP32) return fail(PLK_ERR_ARG, "plk_create: bad precision form");
    if (scaling != PLK_SCALE_NORM && scaling != PLK_SCALE_SAFE) return fail(PLK_ERR_ARG, "plk_create: bad scaling");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(PLK_ERR_HIP, "plk_create: no HIP device visible (the engine has no CPU path)");
    auto *c = new plk_ctx();
    int dev = 0;
    (void)hipGetDevice(&dev);
    {
        // keep freed stream-ordered memory in the device pool for the next context
        static std::once_flag pool_once;
        std::call_once(pool_once, [dev] {
            hipMemPool_t pool = nullptr;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
                uint64_t thr = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
            }
        });
    }
    c->dev = dev;
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) c->ncu = 0;
    c->n = nstates;
    c->ncat = ncat;
    c->block = nstates * ncat;
    c->ntaxa = ntaxa;
    c->nptn = nptn;
    c->nptn_pad = (int)round_up(nptn, 64);
    c->ntiles = c->nptn_pad / 64;
    c->nslots = nslots;
    c->prec = precision;
    c->safe = scaling;
    c->su = nstates == 4 ? 18 : nstates == 20 ? 23 : nstates;
    c->mat_stride = mat_stride_of(nstates, ncat, c->su, precision);
    const size_t rs = c->real_size();
    const int64_t slot = (int64_t)c->nptn_pad * c->block;
    // FP64: uint16 scale_num per pattern (NORM) or per pattern and category (SAFE);
    // FP32: one int32 power-of-two exponent per pattern.
    const int64_t scale_slot = precision == PLK_FP32 ? (int64_t)c->nptn_pad * 2 : (int64_t)c->nptn_pad * (scaling ? ncat : 1);
    auto cleanup = [&](int rc) {
        plk_destroy(c);
        return rc;
    };
#define CREATE_TRY(expr)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return cleanup(fail(e_ == hipErrorOutOfMemory ? PLK_ERR_NOMEM : PLK_ERR_HIP,        \
                                std::string("plk_create: ") + #expr + ": " + hipGetErrorString(e_))); \
    } while (0)
    CREATE_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CREATE_TRY(hipMallocAsync((void **)&c->d_part, rs * slot * nslots, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_scale, 2 * scale_slot * nslots, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_states, (size_t)ntaxa * c->nptn_pad, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_freq, sizeof(double) * c->nptn_pad, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_invar, sizeof(double) * c->nptn_pad, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_theta, sizeof(double) * slot, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_buf_scale, sizeof(double) * c->nptn_pad, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_ptn_lh, sizeof(double) * c->nptn_pad, c->stream));
    // per-tile partial sums; a split launch (single-block units) has up to 4 per tile
    CREATE_TRY(hipMallocAsync((void **)&c->d_tile_sums, sizeof(double) * 2 * 4 * c->ntiles, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_result, sizeof(double) * 2, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_counter, sizeof(unsigned), c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_sink, kSinkBytes, c->stream));
    CREATE_TRY(hipMallocAsync((void **)&c->d_newton, sizeof(NewtonState), c->stream));
    c->cap_side = sizeof(double) * (size_t)c->nptn_pad * (ncat + 2);
    CREATE_TRY(hipMallocAsync((void **)&c->d_side, c->cap_side, c->stream));
    c->cap_part = rs * slot * nslots;
    c->cap_scale = 2 * scale_slot * nslots;
    c->cap_states = (size_t)ntaxa * c->nptn_pad;
    c->cap_pad = sizeof(double) * c->nptn_pad;
    c->cap_theta = sizeof(double) * slot;
    c->cap_tiles = sizeof(double) * 2 * 4 * c->ntiles;
    const int n = nstates;
    const size_t model_doubles = n + 2 * n * n + 2 * ncat + 2 * (size_t)(c->su + 1) * n + n + n * n;
    CREATE_TRY(hipMallocAsync((void **)&c->d_model, sizeof(double) * model_doubles, c->stream));
    c->d_eval = c->d_model;
    c->d_evec = c->d_eval + n;
    c->d_inv_evec = c->d_evec + n * n;
    c->d_rates = c->d_inv_evec + n * n;
    c->d_props = c->d_rates + ncat;
    c->d_tip = c->d_props + ncat;
    c->d_tipvec = c->d_tip + (size_t)(c->su + 1) * n;
    c->d_pi = c->d_tipvec + (size_t)(c->su + 1) * n;
    c->d_piU = c->d_pi + n;
    if (precision == PLK_FP64 && (n == 20 || n == 61))
        CREATE_TRY(hipMallocAsync((void **)&c->d_uinv_frag, sizeof(double) * mfma64_uinv_frag_size(n), c->stream));
    CREATE_TRY(hipHostMalloc((void **)&c->h_result, sizeof(double) * 2, hipHostMallocMapped | hipHostMallocCoherent));
    CREATE_TRY(hipHostGetDevicePointer((void **)&c->h_result_dev, c->h_result, 0));
    CREATE_TRY(hipHostMalloc((void **)&c->h_newton, sizeof(NewtonState), hipHostMallocDefault));
    for (int i = 0; i < 6; i++) CREATE_TRY(hipEventCreate(&c->ev[i]));
    // zero-initialised state: counters, scale counts and partials of unused slots
    CREATE_TRY(hipMemsetAsync(c->d_counter, 0, sizeof(unsigned), c->stream));
    CREATE_TRY(hipMemsetAsync(c->d_scale, 0, 2 * scale_slot * nslots, c->stream));
    CREATE_TRY(hipMemsetAsync(c->d_freq, 0, sizeof(double) * c->nptn_pad, c->stream));
    CREATE_TRY(hipMemsetAsync(c->d_invar, 0, sizeof(double) * c->nptn_pad, c->stream));
    CREATE_TRY(hipStreamSynchronize(c->stream));
    int rc = ensure_op_capacity(*c, std::max(16, ntaxa + 1));
    if (rc) return cleanup(rc);
#undef CREATE_TRY
    *out = c;
    return PLK_OK;
}

extern "C" int plk_destroy(plk_ctx *c) {
    if (!c) return PLK_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->d_part, c->d_scale, c->d_states, c->d_freq, c->d_invar, c->d_mats, c->d_model,
                    c->d_ops, c->d_jobs, c->d_theta, c->d_buf_scale, c->d_ptn_lh, c->d_tile_sums,
                    c->d_result, c->d_counter, c->d_mparams, c->d_mitems, c->d_mtile, c->d_mres,
                    c->d_mcount, c->d_newton, c->d_mnewton, c->d_sink, c->d_uinv_frag, c->d_side};
    for (void *b : bufs)
        if (b) (void)hipFreeAsync(b, c->stream);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->h_result) (void)hipHostFree(c->h_result);
    if (c->h_newton) (void)hipHostFree(c->h_newton);
    if (c->h_mnewton) (void)hipHostFree(c->h_mnewton);
    for (int i = 0; i < 6; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PLK_OK;
}

// Grow one stream-ordered buffer to at least `need` bytes (contents are not kept).
static int grow(Ctx &c, void **buf, size_t &cap, size_t need) {
    if (need <= cap) return PLK_OK;
    if (*buf) (void)hipFreeAsync(*buf, c.stream);
    *buf = nullptr;
    cap = 0;
    HIP_TRY(hipMallocAsync(buf, need, c.stream));
    cap = need;
    return PLK_OK;
}

extern "C" int plk_reshape(plk_ctx *c, int ntaxa, int nptn, int nslots) {
    clear_error();
    if (!c) return fail(PLK_ERR_ARG, "plk_reshape: null context");
    if (ntaxa < 3 || nptn < 1 || nslots < 1) return fail(PLK_ERR_ARG, "plk_reshape: need ntaxa >= 3, nptn >= 1, nslots >= 1");
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int pad = (int)round_up(nptn, 64), ntiles = pad / 64;
    const size_t rs = c->real_size();
    const int64_t slot = (int64_t)pad * c->block;
    const int64_t scale_slot = c->prec == PLK_FP32 ? (int64_t)pad * 2 : (int64_t)pad * (c->safe ? c->ncat : 1);
    int rc = grow(*c, (void **)&c->d_part, c->cap_part, rs * slot * nslots);
    if (!rc) rc = grow(*c, (void **)&c->d_scale, c->cap_scale, 2 * scale_slot * nslots);
    if (!rc) rc = grow(*c, (void **)&c->d_states, c->cap_states, (size_t)ntaxa * pad);
    if (!rc && sizeof(double) * pad > c->cap_pad) {
        size_t cf = c->cap_pad, ci = c->cap_pad, cb = c->cap_pad, cl = c->cap_pad;
        rc = grow(*c, (void **)&c->d_freq, cf, sizeof(double) * pad);
        if (!rc) rc = grow(*c, (void **)&c->d_invar, ci, sizeof(double) * pad);
        if (!rc) rc = grow(*c, (void **)&c->d_buf_scale, cb, sizeof(double) * pad);
        if (!rc) rc = grow(*c, (void **)&c->d_ptn_lh, cl, sizeof(double) * pad);
        if (!rc) c->cap_pad = sizeof(double) * pad;
    }
    if (!rc) rc = grow(*c, (void **)&c->d_theta, c->cap_theta, sizeof(double) * slot);
    if (!rc) rc = grow(*c, (void **)&c->d_tile_sums, c->cap_tiles, sizeof(double) * 2 * 4 * ntiles);
    if (!rc) rc = grow(*c, (void **)&c->d_side, c->cap_side, sizeof(double) * (size_t)pad * (c->ncat + 2));
    if (rc) return rc;
    if (ntiles != c->ntiles && c->mcap) {
        // the many-branch buffers hold (branch, tile) items: rebuilt on next use
        void *old[] = {c->d_mparams, c->d_mitems, c->d_mtile, c->d_mres, c->d_mcount, c->d_mnewton};
        for (void *q : old)
            if (q) (void)hipFreeAsync(q, c->stream);
        if (c->h_mnewton) (void)hipHostFree(c->h_mnewton);
        c->d_mparams = c->d_mitems = nullptr;
        c->d_mtile = c->d_mres = nullptr;
        c->d_mcount = nullptr;
        c->d_mnewton = c->h_mnewton = nullptr;
        c->mcap = 0;
    }
    // the op buffer also holds level bounds and up to ntaxa + 1 tip ids: rebuilt when the
    // taxon count grows
    if (ntaxa > c->ntaxa) c->ops_cap = c->mats_cap = 0;
    c->ntaxa = ntaxa;
    c->nptn = nptn;
    c->nptn_pad = pad;
    c->ntiles = ntiles;
    c->nslots = nslots;
    HIP_TRY(hipMemsetAsync(c->d_scale, 0, 2 * scale_slot * nslots, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_freq, 0, sizeof(double) * pad, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_invar, 0, sizeof(double) * pad, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    rc = ensure_op_capacity(*c, std::max(16, ntaxa + 1));
    if (rc) return rc;
    c->pending.clear();
    c->jobs.clear();
    c->host_freq.clear();
    c->patterns_set = false;
    c->theta_valid = false;
    c->ptn_lh_valid = false;
    return PLK_OK;
}

extern "C" int plk_set_patterns(plk_ctx *c, const uint8_t *states, const double *ptn_freq, const double *ptn_invar) {
    clear_error();
    if (!c || !states || !ptn_freq) return fail(PLK_ERR_ARG, "plk_set_patterns: null argument");
    if (c->su < 0) return fail(PLK_ERR_STATE, "plk_set_patterns: bad context");
    // padded patterns are STATE_UNKNOWN with zero weight (tree/phylokernelnew.h:1706-1708)
    std::vector<uint8_t> st((size_t)c->ntaxa * c->nptn_pad, (uint8_t)c->su);
    for (int t = 0; t < c->ntaxa; t++)
        for (int p = 0; p < c->nptn; p++) {
            uint8_t s = states[(size_t)t * c->nptn + p];
            if (s > c->su) return fail(PLK_ERR_ARG, "plk_set_patterns: state above STATE_UNKNOWN");
            st[(size_t)t * c->nptn_pad + p] = s;
        }
    std::vector<double> f(c->nptn_pad, 0.0), inv(c->nptn_pad, 0.0);
    std::copy(ptn_freq, ptn_freq + c->nptn, f.begin());
    if (ptn_invar) std::copy(ptn_invar, ptn_invar + c->nptn, inv.begin());
    HIP_TRY(hipMemcpyAsync(c->d_states, st.data(), st.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_freq, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_invar, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->host_freq = f;
    c->patterns_set = true;
    c->theta_valid = false;
    return PLK_OK;
}

extern "C" int plk_set_model(plk_ctx *c, const double *eval, const double *evec, const double *inv_evec,
                             const double *rates, const double *props, int state_unknown, const double *tip_lh) {
    clear_error();
    if (!c || !eval || !evec || !inv_evec || !rates || !props || !tip_lh)
        return fail(PLK_ERR_ARG, "plk_set_model: null argument");
    if (state_unknown != c->su) return fail(PLK_ERR_ARG, "plk_set_model: state_unknown does not match the state count");
    const int n = c->n, ncat = c->ncat;
    std::vector<double> m;
    m.insert(m.end(), eval, eval + n);
    m.insert(m.end(), evec, evec + n * n);
    m.insert(m.end(), inv_evec, inv_evec + n * n);
    m.insert(m.end(), rates, rates + ncat);
    m.insert(m.end(), props, props + ncat);
    m.insert(m.end(), tip_lh, tip_lh + (size_t)(c->su + 1) * n);
    // FP32-form extras.  Real tip vectors U * tip_lh; pi is the eigenvector row of the
    // zero eigenvalue of U^-1 (U^-1 = V^T Pi^1/2 with V[:,0] = Pi^1/2), normalised;
    // piU[i][x] = pi_x U[x][i] maps a real-space partial to eigen space.
    std::vector<double> tipvec((size_t)(c->su + 1) * n, 0.0), pi(n, 0.0), piU((size_t)n * n, 0.0);
    for (int s = 0; s <= c->su; s++)
        for (int x = 0; x < n; x++) {
            double v = 0.0;
            for (int i = 0; i < n; i++) v += evec[x * n + i] * tip_lh[s * n + i];
            tipvec[(size_t)s * n + x] = v;
        }
    // the stationary row: among the (near-)zero eigenvalues, the U^-1 row with the most
    // non-zero entries (states of zero frequency also have eigenvalue 0, with a unit row)
    double emax = 1.0;
    for (int i = 0; i < n; i++) emax = std::max(emax, std::fabs(eval[i]));
    int i0 = -1, best_nz = -1;
    for (int i = 0; i < n; i++) {
        if (std::fabs(eval[i]) > 1e-8 * emax) continue;
        int nzc = 0;
        for (int x = 0; x < n; x++) nzc += inv_evec[i * n + x] != 0.0;
        if (nzc > best_nz || (nzc == best_nz && std::fabs(eval[i]) < std::fabs(eval[i0]))) {
            best_nz = nzc;
            i0 = i;
        }
    }
    if (i0 < 0) return fail(PLK_ERR_ARG, "plk_set_model: no zero eigenvalue");
    double psum = 0.0;
    for (int x = 0; x < n; x++) psum += inv_evec[i0 * n + x];
    if (!(std::fabs(psum) > 0.0)) return fail(PLK_ERR_ARG, "plk_set_model: no stationary eigenvector");
    for (int x = 0; x < n; x++) pi[x] = inv_evec[i0 * n + x] / psum;
    for (int i = 0; i < n; i++)
        for (int x = 0; x < n; x++) piU[(size_t)i * n + x] = pi[x] * evec[x * n + i];
    m.insert(m.end(), tipvec.begin(), tipvec.end());
    m.insert(m.end(), pi.begin(), pi.end());
    m.insert(m.end(), piU.begin(), piU.end());
    for (double v : m)
        if (!std::isfinite(v)) return fail(PLK_ERR_ARG, "plk_set_model: non-finite model value");
    HIP_TRY(hipMemcpyAsync(c->d_model, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, c->stream));
    std::vector<double> uf;
    if (c->d_uinv_frag) {
        // [rb][ks][lane] = U^-1[16 rb + lane % 16][4 ks + lane / 16], zero outside n x n
        const int ks_n = (n + 3) / 4;
        uf.assign(mfma64_uinv_frag_size(n), 0.0);
        for (size_t f = 0; f < uf.size(); f++) {
            const int lane = (int)(f % 64), ks = (int)(f / 64) % ks_n, rb = (int)(f / 64) / ks_n;
            const int j = 16 * rb + lane % 16, x = 4 * ks + lane / 16;
            if (j < n && x < n) uf[f] = inv_evec[j * n + x];
        }
        HIP_TRY(hipMemcpyAsync(c->d_uinv_frag, uf.data(), sizeof(double) * uf.size(), hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->model_set = true;
    c->theta_valid = false;
    return PLK_OK;
}

extern "C" int plk_update_partials(plk_ctx *c, const plk_op *ops, int nops) {
    clear_error();
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nops < 0 || (nops > 0 && !ops)) return fail(PLK_ERR_ARG, "plk_update_partials: bad op list");
    for (int i = 0; i < nops; i++) {
        const plk_op &o = ops[i];
        if (o.out_slot < 0 || o.out_slot >= c->nslots || !valid_ref(*c, o.child[0]) || !valid_ref(*c, o.child[1]) ||
            o.child[0] == o.out_slot || o.child[1] == o.out_slot || o.reserved != 0 ||
            !std::isfinite(o.len[0]) || !std::isfinite(o.len[1]) || o.len[0] < 0 || o.len[1] < 0)
            return fail(PLK_ERR_ARG, "plk_update_partials: invalid op " + std::to_string(i));
        if (o.child[0] >= 0 && o.child[0] == o.child[1])
            return fail(PLK_ERR_ARG, "plk_update_partials: both children use one slot (op " + std::to_string(i) + ")");
        int k = (int)c->pending.size();
        DevOp d{};
        d.out_slot = o.out_slot;
        bool t0 = o.child[0] < 0, t1 = o.child[1] < 0;
        int side_of0 = 0, side_of1 = 1;
        if (t0 && t1) {
            d.kind = KIND_TT;
            d.left = -o.child[0] - 1;
            d.right = -o.child[1] - 1;
        } else if (t0 || t1) {
            // the tip is always "left" (tree/phylokernelnew.h:1399-1409)
            d.kind = KIND_TI;
            d.left = t0 ? -o.child[0] - 1 : -o.child[1] - 1;
            d.right = t0 ? o.child[1] : o.child[0];
            if (t1) {
                side_of0 = 1;
                side_of1 = 0;
            }
        } else {
            d.kind = KIND_II;
            d.left = o.child[0];
            d.right = o.child[1];
            // the previous op's output goes right: the 20-state FP32 traversal keeps it in
            // registers for that side (the two factors of the product commute exactly)
            if (k > 0 && c->pending[k - 1].out_slot == o.child[0]) {
                std::swap(d.left, d.right);
                side_of0 = 1;
                side_of1 = 0;
            }
        }
        d.len_left = side_of0 == 0 ? o.len[0] : o.len[1];
        d.len_right = side_of0 == 0 ? o.len[1] : o.len[0];
        c->pending.push_back(d);
        c->jobs.push_back(PJob{k, side_of0, t0 ? 1 : 0, 0, o.len[0]});
        c->jobs.push_back(PJob{k, side_of1, t1 ? 1 : 0, 0, o.len[1]});
    }
    c->theta_valid = false;
    return PLK_OK;
}

extern "C" int plk_flush(plk_ctx *c) {
    clear_error();
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->pending.empty()) return PLK_OK;
    KParams p = make_params(*c);
    bool used[3] = {false, false, false};
    rc = run_traversal(*c, p, MODE_PARTIALS, used);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    timing_collect(*c, used);
    return PLK_OK;
}

static int branch_refs(Ctx &c, int &a, int &b) {
    if (!valid_ref(c, a) || !valid_ref(c, b)) return fail(PLK_ERR_ARG, "branch: node reference out of range");
    if (a < 0 && b < 0) return fail(PLK_ERR_ARG, "branch: both ends are tips");
    if (b < 0) std::swap(a, b);
    if (a == b) return fail(PLK_ERR_ARG, "branch: both ends use one slot");
    return PLK_OK;
}

// sum of f * fixed pattern values (host fallback of tree/phylokernelnew.h:3167-3184)
static double fixup_nonfinite(Ctx &c, std::vector<double> &pl) {
    double s = 0.0;
    for (int p = 0; p < c.nptn; p++) {
        if (!std::isfinite(pl[p])) pl[p] = kLogScalingThreshold * 4;
        s += pl[p] * c.host_freq[p];
    }
    return s;
}

extern "C" int plk_branch_loglik(plk_ctx *c, int a, int b, double t, double *logl, double *ptn_lh) {
    clear_error();
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!logl || !std::isfinite(t) || t < 0) return fail(PLK_ERR_ARG, "plk_branch_loglik: bad arguments");
    rc = branch_refs(*c, a, b);
    if (rc) return rc;
    KParams p = make_params(*c);
    p.br_a = a;
    p.br_b = b;
    p.br_len = t;
    // the logL goes straight to the mapped host result (no copy after the launch)
    p.result = c->h_result_dev;
    bool used[3] = {false, false, false};
    rc = run_traversal(*c, p, MODE_LOGL, used);
    if (rc) return rc;
    c->ptn_lh_valid = true;
    HIP_TRY(hipStreamSynchronize(c->stream));
    timing_collect(*c, used);
    double v = c->h_result[0];
    if (ptn_lh || !std::isfinite(v)) {
        std::vector<double> pl(c->nptn);
        HIP_TRY(hipMemcpy(pl.data(), c->d_ptn_lh, sizeof(double) * c->nptn, hipMemcpyDeviceToHost));
        if (!std::isfinite(v)) v = fixup_nonfinite(*c, pl);
        if (ptn_lh) std::copy(pl.begin(), pl.end(), ptn_lh);
    }
    *logl = v;
    return PLK_OK;
}

extern "C" int plk_branch_derivs(plk_ctx *c, int a, int b, double t, int reuse_theta, double *df, double *ddf) {
    clear_error();
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!df || !ddf || !std::isfinite(t) || t < 0) return fail(PLK_ERR_ARG, "plk_branch_derivs: bad arguments");
    rc = branch_refs(*c, a, b);
    if (rc) return rc;
    KParams p = make_params(*c);
    p.br_a = a;
    p.br_b = b;
    p.br_len = t;
    bool used[3] = {false, false, false};
    const bool reuse = reuse_theta && c->theta_valid && c->theta_a == a && c->theta_b == b && c->pending.empty();
    if (reuse) {
        timing_begin(*c, 2);
        HIP_TRY(f64_theta_eval(p, 1, c->stream));
        timing_end(*c, 2);
        used[2] = true;
    } else {
        rc = run_traversal(*c, p, MODE_THETA, used);
        if (rc) return rc;
        c->theta_valid = true;
        c->theta_a = a;
        c->theta_b = b;
    }
    HIP_TRY(hipMemcpyAsync(c->h_result, c->d_result, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    timing_collect(*c, used);
    double d1 = c->h_result[0], d2 = c->h_result[1];
    if (!std::isfinite(d1)) {
        *df = *ddf = 0.0;
        // NORM: the reference stops with "use -safe" (2594-2596); SAFE warns and zeroes (2646-2649)
        if (!c->safe) return fail(PLK_ERR_NUMERIC, "Numerical underflow (lh-derivative): use SAFE scaling");
        return PLK_OK;
    }
    *df = d1;
    *ddf = d2;
    return PLK_OK;
}

extern "C" int plk_loglik_from_theta(plk_ctx *c, double t, double *logl) {
    clear_error();
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!logl || !std::isfinite(t) || t < 0) return fail(PLK_ERR_ARG, "plk_loglik_from_theta: bad arguments");
    if (!c->theta_valid) return fail(PLK_ERR_STATE, "plk_loglik_from_theta: no theta buffer (call plk_branch_derivs)");
    KParams p = make_params(*c);
    p.br_len = t;
    HIP_TRY(f64_theta_eval(p, 0, c->stream));
    c->ptn_lh_valid = true;
    HIP_TRY(hipMemcpyAsync(c->h_result, c->d_result, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double v = c->h_result[0];
    if (!std::isfinite(v)) {
        if (!c->safe) return fail(PLK_ERR_NUMERIC, "Numerical underflow (lh-from-buffer): use SAFE scaling");
        std::vector<double> pl(c->nptn);
        HIP_TRY(hipMemcpy(pl.data(), c->d_ptn_lh, sizeof(double) * c->nptn, hipMemcpyDeviceToHost));
        v = fixup_nonfinite(*c, pl);
    }
    *logl = v;
    return PLK_OK;
}

extern "C" int plk_newton_branch(plk_ctx *c, int a, int b, double x1, double xguess, double x2, double xacc,
                                 int max_nr, double *out, int *evals) {
    clear_error();
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!out || !std::isfinite(x1) || !std::isfinite(x2) || !std::isfinite(xguess) || x1 < 0 || x2 < x1 ||
        !(xacc > 0) || max_nr < 0)
        return fail(PLK_ERR_ARG, "plk_newton_branch: bad arguments");
    rc = branch_refs(*c, a, b);
    if (rc) return rc;
    NewtonState &h = *c->h_newton;
    h = NewtonState{};
    h.x1 = x1;
    h.x2 = x2;
    h.xacc = xacc;
    h.rts = std::min(std::max(xguess, x1), x2);
    h.max_nr = max_nr;
    h.safe = c->safe;
    HIP_TRY(hipMemcpyAsync(c->d_newton, &h, sizeof(NewtonState), hipMemcpyHostToDevice, c->stream));
    KParams p = make_params(*c);
    p.br_a = a;
    p.br_b = b;
    p.br_len = h.rts;
    bool used[3] = {false, false, false};
    // evaluation 0 builds theta (computeFuncDerv's first call never reuses it)
    rc = run_traversal(*c, p, MODE_THETA, used);
    if (rc) return rc;
    c->theta_valid = true;
    c->theta_a = a;
    c->theta_b = b;
    HIP_TRY(newton_first(c->d_newton, c->d_result, c->stream));
    // queue iterations in growing rounds; a finished solve turns the rest into no-ops
    // (iteration j evaluates at most once and iteration max_nr never does, so max_nr
    // launches always suffice)
    int round = std::min(6, std::max(max_nr, 1)), launched = 0;
    for (;;) {
        for (int k = 0; k < round; k++) HIP_TRY(f64_theta_newton(p, c->d_newton, c->stream));
        launched += round;
        HIP_TRY(hipMemcpyAsync(&h, c->d_newton, sizeof(NewtonState), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (h.done) break;
        if (launched > max_nr) return fail(PLK_ERR_STATE, "plk_newton_branch: solve did not finish");
        round = std::min(2 * round, 48);
    }
    timing_collect(*c, used);
    if (evals) *evals = h.evals;
    switch (h.status) {
        case NEWTON_OK: *out = h.out; return PLK_OK;
        case NEWTON_UNDERFLOW: return fail(PLK_ERR_NUMERIC, "Numerical underflow (lh-derivative): use SAFE scaling");
        case NEWTON_NONFINITE: return fail(PLK_ERR_NUMERIC, "Wrong computeFuncDerv");
        default: return fail(PLK_ERR_NUMERIC, "Maximum number of iterations exceeded in minimizeNewton");
    }
}

// Buffers of the many-branch launches: per-branch parameters, (branch, tile) items,
// 3 reduction values per tile and per branch, counters and Newton states.
static int ensure_multi(Ctx &c, int nbr) {
    if (nbr <= c.mcap) return PLK_OK;
    void *old[] = {c.d_mparams, c.d_mitems, c.d_mtile, c.d_mres, c.d_mcount, c.d_mnewton};
    for (void *q : old)
        if (q) (void)hipFreeAsync(q, c.stream);
    if (c.h_mnewton) (void)hipHostFree(c.h_mnewton);
    c.d_mparams = c.d_mitems = nullptr;
    c.d_mtile = c.d_mres = nullptr;
    c.d_mcount = nullptr;
    c.d_mnewton = c.h_mnewton = nullptr;
    c.mcap = 0;
    const int cap = std::max(nbr, 2 * c.ntaxa);
    HIP_TRY(hipMallocAsync(&c.d_mparams, sizeof(KParams) * cap, c.stream));
    HIP_TRY(hipMallocAsync(&c.d_mitems, sizeof(BatchItem) * (size_t)cap * c.ntiles, c.stream));
    HIP_TRY(hipMallocAsync((void **)&c.d_mtile, sizeof(double) * 3 * (size_t)cap * c.ntiles, c.stream));
    HIP_TRY(hipMallocAsync((void **)&c.d_mres, sizeof(double) * 3 * cap, c.stream));
    HIP_TRY(hipMallocAsync((void **)&c.d_mcount, sizeof(unsigned) * cap, c.stream));
    HIP_TRY(hipMallocAsync((void **)&c.d_mnewton, sizeof(NewtonState) * cap, c.stream));
    HIP_TRY(hipHostMalloc((void **)&c.h_mnewton, sizeof(NewtonState) * cap, hipHostMallocDefault));
    HIP_TRY(hipMemsetAsync(c.d_mcount, 0, sizeof(unsigned) * cap, c.stream));
    std::vector<BatchItem> items((size_t)cap * c.ntiles);
    for (int i = 0; i < cap; i++)
        for (int tt = 0; tt < c.ntiles; tt++) items[(size_t)i * c.ntiles + tt] = BatchItem{i, tt};
    HIP_TRY(hipMemcpyAsync(c.d_mitems, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    c.mcap = cap;
    return PLK_OK;
}

extern "C" int plk_newton_multi(plk_ctx *c, int nbr, const int *a, const int *b, double x1, const double *xguess,
                                double x2, double xacc, int max_nr, double *out, double *logl, int *status) {
    clear_error();
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nbr < 1 || !a || !b || !xguess || !out || !std::isfinite(x1) || !std::isfinite(x2) || x1 < 0 || x2 < x1 ||
        !(xacc > 0) || max_nr < -1)
        return fail(PLK_ERR_ARG, "plk_newton_multi: bad arguments");
    const bool eval_only = max_nr == -1;
    std::vector<int> ra(a, a + nbr), rb(b, b + nbr);
    for (int i = 0; i < nbr; i++) {
        rc = branch_refs(*c, ra[i], rb[i]);
        if (rc) return rc;
        if (!std::isfinite(xguess[i])) return fail(PLK_ERR_ARG, "plk_newton_multi: bad initial length");
    }
    rc = plk_flush(c);
    if (rc) return rc;
    rc = ensure_multi(*c, nbr);
    if (rc) return rc;
    KParams base = make_params(*c);
    base.nops = 0;
    base.theta = nullptr;
    base.buf_scale = nullptr;
    base.g_frac = nullptr;  // per-pattern derivative output: single-branch calls only
    std::vector<KParams> ps(nbr, base);
    NewtonState *h = c->h_mnewton;
    for (int i = 0; i < nbr; i++) {
        ps[i].br_a = ra[i];
        ps[i].br_b = rb[i];
        ps[i].tile_sums = c->d_mtile + (size_t)3 * c->ntiles * i;
        ps[i].counter = c->d_mcount + i;
        ps[i].result = c->d_mres + 3 * i;
        h[i] = NewtonState{};
        h[i].x1 = x1;
        h[i].x2 = x2;
        h[i].xacc = xacc;
        h[i].rts = std::min(std::max(xguess[i], x1), x2);
        h[i].max_nr = std::max(max_nr, 0);
        h[i].safe = c->safe;
        h[i].eval_only = eval_only ? 1 : 0;
    }
    rc = alloc_stage(*c, sizeof(KParams) * nbr);
    if (rc) return rc;
    std::memcpy(c->h_stage, ps.data(), sizeof(KParams) * nbr);
    HIP_TRY(hipMemcpyAsync(c->d_mparams, c->h_stage, sizeof(KParams) * nbr, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_mnewton, h, sizeof(NewtonState) * nbr, hipMemcpyHostToDevice, c->stream));
    const int ncw = std::min(c->ncat, 4);
    auto run_rounds = [&](int limit) -> int {
        // evaluation 0 plus up to max_nr iterations; rounds grow until every branch is done
        int round = std::min(6, std::max(limit, 1)), launched = 0;
        for (;;) {
            for (int k = 0; k < round; k++)
                HIP_TRY(multi_newton(c->prec, ncw, (const KParams *)c->d_mparams, (const BatchItem *)c->d_mitems,
                                     nbr * c->ntiles, c->d_mnewton, c->stream));
            launched += round;
            HIP_TRY(hipMemcpyAsync(h, c->d_mnewton, sizeof(NewtonState) * nbr, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            bool all = true;
            for (int i = 0; i < nbr && all; i++) all = h[i].done != 0;
            if (all) return PLK_OK;
            if (launched > limit + 1) return fail(PLK_ERR_STATE, "plk_newton_multi: solves did not finish");
            round = std::min(2 * round, 48);
        }
    };
    rc = run_rounds(eval_only ? 1 : max_nr);
    if (rc) return rc;
    if (eval_only) {
        for (int i = 0; i < nbr; i++) {
            out[i] = h[i].rts;
            if (logl) logl[i] = h[i].logl_eval;
            if (status) status[i] = PLK_OK;
        }
        c->theta_valid = false;
        return PLK_OK;
    }
    // the logL at the answer is the last evaluation's unless bisection ended on an older
    // bracket point: evaluate those once more
    int more = 0;
    for (int i = 0; i < nbr; i++)
        if (h[i].status == NEWTON_OK && h[i].out != h[i].x_eval) {
            h[i].rts = h[i].out;
            h[i].eval_only = 1;
            h[i].done = 0;
            more++;
        }
    if (more) {
        HIP_TRY(hipMemcpyAsync(c->d_mnewton, h, sizeof(NewtonState) * nbr, hipMemcpyHostToDevice, c->stream));
        rc = run_rounds(1);
        if (rc) return rc;
    }
    int first_err = PLK_OK;
    for (int i = 0; i < nbr; i++) {
        int st = h[i].status == NEWTON_OK ? PLK_OK : PLK_ERR_NUMERIC;
        out[i] = st == PLK_OK ? h[i].out : h[i].rts;
        if (logl) logl[i] = h[i].logl_eval;
        if (status) status[i] = st;
        if (st != PLK_OK && first_err == PLK_OK) first_err = st;
    }
    c->theta_valid = false;
    if (first_err != PLK_OK && !status)
        return fail(first_err, "plk_newton_multi: a solve failed (non-finite derivative or iteration limit)");
    return PLK_OK;
}

extern "C" int plk_get_partial(plk_ctx *c, int slot, double *partial, uint16_t *scale) {
    clear_error();
    if (!c || slot < 0 || slot >= c->nslots) return fail(PLK_ERR_ARG, "plk_get_partial: bad slot");
    int rc = plk_flush(c);
    if (rc) return rc;
    const int64_t slot_len = (int64_t)c->nptn_pad * c->block;
    if (partial) {
        std::vector<double> raw(slot_len);
        if (c->prec == PLK_FP32) {
            std::vector<float> rf(slot_len);
            HIP_TRY(hipMemcpy(rf.data(), (char *)c->d_part + 4 * slot_len * slot, 4 * slot_len, hipMemcpyDeviceToHost));
            std::copy(rf.begin(), rf.end(), raw.begin());
        } else {
            HIP_TRY(hipMemcpy(raw.data(), (char *)c->d_part + 8 * slot_len * slot, 8 * slot_len, hipMemcpyDeviceToHost));
        }
        const bool f32 = c->prec == PLK_FP32;
        for (int p = 0; p < c->nptn; p++)
            for (int i = 0; i < c->block; i++) {
                const int cc = i / c->n, x = i % c->n;
                const int64_t base = ((int64_t)(p / 64) * c->block + (int64_t)cc * c->n) * 64;
                partial[(int64_t)p * c->block + i] =
                    f32 ? raw[base + f32_xoff(c->n, x) + f32_loff(c->n, p % 64)] : raw[base + (int64_t)x * 64 + p % 64];
            }
    }
    if (scale && c->prec == PLK_FP32) return fail(PLK_ERR_ARG, "plk_get_partial: the FP32 form keeps exponents (plk_get_exponents)");
    if (scale) {
        const int64_t ss = (int64_t)c->nptn_pad * (c->safe ? c->ncat : 1);
        std::vector<uint16_t> raw(ss);
        HIP_TRY(hipMemcpy(raw.data(), (char *)c->d_scale + 2 * ss * slot, 2 * ss, hipMemcpyDeviceToHost));
        std::copy(raw.begin(), raw.begin() + (int64_t)c->nptn * (c->safe ? c->ncat : 1), scale);
    }
    return PLK_OK;
}

extern "C" int plk_set_timing(plk_ctx *c, int enable) {
    if (!c) return fail(PLK_ERR_ARG, "null context");
    c->timing = enable != 0;
    for (int k = 0; k < 3; k++) {
        c->t_total[k] = 0;
        c->t_count[k] = 0;
    }
    return PLK_OK;
}

extern "C" int plk_get_timing(plk_ctx *c, int kind, double *total_ms, int64_t *launches) {
    if (!c || kind < 0 || kind > 2) return fail(PLK_ERR_ARG, "plk_get_timing: bad arguments");
    if (total_ms) *total_ms = c->t_total[kind];
    if (launches) *launches = c->t_count[kind];
    return PLK_OK;
}

extern "C" int64_t plk_updates(plk_ctx *c) { return c ? c->updates : 0; }

extern "C" int plk_synchronize(plk_ctx *c) {
    if (!c) return fail(PLK_ERR_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PLK_OK;
}

extern "C" int plk_get_exponents(plk_ctx *c, int slot, int32_t *exponents) {
    clear_error();
    if (!c || slot < 0 || slot >= c->nslots || !exponents) return fail(PLK_ERR_ARG, "plk_get_exponents: bad arguments");
    if (c->prec != PLK_FP32) return fail(PLK_ERR_ARG, "plk_get_exponents: only the FP32 form keeps exponents");
    int rc = plk_flush(c);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(exponents, (char *)c->d_scale + (int64_t)4 * c->nptn_pad * slot, sizeof(int32_t) * c->nptn,
                      hipMemcpyDeviceToHost));
    return PLK_OK;
}

extern "C" int plk_set_tile_shape(plk_ctx *c, int waves_per_category) {
    if (!c || (waves_per_category != 0 && waves_per_category != 1 && waves_per_category != 2 && waves_per_category != 4))
        return fail(PLK_ERR_ARG, "plk_set_tile_shape: shape must be 0 (pattern-block waves) or 1, 2, 4 waves per category");
    if (waves_per_category == 0 && c->ncat > 4)
        return fail(PLK_ERR_ARG, "plk_set_tile_shape: pattern-block waves need ncat <= 4");
    if (c->ncat * waves_per_category > 16) return fail(PLK_ERR_ARG, "plk_set_tile_shape: more than 16 waves per tile");
    c->wpc = waves_per_category;
    return PLK_OK;
}

static int copy_side(plk_ctx *c, size_t off, size_t count, double *out) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(out, c->d_side + off, sizeof(double) * count, hipMemcpyDeviceToHost));
    return PLK_OK;
}

extern "C" int plk_get_pattern_lh_cat(plk_ctx *c, double *out) {
    clear_error();
    if (!c || !out) return fail(PLK_ERR_ARG, "plk_get_pattern_lh_cat: null argument");
    if (c->prec != PLK_FP64) return fail(PLK_ERR_STATE, "plk_get_pattern_lh_cat: FP64 form only");
    if (!c->ptn_lh_valid) return fail(PLK_ERR_STATE, "plk_get_pattern_lh_cat: no logL evaluated yet");
    return copy_side(c, 0, (size_t)c->nptn * c->ncat, out);  // [pattern][category]
}

extern "C" int plk_get_pattern_scaling(plk_ctx *c, double *lh, double *scaling) {
    clear_error();
    if (!c || !lh || !scaling) return fail(PLK_ERR_ARG, "plk_get_pattern_scaling: null argument");
    if (c->prec != PLK_FP64) return fail(PLK_ERR_STATE, "plk_get_pattern_scaling: FP64 form only");
    if (!c->ptn_lh_valid) return fail(PLK_ERR_STATE, "plk_get_pattern_scaling: no logL evaluated yet");
    int rc = copy_side(c, (size_t)c->nptn_pad * c->ncat, c->nptn, scaling);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(lh, c->d_ptn_lh, sizeof(double) * c->nptn, hipMemcpyDeviceToHost));
    // lh = |sum| + invar, recovered from the stored log (one rounding of exp(log))
    for (int p = 0; p < c->nptn; p++) lh[p] = std::exp(lh[p] - scaling[p]);
    return PLK_OK;
}

extern "C" int plk_get_pattern_dlogl(plk_ctx *c, double *out) {
    clear_error();
    if (!c || !out) return fail(PLK_ERR_ARG, "plk_get_pattern_dlogl: null argument");
    return copy_side(c, (size_t)c->nptn_pad * (c->ncat + 1), c->nptn, out);
}

extern "C" int plk_set_f64_form(plk_ctx *c, int form) {
    if (!c || form < -1 || form > 1) return fail(PLK_ERR_ARG, "plk_set_f64_form: form must be -1, 0 or 1");
    c->f64_form = form;
    return PLK_OK;
}

extern "C" int plk_set_level_waves(plk_ctx *c, int waves) {
    if (!c || waves < -1 || waves > 4) return fail(PLK_ERR_ARG, "plk_set_level_waves: slots must be in [-1, 4]");
    c->small_waves = waves;
    return PLK_OK;
}

// Device-to-device copy probe: measured HBM bandwidth (bytes read + written per second)
// for roofline context, and a copy of known size for calibrating traffic counters.
extern "C" int plk_probe_copy(int64_t bytes, int reps, double *gbps) {
    clear_error();
    if (bytes < 1024 || reps < 1 || !gbps) return fail(PLK_ERR_ARG, "plk_probe_copy: bad arguments");
    hipStream_t s = nullptr;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int rc = PLK_OK;
    float best = 1e30f;
    if (hipMallocAsync(&a, bytes, s) != hipSuccess || hipMallocAsync(&b, bytes, s) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemsetAsync(a, 1, bytes, s) != hipSuccess) {
        rc = fail(PLK_ERR_HIP, "plk_probe_copy: allocation failed");
    } else {
        for (int r = 0; r <= reps && rc == PLK_OK; r++) {
            float ms = 0.0f;
            if (hipEventRecord(e0, s) != hipSuccess || hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                rc = fail(PLK_ERR_HIP, "plk_probe_copy: copy failed");
            else if (r > 0)
                best = std::min(best, ms);
        }
    }
    if (a) (void)hipFreeAsync(a, s);
    if (b) (void)hipFreeAsync(b, s);
    (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (rc == PLK_OK) *gbps = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
    return rc;
}

// computeLikelihoodDerv for many branches at once (the per-branch gradient_vector and
// hessian_diagonal of tree/phylokernelnew.h:2567-2568): queued ops first, then one launch
// over (branch, tile) items.  Theta is not kept (plk_loglik_from_theta needs a prior
// plk_branch_derivs).
extern "C" int plk_branch_derivs_multi(plk_ctx *c, int nbr, const int *a, const int *b, const double *t, double *df,
                                       double *ddf) {
    clear_error();
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nbr < 1 || !a || !b || !t || !df || !ddf) return fail(PLK_ERR_ARG, "plk_branch_derivs_multi: bad arguments");
    std::vector<int> ra(a, a + nbr), rb(b, b + nbr);
    for (int i = 0; i < nbr; i++) {
        rc = branch_refs(*c, ra[i], rb[i]);
        if (rc) return rc;
        if (!std::isfinite(t[i]) || t[i] < 0) return fail(PLK_ERR_ARG, "plk_branch_derivs_multi: bad length");
    }
    rc = plk_flush(c);
    if (rc) return rc;
    rc = ensure_multi(*c, nbr);
    if (rc) return rc;
    KParams base = make_params(*c);
    base.nops = 0;
    base.theta = nullptr;
    base.buf_scale = nullptr;
    base.g_frac = nullptr;  // per-pattern derivative output: single-branch calls only
    std::vector<KParams> ps(nbr, base);
    for (int i = 0; i < nbr; i++) {
        ps[i].br_a = ra[i];
        ps[i].br_b = rb[i];
        ps[i].br_len = t[i];
        ps[i].tile_sums = c->d_mtile + (size_t)2 * c->ntiles * i;
        ps[i].counter = c->d_mcount + i;
        ps[i].result = c->d_mres + 2 * i;
    }
    rc = alloc_stage(*c, sizeof(KParams) * nbr + sizeof(double) * 2 * nbr);
    if (rc) return rc;
    std::memcpy(c->h_stage, ps.data(), sizeof(KParams) * nbr);
    double *hres = (double *)((char *)c->h_stage + sizeof(KParams) * nbr);
    HIP_TRY(hipMemcpyAsync(c->d_mparams, c->h_stage, sizeof(KParams) * nbr, hipMemcpyHostToDevice, c->stream));
    timing_begin(*c, 2);
    HIP_TRY(multi_derivs(c->prec, std::min(c->ncat, 4), (const KParams *)c->d_mparams, (const BatchItem *)c->d_mitems,
                         nbr * c->ntiles, c->stream));
    timing_end(*c, 2);
    HIP_TRY(hipMemcpyAsync(hres, c->d_mres, sizeof(double) * 2 * nbr, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const bool used[3] = {false, false, true};
    timing_collect(*c, used);
    int status = PLK_OK;
    for (int i = 0; i < nbr; i++) {
        df[i] = hres[2 * i];
        ddf[i] = hres[2 * i + 1];
        if (!std::isfinite(df[i])) {
            df[i] = ddf[i] = 0.0;
            if (!c->safe) status = fail(PLK_ERR_NUMERIC, "Numerical underflow (lh-derivative): use SAFE scaling");
        }
    }
    c->theta_valid = false;
    return status;
}
