// Shared host/device declarations of the likelihood engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "plk.h"
#include "plk_internal.h"

// Largest category count the GPU kernels hold per pattern in LDS.
#define PLK_GPU_MAX_CATS 8

// FP32-form matrix-core tiling: 16-row x blocks and 4-deep k steps of a state count.
constexpr int f32_xblocks(int n) { return (n + 15) / 16; }
constexpr int f32_ksteps(int n) { return (n + 3) / 4; }
// FP32 leaf-table row length: the matrix-core forms pad a row to whole 16-row blocks
// (zeros past n), so a lane's four output rows are one aligned 16-byte load
constexpr int f32_trow(int n) { return n == 4 ? 4 : f32_xblocks(n) * 16; }
// FP64 matrix-core traversal: U^-1 as A fragments [16-row block][4-state k step][lane]
constexpr int mfma64_uinv_frag_size(int n) { return ((n + 15) / 16) * ((n + 3) / 4) * 64; }
// FP32 partial layout inside one (tile, category) block of 64 patterns.  20 states:
// row groups of 4 states, element (x, p) at ((x / 4) * 64 + p) * 4 + x % 4, so the four
// matrix-core output rows a lane holds (and the four B-operand rows it feeds) are one
// 16-byte access; other state counts: element (x, p) at x * 64 + p.
constexpr bool f32_rg(int n) { return n == 20; }
__host__ __device__ constexpr int f32_xoff(int n, int x) { return f32_rg(n) ? (x >> 2) * 256 + (x & 3) : x * 64; }
__host__ __device__ constexpr int f32_loff(int n, int p) { return f32_rg(n) ? p * 4 : p; }
// A-fragment k-order of the 20-state form: k-step ks < 4 of lane group g reads state
// 4g + ks (row group g, element ks), k-step 4 reads state 16 + g (row group 4)
__host__ __device__ constexpr int f32_kstate(int n, int ks, int g) {
    return f32_rg(n) ? (ks < 4 ? 4 * g + ks : 16 + g) : ks * 4 + g;
}
// Matrix-core output row i (0-15) of x block xb holds state f32_orow(n, xb, i), or -1 for
// a zero row.  20 states: block 1 holds states 16-19 in rows 0, 4, 8, 12, so lane group g
// (output rows 4g..4g+3) has state 16 + g in its first row, which is the entry that lane
// feeds to the next op's k-step 4.  Leaf-table rows follow the same order (row r of a
// table = output row r % 16 of block r / 16).
__host__ __device__ constexpr int f32_orow(int n, int xb, int i) {
    return f32_rg(n) ? (xb == 0 ? i : ((i & 3) == 0 ? 16 + (i >> 2) : -1)) : (xb * 16 + i < n ? xb * 16 + i : -1);
}

namespace plk {

// Device-side op: one directional partial update.  kind: 0 tip-tip, 1 tip-internal
// (left is the tip), 2 internal-internal (tree/phylokernelnew.h:1399-1409, 1636, 1752, 1911).
struct DevOp {
    int32_t out_slot;
    int32_t kind;
    int32_t left;   // taxon (kind 0/1) or slot (kind 2)
    int32_t right;  // taxon (kind 0) or slot (kind 1/2)
    // level-parallel 4-state kernel: LDS buffer of the output / of an internal child
    // (-1: HBM only), or for a tip child its index among the tips staged in LDS
    int32_t lds_out;
    int32_t lds_left;
    int32_t lds_right;
    int32_t level;
    double len_left, len_right;  // branch lengths to the children (the P(t) inputs)
};

// One transition-matrix job of the P(t) kernel: op index, child side, leaf flag, length.
struct PJob {
    int32_t op;
    int32_t side;
    int32_t leaf;
    int32_t pad;
    double len;
};

enum : int { KIND_TT = 0, KIND_TI = 1, KIND_II = 2 };
enum : int { MODE_PARTIALS = 0, MODE_LOGL = 1, MODE_THETA = 2 };

// Parameters of every likelihood kernel (passed by value).
struct KParams {
    int n, ncat, block, nptn, nptn_pad, ntiles, su, safe, ncw;  // ncw = waves per tile
    int wpc;                  // FP32 matrix-core tile shape: waves per category (1, 2, 4)
    void *part;               // slots x slot_stride reals
    int64_t slot_stride;      // reals per slot
    void *scale;              // slots x scale_stride counters (uint16 FP64 / int32 FP32)
    int64_t scale_stride;
    const uint8_t *states;    // ntaxa x nptn_pad
    int ntaxa;
    const double *freq, *invar;
    const void *mats;         // per-op matrices
    int64_t mat_stride;       // reals per op
    const double *inv_evec, *eval, *rates, *props, *tip_lh, *evec;
    const double *pi, *piU, *tipvec;  // FP32 form: frequencies, pi_x U[x][i] as [i][x], real tip vectors
    const DevOp *ops;
    int nops;
    int br_a, br_b;           // branch stage: node refs (br_a may be a tip), br_b a slot
    double br_len;
    double *ptn_lh;           // nptn_pad
    double *ptn_lh_cat;       // nptn_pad x ncat: per-category likelihood of the last logL (FP64 forms)
    double *ptn_scaling;      // nptn_pad: the log scaling of each pattern's likelihood (FP64 forms)
    double *g_frac;           // nptn_pad: L'/L per pattern of the last derivative evaluation (or null)
    void *theta;              // nptn_pad x block reals
    double *buf_scale;        // nptn_pad
    double *tile_sums;        // ntiles x 2
    unsigned *counter;
    double *result;           // 2
    float *sink;              // kSinkBytes that stores of waves without a real target go to
    int ncu;                  // compute units of the device
    int split_full;           // split launches: tiles handled whole (the rest as single blocks)
    // level-parallel 4-state kernel (ops sorted by dependency level)
    const int32_t *lvl;       // nlevels + 1 op-index boundaries
    // a second valid order of the ops (nops indices) that half of the tiles follow, or
    // null: tiles walking the same op sequence in step would alternate chip-wide between
    // a memory burst and a compute phase; two orders interleave the two
    const int32_t *perm;
    int nlevels;
    int nbuf;                 // LDS partial buffers
    int ntip;                 // tips whose states are staged in LDS
    const int32_t *tips;      // their taxon ids
    int br_lds_a, br_lds_b;   // LDS buffer (or staged tip index) of the branch's two sides
    int small_stage;          // level-parallel kernel: op list, level bounds and P(t) exponentials staged in LDS
    const double *uinv_frag;  // FP64 matrix-core traversal: U^-1 as A fragments (null: not built)
    int f64_form;             // FP64 20/61 states: 1 matrix-core traversal, 0 vector-ALU kernels
};

// Device-resident state of one Newton branch-length solve (Optimization::minimizeNewton,
// utils/optimization.cpp:422-499, with computeFuncDerv tree/phylotree.cpp:2555-2562 as the
// function).  The step logic runs on the device after every derivative evaluation, so a
// whole solve is a queue of launches with no host round trip in between.
struct NewtonState {
    double x1, x2, xacc;                            // bounds and tolerance
    double rts, rts_old, xl, xh, dx, f, df, out;    // iterate, bracket, last step, -d1, -d2
    double logl_eval, x_eval, logl0;                // many-branch form: logL at the last evaluated
                                                    // length x_eval, and at evaluation 0
    int32_t max_nr, j, done, status, evals, safe;
    int32_t eval_only, pad;                         // eval_only: evaluate logL at rts and stop
};
enum : int { NEWTON_OK = 0, NEWTON_UNDERFLOW = 1, NEWTON_NONFINITE = 2, NEWTON_MAXITER = 3 };

// Host-side context (opaque plk_ctx).
struct Ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    int n = 0, ncat = 0, block = 0, ntaxa = 0, nptn = 0, nptn_pad = 0, ntiles = 0, nslots = 0;
    int prec = PLK_FP64, safe = 0, su = -1;
    bool model_set = false, patterns_set = false;
    void *d_part = nullptr;
    void *d_scale = nullptr;
    uint8_t *d_states = nullptr;
    double *d_freq = nullptr, *d_invar = nullptr;
    void *d_mats = nullptr;
    int64_t mat_stride = 0;
    int mats_cap = 0;
    double *d_model = nullptr;  // eval | evec | inv_evec | rates | props | tip_lh
    double *d_eval = nullptr, *d_evec = nullptr, *d_inv_evec = nullptr, *d_rates = nullptr, *d_props = nullptr,
           *d_tip = nullptr, *d_pi = nullptr, *d_piU = nullptr, *d_tipvec = nullptr;
    DevOp *d_ops = nullptr;
    PJob *d_jobs = nullptr;
    int ops_cap = 0;
    void *d_theta = nullptr;
    double *d_buf_scale = nullptr, *d_ptn_lh = nullptr, *d_tile_sums = nullptr, *d_result = nullptr;
    double *d_side = nullptr;  // per-pattern side outputs: lh_cat (ncat per pattern) | scaling | L'/L
    size_t cap_side = 0;
    unsigned *d_counter = nullptr;
    float *d_sink = nullptr;
    int ncu = 0;  // compute units of the device
    NewtonState *d_newton = nullptr, *h_newton = nullptr;  // device state, pinned host mirror
    // host staging (pinned)
    void *h_stage = nullptr;
    size_t h_stage_cap = 0;
    double *h_result = nullptr;       // pinned, mapped: the branch logL kernel writes it directly
    double *h_result_dev = nullptr;   // its device address
    std::vector<DevOp> pending;
    std::vector<PJob> jobs;
    int theta_a = 0, theta_b = 0;
    bool theta_valid = false;
    bool ptn_lh_valid = false;  // d_ptn_lh holds the per-pattern logL of the last logL evaluation
    bool timing = false;
    int wpc = 1;  // FP32 matrix-core tile shape (plk_set_tile_shape)
    double *d_uinv_frag = nullptr;  // FP64 20/61 states: U^-1 as matrix-core fragments
    int f64_form = -1;              // FP64 20/61 states, ncat <= 4: 1 matrix cores, 0 vector ALU,
                                    // -1 by state count (61: matrix cores, 20: vector ALU)
    // 4-state models: ops per level run at once by the level-parallel kernel (0: the
    // category-split kernels; -1: chosen by size, see plk_set_level_waves)
    int small_waves = -1;
    // many-branch derivative buffers (plk_branch_derivs_multi)
    void *d_mparams = nullptr, *d_mitems = nullptr;
    double *d_mtile = nullptr, *d_mres = nullptr;  // 3 values per tile / branch
    unsigned *d_mcount = nullptr;
    NewtonState *d_mnewton = nullptr, *h_mnewton = nullptr;  // many-branch Newton states (pinned mirror)
    int mcap = 0;
    // allocated sizes (bytes) of the pattern-sized buffers, for plk_reshape
    size_t cap_part = 0, cap_scale = 0, cap_states = 0, cap_pad = 0, cap_theta = 0, cap_tiles = 0;
    hipEvent_t ev[6] = {};
    double t_total[3] = {0, 0, 0};
    int64_t t_count[3] = {0, 0, 0};
    int64_t updates = 0;
    std::vector<double> host_freq;
    size_t real_size() const { return prec == PLK_FP64 ? 8 : 4; }
};

}  // namespace plk

struct plk_ctx : public plk::Ctx {};

namespace plk {
// Launchers implemented in the kernel translation units.
hipError_t f64_pmatrix(const KParams &p, const PJob *jobs, int njobs, hipStream_t s);
hipError_t f64_traverse(const KParams &p, int mode, hipStream_t s);
hipError_t f64_theta_eval(const KParams &p, int deriv, hipStream_t s);
hipError_t newton_first(NewtonState *st, const double *result, hipStream_t s);
hipError_t f64_theta_newton(const KParams &p, NewtonState *st, hipStream_t s);
hipError_t f32_pmatrix(const KParams &p, const PJob *jobs, int njobs, hipStream_t s);
hipError_t f32_traverse(const KParams &p, int mode, hipStream_t s);
hipError_t f32_theta(const KParams &p, hipStream_t s);
// level-parallel 4-state traversal (form 0 FP64, 1 FP32), lds_bytes of dynamic LDS
// A short op list (ops, level bounds, staged tip ids, as 32-bit words) passed inside the
// kernel arguments of the level-parallel kernel, so that no upload precedes the launch.
constexpr int kSmallInlineWords = 720;
struct SmallInline {
    int32_t nwords, pad;
    uint32_t data[kSmallInlineWords];
};
hipError_t small_traverse(const KParams &p, int form, int mode, int waves, size_t lds_bytes, const SmallInline *inl,
                          hipStream_t s);
size_t small_lds_bytes(int form, int ncat, int safe, int su, int nbuf, int ntip);
// extra dynamic LDS of the level-parallel kernel's staged op list (0 = not staged)
size_t small_stage_bytes(int ncat, int nops, int nlevels);

// Shared host helpers (engine.cpp).
KParams make_params(Ctx &c);
// Write-only scratch of a context: a wave that has no real target for a store (a wave
// without a category of its own, the exponent copies of all but one wave) writes there,
// so that its stores need no branch.  Covers one 20-state tile-category in the FP32
// row-group layout (5 x 64 x 4 floats).
constexpr size_t kSinkBytes = 8192;
bool valid_ref(const Ctx &c, int r);
int ensure_op_capacity(Ctx &c, int nops);

// Batched partitions: one work item = one 64-pattern tile of one partition.
struct BatchItem {
    int32_t part;
    int32_t tile;
};
hipError_t batch_pmatrix(int prec, const KParams *params, const PJob *jobs, int njobs, hipStream_t s);
hipError_t batch_traverse(int prec, const KParams *params, const BatchItem *items, int nitems, int grid,
                          hipStream_t s);
hipError_t multi_derivs(int prec, int ncw, const KParams *params, const BatchItem *items, int nitems, hipStream_t s);
hipError_t multi_newton(int prec, int ncw, const KParams *params, const BatchItem *items, int nitems,
                        NewtonState *states, hipStream_t s);
}  // namespace plk
