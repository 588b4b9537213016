// Level-parallel traversal for 4-state (DNA) models.
//
// The category-split kernels walk a tile's ops one after another, so a 30-op traversal
// costs 30 dependent steps however few patterns there are.  Here the host sorts the op
// list into dependency levels (every op's children come from earlier levels), and a
// workgroup of CW x S wavefronts runs S independent ops of a level at once, CW
// wavefronts (one per category, or two categories each when ncat > 4) per op.  A
// traversal then costs (tree depth) steps.  S <= 4, so a workgroup has at most 16
// wavefronts.  The per-pattern check over categories (NORM
// scaling, FP32 renormalisation) is exchanged through LDS between the op's wavefronts,
// with one barrier per round of S ops.
//
// Partials produced in the launch are also kept in LDS buffers (assigned on the host by
// liveness), so a parent reads its children from LDS instead of re-fetching what the
// workgroup wrote a moment earlier; every partial is still stored to HBM.  The tip
// states of the tile and the tip vectors are staged in LDS at the start.  P(t) of each
// child is formed in the kernel from the eigen system (no P(t) launch, no matrix reads
// from memory), and a tip child's contribution from the P(t) rows and the tip vector
// (the sums the leaf tables hold, in the same order), so no op waits on memory other
// than its children.
//
// Layout of one LDS partial buffer: [c][x][lane] reals (the HBM layout of one tile).
#pragma once

#include "dev_common.h"

namespace plk {
namespace dev {

template <typename REAL>
struct SmallLds {
    REAL *part;      // nbuf x ncat x 4 x 64
    int *scale;      // nbuf x (ncat or 1) x 64  (FP64 counts / FP32 exponents)
    uint8_t *state;  // ntip x 64
    double *tipv;    // (su + 1) x 4 tip vectors (FP64 form: eigen-space tip_lh, FP32 form: real)
    double *val;     // ncat x 4 branch vectors (FP64 logL stage), padded to 3 x
    // staged op list (P.small_stage): nops x 12 words, nlevels + 1 level bounds, and
    // ex[op][side][c][i] = exp(lambda_i r_c t_side)
    int32_t *ops_w;
    int32_t *lvl;
    double *ex;
};

// Carve the dynamic LDS block; the host sizes it with the same formula (small_lds_bytes).
template <typename REAL>
__device__ __forceinline__ SmallLds<REAL> small_lds(const KParams &P, char *base) {
    SmallLds<REAL> s;
    const int scw = (sizeof(REAL) == 8 && P.safe) ? P.ncat : 1;
    size_t off = 0;
    s.part = (REAL *)(base + off);
    off += (size_t)P.nbuf * P.ncat * 4 * 64 * sizeof(REAL);
    s.val = (double *)(base + off);
    off += (size_t)3 * P.ncat * 4 * sizeof(double);
    s.tipv = (double *)(base + off);
    off += (size_t)(P.su + 1) * 4 * sizeof(double);
    off = (off + 15) & ~(size_t)15;
    s.scale = (int *)(base + off);
    off += (size_t)P.nbuf * scw * 64 * sizeof(int);
    s.state = (uint8_t *)(base + off);
    off += (size_t)P.ntip * 64;
    off = (off + 15) & ~(size_t)15;
    s.ops_w = (int32_t *)(base + off);
    off += (size_t)P.nops * sizeof(DevOp);
    s.lvl = (int32_t *)(base + off);
    off += (size_t)(P.nlevels + 1) * 4;
    off = (off + 15) & ~(size_t)15;
    s.ex = (double *)(base + off);
    return s;
}

// One category's 4 values of a child partial (LDS buffer or HBM).
template <typename REAL>
__device__ __forceinline__ void small_child_vals(const KParams &P, const SmallLds<REAL> &L, int ref, int lds, int c,
                                                 int tile, int lane, REAL (&v)[4]) {
    const REAL *src = lds >= 0 ? L.part + ((size_t)lds * P.ncat + c) * 4 * 64 + lane
                               : (const REAL *)P.part + (int64_t)ref * P.slot_stride +
                                     ((int64_t)tile * P.ncat + c) * 4 * 64 + lane;
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = src[i * 64];
}

__device__ __forceinline__ int small_tip_state(const KParams &P, const uint8_t *staged, int taxon, int idx, int tile,
                                               int lane) {
    return idx >= 0 ? (int)staged[idx * 64 + lane] : (int)P.states[(int64_t)taxon * P.nptn_pad + (int64_t)tile * 64 + lane];
}

template <typename REAL>
__device__ __forceinline__ int small_scale(const KParams &P, const SmallLds<REAL> &L, int ref, int lds, int cw, int lane,
                                           int64_t idx) {
    if (lds >= 0) return L.scale[(lds * (sizeof(REAL) == 8 && P.safe ? P.ncat : 1) + cw) * 64 + lane];
    if (sizeof(REAL) == 8) return ((const uint16_t *)P.scale)[(int64_t)ref * P.scale_stride + idx];
    return ((const int *)P.scale)[(int64_t)ref * P.scale_stride + idx];
}

// P(t) is formed inside the kernel (computePartialInfo, tree/phylokernelnew.h:1013-1063,
// with the P(t) kernel's arithmetic): ex[i] = exp(lambda_i (r_c t)); the FP64 form uses
// echild[x][i] = U[x][i] ex[i] (eigen space), the FP32 form P[x][y] = sum_i U[x][i] ex[i]
// Uinv[i][y] in FP64, rounded to FP32.
__device__ __forceinline__ void small_exp(const KParams &P, int c, double t, double (&ex)[4]) {
    const cdptr ev = as_const(P.eval), rt = as_const(P.rates);
#pragma unroll
    for (int i = 0; i < 4; i++) ex[i] = exp(ev[i] * (rt[c] * t));
}

// The exponentials of op k's child `side` in category c: from the staged table, else
// computed here (the same expression either way).
template <typename REAL>
__device__ __forceinline__ void small_exp_op(const KParams &P, const SmallLds<REAL> &L, int k, int side, int c, double t,
                                             double (&ex)[4]) {
    if (P.small_stage) {
        const double *e = L.ex + (((size_t)k * 2 + side) * P.ncat + c) * 4;
#pragma unroll
        for (int i = 0; i < 4; i++) ex[i] = e[i];
    } else {
        small_exp(P, c, t, ex);
    }
}

template <typename REAL>
__device__ __forceinline__ DevOp small_op(const KParams &P, const SmallLds<REAL> &L, int k) {
    if (!P.small_stage) return load_op(P.ops, k);
    const int32_t *w = L.ops_w + (size_t)k * (sizeof(DevOp) / 4);
    DevOp o;
    o.out_slot = __builtin_amdgcn_readfirstlane(w[0]);
    o.kind = __builtin_amdgcn_readfirstlane(w[1]);
    o.left = __builtin_amdgcn_readfirstlane(w[2]);
    o.right = __builtin_amdgcn_readfirstlane(w[3]);
    o.lds_out = __builtin_amdgcn_readfirstlane(w[4]);
    o.lds_left = __builtin_amdgcn_readfirstlane(w[5]);
    o.lds_right = __builtin_amdgcn_readfirstlane(w[6]);
    o.level = __builtin_amdgcn_readfirstlane(w[7]);
    o.len_left = o.len_right = 0.0;  // the exponentials are staged
    return o;
}

__device__ __forceinline__ void small_pmat(const KParams &P, const double (&ex)[4], double (&Pd)[16]) {
    const cdptr U = as_const(P.evec), Ui = as_const(P.inv_evec);
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < 4; i++) v = fma(U[x * 4 + i] * ex[i], Ui[i * 4 + y], v);
            Pd[x * 4 + y] = v;
        }
}

// FP64 side value (echild v)[x]: v a child partial, or for a tip its tip vector (the
// leaf table's row; 1 for the unknown state).
__device__ __forceinline__ double side_f64(const KParams &P, const double (&ex)[4], int x, bool tip, int state, int su,
                                           const double *tipv, const double (&v)[4]) {
    if (tip && state == su) return 1.0;
    const cdptr U = as_const(P.evec);
    double a = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) a = fma(U[x * 4 + i] * ex[i], tip ? tipv[state * 4 + i] : v[i], a);
    return a;
}

// FP32 side value (P v)[x] in FP32; a tip's row is formed in FP64 and rounded once (as
// the FP32 leaf table is).
__device__ __forceinline__ float side_f32(const double (&Pd)[16], int x, bool tip, int state, int su,
                                          const double *tipv, const float (&v)[4]) {
    if (tip) {
        if (state == su) return 1.0f;
        double w = 0.0;
#pragma unroll
        for (int y = 0; y < 4; y++) w = fma(Pd[x * 4 + y], tipv[state * 4 + y], w);
        return (float)w;
    }
    float a = 0.0f;
#pragma unroll
    for (int y = 0; y < 4; y++) a = fmaf((float)Pd[x * 4 + y], v[y], a);
    return a;
}

// Order this wavefront's LDS writes before its later reads (and reads before later
// writes): one wavefront's LDS operations are performed in order, so only the compiler
// has to be kept from moving them across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------------------
// FORM 0: FP64 eigen space, reference arithmetic (per category a[x] = eL[x].l,
// b[x] = eR[x].r, out[j] = Uinv[j].(a*b); NORM / SAFE scale counts,
// tree/phylokernelnew.h:1636-2036).  FORM 1: FP32 real space, out[x] = (P_L l)[x] (P_R r)[x],
// then the pattern is renormalised to a maximum in [0.5, 1) and the exponents add up.
// CPW = categories per wavefront (1 or 2).
template <int FORM, int CPW, int MODE>
__device__ inline void traverse_small_body(const KParams &P, const int tile, char *smem, const uint32_t *inl) {
    using REAL = std::conditional_t<FORM == 0, double, float>;
    const SmallLds<REAL> L = small_lds<REAL>(P, smem);
    const int lane = threadIdx.x & 63;
    const int ncat = P.ncat, su = P.su;
    const int CW = min(ncat, 4);              // category wavefronts per op
    const int w = wave_id(), S = nwaves() / CW;
    const int cw = w % CW, slot = w / CW;     // this wave: categories cw, cw + CW; op slot
    const int64_t p = (int64_t)tile * 64 + lane;
    const double invar = P.invar[p];
    // per (round parity, op slot, category wave, lane): FP64 NORM "all below 2^-256" flag,
    // FP32 category maximum
    __shared__ float s_x[2][4][4][64];
    __shared__ double s_cat[PLK_GPU_MAX_CATS][64];
    __shared__ int s_sc[PLK_GPU_MAX_CATS][64];
    // per wavefront: the two children's rows for the current op and category, formed
    // once by lanes 0-31 (FP64: U[x][i] ex[i]; FP32: P(t)[x][y] in FP64)
    __shared__ double s_rows[16][2][16];

    for (int i = threadIdx.x; i < (su + 1) * 4; i += blockDim.x)
        L.tipv[i] = FORM == 0 ? P.tip_lh[i] : P.tipvec[i];
    // the op list, level bounds and staged tip ids: in device memory, or (inl) in the
    // kernel arguments as words [ops | level bounds | tip ids]
    const int nw = P.nops * (int)(sizeof(DevOp) / 4);
    const uint32_t *ops_src = inl ? inl : (const uint32_t *)P.ops;
    const int32_t *lvl_src = inl ? (const int32_t *)inl + nw : P.lvl;
    const int32_t *tips_src = inl ? (const int32_t *)inl + nw + P.nlevels + 1 : P.tips;
    for (int i = threadIdx.x; i < P.ntip * 64; i += blockDim.x)
        L.state[i] = P.states[(int64_t)tips_src[i >> 6] * P.nptn_pad + (int64_t)tile * 64 + (i & 63)];
    if (P.small_stage) {
        for (int i = threadIdx.x; i < nw; i += blockDim.x) L.ops_w[i] = (int32_t)ops_src[i];
        for (int i = threadIdx.x; i <= P.nlevels; i += blockDim.x) L.lvl[i] = lvl_src[i];
        const int ne = P.nops * 2 * ncat * 4;
        for (int idx = threadIdx.x; idx < ne; idx += blockDim.x) {
            const int i = idx & 3, c = (idx >> 2) % ncat, side = (idx / (4 * ncat)) & 1, k = idx / (8 * ncat);
            // len_left / len_right: the op's words 8-9 / 10-11
            const uint32_t *lw = ops_src + (size_t)k * (sizeof(DevOp) / 4) + 8 + 2 * side;
            const double t = __builtin_bit_cast(double, ((unsigned long long)lw[1] << 32) | lw[0]);
            L.ex[idx] = exp(P.eval[i] * (P.rates[c] * t));
        }
    }
    if (FORM == 0 && MODE == MODE_LOGL) {
        // val[c][i] = exp(lambda_i r_c t) w_c (tree/phylokernelnew.h:2731-2766)
        for (int idx = threadIdx.x; idx < ncat * 4; idx += blockDim.x) {
            const int c = idx / 4, i = idx % 4;
            const double cof = P.eval[i] * P.rates[c];
            L.val[idx] = exp(cof * P.br_len) * P.props[c];
        }
    }
    __syncthreads();

    int par = 0;
    for (int lv = 0; lv < P.nlevels; lv++) {
        const int k0 = P.small_stage ? __builtin_amdgcn_readfirstlane(L.lvl[lv]) : load_int(P.lvl, lv);
        const int k1 = P.small_stage ? __builtin_amdgcn_readfirstlane(L.lvl[lv + 1]) : load_int(P.lvl, lv + 1);
        for (int base = k0; base < k1; base += S, par ^= 1) {
            const int k = base + slot;
            const bool act = k < k1;
            DevOp op{};
            REAL out[CPW][4];
            if (act) {
                op = small_op(P, L, k);
                float flag = 1.0f;  // FP64 NORM: every value of my categories below 2^-256
                float mx = 0.0f;    // FP32: maximum over my categories
                // the op kind is uniform: each kind gets straight-line code
                with_kind(op.kind, [&](auto KC) {
                constexpr int KIND = decltype(KC)::value;
                constexpr bool ltip = KIND <= KIND_TI, rtip = KIND == KIND_TT;
                const int sl = ltip ? small_tip_state(P, L.state, op.left, op.lds_left, tile, lane) : 0;
                const int sr = rtip ? small_tip_state(P, L.state, op.right, op.lds_right, tile, lane) : 0;
                double (*rows)[16] = s_rows[w];
#pragma unroll
                for (int ci = 0; ci < CPW; ci++) {
                    const int c = cw + ci * CW;
                    if (c >= ncat) break;
                    if (ci > 0) wave_lds_sync();  // the previous category's rows are read
                    if (lane < 32) {
                        const int sd = lane >> 4, x = (lane >> 2) & 3, i = lane & 3;
                        double ex[4];
                        small_exp_op(P, L, k, sd, c, sd ? op.len_right : op.len_left, ex);
                        const cdptr U = as_const(P.evec);
                        if constexpr (FORM == 0) {
                            rows[sd][lane & 15] = U[x * 4 + i] * ex[i];
                        } else {
                            const cdptr Ui = as_const(P.inv_evec);
                            double v = 0.0;  // P(t)[x][i], small_pmat's arithmetic
#pragma unroll
                            for (int j = 0; j < 4; j++) v = fma(U[x * 4 + j] * ex[j], Ui[j * 4 + i], v);
                            rows[sd][lane & 15] = v;
                        }
                    }
                    REAL lv4[4] = {0, 0, 0, 0}, rv4[4] = {0, 0, 0, 0};
                    if (!ltip) small_child_vals<REAL>(P, L, op.left, op.lds_left, c, tile, lane, lv4);
                    if (!rtip) small_child_vals<REAL>(P, L, op.right, op.lds_right, c, tile, lane, rv4);
                    wave_lds_sync();
                    if constexpr (FORM == 0) {
                        // side value (echild v)[x] with v the child or its tip vector; 1 for
                        // the unknown state (the leaf table's row)
                        auto side = [&](int sd, int x, bool tip, int st, const REAL(&v)[4]) {
                            double a = 0.0;
#pragma unroll
                            for (int i = 0; i < 4; i++) a = fma(rows[sd][x * 4 + i], tip ? L.tipv[st * 4 + i] : v[i], a);
                            return (tip && st == su) ? 1.0 : a;
                        };
                        const cdptr uinv = as_const(P.inv_evec);
                        double t[4];
#pragma unroll
                        for (int x = 0; x < 4; x++) t[x] = side(0, x, ltip, sl, lv4) * side(1, x, rtip, sr, rv4);
                        double m = 0.0;
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            double o = 0.0;
#pragma unroll
                            for (int x = 0; x < 4; x++) o = fma(uinv[j * 4 + x], t[x], o);
                            out[ci][j] = o;
                            m = fmax(m, fabs(o));
                        }
                        if (P.safe) {
                            // per-category check and counts (1997-2013)
                            const int64_t a = p * ncat + c;
                            unsigned sc = 0;
                            if (KIND == KIND_TI) sc = small_scale<double>(P, L, op.right, op.lds_right, c, lane, a);
                            else if (KIND == KIND_II)
                                sc = small_scale<double>(P, L, op.left, op.lds_left, c, lane, a) +
                                     small_scale<double>(P, L, op.right, op.lds_right, c, lane, a);
                            if (KIND != KIND_TT && m < kST && invar == 0.0) {
#pragma unroll
                                for (int j = 0; j < 4; j++) out[ci][j] *= kTwo256;
                                sc += 1;
                            }
                            ((uint16_t *)P.scale)[(int64_t)op.out_slot * P.scale_stride + a] = (uint16_t)sc;
                            if (op.lds_out >= 0) L.scale[(op.lds_out * ncat + c) * 64 + lane] = (int)sc;
                        } else if (!(m < kST)) {
                            flag = 0.0f;
                        }
                    } else {
                        // (P v)[x] in FP32; a tip's row in FP64, rounded once (the leaf table)
                        auto side = [&](int sd, int x, bool tip, int st, const REAL(&v)[4]) {
                            if (tip) {
                                double wv = 0.0;
#pragma unroll
                                for (int y = 0; y < 4; y++) wv = fma(rows[sd][x * 4 + y], L.tipv[st * 4 + y], wv);
                                return st == su ? 1.0f : (float)wv;
                            }
                            float a = 0.0f;
#pragma unroll
                            for (int y = 0; y < 4; y++) a = fmaf((float)rows[sd][x * 4 + y], v[y], a);
                            return a;
                        };
#pragma unroll
                        for (int x = 0; x < 4; x++) {
                            out[ci][x] = side(0, x, ltip, sl, lv4) * side(1, x, rtip, sr, rv4);
                            mx = fmaxf(mx, fabsf(out[ci][x]));
                        }
                    }
                }
                });
                s_x[par][slot][cw][lane] = FORM == 0 ? flag : mx;
            }
            __syncthreads();
            if (act) {
                int e = 0;
                bool do_scale = false;
                if constexpr (FORM == 0) {
                    if (!P.safe) {
                        // one check over all categories of the pattern (2019-2034)
                        bool all = true;
                        for (int q = 0; q < CW; q++) all = all && s_x[par][slot][q][lane] != 0.0f;
                        do_scale = op.kind != KIND_TT && all && invar == 0.0;
                        if (cw == 0) {
                            unsigned sc = 0;
                            if (op.kind == KIND_TI) sc = small_scale<double>(P, L, op.right, op.lds_right, 0, lane, p);
                            else if (op.kind == KIND_II)
                                sc = small_scale<double>(P, L, op.left, op.lds_left, 0, lane, p) +
                                     small_scale<double>(P, L, op.right, op.lds_right, 0, lane, p);
                            sc += do_scale ? 1u : 0u;
                            ((uint16_t *)P.scale)[(int64_t)op.out_slot * P.scale_stride + p] = (uint16_t)sc;
                            if (op.lds_out >= 0) L.scale[op.lds_out * 64 + lane] = (int)sc;
                        }
                    }
                } else {
                    float pm = 0.0f;
                    for (int q = 0; q < CW; q++) pm = fmaxf(pm, s_x[par][slot][q][lane]);
                    e = norm_exponent(pm);
                    if (cw == 0) {
                        int es = e;
                        if (op.kind == KIND_TI) es += small_scale<float>(P, L, op.right, op.lds_right, 0, lane, p);
                        else if (op.kind == KIND_II)
                            es += small_scale<float>(P, L, op.left, op.lds_left, 0, lane, p) +
                                  small_scale<float>(P, L, op.right, op.lds_right, 0, lane, p);
                        ((int *)P.scale)[(int64_t)op.out_slot * P.scale_stride + p] = es;
                        if (op.lds_out >= 0) L.scale[op.lds_out * 64 + lane] = es;
                    }
                }
#pragma unroll
                for (int ci = 0; ci < CPW; ci++) {
                    const int c = cw + ci * CW;
                    if (c >= ncat) break;
                    REAL *dst = (REAL *)P.part + (int64_t)op.out_slot * P.slot_stride + ((int64_t)tile * ncat + c) * 4 * 64 + lane;
                    REAL *lb = op.lds_out >= 0 ? L.part + ((size_t)op.lds_out * ncat + c) * 4 * 64 + lane : nullptr;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        REAL v;
                        if constexpr (FORM == 0) v = do_scale ? out[ci][j] * kTwo256 : out[ci][j];
                        else v = ldexpf(out[ci][j], -e);
                        dst[j * 64] = v;
                        if (lb) lb[j * 64] = v;
                    }
                }
            }
        }
        __syncthreads();  // this level's LDS partials and counts are visible to the next
    }
    if (MODE == MODE_PARTIALS) return;

    // -------- branch logL: category wavefronts of slot 0, then wavefront 0 sums
    const bool atip = P.br_a < 0;
    if (slot == 0) {
        const int sa = atip ? small_tip_state(P, L.state, -P.br_a - 1, P.br_lds_a, tile, lane) : 0;
#pragma unroll
        for (int ci = 0; ci < CPW; ci++) {
            const int c = cw + ci * CW;
            if (c >= ncat) break;
            REAL av[4] = {0, 0, 0, 0}, bv[4];
            if (!atip) small_child_vals<REAL>(P, L, P.br_a, P.br_lds_a, c, tile, lane, av);
            small_child_vals<REAL>(P, L, P.br_b, P.br_lds_b, c, tile, lane, bv);
            if constexpr (FORM == 0) {
                // lh_c = sum_i val[c][i] a_i b_i (tree/phylokernelnew.h:2903-2914, 3044-3057)
                double acc = 0.0;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const double a = atip ? L.tipv[sa * 4 + i] : av[i];
                    acc = fma(L.val[c * 4 + i] * a, bv[i], acc);
                }
                s_cat[c][lane] = acc;
                if (P.safe)
                    s_sc[c][lane] = small_scale<double>(P, L, P.br_b, P.br_lds_b, c, lane, p * ncat + c) +
                                    (atip ? 0 : small_scale<double>(P, L, P.br_a, P.br_lds_a, c, lane, p * ncat + c));
            } else {
                // lh_c = sum_x pi_x a[x] (P_c(t) b)[x]
                double ex[4], Pt[16];
                small_exp(P, c, P.br_len, ex);
                small_pmat(P, ex, Pt);
                float acc = 0.0f;
#pragma unroll
                for (int x = 0; x < 4; x++) {
                    float d = 0.0f;
#pragma unroll
                    for (int y = 0; y < 4; y++) d = fmaf((float)Pt[x * 4 + y], bv[y], d);
                    const float a = atip ? (float)L.tipv[sa * 4 + x] : av[x];
                    acc = fmaf((float)P.pi[x] * a, d, acc);
                }
                s_cat[c][lane] = (double)acc;
            }
        }
    }
    __syncthreads();
    double r0 = 0.0;
    if (w == 0) {
        double lh = 0.0, sc;
        int se = 0;  // FP32 form: the pattern's power-of-two exponent
        if (FORM == 0 && P.safe) {
            // 3063-3088: categories above the minimum count are scaled by 2^-256 or dropped
            int mn = s_sc[0][lane];
            for (int c = 1; c < ncat; c++) mn = min(mn, s_sc[c][lane]);
            for (int c = 0; c < ncat; c++) {
                const int d = s_sc[c][lane] - mn;
                lh += s_cat[c][lane] * (d == 0 ? 1.0 : (d == 1 ? kST : 0.0));
            }
            sc = (double)mn * kLogST;
        } else {
            for (int c = 0; c < ncat; c++) lh += FORM == 0 ? s_cat[c][lane] : P.props[c] * s_cat[c][lane];
            const int s = small_scale<REAL>(P, L, P.br_b, P.br_lds_b, 0, lane, p) +
                          (atip ? 0 : small_scale<REAL>(P, L, P.br_a, P.br_lds_a, 0, lane, p));
            sc = FORM == 0 ? (double)(unsigned)s * kLogST : 0.0;
            se = s;
        }
        // FP64 counts are 0 on patterns with invariant mass (never rescaled), so the sum is
        // unscaled there; the FP32 exponent is folded in by log_lh_expo.
        const double l = FORM == 0 ? log(fabs(lh) + invar) + sc : log_lh_expo(lh, se, invar);
        if (p < P.nptn) P.ptn_lh[p] = l;
        if (FORM == 0 && P.ptn_lh_cat && p < P.nptn) {
            // _pattern_lh_cat (tree/phylokernelnew.h:3023, SAFE-adjusted 3076-3083) and the
            // save_log_value = false scaling (3103-3104)
            int mn = 0;
            if (P.safe) {
                mn = s_sc[0][lane];
                for (int c = 1; c < ncat; c++) mn = min(mn, s_sc[c][lane]);
            }
            for (int c = 0; c < ncat; c++) {
                const int d = P.safe ? s_sc[c][lane] - mn : 0;
                P.ptn_lh_cat[p * ncat + c] = s_cat[c][lane] * (d == 0 ? 1.0 : (d == 1 ? kST : 0.0));
            }
            P.ptn_scaling[p] = sc;
        }
        r0 = wave_sum(p < P.nptn ? l * P.freq[p] : 0.0);
    }
    finish_tiles(P, r0, 0.0, tile, P.ntiles);
}

}  // namespace dev
}  // namespace plk
