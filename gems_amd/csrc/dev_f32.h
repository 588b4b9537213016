// FP32-form likelihood kernel bodies for gfx950.
//
// Partials are real-space conditional likelihoods in FP32 with one int32 power-of-two
// exponent per pattern and slot: true value = stored * 2^e.  Every op renormalises the
// pattern so that its largest entry lies in [0.5, 1), so FP32 never underflows for
// realistic trees.  Transition matrices P_c(t) = U exp(Lambda r_c t) U^-1, the exponent
// sums, theta, the pattern-weighted logL and the derivative sums stay FP64.
//
// 20- and 61-state models run the pattern x state by state x state contractions on
// the matrix cores (v_mfma_f32_16x16x4f32): for one category, D[x][p] = sum_y P[x][y]
// L[y][p] with x in 16-row blocks, y in 4-deep k-steps and 16 patterns per block.  The
// P fragments are stored pre-swizzled so each lane loads its A operand with one
// coalesced 4-byte load.  DNA uses the vector ALU with scalar-register matrices.
#pragma once

#include "dev_common.h"

namespace plk {
namespace dev {

// ---------------------------------------------------------------------------------------
// P(t) in FP32 form: P_c[x][y] (FP64 math) written as MFMA A fragments
// frag[c][xb][ks][lane] = P_c[f32_orow(xb, lane%16)][f32_kstate(ks, lane/16)] (zero padded),
// or plainly [c][x][y] for the vector-ALU DNA path; leaf tables leaf[s][c][r] =
// (P_c tip_s)[f32_orow(r / 16, r % 16)], rows f32_trow(N) long (zero padded for the
// matrix-core forms).
// Categories c0 .. c1-1 of the job (all by default).
template <int N>
// xbo >= 0 (matrix-core forms): only the rows of fragment block xbo -- its Pm rows, its A
// fragments and its leaf-table rows -- so that a job's blocks run in separate workgroups.
__device__ inline void pmatrix_f32_body(const KParams &P, const PJob &j, int c0 = 0, int c1 = -1, int xbo = -1) {
    constexpr int XB = f32_xblocks(N), KS = f32_ksteps(N), FR = XB * KS * 64, TROW = f32_trow(N);
    const int ncat = P.ncat, su = P.su;
    if (c1 < 0) c1 = ncat;
    float *M = (float *)P.mats + (int64_t)j.op * P.mat_stride;
    float *F = M + (int64_t)j.side * ncat * FR;
    float *T = M + (int64_t)2 * ncat * FR + (int64_t)j.side * (su + 1) * ncat * TROW;
    __shared__ double ex[N];
    __shared__ double Pm[N * N];
    for (int c = c0; c < c1; c++) {
        for (int i = threadIdx.x; i < N; i += blockDim.x) ex[i] = exp(P.eval[i] * (P.rates[c] * j.len));
        __syncthreads();
        const int nrow = xbo >= 0 ? 16 : N;
        for (int idx = threadIdx.x; idx < nrow * N; idx += blockDim.x) {
            const int x = xbo >= 0 ? f32_orow(N, xbo, idx / N) : idx / N, y = idx % N;
            if (x < 0) continue;
            double v = 0.0;
            for (int i = 0; i < N; i++) v = fma(P.evec[x * N + i] * ex[i], P.inv_evec[i * N + y], v);
            Pm[x * N + y] = v;
        }
        __syncthreads();
        if (N == 4) {
            for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) F[(int64_t)c * FR + idx] = (float)Pm[idx];
        } else {
            const int f0 = xbo >= 0 ? xbo * KS * 64 : 0, f1 = xbo >= 0 ? f0 + KS * 64 : FR;
            for (int idx = f0 + threadIdx.x; idx < f1; idx += blockDim.x) {
                const int xb = idx / (KS * 64), ks = (idx / 64) % KS, l = idx % 64;
                const int x = f32_orow(N, xb, l & 15), y = f32_kstate(N, ks, l >> 4);
                F[(int64_t)c * FR + idx] = (x >= 0 && y < N) ? (float)Pm[x * N + y] : 0.0f;
            }
        }
        if (j.leaf) {
            const int rw = xbo >= 0 ? 16 : TROW, r0 = xbo >= 0 ? xbo * 16 : 0;
            for (int idx = threadIdx.x; idx < su * rw; idx += blockDim.x) {
                const int s = idx / rw, r = r0 + idx % rw, x = f32_orow(N, r / 16, r % 16);
                double v = 0.0;
                if (x >= 0)
                    for (int y = 0; y < N; y++) v = fma(Pm[x * N + y], P.tipvec[s * N + y], v);
                T[((int64_t)s * ncat + c) * TROW + r] = (float)v;
            }
            for (int r = r0 + threadIdx.x; r < r0 + rw; r += blockDim.x)
                T[((int64_t)su * ncat + c) * TROW + r] = f32_orow(N, r / 16, r % 16) >= 0 ? 1.0f : 0.0f;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Matrix-core traversal (20 and 61 states).  Workgroup = one 64-pattern tile, one
// wavefront per category.  Lane l of a wave owns pattern column l%16 of each 16-pattern
// block and rows 4*(l/16)..+3 of each 16-row x block of the MFMA output.
//
// Tile shape: WPC wavefronts per category split the tile's four 16-pattern blocks
// (WPC = 1: one wave does all 64 patterns and reuses its A fragments four times;
// WPC = 2 or 4: more waves in flight per tile).  Wave w handles category w / WPC and
// pattern blocks [(w % WPC) * 4/WPC, ...).
//
// Single-block units (blk >= 0, WPC = 4): the workgroup has one wave per category and
// covers only pattern block blk of the tile; `unit` / `nunits` index the workgroup's
// logL partial sum (by default the tile / the tile count).  A launch whose tile count is
// just above a multiple of the compute-unit count runs its last tiles as such units, so
// no compute unit carries a whole extra tile.
template <int N, int MODE, int WPC = 1>
__device__ inline void traverse_f32_mfma_body(const KParams &P, const int tile, const int blk = -1, int unit = -1,
                                              int nunits = -1) {
    constexpr int XB = f32_xblocks(N), KS = f32_ksteps(N), FR = XB * KS * 64, TROW = f32_trow(N);
    constexpr int NPB = 4 / WPC;  // pattern blocks per wave
    const int lane = threadIdx.x & 63, l16 = lane & 15, lg = lane >> 4;
    const int ncat = P.ncat;
    if (unit < 0) {
        unit = tile;
        nunits = P.ntiles;
    }
    // waves beyond ncat * WPC (batched launches) only join barriers
    const int wv = wave_id();
    const int cw = blk >= 0 ? wv : wv / WPC;
    const int pb0 = blk >= 0 ? blk : (wv % WPC) * NPB;
    const bool active = cw < ncat;
    const int c = active ? cw : 0;
    float *__restrict__ part = (float *)P.part;
    int *__restrict__ expo = (int *)P.scale;
    const int64_t cbase = ((int64_t)tile * ncat + c) * N * 64;  // (tile, c, x=0, lane=0)
    // per-pattern maxima of the categories, [pattern][category] so two 16-byte LDS reads
    // give a pattern's eight; entries of categories >= ncat stay zero, waves without a
    // category of their own write s_spare
    static_assert(PLK_GPU_MAX_CATS == 8, "two 16-byte reads of the category maxima");
    __shared__ floatx4 s_max[2][64][2];
    __shared__ float s_spare[64];
    for (int i = threadIdx.x; i < 2 * 64 * 2; i += blockDim.x) s_max[i / 128][(i / 2) % 64][i % 2] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    __shared__ double s_cat[PLK_GPU_MAX_CATS][64];
    __shared__ float s_pi[64];
    __shared__ uint32_t s_tip[kStageTaxa * 16];
    if (MODE == MODE_LOGL && threadIdx.x < N) s_pi[threadIdx.x] = (float)P.pi[threadIdx.x];
    stage_tips(P, tile, s_tip);

    // pattern blocks per renormalisation group (outputs held in registers across the
    // category exchange): two for <= 32 states, whose loads for all of the wave's blocks
    // are issued at the op's start (HOIST); the 61-state model holds its 2 x 64 A
    // fragments in registers, so it loads and works on one block at a time
    constexpr bool HOIST = N <= 32 && NPB <= 2;
    constexpr int PBG = (N > 32) ? 1 : (NPB < 2 ? NPB : 2);
    constexpr int NB = HOIST ? NPB : PBG;  // pattern blocks whose child columns are held
    // Carry (20 states): in a post-order op list nearly every op has the previous op's
    // output as a child (the node's second-visited subtree, or its only internal child),
    // and the host puts that child on the right.  In the row-group layout a lane's output
    // rows 4g..4g+3 are exactly the child column entries it feeds to the next op's matrix
    // instructions (k-steps 0-3), and state 16 + g is the first row of its second x block
    // (f32_orow), which feeds k-step 4; so the previous output is kept in registers, with
    // its exponents, and that child is not read back from memory.  The values are the ones
    // stored, so results are unchanged.
    // The carried output is also written late: op k's stores are issued after op k+1's
    // loads, so that op k+1 does not wait for them (a wavefront's vector memory operations
    // complete in issue order, stores included) and they have a whole op to complete.
    constexpr bool CARRY = f32_rg(N);
    float cv[CARRY ? NPB : 1][KS];
    int ce[CARRY ? NPB : 1];
    int prev_out = -1;  // slot written by this tile's previous op (held in cv, ce)
#pragma unroll
    for (int j = 0; j < (CARRY ? NPB : 1); j++) {
        ce[j] = 0;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) cv[j][ks] = 0.0f;
    }
    // Straight-line stores: a branch around them would make the compiler's wait for the
    // next op's loads assume the path without stores, i.e. wait for these stores too.  So
    // every lane stores, and where there is no real target (before the first op, waves
    // without a category of their own, the exponents of all but the first wave, which keeps
    // them) the address is the context's sink.
    auto store_carried = [&]() {
        if constexpr (CARRY) {
            float *__restrict__ po = (prev_out >= 0 && active) ? part + (int64_t)prev_out * P.slot_stride + cbase : P.sink;
            int *__restrict__ pe = (prev_out >= 0 && cw == 0)
                                       ? expo + (int64_t)prev_out * P.scale_stride + (int64_t)tile * 64
                                       : (int *)P.sink;
#pragma unroll
            for (int j = 0; j < NPB; j++) {
                const int pl = (pb0 + j) * 16 + l16;
                *(floatx4 *)(po + (lg * 64 + pl) * 4) = floatx4{cv[j][0], cv[j][1], cv[j][2], cv[j][3]};
                po[(4 * 64 + pl) * 4 + lg] = cv[j][4];
                pe[pl] = ce[j];
            }
        }
    };
    int sp = 0;  // double-buffered s_max
    for (int k = 0; k < P.nops; k++) {
        const int ko = op_index(P, tile, k);
        const DevOp op = load_op(P.ops, ko);
        // the op kind is uniform over the launch: each kind gets its own straight-line
        // code, so all loads of an op are issued together
        with_kind(op.kind, [&](auto KC) {
        constexpr int KIND = decltype(KC)::value;
        const bool carry_r = CARRY && KIND != KIND_TT && op.right == prev_out;
        if (CARRY && KIND == KIND_II && op.left == prev_out) {
            // the previous output read from memory (an op list whose order does not put
            // it on the right): write it first
            store_carried();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const float *__restrict__ M = (const float *)P.mats + (int64_t)ko * P.mat_stride;
        const float *__restrict__ tabs = M + (int64_t)2 * ncat * FR;
        const float *__restrict__ R = part + (int64_t)op.right * P.slot_stride + cbase;
        const float *__restrict__ L = part + (int64_t)op.left * P.slot_stride + cbase;
        // Every global load of an op (A fragments, the children's columns, tip states and
        // exponents) is issued before the op's first matrix instruction, unconditionally
        // (inactive waves read category 0 and store nothing; rows past N read row N-1 and
        // are zeroed by a select), so an op costs one memory round trip (two with a tip
        // child, whose table rows depend on its states) instead of one per pattern block.
        float aL[XB][KS], aR[XB][KS];
        if (KIND != KIND_TT) {
            const float *__restrict__ FRm = M + (int64_t)(ncat + c) * FR;
#pragma unroll
            for (int xb = 0; xb < XB; xb++)
#pragma unroll
                for (int ks = 0; ks < KS; ks++) aR[xb][ks] = FRm[(xb * KS + ks) * 64 + lane];
            if (KIND == KIND_II) {
                const float *__restrict__ FLm = M + (int64_t)c * FR;
#pragma unroll
                for (int xb = 0; xb < XB; xb++)
#pragma unroll
                    for (int ks = 0; ks < KS; ks++) aL[xb][ks] = FLm[(xb * KS + ks) * 64 + lane];
            }
        }
        float *__restrict__ outp = part + (int64_t)op.out_slot * P.slot_stride + cbase;
        // children's exponents (kept by the wavefronts of category 0, lanes 0-15)
        int ech[NPB];
#pragma unroll
        for (int j = 0; j < NPB; j++) {
            const int64_t pj = (int64_t)tile * 64 + (pb0 + j) * 16 + l16;
            const int jc = CARRY ? j : 0;
            if (KIND == KIND_TI) ech[j] = carry_r ? ce[jc] : expo[(int64_t)op.right * P.scale_stride + pj];
            else if (KIND == KIND_II)
                ech[j] = expo[(int64_t)op.left * P.scale_stride + pj] +
                         (carry_r ? ce[jc] : expo[(int64_t)op.right * P.scale_stride + pj]);
            else ech[j] = 0;
        }
        float bL[NB][KS], bR[NB][KS];
        floatx4 tv[NB][XB];  // tip rows: left table (TI), left x right tables (TT)
        auto load_block = [&](int j, int pb) {
            const int col = pb * 16 + l16;
            if constexpr (f32_rg(N)) {
                // row group lg (k-steps 0-3) as one 16-byte load, state 16 + lg (k-step 4)
                const int jc = CARRY ? pb - pb0 : 0;  // block of the wave (cv index)
                if (KIND != KIND_TT) {
                    if (carry_r) {
#pragma unroll
                        for (int ks = 0; ks < KS; ks++) bR[j][ks] = cv[jc][ks];
                    } else {
                        const floatx4 q = *(const floatx4 *)(R + (lg * 64 + col) * 4);
#pragma unroll
                        for (int ks = 0; ks < 4; ks++) bR[j][ks] = q[ks];
                        bR[j][4] = R[(4 * 64 + col) * 4 + lg];
                    }
                }
                if (KIND == KIND_II) {
                    const floatx4 q = *(const floatx4 *)(L + (lg * 64 + col) * 4);
#pragma unroll
                    for (int ks = 0; ks < 4; ks++) bL[j][ks] = q[ks];
                    bL[j][4] = L[(4 * 64 + col) * 4 + lg];
                }
            } else {
#pragma unroll
                for (int ks = 0; ks < KS; ks++) {
                    const int y = ks * 4 + lg, yr = y < N ? y : N - 1;
                    if (KIND != KIND_TT) {
                        const float v = R[yr * 64 + col];
                        bR[j][ks] = y < N ? v : 0.0f;
                    }
                    if (KIND == KIND_II) {
                        const float v = L[yr * 64 + col];
                        bL[j][ks] = y < N ? v : 0.0f;
                    }
                }
            }
            if (KIND != KIND_II) {
                const int sl = tip_state(P, s_tip, op.left, tile, col);
                const int sr = KIND == KIND_TT ? tip_state(P, s_tip, op.right, tile, col) : 0;
                const floatx4 *tl = (const floatx4 *)(tabs + ((int64_t)sl * ncat + c) * TROW + lg * 4);
                const floatx4 *tr =
                    (const floatx4 *)(tabs + (int64_t)(P.su + 1) * ncat * TROW + ((int64_t)sr * ncat + c) * TROW + lg * 4);
#pragma unroll
                for (int xb = 0; xb < XB; xb++) {
                    if (f32_rg(N) && xb == 1) {
                        // block 1: only the lane group's first row (state 16 + lg) is non-zero
                        const float a = ((const float *)tl)[16];
                        tv[j][1] = floatx4{KIND == KIND_TT ? a * ((const float *)tr)[16] : a, 0.0f, 0.0f, 0.0f};
                    } else {
                        tv[j][xb] = KIND == KIND_TT ? tl[xb * 4] * tr[xb * 4] : tl[xb * 4];
                    }
                }
            }
        };
        if constexpr (HOIST) {
#pragma unroll
            for (int j = 0; j < NPB; j++) load_block(j, pb0 + j);
        }
        // 20-state waves with four blocks load a pair of blocks at a time: the first pair
        // here, the next one as soon as the current pair's matrix instructions have read
        // their operands (before its category exchange and stores)
        constexpr bool EARLY = CARRY && !HOIST;
        if constexpr (EARLY) {
#pragma unroll
            for (int j = 0; j < PBG; j++) load_block(j, pb0 + j);
        }
        if constexpr (CARRY) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            store_carried();
        }
#pragma unroll
        for (int gi = 0; gi < NPB; gi += PBG) {
            const int g = pb0 + gi;
            if constexpr (!HOIST && !EARLY) {
#pragma unroll
                for (int j = 0; j < PBG; j++) load_block(j, g + j);
            }
            floatx4 out[PBG][XB];
            const int jb = HOIST ? gi : 0;  // first held block of this group
#pragma unroll
            for (int j = 0; j < PBG; j++) {
                const int pb = g + j;
                if (KIND == KIND_TT) {
#pragma unroll
                    for (int xb = 0; xb < XB; xb++) out[j][xb] = tv[jb + j][xb];
                } else {
                    // D_R = P_R . R (and D_L = P_L . L) on the matrix cores
                    floatx4 dL[XB], dR[XB];
#pragma unroll
                    for (int xb = 0; xb < XB; xb++) {
                        dL[xb] = floatx4{0, 0, 0, 0};
                        dR[xb] = floatx4{0, 0, 0, 0};
                    }
#pragma unroll
                    for (int ks = 0; ks < KS; ks++) {
#pragma unroll
                        for (int xb = 0; xb < XB; xb++)
                            dR[xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aR[xb][ks], bR[jb + j][ks], dR[xb], 0, 0, 0);
                        if (KIND == KIND_II) {
#pragma unroll
                            for (int xb = 0; xb < XB; xb++)
                                dL[xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aL[xb][ks], bL[jb + j][ks], dL[xb], 0, 0, 0);
                        }
                    }
                    if (KIND == KIND_II) {
#pragma unroll
                        for (int xb = 0; xb < XB; xb++) out[j][xb] = dL[xb] * dR[xb];
                    } else {
#pragma unroll
                        for (int xb = 0; xb < XB; xb++) out[j][xb] = tv[jb + j][xb] * dR[xb];
                    }
                    if constexpr (f32_rg(N)) out[j][1] = floatx4{out[j][1][0], 0.0f, 0.0f, 0.0f};  // zero rows
                }
                // per-pattern maximum over this category (reduced over the 4 row groups);
                // 20 states: rows 1-3 of block 1 are zero rows, left out
                float m = 0.0f;
#pragma unroll
                for (int xb = 0; xb < XB; xb++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (!(f32_rg(N) && xb == 1 && r > 0)) m = fmaxf(m, fabsf(out[j][xb][r]));
                m = fmaxf(m, __shfl_xor(m, 16, 64));
                m = fmaxf(m, __shfl_xor(m, 32, 64));
                // all four row groups store the same value; inactive waves use a spare row
                float *smx = active ? &((float *)&s_max[sp][pb * 16 + l16][0])[c] : &s_spare[pb * 16 + l16];
                *smx = m;
            }
            if constexpr (EARLY) {
                if (gi + PBG < NPB) {
#pragma unroll
                    for (int j = 0; j < PBG; j++) load_block(j, g + PBG + j);
                }
            }
            __syncthreads();
            // maximum over all categories -> power-of-two renormalisation and exponent sum
#pragma unroll
            for (int j = 0; j < PBG; j++) {
                const int pl = (g + j) * 16 + l16;
                const floatx4 c0 = s_max[sp][pl][0], c1 = s_max[sp][pl][1];
                const float pm = fmaxf(fmaxf(fmaxf(c0[0], c0[1]), fmaxf(c0[2], c0[3])),
                                       fmaxf(fmaxf(c1[0], c1[1]), fmaxf(c1[2], c1[3])));
                const int e = norm_exponent(pm);
                if constexpr (f32_rg(N)) {
                    static_assert(!f32_rg(N) || CARRY, "the 20-state form always carries");
                    // rows 4lg..4lg+3 (states 4lg..4lg+3) and state 16 + lg (first row of
                    // block 1): exactly the next op's k-steps 0-3 and 4 in this lane;
                    // store_carried writes them
                    const int jc = gi + j;
#pragma unroll
                    for (int r = 0; r < 4; r++) cv[jc][r] = ldexpf(out[j][0][r], -e);
                    cv[jc][4] = ldexpf(out[j][1][0], -e);
                    ce[jc] = e + ech[gi + j];
                } else {
#pragma unroll
                    for (int xb = 0; xb < XB; xb++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int x = f32_orow(N, xb, lg * 4 + r);
                            if (x >= 0 && active) outp[x * 64 + pl] = ldexpf(out[j][xb][r], -e);
                        }
                }
                if (!CARRY && cw == 0 && lg == 0)  // one wave per pattern block keeps the exponents
                    expo[(int64_t)op.out_slot * P.scale_stride + (int64_t)tile * 64 + pl] = e + ech[gi + j];
            }
            sp ^= 1;
        }
        });
        prev_out = op.out_slot;
        // The next op reads these values (and exponents) in other lanes of this same
        // wavefront: a wavefront's vector memory operations are performed in order, and
        // the wavefront-scope fences keep the compiler from moving the next op's loads
        // above these stores.  No workgroup barrier, so a wave does not wait here for
        // its stores to complete; s_max is double-buffered across the per-op exchange.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (CARRY && prev_out >= 0) {
        store_carried();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (MODE == MODE_PARTIALS) return;
    __syncthreads();  // s_pi visible even when the op list is empty

    // -------- branch logL: lh_c = sum_x pi_x A[x] (P_c(t) B)[x]
    const float *__restrict__ Mt = (const float *)P.mats + (int64_t)P.nops * P.mat_stride + (int64_t)c * FR;
    float at[XB][KS];
#pragma unroll
    for (int xb = 0; xb < XB; xb++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) at[xb][ks] = Mt[(xb * KS + ks) * 64 + lane];
    const bool atip = P.br_a < 0;
    const float *__restrict__ A = atip ? nullptr : part + (int64_t)P.br_a * P.slot_stride + cbase;
    const float *__restrict__ B = part + (int64_t)P.br_b * P.slot_stride + cbase;
    const uint8_t *sta = atip ? P.states + (int64_t)(-P.br_a - 1) * P.nptn_pad + (int64_t)tile * 64 : nullptr;
#pragma unroll
    for (int pb = pb0; pb < pb0 + NPB; pb++) {
        floatx4 d[XB];
#pragma unroll
        for (int xb = 0; xb < XB; xb++) d[xb] = floatx4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int y = f32_kstate(N, ks, lg);
            const float b = (y < N) ? B[f32_xoff(N, y) + f32_loff(N, pb * 16 + l16)] : 0.0f;
#pragma unroll
            for (int xb = 0; xb < XB; xb++) d[xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(at[xb][ks], b, d[xb], 0, 0, 0);
        }
        const double *tv = atip ? P.tipvec + (int64_t)sta[pb * 16 + l16] * N : nullptr;
        float acc = 0.0f;
#pragma unroll
        for (int xb = 0; xb < XB; xb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int x = f32_orow(N, xb, lg * 4 + r);
                if (x >= 0) {
                    const float a = atip ? (float)tv[x] : A[f32_xoff(N, x) + f32_loff(N, pb * 16 + l16)];
                    acc = fmaf(s_pi[x] * a, d[xb][r], acc);
                }
            }
        acc += __shfl_xor(acc, 16, 64);
        acc += __shfl_xor(acc, 32, 64);
        if (lg == 0 && active) s_cat[c][pb * 16 + l16] = (double)acc;
    }
    __syncthreads();
    double r0 = 0.0;
    if (wv == 0) {
        const int64_t p = (int64_t)tile * 64 + lane;
        const bool mine = p < P.nptn && (blk < 0 || (lane >> 4) == blk);  // this workgroup's patterns
        double lh = 0.0;
        for (int cc = 0; cc < ncat; cc++) lh += P.props[cc] * s_cat[cc][lane];
        const int e = expo[(int64_t)P.br_b * P.scale_stride + p] + (atip ? 0 : expo[(int64_t)P.br_a * P.scale_stride + p]);
        const double l = log_lh_expo(lh, e, P.invar[p]);
        if (mine) P.ptn_lh[p] = l;
        r0 = wave_sum(mine ? l * P.freq[p] : 0.0);
    }
    finish_tiles(P, r0, 0.0, unit, nunits);
}

// ---------------------------------------------------------------------------------------
// Matrix-core traversal, pattern-block shape: each wavefront owns one 16-pattern block of
// the tile across ALL categories (ncat <= 4).  The per-pattern maximum over categories is
// then wave-local (registers + two lane shuffles), so wavefronts never wait for each
// other: no LDS exchange and no barrier inside the op loop.  A wave re-reads its own
// stores in other lanes; vector memory operations of one wavefront are performed in
// order, and a wavefront-scope fence keeps the compiler from moving the next op's loads
// above them.
template <int N, int MODE>
__device__ inline void traverse_f32_pbw_body(const KParams &P, const int tile) {
    constexpr int XB = f32_xblocks(N), KS = f32_ksteps(N), FR = XB * KS * 64, TROW = f32_trow(N);
    constexpr int NC = 4;
    const int lane = threadIdx.x & 63, l16 = lane & 15, lg = lane >> 4;
    const int ncat = P.ncat;
    const int pb = wave_id();  // 16-pattern block of the tile
    const bool active = pb < 4;
    const int pbc = active ? pb : 0;
    float *__restrict__ part = (float *)P.part;
    int *__restrict__ expo = (int *)P.scale;
    const int pl = pbc * 16 + l16;                       // pattern within the tile
    const int64_t p = (int64_t)tile * 64 + pl;
    auto cbase = [&](int c) { return ((int64_t)tile * ncat + c) * N * 64 + f32_loff(N, pbc * 16 + l16); };
    __shared__ double s_wsum[16];
    __shared__ uint32_t s_tip[kStageTaxa * 16];
    stage_tips(P, tile, s_tip);

    for (int k = 0; k < P.nops; k++) {
        const int ko = op_index(P, tile, k);
        const DevOp op = load_op(P.ops, ko);
        with_kind(op.kind, [&](auto KC) {
        constexpr int KIND = decltype(KC)::value;
        const float *__restrict__ M = (const float *)P.mats + (int64_t)ko * P.mat_stride;
        const float *__restrict__ tabs = M + (int64_t)2 * ncat * FR;
        int sl = 0, sr = 0;
        if (KIND <= KIND_TI) sl = tip_state(P, s_tip, op.left, tile, pl);
        if (KIND == KIND_TT) sr = tip_state(P, s_tip, op.right, tile, pl);
        floatx4 out[NC][XB];
        float m = 0.0f;
#pragma unroll
        for (int c = 0; c < NC; c++) {
#pragma unroll
            for (int xb = 0; xb < XB; xb++) out[c][xb] = floatx4{0, 0, 0, 0};
            if (c >= ncat || !active) continue;
            if (KIND == KIND_TT) {
                const floatx4 *tl = (const floatx4 *)(tabs + ((int64_t)sl * ncat + c) * TROW + lg * 4);
                const floatx4 *tr =
                    (const floatx4 *)(tabs + (int64_t)(P.su + 1) * ncat * TROW + ((int64_t)sr * ncat + c) * TROW + lg * 4);
#pragma unroll
                for (int xb = 0; xb < XB; xb++) out[c][xb] = tl[xb * 4] * tr[xb * 4];
            } else {
                const float *__restrict__ FRm = M + (int64_t)(ncat + c) * FR;
                const float *__restrict__ FLm = M + (int64_t)c * FR;
                const float *__restrict__ R = part + (int64_t)op.right * P.slot_stride + cbase(c);
                const float *__restrict__ L = part + (int64_t)op.left * P.slot_stride + cbase(c);
                floatx4 dL[XB], dR[XB];
#pragma unroll
                for (int xb = 0; xb < XB; xb++) {
                    dL[xb] = floatx4{0, 0, 0, 0};
                    dR[xb] = floatx4{0, 0, 0, 0};
                }
#pragma unroll
                for (int ks = 0; ks < KS; ks++) {
                    const int y = f32_kstate(N, ks, lg);
                    const float br = (y < N) ? R[f32_xoff(N, y)] : 0.0f;
#pragma unroll
                    for (int xb = 0; xb < XB; xb++)
                        dR[xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(FRm[(xb * KS + ks) * 64 + lane], br, dR[xb], 0, 0, 0);
                    if (KIND == KIND_II) {
                        const float bl = (y < N) ? L[f32_xoff(N, y)] : 0.0f;
#pragma unroll
                        for (int xb = 0; xb < XB; xb++)
                            dL[xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(FLm[(xb * KS + ks) * 64 + lane], bl, dL[xb], 0, 0, 0);
                    }
                }
                if (KIND == KIND_II) {
#pragma unroll
                    for (int xb = 0; xb < XB; xb++) out[c][xb] = dL[xb] * dR[xb];
                } else {
                    const floatx4 *tl = (const floatx4 *)(tabs + ((int64_t)sl * ncat + c) * TROW + lg * 4);
#pragma unroll
                    for (int xb = 0; xb < XB; xb++) out[c][xb] = tl[xb * 4] * dR[xb];
                }
            }
#pragma unroll
            for (int xb = 0; xb < XB; xb++)
#pragma unroll
                for (int r = 0; r < 4; r++) m = fmaxf(m, fabsf(out[c][xb][r]));
        }
        m = fmaxf(m, __shfl_xor(m, 16, 64));
        m = fmaxf(m, __shfl_xor(m, 32, 64));
        const int e = norm_exponent(m);
        if (active) {
            float *__restrict__ outp = part + (int64_t)op.out_slot * P.slot_stride;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                if (c >= ncat) continue;
#pragma unroll
                for (int xb = 0; xb < XB; xb++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int x = f32_orow(N, xb, lg * 4 + r);
                        if (x >= 0) outp[cbase(c) + f32_xoff(N, x)] = ldexpf(out[c][xb][r], -e);
                    }
            }
            if (lg == 0) {
                int es = e;
                if (KIND == KIND_TI) es += expo[(int64_t)op.right * P.scale_stride + p];
                else if (KIND == KIND_II)
                    es += expo[(int64_t)op.left * P.scale_stride + p] + expo[(int64_t)op.right * P.scale_stride + p];
                expo[(int64_t)op.out_slot * P.scale_stride + p] = es;
            }
        }
        });
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (MODE == MODE_PARTIALS) return;

    // -------- branch logL: lh = sum_c w_c sum_x pi_x A_c[x] (P_c(t) B_c)[x]
    const float *__restrict__ Mt = (const float *)P.mats + (int64_t)P.nops * P.mat_stride;
    const bool atip = P.br_a < 0;
    const int sa = (atip && active) ? P.states[(int64_t)(-P.br_a - 1) * P.nptn_pad + p] : 0;
    double lh = 0.0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (c >= ncat || !active) continue;
        const float *__restrict__ Ft = Mt + (int64_t)c * FR;
        const float *__restrict__ A = part + (int64_t)(atip ? 0 : P.br_a) * P.slot_stride + cbase(c);
        const float *__restrict__ B = part + (int64_t)P.br_b * P.slot_stride + cbase(c);
        floatx4 d[XB];
#pragma unroll
        for (int xb = 0; xb < XB; xb++) d[xb] = floatx4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int y = f32_kstate(N, ks, lg);
            const float b = (y < N) ? B[f32_xoff(N, y)] : 0.0f;
#pragma unroll
            for (int xb = 0; xb < XB; xb++) d[xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ft[(xb * KS + ks) * 64 + lane], b, d[xb], 0, 0, 0);
        }
        float acc = 0.0f;
#pragma unroll
        for (int xb = 0; xb < XB; xb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int x = f32_orow(N, xb, lg * 4 + r);
                if (x >= 0) {
                    const float a = atip ? (float)P.tipvec[sa * N + x] : A[f32_xoff(N, x)];
                    acc = fmaf((float)P.pi[x] * a, d[xb][r], acc);
                }
            }
        acc += __shfl_xor(acc, 16, 64);
        acc += __shfl_xor(acc, 32, 64);
        lh += P.props[c] * (double)acc;
    }
    double contrib = 0.0;
    if (active && lg == 0) {
        const int e = expo[(int64_t)P.br_b * P.scale_stride + p] + (atip ? 0 : expo[(int64_t)P.br_a * P.scale_stride + p]);
        const double l = log_lh_expo(lh, e, P.invar[p]);
        if (p < P.nptn) {
            P.ptn_lh[p] = l;
            contrib = l * P.freq[p];
        }
    }
    const double ws = wave_sum(contrib);
    if (lane == 0) s_wsum[wave_id()] = ws;
    __syncthreads();
    double r0 = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < nwaves() && w < 16; w++) r0 += s_wsum[w];
    finish_tiles(P, r0, 0.0, tile, P.ntiles);
}

// ---------------------------------------------------------------------------------------
// Vector-ALU traversal for DNA: thread per (pattern, category), matrices in SGPRs.
template <int N, int MODE>
__device__ inline void traverse_f32_valu_body(const KParams &P, const int tile) {
    constexpr int XB = f32_xblocks(N), KS = f32_ksteps(N), FR = XB * KS * 64;
    const int lane = threadIdx.x & 63;
    const int w = wave_id();
    const int ncw = nwaves(), ncat = P.ncat;
    const int64_t p = (int64_t)tile * 64 + lane;
    float *__restrict__ part = (float *)P.part;
    int *__restrict__ expo = (int *)P.scale;
    __shared__ float s_max[2][4][64];
    __shared__ double s_cat[PLK_GPU_MAX_CATS][64];
    constexpr int CPW = (PLK_GPU_MAX_CATS + 3) / 4;  // categories per wave at most

    for (int k = 0; k < P.nops; k++) {
        const int ko = op_index(P, tile, k);
        const DevOp op = load_op(P.ops, ko);
        const float *__restrict__ M = (const float *)P.mats + (int64_t)ko * P.mat_stride;
        const float *__restrict__ tabs = M + (int64_t)2 * ncat * FR;
        int sl = 0, sr = 0;
        if (op.kind <= KIND_TI) sl = P.states[(int64_t)op.left * P.nptn_pad + p];
        if (op.kind == KIND_TT) sr = P.states[(int64_t)op.right * P.nptn_pad + p];
        float out[CPW][N];
        float wmax = 0.0f;
#pragma unroll
        for (int ci = 0; ci < CPW; ci++) {
            const int c = w + ci * ncw;
            if (c >= ncat) break;
            const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
            const cfptr PL = as_constf(M + (int64_t)c * FR);
            const cfptr PR = as_constf(M + (int64_t)(ncat + c) * FR);
            float l[N], r[N];
            if (op.kind == KIND_II) {
                const float *__restrict__ L = part + (int64_t)op.left * P.slot_stride + off;
#pragma unroll
                for (int y = 0; y < N; y++) l[y] = L[y * 64];
            }
            if (op.kind != KIND_TT) {
                const float *__restrict__ R = part + (int64_t)op.right * P.slot_stride + off;
#pragma unroll
                for (int y = 0; y < N; y++) r[y] = R[y * 64];
            }
            const float *tl = tabs + ((int64_t)sl * ncat + c) * N;
            const float *tr = tabs + (int64_t)(P.su + 1) * ncat * N + ((int64_t)sr * ncat + c) * N;
#pragma unroll
            for (int x = 0; x < N; x++) {
                float a, b;
                if (op.kind == KIND_II) {
                    a = 0.0f;
#pragma unroll
                    for (int y = 0; y < N; y++) a = fmaf(PL[x * N + y], l[y], a);
                } else {
                    a = tl[x];
                }
                if (op.kind == KIND_TT) {
                    b = tr[x];
                } else {
                    b = 0.0f;
#pragma unroll
                    for (int y = 0; y < N; y++) b = fmaf(PR[x * N + y], r[y], b);
                }
                out[ci][x] = a * b;
                wmax = fmaxf(wmax, fabsf(out[ci][x]));
            }
        }
        s_max[k & 1][w][lane] = wmax;
        __syncthreads();
        float pm = 0.0f;
        for (int ww = 0; ww < ncw; ww++) pm = fmaxf(pm, s_max[k & 1][ww][lane]);
        const int e = norm_exponent(pm);
        float *__restrict__ outp = part + (int64_t)op.out_slot * P.slot_stride;
#pragma unroll
        for (int ci = 0; ci < CPW; ci++) {
            const int c = w + ci * ncw;
            if (c >= ncat) break;
            const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
#pragma unroll
            for (int x = 0; x < N; x++) outp[off + x * 64] = ldexpf(out[ci][x], -e);
        }
        if (w == 0) {
            int es = e;
            if (op.kind == KIND_TI) es += expo[(int64_t)op.right * P.scale_stride + p];
            else if (op.kind == KIND_II)
                es += expo[(int64_t)op.left * P.scale_stride + p] + expo[(int64_t)op.right * P.scale_stride + p];
            expo[(int64_t)op.out_slot * P.scale_stride + p] = es;
        }
    }
    if (MODE == MODE_PARTIALS) return;

    const float *__restrict__ Mt = (const float *)P.mats + (int64_t)P.nops * P.mat_stride;
    const bool atip = P.br_a < 0;
    const int sa = atip ? P.states[(int64_t)(-P.br_a - 1) * P.nptn_pad + p] : 0;
    for (int c = w; c < ncat; c += ncw) {
        const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
        const cfptr Pt = as_constf(Mt + (int64_t)c * FR);
        const float *__restrict__ B = (const float *)P.part + (int64_t)P.br_b * P.slot_stride + off;
        const float *__restrict__ A = (const float *)P.part + (int64_t)(atip ? 0 : P.br_a) * P.slot_stride + off;
        float b[N];
#pragma unroll
        for (int y = 0; y < N; y++) b[y] = B[y * 64];
        float acc = 0.0f;
#pragma unroll
        for (int x = 0; x < N; x++) {
            float d = 0.0f;
#pragma unroll
            for (int y = 0; y < N; y++) d = fmaf(Pt[x * N + y], b[y], d);
            const float a = atip ? (float)P.tipvec[sa * N + x] : A[x * 64];
            acc = fmaf((float)P.pi[x] * a, d, acc);
        }
        s_cat[c][lane] = (double)acc;
    }
    __syncthreads();
    double r0 = 0.0;
    if (w == 0) {
        double lh = 0.0;
        for (int cc = 0; cc < ncat; cc++) lh += P.props[cc] * s_cat[cc][lane];
        const int e = expo[(int64_t)P.br_b * P.scale_stride + p] + (atip ? 0 : expo[(int64_t)P.br_a * P.scale_stride + p]);
        const double l = log_lh_expo(lh, e, P.invar[p]);
        if (p < P.nptn) P.ptn_lh[p] = l;
        r0 = wave_sum(p < P.nptn ? l * P.freq[p] : 0.0);
    }
    finish_tiles(P, r0, 0.0, tile, P.ntiles);
}

// ---------------------------------------------------------------------------------------
// theta for the FP32 form, in FP64: theta_c[i] = (sum_x pi_x U[x][i] A_c[x]) *
// (sum_y Uinv[i][y] B_c[y]), the eigen-space product the derivative kernels consume
// (the reference's theta_all, tree/phylokernelnew.h:2187-2193), followed by the first
// derivative evaluation.  buffer_scale_all = (e_A + e_B) ln 2.
// L3: also reduce the logL at the same length (third value), for the many-branch Newton.
template <int N, bool L3 = false>
__device__ inline bool theta_f32_body(const KParams &P, const int tile) {
    const int lane = threadIdx.x & 63;
    const int w = wave_id();
    const int ncw = nwaves(), ncat = P.ncat;
    const int64_t p = (int64_t)tile * 64 + lane;
    const bool atip = P.br_a < 0;
    const int sa = atip ? P.states[(int64_t)(-P.br_a - 1) * P.nptn_pad + p] : 0;
    const float *__restrict__ partf = (const float *)P.part;
    double *__restrict__ theta = (double *)P.theta;
    const cdptr piU = as_const(P.piU);
    const cdptr uinv = as_const(P.inv_evec);
    __shared__ double s_val[3][PLK_GPU_MAX_CATS * N];
    __shared__ double s_cat[3][PLK_GPU_MAX_CATS][64];
    for (int idx = threadIdx.x; idx < N * ncat; idx += blockDim.x) {
        const int c = idx / N, i = idx % N;
        const double cof = P.eval[i] * P.rates[c];
        const double val = exp(cof * P.br_len) * P.props[c];
        const double val1 = cof * val;  // the same products as fill_val (theta kernels)
        s_val[0][idx] = val;
        s_val[1][idx] = val1;
        s_val[2][idx] = cof * val1;
    }
    __syncthreads();
    // A pattern with invariant-site mass gets theta in the true (unscaled) frame and a zero
    // buffer_scale_all, as the reference's never-rescaled +I patterns have
    // (tree/phylokernelnew.h:1999, 2020), so every reader adds ptn_invar to a true value.
    // ldexp may take such a theta into the subnormals or to zero; the invariant mass then
    // dominates that pattern's likelihood.
    const int *expo = (const int *)P.scale;
    const int e_all = expo[(int64_t)P.br_b * P.scale_stride + p] + (atip ? 0 : expo[(int64_t)P.br_a * P.scale_stride + p]);
    const bool has_inv = P.invar[p] != 0.0;
    for (int c = w; c < ncat; c += ncw) {
        const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;  // theta (FP64 layout)
        const int64_t offf = ((int64_t)tile * ncat + c) * N * 64 + f32_loff(N, lane);
        double a[N], b[N];
        const float *__restrict__ B = partf + (int64_t)P.br_b * P.slot_stride + offf;
        const float *__restrict__ A = partf + (int64_t)(atip ? 0 : P.br_a) * P.slot_stride + offf;
#pragma unroll
        for (int x = 0; x < N; x++) {
            b[x] = (double)B[f32_xoff(N, x)];
            a[x] = atip ? P.tipvec[sa * N + x] : (double)A[f32_xoff(N, x)];
        }
        double l0 = 0.0, l1 = 0.0, l2 = 0.0;
        for_states<N>([&](const int i) {
            double ai = 0.0, bi = 0.0;
#pragma unroll
            for (int x = 0; x < N; x++) {
                ai = fma(piU[i * N + x], a[x], ai);
                bi = fma(uinv[i * N + x], b[x], bi);
            }
            const double th = has_inv ? ldexp(ai * bi, e_all) : ai * bi;
            if (theta) theta[off + (int64_t)i * 64] = th;  // null: derivatives only
            l0 = fma(s_val[0][c * N + i], th, l0);
            l1 = fma(s_val[1][c * N + i], th, l1);
            l2 = fma(s_val[2][c * N + i], th, l2);
        });
        s_cat[0][c][lane] = l0;
        s_cat[1][c][lane] = l1;
        s_cat[2][c][lane] = l2;
    }
    __syncthreads();
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    if (w == 0) {
        const int e = has_inv ? 0 : e_all;
        if (P.buf_scale) P.buf_scale[p] = e * kLn2;
        double lh = 0.0, d1 = 0.0, d2 = 0.0;
        for (int c = 0; c < ncat; c++) {
            lh += s_cat[0][c][lane];
            d1 += s_cat[1][c][lane];
            d2 += s_cat[2][c][lane];
        }
        lh = fabs(lh) + P.invar[p];
        const double f = (p < P.nptn) ? P.freq[p] : 0.0;
        const double inv = 1.0 / lh;
        const double dfr = d1 * inv, ddfr = d2 * inv;
        const double t1 = dfr * f, t2 = ddfr * f;
        if (P.g_frac && p < P.nptn) P.g_frac[p] = dfr;  // G_matrix row
        r0 = wave_sum(p < P.nptn ? t1 : 0.0);
        r1 = wave_sum(p < P.nptn ? t2 - t1 * dfr : 0.0);
        if (L3) r2 = wave_sum(p < P.nptn ? (log(lh) + e * kLn2) * f : 0.0);
    }
    if constexpr (L3) return finish_tiles3(P, r0, r1, r2, tile, P.ntiles);
    return finish_tiles(P, r0, r1, tile, P.ntiles);
}

}  // namespace dev
}  // namespace plk
