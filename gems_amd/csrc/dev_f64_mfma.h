// FP64 partial-likelihood traversal on the matrix cores (v_mfma_f64_16x16x4f64), 20 and
// 61 states, at most four rate categories.
//
// Per op and category the eigen-space partial (tree/phylokernelnew.h:1911-2036) is three
// dense products over a block of 16 patterns:
//   a = E_L * L,  b = E_R * R  (E = U exp(Lambda r t), n x n; L, R: n x 16 partials)
//   out = U^-1 (a o b)
// with tip children replaced by leaf-table rows (1729-1747, 1848-1890).  A workgroup owns
// one 16-pattern block; its RB wavefronts each own one 16-row block of the state
// dimension (rows padded to RB*16 with zeros).  A wavefront runs the two E products for
// its rows as chains of 16x16x4 tiles (k = 4 states per step), forms its rows of a o b and
// hands them to the others through LDS, then runs the U^-1 product for its output rows.
// The matrix operands are column-major n x n (E stored [c][i][x], U^-1 pre-swizzled into
// fragments), so each tile operand is one coalesced load; the partials keep the
// pattern-interleaved layout of the other kernels.
//
// Scaling as the reference: NORM checks max |out| over all categories of a pattern,
// SAFE each category separately, against 2^-256, only where ptn_invar is 0 and not for
// tip-tip ops; the maximum is exchanged between the row wavefronts through LDS before
// the outputs are stored.  The sums run in the matrix core's order, so the values differ
// from the vector kernels' at the rounding level (checked against the oracle's 1e-10).
//
// Lane layout of v_mfma_f64_16x16x4f64 (checked on the device): A[i][k] in lane
// i + 16 k, B[k][j] in lane j + 16 k, and D[i][j] in lane j + 16 (i % 4), register i / 4.
//
// A wavefront reads partial rows written by the other wavefronts of its workgroup in the
// previous op, so every op ends with a workgroup barrier.
#pragma once

#include "dev_common.h"

namespace plk {
namespace dev {

typedef double double4_t __attribute__((ext_vector_type(4)));

template <int N>
struct MfmaShape {
    static constexpr int RB = (N + 15) / 16;  // 16-row blocks of the state dimension
    static constexpr int KS = (N + 3) / 4;    // 4-state k steps
};

// Categories c0 .. c1-1 of the block: all of them under NORM scaling (one check over the
// pattern's categories), or one per workgroup under SAFE (each category checked alone).
// NC: categories a workgroup holds at most (1 for the per-category SAFE launch, 4 otherwise).
template <int N, int MODE, int NC>
__device__ inline void traverse_f64_mfma_body(const KParams &P, const int blk, const int c0, const int c1) {
    constexpr int RB = MfmaShape<N>::RB, KS = MfmaShape<N>::KS, ROWS = RB * 16;
    const int lane = threadIdx.x & 63;
    const int rb = wave_id();
    const int g = lane >> 4, col = lane & 15;
    const int tile = blk >> 2, pb = blk & 3;
    const int ncat = P.ncat;
    const int64_t p = (int64_t)tile * 64 + pb * 16 + col;  // this lane's pattern (column)
    const int pl = pb * 16 + col;                         // its position within the tile
    double *__restrict__ part = (double *)P.part;
    uint16_t *__restrict__ scale = (uint16_t *)P.scale;
    const double invar = P.invar[p];
    // U^-1 as A fragments [rb][ks][lane] = U^-1[16 rb + lane % 16][4 ks + lane / 16], zero
    // outside n x n (built on the host by plk_set_model); read per use (L1/L2 resident)
    const double *__restrict__ uf = P.uinv_frag + (int64_t)rb * KS * 64 + lane;

    __shared__ double s_t[NC][ROWS][16];  // a o b of each category, [c - c0][x][pattern]
    __shared__ double s_m[RB][NC][16];    // per row wave: max |out| per category and pattern

    for (int k = 0; k < P.nops; k++) {
        const int ko = op_index(P, tile, k);
        const DevOp op = load_op(P.ops, ko);
        const double *__restrict__ M = (const double *)P.mats + (int64_t)ko * P.mat_stride;
        const double *__restrict__ tabs = M + (int64_t)2 * ncat * N * N;
        with_kind(op.kind, [&](auto KIND_) {
            constexpr int KIND = decltype(KIND_)::value;
            int sl = 0, sr = 0;
            if (KIND <= KIND_TI) sl = P.states[(int64_t)op.left * P.nptn_pad + p];
            if (KIND == KIND_TT) sr = P.states[(int64_t)op.right * P.nptn_pad + p];
            for (int c = c0; c < c1; c++) {
                const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + pl;
                // one side: E * child over the k steps, or the leaf-table rows of a tip
                auto side = [&](const double *__restrict__ E, int slot, const double *__restrict__ tab) {
                    double4_t d = {0.0, 0.0, 0.0, 0.0};
                    if (tab) {
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int x = rb * 16 + 4 * r + g;
                            d[r] = x < N ? tab[x] : 0.0;
                        }
                        return d;
                    }
                    const double *__restrict__ C = part + (int64_t)slot * P.slot_stride + off;
                    const int x = rb * 16 + col;
                    // k steps in chunks of up to 8 (operands of a chunk loaded before its tiles)
                    constexpr int CH = KS < 8 ? KS : 8;
#pragma unroll
                    for (int k0 = 0; k0 < KS; k0 += CH) {
                        double av[CH], bv[CH];
#pragma unroll
                        for (int q = 0; q < CH; q++) {
                            const int i = 4 * (k0 + q) + g;
                            av[q] = (k0 + q < KS && x < N && i < N) ? E[i * N + x] : 0.0;
                            bv[q] = (k0 + q < KS && i < N) ? C[i * 64] : 0.0;
                        }
#pragma unroll
                        for (int q = 0; q < CH; q++)
                            if (k0 + q < KS) d = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], d, 0, 0, 0);
                    }
                    return d;
                };
                const double *EL = M + (int64_t)c * N * N;
                const double *ER = M + (int64_t)(ncat + c) * N * N;
                const double4_t a = side(EL, op.left, KIND <= KIND_TI ? tabs + (int64_t)sl * ncat * N + c * N : nullptr);
                const double4_t b = side(ER, op.right,
                                         KIND == KIND_TT ? tabs + (int64_t)(P.su + 1 + sr) * ncat * N + c * N : nullptr);
#pragma unroll
                for (int r = 0; r < 4; r++) s_t[c - c0][rb * 16 + 4 * r + g][col] = a[r] * b[r];
            }
            __syncthreads();

            // out = U^-1 (a o b): this wave's 16 output rows j = 16 rb + 4 g + r
            double4_t out[NC];
            const double *ufo = uf;  // re-read per op, not held in registers across ops
            asm volatile("" : "+v"(ufo));
#pragma unroll
            for (int cc = 0; cc < NC; cc++) {
                out[cc] = double4_t{0.0, 0.0, 0.0, 0.0};
                if (c0 + cc >= c1) continue;
#pragma unroll
                for (int ks = 0; ks < KS; ks++)
                    out[cc] = __builtin_amdgcn_mfma_f64_16x16x4f64(ufo[ks * 64], s_t[cc][4 * ks + g][col], out[cc], 0, 0, 0);
                double m = 0.0;
#pragma unroll
                for (int r = 0; r < 4; r++) m = fmax(m, fabs(out[cc][r]));
                m = fmax(m, __shfl_xor(m, 16, 64));
                m = fmax(m, __shfl_xor(m, 32, 64));
                if (g == 0) s_m[rb][cc][col] = m;
            }
            __syncthreads();

            // scaling decision per pattern (NORM) or per pattern and category (SAFE)
            bool up[NC];
            if (!P.safe) {
                double pm = 0.0;
                for (int w = 0; w < RB; w++)
                    for (int cc = 0; cc < c1 - c0; cc++) pm = fmax(pm, s_m[w][cc][col]);
                const bool u = KIND != KIND_TT && pm < kST && invar == 0.0;
#pragma unroll
                for (int cc = 0; cc < NC; cc++) up[cc] = u;
            } else {
#pragma unroll
                for (int cc = 0; cc < NC; cc++) {
                    double pm = 0.0;
                    if (c0 + cc < c1)
                        for (int w = 0; w < RB; w++) pm = fmax(pm, s_m[w][cc][col]);
                    up[cc] = KIND != KIND_TT && pm < kST && invar == 0.0;
                }
            }
            double *__restrict__ outp = part + (int64_t)op.out_slot * P.slot_stride;
#pragma unroll
            for (int cc = 0; cc < NC; cc++) {
                if (c0 + cc >= c1) continue;
                const int64_t off = ((int64_t)tile * ncat + c0 + cc) * N * 64 + pl;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int j = rb * 16 + 4 * r + g;
                    if (j < N) outp[off + j * 64] = up[cc] ? out[cc][r] * kTwo256 : out[cc][r];
                }
            }
            // scale counts: s_out = children's counts (+1 where rescaled), 1643-1644,
            // 1757-1760, 1929-1936
            if (rb == 0 && g == 0) {
                if (!P.safe) {
                    unsigned sc = 0;
                    if (KIND == KIND_TI) sc = scale[(int64_t)op.right * P.scale_stride + p];
                    else if (KIND == KIND_II)
                        sc = (unsigned)scale[(int64_t)op.left * P.scale_stride + p] + scale[(int64_t)op.right * P.scale_stride + p];
                    scale[(int64_t)op.out_slot * P.scale_stride + p] = (uint16_t)(sc + (up[0] ? 1u : 0u));
                } else {
#pragma unroll
                    for (int cc = 0; cc < NC; cc++) {
                        if (c0 + cc >= c1) continue;
                        const int64_t a = p * ncat + c0 + cc;
                        unsigned sc = 0;
                        if (KIND == KIND_TI) sc = scale[(int64_t)op.right * P.scale_stride + a];
                        else if (KIND == KIND_II)
                            sc = (unsigned)scale[(int64_t)op.left * P.scale_stride + a] + scale[(int64_t)op.right * P.scale_stride + a];
                        scale[(int64_t)op.out_slot * P.scale_stride + a] = (uint16_t)(sc + (up[cc] ? 1u : 0u));
                    }
                }
            }
        });
        __syncthreads();  // this op's rows (all waves) before the next op reads them
    }
}

}  // namespace dev
}  // namespace plk
