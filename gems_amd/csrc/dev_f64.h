// FP64 likelihood kernel bodies (reference semantics, eigen-space partials).
//
// Layout: a tile is 64 consecutive patterns, one wavefront lane per pattern; a
// workgroup owns one tile and its wavefronts split the categories (wave w handles
// c = w, w + nwaves, ...).  Element (ptn, c, x) of a partial is at
// ((tile*ncat + c)*n + x)*64 + lane, so every load/store of one (c, x) by a wavefront
// is one contiguous 512-byte segment.  All tiles are independent across the whole
// traversal (the reference's packets walk every node for their pattern slice,
// tree/phylokernelnew.h:2851-2853), so one launch runs the full post-order op list with
// no grid-wide barrier; each thread only re-reads values it wrote itself.
#pragma once

#include "dev_common.h"

namespace plk {
namespace dev {

// ---------------------------------------------------------------------------------------
// P(t) kernel (computePartialInfo, tree/phylokernelnew.h:1013-1063): for one child of one
// op, echild[c][x][i] = U[x][i] exp(lambda_i r_c t) (held as [c][i][x] for n <= 20), and
// for a leaf child the table
// leaf[s][c][x] = sum_i echild[c][x][i] tip_lh[s][i] with the unknown-state row = 1.
// ---------------------------------------------------------------------------------------
// 61-state (codon) op with its matrices in LDS.  A 61-double column does not fit the
// scalar registers, and one scalar load per element leaves the single wavefront of a SIMD
// waiting on nearly every FMA.  Here the workgroup copies the op's matrices into LDS,
// one phase at a time: the right child's echild of every category as [c][i][x], then the
// left's, then U^-1 as [j][x], rows padded to 62 doubles so that every row starts 16-byte
// aligned.  Each wavefront (one category) reads them as broadcast 16-byte LDS loads,
// with its child's values streamed from memory a block of 8 ahead.  Same sums in the
// same order as op_partial, so the stored values are identical.  ncat <= 4.
constexpr int kRow61 = 62;                  // padded row length (doubles)
constexpr int kLds61 = 4 * 61 * kRow61;     // doubles: four categories' matrices

// echild of one side for every category, [c][i][x] in memory, into E as padded rows
__device__ __forceinline__ void stage61_echild(const double *__restrict__ src, int ncat, double *E) {
    __syncthreads();  // every wavefront is done with the previous contents
    for (int idx = threadIdx.x; idx < ncat * 61 * 61; idx += blockDim.x) {
        const int c = idx / (61 * 61), rem = idx - c * 61 * 61, i = rem / 61, x = rem - i * 61;
        E[(c * 61 + i) * kRow61 + x] = src[idx];
    }
    __syncthreads();
}

__device__ __forceinline__ void stage61_uinv(const double *__restrict__ uinv, double *E) {
    __syncthreads();
    for (int idx = threadIdx.x; idx < 61 * 61; idx += blockDim.x) {
        const int j = idx / 61, x = idx - j * 61;
        E[j * kRow61 + x] = uinv[idx];
    }
    __syncthreads();
}

// acc[x] = sum_i Ec[i][X0 + x] v[i] over i = 0..60 in ascending order, v[i] = V[i * 64];
// X0 is even, so a row segment starts 16-byte aligned; an odd NX reads one padding value.
template <int X0, int NX>
__device__ __forceinline__ void cols61(const double *Ec, const double *__restrict__ V, double (&acc)[NX]) {
#pragma unroll
    for (int x = 0; x < NX; x++) acc[x] = 0.0;
    double nxt[8];
#pragma unroll
    for (int u = 0; u < 8; u++) nxt[u] = V[u * 64];
    for (int i0 = 0; i0 < 61; i0 += 8) {
        double cur[8];
#pragma unroll
        for (int u = 0; u < 8; u++) cur[u] = nxt[u];
        // the next block (indices clamped into the partial; values past 60 are not used)
#pragma unroll
        for (int u = 0; u < 8; u++) nxt[u] = V[min(i0 + 8 + u, 60) * 64];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (i0 + u >= 61) break;
            const double2 *row = (const double2 *)(Ec + (i0 + u) * kRow61 + X0);
#pragma unroll
            for (int x = 0; x < NX; x += 2) {
                const double2 e = row[x / 2];
                acc[x] = fma(e.x, cur[u], acc[x]);
                if (x + 1 < NX) acc[x + 1] = fma(e.y, cur[u], acc[x + 1]);
            }
        }
    }
}

// out[j] = sum_x Ui[j][x] t[x] (ascending x), four rows at a time, each value stored as
// formed; returns max_j |out[j]|.
__device__ __forceinline__ double rows61(const double *Ui, const double (&t)[61], double *__restrict__ dst) {
    double m = 0.0;
    for (int j0 = 0; j0 < 60; j0 += 4) {
        double o[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int x = 0; x < 61; x += 2) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double2 e = ((const double2 *)(Ui + (j0 + q) * kRow61))[x / 2];
                o[q] = fma(e.x, t[x], o[q]);
                if (x + 1 < 61) o[q] = fma(e.y, t[x + 1], o[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            m = fmax(m, fabs(o[q]));
            dst[(j0 + q) * 64] = o[q];
        }
    }
    double o = 0.0;
#pragma unroll
    for (int x = 0; x < 61; x += 2) {
        const double2 e = ((const double2 *)(Ui + 60 * kRow61))[x / 2];
        o = fma(e.x, t[x], o);
        if (x + 1 < 61) o = fma(e.y, t[x + 1], o);
    }
    dst[60 * 64] = o;
    return fmax(m, fabs(o));
}

// One op for category c of this wavefront (mine = c < ncat); every wavefront of the
// workgroup calls it (the staging has barriers).  Returns max_j |out[j]| (0 if not mine).
__device__ __forceinline__ double op_partial_lds61(const KParams &P, const DevOp &op, const double *__restrict__ M,
                                                   bool mine, int c, int64_t off, int sl, int sr,
                                                   double *__restrict__ dst, double *E) {
    constexpr int N = 61;
    const int ncat = P.ncat, block = N * ncat;
    const double *__restrict__ part = (const double *)P.part;
    const double *__restrict__ tabs = M + (int64_t)2 * ncat * N * N;
    const double *Ec = E + c * 61 * kRow61;
    double t[N];
    if (op.kind == KIND_TT) {
        if (mine) {
            const double *__restrict__ tl = tabs + (int64_t)sl * block + c * N;
            const double *__restrict__ tr = tabs + (int64_t)(P.su + 1) * block + (int64_t)sr * block + c * N;
#pragma unroll
            for (int x = 0; x < N; x++) t[x] = tl[x] * tr[x];
        }
    } else {
        stage61_echild(M + (int64_t)ncat * N * N, ncat, E);  // right child's echild
        if (mine) {
            cols61<0, 61>(Ec, part + (int64_t)op.right * P.slot_stride + off, t);
            if (op.kind == KIND_TI) {
                const double *__restrict__ tl = tabs + (int64_t)sl * block + c * N;
#pragma unroll
                for (int x = 0; x < N; x++) t[x] = tl[x] * t[x];
            }
        }
        if (op.kind == KIND_II) {
            stage61_echild(M, ncat, E);  // left child's echild
            if (mine) {
                const double *__restrict__ L = part + (int64_t)op.left * P.slot_stride + off;
                {
                    double a[32];
                    cols61<0, 32>(Ec, L, a);
#pragma unroll
                    for (int x = 0; x < 32; x++) t[x] = a[x] * t[x];
                }
                {
                    double a[29];
                    cols61<32, 29>(Ec, L, a);
#pragma unroll
                    for (int x = 0; x < 29; x++) t[32 + x] = a[x] * t[32 + x];
                }
            }
        }
    }
    stage61_uinv(P.inv_evec, E);
    return mine ? rows61(E, t, dst) : 0.0;
}

// cat >= 0: category cat only (the job's categories then run in separate workgroups).
__device__ inline void pmatrix_f64_body(const KParams &P, const PJob &j, const int cat = -1) {
    const int n = P.n, ncat = P.ncat, block = n * ncat, su = P.su;
    const int cb = cat >= 0 ? cat : 0, nc = cat >= 0 ? 1 : ncat;  // categories cb .. cb+nc-1
    double *M = (double *)P.mats + (int64_t)j.op * P.mat_stride;
    double *E = M + (int64_t)j.side * ncat * n * n;
    __shared__ double ex[PLK_GPU_MAX_CATS * PLK_MAX_STATES];
    for (int idx = threadIdx.x; idx < nc * n; idx += blockDim.x) {
        const int c = cb + idx / n, i = idx % n;
        ex[c * n + i] = exp(P.eval[i] * (P.rates[c] * j.len));
    }
    __syncthreads();
    // stored [c][i][x]: a column per i, x contiguous (matrix-core fragments and the
    // column-at-a-time kernels read it that way)
    for (int idx = threadIdx.x; idx < nc * n * n; idx += blockDim.x) {
        const int c = cb + idx / (n * n), rem = idx % (n * n), x = rem / n, i = rem % n;
        E[(int64_t)c * n * n + i * n + x] = P.evec[x * n + i] * ex[c * n + i];
    }
    if (j.leaf) {
        double *T = M + (int64_t)2 * ncat * n * n + (int64_t)j.side * (su + 1) * block;
        // leaf[s][c][x]; four entries per thread at a time: independent sums, each in its
        // own order
        const int nt = blockDim.x, per_s = nc * n, total = su * per_s;
        for (int base = threadIdx.x; base < total; base += 4 * nt) {
            int ss[4], cs[4], xs[4];
            double v[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int idx = min(base + q * nt, total - 1), rem = idx % per_s;
                ss[q] = idx / per_s;
                cs[q] = cb + rem / n;
                xs[q] = rem % n;
                v[q] = 0.0;
            }
            for (int i = 0; i < n; i++) {
#pragma unroll
                for (int q = 0; q < 4; q++)
                    v[q] = fma(P.evec[xs[q] * n + i] * ex[cs[q] * n + i], P.tip_lh[ss[q] * n + i], v[q]);
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (base + q * nt < total) T[(int64_t)ss[q] * block + cs[q] * n + xs[q]] = v[q];
        }
        for (int idx = threadIdx.x; idx < nc * n; idx += blockDim.x) T[(int64_t)su * block + cb * n + idx] = 1.0;
    }
}

// ---------------------------------------------------------------------------------------
// One (pattern, category) of one op: tmp[x] = (eL[x].L) * (eR[x].R) with tip children
// replaced by their leaf-table rows, then out = Uinv tmp
// (tree/phylokernelnew.h:1729-1747, 1848-1890, 1979-1994).  Matrices are wave-uniform
// (scalar loads); partials are coalesced 64-lane loads.
template <int N>
__device__ __forceinline__ void op_partial(const KParams &P, const DevOp &op, const double *__restrict__ M, int c,
                                           int64_t off, int sl, int sr, double (&out)[N]) {
    const int ncat = P.ncat, block = N * ncat;
    const double *__restrict__ part = (const double *)P.part;
    const cdptr uinv = as_const(P.inv_evec);
    const cdptr eL = as_const(M + (int64_t)c * N * N);
    const cdptr eR = as_const(M + (int64_t)(ncat + c) * N * N);
    const double *__restrict__ tabs = M + (int64_t)2 * ncat * N * N;
#pragma unroll
    for (int j = 0; j < N; j++) out[j] = 0.0;
    if (op.kind == KIND_II) {
        const double *__restrict__ L = part + (int64_t)op.left * P.slot_stride + off;
        const double *__restrict__ R = part + (int64_t)op.right * P.slot_stride + off;
        double l[N], r[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            l[i] = L[i * 64];
            r[i] = R[i * 64];
        }
        for_states<N>([&](const int x) {
            double a = 0.0, b = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) {
                a = fma(eL[i * N + x], l[i], a);
                b = fma(eR[i * N + x], r[i], b);
            }
            const double t = a * b;
#pragma unroll
            for (int j = 0; j < N; j++) out[j] = fma(uinv[j * N + x], t, out[j]);
        });
    } else if (op.kind == KIND_TI) {
        const double *__restrict__ R = part + (int64_t)op.right * P.slot_stride + off;
        const double *__restrict__ tl = tabs + (int64_t)sl * block + c * N;
        double r[N];
#pragma unroll
        for (int i = 0; i < N; i++) r[i] = R[i * 64];
        for_states<N>([&](const int x) {
            double b = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) b = fma(eR[i * N + x], r[i], b);
            const double t = tl[x] * b;
#pragma unroll
            for (int j = 0; j < N; j++) out[j] = fma(uinv[j * N + x], t, out[j]);
        });
    } else {
        const double *__restrict__ tl = tabs + (int64_t)sl * block + c * N;
        const double *__restrict__ tr = tabs + (int64_t)(P.su + 1) * block + (int64_t)sr * block + c * N;
        for_states<N>([&](const int x) {
            const double t = tl[x] * tr[x];
#pragma unroll
            for (int j = 0; j < N; j++) out[j] = fma(uinv[j * N + x], t, out[j]);
        });
    }
}

// A column of K (2, 4, 10 or 20) wave-uniform doubles read into scalar registers by
// back-to-back scalar loads that are waited on once.  Left to itself the compiler streams
// such values through a single 16-register tuple, one memory round trip per 8 values;
// here a whole column costs one.  The loads and their wait sit in one asm statement, so
// no register is read before its data has arrived; the outputs are early-clobber, since
// later loads of the statement still read the base address while earlier ones are
// writing their destination registers.  Callers fence each column's FMAs
// with scheduling barriers (one column live at a time) and pass an opaque base, so the
// column addresses are not hoisted into registers for the whole loop.
typedef unsigned su4 __attribute__((ext_vector_type(4)));
typedef unsigned su8 __attribute__((ext_vector_type(8)));
typedef unsigned su16 __attribute__((ext_vector_type(16)));

template <int K>
struct SCol;
template <>
struct SCol<20> {
    su16 a, b;
    su8 c;
    __device__ __forceinline__ SCol(const double *p) {
        asm volatile("s_load_dwordx16 %0, %3, 0x0\n\ts_load_dwordx16 %1, %3, 0x40\n\t"
                     "s_load_dwordx8 %2, %3, 0x80\n\ts_waitcnt lgkmcnt(0)"
                     : "=&s"(a), "=&s"(b), "=&s"(c)
                     : "s"(p));
    }
    __device__ __forceinline__ double operator[](int i) const {
        const unsigned lo = i < 8 ? a[2 * i] : i < 16 ? b[2 * i - 16] : c[2 * i - 32];
        const unsigned hi = i < 8 ? a[2 * i + 1] : i < 16 ? b[2 * i - 15] : c[2 * i - 31];
        return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    }
};
template <>
struct SCol<10> {
    su16 a;
    su4 b;
    __device__ __forceinline__ SCol(const double *p) {
        asm volatile("s_load_dwordx16 %0, %2, 0x0\n\ts_load_dwordx4 %1, %2, 0x40\n\ts_waitcnt lgkmcnt(0)"
                     : "=&s"(a), "=&s"(b)
                     : "s"(p));
    }
    __device__ __forceinline__ double operator[](int i) const {
        const unsigned lo = i < 8 ? a[2 * i] : b[2 * i - 16];
        const unsigned hi = i < 8 ? a[2 * i + 1] : b[2 * i - 15];
        return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    }
};
template <>
struct SCol<4> {
    su8 a;
    __device__ __forceinline__ SCol(const double *p) {
        asm volatile("s_load_dwordx8 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=&s"(a) : "s"(p));
    }
    __device__ __forceinline__ double operator[](int i) const {
        return __builtin_bit_cast(double, ((unsigned long long)a[2 * i + 1] << 32) | a[2 * i]);
    }
};
template <>
struct SCol<2> {
    su4 a;
    __device__ __forceinline__ SCol(const double *p) {
        asm volatile("s_load_dwordx4 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=&s"(a) : "s"(p));
    }
    __device__ __forceinline__ double operator[](int i) const {
        return __builtin_bit_cast(double, ((unsigned long long)a[2 * i + 1] << 32) | a[2 * i]);
    }
};

// Same arithmetic as op_partial for N <= 20: echild is held column-major ([i][x]), so
// each step of (1) t[x] = eR[x].R and (2) t[x] *= eL[x].L (in two halves of x) reads one
// contiguous column of wave-uniform values (SCol) and feeds independent FMAs; (3) out[j] =
// Uinv[j].t reads U^-1 a row at a time and is stored as formed.  Every sum runs over its index in
// ascending order from 0, as in op_partial, so the stored values are identical.  Returns
// max_j |out[j]| (the scaling check is applied to the stored values by the caller).
template <int N>
__device__ __forceinline__ double op_partial_store(const KParams &P, const DevOp &op, const double *__restrict__ M,
                                                   int c, int64_t off, int sl, int sr, double *__restrict__ dst) {
    static_assert(N <= 20, "column form is for unrolled state loops");
    const int ncat = P.ncat, block = N * ncat;
    const double *__restrict__ part = (const double *)P.part;
    static_assert(N % 2 == 0, "columns are split in two equal halves");
    const double *uinv = P.inv_evec;
    const double *eL = M + (int64_t)c * N * N;
    const double *eR = M + (int64_t)(ncat + c) * N * N;
    const double *__restrict__ tabs = M + (int64_t)2 * ncat * N * N;
    double t[N];
    if (op.kind == KIND_TT) {
        const double *__restrict__ tl = tabs + (int64_t)sl * block + c * N;
        const double *__restrict__ tr = tabs + (int64_t)(P.su + 1) * block + (int64_t)sr * block + c * N;
#pragma unroll
        for (int x = 0; x < N; x++) t[x] = tl[x] * tr[x];
    } else {
        {
            const double *__restrict__ R = part + (int64_t)op.right * P.slot_stride + off;
            double r[N];
#pragma unroll
            for (int i = 0; i < N; i++) r[i] = R[i * 64];
#pragma unroll
            for (int x = 0; x < N; x++) t[x] = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) {
                const SCol<N> col(opaque(eR) + i * N);
#pragma unroll
                for (int x = 0; x < N; x++) t[x] = fma(col[x], r[i], t[x]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // keep the left child's loads behind phase (1): the right child's registers are
        // free by then
        __builtin_amdgcn_sched_barrier(0);
        if (op.kind == KIND_II) {
            const double *__restrict__ L = part + (int64_t)op.left * P.slot_stride + off;
            double l[N];
#pragma unroll
            for (int i = 0; i < N; i++) l[i] = L[i * 64];
            constexpr int H = N / 2;
#pragma unroll
            for (int h = 0; h < N; h += H) {
                double a[H];
#pragma unroll
                for (int x = 0; x < H; x++) a[x] = 0.0;
#pragma unroll
                for (int i = 0; i < N; i++) {
                    const SCol<H> col(opaque(eL) + i * N + h);
#pragma unroll
                    for (int x = 0; x < H; x++) a[x] = fma(col[x], l[i], a[x]);
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int x = 0; x < H; x++) t[h + x] = a[x] * t[h + x];
            }
        } else {
            const double *__restrict__ tl = tabs + (int64_t)sl * block + c * N;
#pragma unroll
            for (int x = 0; x < N; x++) t[x] = tl[x] * t[x];
        }
    }
    // (3) out[j] = Uinv[j].t, a row of U^-1 at a time (one accumulator per output)
    double m = 0.0;
#pragma unroll
    for (int j = 0; j < N; j++) {
        const SCol<N> row(opaque(uinv) + j * N);
        double o = 0.0;
#pragma unroll
        for (int x = 0; x < N; x++) o = fma(row[x], t[x], o);
        m = fmax(m, fabs(o));
        dst[j * 64] = o;
        __builtin_amdgcn_sched_barrier(0);
    }
    return m;
}

// Fill the branch vectors val0/1/2[c][i] = (1, cof, cof^2) * exp(cof t) * prop_c with
// cof = lambda_i r_c (tree/phylokernelnew.h:2321-2367, 2731-2766).
__device__ __forceinline__ void fill_val(const KParams &P, double len, double *v0, double *v1, double *v2) {
    const int n = P.n, block = n * P.ncat;
    for (int idx = threadIdx.x; idx < block; idx += blockDim.x) {
        int c = idx / n, i = idx % n;
        double cof = P.eval[i] * P.rates[c];
        double val = exp(cof * len) * P.props[c];
        v0[idx] = val;
        if (v1) {
            double val1 = cof * val;
            v1[idx] = val1;
            v2[idx] = cof * val1;
        }
    }
}

// SAFE per-category factor relative to the minimum scale (2918-2944): 1, 2^-256 or 0.
__device__ __forceinline__ double safe_factor(int d) { return d == 0 ? 1.0 : (d == 1 ? kST : 0.0); }

// L3 (with MODE_THETA): also reduce the logL at the same length as a third value (the
// value computeLikelihoodFromBuffer would return there), for the many-branch Newton.
// Returns true in the workgroup that wrote the result.
// E61: the 61-state ops take op_partial_lds61 (ncat <= 4; the workgroup's LDS then holds
// the 121 KB matrix buffer).
template <int N, int MODE, bool L3 = false, bool E61 = false>
__device__ inline bool traverse_f64_body(const KParams &P, const int tile) {
    static_assert(!E61 || N == 61, "the LDS matrix path is the 61-state one");
    const int lane = threadIdx.x & 63;
    const int w = wave_id();
    const int ncw = nwaves(), ncat = P.ncat;
    const int64_t p = (int64_t)tile * 64 + lane;
    double *__restrict__ part = (double *)P.part;
    uint16_t *__restrict__ scale = (uint16_t *)P.scale;
    const double invar = P.invar[p];
    __shared__ double s_max[2][4][64];
    __shared__ double s_val[3][PLK_GPU_MAX_CATS * N];
    __shared__ double s_cat[3][PLK_GPU_MAX_CATS][64];
    __shared__ int s_sc[PLK_GPU_MAX_CATS][64];
    __shared__ __attribute__((aligned(16))) double s_e61[E61 ? kLds61 : 2];

    if (MODE != MODE_PARTIALS) fill_val(P, P.br_len, s_val[0], MODE == MODE_THETA ? s_val[1] : nullptr, s_val[2]);

    // -------- post-order traversal (computePartialLikelihood for every op)
    for (int k = 0; k < P.nops; k++) {
        const int ko = op_index(P, tile, k);
        const DevOp op = load_op(P.ops, ko);
        const double *__restrict__ M = (const double *)P.mats + (int64_t)ko * P.mat_stride;
        double *__restrict__ outp = part + (int64_t)op.out_slot * P.slot_stride;
        int sl = 0, sr = 0;
        if (op.kind <= KIND_TI) sl = P.states[(int64_t)op.left * P.nptn_pad + p];
        if (op.kind == KIND_TT) sr = P.states[(int64_t)op.right * P.nptn_pad + p];
        // children's NORM scale counts, fetched before the partial work so that their
        // latency overlaps it (wave 0 keeps the per-pattern count)
        unsigned sc_norm = 0;
        if (!P.safe && w == 0) {
            if (op.kind == KIND_TI) sc_norm = scale[(int64_t)op.right * P.scale_stride + p];
            else if (op.kind == KIND_II)
                sc_norm = (unsigned)scale[(int64_t)op.left * P.scale_stride + p] + scale[(int64_t)op.right * P.scale_stride + p];
        }
        double wmax = 0.0;
        if constexpr (E61) {
            // one category per wavefront; the SAFE handling of the N <= 20 path below
            const int c = w;
            const bool mine = c < ncat;
            const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
            const int64_t a = p * ncat + c;
            unsigned sc = 0;
            if (P.safe && mine) {
                if (op.kind == KIND_TI) sc = scale[(int64_t)op.right * P.scale_stride + a];
                else if (op.kind == KIND_II)
                    sc = (unsigned)scale[(int64_t)op.left * P.scale_stride + a] + scale[(int64_t)op.right * P.scale_stride + a];
            }
            const double m = op_partial_lds61(P, op, M, mine, c, off, sl, sr, outp + off, s_e61);
            if (P.safe && mine) {
                if (op.kind != KIND_TT && m < kST && invar == 0.0) {
#pragma unroll
                    for (int j = 0; j < N; j++) outp[off + j * 64] *= kTwo256;
                    sc += 1;
                }
                scale[(int64_t)op.out_slot * P.scale_stride + a] = (uint16_t)sc;
            }
            wmax = m;
        }
        for (int c = w; c < (E61 ? 0 : ncat); c += ncw) {
            const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
            if constexpr (N <= 20) {
                // SAFE: this category's children counts, also fetched ahead of the work
                unsigned sc = 0;
                if (P.safe) {
                    const int64_t a = p * ncat + c;
                    if (op.kind == KIND_TI) sc = scale[(int64_t)op.right * P.scale_stride + a];
                    else if (op.kind == KIND_II)
                        sc = (unsigned)scale[(int64_t)op.left * P.scale_stride + a] + scale[(int64_t)op.right * P.scale_stride + a];
                }
                const double m = op_partial_store<N>(P, op, M, c, off, sl, sr, outp + off);
                if (P.safe) {
                    const int64_t a = p * ncat + c;
                    if (op.kind != KIND_TT && m < kST && invar == 0.0) {
#pragma unroll
                        for (int j = 0; j < N; j++) outp[off + j * 64] *= kTwo256;
                        sc += 1;
                    }
                    scale[(int64_t)op.out_slot * P.scale_stride + a] = (uint16_t)sc;
                }
                wmax = fmax(wmax, m);
                continue;
            }
            double out[N];
            op_partial<N>(P, op, M, c, off, sl, sr, out);
            double m = 0.0;
#pragma unroll
            for (int j = 0; j < N; j++) m = fmax(m, fabs(out[j]));
            if (P.safe) {
                // per-category counts and check (1950-2013, 1756-1760, 1871-1886)
                const int64_t a = p * ncat + c;
                unsigned sc = 0;
                if (op.kind == KIND_TI) sc = scale[(int64_t)op.right * P.scale_stride + a];
                else if (op.kind == KIND_II)
                    sc = (unsigned)scale[(int64_t)op.left * P.scale_stride + a] + scale[(int64_t)op.right * P.scale_stride + a];
                if (op.kind != KIND_TT && m < kST && invar == 0.0) {
#pragma unroll
                    for (int j = 0; j < N; j++) out[j] *= kTwo256;
                    sc += 1;
                }
                scale[(int64_t)op.out_slot * P.scale_stride + a] = (uint16_t)sc;
            }
            wmax = fmax(wmax, m);
#pragma unroll
            for (int j = 0; j < N; j++) outp[off + j * 64] = out[j];
        }
        if (!P.safe) {
            // one check over all categories of the pattern (2019-2034, 1893-1907)
            s_max[k & 1][w][lane] = wmax;
            __syncthreads();
            double pm = 0.0;
            for (int ww = 0; ww < ncw; ww++) pm = fmax(pm, s_max[k & 1][ww][lane]);
            const bool do_scale = (op.kind != KIND_TT) && pm < kST && invar == 0.0;
            if (do_scale) {
                for (int c = w; c < ncat; c += ncw) {
                    const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
#pragma unroll
                    for (int j = 0; j < N; j++) outp[off + j * 64] *= kTwo256;
                }
            }
            if (w == 0) scale[(int64_t)op.out_slot * P.scale_stride + p] = (uint16_t)(sc_norm + (do_scale ? 1u : 0u));
        }
    }
    if (MODE == MODE_PARTIALS) return false;

    // -------- branch stage: node refs a (slot or tip) and b (slot)
    __syncthreads();  // s_val complete
    const bool atip = P.br_a < 0;
    const int sa = atip ? P.states[(int64_t)(-P.br_a - 1) * P.nptn_pad + p] : 0;
    const double *__restrict__ A = atip ? nullptr : part + (int64_t)P.br_a * P.slot_stride;
    const double *__restrict__ B = part + (int64_t)P.br_b * P.slot_stride;
    const uint16_t *__restrict__ scA = atip ? nullptr : scale + (int64_t)P.br_a * P.scale_stride;
    const uint16_t *__restrict__ scB = scale + (int64_t)P.br_b * P.scale_stride;
    const double *__restrict__ tip = P.tip_lh + sa * N;
    double *__restrict__ theta = (double *)P.theta;

    for (int c = w; c < ncat; c += ncw) {
        const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
        const double *__restrict__ v0 = s_val[0] + c * N;
        if (MODE == MODE_LOGL) {
            // lh_cat = sum_i val * node * dad (dotProduct3Vec 3044-3057; tip 2903-2914)
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) {
                const double a = atip ? tip[i] : A[off + i * 64];
                acc = fma(v0[i] * a, B[off + i * 64], acc);
            }
            s_cat[0][c][lane] = acc;
        } else {
            // theta = node * dad (2187-2193; tip 2104-2145), then the triple dot (2499-2501)
            const double *__restrict__ v1 = s_val[1] + c * N;
            const double *__restrict__ v2 = s_val[2] + c * N;
            double l0 = 0.0, l1 = 0.0, l2 = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) {
                const double a = atip ? tip[i] : A[off + i * 64];
                const double th = a * B[off + i * 64];
                if (theta) theta[off + i * 64] = th;  // null: derivatives only (many branches)
                l0 = fma(v0[i], th, l0);
                l1 = fma(v1[i], th, l1);
                l2 = fma(v2[i], th, l2);
            }
            s_cat[0][c][lane] = l0;
            s_cat[1][c][lane] = l1;
            s_cat[2][c][lane] = l2;
        }
        if (P.safe) s_sc[c][lane] = (int)scB[p * ncat + c] + (atip ? 0 : (int)scA[p * ncat + c]);
    }
    __syncthreads();

    int sc_min = 0;
    if (P.safe) {
        sc_min = s_sc[0][lane];
        for (int c = 1; c < ncat; c++) sc_min = min(sc_min, s_sc[c][lane]);
        if (MODE == MODE_THETA && theta) {
            // rescale the stored theta of categories above the minimum (2146-2171, 2194-2225)
            for (int c = w; c < ncat; c += ncw) {
                const int d = s_sc[c][lane] - sc_min;
                if (d != 0) {
                    const double f = safe_factor(d);
                    const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
#pragma unroll
                    for (int i = 0; i < N; i++) theta[off + i * 64] *= f;
                }
            }
        }
    }
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    if (w == 0) {
        double sc;
        double lh = 0.0, d1 = 0.0, d2 = 0.0;
        if (P.safe) {
            sc = sc_min;
            for (int c = 0; c < ncat; c++) {
                const double f = safe_factor(s_sc[c][lane] - sc_min);
                lh += s_cat[0][c][lane] * f;
                if (MODE == MODE_THETA) {
                    d1 += s_cat[1][c][lane] * f;
                    d2 += s_cat[2][c][lane] * f;
                }
            }
        } else {
            sc = (double)((unsigned)scB[p] + (atip ? 0u : (unsigned)scA[p]));
            for (int c = 0; c < ncat; c++) {
                lh += s_cat[0][c][lane];
                if (MODE == MODE_THETA) {
                    d1 += s_cat[1][c][lane];
                    d2 += s_cat[2][c][lane];
                }
            }
        }
        lh = fabs(lh) + invar;
        const double f = (p < P.nptn) ? P.freq[p] : 0.0;
        if (MODE == MODE_LOGL) {
            // 3096-3108; side outputs _pattern_lh_cat (3023, SAFE-adjusted 3076-3083) and
            // the save_log_value = false scaling (3103-3104)
            const double l = log(lh) + sc * kLogST;
            if (p < P.nptn) {
                P.ptn_lh[p] = l;
                if (P.ptn_lh_cat) {
                    for (int c = 0; c < ncat; c++)
                        P.ptn_lh_cat[p * ncat + c] = P.safe ? s_cat[0][c][lane] * safe_factor(s_sc[c][lane] - sc_min)
                                                            : s_cat[0][c][lane];
                    P.ptn_scaling[p] = sc * kLogST;
                }
            }
            r0 = wave_sum(p < P.nptn ? l * f : 0.0);
        } else {
            // buffer_scale_all (2178-2179, 2231-2232) and df/ddf (2504-2515)
            if (P.buf_scale) P.buf_scale[p] = sc * kLogST;
            const double inv = 1.0 / lh;
            const double dfr = d1 * inv, ddfr = d2 * inv;
            const double t1 = dfr * f, t2 = ddfr * f;
            if (P.g_frac && p < P.nptn) P.g_frac[p] = dfr;  // G_matrix row (2518)
            r0 = wave_sum(p < P.nptn ? t1 : 0.0);
            r1 = wave_sum(p < P.nptn ? t2 - t1 * dfr : 0.0);
            if (L3) r2 = wave_sum(p < P.nptn ? (log(lh) + sc * kLogST) * f : 0.0);
        }
    }
    if constexpr (L3) return finish_tiles3(P, r0, r1, r2, tile, P.ntiles);
    return finish_tiles(P, r0, r1, tile, P.ntiles);
}

// Newton steps and logL from the resident theta buffer: derivatives
// (dotProductTriple, tree/phylokernelnew.h:2474-2550) or logL
// (computeLikelihoodFromBufferSIMD, 3373-3400).
template <int N, int DERIV>
__device__ inline bool theta_eval_f64_body(const KParams &P, const int tile, const double len) {
    const int lane = threadIdx.x & 63;
    const int w = wave_id();
    const int ncw = nwaves(), ncat = P.ncat;
    const int64_t p = (int64_t)tile * 64 + lane;
    const double *__restrict__ theta = (const double *)P.theta;
    __shared__ double s_val[3][PLK_GPU_MAX_CATS * N];
    __shared__ double s_cat[3][PLK_GPU_MAX_CATS][64];
    fill_val(P, len, s_val[0], DERIV ? s_val[1] : nullptr, s_val[2]);
    __syncthreads();
    for (int c = w; c < ncat; c += ncw) {
        const int64_t off = ((int64_t)tile * ncat + c) * N * 64 + lane;
        const double *__restrict__ v0 = s_val[0] + c * N;
        const double *__restrict__ v1 = s_val[1] + c * N;
        const double *__restrict__ v2 = s_val[2] + c * N;
        double l0 = 0.0, l1 = 0.0, l2 = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const double th = theta[off + i * 64];
            l0 = fma(v0[i], th, l0);
            if (DERIV) {
                l1 = fma(v1[i], th, l1);
                l2 = fma(v2[i], th, l2);
            }
        }
        s_cat[0][c][lane] = l0;
        if (DERIV) {
            s_cat[1][c][lane] = l1;
            s_cat[2][c][lane] = l2;
        }
    }
    __syncthreads();
    double r0 = 0.0, r1 = 0.0;
    if (w == 0) {
        double lh = 0.0, d1 = 0.0, d2 = 0.0;
        for (int c = 0; c < ncat; c++) {
            lh += s_cat[0][c][lane];
            if (DERIV) {
                d1 += s_cat[1][c][lane];
                d2 += s_cat[2][c][lane];
            }
        }
        lh = fabs(lh) + P.invar[p];
        const double f = (p < P.nptn) ? P.freq[p] : 0.0;
        if (DERIV) {
            const double inv = 1.0 / lh;
            const double dfr = d1 * inv, ddfr = d2 * inv;
            const double t1 = dfr * f, t2 = ddfr * f;
            if (P.g_frac && p < P.nptn) P.g_frac[p] = dfr;  // G_matrix row (2518)
            r0 = wave_sum(p < P.nptn ? t1 : 0.0);
            r1 = wave_sum(p < P.nptn ? t2 - t1 * dfr : 0.0);
        } else {
            const double l = log(fabs(lh)) + P.buf_scale[p];
            if (p < P.nptn) P.ptn_lh[p] = l;
            r0 = wave_sum(p < P.nptn ? l * f : 0.0);
        }
    }
    return finish_tiles(P, r0, r1, tile, P.ntiles);
}

// One step of minimizeNewton (utils/optimization.cpp:422-499) after a derivative
// evaluation at s.rts returned (d1, d2) = (dlogL/dt, d2logL/dt2): the function is
// computeFuncDerv's f = -d1, df = -d2 (tree/phylotree.cpp:2555-2562).  Evaluation 0 sets
// up the bracket; every later one updates it, then the loop body of iteration j+1 picks
// the next iterate (bisection or Newton step) or finishes.  Sets done when the solve
// ends; out is the answer when status is NEWTON_OK.  Non-finite derivatives follow the
// engine's derivative call: NORM reports underflow, SAFE continues with zeros.
__device__ inline void newton_after_eval(NewtonState &s, double d1, double d2) {
    s.evals++;
    if (!isfinite(d1)) {
        if (!s.safe) {
            s.status = NEWTON_UNDERFLOW;
            s.done = 1;
            return;
        }
        d1 = d2 = 0.0;
    }
    const double f = -d1, df = -d2;
    s.f = f;
    s.df = df;
    if (!isfinite(f) || !isfinite(df)) {
        s.status = NEWTON_NONFINITE;
        s.done = 1;
        return;
    }
    if (s.j == 0) {
        if (df >= 0.0 && fabs(f) < s.xacc) {
            s.out = s.rts;
            s.done = 1;
            return;
        }
        if (f < 0.0) {
            s.xl = s.rts;
            s.xh = s.x2;
        } else {
            s.xh = s.rts;
            s.xl = s.x1;
        }
        s.dx = fabs(s.xh - s.xl);
    } else {
        if (df > 0.0 && fabs(f) < s.xacc) {
            s.out = s.rts;
            s.done = 1;
            return;
        }
        if (f < 0.0) s.xl = s.rts;
        else if (f > 0.0) s.xh = s.rts;
    }
    s.j++;
    if (s.j > s.max_nr) {
        s.status = NEWTON_MAXITER;
        s.done = 1;
        return;
    }
    const double rts = s.rts, xl = s.xl, xh = s.xh;
    s.rts_old = rts;
    if ((df <= 0.0) || (((rts - xh) * df - f) * ((rts - xl) * df - f) >= 0.0)) {
        s.dx = 0.5 * (xh - xl);
        s.rts = xl + s.dx;
        if (xl == s.rts) {
            s.out = s.rts;
            s.done = 1;
            return;
        }
    } else {
        s.dx = f / df;
        s.rts = rts - s.dx;
        if (rts == s.rts) {
            s.out = s.rts;
            s.done = 1;
            return;
        }
    }
    if (fabs(s.dx) < s.xacc || s.j == s.max_nr) {
        s.out = s.rts_old;
        s.done = 1;
    }
}

// Many-branch form: record the logL at the evaluated length, then take the step (or,
// for an eval_only request, just record and stop).
__device__ inline void newton_record(NewtonState &s, double d1, double d2, double logl) {
    s.logl_eval = logl;
    s.x_eval = s.rts;
    if (s.eval_only) {
        s.eval_only = 0;
        s.done = 1;
        return;
    }
    if (s.evals == 0) s.logl0 = logl;
    newton_after_eval(s, d1, d2);
}

}  // namespace dev
}  // namespace plk
