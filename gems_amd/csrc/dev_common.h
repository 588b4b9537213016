// Device helpers shared by the FP64, FP32 and batched likelihood kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine_common.h"

namespace plk {
namespace dev {

constexpr double kST = 8.636168555094444625386e-78;  // 2^-256 (tree/phylotree.h:74-80)
constexpr double kTwo256 = 0x1p256;
constexpr double kLogST = -177.4456782233459932741;
constexpr double kLn2 = 0.69314718055994530941723212145818;

typedef float floatx4 __attribute__((ext_vector_type(4)));

// log(|s| 2^e + invar): a pattern likelihood held as a mantissa s and a power-of-two exponent
// e (the FP32 form renormalises every pattern) plus the invariant-site mass, which is a true
// (unscaled) value.  The reference adds ptn_invar to an unscaled likelihood
// (tree/phylokernelnew.h:3097) because it never rescales a pattern with ptn_invar != 0
// (:1999, :2020); here the sum is formed in log space, so no 2^e or 2^-e is ever formed.
__device__ inline double log_lh_expo(double s, int e, double invar) {
    const double ls = log(fabs(s)) + e * kLn2;
    if (invar == 0.0) return ls;
    const double li = log(invar);
    const double hi = fmax(ls, li), lo = fmin(ls, li);
    return hi + log1p(exp(lo - hi));
}

// Read-only views in the constant address space: wave-uniform reads through them
// compile to scalar loads whose results feed the FMAs as SGPR operands (the data is
// written by an earlier launch only).
typedef const __attribute__((address_space(4))) double *cdptr;
typedef const __attribute__((address_space(4))) float *cfptr;
__device__ __forceinline__ cdptr as_const(const double *p) { return (cdptr)p; }
__device__ __forceinline__ cfptr as_constf(const float *p) { return (cfptr)p; }

// The same pointer, hidden from loop-invariant code motion: a matrix read through it
// inside a loop is re-read (scalar-cache hits) each iteration instead of being hoisted
// out and parked in VGPR lanes, which would cost a v_readlane per use.
template <typename T>
__device__ __forceinline__ T *opaque(T *p) {
    asm volatile("" : "+s"(p));
    return p;
}

// Op descriptors and level bounds are uploaded before a launch and never written during
// it.  Read through the constant address space they arrive by scalar loads; read as plain
// global memory (which the kernel also stores to) the compiler uses a vector load, whose
// wait (vmcnt) also drains every store the wave has in flight, so each op would first
// wait for the previous op's stores to complete.
__device__ __forceinline__ int load_int(const int32_t *v, int k) {
    return ((const __attribute__((address_space(4))) int32_t *)v)[k];
}
__device__ __forceinline__ DevOp load_op(const DevOp *ops, int k) {
    static_assert(sizeof(DevOp) == 48, "8 int32 fields and 2 doubles");
    const int32_t *w = (const int32_t *)(ops + k);
    const cdptr d = as_const((const double *)(ops + k));
    DevOp o;
    o.out_slot = load_int(w, 0);
    o.kind = load_int(w, 1);
    o.left = load_int(w, 2);
    o.right = load_int(w, 3);
    o.lds_out = load_int(w, 4);
    o.lds_left = load_int(w, 5);
    o.lds_right = load_int(w, 6);
    o.level = load_int(w, 7);
    o.len_left = d[4];
    o.len_right = d[5];
    return o;
}

// Op index of step k for a tile: tiles alternate between the op list's own order and the
// second order in P.perm (when given), by tile parity flipped every 256 tiles, so that both
// orders run on every XCD and on most compute units.
__device__ __forceinline__ int op_index(const KParams &P, int tile, int k) {
    if (P.perm == nullptr || (((tile >> 8) ^ tile) & 1) == 0) return k;
    return load_int(P.perm, k);
}

// Tip states of one tile staged in LDS when a traversal starts (the first kStageTaxa taxa;
// later ones are read from global memory): a tip child's state then costs an LDS read
// instead of a memory round trip that its leaf-table gather would wait behind.  The
// barriers on both sides let persistent (batched) launches restage for every tile.
constexpr int kStageTaxa = 128;
__device__ __forceinline__ void stage_tips(const KParams &P, int tile, uint32_t *s) {
    const int nst = P.ntaxa < kStageTaxa ? P.ntaxa : kStageTaxa;
    __syncthreads();
    for (int i = threadIdx.x; i < nst * 16; i += blockDim.x) {
        const int t = i >> 4, q = i & 15;
        s[i] = ((const uint32_t *)(P.states + (int64_t)t * P.nptn_pad + (int64_t)tile * 64))[q];
    }
    __syncthreads();
}
__device__ __forceinline__ int tip_state(const KParams &P, const uint32_t *s, int taxon, int tile, int col) {
    return taxon < kStageTaxa ? ((const uint8_t *)s)[taxon * 64 + col]
                              : P.states[(int64_t)taxon * P.nptn_pad + (int64_t)tile * 64 + col];
}

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
__device__ __forceinline__ int nwaves() { return blockDim.x >> 6; }

// Deterministic wavefront sum (fixed shuffle tree); result valid in lane 0.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Compile-time loop: f(I) for I = B..E-1, fully unrolled.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Renormalisation exponent of a pattern maximum: m * 2^-e in [0.5, 1).
__device__ __forceinline__ int norm_exponent(float m) {
    int e = 0;
    if (m > 0.0f) (void)frexpf(m, &e);
    return e;
}

// Three-value form of finish_tiles (tile_sums stride 3, result[0..2]).
__device__ inline bool finish_tiles3(const KParams &P, double v0, double v1, double v2, int tile, int ntiles) {
    __shared__ int s_last3;
    if (threadIdx.x == 0) {
        P.tile_sums[3 * tile] = v0;
        P.tile_sums[3 * tile + 1] = v1;
        P.tile_sums[3 * tile + 2] = v2;
        __threadfence();
        unsigned prev = atomicAdd(P.counter, 1u);
        s_last3 = (prev == (unsigned)ntiles - 1);
    }
    __syncthreads();
    if (s_last3) {
        __threadfence();
        if (threadIdx.x < 64) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (int i = threadIdx.x; i < ntiles; i += 64) {
                a0 += __hip_atomic_load(&P.tile_sums[3 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                a1 += __hip_atomic_load(&P.tile_sums[3 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                a2 += __hip_atomic_load(&P.tile_sums[3 * i + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            a0 = wave_sum(a0);
            a1 = wave_sum(a1);
            a2 = wave_sum(a2);
            if (threadIdx.x == 0) {
                P.result[0] = a0;
                P.result[1] = a1;
                P.result[2] = a2;
                __hip_atomic_store(P.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    __syncthreads();
    return s_last3;
}

// f(integral_constant<kind>) for a run-time op kind (uniform over a launch), so each
// kind compiles to its own branch-free body.
template <typename F>
__device__ __forceinline__ void with_kind(int kind, F &&f) {
    if (kind == KIND_II) f(std::integral_constant<int, KIND_II>{});
    else if (kind == KIND_TI) f(std::integral_constant<int, KIND_TI>{});
    else f(std::integral_constant<int, KIND_TT>{});
}

// x-loop over output states: unrolled for small state counts, a plain loop for the
// 61-state codon model.
template <int N, typename F>
__device__ __forceinline__ void for_states(F &&f) {
    if constexpr (N <= 20) {
        static_for<0, N>([&](auto X) { f(decltype(X)::value); });
    } else {
#pragma unroll 1
        for (int x = 0; x < N; x++) f(x);
    }
}

// Per-tile values -> fixed-order total over the ntiles tiles of one alignment.  The
// last workgroup to finish a tile of this alignment sums the tile values in index
// order, so the result does not depend on scheduling (or on other alignments sharing
// the launch); it then re-arms the counter for the next launch (no memset, graph-safe).
// Returns true in the workgroup that wrote the result.
__device__ inline bool finish_tiles(const KParams &P, double v0, double v1, int tile, int ntiles) {
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        P.tile_sums[2 * tile] = v0;
        P.tile_sums[2 * tile + 1] = v1;
        __threadfence();
        unsigned prev = atomicAdd(P.counter, 1u);
        s_last = (prev == (unsigned)ntiles - 1);
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        if (threadIdx.x < 64) {
            double a0 = 0.0, a1 = 0.0;
            for (int i = threadIdx.x; i < ntiles; i += 64) {
                a0 += __hip_atomic_load(&P.tile_sums[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                a1 += __hip_atomic_load(&P.tile_sums[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            a0 = wave_sum(a0);
            a1 = wave_sum(a1);
            if (threadIdx.x == 0) {
                P.result[0] = a0;
                P.result[1] = a1;
                __hip_atomic_store(P.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    __syncthreads();
    return s_last;
}

}  // namespace dev
}  // namespace plk
