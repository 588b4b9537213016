// FP64 kernels for gfx950: thin __global__ wrappers over the bodies in dev_f64.h.
#include "dev_f64.h"
#include "dev_f64_mfma.h"

namespace plk {
namespace {
using namespace dev;

// one workgroup per (job, category)
__global__ void __launch_bounds__(256) k_pmatrix_f64(KParams P, const PJob *__restrict__ jobs) {
    pmatrix_f64_body(P, jobs[blockIdx.x / P.ncat], blockIdx.x % P.ncat);
}

// Occupancy floor: the 20-state partial/logL kernels are held to 128 VGPRs (4 waves per
// SIMD, a few bytes of spill), so that e.g. the 770 tiles x 4 waves of a 49k-pattern
// alignment fit the chip in one round (3 waves per SIMD hold only 768 tiles).  The
// theta variant keeps its full register budget (it would spill far more).
template <int N, int MODE>
constexpr int traverse_min_waves() { return (N == 20 && MODE != MODE_THETA) ? 4 : 1; }

template <int N, int MODE>
__global__ void __launch_bounds__(256, (traverse_min_waves<N, MODE>())) k_traverse_f64(KParams P) {
    traverse_f64_body<N, MODE>(P, blockIdx.x);
}

// 61 states with at most four categories: matrices staged in LDS (one workgroup per CU)
template <int MODE>
__global__ void __launch_bounds__(256, 1) k_traverse_f64_lds61(KParams P) {
    traverse_f64_body<61, MODE, false, true>(P, blockIdx.x);
}

// matrix-core traversal (20, 61 states, ncat <= 4): one workgroup per 16-pattern block,
// one wavefront per 16-row block of the states
// CSPLIT (SAFE scaling only): one workgroup per (block, category), four times the
// wavefronts for the same work
// Register budget measured on the 48-taxon x 3,962-pattern codon traversal: the
// per-category launch runs best at four wavefronts per SIMD (128 VGPRs, a few spilled;
// 0.58 ms against 0.60 at three and 0.62 at two), the all-category (NORM) one at two
// (0.76 ms against 0.79 at three and 1.00 at four).
template <int N, bool CSPLIT>
__global__ void __launch_bounds__(64 * MfmaShape<N>::RB, CSPLIT ? 4 : 2) k_traverse_f64_mfma(KParams P) {
    if constexpr (CSPLIT) {
        const int c = blockIdx.x % P.ncat;
        traverse_f64_mfma_body<N, MODE_PARTIALS, 1>(P, blockIdx.x / P.ncat, c, c + 1);
    } else {
        traverse_f64_mfma_body<N, MODE_PARTIALS, 4>(P, blockIdx.x, 0, P.ncat);
    }
}

template <int N, int DERIV>
__global__ void __launch_bounds__(256) k_theta_eval_f64(KParams P) {
    theta_eval_f64_body<N, DERIV>(P, blockIdx.x, P.br_len);
}

// First Newton step, after the traversal that built theta left (d1, d2) in P.result.
__global__ void k_newton_first(NewtonState *st, const double *result) {
    if (threadIdx.x == 0) newton_after_eval(*st, result[0], result[1]);
}

// One Newton iteration: derivatives at the state's current iterate, then (in the
// workgroup that finishes the reduction) the step.  A finished solve makes every
// later launch of the queue return at once.
template <int N>
__global__ void __launch_bounds__(256) k_theta_newton_f64(KParams P, NewtonState *st) {
    if (st->done) return;
    const double len = st->rts;
    const bool last = theta_eval_f64_body<N, 1>(P, blockIdx.x, len);
    if (last && threadIdx.x == 0) newton_after_eval(*st, P.result[0], P.result[1]);
}

template <int N>
hipError_t traverse_n(const KParams &p0, int mode, hipStream_t s) {
    KParams p = p0;
    dim3 grid(p.ntiles), block(64 * p.ncw);
    if constexpr (N != 4) {
        if (p.f64_form == 1 && p.ncat <= 4 && p.uinv_frag) {
            // the op list on the matrix cores, then the branch stage alone (no ops) in the
            // vector kernel
            if (p.nops > 0) {
                if (p.safe)
                    hipLaunchKernelGGL((k_traverse_f64_mfma<N, true>), dim3(p.ntiles * 4 * p.ncat),
                                       dim3(64 * MfmaShape<N>::RB), 0, s, p);
                else
                    hipLaunchKernelGGL((k_traverse_f64_mfma<N, false>), dim3(p.ntiles * 4), dim3(64 * MfmaShape<N>::RB),
                                       0, s, p);
                if (mode == MODE_PARTIALS) return hipGetLastError();
                p.nops = 0;
            }
            if (mode == MODE_LOGL) hipLaunchKernelGGL((k_traverse_f64<N, MODE_LOGL>), grid, block, 0, s, p);
            else if (mode == MODE_THETA) hipLaunchKernelGGL((k_traverse_f64<N, MODE_THETA>), grid, block, 0, s, p);
            return hipGetLastError();
        }
    }
    if (N == 61 && p.ncat <= 4) {
        if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f64_lds61<MODE_PARTIALS>), grid, block, 0, s, p);
        else if (mode == MODE_LOGL) hipLaunchKernelGGL((k_traverse_f64_lds61<MODE_LOGL>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((k_traverse_f64_lds61<MODE_THETA>), grid, block, 0, s, p);
        return hipGetLastError();
    }
    if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f64<N, MODE_PARTIALS>), grid, block, 0, s, p);
    else if (mode == MODE_LOGL) hipLaunchKernelGGL((k_traverse_f64<N, MODE_LOGL>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((k_traverse_f64<N, MODE_THETA>), grid, block, 0, s, p);
    return hipGetLastError();
}

template <int N>
hipError_t theta_eval_n(const KParams &p, int deriv, hipStream_t s) {
    dim3 grid(p.ntiles), block(64 * p.ncw);
    if (deriv) hipLaunchKernelGGL((k_theta_eval_f64<N, 1>), grid, block, 0, s, p);
    else hipLaunchKernelGGL((k_theta_eval_f64<N, 0>), grid, block, 0, s, p);
    return hipGetLastError();
}

template <int N>
hipError_t theta_newton_n(const KParams &p, NewtonState *st, hipStream_t s) {
    hipLaunchKernelGGL((k_theta_newton_f64<N>), dim3(p.ntiles), dim3(64 * p.ncw), 0, s, p, st);
    return hipGetLastError();
}

}  // namespace

hipError_t newton_first(NewtonState *st, const double *result, hipStream_t s) {
    hipLaunchKernelGGL(k_newton_first, dim3(1), dim3(64), 0, s, st, result);
    return hipGetLastError();
}

hipError_t f64_theta_newton(const KParams &p, NewtonState *st, hipStream_t s) {
    switch (p.n) {
        case 4: return theta_newton_n<4>(p, st, s);
        case 20: return theta_newton_n<20>(p, st, s);
        case 61: return theta_newton_n<61>(p, st, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t f64_pmatrix(const KParams &p, const PJob *jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_pmatrix_f64, dim3(njobs * p.ncat), dim3(256), 0, s, p, jobs);
    return hipGetLastError();
}

hipError_t f64_traverse(const KParams &p, int mode, hipStream_t s) {
    switch (p.n) {
        case 4: return traverse_n<4>(p, mode, s);
        case 20: return traverse_n<20>(p, mode, s);
        case 61: return traverse_n<61>(p, mode, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t f64_theta_eval(const KParams &p, int deriv, hipStream_t s) {
    switch (p.n) {
        case 4: return theta_eval_n<4>(p, deriv, s);
        case 20: return theta_eval_n<20>(p, deriv, s);
        case 61: return theta_eval_n<61>(p, deriv, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace plk
