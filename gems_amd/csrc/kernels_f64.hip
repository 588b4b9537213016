// FP64 kernels for gfx950: thin __global__ wrappers over the bodies in dev_f64.h.
#include "dev_f64.h"

namespace plk {
namespace {
using namespace dev;

__global__ void __launch_bounds__(256) k_pmatrix_f64(KParams P, const PJob *__restrict__ jobs) {
    pmatrix_f64_body(P, jobs[blockIdx.x]);
}

template <int N, int MODE>
__global__ void __launch_bounds__(256) k_traverse_f64(KParams P) {
    traverse_f64_body<N, MODE>(P, blockIdx.x);
}

template <int N, int DERIV>
__global__ void __launch_bounds__(256) k_theta_eval_f64(KParams P) {
    theta_eval_f64_body<N, DERIV>(P, blockIdx.x);
}

template <int N>
hipError_t traverse_n(const KParams &p, int mode, hipStream_t s) {
    dim3 grid(p.ntiles), block(64 * p.ncw);
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

}  // namespace

hipError_t f64_pmatrix(const KParams &p, const PJob *jobs, int njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_pmatrix_f64, dim3(njobs), dim3(256), 0, s, p, jobs);
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
