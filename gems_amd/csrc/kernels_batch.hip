// Batched unlinked partitions (IQ-TREE's PhyloSuperTree loop over partitions,
// tree/phylosupertree.cpp:662-734) in one launch per traversal: every work item is one
// 64-pattern tile of one partition; each partition keeps its own model, P(t), slots,
// scale counters and fixed-order reduction, so its result is bit-identical to running
// it alone.  DNA and protein partitions share the launch (the state count is a
// per-item branch, uniform within a workgroup).
#include "dev_f32.h"
#include "dev_f64.h"

namespace plk {
namespace {
using namespace dev;

// Copy a wave-uniform struct out of device memory through the constant address space
// (scalar loads).
template <typename T>
__device__ __forceinline__ T load_uniform(const T *p) {
    static_assert(sizeof(T) % 8 == 0, "8-byte granular structs only");
    typedef const __attribute__((address_space(4))) uint64_t *cu64;
    T v;
    uint64_t *dst = reinterpret_cast<uint64_t *>(&v);
    const cu64 src = (cu64)p;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 8); i++) dst[i] = src[i];
    return v;
}

__global__ void __launch_bounds__(256) k_batch_pmatrix_f64(const KParams *params, const PJob *jobs) {
    const PJob j = load_uniform(jobs + blockIdx.x);
    const KParams P = load_uniform(params + j.pad);
    pmatrix_f64_body(P, j);
}

__global__ void __launch_bounds__(256) k_batch_pmatrix_f32(const KParams *params, const PJob *jobs) {
    const PJob j = load_uniform(jobs + blockIdx.x);
    const KParams P = load_uniform(params + j.pad);
    switch (P.n) {
        case 4: pmatrix_f32_body<4>(P, j); break;
        case 20: pmatrix_f32_body<20>(P, j); break;
        default: break;
    }
}

// Persistent when gridDim < nitems: workgroup g takes items g, g + grid, ... (the host
// orders items by decreasing cost, so the static stride is a longest-first schedule).
__global__ void __launch_bounds__(256) k_batch_traverse_f64(const KParams *params, const BatchItem *items, int nitems) {
    for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
        const BatchItem it = load_uniform(items + i);
        const KParams P = load_uniform(params + it.part);
        if (P.n == 4) traverse_f64_body<4, MODE_LOGL>(P, it.tile);
        else traverse_f64_body<20, MODE_LOGL>(P, it.tile);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_batch_traverse_f32(const KParams *params, const BatchItem *items, int nitems) {
    for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
        const BatchItem it = load_uniform(items + i);
        const KParams P = load_uniform(params + it.part);
        if (P.n == 4) traverse_f32_valu_body<4, MODE_LOGL>(P, it.tile);
        else traverse_f32_mfma_body<20, MODE_LOGL>(P, it.tile);
        __syncthreads();
    }
}

// df/ddf of many branches of one alignment in one launch: item = (branch, tile); each
// branch has its own KParams (branch ends, length, reduction buffers) and theta is not
// stored.  Same arithmetic as the single-branch kernels, so results are identical.
__global__ void __launch_bounds__(256) k_multi_derivs_f64(const KParams *params, const BatchItem *items, int nitems) {
    const BatchItem it = load_uniform(items + blockIdx.x);
    const KParams P = load_uniform(params + it.part);
    switch (P.n) {
        case 4: traverse_f64_body<4, MODE_THETA>(P, it.tile); break;
        case 20: traverse_f64_body<20, MODE_THETA>(P, it.tile); break;
        case 61: traverse_f64_body<61, MODE_THETA>(P, it.tile); break;
        default: break;
    }
}

__global__ void __launch_bounds__(256) k_multi_derivs_f32(const KParams *params, const BatchItem *items, int nitems) {
    const BatchItem it = load_uniform(items + blockIdx.x);
    const KParams P = load_uniform(params + it.part);
    switch (P.n) {
        case 4: theta_f32_body<4>(P, it.tile); break;
        case 20: theta_f32_body<20>(P, it.tile); break;
        case 61: theta_f32_body<61>(P, it.tile); break;
        default: break;
    }
}

// Many-branch Newton: item = (branch, tile); one launch = one derivative evaluation of
// every unfinished branch at its device-held iterate (with the logL there), then that
// branch's minimizeNewton step in the workgroup finishing its reduction.  A finished
// branch's items return at once.  (A branch's state is written only by its last
// workgroup, after all of its items have read it.)
__global__ void __launch_bounds__(256) k_multi_newton_f64(const KParams *params, const BatchItem *items,
                                                          NewtonState *states) {
    const BatchItem it = load_uniform(items + blockIdx.x);
    NewtonState *st = states + it.part;
    if (st->done) return;
    KParams P = load_uniform(params + it.part);
    P.br_len = st->rts;
    bool last = false;
    switch (P.n) {
        case 4: last = traverse_f64_body<4, MODE_THETA, true>(P, it.tile); break;
        case 20: last = traverse_f64_body<20, MODE_THETA, true>(P, it.tile); break;
        case 61: last = traverse_f64_body<61, MODE_THETA, true>(P, it.tile); break;
        default: break;
    }
    if (last && threadIdx.x == 0) newton_record(*st, P.result[0], P.result[1], P.result[2]);
}

__global__ void __launch_bounds__(256) k_multi_newton_f32(const KParams *params, const BatchItem *items,
                                                          NewtonState *states) {
    const BatchItem it = load_uniform(items + blockIdx.x);
    NewtonState *st = states + it.part;
    if (st->done) return;
    KParams P = load_uniform(params + it.part);
    P.br_len = st->rts;
    bool last = false;
    switch (P.n) {
        case 4: last = theta_f32_body<4, true>(P, it.tile); break;
        case 20: last = theta_f32_body<20, true>(P, it.tile); break;
        case 61: last = theta_f32_body<61, true>(P, it.tile); break;
        default: break;
    }
    if (last && threadIdx.x == 0) newton_record(*st, P.result[0], P.result[1], P.result[2]);
}

}  // namespace

hipError_t multi_newton(int prec, int ncw, const KParams *params, const BatchItem *items, int nitems,
                        NewtonState *states, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    if (prec == PLK_FP64)
        hipLaunchKernelGGL(k_multi_newton_f64, dim3(nitems), dim3(64 * ncw), 0, s, params, items, states);
    else hipLaunchKernelGGL(k_multi_newton_f32, dim3(nitems), dim3(64 * ncw), 0, s, params, items, states);
    return hipGetLastError();
}

hipError_t multi_derivs(int prec, int ncw, const KParams *params, const BatchItem *items, int nitems, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    if (prec == PLK_FP64) hipLaunchKernelGGL(k_multi_derivs_f64, dim3(nitems), dim3(64 * ncw), 0, s, params, items, nitems);
    else hipLaunchKernelGGL(k_multi_derivs_f32, dim3(nitems), dim3(64 * ncw), 0, s, params, items, nitems);
    return hipGetLastError();
}

hipError_t batch_pmatrix(int prec, const KParams *params, const PJob *jobs, int njobs, hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    if (prec == PLK_FP64) hipLaunchKernelGGL(k_batch_pmatrix_f64, dim3(njobs), dim3(256), 0, s, params, jobs);
    else hipLaunchKernelGGL(k_batch_pmatrix_f32, dim3(njobs), dim3(256), 0, s, params, jobs);
    return hipGetLastError();
}

hipError_t batch_traverse(int prec, const KParams *params, const BatchItem *items, int nitems, int grid,
                          hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    const int g = (grid <= 0 || grid > nitems) ? nitems : grid;
    if (prec == PLK_FP64) hipLaunchKernelGGL(k_batch_traverse_f64, dim3(g), dim3(256), 0, s, params, items, nitems);
    else hipLaunchKernelGGL(k_batch_traverse_f32, dim3(g), dim3(256), 0, s, params, items, nitems);
    return hipGetLastError();
}

}  // namespace plk
