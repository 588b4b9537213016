// RELL scoring of bootstrap replicates (UFBoot, tree/iqtree.cpp:3909-3931) on gfx950.
#include "dev_common.h"

namespace plk {
namespace {
using namespace dev;

// Per-pattern logL -> float (saveCurrentTree's pattern_lh, tree/iqtree.cpp:3867-3870);
// non-finite values get the replacement of the branch-logL fallback
// (tree/phylokernelnew.h:3167-3184); padding patterns are 0.
__global__ void k_rell_prep(const double *__restrict__ ptn_lh, float *__restrict__ lhf, int nptn, int npad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npad) return;
    float v = 0.0f;
    if (p < nptn) {
        double l = ptn_lh[p];
        if (!isfinite(l)) l = kLogST * 4;
        v = (float)l;
    }
    lhf[p] = v;
}

// One workgroup per replicate: rell = sum_p lh[p] * count[p] (float x float products are
// exact in FP64; FP64 sums in a fixed order), then, with UPDATE, the replicate's
// best-tree rule in the first thread.
template <bool UPDATE>
__global__ void __launch_bounds__(256) k_rell(const float *__restrict__ lhf, const float *__restrict__ counts,
                                              int64_t stride, int n4, double *__restrict__ rell,
                                              double *__restrict__ boot_logl, int *__restrict__ boot_counts,
                                              double *__restrict__ boot_orig, const double *__restrict__ uniform,
                                              int *__restrict__ improved, double cur_logl, double eps) {
    const int s = blockIdx.x;
    const float4 *__restrict__ c4 = (const float4 *)(counts + (int64_t)s * stride);
    const float4 *__restrict__ l4 = (const float4 *)lhf;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n4; i += 256) {
        const float4 c = c4[i], l = l4[i];
        acc += (double)l.x * (double)c.x;
        acc += (double)l.y * (double)c.y;
        acc += (double)l.z * (double)c.z;
        acc += (double)l.w * (double)c.w;
    }
    __shared__ double ws[4];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double r = (ws[0] + ws[1]) + (ws[2] + ws[3]);
    rell[s] = r;
    if (!UPDATE) return;
    const double best = boot_logl[s];
    bool better = r > best + eps;
    if (!better && r > best - eps) better = uniform[s] <= 1.0 / (boot_counts[s] + 1);
    if (better) {
        if (r <= best + eps) boot_counts[s] += 1;
        else boot_counts[s] = 1;
        boot_logl[s] = fmax(best, r);
        boot_orig[s] = cur_logl;
    }
    improved[s] = better ? 1 : 0;
}

}  // namespace

hipError_t rell_launch(const double *ptn_lh, float *lhf, int nptn, int npad, const float *counts, int nsamples,
                       double *rell, double *boot_logl, int *boot_counts, double *boot_orig, const double *uniform,
                       int *improved, double cur_logl, double eps, bool update, hipStream_t s) {
    hipLaunchKernelGGL(k_rell_prep, dim3((npad + 255) / 256), dim3(256), 0, s, ptn_lh, lhf, nptn, npad);
    if (update)
        hipLaunchKernelGGL(k_rell<true>, dim3(nsamples), dim3(256), 0, s, lhf, counts, (int64_t)npad, npad / 4, rell,
                           boot_logl, boot_counts, boot_orig, uniform, improved, cur_logl, eps);
    else
        hipLaunchKernelGGL(k_rell<false>, dim3(nsamples), dim3(256), 0, s, lhf, counts, (int64_t)npad, npad / 4, rell,
                           boot_logl, boot_counts, boot_orig, uniform, improved, cur_logl, eps);
    return hipGetLastError();
}

}  // namespace plk
