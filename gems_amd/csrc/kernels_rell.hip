// RELL scoring of bootstrap replicates (UFBoot, tree/iqtree.cpp:3909-3931) on gfx950.
#include <mutex>
#include <string>

#include "dev_common.h"
#include "plk_rell.h"

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

hipError_t rell_launch(const double *ptn_lh, float *lhf, int nptn, int npad, const float *counts,
                       int64_t count_stride, int nsamples, double *rell, double *boot_logl, int *boot_counts, double *boot_orig, const double *uniform,
                       int *improved, double cur_logl, double eps, bool update, hipStream_t s) {
    hipLaunchKernelGGL(k_rell_prep, dim3((npad + 255) / 256), dim3(256), 0, s, ptn_lh, lhf, nptn, npad);
    if (update)
        hipLaunchKernelGGL(k_rell<true>, dim3(nsamples), dim3(256), 0, s, lhf, counts, count_stride, npad / 4, rell,
                           boot_logl, boot_counts, boot_orig, uniform, improved, cur_logl, eps);
    else
        hipLaunchKernelGGL(k_rell<false>, dim3(nsamples), dim3(256), 0, s, lhf, counts, count_stride, npad / 4, rell,
                           boot_logl, boot_counts, boot_orig, uniform, improved, cur_logl, eps);
    return hipGetLastError();
}

}  // namespace plk

// ---------------------------------------------------------------------------------------
// RELL scoring that allocates nothing: the float staging of the per-pattern logL lives in
// a file-scope device array instead of a buffer of a plk_rell handle.  Each device holds
// its own copy; its address is looked up once per device (hipGetSymbolAddress) and only
// that address is handed to kernels.
namespace plk {
namespace {
constexpr int kRellScratchPatterns = 1 << 20;  // 4 MiB of floats per device
__device__ __align__(256) float scratch[kRellScratchPatterns];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

int scratch_for_device(int dev, float **out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(PLK_ERR_ARG, "plk_rell_eval_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(PLK_ERR_HIP, std::string("plk_rell_eval_noalloc: hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        g_scratch_ptr[dev] = (float *)p;
    }
    *out = g_scratch_ptr[dev];
    return PLK_OK;
}
}  // namespace
}  // namespace plk

using namespace plk;

extern "C" int plk_rell_eval_noalloc(plk_ctx *c, int nsamples, int nptn, const float *d_counts, int64_t count_stride,
                                     double *d_rell) {
    clear_error();
    if (!c || nsamples < 1 || nptn < 1 || !d_counts || !d_rell)
        return fail(PLK_ERR_ARG, "plk_rell_eval_noalloc: bad arguments");
    const int npad = (nptn + 63) / 64 * 64;
    if (npad > kRellScratchPatterns)
        return fail(PLK_ERR_ARG, "plk_rell_eval_noalloc: " + std::to_string(nptn) + " patterns exceed the scratch "
                                     "maximum of " + std::to_string(kRellScratchPatterns));
    if (count_stride < npad || count_stride % 4 != 0 || ((uintptr_t)d_counts & 15) != 0)
        return fail(PLK_ERR_ARG, "plk_rell_eval_noalloc: count rows must be 16-byte aligned with a stride that is a "
                                 "multiple of 4 and at least nptn rounded up to 64");
    if (c->nptn != nptn) return fail(PLK_ERR_ARG, "plk_rell_eval_noalloc: the context's pattern count differs");
    if (!c->ptn_lh_valid)
        return fail(PLK_ERR_STATE, "plk_rell_eval_noalloc: the context has no per-pattern logL (evaluate a branch first)");
    float *lhf = nullptr;
    int rc = scratch_for_device(c->dev, &lhf);
    if (rc) return rc;
    hipError_t e = rell_launch(c->d_ptn_lh, lhf, nptn, npad, d_counts, count_stride, nsamples, d_rell, nullptr, nullptr, nullptr,
                               nullptr, nullptr, 0.0, 0.0, false, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(PLK_ERR_HIP, std::string("plk_rell_eval_noalloc: ") + hipGetErrorString(e));
    return PLK_OK;
}
