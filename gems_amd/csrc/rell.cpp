// RELL bootstrap state on the device: replicate pattern counts resident in HBM, scored
// against the per-pattern logL that a context's last logL evaluation left on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <vector>

#include "engine_common.h"
#include "plk_rell.h"

namespace plk {

hipError_t rell_launch(const double *ptn_lh, float *lhf, int nptn, int npad, const float *counts,
                       int64_t count_stride, int nsamples, double *rell, double *boot_logl, int *boot_counts, double *boot_orig, const double *uniform,
                       int *improved, double cur_logl, double eps, bool update, hipStream_t s);

#define RHIP_TRY(expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(e_ == hipErrorOutOfMemory ? PLK_ERR_NOMEM : PLK_ERR_HIP,                \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                     \
    } while (0)

struct Rell {
    int nsamples = 0, nptn = 0, npad = 0;
    float *d_counts = nullptr, *d_lhf = nullptr;
    double *d_rell = nullptr, *d_logl = nullptr, *d_orig = nullptr, *d_uniform = nullptr;
    int *d_best_counts = nullptr, *d_improved = nullptr;
};

}  // namespace plk

struct plk_rell : public plk::Rell {};

using namespace plk;

extern "C" int plk_rell_destroy(plk_rell *r) {
    if (!r) return PLK_OK;
    void *bufs[] = {r->d_counts, r->d_lhf, r->d_rell, r->d_logl, r->d_orig, r->d_uniform, r->d_best_counts,
                    r->d_improved};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete r;
    return PLK_OK;
}

extern "C" int plk_rell_create(plk_rell **out, int nsamples, int nptn, const float *counts) {
    clear_error();
    if (!out || nsamples < 1 || nptn < 1 || !counts) return fail(PLK_ERR_ARG, "plk_rell_create: bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(PLK_ERR_HIP, "plk_rell_create: no HIP device visible (the engine has no CPU path)");
    auto *r = new plk_rell();
    r->nsamples = nsamples;
    r->nptn = nptn;
    r->npad = (nptn + 63) / 64 * 64;
    const size_t rows = (size_t)nsamples * r->npad;
    auto cleanup = [&](int rc) {
        plk_rell_destroy(r);
        return rc;
    };
#define CR_TRY(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return cleanup(fail(e_ == hipErrorOutOfMemory ? PLK_ERR_NOMEM : PLK_ERR_HIP,        \
                                std::string("plk_rell_create: ") + #expr + ": " + hipGetErrorString(e_))); \
    } while (0)
    CR_TRY(hipMalloc((void **)&r->d_counts, sizeof(float) * rows));
    CR_TRY(hipMalloc((void **)&r->d_lhf, sizeof(float) * r->npad));
    CR_TRY(hipMalloc((void **)&r->d_rell, sizeof(double) * nsamples));
    CR_TRY(hipMalloc((void **)&r->d_logl, sizeof(double) * nsamples));
    CR_TRY(hipMalloc((void **)&r->d_orig, sizeof(double) * nsamples));
    CR_TRY(hipMalloc((void **)&r->d_uniform, sizeof(double) * nsamples));
    CR_TRY(hipMalloc((void **)&r->d_best_counts, sizeof(int) * nsamples));
    CR_TRY(hipMalloc((void **)&r->d_improved, sizeof(int) * nsamples));
    {
        // padded rows: counts of the padding patterns are 0
        std::vector<float> h(rows, 0.0f);
        for (int s = 0; s < nsamples; s++)
            std::copy(counts + (size_t)s * nptn, counts + (size_t)(s + 1) * nptn, h.begin() + (size_t)s * r->npad);
        CR_TRY(hipMemcpy(r->d_counts, h.data(), sizeof(float) * rows, hipMemcpyHostToDevice));
    }
    {
        // tree/iqtree.cpp:386-390
        std::vector<double> lo(nsamples, -DBL_MAX);
        std::vector<int> zero(nsamples, 0);
        CR_TRY(hipMemcpy(r->d_logl, lo.data(), sizeof(double) * nsamples, hipMemcpyHostToDevice));
        CR_TRY(hipMemcpy(r->d_orig, lo.data(), sizeof(double) * nsamples, hipMemcpyHostToDevice));
        CR_TRY(hipMemcpy(r->d_best_counts, zero.data(), sizeof(int) * nsamples, hipMemcpyHostToDevice));
    }
#undef CR_TRY
    *out = r;
    return PLK_OK;
}

static int rell_run(plk_rell *r, plk_ctx *c, bool update, double cur_logl, double eps, const double *uniform) {
    if (!r || !c) return fail(PLK_ERR_ARG, "plk_rell: null argument");
    if (c->nptn != r->nptn) return fail(PLK_ERR_ARG, "plk_rell: the context's pattern count differs from the replicates'");
    if (!c->ptn_lh_valid) return fail(PLK_ERR_STATE, "plk_rell: the context has no per-pattern logL (evaluate a branch first)");
    if (update) {
        if (!uniform || !std::isfinite(eps) || eps < 0) return fail(PLK_ERR_ARG, "plk_rell_save_tree: bad arguments");
        RHIP_TRY(hipMemcpyAsync(r->d_uniform, uniform, sizeof(double) * r->nsamples, hipMemcpyHostToDevice, c->stream));
    }
    RHIP_TRY(rell_launch(c->d_ptn_lh, r->d_lhf, r->nptn, r->npad, r->d_counts, r->npad, r->nsamples, r->d_rell, r->d_logl,
                         r->d_best_counts, r->d_orig, r->d_uniform, r->d_improved, cur_logl, eps, update, c->stream));
    return PLK_OK;
}

extern "C" int plk_rell_eval(plk_rell *r, plk_ctx *c, double *rell) {
    clear_error();
    int rc = rell_run(r, c, false, 0.0, 0.0, nullptr);
    if (rc) return rc;
    if (rell) RHIP_TRY(hipMemcpyAsync(rell, r->d_rell, sizeof(double) * r->nsamples, hipMemcpyDeviceToHost, c->stream));
    RHIP_TRY(hipStreamSynchronize(c->stream));
    return PLK_OK;
}

extern "C" int plk_rell_save_tree(plk_rell *r, plk_ctx *c, double cur_logl, double eps, const double *uniform,
                                  int32_t *improved, int *nimproved) {
    clear_error();
    int rc = rell_run(r, c, true, cur_logl, eps, uniform);
    if (rc) return rc;
    std::vector<int32_t> imp;
    int32_t *dst = improved;
    if (!dst && nimproved) {
        imp.resize(r->nsamples);
        dst = imp.data();
    }
    if (dst) RHIP_TRY(hipMemcpyAsync(dst, r->d_improved, sizeof(int32_t) * r->nsamples, hipMemcpyDeviceToHost, c->stream));
    RHIP_TRY(hipStreamSynchronize(c->stream));
    if (nimproved) *nimproved = (int)std::count(dst, dst + r->nsamples, 1);
    return PLK_OK;
}

extern "C" int plk_rell_get(plk_rell *r, double *boot_logl, int32_t *boot_counts, double *boot_orig_logl) {
    clear_error();
    if (!r) return fail(PLK_ERR_ARG, "plk_rell_get: null");
    if (boot_logl) RHIP_TRY(hipMemcpy(boot_logl, r->d_logl, sizeof(double) * r->nsamples, hipMemcpyDeviceToHost));
    if (boot_counts) RHIP_TRY(hipMemcpy(boot_counts, r->d_best_counts, sizeof(int32_t) * r->nsamples, hipMemcpyDeviceToHost));
    if (boot_orig_logl) RHIP_TRY(hipMemcpy(boot_orig_logl, r->d_orig, sizeof(double) * r->nsamples, hipMemcpyDeviceToHost));
    return PLK_OK;
}
