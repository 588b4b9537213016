// Level-parallel 4-state traversal kernels: thin __global__ wrappers over dev_small.h.
#include "dev_f32.h"
#include "dev_small.h"

#include <algorithm>

namespace plk {
namespace {
using namespace dev;

template <int FORM, int CPW, int MODE>
__global__ void __launch_bounds__(1024) k_traverse_small(KParams P) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    traverse_small_body<FORM, CPW, MODE>(P, blockIdx.x, smem, nullptr);
}

// the same with the op list in the kernel arguments
template <int FORM, int CPW, int MODE>
__global__ void __launch_bounds__(1024) k_traverse_small_inl(KParams P, SmallInline in) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    traverse_small_body<FORM, CPW, MODE>(P, blockIdx.x, smem, in.data);
}

// slots = ops of a level run at once per tile, each by min(ncat, 4) wavefronts
template <int FORM, int CPW>
hipError_t launch_cpw(const KParams &p, int mode, int slots, size_t lds, const SmallInline *inl, hipStream_t s) {
    dim3 grid(p.ntiles), block(64 * std::min(p.ncat, 4) * slots);
    if (inl) {
        if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_small_inl<FORM, CPW, MODE_PARTIALS>), grid, block, lds, s, p, *inl);
        else hipLaunchKernelGGL((k_traverse_small_inl<FORM, CPW, MODE_LOGL>), grid, block, lds, s, p, *inl);
    } else {
        if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_small<FORM, CPW, MODE_PARTIALS>), grid, block, lds, s, p);
        else hipLaunchKernelGGL((k_traverse_small<FORM, CPW, MODE_LOGL>), grid, block, lds, s, p);
    }
    return hipGetLastError();
}

}  // namespace

size_t small_lds_bytes(int form, int ncat, int safe, int su, int nbuf, int ntip) {
    const size_t rs = form == 0 ? 8 : 4;
    const int scw = (form == 0 && safe) ? ncat : 1;
    size_t off = (size_t)nbuf * ncat * 4 * 64 * rs;
    off += (size_t)3 * ncat * 4 * 8;
    off += (size_t)(su + 1) * 4 * 8;
    off = (off + 15) & ~(size_t)15;
    off += (size_t)nbuf * scw * 64 * 4;
    off += (size_t)ntip * 64;
    return off;
}

size_t small_stage_bytes(int ncat, int nops, int nlevels) {
    // ops (48 B each), level bounds, then exp(lambda_i r_c t) per (op, side, category, i)
    size_t b = (size_t)nops * sizeof(DevOp) + (size_t)(nlevels + 1) * 4;
    b = (b + 15) & ~(size_t)15;
    return b + (size_t)nops * 2 * ncat * 4 * 8;
}

hipError_t small_traverse(const KParams &p, int form, int mode, int slots, size_t lds, const SmallInline *inl,
                          hipStream_t s) {
    if (mode != MODE_PARTIALS && mode != MODE_LOGL) return hipErrorInvalidValue;
    if (slots < 1 || slots > 4 || p.ncat > PLK_GPU_MAX_CATS) return hipErrorInvalidValue;
    if (inl && !p.small_stage) return hipErrorInvalidValue;  // inline lists are read once, into LDS
    if (form == 0)
        return p.ncat <= 4 ? launch_cpw<0, 1>(p, mode, slots, lds, inl, s) : launch_cpw<0, 2>(p, mode, slots, lds, inl, s);
    return p.ncat <= 4 ? launch_cpw<1, 1>(p, mode, slots, lds, inl, s) : launch_cpw<1, 2>(p, mode, slots, lds, inl, s);
}

}  // namespace plk
