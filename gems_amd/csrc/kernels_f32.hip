// FP32-form kernels for gfx950: thin __global__ wrappers over the bodies in dev_f32.h.
#include "dev_f32.h"

namespace plk {
namespace {
using namespace dev;

// one workgroup per (job, category), and for the matrix-core forms per 16-row fragment
// block of the job's matrices as well
template <int N>
__global__ void __launch_bounds__(256) k_pmatrix_f32(KParams P, const PJob *__restrict__ jobs) {
    constexpr int XB = N == 4 ? 1 : f32_xblocks(N);
    const int xb = blockIdx.x % XB, c = (blockIdx.x / XB) % P.ncat;
    pmatrix_f32_body<N>(P, jobs[blockIdx.x / (XB * P.ncat)], c, c + 1, N == 4 ? -1 : xb);
}

// WPC wavefronts per category: 256 threads for up to 4 categories with WPC = 1; WIDE = 1
// allows up to 16 waves (5-8 categories or WPC > 1, 20 states), WIDE = 2 up to 8 (61
// states with 5-8 categories: two waves per SIMD, so half the register file each and
// spills; the 61-state A fragments alone fill 128 registers).  The 20-state forms keep
// four waves per SIMD resident (128 registers): at three, a 49k-pattern alignment's 770
// tiles would overflow the 768 workgroup slots of one round.
template <int N, int WIDE>
constexpr int mfma_min_waves() { return (N <= 32 && !WIDE) ? 4 : 1; }
template <int WIDE>
constexpr int mfma_threads() { return WIDE == 1 ? 64 * 16 : WIDE == 2 ? 64 * 8 : 256; }

template <int N, int MODE, int WIDE, int WPC>
__global__ void __launch_bounds__(mfma_threads<WIDE>(), (mfma_min_waves<N, WIDE>())) k_traverse_f32_mfma(KParams P) {
    traverse_f32_mfma_body<N, MODE, WPC>(P, blockIdx.x);
}

// Whole tiles for blockIdx < P.split_full, then single 16-pattern blocks of the remaining
// tiles (four units per tile), all with one wave per category.
template <int N, int MODE>
__global__ void __launch_bounds__(256, 4) k_traverse_f32_mfma_split(KParams P) {
    const int b = blockIdx.x, full = P.split_full;
    if (b < full) {
        traverse_f32_mfma_body<N, MODE, 1>(P, b, -1, b, gridDim.x);
    } else {
        const int u = b - full;
        traverse_f32_mfma_body<N, MODE, 4>(P, full + u / 4, u % 4, b, gridDim.x);
    }
}

template <int N, int MODE>
__global__ void __launch_bounds__(256) k_traverse_f32_pbw(KParams P) {
    traverse_f32_pbw_body<N, MODE>(P, blockIdx.x);
}

template <int N, int MODE>
__global__ void __launch_bounds__(256) k_traverse_f32_valu(KParams P) {
    traverse_f32_valu_body<N, MODE>(P, blockIdx.x);
}

template <int N>
__global__ void __launch_bounds__(256) k_theta_f32(KParams P) {
    theta_f32_body<N>(P, blockIdx.x);
}

template <int N>
hipError_t traverse_f32_n(const KParams &p, int mode, hipStream_t s) {
    dim3 grid(p.ntiles);
    if constexpr (N == 4) {
        dim3 block(64 * p.ncw);
        if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_valu<N, MODE_PARTIALS>), grid, block, 0, s, p);
        else hipLaunchKernelGGL((k_traverse_f32_valu<N, MODE_LOGL>), grid, block, 0, s, p);
    } else {
        // tile shape chosen by the caller's tuning: 0 = one wave per 16-pattern block
        // over all categories (ncat <= 4); 1, 2, 4 = waves per category
        if (p.wpc == 0 && p.ncat <= 4) {
            dim3 block(256);
            if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_pbw<N, MODE_PARTIALS>), grid, block, 0, s, p);
            else hipLaunchKernelGGL((k_traverse_f32_pbw<N, MODE_LOGL>), grid, block, 0, s, p);
            return hipGetLastError();
        }
        int wpc = p.wpc;
        if (N > 32) {
            // 61 states: one wave per category (a second one would not fit beside it)
            dim3 block(64 * p.ncat);
            if (p.ncat <= 4) {
                if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_PARTIALS, 0, 1>), grid, block, 0, s, p);
                else hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_LOGL, 0, 1>), grid, block, 0, s, p);
            } else {
                if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_PARTIALS, 2, 1>), grid, block, 0, s, p);
                else hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_LOGL, 2, 1>), grid, block, 0, s, p);
            }
            return hipGetLastError();
        }
        if (wpc < 1 || p.ncat * wpc > 16) wpc = 1;
        dim3 block(64 * p.ncat * wpc);
        if constexpr (N <= 32) {
        if (wpc == 1 && p.ncat <= 4 && p.ncu > 0) {
            // tiles beyond a multiple of the compute-unit count (up to a quarter of the
            // compute units) run as single-block units, so that every compute unit holds
            // the same number of whole tiles (four workgroups of this shape fit one)
            const int per = p.ntiles / p.ncu, extra = p.ntiles - per * p.ncu;
            if (per >= 1 && per <= 3 && extra > 0 && 4 * extra <= p.ncu) {
                KParams q = p;
                q.split_full = p.ntiles - extra;
                dim3 g2(q.split_full + 4 * extra);
                if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_mfma_split<N, MODE_PARTIALS>), g2, block, 0, s, q);
                else hipLaunchKernelGGL((k_traverse_f32_mfma_split<N, MODE_LOGL>), g2, block, 0, s, q);
                return hipGetLastError();
            }
        }
        }
        if (wpc == 1 && p.ncat <= 4) {
            if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_PARTIALS, 0, 1>), grid, block, 0, s, p);
            else hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_LOGL, 0, 1>), grid, block, 0, s, p);
        } else if constexpr (N <= 32) {
            if (wpc == 1) {
                if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_PARTIALS, 1, 1>), grid, block, 0, s, p);
                else hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_LOGL, 1, 1>), grid, block, 0, s, p);
            } else if (wpc == 2) {
                if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_PARTIALS, 1, 2>), grid, block, 0, s, p);
                else hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_LOGL, 1, 2>), grid, block, 0, s, p);
            } else {
                if (mode == MODE_PARTIALS) hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_PARTIALS, 1, 4>), grid, block, 0, s, p);
                else hipLaunchKernelGGL((k_traverse_f32_mfma<N, MODE_LOGL, 1, 4>), grid, block, 0, s, p);
            }
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

}  // namespace

hipError_t f32_pmatrix(const KParams &p, const PJob *jobs, int njobs, hipStream_t s) {
    switch (p.n) {
        case 4: hipLaunchKernelGGL(k_pmatrix_f32<4>, dim3(njobs * p.ncat), dim3(256), 0, s, p, jobs); break;
        case 20:
            hipLaunchKernelGGL(k_pmatrix_f32<20>, dim3(njobs * p.ncat * f32_xblocks(20)), dim3(256), 0, s, p, jobs);
            break;
        case 61:
            hipLaunchKernelGGL(k_pmatrix_f32<61>, dim3(njobs * p.ncat * f32_xblocks(61)), dim3(256), 0, s, p, jobs);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t f32_traverse(const KParams &p, int mode, hipStream_t s) {
    switch (p.n) {
        case 4: return traverse_f32_n<4>(p, mode, s);
        case 20: return traverse_f32_n<20>(p, mode, s);
        case 61: return traverse_f32_n<61>(p, mode, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t f32_theta(const KParams &p, hipStream_t s) {
    dim3 grid(p.ntiles), block(64 * p.ncw);
    switch (p.n) {
        case 4: hipLaunchKernelGGL(k_theta_f32<4>, grid, block, 0, s, p); break;
        case 20: hipLaunchKernelGGL(k_theta_f32<20>, grid, block, 0, s, p); break;
        case 61: hipLaunchKernelGGL(k_theta_f32<61>, grid, block, 0, s, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace plk
