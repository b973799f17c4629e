// Building blocks of the per-k solver kernels, shared by loops.hip and
// rtr.hip: one wavefront per k point with the k point's nb x nb blocks in
// LDS, W wavefronts per workgroup, scalar results through per-workgroup
// partials summed by launch_finalize_sums.
#pragma once

#include <algorithm>

#include "loops.h"

namespace hfgpu {
namespace loopk {

constexpr double TINY_REAL = 1.0e-30;
constexpr double LOGIT_CLIP = 1.0e-6;
constexpr double OCC_CLIP = 1.0e-8;

__device__ __forceinline__ double clampd(double v, double lo, double hi) {
    return v < lo ? lo : (hi < v ? hi : v);
}

// complex slots per wavefront of the line-search / retraction kernels:
// 4 nb^2 + nb complex, then nb^2 + nb doubles
__host__ __device__ __forceinline__ int ls_wave_cd(int nb) {
    return 4 * nb * nb + nb + (nb * nb + nb + 1) / 2;
}

__device__ __forceinline__ cd cayley_c(double lam, double tau) {
    const double a = 0.5 * tau * lam;
    const double den = 1.0 + a * a;
    return cd{(1.0 - a * a) / den, (a + a) / den};
}

// C = op(A) op(B) on nb x nb LDS blocks by one wavefront.  With tb the
// lanes of a row start the sum at n = j (the terms in rotated order), so
// their reads of B's rows B[j][n] fall on distinct LDS banks instead of a
// stride of nb complex (all one bank group at nb = 16).
__device__ __forceinline__ void wave_mm(const cd* A, const cd* B, cd* C, int nb, bool ta,
                                        bool tb, int lane) {
    const int nb2 = nb * nb;
    for (int e = lane; e < nb2; e += 64) {
        const int i = e / nb, j = e % nb;
        cd s{0.0, 0.0};
        int n = tb ? j : 0;
        for (int t = 0; t < nb; ++t) {
            const cd a = ta ? cconj(A[n * nb + i]) : A[i * nb + n];
            const cd b = tb ? cconj(B[j * nb + n]) : B[n * nb + j];
            s = cadd(s, cmul(a, b));
            n = n + 1 == nb ? 0 : n + 1;
        }
        C[e] = s;
    }
}

// Sum NS per-thread values over the block and write partial[block * NS + s].
template <int NS, int W>
__device__ __forceinline__ void store_block_partials(double (&v)[NS], double* red,
                                                     double* partial) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double t = wave_sum(v[s]);
        if (lane == 0) red[wave * NS + s] = t;
    }
    __syncthreads();
    if (threadIdx.x < NS) {
        double t = 0.0;
        for (int w = 0; w < W; ++w) t += red[w * NS + threadIdx.x];
        partial[blockIdx.x * NS + threadIdx.x] = t;
    }
}

// Is solve blockIdx.y of the batch active?  (uniform over the workgroup)
__device__ __forceinline__ bool solve_active(unsigned active, int s) {
    return ((active >> s) & 1u) != 0u;
}

inline unsigned loop_blocks(int nk, int W) {
    return std::max(1u, std::min<unsigned>(cdiv(nk, W), 1024u));
}

// The k range of solve blockIdx.y: offset ko (k points before it), count nk
// and the workgroups gx its launch would have alone (loop_blocks).  With
// seg == nullptr every solve has nk_uniform points and the grid is its own.
struct SolveSeg {
    size_t ko;
    int nk, gx;
};
template <int W>
__device__ __forceinline__ SolveSeg solve_seg(const int* seg, int nk_uniform) {
    const int sv = blockIdx.y;
    SolveSeg s;
    if (seg) {
        s.ko = static_cast<size_t>(seg[sv]);
        s.nk = seg[sv + 1] - seg[sv];
        s.gx = max(1, min((s.nk + W - 1) / W, 1024));
    } else {
        s.ko = static_cast<size_t>(sv) * nk_uniform;
        s.nk = nk_uniform;
        s.gx = static_cast<int>(gridDim.x);
    }
    return s;
}

inline void check_nb(int nb) {
    if (nb < 1 || nb > 32) throw std::invalid_argument("solver kernels support 1 <= nb <= 32");
}

// wavefronts per workgroup of HFGPU_LOOP_LAUNCH
inline int waves_for(int nb) { return nb <= 16 ? 4 : 1; }

}  // namespace loopk

// Launch a per-k loop kernel: four wavefronts (k points) per workgroup up to
// nb = 16 with nb a compile-time constant for 2, 4, 8, 12 and 16 (the index
// arithmetic and the small matrix products then unroll), one wavefront and a
// run-time nb above.  LDS is lds_wave bytes per wavefront; S rows of
// workgroups, one per solve of a batch.
#define HFGPU_LOOP_LAUNCH(KERNEL, nb, nk, lds_wave, st, S, ...)                               \
    do {                                                                                     \
        const size_t lw_ = (lds_wave);                                                       \
        const dim3 g4_(loopk::loop_blocks(nk, 4), (S)), b4_(256);                            \
        switch ((nb) > 16 ? -1 : (nb)) {                                                     \
            case -1:                                                                         \
                hipLaunchKernelGGL((KERNEL<1, 0>), dim3(loopk::loop_blocks(nk, 1), (S)),       \
                                   dim3(64), lw_, st, __VA_ARGS__);                          \
                break;                                                                       \
            case 2: hipLaunchKernelGGL((KERNEL<4, 2>), g4_, b4_, 4 * lw_, st, __VA_ARGS__); break;   \
            case 4: hipLaunchKernelGGL((KERNEL<4, 4>), g4_, b4_, 4 * lw_, st, __VA_ARGS__); break;   \
            case 8: hipLaunchKernelGGL((KERNEL<4, 8>), g4_, b4_, 4 * lw_, st, __VA_ARGS__); break;   \
            case 12: hipLaunchKernelGGL((KERNEL<4, 12>), g4_, b4_, 4 * lw_, st, __VA_ARGS__); break; \
            case 16: hipLaunchKernelGGL((KERNEL<4, 16>), g4_, b4_, 4 * lw_, st, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL((KERNEL<4, 0>), g4_, b4_, 4 * lw_, st, __VA_ARGS__); break;  \
        }                                                                                    \
    } while (0)

}  // namespace hfgpu
