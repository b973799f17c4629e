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

// Orders the LDS accesses of one wavefront around it (a lane may then read
// what another lane of its wavefront wrote before it).  The LDS performs the
// accesses of one wavefront in order, so this costs no instruction: the
// fences only keep the compiler from moving accesses across it.  The per-k
// kernels keep each wavefront's blocks in its own LDS region and
// synchronise with it instead of the workgroup barrier, so the wavefronts of
// a workgroup (different k points) do not wait for each other.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Elements e = base + lane + 64 t (t < T) of C = op(A) op(B), nb x nb LDS
// blocks, each summed over n in its own fixed order and handed to
// store(e, value).  The T sums of a lane are independent chains run side by
// side (latency hidden within the lane); lanes past nb^2 compute on row 0
// and store nothing.  With tb the sum for (i, j) starts at n = j (the terms
// in rotated order), so the lanes of a row read B's rows B[j][n] on distinct
// LDS banks instead of a stride of nb complex (one bank group at nb = 16).
template <int T, class Store>
__device__ __forceinline__ void wave_mm_part(const cd* A, const cd* B, int nb, bool ta, bool tb,
                                             int lane, int base, Store&& store) {
    const int nb2 = nb * nb;
    int ii[T], jj[T], nn[T];
    cd s[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int e = base + lane + 64 * t;
        const bool ok = e < nb2;
        ii[t] = ok ? e / nb : 0;
        jj[t] = ok ? e % nb : 0;
        nn[t] = tb ? jj[t] : 0;
        s[t] = cd{0.0, 0.0};
    }
    // (unrolled by two only: a full unroll at a compile-time nb hoists every
    // LDS load and costs occupancy)
#pragma unroll 2
    for (int q = 0; q < nb; ++q) {
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int i = ii[t], j = jj[t], n = nn[t];
            const cd a = ta ? cconj(A[n * nb + i]) : A[i * nb + n];
            const cd b = tb ? cconj(B[j * nb + n]) : B[n * nb + j];
            s[t] = cadd(s[t], cmul(a, b));
            nn[t] = n + 1 == nb ? 0 : n + 1;
        }
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int e = base + lane + 64 * t;
        if (e < nb2) store(e, s[t]);
    }
}

// C = op(A) op(B) for 9..16 bands on the FP64 matrix cores: the blocks
// padded to 16 x 16 with zeros, Re C = Ar Br - Ai Bi and Im C = Ar Bi + Ai Br
// as four v_mfma_f64_16x16x4f64 steps over n each.  Operand lanes (gfx950):
// A lane l holds A[l % 16][l / 16] of a 16 x 4 step, B lane l holds
// B[l / 16][l % 16], result value r of lane l is C[4 r + l / 16][l % 16];
// every lane reads its 2 x 4 operands from LDS once; steps whose n range lies
// wholly in the padding are skipped.  All 64 lanes of the wavefront must be
// active.
typedef double loop_d4_t __attribute__((ext_vector_type(4)));

template <class Store>
__device__ __forceinline__ void wave_mm_mfma(const cd* A, const cd* B, int nb, bool ta, bool tb,
                                             int lane, Store&& store) {
    const int i = lane & 15, q = lane >> 4;
    loop_d4_t cre = {0.0, 0.0, 0.0, 0.0}, cim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (4 * c >= nb) continue;  // an all-zero step (uniform over the wavefront)
        const int n = 4 * c + q;
        const bool ok = i < nb && n < nb;
        cd a{0.0, 0.0}, b{0.0, 0.0};
        if (ok) {
            a = ta ? cconj(A[n * nb + i]) : A[i * nb + n];
            b = tb ? cconj(B[i * nb + n]) : B[n * nb + i];
        }
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, cim, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = 4 * r + q;
        if (row < nb && i < nb) store(row * nb + i, cd{cre[r], cim[r]});
    }
}

// store(e, C_e) for every element of C = op(A) op(B) by one wavefront: on the
// matrix cores for 9..16 bands when mfma, else up to four elements per lane
// at a time.
template <class Store>
__device__ __forceinline__ void wave_mm_each(const cd* A, const cd* B, int nb, bool ta, bool tb,
                                             int lane, Store&& store, bool mfma = false) {
    const int nb2 = nb * nb;
    if (mfma && nb > 8 && nb <= 16) {
        wave_mm_mfma(A, B, nb, ta, tb, lane, store);
    } else if (nb2 <= 64) {
        wave_mm_part<1>(A, B, nb, ta, tb, lane, 0, store);
    } else if (nb2 <= 128) {
        wave_mm_part<2>(A, B, nb, ta, tb, lane, 0, store);
    } else if (nb2 <= 192) {
        wave_mm_part<3>(A, B, nb, ta, tb, lane, 0, store);
    } else {
        for (int base = 0; base < nb2; base += 256)
            wave_mm_part<4>(A, B, nb, ta, tb, lane, base, store);
    }
}

// C = op(A) op(B) on nb x nb LDS blocks by one wavefront.
__device__ __forceinline__ void wave_mm(const cd* A, const cd* B, cd* C, int nb, bool ta,
                                        bool tb, int lane, bool mfma = false) {
    wave_mm_each(A, B, nb, ta, tb, lane, [C](int e, cd v) { C[e] = v; }, mfma);
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
