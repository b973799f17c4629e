// Chemical potential on the GPU: bracket, bisection, Newton, occupations.
//
// Reference: find_mu_bisection (utils.hpp:49-75), solve_mu_inloop
// (utils.hpp:79-116), fermidirac/expit (utils.hpp:18-29).
//
// Each solve keeps its scalar state (bracket, mu, done flag) in device memory,
// so a sequence of launches runs without host round trips.  Every reduction
// is two launches: a wide kernel writes per-workgroup partials, then a
// one-workgroup kernel sums them in workgroup order and updates the state, so
// results are deterministic for a given launch shape and no workgroup ever
// needs a device-scope fence (on a multi-XCD part each such fence writes back
// the XCD's L2, which costs far more than the extra launch).
//
// Bisection runs as a multisection: one launch evaluates the electron count
// at all 2^L - 1 midpoints of the next L bisection levels (the midpoint tree
// is built with the same (lo + hi) / 2 arithmetic as the sequential loop),
// then the finalising workgroup walks the tree with the reference's
// `count > n_e` decisions.  30 bisection steps cost 6 launch pairs instead of
// 30 dependent reductions, with bit-identical brackets.
#include <algorithm>
#include <cmath>

#include "kernels.h"

namespace hfgpu {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxLevels = 5;
constexpr int kMaxNodes = (1 << kMaxLevels) - 1;  // 31
constexpr int kNodeStride = 32;                   // partial doubles per workgroup

__device__ __forceinline__ double expit_d(double x) {
    if (x >= 0.0) return 1.0 / (1.0 + exp(-x));
    const double e = exp(x);
    return e / (1.0 + e);
}

// ---- bracket ---------------------------------------------------------------

__global__ void __launch_bounds__(kBlock)
k_minmax(const double* __restrict__ eps, size_t n, double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock) {
        lo = fmin(lo, eps[i]);
        hi = fmax(hi, eps[i]);
    }
    lo = block_min<kBlock>(lo, red);
    hi = block_max<kBlock>(hi, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = lo;
        partial[2 * blockIdx.x + 1] = hi;
    }
}

// mode 0: bisection bracket; mode 1: Newton bracket + clamped warm start.
__global__ void __launch_bounds__(kBlock)
k_minmax_final(const double* __restrict__ partial, int nparts, int mode, double T, double mu0,
               int mu0_from_state, MuState* state) {
    __shared__ double red[kBlock / 64];
    double lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
        lo = fmin(lo, partial[2 * b]);
        hi = fmax(hi, partial[2 * b + 1]);
    }
    const double emin = block_min<kBlock>(lo, red);
    const double emax = block_max<kBlock>(hi, red);
    if (threadIdx.x != 0) return;
    const double Tj = fmax(T, 1.0e-12);
    state->emin = emin;
    state->emax = emax;
    if (mode == 0) {
        const double span = emax - emin + 10.0 * fmax(Tj, 1.0e-6);
        state->lo = emin - span;
        state->hi = emax + span;
        state->mu = 0.5 * (state->lo + state->hi);
    } else {
        const double l = emin - 50.0 * Tj, h = emax + 50.0 * Tj;
        const double m0 = mu0_from_state ? state->mu : mu0;
        state->lo = l;
        state->hi = h;
        state->mu = m0 < l ? l : (h < m0 ? h : m0);  // std::clamp
    }
    state->done = 0;
}

// ---- bisection -------------------------------------------------------------

// Midpoint tree of the next `levels` bisection steps from [lo, hi]: node nd
// has children 2nd+1 (count > n_e: upper half dropped) and 2nd+2.
__device__ void build_tree(double lo, double hi, int nodes, double* s_mid, double* s_lo,
                           double* s_hi) {
    s_lo[0] = lo;
    s_hi[0] = hi;
    for (int nd = 0; nd < nodes; ++nd) {
        const double m = 0.5 * (s_lo[nd] + s_hi[nd]);
        s_mid[nd] = m;
        if (2 * nd + 2 < nodes) {
            s_lo[2 * nd + 1] = s_lo[nd];
            s_hi[2 * nd + 1] = m;
            s_lo[2 * nd + 2] = m;
            s_hi[2 * nd + 2] = s_hi[nd];
        }
    }
}

__global__ void __launch_bounds__(kBlock)
k_bisect(const double* __restrict__ eps, const double* __restrict__ w, int nb, size_t n,
         double Tj, int levels, const MuState* __restrict__ state, double* __restrict__ partial) {
    __shared__ double s_mid[kMaxNodes], s_lo[kMaxNodes], s_hi[kMaxNodes];
    __shared__ double s_red[kBlock / 64][kNodeStride];
    const int nodes = (1 << levels) - 1;
    if (threadIdx.x == 0) build_tree(state->lo, state->hi, nodes, s_mid, s_lo, s_hi);
    __syncthreads();
    double acc[kMaxNodes];
#pragma unroll
    for (int nd = 0; nd < kMaxNodes; ++nd) acc[nd] = 0.0;
    const double denom = Tj + 1.0e-12;  // fermidirac_scalar(x, Tj)
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock) {
        const double e = eps[i];
        const double wk = w[i / nb];
#pragma unroll
        for (int nd = 0; nd < kMaxNodes; ++nd)
            if (nd < nodes) acc[nd] += wk * expit_d(-(e - s_mid[nd]) / denom);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int nd = 0; nd < kMaxNodes; ++nd) {
        if (nd < nodes) {
            const double s = wave_sum(acc[nd]);
            if (lane == 0) s_red[wave][nd] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < nodes) {
        double s = 0.0;
        for (int wv = 0; wv < kBlock / 64; ++wv) s += s_red[wv][threadIdx.x];
        partial[blockIdx.x * kNodeStride + threadIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kBlock)
k_bisect_final(const double* __restrict__ partial, int nparts, double n_e, int levels,
               MuState* state) {
    __shared__ double s_mid[kMaxNodes], s_lo[kMaxNodes], s_hi[kMaxNodes], s_cnt[kMaxNodes];
    const int nodes = (1 << levels) - 1;
    if (threadIdx.x == 0) build_tree(state->lo, state->hi, nodes, s_mid, s_lo, s_hi);
    // wave v sums nodes v, v + 4, ...; lanes stride over the workgroups
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int nd = wave; nd < nodes; nd += kBlock / 64) {
        double s = 0.0;
        for (int b = lane; b < nparts; b += 64) s += partial[b * kNodeStride + nd];
        s = wave_sum(s);
        if (lane == 0) s_cnt[nd] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double l = s_lo[0], h = s_hi[0];
    int nd = 0;
    for (int lev = 0; lev < levels; ++lev) {
        const double m = s_mid[nd];
        if (s_cnt[nd] > n_e) {
            h = m;
            nd = 2 * nd + 1;
        } else {
            l = m;
            nd = 2 * nd + 2;
        }
    }
    state->lo = l;
    state->hi = h;
    state->mu = 0.5 * (l + h);
}

// ---- Newton ----------------------------------------------------------------

__global__ void __launch_bounds__(kBlock)
k_newton(const double* __restrict__ eps, const double* __restrict__ w_norm, int nb, size_t n,
         double Tj, const MuState* __restrict__ state, double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    if (state->done) return;
    const double mu = state->mu;
    double N = 0.0, Z = 0.0;
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock) {
        const double wk = w_norm[i / nb];
        const double p = expit_d((mu - eps[i]) / Tj);
        N += wk * p;
        Z += wk * p * (1.0 - p) / Tj;
    }
    N = block_sum<kBlock>(N, red);
    Z = block_sum<kBlock>(Z, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = N;
        partial[2 * blockIdx.x + 1] = Z;
    }
}

__global__ void __launch_bounds__(kBlock)
k_newton_final(const double* __restrict__ partial, int nparts, double n_target,
               MuState* state) {
    __shared__ double red[kBlock / 64];
    if (state->done) return;
    double N = 0.0, Z = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
        N += partial[2 * b];
        Z += partial[2 * b + 1];
    }
    const double Ns = block_sum<kBlock>(N, red);
    const double Zs = block_sum<kBlock>(Z, red);
    if (threadIdx.x != 0) return;
    const double mu = state->mu;
    double lo = state->lo, hi = state->hi;
    const double lo0 = lo, hi0 = hi;
    const double g = Ns - n_target;
    if (g < 0.0) lo = mu;
    if (g > 0.0) hi = mu;
    const double Zsafe = fmax(Zs, 1.0e-18);
    double mu_new = mu - g / Zsafe;
    const double mu_bis = 0.5 * (lo + hi);
    if (mu_new <= lo || mu_new >= hi) mu_new = mu_bis;
    mu_new = mu_new < lo ? lo : (hi < mu_new ? hi : mu_new);
    if (!isfinite(mu_new)) mu_new = mu_bis;
    // an unchanged (mu, lo, hi) is a fixed point of the iteration: the
    // remaining iterations would reproduce it exactly.
    if (mu_new == mu && lo == lo0 && hi == hi0) state->done = 1;
    state->lo = lo;
    state->hi = hi;
    state->mu = mu_new;
}

// ---- occupations -----------------------------------------------------------

__global__ void k_occupations(const double* __restrict__ eps, size_t n,
                              const MuState* __restrict__ state, double T, int mode,
                              double* __restrict__ out) {
    const double mu = state->mu;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        out[i] = mode == 0 ? expit_d(-(eps[i] - mu) / (T + 1.0e-12)) : expit_d((mu - eps[i]) / T);
    }
}

__global__ void k_level_shift(const double* __restrict__ eps, size_t n,
                              const MuState* __restrict__ state, double T, double shift,
                              double* __restrict__ out) {
    const double mu = state->mu;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const double occ = expit_d(-(eps[i] - mu) / (T + 1.0e-12));
        out[i] = eps[i] + shift * (1.0 - occ);
    }
}

unsigned reduce_blocks(size_t n) {
    return std::max(1u, std::min<unsigned>(cdiv(n, kBlock), static_cast<unsigned>(kMuBlocks)));
}

void launch_bracket(const double* eps, size_t n, int mode, double T, double mu0,
                    bool mu0_from_state, MuState* state, double* partial, hipStream_t st) {
    const unsigned blocks = reduce_blocks(n);
    hipLaunchKernelGGL(k_minmax, dim3(blocks), dim3(kBlock), 0, st, eps, n, partial);
    HFGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(kBlock), 0, st, partial,
                       static_cast<int>(blocks), mode, T, mu0, mu0_from_state ? 1 : 0, state);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace

void launch_mu_bisection(const double* eps, const double* w, int nk, int nb, double n_e,
                         double T, int maxiter, MuState* state, double* partial,
                         hipStream_t st) {
    if (maxiter <= 0) maxiter = 30;
    const size_t n = static_cast<size_t>(nk) * nb;
    const unsigned blocks = reduce_blocks(n);
    const double Tj = std::max(T, 1.0e-12);
    launch_bracket(eps, n, 0, T, 0.0, false, state, partial, st);
    for (int done = 0; done < maxiter;) {
        const int lv = std::min(kMaxLevels, maxiter - done);
        hipLaunchKernelGGL(k_bisect, dim3(blocks), dim3(kBlock), 0, st, eps, w, nb, n, Tj, lv,
                           state, partial);
        HFGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_bisect_final, dim3(1), dim3(kBlock), 0, st, partial,
                           static_cast<int>(blocks), n_e, lv, state);
        HFGPU_LAUNCH_CHECK();
        done += lv;
    }
}

void launch_mu_newton(const double* eps, const double* w_norm, int nk, int nb,
                      double n_target_norm, double mu0, bool mu0_from_state, double T,
                      int maxiter, MuState* state, double* partial, hipStream_t st) {
    const size_t n = static_cast<size_t>(nk) * nb;
    const unsigned blocks = reduce_blocks(n);
    const double Tj = std::max(T, 1.0e-12);
    launch_bracket(eps, n, 1, T, mu0, mu0_from_state, state, partial, st);
    for (int it = 0; it < maxiter; ++it) {
        hipLaunchKernelGGL(k_newton, dim3(blocks), dim3(kBlock), 0, st, eps, w_norm, nb, n, Tj,
                           state, partial);
        HFGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_newton_final, dim3(1), dim3(kBlock), 0, st, partial,
                           static_cast<int>(blocks), n_target_norm, state);
        HFGPU_LAUNCH_CHECK();
    }
}

void launch_occupations(const double* eps, size_t n, const MuState* state, double T, int mode,
                        double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_occupations, dim3(std::min<unsigned>(cdiv(n, 256), 2048u)), dim3(256),
                       0, st, eps, n, state, T, mode, out);
    HFGPU_LAUNCH_CHECK();
}

void launch_level_shift(const double* eps, size_t n, const MuState* state, double T,
                        double shift, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_level_shift, dim3(std::min<unsigned>(cdiv(n, 256), 2048u)), dim3(256),
                       0, st, eps, n, state, T, shift, out);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
