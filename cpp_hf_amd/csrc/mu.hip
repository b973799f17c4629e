// Chemical potential on the GPU: bracket, bisection, Newton, occupations.
//
// Reference: find_mu_bisection (utils.hpp:49-75), solve_mu_inloop
// (utils.hpp:79-116), fermidirac/expit (utils.hpp:18-29).
//
// Each solve keeps its scalar state (bracket, mu, done flag) in device memory,
// so a sequence of launches runs without host round trips.  Every launch
// writes per-workgroup partial sums; the next launch starts by reducing them
// in every workgroup, in the same fixed order, so all workgroups derive the
// same updated state without any cross-workgroup synchronisation (no
// device-scope fences: on a multi-XCD part each one writes back the XCD's
// L2), and the results are deterministic for a given launch shape.  Partials
// and states ping-pong between two buffers; a one-workgroup launch applies
// the last update.
//
// Bisection runs as a multisection: one launch evaluates the electron count
// at all 2^L - 1 midpoints of the next L bisection levels (the midpoint tree
// is built with the same (lo + hi) / 2 arithmetic as the sequential loop),
// and the next launch walks the tree with the reference's `count > n_e`
// decisions.  Each midpoint costs an expit per eigenvalue and each launch a
// fixed latency; L = 3 (7 midpoints, 30 steps in 10 launches) measured fastest
// at 96 x 96 x 8 (93 us of kernel time against 132 us for L = 5 and 167 us for
// L = 1); HFGPU_BISECT_LEVELS overrides it.  Brackets are bit-identical to the
// sequential loop's for every L.  Newton (solve_mu_inloop) takes one
// launch per iteration and stops computing at an exact fixed point; for at
// most 16384 eigenvalues one workgroup runs all iterations in one launch.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "kernels.h"

namespace hfgpu {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxLevels = 5;
constexpr int kMaxNodes = (1 << kMaxLevels) - 1;  // 31
constexpr int kNodeStride = 32;                   // partial doubles per workgroup
// workspace layout (doubles): partial buffers 0 and 1, then two states
constexpr int kPart0 = 0;
constexpr int kPart1 = 8192;
constexpr int kStates = 16368;
constexpr int kStateDoubles = 8;
static_assert(kPart1 + kMuBlocks * kNodeStride <= kStates, "mu workspace layout");
static_assert(kStates + 2 * kStateDoubles <= kMuPartialDoubles, "mu workspace size");
static_assert(sizeof(MuState) <= kStateDoubles * sizeof(double), "MuState slot");

// expit as in utils.hpp:18-25 (1 / (1 + e^-x) for x >= 0, e^x / (1 + e^x)
// otherwise), written without a branch: both forms share e = exp(-|x|), so
// lanes of either sign run the same instructions and round identically.
__device__ __forceinline__ double expit_d(double x) {
    const double e = exp(-fabs(x));
    return (x >= 0.0 ? 1.0 : e) / (1.0 + e);
}

__device__ __forceinline__ MuState* pp_state(double* ws, int k) {
    return reinterpret_cast<MuState*>(ws + kStates + (k & 1) * kStateDoubles);
}
__device__ __forceinline__ double* pp_part(double* ws, int k) {
    return ws + ((k & 1) ? kPart1 : kPart0);
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

// Bracket from the min/max partials (identical in every workgroup).
//   mode 0: bisection bracket (utils.hpp:53-59);
//   mode 1: Newton bracket + clamped warm start (utils.hpp:84-90).
__device__ MuState bracket_from(const double* __restrict__ part, int nparts, int mode, double T,
                                double mu0, double* red) {
    double lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
        lo = fmin(lo, part[2 * b]);
        hi = fmax(hi, part[2 * b + 1]);
    }
    const double emin = block_min<kBlock>(lo, red);
    const double emax = block_max<kBlock>(hi, red);
    const double Tj = fmax(T, 1.0e-12);
    MuState st{};
    st.emin = emin;
    st.emax = emax;
    if (mode == 0) {
        const double span = emax - emin + 10.0 * fmax(Tj, 1.0e-6);
        st.lo = emin - span;
        st.hi = emax + span;
        st.mu = 0.5 * (st.lo + st.hi);
    } else {
        const double l = emin - 50.0 * Tj, h = emax + 50.0 * Tj;
        st.lo = l;
        st.hi = h;
        st.mu = mu0 < l ? l : (h < mu0 ? h : mu0);  // std::clamp
    }
    st.done = 0;
    return st;
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

// Apply `levels` bisection steps to st from the per-workgroup node counts of
// the launch that evaluated st's midpoint tree.  Counts are summed in a fixed
// order (deterministic, identical in every workgroup).
__device__ MuState bisect_apply(MuState st, const double* __restrict__ part, int nparts,
                                int levels, double n_e) {
    constexpr int kGroups = kBlock / kNodeStride;  // 8 row groups of 32 node slots
    __shared__ double s_mid[kMaxNodes], s_lo[kMaxNodes], s_hi[kMaxNodes];
    __shared__ double s_grp[kGroups][kNodeStride];
    __shared__ double s_cnt[kMaxNodes];
    const int nodes = (1 << levels) - 1;
    if (threadIdx.x == 0) build_tree(st.lo, st.hi, nodes, s_mid, s_lo, s_hi);
    const int nd = threadIdx.x % kNodeStride, g = threadIdx.x / kNodeStride;
    double s = 0.0;
    if (nd < nodes) {
        // eight loads in flight per step
        for (int b0 = g; b0 < nparts; b0 += 8 * kGroups) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int b = b0 + u * kGroups;
                v[u] = b < nparts ? part[b * kNodeStride + nd] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
    }
    s_grp[g][nd] = s;
    __syncthreads();
    if (threadIdx.x < nodes) {
        double t = 0.0;
        for (int gg = 0; gg < kGroups; ++gg) t += s_grp[gg][threadIdx.x];
        s_cnt[threadIdx.x] = t;
    }
    __syncthreads();
    double l = s_lo[0], h = s_hi[0];
    int k = 0;
    for (int lev = 0; lev < levels; ++lev) {
        const double m = s_mid[k];
        if (s_cnt[k] > n_e) {
            h = m;
            k = 2 * k + 1;
        } else {
            l = m;
            k = 2 * k + 2;
        }
    }
    st.lo = l;
    st.hi = h;
    st.mu = 0.5 * (l + h);
    __syncthreads();  // shared arrays are reused by the caller
    return st;
}

// Step j of the bisection: bring the state up to date from the previous
// launch's partials (the min/max partials for j = 0), record it, then
// evaluate the electron count at the NODES midpoints of the next levels.
template <int NODES>
__global__ void __launch_bounds__(kBlock)
k_bisect_step(const double* __restrict__ eps, const double* __restrict__ w, int nb, size_t n,
              double n_e, double T, int j, int prev_levels, int prev_parts, double* ws) {
    __shared__ double red[kBlock / 64];
    __shared__ double s_mid[kMaxNodes], s_lo[kMaxNodes], s_hi[kMaxNodes];
    __shared__ double s_red[kBlock / 64][kNodeStride];
    MuState st;
    if (j == 0)
        st = bracket_from(pp_part(ws, 1), prev_parts, 0, T, 0.0, red);
    else
        st = bisect_apply(*pp_state(ws, j - 1), pp_part(ws, j - 1), prev_parts, prev_levels, n_e);
    if (blockIdx.x == 0 && threadIdx.x == 0) *pp_state(ws, j) = st;
    if (threadIdx.x == 0) build_tree(st.lo, st.hi, NODES, s_mid, s_lo, s_hi);
    __syncthreads();
    double mid[NODES], acc[NODES];
#pragma unroll
    for (int nd = 0; nd < NODES; ++nd) {
        mid[nd] = s_mid[nd];
        acc[nd] = 0.0;
    }
    const double denom = fmax(T, 1.0e-12) + 1.0e-12;  // fermidirac_scalar(x, Tj)
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock) {
        const double e = eps[i];
        const double wk = w[i / nb];
#pragma unroll
        for (int nd = 0; nd < NODES; ++nd) acc[nd] += wk * expit_d(-(e - mid[nd]) / denom);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int nd = 0; nd < NODES; ++nd) {
        const double s = wave_sum(acc[nd]);
        if (lane == 0) s_red[wave][nd] = s;
    }
    __syncthreads();
    if (threadIdx.x < NODES) {
        double s = 0.0;
        for (int wv = 0; wv < kBlock / 64; ++wv) s += s_red[wv][threadIdx.x];
        pp_part(ws, j)[blockIdx.x * kNodeStride + threadIdx.x] = s;
    }
}

// Last bisection levels applied by one workgroup into the caller's state.
__global__ void __launch_bounds__(kBlock)
k_bisect_last(double n_e, int j, int prev_levels, int prev_parts, double* ws, MuState* state) {
    const MuState st =
        bisect_apply(*pp_state(ws, j - 1), pp_part(ws, j - 1), prev_parts, prev_levels, n_e);
    if (threadIdx.x == 0) *state = st;
}

// ---- Newton ----------------------------------------------------------------

// One Newton update of st from summed N, Z (utils.hpp:105-115); marks an
// exact fixed point (the remaining iterations would reproduce it) as done.
__device__ MuState newton_apply(MuState st, const double* __restrict__ part, int nparts,
                                double n_target, double* red) {
    double N = 0.0, Z = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
        N += part[2 * b];
        Z += part[2 * b + 1];
    }
    const double Ns = block_sum<kBlock>(N, red);
    const double Zs = block_sum<kBlock>(Z, red);
    const double mu = st.mu;
    double lo = st.lo, hi = st.hi;
    const double lo0 = lo, hi0 = hi;
    const double g = Ns - n_target;
    if (g < 0.0) lo = mu;
    if (g > 0.0) hi = mu;
    const double Zsafe = fmax(Zs, 1.0e-18);
    double mu_new = mu - g / Zsafe;
    // A nonzero residual whose Newton step rounds away (mu - g/Z == mu) has
    // moved a bracket edge onto mu, and the safeguard below would replace the
    // step by a jump to the bracket middle.  The residual's last bits depend on
    // the summation order (the sequential sum can be exactly zero here, which
    // keeps mu), so mu is taken as converged instead of jumping.
    if (g != 0.0 && mu_new == mu) {
        st.done = 1;
        return st;
    }
    const double mu_bis = 0.5 * (lo + hi);
    if (mu_new <= lo || mu_new >= hi) mu_new = mu_bis;
    mu_new = mu_new < lo ? lo : (hi < mu_new ? hi : mu_new);
    if (!isfinite(mu_new)) mu_new = mu_bis;
    if (mu_new == mu && lo == lo0 && hi == hi0) st.done = 1;
    st.lo = lo;
    st.hi = hi;
    st.mu = mu_new;
    return st;
}

// Step j of solve_mu_inloop: apply iteration j - 1 (or set up the bracket for
// j = 0), record the state, then sum N and Z at the new mu.
__global__ void __launch_bounds__(kBlock)
k_newton_step(const double* __restrict__ eps, const double* __restrict__ w_norm, int nb,
              size_t n, double n_target, double T, double mu0, int mu0_from_state,
              const MuState* __restrict__ warm, int j, int prev_parts, double* ws) {
    __shared__ double red[kBlock / 64];
    MuState st;
    if (j == 0) {
        st = bracket_from(pp_part(ws, 1), prev_parts, 1, T, mu0_from_state ? warm->mu : mu0, red);
    } else {
        st = *pp_state(ws, j - 1);
        if (!st.done) st = newton_apply(st, pp_part(ws, j - 1), prev_parts, n_target, red);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *pp_state(ws, j) = st;
    if (st.done) return;  // uniform over the grid
    const double Tj = fmax(T, 1.0e-12);
    const double mu = st.mu;
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
        double* part = pp_part(ws, j);
        part[2 * blockIdx.x] = N;
        part[2 * blockIdx.x + 1] = Z;
    }
}

// The whole in-loop Newton solve in one workgroup, for few eigenvalues
// (n <= kNewtonSoloMax): the iterations of k_newton_step / k_newton_last
// without a launch each, ending at the exact fixed point (the multi-launch
// path runs every launch and lets the converged ones return at once).
constexpr size_t kNewtonSoloMax = 16384;

__global__ void __launch_bounds__(kBlock)
k_newton_solo(const double* __restrict__ eps, const double* __restrict__ w_norm, int nb,
              size_t n, double n_target, double T, double mu0, int mu0_from_state,
              const MuState* warm, int maxiter, int mm_parts, double* ws,
              MuState* state) {  // warm may be state: read before the first barrier
    __shared__ double red[kBlock / 64];
    __shared__ double part[2];
    MuState st = bracket_from(pp_part(ws, 1), mm_parts, 1, T,
                              mu0_from_state ? warm->mu : mu0, red);
    const double Tj = fmax(T, 1.0e-12);
    for (int j = 0; j < maxiter && !st.done; ++j) {  // st is uniform over the block
        const double mu = st.mu;
        double N = 0.0, Z = 0.0;
        for (size_t i = threadIdx.x; i < n; i += kBlock) {
            const double wk = w_norm[i / nb];
            const double p = expit_d((mu - eps[i]) / Tj);
            N += wk * p;
            Z += wk * p * (1.0 - p) / Tj;
        }
        N = block_sum<kBlock>(N, red);
        Z = block_sum<kBlock>(Z, red);
        if (threadIdx.x == 0) {
            part[0] = N;
            part[1] = Z;
        }
        __syncthreads();
        st = newton_apply(st, part, 1, n_target, red);
        __syncthreads();
    }
    if (threadIdx.x == 0) *state = st;
}

__global__ void __launch_bounds__(kBlock)
k_newton_last(double n_target, int j, int prev_parts, int apply, double* ws, MuState* state) {
    __shared__ double red[kBlock / 64];
    MuState st = *pp_state(ws, j - 1);
    if (apply && !st.done) st = newton_apply(st, pp_part(ws, j - 1), prev_parts, n_target, red);
    if (threadIdx.x == 0) *state = st;
}

// ---- the Python surface's Newton (utils.py:311-360) -------------------------
// One workgroup runs every step: the bracket from the extremes, maxiter
// Newton steps with the bisection safeguard on raw weights, then the final
// count check that falls back to the bracket midpoint.  Every thread derives
// the same scalar update from the block sums.
__global__ void __launch_bounds__(1024)
k_mu_newton_py(const double* __restrict__ eps, const double* __restrict__ w, int nb, size_t n,
               double n_e, double T, int maxiter, double* __restrict__ mu_out) {
    __shared__ double red[16];
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = threadIdx.x; i < n; i += 1024) {
        lo = fmin(lo, eps[i]);
        hi = fmax(hi, eps[i]);
    }
    const double emin = block_min<1024>(lo, red), emax = block_max<1024>(hi, red);
    const double Tj = fmax(T, 1.0e-12);
    const double span = emax - emin + 10.0 * fmax(Tj, 1.0e-6);
    lo = emin - span;
    hi = emax + span;
    double mu = 0.5 * (lo + hi);
    for (int it = 0; it <= maxiter; ++it) {
        double N = 0.0, Z = 0.0;
        for (size_t i = threadIdx.x; i < n; i += 1024) {
            const double wk = w[i / nb];
            const double p = expit_d((mu - eps[i]) / Tj);
            N += wk * p;
            Z += wk * p * (1.0 - p) / Tj;
        }
        N = block_sum<1024>(N, red);
        if (it == maxiter) {  // final check (utils.py:357-359)
            if (fabs(N - n_e) > 1e-12) mu = 0.5 * (lo + hi);
            break;
        }
        Z = block_sum<1024>(Z, red);
        const double g = N - n_e;
        if (g < 0.0) lo = mu;
        if (g > 0.0) hi = mu;
        const double Z_safe = fmax(Z, 1.0e-18);
        double mu_new = mu - g / Z_safe;
        const double mu_bis = 0.5 * (lo + hi);
        if (mu_new <= lo || mu_new >= hi || !isfinite(mu_new)) mu_new = mu_bis;
        mu = fmax(fmin(mu_new, hi), lo);
    }
    if (threadIdx.x == 0) *mu_out = mu;
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

unsigned reduce_blocks(size_t n, unsigned per_thread) {
    return std::max(1u, std::min<unsigned>(cdiv(n, kBlock * per_thread),
                                           static_cast<unsigned>(kMuBlocks)));
}

}  // namespace

void launch_mu_bisection(const double* eps, const double* w, int nk, int nb, double n_e,
                         double T, int maxiter, MuState* state, double* ws, hipStream_t st) {
    if (maxiter <= 0) maxiter = 30;
    const size_t n = static_cast<size_t>(nk) * nb;
    const unsigned blocks = reduce_blocks(n, 1);
    hipLaunchKernelGGL(k_minmax, dim3(blocks), dim3(kBlock), 0, st, eps, n, ws + kPart1);
    HFGPU_LAUNCH_CHECK();
    static const int max_levels = [] {
        const char* e = std::getenv("HFGPU_BISECT_LEVELS");
        const int v = e ? std::atoi(e) : 3;
        return v >= 1 && v <= kMaxLevels ? v : 3;
    }();
    int j = 0, prev_levels = 0;
    for (int done = 0; done < maxiter; ++j) {
        const int lv = std::min(max_levels, maxiter - done);
        const int pp = static_cast<int>(blocks);
        switch (lv) {
            case 1: hipLaunchKernelGGL(k_bisect_step<1>, dim3(blocks), dim3(kBlock), 0, st, eps, w, nb, n, n_e, T, j, prev_levels, pp, ws); break;
            case 2: hipLaunchKernelGGL(k_bisect_step<3>, dim3(blocks), dim3(kBlock), 0, st, eps, w, nb, n, n_e, T, j, prev_levels, pp, ws); break;
            case 3: hipLaunchKernelGGL(k_bisect_step<7>, dim3(blocks), dim3(kBlock), 0, st, eps, w, nb, n, n_e, T, j, prev_levels, pp, ws); break;
            case 4: hipLaunchKernelGGL(k_bisect_step<15>, dim3(blocks), dim3(kBlock), 0, st, eps, w, nb, n, n_e, T, j, prev_levels, pp, ws); break;
            default: hipLaunchKernelGGL(k_bisect_step<31>, dim3(blocks), dim3(kBlock), 0, st, eps, w, nb, n, n_e, T, j, prev_levels, pp, ws); break;
        }
        HFGPU_LAUNCH_CHECK();
        prev_levels = lv;
        done += lv;
    }
    hipLaunchKernelGGL(k_bisect_last, dim3(1), dim3(kBlock), 0, st, n_e, j, prev_levels,
                       static_cast<int>(blocks), ws, state);
    HFGPU_LAUNCH_CHECK();
}

void launch_mu_newton(const double* eps, const double* w_norm, int nk, int nb,
                      double n_target_norm, double mu0, bool mu0_from_state, double T,
                      int maxiter, MuState* state, double* ws, hipStream_t st) {
    const size_t n = static_cast<size_t>(nk) * nb;
    // a few eigenvalues per thread: each step also re-reduces the previous
    // step's partials in every workgroup
    const unsigned blocks = reduce_blocks(n, 4);
    const unsigned mm_blocks = reduce_blocks(n, 1);
    hipLaunchKernelGGL(k_minmax, dim3(mm_blocks), dim3(kBlock), 0, st, eps, n, ws + kPart1);
    HFGPU_LAUNCH_CHECK();
    if (maxiter > 0 && n <= kNewtonSoloMax) {
        hipLaunchKernelGGL(k_newton_solo, dim3(1), dim3(kBlock), 0, st, eps, w_norm, nb, n,
                           n_target_norm, T, mu0, mu0_from_state ? 1 : 0, state, maxiter,
                           static_cast<int>(mm_blocks), ws, state);
        HFGPU_LAUNCH_CHECK();
        return;
    }
    int prev_parts = static_cast<int>(mm_blocks);
    for (int j = 0; j < maxiter; ++j) {
        hipLaunchKernelGGL(k_newton_step, dim3(blocks), dim3(kBlock), 0, st, eps, w_norm, nb, n,
                           n_target_norm, T, mu0, mu0_from_state ? 1 : 0, state, j, prev_parts,
                           ws);
        HFGPU_LAUNCH_CHECK();
        prev_parts = static_cast<int>(blocks);
    }
    if (maxiter > 0) {
        hipLaunchKernelGGL(k_newton_last, dim3(1), dim3(kBlock), 0, st, n_target_norm, maxiter,
                           prev_parts, 1, ws, state);
        HFGPU_LAUNCH_CHECK();
    } else {
        // no iterations: the clamped warm start is the answer (utils.hpp:90)
        hipLaunchKernelGGL(k_newton_step, dim3(1), dim3(kBlock), 0, st, eps, w_norm, nb,
                           size_t(0), n_target_norm, T, mu0, mu0_from_state ? 1 : 0, state, 0,
                           prev_parts, ws);
        HFGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_newton_last, dim3(1), dim3(kBlock), 0, st, n_target_norm, 1,
                           0, 0, ws, state);
        HFGPU_LAUNCH_CHECK();
    }
}

void launch_mu_newton_py(const double* eps, const double* w, int nk, int nb, double n_e,
                         double T, int maxiter, double* mu_out, hipStream_t st) {
    const size_t n = static_cast<size_t>(nk) * nb;
    hipLaunchKernelGGL(k_mu_newton_py, dim3(1), dim3(1024), 0, st, eps, w, nb, n, n_e, T,
                       std::max(0, maxiter), mu_out);
    HFGPU_LAUNCH_CHECK();
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
