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
// most 4096 eigenvalues one workgroup runs all iterations in one launch.
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

__device__ __forceinline__ double expit_d(double x) { return expit_fd(x); }

__device__ __forceinline__ MuState* pp_state(double* ws, int k) {
    return reinterpret_cast<MuState*>(ws + kStates + (k & 1) * kStateDoubles);
}
__device__ __forceinline__ double* pp_part(double* ws, int k) {
    return ws + ((k & 1) ? kPart1 : kPart0);
}

// ---- bracket ---------------------------------------------------------------

// Workgroup blockIdx.x of nblocks: its partial extremes of eps.
__device__ __forceinline__ void minmax_body(const double* __restrict__ eps, size_t n,
                                            double* __restrict__ partial, int nblocks,
                                            double* red) {
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(nblocks) * kBlock) {
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

__global__ void __launch_bounds__(kBlock)
k_minmax(const double* __restrict__ eps, size_t n, double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    minmax_body(eps, n, partial, static_cast<int>(gridDim.x), red);
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
// order (deterministic, identical in every workgroup).  Split in two so a
// caller can issue the partial loads (bisect_load) together with its other
// loads and apply them (bisect_reduce) once the state has arrived.
constexpr int kGroups = kBlock / kNodeStride;  // 8 row groups of 32 node slots
constexpr int kPartPer = kMuBlocks / kGroups;  // partials per thread
static_assert(kPartPer * kGroups == kMuBlocks, "partials per row group");

__device__ __forceinline__ void bisect_load(const double* __restrict__ part, int nparts,
                                            int nodes, double (&v)[kPartPer]) {
    const int nd = threadIdx.x % kNodeStride, g = threadIdx.x / kNodeStride;
#pragma unroll
    for (int u = 0; u < kPartPer; ++u) {
        const int b = g + u * kGroups;
        v[u] = (nd < nodes && b < nparts) ? part[b * kNodeStride + nd] : 0.0;
    }
}

__device__ MuState bisect_reduce(MuState st, const double (&v)[kPartPer], int levels,
                                 double n_e) {
    __shared__ double s_mid[kMaxNodes], s_lo[kMaxNodes], s_hi[kMaxNodes];
    __shared__ double s_grp[kGroups][kNodeStride];
    __shared__ double s_cnt[kMaxNodes];
    const int nodes = (1 << levels) - 1;
    if (threadIdx.x == 0) build_tree(st.lo, st.hi, nodes, s_mid, s_lo, s_hi);
    const int nd = threadIdx.x % kNodeStride, g = threadIdx.x / kNodeStride;
    // the sum in workgroup order b = g, g + 8, ...
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < kPartPer; ++u) s += v[u];
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

__device__ MuState bisect_apply(MuState st, const double* __restrict__ part, int nparts,
                                int levels, double n_e) {
    double v[kPartPer];
    bisect_load(part, nparts, (1 << levels) - 1, v);
    return bisect_reduce(st, v, levels, n_e);
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
    // this thread's first two eigenvalues (all of them at the usual sizes):
    // loaded before the state update so both memory round trips overlap
    const size_t i0 = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x;
    const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
    double pe[2], pw[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const size_t i = i0 + q * stride;
        pe[q] = i < n ? eps[i] : 0.0;
        pw[q] = i < n ? w[i / nb] : 0.0;
    }
    MuState st;
    if (j == 0) {
        st = bracket_from(pp_part(ws, 1), prev_parts, 0, T, 0.0, red);
    } else {
        // the previous launch's partials and state, loads in flight together
        double v[kPartPer];
        bisect_load(pp_part(ws, j - 1), prev_parts, (1 << prev_levels) - 1, v);
        st = bisect_reduce(*pp_state(ws, j - 1), v, prev_levels, n_e);
    }
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
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (i0 + q * stride >= n) break;
#pragma unroll
        for (int nd = 0; nd < NODES; ++nd) acc[nd] += pw[q] * expit_d(-(pe[q] - mid[nd]) / denom);
    }
    for (size_t i = i0 + 2 * stride; i < n; i += stride) {
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
// Above 4096 eigenvalues the persistent multi-workgroup solve is faster: one
// workgroup's four wavefronts already fill its CU's SIMDs with the FP64
// exp work (a 36 x 36 x 8 solve took 87 us in one workgroup against about
// 30 us spread over 41).
constexpr size_t kNewtonSoloMax = 4096;

__device__ __forceinline__ void newton_solo_body(
    const double* __restrict__ eps, const double* __restrict__ w_norm, int nb, size_t n,
    double n_target, double T, double mu0, int mu0_from_state, const MuState* warm, int maxiter,
    int mm_parts, double* ws, MuState* state, double* red, double* part, double* occ) {
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
    // launch_occupations (mode 1) from the final mu
    if (occ)
        for (size_t i = threadIdx.x; i < n; i += kBlock) occ[i] = expit_d((st.mu - eps[i]) / T);
}

__global__ void __launch_bounds__(kBlock)
k_newton_solo(const double* __restrict__ eps, const double* __restrict__ w_norm, int nb,
              size_t n, double n_target, double T, double mu0, int mu0_from_state,
              const MuState* warm, int maxiter, int mm_parts, double* ws,
              MuState* state, double* occ) {  // warm may be state: read before the first barrier
    __shared__ double red[kBlock / 64];
    __shared__ double part[2];
    newton_solo_body(eps, w_norm, nb, n, n_target, T, mu0, mu0_from_state, warm, maxiter,
                     mm_parts, ws, state, red, part, occ);
}

__global__ void __launch_bounds__(kBlock)
k_newton_last(double n_target, int j, int prev_parts, int apply, double* ws, MuState* state) {
    __shared__ double red[kBlock / 64];
    MuState st = *pp_state(ws, j - 1);
    if (apply && !st.done) st = newton_apply(st, pp_part(ws, j - 1), prev_parts, n_target, red);
    if (threadIdx.x == 0) *state = st;
}

// ---- persistent multi-workgroup solves ------------------------------------
// One launch runs a whole bisection or Newton solve on G workgroups.  Each
// round ends with a grid-wide exchange of a few partial sums through
// "slots" in the workspace: a workgroup publishes its values with atomic
// stores at agent scope, and every workgroup gathers all G x K values with
// atomic loads of the same scope, polling each slot until it no longer holds
// the empty pattern (a NaN payload no arithmetic produces) that a fill
// kernel wrote before the launch.  The data is its own flag, so no fence or
// cache write-back is needed; atomics on one location are coherent across
// the XCDs.  Every workgroup then sums the G values in the same fixed order
// and takes the same decision, so all of them follow one trajectory and
// leave the loop together.  Polls are bounded: a timeout (workgroups that
// could not all be resident, which the small grids here make unlikely)
// leaves NaN in the state instead of hanging.

constexpr int kPersistSlotBase = 16384;  // after the multi-launch layout
constexpr int kPersistK = 8;             // slot doubles per workgroup per round
// At most 64 workgroups per persistent solve: a persistent grid only makes
// progress once all its workgroups are resident, and solves run side by side
// on streams (16 of them at 64 x 4 wavefronts fill half the workgroup slots
// of the GPU, so they can all be resident); 256 measured no faster.
constexpr int kPersistMaxG = 64;
constexpr unsigned long long kEmptySlot = 0x7FF4DEADBEEF0001ull;
constexpr long long kSpinMax = 1ll << 24;

__global__ void k_fill_slots(unsigned long long* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = kEmptySlot;
}

__device__ __forceinline__ void slot_put(double* s, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(s),
                       static_cast<unsigned long long>(__double_as_longlong(v)),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long slot_get(const double* s) {
    return __hip_atomic_load(reinterpret_cast<const unsigned long long*>(s), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Publish mine[0..K) (LDS, thread k < K holds the value) for round r and
// gather every workgroup's into all[g * K + k] (LDS).  Returns false after a
// poll timeout (uniform over the workgroup).
template <int K>
__device__ bool slot_exchange(double* slots, int r, int G, const double* mine, double* all,
                              int* s_flag) {
    double* base = slots + static_cast<size_t>(r) * G * kPersistK;
    if (threadIdx.x < K) slot_put(base + blockIdx.x * kPersistK + threadIdx.x, mine[threadIdx.x]);
    if (threadIdx.x == 0) *s_flag = 0;
    constexpr int kPer = (kPersistMaxG * K + kBlock - 1) / kBlock;
    unsigned long long u[kPer];
    // every load of the first poll in flight at once
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int i = threadIdx.x + q * kBlock;
        u[q] = i < G * K ? slot_get(base + (i / K) * kPersistK + i % K) : 0ull;
    }
    bool ok = true;
    for (long long it = 0;; ++it) {
        bool empty = false;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = threadIdx.x + q * kBlock;
            if (i < G * K && u[q] == kEmptySlot) {
                u[q] = slot_get(base + (i / K) * kPersistK + i % K);
                empty = empty || u[q] == kEmptySlot;
            }
        }
        if (!empty) break;
        if (it > kSpinMax) {
            ok = false;
            break;
        }
        if (it < 8)
            __builtin_amdgcn_s_sleep(1);
        else
            __builtin_amdgcn_s_sleep(8);  // back off: polls are loads through the fabric
    }
    __syncthreads();  // s_flag reset before any thread reports
    if (!ok) *s_flag = 1;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int i = threadIdx.x + q * kBlock;
        if (i < G * K) all[i] = __longlong_as_double(static_cast<long long>(u[q]));
    }
    __syncthreads();
    return *s_flag == 0;
}

// Fixed-order reductions over the G gathered values of each k (identical in
// every workgroup): thread k < K walks g = 0 .. G - 1.
template <int K>
__device__ void slot_sum(const double* all, int G, double* out) {
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int g = 0; g < G; ++g) s += all[g * K + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// Extremes of eps over the grid (round 0 of both solves).
__device__ bool persist_minmax(const double* __restrict__ eps, size_t n, double* slots, int G,
                               double* s_mine, double* s_all, double* red, int* s_flag,
                               double& emin, double& emax) {
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(G) * kBlock) {
        lo = fmin(lo, eps[i]);
        hi = fmax(hi, eps[i]);
    }
    lo = block_min<kBlock>(lo, red);
    hi = block_max<kBlock>(hi, red);
    if (threadIdx.x == 0) {
        s_mine[0] = lo;
        s_mine[1] = hi;
    }
    __syncthreads();
    if (!slot_exchange<2>(slots, 0, G, s_mine, s_all, s_flag)) return false;
    emin = INFINITY;
    emax = -INFINITY;
    for (int g = 0; g < G; ++g) {
        emin = fmin(emin, s_all[2 * g]);
        emax = fmax(emax, s_all[2 * g + 1]);
    }
    __syncthreads();
    return true;
}

// find_mu_bisection (utils.hpp:49-75) as a multisection: each round
// evaluates the count at the 2^L - 1 midpoints of the next L levels (tree
// built with the sequential loop's (lo + hi) / 2) and walks them with the
// reference's `count > n_e` decisions.
template <int NODES>
__global__ void __launch_bounds__(kBlock)
k_bisect_persist(const double* __restrict__ eps, const double* __restrict__ w, int nb, size_t n,
                 double n_e, double T, int maxiter, int levels, double* slots,
                 MuState* __restrict__ state) {
    __shared__ double red[kBlock / 64];
    __shared__ double s_mid[kMaxNodes], s_lo[kMaxNodes], s_hi[kMaxNodes];
    __shared__ double s_red[kBlock / 64][kNodeStride];
    __shared__ double s_mine[kPersistK], s_cnt[kPersistK];
    __shared__ double s_all[kPersistMaxG * kPersistK];
    __shared__ int s_flag;
    const int G = gridDim.x;
    double emin, emax;
    MuState st{};
    if (!persist_minmax(eps, n, slots, G, s_mine, s_all, red, &s_flag, emin, emax)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st.mu = NAN;
            st.done = -1;
            *state = st;
        }
        return;
    }
    const double Tj = fmax(T, 1.0e-12);
    const double span = emax - emin + 10.0 * fmax(Tj, 1.0e-6);
    double lo = emin - span, hi = emax + span;
    const double denom = Tj + 1.0e-12;  // fermidirac_scalar(x, Tj)
    int r = 1;
    for (int done = 0; done < maxiter; ++r) {
        const int lv = min(levels, maxiter - done);
        const int nodes = (1 << lv) - 1;
        if (threadIdx.x == 0) build_tree(lo, hi, nodes, s_mid, s_lo, s_hi);
        __syncthreads();
        double mid[NODES], acc[NODES];
#pragma unroll
        for (int nd = 0; nd < NODES; ++nd) {
            mid[nd] = nd < nodes ? s_mid[nd] : 0.0;
            acc[nd] = 0.0;
        }
        for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
             i += static_cast<size_t>(G) * kBlock) {
            const double e = eps[i];
            const double wk = w[i / nb];
#pragma unroll
            for (int nd = 0; nd < NODES; ++nd)
                if (nd < nodes) acc[nd] += wk * expit_d(-(e - mid[nd]) / denom);
        }
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
        for (int nd = 0; nd < NODES; ++nd) {
            const double sv = wave_sum(acc[nd]);
            if (lane == 0) s_red[wave][nd] = sv;
        }
        __syncthreads();
        if (threadIdx.x < NODES) {
            double sv = 0.0;
            for (int wv = 0; wv < kBlock / 64; ++wv) sv += s_red[wv][threadIdx.x];
            s_mine[threadIdx.x] = sv;
        }
        __syncthreads();
        if (!slot_exchange<NODES>(slots, r, G, s_mine, s_all, &s_flag)) {
            lo = hi = NAN;
            break;
        }
        slot_sum<NODES>(s_all, G, s_cnt);
        int k = 0;
        for (int l = 0; l < lv; ++l) {
            const double m = s_mid[k];
            if (s_cnt[k] > n_e) {
                hi = m;
                k = 2 * k + 1;
            } else {
                lo = m;
                k = 2 * k + 2;
            }
        }
        done += lv;
        __syncthreads();  // s_mid / s_cnt are rewritten by the next round
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.emin = emin;
        st.emax = emax;
        st.lo = lo;
        st.hi = hi;
        st.mu = 0.5 * (lo + hi);
        st.done = isnan(st.mu) ? -1 : 0;
        *state = st;
    }
}

// solve_mu_inloop (utils.hpp:79-116): Newton with the bisection safeguard,
// stopping at an exact fixed point as the multi-launch path does.
struct PersistLds {
    double red[kBlock / 64];
    double s_mine[kPersistK], s_sum[kPersistK];
    double s_all[kPersistMaxG * 2];
    int s_flag;
};

// Workgroup blockIdx.x of the G that run one solve.
__device__ __forceinline__ void newton_persist_body(
    const double* __restrict__ eps, const double* __restrict__ w_norm, int nb, size_t n,
    double n_target, double T, double mu0, int mu0_from_state, const MuState* warm, int maxiter,
    double* slots, MuState* state, int G, PersistLds& L, double* occ) {
    double* red = L.red;
    double* s_mine = L.s_mine;
    double* s_sum = L.s_sum;
    double* s_all = L.s_all;
    int& s_flag = L.s_flag;
    // warm may be state: every workgroup reads it before block 0 writes it
    const double m0 = mu0_from_state ? warm->mu : mu0;
    double emin, emax;
    MuState st{};
    if (!persist_minmax(eps, n, slots, G, s_mine, s_all, red, &s_flag, emin, emax)) {
        st.mu = NAN;
        st.done = -1;
    } else {
        const double Tj = fmax(T, 1.0e-12);
        const double l = emin - 50.0 * Tj, h = emax + 50.0 * Tj;
        st.emin = emin;
        st.emax = emax;
        st.lo = l;
        st.hi = h;
        st.mu = m0 < l ? l : (h < m0 ? h : m0);  // std::clamp
        for (int j = 0; j < maxiter && !st.done; ++j) {
            const double mu = st.mu;
            double N = 0.0, Z = 0.0;
            for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
                 i += static_cast<size_t>(G) * kBlock) {
                const double wk = w_norm[i / nb];
                const double p = expit_d((mu - eps[i]) / Tj);
                N += wk * p;
                Z += wk * p * (1.0 - p) / Tj;
            }
            N = block_sum<kBlock>(N, red);
            Z = block_sum<kBlock>(Z, red);
            if (threadIdx.x == 0) {
                s_mine[0] = N;
                s_mine[1] = Z;
            }
            __syncthreads();
            if (!slot_exchange<2>(slots, j + 1, G, s_mine, s_all, &s_flag)) {
                st.mu = NAN;
                st.done = -1;
                break;
            }
            slot_sum<2>(s_all, G, s_sum);
            const double Ns = s_sum[0], Zs = s_sum[1];
            __syncthreads();
            // the update of newton_apply, from the exchanged sums
            double lo = st.lo, hi = st.hi;
            const double lo0 = lo, hi0 = hi;
            const double g = Ns - n_target;
            if (g < 0.0) lo = mu;
            if (g > 0.0) hi = mu;
            double mu_new = mu - g / fmax(Zs, 1.0e-18);
            if (g != 0.0 && mu_new == mu) {
                st.done = 1;
                break;
            }
            const double mu_bis = 0.5 * (lo + hi);
            if (mu_new <= lo || mu_new >= hi) mu_new = mu_bis;
            mu_new = mu_new < lo ? lo : (hi < mu_new ? hi : mu_new);
            if (!isfinite(mu_new)) mu_new = mu_bis;
            if (mu_new == mu && lo == lo0 && hi == hi0) st.done = 1;
            st.lo = lo;
            st.hi = hi;
            st.mu = mu_new;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *state = st;
    // launch_occupations (mode 1) from the final mu, which every workgroup
    // holds (the same sums and updates in each; a workgroup whose exchange
    // timed out has NaN)
    if (occ)
        for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
             i += static_cast<size_t>(G) * kBlock)
            occ[i] = expit_d((st.mu - eps[i]) / T);
}

__global__ void __launch_bounds__(kBlock)
k_newton_persist(const double* __restrict__ eps, const double* __restrict__ w_norm, int nb,
                 size_t n, double n_target, double T, double mu0, int mu0_from_state,
                 const MuState* warm, int maxiter, double* slots, MuState* state, double* occ) {
    __shared__ PersistLds L;
    newton_persist_body(eps, w_norm, nb, n, n_target, T, mu0, mu0_from_state, warm, maxiter,
                        slots, state, static_cast<int>(gridDim.x), L, occ);
}

// ---- one launch pair for the Newton solves of a batch -----------------------
// Solve s = blockIdx.y owns the nb eigenvalues of its k points seg[s] ..
// seg[s + 1] - 1 (weights w_norm + seg[s]), state + s, warm + s and the
// workspace ws + s kMuPartialDoubles.  Each solve runs the path
// launch_mu_newton takes for its size (one workgroup up to kNewtonSoloMax
// eigenvalues, else persist_grid(n) workgroups exchanging through its own
// slots) with the same decomposition, so every mu equals the single call's.

__device__ __forceinline__ int newton_grid_of(size_t n) {
    return static_cast<int>(n > kNewtonSoloMax
                                ? (n + kBlock - 1) / kBlock < static_cast<size_t>(kPersistMaxG)
                                      ? (n + kBlock - 1) / kBlock
                                      : static_cast<size_t>(kPersistMaxG)
                                : 1);
}
__device__ __forceinline__ int mm_blocks_of(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return static_cast<int>(b < 1 ? 1 : (b < static_cast<size_t>(kMuBlocks) ? b : kMuBlocks));
}

// Before k_newton_multi: the extremes' partials of the one-workgroup solves,
// the empty pattern in the slots of the persistent ones (rounds of G slots).
__global__ void __launch_bounds__(kBlock)
k_newton_multi_prep(const double* __restrict__ eps, int nb, const int* __restrict__ seg,
                    int rounds, double* ws, unsigned active) {
    __shared__ double red[kBlock / 64];
    const int sv = blockIdx.y;
    if (!((active >> sv) & 1u)) return;
    const size_t n = static_cast<size_t>(seg[sv + 1] - seg[sv]) * nb;
    double* w = ws + static_cast<size_t>(sv) * kMuPartialDoubles;
    if (n <= kNewtonSoloMax) {
        const int nbk = mm_blocks_of(n);
        if (static_cast<int>(blockIdx.x) >= nbk) return;
        minmax_body(eps + static_cast<size_t>(seg[sv]) * nb, n, pp_part(w, 1), nbk, red);
    } else {
        const int cnt = rounds * newton_grid_of(n) * kPersistK;
        unsigned long long* p = reinterpret_cast<unsigned long long*>(w + kPersistSlotBase);
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < cnt; i += gridDim.x * kBlock)
            p[i] = kEmptySlot;
    }
}

__global__ void __launch_bounds__(kBlock)
k_newton_multi(const double* __restrict__ eps, const double* __restrict__ w_norm, int nb,
               const int* __restrict__ seg, const double* __restrict__ n_target, double T,
               int mu0_from_state, const MuState* warm, int maxiter, double* ws,
               MuState* state, unsigned active, double* occ) {
    __shared__ PersistLds L;
    __shared__ double part[2];
    const int sv = blockIdx.y;
    if (!((active >> sv) & 1u)) return;
    const size_t ko = static_cast<size_t>(seg[sv]);
    const size_t n = static_cast<size_t>(seg[sv + 1] - seg[sv]) * nb;
    const int G = newton_grid_of(n);
    if (static_cast<int>(blockIdx.x) >= G) return;
    double* w = ws + static_cast<size_t>(sv) * kMuPartialDoubles;
    double* o = occ ? occ + ko * nb : nullptr;
    if (n <= kNewtonSoloMax)
        newton_solo_body(eps + ko * nb, w_norm + ko, nb, n, n_target[sv], T, 0.0,
                         mu0_from_state, warm + sv, maxiter, mm_blocks_of(n), w, state + sv,
                         L.red, part, o);
    else
        newton_persist_body(eps + ko * nb, w_norm + ko, nb, n, n_target[sv], T, 0.0,
                            mu0_from_state, warm + sv, maxiter, w + kPersistSlotBase,
                            state + sv, G, L, o);
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

// Row s = blockIdx.y of the grid: solve s of a batch (n values each, or
// with seg the nb values of its k points seg[s] .. seg[s + 1] - 1; its own
// state), skipped when its bit in active is clear.
__global__ void k_occupations(const double* __restrict__ eps, size_t n,
                              const MuState* __restrict__ state, double T, int mode,
                              double* __restrict__ out, unsigned active, const int* seg,
                              int nb) {
    const int sv = blockIdx.y;
    if (!((active >> sv) & 1u)) return;
    if (seg) {
        const size_t o = static_cast<size_t>(seg[sv]) * nb;
        n = static_cast<size_t>(seg[sv + 1] - seg[sv]) * nb;
        eps += o;
        out += o;
    } else {
        eps += sv * n;
        out += sv * n;
    }
    const double mu = state[sv].mu;
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

// Persistent (one-launch) solves: on unless HFGPU_MU_PERSIST=0.  Grid of G
// workgroups with about one eigenvalue per thread, at most 256 (one per CU).
static bool persist_on() {
    static const bool on = [] {
        const char* e = std::getenv("HFGPU_MU_PERSIST");
        return !(e && e[0] == '0');
    }();
    return on;
}
// The bisection's one-launch form is kept behind HFGPU_MU_PERSIST_BISECT=1:
// at 96 x 96 x 8 its 11 exchange rounds took 87 us against about 40 us for
// the multi-launch multisection, whose launch boundaries cost less than a
// grid-wide exchange through atomics.
static bool persist_bisect_on() {
    static const bool on = [] {
        const char* e = std::getenv("HFGPU_MU_PERSIST_BISECT");
        return e && e[0] == '1';
    }();
    return on && persist_on();
}
static int persist_grid(size_t n) {
    return static_cast<int>(std::max<size_t>(1, std::min<size_t>(cdiv(n, kBlock), kPersistMaxG)));
}
static bool persist_fits(int rounds, int G) {
    return static_cast<size_t>(rounds) * G * kPersistK <=
           static_cast<size_t>(kMuPartialDoubles - kPersistSlotBase);
}
static void fill_slots(double* ws, int rounds, int G, hipStream_t st) {
    const int n = rounds * G * kPersistK;
    hipLaunchKernelGGL(k_fill_slots, dim3(std::max(1, std::min(64, (n + 255) / 256))), dim3(256), 0,
                       st, reinterpret_cast<unsigned long long*>(ws + kPersistSlotBase), n);
    HFGPU_LAUNCH_CHECK();
}

void launch_mu_bisection(const double* eps, const double* w, int nk, int nb, double n_e,
                         double T, int maxiter, MuState* state, double* ws, hipStream_t st) {
    if (maxiter <= 0) maxiter = 30;
    const size_t n = static_cast<size_t>(nk) * nb;
    if (persist_bisect_on()) {
        constexpr int kLv = 3;  // 7 midpoints per round (kPersistK slots)
        const int G = persist_grid(n);
        const int rounds = 1 + (maxiter + kLv - 1) / kLv;
        if (persist_fits(rounds, G)) {
            fill_slots(ws, rounds, G, st);
            hipLaunchKernelGGL(k_bisect_persist<7>, dim3(G), dim3(kBlock), 0, st, eps, w, nb, n,
                               n_e, T, maxiter, kLv, ws + kPersistSlotBase, state);
            HFGPU_LAUNCH_CHECK();
            return;
        }
    }
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
                      int maxiter, MuState* state, double* ws, hipStream_t st,
                      const MuState* warm, double* occ) {
    if (!warm) warm = state;
    const size_t n = static_cast<size_t>(nk) * nb;
    // a few eigenvalues per thread: each step also re-reduces the previous
    // step's partials in every workgroup
    const unsigned blocks = reduce_blocks(n, 4);
    const unsigned mm_blocks = reduce_blocks(n, 1);
    if (maxiter > 0 && n > kNewtonSoloMax && persist_on()) {
        const int G = persist_grid(n);
        if (persist_fits(1 + maxiter, G)) {
            fill_slots(ws, 1 + maxiter, G, st);
            hipLaunchKernelGGL(k_newton_persist, dim3(G), dim3(kBlock), 0, st, eps, w_norm, nb, n,
                               n_target_norm, T, mu0, mu0_from_state ? 1 : 0, warm, maxiter,
                               ws + kPersistSlotBase, state, occ);
            HFGPU_LAUNCH_CHECK();
            return;
        }
    }
    hipLaunchKernelGGL(k_minmax, dim3(mm_blocks), dim3(kBlock), 0, st, eps, n, ws + kPart1);
    HFGPU_LAUNCH_CHECK();
    if (maxiter > 0 && n <= kNewtonSoloMax) {
        hipLaunchKernelGGL(k_newton_solo, dim3(1), dim3(kBlock), 0, st, eps, w_norm, nb, n,
                           n_target_norm, T, mu0, mu0_from_state ? 1 : 0, warm, maxiter,
                           static_cast<int>(mm_blocks), ws, state, occ);
        HFGPU_LAUNCH_CHECK();
        return;
    }
    int prev_parts = static_cast<int>(mm_blocks);
    for (int j = 0; j < maxiter; ++j) {
        hipLaunchKernelGGL(k_newton_step, dim3(blocks), dim3(kBlock), 0, st, eps, w_norm, nb, n,
                           n_target_norm, T, mu0, mu0_from_state ? 1 : 0, warm, j, prev_parts,
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
                           size_t(0), n_target_norm, T, mu0, mu0_from_state ? 1 : 0, warm, 0,
                           prev_parts, ws);
        HFGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_newton_last, dim3(1), dim3(kBlock), 0, st, n_target_norm, 1,
                           0, 0, ws, state);
        HFGPU_LAUNCH_CHECK();
    }
    if (occ) launch_occupations(eps, n, state, T, 1, occ, st);
}

bool mu_newton_multi_ok(int nk_max, int nb, int maxiter) {
    if (maxiter <= 0 || !persist_on()) return false;
    const size_t n = static_cast<size_t>(nk_max) * nb;
    return n <= kNewtonSoloMax || persist_fits(1 + maxiter, persist_grid(n));
}

void launch_mu_newton_multi(const double* eps, const double* w_norm, const int* seg, int S,
                            int nk_max, int nb, const double* n_target, double T, int maxiter,
                            bool mu0_from_state, const MuState* warm, MuState* state,
                            double* ws, unsigned active, hipStream_t st, double* occ) {
    if (!mu_newton_multi_ok(nk_max, nb, maxiter))
        throw std::logic_error("launch_mu_newton_multi outside the one-launch Newton paths");
    const size_t n_max = static_cast<size_t>(nk_max) * nb;
    const int rounds = 1 + maxiter;
    // the largest prep grid: extremes of the one-workgroup solves, slot fills
    const unsigned gp = std::max<unsigned>(
        reduce_blocks(std::min<size_t>(n_max, kNewtonSoloMax), 1),
        n_max > kNewtonSoloMax ? static_cast<unsigned>(std::min(64, (rounds * persist_grid(n_max) *
                                                                     kPersistK + 255) / 256))
                               : 1u);
    hipLaunchKernelGGL(k_newton_multi_prep, dim3(gp, static_cast<unsigned>(S)), dim3(kBlock), 0,
                       st, eps, nb, seg, rounds, ws, active);
    HFGPU_LAUNCH_CHECK();
    const unsigned gm = n_max > kNewtonSoloMax ? static_cast<unsigned>(persist_grid(n_max)) : 1u;
    hipLaunchKernelGGL(k_newton_multi, dim3(gm, static_cast<unsigned>(S)), dim3(kBlock), 0, st,
                       eps, w_norm, nb, seg, n_target, T, mu0_from_state ? 1 : 0, warm, maxiter,
                       ws, state, active, occ);
    HFGPU_LAUNCH_CHECK();
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
                       0, st, eps, n, state, T, mode, out, ~0u, static_cast<const int*>(nullptr),
                       0);
    HFGPU_LAUNCH_CHECK();
}

void launch_occupations_batch(const double* eps, size_t n, int S, unsigned active,
                              const MuState* states, double T, int mode, double* out,
                              hipStream_t st, const int* seg, int nb) {
    hipLaunchKernelGGL(k_occupations,
                       dim3(std::min<unsigned>(cdiv(n, 256), 2048u), static_cast<unsigned>(S)),
                       dim3(256), 0, st, eps, n, states, T, mode, out, active, seg, nb);
    HFGPU_LAUNCH_CHECK();
}

void launch_level_shift(const double* eps, size_t n, const MuState* state, double T,
                        double shift, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_level_shift, dim3(std::min<unsigned>(cdiv(n, 256), 2048u)), dim3(256),
                       0, st, eps, n, state, T, shift, out);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
