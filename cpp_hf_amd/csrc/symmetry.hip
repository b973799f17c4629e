// Device symmetry projector (symmetry.py:81-117 averages; the C description
// is hf_symmetry_desc).  One workgroup per k point holds A_k (or A_-k), one
// group element and a product in LDS; each thread owns up to four entries
// of the result.
#include "symmetry.h"

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hfgpu.h"

namespace hfgpu {
namespace {

constexpr int kSymThreads = 256;
constexpr int kSymMaxNb = 32;  // nb^2 <= 4 entries per thread

__device__ __forceinline__ int minus_index(int i, int n, int mod) {
    return mod ? (n - i) % n : n - 1 - i;
}

// T = G A for nb x nb matrices in LDS
__device__ __forceinline__ void lds_gemm(const cd* G, const cd* A, cd* T, int nb) {
    for (int e = threadIdx.x; e < nb * nb; e += kSymThreads) {
        const int i = e / nb, j = e % nb;
        cd s{0.0, 0.0};
        for (int a = 0; a < nb; ++a) s = cadd(s, cmul(G[i * nb + a], A[a * nb + j]));
        T[e] = s;
    }
}

// acc += T G^H, entries e = threadIdx.x + r * kSymThreads
__device__ __forceinline__ void lds_acc_gh(const cd* T, const cd* G, cd* acc, int nb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = threadIdx.x + r * kSymThreads;
        if (e < nb * nb) {
            const int i = e / nb, j = e % nb;
            cd s{0.0, 0.0};
            for (int b = 0; b < nb; ++b) s = cadd(s, cmulc(T[i * nb + b], G[j * nb + b]));
            acc[r] = cadd(acc[r], s);
        }
    }
}

// sum over one set of matrices g of g A g^H (A already in sA)
__device__ void conj_sum(const cd* __restrict__ set, int n, int nb, cd* sA, cd* sT, cd* sG,
                         cd* acc) {
    const int nb2 = nb * nb;
    for (int g = 0; g < n; ++g) {
        for (int e = threadIdx.x; e < nb2; e += kSymThreads)
            sG[e] = set[static_cast<size_t>(g) * nb2 + e];
        __syncthreads();
        lds_gemm(sG, sA, sT, nb);
        __syncthreads();
        lds_acc_gh(sT, sG, acc, nb);
        __syncthreads();
    }
}

// out_k = (sum_u u A_k u^H + sum_s s A_-k s^H) / m
__global__ void __launch_bounds__(kSymThreads)
k_sym_group(int nk2, int nk1, int nb, int n_u, int n_s, int s_mod, int s_axes, double m,
            const cd* __restrict__ U_set, const cd* __restrict__ S_set,
            const cd* __restrict__ A, cd* __restrict__ out) {
    extern __shared__ cd sm[];
    const int nb2 = nb * nb;
    cd* sA = sm;
    cd* sT = sA + nb2;
    cd* sG = sT + nb2;
    const int k = blockIdx.x;
    const int k1 = k / nk2, k2 = k % nk2;
    cd acc_u[4], acc_s[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc_u[r] = acc_s[r] = cd{0.0, 0.0};
    const cd* Ak = A + static_cast<size_t>(k) * nb2;
    for (int e = threadIdx.x; e < nb2; e += kSymThreads) sA[e] = Ak[e];
    conj_sum(U_set, n_u, nb, sA, sT, sG, acc_u);
    if (n_s > 0) {
        const int m1 = (s_axes & 1) ? minus_index(k1, nk1, s_mod) : k1;
        const int m2 = (s_axes & 2) ? minus_index(k2, nk2, s_mod) : k2;
        const cd* Am = A + (static_cast<size_t>(m1) * nk2 + m2) * nb2;
        for (int e = threadIdx.x; e < nb2; e += kSymThreads) sA[e] = Am[e];
        conj_sum(S_set, n_s, nb, sA, sT, sG, acc_s);
    }
    cd* ok = out + static_cast<size_t>(k) * nb2;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = threadIdx.x + r * kSymThreads;
        if (e < nb2) {
            const cd t = n_s > 0 ? cadd(acc_u[r], acc_s[r]) : acc_u[r];
            ok[e] = cd{t.x / m, t.y / m};
        }
    }
}

// out_k = (B_k + U conj(B_-k) U^H) / 2
__global__ void __launch_bounds__(kSymThreads)
k_sym_tr(int nk2, int nk1, int nb, int t_mod, int t_axes, const cd* __restrict__ U,
         const cd* __restrict__ B, cd* __restrict__ out) {
    extern __shared__ cd sm[];
    const int nb2 = nb * nb;
    cd* sA = sm;
    cd* sT = sA + nb2;
    cd* sG = sT + nb2;
    const int k = blockIdx.x;
    const int k1 = k / nk2, k2 = k % nk2;
    const int m1 = (t_axes & 1) ? minus_index(k1, nk1, t_mod) : k1;
    const int m2 = (t_axes & 2) ? minus_index(k2, nk2, t_mod) : k2;
    const cd* Bm = B + (static_cast<size_t>(m1) * nk2 + m2) * nb2;
    for (int e = threadIdx.x; e < nb2; e += kSymThreads) {
        sA[e] = cconj(Bm[e]);
        sG[e] = U[e];
    }
    __syncthreads();
    lds_gemm(sG, sA, sT, nb);
    __syncthreads();
    cd tr[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) tr[r] = cd{0.0, 0.0};
    lds_acc_gh(sT, sG, tr, nb);
    const cd* Bk = B + static_cast<size_t>(k) * nb2;
    cd* ok = out + static_cast<size_t>(k) * nb2;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = threadIdx.x + r * kSymThreads;
        if (e < nb2) ok[e] = cscale(cadd(Bk[e], tr[r]), 0.5);
    }
}

}  // namespace

void SymPlan::init(const hf_symmetry_desc& d, int nk1_, int nk2_, hipStream_t st) {
    nk1 = nk1_;
    nk2 = nk2_;
    nb = static_cast<int>(d.nb);
    if (d.nb == 0 || d.nb > static_cast<size_t>(kSymMaxNb))
        throw std::invalid_argument("the device symmetry projector supports 1 <= nb <= 32");
    if ((d.n_unitary && !d.unitary) || (d.n_spatial && !d.spatial))
        throw std::invalid_argument("symmetry description has a null group array");
    if (d.n_unitary + d.n_spatial > 4096)
        throw std::invalid_argument("symmetry group too large for the device projector");
    if ((d.spatial_axes & ~3) || (d.tr_axes & ~3))
        throw std::invalid_argument("symmetry k axes must be a subset of {0, 1}");
    const size_t nb2 = d.nb * d.nb, bytes = nb2 * sizeof(cd);
    n_s = static_cast<int>(d.n_spatial);
    s_mod = d.spatial_mod ? 1 : 0;
    s_axes = d.spatial_axes;
    t_mod = d.tr_mod ? 1 : 0;
    t_axes = d.tr_axes;
    has_tr = d.tr_U != nullptr;
    if (d.n_unitary > 0) {
        n_u = static_cast<int>(d.n_unitary);
        U_set.alloc(d.n_unitary * bytes);
        h2d(U_set.get(), d.unitary, d.n_unitary * bytes, st);
    } else if (n_s > 0) {
        // symmetry.py:97-99: the spatial set joins the identity
        n_u = 1;
        std::vector<cd> eye(nb2, cd{0.0, 0.0});
        for (size_t i = 0; i < d.nb; ++i) eye[i * d.nb + i] = cd{1.0, 0.0};
        U_set.alloc(bytes);
        h2d(U_set.get(), eye.data(), bytes, st);
        HFGPU_CHECK(hipStreamSynchronize(st));
    }
    if (n_s > 0) {
        S_set.alloc(d.n_spatial * bytes);
        h2d(S_set.get(), d.spatial, d.n_spatial * bytes, st);
    }
    if (has_tr) {
        TR.alloc(bytes);
        h2d(TR.get(), d.tr_U, bytes, st);
    }
    buf.alloc(static_cast<size_t>(nk1) * nk2 * bytes);
    HFGPU_CHECK(hipStreamSynchronize(st));
}

void SymPlan::apply(cd* M, hipStream_t st) {
    const size_t nk = static_cast<size_t>(nk1) * nk2;
    const size_t lds = 3 * static_cast<size_t>(nb) * nb * sizeof(cd);
    const size_t bytes = nk * nb * nb * sizeof(cd);
    if (n_u > 0) {
        const double m = static_cast<double>(n_s > 0 ? n_u + n_s : n_u);
        hipLaunchKernelGGL(k_sym_group, dim3(static_cast<unsigned>(nk)), dim3(kSymThreads), lds,
                           st, nk2, nk1, nb, n_u, n_s, s_mod, s_axes, m, U_set.as<cd>(),
                           S_set.as<cd>(), M, buf.as<cd>());
        HFGPU_LAUNCH_CHECK();
        if (has_tr) {
            hipLaunchKernelGGL(k_sym_tr, dim3(static_cast<unsigned>(nk)), dim3(kSymThreads), lds,
                               st, nk2, nk1, nb, t_mod, t_axes, TR.as<cd>(), buf.as<cd>(), M);
            HFGPU_LAUNCH_CHECK();
        } else {
            d2d(M, buf.get(), bytes, st);
        }
    } else if (has_tr) {
        hipLaunchKernelGGL(k_sym_tr, dim3(static_cast<unsigned>(nk)), dim3(kSymThreads), lds, st,
                           nk2, nk1, nb, t_mod, t_axes, TR.as<cd>(), M, buf.as<cd>());
        HFGPU_LAUNCH_CHECK();
        d2d(M, buf.get(), bytes, st);
    }
}

}  // namespace hfgpu

using namespace hfgpu;

extern "C" int hfgpu_project_symmetry(const hf_symmetry_desc* sym, double* M, size_t nk1,
                                      size_t nk2, size_t nb) {
    return guarded([&] {
        if (!sym || !M) throw std::invalid_argument("null argument");
        if (sym->nb != nb) throw std::invalid_argument("symmetry matrices must be (nb, nb)");
        if (nk1 == 0 || nk2 == 0) return;
        hipStream_t st = stream_acquire();
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)stream_sync_release(s); }
        } guard{st};
        SymPlan plan;
        plan.init(*sym, static_cast<int>(nk1), static_cast<int>(nk2), st);
        const size_t bytes = nk1 * nk2 * nb * nb * sizeof(cd);
        DevBuf dM(bytes);
        h2d(dM.get(), M, bytes, st);
        plan.apply(dM.as<cd>(), st);
        d2h(M, dM.get(), bytes, st);
        HFGPU_CHECK(hipStreamSynchronize(st));
    });
}
