// Cayley spectral helpers and k-grid resampling.
//
// Reference: cayley_factor / cayley_unitary_from_spectrum /
// diag_UFU_from_spectrum (solver_dm.hpp:251-303) and
// resample_axis_periodic_linear / resample_kgrid_2d (utils.hpp:120-176).
#include "kernels.h"
#include "solver_kernels.h"

namespace hfgpu {

namespace {

// c(tau, lam) = (1 + i tau lam / 2) / (1 - i tau lam / 2)
__device__ __forceinline__ cd cayley_factor(double lam, double tau) {
    const double a = 0.5 * tau * lam;
    const double den = 1.0 + a * a;
    return cd{(1.0 - a * a) / den, (a + a) / den};
}

// U_k = V diag(c) V^H.  One wave per k, four per block.
__global__ void __launch_bounds__(256)
k_cayley_unitary(int nk, int nb, const cd* __restrict__ V, const double* __restrict__ lam,
                 double tau, cd* __restrict__ U) {
    extern __shared__ cd smem[];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + wave;
    cd* sv = smem + wave * (nb2 + nb);
    cd* sc = sv + nb2;
    if (k < nk) {
        const size_t kb = static_cast<size_t>(k) * nb2;
        for (int e = lane; e < nb2; e += 64) sv[e] = V[kb + e];
        for (int e = lane; e < nb; e += 64) sc[e] = cayley_factor(lam[static_cast<size_t>(k) * nb + e], tau);
    }
    __syncthreads();
    if (k >= nk) return;
    const size_t kb = static_cast<size_t>(k) * nb2;
    for (int e = lane; e < nb2; e += 64) {
        const int i = e / nb, j = e % nb;
        cd s{0.0, 0.0};
        for (int n = 0; n < nb; ++n) s = cadd(s, cmulc(cmul(sv[i * nb + n], sc[n]), sv[j * nb + n]));
        U[kb + e] = s;
    }
}

// diag_i = Re sum_j (V M)_ij conj(V_ij),  M_ij = conj(c_i) Fe_ij c_j.
__global__ void __launch_bounds__(256)
k_diag_ufu(int nk, int nb, const cd* __restrict__ V, const cd* __restrict__ Fe,
           const double* __restrict__ lam, double tau, double* __restrict__ diag) {
    extern __shared__ cd smem[];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + wave;
    cd* sv = smem + wave * (2 * nb2 + nb);
    cd* sm = sv + nb2;
    cd* sc = sm + nb2;
    if (k < nk) {
        const size_t kb = static_cast<size_t>(k) * nb2;
        for (int e = lane; e < nb; e += 64) sc[e] = cayley_factor(lam[static_cast<size_t>(k) * nb + e], tau);
        for (int e = lane; e < nb2; e += 64) sv[e] = V[kb + e];
    }
    __syncthreads();
    if (k < nk) {
        const size_t kb = static_cast<size_t>(k) * nb2;
        for (int e = lane; e < nb2; e += 64) {
            const int i = e / nb, j = e % nb;
            sm[e] = cmul(cmul(cconj(sc[i]), Fe[kb + e]), sc[j]);
        }
    }
    __syncthreads();
    if (k >= nk) return;
    // lane-per-(i, j) partial products, reduced over j with a row loop
    for (int i = lane; i < nb; i += 64) {
        double s = 0.0;
        for (int j = 0; j < nb; ++j) {
            cd a{0.0, 0.0};
            for (int n = 0; n < nb; ++n) a = cadd(a, cmul(sv[i * nb + n], sm[n * nb + j]));
            s += cmulc(a, sv[i * nb + j]).x;
        }
        diag[static_cast<size_t>(k) * nb + i] = s;
    }
}

// One resampling pass along an axis of length n_old -> n_new with `inner`
// contiguous complex values per index and `outer` independent slabs.
// Bit-exact with the reference's scalar loop: no FMA contraction, the
// complex products are written out exactly as std::complex<double> forms
// them ((1 - frac) + 0i) * x.
__global__ void k_resample_axis(const cd* __restrict__ x, cd* __restrict__ y, int n_old,
                                int n_new, size_t inner, size_t outer) {
#pragma clang fp contract(off)
    const size_t total = outer * static_cast<size_t>(n_new) * inner;
    const double ratio = static_cast<double>(n_old) / static_cast<double>(n_new);
    const long long s_new = n_new / 2, s_old = n_old / 2;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t i = t % inner;
        const size_t j = (t / inner) % static_cast<size_t>(n_new);
        const size_t o = t / (inner * static_cast<size_t>(n_new));
        const double u = (static_cast<double>(j) - static_cast<double>(s_new)) * ratio +
                         static_cast<double>(s_old);
        const double fu = floor(u);
        long long i0 = static_cast<long long>(fu);
        i0 = ((i0 % n_old) + n_old) % n_old;
        const long long i1 = (i0 + 1) % n_old;
        const double frac = u - fu;
        const double ar = 1.0 - frac, ai = 0.0, br = frac, bi = 0.0;
        const cd* xo = x + o * static_cast<size_t>(n_old) * inner;
        const cd x0 = xo[static_cast<size_t>(i0) * inner + i];
        const cd x1 = xo[static_cast<size_t>(i1) * inner + i];
        const double p0r = ar * x0.x - ai * x0.y, p0i = ar * x0.y + ai * x0.x;
        const double p1r = br * x1.x - bi * x1.y, p1i = br * x1.y + bi * x1.x;
        y[t] = cd{p0r + p1r, p0i + p1i};
    }
}

}  // namespace

void launch_cayley_unitary(int nk, int nb, const cd* V, const double* lam, double tau, cd* U,
                           hipStream_t st) {
    const size_t lds = 4 * sizeof(cd) * (nb * nb + nb);
    hipLaunchKernelGGL(k_cayley_unitary, dim3(cdiv(nk, 4)), dim3(256), lds, st, nk, nb, V, lam,
                       tau, U);
    HFGPU_LAUNCH_CHECK();
}

void launch_diag_ufu(int nk, int nb, const cd* V, const cd* Fe, const double* lam, double tau,
                     double* diag, hipStream_t st) {
    const size_t lds = 4 * sizeof(cd) * (2 * nb * nb + nb);
    hipLaunchKernelGGL(k_diag_ufu, dim3(cdiv(nk, 4)), dim3(256), lds, st, nk, nb, V, Fe, lam,
                       tau, diag);
    HFGPU_LAUNCH_CHECK();
}

void launch_resample_axis(const cd* x, cd* y, int n_old, int n_new, size_t inner, size_t outer,
                          hipStream_t st) {
    const size_t total = outer * static_cast<size_t>(n_new) * inner;
    if (total == 0) return;
    hipLaunchKernelGGL(k_resample_axis, dim3(std::min<unsigned>(cdiv(total, 256), 8192u)),
                       dim3(256), 0, st, x, y, n_old, n_new, inner, outer);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
