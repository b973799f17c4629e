// Per-k dense products on nb x nb blocks: density build and hermitisation.
//
// Reference: density_from_fock's P = V diag(f) V^H + hermitise
// (solver_scf.hpp:138-152), density_from_Qp (solver_dm.hpp:117-128),
// hermitize_inplace (linalg.hpp:20-33).
//
// One wavefront per k point, four per workgroup; the k point's V block is
// staged in LDS with coalesced 16 B loads, each lane then forms its own
// output elements.  At nb = 8 the density build moves 2A + a bytes
// (A = 16 nk nb^2) for 8 nb^3 nk flops: HBM-bound.
#include "kernels.h"

namespace hfgpu {

namespace {

__global__ void __launch_bounds__(256)
k_density(int nk, int nb, const cd* __restrict__ V, const double* __restrict__ f,
          cd* __restrict__ P, int hermitize) {
    extern __shared__ cd smem[];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + wave;
    cd* sv = smem + wave * (nb2 + nb);
    double* sf = reinterpret_cast<double*>(sv + nb2);
    if (k < nk) {
        const size_t kb = static_cast<size_t>(k) * nb2;
        for (int e = lane; e < nb2; e += 64) sv[e] = V[kb + e];
        for (int e = lane; e < nb; e += 64) sf[e] = f[static_cast<size_t>(k) * nb + e];
    }
    __syncthreads();
    if (k >= nk) return;
    const size_t kb = static_cast<size_t>(k) * nb2;
    for (int e = lane; e < nb2; e += 64) {
        const int i = e / nb, j = e % nb;
        cd pij{0.0, 0.0};
        for (int n = 0; n < nb; ++n) pij = cadd(pij, cmulc(cscale(sv[i * nb + n], sf[n]), sv[j * nb + n]));
        if (hermitize) {
            cd pji{0.0, 0.0};
            for (int n = 0; n < nb; ++n)
                pji = cadd(pji, cmulc(cscale(sv[j * nb + n], sf[n]), sv[i * nb + n]));
            pij = cd{0.5 * (pij.x + pji.x), 0.5 * (pij.y - pji.y)};
        }
        P[kb + e] = pij;
    }
}

__global__ void k_hermitize(int nk, int nb, cd* __restrict__ M) {
    const int npair = nb * (nb + 1) / 2;
    const size_t total = static_cast<size_t>(nk) * npair;
    const int nb2 = nb * nb;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t k = t / npair;
        int r = static_cast<int>(t % npair);
        int a = 0;
        while (r >= nb - a) { r -= nb - a; ++a; }
        const int b = a + r;
        const size_t iab = k * nb2 + a * nb + b, iba = k * nb2 + b * nb + a;
        const cd x = M[iab], y = M[iba];
        const cd hv{0.5 * (x.x + y.x), 0.5 * (x.y - y.y)};
        M[iab] = hv;
        M[iba] = cconj(hv);
    }
}

}  // namespace

void launch_density(int nk, int nb, const cd* V, const double* f, cd* P, bool hermitize,
                    hipStream_t st) {
    const size_t lds = 4 * (sizeof(cd) * (nb * nb + nb));
    hipLaunchKernelGGL(k_density, dim3(cdiv(nk, 4)), dim3(256), lds, st, nk, nb, V, f, P,
                       hermitize ? 1 : 0);
    HFGPU_LAUNCH_CHECK();
}

void launch_hermitize(int nk, int nb, cd* M, hipStream_t st) {
    const size_t total = static_cast<size_t>(nk) * nb * (nb + 1) / 2;
    hipLaunchKernelGGL(k_hermitize, dim3(std::min<unsigned>(cdiv(total, 256), 4096u)), dim3(256),
                       0, st, nk, nb, M);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
