// Per-k dense products on nb x nb blocks: density build and hermitisation.
//
// Reference: density_from_fock's P = V diag(f) V^H + hermitise
// (solver_scf.hpp:138-152), density_from_Qp (solver_dm.hpp:117-128),
// hermitize_inplace (linalg.hpp:20-33).
//
// One wavefront per k point, four per workgroup; the k point's V block is
// staged in LDS (transposed, padded) with coalesced 16 B loads, each lane
// then forms its own output elements.  At nb = 8 the density build moves 2A + a bytes
// (A = 16 nk nb^2) for 8 nb^3 nk flops: HBM-bound.
#include "kernels.h"

namespace hfgpu {

namespace {

// V is staged transposed with a padded stride, sVt[n * (nb + 1) + j] = V_jn:
// the lanes of an output row read consecutive j (no LDS bank conflicts; the
// row-major block read with stride nb put every lane of a row on one bank
// group at nb = 16), and the row index i is a broadcast.  NB > 0 fixes nb at
// compile time.
//
// Occupations come from f, or (f null) from the eigenvalues and the chemical
// potential as launch_occupations forms them: mode 0 (SCF)
// expit(-(eps - mu) / (T + 1e-12)), mode 1 (DM) expit((mu - eps) / T).
//
// hermitize: P = 1/2 (X + X^H) with X = V diag(f) V^H.  Each lane forms its
// own X_ij once; the mirrored X_ji comes from its lane through LDS (the same
// sum, so the result equals the hermitisation of X bit for bit, and P comes
// out exactly Hermitian).
template <int NB>
__global__ void __launch_bounds__(256)
k_density(int nk, int nb_arg, const cd* __restrict__ V, const double* __restrict__ f,
          const double* __restrict__ eps, const MuState* __restrict__ mus, double T,
          int occ_mode, cd* __restrict__ P, int hermitize) {
    const int nb = NB > 0 ? NB : nb_arg;
    extern __shared__ cd smem[];
    const int nb2 = nb * nb, ld = nb + 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + wave;
    cd* sv = smem + wave * (nb * ld + nb + nb2);
    double* sf = reinterpret_cast<double*>(sv + nb * ld);
    cd* sx = sv + nb * ld + nb;
    if (k < nk) {
        const size_t kb = static_cast<size_t>(k) * nb2;
        for (int e = lane; e < nb2; e += 64) {
            const int j = e / nb, n = e - j * nb;
            sv[n * ld + j] = V[kb + e];
        }
        if (f) {
            for (int e = lane; e < nb; e += 64) sf[e] = f[static_cast<size_t>(k) * nb + e];
        } else {
            const double mu = mus->mu;
            for (int e = lane; e < nb; e += 64) {
                const double x = eps[static_cast<size_t>(k) * nb + e];
                sf[e] = occ_mode == 0 ? expit_fd(-(x - mu) / (T + 1.0e-12)) : expit_fd((mu - x) / T);
            }
        }
    }
    __syncthreads();
    if (k >= nk) return;
    const size_t kb = static_cast<size_t>(k) * nb2;
    for (int e = lane; e < nb2; e += 64) {
        const int i = e / nb, j = e - i * nb;
        cd pij{0.0, 0.0};
        for (int n = 0; n < nb; ++n)
            pij = cadd(pij, cmulc(cscale(sv[n * ld + i], sf[n]), sv[n * ld + j]));
        if (hermitize)
            sx[e] = pij;
        else
            P[kb + e] = pij;
    }
    if (!hermitize) return;
    __syncthreads();
    for (int e = lane; e < nb2; e += 64) {
        const int i = e / nb, j = e - i * nb;
        const cd pij = sx[e], pji = sx[j * nb + i];
        P[kb + e] = cd{0.5 * (pij.x + pji.x), 0.5 * (pij.y - pji.y)};
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

static void density_launch(int nk, int nb, const cd* V, const double* f, const double* eps,
                           const MuState* mus, double T, int mode, cd* P, bool hermitize,
                           hipStream_t st) {
    const size_t lds = 4 * (sizeof(cd) * (nb * (nb + 1) + nb + nb * nb));
    const dim3 grid(cdiv(nk, 4)), block(256);
    const int h = hermitize ? 1 : 0;
    switch (nb) {
        case 2: hipLaunchKernelGGL(k_density<2>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h); break;
        case 4: hipLaunchKernelGGL(k_density<4>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h); break;
        case 8: hipLaunchKernelGGL(k_density<8>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h); break;
        case 12: hipLaunchKernelGGL(k_density<12>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h); break;
        case 16: hipLaunchKernelGGL(k_density<16>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h); break;
        default: hipLaunchKernelGGL(k_density<0>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h); break;
    }
    HFGPU_LAUNCH_CHECK();
}

void launch_density(int nk, int nb, const cd* V, const double* f, cd* P, bool hermitize,
                    hipStream_t st) {
    density_launch(nk, nb, V, f, nullptr, nullptr, 0.0, 0, P, hermitize, st);
}

void launch_density_occ(int nk, int nb, const cd* V, const double* eps, const MuState* mus,
                        double T, int mode, cd* P, bool hermitize, hipStream_t st) {
    density_launch(nk, nb, V, nullptr, eps, mus, T, mode, P, hermitize, st);
}

void launch_hermitize(int nk, int nb, cd* M, hipStream_t st) {
    const size_t total = static_cast<size_t>(nk) * nb * (nb + 1) / 2;
    hipLaunchKernelGGL(k_hermitize, dim3(std::min<unsigned>(cdiv(total, 256), 4096u)), dim3(256),
                       0, st, nk, nb, M);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
