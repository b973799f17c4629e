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
#include <cstdlib>

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
// k point of wavefront index t: t itself, or in a batch (seg) the t-th k
// point of solve blockIdx.y (seg[s] + t; the solve's mu state is mus + s),
// mapped to nk (no k point) past the solve's end or for a solve whose bit in
// active is clear; nk is then the batch's total.
__device__ __forceinline__ int seg_k(int t, int nk, const int* seg, unsigned active) {
    if (!seg) return t;
    const int sv = blockIdx.y;
    const bool on = ((active >> sv) & 1u) && t < seg[sv + 1] - seg[sv];
    return on ? seg[sv] + t : nk;
}

template <int NB>
__global__ void __launch_bounds__(256)
k_density(int nk, int nb_arg, const cd* __restrict__ V, const double* __restrict__ f,
          const double* __restrict__ eps, const MuState* __restrict__ mus, double T,
          int occ_mode, cd* __restrict__ P, int hermitize, const int* __restrict__ seg,
          unsigned active) {
    const int nb = NB > 0 ? NB : nb_arg;
    extern __shared__ cd smem[];
    const int nb2 = nb * nb, ld = nb + 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (seg && mus) mus += blockIdx.y;
    const int k = seg_k(blockIdx.x * 4 + wave, nk, seg, active);
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

// 9 <= nb <= 16: the same product on the FP64 matrix cores.  One wavefront
// per k point; X = V diag(f) V^H is four real 16 x 16 x 16 products (padded
// with zeros past nb),
//   Re X = (f Vr) Vr^T + (f Vi) Vi^T,   Im X = (f Vi) Vr^T - (f Vr) Vi^T,
// each four v_mfma_f64_16x16x4f64 steps over n.  Operand lanes (gfx950): A
// lane l holds A[l % 16][l / 16], B lane l holds B[l / 16][l % 16], and
// result value r of lane l is D[4 r + l / 16][l % 16].  With B = V^T both
// operands of a step are lane l's V[l % 16][4 c + l / 16], so every lane
// loads four elements of V and nothing is staged.  The hermitisation reads
// the transposed element through LDS, as the VALU kernel does.
typedef double d4_t __attribute__((ext_vector_type(4)));

template <int NB>
__global__ void __launch_bounds__(256)
k_density_mfma(int nk, const cd* __restrict__ V, const double* __restrict__ f,
               const double* __restrict__ eps, const MuState* __restrict__ mus, double T,
               int occ_mode, cd* __restrict__ P, int hermitize, const int* __restrict__ seg,
               unsigned active) {
    static_assert(NB > 8 && NB <= 16, "matrix-core density: 9..16 bands");
    __shared__ double sx[4][2][16][17];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (seg && mus) mus += blockIdx.y;
    const int k = seg_k(blockIdx.x * 4 + wave, nk, seg, active);
    if (k >= nk) return;  // no workgroup barrier below: the LDS block is the wavefront's
    const int i = lane & 15, q = lane >> 4;
    const size_t kb = static_cast<size_t>(k) * NB * NB;
    cd v[4];
    double fn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int n = 4 * c + q;
        v[c] = (i < NB && n < NB) ? V[kb + i * NB + n] : cd{0.0, 0.0};
        if (n >= NB) {
            fn[c] = 0.0;
        } else if (f) {
            fn[c] = f[static_cast<size_t>(k) * NB + n];
        } else {
            const double mu = mus->mu, x = eps[static_cast<size_t>(k) * NB + n];
            fn[c] = occ_mode == 0 ? expit_fd(-(x - mu) / (T + 1.0e-12)) : expit_fd((mu - x) / T);
        }
    }
    d4_t cre = {0.0, 0.0, 0.0, 0.0}, cim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double ar = fn[c] * v[c].x, ai = fn[c] * v[c].y;
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, v[c].x, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, v[c].y, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, v[c].x, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar, v[c].y, cim, 0, 0, 0);
    }
    // lane holds X[a][i] for a = 4 r + q
    if (!hermitize) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int a = 4 * r + q;
            if (a < NB && i < NB) P[kb + a * NB + i] = cd{cre[r], cim[r]};
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        sx[wave][0][4 * r + q][i] = cre[r];
        sx[wave][1][4 * r + q][i] = cim[r];
    }
    __builtin_amdgcn_wave_barrier();  // one wavefront's LDS operations complete in order
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int a = 4 * r + q;
        if (a < NB && i < NB) {
            const double tr = sx[wave][0][i][a], ti = sx[wave][1][i][a];
            P[kb + a * NB + i] = cd{0.5 * (cre[r] + tr), 0.5 * (cim[r] - ti)};
        }
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

// The matrix-core product for 9..16 bands unless HFGPU_DENSITY_MFMA=0.
static bool density_mfma_on() {
    static const bool on = [] {
        const char* e = std::getenv("HFGPU_DENSITY_MFMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

// nk k points, or with seg the S solves' k points seg[s] .. seg[s + 1] - 1
// (nk the batch's total, nk_grid the largest solve's count).
static void density_launch(int nk, int nb, const cd* V, const double* f, const double* eps,
                           const MuState* mus, double T, int mode, cd* P, bool hermitize,
                           hipStream_t st, const int* seg = nullptr, int S = 1,
                           int nk_grid = 0, unsigned active = ~0u) {
    const size_t lds = 4 * (sizeof(cd) * (nb * (nb + 1) + nb + nb * nb));
    const dim3 grid(cdiv(seg ? nk_grid : nk, 4), seg ? S : 1), block(256);
    const int h = hermitize ? 1 : 0;
    if (nb > 8 && nb <= 16 && density_mfma_on()) {
        switch (nb) {
            case 9: hipLaunchKernelGGL(k_density_mfma<9>, grid, block, 0, st, nk, V, f, eps, mus, T, mode, P, h, seg, active); break;
            case 10: hipLaunchKernelGGL(k_density_mfma<10>, grid, block, 0, st, nk, V, f, eps, mus, T, mode, P, h, seg, active); break;
            case 11: hipLaunchKernelGGL(k_density_mfma<11>, grid, block, 0, st, nk, V, f, eps, mus, T, mode, P, h, seg, active); break;
            case 12: hipLaunchKernelGGL(k_density_mfma<12>, grid, block, 0, st, nk, V, f, eps, mus, T, mode, P, h, seg, active); break;
            case 13: hipLaunchKernelGGL(k_density_mfma<13>, grid, block, 0, st, nk, V, f, eps, mus, T, mode, P, h, seg, active); break;
            case 14: hipLaunchKernelGGL(k_density_mfma<14>, grid, block, 0, st, nk, V, f, eps, mus, T, mode, P, h, seg, active); break;
            case 15: hipLaunchKernelGGL(k_density_mfma<15>, grid, block, 0, st, nk, V, f, eps, mus, T, mode, P, h, seg, active); break;
            default: hipLaunchKernelGGL(k_density_mfma<16>, grid, block, 0, st, nk, V, f, eps, mus, T, mode, P, h, seg, active); break;
        }
        HFGPU_LAUNCH_CHECK();
        return;
    }
    switch (nb) {
        case 2: hipLaunchKernelGGL(k_density<2>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h, seg, active); break;
        case 4: hipLaunchKernelGGL(k_density<4>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h, seg, active); break;
        case 8: hipLaunchKernelGGL(k_density<8>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h, seg, active); break;
        case 12: hipLaunchKernelGGL(k_density<12>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h, seg, active); break;
        case 16: hipLaunchKernelGGL(k_density<16>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h, seg, active); break;
        default: hipLaunchKernelGGL(k_density<0>, grid, block, lds, st, nk, nb, V, f, eps, mus, T, mode, P, h, seg, active); break;
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

void launch_density_seg(const int* seg, int S, int nk_max, int nk_total, unsigned active,
                        int nb, const cd* V, const double* f, cd* P, bool hermitize,
                        hipStream_t st) {
    density_launch(nk_total, nb, V, f, nullptr, nullptr, 0.0, 0, P, hermitize, st, seg, S,
                   nk_max, active);
}

void launch_hermitize(int nk, int nb, cd* M, hipStream_t st) {
    const size_t total = static_cast<size_t>(nk) * nb * (nb + 1) / 2;
    hipLaunchKernelGGL(k_hermitize, dim3(std::min<unsigned>(cdiv(total, 256), 4096u)), dim3(256),
                       0, st, nk, nb, M);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
