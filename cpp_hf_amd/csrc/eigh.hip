// Batched Hermitian eigensolver: parallel cyclic Jacobi, one wavefront per
// matrix (reference: eigh_one / eigh_batched, linalg.hpp:39-66, and the
// Cayley setup eigh of herm(i d), solver_dm.hpp:206-222).
//
// The nb x nb matrix and its eigenvector accumulator live in LDS.  Each round
// of a sweep applies nb/2 disjoint complex Jacobi rotations at once (circle
// ordering), so every lane updates its own elements A'_ij = sum conj(J_ki)
// A_kl J_lj over the two rotation pairs that touch row i and column j.
// Sweeps stop once the off-diagonal mass falls to (4 nb eps)^2 ||A||_F^2.
#include <stdexcept>

#include "kernels.h"

namespace hfgpu {

namespace {

constexpr int kMaxNb = 32;
constexpr int kMaxSweeps = 40;

template <int MAXE>
__global__ void __launch_bounds__(64)
k_eigh_jacobi(int nmat, int nb, int in_mode, const cd* __restrict__ M, double* __restrict__ w_out,
              cd* __restrict__ V_out) {
    extern __shared__ cd smem[];
    const int nb2 = nb * nb;
    const int nbs = nb + 1;
    cd* A = smem;
    cd* V = A + nb2;
    cd* coefP = V + nb2;
    double* coefI = reinterpret_cast<double*>(coefP + nbs);
    double* newd = coefI + nbs;
    int* partner = reinterpret_cast<int*>(newd + nbs);
    int* rotated = partner + nbs;

    const int mat = blockIdx.x;
    if (mat >= nmat) return;
    const int lane = threadIdx.x;
    const cd* Mk = M + static_cast<size_t>(mat) * nb2;

    // load + hermitise: H = (X + X^H) / 2, X = M (in_mode 0) or i*M (in_mode 1)
    double nrm = 0.0;
    for (int e = lane; e < nb2; e += 64) {
        const int i = e / nb, j = e % nb;
        cd xij = Mk[e], xji = Mk[j * nb + i];
        if (in_mode == 1) {
            xij = cd{-xij.y, xij.x};
            xji = cd{-xji.y, xji.x};
        }
        const cd h{0.5 * (xij.x + xji.x), 0.5 * (xij.y - xji.y)};
        A[e] = h;
        V[e] = cd{i == j ? 1.0 : 0.0, 0.0};
        nrm += cnorm(h);
    }
    nrm = wave_sum(nrm);
    __syncthreads();

    const int nbp = nb + (nb & 1);  // padded to even; index nb is a dummy
    const int m = nbp - 1;
    const int npair = nbp / 2;
    const double eps = 2.220446049250313e-16;
    const double stop = (4.0 * nb * eps) * (4.0 * nb * eps) * nrm;

    int ei[MAXE], ej[MAXE];
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
        const int e = lane + 64 * q;
        ei[q] = e < nb2 ? e / nb : -1;
        ej[q] = e < nb2 ? e % nb : -1;
    }

    for (int sweep = 0; sweep < kMaxSweeps && nrm > 0.0 && nb > 1; ++sweep) {
        double off = 0.0;
#pragma unroll
        for (int q = 0; q < MAXE; ++q)
            if (ei[q] >= 0 && ei[q] < ej[q]) off += cnorm(A[ei[q] * nb + ej[q]]);
        off = wave_sum(off);
        if (off <= stop) break;

        for (int r = 0; r < m; ++r) {
            if (lane < npair) {
                int p, q;
                if (lane == 0) {
                    p = r;
                    q = nbp - 1;
                } else {
                    p = (r + lane) % m;
                    q = (r - lane + m) % m;
                }
                if (p > q) { const int t = p; p = q; q = t; }
                if (q >= nb) {  // paired with the dummy: identity for p
                    coefI[p] = 1.0;
                    coefP[p] = cd{0.0, 0.0};
                    partner[p] = p;
                    rotated[p] = 0;
                    newd[p] = A[p * nb + p].x;
                } else {
                    const cd b = A[p * nb + q];
                    const double ab = sqrt(cnorm(b));
                    const double a = A[p * nb + p].x, d = A[q * nb + q].x;
                    double c = 1.0, s = 0.0, t = 0.0;
                    cd e{1.0, 0.0};
                    if (ab > 0.0) {
                        const double tau = (d - a) / (2.0 * ab);
                        t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        e = cd{b.x / ab, b.y / ab};
                    }
                    // J_pp = J_qq = c, J_pq = s e, J_qp = -s conj(e)
                    coefI[p] = c;
                    coefI[q] = c;
                    coefP[p] = cd{-s * e.x, s * e.y};  // J_qp
                    coefP[q] = cd{s * e.x, s * e.y};   // J_pq
                    partner[p] = q;
                    partner[q] = p;
                    rotated[p] = rotated[q] = (ab > 0.0);
                    newd[p] = a - t * ab;
                    newd[q] = d + t * ab;
                }
            }
            __syncthreads();
            cd na[MAXE], nv[MAXE];
#pragma unroll
            for (int q = 0; q < MAXE; ++q) {
                const int i = ei[q], j = ej[q];
                if (i < 0) continue;
                const int ip = partner[i], jp = partner[j];
                const double ci = coefI[i], cj = coefI[j];
                const cd pi = coefP[i], pj = coefP[j];
                // rows mixed by conj(J_ki): k = i (ci), k = ip (conj pi)
                // cols mixed by J_lj:     l = j (cj), l = jp (pj)
                const cd a_ij = A[i * nb + j];
                cd acc = cscale(a_ij, ci * cj);
                if (jp != j) acc = cadd(acc, cscale(cmul(A[i * nb + jp], pj), ci));
                if (ip != i) {
                    const cd cpi = cconj(pi);
                    acc = cadd(acc, cscale(cmul(cpi, A[ip * nb + j]), cj));
                    if (jp != j) acc = cadd(acc, cmul(cmul(cpi, A[ip * nb + jp]), pj));
                }
                na[q] = acc;
                cd v = cscale(V[i * nb + j], cj);
                if (jp != j) v = cadd(v, cmul(V[i * nb + jp], pj));
                nv[q] = v;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < MAXE; ++q) {
                const int i = ei[q], j = ej[q];
                if (i < 0) continue;
                cd v = na[q];
                if (i == j) {
                    v = cd{newd[i], 0.0};
                } else if (partner[i] == j && rotated[i]) {
                    v = cd{0.0, 0.0};
                }
                A[i * nb + j] = v;
                V[i * nb + j] = nv[q];
            }
            __syncthreads();
        }
    }

    // ascending, stable order of the diagonal
    for (int i = lane; i < nb; i += 64) {
        const double wi = A[i * nb + i].x;
        int rank = 0;
        for (int j = 0; j < nb; ++j) {
            const double wj = A[j * nb + j].x;
            rank += (wj < wi) || (wj == wi && j < i);
        }
        partner[i] = rank;
    }
    __syncthreads();
    double* wk = w_out + static_cast<size_t>(mat) * nb;
    cd* Vk = V_out + static_cast<size_t>(mat) * nb2;
    for (int i = lane; i < nb; i += 64) wk[partner[i]] = A[i * nb + i].x;
    for (int e = lane; e < nb2; e += 64) {
        const int r = e / nb, j = e % nb;
        Vk[r * nb + partner[j]] = V[e];
    }
}

}  // namespace

void launch_eigh(int nmat, int nb, int in_mode, const cd* M, double* w, cd* V, hipStream_t st) {
    if (nb < 1 || nb > kMaxNb)
        throw std::invalid_argument("GPU eigh supports 1 <= nb <= 32");
    if (nmat == 0) return;
    const int nb2 = nb * nb;
    const size_t nbs = static_cast<size_t>(nb) + 1;
    const size_t lds = sizeof(cd) * (2 * static_cast<size_t>(nb2) + nbs) +
                       2 * sizeof(double) * nbs + 2 * sizeof(int) * nbs;
    if (nb2 <= 64)
        hipLaunchKernelGGL(k_eigh_jacobi<1>, dim3(nmat), dim3(64), lds, st, nmat, nb, in_mode, M, w, V);
    else if (nb2 <= 256)
        hipLaunchKernelGGL(k_eigh_jacobi<4>, dim3(nmat), dim3(64), lds, st, nmat, nb, in_mode, M, w, V);
    else
        hipLaunchKernelGGL(k_eigh_jacobi<16>, dim3(nmat), dim3(64), lds, st, nmat, nb, in_mode, M, w, V);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
