// Batched Hermitian eigensolver: parallel cyclic Jacobi, one wavefront per
// matrix (reference: eigh_one / eigh_batched, linalg.hpp:39-66, and the
// Cayley setup eigh of herm(i d), solver_dm.hpp:206-222).
//
// Each round of a sweep applies the nb/2 disjoint complex Jacobi rotations of
// the circle ordering at once; sweeps stop once the off-diagonal mass falls to
// (4 nb eps)^2 ||A||_F^2; the eigenvalues are returned in ascending, stable
// order with the eigenvectors as columns.  nb <= 8 runs the register-row
// kernel (k_eigh_rows), larger nb the LDS kernel (k_eigh_jacobi), where the
// matrix and its eigenvector accumulator live in LDS and every lane updates
// its own elements A'_ij = sum conj(J_ki) A_kl J_lj.
#include <algorithm>
#include <stdexcept>

#include "kernels.h"

namespace hfgpu {

namespace {

constexpr int kMaxNb = 32;
constexpr int kMaxSweeps = 40;

template <int MAXE>
__global__ void __launch_bounds__(64)
k_eigh_jacobi(int nmat, int nb, int in_mode, const cd* __restrict__ M, double* __restrict__ w_out,
              cd* __restrict__ V_out, BlockMap bm, int sort) {
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
        cd h{0.5 * (xij.x + xji.x), 0.5 * (xij.y - xji.y)};
        if (bm.blk[i] != bm.blk[j]) h = cd{0.0, 0.0};  // off-block: ignored
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

    // ascending, stable order of the diagonal (within each block unless sort)
    for (int i = lane; i < nb; i += 64) {
        const double wi = A[i * nb + i].x;
        int rank = sort ? 0 : bm.start[i];
        for (int j = 0; j < nb; ++j) {
            if (!sort && bm.blk[j] != bm.blk[i]) continue;
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

// ---------------------------------------------------------------------------
// Register-row variant for nb <= 8: NP (nb rounded up to 2, 4 or 8) lanes
// per matrix, lane i holding row i of A and of the eigenvector accumulator V
// in registers, 64 / NP matrices per wavefront and no LDS.  A round of the
// circle ordering pairs lane i with lane partner(i, r); both lanes of a pair
// compute the same rotation, the column update A <- A J (and V <- V J) is
// local to each lane, and the row update A <- J^H A swaps the partner's row
// through cross-lane shuffles.  Rounds are unrolled, so every register index
// is a compile-time constant.
// ---------------------------------------------------------------------------

__host__ __device__ constexpr int circle_partner(int i, int r, int np) {
    return i == np - 1 ? r : (i == r ? np - 1 : ((2 * r - i) % (np - 1) + (np - 1)) % (np - 1));
}

// Per-wavefront LDS exchange area of the register-row kernel: each lane's
// row (for the partner's row update), rotation parameters and diagonal.
// Lanes of one wavefront execute LDS operations in order, so a lane may read
// what another lane of its wavefront wrote earlier without a barrier; the
// wave barrier only keeps the compiler from moving LDS accesses across it.
template <int NP>
struct RowLds {
    cd row[64][NP];
    double par[64][4];  // c, Re(s e), Im(s e), t |b|
    double dg[64];
};

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// One round R of the circle ordering.  Every register index below is a
// compile-time constant: lane-dependent choices are made with selects over
// the round's (compile-time) pairs, never by indexing the row with the lane's
// own position, which would push the row out to scratch memory.  dg is this
// row's diagonal element, kept alongside the row.
template <int NP, int R>
__device__ __forceinline__ void jacobi_round(cd (&A)[NP], cd (&V)[NP], double& dg, int i,
                                             int lane, int gbase, int nb, RowLds<NP>& X) {
    const int pi = circle_partner(i, R, NP);
    const bool is_p = i < pi;
    const int pmin = is_p ? i : pi, q = is_p ? pi : i;
    // A_pq on lane p: lane k of pair (k, kp) reads its own column kp
    cd my_off{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int kp = circle_partner(k, R, NP);
        if (k < kp && i == k) my_off = A[kp];
    }
    X.dg[lane] = dg;
    wave_sync();
    const double other_d = X.dg[gbase + pi];
    // rotation of the pair (p, q) as seen from row p; row q takes row p's
    // values, so both rows of a pair apply the same J.  Rutishauser's form:
    // tau = (d - a) / 2|b|, t = sgn(tau) / (|tau| + sqrt(1 + tau^2)),
    // c = 1 / sqrt(1 + t^2), s = t c, e = b / |b|.  An unrotated pair keeps
    // c = 1, s = 0, t|b| = 0: the identity.
    double c = 1.0, tab = 0.0;
    cd se{0.0, 0.0};  // s e
    if (is_p && q < nb) {
        const double a = dg, d = other_d;
        const cd b = my_off;
        const double b2 = cnorm(b);
        if (b2 > 0.0) {
            const double inv_ab = rsqrt(b2);
            const double ab = b2 * inv_ab;
            const double tau = (d - a) * (0.5 * inv_ab);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = rsqrt(1.0 + t * t);
            const double sv = t * c * inv_ab;
            se = cd{sv * b.x, sv * b.y};
            tab = t * ab;
        }
    }
    X.par[lane][0] = c;
    X.par[lane][1] = se.x;
    X.par[lane][2] = se.y;
    X.par[lane][3] = tab;
    wave_sync();
    // column update A <- A J, V <- V J for every pair (k, kp) of the round,
    // with the rotation of lane k
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int kp = circle_partner(k, R, NP);
        if (k < kp) {
            const double* pk = X.par[gbase + k];
            const double ck = pk[0];
            const cd sek{pk[1], pk[2]};
            const cd x = A[k], y = A[kp];
            A[k] = csub(cscale(x, ck), cmul(cconj(sek), y));
            A[kp] = cadd(cmul(sek, x), cscale(y, ck));
            const cd vx = V[k], vy = V[kp];
            V[k] = csub(cscale(vx, ck), cmul(cconj(sek), vy));
            V[kp] = cadd(cmul(sek, vx), cscale(vy, ck));
        }
    }
    const double* pm = X.par[gbase + pmin];
    const double myc = pm[0], mytab = pm[3];
    const cd myse{pm[1], pm[2]};
    // row update A <- J^H A with the partner's (column-updated) row:
    //   row p' = c row p - s e row q,  row q' = s conj(e) row p + c row q
#pragma unroll
    for (int j = 0; j < NP; ++j) X.row[lane][j] = A[j];
    wave_sync();
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const cd other = X.row[gbase + pi][j];
        A[j] = is_p ? csub(cscale(A[j], myc), cmul(myse, other))
                    : cadd(cmul(cconj(myse), other), cscale(A[j], myc));
    }
    wave_sync();
    // exact values on the 2 x 2 block: a - t|b|, d + t|b|, 0
    dg = is_p ? dg - mytab : dg + mytab;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int kp = circle_partner(k, R, NP);
        if (k < kp) {
            if (i == k) {
                A[k] = cd{dg, 0.0};
                A[kp] = cd{0.0, 0.0};
            } else if (i == kp) {
                A[kp] = cd{dg, 0.0};
                A[k] = cd{0.0, 0.0};
            }
        }
    }
}

template <int NP, int R>
struct JacobiSweep {
    static __device__ __forceinline__ void run(cd (&A)[NP], cd (&V)[NP], double& dg, int i,
                                               int lane, int gbase, int nb, RowLds<NP>& X) {
        if constexpr (R < NP - 1) {
            jacobi_round<NP, R>(A, V, dg, i, lane, gbase, nb, X);
            JacobiSweep<NP, R + 1>::run(A, V, dg, i, lane, gbase, nb, X);
        }
    }
};

template <int NP>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = NP / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int NP>
__global__ void __launch_bounds__(256)
k_eigh_rows(int nmat, int nb, int in_mode, const cd* __restrict__ M, double* __restrict__ w_out,
            cd* __restrict__ V_out, BlockMap bm, int sort) {
    constexpr int G = 64 / NP;  // matrices per wavefront
    __shared__ RowLds<NP> lds[4];
    const int lane = threadIdx.x & 63;
    RowLds<NP>& X = lds[threadIdx.x >> 6];
    const int i = lane % NP;
    const int gbase = lane - i;
    const int mat = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * G + lane / NP;
    const bool live = mat < nmat;
    const int nb2 = nb * nb;
    const cd* Mk = M + static_cast<size_t>(live ? mat : 0) * nb2;
    cd A[NP], V[NP];
    double dg = 0.0, nrm = 0.0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        cd h{0.0, 0.0};
        if (live && i < nb && j < nb) {
            cd xij = Mk[i * nb + j], xji = Mk[j * nb + i];
            if (in_mode == 1) {
                xij = cd{-xij.y, xij.x};
                xji = cd{-xji.y, xji.x};
            }
            h = cd{0.5 * (xij.x + xji.x), 0.5 * (xij.y - xji.y)};
            if (bm.blk[i] != bm.blk[j]) h = cd{0.0, 0.0};  // off-block: ignored
        }
        A[j] = h;
        V[j] = cd{j == i ? 1.0 : 0.0, 0.0};
        if (j == i) dg = h.x;
        nrm += cnorm(h);
    }
    nrm = group_sum<NP>(nrm);
    const double eps = 2.220446049250313e-16;
    const double stop = (4.0 * nb * eps) * (4.0 * nb * eps) * nrm;
    for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
        double off = 0.0;
#pragma unroll
        for (int j = 0; j < NP; ++j)
            if (j > i) off += cnorm(A[j]);
        off = group_sum<NP>(off);
        // every matrix of the wavefront converged (or is empty / 1 x 1)
        if (__all(!(nrm > 0.0 && nb > 1) || off <= stop)) break;
        JacobiSweep<NP, 0>::run(A, V, dg, i, lane, gbase, nb, X);
    }
    // ascending, stable order of the diagonal over the real rows (within
    // each block unless sort)
    const double di = dg;
    const int bi = i < nb ? bm.blk[i] : -1;
    int rank = (sort || i >= nb) ? 0 : bm.start[i];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const double dj = __shfl(di, gbase + j, 64);
        const bool same = sort || (j < nb && bm.blk[j] == bi);
        if (j < nb && j != i && same && (dj < di || (dj == di && j < i))) ++rank;
    }
    if (live && i < nb) w_out[static_cast<size_t>(mat) * nb + rank] = di;
    cd* Vk = V_out + static_cast<size_t>(live ? mat : 0) * nb2;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const int rj = __shfl(rank, gbase + j, 64);
        if (live && i < nb && j < nb) Vk[i * nb + rj] = V[j];
    }
}

}  // namespace

void launch_eigh_blocks(int nmat, int nb, int in_mode, const cd* M, double* w, cd* V,
                        const BlockMap& bm, bool sort, hipStream_t st) {
    if (nb < 1 || nb > kMaxNb)
        throw std::invalid_argument("GPU eigh supports 1 <= nb <= 32");
    if (nmat == 0) return;
    const int so = sort ? 1 : 0;
    if (nb <= 8) {
        const int np = nb <= 2 ? 2 : (nb <= 4 ? 4 : 8);
        const int per_block = 4 * (64 / np);
        const dim3 grid(cdiv(nmat, per_block));
        if (np == 2)
            hipLaunchKernelGGL(k_eigh_rows<2>, grid, dim3(256), 0, st, nmat, nb, in_mode, M, w, V, bm, so);
        else if (np == 4)
            hipLaunchKernelGGL(k_eigh_rows<4>, grid, dim3(256), 0, st, nmat, nb, in_mode, M, w, V, bm, so);
        else
            hipLaunchKernelGGL(k_eigh_rows<8>, grid, dim3(256), 0, st, nmat, nb, in_mode, M, w, V, bm, so);
        HFGPU_LAUNCH_CHECK();
        return;
    }
    const int nb2 = nb * nb;
    const size_t nbs = static_cast<size_t>(nb) + 1;
    const size_t lds = sizeof(cd) * (2 * static_cast<size_t>(nb2) + nbs) +
                       2 * sizeof(double) * nbs + 2 * sizeof(int) * nbs;
    if (nb2 <= 256)
        hipLaunchKernelGGL(k_eigh_jacobi<4>, dim3(nmat), dim3(64), lds, st, nmat, nb, in_mode, M, w, V, bm, so);
    else
        hipLaunchKernelGGL(k_eigh_jacobi<16>, dim3(nmat), dim3(64), lds, st, nmat, nb, in_mode, M, w, V, bm, so);
    HFGPU_LAUNCH_CHECK();
}

void launch_eigh(int nmat, int nb, int in_mode, const cd* M, double* w, cd* V, hipStream_t st) {
    BlockMap bm{};  // one block
    launch_eigh_blocks(nmat, nb, in_mode, M, w, V, bm, true, st);
}

namespace {

__global__ void k_max_offblock(int nmat, int nb, const cd* __restrict__ M, BlockMap bm,
                               double* partial) {
    __shared__ double red[4];
    const int nb2 = nb * nb;
    const size_t n = static_cast<size_t>(nmat) * nb2;
    double mx = 0.0;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int e = static_cast<int>(t % nb2);
        const int i = e / nb, j = e - i * nb;
        if (bm.blk[i] != bm.blk[j] || bm.blk[i] < 0) mx = fmax(mx, sqrt(cnorm(M[t])));
    }
    mx = block_max<256>(mx, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = mx;
}

__global__ void k_max_final(const double* __restrict__ partial, int n, double* out) {
    __shared__ double red[4];
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) mx = fmax(mx, partial[i]);
    mx = block_max<256>(mx, red);
    if (threadIdx.x == 0) *out = mx;
}

}  // namespace

void launch_max_offblock(int nmat, int nb, const cd* M, const BlockMap& bm, double* partial,
                         double* out, hipStream_t st) {
    const size_t n = static_cast<size_t>(nmat) * nb * nb;
    const unsigned blocks = std::max(1u, std::min<unsigned>(cdiv(n, 256), 1024u));
    hipLaunchKernelGGL(k_max_offblock, dim3(blocks), dim3(256), 0, st, nmat, nb, M, bm, partial);
    HFGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, st, partial, static_cast<int>(blocks),
                       out);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
