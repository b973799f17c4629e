// Batched Hermitian eigensolver (reference: eigh_one / eigh_batched,
// linalg.hpp:39-66, and the Cayley setup eigh of herm(i d),
// solver_dm.hpp:206-222).  Eigenvalues are returned in ascending, stable
// order with the eigenvectors as columns.
//
// nb <= 16: Householder tridiagonalisation + implicit QR (k_eigh_tri), nb
// rounded up to NP = 2, 4, 8 or 16 lanes per matrix, one row per lane.  At
// 9216 matrices of 8 x 8 it takes 72 us against 168 us for a register-row
// parallel Jacobi: the QR phase needs ~17 shifted steps of at most 7
// rotations, cyclic Jacobi ~56 serial rounds of full-row updates.
// At exactly 16 bands the eigenvector rows also sit in LDS (one wavefront
// per workgroup): 256 VGPRs and two wavefronts per SIMD instead of 256 +
// 102 AGPRs and one, 1701 -> 1473 us at 36864 matrices and 548 -> 505 us at
// 9216, bit-identical (the same arithmetic); at 12 bands the same layout was
// slower (239 -> 288 us at 9216).
// nb 17..32: the same kernel with NP = 32 and the eigenvector rows in LDS
// (registers for both A and Z would spill): at 576 matrices of 26 x 26 it
// takes 542 us against 1281 us for the parallel cyclic Jacobi below, at
// 9216 of 32 x 32 4.3 ms against 19.1 ms.
// HFGPU_EIGH_JACOBI=1 selects the Jacobi kernel (k_eigh_jacobi), one
// wavefront per matrix: each round applies the nb/2 disjoint complex
// rotations of the circle ordering at once, every lane updating its own
// elements A'_ij = sum conj(J_ki) A_kl J_lj; sweeps stop once the
// off-diagonal mass falls to (4 nb eps)^2 ||A||_F^2.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <utility>

#include "hfgpu_core.h"
#include "kernels.h"

namespace hfgpu {

namespace {

constexpr int kMaxNb = 32;
constexpr int kMaxSweeps = 40;

template <int MAXE>
__global__ void __launch_bounds__(64)
k_eigh_jacobi(int nmat, int nb, int in_mode, const cd* __restrict__ M, double* __restrict__ w_out,
              cd* __restrict__ V_out, BlockMap bm, int sort, const double* __restrict__ cond) {
    // the block structure applies unless cond = (off-block max, max) says the
    // batch is not block-diagonal (solver_scf.hpp:111-120): then a dense eigh
    const bool ub = !cond || cond[0] <= 1.0e-10 + 1.0e-10 * cond[1];
    if (!ub) sort = 1;
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
        if (ub && bm.blk[i] != bm.blk[j]) h = cd{0.0, 0.0};  // off-block: ignored
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
// Householder + implicit QR variant for nb <= 16: NP (nb rounded up to 2, 4,
// 8 or 16) lanes per matrix, lane i holding row i of A and of the eigenvector
// accumulator Z.
//
//  1. Tridiagonalisation by Hermitian reflectors H = I - beta u u^H, column k
//     at a time: x = A[k+1:, k] = x0 e1 + x', gamma = -e^{i arg x0} |x|,
//     u = x - gamma e1, beta = 2 / |u|^2 = 1 / (|x| (|x| + |x0|)).  With
//     p = beta A u, K = beta u^H p / 2 and w = p - K u, H A H = A - u w^H - w u^H
//     (each lane updates its own row from the broadcast u and w), and Z <- Z H.
//  2. The complex sub-diagonal t_j = T[j+1][j] becomes real through the
//     diagonal phases phi_0 = 1, phi_{j+1} = phi_j t_j / |t_j|; Z <- Z diag(phi).
//  3. Implicit symmetric QR steps with the Wilkinson shift of the trailing
//     2 x 2 (Golub and Van Loan, Matrix Computations, algorithm 8.3.2) on the
//     bottom unreduced block [l, m], a sub-diagonal entry counting as zero once
//     |e_k| <= eps (|d_k| + |d_{k+1}|).  Every lane of a matrix runs the same
//     scalar recurrence and applies each Givens rotation to its own row of Z.
//
// All loops run over compile-time ranges with predicates: no register array
// is indexed with a run-time value.  A block-diagonal input stays block
// diagonal (a column whose part below the diagonal is zero gets no reflector,
// a zero sub-diagonal entry splits the QR problem), so eigenpair j stays in
// the block of index j.
// Lanes of one wavefront execute LDS operations in order, so a lane may read
// what another lane of its wavefront wrote earlier without a barrier; the
// wave barrier only keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

template <int NP>
struct TriLds {
    cd u[64];
    cd w[64];
    cd t[64];
    double d[64];
};

// Row i of the eigenvector accumulator Z: registers, or (NP = 32, where
// registers for both A and Z would spill) a row of a padded LDS array that
// only its own lane touches.
template <int NL, bool IN_LDS>
struct ZRow {
    cd v[NL];
    __device__ __forceinline__ cd& operator[](int j) { return v[j]; }
};
template <int NL>
struct ZRow<NL, true> {
    cd* p;
    __device__ __forceinline__ cd& operator[](int j) { return p[j]; }
};

template <int NP>
__device__ __forceinline__ cd group_bcast(cd v, int src) {
    return cd{__shfl(v.x, src, 64), __shfl(v.y, src, 64)};
}

// Two wavefronts per SIMD for nb = 12 (its register rows need 279 registers
// at one, a few spills outside the QR loop at two): 9216 matrices 309 -> 240
// us, 14400 418 -> 315 us on MI355X.  Three (8 and 12 bands) and two for the
// 16-lane generic rows (heavy spills) measured slower.
template <int NP, int NBC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NBC == 12 ? 2 : 1)))
k_eigh_tri(int nmat, int nb, int in_mode, const cd* __restrict__ M, double* __restrict__ w_out,
           cd* __restrict__ V_out, BlockMap bm, int sort, const double* __restrict__ cond,
           int mpw, const int* __restrict__ seg, unsigned active) {
    // the block structure applies unless cond = (off-block max, max) says the
    // batch is not block-diagonal (solver_scf.hpp:111-120): then a dense eigh
    const bool ub = !cond || cond[0] <= 1.0e-10 + 1.0e-10 * cond[1];
    if (!ub) sort = 1;
    // lane groups of NP lanes: 64 / NP matrices per full wavefront (mpw used)
    // matrix dimension of the register rows: NBC when the launch fixes nb at
    // compile time (lanes NBC..NP-1 of a group are then padding), else NP
    constexpr int NL = NBC > 0 ? NBC : NP;
    constexpr double kEps = 2.220446049250313e-16;
    // one TriLds per wavefront of the workgroup: 4 (256 threads), or 1 for
    // the launches one wavefront wide whose Z rows take the dynamic LDS
    // (NP = 32, and nb = 16 exactly)
    constexpr bool kZLds = NP == 32 || NBC == 16;
    __shared__ TriLds<NP> lds[kZLds ? 1 : 4];
    const int lane = threadIdx.x & 63;
    TriLds<NP>& X = lds[threadIdx.x >> 6];
    const int i = lane % NP;
    const int gbase = lane - i;
    // mpw <= G matrices per wavefront: fewer than G leaves lane groups idle
    // but puts more wavefronts on each SIMD to hide the latency of the
    // serial QR recurrence
    int mat = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * mpw + lane / NP;
    bool live = lane / NP < mpw && mat < nmat;
    if (seg) {  // solve blockIdx.y of a batch: its matrices seg[s] .. seg[s + 1] - 1
        const int sv = blockIdx.y;
        live = ((active >> sv) & 1u) && lane / NP < mpw && mat < seg[sv + 1] - seg[sv];
        mat += seg[sv];
        // the launch is sized for the largest solve: a wavefront with no
        // matrix of its solve leaves at once (no workgroup barrier below,
        // and the lane-group shuffles stay within the wavefront)
        if (!__any(live)) return;
    }
    const int nb2 = nb * nb;
    const cd* Mk = M + static_cast<size_t>(live ? mat : 0) * nb2;
    cd A[NL];
    ZRow<NL, kZLds> Z;
    if constexpr (kZLds) {
        // dynamic LDS: G matrices per wavefront, NP rows of NP + 1 each
        extern __shared__ cd zlds[];
        Z.p = zlds + ((threadIdx.x >> 6) * 64 + lane) * (NP + 1);
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        cd h{0.0, 0.0};
        if (live && i < nb && j < nb) {
            cd xij = Mk[i * nb + j], xji = Mk[j * nb + i];
            if (in_mode == 1) {
                xij = cd{-xij.y, xij.x};
                xji = cd{-xji.y, xji.x};
            }
            h = cd{0.5 * (xij.x + xji.x), 0.5 * (xij.y - xji.y)};
            if (ub && bm.blk[i] != bm.blk[j]) h = cd{0.0, 0.0};  // off-block: ignored
        }
        A[j] = h;
        Z[j] = cd{j == i ? 1.0 : 0.0, 0.0};
    }

    // ---- 1. tridiagonalisation (a run-time loop over the columns, so only
    // one column's temporaries are live; column k of the row is picked with
    // compile-time selects)
#pragma unroll 1
    for (int k = 0; k < nb - 2; ++k) {
        cd ak{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < NL; ++j)
            if (j == k) ak = A[j];
        const cd xi = (i >= k + 1 && i < nb) ? ak : cd{0.0, 0.0};
        double part = i >= k + 2 ? cnorm(xi) : 0.0;
#pragma unroll
        for (int off = NP / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        const double xn2 = part;
        const cd x0 = group_bcast<NP>(xi, gbase + k + 1);
        const double rho2 = cnorm(x0);
        const double irho = rho2 > 0.0 ? rsqrt(rho2) : 0.0;
        const double rho = rho2 * irho;
        const bool act = xn2 > 0.0;  // x = x0 e1 already: no reflector
        const double sigma = sqrt(rho2 + xn2);
        const cd ph = rho2 > 0.0 ? cd{x0.x * irho, x0.y * irho} : cd{1.0, 0.0};
        const double beta = act ? 1.0 / (sigma * (sigma + rho)) : 0.0;
        const cd ui = i == k + 1 ? cd{ph.x * (rho + sigma), ph.y * (rho + sigma)}
                                 : (i >= k + 2 ? xi : cd{0.0, 0.0});
        X.u[lane] = ui;
        wave_sync();
        cd p{0.0, 0.0}, sz{0.0, 0.0};
#pragma unroll
        for (int j = 1; j < NL; ++j) {
            if (j <= k) continue;
            const cd uj = X.u[gbase + j];
            p = cadd(p, cmul(A[j], uj));
            sz = cadd(sz, cmul(Z[j], uj));
        }
        p = cscale(p, beta);
        double kk = ui.x * p.x + ui.y * p.y;  // Re(conj(u_i) p_i)
#pragma unroll
        for (int off = NP / 2; off > 0; off >>= 1) kk += __shfl_xor(kk, off, 64);
        const double K = 0.5 * beta * kk;
        const cd wi = csub(p, cscale(ui, K));
        X.w[lane] = wi;
        wave_sync();
        const cd bsz = cscale(sz, beta);
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const cd uj = j > k ? X.u[gbase + j] : cd{0.0, 0.0};
            const cd wj = X.w[gbase + j];
            // A_ij -= u_i conj(w_j) + w_i conj(u_j)
            A[j] = csub(A[j], cadd(cmul(ui, cconj(wj)), cmul(wi, cconj(uj))));
            // Z_ij -= beta (Z u)_i conj(u_j)
            Z[j] = csub(Z[j], cmul(bsz, cconj(uj)));
        }
        wave_sync();
        if (act) {
            // exact values of column and row k after the reflection
            const cd gamma{-ph.x * sigma, -ph.y * sigma};
#pragma unroll
            for (int j = 0; j < NL; ++j) {
                if (j == k) {
                    if (i == k + 1) A[j] = gamma;
                    else if (i >= k + 2) A[j] = cd{0.0, 0.0};
                }
                if (i == k && j > k) A[j] = j == k + 1 ? cconj(gamma) : cd{0.0, 0.0};
            }
        }
    }

    // ---- 2. real tridiagonal: d_j, |t_j|, phases
    double dmine = 0.0;
    cd tmine{0.0, 0.0};  // T[i][i-1]
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        if (j == i) dmine = A[j].x;
        if (j + 1 == i) tmine = A[j];
    }
    X.d[lane] = dmine;
    X.t[lane] = tmine;
    wave_sync();
    double d[NL], e[NL];
    cd ph{1.0, 0.0};
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        d[j] = X.d[gbase + j];
        Z[j] = cmul(Z[j], ph);
        if (j + 1 < NL) {
            const cd t = X.t[gbase + j + 1];
            const double a2 = cnorm(t);
            const double ia = a2 > 0.0 ? rsqrt(a2) : 0.0;
            e[j] = j + 1 < nb ? a2 * ia : 0.0;
            if (a2 > 0.0) ph = cmul(ph, cd{t.x * ia, t.y * ia});
        } else {
            e[j] = 0.0;
        }
    }
    wave_sync();

    // ---- 3. implicit QR with Wilkinson shift
    for (int it = 0; it < 30 * NL; ++it) {
#pragma unroll
        for (int k = 0; k + 1 < NL; ++k)
            if (fabs(e[k]) <= kEps * (fabs(d[k]) + fabs(d[k + 1]))) e[k] = 0.0;
        int m = -1;
#pragma unroll
        for (int k = 0; k + 1 < NL; ++k)
            if (e[k] != 0.0) m = k + 1;
        if (__all(m < 0)) break;
        int l = m;
#pragma unroll
        for (int k = NL - 2; k >= 0; --k)
            if (k == l - 1 && e[k] != 0.0) l = k;
        double am = 0.0, am1 = 0.0, bm1 = 0.0, x = 0.0, z = 0.0;
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            if (k == m) am = d[k];
            if (k + 1 == m) {
                am1 = d[k];
                bm1 = e[k];
            }
            if (k == l) {
                x = d[k];
                z = e[k];
            }
        }
        const double dd = 0.5 * (am1 - am);
        const double den = dd + (dd >= 0.0 ? 1.0 : -1.0) * sqrt(dd * dd + bm1 * bm1);
        const double mu = m >= 0 ? am - bm1 * (bm1 / den) : 0.0;
        x -= mu;
#pragma unroll
        for (int k = 0; k + 1 < NL; ++k) {
            if (k < l || k >= m) continue;
            // r = |(x, z)|, c = x / r, s = z / r from one reciprocal square
            // root (a Givens rotation only needs c^2 + s^2 = 1 to rounding)
            const double r2 = x * x + z * z;
            const double ir = r2 > 0.0 ? rsqrt(r2) : 0.0;
            const double r = r2 * ir;
            const double c = r2 > 0.0 ? x * ir : 1.0, s = z * ir;
            if (k > 0 && k > l) e[k - 1] = r;
            const double a0 = d[k], a1 = d[k + 1], b0 = e[k];
            const double cs = c * s, cc = c * c, ss = s * s;
            d[k] = cc * a0 + 2.0 * cs * b0 + ss * a1;
            d[k + 1] = ss * a0 - 2.0 * cs * b0 + cc * a1;
            e[k] = cs * (a1 - a0) + (cc - ss) * b0;
            if (k + 2 < NL && k + 1 < m) {
                const double b1 = e[k + 1];
                x = e[k];
                z = s * b1;
                e[k + 1] = c * b1;
            }
            const cd zk = Z[k], zk1 = Z[k + 1];
            Z[k] = cadd(cscale(zk, c), cscale(zk1, s));
            Z[k + 1] = csub(cscale(zk1, c), cscale(zk, s));
        }
    }

    // ---- ascending, stable order (within each block unless sort)
    double di = 0.0;
#pragma unroll
    for (int j = 0; j < NL; ++j)
        if (j == i) di = d[j];
    const int bi = i < nb ? bm.blk[i] : -1;
    int rank = (sort || i >= nb) ? 0 : bm.start[i];
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const bool same = sort || (j < nb && bm.blk[j] == bi);
        if (j < nb && j != i && same && (d[j] < di || (d[j] == di && j < i))) ++rank;
    }
    if (live && i < nb) w_out[static_cast<size_t>(mat) * nb + rank] = di;
    cd* Vk = V_out + static_cast<size_t>(live ? mat : 0) * nb2;
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const int rj = __shfl(rank, gbase + j, 64);
        if (live && i < nb && j < nb) Vk[i * nb + rj] = Z[j];
    }
}

}  // namespace

// Matrices per wavefront of the register kernels: a full wavefront (g) unless
// the autotuner below kept another count for this (nb, nmat).  Fewer leave
// lane groups idle to put more wavefronts on each SIMD; that measured slower
// (96 x 96 x 8: 72 us at 8 per wavefront, 99 us at 4, 151 us at 2): the
// kernel is bound by VALU issue, not by the latency of its serial
// recurrences.  HFGPU_EIGH_MPW overrides both.  The count only decides which
// lane groups hold a matrix, so results do not depend on it.
namespace {
std::mutex g_mpw_mu;
std::map<std::pair<int, int>, int> g_mpw;  // (nb, nmat) -> matrices per wavefront
}  // namespace

static int pow2_at_most(int want, int g) {
    int p = 1;
    while (p < want && p < g) p <<= 1;
    return std::min(p, g);
}

static int eigh_mpw(int nmat, int nb, int g) {
    static const int env = [] {
        const char* e = std::getenv("HFGPU_EIGH_MPW");
        return e ? std::atoi(e) : 0;
    }();
    if (env > 0) return pow2_at_most(env, g);
    std::lock_guard<std::mutex> lock(g_mpw_mu);
    const auto it = g_mpw.find({nb, nmat});
    return it == g_mpw.end() ? g : pow2_at_most(it->second, g);
}

// The eigh launches: nmat matrices, or with seg the S solves' matrices
// seg[s] .. seg[s + 1] - 1 (nmat the largest count; solves whose bit in
// active is clear are skipped).
static void eigh_launch(int nmat, int nb, int in_mode, const cd* M, double* w, cd* V,
                        const BlockMap& bm, bool sort, hipStream_t st, const double* cond,
                        const int* seg, int S, unsigned active, int mpw_force = 0) {
    if (nb < 1 || nb > kMaxNb)
        throw std::invalid_argument("GPU eigh supports 1 <= nb <= 32");
    if (nmat == 0) return;
    const int so = sort ? 1 : 0;
    const unsigned gy = seg ? static_cast<unsigned>(S) : 1u;
    if (nb <= 16) {
        const int np = nb <= 2 ? 2 : (nb <= 4 ? 4 : (nb <= 8 ? 8 : 16));
        const int mpw = mpw_force > 0 ? pow2_at_most(mpw_force, 64 / np) : eigh_mpw(nmat, nb, 64 / np);
        const int per_block = 4 * mpw;
        const dim3 grid(cdiv(nmat, per_block), gy);
        if (np == 2)
            hipLaunchKernelGGL((k_eigh_tri<2, 0>), grid, dim3(256), 0, st, nmat, nb, in_mode, M, w, V, bm, so, cond, mpw, seg, active);
        else if (np == 4)
            hipLaunchKernelGGL((k_eigh_tri<4, 0>), grid, dim3(256), 0, st, nmat, nb, in_mode, M, w, V, bm, so, cond, mpw, seg, active);
        else if (np == 8)
            hipLaunchKernelGGL((k_eigh_tri<8, 0>), grid, dim3(256), 0, st, nmat, nb, in_mode, M, w, V, bm, so, cond, mpw, seg, active);
        else if (nb == 12)  // the 12-band models: loops sized to 12, not 16
            hipLaunchKernelGGL((k_eigh_tri<16, 12>), grid, dim3(256), 0, st, nmat, nb, in_mode, M, w, V, bm, so, cond, mpw, seg, active);
        else if (nb == 16)  // one wavefront per workgroup, Z rows in LDS (16 rows of 17)
            hipLaunchKernelGGL((k_eigh_tri<16, 16>), dim3(cdiv(nmat, mpw), gy), dim3(64), sizeof(cd) * 64 * 17, st, nmat, nb, in_mode, M, w, V, bm, so, cond, mpw, seg, active);
        else
            hipLaunchKernelGGL((k_eigh_tri<16, 0>), grid, dim3(256), 0, st, nmat, nb, in_mode, M, w, V, bm, so, cond, mpw, seg, active);
        HFGPU_LAUNCH_CHECK();
        return;
    }
    static const bool use_jacobi = [] {
        const char* e = std::getenv("HFGPU_EIGH_JACOBI");
        return e && e[0] == '1';
    }();
    if (!use_jacobi) {
        // 32 lanes per matrix, two matrices per wavefront, one wavefront per
        // workgroup: Z rows in LDS (64 rows of 33 complex)
        const dim3 grid(cdiv(nmat, 2), gy);
        const size_t zl = sizeof(cd) * 64 * 33;
        hipLaunchKernelGGL((k_eigh_tri<32, 0>), grid, dim3(64), zl, st, nmat, nb, in_mode, M, w,
                           V, bm, so, cond, 2, seg, active);
        HFGPU_LAUNCH_CHECK();
        return;
    }
    if (seg) throw std::logic_error("the Jacobi eigh takes no solve segments");
    const int nb2 = nb * nb;
    const size_t nbs = static_cast<size_t>(nb) + 1;
    const size_t lds = sizeof(cd) * (2 * static_cast<size_t>(nb2) + nbs) +
                       2 * sizeof(double) * nbs + 2 * sizeof(int) * nbs;
    if (nb2 <= 256)
        hipLaunchKernelGGL(k_eigh_jacobi<4>, dim3(nmat), dim3(64), lds, st, nmat, nb, in_mode, M, w, V, bm, so, cond);
    else
        hipLaunchKernelGGL(k_eigh_jacobi<16>, dim3(nmat), dim3(64), lds, st, nmat, nb, in_mode, M, w, V, bm, so, cond);
    HFGPU_LAUNCH_CHECK();
}

void launch_eigh_blocks(int nmat, int nb, int in_mode, const cd* M, double* w, cd* V,
                        const BlockMap& bm, bool sort, hipStream_t st, const double* cond) {
    eigh_launch(nmat, nb, in_mode, M, w, V, bm, sort, st, cond, nullptr, 1, ~0u);
}

void launch_eigh(int nmat, int nb, int in_mode, const cd* M, double* w, cd* V, hipStream_t st) {
    BlockMap bm{};  // one block
    launch_eigh_blocks(nmat, nb, in_mode, M, w, V, bm, true, st);
}

void launch_eigh_seg(const int* seg, int S, int nk_max, unsigned active, int nb, int in_mode,
                     const cd* M, double* w, cd* V, hipStream_t st) {
    BlockMap bm{};
    eigh_launch(nk_max, nb, in_mode, M, w, V, bm, true, st, nullptr, seg, S, active);
}

namespace {

// Hermitian test matrices for the autotuner: entry (a, b), a <= b, of matrix
// k from an integer hash (uniform in [-1, 1)), the lower triangle its
// conjugate, a real diagonal.
__device__ __forceinline__ double hash_unit(unsigned long long x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return static_cast<double>(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__global__ void k_fill_hermitian(size_t n, int nb, cd* __restrict__ M) {
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t nb2 = static_cast<size_t>(nb) * nb;
    const size_t k = t / nb2;
    const int e = static_cast<int>(t % nb2), i = e / nb, j = e % nb;
    const int a = min(i, j), b = max(i, j);
    const unsigned long long key = (k * 64 + a) * 64 + b;
    const double re = hash_unit(2 * key), im = a == b ? 0.0 : hash_unit(2 * key + 1);
    M[t] = cd{re, i <= j ? im : -im};
}

}  // namespace

double eigh_autotune(int nmat, int nb, hipStream_t st, int* mpw_out) {
    if (nb < 1 || nb > kMaxNb) throw std::invalid_argument("GPU eigh supports 1 <= nb <= 32");
    if (nmat < 1) throw std::invalid_argument("eigh autotune needs at least one matrix");
    if (nb > 16) {  // the 17..32-band kernel: two matrices per wavefront, fixed
        *mpw_out = 2;
        return 0.0;
    }
    const int np = nb <= 2 ? 2 : (nb <= 4 ? 4 : (nb <= 8 ? 8 : 16));
    const int g = 64 / np;
    const size_t nd = static_cast<size_t>(nmat) * nb * nb;
    DevBuf M(nd * sizeof(cd)), V(nd * sizeof(cd)),
        w(static_cast<size_t>(nmat) * nb * sizeof(double));
    hipLaunchKernelGGL(k_fill_hermitian, dim3(static_cast<unsigned>(cdiv(nd, 256))), dim3(256), 0,
                       st, nd, nb, M.as<cd>());
    HFGPU_LAUNCH_CHECK();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HFGPU_CHECK(hipEventCreate(&e0));
    HFGPU_CHECK(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        hipStream_t s;
        ~EventGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{e0, e1, st};
    BlockMap bm{};
    int reps = 3;
    auto time_mpw = [&](int m) {
        for (int i = 0; i < 2; ++i)
            eigh_launch(nmat, nb, 0, M.as<cd>(), w.as<double>(), V.as<cd>(), bm, true, st,
                        nullptr, nullptr, 1, ~0u, m);
        float best = 1e30f;
        for (int t = 0; t < 3; ++t) {
            HFGPU_CHECK(hipEventRecord(e0, st));
            for (int i = 0; i < reps; ++i)
                eigh_launch(nmat, nb, 0, M.as<cd>(), w.as<double>(), V.as<cd>(), bm, true, st,
                            nullptr, nullptr, 1, ~0u, m);
            HFGPU_CHECK(hipEventRecord(e1, st));
            HFGPU_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HFGPU_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return 1e3 * static_cast<double>(best) / reps;
    };
    int best_m = g;
    double best_us = time_mpw(g);
    // about 1 ms of timed calls per candidate
    reps = std::max(3, std::min(30, static_cast<int>(1000.0 / std::max(best_us, 1.0))));
    best_us = time_mpw(g);
    for (int m = g / 2; m >= 1 && m >= g / 8; m /= 2) {
        const double us = time_mpw(m);
        if (us < best_us) {
            best_us = us;
            best_m = m;
        }
    }
    {
        std::lock_guard<std::mutex> lock(g_mpw_mu);
        g_mpw[{nb, nmat}] = best_m;
    }
    *mpw_out = best_m;
    return best_us;
}

void eigh_autotune_once(int nmat, int nb, hipStream_t st) {
    if (nmat < 1 || nb < 1 || nb > 16) return;
    {
        std::lock_guard<std::mutex> lock(g_mpw_mu);
        if (g_mpw.count({nb, nmat})) return;
    }
    int m = 0;
    (void)eigh_autotune(nmat, nb, st, &m);
}

namespace {

// Per-workgroup maxima of |M| over the off-block entries and over all
// entries (the scale of the block-structure test).
__global__ void k_max_offblock(int nmat, int nb, const cd* __restrict__ M, BlockMap bm,
                               double* partial) {
    __shared__ double red[4];
    const int nb2 = nb * nb;
    const size_t n = static_cast<size_t>(nmat) * nb2;
    double mx = 0.0, ma = 0.0;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int e = static_cast<int>(t % nb2);
        const int i = e / nb, j = e - i * nb;
        const double a = sqrt(cnorm(M[t]));
        if (bm.blk[i] != bm.blk[j] || bm.blk[i] < 0) mx = fmax(mx, a);
        ma = fmax(ma, a);
    }
    mx = block_max<256>(mx, red);
    ma = block_max<256>(ma, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = mx;
        partial[2 * blockIdx.x + 1] = ma;
    }
}

__global__ void k_max_final(const double* __restrict__ partial, int n, double* out) {
    __shared__ double red[4];
    double mx = 0.0, ma = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        mx = fmax(mx, partial[2 * i]);
        ma = fmax(ma, partial[2 * i + 1]);
    }
    mx = block_max<256>(mx, red);
    ma = block_max<256>(ma, red);
    if (threadIdx.x == 0) {
        out[0] = mx;
        out[1] = ma;
    }
}

}  // namespace

void launch_max_offblock(int nmat, int nb, const cd* M, const BlockMap& bm, double* partial,
                         double* out, hipStream_t st) {
    const size_t n = static_cast<size_t>(nmat) * nb * nb;
    const unsigned blocks =
        std::max(1u, std::min<unsigned>(cdiv(n, 256), kMaxOffblockBlocks));
    hipLaunchKernelGGL(k_max_offblock, dim3(blocks), dim3(256), 0, st, nmat, nb, M, bm, partial);
    HFGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, st, partial, static_cast<int>(blocks),
                       out);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
