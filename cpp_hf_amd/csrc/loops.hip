// Fused per-k kernels of the two solver loops.
//
// Direct minimisation (solver_dm.hpp:474-613): each iteration after the Fock
// build is five passes over the k points:
//   orbital_grad  Ft = Q^H F Q, eps, G, g_p          (+ |G|_w^2, sum eps^2)
//   precond       d_Q = G / sqrt(gap^2 + lam^2), h_p (+ the six PR+ products)
//   cg_combine    d += beta d_prev                   (+ |d_Q|_w^2, Omega_0)
//   [eigh of herm(i d_Q)]
//   linesearch    Ft_eig = V^H Ft V, diag(U^H Ft U) at every backtracking
//                 step tau_j at once, frozen free energies (+ pick tau)
//   retract       Q <- Q V c(tau) V^H, eps_work = diag(U^H Ft U)
// The SCF loop (solver_scf.hpp:220-402) adds the residual pass
// (delta = P_new - P, [F, P], both weighted norms).
//
// One wavefront per k point with the k point's nb x nb blocks in LDS; W
// wavefronts per workgroup, grid-stride over k.  Scalar results go through
// per-workgroup partials and a one-workgroup finalising launch that sums them
// in workgroup order into the solver's device scalar array, so no host round
// trip is needed inside an iteration, every sum is deterministic for a fixed
// launch shape, and no workgroup needs a device-scope fence.
#include <cmath>

#include <cstdlib>

#include "loop_common.h"

namespace hfgpu {

namespace {

using namespace loopk;

// Finalising launch: out[s] = sum over workgroups of partial[b * ns + s].
// Thread t sums the rows b = t, t + 256, ... of up to kFinSlots adjacent
// slots at once (every load of a pass in flight together), then a fixed
// wave tree per slot and the four wavefronts in a fixed order.  MODE 1 also
// selects the line-search step from the trial free energies (the reference's
// sequential acceptance test, solver_dm.hpp:561-589).
template <int MODE>
__global__ void __launch_bounds__(256)
k_finalize(const double* __restrict__ partial, int nparts, int ns, double* __restrict__ out,
           double* sc, int nbt, double shrink, double T_r, double max_step, int j_lo, int j_hi,
           int final_stage, int gate, unsigned active, const int* seg, int wv) {
    // one workgroup per solve of a batch (its partials, slots and outputs
    // one per-solve stride apart); with seg, each solve's partials are those
    // of its own grid (wv wavefronts per workgroup of the producing launch)
    const int sv = blockIdx.x;
    if (!solve_active(active, sv)) return;
    if (seg) nparts = max(1, min((seg[sv + 1] - seg[sv] + wv - 1) / wv, 1024));
    partial += sv * kLoopPartialDoubles;
    out += sv * SC_COUNT;
    if (sc) sc += sv * SC_COUNT;
    // MODE 1 (line search): test trials [j_lo, j_hi) in order (solver_dm.hpp:
    // 575-587); the last stage applies the no-acceptance rule; gate: a
    // trial accepted by an earlier stage stands
    if (MODE == 1 && gate && sc[SC_ACCEPTED] > 0.5) return;
    constexpr int kFinSlots = 8;
    __shared__ double s_sum[2 * kMaxBt];
    __shared__ double s_red[4][kFinSlots];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // MODE 1 needs only the slots of the trials it tests
    const int s_lo = MODE == 1 ? 2 * j_lo : 0, s_hi = MODE == 1 ? 2 * j_hi : ns;
    for (int s0 = s_lo; s0 < s_hi; s0 += kFinSlots) {
        const int nsl = min(kFinSlots, s_hi - s0);  // uniform over the workgroup
        double acc[kFinSlots];
#pragma unroll
        for (int j = 0; j < kFinSlots; ++j) acc[j] = 0.0;
#pragma unroll 4
        for (int b = threadIdx.x; b < nparts; b += 256) {
            const double* row = partial + static_cast<size_t>(b) * ns + s0;
#pragma unroll
            for (int j = 0; j < kFinSlots; ++j)
                if (j < nsl) acc[j] += row[j];
        }
#pragma unroll
        for (int j = 0; j < kFinSlots; ++j) {
            if (j < nsl) {
                const double t = wave_sum(acc[j]);
                if (lane == 0) s_red[wave][j] = t;
            }
        }
        __syncthreads();
        if (threadIdx.x < nsl) {
            const int j = threadIdx.x;
            const double t = (s_red[0][j] + s_red[1][j]) + (s_red[2][j] + s_red[3][j]);
            out[s0 + j] = t;
            if (MODE == 1) s_sum[s0 + j] = t;
        }
        __syncthreads();
    }
    if (MODE != 1) return;
    if (threadIdx.x != 0) return;
    const double dqn = sqrt(fmax(0.0, sc[SC_DQN2]));
    const double tau0 = fmin(1.0, max_step / fmax(dqn, TINY_REAL));
    const double Tv = fmax(T_r, 1.0e-14);
    const double om0 = sc[SC_E0] - Tv * sc[SC_S0];
    double t = tau0;
    for (int j = 0; j < j_lo; ++j) t *= shrink;
    bool ok = false;
    for (int j = j_lo; j < j_hi; ++j) {
        const double om = s_sum[2 * j] - Tv * s_sum[2 * j + 1];
        if (om < om0) {
            ok = true;
            break;
        }
        t *= shrink;
    }
    if (!ok && final_stage) t = tau0 * pow(shrink, static_cast<double>(nbt));
    sc[SC_TAU0] = tau0;
    sc[SC_OMEGA0] = om0;
    if (ok || final_stage) sc[SC_TAU] = t;
    sc[SC_ACCEPTED] = ok ? 1.0 : 0.0;
}

// ---------------------------------------------------------------------------
// DM kernels
// ---------------------------------------------------------------------------

// Ft = Q^H F Q (written over F), eps = Re diag Ft, G, g_p.
// Slots: [GN2, EPS_SQ].  grad=false: only the rotation (finalise step).
template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_orbital_grad(int nk, int nb_arg, const cd* __restrict__ Q, cd* __restrict__ F,
               const double* __restrict__ p, const double* __restrict__ w, const MuState* mus,
               double T_r, int grad, double* __restrict__ eps, cd* __restrict__ G,
               double* __restrict__ g_p, double* partial, double* sc, unsigned active,
               const int* seg, int mfma) {
    const int nb = NB > 0 ? NB : nb_arg;  // compile-time for the common sizes
    extern __shared__ cd smem[];
    __shared__ double red[W * 2];
    const int nb2 = nb * nb;
    const SolveSeg sg = solve_seg<W>(seg, nk);
    const int gx = sg.gx;
    {
        const int sv = blockIdx.y;
        if (!solve_active(active, sv) || static_cast<int>(blockIdx.x) >= gx) return;
        nk = sg.nk;
        const size_t od = sg.ko * nb2, ob = sg.ko * nb;
        Q += od;
        F += od;
        if (p) p += ob;
        w += sg.ko;
        mus += sv;
        if (eps) eps += ob;
        if (G) G += od;
        if (g_p) g_p += ob;
        partial += sv * kLoopPartialDoubles;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    cd* sQ = smem + wave * 3 * nb2;
    cd* sF = sQ + nb2;
    cd* sT = sF + nb2;
    const double mu = grad ? mus->mu : 0.0;
    double acc[2] = {0.0, 0.0};
    const int nsteps = (nk + gx * W - 1) / (gx * W);
    for (int step = 0; step < nsteps; ++step) {
        const int k = (step * gx + blockIdx.x) * W + wave;
        const bool act = k < nk;
        const size_t kb = static_cast<size_t>(k) * nb2;
        if (act)
            for (int e = lane; e < nb2; e += 64) {
                sQ[e] = Q[kb + e];
                sF[e] = F[kb + e];
            }
        wave_lds_sync();
        if (act) wave_mm(sQ, sF, sT, nb, true, false, lane, mfma);
        wave_lds_sync();
        if (act) wave_mm(sT, sQ, sF, nb, false, false, lane, mfma);
        wave_lds_sync();
        if (act) {
            for (int e = lane; e < nb2; e += 64) F[kb + e] = sF[e];
            const double* pk = p + static_cast<size_t>(k) * nb;
            for (int i = lane; i < nb; i += 64) {
                const double ei = sF[i * nb + i].x;
                if (eps) eps[static_cast<size_t>(k) * nb + i] = ei;
                if (grad) {
                    const double pv = clampd(pk[i], LOGIT_CLIP, 1.0 - LOGIT_CLIP);
                    g_p[static_cast<size_t>(k) * nb + i] = ei + T_r * (log(pv) - log1p(-pv)) - mu;
                    acc[1] += ei * ei;
                }
            }
            if (grad) {
                const double wk = w[k];
                double gs = 0.0;
                for (int e = lane; e < nb2; e += 64) {
                    const int i = e / nb, j = e % nb;
                    cd g{0.0, 0.0};
                    if (i != j) {
                        const cd aij = cscale(sF[i * nb + j], pk[j] - pk[i]);
                        const cd aji = cscale(sF[j * nb + i], pk[i] - pk[j]);
                        g = cd{0.5 * (aij.x - aji.x), 0.5 * (aij.y + aji.y)};
                    }
                    G[kb + e] = g;
                    gs += cnorm(g);
                }
                acc[0] += wk * gs;
            }
        }
        wave_lds_sync();
    }
    if (!grad) return;
    store_block_partials<2, W>(acc, red, partial);
}

// d_Q = G / sqrt(gap^2 + lam^2), h_p = g_p / curv, plus the PR+ products.
// Slots: [PR_S1 .. PR_S6].
template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_precond(int nk, int nb_arg, const cd* __restrict__ G, const cd* __restrict__ Gp,
          const double* __restrict__ eps, const double* __restrict__ p,
          const double* __restrict__ g_p, const double* __restrict__ g_pp,
          const double* __restrict__ w, double T_r, double denom_scale, cd* __restrict__ dQ,
          double* __restrict__ h_p, double* partial, const double* sc, unsigned active,
          const int* seg) {
    const int nb = NB > 0 ? NB : nb_arg;  // compile-time for the common sizes
    __shared__ double red[W * 6];
    const int nb2 = nb * nb;
    const SolveSeg sg = solve_seg<W>(seg, nk);
    const int gx = sg.gx;
    {
        const int sv = blockIdx.y;
        if (!solve_active(active, sv) || static_cast<int>(blockIdx.x) >= gx) return;
        nk = sg.nk;
        const size_t od = sg.ko * nb2, ob = sg.ko * nb;
        G += od;
        Gp += od;
        eps += ob;
        p += ob;
        g_p += ob;
        g_pp += ob;
        w += sg.ko;
        dQ += od;
        h_p += ob;
        partial += sv * kLoopPartialDoubles;
        sc += sv * SC_COUNT;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t nkb = static_cast<size_t>(nk) * nb;
    const double eps_scale = sqrt(sc[SC_EPS_SQ] / static_cast<double>(nkb > 0 ? nkb : 1) + TINY_REAL);
    const double lam = fmax(T_r, denom_scale * eps_scale);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int k = blockIdx.x * W + wave; k < nk; k += gx * W) {
        const size_t kb = static_cast<size_t>(k) * nb2;
        const double* ek = eps + static_cast<size_t>(k) * nb;
        const double wk = w[k];
        double s1 = 0, s2 = 0, s5 = 0;
        for (int e = lane; e < nb2; e += 64) {
            const int i = e / nb, j = e % nb;
            const double gap = ek[i] - ek[j];
            const double base = sqrt(gap * gap + lam * lam);
            const cd g = G[kb + e];
            const cd d{g.x / base, g.y / base};
            dQ[kb + e] = d;
            const cd gp = Gp[kb + e];
            s1 += g.x * d.x + g.y * d.y;
            s2 += gp.x * d.x + gp.y * d.y;
            s5 += cnorm(gp);
        }
        double s3 = 0, s4 = 0, s6 = 0;
        for (int i = lane; i < nb; i += 64) {
            const size_t idx = static_cast<size_t>(k) * nb + i;
            const double ps = clampd(p[idx], OCC_CLIP, 1.0 - OCC_CLIP);
            const double curv = T_r / (ps * (1.0 - ps));
            const double hp = g_p[idx] / curv;
            h_p[idx] = hp;
            s3 += g_p[idx] * hp;
            s4 += g_pp[idx] * hp;
            s6 += g_pp[idx] * g_pp[idx];
        }
        acc[0] += wk * s1;
        acc[1] += wk * s2;
        acc[2] += wk * s3;
        acc[3] += wk * s4;
        acc[4] += wk * s5;
        acc[5] += wk * s6;
    }
    store_block_partials<6, W>(acc, red, partial);
}

// beta from the PR+ products; d_Q += beta d_Q_prev, d_p = h_p + beta d_p_prev.
// Slots: [DQN2, E0, S0].
template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_cg_combine(int nk, int nb_arg, int restart, cd* __restrict__ dQ, const cd* __restrict__ dQp,
             const double* __restrict__ h_p, const double* __restrict__ d_pp,
             double* __restrict__ d_p, const double* __restrict__ eps,
             const double* __restrict__ p, const double* __restrict__ w, double T_r,
             double* partial, double* sc, unsigned active, const int* seg) {
    const int nb = NB > 0 ? NB : nb_arg;  // compile-time for the common sizes
    __shared__ double red[W * 3];
    const int nb2 = nb * nb;
    const SolveSeg sg = solve_seg<W>(seg, nk);
    const int gx = sg.gx;
    {
        const int sv = blockIdx.y;
        if (!solve_active(active, sv) || static_cast<int>(blockIdx.x) >= gx) return;
        nk = sg.nk;
        const size_t od = sg.ko * nb2, ob = sg.ko * nb;
        dQ += od;
        dQp += od;
        h_p += ob;
        d_pp += ob;
        d_p += ob;
        eps += ob;
        p += ob;
        w += sg.ko;
        partial += sv * kLoopPartialDoubles;
        sc += sv * SC_COUNT;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double num = ((sc[SC_PR1] - sc[SC_PR1 + 1]) + sc[SC_PR1 + 2]) - sc[SC_PR1 + 3];
    const double den = sc[SC_PR1 + 4] + sc[SC_PR1 + 5];
    double beta = (den > TINY_REAL) ? fmax(0.0, num / den) : 0.0;
    beta = fmin(beta, 5.0);
    if (restart) beta = 0.0;
    double acc[3] = {0, 0, 0};
    for (int k = blockIdx.x * W + wave; k < nk; k += gx * W) {
        const size_t kb = static_cast<size_t>(k) * nb2;
        const double wk = w[k];
        double sq = 0.0;
        for (int e = lane; e < nb2; e += 64) {
            const cd a = dQ[kb + e], b = dQp[kb + e];
            const cd d{a.x + beta * b.x, a.y + beta * b.y};
            dQ[kb + e] = d;
            sq += cnorm(d);
        }
        double e0 = 0.0, s0 = 0.0;
        for (int i = lane; i < nb; i += 64) {
            const size_t idx = static_cast<size_t>(k) * nb + i;
            d_p[idx] = h_p[idx] + beta * d_pp[idx];
            const double pv = clampd(p[idx], OCC_CLIP, 1.0 - OCC_CLIP);
            e0 += wk * pv * eps[idx];
            s0 -= wk * (pv * log(pv) + (1.0 - pv) * log1p(-pv));
        }
        acc[0] += wk * sq;
        acc[1] += e0;
        acc[2] += s0;
    }
    store_block_partials<3, W>(acc, red, partial);
    // every workgroup computes the same beta; one of them records it
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_BETA] = beta;
}

// Per-k: Fe = V^H Ft V (stored), then for each tau: diag(U^H Ft U) via the
// spectrum, as in diag_UFU_from_spectrum.  Result for row i in out[i].
__device__ __forceinline__ void diag_at_tau(const cd* sV, const cd* sFe, const double* slam,
                                            cd* sVC, cd* sc_, double* srow, int nb, double tau,
                                            int lane, bool mfma) {
    const int nb2 = nb * nb;
    for (int n = lane; n < nb; n += 64) sc_[n] = cayley_c(slam[n], tau);
    wave_lds_sync();
    for (int e = lane; e < nb2; e += 64) {
        const int n = e % nb;
        sVC[e] = cmulc(sV[e], sc_[n]);  // V_in conj(c_n)
    }
    wave_lds_sync();
    wave_mm_each(
        sVC, sFe, nb, false, false, lane,
        [&](int e, cd a) {
            a = cmul(a, sc_[e % nb]);
            srow[e] = cmulc(a, sV[e]).x;
        },
        mfma);
    wave_lds_sync();
}

// Backtracking trials j in [j_lo, j_hi) of the spectral line search
// (solver_dm.hpp:561-589).  Slots: [2 kMaxBt trial sums (E_j, S_j)], each
// trial reduced over the wavefront as it is evaluated.  Stage 1 (stage2 = 0)
// computes Fe = V^H Ft V, writes it, and evaluates tau0; stage 2 returns at
// once when the finaliser has accepted tau0, else reloads Fe and evaluates
// the remaining trials.
template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_linesearch(int nk, int nb_arg, const cd* __restrict__ V, const double* __restrict__ lam,
             const cd* __restrict__ Ft, cd* __restrict__ Fe, const double* __restrict__ p,
             const double* __restrict__ d_p, const double* __restrict__ w, double T_r,
             double max_step, double shrink, int j_lo, int j_hi, int stage2,
             double* partial, const double* sc, unsigned active, const int* seg, int mfma) {
    const int nb = NB > 0 ? NB : nb_arg;  // compile-time for the common sizes
    const SolveSeg sg = solve_seg<W>(seg, nk);
    const int gx = sg.gx;
    {
        const int sv = blockIdx.y;
        if (!solve_active(active, sv) || static_cast<int>(blockIdx.x) >= gx) return;
        nk = sg.nk;
        const size_t od = sg.ko * nb * nb;
        const size_t ob = sg.ko * nb;
        V += od;
        lam += ob;
        Ft += od;
        Fe += od;
        p += ob;
        d_p += ob;
        w += sg.ko;
        partial += sv * kLoopPartialDoubles;
        sc += sv * SC_COUNT;
    }
    if (stage2 && sc[SC_ACCEPTED] > 0.5) return;  // uniform over the solve's workgroups
    extern __shared__ cd smem[];
    __shared__ double racc[W][2 * kMaxBt];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    cd* sV = smem + wave * ls_wave_cd(nb);
    cd* sF = sV + nb2;
    cd* sT = sF + nb2;
    cd* sVC = sT + nb2;
    cd* scf = sVC + nb2;
    double* srow = reinterpret_cast<double*>(scf + nb);
    double* slam = srow + nb2;
    for (int i = lane; i < 2 * kMaxBt; i += 64) racc[wave][i] = 0.0;
    const double dqn = sqrt(fmax(0.0, sc[SC_DQN2]));
    const double tau0 = fmin(1.0, max_step / fmax(dqn, TINY_REAL));
    double tau_lo = tau0;  // tau0 shrink^j_lo by repeated products, as the reference steps
    for (int j = 0; j < j_lo; ++j) tau_lo *= shrink;
    const int nsteps = (nk + gx * W - 1) / (gx * W);
    for (int step = 0; step < nsteps; ++step) {
        const int k = (step * gx + blockIdx.x) * W + wave;
        const bool act = k < nk;
        const size_t kb = static_cast<size_t>(k) * nb2;
        if (act) {
            for (int e = lane; e < nb2; e += 64) {
                sV[e] = V[kb + e];
                sF[e] = stage2 ? Fe[kb + e] : Ft[kb + e];
            }
            for (int i = lane; i < nb; i += 64) slam[i] = lam[static_cast<size_t>(k) * nb + i];
        }
        wave_lds_sync();
        if (!stage2) {
            if (act) wave_mm(sV, sF, sT, nb, true, false, lane, mfma);
            wave_lds_sync();
            if (act) wave_mm(sT, sV, sF, nb, false, false, lane, mfma);  // sF <- Fe
            wave_lds_sync();
            if (act)
                for (int e = lane; e < nb2; e += 64) Fe[kb + e] = sF[e];
        }
        double tau = tau_lo;
        for (int j = j_lo; j < j_hi; ++j) {
            diag_at_tau(sV, sF, slam, sVC, scf, srow, nb, tau, lane, mfma);
            double ej = 0.0, sj = 0.0;
            if (act) {
                const double wk = w[k];
                for (int i = lane; i < nb; i += 64) {
                    double di = 0.0;
                    for (int jj = 0; jj < nb; ++jj) di += srow[i * nb + jj];
                    const size_t idx = static_cast<size_t>(k) * nb + i;
                    const double pv = clampd(p[idx] - tau * d_p[idx], OCC_CLIP, 1.0 - OCC_CLIP);
                    ej += wk * pv * di;
                    sj -= wk * (pv * log(pv) + (1.0 - pv) * log1p(-pv));
                }
            }
            ej = wave_sum(ej);
            sj = wave_sum(sj);
            if (lane == 0) {
                racc[wave][2 * j] += ej;
                racc[wave][2 * j + 1] += sj;
            }
            wave_lds_sync();
            tau *= shrink;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * kMaxBt) {
        double t = 0.0;
        for (int wv = 0; wv < W; ++wv) t += racc[wv][threadIdx.x];
        partial[blockIdx.x * 2 * kMaxBt + threadIdx.x] = t;
    }
}

// Q_out <- Q V diag(c(tau)) V^H and eps_work = diag(U^H Ft U) at the chosen
// tau (Q_out may be Q: each k point's block is read before it is written).
// use_sc_tau = 0 with U = V (tau ignored): the finalise rotation Q <- Q V.
template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_retract(int nk, int nb_arg, const cd* Q, const cd* __restrict__ V,
          const double* __restrict__ lam, const cd* __restrict__ Fe,
          double* __restrict__ eps_work, int cayley, const double* sc, cd* Q_out,
          unsigned active, const int* seg, int mfma) {
    const int nb = NB > 0 ? NB : nb_arg;  // compile-time for the common sizes
    const SolveSeg sg = solve_seg<W>(seg, nk);
    const int gx = sg.gx;
    {
        const int sv = blockIdx.y;
        if (!solve_active(active, sv) || static_cast<int>(blockIdx.x) >= gx) return;
        nk = sg.nk;
        const size_t od = sg.ko * nb * nb;
        const size_t ob = sg.ko * nb;
        Q += od;
        Q_out += od;
        V += od;
        if (lam) lam += ob;
        if (Fe) Fe += od;
        if (eps_work) eps_work += ob;
        sc += sv * SC_COUNT;
    }
    extern __shared__ cd smem[];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    cd* sV = smem + wave * ls_wave_cd(nb);
    cd* sQ = sV + nb2;
    cd* sU = sQ + nb2;
    cd* sVC = sU + nb2;
    cd* scf = sVC + nb2;
    double* srow = reinterpret_cast<double*>(scf + nb);
    double* slam = srow + nb2;
    cd* sFe = sU;  // reused once U has been applied
    const double tau = cayley ? sc[SC_TAU] : 0.0;
    const int nsteps = (nk + gx * W - 1) / (gx * W);
    for (int step = 0; step < nsteps; ++step) {
        const int k = (step * gx + blockIdx.x) * W + wave;
        const bool act = k < nk;
        const size_t kb = static_cast<size_t>(k) * nb2;
        if (act) {
            for (int e = lane; e < nb2; e += 64) {
                sV[e] = V[kb + e];
                sQ[e] = Q[kb + e];
            }
            for (int i = lane; i < nb; i += 64) {
                slam[i] = cayley ? lam[static_cast<size_t>(k) * nb + i] : 0.0;
                scf[i] = cayley ? cayley_c(slam[i], tau) : cd{1.0, 0.0};
            }
        }
        wave_lds_sync();
        if (act) {
            if (cayley) {
                // U = (V diag(c)) V^H
                for (int e = lane; e < nb2; e += 64) sVC[e] = cmul(sV[e], scf[e % nb]);
            }
        }
        wave_lds_sync();
        if (act) {
            if (cayley) wave_mm(sVC, sV, sU, nb, false, true, lane, mfma);
        }
        wave_lds_sync();
        if (act) {
            // Q_new = Q U (cayley) or Q V (finalise)
            const cd* R = cayley ? sU : sV;
            cd* Qk = Q_out + kb;
            wave_mm_each(sQ, R, nb, false, false, lane, [Qk](int e, cd v) { Qk[e] = v; }, mfma);
        }
        wave_lds_sync();
        if (cayley) {
            if (act)
                for (int e = lane; e < nb2; e += 64) sFe[e] = Fe[kb + e];
            wave_lds_sync();
            diag_at_tau(sV, sFe, slam, sVC, scf, srow, nb, tau, lane, mfma);
            if (act)
                for (int i = lane; i < nb; i += 64) {
                    double di = 0.0;
                    for (int jj = 0; jj < nb; ++jj) di += srow[i * nb + jj];
                    eps_work[static_cast<size_t>(k) * nb + i] = di;
                }
            wave_lds_sync();
        }
    }
}

// ---------------------------------------------------------------------------
// SCF / generic reductions
// ---------------------------------------------------------------------------

// delta = Pn - D; comm = F D - D F (stored when comm != nullptr).
// Slots: [sum w |delta|^2, sum w |comm|^2].
template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_scf_residuals(int nk, int nb_arg, const cd* __restrict__ F, const cd* __restrict__ D,
                const cd* __restrict__ Pn, cd* __restrict__ delta, cd* __restrict__ comm,
                const double* __restrict__ w, double* partial, unsigned active) {
    const int nb = NB > 0 ? NB : nb_arg;  // compile-time for the common sizes
    {
        const int sv = blockIdx.y;
        if (!solve_active(active, sv)) return;
        const size_t od = static_cast<size_t>(sv) * nk * nb * nb;
        F += od;
        D += od;
        Pn += od;
        delta += od;
        if (comm) comm += od;
        w += static_cast<size_t>(sv) * nk;
        partial += sv * kLoopPartialDoubles;
    }
    extern __shared__ cd smem[];
    __shared__ double red[W * 2];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    cd* sF = smem + wave * 2 * nb2;
    cd* sD = sF + nb2;
    double acc[2] = {0.0, 0.0};
    const int nsteps = (nk + gridDim.x * W - 1) / (gridDim.x * W);
    for (int step = 0; step < nsteps; ++step) {
        const int k = (step * gridDim.x + blockIdx.x) * W + wave;
        const bool act = k < nk;
        const size_t kb = static_cast<size_t>(k) * nb2;
        if (act)
            for (int e = lane; e < nb2; e += 64) {
                sF[e] = F[kb + e];
                sD[e] = D[kb + e];
            }
        wave_lds_sync();
        if (act) {
            const double wk = w[k];
            double sd = 0.0, scm = 0.0;
            for (int e = lane; e < nb2; e += 64) {
                const int i = e / nb, j = e % nb;
                const cd dl = csub(Pn[kb + e], sD[e]);
                delta[kb + e] = dl;
                sd += cnorm(dl);
                cd fp{0.0, 0.0}, pf{0.0, 0.0};
                for (int n = 0; n < nb; ++n) {
                    fp = cadd(fp, cmul(sF[i * nb + n], sD[n * nb + j]));
                    pf = cadd(pf, cmul(sD[i * nb + n], sF[n * nb + j]));
                }
                const cd c = csub(fp, pf);
                if (comm) comm[kb + e] = c;
                scm += cnorm(c);
            }
            acc[0] += wk * sd;
            acc[1] += wk * scm;
        }
        wave_lds_sync();
    }
    store_block_partials<2, W>(acc, red, partial);
}

// out = sum_k w_k sum_i Re(conj(X_ki) Y_ki)  (m = elements per k)
__global__ void __launch_bounds__(256)
k_wdot(int nk, int m, const cd* __restrict__ X, const cd* __restrict__ Y,
       const double* __restrict__ w, double* partial) {
    __shared__ double red[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc[1] = {0.0};
    for (int k = blockIdx.x * 4 + wave; k < nk; k += gridDim.x * 4) {
        const size_t kb = static_cast<size_t>(k) * m;
        double s = 0.0;
        for (int e = lane; e < m; e += 64) {
            const cd x = X[kb + e], y = Y[kb + e];
            s += x.x * y.x + x.y * y.y;
        }
        acc[0] += w[k] * s;
    }
    store_block_partials<1, 4>(acc, red, partial);
}

// k_wdot for nd arrays X_d against one Y: the same per-thread terms, k
// order and reduction tree per dot, Y read once.
__global__ void __launch_bounds__(256)
k_wdot_multi(int nk, int m, CdPtrs X, int nd, const cd* __restrict__ Y,
             const double* __restrict__ w, double* partial) {
    __shared__ double red[4 * kMaxDots];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc[kMaxDots];
#pragma unroll
    for (int d = 0; d < kMaxDots; ++d) acc[d] = 0.0;
    for (int k = blockIdx.x * 4 + wave; k < nk; k += gridDim.x * 4) {
        const size_t kb = static_cast<size_t>(k) * m;
        double s[kMaxDots];
#pragma unroll
        for (int d = 0; d < kMaxDots; ++d) s[d] = 0.0;
        for (int e = lane; e < m; e += 64) {
            const cd y = Y[kb + e];
#pragma unroll
            for (int d = 0; d < kMaxDots; ++d) {
                if (d < nd) {
                    const cd x = X.p[d][kb + e];
                    s[d] += x.x * y.x + x.y * y.y;
                }
            }
        }
#pragma unroll
        for (int d = 0; d < kMaxDots; ++d)
            if (d < nd) acc[d] += w[k] * s[d];
    }
    store_block_partials<kMaxDots, 4>(acc, red, partial);
}

__global__ void __launch_bounds__(256)
k_lincomb(size_t n, CdPtrs X, DotCoefs c, int nx, int accumulate, double a, double b,
          const cd* __restrict__ Fcur, cd* __restrict__ out) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        cd y = accumulate ? out[i] : cd{0.0, 0.0};
#pragma unroll
        for (int d = 0; d < kMaxDots; ++d) {
            if (d < nx) {
                const cd v = X.p[d][i];
                y = cd{y.x + c.c[d] * v.x, y.y + c.c[d] * v.y};
            }
        }
        if (Fcur) {
            const cd v = Fcur[i];
            y = cd{a * y.x + b * v.x, a * y.y + b * v.y};
        }
        out[i] = y;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

// The per-k kernels' products run on the matrix cores at 9..16 bands unless
// HFGPU_LOOP_MFMA=0 (read once, on the host; a launch argument).
static bool loop_mfma_on() {
    static const bool on = [] {
        const char* e = std::getenv("HFGPU_LOOP_MFMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

// the launch argument of the per-k kernels' products: matrix cores at 9..16
// bands
static int loop_mfma(int nb) { return loop_mfma_on() && nb > 8 && nb <= 16 ? 1 : 0; }

void launch_finalize_sums(const double* partial, int nparts, int ns, double* out,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_finalize<0>, dim3(1), dim3(256), 0, st, partial, nparts, ns, out,
                       nullptr, 0, 0.0, 0.0, 0.0, 0, 0, 0, 0, ~0u, nullptr, 1);
    HFGPU_LAUNCH_CHECK();
}

// per solve of the scratch's batch: out is solve 0's slot; wv wavefronts per
// workgroup of the launch that wrote the partials
static void finalize(LoopScratch& ls, int nparts, int ns, double* out, hipStream_t st,
                     int wv) {
    hipLaunchKernelGGL(k_finalize<0>, dim3(ls.S), dim3(256), 0, st, ls.partial, nparts, ns, out,
                       nullptr, 0, 0.0, 0.0, 0.0, 0, 0, 0, 0, ls.active, ls.seg, wv);
    HFGPU_LAUNCH_CHECK();
}

void launch_orbital_grad(int nk, int nb, const cd* Q, cd* F, const double* p, const double* w,
                         const MuState* mus, double T_r, bool grad, double* eps, cd* G,
                         double* g_p, LoopScratch& ls, hipStream_t st) {
    check_nb(nb);
    const size_t per_wave = sizeof(cd) * 3 * nb * nb;
    HFGPU_LOOP_LAUNCH(k_orbital_grad, nb, nk, per_wave, st, ls.S, nk, nb, Q, F, p, w, mus, T_r,
                      grad ? 1 : 0, eps, G, g_p, ls.partial, ls.sc, ls.active, ls.seg,
                      loop_mfma(nb));
    HFGPU_LAUNCH_CHECK();
    if (grad) finalize(ls, loop_blocks(nk, waves_for(nb)), 2, ls.sc + SC_GN2, st, waves_for(nb));
}

void launch_precond(int nk, int nb, const cd* G, const cd* Gp, const double* eps,
                    const double* p, const double* g_p, const double* g_pp, const double* w,
                    double T_r, double denom_scale, cd* dQ, double* h_p, LoopScratch& ls,
                    hipStream_t st) {
    check_nb(nb);
    HFGPU_LOOP_LAUNCH(k_precond, nb, nk, 0, st, ls.S, nk, nb, G, Gp, eps, p, g_p, g_pp, w, T_r,
                      denom_scale, dQ, h_p, ls.partial, ls.sc, ls.active, ls.seg);
    HFGPU_LAUNCH_CHECK();
    finalize(ls, loop_blocks(nk, waves_for(nb)), 6, ls.sc + SC_PR1, st, waves_for(nb));
}

void launch_cg_combine(int nk, int nb, bool restart, cd* dQ, const cd* dQp, const double* h_p,
                       const double* d_pp, double* d_p, const double* eps, const double* p,
                       const double* w, double T_r, LoopScratch& ls, hipStream_t st) {
    check_nb(nb);
    HFGPU_LOOP_LAUNCH(k_cg_combine, nb, nk, 0, st, ls.S, nk, nb, restart ? 1 : 0, dQ, dQp, h_p,
                      d_pp, d_p, eps, p, w, T_r, ls.partial, ls.sc, ls.active, ls.seg);
    HFGPU_LAUNCH_CHECK();
    finalize(ls, loop_blocks(nk, waves_for(nb)), 3, ls.sc + SC_DQN2, st, waves_for(nb));
}

static size_t ls_wave_bytes(int nb) { return sizeof(cd) * static_cast<size_t>(ls_wave_cd(nb)); }

void launch_linesearch(int nk, int nb, const cd* V, const double* lam, const cd* Ft, cd* Fe,
                       const double* p, const double* d_p, const double* w, double T_r,
                       double max_step, double shrink, int nbt, LoopScratch& ls,
                       hipStream_t st) {
    check_nb(nb);
    if (nbt < 0 || nbt > kMaxBt)
        throw std::invalid_argument("bt_max above the GPU line-search limit of 16");
    const size_t per_wave = ls_wave_bytes(nb);
    const int nparts = static_cast<int>(loop_blocks(nk, waves_for(nb)));
    // stage 1: Fe and tau0, which the reference's backtracking usually
    // accepts; the later stages (device-gated) only run while no trial has
    // been accepted: trials 1-3, which take most of the rest, then the others
    int bounds[4], nst = 0;
    bounds[nst++] = 0;
    for (int b : {1, 4, nbt})
        if (b > bounds[nst - 1] && b <= nbt) bounds[nst++] = b;
    if (nst == 1) bounds[nst++] = 0;  // nbt = 0: one empty stage applies the no-trial rule
    for (int stage = 0; stage + 1 < nst; ++stage) {
        const int j_lo = bounds[stage], j_hi = bounds[stage + 1];
        const int last = stage + 2 == nst;
        HFGPU_LOOP_LAUNCH(k_linesearch, nb, nk, per_wave, st, ls.S, nk, nb, V, lam, Ft, Fe, p,
                          d_p, w, T_r, max_step, shrink, j_lo, j_hi, stage, ls.partial, ls.sc,
                          ls.active, ls.seg, loop_mfma(nb));
        HFGPU_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_finalize<1>, dim3(ls.S), dim3(256), 0, st, ls.partial, nparts,
                           2 * kMaxBt, ls.sc + SC_TRIAL, ls.sc, nbt, shrink, T_r, max_step, j_lo,
                           j_hi, last, stage, ls.active, ls.seg, waves_for(nb));
        HFGPU_LAUNCH_CHECK();
    }
}

void launch_retract(int nk, int nb, cd* Q, const cd* V, const double* lam, const cd* Fe,
                    double* eps_work, bool cayley, LoopScratch& ls, hipStream_t st, cd* Q_out) {
    check_nb(nb);
    const size_t per_wave = ls_wave_bytes(nb);
    HFGPU_LOOP_LAUNCH(k_retract, nb, nk, per_wave, st, ls.S, nk, nb, Q, V, lam, Fe, eps_work,
                      cayley ? 1 : 0, ls.sc, Q_out ? Q_out : Q, ls.active, ls.seg, loop_mfma(nb));
    HFGPU_LAUNCH_CHECK();
}

void launch_scf_residuals(int nk, int nb, const cd* F, const cd* D, const cd* Pn, cd* delta,
                          cd* comm, const double* w, LoopScratch& ls, double* out,
                          hipStream_t st) {
    check_nb(nb);
    if (ls.seg) throw std::logic_error("launch_scf_residuals takes solves of one k count");
    const size_t per_wave = sizeof(cd) * 2 * nb * nb;
    HFGPU_LOOP_LAUNCH(k_scf_residuals, nb, nk, per_wave, st, ls.S, nk, nb, F, D, Pn, delta,
                      comm, w, ls.partial, ls.active);
    HFGPU_LAUNCH_CHECK();
    finalize(ls, loop_blocks(nk, waves_for(nb)), 2, out, st, waves_for(nb));
}

void launch_wdot(int nk, int m, const cd* X, const cd* Y, const double* w, LoopScratch& ls,
                 double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_wdot, dim3(loop_blocks(nk, 4)), dim3(256), 0, st, nk, m, X, Y, w,
                       ls.partial);
    HFGPU_LAUNCH_CHECK();
    if (ls.seg) throw std::logic_error("launch_wdot takes solves of one k count");
    finalize(ls, loop_blocks(nk, 4), 1, out, st, 4);
}

void launch_wdot_multi(int nk, int m, const CdPtrs& X, int nd, const cd* Y, const double* w,
                       LoopScratch& ls, double* out, hipStream_t st) {
    if (nd < 1 || nd > kMaxDots) throw std::invalid_argument("launch_wdot_multi: 1 <= nd <= 16");
    hipLaunchKernelGGL(k_wdot_multi, dim3(loop_blocks(nk, 4)), dim3(256), 0, st, nk, m, X, nd,
                       Y, w, ls.partial);
    HFGPU_LAUNCH_CHECK();
    if (ls.seg) throw std::logic_error("launch_wdot_multi takes solves of one k count");
    finalize(ls, loop_blocks(nk, 4), kMaxDots, out, st, 4);
}

void launch_lincomb(size_t n, const CdPtrs& X, const DotCoefs& c, int nx, bool accumulate,
                    double a, double b, const cd* Fcur, cd* out, hipStream_t st) {
    if (nx < 0 || nx > kMaxDots) throw std::invalid_argument("launch_lincomb: 0 <= nx <= 16");
    if (!n) return;
    const unsigned grid = std::max(1u, std::min<unsigned>(cdiv(n, 256), 8192u));
    hipLaunchKernelGGL(k_lincomb, dim3(grid), dim3(256), 0, st, n, X, c, nx, accumulate ? 1 : 0,
                       a, b, Fcur, out);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
