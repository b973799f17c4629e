// Per-k kernels of the trust-region Newton solver (solver_rtr.hpp:162-459).
//
// One Steihaug inner iteration on the device is
//   direction  d = P(r) + beta d, dP = herm(Q M(d) Q^H)
//   Fock       F[dP] without the reference density (three exchange passes)
//   hvp        Hd from the response, plus the inner products the host needs
//   update     v, Hv, r and the next preconditioned products
// with the host reading back the scalars between hvp and update to take the
// reference's boundary / step / stop decisions.  Layout and launch shape
// follow loops.hip: one wavefront per k point with its nb x nb blocks in
// LDS, four wavefronts per workgroup up to nb = 16.
#include <cmath>

#include "loop_common.h"
#include "rtr.h"

namespace hfgpu {

namespace {

using namespace loopk;

__device__ __forceinline__ double occ_curv(double pv, double T_r) {
    const double pc = clampd(pv, OCC_CLIP, 1.0 - OCC_CLIP);
    return pc * (1.0 - pc) / T_r;  // inverse of the occupation curvature T / (p (1 - p))
}

template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_rtr_direction(int nk, int nb_arg, int init, double beta, double s_z, double lam_pc,
                double T_r, const cd* __restrict__ Q, const double* __restrict__ p,
                const double* __restrict__ eps, const cd* __restrict__ rX,
                const double* __restrict__ rp, cd* __restrict__ dX, double* __restrict__ dp,
                cd* __restrict__ dP) {
    const int nb = NB > 0 ? NB : nb_arg;
    extern __shared__ cd smem[];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    cd* sQ = smem + wave * 3 * nb2;
    cd* sM = sQ + nb2;
    cd* sT = sM + nb2;
    const double lam2 = lam_pc * lam_pc;
    const int nsteps = (nk + gridDim.x * W - 1) / (gridDim.x * W);
    for (int step = 0; step < nsteps; ++step) {
        const int k = (step * gridDim.x + blockIdx.x) * W + wave;
        const bool act = k < nk;
        const size_t kb = static_cast<size_t>(k) * nb2, kn = static_cast<size_t>(k) * nb;
        if (act) {
            const double* ek = eps + kn;
            const double* pk = p + kn;
            for (int e = lane; e < nb2; e += 64) {
                const int i = e / nb, j = e % nb;
                sQ[e] = Q[kb + e];
                const double gap = ek[i] - ek[j];
                const cd z = cscale(rX[kb + e], 1.0 / sqrt(gap * gap + lam2));
                const cd d = init ? z : cadd(z, cscale(dX[kb + e], beta));
                dX[kb + e] = d;
                if (i != j) sM[e] = cscale(d, pk[j] - pk[i]);
            }
            for (int i = lane; i < nb; i += 64) {
                const double z = rp[kn + i] * occ_curv(pk[i], T_r) - s_z;
                const double d = init ? z : z + beta * dp[kn + i];
                dp[kn + i] = d;
                sM[i * nb + i] = cd{d, 0.0};
            }
        }
        __syncthreads();
        if (act) wave_mm(sQ, sM, sT, nb, false, false, lane);  // Q M
        __syncthreads();
        if (act) wave_mm(sT, sQ, sM, nb, false, true, lane);   // (Q M) Q^H
        __syncthreads();
        if (act)
            for (int e = lane; e < nb2; e += 64) {
                const int i = e / nb, j = e % nb;
                const cd a = sM[e], b = sM[j * nb + i];
                dP[kb + e] = cd{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
            }
        __syncthreads();
    }
}

template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_rtr_hvp(int nk, int nb_arg, double T_r, double pen, const cd* __restrict__ Q,
          const double* __restrict__ p, const cd* __restrict__ Ft, const cd* __restrict__ dX,
          const double* __restrict__ dp, const cd* __restrict__ Fr, const cd* __restrict__ h,
          const cd* __restrict__ dP, const cd* __restrict__ vX, const double* __restrict__ vp,
          const double* __restrict__ w, cd* __restrict__ HdX, double* __restrict__ Hp_raw,
          double* partial) {
    const int nb = NB > 0 ? NB : nb_arg;
    extern __shared__ cd smem[];
    __shared__ double red[W * RH_COUNT];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    cd* sQ = smem + wave * 5 * nb2;
    cd* sF = sQ + nb2;
    cd* sX = sF + nb2;
    cd* sR = sX + nb2;
    cd* sT = sR + nb2;
    double acc[RH_COUNT] = {0, 0, 0, 0, 0, 0, 0};
    const int nsteps = (nk + gridDim.x * W - 1) / (gridDim.x * W);
    for (int step = 0; step < nsteps; ++step) {
        const int k = (step * gridDim.x + blockIdx.x) * W + wave;
        const bool act = k < nk;
        const size_t kb = static_cast<size_t>(k) * nb2, kn = static_cast<size_t>(k) * nb;
        if (act)
            for (int e = lane; e < nb2; e += 64) {
                sQ[e] = Q[kb + e];
                sF[e] = Ft[kb + e];
                sX[e] = dX[kb + e];
                const cd r = csub(Fr[kb + e], h[kb + e]);
                sR[e] = pen != 0.0 ? cadd(r, cscale(dP[kb + e], pen)) : r;
            }
        __syncthreads();
        if (act) wave_mm(sQ, sR, sT, nb, true, false, lane);  // Q^H R
        __syncthreads();
        if (act) wave_mm(sT, sQ, sR, nb, false, false, lane);  // St = Q^H R Q
        __syncthreads();
        if (act)
            for (int e = lane; e < nb2; e += 64) {  // A = Ft X - X Ft + St
                const int i = e / nb, j = e % nb;
                cd s = sR[e];
                for (int n = 0; n < nb; ++n) {
                    s = cadd(s, cmul(sF[i * nb + n], sX[n * nb + j]));
                    s = csub(s, cmul(sX[i * nb + n], sF[n * nb + j]));
                }
                sT[e] = s;
            }
        __syncthreads();
        if (act) {
            const double* pk = p + kn;
            const double* dk = dp + kn;
            double s_dhd = 0.0, s_dd = 0.0, s_vd = 0.0, s_vv = 0.0;
            for (int e = lane; e < nb2; e += 64) {
                const int i = e / nb, j = e % nb;
                cd hv{0.0, 0.0};
                if (i != j) {
                    const cd cij = cadd(cscale(sF[e], dk[j] - dk[i]), cscale(sT[e], pk[j] - pk[i]));
                    const int t = j * nb + i;
                    const cd cji = cadd(cscale(sF[t], dk[i] - dk[j]), cscale(sT[t], pk[i] - pk[j]));
                    hv = cd{0.5 * (cij.x - cji.x), 0.5 * (cij.y + cji.y)};
                }
                HdX[kb + e] = hv;
                const cd x = sX[e], v = vX[kb + e];
                s_dhd += x.x * hv.x + x.y * hv.y;
                s_dd += cnorm(x);
                s_vd += v.x * x.x + v.y * x.y;
                s_vv += cnorm(v);
            }
            double s_dhp = 0.0, s_hp = 0.0, s_dp = 0.0, s_dd_p = 0.0, s_vd_p = 0.0, s_vv_p = 0.0;
            for (int b = lane; b < nb; b += 64) {
                const double pc = clampd(pk[b], OCC_CLIP, 1.0 - OCC_CLIP);
                const double hp = sT[b * nb + b].x + T_r / (pc * (1.0 - pc)) * dk[b];
                Hp_raw[kn + b] = hp;
                const double vb = vp[kn + b];
                s_dhp += dk[b] * hp;
                s_hp += hp;
                s_dp += dk[b];
                s_dd_p += dk[b] * dk[b];
                s_vd_p += vb * dk[b];
                s_vv_p += vb * vb;
            }
            const double wk = w[k];
            acc[RH_DHDX] += s_dhd;
            acc[RH_DHP] += wk * s_dhp;
            acc[RH_HP] += wk * s_hp;
            acc[RH_DP] += wk * s_dp;
            acc[RH_DD] += s_dd + wk * s_dd_p;
            acc[RH_VD] += s_vd + wk * s_vd_p;
            acc[RH_VV] += s_vv + wk * s_vv_p;
        }
        __syncthreads();
    }
    store_block_partials<RH_COUNT, W>(acc, red, partial);
}

template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_rtr_update(int nk, int nb_arg, int mode, double a, double s_H, double lam_pc, double T_r,
             const double* __restrict__ eps, const double* __restrict__ p,
             const cd* __restrict__ G, const double* __restrict__ w, const cd* __restrict__ dX,
             const double* __restrict__ dp, const cd* __restrict__ HdX,
             const double* __restrict__ Hp_raw, cd* __restrict__ vX, double* __restrict__ vp,
             cd* __restrict__ HvX, double* __restrict__ Hvp, cd* __restrict__ rX,
             double* __restrict__ rp, double* partial, const double* __restrict__ step) {
    const int nb = NB > 0 ? NB : nb_arg;
    __shared__ double red[W * RU_COUNT];
    if (step) {  // the device's step decision (uniform)
        mode = static_cast<int>(step[TS_MODE]);
        a = step[TS_A];
        s_H = step[TS_SH];
    }
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double lam2 = lam_pc * lam_pc;
    double acc[RU_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = blockIdx.x * W + wave; k < nk; k += gridDim.x * W) {
        const size_t kb = static_cast<size_t>(k) * nb2, kn = static_cast<size_t>(k) * nb;
        const double* ek = eps + kn;
        double gv = 0.0, vhv = 0.0, vv = 0.0, rzx = 0.0, rn = 0.0;
        for (int e = lane; e < nb2; e += 64) {
            const int i = e / nb, j = e % nb;
            cd v, hv, r;
            const cd g = G[kb + e];
            if (mode == 0) {
                v = cd{0.0, 0.0};
                hv = v;
                r = cneg(g);
                rX[kb + e] = r;
            } else {
                const cd d = dX[kb + e], hd = HdX[kb + e];
                v = cadd(vX[kb + e], cscale(d, a));
                hv = cadd(HvX[kb + e], cscale(hd, a));
                if (mode == 1) {
                    r = csub(rX[kb + e], cscale(hd, a));
                    rX[kb + e] = r;
                } else {
                    r = rX[kb + e];
                }
            }
            vX[kb + e] = v;
            HvX[kb + e] = hv;
            gv += g.x * v.x + g.y * v.y;
            vhv += v.x * hv.x + v.y * hv.y;
            vv += cnorm(v);
            const double gap = ek[i] - ek[j];
            rzx += cnorm(r) / sqrt(gap * gap + lam2);
            rn += cnorm(r);
        }
        double vhv_p = 0.0, vv_p = 0.0, rz_p = 0.0, r_p = 0.0, z_p = 0.0, rn_p = 0.0;
        for (int b = lane; b < nb; b += 64) {
            double v, hv, r;
            if (mode == 0) {
                v = 0.0;
                hv = 0.0;
                r = 0.0;
                rp[kn + b] = r;
            } else {
                const double hd = Hp_raw[kn + b] - s_H;
                v = vp[kn + b] + a * dp[kn + b];
                hv = Hvp[kn + b] + a * hd;
                if (mode == 1) {
                    r = rp[kn + b] - a * hd;
                    rp[kn + b] = r;
                } else {
                    r = rp[kn + b];
                }
            }
            vp[kn + b] = v;
            Hvp[kn + b] = hv;
            const double z = r * occ_curv(p[kn + b], T_r);
            vhv_p += v * hv;
            vv_p += v * v;
            rz_p += r * z;
            r_p += r;
            z_p += z;
            rn_p += r * r;
        }
        const double wk = w[k];
        acc[RU_GV] += wk * gv;
        acc[RU_VHV] += wk * (vhv + vhv_p);
        acc[RU_VV] += vv + wk * vv_p;
        acc[RU_RZX] += rzx;
        acc[RU_RZP] += wk * rz_p;
        acc[RU_RP] += wk * r_p;
        acc[RU_ZP] += wk * z_p;
        acc[RU_RN2] += rn + wk * rn_p;
    }
    store_block_partials<RU_COUNT, W>(acc, red, partial);
}

template <int W, int NB>
__global__ void __launch_bounds__(W * 64)
k_rtr_retract(int nk, int nb_arg, double tau, const cd* __restrict__ V,
              const double* __restrict__ lam, const cd* __restrict__ Q,
              const cd* __restrict__ Ft, cd* __restrict__ Qt, double* __restrict__ eps_t) {
    const int nb = NB > 0 ? NB : nb_arg;
    extern __shared__ cd smem[];
    const int nb2 = nb * nb;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    cd* sV = smem + wave * (4 * nb2 + nb);
    cd* sU = sV + nb2;
    cd* sQ = sU + nb2;
    cd* sF = sQ + nb2;
    cd* sc = sF + nb2;
    const int nsteps = (nk + gridDim.x * W - 1) / (gridDim.x * W);
    for (int step = 0; step < nsteps; ++step) {
        const int k = (step * gridDim.x + blockIdx.x) * W + wave;
        const bool act = k < nk;
        const size_t kb = static_cast<size_t>(k) * nb2, kn = static_cast<size_t>(k) * nb;
        if (act) {
            for (int e = lane; e < nb2; e += 64) {
                sV[e] = V[kb + e];
                sQ[e] = Q[kb + e];
                sF[e] = Ft[kb + e];
            }
            for (int n = lane; n < nb; n += 64) sc[n] = cayley_c(lam[kn + n], tau);
        }
        __syncthreads();
        if (act)
            for (int e = lane; e < nb2; e += 64) {  // U = V diag(c) V^H
                const int i = e / nb, j = e % nb;
                cd s{0.0, 0.0};
                for (int n = 0; n < nb; ++n)
                    s = cadd(s, cmulc(cmul(sV[i * nb + n], sc[n]), sV[j * nb + n]));
                sU[e] = s;
            }
        __syncthreads();
        if (act)
            for (int e = lane; e < nb2; e += 64) {  // Qt = Q U
                const int i = e / nb, j = e % nb;
                cd s{0.0, 0.0};
                for (int n = 0; n < nb; ++n) s = cadd(s, cmul(sQ[i * nb + n], sU[n * nb + j]));
                Qt[kb + e] = s;
            }
        __syncthreads();
        if (act) wave_mm(sF, sU, sQ, nb, false, false, lane);  // Ft U (over Q)
        __syncthreads();
        if (act)
            for (int i = lane; i < nb; i += 64) {  // diag(U^H (Ft U))
                double s = 0.0;
                for (int n = 0; n < nb; ++n) {
                    const cd u = sU[n * nb + i], t = sQ[n * nb + i];
                    s += u.x * t.x + u.y * t.y;
                }
                eps_t[kn + i] = s;
            }
        __syncthreads();
    }
}

// out = sum_k w_k sum_e |P - T|^2, one k point per wavefront
__global__ void __launch_bounds__(256)
k_wdist2(int nk, int m, const cd* __restrict__ P, const cd* __restrict__ T,
         const double* __restrict__ w, double* partial) {
    __shared__ double red[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc[1] = {0.0};
    for (int k = blockIdx.x * 4 + wave; k < nk; k += gridDim.x * 4) {
        const size_t kb = static_cast<size_t>(k) * m;
        double s = 0.0;
        for (int e = lane; e < m; e += 64) s += cnorm(csub(P[kb + e], T[kb + e]));
        acc[0] += w[k] * s;
    }
    store_block_partials<1, 4>(acc, red, partial);
}

// F += sum_t 2 phi'(d2_t) (P - T_t), accumulated over t in order first
__global__ void __launch_bounds__(256)
k_deflation_apply(size_t n, int n_defl, const cd* __restrict__ P, const cd* __restrict__ T,
                  const double* __restrict__ d2, double sigma, double inv2L2, cd* __restrict__ F,
                  double* __restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double Phi = 0.0, diag = 0.0;
        for (int t = 0; t < n_defl; ++t) {
            const double phi = sigma * exp(-d2[t] * inv2L2);
            Phi += phi;
            diag += -phi * inv2L2;
        }
        out[0] = Phi;
        out[1] = 2.0 * diag;
    }
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const cd pv = P[i];
        cd s{0.0, 0.0};
        for (int t = 0; t < n_defl; ++t) {
            const double phi = sigma * exp(-d2[t] * inv2L2);
            const double c = 2.0 * (-phi * inv2L2);
            s = cadd(s, cscale(csub(pv, T[static_cast<size_t>(t) * n + i]), c));
        }
        F[i] = cadd(F[i], s);
    }
}

}  // namespace

void launch_rtr_direction(int nk, int nb, bool init, double beta, double s_z, double lam_pc,
                          double T_r, const cd* Q, const double* p, const double* eps,
                          const cd* rX, const double* rp, cd* dX, double* dp, cd* dP,
                          hipStream_t st) {
    check_nb(nb);
    const size_t per_wave = sizeof(cd) * 3 * nb * nb;
    HFGPU_LOOP_LAUNCH(k_rtr_direction, nb, nk, per_wave, st, 1, nk, nb, init ? 1 : 0, beta, s_z,
                      lam_pc, T_r, Q, p, eps, rX, rp, dX, dp, dP);
    HFGPU_LAUNCH_CHECK();
}

void launch_rtr_hvp(int nk, int nb, double T_r, double pen, const cd* Q, const double* p,
                    const cd* Ft, const cd* dX, const double* dp, const cd* Fr, const cd* h,
                    const cd* dP, const cd* vX, const double* vp, const double* w, cd* HdX,
                    double* Hp_raw, LoopScratch& ls, double* out, hipStream_t st) {
    check_nb(nb);
    const size_t per_wave = sizeof(cd) * 5 * nb * nb;
    HFGPU_LOOP_LAUNCH(k_rtr_hvp, nb, nk, per_wave, st, 1, nk, nb, T_r, pen, Q, p, Ft, dX, dp, Fr,
                      h, dP, vX, vp, w, HdX, Hp_raw, ls.partial);
    HFGPU_LAUNCH_CHECK();
    launch_finalize_sums(ls.partial, loop_blocks(nk, waves_for(nb)), RH_COUNT, out, st);
}

void launch_rtr_update(int nk, int nb, int mode, double a, double s_H, double lam_pc,
                       double T_r, const double* eps, const double* p, const cd* G,
                       const double* w, const cd* dX, const double* dp, const cd* HdX,
                       const double* Hp_raw, cd* vX, double* vp, cd* HvX, double* Hvp, cd* rX,
                       double* rp, LoopScratch& ls, double* out, hipStream_t st,
                       const double* step) {
    check_nb(nb);
    if (mode < 0 || mode > 2) throw std::invalid_argument("rtr update mode must be 0, 1 or 2");
    HFGPU_LOOP_LAUNCH(k_rtr_update, nb, nk, 0, st, 1, nk, nb, mode, a, s_H, lam_pc, T_r, eps, p, G,
                      w, dX, dp, HdX, Hp_raw, vX, vp, HvX, Hvp, rX, rp, ls.partial, step);
    HFGPU_LAUNCH_CHECK();
    launch_finalize_sums(ls.partial, loop_blocks(nk, waves_for(nb)), RU_COUNT, out, st);
}

namespace {

// std::max's selection
__device__ __forceinline__ double smax(double a, double b) { return a < b ? b : a; }

__global__ void k_tcg_step(const double* __restrict__ H, const double* __restrict__ U, int first,
                           double rz_host, double occ_den, double Delta, double* step) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double rz = first ? U[RU_RZX] : rz_host;
    const double s_H = H[RH_HP] / occ_den;
    const double dHd = H[RH_DHDX] + (H[RH_DHP] - s_H * H[RH_DP]);
    const double dd = H[RH_DD], vd = H[RH_VD], vv0 = H[RH_VV];
    bool boundary = !(dHd > 1e-14 * dd);
    double alpha = 0.0;
    if (!boundary) {
        alpha = rz / dHd;
        const double s2 = vv0 + 2.0 * alpha * vd + alpha * alpha * dd;
        boundary = sqrt(smax(0.0, s2)) >= Delta;
    }
    if (boundary) {
        const double a = dd, b = 2.0 * vd, c2 = smax(0.0, vv0);
        const double cc = c2 - Delta * Delta;
        const double t = (-b + sqrt(smax(b * b - 4.0 * a * cc, 0.0))) / (2.0 * a + 1e-30);
        step[TS_MODE] = 2.0;
        step[TS_A] = t;
    } else {
        step[TS_MODE] = 1.0;
        step[TS_A] = alpha;
    }
    step[TS_SH] = s_H;
    step[TS_RZ] = rz;
    step[TS_BOUND] = boundary ? 1.0 : 0.0;
}

}  // namespace

void launch_tcg_step(const double* H, const double* U, int first, double rz_host,
                     double occ_den, double Delta, double* step, hipStream_t st) {
    hipLaunchKernelGGL(k_tcg_step, dim3(1), dim3(64), 0, st, H, U, first, rz_host, occ_den,
                       Delta, step);
    HFGPU_LAUNCH_CHECK();
}

void launch_rtr_retract(int nk, int nb, double tau, const cd* V, const double* lam,
                        const cd* Q, const cd* Ft, cd* Qt, double* eps_t, hipStream_t st) {
    check_nb(nb);
    const size_t per_wave = sizeof(cd) * (4 * nb * nb + nb);
    HFGPU_LOOP_LAUNCH(k_rtr_retract, nb, nk, per_wave, st, 1, nk, nb, tau, V, lam, Q, Ft, Qt,
                      eps_t);
    HFGPU_LAUNCH_CHECK();
}

void launch_deflation(int nk, int nb, int n_defl, const cd* P, const cd* targets,
                      const double* w, double sigma, double length, double* d2, cd* F,
                      double* out, LoopScratch& ls, hipStream_t st) {
    const int m = nb * nb;
    const size_t n = static_cast<size_t>(nk) * m;
    for (int t = 0; t < n_defl; ++t) {
        hipLaunchKernelGGL(k_wdist2, dim3(loop_blocks(nk, 4)), dim3(256), 0, st, nk, m, P,
                           targets + static_cast<size_t>(t) * n, w, ls.partial);
        HFGPU_LAUNCH_CHECK();
        launch_finalize_sums(ls.partial, loop_blocks(nk, 4), 1, d2 + t, st);
    }
    const double inv2L2 = 1.0 / (2.0 * length * length);
    hipLaunchKernelGGL(k_deflation_apply, dim3(std::max(1u, std::min<unsigned>(cdiv(n, 256), 4096u))),
                       dim3(256), 0, st, n, n_defl, P, targets, d2, sigma, inv2L2, F, out);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
