// Kernels of the trust-region Newton solver (rtr.hip): launchers.
//
// Tangent vectors are pairs (X, x): X (nk, nb, nb) skew-Hermitian with a zero
// diagonal in the orbital basis, x (nk, nb) occupation changes.  The joint
// inner product is sum_k (Re <X_k, Y_k> + w_k <x_k, y_k>) with the raw k
// weights w (solver_rtr.hpp:211-217).  Every launcher that reduces writes
// its sums into `out` through the per-workgroup partials of ls.
#pragma once

#include "common.h"
#include "loops.h"

namespace hfgpu {

// Slots written by launch_rtr_update (out[0..8)).
enum RtrUpdateSlot : int {
    RU_GV = 0,    // sum_k w_k Re <G, vX>
    RU_VHV = 1,   // sum_k w_k (Re <vX, HvX> + <vp, Hvp>)
    RU_VV = 2,    // joint |v|^2
    RU_RZX = 3,   // sum Re <rX, zX>  (orbital part of <r, z>)
    RU_RZP = 4,   // sum_k w_k <rp, zp_raw>
    RU_RP = 5,    // sum_k w_k sum_b rp
    RU_ZP = 6,    // sum_k w_k sum_b zp_raw
    RU_RN2 = 7,   // joint |r|^2
    RU_COUNT = 8
};

// Slots written by launch_tcg_step (step[0..5)).
enum TcgStepSlot : int {
    TS_MODE = 0,   // 1: interior step alpha, 2: step t to the trust-region boundary
    TS_A = 1,      // alpha or t
    TS_SH = 2,     // the particle-conserving shift s_H of this direction
    TS_RZ = 3,     // <r, z> the step used
    TS_BOUND = 4,  // 1 when the step reached the boundary
    TS_COUNT = 5
};

// Slots written by launch_rtr_hvp (out[0..7)).
enum RtrHvpSlot : int {
    RH_DHDX = 0,  // sum Re <dX, HdX>
    RH_DHP = 1,   // sum_k w_k <dp, Hp_raw>
    RH_HP = 2,    // sum_k w_k sum_b Hp_raw
    RH_DP = 3,    // sum_k w_k sum_b dp
    RH_DD = 4,    // joint |d|^2
    RH_VD = 5,    // joint <v, d>
    RH_VV = 6,    // joint |v|^2
    RH_COUNT = 7
};

// Preconditioned direction and its density variation (solver_rtr.hpp:
// 218-237, 55-77): z = (rX / sqrt(gap^2 + lam^2), rp p (1 - p) / T - s_z),
// d = z (init) or z + beta d, then dP_k = herm(Q_k M_k Q_k^H) with
// M = offdiag(dX_ij (p_j - p_i)) + diag(dp).
void launch_rtr_direction(int nk, int nb, bool init, double beta, double s_z, double lam_pc,
                          double T_r, const cd* Q, const double* p, const double* eps,
                          const cd* rX, const double* rp, cd* dX, double* dp, cd* dP,
                          hipStream_t st);

// Assembly of the Hessian-vector product (solver_rtr.hpp:78-111) from the
// interaction response Fr = F[dP] without the reference density:
// R = Fr - h + pen dP, St = Q^H R Q, A = [Ft, dX] + St,
// C_ij = (dp_j - dp_i) Ft_ij + (p_j - p_i) A_ij, HdX = offdiag((C - C^H) / 2),
// Hp_raw = Re diag(A) + T dp / (p (1 - p)) before the particle-conserving
// shift, which the caller applies from out[RH_HP].
void launch_rtr_hvp(int nk, int nb, double T_r, double pen, const cd* Q, const double* p,
                    const cd* Ft, const cd* dX, const double* dp, const cd* Fr, const cd* h,
                    const cd* dP, const cd* vX, const double* vp, const double* w, cd* HdX,
                    double* Hp_raw, LoopScratch& ls, double* out, hipStream_t st);

// Steihaug CG updates (solver_rtr.hpp:283-346).  mode 0: v = Hv = 0,
// r = (-G, 0); mode 1: v += a d, Hv += a Hd, r -= a Hd; mode 2: v += a d,
// Hv += a Hd.  Hd's occupation part is Hp_raw - s_H.
void launch_rtr_update(int nk, int nb, int mode, double a, double s_H, double lam_pc,
                       double T_r, const double* eps, const double* p, const cd* G,
                       const double* w, const cd* dX, const double* dp, const cd* HdX,
                       const double* Hp_raw, cd* vX, double* vp, cd* HvX, double* Hvp, cd* rX,
                       double* rp, LoopScratch& ls, double* out, hipStream_t st,
                       const double* step = nullptr);

// One step decision of Steihaug's truncated CG (solver_rtr.hpp:296-330) on
// the device, from the HVP slots H (RtrHvpSlot) of direction d and, on the
// first step, <r, z> from the update slots U; rz_host otherwise.  Writes
// step[TcgStepSlot], which launch_rtr_update reads when given step: the
// interior step alpha = <r, z> / <d, Hd> unless the curvature is not
// positive or |v + alpha d| >= Delta, else the step t to the boundary.
// The arithmetic is the host's, uncontracted.
void launch_tcg_step(const double* H, const double* U, int first, double rz_host,
                     double occ_den, double Delta, double* step, hipStream_t st);

// Trial point (solver_rtr.hpp:366-381): U = V diag(c(lam, tau)) V^H,
// Qt = Q U, eps_t = diag(U^H Ft U).
void launch_rtr_retract(int nk, int nb, double tau, const cd* V, const double* lam,
                        const cd* Q, const cd* Ft, cd* Qt, double* eps_t, hipStream_t st);

// Deflation bias (solver_rtr.hpp:127-160): d2[t] = sum_k w_k |P_k - T_t,k|^2
// for each of n_defl targets, then F += sum_t 2 phi'(d2_t) (P - T_t) with
// phi(x) = sigma exp(-x / (2 L^2)); out[0] = Phi = sum_t phi, out[1] =
// 2 sum_t phi'.  d2 holds n_defl doubles.
void launch_deflation(int nk, int nb, int n_defl, const cd* P, const cd* targets,
                      const double* w, double sigma, double length, double* d2, cd* F,
                      double* out, LoopScratch& ls, hipStream_t st);

}  // namespace hfgpu
