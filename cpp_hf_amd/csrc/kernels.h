// Launchers of the per-k dense kernels, the eigensolver and the chemical
// potential solvers.  All take device pointers and a stream.
#pragma once

#include "common.h"

namespace hfgpu {

// ---- eigh (eigh.hip) -------------------------------------------------------
// in_mode 0: eigh(herm(M)); in_mode 1: eigh(herm(i M)) (Cayley setup).
void launch_eigh(int nmat, int nb, int in_mode, const cd* M, double* w, cd* V, hipStream_t st);

// launch_eigh over the solves of a batch: solve s owns the matrices
// seg[s] .. seg[s + 1] - 1 (seg on the device, nk_max the largest count);
// solves whose bit in active is clear are skipped.  Same results per matrix.
void launch_eigh_seg(const int* seg, int S, int nk_max, unsigned active, int nb, int in_mode,
                     const cd* M, double* w, cd* V, hipStream_t st);

// Block-diagonal structure of an nb x nb matrix (nb <= 32): block of each
// index and the first index of that block.  One block: the plain eigh.
struct BlockMap {
    signed char blk[32];
    signed char start[32];
};
// Block-diagonal eigh (linalg.hpp:72-120): the off-block entries are ignored,
// each block is diagonalised on its own; sort orders all eigenpairs
// ascending, otherwise each block's ascending pairs stay at the block's
// position.  With cond (two device doubles from launch_max_offblock: the
// off-block maximum and the overall maximum of M), the blocks apply only when
// off <= 1e-10 + 1e-10 max (solver_scf.hpp:111-120); otherwise the call is
// the plain sorted eigh.  The test runs on the device: no host round trip.
void launch_eigh_blocks(int nmat, int nb, int in_mode, const cd* M, double* w, cd* V,
                        const BlockMap& bm, bool sort, hipStream_t st,
                        const double* cond = nullptr);
// Matrices-per-wavefront autotuner of the nb <= 16 eigh kernels: times nmat
// Hermitian test matrices of nb bands with 64/NP, 32/NP, ... matrices per
// wavefront (down to an eighth), keeps the fastest for (nb, nmat) in the
// process and returns microseconds per call; results of later calls do not
// change.  nb 17..32: the fixed two per wavefront, 0 us.
double eigh_autotune(int nmat, int nb, hipStream_t st, int* mpw_out);
// eigh_autotune unless (nb, nmat) is already tuned in this process (nb <= 16).
void eigh_autotune_once(int nmat, int nb, hipStream_t st);

// out[0] = max |M_ij| over the off-block entries of all nmat matrices
// (linalg.hpp:123-140, max_offblock_sizes), out[1] = max |M_ij| over all
// entries (linalg.hpp:142-146, max_abs); partial needs
// 2 kMaxOffblockBlocks doubles.
constexpr unsigned kMaxOffblockBlocks = 1024;
void launch_max_offblock(int nmat, int nb, const cd* M, const BlockMap& bm, double* partial,
                         double* out, hipStream_t st);

// ---- chemical potential (mu.hip) -------------------------------------------
// Device-side state shared by the mu kernels of one solve.
struct MuState {
    double lo, hi, mu;
    double emin, emax;
    int done;
    int pad;
};

// Workspace of the mu solvers (doubles): two partial-sum buffers of up to
// kMuBlocks workgroups x 32 slots and two ping-pong states (the first 16384),
// then the exchange slots of the one-launch solves (rounds x 256 x 8).
constexpr int kMuBlocks = 240;
constexpr int kMuPartialDoubles = 16384 + 28 * 256 * 8;

// find_mu_bisection (utils.hpp:49-75): result mu = (lo+hi)/2 left in
// state->mu.  ws: kMuPartialDoubles of workspace.
void launch_mu_bisection(const double* eps, const double* w, int nk, int nb, double n_e,
                         double T, int maxiter, MuState* state, double* ws, hipStream_t st);
// solve_mu_inloop (utils.hpp:79-116); warm start read from warm->mu (warm
// defaults to state) when mu0_from_state, else mu0; the result goes to state.
// With occ, also the occupations occ = expit((mu - eps) / T) of
// launch_occupations (mode 1), formed by the Newton launch itself where it
// runs in one launch.
void launch_mu_newton(const double* eps, const double* w_norm, int nk, int nb,
                      double n_target_norm, double mu0, bool mu0_from_state, double T,
                      int maxiter, MuState* state, double* ws, hipStream_t st,
                      const MuState* warm = nullptr, double* occ = nullptr);

// launch_mu_newton for the S solves of a batch in two launches: solve s has
// the nb eigenvalues of its k points seg[s] .. seg[s + 1] - 1 (seg on the
// device, nk_max the largest count), weights w_norm + seg[s], target
// n_target[s] (device), warm start warm[s] (mu0 = 0 without
// mu0_from_state), result state[s], workspace ws + s kMuPartialDoubles;
// solves whose bit in active is clear are skipped.  Each mu equals the
// single call's bit for bit.  Only where mu_newton_multi_ok holds (the
// one-launch Newton paths); elsewhere call launch_mu_newton per solve.
bool mu_newton_multi_ok(int nk_max, int nb, int maxiter);
// With occ, also each active solve's occupations (mode 1, occ + seg[s] nb).
void launch_mu_newton_multi(const double* eps, const double* w_norm, const int* seg, int S,
                            int nk_max, int nb, const double* n_target, double T, int maxiter,
                            bool mu0_from_state, const MuState* warm, MuState* state,
                            double* ws, unsigned active, hipStream_t st, double* occ = nullptr);

// The Python surface's Newton (src/cpp_hf/utils.py:311-360) on raw weights:
// one workgroup, result in *mu_out.
void launch_mu_newton_py(const double* eps, const double* w, int nk, int nb, double n_e,
                         double T, int maxiter, double* mu_out, hipStream_t st);

// occupations from state->mu.
//   mode 0 (SCF): occ = expit(-(eps - mu) / (T + 1e-12))
//   mode 1 (DM):  p   = expit((mu - eps) / T)
void launch_occupations(const double* eps, size_t n, const MuState* state, double T, int mode,
                        double* out, hipStream_t st);
// launch_occupations for S solves of n values each stored one after another,
// each with its own state (solves whose bit in active is clear are skipped);
// with seg (device, S + 1 k offsets) solve s has the nb values of its k
// points seg[s] .. seg[s + 1] - 1 and n is the largest solve's count.
void launch_occupations_batch(const double* eps, size_t n, int S, unsigned active,
                              const MuState* states, double T, int mode, double* out,
                              hipStream_t st, const int* seg = nullptr, int nb = 0);

// level shift: out = eps + shift * (1 - expit(-(eps - mu)/(T + 1e-12)))
void launch_level_shift(const double* eps, size_t n, const MuState* state, double T,
                        double shift, double* out, hipStream_t st);

// ---- dense per-k products (dense.hip) --------------------------------------
// P_k = herm(V_k diag(f_k) V_k^H) when hermitize, else V diag(f) V^H.
void launch_density(int nk, int nb, const cd* V, const double* f, cd* P, bool hermitize,
                    hipStream_t st);
// The same with the occupations formed from eigenvalues and mu as
// launch_occupations does (mode 0 SCF, mode 1 DM), without storing them.
void launch_density_occ(int nk, int nb, const cd* V, const double* eps, const MuState* mus,
                        double T, int mode, cd* P, bool hermitize, hipStream_t st);
// launch_density over the solves of a batch: solve s owns the k points
// seg[s] .. seg[s + 1] - 1 of nk_total (seg on the device, nk_max the
// largest count); solves whose bit in active is clear are skipped.
void launch_density_seg(const int* seg, int S, int nk_max, int nk_total, unsigned active,
                        int nb, const cd* V, const double* f, cd* P, bool hermitize,
                        hipStream_t st);
// in-place hermitisation of (nk, nb, nb)
void launch_hermitize(int nk, int nb, cd* M, hipStream_t st);

}  // namespace hfgpu
