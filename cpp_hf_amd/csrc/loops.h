// Solver-loop kernels (loops.hip): scalar slots and launchers.
#pragma once

#include "common.h"
#include "kernels.h"

namespace hfgpu {

constexpr int kMaxBt = 16;  // backtracking trials evaluated in one pass

// Slots of the solver's device scalar array.
enum ScSlot : int {
    SC_E = 0,
    SC_GN2 = 1,
    SC_EPS_SQ = 2,
    SC_PR1 = 3,  // 3..8: the six PR+ inner products
    SC_BETA = 9,
    SC_DQN2 = 10,
    SC_E0 = 11,
    SC_S0 = 12,
    SC_TAU0 = 13,
    SC_TAU = 14,
    SC_OMEGA0 = 15,
    SC_ACCEPTED = 16,
    SC_RES = 17,  // 17, 18: SCF weighted |delta|^2, |[F, P]|^2
    SC_DOT = 19,
    SC_E2 = 20,
    SC_TRIAL = 24,  // 24 .. 24 + 2 kMaxBt
    SC_COUNT = 64
};

struct LoopScratch {
    double* partial = nullptr;  // kLoopPartialDoubles per solve
    double* sc = nullptr;       // SC_COUNT doubles per solve
    // A batch of S solves of one shape whose arrays are stored solve after
    // solve: workgroup row s (blockIdx.y; blockIdx.x of a finalising launch)
    // works on solve s, whose arrays, scalar slots and partials start s
    // per-solve strides in.  Solves whose bit in active is clear are
    // skipped.  S = 1 is a single solve.
    int S = 1;
    unsigned active = ~0u;
    // Solves of different k counts: seg (device, S + 1 ints) holds the
    // prefix offsets of the solves' k ranges, solve s owning k points
    // seg[s] .. seg[s + 1] - 1 of arrays stored back to back; every launch is
    // sized for the largest solve and each solve keeps the workgroup
    // decomposition it has alone.  nullptr: S solves of nk points each.
    const int* seg = nullptr;
};
constexpr size_t kLoopPartialDoubles = 1024 * 2 * kMaxBt;
constexpr int kMaxBatch = 32;  // bits of LoopScratch::active

// out[s] = sum over nparts workgroups of partial[b * ns + s] (fixed order).
void launch_finalize_sums(const double* partial, int nparts, int ns, double* out,
                          hipStream_t st);

void launch_orbital_grad(int nk, int nb, const cd* Q, cd* F, const double* p, const double* w,
                         const MuState* mus, double T_r, bool grad, double* eps, cd* G,
                         double* g_p, LoopScratch& ls, hipStream_t st);
void launch_precond(int nk, int nb, const cd* G, const cd* Gp, const double* eps,
                    const double* p, const double* g_p, const double* g_pp, const double* w,
                    double T_r, double denom_scale, cd* dQ, double* h_p, LoopScratch& ls,
                    hipStream_t st);
void launch_cg_combine(int nk, int nb, bool restart, cd* dQ, const cd* dQp, const double* h_p,
                       const double* d_pp, double* d_p, const double* eps, const double* p,
                       const double* w, double T_r, LoopScratch& ls, hipStream_t st);
void launch_linesearch(int nk, int nb, const cd* V, const double* lam, const cd* Ft, cd* Fe,
                       const double* p, const double* d_p, const double* w, double T_r,
                       double max_step, double shrink, int nbt, LoopScratch& ls,
                       hipStream_t st);
// Q_out (default Q itself) receives the rotated orbitals.
void launch_retract(int nk, int nb, cd* Q, const cd* V, const double* lam, const cd* Fe,
                    double* eps_work, bool cayley, LoopScratch& ls, hipStream_t st,
                    cd* Q_out = nullptr);
void launch_scf_residuals(int nk, int nb, const cd* F, const cd* D, const cd* Pn, cd* delta,
                          cd* comm, const double* w, LoopScratch& ls, double* out,
                          hipStream_t st);
void launch_wdot(int nk, int m, const cd* X, const cd* Y, const double* w, LoopScratch& ls,
                 double* out, hipStream_t st);

// Up to kMaxDots arrays (or coefficients) passed by value to one launch.
constexpr int kMaxDots = 16;
struct CdPtrs {
    const cd* p[kMaxDots];
};
struct DotCoefs {
    double c[kMaxDots];
};

// out[d] = launch_wdot(X_d, Y) for d < nd <= kMaxDots, bit for bit, in one
// pass over Y; out holds kMaxDots slots.
void launch_wdot_multi(int nk, int m, const CdPtrs& X, int nd, const cd* Y, const double* w,
                       LoopScratch& ls, double* out, hipStream_t st);

// out = sum_{i < nx} c_i X_i (starting from zero, or from out with
// accumulate), then, with Fcur, out = a out + b Fcur: the arithmetic of a
// zero fill, one launch_axpy_c per term and an axpby, in one pass.
void launch_lincomb(size_t n, const CdPtrs& X, const DotCoefs& c, int nx, bool accumulate,
                    double a, double b, const cd* Fcur, cd* out, hipStream_t st);

}  // namespace hfgpu
