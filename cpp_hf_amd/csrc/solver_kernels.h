// Launchers of the solver-loop kernels (Cayley retraction, line search,
// gradients, residuals) and of the k-grid resampler.
#pragma once

#include "common.h"

namespace hfgpu {

// ---- cayley.hip ------------------------------------------------------------
void launch_cayley_unitary(int nk, int nb, const cd* V, const double* lam, double tau, cd* U,
                           hipStream_t st);
void launch_diag_ufu(int nk, int nb, const cd* V, const cd* Fe, const double* lam, double tau,
                     double* diag, hipStream_t st);
void launch_resample_axis(const cd* x, cd* y, int n_old, int n_new, size_t inner, size_t outer,
                          hipStream_t st);

}  // namespace hfgpu
