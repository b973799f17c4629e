/* ORACLE — test infrastructure only: C ABI of the CPU restatement of the
 * reference hot path (see hfcpu.cpp).  Only tests/, bench.py's cpu_baseline
 * leg and __graft_entry__.smoke() load it, as a checker or baseline. */
#ifndef HFCPU_H
#define HFCPU_H

#include "../include/hf_desc.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* hfcpu_last_error(void);
int hfcpu_num_threads(void);
int hfcpu_fft2d(double* buf, size_t nk1, size_t nk2, size_t batch, int inverse);
int hfcpu_selfenergy_fft_full(const double* VR, const double* P, double* sigma, size_t nk1,
                              size_t nk2, size_t nb, size_t dv1, size_t dv2,
                              int apply_ifftshift, int hcp);
int hfcpu_eigh_batched(const double* M, double* w, double* V, size_t nk, size_t nb);
int hfcpu_eigh_block_sizes(const double* M, double* w, double* V, size_t nk, size_t nb,
                           const size_t* sizes, size_t nsizes, int sort);
int hfcpu_max_offblock_sizes(const double* M, size_t nk, size_t nb, const size_t* sizes,
                             size_t nsizes, double* out);
int hfcpu_find_mu_bisection(const double* bands, const double* weights, size_t nk, size_t nb,
                            double n_e, double T, int maxiter, double* mu);
int hfcpu_find_mu_newton(const double* bands, const double* weights, size_t nk, size_t nb,
                         double n_e, double T, int maxiter, double* mu);
int hfcpu_solve_mu_inloop(const double* eps, const double* w_norm, size_t nk, size_t nb,
                          double n_target_norm, double mu0, double T, int maxiter, double* mu);
int hfcpu_validate_electron_count(const double* weights, size_t nk, size_t nb, double n_e);
int hfcpu_build_fock_apply(const hf_kernel_desc* kd, const double* P, double* Sigma, double* Hh,
                           double* F, double* E, hf_project_fn pfn, void* puser);
int hfcpu_density_from_fock(const double* F, size_t nk, size_t nb, const double* w, double n_e,
                            double T, double level_shift, double* P_new, double* mu);
int hfcpu_cayley_spectral_setup(const double* d, double* V, double* lam, size_t nk, size_t nb);
int hfcpu_cayley_unitary_from_spectrum(const double* V, const double* lam, double tau,
                                       double* U, size_t nk, size_t nb);
int hfcpu_diag_UFU_from_spectrum(const double* V, const double* Fe, const double* lam,
                                 double tau, double* diag, size_t nk, size_t nb);
int hfcpu_solve_scf(const hf_kernel_desc* kd, const double* P0, double n_e,
                    const hf_scf_config* cfg, hf_project_fn pfn, void* puser,
                    hf_scf_result* out);
int hfcpu_solve_dm(const hf_kernel_desc* kd, const double* P0, double n_e,
                   const hf_dm_config* cfg, hf_project_fn pfn, void* puser, hf_dm_result* out);
int hfcpu_occupation_entropy(const double* p, const double* w_norm, size_t nk, size_t nb,
                             double* S);
int hfcpu_resample_kgrid_2d(const double* x, double* out, size_t nk_old, size_t nk_new,
                            size_t inner);

#ifdef __cplusplus
}
#endif
#endif
