/*
 * Plain-C descriptors shared by the GPU library (libhfgpu) and the CPU oracle
 * (liboracle_hfcpu).  They carry exactly the information of the reference's
 * `kernel_args` dict (src/cpp_hf/solver.py:87-103) and of its two solver
 * configs (cpp/include/cpp_hf/solver_dm.hpp:22-63,
 * cpp/include/cpp_hf/solver_scf.hpp:20-50), as sizes, flags and raw pointers.
 *
 * Complex arrays are interleaved (re, im) doubles in C order, so a
 * complex128 numpy array can be passed as `const double*` without a copy.
 */
#ifndef HF_DESC_H
#define HF_DESC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by every entry point. */
enum hf_status {
    HF_OK = 0,
    HF_EINVAL = 1,   /* bad argument: maps to Python ValueError     */
    HF_ERUNTIME = 2, /* runtime failure: maps to Python RuntimeError */
    HF_EDEVICE = 3   /* HIP error: maps to Python RuntimeError       */
};

/*
 * Superlattice mode of a kernel (reference: the superlattice fields of
 * HFKernel, kernel.hpp:36-69, filled by inject_superlattice_into_kernel,
 * cpp_hf_native.cpp:104-165).  With fock_active the exchange is the
 * extended-grid convolution of selfenergy_superlattice_streamed
 * (superlattice.hpp:53-246); with hartree_active the Hartree term is the
 * k-independent G-basis shift of fock.hpp:66-173, added to Sigma.
 *   V_lag_fft         : (N_ext_x, N_ext_y) complex
 *   V_lag_fft_orbital : (N_ext_x, N_ext_y, dim_orb, dim_orb) complex or NULL
 *   g_a_off           : (n_G, 2); pair_i, pair_j: (n_pairs); pair_start:
 *                       (n_delta + 1); pair_to_delta: (n_G, n_G)
 *   HH_GG             : (n_G, n_G) real; HH_GG_orbital: (n_G, n_G, dim_orb,
 *                       dim_orb) real or NULL
 * n_G * dim_orb must equal the kernel's nb.
 */
typedef struct hf_superlattice_desc {
    int fock_active;
    int hartree_active;
    double hartree_degeneracy;
    size_t n_G, dim_orb, n_delta, N_ext_x, N_ext_y, n_pairs;
    const double* V_lag_fft;
    const double* V_lag_fft_orbital;
    const int64_t* g_a_off;
    const int64_t* pair_i;
    const int64_t* pair_j;
    const int64_t* pair_start;
    const int64_t* pair_to_delta;
    const double* HH_GG;
    const double* HH_GG_orbital;
} hf_superlattice_desc;

/*
 * Hartree-Fock kernel description (reference: HFKernel, kernel.hpp:10-34,
 * filled by make_kernel, cpp_hf_native.cpp:42-102).  All pointers are host
 * pointers, borrowed for the duration of the call that receives them.
 *   h, refP     : (nk1, nk2, nb, nb) complex
 *   VR          : (nk1, nk2, dv1, dv2) complex, dv = 1 or nb
 *   w2d         : (nk1, nk2) real
 *   HH          : (nb, nb) real
 *   contact_g   : (n_contact,) real; contact_Oi/Oj: (n_contact, nb, nb) complex
 */
typedef struct hf_kernel_desc {
    size_t nk1, nk2, nb, dv1, dv2, n_contact;
    int include_hartree;
    int include_exchange;
    int exchange_hcp;
    int has_refP;
    double T;
    double weight_sum;
    const double* h;
    const double* VR;
    const double* refP;
    const double* w2d;
    const double* HH;
    const double* contact_g;
    const double* contact_Oi;
    const double* contact_Oj;
    const hf_superlattice_desc* superlattice; /* NULL: the mesh exchange */
} hf_kernel_desc;

/*
 * Symmetry projector on the device (no counterpart in the reference, whose
 * ProjectFn is a host callback; the averages are those of make_project_fn,
 * symmetry.py:81-117).  On a (nk1, nk2, nb, nb) complex array A:
 *   with n_spatial > 0: A_k <- (sum_g g A_k g^H + sum_s s A_-k s^H) /
 *       (n_unitary' + n_spatial), where the unitary set is the identity
 *       alone (n_unitary' = 1) when n_unitary == 0;
 *   else with n_unitary > 0: A_k <- sum_g g A_k g^H / n_unitary;
 *   then with tr_U: A_k <- (A_k + U conj(A_-k) U^H) / 2.
 * -k maps index i to (-i) mod n (mod != 0) or n - 1 - i (flip) along the
 * axes whose bit is set in axes (bit 0: k1, bit 1: k2).
 *   unitary : (n_unitary, nb, nb) complex;  spatial : (n_spatial, nb, nb)
 *   tr_U    : (nb, nb) complex or NULL
 */
typedef struct hf_symmetry_desc {
    size_t nb;
    size_t n_unitary;
    const double* unitary;
    size_t n_spatial;
    const double* spatial;
    int spatial_mod;
    int spatial_axes;
    const double* tr_U;
    int tr_mod;
    int tr_axes;
} hf_symmetry_desc;

/* Reference: SCFConfig, solver_scf.hpp:20-50. acceleration: 0 linear, 1 diis, 2 oda. */
typedef struct hf_scf_config {
    size_t max_iter;
    double density_tol;
    double comm_tol;
    double mixing;
    double level_shift;
    int acceleration;
    size_t diis_size;
    size_t diis_start;
    double diis_damping;
    double trust_radius;
    const size_t* block_sizes; /* may be NULL when n_blocks == 0 */
    size_t n_blocks;
    /* Mixed precision (no counterpart in the reference): exchange_fp32 != 0
     * runs the Fock builds with the FP32 exchange workspace while the
     * density residual is above fp32_switch (1e-6 when <= 0), then FP64
     * builds to convergence (the FP64 correction phase).  0: FP64 only. */
    int exchange_fp32;
    double fp32_switch;
    /* Device symmetry projector in place of the project_fn callback (no
     * counterpart in the reference); NULL: none.  Passing both is an error. */
    const hf_symmetry_desc* symmetry;
} hf_scf_config;

/* Reference: SolverConfig, solver_dm.hpp:22-63.  optimizer 0: Riemannian CG
 * (solver_dm.hpp:392-653); 1: trust-region Newton (solver_rtr.hpp:162-459)
 * with tr_delta0, tr_cg_max and the deflation bias around n_deflation
 * densities (each nk1 x nk2 x nb x nb complex, interleaved) when
 * deflation_sigma > 0.  Zero-initialised trailing fields select CG without
 * deflation. */
typedef struct hf_dm_config {
    size_t max_iter;
    double tol_E;
    double tol_grad;
    size_t plateau_window;
    double denom_scale;
    double max_step;
    size_t cg_restart;
    double bt_shrink;
    size_t bt_max;
    int mu_maxiter;
    const size_t* block_sizes;
    size_t n_blocks;
    int optimizer;
    double tr_delta0;
    size_t tr_cg_max;
    const double* deflation_targets; /* may be NULL when n_deflation == 0 */
    size_t n_deflation;
    double deflation_sigma;
    double deflation_length;
    /* Mixed precision for the CG loop (no counterpart in the reference): as
     * in hf_scf_config, switching to FP64 builds once the gradient norm is
     * at most fp32_switch (1e-6 when <= 0).  The trust-region solver always
     * builds in FP64. */
    int exchange_fp32;
    double fp32_switch;
    /* as in hf_scf_config */
    const hf_symmetry_desc* symmetry;
} hf_dm_config;

/*
 * Host callback standing in for the reference's ProjectFn
 * (fock.hpp:21-23; wrapped from Python in cpp_hf_native.cpp:215-238).
 * It rewrites the (nk1, nk2, nb, nb) complex array M in place and returns 0,
 * or returns non-zero to abort the caller with HF_ERUNTIME.
 */
typedef int (*hf_project_fn)(void* user, double* M, size_t nk1, size_t nk2,
                             size_t nb);

/* SCF outputs (reference: SCFResult, solver_scf.hpp:52-62).  density/fock may
 * be NULL; hist_* hold max_iter entries, the first *iterations are valid. */
typedef struct hf_scf_result {
    double* density;
    double* fock;
    double energy;
    double mu;
    size_t iterations;
    int converged;
    double* hist_E;
    double* hist_density;
    double* hist_comm;
} hf_scf_result;

/* DM outputs (reference: DMResult, solver_dm.hpp:65-76).  Q/density/fock may
 * be NULL; p is required; hist_* hold max_iter entries. */
typedef struct hf_dm_result {
    double* Q;
    double* p;
    double* density;
    double* fock;
    double mu;
    double energy;
    size_t n_iter;
    int converged;
    double* hist_E;
    double* hist_grad;
    /* line-search trials evaluated over the CG solve (frozen free-energy
     * evaluations of solver_dm.hpp:561-589); 0 for the Newton solver.  No
     * counterpart in the reference's result. */
    size_t n_ls_trials;
} hf_dm_result;

#ifdef __cplusplus
}
#endif

#endif /* HF_DESC_H */
