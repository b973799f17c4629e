/*
 * libhfgpu — C ABI of the MI355X (gfx950) Hartree-Fock hot path.
 *
 * Every entry point here replaces one function of the reference's pybind11
 * module `cpp_hf._native` (cpp/cpp_hf_native.cpp) or the C++ core it calls;
 * the citation on each names it.  The pybind11 module
 * `cpp_hf_amd._native` binds these with the reference's Python signatures,
 * and any other host language (ctypes, cffi, JNI, cgo) can bind them too:
 * plain pointers and sizes, no torch or HIP types in host-pointer calls.
 *
 * Conventions
 *  - Complex arrays are interleaved (re, im) doubles in C order (numpy
 *    complex128 memory).  Real arrays are doubles.
 *  - Host-pointer entry points copy inputs to the GPU, run, copy outputs back
 *    and synchronise before returning.  Output buffers are caller-allocated.
 *  - Return value: HF_OK (0) or an hf_status code; hfgpu_last_error() gives
 *    the message of the last failure on the calling thread.  HF_EINVAL maps
 *    to Python ValueError (the reference's std::invalid_argument), the others
 *    to RuntimeError.
 *  - One GPU: the library runs on the current HIP device of the caller.
 */
#ifndef HFGPU_H
#define HFGPU_H

#include <stddef.h>

#include "hf_desc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing call on this thread ("" if none). */
const char* hfgpu_last_error(void);

/* Number of visible GPUs (0 when none); never fails. */
int hfgpu_device_count(void);

/* Drop the cached FFT geometries of kernel-less self-energy calls.
 * Replaces: _native.clear_fft_plan_cache (cpp/cpp_hf_native.cpp:280-284) ->
 * FftPlanCache::clear (cpp/include/cpp_hf/selfenergy.hpp:39-42). */
void hfgpu_clear_plan_cache(void);

/* ------------------------------------------------------------------------
 * Device-resident kernel handle.
 * Replaces: make_kernel (cpp/cpp_hf_native.cpp:42-102) building an HFKernel
 * (cpp/include/cpp_hf/kernel.hpp:10-84).  Uploads h, VR, refP, w2d, HH and
 * the contact terms once; validates shapes with the same messages.
 * ---------------------------------------------------------------------- */
typedef struct hfgpu_kernel hfgpu_kernel;

int hfgpu_kernel_create(const hf_kernel_desc* desc, hfgpu_kernel** out);
int hfgpu_kernel_destroy(hfgpu_kernel* k);
/* 1 when the Hermitian-packed exchange (nb(nb+1)/2 channels) is usable for
 * Hermitian densities: scalar V with exchange_hcp, or (nb, nb) V whose
 * index-space kernel is real and symmetric in the band pair. */
int hfgpu_kernel_packable(const hfgpu_kernel* k);
/* Launch-shape autotuner (no counterpart in the reference): times the Fock
 * build of this kernel (Hermitian or general density) over the two-step
 * exchange's launch shapes and keeps the fastest for this kernel and for
 * later kernels of the same shape in the process (HFGPU_AUTOTUNE=1 does this
 * on every new geometry).  shape: 5 ints (tiles per row workgroup, 0 = whole
 * row; tiles per column workgroup; wavefronts per tile in the forward-row,
 * column and inverse-row passes); best_us: microseconds per build (0 when
 * the mesh has no two-step transform).  Results of builds do not change. */
int hfgpu_kernel_autotune(hfgpu_kernel* k, int p_hermitian, int* shape, double* best_us);
/* Matrices-per-wavefront autotuner of the batched eigh (no counterpart in the
 * reference): times nmat Hermitian test matrices of nb bands (1..16) with
 * full and with half, quarter and eighth wavefronts of matrices, keeps the
 * fastest for later eigh calls of that (nb, nmat) in the process (the
 * solvers' included; HFGPU_AUTOTUNE=1 does this for every new kernel's nk),
 * mpw: matrices per wavefront kept, best_us: microseconds per call.  nb
 * 17..32: the fixed two per wavefront, 0 us.  Results of later calls do not
 * change. */
int hfgpu_eigh_autotune(size_t nmat, size_t nb, int* mpw, double* best_us);
/* Measurement of the Fock build's dataflow floor (no counterpart in the
 * reference): three dependent streaming launches per build that read and
 * write exactly the bytes of this kernel's two-step exchange passes (P into
 * the tile-major workspace; workspace and interaction into the workspace;
 * workspace and packed h into Sigma and F) with no arithmetic; us:
 * microseconds per build (best of three runs of reps builds), bytes: bytes
 * one build moves.  HF_EINVAL when the mesh has no two-step transform. */
int hfgpu_kernel_dataflow_floor(hfgpu_kernel* k, int p_hermitian, int reps, double* us,
                                double* bytes);

/* ------------------------------------------------------------------------
 * Exchange self-energy.
 * Replaces: _native.selfenergy_fft_full (cpp/cpp_hf_native.cpp:246-278) ->
 * selfenergy_fft_full / selfenergy_fft_full_hcp / ifftshift_2d_batch
 * (cpp/include/cpp_hf/selfenergy.hpp:91-190).
 *   VR: (nk1, nk2, dv1, dv2), P and sigma: (nk1, nk2, nb, nb).
 * ---------------------------------------------------------------------- */
int hfgpu_selfenergy_fft_full(const double* VR, const double* P, double* sigma, size_t nk1,
                              size_t nk2, size_t nb, size_t dv1, size_t dv2,
                              int apply_ifftshift, int hermitian_channel_packing);

/* ------------------------------------------------------------------------
 * Symmetry projector (no counterpart in the reference, whose ProjectFn is a
 * host callback; the averages are make_project_fn's, symmetry.py:81-117).
 * For callers who want the device averaging outside a solve; the solvers
 * take the same description through hf_scf_config / hf_dm_config.symmetry,
 * which keeps the projections on the device inside their loops.
 *   M: (nk1, nk2, nb, nb) complex host array, rewritten in place.
 * ---------------------------------------------------------------------- */
int hfgpu_project_symmetry(const hf_symmetry_desc* sym, double* M, size_t nk1, size_t nk2,
                           size_t nb);

/* ------------------------------------------------------------------------
 * Superlattice exchange on the extended momentum grid.
 * Replaces: _native.selfenergy_superlattice_streamed
 * (cpp/cpp_hf_native.cpp:719-790) -> selfenergy_superlattice_streamed
 * (cpp/include/cpp_hf/superlattice.hpp:50-246).
 *   rho, sigma     : (nkx, nky, n_G, dim_orb, n_G, dim_orb) complex
 *   VR_fft         : (N_ext_x, N_ext_y) complex lag-kernel transform
 *   VR_fft_orbital : (N_ext_x, N_ext_y, dim_orb, dim_orb) complex or NULL
 *   g_a_off        : (n_G, 2) tile offsets in units of (nkx, nky)
 *   pair_i, pair_j : (n_pairs) pairs grouped by channel, pair_start
 *                    (n_delta + 1) the CSR row pointer
 * n_pairs is an addition of this interface: the length of pair_i / pair_j,
 * against which pair_start is checked.  Every index and every tile is
 * checked against the grid (HF_EINVAL otherwise).
 * ---------------------------------------------------------------------- */
int hfgpu_selfenergy_superlattice_streamed(
    const double* rho, double* sigma, const double* VR_fft, const double* VR_fft_orbital,
    const int64_t* g_a_off, const int64_t* pair_i, const int64_t* pair_j,
    const int64_t* pair_start, size_t n_pairs, size_t n_delta, size_t N_ext_x, size_t N_ext_y,
    size_t nkx, size_t nky, size_t n_G, size_t dim_orb, int hermitian_rho);

/* ------------------------------------------------------------------------
 * Batched Hermitian eigendecomposition (hermitises its input first).
 * Replaces: _native.eigh_batched (cpp/cpp_hf_native.cpp:287-312) ->
 * eigh_batched / eigh_one (cpp/include/cpp_hf/linalg.hpp:39-66).
 *   M, V: (nmat, nb, nb); w: (nmat, nb) ascending; V columns = eigenvectors.
 * ---------------------------------------------------------------------- */
int hfgpu_eigh_batched(const double* M, double* w, double* V, size_t nmat, size_t nb);

/* Block-diagonal eigh: each diagonal block of the given sizes diagonalised
 * on its own (off-block entries ignored); sort != 0 orders all eigenpairs
 * ascending, else each block's ascending pairs stay at the block's position.
 * Replaces: _native.eigh_block_sizes (cpp/cpp_hf_native.cpp:314-343) ->
 * eigh_block_sizes_batched (cpp/include/cpp_hf/linalg.hpp:72-113). */
int hfgpu_eigh_block_sizes(const double* M, double* w, double* V, size_t nmat, size_t nb,
                           const size_t* sizes, size_t nsizes, int sort);

/* max |M_ij| over the off-block entries (blocks of the given sizes) of all
 * nmat matrices.  Replaces: _native.max_offblock_sizes
 * (cpp/cpp_hf_native.cpp:345-358) -> max_offblock_sizes (linalg.hpp:115-130). */
int hfgpu_max_offblock_sizes(const double* M, size_t nmat, size_t nb, const size_t* sizes,
                             size_t nsizes, double* out);

/* ------------------------------------------------------------------------
 * Chemical potential by bisection (maxiter <= 0 means 30).
 * Replaces: _native.find_mu_bisection (cpp/cpp_hf_native.cpp:361-377) ->
 * find_mu_bisection (cpp/include/cpp_hf/utils.hpp:49-75).
 * ---------------------------------------------------------------------- */
int hfgpu_find_mu_bisection(const double* bands, const double* weights, size_t nk, size_t nb,
                            double n_e, double T, int maxiter, double* mu);

/* Chemical potential by Newton with bisection safeguard on raw weights,
 * started at the middle of [emin - span, emax + span], with the bracket
 * midpoint returned when the final count misses n_e by more than 1e-12.
 * Replaces: cpp_hf.utils._find_mu_newton (src/cpp_hf/utils.py:311-360),
 * reached through find_chemical_potential(method="newton") (utils.py:254-274). */
int hfgpu_find_mu_newton(const double* bands, const double* weights, size_t nk, size_t nb,
                         double n_e, double T, int maxiter, double* mu);

/* In-loop Newton/bisection chemical potential on normalised weights.
 * Replaces: solve_mu_inloop (cpp/include/cpp_hf/utils.hpp:79-116). */
int hfgpu_solve_mu_inloop(const double* eps, const double* w_norm, size_t nk, size_t nb,
                          double n_target_norm, double mu0, double T, int maxiter,
                          double* mu);

/* ------------------------------------------------------------------------
 * Fock build + energy at density P.
 * Replaces: _native.build_fock_apply (cpp/cpp_hf_native.cpp:380-415) ->
 * build_fock + hf_energy (cpp/include/cpp_hf/fock.hpp:302-351).
 *   Outputs Sigma, Hh (dense Hartree diagonal), F: (nk1, nk2, nb, nb); E.
 *   project (nullable) is applied to F, followed by hermitisation, like the
 *   reference's ProjectFn (fock.hpp:295-298).
 * ---------------------------------------------------------------------- */
int hfgpu_build_fock_apply(hfgpu_kernel* k, const double* P, double* Sigma, double* Hh,
                           double* F, double* E, hf_project_fn project, void* user);

/* Device-pointer Fock build for callers that keep P resident in HBM (torch
 * tensors, benchmarks).  Replaces: build_fock_compact (fock.hpp:34-299).
 *   dP, dSigma, dF: device (nk1, nk2, nb, nb) complex; dHdiag: device (nb).
 *   p_hermitian != 0 asserts P == P^H (enables the packed exchange).
 *   stream: the hipStream_t to run on (NULL: the null stream).  Asynchronous.
 *   With HFGPU_GRAPH=1 in the environment when the kernel is created, the
 *   launches are captured into a hipGraph on the first call with these
 *   pointers and replayed afterwards. */
int hfgpu_build_fock_device(hfgpu_kernel* k, const void* dP, void* dSigma, void* dF,
                            void* dHdiag, int p_hermitian, void* stream);

/* Mixed-precision Fock build: the same outputs as hfgpu_build_fock_apply /
 * hfgpu_build_fock_device with the exchange workspace and the interaction
 * held in complex float between the transform passes (the transforms'
 * arithmetic, P, Sigma, F, the Hartree and contact terms stay FP64), so
 * Sigma carries a relative error of order 1e-7.  No counterpart in the
 * reference (FP64 only); for the early iterations of a solve, whose FP64
 * correction phase finishes it (hf_scf_config / hf_dm_config
 * exchange_fp32).  Meshes without the two-step transforms (both axes in
 * 12, 16, 24, 32, 36, 48, 60, 64, 72, 96, 120, 144, 192) run the FP64
 * workspace; hfgpu_kernel_mixed_capable says which applies. */
int hfgpu_build_fock_mixed(hfgpu_kernel* k, const double* P, double* Sigma, double* Hh,
                           double* F, double* E, hf_project_fn project, void* user);
int hfgpu_build_fock_mixed_device(hfgpu_kernel* k, const void* dP, void* dSigma, void* dF,
                                  void* dHdiag, int p_hermitian, void* stream);
int hfgpu_kernel_mixed_capable(hfgpu_kernel* k, int p_hermitian);

/* Vector-Jacobian product of the Fock build.  No counterpart in the
 * reference, which has no derivatives; for callers that differentiate
 * through build_fock (autograd, parameter gradients at a converged state).
 * It is the adjoint of build_fock_compact + build_fock (fock.hpp:34-321)
 * with the interaction transform VR = phase * fftn(w Vq) of
 * HartreeFockKernel (src/cpp_hf/problem.py:84-107):
 *   P: (nk1, nk2, nb, nb) complex density;
 *   cot_sigma, cot_hartree, cot_F: cotangents of the outputs Sigma, Hh and F
 *     (same shape; NULL = zero);
 *   gP, gh: (nk1, nk2, nb, nb) complex gradients with respect to P and h;
 *   gVq: (nk1, nk2, dv1, dv2) complex gradient with respect to coulomb_q
 *     (its real part for a real interaction); gHH: (nb, nb) real.
 * Convention dL = Re sum conj(g) dX.  gVq, gh, gHH may be NULL.  The
 * transforms are recomputed, not stored. */
int hfgpu_build_fock_vjp(hfgpu_kernel* k, const double* P, const double* cot_sigma,
                         const double* cot_hartree, const double* cot_F, double* gP,
                         double* gVq, double* gh, double* gHH);
/* The same on device pointers and a hipStream_t (NULL: the null stream);
 * returns after the stream has drained the call's work. */
int hfgpu_build_fock_vjp_device(hfgpu_kernel* k, const void* dP, const void* dGs,
                                const void* dGh, const void* dGF, void* dgP, void* dgVq,
                                void* dgh, void* dgHH, void* stream);

/* Device-pointer energy: E = hf_energy_with_hartree_diag (fock.hpp:355-383)
 * written to device dE (one double).  stream as above.  Asynchronous. */
int hfgpu_energy_device(hfgpu_kernel* k, const void* dP, const void* dSigma,
                        const void* dHdiag, void* dE, void* stream);

/* ------------------------------------------------------------------------
 * Public building blocks of the reference's Python layer, on the GPU.
 * ---------------------------------------------------------------------- */

/* E = sum_k w_k Re[Tr(h_k P_k) + 1/2 Tr((Sigma_k + H_k)(P_k - refP_k))].
 * Replaces: fock.hf_energy (src/cpp_hf/fock.py:91-107) and hf_energy
 * (cpp/include/cpp_hf/fock.hpp:324-351).  H has n_H k-blocks: 1 (broadcast,
 * the (1, 1, nb, nb) Hartree matrix) or nk; refP may be NULL. */
int hfgpu_hf_energy(const double* h, const double* P, const double* Sigma, const double* H,
                    size_t n_H, const double* w, const double* refP, size_t nk, size_t nb,
                    double* E);

/* S = -sum_k w_k sum_b [p ln p + (1 - p) ln(1 - p)], p clipped to
 * [1e-14, 1 - 1e-14].  Replaces: occupation_entropy
 * (cpp/include/cpp_hf/fock.hpp:385-401; src/cpp_hf/fock.py:110-113). */
int hfgpu_occupation_entropy(const double* p, const double* w_norm, size_t nk, size_t nb,
                             double* S);

/* P = herm(V diag(f) V^H) with V, eps = eigh(herm(F)), mu by bisection with
 * maxiter steps (<= 0: 30) and f = fermidirac(eps - mu, T).
 * Replaces: density_from_fock (cpp/include/cpp_hf/solver_scf.hpp:96-154,
 * level_shift 0) and utils.density_matrix_from_fock
 * (src/cpp_hf/utils.py:363-389). */
int hfgpu_density_from_fock(const double* F, const double* weights, size_t nk, size_t nb,
                            double n_e, double T, int maxiter, double* P, double* mu);

/* ------------------------------------------------------------------------
 * Solvers.
 * Replaces: _native.solve_scf (cpp/cpp_hf_native.cpp:418-485) -> solve_scf
 * (cpp/include/cpp_hf/solver_scf.hpp:156-444), and _native.solve_dm
 * (cpp/cpp_hf_native.cpp:488-582) -> solve_dm
 * (cpp/include/cpp_hf/solver_dm.hpp:392-653) for cfg->optimizer != 1, solve_rtr
 * (cpp/include/cpp_hf/solver_rtr.hpp:162-459, trust-region Newton with the
 * deflation bias) for cfg->optimizer == 1.
 * ---------------------------------------------------------------------- */
int hfgpu_solve_scf(hfgpu_kernel* k, const double* P0, double n_e, const hf_scf_config* cfg,
                    hf_project_fn project, void* user, hf_scf_result* out);
int hfgpu_solve_dm(hfgpu_kernel* k, const double* P0, double n_e, const hf_dm_config* cfg,
                   hf_project_fn project, void* user, hf_dm_result* out);

/* n (1..32) direct-minimisation solves advanced together, one per kernel
 * (n distinct handles: a handle appearing twice is HF_EINVAL):
 * every kernel with the same band count and temperature (meshes may
 * differ), one shared configuration (CG optimizer; no blocks, symmetry,
 * deflation or FP32 builds).  Each launch of an iteration covers all solves
 * (the exchange passes one launch per mesh) and each solve stops on its own
 * test; outs[i] equals what hfgpu_solve_dm
 * returns for kernel i, P0s[i], n_es[i] bit for bit.  No counterpart in the
 * reference, which runs sweeps one solve after another or in separate
 * processes (examples/tetralayer_quartermetal/deprecated/quartermetal_grid.py:
 * 26): for parameter sweeps, where one small solve leaves most of the GPU
 * idle. */
int hfgpu_solve_dm_batch(hfgpu_kernel* const* ks, size_t n, const double* const* P0s,
                         const double* n_es, const hf_dm_config* cfg, hf_dm_result* outs);

/* density_from_fock on device pointers that allocates nothing: F (nk, nb,
 * nb) complex and weights (nk) in, P (nk, nb, nb) complex and mu (one
 * double) out, all device memory, on `stream` (a hipStream_t; NULL: the null
 * stream), asynchronous.  The eigenvectors, eigenvalues, occupations and the
 * bisection workspace live in a fixed 16 MiB device array of the library
 * (nk nb^2 up to about 10^6, e.g. 96 x 96 k with 8 bands); larger inputs are
 * rejected with HF_EINVAL.  The array is shared by every call on a GPU:
 * calls that overlap on one GPU race.  Results equal hfgpu_density_from_fock
 * bit for bit.  No counterpart in the reference: for callers that cannot
 * allocate device memory (HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). */
int hfgpu_density_from_fock_noalloc(const double* dF, const double* dweights, size_t nk,
                                    size_t nb, double n_e, double T, int maxiter, double* dP,
                                    double* dmu, void* stream);

/* ------------------------------------------------------------------------
 * Cayley spectral helpers (test hooks of the reference).
 * Replaces: _native._cayley_spectral_setup / _cayley_unitary_from_spectrum /
 * _diag_UFU_from_spectrum (cpp/cpp_hf_native.cpp:585-651) ->
 * cpp/include/cpp_hf/solver_dm.hpp:206-222, 258-272, 277-303.
 * ---------------------------------------------------------------------- */
int hfgpu_cayley_spectral_setup(const double* d, double* V, double* lam, size_t nk,
                                size_t nb);
int hfgpu_cayley_unitary_from_spectrum(const double* V, const double* lam, double tau,
                                       double* U, size_t nk, size_t nb);
int hfgpu_diag_UFU_from_spectrum(const double* V, const double* Ft_eig, const double* lam,
                                 double tau, double* diag, size_t nk, size_t nb);

/* ------------------------------------------------------------------------
 * Periodic linear k-grid resampling (bit-exact index map).
 * Replaces: _native.resample_kgrid_2d (cpp/cpp_hf_native.cpp:819-840) ->
 * resample_kgrid_2d (cpp/include/cpp_hf/utils.hpp:120-176).
 *   x: (nk_old, nk_old, inner), out: (nk_new, nk_new, inner) complex.
 * ---------------------------------------------------------------------- */
int hfgpu_resample_kgrid_2d(const double* x, double* out, size_t nk_old, size_t nk_new,
                            size_t inner);

#ifdef __cplusplus
}
#endif

#endif /* HFGPU_H */
