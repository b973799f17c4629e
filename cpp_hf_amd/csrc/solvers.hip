// Solver loops on the GPU: Roothaan SCF (solver_scf.hpp:156-444) and
// direct minimisation with preconditioned Riemannian CG (solver_dm.hpp:392-653).
//
// All arrays live in HBM for the whole solve.  The host only issues kernels
// and, once per iteration, reads back a handful of scalars (energy, residual
// norms) to take the same stopping decisions as the reference.  Inside an
// iteration every data-dependent scalar (mu, beta, tau) is produced and
// consumed on the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hfgpu.h"
#include "hfgpu_core.h"
#include "loops.h"

namespace hfgpu {

namespace {

constexpr double TINY_REAL = 1.0e-30;

struct Projector {
    hf_project_fn fn = nullptr;
    void* user = nullptr;
    int nk1 = 0, nk2 = 0, nb = 0;
    size_t nd = 0;
    std::vector<double> host;

    explicit operator bool() const { return fn != nullptr; }
    // Host round trip: the callback sees and rewrites a host copy.
    void apply(cd* dM, hipStream_t st) {
        host.resize(2 * nd);
        d2h(host.data(), dM, nd * sizeof(cd), st);
        HFGPU_CHECK(hipStreamSynchronize(st));
        if (fn(user, host.data(), static_cast<size_t>(nk1), static_cast<size_t>(nk2),
               static_cast<size_t>(nb)) != 0)
            throw std::runtime_error("project_fn failed");
        h2d(dM, host.data(), nd * sizeof(cd), st);
    }
};

// electron-count check (utils.hpp:31-46)
void validate_electron_count(const std::vector<double>& w, size_t nb, double n_e) {
    double total = 0.0;
    for (double x : w) total += x;
    const double max_e = total * static_cast<double>(nb);
    const double tol = std::max(1e-8, 1e-6 * std::max(1.0, std::fabs(max_e)));
    if (!std::isfinite(n_e)) throw std::invalid_argument("n_electrons must be finite.");
    if (n_e < -tol || n_e > max_e + tol)
        throw std::invalid_argument("n_electrons=" + std::to_string(n_e) +
                                    " is outside the physically reachable range [0, " +
                                    std::to_string(max_e) + "].");
}

// block_sizes validation of the direct-minimisation solver
// (solver_dm.hpp:80-90).
void validate_block_sizes(const size_t* sizes, size_t n, size_t nb) {
    if (n == 0) return;
    size_t total = 0;
    for (size_t i = 0; i < n; ++i) {
        if (sizes[i] == 0) throw std::invalid_argument("block_sizes entries must be positive.");
        total += sizes[i];
    }
    if (total != nb) throw std::invalid_argument("block_sizes must sum to nb.");
}

// Block structure of a batch of nb x nb matrices (solver_dm.hpp:92-98): the
// off-block maximum at most 1e-10 + 1e-10 max |M|.  Reduced on the device,
// decided on the host (a one-time check at the start of a solve).
bool block_ok_host(const cd* M, int nmat, int nb, const BlockMap& bm, hipStream_t st) {
    DevBuf part(2 * kMaxOffblockBlocks * sizeof(double)), out(2 * sizeof(double));
    launch_max_offblock(nmat, nb, M, bm, part.as<double>(), out.as<double>(), st);
    double h[2];
    d2h(h, out.get(), sizeof(h), st);
    HFGPU_CHECK(hipStreamSynchronize(st));
    return h[0] <= 1.0e-10 + 1.0e-10 * h[1];
}

// Workspace shared by both solvers.
struct SolveWs {
    DevBuf mus, mu_partial, ls_partial, sc;
    LoopScratch ls;
    explicit SolveWs(hipStream_t st) {
        mus.alloc(sizeof(MuState));
        mu_partial.alloc(kMuPartialDoubles * sizeof(double));
        ls_partial.alloc(kLoopPartialDoubles * sizeof(double));
        sc.alloc(SC_COUNT * sizeof(double));
        fill_zero(mus.get(), sizeof(MuState), st);
        fill_zero(sc.get(), SC_COUNT * sizeof(double), st);
        ls.partial = ls_partial.as<double>();
        ls.sc = sc.as<double>();
    }
    MuState* mu() { return mus.as<MuState>(); }
};

// Full-pivot LU solve for the DIIS system; false when rank deficient.
bool fullpiv_solve(std::vector<double> B, std::vector<double> rhs, size_t n,
                   std::vector<double>& sol) {
    std::vector<size_t> colp(n);
    for (size_t i = 0; i < n; ++i) colp[i] = i;
    double maxpiv = 0.0;
    size_t rank = 0;
    const double thr = std::numeric_limits<double>::epsilon() * static_cast<double>(n);
    for (size_t s = 0; s < n; ++s) {
        size_t pr = s, pc = s;
        double best = -1.0;
        for (size_t i = s; i < n; ++i)
            for (size_t j = s; j < n; ++j)
                if (std::fabs(B[i * n + j]) > best) {
                    best = std::fabs(B[i * n + j]);
                    pr = i;
                    pc = j;
                }
        if (s == 0) maxpiv = best;
        if (best <= thr * maxpiv || best == 0.0) break;
        ++rank;
        if (pr != s) {
            for (size_t j = 0; j < n; ++j) std::swap(B[s * n + j], B[pr * n + j]);
            std::swap(rhs[s], rhs[pr]);
        }
        if (pc != s) {
            for (size_t i = 0; i < n; ++i) std::swap(B[i * n + s], B[i * n + pc]);
            std::swap(colp[s], colp[pc]);
        }
        for (size_t i = s + 1; i < n; ++i) {
            const double f = B[i * n + s] / B[s * n + s];
            for (size_t j = s; j < n; ++j) B[i * n + j] -= f * B[s * n + j];
            rhs[i] -= f * rhs[s];
        }
    }
    if (rank < n) return false;
    std::vector<double> y(n);
    for (size_t i = n; i-- > 0;) {
        double s = rhs[i];
        for (size_t j = i + 1; j < n; ++j) s -= B[i * n + j] * y[j];
        y[i] = s / B[i * n + i];
    }
    sol.assign(n, 0.0);
    for (size_t i = 0; i < n; ++i) sol[colp[i]] = y[i];
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// Roothaan SCF
// ---------------------------------------------------------------------------

namespace {

struct ScfState {
    GpuKernel& K;
    hipStream_t st;
    SolveWs ws;
    DevBuf eps, shifted, occ, V;
    Projector& proj;
    double n_e, level_shift;
    // block_sizes summing to nb with more than one entry (solver_scf.hpp:
    // 107-110; other values leave the dense eigh, without an error)
    bool blocks = false;
    BlockMap bm{};
    DevBuf bpart, bcond;

    ScfState(GpuKernel& K_, Projector& pr, double ne, double ls, const size_t* sizes,
             size_t nsizes)
        : K(K_), st(K_.stream), ws(K_.stream), proj(pr), n_e(ne), level_shift(ls) {
        const size_t nkb = K.nk * K.nb;
        eps.alloc(nkb * sizeof(double));
        shifted.alloc(nkb * sizeof(double));
        occ.alloc(nkb * sizeof(double));
        V.alloc(K.nd * sizeof(cd));
        size_t total = 0;
        for (size_t i = 0; i < nsizes; ++i) total += sizes[i];
        if (nsizes > 1 && total == static_cast<size_t>(K.nb)) {
            blocks = true;
            bm = block_map(static_cast<size_t>(K.nb), sizes, nsizes, true);
            bpart.alloc(2 * kMaxOffblockBlocks * sizeof(double));
            bcond.alloc(2 * sizeof(double));
        }
    }

    // density_from_fock (solver_scf.hpp:96-154).  With blocks, F's off-block
    // test is reduced on the device and read there by the eigh, which treats
    // the blocks separately (each block ascending at its position) only when
    // F passes it.
    void density_from_fock(const cd* F, cd* P_out) {
        const int nk = static_cast<int>(K.nk), nb = K.nb;
        const size_t nkb = K.nk * K.nb;
        if (blocks) {
            launch_max_offblock(nk, nb, F, bm, bpart.as<double>(), bcond.as<double>(), st);
            launch_eigh_blocks(nk, nb, 0, F, eps.as<double>(), V.as<cd>(), bm, false, st,
                               bcond.as<double>());
        } else {
            launch_eigh(nk, nb, 0, F, eps.as<double>(), V.as<cd>(), st);
        }
        launch_mu_bisection(eps.as<double>(), K.w2d.as<double>(), nk, nb, n_e, K.T, 0,
                            ws.mu(), ws.mu_partial.as<double>(), st);
        const double* e_used = eps.as<double>();
        if (level_shift != 0.0) {
            launch_level_shift(eps.as<double>(), nkb, ws.mu(), K.T, level_shift,
                               shifted.as<double>(), st);
            e_used = shifted.as<double>();
            launch_mu_bisection(e_used, K.w2d.as<double>(), nk, nb, n_e, K.T, 0, ws.mu(),
                                ws.mu_partial.as<double>(), st);
        }
        // (with level_shift == 0 the reference's second bisection repeats the
        // first on identical input, so its result is the same mu)
        launch_occupations(e_used, nkb, ws.mu(), K.T, 0, occ.as<double>(), st);
        launch_density(nk, nb, V.as<cd>(), occ.as<double>(), P_out, true, st);
    }

    // build_and_occupy (solver_scf.hpp:179-197); energy -> sc[E].
    void build_and_occupy(cd* d, cd* Sigma, cd* F, double* hdiag, cd* P_out, double* E_slot) {
        if (proj) proj.apply(d, st);
        launch_hermitize(static_cast<int>(K.nk), K.nb, d, st);
        K.build_fock(d, Sigma, F, hdiag, true, st);
        if (proj) {
            proj.apply(F, st);
            launch_hermitize(static_cast<int>(K.nk), K.nb, F, st);
        }
        K.energy(d, Sigma, hdiag, E_slot, st);
        density_from_fock(F, P_out);
        if (proj) proj.apply(P_out, st);
        launch_hermitize(static_cast<int>(K.nk), K.nb, P_out, st);
    }
};

__global__ void k_axpby(size_t n, double a, cd* __restrict__ y, double b, const cd* __restrict__ x) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const cd u = y[i], v = x[i];
        y[i] = cd{a * u.x + b * v.x, a * u.y + b * v.y};
    }
}

void launch_axpby(size_t n, double a, cd* y, double b, const cd* x, hipStream_t st) {
    hipLaunchKernelGGL(k_axpby, dim3(std::max(1u, std::min<unsigned>(cdiv(n, 256), 8192u))),
                       dim3(256), 0, st, n, a, y, b, x);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace

void solve_scf_gpu(GpuKernel& K, const double* P0, double n_e, const hf_scf_config& cfg,
                   Projector& proj, hf_scf_result* out) {
    if (cfg.acceleration < 0 || cfg.acceleration > 2)
        throw std::invalid_argument("SCFConfig.acceleration must be 'linear', 'diis', or 'oda'");
    const bool use_diis = cfg.acceleration == 1, use_oda = cfg.acceleration == 2;
    const size_t nd = K.nd, bytes = nd * sizeof(cd);
    const int nk = static_cast<int>(K.nk), nb = K.nb;
    hipStream_t st = K.stream;
    ScfState S(K, proj, n_e, cfg.level_shift, cfg.block_sizes, cfg.n_blocks);
    DevBuf density(bytes), Sigma(bytes), F(bytes), P_new(bytes), delta(bytes), hdiag(nb * 8);
    DevBuf comm(use_diis ? bytes : 0), Sigma_new(use_oda ? bytes : 0), F_new(use_oda ? bytes : 0),
        hdiag_new(use_oda ? nb * 8 : 0), pre(cfg.trust_radius > 0.0 ? bytes : 0),
        diff(cfg.trust_radius > 0.0 ? bytes : 0);
    std::vector<DevBuf> diis_F, diis_R;
    h2d(density.get(), P0, bytes, st);
    double* sc = S.ws.sc.as<double>();

    std::vector<double> hE, hD, hC;
    bool converged = false;
    size_t k = 0;
    const double ws = std::max(K.weight_sum, 1.0e-30);
    double hsc[SC_COUNT];
    MuState hmu;
    while (k < cfg.max_iter && !converged) {
        S.build_and_occupy(density.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(),
                           P_new.as<cd>(), sc + SC_E);
        launch_scf_residuals(nk, nb, F.as<cd>(), density.as<cd>(), P_new.as<cd>(),
                             delta.as<cd>(), use_diis ? comm.as<cd>() : nullptr,
                             K.w2d.as<double>(), S.ws.ls, sc + SC_RES, st);
        d2h(hsc, sc, sizeof(hsc), st);
        d2h(&hmu, S.ws.mu(), sizeof(MuState), st);
        HFGPU_CHECK(hipStreamSynchronize(st));
        const double E_iter = hsc[SC_E];
        const double d_res = std::sqrt(hsc[SC_RES] / ws);
        const double c_res = std::sqrt(hsc[SC_RES + 1] / ws);
        converged = (d_res <= cfg.density_tol) && (c_res <= cfg.comm_tol);
        const bool clip = cfg.trust_radius > 0.0 && !converged;
        if (clip) d2d(pre.get(), density.get(), bytes, st);

        auto linear_mix = [&](double alpha) {
            launch_axpy_c(nd, alpha, delta.as<cd>(), density.as<cd>(), st);
            if (proj) proj.apply(density.as<cd>(), st);
            launch_hermitize(nk, nb, density.as<cd>(), st);
        };

        if (converged) {
            d2d(density.get(), P_new.get(), bytes, st);
        } else if (use_diis) {
            if (diis_F.size() >= cfg.diis_size && !diis_F.empty()) {
                diis_F.erase(diis_F.begin());
                diis_R.erase(diis_R.begin());
            }
            diis_F.emplace_back(bytes);
            diis_R.emplace_back(bytes);
            d2d(diis_F.back().get(), F.get(), bytes, st);
            d2d(diis_R.back().get(), comm.get(), bytes, st);
            const size_t m = diis_F.size();
            if (m >= std::max<size_t>(2, cfg.diis_start)) {
                const size_t n = m + 1;
                DevBuf dB(m * m * sizeof(double));
                for (size_t i = 0; i < m; ++i)
                    for (size_t j = i; j < m; ++j)
                        launch_wdot(nk, static_cast<int>(K.nb2), diis_R[i].as<cd>(),
                                    diis_R[j].as<cd>(), K.w2d.as<double>(), S.ws.ls,
                                    dB.as<double>() + i * m + j, st);
                std::vector<double> hb(m * m);
                d2h(hb.data(), dB.get(), m * m * sizeof(double), st);
                HFGPU_CHECK(hipStreamSynchronize(st));
                std::vector<double> B(n * n, 0.0), rhs(n, 0.0), sol;
                for (size_t i = 0; i < m; ++i) {
                    for (size_t j = i; j < m; ++j) {
                        B[i * n + j] = hb[i * m + j];
                        B[j * n + i] = hb[i * m + j];
                    }
                    B[i * n + m] = -1.0;
                    B[m * n + i] = -1.0;
                }
                rhs[m] = -1.0;
                bool ok = fullpiv_solve(B, rhs, n, sol);
                for (size_t i = 0; ok && i < m; ++i) ok = std::isfinite(sol[i]);
                if (ok) {
                    DevBuf Fx(bytes), Pd(bytes);
                    fill_zero(Fx.get(), bytes, st);
                    for (size_t i = 0; i < m; ++i)
                        launch_axpy_c(nd, sol[i], diis_F[i].as<cd>(), Fx.as<cd>(), st);
                    const double damp = std::clamp(cfg.diis_damping, 0.0, 1.0);
                    if (damp < 1.0) launch_axpby(nd, damp, Fx.as<cd>(), 1.0 - damp, F.as<cd>(), st);
                    launch_hermitize(nk, nb, Fx.as<cd>(), st);
                    S.density_from_fock(Fx.as<cd>(), Pd.as<cd>());
                    if (proj) proj.apply(Pd.as<cd>(), st);
                    launch_hermitize(nk, nb, Pd.as<cd>(), st);
                    d2d(density.get(), Pd.get(), bytes, st);
                    HFGPU_CHECK(hipStreamSynchronize(st));
                } else {
                    linear_mix(cfg.mixing);
                }
            } else {
                linear_mix(cfg.mixing);
            }
        } else if (use_oda) {
            K.build_fock(P_new.as<cd>(), Sigma_new.as<cd>(), F_new.as<cd>(),
                         hdiag_new.as<double>(), true, st);
            if (proj) {
                proj.apply(F_new.as<cd>(), st);
                launch_hermitize(nk, nb, F_new.as<cd>(), st);
            }
            K.energy(P_new.as<cd>(), Sigma_new.as<cd>(), hdiag_new.as<double>(), sc + SC_E2, st);
            // c1 = sum_k w_k Re tr(F_k delta_k); delta is exactly Hermitian
            launch_wdot(nk, static_cast<int>(K.nb2), delta.as<cd>(), F.as<cd>(),
                        K.w2d.as<double>(), S.ws.ls, sc + SC_DOT, st);
            d2h(hsc, sc, sizeof(hsc), st);
            HFGPU_CHECK(hipStreamSynchronize(st));
            const double E_new = hsc[SC_E2], c1 = hsc[SC_DOT];
            const double c2 = (E_new - E_iter) - c1;
            double lam = 1.0;
            if (c2 > 1.0e-12) lam = std::clamp(-c1 / (2.0 * c2), 0.0, 1.0);
            else if (c1 < 0.0) lam = 1.0;
            else lam = 0.0;
            linear_mix(lam);
        } else {
            linear_mix(cfg.mixing);
        }

        if (clip) {
            launch_sub_c(nd, density.as<cd>(), pre.as<cd>(), diff.as<cd>(), st);
            launch_wdot(nk, static_cast<int>(K.nb2), diff.as<cd>(), diff.as<cd>(),
                        K.w2d.as<double>(), S.ws.ls, sc + SC_DOT, st);
            double dot = 0.0;
            d2h(&dot, sc + SC_DOT, sizeof(double), st);
            HFGPU_CHECK(hipStreamSynchronize(st));
            const double step = std::sqrt(dot / ws);
            if (step > cfg.trust_radius) {
                const double scale = cfg.trust_radius / step;
                d2d(density.get(), pre.get(), bytes, st);
                launch_axpy_c(nd, scale, diff.as<cd>(), density.as<cd>(), st);
                if (proj) proj.apply(density.as<cd>(), st);
                launch_hermitize(nk, nb, density.as<cd>(), st);
            }
        }
        hE.push_back(E_iter);
        hD.push_back(d_res);
        hC.push_back(c_res);
        (void)hmu;
        ++k;
    }
    S.build_and_occupy(density.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(),
                       P_new.as<cd>(), sc + SC_E);
    d2h(hsc, sc, sizeof(hsc), st);
    d2h(&hmu, S.ws.mu(), sizeof(MuState), st);
    if (out->density) d2h(out->density, density.get(), bytes, st);
    if (out->fock) d2h(out->fock, F.get(), bytes, st);
    HFGPU_CHECK(hipStreamSynchronize(st));
    out->energy = hsc[SC_E];
    out->mu = hmu.mu;
    out->iterations = k;
    out->converged = converged ? 1 : 0;
    for (size_t i = 0; i < k; ++i) {
        if (out->hist_E) out->hist_E[i] = hE[i];
        if (out->hist_density) out->hist_density[i] = hD[i];
        if (out->hist_comm) out->hist_comm[i] = hC[i];
    }
}

// ---------------------------------------------------------------------------
// Direct minimisation
// ---------------------------------------------------------------------------

void solve_dm_gpu(GpuKernel& K, const double* P0, double n_e, const hf_dm_config& cfg,
                  Projector& proj, hf_dm_result* out) {
    validate_block_sizes(cfg.block_sizes, cfg.n_blocks, static_cast<size_t>(K.nb));
    if (cfg.bt_max > static_cast<size_t>(kMaxBt))
        throw std::invalid_argument("bt_max above the GPU line-search limit of 16");
    const int nk = static_cast<int>(K.nk), nb = K.nb;
    const size_t nd = K.nd, bytes = nd * sizeof(cd), nkb = K.nk * K.nb;
    const size_t rb = nkb * sizeof(double);
    hipStream_t st = K.stream;
    SolveWs W(st);
    double* sc = W.sc.as<double>();

    const double wsum = std::max(K.weight_sum, TINY_REAL);
    std::vector<double> wn(K.w_host.size());
    for (size_t i = 0; i < wn.size(); ++i) wn[i] = K.w_host[i] / wsum;
    const double n_target = n_e / wsum;
    const double T_r = std::max(K.T, 1.0e-12);

    DevBuf w_norm(K.nk * sizeof(double));
    h2d(w_norm.get(), wn.data(), K.nk * sizeof(double), st);
    DevBuf Q(bytes), P(bytes), Sigma(bytes), F(bytes), G(bytes), Gp(bytes), dQ(bytes),
        dQp(bytes), Vd(bytes);
    DevBuf p(rb), eps(rb), g_p(rb), g_pp(rb), h_p(rb), d_p(rb), d_pp(rb), lam(rb), eps_w(rb),
        hdiag(nb * sizeof(double));
    fill_zero(Gp.get(), bytes, st);
    fill_zero(dQp.get(), bytes, st);
    fill_zero(g_pp.get(), rb, st);
    fill_zero(d_pp.get(), rb, st);
    const double* wd = w_norm.as<double>();

    auto mu_solve = [&](const double* e, bool warm, double mu0) {
        launch_mu_newton(e, wd, nk, nb, n_target, mu0, warm, T_r, cfg.mu_maxiter, W.mu(),
                         W.mu_partial.as<double>(), st);
        launch_occupations(e, nkb, W.mu(), T_r, 1, p.as<double>(), st);
    };
    auto density_step = [&]() {
        launch_density(nk, nb, Q.as<cd>(), p.as<double>(), P.as<cd>(), false, st);
        if (proj) proj.apply(P.as<cd>(), st);
        launch_hermitize(nk, nb, P.as<cd>(), st);
    };

    // ---- block structure (solver_dm.hpp:100-113, 403-405): more than one
    // block, no projector, and h, refP and the active contact operators
    // block-diagonal; then F at P0 must be too (446-447).  The blocks then
    // stay separate in every eigh of the solve: each block's ascending pairs
    // at the block's position.
    BlockMap bm{};
    bool use_block = false;
    if (cfg.n_blocks > 1) {
        bm = block_map(static_cast<size_t>(nb), cfg.block_sizes, cfg.n_blocks, true);
        use_block = !proj && block_ok_host(K.h.as<cd>(), nk, nb, bm, st) &&
                    (!K.has_refP || block_ok_host(K.refP.as<cd>(), nk, nb, bm, st));
        for (int t = 0; use_block && t < K.n_contact; ++t) {
            if (K.cg_host[t] == 0.0) continue;
            use_block = block_ok_host(K.Oi.as<cd>() + static_cast<size_t>(t) * K.nb2, 1, nb, bm,
                                      st) &&
                        block_ok_host(K.Oj.as<cd>() + static_cast<size_t>(t) * K.nb2, 1, nb, bm,
                                      st);
        }
    }
    auto eigh = [&](int in_mode, const cd* M, double* w, cd* V) {
        if (use_block)
            launch_eigh_blocks(nk, nb, in_mode, M, w, V, bm, false, st);
        else
            launch_eigh(nk, nb, in_mode, M, w, V, st);
    };

    // ---- init (solver_dm.hpp:436-459)
    h2d(P.get(), P0, bytes, st);
    launch_hermitize(nk, nb, P.as<cd>(), st);
    K.build_fock(P.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(), true, st);
    if (use_block && !block_ok_host(F.as<cd>(), nk, nb, bm, st)) use_block = false;
    eigh(0, F.as<cd>(), eps.as<double>(), Q.as<cd>());
    mu_solve(eps.as<double>(), false, 0.0);

    std::vector<double> hist_E, hist_G;
    double grad_norm = std::numeric_limits<double>::infinity();
    double E_prev = std::numeric_limits<double>::infinity();
    double dE = std::numeric_limits<double>::infinity();
    size_t k = 0;
    double hsc[SC_COUNT];
    while (k < cfg.max_iter) {
        bool e_nc;
        if (cfg.plateau_window > 0 && k > cfg.plateau_window)
            e_nc = (hist_E[k - 1 - cfg.plateau_window] - hist_E[k - 1]) > cfg.tol_E;
        else
            e_nc = dE > cfg.tol_E;
        const bool g_nc = (cfg.tol_grad > 0.0) && (grad_norm > cfg.tol_grad);
        if (!(e_nc || g_nc)) break;

        density_step();
        K.build_fock(P.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(), true, st);
        K.energy(P.as<cd>(), Sigma.as<cd>(), hdiag.as<double>(), sc + SC_E, st);
        launch_orbital_grad(nk, nb, Q.as<cd>(), F.as<cd>(), p.as<double>(), wd, W.mu(), T_r, true,
                            eps.as<double>(), G.as<cd>(), g_p.as<double>(), W.ls, st);
        launch_precond(nk, nb, G.as<cd>(), Gp.as<cd>(), eps.as<double>(), p.as<double>(),
                       g_p.as<double>(), g_pp.as<double>(), wd, T_r, cfg.denom_scale, dQ.as<cd>(),
                       h_p.as<double>(), W.ls, st);
        const bool restart = (k == 0) || (cfg.cg_restart > 0 && k % cfg.cg_restart == 0);
        launch_cg_combine(nk, nb, restart, dQ.as<cd>(), dQp.as<cd>(), h_p.as<double>(),
                          d_pp.as<double>(), d_p.as<double>(), eps.as<double>(), p.as<double>(),
                          wd, T_r, W.ls, st);
        eigh(1, dQ.as<cd>(), lam.as<double>(), Vd.as<cd>());
        launch_linesearch(nk, nb, Vd.as<cd>(), lam.as<double>(), F.as<cd>(), Sigma.as<cd>(),
                          p.as<double>(), d_p.as<double>(), wd, T_r, cfg.max_step, cfg.bt_shrink,
                          static_cast<int>(cfg.bt_max), W.ls, st);
        launch_retract(nk, nb, Q.as<cd>(), Vd.as<cd>(), lam.as<double>(), Sigma.as<cd>(),
                       eps_w.as<double>(), true, W.ls, st);
        mu_solve(eps_w.as<double>(), true, 0.0);

        d2h(hsc, sc, sizeof(hsc), st);
        HFGPU_CHECK(hipStreamSynchronize(st));
        const double E = hsc[SC_E];
        grad_norm = std::sqrt(std::max(0.0, hsc[SC_GN2]));
        hist_E.push_back(E);
        hist_G.push_back(grad_norm);
        dE = std::fabs(E - E_prev);
        std::swap(dQ, dQp);
        std::swap(d_p, d_pp);
        std::swap(G, Gp);
        std::swap(g_p, g_pp);
        E_prev = E;
        ++k;
    }

    // ---- finalise (solver_dm.hpp:615-652)
    density_step();
    K.build_fock(P.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(), true, st);
    launch_orbital_grad(nk, nb, Q.as<cd>(), F.as<cd>(), nullptr, wd, W.mu(), T_r, false, nullptr,
                        nullptr, nullptr, W.ls, st);
    eigh(0, F.as<cd>(), eps_w.as<double>(), Vd.as<cd>());
    launch_retract(nk, nb, Q.as<cd>(), Vd.as<cd>(), nullptr, nullptr, nullptr, false, W.ls, st);
    mu_solve(eps_w.as<double>(), true, 0.0);
    density_step();
    K.build_fock(P.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(), true, st);
    K.energy(P.as<cd>(), Sigma.as<cd>(), hdiag.as<double>(), sc + SC_E, st);

    MuState hmu;
    d2h(hsc, sc, sizeof(hsc), st);
    d2h(&hmu, W.mu(), sizeof(MuState), st);
    if (out->Q) d2h(out->Q, Q.get(), bytes, st);
    if (out->p) d2h(out->p, p.get(), rb, st);
    if (out->density) d2h(out->density, P.get(), bytes, st);
    if (out->fock) d2h(out->fock, F.get(), bytes, st);
    HFGPU_CHECK(hipStreamSynchronize(st));
    out->mu = hmu.mu;
    out->energy = hsc[SC_E];
    out->n_iter = k;
    out->converged = k < cfg.max_iter ? 1 : 0;
    for (size_t i = 0; i < k; ++i) {
        if (out->hist_E) out->hist_E[i] = hist_E[i];
        if (out->hist_grad) out->hist_grad[i] = hist_G[i];
    }
}

}  // namespace hfgpu

using namespace hfgpu;

extern "C" {

int hfgpu_solve_scf(hfgpu_kernel* k, const double* P0, double n_e, const hf_scf_config* cfg,
                    hf_project_fn project, void* user, hf_scf_result* out) {
    return guarded([&] {
        if (!k || !P0 || !cfg || !out) throw std::invalid_argument("null argument");
        GpuKernel& K = *k->impl;
        Projector pr;
        pr.fn = project;
        pr.user = user;
        pr.nk1 = K.nk1;
        pr.nk2 = K.nk2;
        pr.nb = K.nb;
        pr.nd = K.nd;
        solve_scf_gpu(K, P0, n_e, *cfg, pr, out);
    });
}

int hfgpu_solve_dm(hfgpu_kernel* k, const double* P0, double n_e, const hf_dm_config* cfg,
                   hf_project_fn project, void* user, hf_dm_result* out) {
    return guarded([&] {
        if (!k || !P0 || !cfg || !out || !out->p) throw std::invalid_argument("null argument");
        GpuKernel& K = *k->impl;
        validate_electron_count(K.w_host, static_cast<size_t>(K.nb), n_e);
        Projector pr;
        pr.fn = project;
        pr.user = user;
        pr.nk1 = K.nk1;
        pr.nk2 = K.nk2;
        pr.nb = K.nb;
        pr.nd = K.nd;
        solve_dm_gpu(K, P0, n_e, *cfg, pr, out);
    });
}

}  // extern "C"
