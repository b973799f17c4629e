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
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hfgpu.h"
#include "hfgpu_core.h"
#include "loops.h"
#include "rtr.h"
#include "symmetry.h"

namespace hfgpu {

namespace {

constexpr double TINY_REAL = 1.0e-30;

struct Projector {
    hf_project_fn fn = nullptr;
    void* user = nullptr;
    SymPlan* sym = nullptr;  // device symmetry projector in place of fn
    int nk1 = 0, nk2 = 0, nb = 0;
    size_t nd = 0;
    std::vector<double> host;

    explicit operator bool() const { return fn != nullptr || sym != nullptr; }
    // Host round trip: the callback sees and rewrites a host copy; the
    // device projector runs on the stream.
    void apply(cd* dM, hipStream_t st) {
        if (sym) {
            sym->apply(dM, st);
            return;
        }
        host.resize(2 * nd);
        d2h(host.data(), dM, nd * sizeof(cd), st);
        HFGPU_CHECK(hipStreamSynchronize(st));
        if (fn(user, host.data(), static_cast<size_t>(nk1), static_cast<size_t>(nk2),
               static_cast<size_t>(nb)) != 0)
            throw std::runtime_error("project_fn failed");
        h2d(dM, host.data(), nd * sizeof(cd), st);
    }
};

// Per-iteration launch sequences of the solver loops replayed from captured
// hipGraphs with HFGPU_SOLVER_GRAPH=1: one host call per iteration instead of
// one per kernel.  Off by default: with the next iteration enqueued ahead of
// the host's tests the launches are off the critical path, and a replay
// measured slower than the same kernels launched one by one (direct
// minimisation of the 24 x 24 QWZ model 3.0 against 2.4 ms, of the 96 x 96
// layered model 14.5 against 13.8 ms; the 48-solve sweep unchanged).  The captured work allocates nothing and clears memory
// only with fill kernels (no memset nodes).
bool solver_graphs_on() {
    const char* e = std::getenv("HFGPU_SOLVER_GRAPH");  // read per solve
    return e && e[0] == '1';
}

struct IterGraphs {
    hipStream_t st;
    bool on;
    std::vector<std::pair<int, hipGraphExec_t>> execs;
    IterGraphs(hipStream_t s, bool enable) : st(s), on(enable) {}
    ~IterGraphs() {
        for (auto& e : execs) (void)hipGraphExecDestroy(e.second);
    }
    IterGraphs(const IterGraphs&) = delete;
    IterGraphs& operator=(const IterGraphs&) = delete;

    template <class Fn>
    void run(int key, Fn&& body) {
        for (auto& e : execs)
            if (e.first == key) {
                HFGPU_CHECK(hipGraphLaunch(e.second, st));
                return;
            }
        hipGraph_t graph = nullptr;
        // relaxed: other solves' threads may allocate or copy synchronously
        // (new geometries, pool misses) while this one captures
        HFGPU_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HFGPU_CHECK(hipStreamEndCapture(st, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HFGPU_CHECK(ie);
        execs.emplace_back(key, exec);
        HFGPU_CHECK(hipGraphLaunch(exec, st));
    }
};

// The SCF loop with linear mixing runs one iteration ahead of the host's
// convergence test (HFGPU_SCF_PIPELINE=0: the test before every launch).
bool scf_pipeline_on() {
    const char* e = std::getenv("HFGPU_SCF_PIPELINE");  // read per solve
    return !(e && e[0] == '0');
}

// The direct-minimisation loop enqueues each iteration before the previous
// one's stop test (HFGPU_DM_PIPELINE=0: the test before every launch).
bool dm_pipeline_on() {
    const char* e = std::getenv("HFGPU_DM_PIPELINE");  // read per solve
    return !(e && e[0] == '0');
}

// Pinned host memory for scalars read back while later work is in flight: an
// asynchronous copy into it completes without the host waiting.  One page
// per host thread, kept for the thread's lifetime (freeing pinned memory
// synchronises the device, which would stall concurrent solves).
double* pinned_scalars() {
    thread_local double* p = nullptr;
    if (!p) {
        void* q = nullptr;
        HFGPU_CHECK(hipHostMalloc(&q, 4096, hipHostMallocDefault));
        p = static_cast<double*>(q);
    }
    return p;
}

struct StreamEvent {
    hipEvent_t e = nullptr;
    StreamEvent() = default;
    void create() {
        if (!e) HFGPU_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    ~StreamEvent() {
        if (e) (void)hipEventDestroy(e);
    }
    StreamEvent(const StreamEvent&) = delete;
    StreamEvent& operator=(const StreamEvent&) = delete;
};

// A device symmetry projector from the solver configuration, if any.
void attach_symmetry(Projector& pr, SymPlan& sp, const hf_symmetry_desc* sym, GpuKernel& K) {
    if (!sym) return;
    if (pr.fn)
        throw std::invalid_argument(
            "pass either project_fn or a device symmetry projector, not both");
    if (sym->nb != static_cast<size_t>(K.nb))
        throw std::invalid_argument("symmetry matrices must be (nb, nb) with the kernel's nb");
    sp.init(*sym, K.nk1, K.nk2, K.stream);
    pr.sym = &sp;
}

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
    DevBuf eps, shifted, V;
    Projector& proj;
    double n_e, level_shift;
    // Fock builds with the FP32 exchange workspace (mixed precision) until
    // the solver switches to the FP64 correction phase
    bool mixed = false;
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
        // occupations formed inside the density kernel (solver_scf.hpp:138-152)
        launch_density_occ(nk, nb, V.as<cd>(), e_used, ws.mu(), K.T, 0, P_out, true, st);
    }

    // build_and_occupy (solver_scf.hpp:179-197); energy -> sc[E].
    //
    // Hermitisation is idempotent bit for bit (herm(herm(X)) = herm(X)), and
    // real linear combinations of exactly Hermitian matrices are exactly
    // Hermitian; without a projector the launches below skip the
    // hermitisations whose input is already exactly Hermitian: the density
    // after the first iteration (every update ends in a hermitisation or a
    // real combination of Hermitian matrices) and the density the eigenvectors
    // give (the density kernel hermitises).  d_herm: d is exactly Hermitian.
    void build_and_occupy(cd* d, cd* Sigma, cd* F, double* hdiag, cd* P_out, double* E_slot,
                          bool d_herm = false) {
        if (proj) proj.apply(d, st);
        if (proj || !d_herm) launch_hermitize(static_cast<int>(K.nk), K.nb, d, st);
        K.build_fock(d, Sigma, F, hdiag, true, st, true, mixed);
        if (proj) {
            proj.apply(F, st);
            launch_hermitize(static_cast<int>(K.nk), K.nb, F, st);
        }
        K.energy(d, Sigma, hdiag, E_slot, st);
        density_from_fock(F, P_out);
        if (proj) {
            proj.apply(P_out, st);
            launch_hermitize(static_cast<int>(K.nk), K.nb, P_out, st);
        }
    }
};

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
    // DIIS: B entries of the kept residual pairs (row-major, history order),
    // the new residual's products, and the extrapolated Fock matrix
    std::vector<double> Bc, hdots;
    const size_t dot_slots = use_diis ? kMaxDots * (cfg.diis_size / kMaxDots + 1) : 0;
    DevBuf dots(dot_slots * sizeof(double)), Fx(use_diis ? bytes : 0);
    hdots.resize(dot_slots);
    S.mixed = cfg.exchange_fp32 != 0;
    const double fp32_switch = cfg.fp32_switch > 0.0 ? cfg.fp32_switch : 1e-6;
    h2d(density.get(), P0, bytes, st);
    double* sc = S.ws.sc.as<double>();

    std::vector<double> hE, hD, hC;
    bool converged = false;
    size_t k = 0;
    const double ws = std::max(K.weight_sum, 1.0e-30);
    double hsc[SC_COUNT];
    MuState hmu;
    IterGraphs graphs(st, !proj && solver_graphs_on());

    // Linear mixing with no trust radius, projector or FP32 phase: iteration
    // j + 1 depends on the host only through the convergence test of
    // iteration j (converged: the density becomes P_new; otherwise it takes
    // the mixing step), so iteration j + 1 is enqueued, mixing step first,
    // before the residuals of iteration j are read back.  Its density,
    // energy and residuals go to the other of two slots; when iteration j
    // converges the speculative one is discarded, so every result equals the
    // synchronous loop's bit for bit.  The device then runs the iterations
    // back to back instead of idling through each host round trip.
    const bool pipelined = !use_diis && !use_oda && cfg.trust_radius <= 0.0 && !proj &&
                           !S.mixed && cfg.max_iter > 0 && scf_pipeline_on();
    // the second slot and the events live until the solve's final stream
    // synchronisation: a discarded speculative iteration may still be running
    // when the loop ends, and a buffer back in the pool could reach another
    // thread's solve before this stream drains
    DevBuf P_alt, sc_alt;
    StreamEvent ev[2];
    if (pipelined) {
        P_alt.alloc(bytes);
        sc_alt.alloc(SC_COUNT * sizeof(double));
        ev[0].create();
        ev[1].create();
        fill_zero(sc_alt.get(), SC_COUNT * sizeof(double), st);
        cd* const Pn[2] = {P_new.as<cd>(), P_alt.as<cd>()};
        double* const scs[2] = {sc, sc_alt.as<double>()};
        constexpr size_t kSlot = 32;  // doubles per host slot (>= SC_RES + 2)
        static_assert(SC_RES + 2 <= kSlot && SC_E < SC_RES, "host slot layout");
        double* const hs = pinned_scalars();
        auto iteration = [&](size_t j) {
            const int s = static_cast<int>(j & 1);
            if (j > 0) launch_axpy_c(nd, cfg.mixing, delta.as<cd>(), density.as<cd>(), st);
            S.build_and_occupy(density.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(),
                               Pn[s], scs[s] + SC_E, j > 0);
            launch_scf_residuals(nk, nb, F.as<cd>(), density.as<cd>(), Pn[s], delta.as<cd>(),
                                 nullptr, K.w2d.as<double>(), S.ws.ls, scs[s] + SC_RES, st);
        };
        auto enqueue = [&](size_t j) {
            const int s = static_cast<int>(j & 1);
            if (graphs.on && j > 0)
                graphs.run(2 | s, [&] { iteration(j); });
            else
                iteration(j);
            d2h(hs + s * kSlot, scs[s], (SC_RES + 2) * sizeof(double), st);
            HFGPU_CHECK(hipEventRecord(ev[s].e, st));
        };
        enqueue(0);
        for (size_t j = 0;; ++j) {
            const bool ahead = j + 1 < cfg.max_iter;
            if (ahead) enqueue(j + 1);
            HFGPU_CHECK(hipEventSynchronize(ev[j & 1].e));
            const double* h = hs + (j & 1) * kSlot;
            const double d_res = std::sqrt(h[SC_RES] / ws);
            const double c_res = std::sqrt(h[SC_RES + 1] / ws);
            converged = (d_res <= cfg.density_tol) && (c_res <= cfg.comm_tol);
            hE.push_back(h[SC_E]);
            hD.push_back(d_res);
            hC.push_back(c_res);
            k = j + 1;
            if (converged) {
                d2d(density.get(), Pn[j & 1], bytes, st);
                break;
            }
            if (!ahead) {
                launch_axpy_c(nd, cfg.mixing, delta.as<cd>(), density.as<cd>(), st);
                break;
            }
        }
    }

    while (!pipelined && k < cfg.max_iter && !converged) {
        auto body = [&]() {
            S.build_and_occupy(density.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(),
                               P_new.as<cd>(), sc + SC_E, k > 0);
            launch_scf_residuals(nk, nb, F.as<cd>(), density.as<cd>(), P_new.as<cd>(),
                                 delta.as<cd>(), use_diis ? comm.as<cd>() : nullptr,
                                 K.w2d.as<double>(), S.ws.ls, sc + SC_RES, st);
        };
        // the build, occupation and residual launches of every iteration after
        // the first are one captured graph per precision (the host decisions
        // that follow read the residuals back)
        if (graphs.on && k > 0)
            graphs.run(S.mixed ? 1 : 0, body);
        else
            body();
        // DIIS: the products of the new commutator residual with the
        // residuals the history keeps (its oldest leaves when it is full) and
        // with itself, in the read-back of the convergence test.  The
        // reference recomputes every B entry each iteration
        // (solver_scf.hpp:263-290); an entry depends only on its two stored
        // residuals, so the older ones are kept on the host instead.
        const size_t drop = (use_diis && !diis_R.empty() && diis_R.size() >= cfg.diis_size) ? 1 : 0;
        const size_t ndots = use_diis ? diis_R.size() - drop + 1 : 0;
        for (size_t c0 = 0; c0 < ndots; c0 += kMaxDots) {
            CdPtrs X{};
            const size_t cnt = std::min<size_t>(kMaxDots, ndots - c0);
            for (size_t d = 0; d < cnt; ++d) {
                const size_t i = c0 + d + drop;
                X.p[d] = i < diis_R.size() ? diis_R[i].as<cd>() : comm.as<cd>();
            }
            launch_wdot_multi(nk, static_cast<int>(K.nb2), X, static_cast<int>(cnt), comm.as<cd>(),
                              K.w2d.as<double>(), S.ws.ls, dots.as<double>() + c0, st);
        }
        if (ndots) d2h(hdots.data(), dots.get(), dots.size(), st);
        d2h(hsc, sc, sizeof(hsc), st);
        d2h(&hmu, S.ws.mu(), sizeof(MuState), st);
        HFGPU_CHECK(hipStreamSynchronize(st));
        const double E_iter = hsc[SC_E];
        const double d_res = std::sqrt(hsc[SC_RES] / ws);
        const double c_res = std::sqrt(hsc[SC_RES + 1] / ws);
        converged = (d_res <= cfg.density_tol) && (c_res <= cfg.comm_tol);
        // mixed precision: the FP64 correction phase starts below the switch
        // residual, and convergence counts only on FP64 builds
        if (S.mixed && (converged || d_res <= fp32_switch)) {
            S.mixed = false;
            converged = false;
        }
        const bool clip = cfg.trust_radius > 0.0 && !converged;
        if (clip) d2d(pre.get(), density.get(), bytes, st);

        // density + alpha delta of two exactly Hermitian matrices is exactly
        // Hermitian (negation and the real scaling commute with rounding)
        auto linear_mix = [&](double alpha) {
            launch_axpy_c(nd, alpha, delta.as<cd>(), density.as<cd>(), st);
            if (proj) {
                proj.apply(density.as<cd>(), st);
                launch_hermitize(nk, nb, density.as<cd>(), st);
            }
        };

        if (converged) {
            d2d(density.get(), P_new.get(), bytes, st);
        } else if (use_diis) {
            if (drop) {
                diis_F.erase(diis_F.begin());
                diis_R.erase(diis_R.begin());
                const size_t mo = diis_R.size() + 1;
                std::vector<double> Bn((mo - 1) * (mo - 1));
                for (size_t i = 1; i < mo; ++i)
                    for (size_t j = 1; j < mo; ++j) Bn[(i - 1) * (mo - 1) + j - 1] = Bc[i * mo + j];
                Bc.swap(Bn);
            }
            diis_F.emplace_back(bytes);
            diis_R.emplace_back(bytes);
            d2d(diis_F.back().get(), F.get(), bytes, st);
            d2d(diis_R.back().get(), comm.get(), bytes, st);
            const size_t m = diis_F.size();
            {
                std::vector<double> Bn(m * m);
                for (size_t i = 0; i + 1 < m; ++i)
                    for (size_t j = 0; j + 1 < m; ++j) Bn[i * m + j] = Bc[i * (m - 1) + j];
                for (size_t i = 0; i < m; ++i) {
                    Bn[i * m + (m - 1)] = hdots[i];
                    Bn[(m - 1) * m + i] = hdots[i];
                }
                Bc.swap(Bn);
            }
            if (m >= std::max<size_t>(2, cfg.diis_start)) {
                const size_t n = m + 1;
                std::vector<double> B(n * n, 0.0), rhs(n, 0.0), sol;
                for (size_t i = 0; i < m; ++i) {
                    for (size_t j = 0; j < m; ++j) B[i * n + j] = Bc[i * m + j];
                    B[i * n + m] = -1.0;
                    B[m * n + i] = -1.0;
                }
                rhs[m] = -1.0;
                bool ok = fullpiv_solve(B, rhs, n, sol);
                for (size_t i = 0; ok && i < m; ++i) ok = std::isfinite(sol[i]);
                if (ok) {
                    // F_extrap = sum_i c_i F_i, then the damping blend, in one
                    // pass per 16 terms (solver_scf.hpp:305-324).  Real
                    // combinations of exactly Hermitian Fock matrices are
                    // exactly Hermitian and the density kernel hermitises,
                    // so without a projector the reference's two
                    // hermitisations are no-ops and are skipped.
                    const double damp = std::clamp(cfg.diis_damping, 0.0, 1.0);
                    for (size_t c0 = 0; c0 < m; c0 += kMaxDots) {
                        CdPtrs X{};
                        DotCoefs cf{};
                        const size_t cnt = std::min<size_t>(kMaxDots, m - c0);
                        for (size_t d = 0; d < cnt; ++d) {
                            X.p[d] = diis_F[c0 + d].as<cd>();
                            cf.c[d] = sol[c0 + d];
                        }
                        const bool last = c0 + cnt == m;
                        launch_lincomb(nd, X, cf, static_cast<int>(cnt), c0 > 0, damp, 1.0 - damp,
                                       last && damp < 1.0 ? F.as<cd>() : nullptr, Fx.as<cd>(), st);
                    }
                    if (proj) launch_hermitize(nk, nb, Fx.as<cd>(), st);
                    S.density_from_fock(Fx.as<cd>(), density.as<cd>());
                    if (proj) {
                        proj.apply(density.as<cd>(), st);
                        launch_hermitize(nk, nb, density.as<cd>(), st);
                    }
                } else {
                    linear_mix(cfg.mixing);
                }
            } else {
                linear_mix(cfg.mixing);
            }
        } else if (use_oda) {
            K.build_fock(P_new.as<cd>(), Sigma_new.as<cd>(), F_new.as<cd>(),
                         hdiag_new.as<double>(), true, st, true, S.mixed);
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
    S.mixed = false;
    S.build_and_occupy(density.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(),
                       P_new.as<cd>(), sc + SC_E, k > 0);
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
                         W.mu_partial.as<double>(), st, nullptr, p.as<double>());
    };
    auto density_step = [&]() {
        // the hermitising density kernel equals X followed by a hermitisation
        // bit for bit; a projector acts on X itself (solver_dm.hpp:117-128)
        launch_density(nk, nb, Q.as<cd>(), p.as<double>(), P.as<cd>(), !proj, st);
        if (proj) {
            proj.apply(P.as<cd>(), st);
            launch_hermitize(nk, nb, P.as<cd>(), st);
        }
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

    // mixed precision: FP32 exchange workspace until the gradient norm
    // reaches the switch, FP64 builds afterwards and in the finalisation
    bool mixed = cfg.exchange_fp32 != 0;
    const double fp32_switch = cfg.fp32_switch > 0.0 ? cfg.fp32_switch : 1e-6;

    // ---- init (solver_dm.hpp:436-459)
    h2d(P.get(), P0, bytes, st);
    launch_hermitize(nk, nb, P.as<cd>(), st);
    K.build_fock(P.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(), true, st, true,
                 mixed);
    if (use_block && !block_ok_host(F.as<cd>(), nk, nb, bm, st)) use_block = false;
    eigh(0, F.as<cd>(), eps.as<double>(), Q.as<cd>());
    mu_solve(eps.as<double>(), false, 0.0);

    IterGraphs graphs(st, !proj && solver_graphs_on());
    std::vector<double> hist_E, hist_G;
    size_t ls_trials = 0;
    double grad_norm = std::numeric_limits<double>::infinity();
    double E_prev = std::numeric_limits<double>::infinity();
    double dE = std::numeric_limits<double>::infinity();
    size_t k = 0;
    double hsc[SC_COUNT];
    // the reference's test at the top of iteration k (solver_dm.hpp:476-490)
    auto keep_going = [&]() {
        bool e_nc;
        if (cfg.plateau_window > 0 && k > cfg.plateau_window)
            e_nc = (hist_E[k - 1 - cfg.plateau_window] - hist_E[k - 1]) > cfg.tol_E;
        else
            e_nc = dE > cfg.tol_E;
        const bool g_nc = (cfg.tol_grad > 0.0) && (grad_norm > cfg.tol_grad);
        return e_nc || g_nc;
    };
    // the host's record of an iteration from its scalar slots
    auto record = [&](const double* h) {
        const double E = h[SC_E];
        grad_norm = std::sqrt(std::max(0.0, h[SC_GN2]));
        hist_E.push_back(E);
        hist_G.push_back(grad_norm);
        // trials the reference's sequential search evaluates: up to and
        // including the accepted tau0 shrink^j, all bt_max when none is
        if (h[SC_ACCEPTED] > 0.5 && h[SC_TAU0] > 0.0 && cfg.bt_shrink > 0.0 &&
            cfg.bt_shrink < 1.0) {
            const double j = std::log(h[SC_TAU] / h[SC_TAU0]) / std::log(cfg.bt_shrink);
            ls_trials += static_cast<size_t>(std::llround(std::max(0.0, j))) + 1;
        } else {
            ls_trials += cfg.bt_max;
        }
        dE = std::fabs(E - E_prev);
        E_prev = E;
    };

    // Without a projector or FP32 builds, iteration k + 1 depends on the host
    // only through the stop test at its top, which reads iteration k's energy
    // and gradient norm: it is enqueued before those are read back.  Its new
    // orbitals, occupations and chemical potential go to the other of two
    // buffers (the retraction and the Newton solve write out of place) and
    // its scalars to the other of two slot arrays, so when the test stops the
    // loop the speculative iteration is discarded and the state is that of
    // the test-first loop bit for bit.  HFGPU_DM_PIPELINE=0 tests first.
    const bool spec = !proj && !mixed && cfg.max_iter > 0 && dm_pipeline_on();
    // the second buffers and the events live until the solve's final stream
    // synchronisation (a discarded iteration may still be running)
    DevBuf Q_alt, p_alt, mu_alt, sc_alt;
    StreamEvent ev[2];
    if (spec) {
        Q_alt.alloc(bytes);
        p_alt.alloc(rb);
        mu_alt.alloc(sizeof(MuState));
        sc_alt.alloc(SC_COUNT * sizeof(double));
        fill_zero(mu_alt.get(), sizeof(MuState), st);
        fill_zero(sc_alt.get(), SC_COUNT * sizeof(double), st);
        ev[0].create();
        ev[1].create();
        cd* const Qs[2] = {Q.as<cd>(), Q_alt.as<cd>()};
        double* const ps[2] = {p.as<double>(), p_alt.as<double>()};
        MuState* const ms[2] = {W.mu(), mu_alt.as<MuState>()};
        LoopScratch lss[2] = {W.ls, W.ls};
        lss[1].sc = sc_alt.as<double>();
        constexpr size_t kSlot = 32;  // doubles per host slot
        static_assert(SC_ACCEPTED < kSlot, "host slot layout");
        double* const hs = pinned_scalars();
        // iteration j reads state j & 1 and writes state (j + 1) & 1
        auto iteration = [&](size_t j, bool restart) {
            const int s0 = static_cast<int>(j & 1), s1 = s0 ^ 1;
            LoopScratch& L = lss[s0];
            launch_density(nk, nb, Qs[s0], ps[s0], P.as<cd>(), true, st);
            K.build_fock(P.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(), true, st,
                         true, false);
            K.energy(P.as<cd>(), Sigma.as<cd>(), hdiag.as<double>(), L.sc + SC_E, st);
            launch_orbital_grad(nk, nb, Qs[s0], F.as<cd>(), ps[s0], wd, ms[s0], T_r, true,
                                eps.as<double>(), G.as<cd>(), g_p.as<double>(), L, st);
            launch_precond(nk, nb, G.as<cd>(), Gp.as<cd>(), eps.as<double>(), ps[s0],
                           g_p.as<double>(), g_pp.as<double>(), wd, T_r, cfg.denom_scale,
                           dQ.as<cd>(), h_p.as<double>(), L, st);
            launch_cg_combine(nk, nb, restart, dQ.as<cd>(), dQp.as<cd>(), h_p.as<double>(),
                              d_pp.as<double>(), d_p.as<double>(), eps.as<double>(), ps[s0], wd,
                              T_r, L, st);
            eigh(1, dQ.as<cd>(), lam.as<double>(), Vd.as<cd>());
            launch_linesearch(nk, nb, Vd.as<cd>(), lam.as<double>(), F.as<cd>(), Sigma.as<cd>(),
                              ps[s0], d_p.as<double>(), wd, T_r, cfg.max_step, cfg.bt_shrink,
                              static_cast<int>(cfg.bt_max), L, st);
            launch_retract(nk, nb, Qs[s0], Vd.as<cd>(), lam.as<double>(), Sigma.as<cd>(),
                           eps_w.as<double>(), true, L, st, Qs[s1]);
            launch_mu_newton(eps_w.as<double>(), wd, nk, nb, n_target, 0.0, true, T_r,
                             cfg.mu_maxiter, ms[s1], W.mu_partial.as<double>(), st, ms[s0],
                             ps[s1]);
        };
        auto enqueue = [&](size_t j) {
            const bool restart = (j == 0) || (cfg.cg_restart > 0 && j % cfg.cg_restart == 0);
            const int s0 = static_cast<int>(j & 1);
            if (graphs.on && j > 0)
                graphs.run(8 | s0 | (restart ? 2 : 0), [&] { iteration(j, restart); });
            else
                iteration(j, restart);
            d2h(hs + s0 * kSlot, lss[s0].sc, (SC_ACCEPTED + 1) * sizeof(double), st);
            HFGPU_CHECK(hipEventRecord(ev[s0].e, st));
            // the buffer pairs the next iteration reads as its history
            std::swap(dQ, dQp);
            std::swap(d_p, d_pp);
            std::swap(G, Gp);
            std::swap(g_p, g_pp);
        };
        if (keep_going()) {
            enqueue(0);
            for (;;) {
                const bool ahead = k + 1 < cfg.max_iter;
                if (ahead) enqueue(k + 1);
                HFGPU_CHECK(hipEventSynchronize(ev[k & 1].e));
                record(hs + (k & 1) * kSlot);
                ++k;
                if (!ahead || !keep_going()) break;  // an enqueued iteration k is discarded
            }
        }
        // the state after k iterations is state k & 1
        if (k & 1) {
            std::swap(Q, Q_alt);
            std::swap(p, p_alt);
            std::swap(W.mus, mu_alt);
        }
    }

    while (!spec && k < cfg.max_iter) {
        if (!keep_going()) {
            if (!mixed) break;
            mixed = false;  // stop only on FP64 builds
        }

        const bool restart = (k == 0) || (cfg.cg_restart > 0 && k % cfg.cg_restart == 0);
        auto body = [&]() {
            density_step();
            K.build_fock(P.as<cd>(), Sigma.as<cd>(), F.as<cd>(), hdiag.as<double>(), true, st,
                         true, mixed);
            K.energy(P.as<cd>(), Sigma.as<cd>(), hdiag.as<double>(), sc + SC_E, st);
            launch_orbital_grad(nk, nb, Q.as<cd>(), F.as<cd>(), p.as<double>(), wd, W.mu(), T_r,
                                true, eps.as<double>(), G.as<cd>(), g_p.as<double>(), W.ls, st);
            launch_precond(nk, nb, G.as<cd>(), Gp.as<cd>(), eps.as<double>(), p.as<double>(),
                           g_p.as<double>(), g_pp.as<double>(), wd, T_r, cfg.denom_scale,
                           dQ.as<cd>(), h_p.as<double>(), W.ls, st);
            launch_cg_combine(nk, nb, restart, dQ.as<cd>(), dQp.as<cd>(), h_p.as<double>(),
                              d_pp.as<double>(), d_p.as<double>(), eps.as<double>(),
                              p.as<double>(), wd, T_r, W.ls, st);
            eigh(1, dQ.as<cd>(), lam.as<double>(), Vd.as<cd>());
            launch_linesearch(nk, nb, Vd.as<cd>(), lam.as<double>(), F.as<cd>(), Sigma.as<cd>(),
                              p.as<double>(), d_p.as<double>(), wd, T_r, cfg.max_step,
                              cfg.bt_shrink, static_cast<int>(cfg.bt_max), W.ls, st);
            launch_retract(nk, nb, Q.as<cd>(), Vd.as<cd>(), lam.as<double>(), Sigma.as<cd>(),
                           eps_w.as<double>(), true, W.ls, st);
            mu_solve(eps_w.as<double>(), true, 0.0);
        };
        // After the first iteration (which makes every lazy allocation) the
        // body is the same launch sequence each time up to the CG restart
        // flag, the FP32/FP64 build and the parity of the swapped buffer
        // pairs below: one captured graph per variant, replayed.
        if (graphs.on && k > 0)
            graphs.run(((k & 1) ? 1 : 0) | (restart ? 2 : 0) | (mixed ? 4 : 0), body);
        else
            body();

        d2h(hsc, sc, sizeof(hsc), st);
        HFGPU_CHECK(hipStreamSynchronize(st));
        record(hsc);
        if (mixed && grad_norm <= fp32_switch) mixed = false;
        std::swap(dQ, dQp);
        std::swap(d_p, d_pp);
        std::swap(G, Gp);
        std::swap(g_p, g_pp);
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
    out->n_ls_trials = ls_trials;
    for (size_t i = 0; i < k; ++i) {
        if (out->hist_E) out->hist_E[i] = hist_E[i];
        if (out->hist_grad) out->hist_grad[i] = hist_G[i];
    }
}

// ---------------------------------------------------------------------------
// Direct minimisation of several solves of one shape at once
// ---------------------------------------------------------------------------

namespace {

// Pinned host scalars for a batch: grows per host thread, kept for the
// thread's lifetime (as pinned_scalars).
double* pinned_batch_scalars(size_t doubles) {
    thread_local double* p = nullptr;
    thread_local size_t cap = 0;
    if (doubles > cap) {
        void* q = nullptr;
        const size_t want = std::max<size_t>(doubles, 4096);
        HFGPU_CHECK(hipHostMalloc(&q, want * sizeof(double), hipHostMallocDefault));
        p = static_cast<double*>(q);  // the previous block stays allocated
        cap = want;
    }
    return p;
}

}  // namespace

// S independent direct-minimisation solves (one Fock kernel each) with one
// band count and temperature, on any meshes, advanced together: every launch
// of an iteration covers all S solves.  The solves' k points are stored back
// to back (solve s owns k points seg[s] .. seg[s + 1] - 1 of every array);
// the per-k kernels (density, orbital gradient, preconditioner, CG
// combination, eigh of the direction, line search, retraction, occupations)
// give solve s workgroup row s with the workgroup decomposition it has alone,
// the Fock builds of solves on one mesh share one launch per exchange pass
// (FockBatch), the energies one launch pair and the chemical potentials one
// launch pair (launch_mu_newton_multi).  Each solve keeps the reference's
// stop test (solver_dm.hpp:476-490) on its own history: a solve that stops
// is masked out of the launches that change its state (retraction,
// occupations, chemical potential), so its orbitals stay those of its last
// iteration, and the finalisation runs for all solves at the end.  Every
// per-solve value is computed by the same kernels with the same
// decomposition as a single solve, so each result equals solve_dm_gpu's bit
// for bit.  Iteration k + 1 is enqueued before the scalars of iteration k
// are read back, as in solve_dm_gpu.
void solve_dm_batch_gpu(const std::vector<GpuKernel*>& Ks, const std::vector<const double*>& P0s,
                        const std::vector<double>& n_es, const hf_dm_config& cfg,
                        hf_dm_result* outs) {
    const int S = static_cast<int>(Ks.size());
    if (S < 1 || S > kMaxBatch) throw std::invalid_argument("batch of 1..32 solves");
    GpuKernel& K0 = *Ks[0];
    for (GpuKernel* Kp : Ks) {
        const GpuKernel& K = *Kp;
        if (K.nb != K0.nb || K.T != K0.T)
            throw std::invalid_argument(
                "a batch needs one band count and temperature for every kernel");
    }
    if (cfg.n_blocks > 0 || cfg.symmetry || cfg.exchange_fp32 || cfg.optimizer == 1 ||
        cfg.n_deflation > 0)
        throw std::invalid_argument(
            "batched solves take the CG optimizer without blocks, symmetry, deflation or FP32 "
            "builds");
    if (cfg.bt_max > static_cast<size_t>(kMaxBt))
        throw std::invalid_argument("bt_max above the GPU line-search limit of 16");
    const int nb = K0.nb;
    const size_t SS = static_cast<size_t>(S);
    const double T_r = std::max(K0.T, 1.0e-12);
    hipStream_t st = K0.stream;

    // the solves' k ranges
    std::vector<int> koff(SS + 1, 0);
    int nk_max = 0;
    for (int s = 0; s < S; ++s) {
        const size_t nks = Ks[s]->nk;
        if (static_cast<size_t>(koff[s]) + nks > static_cast<size_t>(std::numeric_limits<int>::max() / 64))
            throw std::invalid_argument("batched solves hold too many k points");
        koff[s + 1] = koff[s] + static_cast<int>(nks);
        nk_max = std::max(nk_max, static_cast<int>(nks));
    }
    const size_t nk_tot = static_cast<size_t>(koff[S]);
    const size_t nb2 = static_cast<size_t>(nb) * nb;
    const size_t bytes_all = nk_tot * nb2 * sizeof(cd), rb_all = nk_tot * nb * sizeof(double);
    auto nd_of = [&](int s) { return static_cast<size_t>(koff[s + 1] - koff[s]) * nb2; };
    auto nkb_of = [&](int s) { return static_cast<size_t>(koff[s + 1] - koff[s]) * nb; };

    std::vector<double> n_target(SS);
    std::vector<double> wn(nk_tot);
    for (int s = 0; s < S; ++s) {
        const GpuKernel& K = *Ks[s];
        validate_electron_count(K.w_host, static_cast<size_t>(nb), n_es[s]);
        const double wsum = std::max(K.weight_sum, TINY_REAL);
        for (size_t i = 0; i < K.nk; ++i) wn[koff[s] + i] = K.w_host[i] / wsum;
        n_target[s] = n_es[s] / wsum;
    }
    DevBuf w_norm(nk_tot * sizeof(double)), seg_d((SS + 1) * sizeof(int)),
        nt_d(SS * sizeof(double));
    h2d(w_norm.get(), wn.data(), nk_tot * sizeof(double), st);
    h2d(seg_d.get(), koff.data(), (SS + 1) * sizeof(int), st);
    h2d(nt_d.get(), n_target.data(), SS * sizeof(double), st);
    const double* wd = w_norm.as<double>();

    DevBuf Q(bytes_all), P(bytes_all), Sigma(bytes_all), F(bytes_all), G(bytes_all),
        Gp(bytes_all), dQ(bytes_all), dQp(bytes_all), Vd(bytes_all), Q_alt(bytes_all);
    DevBuf p(rb_all), p_alt(rb_all), eps(rb_all), g_p(rb_all), g_pp(rb_all), h_p(rb_all),
        d_p(rb_all), d_pp(rb_all), lam(rb_all), eps_w(rb_all),
        hdiag(SS * std::max(nb, 1) * sizeof(double));
    DevBuf mus(SS * sizeof(MuState)), mus_alt(SS * sizeof(MuState));
    const bool mu_multi = mu_newton_multi_ok(nk_max, nb, cfg.mu_maxiter);
    DevBuf mu_ws((mu_multi ? SS : 1) * kMuPartialDoubles * sizeof(double));
    DevBuf part(SS * kLoopPartialDoubles * sizeof(double));
    DevBuf scb[2] = {DevBuf(SS * SC_COUNT * sizeof(double)), DevBuf(SS * SC_COUNT * sizeof(double))};
    fill_zero(Gp.get(), bytes_all, st);
    fill_zero(dQp.get(), bytes_all, st);
    fill_zero(g_pp.get(), rb_all, st);
    fill_zero(d_pp.get(), rb_all, st);
    fill_zero(mus.get(), SS * sizeof(MuState), st);
    fill_zero(mus_alt.get(), SS * sizeof(MuState), st);
    fill_zero(scb[0].get(), SS * SC_COUNT * sizeof(double), st);
    fill_zero(scb[1].get(), SS * SC_COUNT * sizeof(double), st);

    auto Pd = [&](const DevBuf& b, int s) { return b.as<cd>() + static_cast<size_t>(koff[s]) * nb2; };
    auto Rd = [&](const DevBuf& b, int s) {
        return b.as<double>() + static_cast<size_t>(koff[s]) * nb;
    };
    // chemical potentials: one launch pair for the batch where the one-launch
    // Newton paths apply, else one launch sequence per solve
    // (with the occupations of the active solves in occ, as launch_occupations
    // mode 1)
    auto mu_solve = [&](const double* e, MuState* out, const MuState* warm, bool use_warm,
                        unsigned act, double* occ) {
        if (mu_multi) {
            launch_mu_newton_multi(e, wd, seg_d.as<int>(), S, nk_max, nb, nt_d.as<double>(), T_r,
                                   cfg.mu_maxiter, use_warm, warm, out, mu_ws.as<double>(), act,
                                   st, occ);
            return;
        }
        for (int s = 0; s < S; ++s)
            if ((act >> s) & 1u)
                launch_mu_newton(e + static_cast<size_t>(koff[s]) * nb, wd + koff[s],
                                 koff[s + 1] - koff[s], nb, n_target[s], 0.0, use_warm, T_r,
                                 cfg.mu_maxiter, out + s, mu_ws.as<double>(), st, warm + s,
                                 occ + static_cast<size_t>(koff[s]) * nb);
    };
    // Fock builds and energies: the scalar slot array decides where the
    // energies land, so there is one FockBatch per slot array
    FockBatch fb[2];
    for (int b = 0; b < 2; ++b) {
        std::vector<const cd*> Pp(SS);
        std::vector<cd*> Sp(SS), Fp(SS);
        std::vector<double*> hp(SS), Ep(SS);
        for (int s = 0; s < S; ++s) {
            Pp[s] = Pd(P, s);
            Sp[s] = Pd(Sigma, s);
            Fp[s] = Pd(F, s);
            hp[s] = hdiag.as<double>() + static_cast<size_t>(s) * nb;
            Ep[s] = scb[b].as<double>() + static_cast<size_t>(s) * SC_COUNT + SC_E;
        }
        fb[b].setup(Ks, Pp, Sp, Fp, hp, Ep, st);
    }
    const unsigned all = S == 32 ? ~0u : ((1u << S) - 1u);
    LoopScratch L;
    L.partial = part.as<double>();
    L.S = S;
    L.seg = seg_d.as<int>();

    // ---- init (solver_dm.hpp:436-459), every solve
    for (int s = 0; s < S; ++s) h2d(Pd(P, s), P0s[s], nd_of(s) * sizeof(cd), st);
    launch_hermitize(static_cast<int>(nk_tot), nb, P.as<cd>(), st);
    fb[0].build(all, false, st);
    launch_eigh(static_cast<int>(nk_tot), nb, 0, F.as<cd>(), eps.as<double>(), Q.as<cd>(), st);
    mu_solve(eps.as<double>(), mus.as<MuState>(), mus.as<MuState>(), false, all, p.as<double>());

    // per-solve host state
    std::vector<std::vector<double>> hist_E(SS), hist_G(SS);
    std::vector<size_t> ls_trials(SS, 0), k_s(SS, 0);
    std::vector<double> grad_norm(SS, std::numeric_limits<double>::infinity());
    std::vector<double> E_prev(SS, std::numeric_limits<double>::infinity());
    std::vector<double> dE(SS, std::numeric_limits<double>::infinity());
    auto keep_going = [&](int s, size_t k) {
        bool e_nc;
        const std::vector<double>& h = hist_E[s];
        if (cfg.plateau_window > 0 && k > cfg.plateau_window)
            e_nc = (h[k - 1 - cfg.plateau_window] - h[k - 1]) > cfg.tol_E;
        else
            e_nc = dE[s] > cfg.tol_E;
        const bool g_nc = (cfg.tol_grad > 0.0) && (grad_norm[s] > cfg.tol_grad);
        return e_nc || g_nc;
    };
    auto record = [&](int s, const double* h) {
        const double E = h[SC_E];
        grad_norm[s] = std::sqrt(std::max(0.0, h[SC_GN2]));
        hist_E[s].push_back(E);
        hist_G[s].push_back(grad_norm[s]);
        if (h[SC_ACCEPTED] > 0.5 && h[SC_TAU0] > 0.0 && cfg.bt_shrink > 0.0 &&
            cfg.bt_shrink < 1.0) {
            const double j = std::log(h[SC_TAU] / h[SC_TAU0]) / std::log(cfg.bt_shrink);
            ls_trials[s] += static_cast<size_t>(std::llround(std::max(0.0, j))) + 1;
        } else {
            ls_trials[s] += cfg.bt_max;
        }
        dE[s] = std::fabs(E - E_prev[s]);
        E_prev[s] = E;
    };

    cd* const Qs[2] = {Q.as<cd>(), Q_alt.as<cd>()};
    double* const ps[2] = {p.as<double>(), p_alt.as<double>()};
    MuState* const ms[2] = {mus.as<MuState>(), mus_alt.as<MuState>()};
    double* const hs = pinned_batch_scalars(2 * SS * SC_COUNT);
    StreamEvent ev[2];
    ev[0].create();
    ev[1].create();
    const int nkm = nk_max;
    // iteration j over the solves in act: reads state j & 1, writes (j + 1) & 1
    auto enqueue = [&](size_t j, unsigned act) {
        const bool restart = (j == 0) || (cfg.cg_restart > 0 && j % cfg.cg_restart == 0);
        const int s0 = static_cast<int>(j & 1), s1 = s0 ^ 1;
        L.sc = scb[s0].as<double>();
        L.active = act;
        launch_density_seg(L.seg, S, nk_max, static_cast<int>(nk_tot), act, nb, Qs[s0], ps[s0],
                           P.as<cd>(), true, st);
        fb[s0].build(act, true, st);
        launch_orbital_grad(nkm, nb, Qs[s0], F.as<cd>(), ps[s0], wd, ms[s0], T_r, true,
                            eps.as<double>(), G.as<cd>(), g_p.as<double>(), L, st);
        launch_precond(nkm, nb, G.as<cd>(), Gp.as<cd>(), eps.as<double>(), ps[s0],
                       g_p.as<double>(), g_pp.as<double>(), wd, T_r, cfg.denom_scale, dQ.as<cd>(),
                       h_p.as<double>(), L, st);
        launch_cg_combine(nkm, nb, restart, dQ.as<cd>(), dQp.as<cd>(), h_p.as<double>(),
                          d_pp.as<double>(), d_p.as<double>(), eps.as<double>(), ps[s0], wd, T_r,
                          L, st);
        launch_eigh_seg(L.seg, S, nk_max, act, nb, 1, dQ.as<cd>(), lam.as<double>(), Vd.as<cd>(),
                        st);
        launch_linesearch(nkm, nb, Vd.as<cd>(), lam.as<double>(), F.as<cd>(), Sigma.as<cd>(),
                          ps[s0], d_p.as<double>(), wd, T_r, cfg.max_step, cfg.bt_shrink,
                          static_cast<int>(cfg.bt_max), L, st);
        launch_retract(nkm, nb, Qs[s0], Vd.as<cd>(), lam.as<double>(), Sigma.as<cd>(),
                       eps_w.as<double>(), true, L, st, Qs[s1]);
        mu_solve(eps_w.as<double>(), ms[s1], ms[s0], true, act, ps[s1]);
        d2h(hs + static_cast<size_t>(s0) * SS * SC_COUNT, L.sc, SS * SC_COUNT * sizeof(double),
            st);
        HFGPU_CHECK(hipEventRecord(ev[s0].e, st));
        std::swap(dQ, dQp);
        std::swap(d_p, d_pp);
        std::swap(G, Gp);
        std::swap(g_p, g_pp);
    };

    unsigned act = 0;
    for (int s = 0; s < S; ++s)
        if (cfg.max_iter > 0 && keep_going(s, 0)) act |= 1u << s;
    size_t k = 0;
    if (act) {
        enqueue(0, act);
        for (;;) {
            const bool ahead = k + 1 < cfg.max_iter;
            // the solves active at the top of k run iteration k + 1 at once;
            // those that stop at its top are masked from k + 2 on
            if (ahead) enqueue(k + 1, act);
            HFGPU_CHECK(hipEventSynchronize(ev[k & 1].e));
            const double* h = hs + static_cast<size_t>(k & 1) * SS * SC_COUNT;
            for (int s = 0; s < S; ++s) {
                if (!((act >> s) & 1u)) continue;
                record(s, h + static_cast<size_t>(s) * SC_COUNT);
                k_s[s] = k + 1;
                if (!ahead || !keep_going(s, k + 1)) act &= ~(1u << s);
            }
            ++k;
            if (!ahead || !act) break;
        }
    }

    // each solve's state after its k_s iterations is state k_s & 1: gather
    // the odd ones into buffer 0
    for (int s = 0; s < S; ++s) {
        if (!(k_s[s] & 1)) continue;
        d2d(Pd(Q, s), Pd(Q_alt, s), nd_of(s) * sizeof(cd), st);
        d2d(Rd(p, s), Rd(p_alt, s), nkb_of(s) * sizeof(double), st);
        d2d(mus.as<MuState>() + s, mus_alt.as<MuState>() + s, sizeof(MuState), st);
    }

    // ---- finalise (solver_dm.hpp:615-652), every solve
    L.sc = scb[0].as<double>();
    L.active = all;
    launch_density(static_cast<int>(nk_tot), nb, Q.as<cd>(), p.as<double>(), P.as<cd>(), true, st);
    fb[0].build(all, false, st);
    launch_orbital_grad(nkm, nb, Q.as<cd>(), F.as<cd>(), nullptr, wd, mus.as<MuState>(), T_r,
                        false, nullptr, nullptr, nullptr, L, st);
    launch_eigh(static_cast<int>(nk_tot), nb, 0, F.as<cd>(), eps_w.as<double>(), Vd.as<cd>(), st);
    launch_retract(nkm, nb, Q.as<cd>(), Vd.as<cd>(), nullptr, nullptr, nullptr, false, L, st);
    mu_solve(eps_w.as<double>(), mus.as<MuState>(), mus.as<MuState>(), true, all,
             p.as<double>());
    launch_density(static_cast<int>(nk_tot), nb, Q.as<cd>(), p.as<double>(), P.as<cd>(), true, st);
    fb[0].build(all, true, st);

    std::vector<double> hsc(SS * SC_COUNT);
    std::vector<MuState> hmu(SS);
    d2h(hsc.data(), L.sc, SS * SC_COUNT * sizeof(double), st);
    d2h(hmu.data(), mus.get(), SS * sizeof(MuState), st);
    for (int s = 0; s < S; ++s) {
        hf_dm_result* out = outs + s;
        if (out->Q) d2h(out->Q, Pd(Q, s), nd_of(s) * sizeof(cd), st);
        if (out->p) d2h(out->p, Rd(p, s), nkb_of(s) * sizeof(double), st);
        if (out->density) d2h(out->density, Pd(P, s), nd_of(s) * sizeof(cd), st);
        if (out->fock) d2h(out->fock, Pd(F, s), nd_of(s) * sizeof(cd), st);
    }
    HFGPU_CHECK(hipStreamSynchronize(st));
    for (int s = 0; s < S; ++s) {
        hf_dm_result* out = outs + s;
        out->mu = hmu[s].mu;
        out->energy = hsc[static_cast<size_t>(s) * SC_COUNT + SC_E];
        out->n_iter = k_s[s];
        out->converged = k_s[s] < cfg.max_iter ? 1 : 0;
        out->n_ls_trials = ls_trials[s];
        for (size_t i = 0; i < k_s[s]; ++i) {
            if (out->hist_E) out->hist_E[i] = hist_E[s][i];
            if (out->hist_grad) out->hist_grad[i] = hist_G[s][i];
        }
    }
}

// ---------------------------------------------------------------------------
// Trust-region Newton (solver_rtr.hpp:162-459)
// ---------------------------------------------------------------------------

namespace {

// One point (Q, p) of the trust-region iteration and what eval_state derives
// from it: Ft = Q^H F Q, eps = Re diag Ft, G, and mu.  The solver keeps the
// accepted point and a trial point and swaps them on acceptance, so a
// rejected step costs no copies (solver_rtr.hpp:383-428).
struct RtrPoint {
    DevBuf Q, p, Ft, G, eps, mus;
    RtrPoint(size_t bytes, size_t rb)
        : Q(bytes), p(rb), Ft(bytes), G(bytes), eps(rb), mus(sizeof(MuState)) {}
};

// Slots of the solver's own scalar array.
enum RtrSlot : int {
    RS_UPD = 0,
    RS_HVP = 8,
    RS_DEFL = 16,
    RS_STEP = 18,  // TcgStepSlot
    RS_D2 = 24,
    RS_FIXED = 24
};
static_assert(RS_UPD + RU_COUNT <= RS_HVP && RS_HVP + RH_COUNT <= RS_DEFL &&
                  RS_STEP + TS_COUNT <= RS_D2,
              "trust-region slot layout");

}  // namespace

void solve_rtr_gpu(GpuKernel& K, const double* P0, double n_e, const hf_dm_config& cfg,
                   Projector& proj, hf_dm_result* out) {
    const int nk = static_cast<int>(K.nk), nb = K.nb;
    const size_t nd = K.nd, bytes = nd * sizeof(cd), nkb = K.nk * K.nb;
    const size_t rb = nkb * sizeof(double);
    hipStream_t st = K.stream;
    SolveWs W(st);
    double* sc = W.sc.as<double>();

    const double wsum = std::max(K.weight_sum, TINY_REAL);
    std::vector<double> wn(K.w_host.size());
    for (size_t i = 0; i < wn.size(); ++i) wn[i] = K.w_host[i] / wsum;
    double w_total = 0.0;  // project_occ's denominator / nb (solver_rtr.hpp:42-50)
    for (double x : K.w_host) w_total += x;
    const double occ_den = w_total * static_cast<double>(nb);
    const double n_target = n_e / wsum;
    const double T_r = std::max(K.T, 1.0e-12);
    const double tol_g = cfg.tol_grad > 0.0 ? cfg.tol_grad : 1.0e-6;

    const size_t n_defl = cfg.n_deflation;
    const bool defl_on = n_defl > 0 && cfg.deflation_sigma > 0.0 && cfg.deflation_targets;
    const double defl_len = std::max(cfg.deflation_length, 1.0e-12);

    DevBuf w_norm(K.nk * sizeof(double));
    h2d(w_norm.get(), wn.data(), K.nk * sizeof(double), st);
    const double* wd = w_norm.as<double>();
    const double* w2d = K.w2d.as<double>();
    RtrPoint A(bytes, rb), B(bytes, rb);
    RtrPoint* cur = &A;
    RtrPoint* trial = &B;
    DevBuf P(bytes), Sigma(bytes), hdiag(nb * sizeof(double)), gp(rb);
    DevBuf vX(bytes), HvX(bytes), rX(bytes), dX(bytes), HdX(bytes);
    DevBuf vp(rb), Hvp(rb), rp(rb), dpv(rb), Hdp(rb);
    DevBuf dPb(bytes), Fb(bytes), Sb(bytes), hb(nb * sizeof(double));
    DevBuf Vd(bytes), lam(rb), eps_t(rb);
    DevBuf rsc((RS_FIXED + (defl_on ? n_defl : 0)) * sizeof(double));
    DevBuf tgt(defl_on ? n_defl * bytes : 0);
    if (defl_on) h2d(tgt.get(), cfg.deflation_targets, n_defl * bytes, st);
    double* rs = rsc.as<double>();
    fill_zero(rs, rsc.size(), st);

    std::vector<double> hrs(rsc.size() / sizeof(double));
    double hsc[SC_COUNT];
    auto read_back = [&]() {
        d2h(hsc, sc, sizeof(hsc), st);
        d2h(hrs.data(), rs, rsc.size(), st);
        HFGPU_CHECK(hipStreamSynchronize(st));
    };

    // density, Fock (+ deflation bias), Ft, eps and G at a point
    // (solver_rtr.hpp:275-293); energy, |G|_w^2 and sum eps^2 land in sc
    auto eval_state = [&](RtrPoint& s) {
        // the hermitising density kernel equals X followed by a hermitisation
        // bit for bit; a projector acts on X itself (solver_dm.hpp:117-128)
        launch_density(nk, nb, s.Q.as<cd>(), s.p.as<double>(), P.as<cd>(), !proj, st);
        if (proj) {
            proj.apply(P.as<cd>(), st);
            launch_hermitize(nk, nb, P.as<cd>(), st);
        }
        K.build_fock(P.as<cd>(), Sigma.as<cd>(), s.Ft.as<cd>(), hdiag.as<double>(), true, st);
        K.energy(P.as<cd>(), Sigma.as<cd>(), hdiag.as<double>(), sc + SC_E, st);
        if (defl_on)
            launch_deflation(nk, nb, static_cast<int>(n_defl), P.as<cd>(), tgt.as<cd>(), w2d,
                             cfg.deflation_sigma, defl_len, rs + RS_D2, s.Ft.as<cd>(),
                             rs + RS_DEFL, W.ls, st);
        launch_orbital_grad(nk, nb, s.Q.as<cd>(), s.Ft.as<cd>(), s.p.as<double>(), wd,
                            s.mus.as<MuState>(), T_r, true, s.eps.as<double>(), s.G.as<cd>(),
                            gp.as<double>(), W.ls, st);
    };
    struct EvalOut {
        double E, gn, eps_sq, pen;
    };
    auto eval_result = [&]() {
        read_back();
        EvalOut r;
        r.E = hsc[SC_E] + (defl_on ? hrs[RS_DEFL] : 0.0);
        r.gn = std::sqrt(std::max(0.0, hsc[SC_GN2]));
        r.eps_sq = hsc[SC_EPS_SQ];
        r.pen = defl_on ? hrs[RS_DEFL + 1] : 0.0;
        return r;
    };

    // ---- init (solver_rtr.hpp:262-272)
    h2d(P.get(), P0, bytes, st);
    launch_hermitize(nk, nb, P.as<cd>(), st);
    K.build_fock(P.as<cd>(), Sigma.as<cd>(), Fb.as<cd>(), hdiag.as<double>(), true, st);
    launch_eigh(nk, nb, 0, Fb.as<cd>(), cur->eps.as<double>(), cur->Q.as<cd>(), st);
    fill_zero(cur->mus.get(), sizeof(MuState), st);
    launch_mu_newton(cur->eps.as<double>(), wd, nk, nb, n_target, 0.0, false, T_r,
                     cfg.mu_maxiter, cur->mus.as<MuState>(), W.mu_partial.as<double>(), st,
                     nullptr, cur->p.as<double>());
    eval_state(*cur);
    EvalOut now = eval_result();

    std::vector<double> hist_E, hist_G;
    double Delta = cfg.tr_delta0;
    size_t k = 0;
    bool converged = false;
    while (k < cfg.max_iter) {
        hist_E.push_back(now.E);
        hist_G.push_back(now.gn);
        if (now.gn < tol_g) {
            converged = true;
            ++k;
            break;
        }
        const double lam_pc = std::max(
            T_r, cfg.denom_scale *
                     std::sqrt(now.eps_sq / static_cast<double>(std::max<size_t>(nkb, 1)) +
                               TINY_REAL));
        const cd* Qc = cur->Q.as<cd>();
        const double* pc = cur->p.as<double>();
        const cd* Ftc = cur->Ft.as<cd>();
        const double* epc = cur->eps.as<double>();
        // step: the device's decision in rs + RS_STEP (mode, a and s_H are
        // then read there), or nullptr with the host's values
        auto update = [&](int mode, double a, double s_H, const double* step) {
            launch_rtr_update(nk, nb, mode, a, s_H, lam_pc, T_r, epc, pc, cur->G.as<cd>(), w2d,
                              dX.as<cd>(), dpv.as<double>(), HdX.as<cd>(), Hdp.as<double>(),
                              vX.as<cd>(), vp.as<double>(), HvX.as<cd>(), Hvp.as<double>(),
                              rX.as<cd>(), rp.as<double>(), W.ls, rs + RS_UPD, st, step);
        };
        auto hvp = [&](bool init, double beta, double s_z) {
            launch_rtr_direction(nk, nb, init, beta, s_z, lam_pc, T_r, Qc, pc, epc, rX.as<cd>(),
                                 rp.as<double>(), dX.as<cd>(), dpv.as<double>(), dPb.as<cd>(),
                                 st);
            K.build_fock(dPb.as<cd>(), Sb.as<cd>(), Fb.as<cd>(), hb.as<double>(), true, st,
                         false);
            launch_rtr_hvp(nk, nb, T_r, now.pen, Qc, pc, Ftc, dX.as<cd>(), dpv.as<double>(),
                           Fb.as<cd>(), K.h.as<cd>(), dPb.as<cd>(), vX.as<cd>(),
                           vp.as<double>(), w2d, HdX.as<cd>(), Hdp.as<double>(), W.ls,
                           rs + RS_HVP, st);
        };

        // ---- Steihaug truncated CG (solver_rtr.hpp:283-346).  Each step's
        // decision (interior alpha or the boundary step t) is taken on the
        // device from the HVP's sums, with the host's arithmetic, and the
        // update applies it at once: one host read-back per CG step.
        update(0, 0.0, 0.0, nullptr);  // v = Hv = 0, r = -G
        double rz = 0.0, gv = 0.0, vhv = 0.0, vv = 0.0;
        double beta = 0.0, s_z = 0.0;
        for (size_t it = 0; it < cfg.tr_cg_max; ++it) {
            hvp(it == 0, beta, s_z);
            // on the first step <r, z> (r = (-G, 0)) comes from update 0
            launch_tcg_step(rs + RS_HVP, rs + RS_UPD, it == 0 ? 1 : 0, rz, occ_den, Delta,
                            rs + RS_STEP, st);
            update(1, 0.0, 0.0, rs + RS_STEP);
            read_back();
            const double* U = hrs.data() + RS_UPD;
            rz = hrs[RS_STEP + TS_RZ];
            gv = U[RU_GV];
            vhv = U[RU_VHV];
            vv = U[RU_VV];
            if (hrs[RS_STEP + TS_BOUND] != 0.0) break;
            if (std::sqrt(std::max(0.0, U[RU_RN2])) < 0.05 * now.gn) break;
            s_z = U[RU_ZP] / occ_den;
            const double rz_new = U[RU_RZX] + (U[RU_RZP] - s_z * U[RU_RP]);
            beta = rz_new / rz;
            rz = rz_new;
        }

        // predicted reduction in energy units (solver_rtr.hpp:348-361)
        const double pred = -(gv + 0.5 * vhv);
        if (Delta < 1e-10) break;

        // ---- trial point: Q Cayley(tau = -1, v), frozen-Ft energies, mu
        // (solver_rtr.hpp:363-385), then eval_state at it
        launch_eigh(nk, nb, 1, vX.as<cd>(), lam.as<double>(), Vd.as<cd>(), st);
        launch_rtr_retract(nk, nb, -1.0, Vd.as<cd>(), lam.as<double>(), Qc, Ftc,
                           trial->Q.as<cd>(), eps_t.as<double>(), st);
        d2d(trial->mus.get(), cur->mus.get(), sizeof(MuState), st);
        launch_mu_newton(eps_t.as<double>(), wd, nk, nb, n_target, 0.0, true, T_r,
                         cfg.mu_maxiter, trial->mus.as<MuState>(), W.mu_partial.as<double>(), st,
                         nullptr, trial->p.as<double>());
        eval_state(*trial);
        const EvalOut tr = eval_result();

        // ---- acceptance and trust radius (solver_rtr.hpp:398-420)
        const double noise = 1e-12 * std::max(std::fabs(now.E), 1.0);
        bool accept;
        if (pred <= noise) {
            accept = tr.gn <= now.gn && tr.E <= now.E + noise;
            if (!accept) Delta *= 0.5;
        } else {
            const double ratio = (now.E - tr.E) / pred;
            if (ratio < 0.25)
                Delta *= 0.25;
            else if (ratio > 0.75 && std::sqrt(std::max(0.0, vv)) > 0.9 * Delta)
                Delta = std::min(2.0 * Delta, 5.0);
            accept = ratio > 0.1;
        }
        if (accept) {
            std::swap(cur, trial);
            now = tr;
        }
        ++k;
    }

    // ---- finalise (solver_rtr.hpp:435-441)
    // the hermitising density kernel equals X followed by a hermitisation
    // bit for bit; a projector acts on X itself (solver_dm.hpp:117-128)
    launch_density(nk, nb, cur->Q.as<cd>(), cur->p.as<double>(), P.as<cd>(), !proj, st);
    if (proj) {
        proj.apply(P.as<cd>(), st);
        launch_hermitize(nk, nb, P.as<cd>(), st);
    }
    K.build_fock(P.as<cd>(), Sigma.as<cd>(), Fb.as<cd>(), hdiag.as<double>(), true, st);
    K.energy(P.as<cd>(), Sigma.as<cd>(), hdiag.as<double>(), sc + SC_E, st);

    MuState hmu;
    d2h(hsc, sc, sizeof(hsc), st);
    d2h(&hmu, cur->mus.get(), sizeof(MuState), st);
    if (out->Q) d2h(out->Q, cur->Q.get(), bytes, st);
    if (out->p) d2h(out->p, cur->p.get(), rb, st);
    if (out->density) d2h(out->density, P.get(), bytes, st);
    if (out->fock) d2h(out->fock, Fb.get(), bytes, st);
    HFGPU_CHECK(hipStreamSynchronize(st));
    out->mu = hmu.mu;
    out->energy = hsc[SC_E];
    out->n_iter = k;
    out->converged = converged ? 1 : 0;
    out->n_ls_trials = 0;
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
        SymPlan sp;
        attach_symmetry(pr, sp, cfg->symmetry, K);
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
        SymPlan sp;
        attach_symmetry(pr, sp, cfg->symmetry, K);
        if (cfg->optimizer == 1)
            solve_rtr_gpu(K, P0, n_e, *cfg, pr, out);
        else
            solve_dm_gpu(K, P0, n_e, *cfg, pr, out);
    });
}

int hfgpu_solve_dm_batch(hfgpu_kernel* const* ks, size_t n, const double* const* P0s,
                         const double* n_es, const hf_dm_config* cfg, hf_dm_result* outs) {
    return guarded([&] {
        if (!ks || !P0s || !n_es || !cfg || !outs) throw std::invalid_argument("null argument");
        if (n < 1 || n > static_cast<size_t>(kMaxBatch))
            throw std::invalid_argument("hfgpu_solve_dm_batch takes 1 to 32 solves");
        // each solve builds in its kernel's workspaces, so one handle cannot
        // serve two solves of a batch
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < i; ++j)
                if (ks[i] && ks[i] == ks[j])
                    throw std::invalid_argument(
                        "hfgpu_solve_dm_batch: a kernel handle appears twice in the batch");
        std::vector<GpuKernel*> Ks(n);
        std::vector<const double*> P0(n);
        std::vector<double> ne(n);
        for (size_t i = 0; i < n; ++i) {
            if (!ks[i] || !P0s[i] || !outs[i].p) throw std::invalid_argument("null argument");
            Ks[i] = ks[i]->impl.get();
            P0[i] = P0s[i];
            ne[i] = n_es[i];
        }
        solve_dm_batch_gpu(Ks, P0, ne, *cfg, outs);
    });
}

}  // extern "C"
