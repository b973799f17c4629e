// ORACLE — test infrastructure only.  A standalone driver that calls every
// mesh entry point of the CPU restatement (hfcpu.h) on small, seeded inputs:
// built with AddressSanitizer and UndefinedBehaviorSanitizer (the Makefile's
// `asan` target), so an out-of-bounds access, a use after free or undefined
// arithmetic anywhere on those paths aborts the run.  Odd meshes (5 x 6)
// exercise the mixed-radix transforms, 3 bands the dense and packed
// channels, a contact term and a reference density every Fock term.
// Prints "asan check ok" and exits 0 when every call returned its expected
// status.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "hfcpu.h"

namespace {

using c64 = std::complex<double>;

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 1) {}
    double uniform() {  // xorshift64*, in [-1, 1)
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return static_cast<double>((s * 0x2545F4914F6CDD1Dull) >> 11) * (2.0 / 9007199254740992.0) -
               1.0;
    }
};

int g_fail = 0;

void expect(int rc, int want, const char* what) {
    if (rc != want) {
        std::fprintf(stderr, "%s: status %d (want %d): %s\n", what, rc, want, hfcpu_last_error());
        ++g_fail;
    }
}

const double* D(const std::vector<c64>& v) { return reinterpret_cast<const double*>(v.data()); }
double* D(std::vector<c64>& v) { return reinterpret_cast<double*>(v.data()); }

// nk Hermitian nb x nb blocks: scale * (X + X^H) / 2 plus shift on the diagonal
std::vector<c64> hermitian(Rng& r, size_t nk, size_t nb, double scale, double shift) {
    std::vector<c64> a(nk * nb * nb);
    for (size_t k = 0; k < nk; ++k)
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = i; j < nb; ++j) {
                const c64 x = i == j ? c64(r.uniform(), 0.0) : c64(r.uniform(), r.uniform());
                a[(k * nb + i) * nb + j] = scale * x;
                a[(k * nb + j) * nb + i] = scale * std::conj(x);
                if (i == j) a[(k * nb + i) * nb + j] += shift;
            }
    return a;
}

}  // namespace

int main() {
    Rng rng(7);
    const size_t nk1 = 5, nk2 = 6, nb = 3, nk = nk1 * nk2, nd = nk * nb * nb;

    // ---- model: dense V, Hartree, reference density, one contact term
    std::vector<c64> h = hermitian(rng, nk, nb, 0.5, 0.0);
    std::vector<c64> VR(nd), VRs(nk);
    for (size_t q = 0; q < nk; ++q) {
        const double v = 1.0 + 0.3 * rng.uniform();
        VRs[q] = c64(v, 0.0);
        for (size_t a = 0; a < nb; ++a)
            for (size_t b = 0; b < nb; ++b)
                VR[(q * nb + a) * nb + b] = c64(v * (1.0 - 0.1 * std::abs(double(a) - double(b))), 0.0);
    }
    std::vector<c64> refP(nd, c64(0.0, 0.0));
    for (size_t k = 0; k < nk; ++k)
        for (size_t a = 0; a < nb; ++a) refP[(k * nb + a) * nb + a] = 0.5;
    std::vector<double> w(nk, 1.0 / nk), HH(nb * nb, 0.7), cg{0.2};
    std::vector<c64> Oi(nb * nb, 0.0), Oj(nb * nb, 0.0);
    for (size_t a = 0; a < nb; ++a) {
        Oi[a * nb + a] = a == 0 ? 1.0 : 0.0;
        Oj[a * nb + a] = a == 1 ? 1.0 : 0.0;
    }
    hf_kernel_desc kd{};
    kd.nk1 = nk1;
    kd.nk2 = nk2;
    kd.nb = nb;
    kd.dv1 = nb;
    kd.dv2 = nb;
    kd.n_contact = 1;
    kd.include_hartree = 1;
    kd.include_exchange = 1;
    kd.exchange_hcp = 0;
    kd.has_refP = 1;
    kd.T = 0.1;
    kd.weight_sum = 1.0;
    kd.h = D(h);
    kd.VR = D(VR);
    kd.refP = D(refP);
    kd.w2d = w.data();
    kd.HH = HH.data();
    kd.contact_g = cg.data();
    kd.contact_Oi = D(Oi);
    kd.contact_Oj = D(Oj);
    kd.superlattice = nullptr;
    hf_kernel_desc kh = kd;  // scalar V on the packed (HCP) channels, no Hartree
    kh.dv1 = kh.dv2 = 1;
    kh.VR = D(VRs);
    kh.exchange_hcp = 1;
    kh.include_hartree = 0;
    kh.has_refP = 0;
    kh.refP = nullptr;
    kh.n_contact = 0;

    std::vector<c64> P = hermitian(rng, nk, nb, 0.1, 0.5);
    const double n_e = 1.4;

    // ---- transforms and the exchange
    {
        std::vector<c64> buf = P;
        expect(hfcpu_fft2d(D(buf), nk1, nk2, nb * nb, 0), 0, "fft2d forward");
        expect(hfcpu_fft2d(D(buf), nk1, nk2, nb * nb, 1), 0, "fft2d inverse");
        std::vector<c64> sig(nd);
        expect(hfcpu_selfenergy_fft_full(D(VR), D(P), D(sig), nk1, nk2, nb, nb, nb, 1, 0), 0,
               "selfenergy dense");
        expect(hfcpu_selfenergy_fft_full(D(VRs), D(P), D(sig), nk1, nk2, nb, 1, 1, 0, 1), 0,
               "selfenergy hcp");
    }
    std::vector<c64> Sg(nd), Hh(nd), F(nd);
    double E = 0.0;
    expect(hfcpu_build_fock_apply(&kd, D(P), D(Sg), D(Hh), D(F), &E, nullptr, nullptr), 0,
           "build_fock dense");
    expect(hfcpu_build_fock_apply(&kh, D(P), D(Sg), D(Hh), D(F), &E, nullptr, nullptr), 0,
           "build_fock hcp");

    // ---- eigh, mu, density, Cayley hooks
    std::vector<double> ev(nk * nb), lam(nk * nb), diag(nk * nb), occ(nk * nb, 0.5);
    std::vector<c64> V(nd), U(nd);
    expect(hfcpu_eigh_batched(D(F), ev.data(), D(V), nk, nb), 0, "eigh");
    const size_t sizes[2] = {1, 2};
    expect(hfcpu_eigh_block_sizes(D(F), ev.data(), D(V), nk, nb, sizes, 2, 1), 0,
           "eigh blocks");
    double off[2];
    expect(hfcpu_max_offblock_sizes(D(F), nk, nb, sizes, 2, off), 0, "max offblock");
    double mu = 0.0;
    expect(hfcpu_find_mu_bisection(ev.data(), w.data(), nk, nb, n_e, 0.1, 30, &mu), 0,
           "bisection");
    expect(hfcpu_find_mu_newton(ev.data(), w.data(), nk, nb, n_e, 0.1, 15, &mu), 0, "newton");
    expect(hfcpu_solve_mu_inloop(ev.data(), w.data(), nk, nb, n_e, 0.0, 0.1, 25, &mu), 0,
           "newton in loop");
    expect(hfcpu_validate_electron_count(w.data(), nk, nb, n_e), 0, "electron count");
    expect(hfcpu_validate_electron_count(w.data(), nk, nb, 10.0 * nb) != 0 ? 1 : 0, 1,
           "electron count out of range");
    std::vector<c64> Pn(nd);
    expect(hfcpu_density_from_fock(D(F), nk, nb, w.data(), n_e, 0.1, 0.2, D(Pn), &mu), 0,
           "density_from_fock");
    std::vector<c64> d = hermitian(rng, nk, nb, 0.2, 0.0);
    expect(hfcpu_cayley_spectral_setup(D(d), D(V), lam.data(), nk, nb), 0, "cayley setup");
    expect(hfcpu_cayley_unitary_from_spectrum(D(V), lam.data(), 0.3, D(U), nk, nb), 0,
           "cayley unitary");
    expect(hfcpu_diag_UFU_from_spectrum(D(V), D(F), lam.data(), 0.3, diag.data(), nk, nb), 0,
           "diag UFU");
    double S = 0.0;
    expect(hfcpu_occupation_entropy(occ.data(), w.data(), nk, nb, &S), 0, "entropy");
    {
        std::vector<c64> x = hermitian(rng, nk1 * nk1, nb, 1.0, 0.0), y(9 * 9 * nb * nb);
        expect(hfcpu_resample_kgrid_2d(D(x), D(y), nk1, 9, nb * nb), 0, "resample");
    }

    // ---- solvers: SCF linear / DIIS / ODA (trust radius), DM CG / Newton /
    // Newton with a deflation target
    std::vector<double> hE(16), hD(16), hC(16);
    for (int acc = 0; acc < 3; ++acc) {
        hf_scf_config c{};
        c.max_iter = 12;
        c.density_tol = 1e-9;
        c.comm_tol = 1e-9;
        c.mixing = 0.5;
        c.acceleration = acc;
        c.diis_size = 4;
        c.diis_start = 1;
        c.diis_damping = 0.1;
        c.trust_radius = acc == 2 ? 0.05 : 0.0;
        hf_scf_result r{};
        std::vector<c64> Dout(nd), Fout(nd);
        r.density = D(Dout);
        r.fock = D(Fout);
        r.hist_E = hE.data();
        r.hist_density = hD.data();
        r.hist_comm = hC.data();
        expect(hfcpu_solve_scf(&kd, D(P), n_e, &c, nullptr, nullptr, &r), 0, "solve_scf");
    }
    std::vector<c64> target = Pn;
    for (int mode = 0; mode < 3; ++mode) {
        hf_dm_config c{};
        c.max_iter = mode == 0 ? 12 : 4;
        c.tol_E = 1e-12;
        c.plateau_window = 5;
        c.denom_scale = 1e-3;
        c.max_step = 0.6;
        c.cg_restart = 10;
        c.bt_shrink = 0.5;
        c.bt_max = 8;
        c.mu_maxiter = 25;
        c.optimizer = mode == 0 ? 0 : 1;
        c.tr_delta0 = 0.5;
        c.tr_cg_max = 6;
        if (mode == 2) {
            c.deflation_targets = D(target);
            c.n_deflation = 1;
            c.deflation_sigma = 0.1;
            c.deflation_length = 0.5;
        }
        hf_dm_result r{};
        std::vector<c64> Q(nd), Dout(nd), Fout(nd);
        std::vector<double> p(nk * nb);
        r.Q = D(Q);
        r.p = p.data();
        r.density = D(Dout);
        r.fock = D(Fout);
        r.hist_E = hE.data();
        r.hist_grad = hD.data();
        expect(hfcpu_solve_dm(&kd, D(P), n_e, &c, nullptr, nullptr, &r), 0, "solve_dm");
    }

    if (g_fail) {
        std::fprintf(stderr, "asan check: %d call(s) failed\n", g_fail);
        return 1;
    }
    std::printf("asan check ok\n");
    return 0;
}
