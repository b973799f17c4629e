// ORACLE — test infrastructure only.  CPU FP64 restatement of the reference's
// Hartree-Fock hot path (skilledwolf/cpp_hf, cpp/include/cpp_hf/*.hpp), used
// by tests/ as the parity checker and by bench.py as the CPU baseline.  It is
// never linked into, loaded by, or called from the cpp_hf_amd product path.
//
// The reference calls FFTW (fft.hpp) and Eigen (linalg.hpp); neither exists in
// this image, so this file carries its own mixed-radix Stockham FFT and its
// own cyclic Jacobi Hermitian eigensolver.  Those two are pinned by the tests
// against numpy.fft and numpy.linalg.eigh (LAPACK), and by brute-force
// O(N^2) convolution; everything else follows the reference line by line:
// same constants, same clip points, same hermitisation points and the same
// serial reduction order (parallel.hpp:3-10: only disjoint-output loops are
// threaded, every reduction is serial).
//
// Functions whose parity is pinned only by following the source (no golden
// vector, no closed form): the iteration trajectories of solve_dm and
// solve_scf (n_iter, histories).  Their fixed points are pinned by
// stationarity tests and by the meshgrid regression fixture.

#include "hfcpu.h"

#include <algorithm>
#include <map>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using c64 = std::complex<double>;

// types.hpp:13-16
constexpr double TINY_REAL = 1.0e-30;
constexpr double LOGIT_CLIP = 1.0e-6;
constexpr double ENTROPY_CLIP = 1.0e-14;
constexpr double OCC_CLIP = 1.0e-8;

thread_local std::string g_last_error;

// parallel.hpp:22-31 (static partition over disjoint outputs)
template <class F>
void parallel_for(size_t n, F&& fn) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
    for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(n); ++i)
        fn(static_cast<size_t>(i));
#else
    for (size_t i = 0; i < n; ++i) fn(i);
#endif
}

// ---------------------------------------------------------------------------
// FFT: mixed-radix Stockham autosort, applied along the leading axis of an
// (n, vw) array for all vw columns at once (the vw axis is contiguous).
// Replaces FftBatched2D (fft.hpp:25-142): forward sign -1, inverse sign +1
// with 1/(nk1*nk2) normalisation (fft.hpp:102-111).
// ---------------------------------------------------------------------------

struct Plan1D {
    size_t n = 0;
    std::vector<int> radices;
    std::vector<c64> tw;  // tw[m] = exp(-2 pi i m / n)
};

Plan1D make_plan(size_t n) {
    Plan1D p;
    p.n = n;
    size_t m = n;
    while (m % 4 == 0) { p.radices.push_back(4); m /= 4; }
    while (m % 2 == 0) { p.radices.push_back(2); m /= 2; }
    while (m % 3 == 0) { p.radices.push_back(3); m /= 3; }
    while (m % 5 == 0) { p.radices.push_back(5); m /= 5; }
    for (size_t f = 7; m > 1; f += 2) {
        while (m % f == 0) { p.radices.push_back(static_cast<int>(f)); m /= f; }
    }
    p.tw.resize(n);
    for (size_t k = 0; k < n; ++k) {
        const long double a = -2.0L * 3.141592653589793238462643383279502884L *
                              static_cast<long double>(k) / static_cast<long double>(n);
        p.tw[k] = c64(static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a)));
    }
    return p;
}

// Small forward DFT of length r on v[0..r) (sign -1).
inline void dft_small(c64* v, int r, const Plan1D& plan) {
    if (r == 1) return;
    if (r == 2) {
        const c64 a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
        return;
    }
    if (r == 4) {
        const c64 a0 = v[0] + v[2], a1 = v[0] - v[2];
        const c64 b0 = v[1] + v[3], b1 = v[1] - v[3];
        const c64 b1m(b1.imag(), -b1.real());  // -i * b1
        v[0] = a0 + b0;
        v[2] = a0 - b0;
        v[1] = a1 + b1m;
        v[3] = a1 - b1m;
        return;
    }
    if (r == 3) {
        const double c3 = -0.5;
        const double s3 = -0.86602540378443864676372317075293618;  // -sin(2pi/3)
        const c64 t1 = v[1] + v[2];
        const c64 t2 = v[1] - v[2];
        const c64 m1 = v[0] + c3 * t1;
        const c64 m2(-s3 * t2.imag(), s3 * t2.real());  // i*s3*t2
        v[0] = v[0] + t1;
        v[1] = m1 + m2;
        v[2] = m1 - m2;
        return;
    }
    // generic r: direct DFT using the plan's n-th roots (r divides n)
    c64 out[64];
    const size_t step = plan.n / static_cast<size_t>(r);
    for (int m = 0; m < r; ++m) {
        c64 s(0.0, 0.0);
        for (int t = 0; t < r; ++t)
            s += v[t] * plan.tw[(static_cast<size_t>(t) * m % r) * step];
        out[m] = s;
    }
    for (int m = 0; m < r; ++m) v[m] = out[m];
}

// Forward FFT along axis 0 of x (n, vw), result left in x.  work: n*vw.
void fft_axis0(c64* x, size_t n, size_t vw, const Plan1D& plan, c64* work) {
    if (n == 1) return;
    c64* src = x;
    c64* dst = work;
    size_t Ns = 1;
    c64 v[64];
    for (int r : plan.radices) {
        if (r > 64) throw std::invalid_argument("FFT radix > 64 is not supported");
        const size_t nr = n / static_cast<size_t>(r);
        const size_t tw_step = n / (Ns * static_cast<size_t>(r));
        for (size_t j = 0; j < nr; ++j) {
            const size_t k = j % Ns;
            const size_t idxD = (j / Ns) * Ns * static_cast<size_t>(r) + k;
            for (size_t col = 0; col < vw; ++col) {
                for (int t = 0; t < r; ++t) {
                    c64 val = src[(j + static_cast<size_t>(t) * nr) * vw + col];
                    if (t > 0 && k > 0)
                        val *= plan.tw[(static_cast<size_t>(t) * k * tw_step) % n];
                    v[t] = val;
                }
                dft_small(v, r, plan);
                for (int t = 0; t < r; ++t)
                    dst[(idxD + static_cast<size_t>(t) * Ns) * vw + col] = v[t];
            }
        }
        std::swap(src, dst);
        Ns *= static_cast<size_t>(r);
    }
    if (src != x) std::memcpy(x, src, n * vw * sizeof(c64));
}

// In-place 2-D FFT over axes (0, 1) of a (nk1, nk2, batch) array.
// inverse: conj(FFT(conj(x))) / (nk1*nk2)  == numpy.fft.ifftn semantics.
void fft2d(c64* buf, size_t nk1, size_t nk2, size_t batch, bool inverse) {
    const size_t n_tot = nk1 * nk2 * batch;
    if (inverse)
        for (size_t i = 0; i < n_tot; ++i) buf[i] = std::conj(buf[i]);
    const Plan1D p2 = make_plan(nk2);
    const Plan1D p1 = make_plan(nk1);
    // axis 1: per k1 slab (nk2, batch)
    parallel_for(nk1, [&](size_t k1) {
        std::vector<c64> work(nk2 * batch);
        fft_axis0(buf + k1 * nk2 * batch, nk2, batch, p2, work.data());
    });
    // axis 0: (nk1, nk2*batch), split the column range across threads
    const size_t row = nk2 * batch;
    const size_t chunk = 64;
    const size_t n_chunks = (row + chunk - 1) / chunk;
    parallel_for(n_chunks, [&](size_t ci) {
        const size_t c0 = ci * chunk;
        const size_t cw = std::min(chunk, row - c0);
        std::vector<c64> tmp(nk1 * cw), work(nk1 * cw);
        for (size_t i = 0; i < nk1; ++i)
            for (size_t c = 0; c < cw; ++c) tmp[i * cw + c] = buf[i * row + c0 + c];
        fft_axis0(tmp.data(), nk1, cw, p1, work.data());
        for (size_t i = 0; i < nk1; ++i)
            for (size_t c = 0; c < cw; ++c) buf[i * row + c0 + c] = tmp[i * cw + c];
    });
    if (inverse) {
        const double s = 1.0 / static_cast<double>(nk1 * nk2);
        for (size_t i = 0; i < n_tot; ++i) buf[i] = std::conj(buf[i]) * s;
    }
}

// ---------------------------------------------------------------------------
// Exchange self-energy (selfenergy.hpp:60-190)
// ---------------------------------------------------------------------------

// selfenergy.hpp:60-81
void multiply_VR_inplace(c64* P, const c64* VR, size_t nk1, size_t nk2, size_t nb,
                         size_t dv1, size_t dv2) {
    const size_t nk = nk1 * nk2, nb2 = nb * nb;
    if (dv1 == 1 && dv2 == 1) {
        for (size_t k = 0; k < nk; ++k) {
            const c64 v = VR[k];
            for (size_t i = 0; i < nb2; ++i) P[k * nb2 + i] *= v;
        }
    } else if (dv1 == nb && dv2 == nb) {
        for (size_t k = 0; k < nk; ++k)
            for (size_t i = 0; i < nb2; ++i) P[k * nb2 + i] *= VR[k * nb2 + i];
    } else {
        throw std::invalid_argument("VR shape must be (..., 1, 1) or (..., nb, nb).");
    }
}

// selfenergy.hpp:105-117
void selfenergy_fft_full_inplace(c64* sigma, const c64* VR, size_t nk1, size_t nk2,
                                 size_t nb, size_t dv1, size_t dv2) {
    const size_t nb2 = nb * nb, n_tot = nk1 * nk2 * nb2;
    fft2d(sigma, nk1, nk2, nb2, false);
    multiply_VR_inplace(sigma, VR, nk1, nk2, nb, dv1, dv2);
    fft2d(sigma, nk1, nk2, nb2, true);
    for (size_t i = 0; i < n_tot; ++i) sigma[i] = -sigma[i];
}

// selfenergy.hpp:91-100
void selfenergy_fft_full(const c64* P, c64* sigma, const c64* VR, size_t nk1,
                         size_t nk2, size_t nb, size_t dv1, size_t dv2) {
    std::memcpy(sigma, P, nk1 * nk2 * nb * nb * sizeof(c64));
    selfenergy_fft_full_inplace(sigma, VR, nk1, nk2, nb, dv1, dv2);
}

// selfenergy.hpp:123-171 (Hermitian channel packing; assumes P Hermitian)
void selfenergy_fft_full_hcp(const c64* P, c64* sigma, const c64* VR, size_t nk1,
                             size_t nk2, size_t nb) {
    const size_t nb2 = nb * nb, nk = nk1 * nk2;
    std::vector<size_t> ti, tj;
    for (size_t i = 0; i < nb; ++i)
        for (size_t j = i; j < nb; ++j) { ti.push_back(i); tj.push_back(j); }
    const size_t np = ti.size();
    std::vector<c64> packed(nk * np);
    for (size_t k = 0; k < nk; ++k)
        for (size_t t = 0; t < np; ++t) packed[k * np + t] = P[k * nb2 + ti[t] * nb + tj[t]];
    fft2d(packed.data(), nk1, nk2, np, false);
    for (size_t k = 0; k < nk; ++k)
        for (size_t t = 0; t < np; ++t) packed[k * np + t] *= VR[k];
    fft2d(packed.data(), nk1, nk2, np, true);
    for (size_t k = 0; k < nk; ++k) {
        for (size_t t = 0; t < np; ++t) {
            const c64 val = -packed[k * np + t];
            sigma[k * nb2 + ti[t] * nb + tj[t]] = val;
            if (ti[t] != tj[t]) sigma[k * nb2 + tj[t] * nb + ti[t]] = std::conj(val);
        }
    }
}

// selfenergy.hpp:174-190
void ifftshift_2d_batch(c64* buf, size_t nk1, size_t nk2, size_t batch) {
    const size_t row = nk2 * batch, s0 = nk1 / 2, s1 = nk2 / 2;
    std::vector<c64> tmp(nk1 * row);
    for (size_t i = 0; i < nk1; ++i) {
        const size_t is = (i + s0) % nk1;
        for (size_t j = 0; j < nk2; ++j) {
            const size_t js = (j + s1) % nk2;
            std::memcpy(&tmp[(i * nk2 + j) * batch], &buf[(is * nk2 + js) * batch],
                        batch * sizeof(c64));
        }
    }
    std::memcpy(buf, tmp.data(), nk1 * row * sizeof(c64));
}

// ---------------------------------------------------------------------------
// Linear algebra (linalg.hpp)
// ---------------------------------------------------------------------------

// linalg.hpp:20-33
void hermitize_inplace(c64* M, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        c64* mk = M + k * nb2;
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = i; j < nb; ++j) {
                const c64 h = 0.5 * (mk[i * nb + j] + std::conj(mk[j * nb + i]));
                mk[i * nb + j] = h;
                mk[j * nb + i] = std::conj(h);
            }
    });
}

// Cyclic Jacobi eigensolver for a Hermitian n x n matrix (row-major A,
// overwritten).  Replaces Eigen::SelfAdjointEigenSolver (linalg.hpp:42).
// Outputs ascending w and row-major V whose columns are eigenvectors.
void jacobi_eigh(std::vector<c64>& A, size_t n, double* w_out, c64* V_out) {
    std::vector<c64> V(n * n, c64(0.0, 0.0));
    for (size_t i = 0; i < n; ++i) V[i * n + i] = 1.0;
    double norm2 = 0.0;
    for (size_t i = 0; i < n * n; ++i) norm2 += std::norm(A[i]);
    const double eps = std::numeric_limits<double>::epsilon();
    const double stop = (4.0 * static_cast<double>(n) * eps) *
                        (4.0 * static_cast<double>(n) * eps) * norm2;
    for (int sweep = 0; sweep < 60 && norm2 > 0.0; ++sweep) {
        double off = 0.0;
        for (size_t i = 0; i < n; ++i)
            for (size_t j = i + 1; j < n; ++j) off += std::norm(A[i * n + j]);
        if (off <= stop) break;
        for (size_t p = 0; p < n; ++p) {
            for (size_t q = p + 1; q < n; ++q) {
                const c64 b = A[p * n + q];
                const double ab = std::abs(b);
                if (ab == 0.0) continue;
                const double a = A[p * n + p].real(), d = A[q * n + q].real();
                const double tau = (d - a) / (2.0 * ab);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) /
                                 (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double s = t * c;
                const c64 e = b / ab;
                const c64 se = s * e, sec = s * std::conj(e);
                // A <- A J   (columns p, q)
                for (size_t i = 0; i < n; ++i) {
                    const c64 aip = A[i * n + p], aiq = A[i * n + q];
                    A[i * n + p] = aip * c - aiq * sec;
                    A[i * n + q] = aip * se + aiq * c;
                }
                // A <- J^H A (rows p, q)
                for (size_t j = 0; j < n; ++j) {
                    const c64 apj = A[p * n + j], aqj = A[q * n + j];
                    A[p * n + j] = c * apj - se * aqj;
                    A[q * n + j] = sec * apj + c * aqj;
                }
                A[p * n + q] = 0.0;
                A[q * n + p] = 0.0;
                A[p * n + p] = a - t * ab;
                A[q * n + q] = d + t * ab;
                for (size_t i = 0; i < n; ++i) {
                    const c64 vip = V[i * n + p], viq = V[i * n + q];
                    V[i * n + p] = vip * c - viq * sec;
                    V[i * n + q] = vip * se + viq * c;
                }
            }
        }
    }
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return A[x * n + x].real() < A[y * n + y].real();
    });
    for (size_t i = 0; i < n; ++i) {
        w_out[i] = A[order[i] * n + order[i]].real();
        for (size_t r = 0; r < n; ++r) V_out[r * n + i] = V[r * n + order[i]];
    }
}

// linalg.hpp:39-55 (hermitise first, then solve)
void eigh_one(const c64* M, size_t nb, double* w, c64* V) {
    std::vector<c64> H(nb * nb);
    for (size_t i = 0; i < nb; ++i)
        for (size_t j = 0; j < nb; ++j)
            H[i * nb + j] = 0.5 * (M[i * nb + j] + std::conj(M[j * nb + i]));
    jacobi_eigh(H, nb, w, V);
}

// linalg.hpp:59-66
void eigh_batched(const c64* M, double* w, c64* V, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) { eigh_one(M + k * nb2, nb, w + k * nb, V + k * nb2); });
}

// linalg.hpp:72-104: each diagonal block hermitised and diagonalised on its
// own, placed at its position; sort: all pairs ascending (ties by position).
void eigh_block_sizes_one(const c64* M, size_t nb, const std::vector<size_t>& sizes, bool sort,
                          double* w, c64* V) {
    for (size_t e = 0; e < nb * nb; ++e) V[e] = 0.0;
    size_t start = 0;
    for (size_t s : sizes) {
        if (s == 0) continue;
        std::vector<c64> sub(s * s), Vs(s * s);
        for (size_t i = 0; i < s; ++i)
            for (size_t j = 0; j < s; ++j) sub[i * s + j] = M[(start + i) * nb + start + j];
        eigh_one(sub.data(), s, w + start, Vs.data());
        for (size_t i = 0; i < s; ++i)
            for (size_t j = 0; j < s; ++j) V[(start + i) * nb + start + j] = Vs[i * s + j];
        start += s;
    }
    if (!sort) return;
    std::vector<size_t> order(nb);
    for (size_t i = 0; i < nb; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return w[a] < w[b]; });
    std::vector<double> wt(nb);
    std::vector<c64> Vt(nb * nb);
    for (size_t c = 0; c < nb; ++c) {
        wt[c] = w[order[c]];
        for (size_t r = 0; r < nb; ++r) Vt[r * nb + c] = V[r * nb + order[c]];
    }
    std::copy(wt.begin(), wt.end(), w);
    std::copy(Vt.begin(), Vt.end(), V);
}

// linalg.hpp:115-130
double max_offblock_sizes(const c64* M, size_t nk, size_t nb, const std::vector<size_t>& sizes) {
    std::vector<char> mask(nb * nb, 1);
    size_t start = 0;
    for (size_t s : sizes) {
        for (size_t i = start; i < start + s; ++i)
            for (size_t j = start; j < start + s; ++j) mask[i * nb + j] = 0;
        start += s;
    }
    double mx = 0.0;
    for (size_t k = 0; k < nk; ++k)
        for (size_t e = 0; e < nb * nb; ++e)
            if (mask[e]) mx = std::max(mx, std::abs(M[k * nb * nb + e]));
    return mx;
}

// ---------------------------------------------------------------------------
// Physics utilities (utils.hpp)
// ---------------------------------------------------------------------------

// utils.hpp:18-25
inline double expit_scalar(double x) {
    if (x >= 0.0) return 1.0 / (1.0 + std::exp(-x));
    const double e = std::exp(x);
    return e / (1.0 + e);
}

// utils.hpp:27-29
inline double fermidirac_scalar(double x, double T) { return expit_scalar(-x / (T + 1.0e-12)); }

// utils.hpp:31-46
void validate_electron_count(const double* weights, size_t nk, size_t nb, double n_e) {
    double total = 0.0;
    for (size_t i = 0; i < nk; ++i) total += weights[i];
    const double max_e = total * static_cast<double>(nb);
    const double tol = std::max(1e-8, 1e-6 * std::max(1.0, std::fabs(max_e)));
    if (!std::isfinite(n_e)) throw std::invalid_argument("n_electrons must be finite.");
    if (n_e < -tol || n_e > max_e + tol)
        throw std::invalid_argument("n_electrons=" + std::to_string(n_e) +
                                    " is outside the physically reachable range [0, " +
                                    std::to_string(max_e) + "].");
}

// utils.hpp:49-75
double find_mu_bisection(const double* eps, size_t n_eps, const double* weights, size_t nk,
                         size_t nb, double n_e, double T, int maxiter) {
    if (maxiter <= 0) maxiter = 30;
    const double Tj = std::max(T, 1.0e-12);
    double emin = eps[0], emax = eps[0];
    for (size_t i = 1; i < n_eps; ++i) {
        emin = std::min(emin, eps[i]);
        emax = std::max(emax, eps[i]);
    }
    const double span = emax - emin + 10.0 * std::max(Tj, 1.0e-6);
    double lo = emin - span, hi = emax + span;
    for (int it = 0; it < maxiter; ++it) {
        const double mid = 0.5 * (lo + hi);
        double count = 0.0;
        for (size_t k = 0; k < nk; ++k)
            for (size_t b = 0; b < nb; ++b)
                count += weights[k] * fermidirac_scalar(eps[k * nb + b] - mid, Tj);
        if (count > n_e) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// utils.hpp:79-116
double solve_mu_inloop(const double* eps, size_t n_eps, const double* w_norm, size_t nk,
                       size_t nb, double n_target_norm, double mu0, double T, int maxiter) {
    const double Tj = std::max(T, 1.0e-12);
    double emin = eps[0], emax = eps[0];
    for (size_t i = 1; i < n_eps; ++i) {
        emin = std::min(emin, eps[i]);
        emax = std::max(emax, eps[i]);
    }
    double lo = emin - 50.0 * Tj, hi = emax + 50.0 * Tj;
    double mu = std::clamp(mu0, lo, hi);
    for (int it = 0; it < maxiter; ++it) {
        double N = 0.0, Z = 0.0;
        for (size_t k = 0; k < nk; ++k) {
            const double wk = w_norm[k];
            for (size_t b = 0; b < nb; ++b) {
                const double p = expit_scalar((mu - eps[k * nb + b]) / Tj);
                N += wk * p;
                Z += wk * p * (1.0 - p) / Tj;
            }
        }
        const double g = N - n_target_norm;
        if (g < 0.0) lo = mu;
        if (g > 0.0) hi = mu;
        const double Z_safe = std::max(Z, 1.0e-18);
        double mu_new = mu - g / Z_safe;
        const double mu_bis = 0.5 * (lo + hi);
        if (mu_new <= lo || mu_new >= hi) mu_new = mu_bis;
        mu_new = std::clamp(mu_new, lo, hi);
        if (!std::isfinite(mu_new)) mu_new = mu_bis;
        mu = mu_new;
    }
    return mu;
}

// src/cpp_hf/utils.py:311-360 (_find_mu_newton, the Python surface's
// method="newton"): raw weights and target, bracket [emin - span,
// emax + span] with span = emax - emin + 10 max(Tj, 1e-6), start at its
// midpoint, maxiter Newton steps with the bisection safeguard, and the
// bracket midpoint when the final count misses the target by more than 1e-12.
double find_mu_newton_py(const double* eps, const double* weights, size_t nk, size_t nb,
                         double n_e, double T, int maxiter) {
    const double Tj = std::max(T, 1.0e-12);
    const size_t n_eps = nk * nb;
    double emin = eps[0], emax = eps[0];
    for (size_t i = 1; i < n_eps; ++i) {
        emin = std::min(emin, eps[i]);
        emax = std::max(emax, eps[i]);
    }
    const double span = emax - emin + 10.0 * std::max(Tj, 1.0e-6);
    double lo = emin - span, hi = emax + span;
    double mu = 0.5 * (lo + hi);
    auto count_and_slope = [&](double m, double& N, double& Z) {
        N = 0.0;
        Z = 0.0;
        for (size_t k = 0; k < nk; ++k)
            for (size_t b = 0; b < nb; ++b) {
                const double p = expit_scalar((m - eps[k * nb + b]) / Tj);
                N += weights[k] * p;
                Z += weights[k] * p * (1.0 - p) / Tj;
            }
    };
    for (int it = 0; it < maxiter; ++it) {
        double N, Z;
        count_and_slope(mu, N, Z);
        const double g = N - n_e;
        if (g < 0.0) lo = mu;
        if (g > 0.0) hi = mu;
        const double Z_safe = std::max(Z, 1.0e-18);
        double mu_new = mu - g / Z_safe;
        const double mu_bis = 0.5 * (lo + hi);
        if (mu_new <= lo || mu_new >= hi || !std::isfinite(mu_new)) mu_new = mu_bis;
        mu_new = std::max(std::min(mu_new, hi), lo);
        mu = mu_new;
    }
    double N_fin, Z_fin;
    count_and_slope(mu, N_fin, Z_fin);
    if (std::fabs(N_fin - n_e) > 1e-12) mu = 0.5 * (lo + hi);
    return mu;
}

// utils.hpp:120-146 (no FMA contraction: the build uses -ffp-contract=off)
void resample_axis(const c64* x, c64* y, size_t n_old, size_t n_new, size_t inner) {
    if (n_old == n_new) {
        std::memcpy(y, x, n_old * inner * sizeof(c64));
        return;
    }
    const double ratio = static_cast<double>(n_old) / static_cast<double>(n_new);
    const size_t s_new = n_new / 2, s_old = n_old / 2;
    for (size_t j = 0; j < n_new; ++j) {
        const double u = (static_cast<double>(j) - static_cast<double>(s_new)) * ratio +
                         static_cast<double>(s_old);
        const double fu = std::floor(u);
        long long i0 = static_cast<long long>(fu);
        const long long no = static_cast<long long>(n_old);
        i0 = ((i0 % no) + no) % no;
        const size_t i1 = (static_cast<size_t>(i0) + 1) % n_old;
        const double frac = u - fu;
        const c64 a(1.0 - frac), b(frac);
        for (size_t i = 0; i < inner; ++i)
            y[j * inner + i] = a * x[static_cast<size_t>(i0) * inner + i] + b * x[i1 * inner + i];
    }
}

// utils.hpp:151-176
void resample_kgrid_2d(const c64* x, c64* out, size_t nk_old, size_t nk_new, size_t inner) {
    if (nk_old == nk_new) {
        std::memcpy(out, x, nk_old * nk_old * inner * sizeof(c64));
        return;
    }
    std::vector<c64> tmp(nk_new * nk_old * inner);
    resample_axis(x, tmp.data(), nk_old, nk_new, nk_old * inner);
    for (size_t i = 0; i < nk_new; ++i)
        resample_axis(tmp.data() + i * nk_old * inner, out + i * nk_new * inner, nk_old,
                      nk_new, inner);
}

// ---------------------------------------------------------------------------
// Fock build and energy (fock.hpp)
// ---------------------------------------------------------------------------

struct Kernel {
    size_t nk1, nk2, nb, dv1, dv2, n_contact;
    bool include_hartree, include_exchange, exchange_hcp, has_refP;
    double T, weight_sum;
    const c64 *h, *VR, *refP;
    const double *w2d, *HH, *contact_g;
    const c64 *contact_Oi, *contact_Oj;
    const hf_superlattice_desc* sl = nullptr;  // superlattice mode (kernel.hpp:36-69)
    size_t nk() const { return nk1 * nk2; }
    size_t nb2() const { return nb * nb; }
    size_t n_dense() const { return nk() * nb2(); }
};

// make_kernel's validation (cpp_hf_native.cpp:58-87)
Kernel make_kernel(const hf_kernel_desc* d) {
    if (!d) throw std::invalid_argument("kernel description is null");
    if (!((d->dv1 == 1 && d->dv2 == 1) || (d->dv1 == d->nb && d->dv2 == d->nb)))
        throw std::invalid_argument("VR last two dims must be (1,1) or (nb,nb)");
    if (d->has_refP && !d->refP)
        throw std::invalid_argument("refP must match h shape when has_refP is true");
    Kernel K;
    K.nk1 = d->nk1; K.nk2 = d->nk2; K.nb = d->nb; K.dv1 = d->dv1; K.dv2 = d->dv2;
    K.n_contact = d->n_contact;
    K.include_hartree = d->include_hartree != 0;
    K.include_exchange = d->include_exchange != 0;
    K.exchange_hcp = d->exchange_hcp != 0;
    K.has_refP = d->has_refP != 0;
    K.T = d->T; K.weight_sum = d->weight_sum;
    K.h = reinterpret_cast<const c64*>(d->h);
    K.VR = reinterpret_cast<const c64*>(d->VR);
    K.refP = K.has_refP ? reinterpret_cast<const c64*>(d->refP) : nullptr;
    K.w2d = d->w2d; K.HH = d->HH; K.contact_g = d->contact_g;
    K.contact_Oi = reinterpret_cast<const c64*>(d->contact_Oi);
    K.contact_Oj = reinterpret_cast<const c64*>(d->contact_Oj);
    const hf_superlattice_desc* sl = d->superlattice;
    if (sl && (sl->fock_active || sl->hartree_active)) {
        // inject_superlattice_into_kernel (cpp_hf_native.cpp:109-165)
        if (sl->n_G * sl->dim_orb != d->nb)
            throw std::invalid_argument("superlattice mode requires n_G × dim_orb == nb");
        if (!sl->V_lag_fft || !sl->g_a_off || !sl->pair_i || !sl->pair_j || !sl->pair_start ||
            !sl->pair_to_delta || !sl->HH_GG)
            throw std::invalid_argument("superlattice description has a null array");
        K.sl = sl;
    }
    return K;
}

// superlattice.hpp:53-246: the (alpha, beta, dG) channels of the
// superlattice exchange, each scattered onto the extended grid, convolved
// with the lag kernel and gathered back; with hermitian, one channel of
// each (+dG, -dG) partner pair is computed and the other filled by
// Sigma[k, j b, i a] = conj(Sigma[k, i a, j b]).
void selfenergy_superlattice(const c64* rho, c64* sigma, const c64* V, const c64* V_orb,
                             const int64_t* g_a_off, const int64_t* pair_i,
                             const int64_t* pair_j, const int64_t* pair_start, size_t n_delta,
                             size_t Nx, size_t Ny, size_t nkx, size_t nky, size_t n_G,
                             size_t orb, bool hermitian) {
    const size_t orb2 = orb * orb, chan = Nx * Ny * orb2;
    const size_t D = n_G * orb, row_rho = nky * D * D, row_dg = Ny * orb2;
    std::fill(sigma, sigma + nkx * nky * D * D, c64(0.0, 0.0));
    std::vector<int64_t> partner(n_delta, -1);
    if (hermitian) {
        // channel whose (i, j) pair set is this channel's reversed
        std::map<std::vector<int64_t>, size_t> by_key;
        std::vector<std::vector<int64_t>> reversed(n_delta);
        for (size_t d = 0; d < n_delta; ++d) {
            std::vector<int64_t> key, rkey;
            for (int64_t p = pair_start[d]; p < pair_start[d + 1]; ++p) {
                key.push_back(pair_i[p] * static_cast<int64_t>(n_G) + pair_j[p]);
                rkey.push_back(pair_j[p] * static_cast<int64_t>(n_G) + pair_i[p]);
            }
            std::sort(key.begin(), key.end());
            std::sort(rkey.begin(), rkey.end());
            by_key.emplace(std::move(key), d);
            reversed[d] = std::move(rkey);
        }
        for (size_t d = 0; d < n_delta; ++d) {
            auto it = by_key.find(reversed[d]);
            if (it != by_key.end()) partner[d] = static_cast<int64_t>(it->second);
        }
    }
    for (size_t d = 0; d < n_delta; ++d) {
        const int64_t pr = partner[d];
        const bool fill = pr >= 0 && static_cast<size_t>(pr) != d;
        if (fill && static_cast<size_t>(pr) < d) continue;  // filled by its partner
        const int64_t a = pair_start[d], b = pair_start[d + 1];
        if (a == b) continue;
        std::vector<c64> g(chan, c64(0.0, 0.0));
        for (int64_t p = a; p < b; ++p) {
            const size_t i = pair_i[p], j = pair_j[p];
            const size_t ox = nkx * g_a_off[2 * i], oy = nky * g_a_off[2 * i + 1];
            for (size_t kx = 0; kx < nkx; ++kx)
                for (size_t ky = 0; ky < nky; ++ky)
                    for (size_t al = 0; al < orb; ++al)
                        for (size_t be = 0; be < orb; ++be)
                            g[(ox + kx) * row_dg + (oy + ky) * orb2 + al * orb + be] =
                                rho[kx * row_rho + (ky * D + i * orb + al) * D + j * orb + be];
        }
        bool nonzero = false;
        for (size_t t = 0; t < chan && !nonzero; ++t) nonzero = g[t] != c64(0.0, 0.0);
        if (!nonzero) continue;  // zero in, zero out
        if (V_orb)
            selfenergy_fft_full_inplace(g.data(), V_orb, Nx, Ny, orb, orb, orb);
        else
            selfenergy_fft_full_inplace(g.data(), V, Nx, Ny, orb, 1, 1);
        for (int64_t p = a; p < b; ++p) {
            const size_t i = pair_i[p], j = pair_j[p];
            const size_t ox = nkx * g_a_off[2 * i], oy = nky * g_a_off[2 * i + 1];
            for (size_t kx = 0; kx < nkx; ++kx)
                for (size_t ky = 0; ky < nky; ++ky)
                    for (size_t al = 0; al < orb; ++al)
                        for (size_t be = 0; be < orb; ++be) {
                            const c64 v = g[(ox + kx) * row_dg + (oy + ky) * orb2 + al * orb + be];
                            sigma[kx * row_rho + (ky * D + i * orb + al) * D + j * orb + be] = v;
                            if (fill)
                                sigma[kx * row_rho + (ky * D + j * orb + be) * D + i * orb + al] =
                                    std::conj(v);
                        }
        }
    }
}

// fock.hpp:66-173: the k-independent superlattice Hartree shift from
// rho_bar = sum_k w_k (P - refP)_k
std::vector<c64> superlattice_hartree(const Kernel& K, const c64* dP) {
    const hf_superlattice_desc& L = *K.sl;
    const size_t nb = K.nb, nb2 = nb * nb, nG = L.n_G, orb = L.dim_orb;
    std::vector<c64> rho_bar(nb2, c64(0.0, 0.0));
    for (size_t k = 0; k < K.nk(); ++k)
        for (size_t i = 0; i < nb2; ++i) rho_bar[i] += c64(K.w2d[k]) * dP[k * nb2 + i];
    std::vector<c64> out(nb2, c64(0.0, 0.0));
    const double deg = L.hartree_degeneracy;
    if (!L.HH_GG_orbital) {
        std::vector<c64> per_delta(L.n_delta, c64(0.0, 0.0));
        for (size_t A = 0; A < nG; ++A)
            for (size_t B = 0; B < nG; ++B) {
                c64 s(0.0, 0.0);
                for (size_t x = 0; x < orb; ++x) s += rho_bar[(A * orb + x) * nb + B * orb + x];
                per_delta[L.pair_to_delta[A * nG + B]] += s;
            }
        for (size_t A = 0; A < nG; ++A)
            for (size_t B = 0; B < nG; ++B) {
                const c64 s = c64(deg * L.HH_GG[A * nG + B]) * per_delta[L.pair_to_delta[A * nG + B]];
                for (size_t x = 0; x < orb; ++x) out[(A * orb + x) * nb + B * orb + x] = s;
            }
    } else {
        std::vector<c64> per_delta(L.n_delta * orb, c64(0.0, 0.0));
        for (size_t A = 0; A < nG; ++A)
            for (size_t B = 0; B < nG; ++B)
                for (size_t g = 0; g < orb; ++g)
                    per_delta[L.pair_to_delta[A * nG + B] * orb + g] +=
                        rho_bar[(A * orb + g) * nb + B * orb + g];
        for (size_t A = 0; A < nG; ++A)
            for (size_t B = 0; B < nG; ++B) {
                const size_t d = L.pair_to_delta[A * nG + B];
                for (size_t a = 0; a < orb; ++a) {
                    c64 s(0.0, 0.0);
                    for (size_t g = 0; g < orb; ++g)
                        s += c64(deg * L.HH_GG_orbital[((A * nG + B) * orb + a) * orb + g]) *
                             per_delta[d * orb + g];
                    out[(A * orb + a) * nb + B * orb + a] = s;
                }
            }
    }
    return out;
}

struct Projector {
    hf_project_fn fn = nullptr;
    void* user = nullptr;
    explicit operator bool() const { return fn != nullptr; }
    void operator()(c64* M, const Kernel& K) const {
        if (fn(user, reinterpret_cast<double*>(M), K.nk1, K.nk2, K.nb) != 0)
            throw std::runtime_error("project_fn failed");
    }
};

// fock.hpp:25-29
bool has_contact_terms(const Kernel& K) {
    for (size_t t = 0; t < K.n_contact; ++t)
        if (K.contact_g[t] != 0.0) return true;
    return K.n_contact > 1;
}

// fock.hpp:34-299
void build_fock_compact(const Kernel& K, const c64* P, c64* Sigma, c64* F, double* hdiag,
                        const Projector* proj) {
    const size_t nk = K.nk(), nb = K.nb, nb2 = K.nb2(), n_tot = K.n_dense();
    const bool sl_fock = K.sl && K.sl->fock_active;
    const bool sl_hartree = K.sl && K.sl->hartree_active;
    std::vector<c64> sl_shift;
    if (K.include_exchange || sl_hartree) {
        if (K.has_refP) {
            parallel_for(nk, [&](size_t k) {
                for (size_t i = 0; i < nb2; ++i)
                    Sigma[k * nb2 + i] = P[k * nb2 + i] - K.refP[k * nb2 + i];
            });
        } else {
            std::memcpy(Sigma, P, n_tot * sizeof(c64));
        }
        if (K.include_hartree && sl_hartree) sl_shift = superlattice_hartree(K, Sigma);
    }
    if (K.include_exchange) {
        if (sl_fock) {
            const hf_superlattice_desc& L = *K.sl;
            std::vector<c64> out(n_tot);
            selfenergy_superlattice(Sigma, out.data(), reinterpret_cast<const c64*>(L.V_lag_fft),
                                    reinterpret_cast<const c64*>(L.V_lag_fft_orbital), L.g_a_off,
                                    L.pair_i, L.pair_j, L.pair_start, L.n_delta, L.N_ext_x,
                                    L.N_ext_y, K.nk1, K.nk2, L.n_G, L.dim_orb, true);
            std::memcpy(Sigma, out.data(), n_tot * sizeof(c64));
        } else if (K.exchange_hcp && K.dv1 == 1 && K.dv2 == 1)
            selfenergy_fft_full_hcp(Sigma, Sigma, K.VR, K.nk1, K.nk2, nb);
        else
            selfenergy_fft_full_inplace(Sigma, K.VR, K.nk1, K.nk2, nb, K.dv1, K.dv2);
    } else {
        std::fill(Sigma, Sigma + n_tot, c64(0.0, 0.0));
    }
    if (!sl_shift.empty())
        parallel_for(nk, [&](size_t k) {
            for (size_t i = 0; i < nb2; ++i) Sigma[k * nb2 + i] += sl_shift[i];
        });

    std::fill(hdiag, hdiag + nb, 0.0);
    if (K.include_hartree && !sl_hartree) {
        std::vector<double> n_vec(nb, 0.0);
        for (size_t k = 0; k < nk; ++k) {
            const c64* pk = P + k * nb2;
            const c64* rk = K.has_refP ? K.refP + k * nb2 : nullptr;
            for (size_t i = 0; i < nb; ++i) {
                const double rr = rk ? rk[i * nb + i].real() : 0.0;
                n_vec[i] += K.w2d[k] * (pk[i * nb + i].real() - rr);
            }
        }
        for (size_t i = 0; i < nb; ++i) {
            double s = 0.0;
            for (size_t j = 0; j < nb; ++j) s += K.HH[i * nb + j] * n_vec[j];
            hdiag[i] = s;
        }
    }

    if (has_contact_terms(K)) {
        std::vector<c64> rho(nb2, c64(0.0, 0.0));
        for (size_t k = 0; k < nk; ++k) {
            const c64* pk = P + k * nb2;
            const c64* rk = K.has_refP ? K.refP + k * nb2 : nullptr;
            for (size_t i = 0; i < nb2; ++i)
                rho[i] += c64(K.w2d[k]) * (pk[i] - (rk ? rk[i] : c64(0.0, 0.0)));
        }
        std::vector<c64> SC(nb2, c64(0.0, 0.0));
        for (size_t t = 0; t < K.n_contact; ++t) {
            const double g = K.contact_g[t];
            if (g == 0.0) continue;
            const c64* Oi = K.contact_Oi + t * nb2;
            const c64* Oj = K.contact_Oj + t * nb2;
            c64 tr(0.0, 0.0);
            for (size_t i = 0; i < nb; ++i)
                for (size_t j = 0; j < nb; ++j) tr += Oj[i * nb + j] * rho[j * nb + i];
            // Oi * rho * Oj
            std::vector<c64> OR(nb2, c64(0.0, 0.0)), ORO(nb2, c64(0.0, 0.0));
            for (size_t i = 0; i < nb; ++i)
                for (size_t k2 = 0; k2 < nb; ++k2)
                    for (size_t j = 0; j < nb; ++j) OR[i * nb + j] += Oi[i * nb + k2] * rho[k2 * nb + j];
            for (size_t i = 0; i < nb; ++i)
                for (size_t k2 = 0; k2 < nb; ++k2)
                    for (size_t j = 0; j < nb; ++j) ORO[i * nb + j] += OR[i * nb + k2] * Oj[k2 * nb + j];
            for (size_t i = 0; i < nb2; ++i)
                SC[i] += Oi[i] * tr * c64(g) + (-c64(g)) * ORO[i];
        }
        parallel_for(nk, [&](size_t k) {
            for (size_t i = 0; i < nb2; ++i) Sigma[k * nb2 + i] += SC[i];
        });
    }

    parallel_for(nk, [&](size_t k) {
        for (size_t i = 0; i < nb2; ++i) F[k * nb2 + i] = K.h[k * nb2 + i] + Sigma[k * nb2 + i];
        if (K.include_hartree)
            for (size_t i = 0; i < nb; ++i) F[k * nb2 + i * nb + i] += c64(hdiag[i], 0.0);
    });
    hermitize_inplace(F, nk, nb);
    if (proj && *proj) {
        (*proj)(F, K);
        hermitize_inplace(F, nk, nb);
    }
}

// fock.hpp:302-321
void build_fock(const Kernel& K, const c64* P, c64* Sigma, c64* Hh, c64* F,
                const Projector* proj) {
    const size_t nk = K.nk(), nb = K.nb, nb2 = K.nb2();
    std::vector<double> hdiag(nb, 0.0);
    build_fock_compact(K, P, Sigma, F, hdiag.data(), proj);
    std::fill(Hh, Hh + K.n_dense(), c64(0.0, 0.0));
    if (K.include_hartree)
        for (size_t k = 0; k < nk; ++k)
            for (size_t i = 0; i < nb; ++i) Hh[k * nb2 + i * nb + i] = c64(hdiag[i], 0.0);
}

// fock.hpp:324-351
double hf_energy(const Kernel& K, const c64* P, const c64* Sigma, const c64* Hh) {
    const size_t nk = K.nk(), nb = K.nb, nb2 = K.nb2();
    double total = 0.0;
    for (size_t k = 0; k < nk; ++k) {
        const c64 *hk = K.h + k * nb2, *sk = Sigma + k * nb2, *hh = Hh + k * nb2,
                  *pk = P + k * nb2;
        const c64* rk = K.has_refP ? K.refP + k * nb2 : nullptr;
        double s = 0.0;
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j) {
                const c64 p = pk[j * nb + i];
                const c64 dp = p - (rk ? rk[j * nb + i] : c64(0.0, 0.0));
                const c64 inter = 0.5 * (sk[i * nb + j] + hh[i * nb + j]);
                s += (hk[i * nb + j] * p).real();
                s += (inter * dp).real();
            }
        total += K.w2d[k] * s;
    }
    return total;
}

// fock.hpp:355-383
double hf_energy_with_hartree_diag(const Kernel& K, const c64* P, const c64* Sigma,
                                   const double* hdiag) {
    const size_t nk = K.nk(), nb = K.nb, nb2 = K.nb2();
    double total = 0.0;
    for (size_t k = 0; k < nk; ++k) {
        const c64 *hk = K.h + k * nb2, *sk = Sigma + k * nb2, *pk = P + k * nb2;
        const c64* rk = K.has_refP ? K.refP + k * nb2 : nullptr;
        double s = 0.0;
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j) {
                c64 inter = 0.5 * sk[i * nb + j];
                if (K.include_hartree && i == j) inter += c64(0.5 * hdiag[i], 0.0);
                const c64 p = pk[j * nb + i];
                const c64 dp = p - (rk ? rk[j * nb + i] : c64(0.0, 0.0));
                s += (hk[i * nb + j] * p).real();
                s += (inter * dp).real();
            }
        total += K.w2d[k] * s;
    }
    return total;
}

// fock.hpp:385-401
double occupation_entropy(const double* p, const double* w_norm, size_t nk, size_t nb) {
    double S = 0.0;
    for (size_t k = 0; k < nk; ++k)
        for (size_t b = 0; b < nb; ++b) {
            double pv = p[k * nb + b];
            if (pv < ENTROPY_CLIP) pv = ENTROPY_CLIP;
            if (pv > 1.0 - ENTROPY_CLIP) pv = 1.0 - ENTROPY_CLIP;
            S -= w_norm[k] * (pv * std::log(pv) + (1.0 - pv) * std::log1p(-pv));
        }
    return S;
}

// ---------------------------------------------------------------------------
// Direct minimisation helpers (solver_dm.hpp:117-388)
// ---------------------------------------------------------------------------

// C = A * B (n x n, row-major); mode bits: 1 -> use A^H, 2 -> use B^H
void matmul(const c64* A, const c64* B, c64* C, size_t n, int mode) {
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            c64 s(0.0, 0.0);
            for (size_t k = 0; k < n; ++k) {
                const c64 a = (mode & 1) ? std::conj(A[k * n + i]) : A[i * n + k];
                const c64 b = (mode & 2) ? std::conj(B[j * n + k]) : B[k * n + j];
                s += a * b;
            }
            C[i * n + j] = s;
        }
}

// solver_dm.hpp:117-128
void density_from_Qp(const c64* Q, const double* p, c64* P, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> Qp(Q + k * nb2, Q + (k + 1) * nb2);
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j) Qp[i * nb + j] *= p[k * nb + j];
        matmul(Qp.data(), Q + k * nb2, P + k * nb2, nb, 2);
    });
}

// solver_dm.hpp:130-139 (in-place safe: output written after the product)
void fock_in_orbital_basis(const c64* Q, const c64* F, c64* Ft, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> T1(nb2), T2(nb2);
        matmul(Q + k * nb2, F + k * nb2, T1.data(), nb, 1);
        matmul(T1.data(), Q + k * nb2, T2.data(), nb, 0);
        std::memcpy(Ft + k * nb2, T2.data(), nb2 * sizeof(c64));
    });
}

// solver_dm.hpp:142-160
void compute_orbital_gradient(const c64* Ft, const double* p, c64* G, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        const c64* Fk = Ft + k * nb2;
        const double* pk = p + k * nb;
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j) {
                if (i == j) { G[k * nb2 + i * nb + j] = 0.0; continue; }
                const c64 aij = (pk[j] - pk[i]) * Fk[i * nb + j];
                const c64 aji = (pk[i] - pk[j]) * Fk[j * nb + i];
                G[k * nb2 + i * nb + j] = 0.5 * (aij - std::conj(aji));
            }
    });
}

// solver_dm.hpp:163-177
double ip_matrix(const c64* X, const c64* Y, const double* w, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    double total = 0.0;
    for (size_t k = 0; k < nk; ++k) {
        double s = 0.0;
        for (size_t i = 0; i < nb2; ++i) s += (std::conj(X[k * nb2 + i]) * Y[k * nb2 + i]).real();
        total += w[k] * s;
    }
    return total;
}

// solver_dm.hpp:179-191
double ip_vec(const double* x, const double* y, const double* w, size_t nk, size_t nb) {
    double total = 0.0;
    for (size_t k = 0; k < nk; ++k) {
        double s = 0.0;
        for (size_t b = 0; b < nb; ++b) s += x[k * nb + b] * y[k * nb + b];
        total += w[k] * s;
    }
    return total;
}

double norm_matrix(const c64* X, const double* w, size_t nk, size_t nb) {
    return std::sqrt(std::max(0.0, ip_matrix(X, X, w, nk, nb)));
}

// solver_dm.hpp:206-222
void cayley_spectral_setup(const c64* d, c64* V, double* lam, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> iA(nb2), H(nb2);
        for (size_t i = 0; i < nb2; ++i) iA[i] = c64(0.0, 1.0) * d[k * nb2 + i];
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j)
                H[i * nb + j] = 0.5 * (iA[i * nb + j] + std::conj(iA[j * nb + i]));
        jacobi_eigh(H, nb, lam + k * nb, V + k * nb2);
    });
}

// solver_dm.hpp:251-255
inline c64 cayley_factor(double lam, double tau) {
    const double arg = 0.5 * tau * lam;
    const c64 ie(0.0, arg);
    return (c64(1.0, 0.0) + ie) / (c64(1.0, 0.0) - ie);
}

// solver_dm.hpp:258-272
void cayley_unitary_from_spectrum(const c64* V, const double* lam, double tau, c64* U,
                                  size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> Vc(V + k * nb2, V + (k + 1) * nb2);
        for (size_t j = 0; j < nb; ++j) {
            const c64 c = cayley_factor(lam[k * nb + j], tau);
            for (size_t i = 0; i < nb; ++i) Vc[i * nb + j] *= c;
        }
        matmul(Vc.data(), V + k * nb2, U + k * nb2, nb, 2);
    });
}

// solver_dm.hpp:277-303
void diag_UFU_from_spectrum(const c64* V, const c64* Fe, const double* lam, double tau,
                            double* diag, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> c(nb), M(nb2), A(nb2);
        for (size_t i = 0; i < nb; ++i) c[i] = cayley_factor(lam[k * nb + i], tau);
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j)
                M[i * nb + j] = std::conj(c[i]) * Fe[k * nb2 + i * nb + j] * c[j];
        matmul(V + k * nb2, M.data(), A.data(), nb, 0);
        for (size_t i = 0; i < nb; ++i) {
            double s = 0.0;
            for (size_t j = 0; j < nb; ++j)
                s += (A[i * nb + j] * std::conj(V[k * nb2 + i * nb + j])).real();
            diag[k * nb + i] = s;
        }
    });
}

// solver_dm.hpp:305-314
void compute_Ft_eig(const c64* Ft, const c64* Vd, c64* Fe, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> T1(nb2);
        matmul(Vd + k * nb2, Ft + k * nb2, T1.data(), nb, 1);
        matmul(T1.data(), Vd + k * nb2, Fe + k * nb2, nb, 0);
    });
}

// solver_dm.hpp:317-326 (in-place safe)
void Q_times_U(const c64* Q, const c64* U, c64* Qn, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> out(nb2);
        matmul(Q + k * nb2, U + k * nb2, out.data(), nb, 0);
        std::memcpy(Qn + k * nb2, out.data(), nb2 * sizeof(c64));
    });
}

// solver_dm.hpp:328-344
void Q_times_cayley_spectrum_inplace(c64* Q, const c64* V, const double* lam, double tau,
                                     size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> Vc(V + k * nb2, V + (k + 1) * nb2), U(nb2), out(nb2);
        for (size_t j = 0; j < nb; ++j) {
            const c64 c = cayley_factor(lam[k * nb + j], tau);
            for (size_t i = 0; i < nb; ++i) Vc[i * nb + j] *= c;
        }
        matmul(Vc.data(), V + k * nb2, U.data(), nb, 2);
        matmul(Q + k * nb2, U.data(), out.data(), nb, 0);
        std::memcpy(Q + k * nb2, out.data(), nb2 * sizeof(c64));
    });
}

// solver_dm.hpp:346-365 and 367-388 (trial form with d_p, tau)
double frozen_free_energy(const double* eps, const double* p, const double* d_p, double tau,
                          const double* w, size_t nk, size_t nb, double T) {
    double E = 0.0, S = 0.0;
    for (size_t k = 0; k < nk; ++k)
        for (size_t b = 0; b < nb; ++b) {
            const size_t idx = k * nb + b;
            const double raw = d_p ? p[idx] - tau * d_p[idx] : p[idx];
            const double pv = std::clamp(raw, OCC_CLIP, 1.0 - OCC_CLIP);
            E += w[k] * pv * eps[idx];
            S -= w[k] * (pv * std::log(pv) + (1.0 - pv) * std::log1p(-pv));
        }
    return E - std::max(T, 1.0e-14) * S;
}

// solver_dm.hpp:92-98: every matrix of the batch block-diagonal in `sizes` up
// to 1e-10 + 1e-10 max|M|
bool block_structure_ok(const c64* M, size_t nmat, size_t nb, const std::vector<size_t>& sizes) {
    double scale = 0.0;
    for (size_t i = 0; i < nmat * nb * nb; ++i) scale = std::max(scale, std::abs(M[i]));
    return max_offblock_sizes(M, nmat, nb, sizes) <= 1.0e-10 + 1.0e-10 * scale;
}

// solver_dm.hpp:100-113: h, refP and the active contact operators keep the
// block structure, and no projector may mix the blocks
bool block_problem_invariant(const Kernel& K, const Projector* proj,
                             const std::vector<size_t>& sizes) {
    if (proj && *proj) return false;
    const size_t nk = K.nk(), nb = K.nb, nb2 = nb * nb;
    if (!block_structure_ok(K.h, nk, nb, sizes)) return false;
    if (K.has_refP && !block_structure_ok(K.refP, nk, nb, sizes)) return false;
    for (size_t t = 0; t < K.n_contact; ++t) {
        if (K.contact_g[t] == 0.0) continue;
        if (!block_structure_ok(K.contact_Oi + t * nb2, 1, nb, sizes)) return false;
        if (!block_structure_ok(K.contact_Oj + t * nb2, 1, nb, sizes)) return false;
    }
    return true;
}

// linalg.hpp:110-121 with sort = false: each block's ascending pairs at the
// block's position
void eigh_blocks_batched(const c64* M, double* w, c64* V, size_t nk, size_t nb,
                         const std::vector<size_t>& sizes) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        eigh_block_sizes_one(M + k * nb2, nb, sizes, false, w + k * nb, V + k * nb2);
    });
}

// solver_dm.hpp:224-249: the Cayley spectrum block by block
void cayley_spectral_setup_blocks(const c64* d, c64* V, double* lam, size_t nk, size_t nb,
                                  const std::vector<size_t>& sizes) {
    const size_t nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> iA(nb2);
        for (size_t i = 0; i < nb2; ++i) iA[i] = c64(0.0, 1.0) * d[k * nb2 + i];
        eigh_block_sizes_one(iA.data(), nb, sizes, false, lam + k * nb, V + k * nb2);
    });
}

// solver_dm.hpp:392-653 (optimizer "cg")
void solve_dm(const Kernel& K, const c64* P0, double n_e, const hf_dm_config& cfg,
              const Projector* proj, hf_dm_result* out) {
    size_t ls_trials = 0;  // frozen free-energy evaluations of the line search
    const size_t nk1 = K.nk1, nk2 = K.nk2, nb = K.nb, nk = nk1 * nk2, nb2 = nb * nb;
    const size_t n_tot = nk * nb2, nkb = nk * nb;
    (void)nk2;
    std::vector<double> w_norm(nk);
    const double ws = std::max(K.weight_sum, TINY_REAL);
    for (size_t i = 0; i < nk; ++i) w_norm[i] = K.w2d[i] / ws;
    const double n_target_norm = n_e / ws;
    const double T_r = std::max(K.T, 1.0e-12);

    std::vector<c64> Q(n_tot), P_cur(n_tot), Sigma(n_tot), F(n_tot), G_Q(n_tot),
        G_Q_prev(n_tot, 0.0), d_Q(n_tot), d_Q_prev(n_tot, 0.0), V_d(n_tot);
    std::vector<double> p(nkb), hdiag(nb, 0.0), eps(nkb), g_p(nkb), h_p(nkb),
        g_p_prev(nkb, 0.0), d_p(nkb), d_p_prev(nkb, 0.0), lam_d(nkb), eps_work(nkb);
    std::vector<double> hist_E(cfg.max_iter, 0.0), hist_G(cfg.max_iter, 0.0);

    // block structure (80-90, 403-405, 446-447)
    const std::vector<size_t> sizes(cfg.block_sizes, cfg.block_sizes + cfg.n_blocks);
    bool use_block = false;
    if (!sizes.empty()) {
        size_t total = 0;
        for (size_t s : sizes) {
            if (s == 0) throw std::invalid_argument("block_sizes entries must be positive.");
            total += s;
        }
        if (total != nb) throw std::invalid_argument("block_sizes must sum to nb.");
        use_block = sizes.size() > 1 && block_problem_invariant(K, proj, sizes);
    }

    // init (436-459)
    std::memcpy(P_cur.data(), P0, n_tot * sizeof(c64));
    hermitize_inplace(P_cur.data(), nk, nb);
    build_fock_compact(K, P_cur.data(), Sigma.data(), F.data(), hdiag.data(), nullptr);
    if (use_block && !block_structure_ok(F.data(), nk, nb, sizes)) use_block = false;
    if (use_block)
        eigh_blocks_batched(F.data(), eps.data(), Q.data(), nk, nb, sizes);
    else
        eigh_batched(F.data(), eps.data(), Q.data(), nk, nb);
    const double mu0 = solve_mu_inloop(eps.data(), nkb, w_norm.data(), nk, nb, n_target_norm,
                                       0.0, T_r, cfg.mu_maxiter);
    for (size_t i = 0; i < nkb; ++i) p[i] = expit_scalar((mu0 - eps[i]) / T_r);

    auto project = [&](c64* M) {
        if (proj && *proj) (*proj)(M, K);
    };

    double grad_norm = std::numeric_limits<double>::infinity();
    double E_prev = std::numeric_limits<double>::infinity();
    double dE = std::numeric_limits<double>::infinity();
    double E = 0.0, mu = mu0;
    size_t k = 0;
    while (k < cfg.max_iter) {
        bool e_nc;
        if (cfg.plateau_window > 0 && k > cfg.plateau_window)
            e_nc = (hist_E[k - 1 - cfg.plateau_window] - hist_E[k - 1]) > cfg.tol_E;
        else
            e_nc = dE > cfg.tol_E;
        const bool g_nc = (cfg.tol_grad > 0.0) && (grad_norm > cfg.tol_grad);
        if (!(e_nc || g_nc)) break;

        density_from_Qp(Q.data(), p.data(), P_cur.data(), nk, nb);
        project(P_cur.data());
        hermitize_inplace(P_cur.data(), nk, nb);
        build_fock_compact(K, P_cur.data(), Sigma.data(), F.data(), hdiag.data(), nullptr);
        E = hf_energy_with_hartree_diag(K, P_cur.data(), Sigma.data(), hdiag.data());
        fock_in_orbital_basis(Q.data(), F.data(), F.data(), nk, nb);
        for (size_t kk = 0; kk < nk; ++kk)
            for (size_t b = 0; b < nb; ++b) eps[kk * nb + b] = F[kk * nb2 + b * nb + b].real();

        compute_orbital_gradient(F.data(), p.data(), G_Q.data(), nk, nb);
        for (size_t i = 0; i < nkb; ++i) {
            const double pv = std::clamp(p[i], LOGIT_CLIP, 1.0 - LOGIT_CLIP);
            g_p[i] = eps[i] + T_r * (std::log(pv) - std::log1p(-pv)) - mu;
        }
        grad_norm = norm_matrix(G_Q.data(), w_norm.data(), nk, nb);

        double eps_sq = 0.0;
        for (size_t i = 0; i < nkb; ++i) eps_sq += eps[i] * eps[i];
        const double eps_scale =
            std::sqrt(eps_sq / static_cast<double>(std::max<size_t>(nkb, 1)) + TINY_REAL);
        const double lam_pc = std::max(T_r, cfg.denom_scale * eps_scale);
        for (size_t kk = 0; kk < nk; ++kk)
            for (size_t i = 0; i < nb; ++i)
                for (size_t j = 0; j < nb; ++j) {
                    const double gap = eps[kk * nb + i] - eps[kk * nb + j];
                    const double base = std::sqrt(gap * gap + lam_pc * lam_pc);
                    d_Q[kk * nb2 + i * nb + j] = G_Q[kk * nb2 + i * nb + j] / base;
                }
        for (size_t i = 0; i < nkb; ++i) {
            const double ps = std::clamp(p[i], OCC_CLIP, 1.0 - OCC_CLIP);
            const double curv = T_r / (ps * (1.0 - ps));
            h_p[i] = g_p[i] / curv;
        }

        const double pr_num = ip_matrix(G_Q.data(), d_Q.data(), w_norm.data(), nk, nb) -
                              ip_matrix(G_Q_prev.data(), d_Q.data(), w_norm.data(), nk, nb) +
                              ip_vec(g_p.data(), h_p.data(), w_norm.data(), nk, nb) -
                              ip_vec(g_p_prev.data(), h_p.data(), w_norm.data(), nk, nb);
        const double pr_den = ip_matrix(G_Q_prev.data(), G_Q_prev.data(), w_norm.data(), nk, nb) +
                              ip_vec(g_p_prev.data(), g_p_prev.data(), w_norm.data(), nk, nb);
        double beta = (pr_den > TINY_REAL) ? std::max(0.0, pr_num / pr_den) : 0.0;
        beta = std::min(beta, 5.0);
        if (k == 0 || (cfg.cg_restart > 0 && k % cfg.cg_restart == 0)) beta = 0.0;
        for (size_t i = 0; i < n_tot; ++i) d_Q[i] = d_Q[i] + beta * d_Q_prev[i];
        for (size_t i = 0; i < nkb; ++i) d_p[i] = h_p[i] + beta * d_p_prev[i];

        if (use_block)
            cayley_spectral_setup_blocks(d_Q.data(), V_d.data(), lam_d.data(), nk, nb, sizes);
        else
            cayley_spectral_setup(d_Q.data(), V_d.data(), lam_d.data(), nk, nb);
        compute_Ft_eig(F.data(), V_d.data(), Sigma.data(), nk, nb);
        const double dQn = norm_matrix(d_Q.data(), w_norm.data(), nk, nb);
        const double tau0 = std::min(1.0, cfg.max_step / std::max(dQn, TINY_REAL));
        const double Om0 = frozen_free_energy(eps.data(), p.data(), nullptr, 0.0,
                                              w_norm.data(), nk, nb, T_r);
        double tau = tau0;
        bool ok = false;
        for (size_t bt = 0; bt < cfg.bt_max; ++bt) {
            ++ls_trials;
            diag_UFU_from_spectrum(V_d.data(), Sigma.data(), lam_d.data(), tau,
                                   eps_work.data(), nk, nb);
            const double Om = frozen_free_energy(eps_work.data(), p.data(), d_p.data(), tau,
                                                 w_norm.data(), nk, nb, T_r);
            if (Om < Om0) { ok = true; break; }
            tau *= cfg.bt_shrink;
        }
        if (!ok) tau = tau0 * std::pow(cfg.bt_shrink, static_cast<double>(cfg.bt_max));
        Q_times_cayley_spectrum_inplace(Q.data(), V_d.data(), lam_d.data(), tau, nk, nb);
        diag_UFU_from_spectrum(V_d.data(), Sigma.data(), lam_d.data(), tau, eps_work.data(),
                               nk, nb);
        const double mu_new = solve_mu_inloop(eps_work.data(), nkb, w_norm.data(), nk, nb,
                                              n_target_norm, mu, T_r, cfg.mu_maxiter);
        for (size_t i = 0; i < nkb; ++i) p[i] = expit_scalar((mu_new - eps_work[i]) / T_r);
        mu = mu_new;

        hist_E[k] = E;
        hist_G[k] = grad_norm;
        dE = std::fabs(E - E_prev);
        d_Q_prev = d_Q;
        d_p_prev = d_p;
        G_Q_prev = G_Q;
        g_p_prev = g_p;
        E_prev = E;
        ++k;
    }

    // finalize (615-652)
    density_from_Qp(Q.data(), p.data(), P_cur.data(), nk, nb);
    project(P_cur.data());
    hermitize_inplace(P_cur.data(), nk, nb);
    build_fock_compact(K, P_cur.data(), Sigma.data(), F.data(), hdiag.data(), nullptr);
    fock_in_orbital_basis(Q.data(), F.data(), F.data(), nk, nb);
    if (use_block)
        eigh_blocks_batched(F.data(), eps_work.data(), V_d.data(), nk, nb, sizes);
    else
        eigh_batched(F.data(), eps_work.data(), V_d.data(), nk, nb);
    Q_times_U(Q.data(), V_d.data(), Q.data(), nk, nb);
    mu = solve_mu_inloop(eps_work.data(), nkb, w_norm.data(), nk, nb, n_target_norm, mu, T_r,
                         cfg.mu_maxiter);
    for (size_t i = 0; i < nkb; ++i) p[i] = expit_scalar((mu - eps_work[i]) / T_r);
    density_from_Qp(Q.data(), p.data(), P_cur.data(), nk, nb);
    project(P_cur.data());
    hermitize_inplace(P_cur.data(), nk, nb);
    build_fock_compact(K, P_cur.data(), Sigma.data(), F.data(), hdiag.data(), nullptr);
    const double E_fin = hf_energy_with_hartree_diag(K, P_cur.data(), Sigma.data(), hdiag.data());

    if (out->Q) std::memcpy(out->Q, Q.data(), n_tot * sizeof(c64));
    std::memcpy(out->p, p.data(), nkb * sizeof(double));
    if (out->density) std::memcpy(out->density, P_cur.data(), n_tot * sizeof(c64));
    if (out->fock) std::memcpy(out->fock, F.data(), n_tot * sizeof(c64));
    out->mu = mu;
    out->energy = E_fin;
    out->n_iter = k;
    out->converged = k < cfg.max_iter ? 1 : 0;
    out->n_ls_trials = ls_trials;
    if (out->hist_E) std::memcpy(out->hist_E, hist_E.data(), k * sizeof(double));
    if (out->hist_grad) std::memcpy(out->hist_grad, hist_G.data(), k * sizeof(double));
}

// ---------------------------------------------------------------------------
// Roothaan SCF (solver_scf.hpp)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Trust-region Newton (solver_rtr.hpp)
// ---------------------------------------------------------------------------

// solver_rtr.hpp:32-39: R[dP] = F[dP] - h with the reference density switched
// off (the self-energy, Hartree and contact terms are linear in the density)
void fock_response(const Kernel& K, const c64* dP, c64* resp, c64* Sg, c64* Fb, double* hb) {
    Kernel Kr = K;
    Kr.has_refP = false;
    Kr.refP = nullptr;
    build_fock_compact(Kr, dP, Sg, Fb, hb, nullptr);
    for (size_t i = 0; i < K.n_dense(); ++i) resp[i] = Fb[i] - K.h[i];
}

// solver_rtr.hpp:42-50: remove the w-weighted mean so sum_k w_k sum_b v = 0
void project_occ(double* v, const double* w, size_t nk, size_t nb) {
    double num = 0.0, den = 0.0;
    for (size_t k = 0; k < nk; ++k) {
        for (size_t b = 0; b < nb; ++b) num += w[k] * v[k * nb + b];
        den += w[k] * static_cast<double>(nb);
    }
    const double shift = den > 0.0 ? num / den : 0.0;
    for (size_t i = 0; i < nk * nb; ++i) v[i] -= shift;
}

// solver_rtr.hpp:55-112: joint (Q, p) Hessian-vector product, one Fock
// response; pen adds the deflation curvature pen * dP to the response.
void joint_hvp(const Kernel& K, const c64* X, const double* dp, const c64* Q, const double* p,
               const c64* Ft, double T_r, c64* HX, double* Hp, c64* dPb, c64* rb, c64* Sb,
               c64* Fb, double* hb, double pen) {
    const size_t nk = K.nk(), nb = K.nb, nb2 = nb * nb;
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> M(nb2), T1(nb2), D(nb2);
        const double* pk = p + k * nb;
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j)
                M[i * nb + j] = i == j ? c64(dp[k * nb + i], 0.0)
                                       : X[k * nb2 + i * nb + j] * (pk[j] - pk[i]);
        matmul(Q + k * nb2, M.data(), T1.data(), nb, 0);
        matmul(T1.data(), Q + k * nb2, D.data(), nb, 2);
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j)
                dPb[k * nb2 + i * nb + j] = 0.5 * (D[i * nb + j] + std::conj(D[j * nb + i]));
    });
    fock_response(K, dPb, rb, Sb, Fb, hb);
    if (pen != 0.0)
        for (size_t i = 0; i < K.n_dense(); ++i) rb[i] += c64(pen, 0.0) * dPb[i];
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> T1(nb2), St(nb2), A(nb2), C(nb2);
        const c64* Fk = Ft + k * nb2;
        const c64* Xk = X + k * nb2;
        matmul(Q + k * nb2, rb + k * nb2, T1.data(), nb, 1);
        matmul(T1.data(), Q + k * nb2, St.data(), nb, 0);
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j) {
                c64 fx(0.0, 0.0), xf(0.0, 0.0);
                for (size_t m = 0; m < nb; ++m) {
                    fx += Fk[i * nb + m] * Xk[m * nb + j];
                    xf += Xk[i * nb + m] * Fk[m * nb + j];
                }
                A[i * nb + j] = fx - xf + St[i * nb + j];
            }
        const double* pk = p + k * nb;
        const double* dk = dp + k * nb;
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j)
                C[i * nb + j] = (dk[j] - dk[i]) * Fk[i * nb + j] + (pk[j] - pk[i]) * A[i * nb + j];
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j)
                HX[k * nb2 + i * nb + j] =
                    i == j ? c64(0.0, 0.0) : 0.5 * (C[i * nb + j] - std::conj(C[j * nb + i]));
        for (size_t b = 0; b < nb; ++b) {
            const double pc = std::clamp(pk[b], OCC_CLIP, 1.0 - OCC_CLIP);
            Hp[k * nb + b] = A[b * nb + b].real() + T_r / (pc * (1.0 - pc)) * dk[b];
        }
    });
    project_occ(Hp, K.w2d, nk, nb);
}

// solver_rtr.hpp:127-160: Gaussian repulsion around found densities;
// returns Phi, writes S_pen (dPhi/dP_k / w_k) and pen = 2 sum_i phi'(d_i^2).
double deflation_bias(const c64* P, const c64* targets, size_t n_defl, const double* w,
                      size_t nk, size_t nb, double sigma, double length, c64* S_pen,
                      double& pen) {
    const size_t nb2 = nb * nb, n_tot = nk * nb2;
    std::fill(S_pen, S_pen + n_tot, c64(0.0, 0.0));
    pen = 0.0;
    if (n_defl == 0 || sigma == 0.0 || !targets) return 0.0;
    const double inv2L2 = 1.0 / (2.0 * length * length);
    double Phi = 0.0, diag = 0.0;
    for (size_t t = 0; t < n_defl; ++t) {
        const c64* Tt = targets + t * n_tot;
        double d2 = 0.0;
        for (size_t k = 0; k < nk; ++k) {
            double per = 0.0;
            for (size_t a = 0; a < nb2; ++a) per += std::norm(P[k * nb2 + a] - Tt[k * nb2 + a]);
            d2 += w[k] * per;
        }
        const double phi = sigma * std::exp(-d2 * inv2L2);
        const double dphi = -phi * inv2L2;
        Phi += phi;
        diag += dphi;
        for (size_t i = 0; i < n_tot; ++i) S_pen[i] += c64(2.0 * dphi, 0.0) * (P[i] - Tt[i]);
    }
    pen = 2.0 * diag;
    return Phi;
}

// solver_rtr.hpp:162-459
void solve_rtr(const Kernel& K, const c64* P0, double n_e, const hf_dm_config& cfg,
               const Projector* proj, hf_dm_result* out) {
    const size_t nk = K.nk(), nb = K.nb, nb2 = nb * nb, n_tot = nk * nb2, npb = nk * nb;
    const double T_r = std::max(K.T, 1.0e-12);
    const double tol_g = cfg.tol_grad > 0.0 ? cfg.tol_grad : 1.0e-6;
    std::vector<double> w_norm(nk), ones(nk, 1.0);
    const double ws = std::max(K.weight_sum, TINY_REAL);
    for (size_t i = 0; i < nk; ++i) w_norm[i] = K.w2d[i] / ws;
    const double n_target = n_e / ws;

    std::vector<c64> Q(n_tot), P(n_tot), Sigma(n_tot), F(n_tot), Ft(n_tot), G(n_tot);
    std::vector<c64> vX(n_tot), dPb(n_tot), rb(n_tot), Fb(n_tot), Sb(n_tot);
    std::vector<c64> rX(n_tot), dX(n_tot), zX(n_tot), HvX(n_tot), HdX(n_tot);
    std::vector<c64> Vd(n_tot), Qt(n_tot), U(n_tot), Gs(n_tot), Fts(n_tot);
    std::vector<double> p(npb), eps(npb), hdiag(nb, 0.0), hb(nb, 0.0), vp(npb), rp(npb),
        dpv(npb), zp(npb), Hvp(npb), Hdp(npb), lam(npb), eps_s(npb), eps_t(npb), p_t(npb);
    std::vector<double> hist_E(cfg.max_iter, 0.0), hist_G(cfg.max_iter, 0.0);

    const size_t n_defl = cfg.n_deflation;
    const c64* targets = reinterpret_cast<const c64*>(cfg.deflation_targets);
    const double d_sigma = cfg.deflation_sigma;
    const double d_len = std::max(cfg.deflation_length, 1.0e-12);
    const bool defl_on = n_defl > 0 && d_sigma > 0.0 && targets;
    std::vector<c64> S_pen(defl_on ? n_tot : 0);
    double pen = 0.0;

    // joint metric: orbital part unweighted, occupation part w2d-weighted
    auto jip = [&](const c64* aX, const double* ap, const c64* bX, const double* bp) {
        return ip_matrix(aX, bX, ones.data(), nk, nb) + ip_vec(ap, bp, K.w2d, nk, nb);
    };
    auto jnorm = [&](const c64* aX, const double* ap) {
        return std::sqrt(std::max(0.0, jip(aX, ap, aX, ap)));
    };
    // block preconditioner: orbital 1 / sqrt(gap^2 + lam^2), occupation p (1 - p) / T
    auto precond = [&](const c64* aX, const double* ap, c64* oX, double* op, double lam_pc) {
        for (size_t k = 0; k < nk; ++k)
            for (size_t i = 0; i < nb; ++i)
                for (size_t j = 0; j < nb; ++j) {
                    const double gap = eps[k * nb + i] - eps[k * nb + j];
                    oX[k * nb2 + i * nb + j] =
                        aX[k * nb2 + i * nb + j] / std::sqrt(gap * gap + lam_pc * lam_pc);
                }
        for (size_t i = 0; i < npb; ++i) {
            const double pc = std::clamp(p[i], OCC_CLIP, 1.0 - OCC_CLIP);
            op[i] = ap[i] * pc * (1.0 - pc) / T_r;
        }
        project_occ(op, K.w2d, nk, nb);
    };
    auto project = [&](c64* M) {
        if (proj && *proj) (*proj)(M, K);
    };
    // density, Fock (+ deflation), Ft, eps and G at (Qs, ps); returns E (+ Phi)
    auto eval_state = [&](const c64* Qs, const double* ps) {
        density_from_Qp(Qs, ps, P.data(), nk, nb);
        project(P.data());
        hermitize_inplace(P.data(), nk, nb);
        build_fock_compact(K, P.data(), Sigma.data(), F.data(), hdiag.data(), nullptr);
        double E = hf_energy_with_hartree_diag(K, P.data(), Sigma.data(), hdiag.data());
        if (defl_on) {
            E += deflation_bias(P.data(), targets, n_defl, K.w2d, nk, nb, d_sigma, d_len,
                                S_pen.data(), pen);
            for (size_t i = 0; i < n_tot; ++i) F[i] += S_pen[i];
        }
        fock_in_orbital_basis(Qs, F.data(), Ft.data(), nk, nb);
        for (size_t k = 0; k < nk; ++k)
            for (size_t b = 0; b < nb; ++b) eps[k * nb + b] = Ft[k * nb2 + b * nb + b].real();
        compute_orbital_gradient(Ft.data(), ps, G.data(), nk, nb);
        return E;
    };

    // init (solver_rtr.hpp:262-272)
    std::memcpy(P.data(), P0, n_tot * sizeof(c64));
    hermitize_inplace(P.data(), nk, nb);
    build_fock_compact(K, P.data(), Sigma.data(), F.data(), hdiag.data(), nullptr);
    eigh_batched(F.data(), eps.data(), Q.data(), nk, nb);
    double mu = solve_mu_inloop(eps.data(), npb, w_norm.data(), nk, nb, n_target, 0.0, T_r,
                                cfg.mu_maxiter);
    for (size_t i = 0; i < npb; ++i) p[i] = expit_scalar((mu - eps[i]) / T_r);

    double E = eval_state(Q.data(), p.data());
    double Delta = cfg.tr_delta0;
    size_t k = 0;
    bool converged = false;
    while (k < cfg.max_iter) {
        const double gn = norm_matrix(G.data(), w_norm.data(), nk, nb);
        hist_E[k] = E;
        hist_G[k] = gn;
        if (gn < tol_g) {
            converged = true;
            ++k;
            break;
        }
        double eps_sq = 0.0;
        for (size_t i = 0; i < npb; ++i) eps_sq += eps[i] * eps[i];
        const double lam_pc = std::max(
            T_r, cfg.denom_scale *
                     std::sqrt(eps_sq / static_cast<double>(std::max<size_t>(npb, 1)) + TINY_REAL));

        // Steihaug truncated CG on the joint quadratic model (288-347)
        std::fill(vX.begin(), vX.end(), c64(0.0, 0.0));
        std::fill(vp.begin(), vp.end(), 0.0);
        std::fill(HvX.begin(), HvX.end(), c64(0.0, 0.0));
        std::fill(Hvp.begin(), Hvp.end(), 0.0);
        for (size_t i = 0; i < n_tot; ++i) rX[i] = -G[i];
        std::fill(rp.begin(), rp.end(), 0.0);
        precond(rX.data(), rp.data(), zX.data(), zp.data(), lam_pc);
        dX = zX;
        dpv = zp;
        double rz = jip(rX.data(), rp.data(), zX.data(), zp.data());
        for (size_t it = 0; it < cfg.tr_cg_max; ++it) {
            joint_hvp(K, dX.data(), dpv.data(), Q.data(), p.data(), Ft.data(), T_r, HdX.data(),
                      Hdp.data(), dPb.data(), rb.data(), Sb.data(), Fb.data(), hb.data(), pen);
            const double dHd = jip(dX.data(), dpv.data(), HdX.data(), Hdp.data());
            const double dd = jip(dX.data(), dpv.data(), dX.data(), dpv.data());
            auto to_boundary = [&]() {
                const double a = dd;
                const double b = 2.0 * jip(vX.data(), vp.data(), dX.data(), dpv.data());
                const double c = jnorm(vX.data(), vp.data());
                const double cc = c * c - Delta * Delta;
                const double t =
                    (-b + std::sqrt(std::max(b * b - 4.0 * a * cc, 0.0))) / (2.0 * a + 1e-30);
                for (size_t i = 0; i < n_tot; ++i) {
                    vX[i] += t * dX[i];
                    HvX[i] += t * HdX[i];
                }
                for (size_t i = 0; i < npb; ++i) {
                    vp[i] += t * dpv[i];
                    Hvp[i] += t * Hdp[i];
                }
            };
            if (dHd <= 1e-14 * dd) {  // negative curvature
                to_boundary();
                break;
            }
            const double alpha = rz / dHd;
            double s2 = 0.0;
            for (size_t i = 0; i < n_tot; ++i) s2 += std::norm(vX[i] + alpha * dX[i]);
            for (size_t kk = 0; kk < nk; ++kk) {
                double per = 0.0;
                for (size_t b = 0; b < nb; ++b) {
                    const double e = vp[kk * nb + b] + alpha * dpv[kk * nb + b];
                    per += e * e;
                }
                s2 += K.w2d[kk] * per;
            }
            if (std::sqrt(std::max(0.0, s2)) >= Delta) {
                to_boundary();
                break;
            }
            for (size_t i = 0; i < n_tot; ++i) {
                vX[i] += alpha * dX[i];
                HvX[i] += alpha * HdX[i];
                rX[i] -= alpha * HdX[i];
            }
            for (size_t i = 0; i < npb; ++i) {
                vp[i] += alpha * dpv[i];
                Hvp[i] += alpha * Hdp[i];
                rp[i] -= alpha * Hdp[i];
            }
            const double rn = jnorm(rX.data(), rp.data());
            if (rn < 0.05 * gn) break;
            precond(rX.data(), rp.data(), zX.data(), zp.data(), lam_pc);
            const double rz_new = jip(rX.data(), rp.data(), zX.data(), zp.data());
            const double beta = rz_new / rz;
            for (size_t i = 0; i < n_tot; ++i) dX[i] = zX[i] + beta * dX[i];
            for (size_t i = 0; i < npb; ++i) dpv[i] = zp[i] + beta * dpv[i];
            rz = rz_new;
        }

        // predicted reduction in energy units (w2d-weighted, 349-361)
        const double gv = ip_matrix(G.data(), vX.data(), K.w2d, nk, nb);
        const double vHv = ip_matrix(vX.data(), HvX.data(), K.w2d, nk, nb) +
                           ip_vec(vp.data(), Hvp.data(), K.w2d, nk, nb);
        const double pred = -(gv + 0.5 * vHv);
        if (Delta < 1e-10) break;

        // retraction along +vX (Cayley at tau = -1), frozen-Ft trial energies,
        // occupations at the re-solved mu (363-385)
        cayley_spectral_setup(vX.data(), Vd.data(), lam.data(), nk, nb);
        cayley_unitary_from_spectrum(Vd.data(), lam.data(), -1.0, U.data(), nk, nb);
        Q_times_U(Q.data(), U.data(), Qt.data(), nk, nb);
        parallel_for(nk, [&](size_t kk) {
            std::vector<c64> T1(nb2), T2(nb2);
            matmul(U.data() + kk * nb2, Ft.data() + kk * nb2, T1.data(), nb, 1);
            matmul(T1.data(), U.data() + kk * nb2, T2.data(), nb, 0);
            for (size_t b = 0; b < nb; ++b) eps_t[kk * nb + b] = T2[b * nb + b].real();
        });
        const double mu_t = solve_mu_inloop(eps_t.data(), npb, w_norm.data(), nk, nb, n_target,
                                            mu, T_r, cfg.mu_maxiter);
        for (size_t i = 0; i < npb; ++i) p_t[i] = expit_scalar((mu_t - eps_t[i]) / T_r);

        Gs = G;
        Fts = Ft;
        eps_s = eps;
        const double pen_s = pen;
        const double E_t = eval_state(Qt.data(), p_t.data());
        const double g_t = norm_matrix(G.data(), w_norm.data(), nk, nb);

        // acceptance and trust radius (399-420)
        const double noise = 1e-12 * std::max(std::fabs(E), 1.0);
        bool accept;
        if (pred <= noise) {
            accept = g_t <= gn && E_t <= E + noise;
            if (!accept) Delta *= 0.5;
        } else {
            const double ratio = (E - E_t) / pred;
            if (ratio < 0.25)
                Delta *= 0.25;
            else if (ratio > 0.75 && jnorm(vX.data(), vp.data()) > 0.9 * Delta)
                Delta = std::min(2.0 * Delta, 5.0);
            accept = ratio > 0.1;
        }
        if (std::getenv("HFCPU_RTR_DEBUG"))
            std::fprintf(stderr,
                         "[rtr] k=%zu g=%.3e |v|=%.3e Delta=%.3e pred=%.3e dE=%+.3e gtr=%.3e "
                         "accept=%d\n",
                         k, gn, jnorm(vX.data(), vp.data()), Delta, pred, E_t - E, g_t,
                         static_cast<int>(accept));
        if (accept) {
            Q = Qt;
            p = p_t;
            E = E_t;
            mu = mu_t;
        } else {
            G = Gs;
            Ft = Fts;
            eps = eps_s;
            pen = pen_s;
        }
        ++k;
    }

    // finalize (436-441)
    density_from_Qp(Q.data(), p.data(), P.data(), nk, nb);
    project(P.data());
    hermitize_inplace(P.data(), nk, nb);
    build_fock_compact(K, P.data(), Sigma.data(), F.data(), hdiag.data(), nullptr);
    const double E_fin = hf_energy_with_hartree_diag(K, P.data(), Sigma.data(), hdiag.data());

    if (out->Q) std::memcpy(out->Q, Q.data(), n_tot * sizeof(c64));
    std::memcpy(out->p, p.data(), npb * sizeof(double));
    if (out->density) std::memcpy(out->density, P.data(), n_tot * sizeof(c64));
    if (out->fock) std::memcpy(out->fock, F.data(), n_tot * sizeof(c64));
    out->mu = mu;
    out->energy = E_fin;
    out->n_iter = k;
    out->converged = converged ? 1 : 0;
    if (out->hist_E) std::memcpy(out->hist_E, hist_E.data(), k * sizeof(double));
    if (out->hist_grad) std::memcpy(out->hist_grad, hist_G.data(), k * sizeof(double));
}

// solver_scf.hpp:64-75
double weighted_matrix_norm(const c64* M, const double* w, size_t nk, size_t nb, double ws) {
    const size_t nb2 = nb * nb;
    double total = 0.0;
    for (size_t k = 0; k < nk; ++k) {
        double s = 0.0;
        for (size_t i = 0; i < nb2; ++i) s += std::norm(M[k * nb2 + i]);
        total += w[k] * s;
    }
    return std::sqrt(total / std::max(ws, 1.0e-30));
}

// solver_scf.hpp:96-154; per-block eigh when block_sizes (summing to nb, more
// than one entry) describes F's structure up to 1e-10 + 1e-10 max|F|
double density_from_fock(const c64* F, size_t nk, size_t nb, const double* w, double n_e,
                         double T, double level_shift, c64* P_new,
                         const std::vector<size_t>& sizes = {}) {
    const size_t nb2 = nb * nb, nkb = nk * nb;
    std::vector<double> eps(nkb), shifted(nkb), occ(nkb);
    std::vector<c64> V(nk * nb2);
    bool used_blocks = false;
    if (!sizes.empty()) {
        size_t total = 0;
        for (size_t s : sizes) total += s;
        if (total == nb && sizes.size() > 1 && block_structure_ok(F, nk, nb, sizes)) {
            eigh_blocks_batched(F, eps.data(), V.data(), nk, nb, sizes);
            used_blocks = true;
        }
    }
    if (!used_blocks) eigh_batched(F, eps.data(), V.data(), nk, nb);
    const double mu_raw = find_mu_bisection(eps.data(), nkb, w, nk, nb, n_e, T, 0);
    if (level_shift != 0.0) {
        for (size_t i = 0; i < nkb; ++i)
            shifted[i] = eps[i] + level_shift * (1.0 - fermidirac_scalar(eps[i] - mu_raw, T));
    } else {
        shifted = eps;
    }
    const double mu = find_mu_bisection(shifted.data(), nkb, w, nk, nb, n_e, T, 0);
    for (size_t i = 0; i < nkb; ++i) occ[i] = fermidirac_scalar(shifted[i] - mu, T);
    parallel_for(nk, [&](size_t k) {
        std::vector<c64> Vo(V.begin() + k * nb2, V.begin() + (k + 1) * nb2), P(nb2);
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j) Vo[i * nb + j] *= occ[k * nb + j];
        matmul(Vo.data(), V.data() + k * nb2, P.data(), nb, 2);
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = 0; j < nb; ++j)
                P_new[k * nb2 + i * nb + j] = 0.5 * (P[i * nb + j] + std::conj(P[j * nb + i]));
    });
    return mu;
}

// Full-pivot LU solve of the (m+1)x(m+1) DIIS system; returns false when the
// matrix is rank deficient (Eigen::FullPivLU::isInvertible, threshold
// epsilon * size * |max pivot|).
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
                if (std::fabs(B[i * n + j]) > best) { best = std::fabs(B[i * n + j]); pr = i; pc = j; }
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

// solver_scf.hpp:156-444
void solve_scf(const Kernel& K, const c64* P0, double n_e, const hf_scf_config& cfg,
               const Projector* proj, hf_scf_result* out) {
    const size_t nk = K.nk(), nb = K.nb, nb2 = K.nb2(), n_tot = K.n_dense();
    std::vector<c64> density(P0, P0 + n_tot), Sigma(n_tot), F(n_tot), P_new(n_tot),
        delta(n_tot), comm(n_tot);
    std::vector<double> hdiag(nb, 0.0);
    std::vector<double> hE(cfg.max_iter, 0.0), hD(cfg.max_iter, 0.0), hC(cfg.max_iter, 0.0);
    const std::vector<size_t> sizes(cfg.block_sizes, cfg.block_sizes + cfg.n_blocks);
    double mu = 0.0, E = 0.0;
    bool converged = false;
    size_t k = 0;
    (void)E;
    (void)mu;

    auto project = [&](c64* M) {
        if (proj && *proj) (*proj)(M, K);
    };
    auto build_and_occupy = [&](std::vector<c64>& d, c64* P_out, double& E_out,
                                double& mu_out) {
        project(d.data());
        hermitize_inplace(d.data(), nk, nb);
        build_fock_compact(K, d.data(), Sigma.data(), F.data(), hdiag.data(), proj);
        E_out = hf_energy_with_hartree_diag(K, d.data(), Sigma.data(), hdiag.data());
        mu_out = density_from_fock(F.data(), nk, nb, K.w2d, n_e, K.T, cfg.level_shift, P_out,
                                   sizes);
        project(P_out);
        hermitize_inplace(P_out, nk, nb);
    };

    const bool use_diis = cfg.acceleration == 1, use_oda = cfg.acceleration == 2;
    if (cfg.acceleration < 0 || cfg.acceleration > 2)
        throw std::invalid_argument("SCFConfig.acceleration must be 'linear', 'diis', or 'oda'");
    std::vector<std::vector<c64>> diis_F, diis_R;
    std::vector<c64> Sigma_new(use_oda ? n_tot : 0), F_new(use_oda ? n_tot : 0);
    std::vector<double> hdiag_new(use_oda ? nb : 0);

    while (k < cfg.max_iter && !converged) {
        double E_iter = 0.0, mu_iter = 0.0;
        build_and_occupy(density, P_new.data(), E_iter, mu_iter);
        parallel_for(nk, [&](size_t kk) {
            for (size_t i = 0; i < nb2; ++i)
                delta[kk * nb2 + i] = P_new[kk * nb2 + i] - density[kk * nb2 + i];
        });
        parallel_for(nk, [&](size_t kk) {
            std::vector<c64> FP(nb2), PF(nb2);
            matmul(F.data() + kk * nb2, density.data() + kk * nb2, FP.data(), nb, 0);
            matmul(density.data() + kk * nb2, F.data() + kk * nb2, PF.data(), nb, 0);
            for (size_t i = 0; i < nb2; ++i) comm[kk * nb2 + i] = FP[i] - PF[i];
        });
        const double d_res = weighted_matrix_norm(delta.data(), K.w2d, nk, nb, K.weight_sum);
        const double c_res = weighted_matrix_norm(comm.data(), K.w2d, nk, nb, K.weight_sum);
        converged = (d_res <= cfg.density_tol) && (c_res <= cfg.comm_tol);

        std::vector<c64> density_pre;
        if (cfg.trust_radius > 0.0 && !converged) density_pre = density;

        auto linear_mix = [&](double alpha) {
            for (size_t i = 0; i < n_tot; ++i) density[i] = density[i] + alpha * delta[i];
            project(density.data());
            hermitize_inplace(density.data(), nk, nb);
        };

        if (converged) {
            density = P_new;
        } else if (use_diis) {
            if (diis_F.size() >= cfg.diis_size) {
                diis_F.erase(diis_F.begin());
                diis_R.erase(diis_R.begin());
            }
            diis_F.emplace_back(F.begin(), F.end());
            diis_R.emplace_back(comm.begin(), comm.end());
            const size_t m = diis_F.size();
            if (m >= std::max<size_t>(2, cfg.diis_start)) {
                const size_t n = m + 1;
                std::vector<double> B(n * n, 0.0), rhs(n, 0.0), sol;
                for (size_t i = 0; i < m; ++i) {
                    for (size_t j = i; j < m; ++j) {
                        double total = 0.0;
                        for (size_t kk = 0; kk < nk; ++kk) {
                            double s = 0.0;
                            for (size_t a = 0; a < nb2; ++a)
                                s += (std::conj(diis_R[i][kk * nb2 + a]) * diis_R[j][kk * nb2 + a]).real();
                            total += K.w2d[kk] * s;
                        }
                        B[i * n + j] = total;
                        B[j * n + i] = total;
                    }
                    B[i * n + m] = -1.0;
                    B[m * n + i] = -1.0;
                }
                rhs[m] = -1.0;
                bool ok = fullpiv_solve(B, rhs, n, sol);
                for (size_t i = 0; ok && i < m; ++i) ok = std::isfinite(sol[i]);
                if (ok) {
                    std::vector<c64> Fx(n_tot, c64(0.0, 0.0));
                    for (size_t i = 0; i < m; ++i)
                        for (size_t j = 0; j < n_tot; ++j) Fx[j] += c64(sol[i], 0.0) * diis_F[i][j];
                    const double damp = std::clamp(cfg.diis_damping, 0.0, 1.0);
                    if (damp < 1.0)
                        for (size_t j = 0; j < n_tot; ++j)
                            Fx[j] = c64(damp, 0.0) * Fx[j] + c64(1.0 - damp, 0.0) * F[j];
                    hermitize_inplace(Fx.data(), nk, nb);
                    std::vector<c64> Pd(n_tot);
                    density_from_fock(Fx.data(), nk, nb, K.w2d, n_e, K.T, cfg.level_shift,
                                      Pd.data(), sizes);
                    project(Pd.data());
                    hermitize_inplace(Pd.data(), nk, nb);
                    density = std::move(Pd);
                } else {
                    linear_mix(cfg.mixing);
                }
            } else {
                linear_mix(cfg.mixing);
            }
        } else if (use_oda) {
            build_fock_compact(K, P_new.data(), Sigma_new.data(), F_new.data(), hdiag_new.data(), proj);
            const double E_new = hf_energy_with_hartree_diag(K, P_new.data(), Sigma_new.data(),
                                                             hdiag_new.data());
            double c1 = 0.0;
            for (size_t kk = 0; kk < nk; ++kk) {
                double s = 0.0;
                for (size_t a = 0; a < nb; ++a)
                    for (size_t b = 0; b < nb; ++b)
                        s += (F[kk * nb2 + a * nb + b] * delta[kk * nb2 + b * nb + a]).real();
                c1 += K.w2d[kk] * s;
            }
            const double c2 = (E_new - E_iter) - c1;
            double lam = 1.0;
            if (c2 > 1.0e-12) lam = std::clamp(-c1 / (2.0 * c2), 0.0, 1.0);
            else if (c1 < 0.0) lam = 1.0;
            else lam = 0.0;
            for (size_t i = 0; i < n_tot; ++i) density[i] = density[i] + c64(lam, 0.0) * delta[i];
            project(density.data());
            hermitize_inplace(density.data(), nk, nb);
        } else {
            linear_mix(cfg.mixing);
        }

        if (cfg.trust_radius > 0.0 && !converged && !density_pre.empty()) {
            double total = 0.0;
            for (size_t kk = 0; kk < nk; ++kk) {
                double s = 0.0;
                for (size_t i = 0; i < nb2; ++i)
                    s += std::norm(density[kk * nb2 + i] - density_pre[kk * nb2 + i]);
                total += K.w2d[kk] * s;
            }
            const double step = std::sqrt(total / std::max(K.weight_sum, 1.0e-30));
            if (step > cfg.trust_radius) {
                const double sc = cfg.trust_radius / step;
                for (size_t i = 0; i < n_tot; ++i)
                    density[i] = density_pre[i] + c64(sc, 0.0) * (density[i] - density_pre[i]);
                project(density.data());
                hermitize_inplace(density.data(), nk, nb);
            }
        }
        hE[k] = E_iter;
        hD[k] = d_res;
        hC[k] = c_res;
        ++k;
    }
    double E_fin = 0.0, mu_fin = 0.0;
    build_and_occupy(density, P_new.data(), E_fin, mu_fin);
    if (out->density) std::memcpy(out->density, density.data(), n_tot * sizeof(c64));
    if (out->fock) std::memcpy(out->fock, F.data(), n_tot * sizeof(c64));
    out->energy = E_fin;
    out->mu = mu_fin;
    out->iterations = k;
    out->converged = converged ? 1 : 0;
    if (out->hist_E) std::memcpy(out->hist_E, hE.data(), k * sizeof(double));
    if (out->hist_density) std::memcpy(out->hist_density, hD.data(), k * sizeof(double));
    if (out->hist_comm) std::memcpy(out->hist_comm, hC.data(), k * sizeof(double));
}

template <class Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        return HF_OK;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return HF_EINVAL;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return HF_ERUNTIME;
    }
}

inline c64* C(double* p) { return reinterpret_cast<c64*>(p); }
inline const c64* C(const double* p) { return reinterpret_cast<const c64*>(p); }

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* hfcpu_last_error(void) { return g_last_error.c_str(); }

int hfcpu_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int hfcpu_fft2d(double* buf, size_t nk1, size_t nk2, size_t batch, int inverse) {
    return guarded([&] { fft2d(C(buf), nk1, nk2, batch, inverse != 0); });
}

int hfcpu_selfenergy_fft_full(const double* VR, const double* P, double* sigma, size_t nk1,
                              size_t nk2, size_t nb, size_t dv1, size_t dv2,
                              int apply_ifftshift, int hcp) {
    return guarded([&] {
        if (hcp) {
            if (dv1 != 1 || dv2 != 1)
                throw std::invalid_argument(
                    "hermitian_channel_packing requires VR with shape (...,1,1)");
            selfenergy_fft_full_hcp(C(P), C(sigma), C(VR), nk1, nk2, nb);
        } else {
            selfenergy_fft_full(C(P), C(sigma), C(VR), nk1, nk2, nb, dv1, dv2);
        }
        if (apply_ifftshift) ifftshift_2d_batch(C(sigma), nk1, nk2, nb * nb);
    });
}

int hfcpu_eigh_batched(const double* M, double* w, double* V, size_t nk, size_t nb) {
    return guarded([&] { eigh_batched(C(M), w, C(V), nk, nb); });
}

int hfcpu_eigh_block_sizes(const double* M, double* w, double* V, size_t nk, size_t nb,
                           const size_t* sizes, size_t nsizes, int sort) {
    return guarded([&] {
        const std::vector<size_t> sz(sizes, sizes + nsizes);
        size_t total = 0;
        for (size_t x : sz) total += x;
        if (total != nb) throw std::invalid_argument("block sizes must sum to nb");
        const size_t nb2 = nb * nb;
        parallel_for(nk, [&](size_t k) {
            eigh_block_sizes_one(C(M) + k * nb2, nb, sz, sort != 0, w + k * nb, C(V) + k * nb2);
        });
    });
}

int hfcpu_max_offblock_sizes(const double* M, size_t nk, size_t nb, const size_t* sizes,
                             size_t nsizes, double* out) {
    return guarded([&] {
        const std::vector<size_t> sz(sizes, sizes + nsizes);
        size_t total = 0;
        for (size_t x : sz) total += x;
        if (total > nb) throw std::invalid_argument("block sizes must sum to nb");
        *out = max_offblock_sizes(C(M), nk, nb, sz);
    });
}

int hfcpu_find_mu_bisection(const double* bands, const double* weights, size_t nk, size_t nb,
                            double n_e, double T, int maxiter, double* mu) {
    return guarded([&] { *mu = find_mu_bisection(bands, nk * nb, weights, nk, nb, n_e, T, maxiter); });
}

int hfcpu_find_mu_newton(const double* bands, const double* weights, size_t nk, size_t nb,
                         double n_e, double T, int maxiter, double* mu) {
    return guarded([&] { *mu = find_mu_newton_py(bands, weights, nk, nb, n_e, T, maxiter); });
}

int hfcpu_solve_mu_inloop(const double* eps, const double* w_norm, size_t nk, size_t nb,
                          double n_target_norm, double mu0, double T, int maxiter, double* mu) {
    return guarded([&] {
        *mu = solve_mu_inloop(eps, nk * nb, w_norm, nk, nb, n_target_norm, mu0, T, maxiter);
    });
}

int hfcpu_validate_electron_count(const double* weights, size_t nk, size_t nb, double n_e) {
    return guarded([&] { validate_electron_count(weights, nk, nb, n_e); });
}

int hfcpu_build_fock_apply(const hf_kernel_desc* kd, const double* P, double* Sigma, double* Hh,
                           double* F, double* E, hf_project_fn pfn, void* puser) {
    return guarded([&] {
        const Kernel K = make_kernel(kd);
        Projector pr{pfn, puser};
        build_fock(K, C(P), C(Sigma), C(Hh), C(F), pfn ? &pr : nullptr);
        *E = hf_energy(K, C(P), C(Sigma), C(Hh));
    });
}

int hfcpu_density_from_fock(const double* F, size_t nk, size_t nb, const double* w, double n_e,
                            double T, double level_shift, double* P_new, double* mu) {
    return guarded([&] { *mu = density_from_fock(C(F), nk, nb, w, n_e, T, level_shift, C(P_new)); });
}

int hfcpu_cayley_spectral_setup(const double* d, double* V, double* lam, size_t nk, size_t nb) {
    return guarded([&] { cayley_spectral_setup(C(d), C(V), lam, nk, nb); });
}

int hfcpu_cayley_unitary_from_spectrum(const double* V, const double* lam, double tau,
                                       double* U, size_t nk, size_t nb) {
    return guarded([&] { cayley_unitary_from_spectrum(C(V), lam, tau, C(U), nk, nb); });
}

int hfcpu_diag_UFU_from_spectrum(const double* V, const double* Fe, const double* lam,
                                 double tau, double* diag, size_t nk, size_t nb) {
    return guarded([&] { diag_UFU_from_spectrum(C(V), C(Fe), lam, tau, diag, nk, nb); });
}

int hfcpu_solve_scf(const hf_kernel_desc* kd, const double* P0, double n_e,
                    const hf_scf_config* cfg, hf_project_fn pfn, void* puser,
                    hf_scf_result* out) {
    return guarded([&] {
        const Kernel K = make_kernel(kd);
        Projector pr{pfn, puser};
        solve_scf(K, C(P0), n_e, *cfg, pfn ? &pr : nullptr, out);
    });
}

int hfcpu_solve_dm(const hf_kernel_desc* kd, const double* P0, double n_e,
                   const hf_dm_config* cfg, hf_project_fn pfn, void* puser, hf_dm_result* out) {
    return guarded([&] {
        const Kernel K = make_kernel(kd);
        validate_electron_count(K.w2d, K.nk(), K.nb, n_e);
        Projector pr{pfn, puser};
        if (cfg->optimizer == 1)
            solve_rtr(K, C(P0), n_e, *cfg, pfn ? &pr : nullptr, out);
        else
            solve_dm(K, C(P0), n_e, *cfg, pfn ? &pr : nullptr, out);
    });
}

int hfcpu_selfenergy_superlattice(const double* rho, double* sigma, const double* V,
                                  const double* V_orb, const int64_t* g_a_off,
                                  const int64_t* pair_i, const int64_t* pair_j,
                                  const int64_t* pair_start, size_t n_delta, size_t Nx, size_t Ny,
                                  size_t nkx, size_t nky, size_t n_G, size_t orb, int hermitian) {
    return guarded([&] {
        selfenergy_superlattice(C(rho), C(sigma), C(V), V_orb ? C(V_orb) : nullptr, g_a_off,
                                pair_i, pair_j, pair_start, n_delta, Nx, Ny, nkx, nky, n_G, orb,
                                hermitian != 0);
    });
}

int hfcpu_rtr_joint_hvp(const hf_kernel_desc* kd, const double* X, const double* dp,
                        const double* Q, const double* p, const double* Ft, double* HX,
                        double* Hp) {
    return guarded([&] {
        const Kernel K = make_kernel(kd);
        const size_t n_tot = K.n_dense();
        std::vector<c64> dPb(n_tot), rb(n_tot), Sb(n_tot), Fb(n_tot);
        std::vector<double> hb(K.nb, 0.0);
        joint_hvp(K, C(X), dp, C(Q), p, C(Ft), std::max(K.T, 1.0e-12), C(HX), Hp, dPb.data(),
                  rb.data(), Sb.data(), Fb.data(), hb.data(), 0.0);
    });
}

int hfcpu_deflation_bias(const double* P, const double* targets, size_t n_defl,
                         const double* w2d, size_t nk, size_t nb, double sigma, double length,
                         double* S_pen, double* Phi, double* diag) {
    return guarded([&] {
        *Phi = deflation_bias(C(P), n_defl ? C(targets) : nullptr, n_defl, w2d, nk, nb, sigma,
                              length, C(S_pen), *diag);
    });
}

int hfcpu_occupation_entropy(const double* p, const double* w_norm, size_t nk, size_t nb,
                             double* S) {
    return guarded([&] { *S = occupation_entropy(p, w_norm, nk, nb); });
}

int hfcpu_resample_kgrid_2d(const double* x, double* out, size_t nk_old, size_t nk_new,
                            size_t inner) {
    return guarded([&] { resample_kgrid_2d(C(x), C(out), nk_old, nk_new, inner); });
}

}  // extern "C"
