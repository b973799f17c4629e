// Exchange self-energy and Fock assembly on the GPU (host-side interface).
//
// Reference: selfenergy.hpp:60-190 (FFT convolution, Hermitian channel
// packing, ifftshift) and fock.hpp:34-299 (build_fock_compact).
#pragma once

#include <vector>

#include "common.h"
#include "fft_lds.h"

namespace hfgpu {

constexpr int kFftThreads = 256;      // every exchange pass runs 256 threads
constexpr int kTileMaxPerThread = 16; // tile elements per thread (n * C <= 4096)

// Twiddle table + radix schedule for one FFT length (device resident).
struct AxisPlan {
    int n = 0;
    std::vector<int> radices;
    cd* d_tw = nullptr;
    FftAxis axis() const;
    void init(int n_);
    void release();
};

// Channel layout of the exchange transform for one (nk1, nk2, nb) geometry.
//   packed: channels are the upper triangle (a <= b) of the nb x nb block,
//           valid when the transformed density is Hermitian and the
//           interaction is symmetric in (a, b).
//   full:   all nb*nb channels in row-major order.
struct ExchangeGeom {
    int nk1 = 0, nk2 = 0, nb = 0;
    int nch = 0;
    bool packed = false;
    int clog = 0;   // log2 of channels per workgroup tile
    int nblk = 0;   // channel tiles
    size_t lds_rows = 0, lds_cols = 0;
    int* d_chan_src = nullptr;  // nch: a * nb + b
    int* d_chan_a = nullptr;    // nch
    int* d_chan_b = nullptr;    // nch
    AxisPlan ax1, ax2;          // along k1 (nk1) and along k2 (nk2)

    void init(int nk1_, int nk2_, int nb_, bool packed_);
    void release();
};

// Epilogue of the inverse row pass.
enum ExchEpilogue : int {
    EPI_RAW_FULL = 0,    // Sigma_x for all channels
    EPI_RAW_PACKED = 1,  // Sigma_x upper triangle + conj mirror
    EPI_FOCK_PACKED = 2  // Sigma = Sigma_x + SC, F = herm(h + Sigma + diag(H))
};

// k-independent terms derived from rho_bar = sum_k w_k (P_k - refP_k):
// the Hartree diagonal HH n (fock.hpp:222-241) and the contact self-energy
// (fock.hpp:244-279).  Computed by the last workgroup of pass A.
struct SmallTerms {
    int enabled = 0;
    int include_hartree = 0;
    int has_contact = 0;
    const double* HH = nullptr;
    int n_contact = 0;
    const double* contact_g = nullptr;
    const cd* Oi = nullptr;
    const cd* Oj = nullptr;
    double* hdiag = nullptr;  // out (nb)
    cd* SC = nullptr;         // out (nb, nb)
};

struct ExchangeArgs {
    const cd* P = nullptr;     // (nk, nb, nb)
    const cd* refP = nullptr;  // nullable
    const cd* VR = nullptr;    // (nk, nvr) in the geometry's channel order, nvr = 1 or nch
    int nvr = 1;
    cd* T = nullptr;           // workspace (nk, nch)
    cd* partial = nullptr;     // (nk1, nch) weighted row sums of P - refP, nullable
    const double* w2d = nullptr;
    SmallTerms small;
    int epilogue = EPI_RAW_FULL;
    bool ifftshift = false;
    cd* Sigma = nullptr;       // output
    cd* F = nullptr;           // EPI_FOCK_PACKED
    const cd* h = nullptr;     // EPI_FOCK_PACKED
    const cd* SC = nullptr;    // (nb, nb) contact self-energy, nullable
    const double* hdiag = nullptr;  // (nb) Hartree diagonal, nullable
};

// Pass A without the transform: weighted row sums (+ small terms) only, for
// kernels with exchange off but Hartree/contact terms on.
void launch_row_sums(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st);
// Passes A (rows fwd, + row sums and small terms) and B (cols fwd, x VR,
// cols inv).
void launch_exchange_fwd(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st);
// Pass C (rows inv + epilogue).
void launch_exchange_inv(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st);

// Full-path assembly: Sigma += SC, F = herm(h + Sigma + diag(hdiag)).
void launch_fock_assemble(int nk, int nb, const cd* h, cd* Sigma, cd* F, const cd* SC,
                          const double* hdiag, hipStream_t st);

// E = sum_k w_k Re[Tr(h P) + 1/2 Tr((Sigma + H)(P - refP))]  (two-stage,
// deterministic).  partial needs cdiv(nk, 4) doubles.
void launch_energy(int nk, int nb, const cd* h, const cd* P, const cd* refP, const cd* Sigma,
                   const double* hdiag, const double* w2d, double* partial, double* E_out,
                   hipStream_t st);
// Same with a full Hartree matrix Hm (k stride hk_stride, 0 = broadcast).
void launch_energy_fullH(int nk, int nb, const cd* h, const cd* P, const cd* refP,
                         const cd* Sigma, const cd* Hm, size_t hk_stride, const double* w2d,
                         double* partial, double* E_out, hipStream_t st);
// S = -sum w [p ln p + (1-p) ln(1-p)] with the 1e-14 clip (fock.hpp:385-401);
// partial needs min(cdiv(nk nb, 256), 1024) doubles.
void launch_entropy(const double* p, const double* w, int nk, int nb, double* partial,
                    double* S_out, hipStream_t st);

}  // namespace hfgpu
