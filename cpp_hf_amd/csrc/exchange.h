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
//   packed: channels are the upper triangle (a <= b) of the nb x nb block in
//           row-major order, valid when the transformed density is Hermitian
//           and the interaction is symmetric in (a, b).
//   full:   all nb*nb channels in row-major order.
// Channels are processed in tiles of C = 2^clog; the last tile is padded
// with zero channels.
//
// Workspace layout (tile-major): T[blk][k1][k2][c], so the row passes move
// one contiguous nk2 * C block per workgroup and the column pass reads C
// contiguous channels per k point.  The per-geometry constant operands use
// the layout of the pass that reads them:
//   VR tile-major by column: VRb[blk][k2][k1][c]   (pass B)
//   herm(h) tile-major by row: hhb[blk][k1][k2][c] (pass C, packed Fock)
struct ExchangeGeom {
    int nk1 = 0, nk2 = 0, nb = 0;
    int nch = 0;
    bool packed = false;
    int clog = 0;     // log2 of channels per workgroup tile
    int nblk = 0;     // channel tiles
    int slots_row = 0, slots_col = 0;  // tile elements per thread (power of 2)
    bool generic = false;              // a radix 7, 11 or 13 stage is present
    // two-step register transforms (tile of 8 channels, N = N1 x N2 with
    // both factors at most 16): rows along k2, cols along k1
    bool fast_rows = false, fast_cols = false;
    size_t lds_rows = 0, lds_cols = 0;
    // two-step launch shape: wavefronts (channel tiles) per row workgroup
    // (0: the whole row), wavefronts per column workgroup
    int row_waves = 1, col_waves = 2;
    // wavefronts sharing one channel tile in the two-step passes (1 or 2):
    // rows forward, columns, rows inverse
    int split_fwd = 2, split_col = 2, split_inv = 1;
    int* d_chan_a = nullptr;  // nch
    int* d_chan_b = nullptr;  // nch
    AxisPlan ax1, ax2;        // along k1 (nk1) and along k2 (nk2)

    void init(int nk1_, int nk2_, int nb_, bool packed_);
    void release();
    int tile() const { return 1 << clog; }
    // complex elements of the workspace and of the tile-major operands
    size_t tile_elems() const {
        return static_cast<size_t>(nblk) * tile() * nk1 * nk2;
    }
};

// Epilogue of the inverse row pass.
enum ExchEpilogue : int {
    EPI_RAW_FULL = 0,    // Sigma_x for all channels
    EPI_RAW_PACKED = 1,  // Sigma_x upper triangle + conj mirror
    EPI_FOCK_PACKED = 2  // Sigma = Sigma_x + SC, F = herm(h) + herm(Sigma) + diag(H)
};

// k-independent terms derived from rho_bar = sum_k w_k (P_k - refP_k):
// the Hartree diagonal HH n (fock.hpp:222-241) and the contact self-energy
// (fock.hpp:244-279).  Computed by one workgroup after pass A.
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
    cd* rho_out = nullptr;    // out (nb, nb): rho_bar itself, nullable
};

struct ExchangeArgs {
    const cd* P = nullptr;      // (nk, nb, nb)
    const cd* refP = nullptr;   // (nk, nb, nb), nullable
    const cd* refP0 = nullptr;  // (nb, nb) k-independent reference density, nullable
    const double* wrow = nullptr;  // (nk1) row sums of w2d, needed with refP0
    const cd* VR = nullptr;     // VRb (tile-major) or scalar (nk) when vr_scalar
    bool vr_scalar = false;
    cd* T = nullptr;            // workspace, tile_elems()
    cd* partial = nullptr;      // (nk1, nch) weighted row sums of P - refP, nullable
    const double* w2d = nullptr;
    SmallTerms small;
    int epilogue = EPI_RAW_FULL;
    bool ifftshift = false;
    cd* Sigma = nullptr;        // output
    cd* F = nullptr;            // EPI_FOCK_PACKED
    const cd* hh = nullptr;     // EPI_FOCK_PACKED: herm(h), tile-major by row
    const cd* SC = nullptr;     // (nb, nb) contact self-energy, nullable
    const double* hdiag = nullptr;  // (nb) Hartree diagonal, nullable
    // mixed precision: workspace and interaction stored as complex float
    // (two-step meshes only); the transforms' arithmetic stays FP64
    bool ws32 = false;
    cf* T32 = nullptr;          // workspace, tile_elems()
    const cf* VR32 = nullptr;   // VRb or scalar VR in FP32
};

// The operands of one solve in a batch of Fock builds on one geometry: the
// ExchangeArgs fields that differ from solve to solve.  The two-step passes
// take a device array of them, entry s for workgroup row s (blockIdx.y).
struct ExchBatchEntry {
    int sv = 0;  // the solve's bit in the launch's active mask
    const cd* P = nullptr;
    const cd* refP = nullptr;
    const cd* refP0 = nullptr;
    const double* wrow = nullptr;
    cd* T = nullptr;
    cd* partial = nullptr;
    const double* w2d = nullptr;
    const cd* VR = nullptr;
    SmallTerms small;
    cd* Sigma = nullptr;
    cd* F = nullptr;
    const cd* hh = nullptr;
    const cd* SC = nullptr;
    const double* hdiag = nullptr;
};

// Pass A without the transform: weighted row sums (+ small terms) only, for
// kernels with exchange off but Hartree/contact terms on.
void launch_row_sums(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st);
// Passes A (rows fwd, + row sums) and B (cols fwd, x VR, cols inv; + the
// small terms).
void launch_exchange_fwd(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st);
// Pass C (rows inv + epilogue).
void launch_exchange_inv(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st);
// The same passes for S solves on one two-step geometry (fast_rows and
// fast_cols, FP64 workspace) in one launch each: a gives the flags
// (vr_scalar, small.enabled, epilogue, ifftshift), tab (device, S entries)
// each solve's operands; entries whose bit sv in active is clear are
// skipped.  Every solve's outputs equal its single call's.
void launch_exchange_fwd_batch(const ExchangeGeom& g, const ExchangeArgs& a,
                               const ExchBatchEntry* tab, int S, unsigned active,
                               hipStream_t st);
void launch_exchange_inv_batch(const ExchangeGeom& g, const ExchangeArgs& a,
                               const ExchBatchEntry* tab, int S, unsigned active,
                               hipStream_t st);

// Tile-major operands of a geometry:
//   VRb from a (nk, nb*nb) interaction (channel (a, b) <- VR[q][a * nb + b]);
//   hhb = herm(h) over the geometry's channels.
void launch_pack_vr(const ExchangeGeom& g, const cd* VR_full, cd* VRb, hipStream_t st);
void launch_pack_herm_h(const ExchangeGeom& g, const cd* h, cd* hhb, hipStream_t st);

// Full-path assembly: Sigma += SC, F = herm(h + Sigma + diag(hdiag)).
void launch_fock_assemble(int nk, int nb, const cd* h, cd* Sigma, cd* F, const cd* SC,
                          const double* hdiag, hipStream_t st);

// E = sum_k w_k Re[Tr(h P) + 1/2 Tr((Sigma + H)(P - refP))]  (two-stage,
// deterministic).  partial needs cdiv(nk, 4) doubles.
void launch_energy(int nk, int nb, const cd* h, const cd* P, const cd* refP, const cd* Sigma,
                   const double* hdiag, const double* w2d, double* partial, double* E_out,
                   hipStream_t st);
// launch_energy for the S solves of a batch in two launches: tab (device)
// holds each solve's operands, nk and output; nk_max is the largest nk.
// Solves whose bit in active is clear are skipped.  Each energy equals its
// single call's bit for bit.
struct EnergyBatchEntry {
    const cd* h = nullptr;
    const cd* P = nullptr;
    const cd* refP = nullptr;
    const cd* Sigma = nullptr;
    const double* hdiag = nullptr;
    const double* w2d = nullptr;
    double* partial = nullptr;  // cdiv(nk, 4) doubles
    double* E = nullptr;
    int nk = 0;
};
void launch_energy_batch(const EnergyBatchEntry* tab, int S, int nk_max, int nb,
                         unsigned active, hipStream_t st);
// Same with a full Hartree matrix Hm (k stride hk_stride, 0 = broadcast).
void launch_energy_fullH(int nk, int nb, const cd* h, const cd* P, const cd* refP,
                         const cd* Sigma, const cd* Hm, size_t hk_stride, const double* w2d,
                         double* partial, double* E_out, hipStream_t st);
// S = -sum w [p ln p + (1-p) ln(1-p)] with the 1e-14 clip (fock.hpp:385-401);
// partial needs min(cdiv(nk nb, 256), 1024) doubles.
void launch_entropy(const double* p, const double* w, int nk, int nb, double* partial,
                    double* S_out, hipStream_t st);


// out = (complex float) in, n elements.
void launch_to_cf(size_t n, const cd* in, cf* out, hipStream_t st);

// ---- adjoint building blocks (natural layout X[k1][k2][ch]) ----------------
// Unnormalised 2-D transform of nch <= g.nch channels of a full-channel
// geometry (forward exp(-2 pi i / n) or inverse); in == out allowed.
void launch_fft2_plain(const ExchangeGeom& g, const cd* in, cd* out, int nch, bool inverse,
                       hipStream_t st);
// Source of the superlattice convolution's forward rows: grid row px of
// slot ls holds, at py, the (a, b) entry of the pair i n_G + j that
// tiles[(ls ntx + px / nkx) nty + py / nky] names (-1: zero) at k =
// (px mod nkx, py mod nky), minus the reference density; rows load it
// straight from P (the scatter folded into the row pass) and set
// flags[ls] when a value is nonzero.
struct SlRowsSrc {
    const cd* P = nullptr;
    const cd* refP = nullptr;   // (nk, D, D), nullable
    const cd* refP0 = nullptr;  // (D, D) for every k, nullable
    const int* tiles = nullptr;  // the batch's first slot
    int* flags = nullptr;
    int nkx = 0, nky = 0, ntx = 0, nty = 0, D = 0, n_G = 0, orb = 0, orb2 = 0;
    // inverse rows storing Sigma directly (nullable Sigma: store X and let
    // the caller gather): srct is the (n_G, n_G) table of 4 ints {slot, G
    // whose tile holds the value, conj, 0} naming each Sigma block's source,
    // s0 the batch's first slot
    const int* srct = nullptr;
    cd* Sigma = nullptr;
    int s0 = 0;
};
// 2-D convolution of the nch <= g.nch channels of one superlattice batch
// in natural layout X[px][py][ch]: forward rows from the density (only the
// rows that rowmask, (channel tiles, nk1) ints, marks as holding data), the
// columns forward x scale V x inverse (V[q * period + ch % period]; rows not
// marked read as zero; channels whose flags[ch / per_slot] is 0 give exactly
// zero), inverse rows of the marked rows (the only ones read afterwards),
// stored into X, or with src.Sigma straight into the Sigma blocks whose
// source (src.srct) is this batch's value.
void launch_sl_convolve(const ExchangeGeom& g, const SlRowsSrc& src, cd* X, int nch,
                        const cd* V, int period, const int* flags, int per_slot, double scale,
                        const int* rowmask, hipStream_t st);
// out = scale A conj(B) per element, or summed over channels per q.
void launch_cross_spectrum(size_t nq, int nch, const cd* A, const cd* B, double scale,
                           bool sum_channels, cd* out, hipStream_t st);
// X *= scale op(V), V per (q, ch) or per q (v_scalar), op = conj with conj_v.
void launch_mul_spectrum(size_t nq, int nch, cd* X, const cd* V, bool v_scalar, bool conj_v,
                         double scale, hipStream_t st);
// X *= conj(ifftshift phase of q) (the phase folded into VR, problem.py:95-104).
void launch_unphase(int nk1, int nk2, int nch, cd* X, hipStream_t st);
// out[e] = sum_k (w[k]) X[k][e], e < m; partial needs kKsumParts * m complex.
constexpr int kKsumParts = 256;
void launch_ksum(int nk, int m, const cd* X, const double* w, cd* partial, cd* out,
                 hipStream_t st);

}  // namespace hfgpu
