// Superlattice (plane-wave basis) exchange and Hartree shift on the GPU.
//
// Reference: superlattice.hpp:50-246 (selfenergy_superlattice_streamed) and
// fock.hpp:66-175 (the superlattice Hartree shift of build_fock_compact).
//
// The density of a G-basis Hamiltonian, rho[k][(G_a, a)][(G_b, b)], is a
// function of the absolute momenta p = k - G_a and p' = k - G_b; every
// channel (dG = G_a - G_b, a, b) is one 2-D convolution in p with the lag
// kernel V.  The channels of all transformed dG are laid out side by side
// on the extended grid, X[px][py][slot][a][b] (slot = transformed dG in
// order, channel index fastest), and transformed together:
//   rows     forward transform along py of the grid rows that hold data,
//            loading P - refP straight into the tile of G_a (the scatter)
//   columns  forward transform along px (rows without data read as zero),
//            X *= -V(q) / (N_ext_x N_ext_y) (V per orbital pair optional),
//            inverse transform, in one pass
//   rows     inverse transform along py of the rows with data
//   gather   Sigma[k][(G_a, a)][(G_b, b)] = X at the tile of G_a, or, for a
//            channel filled from its (-dG) partner, conj of the partner's
//            value at the tile of G_b with (a, b) swapped.
// Channels are processed in batches of at most 1024 (the plain transform's
// channel limit) and about 2 GiB of grid.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "hfgpu_core.h"

namespace hfgpu {

// Host description of one superlattice exchange (all host pointers).
struct SlInput {
    const cd* V = nullptr;        // (Nx, Ny) lag kernel transform
    const cd* V_orb = nullptr;    // (Nx, Ny, orb, orb), nullable: replaces V
    const int64_t* g_a_off = nullptr;     // (n_G, 2) tile offsets, units of (nkx, nky)
    const int64_t* pair_i = nullptr;      // (n_pairs)
    const int64_t* pair_j = nullptr;      // (n_pairs)
    const int64_t* pair_start = nullptr;  // (n_delta + 1) CSR row pointer
    size_t n_pairs = 0, n_delta = 0, Nx = 0, Ny = 0, nkx = 0, nky = 0, n_G = 0, orb = 0;
    bool hermitian = false;  // fill each (-dG) partner channel from its (+dG) one
};

// Device-resident plan of the superlattice exchange for one layout.
struct SlPlan {
    int nkx = 0, nky = 0, n_G = 0, orb = 0, orb2 = 0, D = 0, Nx = 0, Ny = 0;
    int ntx = 0, nty = 0;           // tiles per grid axis
    int n_slots = 0;                // transformed channels
    int slots_per_batch = 0;
    bool orbital = false;
    DevBuf V;          // (Nx Ny period) complex, period = orb^2 or 1
    DevBuf tiles;      // (n_slots, ntx, nty) int: i * n_G + j of the pair placed there, or -1
    DevBuf src;        // (n_G, n_G) int4 {slot, G whose tile is read, conj, 0}; slot -1: zero
    DevBuf goff;       // (n_G, 2) int
    DevBuf X;          // grid of one batch
    DevBuf flags;      // (slots_per_batch) int: channel holds a nonzero value
    DevBuf rowmask;    // (batches, channel tiles, Nx) int: grid row px of a tile holds data
    int tiles_per_batch = 0;
    // every Sigma block has a source and no two pairs of a slot share a
    // tile: the inverse rows store Sigma themselves (no gather pass)
    bool direct_store = false;
    std::unique_ptr<ExchangeGeom> geo;  // Nx x Ny plain transform

    SlPlan() = default;
    SlPlan(const SlPlan&) = delete;
    SlPlan& operator=(const SlPlan&) = delete;
    ~SlPlan();
    // Validates the layout (every index in bounds, every tile inside the grid)
    // and uploads the tables; throws std::invalid_argument on bad input.
    void init(const SlInput& in, hipStream_t st);
    // Sigma (nkx nky, D, D) from P - refP (refP nullable; refP0 nullable: one
    // (D, D) reference for every k).  Device pointers on stream st.
    void apply(const cd* P, const cd* refP, const cd* refP0, cd* Sigma, hipStream_t st);
};

// Device tables of the superlattice Hartree shift (fock.hpp:66-175).
struct SlHartree {
    int n_G = 0, orb = 0, nb = 0, n_used = 0;  // n_used: channels referenced by pair_to_delta
    double deg = 1.0;
    bool orbital = false;
    DevBuf start;   // (n_used + 1) int: pairs of each used channel
    DevBuf pairs;   // (n_G^2) int: A * n_G + B in row-major order within a channel
    DevBuf cidx;    // (n_G^2) int: used-channel index of each pair
    DevBuf HH;      // (n_G^2) or (n_G^2 orb^2) real
    void init(const hf_superlattice_desc& d, hipStream_t st);
    // SC = (accumulate ? SC : 0) + shift(rho_bar); rho_bar (nb, nb) on device.
    void launch(const cd* rho_bar, cd* SC, bool accumulate, hipStream_t st) const;
};

}  // namespace hfgpu
