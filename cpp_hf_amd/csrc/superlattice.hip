// Superlattice exchange (superlattice.hpp:50-246) and the superlattice
// Hartree shift (fock.hpp:66-175) on the GPU.  Layout and passes are
// described in superlattice.h.
#include "superlattice.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>

#include "../../include/hfgpu.h"

namespace hfgpu {
namespace {

unsigned grid_for(size_t n) { return std::max(1u, std::min<unsigned>(cdiv(n, 256), 8192u)); }

struct SlSrc {
    int slot;  // transformed channel, -1: no channel writes this pair
    int g;     // G whose tile holds the value
    int conj;  // 1: conj of the (b, a) entry (the -dG partner's fill)
    int pad;
};

constexpr size_t kSlGridBytes = size_t(2) << 30;  // grid of one batch, at most
constexpr int kSlMaxChannels = 1024;              // plain transform channel limit

// Sigma entries whose channel lies in slots [s0, s0 + nslots); with first,
// also the zero entries of pairs that no channel writes
__global__ void __launch_bounds__(256)
k_sl_gather(size_t n, int D, int orb, int orb2, int n_G, int nkx, int nky, int Ny, int nch,
            const SlSrc* __restrict__ src, const int* __restrict__ goff, int s0, int nslots,
            int first, const cd* __restrict__ X, cd* __restrict__ Sigma) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int col = static_cast<int>(e % D);
        const int r = static_cast<int>((e / D) % D);
        const int k = static_cast<int>(e / (static_cast<size_t>(D) * D));
        const int R = r / orb, ro = r % orb, Cg = col / orb, co = col % orb;
        const SlSrc s = src[R * n_G + Cg];
        if (s.slot < 0) {
            if (first) Sigma[e] = cd{0.0, 0.0};
            continue;
        }
        const int ls = s.slot - s0;
        if (ls < 0 || ls >= nslots) continue;
        const int kx = k / nky, ky = k % nky;
        const int px = goff[2 * s.g] * nkx + kx, py = goff[2 * s.g + 1] * nky + ky;
        const int ch = ls * orb2 + (s.conj ? co * orb + ro : ro * orb + co);
        const cd v = X[(static_cast<size_t>(px) * Ny + py) * nch + ch];
        Sigma[e] = s.conj ? cconj(v) : v;
    }
}

// One workgroup: per-channel sums of rho_bar over the (G_a, G_b) pairs of
// each channel (row-major pair order), then the shift
//   scalar:  S[(A x), (B x)] = deg HH[A, B] sum_pairs(d) sum_x rho_bar[(A' x), (B' x)]
//   orbital: S[(A a), (B a)] = sum_g deg HHo[A, B, a, g] sum_pairs(d) rho_bar[(A' g), (B' g)]
// with d the channel of (A, B).
__global__ void __launch_bounds__(256)
k_sl_hartree(int n_G, int orb, int nb, int n_used, int orbital, double deg,
             const int* __restrict__ start, const int* __restrict__ pairs,
             const int* __restrict__ cidx, const double* __restrict__ HH,
             const cd* __restrict__ rho, cd* __restrict__ SC, int accumulate) {
    extern __shared__ cd pd[];
    const int per = orbital ? orb : 1;
    for (int t = threadIdx.x; t < n_used * per; t += blockDim.x) {
        const int c = t / per, g = t % per;
        cd acc{0.0, 0.0};
        for (int p = start[c]; p < start[c + 1]; ++p) {
            const int A = pairs[p] / n_G, B = pairs[p] % n_G;
            if (orbital) {
                acc = cadd(acc, rho[static_cast<size_t>(A * orb + g) * nb + B * orb + g]);
            } else {
                cd s{0.0, 0.0};
                for (int x = 0; x < orb; ++x)
                    s = cadd(s, rho[static_cast<size_t>(A * orb + x) * nb + B * orb + x]);
                acc = cadd(acc, s);
            }
        }
        pd[t] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        const int r = e / nb, cc = e % nb;
        const int A = r / orb, a = r % orb, B = cc / orb, b = cc % orb;
        cd v{0.0, 0.0};
        if (a == b) {
            const int c = cidx[A * n_G + B];
            if (orbital) {
                for (int g = 0; g < orb; ++g)
                    v = cadd(v, cscale(pd[c * orb + g],
                                       deg * HH[(static_cast<size_t>(A * n_G + B) * orb + a) * orb + g]));
            } else {
                v = cscale(pd[c], deg * HH[A * n_G + B]);
            }
        }
        SC[e] = accumulate ? cadd(SC[e], v) : v;
    }
}

void check_layout(const SlInput& in) {
    if (!in.V || !in.g_a_off || !in.pair_i || !in.pair_j || !in.pair_start)
        throw std::invalid_argument("superlattice layout has a null array");
    if (!in.nkx || !in.nky || !in.n_G || !in.orb || !in.Nx || !in.Ny)
        throw std::invalid_argument("superlattice sizes must be non-zero");
    if (in.Nx > 4096 || in.Ny > 4096 || in.n_G * in.orb > 4096 || in.nkx * in.nky > (1u << 24))
        throw std::invalid_argument("superlattice grid too large for the GPU exchange");
    if (in.pair_start[0] < 0) throw std::invalid_argument("pair_start must be non-negative");
    for (size_t d = 0; d < in.n_delta; ++d)
        if (in.pair_start[d + 1] < in.pair_start[d])
            throw std::invalid_argument("pair_start must be non-decreasing");
    if (static_cast<size_t>(in.pair_start[in.n_delta]) > in.n_pairs)
        throw std::invalid_argument("pair_start runs past the end of pair_i / pair_j");
    const int64_t nG = static_cast<int64_t>(in.n_G);
    for (int64_t p = in.pair_start[0]; p < in.pair_start[in.n_delta]; ++p)
        if (in.pair_i[p] < 0 || in.pair_i[p] >= nG || in.pair_j[p] < 0 || in.pair_j[p] >= nG)
            throw std::invalid_argument("pair_i / pair_j entries must lie in [0, n_G)");
    for (size_t g = 0; g < in.n_G; ++g) {
        const int64_t ox = in.g_a_off[2 * g], oy = in.g_a_off[2 * g + 1];
        if (ox < 0 || oy < 0 || static_cast<size_t>(ox + 1) * in.nkx > in.Nx ||
            static_cast<size_t>(oy + 1) * in.nky > in.Ny)
            throw std::invalid_argument(
                "g_a_off places a G tile outside the (N_ext_x, N_ext_y) grid");
    }
}

}  // namespace

SlPlan::~SlPlan() {
    if (geo) geo->release();
}

void SlPlan::init(const SlInput& in, hipStream_t st) {
    check_layout(in);
    nkx = static_cast<int>(in.nkx);
    nky = static_cast<int>(in.nky);
    n_G = static_cast<int>(in.n_G);
    orb = static_cast<int>(in.orb);
    orb2 = orb * orb;
    D = n_G * orb;
    Nx = static_cast<int>(in.Nx);
    Ny = static_cast<int>(in.Ny);
    ntx = static_cast<int>(cdiv(in.Nx, in.nkx));
    nty = static_cast<int>(cdiv(in.Ny, in.nky));
    orbital = in.V_orb != nullptr;
    const size_t nd = in.n_delta;
    auto lo = [&](size_t d) { return in.pair_start[d]; };
    auto hi = [&](size_t d) { return in.pair_start[d + 1]; };

    // -dG partner of each channel: the channel whose pair set is this one's
    // with (i, j) reversed (first such channel)
    std::vector<int64_t> partner(nd, -1);
    if (in.hermitian) {
        std::map<std::vector<int64_t>, int64_t> first_with;
        std::vector<std::vector<int64_t>> rev(nd);
        for (size_t d = 0; d < nd; ++d) {
            std::vector<int64_t> fwd;
            for (int64_t p = lo(d); p < hi(d); ++p) {
                fwd.push_back(in.pair_i[p] * n_G + in.pair_j[p]);
                rev[d].push_back(in.pair_j[p] * n_G + in.pair_i[p]);
            }
            std::sort(fwd.begin(), fwd.end());
            std::sort(rev[d].begin(), rev[d].end());
            first_with.emplace(std::move(fwd), static_cast<int64_t>(d));
        }
        for (size_t d = 0; d < nd; ++d) {
            const auto it = first_with.find(rev[d]);
            if (it != first_with.end()) partner[d] = it->second;
        }
    }

    // transformed channels, their tile tables and the source of every
    // (G_a, G_b) block of Sigma (later writes win, as in the channel loop)
    std::vector<int> tile_tab;
    std::vector<SlSrc> src_tab(static_cast<size_t>(n_G) * n_G, SlSrc{-1, 0, 0, 0});
    const size_t tiles_per = static_cast<size_t>(ntx) * nty;
    n_slots = 0;
    bool collision = false;
    for (size_t d = 0; d < nd; ++d) {
        const bool fill = partner[d] >= 0 && static_cast<size_t>(partner[d]) != d;
        if (fill && static_cast<size_t>(partner[d]) < d) continue;
        if (lo(d) == hi(d)) continue;
        const int s = n_slots++;
        tile_tab.resize(static_cast<size_t>(n_slots) * tiles_per, -1);
        int* tt = tile_tab.data() + static_cast<size_t>(s) * tiles_per;
        for (int64_t p = lo(d); p < hi(d); ++p) {
            const int i = static_cast<int>(in.pair_i[p]), j = static_cast<int>(in.pair_j[p]);
            const int tx = static_cast<int>(in.g_a_off[2 * i]);
            const int ty = static_cast<int>(in.g_a_off[2 * i + 1]);
            int& cell = tt[static_cast<size_t>(tx) * nty + ty];
            if (cell >= 0) collision = true;
            cell = i * n_G + j;
            src_tab[static_cast<size_t>(i) * n_G + j] = SlSrc{s, i, 0, 0};
            if (fill) src_tab[static_cast<size_t>(j) * n_G + i] = SlSrc{s, i, 1, 0};
        }
    }

    bool all_sourced = true;
    for (const SlSrc& e : src_tab) all_sourced = all_sourced && e.slot >= 0;
    direct_store = !collision && all_sourced;
    const size_t nq = static_cast<size_t>(Nx) * Ny;
    const size_t period = orbital ? static_cast<size_t>(orb2) : 1;
    V.alloc(nq * period * sizeof(cd));
    h2d(V.get(), orbital ? in.V_orb : in.V, nq * period * sizeof(cd), st);
    src.alloc(src_tab.size() * sizeof(SlSrc));
    h2d(src.get(), src_tab.data(), src_tab.size() * sizeof(SlSrc), st);
    std::vector<int> off(2 * static_cast<size_t>(n_G));
    for (size_t t = 0; t < off.size(); ++t) off[t] = static_cast<int>(in.g_a_off[t]);
    goff.alloc(off.size() * sizeof(int));
    h2d(goff.get(), off.data(), off.size() * sizeof(int), st);

    slots_per_batch = 0;
    if (n_slots > 0) {
        const size_t by_bytes = kSlGridBytes / (nq * sizeof(cd));
        const size_t max_ch = std::min<size_t>(kSlMaxChannels, by_bytes);
        if (max_ch < static_cast<size_t>(orb2))
            throw std::invalid_argument(
                "superlattice grid too large: one channel's orbital block does not fit a batch");
        slots_per_batch = static_cast<int>(std::min<size_t>(n_slots, max_ch / orb2));
        const int nch = slots_per_batch * orb2;
        int nbg = 1;
        while (nbg * nbg < nch) ++nbg;
        geo = std::make_unique<ExchangeGeom>();
        geo->init(Nx, Ny, nbg, false);
        // rows px of each batch's channel tiles that hold data: the rows of
        // the tiles that the scatter fills, which are also the only rows
        // the gather reads (both follow the pairs' G_a tiles)
        const int C = geo->tile();
        tiles_per_batch = (nch + C - 1) / C;
        const int n_batches = (n_slots + slots_per_batch - 1) / slots_per_batch;
        std::vector<int> mask(static_cast<size_t>(n_batches) * tiles_per_batch * Nx, 0);
        for (int b = 0; b < n_batches; ++b)
            for (int t = 0; t < tiles_per_batch; ++t) {
                int* m = mask.data() + (static_cast<size_t>(b) * tiles_per_batch + t) * Nx;
                const int c0 = t * C, c1 = std::min(nch, c0 + C);
                for (int c = c0; c < c1; ++c) {
                    const int s = b * slots_per_batch + c / orb2;
                    if (s >= n_slots) break;
                    const int* tt = tile_tab.data() + static_cast<size_t>(s) * tiles_per;
                    for (int tx = 0; tx < ntx; ++tx)
                        for (int ty = 0; ty < nty; ++ty)
                            if (tt[static_cast<size_t>(tx) * nty + ty] >= 0)
                                for (int px = tx * nkx; px < std::min(Nx, (tx + 1) * nkx); ++px)
                                    m[px] = 1;
                }
            }
        rowmask.alloc(mask.size() * sizeof(int));
        h2d(rowmask.get(), mask.data(), mask.size() * sizeof(int), st);
        tiles.alloc(tile_tab.size() * sizeof(int));
        h2d(tiles.get(), tile_tab.data(), tile_tab.size() * sizeof(int), st);
        X.alloc(nq * nch * sizeof(cd));
        flags.alloc(static_cast<size_t>(slots_per_batch + 1) / 2 * 2 * sizeof(int));
    }
    HFGPU_CHECK(hipStreamSynchronize(st));
}

void SlPlan::apply(const cd* P, const cd* refP, const cd* refP0, cd* Sigma, hipStream_t st) {
    const size_t nq = static_cast<size_t>(Nx) * Ny;
    const size_t n_out = static_cast<size_t>(nkx) * nky * D * D;
    const int period = orbital ? orb2 : 1;
    const double scale = -1.0 / static_cast<double>(nq);
    const size_t tiles_per = static_cast<size_t>(ntx) * nty;
    int s0 = 0;
    do {
        const int nslots = std::min(slots_per_batch, n_slots - s0);
        const int nch = nslots * orb2;
        if (nslots > 0) {
            fill_zero(flags.get(), flags.size(), st);
            SlRowsSrc src;
            src.P = P;
            src.refP = refP;
            src.refP0 = refP0;
            src.tiles = tiles.as<int>() + static_cast<size_t>(s0) * tiles_per;
            src.flags = flags.as<int>();
            src.nkx = nkx;
            src.nky = nky;
            src.ntx = ntx;
            src.nty = nty;
            src.D = D;
            src.n_G = n_G;
            src.orb = orb;
            src.orb2 = orb2;
            if (direct_store) {
                src.srct = this->src.as<int>();
                src.Sigma = Sigma;
                src.s0 = s0;
            }
            launch_sl_convolve(
                *geo, src, X.as<cd>(), nch, V.as<cd>(), period, flags.as<int>(), orb2, scale,
                rowmask.as<int>() + static_cast<size_t>(s0 / slots_per_batch) * tiles_per_batch * Nx,
                st);
        }
        if (!direct_store || nslots == 0) {
            hipLaunchKernelGGL(k_sl_gather, dim3(grid_for(n_out)), dim3(256), 0, st, n_out, D,
                               orb, orb2, n_G, nkx, nky, Ny, std::max(nch, 1), src.as<SlSrc>(),
                               goff.as<int>(), s0, nslots, s0 == 0 ? 1 : 0, X.as<cd>(), Sigma);
            HFGPU_LAUNCH_CHECK();
        }
        s0 += std::max(nslots, 1);
    } while (s0 < n_slots);
}

void SlHartree::init(const hf_superlattice_desc& d, hipStream_t st) {
    n_G = static_cast<int>(d.n_G);
    orb = static_cast<int>(d.dim_orb);
    nb = n_G * orb;
    deg = d.hartree_degeneracy;
    orbital = d.HH_GG_orbital != nullptr;
    const size_t nG2 = static_cast<size_t>(n_G) * n_G;
    // channels referenced by pair_to_delta, numbered in order of first use,
    // each with its pairs in row-major order
    std::vector<int> used(d.n_delta, -1), cid(nG2);
    std::vector<std::vector<int>> members;
    for (size_t p = 0; p < nG2; ++p) {
        const int64_t dl = d.pair_to_delta[p];
        if (dl < 0 || static_cast<size_t>(dl) >= d.n_delta)
            throw std::invalid_argument("pair_to_delta entries must lie in [0, n_delta)");
        if (used[dl] < 0) {
            used[dl] = static_cast<int>(members.size());
            members.emplace_back();
        }
        cid[p] = used[dl];
        members[used[dl]].push_back(static_cast<int>(p));
    }
    n_used = static_cast<int>(members.size());
    if (static_cast<size_t>(n_used) * (orbital ? orb : 1) * sizeof(cd) > 64 * 1024)
        throw std::invalid_argument("too many superlattice channels for the GPU Hartree shift");
    std::vector<int> st_tab(1, 0), pr_tab;
    for (const auto& m : members) {
        pr_tab.insert(pr_tab.end(), m.begin(), m.end());
        st_tab.push_back(static_cast<int>(pr_tab.size()));
    }
    start.alloc(st_tab.size() * sizeof(int));
    h2d(start.get(), st_tab.data(), st_tab.size() * sizeof(int), st);
    pairs.alloc(pr_tab.size() * sizeof(int));
    h2d(pairs.get(), pr_tab.data(), pr_tab.size() * sizeof(int), st);
    cidx.alloc(nG2 * sizeof(int));
    h2d(cidx.get(), cid.data(), nG2 * sizeof(int), st);
    const size_t nh = orbital ? nG2 * orb * orb : nG2;
    HH.alloc(nh * sizeof(double));
    h2d(HH.get(), orbital ? d.HH_GG_orbital : d.HH_GG, nh * sizeof(double), st);
    HFGPU_CHECK(hipStreamSynchronize(st));
}

void SlHartree::launch(const cd* rho_bar, cd* SC, bool accumulate, hipStream_t st) const {
    const size_t lds = static_cast<size_t>(std::max(n_used, 1)) * (orbital ? orb : 1) * sizeof(cd);
    hipLaunchKernelGGL(k_sl_hartree, dim3(1), dim3(256), lds, st, n_G, orb, nb, n_used,
                       orbital ? 1 : 0, deg, start.as<int>(), pairs.as<int>(), cidx.as<int>(),
                       HH.as<double>(), rho_bar, SC, accumulate ? 1 : 0);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu

using namespace hfgpu;

extern "C" int hfgpu_selfenergy_superlattice_streamed(
    const double* rho, double* sigma, const double* VR_fft, const double* VR_fft_orbital,
    const int64_t* g_a_off, const int64_t* pair_i, const int64_t* pair_j,
    const int64_t* pair_start, size_t n_pairs, size_t n_delta, size_t N_ext_x, size_t N_ext_y,
    size_t nkx, size_t nky, size_t n_G, size_t dim_orb, int hermitian_rho) {
    return guarded([&] {
        if (!rho || !sigma) throw std::invalid_argument("rho and sigma must be non-null");
        SlInput in;
        in.V = reinterpret_cast<const cd*>(VR_fft);
        in.V_orb = reinterpret_cast<const cd*>(VR_fft_orbital);
        in.g_a_off = g_a_off;
        in.pair_i = pair_i;
        in.pair_j = pair_j;
        in.pair_start = pair_start;
        in.n_pairs = n_pairs;
        in.n_delta = n_delta;
        in.Nx = N_ext_x;
        in.Ny = N_ext_y;
        in.nkx = nkx;
        in.nky = nky;
        in.n_G = n_G;
        in.orb = dim_orb;
        in.hermitian = hermitian_rho != 0;
        check_layout(in);  // before any device call: bad layouts fail on any host
        hipStream_t st = stream_acquire();
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)stream_sync_release(s); }
        } guard{st};
        SlPlan plan;
        plan.init(in, st);
        const size_t bytes = nkx * nky * n_G * dim_orb * n_G * dim_orb * sizeof(cd);
        DevBuf dR(bytes), dS(bytes);
        h2d(dR.get(), rho, bytes, st);
        plan.apply(dR.as<cd>(), nullptr, nullptr, dS.as<cd>(), st);
        d2h(sigma, dS.get(), bytes, st);
        HFGPU_CHECK(hipStreamSynchronize(st));
    });
}
