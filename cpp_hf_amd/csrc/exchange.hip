// Exchange self-energy Sigma = -IFFT2[FFT2(P - refP) * VR] on gfx950, and the
// Fock assembly that follows it (reference: selfenergy.hpp:91-171,
// fock.hpp:34-299).
//
// Inputs and outputs keep the reference's C order, (nk1, nk2, nb, nb) with
// the band pair innermost.  The 2-D transform runs as three passes over a
// tile-major workspace T[blk][k1][k2][c] (C channels per tile):
//   A  rows:  gather P - refP for a tile of C channels of one k1 row (each
//             thread owns one channel, so its (a, b) is fixed), FFT along k2
//             in LDS, store the tile as one contiguous block; optionally the
//             weighted row sums feeding the Hartree/contact terms.
//   B  cols:  load one k2 column of T together with the matching tile-major
//             VR block, FFT along k1, multiply by VR, inverse FFT along k1,
//             store T.  The V(q) product is fused between the forward and
//             inverse transforms, so the k-space product never makes its own
//             trip to HBM.  One extra workgroup reduces the row sums to the
//             Hartree diagonal and contact self-energy meanwhile.
//   C  rows:  inverse FFT along k2 of a contiguous T block, scale by
//             -1/(nk1 nk2), and write Sigma (plus F for the fused Fock
//             epilogue, reading herm(h) in the same tile-major order).
// Every operand a pass reads is requested before its transform starts, so a
// workgroup pays one memory round trip per pass.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "dft_reg.h"
#include "exchange.h"

namespace hfgpu {

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------

FftAxis AxisPlan::axis() const {
    FftAxis a{};
    a.n = n;
    a.nstage = static_cast<int>(radices.size());
    for (size_t i = 0; i < radices.size(); ++i) a.radix[i] = radices[i];
    a.tw = d_tw;
    return a;
}

// Synchronous upload on a private non-blocking (pooled) stream, never the
// null stream: a null-stream copy orders against the process's other
// streams, which breaks hipGraph captures running in other host threads.
static void upload_now(void* dst, const void* src, size_t bytes) {
    hipStream_t s = stream_acquire();
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    const hipError_t w = stream_sync_release(s);
    HFGPU_CHECK(e);
    HFGPU_CHECK(w);
}

void AxisPlan::init(int n_) {
    release();
    n = n_;
    radices.clear();
    int m = n;
    while (m % 8 == 0) { radices.push_back(8); m /= 8; }
    while (m % 4 == 0) { radices.push_back(4); m /= 4; }
    while (m % 2 == 0) { radices.push_back(2); m /= 2; }
    for (int f : {3, 5, 7, 11, 13})
        while (m % f == 0) { radices.push_back(f); m /= f; }
    if (m != 1)
        throw std::invalid_argument("k-mesh size " + std::to_string(n) +
                                    " has a prime factor above 13, which the GPU FFT does not support");
    if (radices.size() > static_cast<size_t>(kMaxStages))
        throw std::invalid_argument("k-mesh size too large for the GPU FFT");
    std::vector<cd> tw(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) {
        const long double a = -2.0L * 3.141592653589793238462643383279502884L *
                              static_cast<long double>(k) / static_cast<long double>(n);
        tw[k] = cd{static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a))};
    }
    d_tw = static_cast<cd*>(dev_alloc(sizeof(cd) * tw.size()));
    upload_now(d_tw, tw.data(), sizeof(cd) * tw.size());
}

void AxisPlan::release() {
    if (d_tw) dev_free(d_tw, sizeof(cd) * static_cast<size_t>(n));
    d_tw = nullptr;
}

static bool tile_fits(const AxisPlan& ax, int C) {
    if (static_cast<long>(ax.n) * C > static_cast<long>(kFftThreads) * kTileMaxPerThread)
        return false;
    for (int r : ax.radices) {
        const long need = static_cast<long>(ax.n / r) * C;
        if (need > static_cast<long>(maxb_of(r)) * kFftThreads) return false;
    }
    return true;
}

// complex LDS slots of the small-terms workgroup
static size_t small_terms_cd(int nb) { return 3 * static_cast<size_t>(nb) * nb + 257; }

// Mesh sizes with a two-step register transform, as (N, N1, N2).
#define HFGPU_FAST_SIZES(X)                                                               \
    X(12, 4, 3) X(16, 4, 4) X(24, 4, 6) X(32, 4, 8) X(36, 6, 6) X(48, 8, 6) X(60, 6, 10) \
    X(64, 8, 8) X(72, 8, 9) X(96, 8, 12) X(120, 8, 15) X(144, 12, 12) X(192, 16, 12)

static bool fast_size(int n) {
#define HFGPU_FAST_CASE(N, A, B) if (n == N) return true;
    HFGPU_FAST_SIZES(HFGPU_FAST_CASE)
#undef HFGPU_FAST_CASE
    return false;
}

static int pow2_slots(long elems) {
    int s = 2;  // smallest instantiated slot count
    while (static_cast<long>(s) * kFftThreads < elems) s <<= 1;
    return s;
}

void ExchangeGeom::init(int nk1_, int nk2_, int nb_, bool packed_) {
    release();
    nk1 = nk1_;
    nk2 = nk2_;
    nb = nb_;
    packed = packed_;
    if (nb > 32) throw std::invalid_argument("the GPU Fock build supports nb <= 32");
    std::vector<int> ca, cb;
    for (int a = 0; a < nb; ++a)
        for (int b = packed ? a : 0; b < nb; ++b) {
            ca.push_back(a);
            cb.push_back(b);
        }
    nch = static_cast<int>(ca.size());
    ax1.init(nk1);
    ax2.init(nk2);
    // Channel tile: 8 channels (128 B per k point, one cache line of the
    // workspace) by default; HFGPU_EXCH_TILE overrides it (1, 2, 4, 8 or
    // 16).  Shrunk until every FFT stage fits the per-thread butterfly budget
    // of a 256-thread workgroup.
    int want = 8;
    if (const char* e = std::getenv("HFGPU_EXCH_TILE")) {
        const int v = std::atoi(e);
        if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) want = v;
    }
    int C = 1;
    while (C < nch && C < want) C <<= 1;
    while (C > 1 && !(tile_fits(ax1, C) && tile_fits(ax2, C))) C >>= 1;
    if (!(tile_fits(ax1, C) && tile_fits(ax2, C)))
        throw std::invalid_argument("k-mesh too large for the GPU FFT tile");
    clog = 0;
    while ((1 << clog) < C) ++clog;
    nblk = (nch + C - 1) / C;
    slots_row = pow2_slots(static_cast<long>(nk2) * C);
    slots_col = pow2_slots(static_cast<long>(nk1) * C);
    generic = false;
    for (const AxisPlan* ax : {&ax1, &ax2})
        for (int r : ax->radices)
            if (r == 7 || r == 11 || r == 13) generic = true;
    const char* fe = std::getenv("HFGPU_EXCH_FAST");
    const bool fast_ok = C == 8 && !(fe && fe[0] == '0');
    fast_rows = fast_ok && fast_size(nk2);
    fast_cols = fast_ok && fast_size(nk1);
    // launch shape of the two-step passes (measured best at 96 x 96, nb = 8:
    // one channel tile per row workgroup, two per column workgroup);
    // HFGPU_ROW_WAVES (0: whole rows) and HFGPU_COL_WAVES override it
    row_waves = 1;
    col_waves = 2;
    // wavefronts per channel tile of the two-step passes (measured at 96 x 96
    // and 192 x 192: two help the forward row and the column passes, whose
    // lanes then hold half the registers; one for the inverse rows, whose
    // epilogue stores dominate); HFGPU_SPLIT = 1 or 2 sets all three
    split_fwd = 2;
    split_col = 2;
    split_inv = 1;
    if (const char* e = std::getenv("HFGPU_SPLIT"))
        split_fwd = split_col = split_inv = std::atoi(e) == 2 ? 2 : 1;
    if (const char* e = std::getenv("HFGPU_ROW_WAVES")) row_waves = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("HFGPU_COL_WAVES")) {
        const int v = std::atoi(e);
        if (v == 1 || v == 2 || v == 4) col_waves = v;
    }
    lds_rows = sizeof(cd) * (static_cast<size_t>(nk2) * C + nk2);
    lds_cols = sizeof(cd) * (static_cast<size_t>(nk1) * C + nk1);
    if (lds_rows > 160 * 1024 || lds_cols > 160 * 1024)
        throw std::invalid_argument("k-mesh too large for an LDS-resident FFT line");
    d_chan_a = static_cast<int*>(dev_alloc(sizeof(int) * nch));
    d_chan_b = static_cast<int*>(dev_alloc(sizeof(int) * nch));
    upload_now(d_chan_a, ca.data(), sizeof(int) * nch);
    upload_now(d_chan_b, cb.data(), sizeof(int) * nch);
}

void ExchangeGeom::release() {
    if (d_chan_a) dev_free(d_chan_a, sizeof(int) * nch);
    if (d_chan_b) dev_free(d_chan_b, sizeof(int) * nch);
    d_chan_a = d_chan_b = nullptr;
    ax1.release();
    ax2.release();
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

namespace {

__device__ __forceinline__ void load_twiddles(cd* s_tw, const cd* tw, int n) {
    for (int i = threadIdx.x; i < n; i += kFftThreads) s_tw[i] = tw[i];
}

// Band pair of channel ch: row-major upper triangle (packed) or row-major.
__device__ __forceinline__ void chan_ab(int ch, int nb, int packed, int& a, int& b) {
    if (!packed) {
        a = ch / nb;
        b = ch - a * nb;
        return;
    }
    int r = ch;
    a = 0;
    while (r >= nb - a) {
        r -= nb - a;
        ++a;
    }
    b = a + r;
}

// Workgroup L of a 1-D grid -> (line, channel tile).  Workgroups are dealt
// round-robin to the 8 XCDs (L % 8); when the line count is a multiple of 8,
// all channel tiles of one line go to the same XCD, so the 128 B lines they
// share are fetched into one L2.
__device__ __forceinline__ void tile_of(int L, int nlines, int nblk, int& line, int& blk) {
    if ((nlines & 7) == 0) {
        const int x = L & 7, slot = L >> 3;
        line = x + 8 * (slot / nblk);
        blk = slot % nblk;
    } else {
        line = L / nblk;
        blk = L % nblk;
    }
}

// rho_bar from pass A's row sums, then the Hartree diagonal and the contact
// self-energy; run by one workgroup.
__device__ void small_terms_block(const cd* __restrict__ partial, int nrows, int nch,
                                  const int* __restrict__ chan_a, const int* __restrict__ chan_b,
                                  int nb, int packed, const SmallTerms& st, cd* smem) {
    const int nb2 = nb * nb;
    cd* rho = smem;
    cd* tmp = rho + nb2;
    cd* acc = tmp + nb2;
    cd* grp = acc + nb2;   // <= 256 group partials (G nch <= blockDim.x <= 256)
    cd* trv = grp + 256;
    const int nt = static_cast<int>(blockDim.x);
    const int G = nch >= nt ? 1 : nt / nch;
    // rows split over G groups per channel (loads batched 16 at a time), then
    // the groups summed in order; with one group a thread owns whole channels
    for (int t = threadIdx.x; t < G * nch; t += nt) {
        const int ch = t % nch, g = t / nch;
        cd s{0.0, 0.0};
        for (int r0 = g; r0 < nrows; r0 += 16 * G) {
            cd v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int r = r0 + u * G;
                v[u] = r < nrows ? partial[static_cast<size_t>(r) * nch + ch] : cd{0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) s = cadd(s, v[u]);
        }
        if (G == 1) {
            const int a = chan_a[ch], b = chan_b[ch];
            rho[a * nb + b] = s;
            if (packed && a != b) rho[b * nb + a] = cconj(s);
        } else {
            grp[t] = s;
        }
    }
    __syncthreads();
    if (G > 1) {
        for (int ch = threadIdx.x; ch < nch; ch += nt) {
            cd s{0.0, 0.0};
            for (int g = 0; g < G; ++g) s = cadd(s, grp[g * nch + ch]);
            const int a = chan_a[ch], b = chan_b[ch];
            rho[a * nb + b] = s;
            if (packed && a != b) rho[b * nb + a] = cconj(s);
        }
    }
    for (int e = threadIdx.x; e < nb2; e += nt) acc[e] = cd{0.0, 0.0};
    __syncthreads();
    if (st.rho_out)
        for (int e = threadIdx.x; e < nb2; e += nt) st.rho_out[e] = rho[e];
    if (st.include_hartree) {
        for (int a = threadIdx.x; a < nb; a += nt) {
            double s = 0.0;
            for (int b = 0; b < nb; ++b) s += st.HH[a * nb + b] * rho[b * nb + b].x;
            st.hdiag[a] = s;
        }
    } else {
        for (int a = threadIdx.x; a < nb; a += nt) st.hdiag[a] = 0.0;
    }
    if (st.has_contact) {
        for (int t = 0; t < st.n_contact; ++t) {
            const double g = st.contact_g[t];
            if (g == 0.0) continue;  // uniform across the block
            const cd* oi = st.Oi + static_cast<size_t>(t) * nb2;
            const cd* oj = st.Oj + static_cast<size_t>(t) * nb2;
            if (threadIdx.x == 0) {
                cd tr{0.0, 0.0};
                for (int i = 0; i < nb; ++i)
                    for (int j = 0; j < nb; ++j) tr = cadd(tr, cmul(oj[i * nb + j], rho[j * nb + i]));
                trv[0] = tr;
            }
            for (int e = threadIdx.x; e < nb2; e += nt) {
                const int i = e / nb, j = e % nb;
                cd s{0.0, 0.0};
                for (int k = 0; k < nb; ++k) s = cadd(s, cmul(oi[i * nb + k], rho[k * nb + j]));
                tmp[e] = s;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < nb2; e += nt) {
                const int i = e / nb, j = e % nb;
                cd s{0.0, 0.0};
                for (int k = 0; k < nb; ++k) s = cadd(s, cmul(tmp[i * nb + k], oj[k * nb + j]));
                // g Oi tr(Oj rho) - g Oi rho Oj
                const cd herm = cmul(cmul(oi[e], trv[0]), cd{g, 0.0});
                const cd fk = cmul(cd{-g, 0.0}, s);
                acc[e] = cadd(acc[e], cadd(herm, fk));
            }
            __syncthreads();
        }
        for (int e = threadIdx.x; e < nb2; e += nt) st.SC[e] = acc[e];
    }
}

// Weighted sum over the line of each channel of the tile, by all 256
// threads: TPC = 256 / C threads per channel, fixed reduction tree.  wv holds
// the weights this thread needs (prefetched with the tile).  With a
// k-independent reference density the row sum of P is corrected by
// refP0_ab * sum_k2 w(k1, k2).
template <int S>
__device__ __forceinline__ void tile_row_sums(const cd* s, const double (&wv)[S], int n,
                                              int clog, int c0, int nch, int line, int nb,
                                              int packed, const cd* __restrict__ refP0,
                                              double wrow, cd* __restrict__ partial, cd* red) {
    const int C = 1 << clog;
    const int tpc = kFftThreads >> clog;
    const int c = threadIdx.x / tpc, g = threadIdx.x % tpc;
    cd acc{0.0, 0.0};
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int i = g + r * tpc;
        if (i < n) {
            const cd v = s[(i << clog) + c];
            acc.x += wv[r] * v.x;
            acc.y += wv[r] * v.y;
        }
    }
    const int width = tpc < 64 ? tpc : 64;
    for (int off = width >> 1; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, 64);
        acc.y += __shfl_xor(acc.y, off, 64);
    }
    int cw = -1;  // channel this thread writes, if any
    if (tpc <= 64) {
        if (g == 0) cw = c;
    } else {
        const int waves_per_c = tpc >> 6;
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x < C) {
            acc = cd{0.0, 0.0};
            for (int w = 0; w < waves_per_c; ++w) acc = cadd(acc, red[threadIdx.x * waves_per_c + w]);
            cw = threadIdx.x;
        }
    }
    if (cw >= 0 && c0 + cw < nch) {
        if (refP0) {
            int a, b;
            chan_ab(c0 + cw, nb, packed, a, b);
            const cd r = refP0[a * nb + b];
            acc.x -= r.x * wrow;
            acc.y -= r.y * wrow;
        }
        partial[static_cast<size_t>(line) * nch + c0 + cw] = acc;
    }
}

// Pass A.  1-D grid of nk1 * nblk tiles; S tile elements per thread.
template <bool GENERIC, int S>
__global__ void __launch_bounds__(kFftThreads)
k_rows_fwd(FftAxis ax2, int nk1, int nk2, int nb, int nch, int clog, int nblk, int packed,
           const cd* __restrict__ P, const cd* __restrict__ refP, const cd* __restrict__ refP0,
           const double* __restrict__ wrow, cd* __restrict__ T, cd* __restrict__ partial,
           const double* __restrict__ w2d, int do_fft) {
    extern __shared__ cd smem[];
    __shared__ cd red[kFftThreads / 64];
    const int C = 1 << clog;
    const int tot = nk2 << clog;
    cd* s = smem;
    cd* s_tw = smem + tot;
    int k1, blk;
    tile_of(blockIdx.x, nk1, nblk, k1, blk);
    // thread -> channel is fixed: element e = tid + 256 r has c = tid % C
    const int c = threadIdx.x & (C - 1);
    const int ch = blk * C + c;
    const bool valid = ch < nch;
    int a = 0, b = 0;
    if (valid) chan_ab(ch, nb, packed, a, b);
    const int nb2 = nb * nb;
    const size_t base = static_cast<size_t>(k1) * nk2 * nb2 + a * nb + b;
    if (do_fft) load_twiddles(s_tw, ax2.tw, nk2);
    // weights of this thread's row-sum elements, requested with the tile
    double wv[S];
    if (partial) {
        const int tpc = kFftThreads >> clog, g = threadIdx.x % tpc;
        const double* wline = w2d + static_cast<size_t>(k1) * nk2;
#pragma unroll
        for (int r = 0; r < S; ++r) {
            const int i = g + r * tpc;
            wv[r] = i < nk2 ? wline[i] : 0.0;
        }
    }
    cd v[S];
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        v[r] = cd{0.0, 0.0};
        if (valid && e < tot) {
            const size_t off = base + static_cast<size_t>(e >> clog) * nb2;
            v[r] = P[off];
            if (refP) v[r] = csub(v[r], refP[off]);
        }
    }
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) s[e] = v[r];
    }
    __syncthreads();
    if (partial)
        tile_row_sums<S>(s, wv, nk2, clog, blk * C, nch, k1, nb, packed, refP0,
                         refP0 ? wrow[k1] : 0.0, partial, red);
    if (!do_fft) return;
    __syncthreads();
    fft_tile<false, GENERIC>(s, s_tw, ax2, clog);
    // a k-independent reference density only reaches q2 = 0: nk2 * refP0
    if (refP0 && valid && threadIdx.x < C) {
        const cd r = refP0[a * nb + b];
        s[c] = cd{s[c].x - nk2 * r.x, s[c].y - nk2 * r.y};
    }
    cd* Tt = T + (static_cast<size_t>(blk) * nk1 + k1) * tot;
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) Tt[e] = s[e];
    }
}

// Small terms on their own, for geometries whose pass B cannot host them.
__global__ void __launch_bounds__(kFftThreads)
k_small_terms(const cd* __restrict__ partial, int nrows, int nch, const int* __restrict__ chan_a,
              const int* __restrict__ chan_b, int nb, int packed, SmallTerms st) {
    extern __shared__ cd smem[];
    small_terms_block(partial, nrows, nch, chan_a, chan_b, nb, packed, st, smem);
}

// Pass B.  1-D grid of nk2 * nblk tiles, plus workgroup 0 for the small
// terms when has_small.
template <bool GENERIC, int S>
__global__ void __launch_bounds__(kFftThreads)
k_cols(FftAxis ax1, int nk1, int nk2, int nch, int clog, int nblk, cd* __restrict__ T,
       const cd* __restrict__ VR, int vr_scalar, int has_small, const cd* __restrict__ partial,
       const int* __restrict__ chan_a, const int* __restrict__ chan_b, int nb, int packed,
       SmallTerms st) {
    extern __shared__ cd smem[];
    // the first workgroup reduces pass A's row sums to the Hartree diagonal
    // and contact self-energy, off the transform's critical path
    if (has_small && blockIdx.x == 0) {
        small_terms_block(partial, nk1, nch, chan_a, chan_b, nb, packed, st, smem);
        return;
    }
    const int C = 1 << clog;
    const int tot = nk1 << clog;
    cd* s = smem;
    cd* s_tw = smem + tot;
    int k2, blk;
    tile_of(static_cast<int>(blockIdx.x) - has_small, nk2, nblk, k2, blk);
    const int c = threadIdx.x & (C - 1);
    cd* Tb = T + static_cast<size_t>(blk) * nk1 * nk2 * C;   // tile block of T
    const cd* Vb = VR + (static_cast<size_t>(blk) * nk2 + k2) * tot;  // contiguous VR tile
    load_twiddles(s_tw, ax1.tw, nk1);
    cd v[S], w[S];
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) {
            const int i = e >> clog;
            v[r] = Tb[(static_cast<size_t>(i) * nk2 + k2) * C + c];
            w[r] = vr_scalar ? VR[static_cast<size_t>(i) * nk2 + k2] : Vb[e];
        }
    }
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) s[e] = v[r];
    }
    __syncthreads();
    fft_tile<false, GENERIC>(s, s_tw, ax1, clog);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) s[e] = cmul(s[e], w[r]);
    }
    __syncthreads();
    fft_tile<true, GENERIC>(s, s_tw, ax1, clog);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) Tb[(static_cast<size_t>(e >> clog) * nk2 + k2) * C + c] = s[e];
    }
}

// Pass C.  1-D grid of nk1 * nblk tiles.
template <bool GENERIC, int S>
__global__ void __launch_bounds__(kFftThreads)
k_rows_inv(FftAxis ax2, int nk1, int nk2, int nb, int nch, int clog, int nblk, int packed,
           const cd* __restrict__ T, int epilogue, int shift, double inv_n,
           cd* __restrict__ Sigma, cd* __restrict__ F, const cd* __restrict__ hh,
           const cd* __restrict__ SC, const double* __restrict__ hdiag) {
    extern __shared__ cd smem[];
    const int C = 1 << clog;
    const int tot = nk2 << clog;
    cd* s = smem;
    cd* s_tw = smem + tot;
    int k1, blk;
    tile_of(blockIdx.x, nk1, nblk, k1, blk);
    const int c = threadIdx.x & (C - 1);
    const int ch = blk * C + c;
    const bool valid = ch < nch;
    int a = 0, b = 0;
    if (valid) chan_ab(ch, nb, packed, a, b);
    const size_t toff = (static_cast<size_t>(blk) * nk1 + k1) * tot;
    const bool fock = epilogue == EPI_FOCK_PACKED;
    load_twiddles(s_tw, ax2.tw, nk2);
    cd v[S], hv[S];
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) {
            v[r] = T[toff + e];
            if (fock) hv[r] = hh[toff + e];
        }
    }
    // per-thread constants of the Fock epilogue
    cd sc_ab{0.0, 0.0}, sc_ba{0.0, 0.0};
    double hd = 0.0;
    if (fock && valid) {
        if (SC) {
            sc_ab = SC[a * nb + b];
            sc_ba = SC[b * nb + a];
        }
        if (hdiag && a == b) hd = hdiag[a];
    }
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) s[e] = v[r];
    }
    __syncthreads();
    fft_tile<true, GENERIC>(s, s_tw, ax2, clog);
    if (!valid) return;
    // destination row (ifftshift remaps source (i, j) -> (i - n/2, j - n/2))
    const int d1 = shift ? (k1 + nk1 - nk1 / 2) % nk1 : k1;
    const int nb2 = nb * nb;
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e >= tot) continue;
        const int i = e >> clog;
        const cd y = s[e];
        // reference order: ifftn = backward * (1/N), then negate
        const cd sig{-(y.x * inv_n), -(y.y * inv_n)};
        const int d2 = shift ? (i + nk2 - nk2 / 2) % nk2 : i;
        const size_t kb = static_cast<size_t>(d1 * nk2 + d2) * nb2;
        const size_t iab = kb + a * nb + b, iba = kb + b * nb + a;
        if (epilogue == EPI_RAW_FULL) {
            Sigma[iab] = sig;
        } else if (epilogue == EPI_RAW_PACKED) {
            Sigma[iab] = sig;
            if (a != b) Sigma[iba] = cconj(sig);
        } else {
            // a == b is one element: its mirror is itself, not its conjugate
            const cd s_ab = cadd(sig, sc_ab);
            const cd s_ba = cadd(a == b ? sig : cconj(sig), sc_ba);
            // F = herm(h) + herm(Sigma) + diag(H)
            cd f{hv[r].x + 0.5 * (s_ab.x + s_ba.x), hv[r].y + 0.5 * (s_ab.y - s_ba.y)};
            f.x += hd;
            Sigma[iab] = s_ab;
            F[iab] = f;
            if (a != b) {
                Sigma[iba] = s_ba;
                F[iba] = cconj(f);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Two-step line transforms (N = N1 x N2, both at most 16): one wavefront per
// tile of kFastLines channels.
//
// Forward, x[i1 + N1 i2] -> X[k2 + N2 k1]:
//   F1  task (l, i1): N2-point DFT over i2 in registers, times W_N^(i1 k2),
//       into Y[i1][k2][l] (LDS);
//   F2  task (l, k2): N1-point DFT over i1 -> X[k2 + N2 k1], k1 = 0..N1-1.
// Inverse, Z[k2 + N2 k1] -> z[i1 + N1 i2]:
//   I1  task (l, k2): N1-point inverse DFT over k1, times conj W_N^(i1 k2),
//       into Y[i1][k2][l];
//   I2  task (l, i1): N2-point inverse DFT over k2 -> z[i1 + N1 i2].
// Each transform crosses LDS once; lanes are ordered channel-fastest, so
// every LDS access of eight consecutive lanes covers 128 contiguous bytes and
// every global access of a tile is 128 B per k point.  In pass B the forward
// F2, the V(q) product and the inverse I1 run on the same registers.
// ---------------------------------------------------------------------------

// Workgroup g of a row pass whose k1 rows are each covered by J workgroups
// -> (row, part).  Workgroups are dealt round-robin to the 8 XCDs (g % 8); the
// J parts of a row get ids congruent mod 8, so they run on one XCD: the 128 B
// lines of P, Sigma and F that several parts share meet in one L2.
__device__ __forceinline__ void split_of(int g, int J, int& row, int& part) {
    if (J == 1) {
        row = g;
        part = 0;
        return;
    }
    const int s = g >> 3;
    part = s % J;
    row = (s / J) * 8 + (g & 7);
}

constexpr int kFastLines = 8;  // channels per tile (the geometry's C)
constexpr int kFastWaves = 4;  // most tiles per workgroup of the column pass
// Most tiles per workgroup of a row pass.  Fewer tiles per workgroup spread a
// row over more CUs: the epilogues are FP64/integer VALU work at 4 cycles per
// wavefront instruction, and five tiles on a CU's four SIMDs double the
// critical path of one of them.  The parts of a row stay on one XCD.
constexpr int kFastRowWavesMax = 8;

// A tile's tasks are dealt to S wavefronts (S = 1 or 2): task t of the tile
// runs on wavefront (t / 64) % S, lane t % 64; two wavefronts per tile halve
// the registers of a lane and double the wavefronts of a pass.
template <int N1, int N2, int S = 1>
struct FastShape {
    static constexpr int N = N1 * N2;
    static constexpr int L = kFastLines;
    static constexpr int R1 = (L * N1 + 63) / 64;  // F1 / I2 task rounds per tile
    static constexpr int R2 = (L * N2 + 63) / 64;  // F2 / I1 task rounds per tile
    static constexpr int R1S = (R1 + S - 1) / S;   // ... per wavefront
    static constexpr int R2S = (R2 + S - 1) / S;
    static constexpr int Y = N * L;                // LDS complex slots per tile
};

// Pass A (rows, forward) on the two-step transform.
// TS: storage type of the workspace (cd, or cf for the mixed-precision build).
template <int N1, int N2, class TS, int S>
__global__ void __launch_bounds__(64 * kFastRowWavesMax)
k_rows_fwd_fast(const cd* __restrict__ tw, int nk1, int nb, int nch, int nblk, int packed,
                const cd* __restrict__ P, const cd* __restrict__ refP,
                const cd* __restrict__ refP0, const double* __restrict__ wrow,
                TS* __restrict__ T, cd* __restrict__ partial, const double* __restrict__ w2d,
                int rpw, const ExchBatchEntry* __restrict__ bt, unsigned bt_active) {
    using Sh = FastShape<N1, N2, S>;
    constexpr int N = Sh::N, L = Sh::L;
    if (bt) {  // solve blockIdx.y of a batch (FP64 workspace)
        const ExchBatchEntry& e = bt[blockIdx.y];
        if (!((bt_active >> e.sv) & 1u)) return;
        P = e.P;
        refP = e.refP;
        refP0 = e.refP0;
        wrow = e.wrow;
        T = reinterpret_cast<TS*>(e.T);
        partial = e.partial;
        w2d = e.w2d;
    }
    extern __shared__ cd smem[];
    cd* s_tw = smem;
    const int nw = (blockDim.x >> 6) / S;  // tiles per workgroup
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tix = wave / S, sub = wave % S;
    cd* Y = smem + N + tix * Sh::Y;
    cd* s_part = smem + N + nw * Sh::Y + tix * N1 * L;
    // rpw workgroups per k1 row, nw channel tiles each
    int row, part;
    split_of(static_cast<int>(blockIdx.x), rpw, row, part);
    const bool active = row < nk1 && part * nw + tix < nblk;
    const int k1 = active ? row : 0, blk = active ? part * nw + tix : 0;
    const int nb2 = nb * nb;
    for (int i = threadIdx.x; i < N; i += blockDim.x) s_tw[i] = tw[i];
    cd u[Sh::R1S][N2];
    double wv[Sh::R1S][N2];
#pragma unroll
    for (int r = 0; r < Sh::R1S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        const int l = t & (L - 1), i1 = t >> 3;
        const int ch = blk * L + l;
        const bool ok = active && t < L * N1 && ch < nch;
        int a = 0, b = 0;
        if (ok) chan_ab(ch, nb, packed, a, b);
        const size_t base = static_cast<size_t>(k1) * N * nb2 + a * nb + b;
#pragma unroll
        for (int i2 = 0; i2 < N2; ++i2) {
            const int k2 = i1 + N1 * i2;
            cd x{0.0, 0.0};
            if (ok) {
                x = P[base + static_cast<size_t>(k2) * nb2];
                if (refP) x = csub(x, refP[base + static_cast<size_t>(k2) * nb2]);
            }
            u[r][i2] = x;
            wv[r][i2] = (ok && partial) ? w2d[static_cast<size_t>(k1) * N + k2] : 0.0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < Sh::R1S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        if (t < L * N1) {
            const int l = t & (L - 1), i1 = t >> 3;
            if (partial) {
                cd acc{0.0, 0.0};
#pragma unroll
                for (int i2 = 0; i2 < N2; ++i2) {
                    acc.x += wv[r][i2] * u[r][i2].x;
                    acc.y += wv[r][i2] * u[r][i2].y;
                }
                s_part[i1 * L + l] = acc;
            }
            Dft<N2, false>::run(u[r]);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                cd x = u[r][k2];
                if (k2 > 0) x = cmul(x, s_tw[i1 * k2]);
                Y[(i1 * N2 + k2) * L + l] = x;
            }
        }
    }
    __syncthreads();
    if (partial && active && sub == 0 && lane < L) {
        const int ch = blk * L + lane;
        if (ch < nch) {
            cd acc{0.0, 0.0};
            for (int i1 = 0; i1 < N1; ++i1) acc = cadd(acc, s_part[i1 * L + lane]);
            if (refP0) {
                int a, b;
                chan_ab(ch, nb, packed, a, b);
                const cd rv = refP0[a * nb + b];
                const double wr = wrow[k1];
                acc.x -= rv.x * wr;
                acc.y -= rv.y * wr;
            }
            partial[static_cast<size_t>(k1) * nch + ch] = acc;
        }
    }
    TS* Tt = T + (static_cast<size_t>(blk) * nk1 + k1) * N * L;
#pragma unroll
    for (int r = 0; r < Sh::R2S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        if (active && t < L * N2) {
            const int l = t & (L - 1), k2 = t >> 3;
            cd v[N1];
#pragma unroll
            for (int i1 = 0; i1 < N1; ++i1) v[i1] = Y[(i1 * N2 + k2) * L + l];
            Dft<N1, false>::run(v);
            // a k-independent reference density only reaches q2 = 0: N refP0
            if (refP0 && k2 == 0 && blk * L + l < nch) {
                int a, b;
                chan_ab(blk * L + l, nb, packed, a, b);
                const cd rv = refP0[a * nb + b];
                v[0] = cd{v[0].x - N * rv.x, v[0].y - N * rv.y};
            }
#pragma unroll
            for (int k1p = 0; k1p < N1; ++k1p) store_stream(Tt + (k2 + N2 * k1p) * L + l, v[k1p]);
        }
    }
}

// Pass B (columns: forward, x VR, inverse) on the two-step transform;
// workgroup 0 hosts the small terms when has_small.
template <int N1, int N2, class TS, int S>
__global__ void __launch_bounds__(64 * kFastWaves)
k_cols_fast(const cd* __restrict__ tw, int nk2, int nch, int nblk, TS* __restrict__ T,
            const TS* __restrict__ VR, int vr_scalar, int has_small,
            const cd* __restrict__ partial, const int* __restrict__ chan_a,
            const int* __restrict__ chan_b, int nb, int packed, SmallTerms st,
            const ExchBatchEntry* __restrict__ bt, unsigned bt_active) {
    using Sh = FastShape<N1, N2, S>;
    constexpr int N = Sh::N, L = Sh::L;
    if (bt) {  // solve blockIdx.y of a batch (FP64 workspace)
        const ExchBatchEntry& e = bt[blockIdx.y];
        if (!((bt_active >> e.sv) & 1u)) return;
        T = reinterpret_cast<TS*>(e.T);
        VR = reinterpret_cast<const TS*>(e.VR);
        partial = e.partial;
        st = e.small;
    }
    extern __shared__ cd smem[];
    if (has_small && blockIdx.x == 0) {
        small_terms_block(partial, N, nch, chan_a, chan_b, nb, packed, st, smem);
        return;
    }
    cd* s_tw = smem;
    const int nw = (blockDim.x >> 6) / S;  // tiles per workgroup
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tix = wave / S, sub = wave % S;
    cd* Y = smem + N + tix * Sh::Y;
    // the tiles of a workgroup are adjacent columns of one channel tile, so
    // each k1 row of the workspace is read and written as nw * 128
    // contiguous bytes
    const int tile = (static_cast<int>(blockIdx.x) - has_small) * nw + tix;
    const bool active = tile < nk2 * nblk;
    const int kc = active ? tile % nk2 : 0, blk = active ? tile / nk2 : 0;
    TS* Tb = T + static_cast<size_t>(blk) * N * nk2 * L;                 // [k1][k2][c]
    const TS* Vb = VR + (static_cast<size_t>(blk) * nk2 + kc) * N * L;   // [q1][c]
    for (int i = threadIdx.x; i < N; i += blockDim.x) s_tw[i] = tw[i];
    // loads land in registers of the storage type and are widened after the
    // barrier, so every load of the wave is in flight before the first use
    TS u[Sh::R1S][N2];
    TS vv[Sh::R2S][N1];
#pragma unroll
    for (int r = 0; r < Sh::R1S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        const int l = t & (L - 1), i1 = t >> 3;
        const bool ok = active && t < L * N1;
#pragma unroll
        for (int i2 = 0; i2 < N2; ++i2)
            u[r][i2] = ok ? Tb[(static_cast<size_t>(i1 + N1 * i2) * nk2 + kc) * L + l] : TS{};
    }
#pragma unroll
    for (int r = 0; r < Sh::R2S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        const int l = t & (L - 1), k2 = t >> 3;
        const bool ok = active && t < L * N2;
#pragma unroll
        for (int k1p = 0; k1p < N1; ++k1p) {
            const int q = k2 + N2 * k1p;
            vv[r][k1p] = !ok ? TS{}
                             : (vr_scalar ? VR[static_cast<size_t>(q) * nk2 + kc] : Vb[q * L + l]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < Sh::R1S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        if (t < L * N1) {
            const int l = t & (L - 1), i1 = t >> 3;
            cd w[N2];
#pragma unroll
            for (int i2 = 0; i2 < N2; ++i2) w[i2] = to_cd(u[r][i2]);
            Dft<N2, false>::run(w);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                cd x = w[k2];
                if (k2 > 0) x = cmul(x, s_tw[i1 * k2]);
                Y[(i1 * N2 + k2) * L + l] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < Sh::R2S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        if (t < L * N2) {
            const int l = t & (L - 1), k2 = t >> 3;
            cd v[N1];
#pragma unroll
            for (int i1 = 0; i1 < N1; ++i1) v[i1] = Y[(i1 * N2 + k2) * L + l];
            Dft<N1, false>::run(v);
#pragma unroll
            for (int k1p = 0; k1p < N1; ++k1p) v[k1p] = cmul(v[k1p], to_cd(vv[r][k1p]));
            Dft<N1, true>::run(v);
            // this task reads and rewrites the same Y slots: no barrier needed
#pragma unroll
            for (int i1 = 0; i1 < N1; ++i1) {
                cd x = v[i1];
                if (i1 > 0 && k2 > 0) x = cmulc(x, s_tw[i1 * k2]);
                Y[(i1 * N2 + k2) * L + l] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < Sh::R1S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        if (active && t < L * N1) {
            const int l = t & (L - 1), i1 = t >> 3;
            cd z[N2];
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) z[k2] = Y[(i1 * N2 + k2) * L + l];
            Dft<N2, true>::run(z);
#pragma unroll
            for (int i2 = 0; i2 < N2; ++i2)
                store_col_ws(Tb + (static_cast<size_t>(i1 + N1 * i2) * nk2 + kc) * L + l, z[i2]);
        }
    }
}

// Pass C (rows, inverse + epilogue) on the two-step transform.
template <int N1, int N2, class TS, int S>
__global__ void __launch_bounds__(64 * kFastRowWavesMax)
k_rows_inv_fast(const cd* __restrict__ tw, int nk1, int nb, int nch, int nblk, int packed,
                const TS* __restrict__ T, int epilogue, int shift, double inv_n,
                cd* __restrict__ Sigma, cd* __restrict__ F, const cd* __restrict__ hh,
                const cd* __restrict__ SC, const double* __restrict__ hdiag, int rpw,
                const ExchBatchEntry* __restrict__ bt, unsigned bt_active) {
    using Sh = FastShape<N1, N2, S>;
    constexpr int N = Sh::N, L = Sh::L;
    if (bt) {  // solve blockIdx.y of a batch (FP64 workspace)
        const ExchBatchEntry& e = bt[blockIdx.y];
        if (!((bt_active >> e.sv) & 1u)) return;
        T = reinterpret_cast<const TS*>(e.T);
        Sigma = e.Sigma;
        F = e.F;
        hh = e.hh;
        SC = e.SC;
        hdiag = e.hdiag;
    }
    extern __shared__ cd smem[];
    cd* s_tw = smem;
    const int nw = (blockDim.x >> 6) / S;  // tiles per workgroup
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tix = wave / S, sub = wave % S;
    cd* Y = smem + N + tix * Sh::Y;
    // rpw workgroups per k1 row, nw channel tiles each
    int row, part;
    split_of(static_cast<int>(blockIdx.x), rpw, row, part);
    const bool active = row < nk1 && part * nw + tix < nblk;
    const int k1 = active ? row : 0, blk = active ? part * nw + tix : 0;
    const size_t toff = (static_cast<size_t>(blk) * nk1 + k1) * N * L;
    const bool fock = epilogue == EPI_FOCK_PACKED;
    for (int i = threadIdx.x; i < N; i += blockDim.x) s_tw[i] = tw[i];
    TS v[Sh::R2S][N1];  // storage type, widened after the barrier
#pragma unroll
    for (int r = 0; r < Sh::R2S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        const int l = t & (L - 1), k2 = t >> 3;
        const bool ok = active && t < L * N2;
#pragma unroll
        for (int k1p = 0; k1p < N1; ++k1p)
            v[r][k1p] = ok ? T[toff + (k2 + N2 * k1p) * L + l] : TS{};
    }
    cd hv[Sh::R1S][N2];
    if (fock) {
#pragma unroll
        for (int r = 0; r < Sh::R1S; ++r) {
            const int t = lane + 64 * (r * S + sub);
            const int l = t & (L - 1), i1 = t >> 3;
            const bool ok = active && t < L * N1;
#pragma unroll
            for (int i2 = 0; i2 < N2; ++i2)
                hv[r][i2] = ok ? hh[toff + (i1 + N1 * i2) * L + l] : cd{0.0, 0.0};
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < Sh::R2S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        if (t < L * N2) {
            const int l = t & (L - 1), k2 = t >> 3;
            cd w[N1];
#pragma unroll
            for (int k1p = 0; k1p < N1; ++k1p) w[k1p] = to_cd(v[r][k1p]);
            Dft<N1, true>::run(w);
#pragma unroll
            for (int i1 = 0; i1 < N1; ++i1) {
                cd x = w[i1];
                if (i1 > 0 && k2 > 0) x = cmulc(x, s_tw[i1 * k2]);
                Y[(i1 * N2 + k2) * L + l] = x;
            }
        }
    }
    __syncthreads();
    const int d1 = shift ? (k1 + nk1 - nk1 / 2) % nk1 : k1;
    const int nb2 = nb * nb;
#pragma unroll
    for (int r = 0; r < Sh::R1S; ++r) {
        const int t = lane + 64 * (r * S + sub);
        const int l = t & (L - 1), i1 = t >> 3;
        const int ch = blk * L + l;
        if (!(active && t < L * N1 && ch < nch)) continue;
        int a, b;
        chan_ab(ch, nb, packed, a, b);
        cd sc_ab{0.0, 0.0}, sc_ba{0.0, 0.0};
        double hd = 0.0;
        if (fock) {
            if (SC) {
                sc_ab = SC[a * nb + b];
                sc_ba = SC[b * nb + a];
            }
            if (hdiag && a == b) hd = hdiag[a];
        }
        cd z[N2];
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) z[k2] = Y[(i1 * N2 + k2) * L + l];
        Dft<N2, true>::run(z);
#pragma unroll
        for (int i2 = 0; i2 < N2; ++i2) {
            const int k2 = i1 + N1 * i2;
            const cd y = z[i2];
            // reference order: ifftn = backward * (1/N), then negate
            const cd sig{-(y.x * inv_n), -(y.y * inv_n)};
            const int d2 = shift ? (k2 + N - N / 2) % N : k2;
            const size_t kb = static_cast<size_t>(d1 * N + d2) * nb2;
            const size_t iab = kb + a * nb + b, iba = kb + b * nb + a;
            if (epilogue == EPI_RAW_FULL) {
                Sigma[iab] = sig;
            } else if (epilogue == EPI_RAW_PACKED) {
                Sigma[iab] = sig;
                if (a != b) Sigma[iba] = cconj(sig);
            } else {
                const cd s_ab = cadd(sig, sc_ab);
                const cd s_ba = cadd(a == b ? sig : cconj(sig), sc_ba);
                cd f{hv[r][i2].x + 0.5 * (s_ab.x + s_ba.x), hv[r][i2].y + 0.5 * (s_ab.y - s_ba.y)};
                f.x += hd;
                // regular stores: the mirrored entries are single 16 B
                // pieces of 128 B lines, which streaming stores handle badly
                Sigma[iab] = s_ab;
                F[iab] = f;
                if (a != b) {
                    Sigma[iba] = s_ba;
                    F[iba] = cconj(f);
                }
            }
        }
    }
}

// Tile-major operand packing (run once per kernel geometry).
__global__ void k_pack_vr(int nk1, int nk2, int nb, int nch, int clog, int nblk, int packed,
                          const cd* __restrict__ VR, cd* __restrict__ VRb) {
    const int C = 1 << clog;
    const size_t total = static_cast<size_t>(nblk) * C * nk1 * nk2;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        // t = ((blk * nk2 + k2) * nk1 + k1) * C + c
        const int c = static_cast<int>(t & (C - 1));
        size_t r = t >> clog;
        const int k1 = static_cast<int>(r % nk1);
        r /= nk1;
        const int k2 = static_cast<int>(r % nk2);
        const int blk = static_cast<int>(r / nk2);
        const int ch = blk * C + c;
        cd v{0.0, 0.0};
        if (ch < nch) {
            int a, b;
            chan_ab(ch, nb, packed, a, b);
            v = VR[(static_cast<size_t>(k1) * nk2 + k2) * nb * nb + a * nb + b];
        }
        VRb[t] = v;
    }
}

__global__ void k_pack_herm_h(int nk1, int nk2, int nb, int nch, int clog, int nblk,
                              int packed, const cd* __restrict__ h, cd* __restrict__ hhb) {
    const int C = 1 << clog;
    const size_t total = static_cast<size_t>(nblk) * C * nk1 * nk2;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        // t = ((blk * nk1 + k1) * nk2 + k2) * C + c
        const int c = static_cast<int>(t & (C - 1));
        size_t r = t >> clog;
        const int k2 = static_cast<int>(r % nk2);
        r /= nk2;
        const int k1 = static_cast<int>(r % nk1);
        const int blk = static_cast<int>(r / nk1);
        const int ch = blk * C + c;
        cd v{0.0, 0.0};
        if (ch < nch) {
            int a, b;
            chan_ab(ch, nb, packed, a, b);
            const size_t kb = (static_cast<size_t>(k1) * nk2 + k2) * nb * nb;
            const cd x = h[kb + a * nb + b], y = h[kb + b * nb + a];
            v = cd{0.5 * (x.x + y.x), 0.5 * (x.y - y.y)};
        }
        hhb[t] = v;
    }
}

// grid-stride over (k, a <= b) pairs.
__global__ void k_fock_assemble(int nk, int nb, const cd* __restrict__ h,
                                cd* __restrict__ Sigma, cd* __restrict__ F,
                                const cd* __restrict__ SC, const double* __restrict__ hdiag) {
    const int npair = nb * (nb + 1) / 2;
    const size_t total = static_cast<size_t>(nk) * npair;
    const int nb2 = nb * nb;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < total;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t k = t / npair;
        int r = static_cast<int>(t % npair);
        int a = 0;
        while (r >= nb - a) { r -= nb - a; ++a; }
        const int b = a + r;
        const size_t kb = k * nb2;
        const size_t iab = kb + a * nb + b, iba = kb + b * nb + a;
        cd s_ab = Sigma[iab], s_ba = Sigma[iba];
        if (SC) {
            s_ab = cadd(s_ab, SC[a * nb + b]);
            s_ba = cadd(s_ba, SC[b * nb + a]);
            Sigma[iab] = s_ab;
            if (a != b) Sigma[iba] = s_ba;
        }
        cd x_ab = cadd(h[iab], s_ab), x_ba = cadd(h[iba], s_ba);
        if (hdiag && a == b) {
            x_ab.x += hdiag[a];
            x_ba.x += hdiag[a];
        }
        const cd f{0.5 * (x_ab.x + x_ba.x), 0.5 * (x_ab.y - x_ba.y)};
        F[iab] = f;
        if (a != b) F[iba] = cconj(f);
    }
}

// One wave per k point, four per block; partial[block] = sum over its k.
// Hm (nullable): full Hartree matrix with k stride hk_stride (0: broadcast).
__device__ __forceinline__ void energy_partial_body(
    int nk, int nb, const cd* __restrict__ h, const cd* __restrict__ P,
    const cd* __restrict__ refP, const cd* __restrict__ Sigma, const double* __restrict__ hdiag,
    const cd* __restrict__ Hm, size_t hk_stride, const double* __restrict__ w2d,
    double* __restrict__ partial, double* red) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * 4 + wave;
    const int nb2 = nb * nb;
    double s = 0.0;
    if (k < nk) {
        const size_t kb = static_cast<size_t>(k) * nb2;
        for (int e = lane; e < nb2; e += 64) {
            const int i = e / nb, j = e % nb;
            const cd p = P[kb + j * nb + i];
            cd dp = p;
            if (refP) dp = csub(dp, refP[kb + j * nb + i]);
            cd inter;
            if (Hm) {
                inter = cscale(cadd(Sigma[kb + e], Hm[k * hk_stride + e]), 0.5);
            } else {
                inter = cscale(Sigma[kb + e], 0.5);
                if (hdiag && i == j) inter.x += 0.5 * hdiag[i];
            }
            s += cmul(h[kb + e], p).x;
            s += cmul(inter, dp).x;
        }
        s = wave_sum(s) * w2d[k];
    }
    if (lane == 0) red[wave] = (k < nk) ? s : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_energy_partial(int nk, int nb, const cd* __restrict__ h,
                                 const cd* __restrict__ P, const cd* __restrict__ refP,
                                 const cd* __restrict__ Sigma, const double* __restrict__ hdiag,
                                 const cd* __restrict__ Hm, size_t hk_stride,
                                 const double* __restrict__ w2d, double* __restrict__ partial) {
    __shared__ double red[4];
    energy_partial_body(nk, nb, h, P, refP, Sigma, hdiag, Hm, hk_stride, w2d, partial, red);
}

__device__ __forceinline__ void sum_partials_body(const double* __restrict__ partial, int n,
                                                  double* out, double* red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* out) {
    __shared__ double red[4];
    sum_partials_body(partial, n, out, red);
}

// Energies of a batch: solve blockIdx.y (first launch), blockIdx.x (second),
// each with the decomposition of its single launch_energy.
__global__ void k_energy_partial_batch(const EnergyBatchEntry* __restrict__ tab, int nb,
                                       const unsigned active) {
    __shared__ double red[4];
    if (!((active >> blockIdx.y) & 1u)) return;
    const EnergyBatchEntry& e = tab[blockIdx.y];
    if (static_cast<int>(blockIdx.x) >= (e.nk + 3) / 4) return;
    energy_partial_body(e.nk, nb, e.h, e.P, e.refP, e.Sigma, e.hdiag, nullptr, 0, e.w2d,
                        e.partial, red);
}

__global__ void k_sum_partials_batch(const EnergyBatchEntry* __restrict__ tab,
                                     const unsigned active) {
    __shared__ double red[4];
    if (!((active >> blockIdx.x) & 1u)) return;
    const EnergyBatchEntry& e = tab[blockIdx.x];
    sum_partials_body(e.partial, (e.nk + 3) / 4, e.E, red);
}


// ---------------------------------------------------------------------------
// Plain batched 2-D transforms and spectral products (the adjoint of the
// exchange).  Arrays in natural layout X[k1][k2][ch], nch channels
// innermost, processed in tiles of C = 2^clog channels per workgroup;
// transforms unnormalised (forward exp(-2 pi i / n)).
// ---------------------------------------------------------------------------

// Along k2 (rows: one k1 line and one channel tile per workgroup).  With a
// row mask, lines whose rowmask[blk * nk1 + k1] is 0 are left untouched.
template <bool GENERIC, int S>
__global__ void __launch_bounds__(kFftThreads)
k_plain_rows(FftAxis ax2, int nk1, int nk2, int nch, int clog, int nblk, int inv,
             const cd* __restrict__ in, cd* __restrict__ out, const int* __restrict__ rowmask) {
    extern __shared__ cd smem[];
    const int C = 1 << clog;
    const int tot = nk2 << clog;
    cd* s = smem;
    cd* s_tw = smem + tot;
    int k1, blk;
    tile_of(static_cast<int>(blockIdx.x), nk1, nblk, k1, blk);
    if (rowmask && rowmask[static_cast<size_t>(blk) * nk1 + k1] == 0) return;  // uniform
    const int ch = blk * C + (threadIdx.x & (C - 1));
    const bool valid = ch < nch;
    const size_t base = static_cast<size_t>(k1) * nk2 * nch + ch;
    load_twiddles(s_tw, ax2.tw, nk2);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot)
            s[e] = valid ? in[base + static_cast<size_t>(e >> clog) * nch] : cd{0.0, 0.0};
    }
    __syncthreads();
    if (inv)
        fft_tile<true, GENERIC>(s, s_tw, ax2, clog);
    else
        fft_tile<false, GENERIC>(s, s_tw, ax2, clog);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot && valid) out[base + static_cast<size_t>(e >> clog) * nch] = s[e];
    }
}

// Along k1 (columns), in place.
template <bool GENERIC, int S>
__global__ void __launch_bounds__(kFftThreads)
k_plain_cols(FftAxis ax1, int nk1, int nk2, int nch, int clog, int nblk, int inv,
             cd* __restrict__ X) {
    extern __shared__ cd smem[];
    const int C = 1 << clog;
    const int tot = nk1 << clog;
    cd* s = smem;
    cd* s_tw = smem + tot;
    int k2, blk;
    tile_of(static_cast<int>(blockIdx.x), nk2, nblk, k2, blk);
    const int ch = blk * C + (threadIdx.x & (C - 1));
    const bool valid = ch < nch;
    const size_t base = static_cast<size_t>(k2) * nch + ch;
    const size_t stride = static_cast<size_t>(nk2) * nch;
    load_twiddles(s_tw, ax1.tw, nk1);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) s[e] = valid ? X[base + (e >> clog) * stride] : cd{0.0, 0.0};
    }
    __syncthreads();
    if (inv)
        fft_tile<true, GENERIC>(s, s_tw, ax1, clog);
    else
        fft_tile<false, GENERIC>(s, s_tw, ax1, clog);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot && valid) X[base + (e >> clog) * stride] = s[e];
    }
}

// Superlattice forward rows with the scatter folded in (SlRowsSrc): row
// k1 = px of channel tile blk, only when rowmask marks it.
template <bool GENERIC, int S>
__global__ void __launch_bounds__(kFftThreads)
k_sl_rows_fwd(FftAxis ax2, int nk1, int nk2, int nch, int clog, int nblk, SlRowsSrc src,
              const int* __restrict__ rowmask, cd* __restrict__ X) {
    extern __shared__ cd smem[];
    const int C = 1 << clog;
    const int tot = nk2 << clog;
    cd* s = smem;
    cd* s_tw = smem + tot;
    int k1, blk;
    tile_of(static_cast<int>(blockIdx.x), nk1, nblk, k1, blk);
    if (rowmask[static_cast<size_t>(blk) * nk1 + k1] == 0) return;  // uniform
    const int ch = blk * C + (threadIdx.x & (C - 1));
    const bool valid = ch < nch;
    const int ls = valid ? ch / src.orb2 : 0, ab = valid ? ch % src.orb2 : 0;
    const int a = ab / src.orb, b = ab % src.orb;
    const int tx = k1 / src.nkx, kx = k1 - tx * src.nkx;
    const size_t base = static_cast<size_t>(k1) * nk2 * nch + ch;
    const size_t DD = static_cast<size_t>(src.D) * src.D;
    load_twiddles(s_tw, ax2.tw, nk2);
    bool nz = false;
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e >= tot) continue;
        cd v{0.0, 0.0};
        if (valid) {
            const int py = e >> clog;
            const int ty = py / src.nky, ky = py - ty * src.nky;
            const int pair = src.tiles[(static_cast<size_t>(ls) * src.ntx + tx) * src.nty + ty];
            if (pair >= 0) {
                const int i = pair / src.n_G, j = pair % src.n_G;
                const size_t rc = static_cast<size_t>(i * src.orb + a) * src.D + j * src.orb + b;
                const size_t idx = (static_cast<size_t>(kx) * src.nky + ky) * DD + rc;
                v = src.P[idx];
                if (src.refP)
                    v = csub(v, src.refP[idx]);
                else if (src.refP0)
                    v = csub(v, src.refP0[rc]);
                nz = nz || v.x != 0.0 || v.y != 0.0;
            }
        }
        s[e] = v;
    }
    if (nz) src.flags[ls] = 1;  // every writer stores the same value
    __syncthreads();
    fft_tile<false, GENERIC>(s, s_tw, ax2, clog);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot && valid) X[base + static_cast<size_t>(e >> clog) * nch] = s[e];
    }
}

// Superlattice inverse rows with the gather folded in: after the transform,
// the value of pair (i, j) at (k, a, b) goes to Sigma[k][(i a)][(j b)] and
// its conjugate to Sigma[k][(j b)][(i a)] where the source table names this
// slot and G_i's tile for that block (direct and partner fill).
template <bool GENERIC, int S>
__global__ void __launch_bounds__(kFftThreads)
k_sl_rows_inv(FftAxis ax2, int nk1, int nk2, int nch, int clog, int nblk, SlRowsSrc src,
              const int* __restrict__ rowmask, const cd* __restrict__ X) {
    extern __shared__ cd smem[];
    const int C = 1 << clog;
    const int tot = nk2 << clog;
    cd* s = smem;
    cd* s_tw = smem + tot;
    int k1, blk;
    tile_of(static_cast<int>(blockIdx.x), nk1, nblk, k1, blk);
    if (rowmask[static_cast<size_t>(blk) * nk1 + k1] == 0) return;  // uniform
    const int ch = blk * C + (threadIdx.x & (C - 1));
    const bool valid = ch < nch;
    const int ls = valid ? ch / src.orb2 : 0, ab = valid ? ch % src.orb2 : 0;
    const int a = ab / src.orb, b = ab % src.orb;
    const int tx = k1 / src.nkx, kx = k1 - tx * src.nkx;
    const size_t base = static_cast<size_t>(k1) * nk2 * nch + ch;
    const size_t DD = static_cast<size_t>(src.D) * src.D;
    const int slot = src.s0 + ls;
    load_twiddles(s_tw, ax2.tw, nk2);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) s[e] = valid ? X[base + static_cast<size_t>(e >> clog) * nch] : cd{0.0, 0.0};
    }
    __syncthreads();
    fft_tile<true, GENERIC>(s, s_tw, ax2, clog);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e >= tot || !valid) continue;
        const int py = e >> clog;
        const int ty = py / src.nky, ky = py - ty * src.nky;
        const int pair = src.tiles[(static_cast<size_t>(ls) * src.ntx + tx) * src.nty + ty];
        if (pair < 0) continue;
        const int i = pair / src.n_G, j = pair % src.n_G;
        const cd v = s[e];
        const size_t kb = (static_cast<size_t>(kx) * src.nky + ky) * DD;
        const int* d = src.srct + 4 * (static_cast<size_t>(i) * src.n_G + j);
        if (d[0] == slot && d[1] == i && d[2] == 0)
            src.Sigma[kb + static_cast<size_t>(i * src.orb + a) * src.D + j * src.orb + b] = v;
        const int* c = src.srct + 4 * (static_cast<size_t>(j) * src.n_G + i);
        if (c[0] == slot && c[1] == i && c[2] == 1)
            src.Sigma[kb + static_cast<size_t>(j * src.orb + b) * src.D + i * src.orb + a] =
                cconj(v);
    }
}

// Along k1 (columns), in place: forward transform, times scale V, inverse
// transform (the convolution's column half; one HBM read and write).  V is
// V[q * period + ch % period] at q = q1 nk2 + k2; channels whose flag
// (flags[ch / per_slot]) is 0 come out exactly zero.
template <bool GENERIC, int S>
__global__ void __launch_bounds__(kFftThreads)
k_plain_cols_conv(FftAxis ax1, int nk1, int nk2, int nch, int clog, int nblk,
                  cd* __restrict__ X, const cd* __restrict__ V, int period,
                  const int* __restrict__ flags, int per_slot, double scale,
                  const int* __restrict__ rowmask) {
    extern __shared__ cd smem[];
    const int C = 1 << clog;
    const int tot = nk1 << clog;
    cd* s = smem;
    cd* s_tw = smem + tot;
    int k2, blk;
    tile_of(static_cast<int>(blockIdx.x), nk2, nblk, k2, blk);
    const int ch = blk * C + (threadIdx.x & (C - 1));
    const bool valid = ch < nch;
    const bool live = valid && flags[valid ? ch / per_slot : 0] != 0;
    const size_t base = static_cast<size_t>(k2) * nch + ch;
    const size_t stride = static_cast<size_t>(nk2) * nch;
    load_twiddles(s_tw, ax1.tw, nk1);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) {
            const int q1 = e >> clog;
            const bool has = valid && (!rowmask || rowmask[static_cast<size_t>(blk) * nk1 + q1]);
            s[e] = has ? X[base + q1 * stride] : cd{0.0, 0.0};
        }
    }
    __syncthreads();
    fft_tile<false, GENERIC>(s, s_tw, ax1, clog);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot) {
            const size_t q = static_cast<size_t>(e >> clog) * nk2 + k2;
            s[e] = live ? cscale(cmul(s[e], V[q * period + (valid ? ch % period : 0)]), scale)
                        : cd{0.0, 0.0};
        }
    }
    __syncthreads();
    fft_tile<true, GENERIC>(s, s_tw, ax1, clog);
#pragma unroll
    for (int r = 0; r < S; ++r) {
        const int e = threadIdx.x + r * kFftThreads;
        if (e < tot && valid) X[base + (e >> clog) * stride] = s[e];
    }
}

// out[q][ch] = scale A[q][ch] conj(B[q][ch]), or with sum_ch the sum over
// channels into out[q] (channel order fixed: deterministic)
__global__ void __launch_bounds__(256)
k_cross_spectrum(size_t nq, int nch, const cd* __restrict__ A, const cd* __restrict__ B,
                 double scale, int sum_ch, cd* __restrict__ out) {
    const size_t n = sum_ch ? nq : nq * nch;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        if (sum_ch) {
            cd s{0.0, 0.0};
            for (int c = 0; c < nch; ++c) {
                const size_t j = i * nch + c;
                s = cadd(s, cmulc(A[j], B[j]));
            }
            out[i] = cscale(s, scale);
        } else {
            out[i] = cscale(cmulc(A[i], B[i]), scale);
        }
    }
}

// X[q][ch] *= scale op(V[q][ch]) (or V[q] when v_scalar); op = conj with conj_v
__global__ void __launch_bounds__(256)
k_mul_spectrum(size_t nq, int nch, cd* __restrict__ X, const cd* __restrict__ V, int v_scalar,
               int conj_v, double scale) {
    const size_t n = nq * nch;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        cd v = v_scalar ? V[i / nch] : V[i];
        if (conj_v) v.y = -v.y;
        X[i] = cscale(cmul(X[i], v), scale);
    }
}

// X[q1][q2][ch] *= conj(exp(2 pi i (q1 s1 / n1 + q2 s2 / n2))), s = n / 2: the
// adjoint of the ifftshift phase folded into VR
__global__ void __launch_bounds__(256)
k_unphase(int nk1, int nk2, int nch, cd* __restrict__ X) {
    const size_t n = static_cast<size_t>(nk1) * nk2 * nch;
    const int s1 = nk1 / 2, s2 = nk2 / 2;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t q = i / nch;
        const int q1 = static_cast<int>(q / nk2), q2 = static_cast<int>(q % nk2);
        double s_a, c_a, s_b, c_b;
        sincospi(2.0 * static_cast<double>((static_cast<long>(q1) * s1) % nk1) / nk1, &s_a, &c_a);
        sincospi(2.0 * static_cast<double>((static_cast<long>(q2) * s2) % nk2) / nk2, &s_b, &c_b);
        const cd ph = cmul(cd{c_a, s_a}, cd{c_b, s_b});
        X[i] = cmulc(X[i], ph);
    }
}

// Deterministic sum over k of X[k][e] (m entries per k), weighted by w[k]
// when w is given: per-workgroup partials over k chunks, then one
// workgroup adds them in order.
__global__ void __launch_bounds__(256)
k_ksum_partial(int nk, int m, const cd* __restrict__ X, const double* __restrict__ w,
               cd* __restrict__ partial) {
    const int chunk = (nk + gridDim.x - 1) / gridDim.x;
    const int k0 = blockIdx.x * chunk, k1 = min(nk, k0 + chunk);
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
        cd s{0.0, 0.0};
        // loads issued eight at a time, summed in k order
        for (int kb = k0; kb < k1; kb += 8) {
            cd v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = kb + u;
                v[u] = cd{0.0, 0.0};
                if (k < k1) {
                    v[u] = X[static_cast<size_t>(k) * m + e];
                    if (w) v[u] = cscale(v[u], w[k]);
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (kb + u < k1) s = cadd(s, v[u]);
        }
        partial[static_cast<size_t>(blockIdx.x) * m + e] = s;
    }
}

// out[e] = sum_b partial[b][e]: the partials are split into G = 256 / m
// contiguous ranges (one per thread group, loads eight at a time), and the
// group sums are added in group order: deterministic for given nparts, m.
__global__ void __launch_bounds__(256)
k_ksum_final(int nparts, int m, const cd* __restrict__ partial, cd* __restrict__ out) {
    __shared__ cd red[256];
    const int G = m >= 256 ? 1 : 256 / m;
    const int span = (nparts + G - 1) / G;
    for (int t = threadIdx.x; t < G * m; t += blockDim.x) {
        const int e = t % m, g = t / m;
        const int b0 = g * span, b1 = min(nparts, b0 + span);
        cd s{0.0, 0.0};
        for (int bb = b0; bb < b1; bb += 8) {
            cd v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                v[u] = bb + u < b1 ? partial[static_cast<size_t>(bb + u) * m + e] : cd{0.0, 0.0};
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (bb + u < b1) s = cadd(s, v[u]);
        }
        if (G == 1)
            out[e] = s;
        else
            red[t] = s;
    }
    if (G == 1) return;  // uniform over the block
    __syncthreads();
    for (int e = threadIdx.x; e < m; e += blockDim.x) {
        cd s{0.0, 0.0};
        for (int g = 0; g < G; ++g) s = cadd(s, red[g * m + e]);
        out[e] = s;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

// Launch kern<generic, S> with S the geometry's slot count (2, 4, 8 or 16).
#define HFGPU_EXCH_LAUNCH(kern, gen, slots, grid, lds, st, ...)                          \
    do {                                                                                   \
        const dim3 blk_(kFftThreads);                                                      \
        if (gen) {                                                                         \
            if ((slots) <= 2) hipLaunchKernelGGL((kern<true, 2>), grid, blk_, lds, st, __VA_ARGS__);       \
            else if ((slots) <= 4) hipLaunchKernelGGL((kern<true, 4>), grid, blk_, lds, st, __VA_ARGS__);  \
            else if ((slots) <= 8) hipLaunchKernelGGL((kern<true, 8>), grid, blk_, lds, st, __VA_ARGS__);  \
            else hipLaunchKernelGGL((kern<true, 16>), grid, blk_, lds, st, __VA_ARGS__);                   \
        } else {                                                                           \
            if ((slots) <= 2) hipLaunchKernelGGL((kern<false, 2>), grid, blk_, lds, st, __VA_ARGS__);      \
            else if ((slots) <= 4) hipLaunchKernelGGL((kern<false, 4>), grid, blk_, lds, st, __VA_ARGS__); \
            else if ((slots) <= 8) hipLaunchKernelGGL((kern<false, 8>), grid, blk_, lds, st, __VA_ARGS__); \
            else hipLaunchKernelGGL((kern<false, 16>), grid, blk_, lds, st, __VA_ARGS__);                  \
        }                                                                                  \
        HFGPU_LAUNCH_CHECK();                                                              \
    } while (0)

static void launch_small_terms(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_small_terms, dim3(1), dim3(kFftThreads),
                       sizeof(cd) * small_terms_cd(g.nb), st, a.partial, g.nk1, g.nch,
                       g.d_chan_a, g.d_chan_b, g.nb, g.packed ? 1 : 0, a.small);
    HFGPU_LAUNCH_CHECK();
}

// LDS bytes of the two-step passes for a line of n points, w wavefronts
// (+ row-sum slots)
static size_t fast_lds(int n, int n1, int w, bool row_sums) {
    return sizeof(cd) * (static_cast<size_t>(n) * (1 + static_cast<size_t>(w) * kFastLines) +
                         (row_sums ? static_cast<size_t>(w) * kFastLines * n1 : 0));
}

static unsigned fast_grid(int lines, int nblk, int w) {
    return cdiv(static_cast<size_t>(lines) * nblk, w);
}

// grid of a row pass: rpw workgroups for each of nk1 rows (split_of order)
static unsigned row_grid(int nk1, int rpw) {
    return rpw == 1 ? static_cast<unsigned>(nk1)
                    : static_cast<unsigned>(cdiv(static_cast<size_t>(nk1), 8) * 8 * rpw);
}

// wavefronts (channel tiles) per workgroup of a row pass: the whole row if it
// fits (or g.row_waves when set), at most kFastRowWavesMax and the LDS
static int row_waves(const ExchangeGeom& g, int n1, bool row_sums, int split) {
    int w = g.row_waves > 0 ? g.row_waves : g.nblk;
    w = std::min({w, g.nblk, kFastRowWavesMax / split});
    while (w > 1 && fast_lds(g.nk2, n1, w, row_sums) > 160 * 1024) --w;
    return w;
}

static void launch_rows_fwd_fast(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st,
                                 const ExchBatchEntry* bt = nullptr, int nS = 1,
                                 unsigned bt_active = ~0u) {
#define HFGPU_FAST_CASE(N, A, B)                                                              \
    if (g.nk2 == N) {                                                                         \
        const int w = row_waves(g, A, true, g.split_fwd), rpw = (g.nblk + w - 1) / w;         \
        const dim3 grid(row_grid(g.nk1, rpw), nS), block(64 * w * g.split_fwd);               \
        const size_t lds = fast_lds(N, A, w, true);                                           \
        if (a.ws32 && g.split_fwd == 2)                                                       \
            hipLaunchKernelGGL((k_rows_fwd_fast<A, B, cf, 2>), grid, block, lds, st,          \
                               g.ax2.d_tw, g.nk1, g.nb, g.nch, g.nblk, g.packed ? 1 : 0, a.P, \
                               a.refP, a.refP0, a.wrow, a.T32, a.partial, a.w2d, rpw, bt, bt_active);    \
        else if (a.ws32)                                                                      \
            hipLaunchKernelGGL((k_rows_fwd_fast<A, B, cf, 1>), grid, block, lds, st,          \
                               g.ax2.d_tw, g.nk1, g.nb, g.nch, g.nblk, g.packed ? 1 : 0, a.P, \
                               a.refP, a.refP0, a.wrow, a.T32, a.partial, a.w2d, rpw, bt, bt_active);    \
        else if (g.split_fwd == 2)                                                            \
            hipLaunchKernelGGL((k_rows_fwd_fast<A, B, cd, 2>), grid, block, lds, st,          \
                               g.ax2.d_tw, g.nk1, g.nb, g.nch, g.nblk, g.packed ? 1 : 0, a.P, \
                               a.refP, a.refP0, a.wrow, a.T, a.partial, a.w2d, rpw, bt, bt_active);      \
        else                                                                                  \
            hipLaunchKernelGGL((k_rows_fwd_fast<A, B, cd, 1>), grid, block, lds, st,          \
                               g.ax2.d_tw, g.nk1, g.nb, g.nch, g.nblk, g.packed ? 1 : 0, a.P, \
                               a.refP, a.refP0, a.wrow, a.T, a.partial, a.w2d, rpw, bt, bt_active);      \
        HFGPU_LAUNCH_CHECK();                                                                 \
        return;                                                                               \
    }
    HFGPU_FAST_SIZES(HFGPU_FAST_CASE)
#undef HFGPU_FAST_CASE
    throw std::logic_error("no two-step transform for this row length");
}

static void launch_cols_fast(const ExchangeGeom& g, const ExchangeArgs& a, int hs,
                             hipStream_t st, const ExchBatchEntry* bt = nullptr, int nS = 1,
                             unsigned bt_active = ~0u) {
    const int cw = std::max(1, std::min(g.col_waves, kFastWaves / g.split_col));
    const dim3 grid(fast_grid(g.nk2, g.nblk, cw) + hs, nS), block(64 * cw * g.split_col);
#define HFGPU_FAST_CASE(N, A, B)                                                              \
    if (g.nk1 == N) {                                                                         \
        const size_t lds = std::max(fast_lds(N, A, cw, false),                                \
                                    hs ? sizeof(cd) * small_terms_cd(g.nb) : size_t(0));      \
        if (a.ws32 && g.split_col == 2)                                                       \
            hipLaunchKernelGGL((k_cols_fast<A, B, cf, 2>), grid, block, lds, st, g.ax1.d_tw,  \
                               g.nk2, g.nch, g.nblk, a.T32, a.VR32, a.vr_scalar ? 1 : 0, hs,  \
                               a.partial, g.d_chan_a, g.d_chan_b, g.nb, g.packed ? 1 : 0,     \
                               a.small, bt, bt_active);                                                  \
        else if (a.ws32)                                                                      \
            hipLaunchKernelGGL((k_cols_fast<A, B, cf, 1>), grid, block, lds, st, g.ax1.d_tw,  \
                               g.nk2, g.nch, g.nblk, a.T32, a.VR32, a.vr_scalar ? 1 : 0, hs,  \
                               a.partial, g.d_chan_a, g.d_chan_b, g.nb, g.packed ? 1 : 0,     \
                               a.small, bt, bt_active);                                                  \
        else if (g.split_col == 2)                                                            \
            hipLaunchKernelGGL((k_cols_fast<A, B, cd, 2>), grid, block, lds, st, g.ax1.d_tw,  \
                               g.nk2, g.nch, g.nblk, a.T, a.VR, a.vr_scalar ? 1 : 0, hs,      \
                               a.partial, g.d_chan_a, g.d_chan_b, g.nb, g.packed ? 1 : 0,     \
                               a.small, bt, bt_active);                                                  \
        else                                                                                  \
            hipLaunchKernelGGL((k_cols_fast<A, B, cd, 1>), grid, block, lds, st, g.ax1.d_tw,  \
                               g.nk2, g.nch, g.nblk, a.T, a.VR, a.vr_scalar ? 1 : 0, hs,      \
                               a.partial, g.d_chan_a, g.d_chan_b, g.nb, g.packed ? 1 : 0,     \
                               a.small, bt, bt_active);                                                  \
        HFGPU_LAUNCH_CHECK();                                                                 \
        return;                                                                               \
    }
    HFGPU_FAST_SIZES(HFGPU_FAST_CASE)
#undef HFGPU_FAST_CASE
    throw std::logic_error("no two-step transform for this column length");
}

static void launch_rows_inv_fast(const ExchangeGeom& g, const ExchangeArgs& a, double inv_n,
                                 hipStream_t st, const ExchBatchEntry* bt = nullptr,
                                 int nS = 1, unsigned bt_active = ~0u) {
#define HFGPU_FAST_CASE(N, A, B)                                                              \
    if (g.nk2 == N) {                                                                         \
        const int w = row_waves(g, A, false, g.split_inv), rpw = (g.nblk + w - 1) / w;        \
        const dim3 grid(row_grid(g.nk1, rpw), nS), block(64 * w * g.split_inv);               \
        const size_t lds = fast_lds(N, A, w, false);                                          \
        if (a.ws32 && g.split_inv == 2)                                                       \
            hipLaunchKernelGGL((k_rows_inv_fast<A, B, cf, 2>), grid, block, lds, st,          \
                               g.ax2.d_tw, g.nk1, g.nb, g.nch, g.nblk, g.packed ? 1 : 0,      \
                               a.T32, a.epilogue, a.ifftshift ? 1 : 0, inv_n, a.Sigma, a.F,   \
                               a.hh, a.SC, a.hdiag, rpw, bt, bt_active);                                 \
        else if (a.ws32)                                                                      \
            hipLaunchKernelGGL((k_rows_inv_fast<A, B, cf, 1>), grid, block, lds, st,          \
                               g.ax2.d_tw, g.nk1, g.nb, g.nch, g.nblk, g.packed ? 1 : 0,      \
                               a.T32, a.epilogue, a.ifftshift ? 1 : 0, inv_n, a.Sigma, a.F,   \
                               a.hh, a.SC, a.hdiag, rpw, bt, bt_active);                                 \
        else if (g.split_inv == 2)                                                            \
            hipLaunchKernelGGL((k_rows_inv_fast<A, B, cd, 2>), grid, block, lds, st,          \
                               g.ax2.d_tw, g.nk1, g.nb, g.nch, g.nblk, g.packed ? 1 : 0, a.T, \
                               a.epilogue, a.ifftshift ? 1 : 0, inv_n, a.Sigma, a.F, a.hh,    \
                               a.SC, a.hdiag, rpw, bt, bt_active);                                       \
        else                                                                                  \
            hipLaunchKernelGGL((k_rows_inv_fast<A, B, cd, 1>), grid, block, lds, st,          \
                               g.ax2.d_tw, g.nk1, g.nb, g.nch, g.nblk, g.packed ? 1 : 0, a.T, \
                               a.epilogue, a.ifftshift ? 1 : 0, inv_n, a.Sigma, a.F, a.hh,    \
                               a.SC, a.hdiag, rpw, bt, bt_active);                                       \
        HFGPU_LAUNCH_CHECK();                                                                 \
        return;                                                                               \
    }
    HFGPU_FAST_SIZES(HFGPU_FAST_CASE)
#undef HFGPU_FAST_CASE
    throw std::logic_error("no two-step transform for this row length");
}

static void launch_rows_fwd(const ExchangeGeom& g, const ExchangeArgs& a, int do_fft,
                            hipStream_t st) {
    if (a.refP0 && !a.wrow) throw std::logic_error("refP0 without row weights");
    if (do_fft && g.fast_rows) {
        launch_rows_fwd_fast(g, a, st);
        return;
    }
    HFGPU_EXCH_LAUNCH(k_rows_fwd, g.generic, g.slots_row, dim3(g.nblk * g.nk1), g.lds_rows, st,
                      g.ax2.axis(), g.nk1, g.nk2, g.nb, g.nch, g.clog, g.nblk, g.packed ? 1 : 0,
                      a.P, a.refP, a.refP0, a.wrow, a.T, a.partial, a.w2d, do_fft);
}

void launch_row_sums(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st) {
    launch_rows_fwd(g, a, 0, st);
    if (a.small.enabled) launch_small_terms(g, a, st);
}

void launch_exchange_fwd(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st) {
    if (a.ws32 && !(g.fast_rows && g.fast_cols))
        throw std::logic_error("the FP32 workspace needs the two-step transforms");
    launch_rows_fwd(g, a, 1, st);
    if (g.fast_cols) {
        // workgroup 0 of the column pass hosts the small terms
        launch_cols_fast(g, a, a.small.enabled ? 1 : 0, st);
        return;
    }
    // the small terms ride in pass B as one extra workgroup when its LDS fits
    const bool fuse_small = a.small.enabled && sizeof(cd) * small_terms_cd(g.nb) <= g.lds_cols;
    if (a.small.enabled && !fuse_small) launch_small_terms(g, a, st);
    const int hs = fuse_small ? 1 : 0;
    HFGPU_EXCH_LAUNCH(k_cols, g.generic, g.slots_col, dim3(g.nblk * g.nk2 + hs), g.lds_cols, st,
                      g.ax1.axis(), g.nk1, g.nk2, g.nch, g.clog, g.nblk, a.T, a.VR,
                      a.vr_scalar ? 1 : 0, hs, a.partial, g.d_chan_a, g.d_chan_b, g.nb,
                      g.packed ? 1 : 0, a.small);
}

void launch_exchange_fwd_batch(const ExchangeGeom& g, const ExchangeArgs& a,
                               const ExchBatchEntry* tab, int S, unsigned active,
                               hipStream_t st) {
    if (!(g.fast_rows && g.fast_cols) || a.ws32)
        throw std::logic_error("batched exchange needs the two-step transforms, FP64 workspace");
    launch_rows_fwd_fast(g, a, st, tab, S, active);
    launch_cols_fast(g, a, a.small.enabled ? 1 : 0, st, tab, S, active);
}

void launch_exchange_inv_batch(const ExchangeGeom& g, const ExchangeArgs& a,
                               const ExchBatchEntry* tab, int S, unsigned active,
                               hipStream_t st) {
    if (!(g.fast_rows && g.fast_cols) || a.ws32)
        throw std::logic_error("batched exchange needs the two-step transforms, FP64 workspace");
    if (a.epilogue != EPI_RAW_FULL && !g.packed)
        throw std::logic_error("packed epilogue on a full-channel geometry");
    if (a.epilogue == EPI_RAW_FULL && g.packed)
        throw std::logic_error("full epilogue on a packed geometry");
    const double inv_n = 1.0 / static_cast<double>(static_cast<size_t>(g.nk1) * g.nk2);
    launch_rows_inv_fast(g, a, inv_n, st, tab, S, active);
}

void launch_exchange_inv(const ExchangeGeom& g, const ExchangeArgs& a, hipStream_t st) {
    if (a.epilogue != EPI_RAW_FULL && !g.packed)
        throw std::logic_error("packed epilogue on a full-channel geometry");
    if (a.epilogue == EPI_RAW_FULL && g.packed)
        throw std::logic_error("full epilogue on a packed geometry");
    if (a.epilogue == EPI_FOCK_PACKED && !a.hh)
        throw std::logic_error("Fock epilogue without herm(h)");
    const double inv_n = 1.0 / static_cast<double>(static_cast<size_t>(g.nk1) * g.nk2);
    if (g.fast_rows) {
        launch_rows_inv_fast(g, a, inv_n, st);
        return;
    }
    if (a.ws32) throw std::logic_error("the FP32 workspace needs the two-step transforms");
    HFGPU_EXCH_LAUNCH(k_rows_inv, g.generic, g.slots_row, dim3(g.nblk * g.nk1), g.lds_rows, st,
                      g.ax2.axis(), g.nk1, g.nk2, g.nb, g.nch, g.clog, g.nblk, g.packed ? 1 : 0,
                      a.T, a.epilogue, a.ifftshift ? 1 : 0, inv_n, a.Sigma, a.F, a.hh, a.SC,
                      a.hdiag);
}

static unsigned pack_blocks(size_t total) {
    return static_cast<unsigned>(std::min<size_t>(cdiv(total, 256), 8192));
}

void launch_pack_vr(const ExchangeGeom& g, const cd* VR_full, cd* VRb, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_vr, dim3(pack_blocks(g.tile_elems())), dim3(256), 0, st, g.nk1,
                       g.nk2, g.nb, g.nch, g.clog, g.nblk, g.packed ? 1 : 0, VR_full, VRb);
    HFGPU_LAUNCH_CHECK();
}

void launch_pack_herm_h(const ExchangeGeom& g, const cd* h, cd* hhb, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_herm_h, dim3(pack_blocks(g.tile_elems())), dim3(256), 0, st,
                       g.nk1, g.nk2, g.nb, g.nch, g.clog, g.nblk, g.packed ? 1 : 0, h, hhb);
    HFGPU_LAUNCH_CHECK();
}

void launch_fock_assemble(int nk, int nb, const cd* h, cd* Sigma, cd* F, const cd* SC,
                          const double* hdiag, hipStream_t st) {
    const size_t total = static_cast<size_t>(nk) * nb * (nb + 1) / 2;
    const unsigned blocks = std::min<unsigned>(cdiv(total, 256), 4096u);
    hipLaunchKernelGGL(k_fock_assemble, dim3(blocks), dim3(256), 0, st, nk, nb, h, Sigma, F,
                       SC, hdiag);
    HFGPU_LAUNCH_CHECK();
}

void launch_energy_batch(const EnergyBatchEntry* tab, int S, int nk_max, int nb,
                         unsigned active, hipStream_t st) {
    hipLaunchKernelGGL(k_energy_partial_batch, dim3(cdiv(nk_max, 4), S), dim3(256), 0, st, tab,
                       nb, active);
    HFGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials_batch, dim3(S), dim3(256), 0, st, tab, active);
    HFGPU_LAUNCH_CHECK();
}

void launch_energy(int nk, int nb, const cd* h, const cd* P, const cd* refP, const cd* Sigma,
                   const double* hdiag, const double* w2d, double* partial, double* E_out,
                   hipStream_t st) {
    const unsigned blocks = cdiv(nk, 4);
    hipLaunchKernelGGL(k_energy_partial, dim3(blocks), dim3(256), 0, st, nk, nb, h, P, refP,
                       Sigma, hdiag, static_cast<const cd*>(nullptr), size_t(0), w2d, partial);
    HFGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial,
                       static_cast<int>(blocks), E_out);
    HFGPU_LAUNCH_CHECK();
}

void launch_energy_fullH(int nk, int nb, const cd* h, const cd* P, const cd* refP,
                         const cd* Sigma, const cd* Hm, size_t hk_stride, const double* w2d,
                         double* partial, double* E_out, hipStream_t st) {
    const unsigned blocks = cdiv(nk, 4);
    hipLaunchKernelGGL(k_energy_partial, dim3(blocks), dim3(256), 0, st, nk, nb, h, P, refP,
                       Sigma, static_cast<const double*>(nullptr), Hm, hk_stride, w2d, partial);
    HFGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial,
                       static_cast<int>(blocks), E_out);
    HFGPU_LAUNCH_CHECK();
}

namespace {
__global__ void k_entropy(const double* __restrict__ p, const double* __restrict__ w, int nb,
                          size_t n, double* partial) {
    __shared__ double red[4];
    double s = 0.0;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        double pv = p[i];
        if (pv < 1.0e-14) pv = 1.0e-14;
        if (pv > 1.0 - 1.0e-14) pv = 1.0 - 1.0e-14;
        s -= w[i / nb] * (pv * log(pv) + (1.0 - pv) * log1p(-pv));
    }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
}  // namespace

void launch_entropy(const double* p, const double* w, int nk, int nb, double* partial,
                    double* S_out, hipStream_t st) {
    const size_t n = static_cast<size_t>(nk) * nb;
    const unsigned blocks = std::max(1u, std::min<unsigned>(cdiv(n, 256), 1024u));
    hipLaunchKernelGGL(k_entropy, dim3(blocks), dim3(256), 0, st, p, w, nb, n, partial);
    HFGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial,
                       static_cast<int>(blocks), S_out);
    HFGPU_LAUNCH_CHECK();
}


// ---------------------------------------------------------------------------
// adjoint building blocks
// ---------------------------------------------------------------------------

namespace {
__global__ void __launch_bounds__(256) k_to_cf(size_t n, const cd* __restrict__ in, cf* __restrict__ out) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[i] = cf{static_cast<float>(in[i].x), static_cast<float>(in[i].y)};
}
}  // namespace

void launch_to_cf(size_t n, const cd* in, cf* out, hipStream_t st) {
    hipLaunchKernelGGL(k_to_cf, dim3(std::max(1u, std::min<unsigned>(cdiv(n, 256), 8192u))),
                       dim3(256), 0, st, n, in, out);
    HFGPU_LAUNCH_CHECK();
}

void launch_fft2_plain(const ExchangeGeom& g, const cd* in, cd* out, int nch, bool inverse,
                       hipStream_t st) {
    if (g.packed) throw std::logic_error("plain transforms need a full-channel geometry");
    if (nch > g.nch) throw std::logic_error("more channels than the geometry holds");
    const int nblk = (nch + g.tile() - 1) / g.tile();
    HFGPU_EXCH_LAUNCH(k_plain_rows, g.generic, g.slots_row, dim3(nblk * g.nk1), g.lds_rows, st,
                      g.ax2.axis(), g.nk1, g.nk2, nch, g.clog, nblk, inverse ? 1 : 0, in, out,
                      static_cast<const int*>(nullptr));
    HFGPU_EXCH_LAUNCH(k_plain_cols, g.generic, g.slots_col, dim3(nblk * g.nk2), g.lds_cols, st,
                      g.ax1.axis(), g.nk1, g.nk2, nch, g.clog, nblk, inverse ? 1 : 0, out);
}

void launch_sl_convolve(const ExchangeGeom& g, const SlRowsSrc& src, cd* X, int nch,
                        const cd* V, int period, const int* flags, int per_slot, double scale,
                        const int* rowmask, hipStream_t st) {
    if (g.packed) throw std::logic_error("plain transforms need a full-channel geometry");
    if (nch > g.nch) throw std::logic_error("more channels than the geometry holds");
    if (!rowmask) throw std::logic_error("the superlattice rows need a row mask");
    const int nblk = (nch + g.tile() - 1) / g.tile();
    HFGPU_EXCH_LAUNCH(k_sl_rows_fwd, g.generic, g.slots_row, dim3(nblk * g.nk1), g.lds_rows, st,
                      g.ax2.axis(), g.nk1, g.nk2, nch, g.clog, nblk, src, rowmask, X);
    HFGPU_EXCH_LAUNCH(k_plain_cols_conv, g.generic, g.slots_col, dim3(nblk * g.nk2), g.lds_cols,
                      st, g.ax1.axis(), g.nk1, g.nk2, nch, g.clog, nblk, X, V, period, flags,
                      per_slot, scale, rowmask);
    if (src.Sigma)
        HFGPU_EXCH_LAUNCH(k_sl_rows_inv, g.generic, g.slots_row, dim3(nblk * g.nk1), g.lds_rows,
                          st, g.ax2.axis(), g.nk1, g.nk2, nch, g.clog, nblk, src, rowmask,
                          static_cast<const cd*>(X));
    else
        HFGPU_EXCH_LAUNCH(k_plain_rows, g.generic, g.slots_row, dim3(nblk * g.nk1), g.lds_rows,
                          st, g.ax2.axis(), g.nk1, g.nk2, nch, g.clog, nblk, 1, X, X, rowmask);
}

static unsigned ew_blocks(size_t n) {
    return std::max(1u, std::min<unsigned>(cdiv(n, 256), 8192u));
}

void launch_cross_spectrum(size_t nq, int nch, const cd* A, const cd* B, double scale,
                           bool sum_channels, cd* out, hipStream_t st) {
    hipLaunchKernelGGL(k_cross_spectrum, dim3(ew_blocks(sum_channels ? nq : nq * nch)),
                       dim3(256), 0, st, nq, nch, A, B, scale, sum_channels ? 1 : 0, out);
    HFGPU_LAUNCH_CHECK();
}

void launch_mul_spectrum(size_t nq, int nch, cd* X, const cd* V, bool v_scalar, bool conj_v,
                         double scale, hipStream_t st) {
    hipLaunchKernelGGL(k_mul_spectrum, dim3(ew_blocks(nq * nch)), dim3(256), 0, st, nq, nch, X,
                       V, v_scalar ? 1 : 0, conj_v ? 1 : 0, scale);
    HFGPU_LAUNCH_CHECK();
}

void launch_unphase(int nk1, int nk2, int nch, cd* X, hipStream_t st) {
    hipLaunchKernelGGL(k_unphase, dim3(ew_blocks(static_cast<size_t>(nk1) * nk2 * nch)),
                       dim3(256), 0, st, nk1, nk2, nch, X);
    HFGPU_LAUNCH_CHECK();
}

void launch_ksum(int nk, int m, const cd* X, const double* w, cd* partial, cd* out,
                 hipStream_t st) {
    const int nparts = std::max(1, std::min(kKsumParts, nk));
    hipLaunchKernelGGL(k_ksum_partial, dim3(nparts), dim3(256), 0, st, nk, m, X, w, partial);
    HFGPU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_ksum_final, dim3(1), dim3(256), 0, st, nparts, m, partial, out);
    HFGPU_LAUNCH_CHECK();
}

}  // namespace hfgpu
