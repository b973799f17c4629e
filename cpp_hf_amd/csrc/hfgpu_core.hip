// libhfgpu core: device-resident HF kernel, Fock build orchestration and the
// C ABI entry points that are not solver loops.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/hfgpu.h"
#include "hfgpu_core.h"
#include "solver_kernels.h"
#include "superlattice.h"

namespace hfgpu {

thread_local std::string g_last_error;

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------

namespace {

__global__ void k_fill_zero(unsigned long long* p, size_t n8) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n8;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = 0ull;
}

__global__ void k_axpy_c(size_t n, double alpha, const cd* __restrict__ x, cd* __restrict__ y) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const cd a = y[i], b = x[i];
        y[i] = cd{a.x + alpha * b.x, a.y + alpha * b.y};
    }
}

__global__ void k_sub_c(size_t n, const cd* __restrict__ a, const cd* __restrict__ b,
                        cd* __restrict__ out) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[i] = csub(a[i], b[i]);
}

unsigned grid_for(size_t n) { return std::max(1u, std::min<unsigned>(cdiv(n, 256), 8192u)); }

}  // namespace

void fill_zero(void* p, size_t bytes, hipStream_t st) {
    if (!bytes) return;
    if (bytes % 8) throw std::logic_error("fill_zero needs a multiple of 8 bytes");
    const size_t n8 = bytes / 8;
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n8)), dim3(256), 0, st,
                       static_cast<unsigned long long*>(p), n8);
    HFGPU_LAUNCH_CHECK();
}

void launch_axpy_c(size_t n, double alpha, const cd* x, cd* y, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_axpy_c, dim3(grid_for(n)), dim3(256), 0, st, n, alpha, x, y);
    HFGPU_LAUNCH_CHECK();
}

void launch_sub_c(size_t n, const cd* a, const cd* b, cd* out, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_sub_c, dim3(grid_for(n)), dim3(256), 0, st, n, a, b, out);
    HFGPU_LAUNCH_CHECK();
}

bool host_is_hermitian(const double* M, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    for (size_t k = 0; k < nk; ++k) {
        const double* m = M + 2 * k * nb2;
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = i; j < nb; ++j) {
                const double* a = m + 2 * (i * nb + j);
                const double* b = m + 2 * (j * nb + i);
                if (a[0] != b[0] || a[1] != -b[1]) return false;
            }
    }
    return true;
}

void host_hermitize(double* M, size_t nk, size_t nb) {
    const size_t nb2 = nb * nb;
    for (size_t k = 0; k < nk; ++k) {
        double* m = M + 2 * k * nb2;
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = i; j < nb; ++j) {
                double* a = m + 2 * (i * nb + j);
                double* b = m + 2 * (j * nb + i);
                const double hr = 0.5 * (a[0] + b[0]);
                const double hi = 0.5 * (a[1] - b[1]);
                a[0] = hr;
                a[1] = hi;
                b[0] = hr;
                b[1] = -hi;
            }
    }
}

// ---------------------------------------------------------------------------
// Device memory pool
// ---------------------------------------------------------------------------

namespace {

struct DevPool {
    std::mutex mu;
    std::map<std::pair<int, size_t>, std::vector<void*>> free_lists;  // (device, class)
};

DevPool& dev_pool() {
    static DevPool* p = new DevPool;  // outlives every static DevBuf
    return *p;
}

size_t size_class(size_t bytes) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
}

}  // namespace

void* dev_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    int dev = 0;
    HFGPU_CHECK(hipGetDevice(&dev));
    const size_t cls = size_class(bytes);
    DevPool& pool = dev_pool();
    {
        std::lock_guard<std::mutex> lock(pool.mu);
        auto it = pool.free_lists.find({dev, cls});
        if (it != pool.free_lists.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            return p;
        }
    }
    void* p = nullptr;
    HFGPU_CHECK(hipMalloc(&p, cls));
    return p;
}

namespace {

struct StreamPool {
    std::mutex mu;
    std::map<int, std::vector<hipStream_t>> idle;  // device -> streams
};

StreamPool& stream_pool() {
    static StreamPool* p = new StreamPool;  // streams outlive static destructors
    return *p;
}

}  // namespace

hipStream_t stream_acquire() {
    int dev = 0;
    HFGPU_CHECK(hipGetDevice(&dev));
    StreamPool& pool = stream_pool();
    {
        std::lock_guard<std::mutex> lock(pool.mu);
        auto it = pool.idle.find(dev);
        if (it != pool.idle.end() && !it->second.empty()) {
            hipStream_t s = it->second.back();
            it->second.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    HFGPU_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void stream_release(hipStream_t s) {
    if (!s) return;
    int dev = 0;
    if (hipStreamGetDevice(s, &dev) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return;
    }
    StreamPool& pool = stream_pool();
    std::lock_guard<std::mutex> lock(pool.mu);
    pool.idle[dev].push_back(s);
}

hipError_t stream_sync_release(hipStream_t s) {
    if (!s) return hipSuccess;
    const hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess)
        stream_release(s);
    else
        (void)hipStreamDestroy(s);
    return e;
}

void dev_free(void* p, size_t bytes) {
    if (!p) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    // A buffer freed while an exception unwinds may still be read or written
    // by work queued on a stream that has not drained (the stream objects of
    // the entry points may be destroyed after their buffers): drain the
    // device before the memory can reach another thread's allocation.
    if (std::uncaught_exceptions() > 0) (void)hipDeviceSynchronize();
    DevPool& pool = dev_pool();
    std::lock_guard<std::mutex> lock(pool.mu);
    pool.free_lists[{dev, size_class(bytes)}].push_back(p);
}

// ---------------------------------------------------------------------------
// GpuKernel
// ---------------------------------------------------------------------------

namespace {

// std::max's selection (NaN operands ignored as the host loop ignored them)
__device__ __forceinline__ double hmax(double a, double b) { return a < b ? b : a; }

constexpr int kCheckBlocks = 256;

// Packed exchange for (nb, nb) V needs the index-space kernel W_ab =
// IFFT(VR_ab) to be real and symmetric in (a, b):
//   VR_ab(q) == VR_ba(q)  and  VR_ab(-q) == conj(VR_ab(q)),
// each to 64 eps of max |VR| (compared as squared magnitudes).  Per
// workgroup: max |VR|^2 and the larger squared defect, without contraction
// (maxima are exact, so the decision does not depend on the order).
__global__ void __launch_bounds__(256)
k_vr_check(const double* __restrict__ VR, int nk1, int nk2, int nb, double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double red[2][4];
    const size_t nb2 = static_cast<size_t>(nb) * nb;
    const size_t n = static_cast<size_t>(nk1) * nk2 * nb2;
    double vmax2 = 0.0, err2 = 0.0;
    for (size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; idx < n;
         idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t q = idx / nb2, ab = idx % nb2;
        const size_t a = ab / nb, b = ab % nb;
        const size_t q1 = q / nk2, q2 = q % nk2;
        const size_t mq = ((nk1 - q1) % nk1) * nk2 + (nk2 - q2) % nk2;
        const double* v = VR + 2 * (q * nb2 + a * nb + b);
        const double* vt = VR + 2 * (q * nb2 + b * nb + a);
        const double* vm = VR + 2 * (mq * nb2 + a * nb + b);
        const double d1r = v[0] - vt[0], d1i = v[1] - vt[1];
        const double d2r = vm[0] - v[0], d2i = vm[1] + v[1];
        vmax2 = hmax(vmax2, v[0] * v[0] + v[1] * v[1]);
        err2 = hmax(err2, hmax(d1r * d1r + d1i * d1i, d2r * d2r + d2i * d2i));
    }
    for (int off = 32; off > 0; off >>= 1) {
        vmax2 = hmax(vmax2, __shfl_xor(vmax2, off, 64));
        err2 = hmax(err2, __shfl_xor(err2, off, 64));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wave] = vmax2;
        red[1][wave] = err2;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double m = red[threadIdx.x][0];
        for (int w = 1; w < 4; ++w) m = hmax(m, red[threadIdx.x][w]);
        out[2 * blockIdx.x + threadIdx.x] = m;
    }
}

// 1 in out[block] when some k point's reference block differs bit for bit
// from the first (the memcmp test the host ran before).
__global__ void __launch_bounds__(256)
k_refp_const(const unsigned long long* __restrict__ R, size_t n_words, size_t blk_words,
             int* __restrict__ out) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    int diff = 0;
    for (size_t i = blk_words + blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
         i < n_words; i += static_cast<size_t>(gridDim.x) * blockDim.x)
        diff |= R[i] != R[i % blk_words] ? 1 : 0;
    if (diff) any = 1;  // every writer writes 1
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = any;
}

}  // namespace

// The two set-up tests on the uploaded arrays, on the device: a host loop
// over VR's transposed and mirrored entries took milliseconds at 120 x 120
// k with 12 bands.
static void device_setup_checks(const cd* VR, bool vr_test, const cd* refP, bool refp_test,
                                int nk1, int nk2, int nb, hipStream_t st, bool* packable,
                                bool* refp_const) {
    const size_t nb2 = static_cast<size_t>(nb) * nb, nd = static_cast<size_t>(nk1) * nk2 * nb2;
    const unsigned blocks = std::max(1u, std::min<unsigned>(cdiv(nd, 256), kCheckBlocks));
    DevBuf vr_part(2 * kCheckBlocks * sizeof(double)), rp_part(kCheckBlocks * sizeof(int));
    if (vr_test)
        hipLaunchKernelGGL(k_vr_check, dim3(blocks), dim3(256), 0, st,
                           reinterpret_cast<const double*>(VR), nk1, nk2, nb,
                           vr_part.as<double>());
    if (refp_test)
        hipLaunchKernelGGL(k_refp_const, dim3(blocks), dim3(256), 0, st,
                           reinterpret_cast<const unsigned long long*>(refP), 2 * nd, 2 * nb2,
                           rp_part.as<int>());
    HFGPU_LAUNCH_CHECK();
    std::vector<double> hv(2 * blocks);
    std::vector<int> hr(blocks);
    if (vr_test) d2h(hv.data(), vr_part.get(), hv.size() * sizeof(double), st);
    if (refp_test) d2h(hr.data(), rp_part.get(), hr.size() * sizeof(int), st);
    HFGPU_CHECK(hipStreamSynchronize(st));
    if (vr_test) {
        double vmax2 = 0.0, err2 = 0.0;
        for (unsigned b = 0; b < blocks; ++b) {
            vmax2 = std::max(vmax2, hv[2 * b]);
            err2 = std::max(err2, hv[2 * b + 1]);
        }
        const double tol = 64.0 * 2.220446049250313e-16;
        *packable = err2 <= tol * tol * vmax2;
    }
    if (refp_test) {
        bool c = true;
        for (unsigned b = 0; b < blocks; ++b) c = c && hr[b] == 0;
        *refp_const = c;
    }
}

GpuKernel::GpuKernel(const hf_kernel_desc& d) {
    if (d.nk1 == 0 || d.nk2 == 0 || d.nb == 0)
        throw std::invalid_argument("h must be (nk1,nk2,nb,nb) with non-zero sizes");
    if (!((d.dv1 == 1 && d.dv2 == 1) || (d.dv1 == d.nb && d.dv2 == d.nb)))
        throw std::invalid_argument("VR last two dims must be (1,1) or (nb,nb)");
    if (d.has_refP && !d.refP)
        throw std::invalid_argument("refP must match h shape when has_refP is true");
    if (!d.h || !d.VR || !d.w2d || !d.HH || !d.contact_g || !d.contact_Oi || !d.contact_Oj)
        throw std::invalid_argument("kernel description has a null array");
    nk1 = static_cast<int>(d.nk1);
    nk2 = static_cast<int>(d.nk2);
    nb = static_cast<int>(d.nb);
    dv = static_cast<int>(d.dv1);
    n_contact = static_cast<int>(d.n_contact);
    nk = d.nk1 * d.nk2;
    nb2 = d.nb * d.nb;
    nd = nk * nb2;
    hartree = d.include_hartree != 0;
    exchange = d.include_exchange != 0;
    hcp = d.exchange_hcp != 0 && dv == 1;
    has_refP = d.has_refP != 0;
    T = d.T;
    weight_sum = d.weight_sum;
    w_host.assign(d.w2d, d.w2d + nk);
    // fock.hpp:25-29
    contact = n_contact > 1;
    for (int t = 0; t < n_contact; ++t)
        if (d.contact_g[t] != 0.0) contact = true;

    // Device Fock builds launch eagerly by default: on MI355X the four
    // kernels on one stream measured faster than the replay of a captured
    // hipGraph of the same launches.  HFGPU_GRAPH=1 selects the graph replay.
    if (const char* e = std::getenv("HFGPU_GRAPH")) use_graph = e[0] == '1';
    if (const char* e = std::getenv("HFGPU_AUTOTUNE")) autotune_env = e[0] == '1';
    // the kernel's own stream (a solve's) is created, not pooled: HIP deals
    // new streams to its hardware queues in turn, so solves running at
    // the same time spread over the queues; a pooled stream keeps whatever
    // queue it was created on
    HFGPU_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const size_t cz = sizeof(cd);
    h.alloc(nd * cz);
    h2d(h.get(), d.h, nd * cz, stream);
    if (has_refP) {
        refP.alloc(nd * cz);
        h2d(refP.get(), d.refP, nd * cz, stream);
    }
    w2d.alloc(nk * sizeof(double));
    h2d(w2d.get(), d.w2d, nk * sizeof(double), stream);
    HH.alloc(nb2 * sizeof(double));
    h2d(HH.get(), d.HH, nb2 * sizeof(double), stream);
    if (n_contact > 0) {
        cg.alloc(n_contact * sizeof(double));
        h2d(cg.get(), d.contact_g, n_contact * sizeof(double), stream);
        cg_host.assign(d.contact_g, d.contact_g + n_contact);
        Oi.alloc(n_contact * nb2 * cz);
        Oj.alloc(n_contact * nb2 * cz);
        h2d(Oi.get(), d.contact_Oi, n_contact * nb2 * cz, stream);
        h2d(Oj.get(), d.contact_Oj, n_contact * nb2 * cz, stream);
    }
    const size_t nvr = static_cast<size_t>(dv) * dv;
    VR_full.alloc(nk * nvr * cz);
    h2d(VR_full.get(), d.VR, nk * nvr * cz, stream);
    const bool vr_test = exchange && dv > 1 && nb > 1;
    if (vr_test || has_refP)
        device_setup_checks(VR_full.as<cd>(), vr_test, has_refP ? refP.as<cd>() : nullptr,
                            has_refP, nk1, nk2, nb, stream, &dense_packable, &refP_const);
    if (has_refP) {
        if (refP_const) {
            refP0.alloc(nb2 * cz);
            h2d(refP0.get(), d.refP, nb2 * cz, stream);
            std::vector<double> wr(d.nk1, 0.0);
            for (size_t i = 0; i < d.nk1; ++i)
                for (size_t j = 0; j < d.nk2; ++j) wr[i] += d.w2d[i * d.nk2 + j];
            wrow.alloc(d.nk1 * sizeof(double));
            h2d(wrow.get(), wr.data(), d.nk1 * sizeof(double), stream);
            HFGPU_CHECK(hipStreamSynchronize(stream));
        }
    }
    // workspace of either geometry: at most nb^2 + 15 padded channels
    T_ws.alloc((nb2 + 16) * nk * cz);
    partial.alloc(static_cast<size_t>(nk1) * nb2 * cz);
    SC.alloc(nb2 * cz);
    hdiag.alloc(std::max<size_t>(nb, 1) * sizeof(double));
    e_partial.alloc((cdiv(nk, 4) + 1) * sizeof(double));
    e_out.alloc(sizeof(double));
    fill_zero(SC.get(), nb2 * cz, stream);
    fill_zero(hdiag.get(), nb * sizeof(double), stream);
    HFGPU_CHECK(hipStreamSynchronize(stream));

    const hf_superlattice_desc* L = d.superlattice;
    if (L && (L->fock_active || L->hartree_active)) {
        // inject_superlattice_into_kernel's checks (cpp_hf_native.cpp:109-165)
        if (L->n_G * L->dim_orb != d.nb)
            throw std::invalid_argument("superlattice mode requires n_G × dim_orb == nb");
        if (!L->V_lag_fft || !L->g_a_off || !L->pair_i || !L->pair_j || !L->pair_start ||
            !L->pair_to_delta || !L->HH_GG)
            throw std::invalid_argument("superlattice description has a null array");
        // fock.hpp:176-199: with the exchange on, superlattice mode replaces
        // the mesh exchange; the mesh exchange beside a superlattice Hartree
        // shift is not a combination the superlattice kernel makes
        if (exchange && !L->fock_active)
            throw std::invalid_argument(
                "the GPU superlattice kernel needs superlattice_fock_active when the exchange is on");
        try {
            if (exchange) {
                SlInput in;
                in.V = reinterpret_cast<const cd*>(L->V_lag_fft);
                in.V_orb = reinterpret_cast<const cd*>(L->V_lag_fft_orbital);
                in.g_a_off = L->g_a_off;
                in.pair_i = L->pair_i;
                in.pair_j = L->pair_j;
                in.pair_start = L->pair_start;
                in.n_pairs = L->n_pairs;
                in.n_delta = L->n_delta;
                in.Nx = L->N_ext_x;
                in.Ny = L->N_ext_y;
                in.nkx = d.nk1;
                in.nky = d.nk2;
                in.n_G = L->n_G;
                in.orb = L->dim_orb;
                // fock.hpp:180-188: P - refP is Hermitian, one channel of
                // each (+dG, -dG) pair is transformed
                in.hermitian = true;
                sl_plan = std::make_unique<SlPlan>();
                sl_plan->init(in, stream);
            }
            if (L->hartree_active && hartree) {
                sl_hartree = std::make_unique<SlHartree>();
                sl_hartree->init(*L, stream);
                rho_bar.alloc(nb2 * cz);
            }
        } catch (...) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
            stream = nullptr;
            throw;
        }
        // fock.hpp:222: the mesh Hartree diagonal only without the shift
        if (L->hartree_active) hartree = false;
    }
}

GpuKernel::~GpuKernel() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    if (geo_full.g) geo_full.g->release();
    if (geo_packed.g) geo_packed.g->release();
    if (adj_geo) adj_geo->release();
    if (adj_geo_v) adj_geo_v->release();
}

// launch shapes of the two-step exchange passes (autotuner below)
namespace {

struct LaunchShape {
    int row_waves, col_waves, split_fwd, split_col, split_inv;
};

LaunchShape shape_of(const ExchangeGeom& g) {
    return {g.row_waves, g.col_waves, g.split_fwd, g.split_col, g.split_inv};
}

void set_shape(ExchangeGeom& g, const LaunchShape& s) {
    g.row_waves = s.row_waves;
    g.col_waves = s.col_waves;
    g.split_fwd = s.split_fwd;
    g.split_col = s.split_col;
    g.split_inv = s.split_inv;
}

// (nk1, nk2, nb, packed, dense V, small terms)
using TuneKey = std::tuple<int, int, int, bool, bool, bool>;
std::mutex g_tune_mu;
std::map<TuneKey, LaunchShape> g_tuned;

}  // namespace

GpuKernel::Geom& GpuKernel::geom(bool packed) {
    Geom& G = packed ? geo_packed : geo_full;
    if (!G.g) {
        auto g = std::make_unique<ExchangeGeom>();
        g->init(nk1, nk2, nb, packed);
        const size_t bytes = g->tile_elems() * sizeof(cd);
        if (bytes > T_ws.size()) throw std::logic_error("exchange workspace too small");
        if (exchange && dv > 1) {
            G.VRb.alloc(bytes);
            launch_pack_vr(*g, VR_full.as<cd>(), G.VRb.as<cd>(), stream);
        }
        if (packed) {
            G.hhb.alloc(bytes);
            launch_pack_herm_h(*g, h.as<cd>(), G.hhb.as<cd>(), stream);
        }
        HFGPU_CHECK(hipStreamSynchronize(stream));
        G.g = std::move(g);
        // the solvers' eigh of this kernel's nk matrices (matrices per wavefront)
        if (autotune_env) eigh_autotune_once(static_cast<int>(nk), nb, stream);
        if (autotune_env && exchange && !superlattice() && packed == packed_for(packed)) {
            const TuneKey key{nk1, nk2, nb, packed, dv > 1, hartree || contact};
            bool cached = false;
            {
                std::lock_guard<std::mutex> lock(g_tune_mu);
                const auto it = g_tuned.find(key);
                if (it != g_tuned.end()) {
                    set_shape(*G.g, it->second);
                    cached = true;
                }
            }
            if (!cached) {
                int shape[5];
                (void)autotune(packed, shape);
            }
        }
    }
    return G;
}

// ---------------------------------------------------------------------------
// dataflow floor: the two-step build's passes as plain streaming launches
// ---------------------------------------------------------------------------

namespace {

// Reads a[0..na) and b[0..nb) and writes o1[0..n1) and o2[0..n2) once each,
// 16 B per access, coalesced; the stores depend on the loads so that
// neither is dropped.  No transform, no LDS, no gather.
__global__ void __launch_bounds__(256)
k_stream_pass(const cd* __restrict__ a, size_t na, const cd* __restrict__ b, size_t nbb,
              cd* __restrict__ o1, size_t n1, cd* __restrict__ o2, size_t n2) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t t0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    cd acc{0.0, 0.0};
    for (size_t i = t0; i < na; i += stride) acc = cadd(acc, a[i]);
    for (size_t i = t0; i < nbb; i += stride) acc = cadd(acc, b[i]);
    for (size_t i = t0; i < n1; i += stride) o1[i] = cd{acc.x + static_cast<double>(i), acc.y};
    for (size_t i = t0; i < n2; i += stride) o2[i] = cd{acc.y, acc.x - static_cast<double>(i)};
}

}  // namespace

double GpuKernel::dataflow_floor(bool p_hermitian, int reps, double* bytes_per_build) {
    if (reps < 1) throw std::invalid_argument("dataflow_floor needs reps >= 1");
    const bool packed = exchange && packed_for(p_hermitian);
    Geom& G = geom(packed);
    if (!exchange || superlattice() || !(G.g->fast_rows && G.g->fast_cols))
        throw std::invalid_argument("dataflow_floor needs the two-step exchange");
    // pass A: P -> T; pass B: T, V -> T; pass C: T, h -> Sigma, F (the
    // k-independent reference density and the small terms are not read)
    const size_t nt = G.g->tile_elems();
    DevBuf P(nd * sizeof(cd)), T(nt * sizeof(cd)), T2(nt * sizeof(cd)), V(nt * sizeof(cd)),
        Hh(nt * sizeof(cd)), S(nd * sizeof(cd)), F(nd * sizeof(cd));
    fill_zero(P.get(), nd * sizeof(cd), stream);
    fill_zero(V.get(), nt * sizeof(cd), stream);
    fill_zero(Hh.get(), nt * sizeof(cd), stream);
    *bytes_per_build = static_cast<double>(sizeof(cd)) *
                       static_cast<double>(nd + nt + 2 * nt + nt + 2 * nt + 2 * nd);
    const dim3 grid(2048), block(256);
    auto build = [&] {
        hipLaunchKernelGGL(k_stream_pass, grid, block, 0, stream, P.as<cd>(), nd,
                           static_cast<const cd*>(nullptr), size_t(0), T.as<cd>(), nt,
                           static_cast<cd*>(nullptr), size_t(0));
        hipLaunchKernelGGL(k_stream_pass, grid, block, 0, stream, T.as<cd>(), nt, V.as<cd>(), nt,
                           T2.as<cd>(), nt, static_cast<cd*>(nullptr), size_t(0));
        hipLaunchKernelGGL(k_stream_pass, grid, block, 0, stream, T2.as<cd>(), nt, Hh.as<cd>(),
                           nt, S.as<cd>(), nd, F.as<cd>(), nd);
        HFGPU_LAUNCH_CHECK();
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HFGPU_CHECK(hipEventCreate(&e0));
    HFGPU_CHECK(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        hipStream_t s;
        ~EventGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{e0, e1, stream};
    for (int i = 0; i < 10; ++i) build();
    float best = 1e30f;
    for (int t = 0; t < 3; ++t) {
        HFGPU_CHECK(hipEventRecord(e0, stream));
        for (int i = 0; i < reps; ++i) build();
        HFGPU_CHECK(hipEventRecord(e1, stream));
        HFGPU_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HFGPU_CHECK(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    return 1e3 * static_cast<double>(best) / reps;
}

// ---------------------------------------------------------------------------
// launch-shape autotuner
// ---------------------------------------------------------------------------


double GpuKernel::autotune(bool p_hermitian, int shape_out[5]) {
    const bool packed = exchange && packed_for(p_hermitian);
    Geom& G = geom(packed);
    ExchangeGeom& g = *G.g;
    auto report = [&](const LaunchShape& s) {
        shape_out[0] = s.row_waves;
        shape_out[1] = s.col_waves;
        shape_out[2] = s.split_fwd;
        shape_out[3] = s.split_col;
        shape_out[4] = s.split_inv;
    };
    report(shape_of(g));
    if (!exchange || superlattice() || !(g.fast_rows && g.fast_cols)) return 0.0;
    const TuneKey key{nk1, nk2, nb, packed, dv > 1, hartree || contact};
    DevBuf P(nd * sizeof(cd)), S(nd * sizeof(cd)), F(nd * sizeof(cd)),
        hd(std::max<size_t>(nb, 1) * sizeof(double));
    fill_zero(P.get(), nd * sizeof(cd), stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HFGPU_CHECK(hipEventCreate(&e0));
    HFGPU_CHECK(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{e0, e1};
    int reps = 5;
    auto time_shape = [&](const LaunchShape& s) {
        set_shape(g, s);
        for (int i = 0; i < 2; ++i)
            build_fock(P.as<cd>(), S.as<cd>(), F.as<cd>(), hd.as<double>(), p_hermitian, stream);
        float best = 1e30f;
        for (int t = 0; t < 2; ++t) {
            HFGPU_CHECK(hipEventRecord(e0, stream));
            for (int i = 0; i < reps; ++i)
                build_fock(P.as<cd>(), S.as<cd>(), F.as<cd>(), hd.as<double>(), p_hermitian,
                           stream);
            HFGPU_CHECK(hipEventRecord(e1, stream));
            HFGPU_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HFGPU_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        return 1e3 * static_cast<double>(best) / reps;
    };
    LaunchShape best = shape_of(g);
    double best_us = time_shape(best);
    // about 2 ms of timed builds per candidate
    reps = std::max(5, std::min(50, static_cast<int>(2000.0 / std::max(best_us, 1.0))));
    best_us = time_shape(best);
    auto consider = [&](LaunchShape s) {
        const double us = time_shape(s);
        if (us < best_us) {
            best_us = us;
            best = s;
        }
    };
    for (int sf : {1, 2}) {
        LaunchShape s = best;
        s.split_fwd = sf;
        if (sf != best.split_fwd) consider(s);
    }
    for (int sc : {1, 2})
        for (int cw : {1, 2, 4}) {
            if (sc * cw > 4) continue;
            LaunchShape s = best;
            s.split_col = sc;
            s.col_waves = cw;
            if (sc != best.split_col || cw != best.col_waves) consider(s);
        }
    for (int si : {1, 2}) {
        LaunchShape s = best;
        s.split_inv = si;
        if (si != best.split_inv) consider(s);
    }
    for (int rw : {1, 2, 0}) {
        if (rw == 2 && g.nblk < 3) continue;
        LaunchShape s = best;
        s.row_waves = rw;
        if (rw != best.row_waves) consider(s);
    }
    set_shape(g, best);
    HFGPU_CHECK(hipStreamSynchronize(stream));
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tuned[key] = best;
    }
    report(best);
    return best_us;
}

bool GpuKernel::fp32_capable(bool p_hermitian) {
    if (!exchange || sl_plan) return false;
    const ExchangeGeom& g = *geom(packed_for(p_hermitian)).g;
    return g.fast_rows && g.fast_cols;
}

void GpuKernel::build_fock(const cd* P, cd* Sigma, cd* F, double* hdiag_out, bool p_hermitian,
                           hipStream_t st, bool use_refP, bool fp32_ws) {
    const bool sl_x = exchange && sl_plan;  // superlattice exchange
    const bool need_partials = hartree || contact || sl_hartree;
    const bool mesh_x = exchange && !sl_x;
    const bool packed = mesh_x && packed_for(p_hermitian);
    Geom& G = geom(packed);
    const ExchangeGeom& g = *G.g;
    ExchangeArgs a;
    if (fp32_ws && mesh_x && g.fast_rows && g.fast_cols) {
        // FP32 copies of the workspace and the interaction, made once
        if (T32_ws.size() < g.tile_elems() * sizeof(cf)) T32_ws.alloc(g.tile_elems() * sizeof(cf));
        if (!G.VR32.get()) {
            const size_t n = dv > 1 ? g.tile_elems() : nk;
            G.VR32.alloc(n * sizeof(cf));
            launch_to_cf(n, dv > 1 ? G.VRb.as<cd>() : VR_full.as<cd>(), G.VR32.as<cf>(), stream);
            HFGPU_CHECK(hipStreamSynchronize(stream));
        }
        a.ws32 = true;
        a.T32 = T32_ws.as<cf>();
        a.VR32 = G.VR32.as<cf>();
    }
    a.P = P;
    if (has_refP && use_refP) {
        if (refP_const) {
            a.refP0 = refP0.as<cd>();
            a.wrow = wrow.as<double>();
        } else {
            a.refP = refP.as<cd>();
        }
    }
    a.T = T_ws.as<cd>();
    a.partial = need_partials ? partial.as<cd>() : nullptr;
    a.w2d = w2d.as<double>();
    if (dv > 1) {
        a.VR = G.VRb.as<cd>();
        a.vr_scalar = false;
    } else {
        a.VR = VR_full.as<cd>();
        a.vr_scalar = true;
    }
    double* hd = hdiag_out ? hdiag_out : hdiag.as<double>();
    if (need_partials) {
        SmallTerms& s = a.small;
        s.enabled = 1;
        s.include_hartree = hartree ? 1 : 0;
        s.has_contact = contact ? 1 : 0;
        s.HH = HH.as<double>();
        s.n_contact = n_contact;
        s.contact_g = cg.as<double>();
        s.Oi = Oi.as<cd>();
        s.Oj = Oj.as<cd>();
        s.hdiag = hd;
        s.SC = SC.as<cd>();
        s.rho_out = sl_hartree ? rho_bar.as<cd>() : nullptr;
    }
    if (mesh_x)
        launch_exchange_fwd(g, a, st);
    else if (need_partials)
        launch_row_sums(g, a, st);
    if (!need_partials) fill_zero(hd, nb * sizeof(double), st);
    // fock.hpp:66-175, 200-212: the superlattice Hartree shift joins the
    // contact self-energy added to Sigma
    if (sl_hartree) sl_hartree->launch(rho_bar.as<cd>(), SC.as<cd>(), contact, st);
    const cd* sc = contact || sl_hartree ? SC.as<cd>() : nullptr;
    const double* hdp = hartree ? hd : nullptr;
    if (sl_x) {
        const bool ref = has_refP && use_refP;
        sl_plan->apply(P, ref && !refP_const ? refP.as<cd>() : nullptr,
                       ref && refP_const ? refP0.as<cd>() : nullptr, Sigma, st);
        launch_fock_assemble(static_cast<int>(nk), nb, h.as<cd>(), Sigma, F, sc, hdp, st);
    } else if (exchange) {
        a.Sigma = Sigma;
        a.ifftshift = false;
        if (packed) {
            a.epilogue = EPI_FOCK_PACKED;
            a.F = F;
            a.hh = G.hhb.as<cd>();
            a.SC = sc;
            a.hdiag = hdp;
            launch_exchange_inv(g, a, st);
        } else {
            a.epilogue = EPI_RAW_FULL;
            launch_exchange_inv(g, a, st);
            launch_fock_assemble(static_cast<int>(nk), nb, h.as<cd>(), Sigma, F, sc, hdp, st);
        }
    } else {
        fill_zero(Sigma, nd * sizeof(cd), st);
        launch_fock_assemble(static_cast<int>(nk), nb, h.as<cd>(), Sigma, F, sc, hdp, st);
    }
}

void GpuKernel::build_fock_graph(const cd* P, cd* Sigma, cd* F, double* hdiag_out,
                                 bool p_hermitian, hipStream_t st) {
    if (!use_graph) {
        build_fock(P, Sigma, F, hdiag_out, p_hermitian, st);
        return;
    }
    for (auto& g : graphs)
        if (g.P == P && g.Sigma == Sigma && g.F == F && g.hdiag == hdiag_out &&
            g.herm == p_hermitian) {
            HFGPU_CHECK(hipGraphLaunch(g.exec, st));
            return;
        }
    // geometry set-up allocates and copies: never inside a capture
    (void)geom(exchange && packed_for(p_hermitian));
    HFGPU_CHECK(hipStreamSynchronize(stream));
    hipGraph_t graph = nullptr;
    HFGPU_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
        build_fock(P, Sigma, F, hdiag_out, p_hermitian, stream);
    } catch (...) {
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HFGPU_CHECK(hipStreamEndCapture(stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HFGPU_CHECK(ie);
    if (graphs.size() >= 4) {
        (void)hipGraphExecDestroy(graphs.front().exec);
        graphs.erase(graphs.begin());
    }
    graphs.push_back(GraphEntry{P, Sigma, F, hdiag_out, p_hermitian, exec});
    HFGPU_CHECK(hipGraphLaunch(exec, st));
}

// ---------------------------------------------------------------------------
// batched Fock builds
// ---------------------------------------------------------------------------

void FockBatch::setup(const std::vector<GpuKernel*>& ks, const std::vector<const cd*>& P_,
                      const std::vector<cd*>& Sigma_, const std::vector<cd*>& F_,
                      const std::vector<double*>& hdiag_, const std::vector<double*>& E_,
                      hipStream_t st) {
    const int S = static_cast<int>(ks.size());
    if (S < 1 || S > 32) throw std::invalid_argument("FockBatch takes 1 to 32 kernels");
    Ks = ks;
    P = P_;
    Sigma = Sigma_;
    F = F_;
    hdiag = hdiag_;
    E = E_;
    groups.clear();
    single.clear();
    nk_max = 0;
    for (GpuKernel* K : Ks) nk_max = std::max(nk_max, static_cast<int>(K->nk));
    // the kind of a kernel's build (build_fock's branches with a Hermitian
    // P): solves of one kind and mesh share the passes
    auto kind_of = [](GpuKernel& K) {
        const bool need_partials = K.hartree || K.contact;
        return std::make_tuple(K.nk1, K.nk2, K.nb, K.dv > 1, K.has_refP, K.refP_const,
                               K.hartree, K.contact, need_partials);
    };
    std::vector<std::pair<decltype(kind_of(*Ks[0])), std::vector<int>>> kinds;
    for (int s = 0; s < S; ++s) {
        GpuKernel& K = *Ks[s];
        bool ok = K.exchange && !K.superlattice() && K.packed_for(true);
        if (ok) {
            const ExchangeGeom& g = *K.geom(true).g;
            ok = g.fast_rows && g.fast_cols;
        }
        if (!ok) {
            single.push_back(s);
            continue;
        }
        const auto key = kind_of(K);
        bool placed = false;
        for (auto& kv : kinds)
            if (kv.first == key) {
                kv.second.push_back(s);
                placed = true;
                break;
            }
        if (!placed) kinds.emplace_back(key, std::vector<int>{s});
    }
    for (auto& kv : kinds) {
        Group gr;
        gr.members = kv.second;
        GpuKernel& K0 = *Ks[gr.members[0]];
        GpuKernel::Geom& G0 = K0.geom(true);
        gr.g = G0.g.get();
        const bool need_partials = K0.hartree || K0.contact;
        gr.zero_hdiag = !need_partials;
        ExchangeArgs& a = gr.a;
        a.vr_scalar = K0.dv <= 1;
        a.small.enabled = need_partials ? 1 : 0;
        a.epilogue = EPI_FOCK_PACKED;
        a.ifftshift = false;
        a.hh = G0.hhb.as<cd>();  // non-null marker; each solve's comes from the table
        std::vector<ExchBatchEntry> tab(gr.members.size());
        for (size_t i = 0; i < gr.members.size(); ++i) {
            const int s = gr.members[i];
            GpuKernel& K = *Ks[s];
            GpuKernel::Geom& G = K.geom(true);
            ExchBatchEntry& e = tab[i];
            e.sv = s;
            e.P = P[s];
            if (K.has_refP) {
                if (K.refP_const) {
                    e.refP0 = K.refP0.as<cd>();
                    e.wrow = K.wrow.as<double>();
                } else {
                    e.refP = K.refP.as<cd>();
                }
            }
            e.T = K.T_ws.as<cd>();
            e.partial = need_partials ? K.partial.as<cd>() : nullptr;
            e.w2d = K.w2d.as<double>();
            e.VR = K.dv > 1 ? G.VRb.as<cd>() : K.VR_full.as<cd>();
            if (need_partials) {
                SmallTerms& sm = e.small;
                sm.enabled = 1;
                sm.include_hartree = K.hartree ? 1 : 0;
                sm.has_contact = K.contact ? 1 : 0;
                sm.HH = K.HH.as<double>();
                sm.n_contact = K.n_contact;
                sm.contact_g = K.cg.as<double>();
                sm.Oi = K.Oi.as<cd>();
                sm.Oj = K.Oj.as<cd>();
                sm.hdiag = hdiag[s];
                sm.SC = K.SC.as<cd>();
                sm.rho_out = nullptr;
            }
            e.Sigma = Sigma[s];
            e.F = F[s];
            e.hh = G.hhb.as<cd>();
            e.SC = K.contact ? K.SC.as<cd>() : nullptr;
            e.hdiag = K.hartree ? hdiag[s] : nullptr;
        }
        gr.tab.alloc(tab.size() * sizeof(ExchBatchEntry));
        h2d(gr.tab.get(), tab.data(), tab.size() * sizeof(ExchBatchEntry), st);
        HFGPU_CHECK(hipStreamSynchronize(st));  // tab is a stack vector
        groups.push_back(std::move(gr));
    }
    std::vector<EnergyBatchEntry> et(static_cast<size_t>(S));
    for (int s = 0; s < S; ++s) {
        GpuKernel& K = *Ks[s];
        EnergyBatchEntry& e = et[s];
        e.h = K.h.as<cd>();
        e.P = P[s];
        e.refP = K.has_refP ? K.refP.as<cd>() : nullptr;
        e.Sigma = Sigma[s];
        e.hdiag = K.hartree ? hdiag[s] : nullptr;
        e.w2d = K.w2d.as<double>();
        e.partial = K.e_partial.as<double>();
        e.E = E[s];
        e.nk = static_cast<int>(K.nk);
    }
    etab.alloc(et.size() * sizeof(EnergyBatchEntry));
    h2d(etab.get(), et.data(), et.size() * sizeof(EnergyBatchEntry), st);
    HFGPU_CHECK(hipStreamSynchronize(st));
}

void FockBatch::build(unsigned active, bool energy, hipStream_t st) {
    const int S = static_cast<int>(Ks.size());
    for (const Group& gr : groups) {
        // a group's launches cover its members; inactive ones return at once
        bool any = false;
        for (int s : gr.members) any = any || ((active >> s) & 1u);
        if (!any) continue;
        const int n = static_cast<int>(gr.members.size());
        if (gr.zero_hdiag)
            for (int s : gr.members)
                if ((active >> s) & 1u) fill_zero(hdiag[s], Ks[s]->nb * sizeof(double), st);
        const ExchBatchEntry* tab = gr.tab.as<ExchBatchEntry>();
        launch_exchange_fwd_batch(*gr.g, gr.a, tab, n, active, st);
        launch_exchange_inv_batch(*gr.g, gr.a, tab, n, active, st);
    }
    for (int s : single)
        if ((active >> s) & 1u)
            Ks[s]->build_fock(P[s], Sigma[s], F[s], hdiag[s], true, st, true, false);
    if (energy)
        launch_energy_batch(etab.as<EnergyBatchEntry>(), S, nk_max, Ks[0]->nb, active, st);
}

void GpuKernel::energy(const cd* P, const cd* Sigma, const double* hdiag_in, double* E,
                       hipStream_t st) {
    launch_energy(static_cast<int>(nk), nb, h.as<cd>(), P, has_refP ? refP.as<cd>() : nullptr,
                  Sigma, hartree ? hdiag_in : nullptr, w2d.as<double>(), e_partial.as<double>(),
                  E, st);
}

// ---------------------------------------------------------------------------
// geometry cache for kernel-less self-energy calls (the counterpart of the
// reference's FftPlanCache, selfenergy.hpp:20-55: 4 entries, FIFO)
// ---------------------------------------------------------------------------

namespace {

struct GeomCache {
    using Ptr = std::shared_ptr<ExchangeGeom>;
    std::mutex mu;
    std::vector<std::pair<std::tuple<int, int, int, bool>, Ptr>> entries;

    // Shared ownership: a caller keeps its geometry alive (device tables
    // included) until it drops the pointer, after its stream has drained,
    // even when another thread evicts the entry or clears the cache meanwhile.
    Ptr get(int nk1, int nk2, int nb, bool packed) {
        std::lock_guard<std::mutex> lock(mu);
        const auto key = std::make_tuple(nk1, nk2, nb, packed);
        for (auto& e : entries)
            if (e.first == key) return e.second;
        while (entries.size() >= 4) entries.erase(entries.begin());
        Ptr g(new ExchangeGeom(), [](ExchangeGeom* x) {
            x->release();
            delete x;
        });
        g->init(nk1, nk2, nb, packed);
        entries.emplace_back(key, g);
        return g;
    }
    void clear() {
        std::lock_guard<std::mutex> lock(mu);
        entries.clear();
    }
    size_t size() {
        std::lock_guard<std::mutex> lock(mu);
        return entries.size();
    }
};

GeomCache& geom_cache() {
    static GeomCache c;
    return c;
}

struct Stream {
    hipStream_t s = nullptr;
    Stream() : s(stream_acquire()) {}
    ~Stream() { (void)stream_sync_release(s); }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    void sync() { HFGPU_CHECK(hipStreamSynchronize(s)); }
};

}  // namespace

void clear_geometry_cache() { geom_cache().clear(); }

// Block map of the given sizes.  exact: the sizes must sum to nb; else
// indices past their sum get block -1 (off-block to everything, itself
// included), which is the reference mask's reading of short size lists.
BlockMap block_map(size_t nb, const size_t* sizes, size_t nsizes, bool exact) {
    if (nb > 32) throw std::invalid_argument("GPU eigh supports 1 <= nb <= 32");
    size_t total = 0;
    for (size_t b = 0; b < nsizes; ++b) total += sizes[b];
    if (exact ? total != nb : total > nb)
        throw std::invalid_argument("block sizes must sum to nb");
    BlockMap bm{};
    for (size_t i = total; i < nb; ++i) bm.blk[i] = -1;
    size_t i = 0;
    for (size_t b = 0; b < nsizes; ++b)
        for (size_t t = 0; t < sizes[b]; ++t, ++i) {
            bm.blk[i] = static_cast<signed char>(b);
            bm.start[i] = static_cast<signed char>(i - t);
        }
    return bm;
}

}  // namespace hfgpu

using namespace hfgpu;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* hfgpu_last_error(void) { return g_last_error.c_str(); }

int hfgpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void hfgpu_clear_plan_cache(void) { clear_geometry_cache(); }

int hfgpu_kernel_create(const hf_kernel_desc* desc, hfgpu_kernel** out) {
    return guarded([&] {
        if (!desc || !out) throw std::invalid_argument("null argument");
        auto* k = new hfgpu_kernel;
        try {
            k->impl = std::make_unique<GpuKernel>(*desc);
        } catch (...) {
            delete k;
            throw;
        }
        *out = k;
    });
}

int hfgpu_kernel_destroy(hfgpu_kernel* k) {
    return guarded([&] { delete k; });
}

int hfgpu_kernel_packable(const hfgpu_kernel* k) {
    if (!k) return 0;
    return (k->impl->hcp || k->impl->dense_packable) ? 1 : 0;
}

int hfgpu_kernel_autotune(hfgpu_kernel* k, int p_hermitian, int* shape, double* best_us) {
    return guarded([&] {
        if (!k || !shape || !best_us) throw std::invalid_argument("null argument");
        *best_us = k->impl->autotune(p_hermitian != 0, shape);
    });
}

int hfgpu_kernel_dataflow_floor(hfgpu_kernel* k, int p_hermitian, int reps, double* us,
                                double* bytes) {
    return guarded([&] {
        if (!k || !us || !bytes) throw std::invalid_argument("null argument");
        *us = k->impl->dataflow_floor(p_hermitian != 0, reps, bytes);
    });
}

int hfgpu_eigh_autotune(size_t nmat, size_t nb, int* mpw, double* best_us) {
    return guarded([&] {
        if (!mpw || !best_us) throw std::invalid_argument("null argument");
        if (nmat == 0 || nmat > static_cast<size_t>(INT_MAX))
            throw std::invalid_argument("hfgpu_eigh_autotune: nmat must be in 1 .. 2^31 - 1");
        Stream st;
        *best_us = eigh_autotune(static_cast<int>(nmat), static_cast<int>(nb), st.s, mpw);
        st.sync();
    });
}

int hfgpu_selfenergy_fft_full(const double* VR, const double* P, double* sigma, size_t nk1,
                              size_t nk2, size_t nb, size_t dv1, size_t dv2,
                              int apply_ifftshift, int hcp) {
    return guarded([&] {
        if (hcp && (dv1 != 1 || dv2 != 1))
            throw std::invalid_argument(
                "hermitian_channel_packing requires VR with shape (...,1,1)");
        if (!((dv1 == 1 && dv2 == 1) || (dv1 == nb && dv2 == nb)))
            throw std::invalid_argument("VR shape must be (..., 1, 1) or (..., nb, nb).");
        const size_t nk = nk1 * nk2, nd = nk * nb * nb;
        if (nd == 0) return;
        // held past the stream's drain (st is destroyed first)
        const auto gp = geom_cache().get(static_cast<int>(nk1), static_cast<int>(nk2),
                                         static_cast<int>(nb), hcp != 0);
        const ExchangeGeom& g = *gp;
        const bool dense = dv1 > 1;
        Stream st;
        DevBuf dP(nd * sizeof(cd)), dS(nd * sizeof(cd)), dT(g.tile_elems() * sizeof(cd)),
            dV(nk * dv1 * dv2 * sizeof(cd)), dVb(dense ? g.tile_elems() * sizeof(cd) : 0);
        h2d(dP.get(), P, nd * sizeof(cd), st.s);
        h2d(dV.get(), VR, nk * dv1 * dv2 * sizeof(cd), st.s);
        if (dense) launch_pack_vr(g, dV.as<cd>(), dVb.as<cd>(), st.s);
        ExchangeArgs a;
        a.P = dP.as<cd>();
        a.VR = dense ? dVb.as<cd>() : dV.as<cd>();
        a.vr_scalar = !dense;
        a.T = dT.as<cd>();
        a.epilogue = hcp ? EPI_RAW_PACKED : EPI_RAW_FULL;
        a.ifftshift = apply_ifftshift != 0;
        a.Sigma = dS.as<cd>();
        launch_exchange_fwd(g, a, st.s);
        launch_exchange_inv(g, a, st.s);
        d2h(sigma, dS.get(), nd * sizeof(cd), st.s);
        st.sync();
    });
}

int hfgpu_eigh_batched(const double* M, double* w, double* V, size_t nmat, size_t nb) {
    return guarded([&] {
        if (nmat == 0 || nb == 0) return;
        Stream st;
        const size_t n = nmat * nb * nb;
        DevBuf dM(n * sizeof(cd)), dV(n * sizeof(cd)), dw(nmat * nb * sizeof(double));
        h2d(dM.get(), M, n * sizeof(cd), st.s);
        launch_eigh(static_cast<int>(nmat), static_cast<int>(nb), 0, dM.as<cd>(), dw.as<double>(),
                    dV.as<cd>(), st.s);
        d2h(w, dw.get(), nmat * nb * sizeof(double), st.s);
        d2h(V, dV.get(), n * sizeof(cd), st.s);
        st.sync();
    });
}


int hfgpu_eigh_block_sizes(const double* M, double* w, double* V, size_t nmat, size_t nb,
                           const size_t* sizes, size_t nsizes, int sort) {
    return guarded([&] {
        const BlockMap bm = block_map(nb, sizes, nsizes);
        if (nmat == 0 || nb == 0) return;
        Stream st;
        const size_t n = nmat * nb * nb;
        DevBuf dM(n * sizeof(cd)), dV(n * sizeof(cd)), dw(nmat * nb * sizeof(double));
        h2d(dM.get(), M, n * sizeof(cd), st.s);
        launch_eigh_blocks(static_cast<int>(nmat), static_cast<int>(nb), 0, dM.as<cd>(),
                           dw.as<double>(), dV.as<cd>(), bm, sort != 0, st.s);
        d2h(w, dw.get(), nmat * nb * sizeof(double), st.s);
        d2h(V, dV.get(), n * sizeof(cd), st.s);
        st.sync();
    });
}

int hfgpu_max_offblock_sizes(const double* M, size_t nmat, size_t nb, const size_t* sizes,
                             size_t nsizes, double* out) {
    return guarded([&] {
        const BlockMap bm = block_map(nb, sizes, nsizes, false);
        *out = 0.0;
        if (nmat == 0 || nb == 0) return;
        Stream st;
        const size_t n = nmat * nb * nb;
        DevBuf dM(n * sizeof(cd)), dp(2 * kMaxOffblockBlocks * sizeof(double)),
            dout(2 * sizeof(double));
        h2d(dM.get(), M, n * sizeof(cd), st.s);
        launch_max_offblock(static_cast<int>(nmat), static_cast<int>(nb), dM.as<cd>(), bm,
                            dp.as<double>(), dout.as<double>(), st.s);
        d2h(out, dout.get(), sizeof(double), st.s);
        st.sync();
    });
}

int hfgpu_find_mu_bisection(const double* bands, const double* weights, size_t nk, size_t nb,
                            double n_e, double T, int maxiter, double* mu) {
    return guarded([&] {
        if (nk * nb == 0) throw std::invalid_argument("bands must not be empty");
        Stream st;
        DevBuf de(nk * nb * sizeof(double)), dw(nk * sizeof(double)), ds(sizeof(MuState)),
            dp(kMuPartialDoubles * sizeof(double));
        h2d(de.get(), bands, nk * nb * sizeof(double), st.s);
        h2d(dw.get(), weights, nk * sizeof(double), st.s);
        fill_zero(ds.get(), sizeof(MuState), st.s);
        launch_mu_bisection(de.as<double>(), dw.as<double>(), static_cast<int>(nk),
                            static_cast<int>(nb), n_e, T, maxiter, ds.as<MuState>(),
                            dp.as<double>(), st.s);
        MuState hs;
        d2h(&hs, ds.get(), sizeof(MuState), st.s);
        st.sync();
        *mu = hs.mu;
    });
}

int hfgpu_find_mu_newton(const double* bands, const double* weights, size_t nk, size_t nb,
                         double n_e, double T, int maxiter, double* mu) {
    return guarded([&] {
        if (nk * nb == 0) throw std::invalid_argument("bands must not be empty");
        Stream st;
        DevBuf de(nk * nb * sizeof(double)), dw(nk * sizeof(double)), dm(sizeof(double));
        h2d(de.get(), bands, nk * nb * sizeof(double), st.s);
        h2d(dw.get(), weights, nk * sizeof(double), st.s);
        launch_mu_newton_py(de.as<double>(), dw.as<double>(), static_cast<int>(nk),
                            static_cast<int>(nb), n_e, T, maxiter, dm.as<double>(), st.s);
        d2h(mu, dm.get(), sizeof(double), st.s);
        st.sync();
    });
}

int hfgpu_solve_mu_inloop(const double* eps, const double* w_norm, size_t nk, size_t nb,
                          double n_target_norm, double mu0, double T, int maxiter, double* mu) {
    return guarded([&] {
        if (nk * nb == 0) throw std::invalid_argument("eps must not be empty");
        Stream st;
        DevBuf de(nk * nb * sizeof(double)), dw(nk * sizeof(double)), ds(sizeof(MuState)),
            dp(kMuPartialDoubles * sizeof(double));
        h2d(de.get(), eps, nk * nb * sizeof(double), st.s);
        h2d(dw.get(), w_norm, nk * sizeof(double), st.s);
        fill_zero(ds.get(), sizeof(MuState), st.s);
        launch_mu_newton(de.as<double>(), dw.as<double>(), static_cast<int>(nk),
                         static_cast<int>(nb), n_target_norm, mu0, false, T, maxiter,
                         ds.as<MuState>(), dp.as<double>(), st.s);
        MuState hs;
        d2h(&hs, ds.get(), sizeof(MuState), st.s);
        st.sync();
        *mu = hs.mu;
    });
}

static int build_fock_apply_impl(hfgpu_kernel* k, const double* P, double* Sigma, double* Hh,
                                 double* F, double* E, hf_project_fn project, void* user,
                                 bool fp32_ws) {
    return guarded([&] {
        if (!k) throw std::invalid_argument("null kernel");
        GpuKernel& K = *k->impl;
        hipStream_t s = K.stream;
        const size_t bytes = K.nd * sizeof(cd);
        DevBuf dP(bytes), dS(bytes), dF(bytes), dH(K.nb * sizeof(double)), dE(sizeof(double));
        h2d(dP.get(), P, bytes, s);
        const bool herm = K.dense_packable ? host_is_hermitian(P, K.nk, K.nb) : false;
        K.build_fock(dP.as<cd>(), dS.as<cd>(), dF.as<cd>(), dH.as<double>(), herm, s, true,
                     fp32_ws);
        // hf_energy (fock.hpp:324-351) with the dense Hartree array equals the
        // compact-diagonal form exactly: 0.5 * (s + h) == 0.5 * s + 0.5 * h.
        K.energy(dP.as<cd>(), dS.as<cd>(), dH.as<double>(), dE.as<double>(), s);
        std::vector<double> hd(K.nb);
        d2h(Sigma, dS.get(), bytes, s);
        d2h(F, dF.get(), bytes, s);
        d2h(hd.data(), dH.get(), K.nb * sizeof(double), s);
        d2h(E, dE.get(), sizeof(double), s);
        HFGPU_CHECK(hipStreamSynchronize(s));
        std::memset(Hh, 0, bytes);
        if (K.hartree)
            for (size_t kk = 0; kk < K.nk; ++kk)
                for (int i = 0; i < K.nb; ++i) Hh[2 * (kk * K.nb2 + i * K.nb + i)] = hd[i];
        if (project) {
            if (project(user, F, K.nk1, K.nk2, K.nb) != 0)
                throw std::runtime_error("project_fn failed");
            host_hermitize(F, K.nk, K.nb);
        }
    });
}

int hfgpu_build_fock_apply(hfgpu_kernel* k, const double* P, double* Sigma, double* Hh,
                           double* F, double* E, hf_project_fn project, void* user) {
    return build_fock_apply_impl(k, P, Sigma, Hh, F, E, project, user, false);
}

int hfgpu_build_fock_mixed(hfgpu_kernel* k, const double* P, double* Sigma, double* Hh,
                           double* F, double* E, hf_project_fn project, void* user) {
    return build_fock_apply_impl(k, P, Sigma, Hh, F, E, project, user, true);
}

int hfgpu_build_fock_mixed_device(hfgpu_kernel* k, const void* dP, void* dSigma, void* dF,
                                  void* dHdiag, int p_hermitian, void* stream) {
    return guarded([&] {
        if (!k) throw std::invalid_argument("null kernel");
        GpuKernel& K = *k->impl;
        K.build_fock(static_cast<const cd*>(dP), static_cast<cd*>(dSigma), static_cast<cd*>(dF),
                     static_cast<double*>(dHdiag), p_hermitian != 0,
                     static_cast<hipStream_t>(stream), true, true);
    });
}

int hfgpu_kernel_mixed_capable(hfgpu_kernel* k, int p_hermitian) {
    if (!k) return 0;
    int r = 0;
    const int rc = guarded([&] { r = k->impl->fp32_capable(p_hermitian != 0) ? 1 : 0; });
    return rc == HF_OK ? r : 0;
}

int hfgpu_build_fock_device(hfgpu_kernel* k, const void* dP, void* dSigma, void* dF,
                            void* dHdiag, int p_hermitian, void* stream) {
    return guarded([&] {
        if (!k) throw std::invalid_argument("null kernel");
        GpuKernel& K = *k->impl;
        hipStream_t s = static_cast<hipStream_t>(stream);
        K.build_fock_graph(static_cast<const cd*>(dP), static_cast<cd*>(dSigma),
                           static_cast<cd*>(dF), static_cast<double*>(dHdiag), p_hermitian != 0,
                           s);
    });
}

int hfgpu_energy_device(hfgpu_kernel* k, const void* dP, const void* dSigma,
                        const void* dHdiag, void* dE, void* stream) {
    return guarded([&] {
        if (!k) throw std::invalid_argument("null kernel");
        GpuKernel& K = *k->impl;
        hipStream_t s = static_cast<hipStream_t>(stream);
        K.energy(static_cast<const cd*>(dP), static_cast<const cd*>(dSigma),
                 static_cast<const double*>(dHdiag), static_cast<double*>(dE), s);
    });
}

int hfgpu_hf_energy(const double* h, const double* P, const double* Sigma, const double* H,
                    size_t n_H, const double* w, const double* refP, size_t nk, size_t nb,
                    double* E) {
    return guarded([&] {
        if (n_H != 1 && n_H != nk) throw std::invalid_argument("H must be broadcast (1) or per k");
        if (nk == 0 || nb == 0) {
            *E = 0.0;
            return;
        }
        Stream st;
        const size_t nd = nk * nb * nb, hb = n_H * nb * nb;
        DevBuf dh(nd * sizeof(cd)), dP(nd * sizeof(cd)), dS(nd * sizeof(cd)),
            dH(hb * sizeof(cd)), dw(nk * sizeof(double)), dr(refP ? nd * sizeof(cd) : 0),
            dpart((cdiv(nk, 4) + 1) * sizeof(double)), dE(sizeof(double));
        h2d(dh.get(), h, nd * sizeof(cd), st.s);
        h2d(dP.get(), P, nd * sizeof(cd), st.s);
        h2d(dS.get(), Sigma, nd * sizeof(cd), st.s);
        h2d(dH.get(), H, hb * sizeof(cd), st.s);
        h2d(dw.get(), w, nk * sizeof(double), st.s);
        if (refP) h2d(dr.get(), refP, nd * sizeof(cd), st.s);
        launch_energy_fullH(static_cast<int>(nk), static_cast<int>(nb), dh.as<cd>(), dP.as<cd>(),
                            refP ? dr.as<cd>() : nullptr, dS.as<cd>(), dH.as<cd>(),
                            n_H == 1 ? 0 : nb * nb, dw.as<double>(), dpart.as<double>(),
                            dE.as<double>(), st.s);
        d2h(E, dE.get(), sizeof(double), st.s);
        st.sync();
    });
}

int hfgpu_occupation_entropy(const double* p, const double* w_norm, size_t nk, size_t nb,
                             double* S) {
    return guarded([&] {
        if (nk * nb == 0) {
            *S = 0.0;
            return;
        }
        Stream st;
        DevBuf dp(nk * nb * sizeof(double)), dw(nk * sizeof(double)),
            dpart(1025 * sizeof(double)), dS(sizeof(double));
        h2d(dp.get(), p, nk * nb * sizeof(double), st.s);
        h2d(dw.get(), w_norm, nk * sizeof(double), st.s);
        launch_entropy(dp.as<double>(), dw.as<double>(), static_cast<int>(nk),
                       static_cast<int>(nb), dpart.as<double>(), dS.as<double>(), st.s);
        d2h(S, dS.get(), sizeof(double), st.s);
        st.sync();
    });
}

int hfgpu_density_from_fock(const double* F, const double* weights, size_t nk, size_t nb,
                            double n_e, double T, int maxiter, double* P, double* mu) {
    return guarded([&] {
        if (nk * nb == 0) throw std::invalid_argument("F must not be empty");
        Stream st;
        const size_t nd = nk * nb * nb, nkb = nk * nb;
        DevBuf dF(nd * sizeof(cd)), dV(nd * sizeof(cd)), dP(nd * sizeof(cd)),
            de(nkb * sizeof(double)), docc(nkb * sizeof(double)), dw(nk * sizeof(double)),
            ds(sizeof(MuState)), dpart(kMuPartialDoubles * sizeof(double));
        h2d(dF.get(), F, nd * sizeof(cd), st.s);
        h2d(dw.get(), weights, nk * sizeof(double), st.s);
        fill_zero(ds.get(), sizeof(MuState), st.s);
        launch_eigh(static_cast<int>(nk), static_cast<int>(nb), 0, dF.as<cd>(), de.as<double>(),
                    dV.as<cd>(), st.s);
        launch_mu_bisection(de.as<double>(), dw.as<double>(), static_cast<int>(nk),
                            static_cast<int>(nb), n_e, T, maxiter, ds.as<MuState>(),
                            dpart.as<double>(), st.s);
        launch_occupations(de.as<double>(), nkb, ds.as<MuState>(), T, 0, docc.as<double>(), st.s);
        launch_density(static_cast<int>(nk), static_cast<int>(nb), dV.as<cd>(), docc.as<double>(),
                       dP.as<cd>(), true, st.s);
        MuState hs;
        d2h(P, dP.get(), nd * sizeof(cd), st.s);
        d2h(&hs, ds.get(), sizeof(MuState), st.s);
        st.sync();
        *mu = hs.mu;
    });
}

int hfgpu_cayley_spectral_setup(const double* d, double* V, double* lam, size_t nk, size_t nb) {
    return guarded([&] {
        if (nk == 0 || nb == 0) return;
        Stream st;
        const size_t n = nk * nb * nb;
        DevBuf dd(n * sizeof(cd)), dV(n * sizeof(cd)), dl(nk * nb * sizeof(double));
        h2d(dd.get(), d, n * sizeof(cd), st.s);
        launch_eigh(static_cast<int>(nk), static_cast<int>(nb), 1, dd.as<cd>(), dl.as<double>(),
                    dV.as<cd>(), st.s);
        d2h(V, dV.get(), n * sizeof(cd), st.s);
        d2h(lam, dl.get(), nk * nb * sizeof(double), st.s);
        st.sync();
    });
}

int hfgpu_cayley_unitary_from_spectrum(const double* V, const double* lam, double tau,
                                       double* U, size_t nk, size_t nb) {
    return guarded([&] {
        if (nk == 0 || nb == 0) return;
        Stream st;
        const size_t n = nk * nb * nb;
        DevBuf dV(n * sizeof(cd)), dU(n * sizeof(cd)), dl(nk * nb * sizeof(double));
        h2d(dV.get(), V, n * sizeof(cd), st.s);
        h2d(dl.get(), lam, nk * nb * sizeof(double), st.s);
        launch_cayley_unitary(static_cast<int>(nk), static_cast<int>(nb), dV.as<cd>(),
                              dl.as<double>(), tau, dU.as<cd>(), st.s);
        d2h(U, dU.get(), n * sizeof(cd), st.s);
        st.sync();
    });
}

int hfgpu_diag_UFU_from_spectrum(const double* V, const double* Ft_eig, const double* lam,
                                 double tau, double* diag, size_t nk, size_t nb) {
    return guarded([&] {
        if (nk == 0 || nb == 0) return;
        Stream st;
        const size_t n = nk * nb * nb;
        DevBuf dV(n * sizeof(cd)), dF(n * sizeof(cd)), dl(nk * nb * sizeof(double)),
            dd(nk * nb * sizeof(double));
        h2d(dV.get(), V, n * sizeof(cd), st.s);
        h2d(dF.get(), Ft_eig, n * sizeof(cd), st.s);
        h2d(dl.get(), lam, nk * nb * sizeof(double), st.s);
        launch_diag_ufu(static_cast<int>(nk), static_cast<int>(nb), dV.as<cd>(), dF.as<cd>(),
                        dl.as<double>(), tau, dd.as<double>(), st.s);
        d2h(diag, dd.get(), nk * nb * sizeof(double), st.s);
        st.sync();
    });
}

int hfgpu_resample_kgrid_2d(const double* x, double* out, size_t nk_old, size_t nk_new,
                            size_t inner) {
    return guarded([&] {
        const size_t n_in = nk_old * nk_old * inner, n_out = nk_new * nk_new * inner;
        if (nk_old == nk_new) {
            std::memcpy(out, x, n_in * sizeof(cd));
            return;
        }
        if (n_out == 0) return;
        if (nk_old == 0) throw std::invalid_argument("resample_kgrid_2d: empty input grid");
        Stream st;
        DevBuf dx(n_in * sizeof(cd)), dt(nk_new * nk_old * inner * sizeof(cd)),
            dy(n_out * sizeof(cd));
        h2d(dx.get(), x, n_in * sizeof(cd), st.s);
        // axis 0: (nk_old, nk_old*inner) -> (nk_new, nk_old*inner)
        launch_resample_axis(dx.as<cd>(), dt.as<cd>(), static_cast<int>(nk_old),
                             static_cast<int>(nk_new), nk_old * inner, 1, st.s);
        // axis 1 of each of the nk_new rows
        launch_resample_axis(dt.as<cd>(), dy.as<cd>(), static_cast<int>(nk_old),
                             static_cast<int>(nk_new), inner, nk_new, st.s);
        d2h(out, dy.get(), n_out * sizeof(cd), st.s);
        st.sync();
    });
}

}  // extern "C"
