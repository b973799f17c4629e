// Host-side core of libhfgpu: device buffers and the device-resident HF
// kernel (the GPU counterpart of the reference's HFKernel, kernel.hpp:10-84).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/hf_desc.h"
#include "common.h"
#include "exchange.h"
#include "kernels.h"

namespace hfgpu {
struct SlPlan;
struct SlHartree;
}  // namespace hfgpu

namespace hfgpu {

// Owning device allocation.
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void alloc(size_t bytes) {
        release();
        if (bytes) p_ = dev_alloc(bytes);
        n_ = bytes;
    }
    void ensure(size_t bytes) { if (bytes > n_) alloc(bytes); }
    void release() {
        if (p_) dev_free(p_, n_);
        p_ = nullptr;
        n_ = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p_); }
    void* get() const { return p_; }
    size_t size() const { return n_; }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

inline void h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes) HFGPU_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
}
inline void d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes) HFGPU_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
}
inline void d2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes) HFGPU_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
}

// Zero-fill through a kernel (safe under graph capture, unlike memset nodes).
void fill_zero(void* p, size_t bytes, hipStream_t st);
// Device-side elementwise helpers used by the solvers.
void launch_axpy_c(size_t n, double alpha, const cd* x, cd* y, hipStream_t st);  // y += a x
void launch_sub_c(size_t n, const cd* a, const cd* b, cd* out, hipStream_t st);    // a - b

// Device-resident HF kernel.
struct GpuKernel {
    int nk1 = 0, nk2 = 0, nb = 0, dv = 1, n_contact = 0;
    size_t nk = 0, nb2 = 0, nd = 0;
    bool hartree = false, exchange = true, hcp = false, has_refP = false, contact = false;
    bool dense_packable = false;  // (nb, nb) V with a real, band-symmetric kernel
    double T = 0.0, weight_sum = 0.0;
    std::vector<double> w_host;   // (nk) weights, host copy
    std::vector<double> cg_host;  // (n_contact) contact strengths, host copy
    hipStream_t stream = nullptr;

    DevBuf h, refP, w2d, HH, cg, Oi, Oj;
    DevBuf VR_full;  // (nk, dv*dv) as given
    // k-independent reference density (refP_k identical for every k): its
    // transform is a single q = 0 term, so pass A reads P alone
    bool refP_const = false;
    DevBuf refP0, wrow;  // (nb, nb) and (nk1) row sums of w2d

    // An exchange geometry with its tile-major constant operands.
    struct Geom {
        std::unique_ptr<ExchangeGeom> g;
        DevBuf VRb;  // VR in pass B order (dense V only)
        DevBuf hhb;  // herm(h) in pass C order (packed geometry only)
        DevBuf VR32;  // VRb (dense V) or VR (scalar V) as complex float, on first use
    };
    Geom geo_full, geo_packed;

    // Fock workspaces (T32_ws: the mixed-precision build's, on first use)
    DevBuf T_ws, T32_ws, partial, SC, hdiag, e_partial, e_out;

    // Full-channel geometries of the adjoint's plain transforms (nb^2
    // channels, and one channel for a scalar interaction), made on first use.
    std::unique_ptr<ExchangeGeom> adj_geo, adj_geo_v;

    // Superlattice mode (kernel.hpp:36-69): the exchange on the extended
    // momentum grid in place of the mesh exchange, and the superlattice
    // Hartree shift in place of the mesh Hartree diagonal (hartree then
    // stands for the mesh Hartree only).
    std::unique_ptr<SlPlan> sl_plan;       // superlattice exchange, nullable
    std::unique_ptr<SlHartree> sl_hartree;  // superlattice Hartree shift, nullable
    DevBuf rho_bar;                          // (nb, nb) sum_k w_k (P - refP)_k
    bool superlattice() const { return sl_plan || sl_hartree; }

    explicit GpuKernel(const hf_kernel_desc& d);
    ~GpuKernel();

    // Launch-shape autotuner of the two-step exchange passes: coordinate
    // descent over wavefronts per tile (split) and tiles per workgroup,
    // timing whole Fock builds with HIP events; the best shape is kept in
    // the geometry and in a per-process cache keyed by the build's shape.
    // Launch shapes only move tasks between lanes: results are unchanged.
    // Returns the best microseconds per build (0 when nothing is tunable).
    double autotune(bool p_hermitian, int shape_out[5]);
    // Three dependent streaming launches with the byte volumes of this
    // kernel's two-step Fock build (no arithmetic): microseconds per build
    // over reps builds, and the bytes one build moves.
    double dataflow_floor(bool p_hermitian, int reps, double* bytes_per_build);
    bool autotune_env = false;  // HFGPU_AUTOTUNE=1: tune each new geometry

    bool packed_for(bool p_hermitian) const { return hcp || (dense_packable && p_hermitian); }
    Geom& geom(bool packed);

    // build_fock_compact on device: Sigma, F, hdiag (nb) from P.  use_refP =
    // false drops the reference density (the linear response to a density
    // variation, solver_rtr.hpp:32-39).  fp32_ws: the exchange workspace and
    // interaction held in complex float (the mixed-precision build) on
    // meshes of the two-step transforms; FP64 elsewhere.
    void build_fock(const cd* P, cd* Sigma, cd* F, double* hdiag_out, bool p_hermitian,
                    hipStream_t st, bool use_refP = true, bool fp32_ws = false);
    // true when build_fock(..., fp32_ws = true) runs the FP32 workspace
    bool fp32_capable(bool p_hermitian);
    // hf_energy_with_hartree_diag on device into E (one double).
    void energy(const cd* P, const cd* Sigma, const double* hdiag_in, double* E, hipStream_t st);

    // build_fock replayed from a captured hipGraph (one per pointer set,
    // small FIFO) when use_graph, else launched eagerly.
    bool use_graph = false;
    void build_fock_graph(const cd* P, cd* Sigma, cd* F, double* hdiag_out, bool p_hermitian,
                          hipStream_t st);
    struct GraphEntry {
        const void *P, *Sigma, *F, *hdiag;
        bool herm;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
};

// build_fock (+ energy) of several kernels in one launch sequence: the
// exchange passes of solves on one two-step geometry take one launch each
// (their operands through a device table, ExchBatchEntry) and the energies
// one pair of launches for all solves.  Solves whose kernels cannot share
// the passes (another mesh, packing, interaction or reference-density kind,
// superlattice, Stockham meshes) are grouped by geometry; a kernel outside
// every group builds on its own.  Each solve's Sigma, F, hdiag and energy
// equal its single build_fock / energy bit for bit.
struct FockBatch {
    std::vector<GpuKernel*> Ks;
    std::vector<const cd*> P;
    std::vector<cd*> Sigma, F;
    std::vector<double*> hdiag, E;
    struct Group {
        std::vector<int> members;  // solve indices
        ExchangeArgs a;            // flags of the group's passes
        const ExchangeGeom* g = nullptr;
        DevBuf tab;                // members.size() ExchBatchEntry
        bool zero_hdiag = false;   // no small terms: hdiag is zero
    };
    std::vector<Group> groups;
    std::vector<int> single;       // solves built one by one
    DevBuf etab;                   // S EnergyBatchEntry
    int nk_max = 0;
    // Device pointers per solve; E[s] receives solve s's energy.  Uploads
    // the tables on st (and waits for them).
    void setup(const std::vector<GpuKernel*>& ks, const std::vector<const cd*>& P_,
               const std::vector<cd*>& Sigma_, const std::vector<cd*>& F_,
               const std::vector<double*>& hdiag_, const std::vector<double*>& E_,
               hipStream_t st);
    // Fock builds of the solves whose bit in active is set (Hermitian P, the
    // reference density applied), then their energies when energy.
    void build(unsigned active, bool energy, hipStream_t st);
};

// Vector-Jacobian product of build_fock (fock.hpp:34-321) at density P for
// cotangents on its outputs Sigma, Hh and F (nullable = zero), in the
// convention dL = Re sum conj(g) dX: gP (nd), gVq (nk * dv * dv, the
// gradient with respect to the interaction coulomb_q before the VR
// transform, complex; its real part for a real Vq), gh (nd), gHH (nb * nb,
// real).  All device pointers on stream st; gVq, gh, gHH nullable.
void fock_vjp(GpuKernel& K, const cd* P, const cd* Gs, const cd* Gh, const cd* GF, cd* gP,
              cd* gVq, cd* gh, double* gHH, hipStream_t st);

// Block map of block sizes summing to nb (exact) or at most nb (indices past
// the last block belong to no block); throws std::invalid_argument otherwise.
BlockMap block_map(size_t nb, const size_t* sizes, size_t nsizes, bool exact = true);

// Last error message of the calling thread (hfgpu_last_error).
extern thread_local std::string g_last_error;

// Run fn, mapping exceptions to hf_status codes: std::invalid_argument ->
// HF_EINVAL, DeviceError -> HF_EDEVICE, anything else -> HF_ERUNTIME.  On a
// failure the device is drained before returning, so workspaces that the
// unwinding returned to the memory pool are idle before anyone reuses them.
template <class Fn>
int guarded(Fn&& fn) {
    int rc = HF_OK;
    try {
        fn();
        return HF_OK;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        rc = HF_EINVAL;
    } catch (const DeviceError& e) {
        g_last_error = e.what();
        rc = HF_EDEVICE;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        rc = HF_ERUNTIME;
    }
    (void)hipDeviceSynchronize();
    return rc;
}

// True when M (nk, nb, nb) equals its conjugate transpose bit for bit.
bool host_is_hermitian(const double* M, size_t nk, size_t nb);
// In-place hermitisation on the host (linalg.hpp:20-33).
void host_hermitize(double* M, size_t nk, size_t nb);

}  // namespace hfgpu

struct hfgpu_kernel {
    std::unique_ptr<hfgpu::GpuKernel> impl;
};
