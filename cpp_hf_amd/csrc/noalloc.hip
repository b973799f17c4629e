// density_from_fock on device pointers with no allocation: the eigenvectors,
// eigenvalues, occupations and the chemical-potential workspace live in a
// fixed device array of this library instead of memory the call allocates.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  Every
// call on a GPU shares that GPU's array, so calls that overlap on one GPU
// (two streams, two threads) race; the array keeps the last call's contents.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/hfgpu.h"
#include "hfgpu_core.h"
#include "kernels.h"

namespace hfgpu {
namespace {

// 16 MiB: a 96 x 96 mesh with 8 bands (nk nb^2 = 589824) fits with room.
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Layout: MuState (256 B), the mu workspace (kMuPartialDoubles doubles),
// then eigenvalues and occupations (nk nb doubles each) and eigenvectors
// (nk nb^2 complex), each 256-byte aligned.
constexpr size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
constexpr size_t kStateOff = 0;
constexpr size_t kWsOff = 256;
constexpr size_t kEpsOff = kWsOff + align256(kMuPartialDoubles * sizeof(double));

size_t scratch_needed(size_t nk, size_t nb) {
    const size_t nkb = nk * nb, nd = nkb * nb;
    return kEpsOff + 2 * align256(nkb * sizeof(double)) + nd * sizeof(cd);
}

// Device address of this device's copy of the array, looked up once per
// device with that device current (the array's name is a host-side stand-in,
// not a device address).
unsigned char* scratch_ptr() {
    static std::mutex mu;
    static std::array<void*, 64> ptrs{};
    int dev = 0;
    HFGPU_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= static_cast<int>(ptrs.size()))
        throw std::runtime_error("device index outside the scratch table");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        HFGPU_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw DeviceError("hipGetSymbolAddress returned no address for the scratch array");
        ptrs[dev] = p;
    }
    return static_cast<unsigned char*>(ptrs[dev]);
}

}  // namespace
}  // namespace hfgpu

using namespace hfgpu;

int hfgpu_density_from_fock_noalloc(const double* dF, const double* dweights, size_t nk,
                                    size_t nb, double n_e, double T, int maxiter, double* dP,
                                    double* dmu, void* stream) {
    return guarded([&] {
        if (nk * nb == 0) throw std::invalid_argument("F must not be empty");
        if (!dF || !dweights || !dP || !dmu) throw std::invalid_argument("null argument");
        if (nb > 32) throw std::invalid_argument("GPU eigh supports 1 <= nb <= 32");
        const size_t need = scratch_needed(nk, nb);
        if (need > kScratchBytes)
            throw std::invalid_argument(
                "density_from_fock_noalloc: nk * nb^2 = " + std::to_string(nk * nb * nb) +
                " needs " + std::to_string(need) + " bytes of scratch, more than the " +
                std::to_string(kScratchBytes) + " bytes this library reserves");
        const hipStream_t st = static_cast<hipStream_t>(stream);
        unsigned char* base = scratch_ptr();
        const size_t nkb = nk * nb;
        auto* state = reinterpret_cast<MuState*>(base + kStateOff);
        auto* ws = reinterpret_cast<double*>(base + kWsOff);
        auto* eps = reinterpret_cast<double*>(base + kEpsOff);
        auto* occ = reinterpret_cast<double*>(base + kEpsOff + align256(nkb * sizeof(double)));
        auto* V = reinterpret_cast<cd*>(base + kEpsOff + 2 * align256(nkb * sizeof(double)));
        // the bracket state starts from zero (a fill kernel, safe under capture)
        fill_zero(state, sizeof(MuState), st);
        launch_eigh(static_cast<int>(nk), static_cast<int>(nb), 0,
                    reinterpret_cast<const cd*>(dF), eps, V, st);
        launch_mu_bisection(eps, dweights, static_cast<int>(nk), static_cast<int>(nb), n_e, T,
                            maxiter, state, ws, st);
        launch_occupations(eps, nkb, state, T, 0, occ, st);
        launch_density(static_cast<int>(nk), static_cast<int>(nb), V, occ, reinterpret_cast<cd*>(dP),
                       true, st);
        HFGPU_CHECK(hipMemcpyAsync(dmu, &state->mu, sizeof(double), hipMemcpyDeviceToDevice, st));
    });
}
