// Shared device/host helpers for libhfgpu (gfx950, FP64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace hfgpu {

// Interleaved complex double, layout-compatible with numpy complex128.
struct cd {
    double x, y;
};

__host__ __device__ __forceinline__ cd mk(double x, double y) { return cd{x, y}; }
__host__ __device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__host__ __device__ __forceinline__ cd csub(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__host__ __device__ __forceinline__ cd cmul(cd a, cd b) {
    return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ __forceinline__ cd cconj(cd a) { return cd{a.x, -a.y}; }
__host__ __device__ __forceinline__ cd cscale(cd a, double s) { return cd{a.x * s, a.y * s}; }
__host__ __device__ __forceinline__ cd cneg(cd a) { return cd{-a.x, -a.y}; }
// a * conj(b)
__host__ __device__ __forceinline__ cd cmulc(cd a, cd b) {
    return cd{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
__host__ __device__ __forceinline__ double cnorm(cd a) { return a.x * a.x + a.y * a.y; }

// Streaming store of one complex value (one 16-byte non-temporal store): for
// data the writing kernel never reads back, so it drains during the kernel
// instead of sitting dirty in L2 until the kernel-end write-back.
typedef double cd_vec __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream(cd* p, cd v) {
    cd_vec t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<cd_vec*>(p));
}

// Interleaved complex float: the FP32 storage of the exchange workspace and
// interaction in the mixed-precision Fock build (arithmetic stays FP64).
struct alignas(8) cf {  // 8-byte aligned: one dwordx2 access per element
    float x, y;
};

// workspace loads and streaming stores for either storage precision
__device__ __forceinline__ cd ld_ws(const cd* p) { return *p; }
__device__ __forceinline__ cd ld_ws(const cf* p) {
    const cf v = *p;
    return cd{static_cast<double>(v.x), static_cast<double>(v.y)};
}
typedef float cf_vec __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream(cf* p, cd v) {
    cf_vec t;
    t.x = static_cast<float>(v.x);
    t.y = static_cast<float>(v.y);
    __builtin_nontemporal_store(t, reinterpret_cast<cf_vec*>(p));
}

__device__ __forceinline__ cd to_cd(cd v) { return v; }
__device__ __forceinline__ cd to_cd(cf v) {
    return cd{static_cast<double>(v.x), static_cast<double>(v.y)};
}

// store of a workspace element where eight lanes cover half a 128 B line
// (FP32 columns): a regular store, so the line's two halves from the two
// wavefronts of a workgroup merge in L2; full-line FP64 stores stream
__device__ __forceinline__ void store_col_ws(cd* p, cd v) { store_stream(p, v); }
__device__ __forceinline__ void store_col_ws(cf* p, cd v) {
    *p = cf{static_cast<float>(v.x), static_cast<float>(v.y)};
}

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Device memory from the library's caching pool (power-of-two size classes
// per device, never returned to the runtime until the process ends).  A
// hipFree synchronises the whole device, so a solve that allocates and frees
// its workspaces would serialise every stream; the pool keeps solves on
// separate streams independent.  Thread-safe.
void* dev_alloc(size_t bytes);
void dev_free(void* p, size_t bytes);

// Non-blocking HIP streams (default priority) from a per-device pool:
// creating and destroying a stream costs ~100 + 60 us of host time, which
// was a large share of a small solve's set-up (about ten per solve).  Used
// for the short-lived streams (entry points, synchronous uploads); a kernel
// handle creates its own.
// stream_acquire takes a stream of the current device; stream_release takes
// back an idle stream, one the caller has synchronised without error (a
// stream whose work failed is destroyed instead).  Pooled streams live until
// the process ends.  Thread-safe.
hipStream_t stream_acquire();
void stream_release(hipStream_t s);
// Synchronise s, then return it to the pool (or destroy it if the wait
// failed); the wait's status is returned.
hipError_t stream_sync_release(hipStream_t s);

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess)
        throw DeviceError(std::string(what) + " failed: " + hipGetErrorString(e) + " (" +
                          file + ":" + std::to_string(line) + ")");
}

#define HFGPU_CHECK(expr) ::hfgpu::hip_check((expr), #expr, __FILE__, __LINE__)
#define HFGPU_LAUNCH_CHECK() ::hfgpu::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

inline unsigned cdiv(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// Block-wide sum of one double, result valid in every thread.  Fixed
// reduction tree, so the result is deterministic for a given launch shape.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* red /* BLOCK/64 slots */) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) s += red[w];
    __syncthreads();
    return s;
}

template <int BLOCK>
__device__ __forceinline__ double block_max(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) s = fmax(s, red[w]);
    __syncthreads();
    return s;
}

template <int BLOCK>
__device__ __forceinline__ double block_min(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) s = fmin(s, red[w]);
    __syncthreads();
    return s;
}

// expit as in utils.hpp:18-25 (1 / (1 + e^-x) for x >= 0, e^x / (1 + e^x)
// otherwise), written without a branch: both forms share e = exp(-|x|), so
// lanes of either sign run the same instructions and round identically.
__device__ __forceinline__ double expit_fd(double x) {
    const double e = exp(-fabs(x));
    return (x >= 0.0 ? 1.0 : e) / (1.0 + e);
}

// Wave-wide sum (64 lanes), result valid in every lane.
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace hfgpu
