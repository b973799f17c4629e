// Mixed-radix Stockham FFT on a tile held in LDS, batched over channels.
//
// Tile layout: s[i * C + c], i in [0, n) along the transform axis, c in
// [0, C) over channels (C a power of two).  The whole workgroup cooperates;
// each radix-R stage reads its butterfly inputs into registers, barriers,
// then writes the autosorted outputs back in place.
//
// Forward transform uses exp(-2 pi i / n), inverse the conjugate, both
// unnormalised (the caller scales), matching FFTW's FORWARD/BACKWARD.
#pragma once

#include "common.h"

namespace hfgpu {

constexpr int kMaxStages = 16;

struct FftAxis {
    int n;
    int nstage;
    int radix[kMaxStages];
    const cd* tw;  // device table, tw[m] = exp(-2 pi i m / n), m < n
};

// Butterflies held per thread per stage.  Kept small: the exchange passes
// are latency-bound, so registers go to occupancy (more resident tiles per
// CU) rather than to more butterflies per thread; the host sizes the channel
// tile so every stage fits.
template <int R>
struct MaxB {
    static constexpr int v = R <= 8 ? 2 : 1;
};

__host__ __device__ constexpr int maxb_of(int r) { return r <= 8 ? 2 : 1; }

// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ cd rot90(cd a) {
    return INV ? cd{-a.y, a.x} : cd{a.y, -a.x};
}

template <int R, bool INV>
__device__ __forceinline__ void butterfly(cd* v, const cd* tw, int n);

template <>
__device__ __forceinline__ void butterfly<2, false>(cd* v, const cd*, int) {
    const cd a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void butterfly<2, true>(cd* v, const cd*, int) {
    const cd a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void bfly3(cd* v) {
    const double s = 0.86602540378443864676372317075293618;  // sin(2 pi / 3)
    const cd t1 = cadd(v[1], v[2]);
    const cd t2 = csub(v[1], v[2]);
    const cd m1 = cd{v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y};
    const cd j = rot90<INV>(cscale(t2, s));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m1, j);
    v[2] = csub(m1, j);
}
template <> __device__ __forceinline__ void butterfly<3, false>(cd* v, const cd*, int) { bfly3<false>(v); }
template <> __device__ __forceinline__ void butterfly<3, true>(cd* v, const cd*, int) { bfly3<true>(v); }

template <bool INV>
__device__ __forceinline__ void bfly4(cd* v) {
    const cd t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const cd t2 = cadd(v[1], v[3]), t3 = rot90<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}
template <> __device__ __forceinline__ void butterfly<4, false>(cd* v, const cd*, int) { bfly4<false>(v); }
template <> __device__ __forceinline__ void butterfly<4, true>(cd* v, const cd*, int) { bfly4<true>(v); }

template <bool INV>
__device__ __forceinline__ void bfly5(cd* v) {
    const double c1 = 0.30901699437494742410229341718281906;   // cos(2 pi / 5)
    const double c2 = -0.80901699437494742410229341718281906;  // cos(4 pi / 5)
    const double s1 = 0.95105651629515357211643933337938214;   // sin(2 pi / 5)
    const double s2 = 0.58778525229247312916870595463907277;   // sin(4 pi / 5)
    const cd t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const cd t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const cd m1 = cd{v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
    const cd m2 = cd{v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
    const cd n1 = rot90<INV>(cd{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y});
    const cd n2 = rot90<INV>(cd{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y});
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
}
template <> __device__ __forceinline__ void butterfly<5, false>(cd* v, const cd*, int) { bfly5<false>(v); }
template <> __device__ __forceinline__ void butterfly<5, true>(cd* v, const cd*, int) { bfly5<true>(v); }

template <bool INV>
__device__ __forceinline__ void bfly8(cd* v) {
    const double h = 0.70710678118654752440084436210484904;  // sqrt(1/2)
    cd b[4], c[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        b[t] = cadd(v[t], v[t + 4]);
        c[t] = csub(v[t], v[t + 4]);
    }
    // c[t] *= W8^t, W8 = exp(-+ i pi / 4)
    c[1] = INV ? cd{h * (c[1].x - c[1].y), h * (c[1].x + c[1].y)}
               : cd{h * (c[1].x + c[1].y), h * (c[1].y - c[1].x)};
    c[2] = rot90<INV>(c[2]);
    c[3] = INV ? cd{-h * (c[3].x + c[3].y), h * (c[3].x - c[3].y)}
               : cd{h * (c[3].y - c[3].x), -h * (c[3].x + c[3].y)};
    bfly4<INV>(b);
    bfly4<INV>(c);
    v[0] = b[0]; v[2] = b[1]; v[4] = b[2]; v[6] = b[3];
    v[1] = c[0]; v[3] = c[1]; v[5] = c[2]; v[7] = c[3];
}
template <> __device__ __forceinline__ void butterfly<8, false>(cd* v, const cd*, int) { bfly8<false>(v); }
template <> __device__ __forceinline__ void butterfly<8, true>(cd* v, const cd*, int) { bfly8<true>(v); }

// Generic odd prime radix: direct DFT with the n-th root table.
template <int R, bool INV>
__device__ __forceinline__ void bfly_generic(cd* v, const cd* tw, int n) {
    cd out[R];
    const int step = n / R;
#pragma unroll
    for (int m = 0; m < R; ++m) {
        cd s = v[0];
#pragma unroll
        for (int t = 1; t < R; ++t) {
            cd w = tw[((t * m) % R) * step];
            if (INV) w.y = -w.y;
            s = cadd(s, cmul(v[t], w));
        }
        out[m] = s;
    }
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = out[m];
}
template <> __device__ __forceinline__ void butterfly<7, false>(cd* v, const cd* tw, int n) { bfly_generic<7, false>(v, tw, n); }
template <> __device__ __forceinline__ void butterfly<7, true>(cd* v, const cd* tw, int n) { bfly_generic<7, true>(v, tw, n); }
template <> __device__ __forceinline__ void butterfly<11, false>(cd* v, const cd* tw, int n) { bfly_generic<11, false>(v, tw, n); }
template <> __device__ __forceinline__ void butterfly<11, true>(cd* v, const cd* tw, int n) { bfly_generic<11, true>(v, tw, n); }
template <> __device__ __forceinline__ void butterfly<13, false>(cd* v, const cd* tw, int n) { bfly_generic<13, false>(v, tw, n); }
template <> __device__ __forceinline__ void butterfly<13, true>(cd* v, const cd* tw, int n) { bfly_generic<13, true>(v, tw, n); }

// One Stockham stage of radix R over the whole tile.  Ns = product of the
// radices of the earlier stages.  s_tw is the twiddle table in LDS.
template <int R, bool INV>
__device__ __forceinline__ void fft_stage(cd* s, const cd* s_tw, int n, int clog, int Ns) {
    constexpr int MB = MaxB<R>::v;
    const int C = 1 << clog;
    const int nr = n / R;
    const int nbf = nr << clog;
    const int tw_step = n / (Ns * R);
    cd v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int u = threadIdx.x + b * blockDim.x;
        if (u < nbf) {
            const int c = u & (C - 1), j = u >> clog;
#pragma unroll
            for (int t = 0; t < R; ++t) v[b][t] = s[((j + t * nr) << clog) + c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int u = threadIdx.x + b * blockDim.x;
        if (u < nbf) {
            const int c = u & (C - 1), j = u >> clog;
            const int k = j % Ns;
            if (k != 0) {
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    cd w = s_tw[t * k * tw_step];
                    if (INV) w.y = -w.y;
                    v[b][t] = cmul(v[b][t], w);
                }
            }
            butterfly<R, INV>(v[b], s_tw, n);
            const int idxD = (j - k) * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t) s[((idxD + t * Ns) << clog) + c] = v[b][t];
        }
    }
    __syncthreads();
}

// Full transform of the tile along its leading axis.  GENERIC = false
// compiles only the radix 2/3/4/5/8 stages (the meshes of every production
// configuration); GENERIC = true adds the direct radix 7/11/13 stages, whose
// register footprint would otherwise halve the occupancy of every mesh.
template <bool INV, bool GENERIC>
__device__ void fft_tile(cd* s, const cd* s_tw, const FftAxis& ax, int clog) {
    int Ns = 1;
    for (int st = 0; st < ax.nstage; ++st) {
        const int r = ax.radix[st];
        switch (r) {
            case 2: fft_stage<2, INV>(s, s_tw, ax.n, clog, Ns); break;
            case 3: fft_stage<3, INV>(s, s_tw, ax.n, clog, Ns); break;
            case 4: fft_stage<4, INV>(s, s_tw, ax.n, clog, Ns); break;
            case 5: fft_stage<5, INV>(s, s_tw, ax.n, clog, Ns); break;
            case 8: fft_stage<8, INV>(s, s_tw, ax.n, clog, Ns); break;
            default:
                if constexpr (GENERIC) {
                    if (r == 7) fft_stage<7, INV>(s, s_tw, ax.n, clog, Ns);
                    else if (r == 11) fft_stage<11, INV>(s, s_tw, ax.n, clog, Ns);
                    else if (r == 13) fft_stage<13, INV>(s, s_tw, ax.n, clog, Ns);
                }
                break;  // other radices are rejected on the host
        }
        Ns *= r;
    }
}

}  // namespace hfgpu
