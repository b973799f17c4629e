// Small DFTs held entirely in registers, for the two-step (N = N1 x N2) line
// transforms of the exchange passes.
//
// Dft<N, INV>::run(v) replaces v[0..N) by its DFT in natural order, forward
// with exp(-2 pi i / N) and inverse with the conjugate, both unnormalised.
// Sizes 2, 3, 4, 5 and 8 use the closed-form butterflies of fft_lds.h, the
// primes 7, 11 and 13 a direct sum, and every other size a Cooley-Tukey
// split N = A x B carried out in registers.  All indices and twiddles are
// compile-time constants once the loops are unrolled; the twiddles come from
// a constexpr sine/cosine (Taylor series after octant reduction), accurate to
// an ulp or two.
#pragma once

#include "fft_lds.h"

namespace hfgpu {

namespace dftc {

constexpr double kHalfPi = 1.57079632679489661923132169163975144;

// sin and cos of x for |x| <= pi/4 (terms to x^25: below 1e-20 there)
constexpr double taylor_sin(double x) {
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i <= 12; ++i) {
        term *= -x2 / static_cast<double>((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i <= 12; ++i) {
        term *= -x2 / static_cast<double>((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}

// cos(2 pi m / N) and sin(2 pi m / N): quadrant q = floor(4 m / N) and the
// remainder angle phi = (pi/2) r / N in [0, pi/2), reflected to [0, pi/4].
constexpr double root_cs(int m, int N, bool want_sin) {
    m %= N;
    if (m < 0) m += N;
    const int q = (4 * m) / N;
    const int r = 4 * m - q * N;  // 0 <= r < N
    double c = 0.0, s = 0.0;
    if (2 * r <= N) {
        const double phi = kHalfPi * static_cast<double>(r) / static_cast<double>(N);
        c = taylor_cos(phi);
        s = taylor_sin(phi);
    } else {
        const double phi = kHalfPi * static_cast<double>(N - r) / static_cast<double>(N);
        c = taylor_sin(phi);
        s = taylor_cos(phi);
    }
    // rotate by q quarter turns: (c, s) -> (-s, c)
    for (int k = 0; k < q; ++k) {
        const double t = c;
        c = -s;
        s = t;
    }
    return want_sin ? s : c;
}

// W_N^m = exp(-2 pi i m / N) (forward) or its conjugate (inverse), as
// compile-time constants
template <int N, int M, bool INV>
struct Root {
    static constexpr double re = root_cs(M, N, false);
    static constexpr double im = INV ? root_cs(M, N, true) : -root_cs(M, N, true);
};

constexpr bool is_small_prime(int n) { return n == 7 || n == 11 || n == 13; }

// split factor of a composite size (4 first, then 2, 3, 5)
constexpr int split_of(int n) {
    return n % 4 == 0 ? 4 : (n % 2 == 0 ? 2 : (n % 3 == 0 ? 3 : 5));
}

}  // namespace dftc

template <int N, int M, bool INV>
__device__ __forceinline__ cd mul_root(cd v) {
    if constexpr (M == 0) {
        return v;
    } else {
        return cmul(v, cd{dftc::Root<N, M, INV>::re, dftc::Root<N, M, INV>::im});
    }
}

// w[a B + KB] *= W_N^(a KB) for a = A_I .. A-1
template <int N, int A, int B, int KB, int A_I, bool INV>
struct TwiddleCol {
    static __device__ __forceinline__ void run(cd* w) {
        if constexpr (A_I < A) {
            w[A_I * B + KB] = mul_root<N, (A_I * KB) % N, INV>(w[A_I * B + KB]);
            TwiddleCol<N, A, B, KB, A_I + 1, INV>::run(w);
        }
    }
};
template <int N, int A, int B, int KB, bool INV>
struct TwiddleAll {
    static __device__ __forceinline__ void run(cd* w) {
        if constexpr (KB < B) {
            TwiddleCol<N, A, B, KB, 1, INV>::run(w);
            TwiddleAll<N, A, B, KB + 1, INV>::run(w);
        }
    }
};

// out[m] += v[T] W_N^(T m) for m = M .. N-1, then the next T
template <int N, int T, int M, bool INV>
struct DirectRow {
    static __device__ __forceinline__ void run(const cd* v, cd* out) {
        if constexpr (M < N) {
            out[M] = cadd(out[M], mul_root<N, (T * M) % N, INV>(v[T]));
            DirectRow<N, T, M + 1, INV>::run(v, out);
        }
    }
};
template <int N, int T, bool INV>
struct DirectAll {
    static __device__ __forceinline__ void run(const cd* v, cd* out) {
        if constexpr (T < N) {
            DirectRow<N, T, 0, INV>::run(v, out);
            DirectAll<N, T + 1, INV>::run(v, out);
        }
    }
};

template <int N, bool INV>
struct Dft {
    static __device__ __forceinline__ void run(cd* v) {
        if constexpr (N == 1) {
            return;
        } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 8) {
            butterfly<N, INV>(v, nullptr, 0);
        } else if constexpr (dftc::is_small_prime(N)) {
            cd out[N];
#pragma unroll
            for (int m = 0; m < N; ++m) out[m] = v[0];
            DirectAll<N, 1, INV>::run(v, out);
#pragma unroll
            for (int m = 0; m < N; ++m) v[m] = out[m];
        } else {
            constexpr int A = dftc::split_of(N);
            constexpr int B = N / A;
            // x[a + A b]: a B-point DFT over b for each a, then the twiddles
            // W_N^(a kb), then an A-point DFT over a for each kb:
            // X[kb + B ka]
            cd w[N];
#pragma unroll
            for (int a = 0; a < A; ++a) {
                cd t[B];
#pragma unroll
                for (int b = 0; b < B; ++b) t[b] = v[a + A * b];
                Dft<B, INV>::run(t);
#pragma unroll
                for (int kb = 0; kb < B; ++kb) w[a * B + kb] = t[kb];
            }
            TwiddleAll<N, A, B, 1, INV>::run(w);
#pragma unroll
            for (int kb = 0; kb < B; ++kb) {
                cd t[A];
#pragma unroll
                for (int a = 0; a < A; ++a) t[a] = w[a * B + kb];
                Dft<A, INV>::run(t);
#pragma unroll
                for (int ka = 0; ka < A; ++ka) v[kb + B * ka] = t[ka];
            }
        }
    }
};

}  // namespace hfgpu
