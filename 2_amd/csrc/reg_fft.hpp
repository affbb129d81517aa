// reg_fft.hpp -- register-resident FFT building blocks for the fast passes.
//
// A line of L = T*E points is held by T threads with E registers each:
// thread t owns x[t + T q], q < E.  The forward transform is the four-step
// split L = E x T: an E-point FFT in registers, the twiddle W_L^(t k1), one
// transpose through LDS, and T-point FFTs in registers; afterwards thread t
// owns X[k1 + E k2] for k1 = t G + g (g < G = E/T) and all k2 < T.  The
// inverse runs the same steps backwards from that layout and returns to the
// input layout, so forward + pointwise work + inverse costs two LDS
// transposes.  Small FFTs are unrolled radix-2 Stockham with compile-time
// twiddle constants.
#pragma once

#include <type_traits>
#include <utility>

#include "mmx_internal.hpp"

namespace mmx {

// ------------------------------------------------ compile-time twiddles

constexpr double cx_sin_small(double x)
{
    const double x2 = x * x;
    double term = x, sum = x;
    for (int n = 1; n < 14; n++) {
        term *= -x2 / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos_small(double x)
{
    const double x2 = x * x;
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 14; n++) {
        term *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
// cos and sin of 2 pi k / n, reduced to [0, pi/4] by symmetry
constexpr double cx_cs(long k, long n, bool want_sin)
{
    long kk = ((k % n) + n) % n;
    // angle = 2 pi kk / n = quadrant * pi/2 + rest
    const long num = 4 * kk;
    const long quad = num / n;
    const double frac = (double)(num - quad * n) / (double)n; // in [0,1)
    const double hp = 1.57079632679489661923132169163975144;
    double c = 0.0, s = 0.0;
    if (frac <= 0.5) {
        c = cx_cos_small(frac * hp);
        s = cx_sin_small(frac * hp);
    } else {
        c = cx_sin_small((1.0 - frac) * hp);
        s = cx_cos_small((1.0 - frac) * hp);
    }
    double C = c, S = s;
    switch (quad) {
    case 1: C = -s; S = c; break;
    case 2: C = -c; S = -s; break;
    case 3: C = s; S = -c; break;
    default: break;
    }
    return want_sin ? S : C;
}

// x * W_M^K with W_M = exp(DIR * 2 pi i / M); trivial factors avoided
template <int M, int K, int DIR> __device__ __forceinline__ float2 twmul(float2 a)
{
    constexpr int k = ((K % M) + M) % M;
    if constexpr (k == 0) {
        return a;
    } else if constexpr (4 * k == M) {
        return rot90<float2, DIR>(a);
    } else if constexpr (2 * k == M) {
        return float2{-a.x, -a.y};
    } else if constexpr (4 * k == 3 * M) {
        return rot90<float2, -DIR>(a);
    } else {
        constexpr float c = (float)cx_cs(k, M, false);
        constexpr float s = (float)((double)DIR * cx_cs(k, M, true));
        return float2{a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

// --------------------------------------------- unrolled radix-2 Stockham

template <int N, int DIR, int Ns> struct R2Stages {
    template <int J> __device__ __forceinline__ static void stage(const float2 *a, float2 *b)
    {
        if constexpr (J < N / 2) {
            constexpr int k = J % Ns;
            const float2 u0 = a[J];
            const float2 u1 = twmul<2 * Ns, k, DIR>(a[J + N / 2]);
            constexpr int base = (J / Ns) * Ns * 2 + k;
            b[base] = float2{u0.x + u1.x, u0.y + u1.y};
            b[base + Ns] = float2{u0.x - u1.x, u0.y - u1.y};
            stage<J + 1>(a, b);
        }
    }
    // returns true if the result ended in b
    __device__ __forceinline__ static void run(float2 *a, float2 *b, bool &in_b)
    {
        if constexpr (Ns < N) {
            stage<0>(a, b);
            in_b = !in_b;
            R2Stages<N, DIR, Ns * 2>::run(b, a, in_b);
        }
    }
};

// ------------------------------------- mixed radix (N = 2^a * {3, 5, 7})
//
// The same Stockham stage for radix R at span Ns: inputs a[J + r N/R],
// r < R, times W_{Ns R}^{r k} (k = J mod Ns), an R-point DFT, outputs
// b[(J/Ns) Ns R + k + r Ns].  The radix-2 stages run first (so that a line
// whose upper half is zero starts with copies), the odd radix last.
template <int N, int DIR, int Ns, int R> struct MRStage {
    template <int J> __device__ __forceinline__ static void step(const float2 *a, float2 *b)
    {
        if constexpr (J < N / R) {
            constexpr int k = J % Ns;
            float2 v[R];
            static_for_r<R>([&](auto r) {
                constexpr int q = decltype(r)::value;
                v[q] = twmul<Ns * R, k * q, DIR>(a[J + q * (N / R)]);
            });
            SmallDFT<float2, R, DIR>::run(v);
            constexpr int base = (J / Ns) * Ns * R + k;
            static_for_r<R>([&](auto r) {
                constexpr int q = decltype(r)::value;
                b[base + q * Ns] = v[q];
            });
            step<J + 1>(a, b);
        }
    }
    template <typename F, int... I>
    __device__ __forceinline__ static void static_for_r_impl(F &&f, std::integer_sequence<int, I...>)
    {
        (f(std::integral_constant<int, I>{}), ...);
    }
    template <int M, typename F> __device__ __forceinline__ static void static_for_r(F &&f)
    {
        static_for_r_impl(f, std::make_integer_sequence<int, M>{});
    }
};

// the radix of the stage at span Ns: 2 while N/Ns is even, then the odd rest
template <int N, int Ns> constexpr int mr_radix() { return ((N / Ns) % 2 == 0) ? 2 : N / Ns; }

template <int N, int DIR, int Ns>
__device__ __forceinline__ void mr_stages(float2 *a, float2 *b, bool &in_b)
{
    if constexpr (Ns < N) {
        constexpr int R = mr_radix<N, Ns>();
        static_assert(R == 2 || R == 3 || R == 5 || R == 7, "lengths 2^a * {1, 3, 5, 7} only");
        MRStage<N, DIR, Ns, R>::template step<0>(a, b);
        in_b = !in_b;
        mr_stages<N, DIR, Ns * R>(b, a, in_b);
    }
}

constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// In-register FFT of N points, natural order in and out: unrolled radix-2
// Stockham for powers of two, mixed radix for N = 2^a * {3, 5, 7}.
template <int N, int DIR> __device__ __forceinline__ void fft_reg(float2 (&v)[N])
{
    if constexpr (!is_pow2(N)) {
        float2 tmp[N];
        bool in_b = false;
        mr_stages<N, DIR, 1>(v, tmp, in_b);
        if (in_b) {
#pragma unroll
            for (int i = 0; i < N; i++) v[i] = tmp[i];
        }
    } else if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = float2{a.x + b.x, a.y + b.y};
        v[1] = float2{a.x - b.x, a.y - b.y};
    } else if constexpr (N == 4) {
        SmallDFT<float2, 4, DIR>::run(v);
    } else if constexpr (N == 8) {
        SmallDFT<float2, 8, DIR>::run(v);
    } else {
        float2 tmp[N];
        bool in_b = false;
        R2Stages<N, DIR, 1>::run(v, tmp, in_b);
        // after log2(N) ping-pong stages the result is in tmp when log2 N is odd
        constexpr int lg = __builtin_ctz(N);
        if constexpr (lg & 1) {
#pragma unroll
            for (int i = 0; i < N; i++) v[i] = tmp[i];
        }
    }
}

// In-register FFT of N points whose upper half v[N/2 .. N-1] is zero (those
// registers are not read): the first radix-2 stage reduces to copies.
template <int N, int DIR> __device__ __forceinline__ void fft_reg_halfin(float2 (&v)[N])
{
    if constexpr (!is_pow2(N)) {
        // first radix-2 stage (Ns = 1) on a zero upper half: copies
        static_assert(N % 2 == 0, "a zero-padded line has an even length");
        float2 tmp[N];
#pragma unroll
        for (int j = 0; j < N / 2; j++) {
            tmp[2 * j] = v[j];
            tmp[2 * j + 1] = v[j];
        }
        bool in_b = true;
        mr_stages<N, DIR, 2>(tmp, v, in_b);
        if (in_b) {
#pragma unroll
            for (int i = 0; i < N; i++) v[i] = tmp[i];
        }
    } else if constexpr (N <= 8) {
#pragma unroll
        for (int i = N / 2; i < N; i++) v[i] = float2{0.0f, 0.0f};
        fft_reg<N, DIR>(v);
    } else {
        float2 tmp[N];
#pragma unroll
        for (int j = 0; j < N / 2; j++) {
            tmp[2 * j] = v[j];
            tmp[2 * j + 1] = v[j];
        }
        bool in_b = true;
        R2Stages<N, DIR, 2>::run(tmp, v, in_b);
        // the remaining log2(N) - 1 stages alternate v, tmp, v, ...: the result
        // is in tmp when log2 N is odd
        constexpr int lg = __builtin_ctz(N);
        if constexpr (lg & 1) {
#pragma unroll
            for (int i = 0; i < N; i++) v[i] = tmp[i];
        }
    }
}

// Calls f(std::integral_constant<int, I>) for I = 0 .. N-1 (compile-time loop).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// -------------------------------------------------- four-step geometry

// threads per line T and registers per thread E = L/T, E >= T, T | E
template <int L> struct Split;
template <> struct Split<8> { static constexpr int T = 2, E = 4; };
template <> struct Split<16> { static constexpr int T = 4, E = 4; };
template <> struct Split<32> { static constexpr int T = 4, E = 8; };
template <> struct Split<64> { static constexpr int T = 8, E = 8; };
template <> struct Split<128> { static constexpr int T = 8, E = 16; };
template <> struct Split<256> { static constexpr int T = 16, E = 16; };
template <> struct Split<512> { static constexpr int T = 16, E = 32; };
template <> struct Split<1024> { static constexpr int T = 32, E = 32; };
template <> struct Split<2048> { static constexpr int T = 32, E = 64; };
// lengths with an odd factor 3, 5 or 7: the odd factor is in E (T must divide
// E and be a power of two)
template <> struct Split<24> { static constexpr int T = 2, E = 12; };
template <> struct Split<48> { static constexpr int T = 4, E = 12; };
template <> struct Split<96> { static constexpr int T = 4, E = 24; };
template <> struct Split<192> { static constexpr int T = 8, E = 24; };
template <> struct Split<384> { static constexpr int T = 8, E = 48; };
template <> struct Split<40> { static constexpr int T = 2, E = 20; };
template <> struct Split<80> { static constexpr int T = 4, E = 20; };
template <> struct Split<160> { static constexpr int T = 4, E = 40; };
template <> struct Split<320> { static constexpr int T = 8, E = 40; };
template <> struct Split<56> { static constexpr int T = 2, E = 28; };
template <> struct Split<112> { static constexpr int T = 4, E = 28; };
template <> struct Split<224> { static constexpr int T = 4, E = 56; };
template <> struct Split<448> { static constexpr int T = 8, E = 56; };

// Lengths of the register-FFT passes.  Strided y passes of 3-D grids take
// every length above; the passes that hold three components per thread (x
// passes, film y, and z whose LDS holds three components) only those with at
// most 28 registers per component.
inline bool fast_pow2_length(int L)
{
    return L == 8 || L == 16 || L == 32 || L == 64 || L == 128 || L == 256 || L == 512 ||
           L == 1024 || L == 2048;
}
inline bool fast_odd_short(int L)
{
    return L == 24 || L == 48 || L == 96 || L == 192 || L == 40 || L == 80 || L == 56 || L == 112;
}
inline bool fast_odd_long(int L) { return L == 384 || L == 160 || L == 320 || L == 224 || L == 448; }
inline bool fast_length_impl(int L) { return fast_pow2_length(L) || fast_odd_short(L); }
inline bool fast_ylength_impl(int L) { return fast_length_impl(L) || fast_odd_long(L); }

// LDS row pitch (in float2) of a column-tile transpose: rows of T*W points,
// read back by lanes (t, w) at row t*G + g.  A pitch with G*S = W (mod 32)
// puts the lanes' 64-bit words on distinct bank pairs (64 lanes cover the
// 64 four-byte banks twice, each half-wave once).
constexpr int lds_pitch(int T, int W, int G)
{
    for (int pad = 0; pad < 32; pad++)
        if (((G * (T * W + pad) - W) % 32 + 32) % 32 == 0) return T * W + pad;
    return T * W + 1;
}

// Blocks are dealt to the 8 XCDs round-robin; this maps a block id to a tile
// id so that each XCD walks one contiguous range of tiles (its own L2).
__device__ __forceinline__ int xcd_tile(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = bid & 7, l = bid >> 3;
    return (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// Forward step 1-2 of the four-step: E-point FFT, then v[k1] *= W_L^(t k1)
// (twiddle table tw[L] in LDS, forward sign).  HALFIN: the line's upper half
// (v[E/2 ..]) is zero and not read.
template <int L, int DIR, bool HALFIN = false>
__device__ __forceinline__ void fs_first(float2 (&v)[Split<L>::E], int t, const float2 *stw)
{
    constexpr int E = Split<L>::E;
    if constexpr (HALFIN)
        fft_reg_halfin<E, DIR>(v);
    else
        fft_reg<E, DIR>(v);
#pragma unroll
    for (int k1 = 1; k1 < E; k1++) {
        const float2 w = stw[t * k1];
        v[k1] = (DIR < 0) ? cmul(v[k1], w) : cmulc(v[k1], w);
    }
}

} // namespace mmx
