// mmx_internal.hpp -- shared host/device helpers of the gfx950 solver core.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mmx.h"

namespace mmx {

// ------------------------------------------------------------------ errors

void set_error(const char *fmt, ...);

#define MMX_HIP(call)                                                            \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess) {                                                  \
            ::mmx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,          \
                             hipGetErrorString(_e));                             \
            return MMX_ERR_HIP;                                                  \
        }                                                                        \
    } while (0)

#define MMX_CHECK_LAUNCH()                                                       \
    do {                                                                         \
        hipError_t _e = hipGetLastError();                                       \
        if (_e != hipSuccess) {                                                  \
            ::mmx::set_error("%s:%d launch: %s", __FILE__, __LINE__,             \
                             hipGetErrorString(_e));                             \
            return MMX_ERR_HIP;                                                  \
        }                                                                        \
    } while (0)

#define MMX_REQUIRE(cond, code, ...)                                             \
    do {                                                                         \
        if (!(cond)) {                                                           \
            ::mmx::set_error(__VA_ARGS__);                                       \
            return (code);                                                       \
        }                                                                        \
    } while (0)

#define MMX_TRY(expr)                                                            \
    do {                                                                         \
        int _rc = (expr);                                                        \
        if (_rc != MMX_OK) return _rc;                                           \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Looks up the current device's copy of the allocation-free reductions'
// partials array (local.hip); called by mmx_init and on first use.
int noalloc_setup(int dev);

// Physical constants, src/mumax/common/constants.go:16-26.
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kMu0 = 4.0 * kPi * 1e-7;
constexpr double kGamma0 = 2.211e5;
constexpr double kMuB = 9.2740091523e-24;
constexpr double kQe = 1.60217646e-19;
constexpr double kKb = 1.380650424e-23; // common/constants.go:21

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// Grids are launched 1-D with this many threads per workgroup (4 waves).
constexpr int kBlock = 256;

// ----------------------------------------------------------- complex types

template <typename T> struct cpx_t;
template <> struct cpx_t<float> { using type = float2; };
template <> struct cpx_t<double> { using type = double2; };

template <typename C> __host__ __device__ inline C cmul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename C> __host__ __device__ inline C cmulc(C a, C b)
{
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename C> __host__ __device__ inline C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __host__ __device__ inline C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __host__ __device__ inline C cconj(C a) { return C{a.x, -a.y}; }

// ---------------------------------------------------------------- FFT plan

constexpr int kMaxStages = 12;

// One transform length: radix sequence (product = L) and twiddle tables in
// device memory: tw[t] = exp(-2 pi i t / L) for t < L; rtw[k] = exp(-pi i k / L)
// for k <= L (the R2C/C2R split factors of a 2L-point real transform).
template <typename T> struct LineFFT {
    using C = typename cpx_t<T>::type;
    int L = 1;
    int nstages = 0;
    int radix[kMaxStages] = {};
    const C *tw = nullptr;
    const C *rtw = nullptr;
};

// Splits L into radices from {8,4,2,3,5,7}; false if L has another factor.
bool factor_radices(int L, int *radix, int *nstages);

// ---------------------------------------------------- small DFTs (device)

template <typename C, int R, int DIR> struct SmallDFT;

template <typename C, int DIR> struct SmallDFT<C, 2, DIR> {
    __device__ static inline void run(C *v)
    {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

// multiply by -i (DIR = -1) or +i (DIR = +1)
template <typename C, int DIR> __device__ inline C rot90(C a)
{
    if (DIR < 0) return C{a.y, -a.x};
    return C{-a.y, a.x};
}

template <typename C, int DIR> struct SmallDFT<C, 4, DIR> {
    __device__ static inline void run(C *v)
    {
        C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        C b0 = cadd(v[1], v[3]), b1 = rot90<C, DIR>(csub(v[1], v[3]));
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    }
};

template <typename C, int DIR> struct SmallDFT<C, 8, DIR> {
    __device__ static inline void run(C *v)
    {
        using T = decltype(v[0].x);
        const T r = (T)0.70710678118654752440084436210484903928;
        // radix-2 on pairs (v[j], v[j+4]), twiddle W8^j, then two radix-4
        C e[4], o[4];
        for (int j = 0; j < 4; j++) {
            e[j] = cadd(v[j], v[j + 4]);
            o[j] = csub(v[j], v[j + 4]);
        }
        // o[j] *= W8^(j) with W8 = exp(DIR * 2 pi i / 8)
        o[1] = C{(o[1].x - DIR * o[1].y) * r, (o[1].y + DIR * o[1].x) * r};
        o[2] = rot90<C, DIR>(o[2]);
        o[3] = C{(-o[3].x - DIR * o[3].y) * r, (-o[3].y + DIR * o[3].x) * r};
        SmallDFT<C, 4, DIR>::run(e);
        SmallDFT<C, 4, DIR>::run(o);
        for (int j = 0; j < 4; j++) {
            v[2 * j] = e[j];
            v[2 * j + 1] = o[j];
        }
    }
};

// odd radices: direct DFT with exact-rounded constants
template <typename C, int R, int DIR> struct OddDFT {
    __device__ static inline void run(C *v)
    {
        using T = decltype(v[0].x);
        C out[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            T re = v[0].x, im = v[0].y;
#pragma unroll
            for (int n = 1; n < R; n++) {
                const int t = (n * k) % R;
                const double ang = 2.0 * kPi * (double)t / (double)R;
                const T c = (T)__builtin_cos(ang);
                const T s = (T)(DIR * __builtin_sin(ang));
                re += v[n].x * c - v[n].y * s;
                im += v[n].x * s + v[n].y * c;
            }
            out[k] = C{re, im};
        }
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = out[k];
    }
};
template <typename C, int DIR> struct SmallDFT<C, 3, DIR> : OddDFT<C, 3, DIR> {};
template <typename C, int DIR> struct SmallDFT<C, 5, DIR> : OddDFT<C, 5, DIR> {};
template <typename C, int DIR> struct SmallDFT<C, 7, DIR> : OddDFT<C, 7, DIR> {};

// ------------------------------------------- Stockham stage over LDS lines
//
// Lines are stored in LDS as buf[line * ls + elem * es].  Each thread of the
// workgroup takes butterflies b = tid, tid + nthreads, ...; butterfly b of a
// stage with radix R works on line b / (L/R), position j = b % (L/R): it
// reads x[j + r L/R], applies W_{Ns R}^{r k} (k = j mod Ns), does a radix-R DFT
// and writes y[(j/Ns) Ns R + k + r Ns].  After all stages the data is in
// natural order (Stockham autosort).
template <typename C, int R, int DIR>
__device__ inline void stockham_stage(const C *__restrict__ A, C *__restrict__ B,
                                      int nl, int ls, int es, int L, int Ns,
                                      const C *__restrict__ tw)
{
    const int M = L / R;
    const int nb = nl * M;
    const int tstep = L / (Ns * R);
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const int l = b / M;
        const int j = b - l * M;
        const int k = j % Ns;
        C v[R];
        const C *a = A + l * ls;
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = a[(j + r * M) * es];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                C w = tw[r * k * tstep];
                v[r] = (DIR < 0) ? cmul(v[r], w) : cmulc(v[r], w);
            }
        }
        SmallDFT<C, R, DIR>::run(v);
        const int base = (j / Ns) * Ns * R + k;
        C *o = B + l * ls;
#pragma unroll
        for (int r = 0; r < R; r++) o[(base + r * Ns) * es] = v[r];
    }
}

// The same stage for a prime radix R > 7 known only at run time: each output
// of a butterfly is its R-point DFT evaluated directly from the twiddled
// inputs in A (R^2 complex multiply-adds per butterfly, the DFT's roots
// W_R^(rq) = W_L^((rq mod R) L/R) taken from the same table), so that every
// transform length has a plan, as cuFFT gives the reference
// (fftplanX.go:59-64).
template <typename C, int DIR>
__device__ inline void stockham_stage_any(const C *__restrict__ A, C *__restrict__ B, int nl,
                                          int ls, int es, int L, int Ns, int R,
                                          const C *__restrict__ tw)
{
    const int M = L / R;
    const int nb = nl * M;
    const int tstep = L / (Ns * R);
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const int l = b / M;
        const int j = b - l * M;
        const int k = j % Ns;
        const C *a = A + l * ls;
        const int base = (j / Ns) * Ns * R + k;
        C *o = B + l * ls;
        for (int q = 0; q < R; q++) {
            C acc = a[j * es];
            int rq = 0; // r q mod R
            for (int r = 1; r < R; r++) {
                rq += q;
                if (rq >= R) rq -= R;
                C v = a[(j + r * M) * es];
                if (Ns > 1) {
                    const C w = tw[r * k * tstep];
                    v = (DIR < 0) ? cmul(v, w) : cmulc(v, w);
                }
                const C w = tw[rq * M];
                v = (DIR < 0) ? cmul(v, w) : cmulc(v, w);
                acc.x += v.x;
                acc.y += v.y;
            }
            o[(base + q * Ns) * es] = acc;
        }
    }
}

// Full transform of nl lines; ping-pongs between A and B and returns the
// buffer that holds the result.  Ends with a barrier.
template <typename T, int DIR>
__device__ inline typename cpx_t<T>::type *lds_fft(typename cpx_t<T>::type *A,
                                                   typename cpx_t<T>::type *B,
                                                   int nl, int ls, int es,
                                                   const LineFFT<T> &f)
{
    using C = typename cpx_t<T>::type;
    int Ns = 1;
    for (int s = 0; s < f.nstages; s++) {
        const int R = f.radix[s];
        switch (R) {
        case 8: stockham_stage<C, 8, DIR>(A, B, nl, ls, es, f.L, Ns, f.tw); break;
        case 4: stockham_stage<C, 4, DIR>(A, B, nl, ls, es, f.L, Ns, f.tw); break;
        case 2: stockham_stage<C, 2, DIR>(A, B, nl, ls, es, f.L, Ns, f.tw); break;
        case 3: stockham_stage<C, 3, DIR>(A, B, nl, ls, es, f.L, Ns, f.tw); break;
        case 5: stockham_stage<C, 5, DIR>(A, B, nl, ls, es, f.L, Ns, f.tw); break;
        case 7: stockham_stage<C, 7, DIR>(A, B, nl, ls, es, f.L, Ns, f.tw); break;
        default: stockham_stage_any<C, DIR>(A, B, nl, ls, es, f.L, Ns, R, f.tw); break;
        }
        __syncthreads();
        C *t = A;
        A = B;
        B = t;
        Ns *= R;
    }
    return A;
}

// ------------------------------------------------------- device utilities

__device__ inline float mask_unity(const float *__restrict__ m, long i)
{
    return m ? m[i] : 1.0f;
}
__device__ inline float div_zero(float a, float b) { return (b == 0.0f) ? 0.0f : a / b; }
__device__ inline float avg_geom_zero(float a, float b)
{
    float s = a + b;
    return (s == 0.0f) ? 0.0f : 2.0f * a * b / s;
}

// wave64 reductions
__device__ inline float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide max of non-negative floats, merged into *dst with an atomic on
// the bit pattern (ordered like unsigned ints for x >= 0).  NaN is mapped to
// +inf so that it can never hide below a finite maximum.
__device__ inline void block_max_to(float v, unsigned int *dst)
{
    __shared__ float part[kBlock / 64];
    if (v != v) v = __builtin_inff();
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = part[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmaxf(m, part[w]);
        atomicMax(dst, __float_as_uint(m));
    }
}

} // namespace mmx
