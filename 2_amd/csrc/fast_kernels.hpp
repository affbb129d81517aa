// fast_kernels.hpp -- the demag pass kernels for power-of-two transform
// lengths (register four-step FFTs), shared by demag_fast.hip and xi_gen.hip.
//
// Same five passes as demag.hip (xf, yf, z, yi, xi; films: xf, the y
// transform with the kernel multiply on z_fast, xi) with
// the FFT held in registers (four-step split, reg_fft.hpp): each pass reads its
// input once and writes its output once, with one LDS transpose per transform
// direction.  Strided passes work on tiles of W spectrum columns (W >= 16, so
// each row segment is at least 128 bytes on 128-byte aligned rows); blocks are
// mapped to tiles XCD by XCD (xcd_tile).
#pragma once

#include "epilogue.hpp"
#include "reg_fft.hpp"

namespace mmx {

// Work split of a strided (y or z) pass: T threads per line, E registers, W
// lines (spectrum columns) per workgroup, LDS row pitch S.
template <int L> struct ColTile {
    static constexpr int T = Split<L>::T;
    static constexpr int E = Split<L>::E;
    static constexpr int G = E / T;
    // length 2048: 2 columns (3 % faster than 4 at 64x2048x2048: 49 KB of LDS
    // per workgroup instead of 83, three workgroups per CU instead of one)
    // lengths with an odd factor and more than 32 registers: half as many
    // columns (LDS below 64 KB)
    static constexpr int W = (L == 2048) ? 2 : (L == 1024 ? 8 : (E > 32 ? 128 / T : 256 / T));
    static constexpr int S = lds_pitch(T, W, G);
    static constexpr int NT = T * W;                     // threads
    static constexpr int LDS = (E * S + L) * 8;           // transpose + twiddles
};

// Row pass (x): R rows per workgroup, T threads per row.
template <int LX> struct RowTile {
    static constexpr int T = Split<LX>::T;
    static constexpr int E = Split<LX>::E;
    static constexpr int G = E / T;
    static constexpr int R = (LX == 2048) ? 4 : (LX == 1024 ? 8 : 256 / T);
    // pitch = 1 mod 32 complex: the transposed reads of 16-32 rows x T lanes
    // hit distinct bank pairs
    static constexpr int S = T * R + 1;
    static constexpr int NT = T * R;
    static constexpr int LDS = (E * S + 2 * LX + 1) * 8;
};

// ---------------------------------------------------- column transforms
//
// Thread (w, t) of a column tile: w = tid % W (column), t = tid / W.
// Transpose buffer index of Y[t][k1] of column w: k1*S + t*W + w.

template <int L, int W, int S, bool HALFIN = false>
__device__ __forceinline__ void col_forward(float2 (&v)[Split<L>::E],
                                            float2 (&u)[Split<L>::E / Split<L>::T][Split<L>::T],
                                            int w, int t, float2 *sbuf, const float2 *stw)
{
    constexpr int T = Split<L>::T, E = Split<L>::E, G = E / T;
    fs_first<L, -1, HALFIN>(v, t, stw);
#pragma unroll
    for (int k1 = 0; k1 < E; k1++) sbuf[k1 * S + t * W + w] = v[k1];
    __syncthreads();
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
#pragma unroll
        for (int t2 = 0; t2 < T; t2++) u[gg][t2] = sbuf[(t * G + gg) * S + t2 * W + w];
        fft_reg<T, -1>(u[gg]);
    }
    __syncthreads();
}

template <int L, int W, int S>
__device__ __forceinline__ void col_inverse(float2 (&u)[Split<L>::E / Split<L>::T][Split<L>::T],
                                            float2 (&v)[Split<L>::E], int w, int t,
                                            float2 *sbuf, const float2 *stw)
{
    constexpr int T = Split<L>::T, E = Split<L>::E, G = E / T;
#pragma unroll
    for (int gg = 0; gg < G; gg++) {
        fft_reg<T, +1>(u[gg]);
        const int k1 = t * G + gg;
#pragma unroll
        for (int t2 = 0; t2 < T; t2++) {
            const float2 x = (t2 == 0) ? u[gg][0] : cmulc(u[gg][t2], stw[t2 * k1]);
            sbuf[k1 * S + t2 * W + w] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < E; k1++) v[k1] = sbuf[k1 * S + t * W + w];
    fft_reg<E, +1>(v);
    __syncthreads();
}

// *(T *)(base + off) with a 32-bit byte offset: with a wave-uniform base this
// is one global access with a scalar base and a vector offset.
template <typename T> __device__ __forceinline__ T ld_off(const void *base, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T> __device__ __forceinline__ void st_off(void *base, unsigned off, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}

// yf: FFT along y of the n0 nonzero planes.  FULL: periodic y (n1 == L);
// otherwise n1 == L/2 and rows n1 .. L-1 are zero padding (not read).
template <int L, bool FULL>
__global__ void __launch_bounds__(ColTile<L>::NT)
yf_fast(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float2 *__restrict__ xb,
        float2 *__restrict__ yb)
{
    using C = ColTile<L>;
    constexpr int NQ = FULL ? C::E : C::E / 2; // rows t + T q held, q < NQ
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, rest = tile / ntx, i = rest % g.n0, c = rest / g.n0;
    const int kx0 = tx * C::W;
    const int w = threadIdx.x % C::W, t = threadIdx.x / C::W;
    for (int q = threadIdx.x; q < L; q += C::NT) stw[q] = tw[q];
    const bool ok = kx0 + w < g.h2;
    // columns past h2 load column h2 - 1 (in bounds) and store nothing
    const int col = ok ? kx0 + w : g.h2 - 1;
    const float2 *src = xb + ((size_t)c * g.nrows + (size_t)i * g.n1) * g.sx;
    const unsigned roff = (unsigned)(((size_t)t * g.sx + col) * sizeof(float2));
    float2 v[C::E];
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = ld_off<float2>(src + (size_t)q * C::T * g.sx, roff);
    __syncthreads();
    float2 u[C::G][C::T];
    col_forward<L, C::W, C::S, !FULL>(v, u, w, t, sbuf, stw);
    if (!ok) return;
    float2 *dst = yb + ((size_t)c * g.n0 + i) * (size_t)g.p1 * g.sx;
    const unsigned woff = (unsigned)(((size_t)t * C::G * g.sx + col) * sizeof(float2));
#pragma unroll
    for (int gg = 0; gg < C::G; gg++)
#pragma unroll
        for (int k2 = 0; k2 < C::T; k2++)
            st_off<float2>(dst + (size_t)(gg + C::E * k2) * g.sx, woff, u[gg][k2]);
}

// yi: inverse FFT along y, keep j < n1 (FULL: periodic y, n1 == L)
template <int L, bool FULL>
__global__ void __launch_bounds__(ColTile<L>::NT)
yi_fast(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float2 *__restrict__ yb,
        float2 *__restrict__ xb)
{
    using C = ColTile<L>;
    constexpr int NQ = FULL ? C::E : C::E / 2;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, rest = tile / ntx, i = rest % g.n0, c = rest / g.n0;
    const int kx0 = tx * C::W;
    const int w = threadIdx.x % C::W, t = threadIdx.x / C::W;
    for (int q = threadIdx.x; q < L; q += C::NT) stw[q] = tw[q];
    const bool ok = kx0 + w < g.h2;
    const int col = ok ? kx0 + w : g.h2 - 1;
    const float2 *src = yb + ((size_t)c * g.n0 + i) * (size_t)g.p1 * g.sx;
    const unsigned roff = (unsigned)(((size_t)t * C::G * g.sx + col) * sizeof(float2));
    float2 u[C::G][C::T];
#pragma unroll
    for (int gg = 0; gg < C::G; gg++)
#pragma unroll
        for (int k2 = 0; k2 < C::T; k2++)
            u[gg][k2] = ld_off<float2>(src + (size_t)(gg + C::E * k2) * g.sx, roff);
    __syncthreads();
    float2 v[C::E];
    col_inverse<L, C::W, C::S>(u, v, w, t, sbuf, stw);
    if (!ok) return;
    float2 *dst = xb + ((size_t)c * g.nrows + (size_t)i * g.n1) * g.sx;
    const unsigned woff = (unsigned)(((size_t)t * g.sx + col) * sizeof(float2));
#pragma unroll
    for (int q = 0; q < NQ; q++) st_off<float2>(dst + (size_t)q * C::T * g.sx, woff, v[q]);
}

// ------------------------------------------- y transforms of length 2048
//
// L = T E with T = 64 threads per column and E = L / 64 = 32 registers each
// (the four-step split of ColTile needs T | E: 32 x 64, whose 64 registers and
// 16 KB twiddle table leave room for three 32-thread columns per CU).  Step
// one as in col_forward (thread t: E-point DFT of the rows t + T q, times
// W_L^(t k1)); after the LDS exchange, the T-point DFT of each k1 runs on a
// lane pair (p = k1, s): lane s takes the rows t = 2 m + s, an E-point DFT
// Y_s, and the pair combines X[k1 + E k'] = Y_0 + W_T^k' Y_1 (lane 0) and
// X[k1 + E (k' + T/2)] = Y_0 - W_T^k' Y_1 (lane 1) with one lane swap per
// value.  Lanes: tid = s + 2 (w + W p), t = 2 p + s, so the pair is adjacent
// (a DPP swap) and lanes of one row still cover W neighbouring columns.
// Twiddles W_L^j = A[j / E] B[j % E] from two tables of T and E entries.

template <int L> struct PairTile {
    static constexpr int T = 64;
    static constexpr int E = L / T;
    static constexpr int W = 4;                // columns per workgroup
    static constexpr int NT = T * W;           // 256 threads
    // row pitch of the exchange: the 32 lanes of a b64 access hit distinct
    // bank pairs in both directions (T W + 2 W = 2 W mod 32)
    static constexpr int S = T * W + 2 * W;
    static constexpr int LDS = (E * S + T + E) * 8;
    static_assert(L == 2048 && E == T / 2, "pair split: E = T / 2");
};

// swap values between lanes 2i and 2i+1 (DPP quad_perm [1, 0, 3, 2])
__device__ __forceinline__ float2 pair_swap(float2 v)
{
    float2 r;
    r.x = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.x), 0xB1, 0xF, 0xF, false));
    r.y = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.y), 0xB1, 0xF, 0xF, false));
    return r;
}

// W_L^j (forward sign), 0 <= j < L
template <int L>
__device__ __forceinline__ float2 pair_tw(const float2 *sa, const float2 *sb, int j)
{
    constexpr int E = PairTile<L>::E;
    return cmul(sa[j / E], sb[j % E]);
}

template <int L>
__device__ __forceinline__ void pair_tables(const float2 *__restrict__ tw, float2 *sa, float2 *sb)
{
    using C = PairTile<L>;
    for (int q = threadIdx.x; q < C::T; q += C::NT) sa[q] = tw[C::E * q];
    for (int q = threadIdx.x; q < C::E; q += C::NT) sb[q] = tw[q];
}

// yf of length 2048 on the pair split; FULL and the zero-padded rows as yf_fast
template <int L, bool FULL>
__global__ void __launch_bounds__(PairTile<L>::NT)
yf_pair(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float2 *__restrict__ xb,
        float2 *__restrict__ yb)
{
    using C = PairTile<L>;
    constexpr int T = C::T, E = C::E, NQ = FULL ? E : E / 2;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *sa = sm + E * C::S, *sb = sa + T;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, rest = tile / ntx, i = rest % g.n0, c = rest / g.n0;
    const int s = threadIdx.x & 1, w = (threadIdx.x >> 1) % C::W, p = (threadIdx.x >> 1) / C::W;
    const int t = 2 * p + s;
    pair_tables<L>(tw, sa, sb);
    const int kx0 = tx * C::W;
    const bool ok = kx0 + w < g.h2;
    const int col = ok ? kx0 + w : g.h2 - 1; // past h2: load in bounds, store nothing
    const float2 *src = xb + ((size_t)c * g.nrows + (size_t)i * g.n1) * g.sx;
    const unsigned roff = (unsigned)(((size_t)t * g.sx + col) * sizeof(float2));
    float2 v[E];
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = ld_off<float2>(src + (size_t)q * T * g.sx, roff);
    __syncthreads();
    if constexpr (FULL)
        fft_reg<E, -1>(v);
    else
        fft_reg_halfin<E, -1>(v);
#pragma unroll
    for (int k1 = 1; k1 < E; k1++) v[k1] = cmul(v[k1], pair_tw<L>(sa, sb, t * k1));
#pragma unroll
    for (int k1 = 0; k1 < E; k1++) sbuf[k1 * C::S + t * C::W + w] = v[k1];
    __syncthreads();
    float2 u[E];
#pragma unroll
    for (int m = 0; m < E; m++) u[m] = sbuf[p * C::S + (2 * m + s) * C::W + w];
    fft_reg<E, -1>(u);
#pragma unroll
    for (int k = 0; k < E; k++) {
        const float2 own = s ? cmul(u[k], sa[k]) : u[k]; // lane 1: W_T^k Y_1[k]
        const float2 oth = pair_swap(own);
        u[k] = s ? float2{oth.x - own.x, oth.y - own.y} : float2{own.x + oth.x, own.y + oth.y};
    }
    if (!ok) return;
    float2 *dst = yb + ((size_t)c * g.n0 + i) * (size_t)g.p1 * g.sx;
    const unsigned woff = (unsigned)(((size_t)(p + E * (T / 2) * s) * g.sx + col) * sizeof(float2));
#pragma unroll
    for (int k = 0; k < E; k++) st_off<float2>(dst + (size_t)E * k * g.sx, woff, u[k]);
}

// yi of length 2048 on the pair split: the steps of yf_pair reversed
template <int L, bool FULL>
__global__ void __launch_bounds__(PairTile<L>::NT)
yi_pair(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float2 *__restrict__ yb,
        float2 *__restrict__ xb)
{
    using C = PairTile<L>;
    constexpr int T = C::T, E = C::E, NQ = FULL ? E : E / 2;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *sa = sm + E * C::S, *sb = sa + T;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, rest = tile / ntx, i = rest % g.n0, c = rest / g.n0;
    const int s = threadIdx.x & 1, w = (threadIdx.x >> 1) % C::W, p = (threadIdx.x >> 1) / C::W;
    const int t = 2 * p + s;
    pair_tables<L>(tw, sa, sb);
    const int kx0 = tx * C::W;
    const bool ok = kx0 + w < g.h2;
    const int col = ok ? kx0 + w : g.h2 - 1;
    const float2 *src = yb + ((size_t)c * g.n0 + i) * (size_t)g.p1 * g.sx;
    const unsigned roff = (unsigned)(((size_t)(p + E * (T / 2) * s) * g.sx + col) * sizeof(float2));
    float2 u[E]; // lane s: X[p + E (k + s T/2)]
#pragma unroll
    for (int k = 0; k < E; k++) u[k] = ld_off<float2>(src + (size_t)E * k * g.sx, roff);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < E; k++) {
        const float2 oth = pair_swap(u[k]);
        // lane 0: X0 + X1; lane 1: (X0 - X1) W_T^-k
        u[k] = s ? cmulc(float2{oth.x - u[k].x, oth.y - u[k].y}, sa[k])
                 : float2{u[k].x + oth.x, u[k].y + oth.y};
    }
    fft_reg<E, +1>(u); // u[m]: row t = 2 m + s of step one's input, at k1 = p
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int tt = 2 * m + s;
        const float2 x = (m == 0 && s == 0) ? u[m] : cmulc(u[m], pair_tw<L>(sa, sb, tt * p));
        sbuf[p * C::S + tt * C::W + w] = x;
    }
    __syncthreads();
    float2 v[E];
#pragma unroll
    for (int k1 = 0; k1 < E; k1++) v[k1] = sbuf[k1 * C::S + t * C::W + w];
    fft_reg<E, +1>(v);
    if (!ok) return;
    float2 *dst = xb + ((size_t)c * g.nrows + (size_t)i * g.n1) * g.sx;
    const unsigned woff = (unsigned)(((size_t)t * g.sx + col) * sizeof(float2));
#pragma unroll
    for (int q = 0; q < NQ; q++) st_off<float2>(dst + (size_t)q * T * g.sx, woff, v[q]);
}

// z pass: three thread groups, one per field component, each transforms its
// component of W spectrum columns along z; the 3x3 kernel multiply runs
// through LDS, one octant kernel point for each pair kz, p0-kz.
template <int L> struct ZTile {
    static constexpr int T = Split<L>::T;
    static constexpr int E = Split<L>::E;
    static constexpr int G = E / T;
    // 8 columns for L = 128 (measured: 141 us vs 145 us with 16 at 256x256x64)
    // lengths with an odd factor and more than 16 registers: one wave per
    // component group (the three groups' LDS below 64 KB)
    static constexpr int W =
        (L == 2048) ? 2 : (L == 128 ? 8 : ((!is_pow2(L) && E > 16) ? 64 / T : 128 / T));
    static constexpr int S = lds_pitch(T, W, G);
    static constexpr int NG = T * W;          // threads per component group
    static constexpr int NT = 3 * NG;
    static constexpr int LDS = (3 * E * S + L) * 8;
    // spectrum rows (one kz, W columns) are exchanged by lanes (t, w) at rows
    // t*G + g: with G = 2 only every other row's banks would be used, so bit
    // 0 of the row is flipped by the row bit that advances one bank span
    static constexpr int SWZ = (G == 2 && W < 32) ? __builtin_ctz(32 / W) : -1;
    static_assert(NG % 64 == 0, "component groups must be whole waves");
    static_assert(NT % W == 0, "");
};

template <int L> __device__ __forceinline__ int zrow(int kz)
{
    if constexpr (ZTile<L>::SWZ >= 0)
        return kz ^ ((kz >> ZTile<L>::SWZ) & 1);
    else
        return kz;
}

// ky visiting order 0, p/2, 1, p-1, 2, p-2, ...: tiles of ky and p - ky
// (which read the same kernel octant rows) are neighbours
__device__ __forceinline__ int mirror_order(int i, int p)
{
    if (i == 0) return 0;
    if (p & 1) { // odd p: pairs (u, p-u) for u = 1 .. (p-1)/2
        const int u = (i - 1) / 2 + 1;
        return (i & 1) ? u : p - u;
    }
    if (i == 1) return p / 2;
    const int u = (i - 2) / 2 + 1;
    return (i & 1) ? p - u : u;
}

// FULL: periodic z (n0 == L, every plane holds data); otherwise n0 == L/2
// and planes n0 .. L-1 are zero padding (not read, not written).  Byte
// offsets from the wave-uniform bases stay below 4 GiB (fast_z_addressable).
template <int L, bool FULL>
__global__ void __launch_bounds__(ZTile<L>::NT)
z_fast(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float *__restrict__ K,
       float2 *__restrict__ yb, int film_y)
{
    using Z = ZTile<L>;
    constexpr int NQ = FULL ? Z::E : Z::E / 2; // planes t + T q held, q < NQ
    extern __shared__ float2 sm[];
    float2 *stw = sm + 3 * Z::E * Z::S;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, ky = mirror_order(tile / ntx, g.p1);
    const int kx0 = tx * Z::W;
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x / Z::NG);
    const int lt = threadIdx.x % Z::NG;
    const int w = lt % Z::W, t = lt / Z::W;
    float2 *sbuf = sm + c * Z::E * Z::S;
    for (int q = threadIdx.x; q < L; q += Z::NT) stw[q] = tw[q];
    const bool ok = kx0 + w < g.h2;
    // columns past h2 load column h2 - 1 (in bounds) and store nothing
    const int col = ok ? kx0 + w : g.h2 - 1;
    const size_t istride = (size_t)g.p1 * g.sx;
    float2 *cbase = yb + (size_t)c * g.n0 * istride + (size_t)ky * g.sx;
    const unsigned voff = (unsigned)(((size_t)t * istride + col) * sizeof(float2));
    const bool my = ky > g.p1 / 2;
    const int kyp = my ? g.p1 - ky : ky;
    const float sy = my ? -1.0f : 1.0f;
    // the kernel octant points this thread multiplies, loaded before the
    // spectrum so that their latency hides behind the loads and the forward
    // transform (when they fit in a few registers)
    constexpr int NITEM = ((L / 2 + 1) * Z::W + Z::NT - 1) / Z::NT;
    constexpr int NPF = (NITEM <= 4) ? NITEM : 1;
    float2 kpf[NPF][3];
    const float *kbase = K + 6 * ((size_t)kyp * g.h2 + kx0);
    const size_t kzs = (size_t)g.ko1 * g.h2;   // octant points per kz
    if constexpr (NITEM <= 4) {
        const int w2 = threadIdx.x % Z::W;
        const unsigned koff = (unsigned)(((threadIdx.x / Z::W) * kzs + w2) * 6 * sizeof(float));
#pragma unroll
        for (int r = 0; r < NITEM; r++) {
            const int item = threadIdx.x + r * Z::NT;
            const float *kb = kbase + 6 * (size_t)(r * (Z::NT / Z::W)) * kzs;
            if (item < (L / 2 + 1) * Z::W && kx0 + w2 < g.h2) {
                kpf[r][0] = ld_off<float2>(kb, koff);
                kpf[r][1] = ld_off<float2>(kb, koff + 8);
                kpf[r][2] = ld_off<float2>(kb, koff + 16);
            } else {
                kpf[r][0] = kpf[r][1] = kpf[r][2] = float2{0.0f, 0.0f};
            }
        }
    }
    float2 v[Z::E];
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = ld_off<float2>(cbase + (size_t)q * Z::T * istride, voff);
    __syncthreads();
    float2 u[Z::G][Z::T];
    col_forward<L, Z::W, Z::S, !FULL>(v, u, w, t, sbuf, stw);

    // spectrum of all three components through LDS: spec[c][zrow(kz)][w]
#pragma unroll
    for (int gg = 0; gg < Z::G; gg++)
#pragma unroll
        for (int k2 = 0; k2 < Z::T; k2++)
            sm[(c * L + zrow<L>(t * Z::G + gg + Z::E * k2)) * Z::W + w] = u[gg][k2];
    __syncthreads();
    auto multiply = [&](const Kern6 &k, int kzp, int w2) {
        const int r1 = zrow<L>(kzp);
        kern_apply(k, sm[(0 * L + r1) * Z::W + w2], sm[(1 * L + r1) * Z::W + w2],
                   sm[(2 * L + r1) * Z::W + w2]);
        if (kzp > 0 && kzp < L / 2) {
            // K at L - k: odd in the transformed axis, which holds XZ and XY
            // (internal axis 0) or, for a film's y pass (film_y), YZ and XY
            Kern6 km = k;
            if (film_y)
                km.yz = -km.yz;
            else
                km.xz = -km.xz;
            km.xy = -km.xy;
            const int r2 = zrow<L>(L - kzp);
            kern_apply(km, sm[(0 * L + r2) * Z::W + w2], sm[(1 * L + r2) * Z::W + w2],
                       sm[(2 * L + r2) * Z::W + w2]);
        }
    };
    if constexpr (NITEM <= 4) {
#pragma unroll
        for (int r = 0; r < NITEM; r++) {
            const int item = threadIdx.x + r * Z::NT;
            const int w2 = item % Z::W, kzp = item / Z::W;
            if (item >= (L / 2 + 1) * Z::W || kx0 + w2 >= g.h2) continue;
            // the layout of kern_octant (XX,YY | ZZ,YZ | XZ,XY), z sign +1
            Kern6 k;
            k.xx = kpf[r][0].x;
            k.yy = kpf[r][0].y;
            k.zz = kpf[r][1].x;
            k.yz = sy * kpf[r][1].y;
            k.xz = kpf[r][2].x;
            k.xy = sy * kpf[r][2].y;
            multiply(k, kzp, w2);
        }
    } else {
        for (int item = threadIdx.x; item < (L / 2 + 1) * Z::W; item += Z::NT) {
            const int w2 = item % Z::W, kzp = item / Z::W;
            if (kx0 + w2 >= g.h2) continue;
            multiply(kern_octant(g, K, kzp, kyp, kx0 + w2, 1.0f, sy), kzp, w2);
        }
    }
    __syncthreads();
#pragma unroll
    for (int gg = 0; gg < Z::G; gg++)
#pragma unroll
        for (int k2 = 0; k2 < Z::T; k2++)
            u[gg][k2] = sm[(c * L + zrow<L>(t * Z::G + gg + Z::E * k2)) * Z::W + w];
    __syncthreads();
    col_inverse<L, Z::W, Z::S>(u, v, w, t, sbuf, stw);
    if (ok) {
#pragma unroll
        for (int q = 0; q < NQ; q++) st_off<float2>(cbase + (size_t)q * Z::T * istride, voff, v[q]);
    }
}

// ---------------------------------------------------------- row transforms
//
// Thread (r, t): r = tid / T (row of the tile), t = tid % T.  Transpose
// buffer index of Y[t][k1] of row r: k1*S + r*T + t.

template <int LX, bool HALFIN = false>
__device__ __forceinline__ void row_forward(float2 (&v)[RowTile<LX>::E],
                                            float2 (&u)[RowTile<LX>::G][RowTile<LX>::T], int r,
                                            int t, float2 *sbuf, const float2 *stw)
{
    using C = RowTile<LX>;
    fs_first<LX, -1, HALFIN>(v, t, stw);
#pragma unroll
    for (int k1 = 0; k1 < C::E; k1++) sbuf[k1 * C::S + r * C::T + t] = v[k1];
    __syncthreads();
#pragma unroll
    for (int gg = 0; gg < C::G; gg++) {
#pragma unroll
        for (int t2 = 0; t2 < C::T; t2++) u[gg][t2] = sbuf[(t * C::G + gg) * C::S + r * C::T + t2];
        fft_reg<C::T, -1>(u[gg]);
    }
    __syncthreads();
}

template <int LX>
__device__ __forceinline__ void row_inverse(float2 (&u)[RowTile<LX>::G][RowTile<LX>::T],
                                            float2 (&v)[RowTile<LX>::E], int r, int t,
                                            float2 *sbuf, const float2 *stw)
{
    using C = RowTile<LX>;
#pragma unroll
    for (int gg = 0; gg < C::G; gg++) {
        fft_reg<C::T, +1>(u[gg]);
        const int k1 = t * C::G + gg;
#pragma unroll
        for (int t2 = 0; t2 < C::T; t2++) {
            const float2 x = (t2 == 0) ? u[gg][0] : cmulc(u[gg][t2], stw[t2 * k1]);
            sbuf[k1 * C::S + r * C::T + t2] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < C::E; k1++) v[k1] = sbuf[k1 * C::S + r * C::T + t];
    fft_reg<C::E, +1>(v);
    __syncthreads();
}

__device__ __forceinline__ float2 shfl2(float2 a, int lane)
{
    return float2{__shfl(a.x, lane, 64), __shfl(a.y, lane, 64)};
}

// conj(Z[LX - k]) for the value u[g][k2] held at k = t G + g + E k2 (all lanes
// of the wave must call this: it shuffles)
template <int LX, int GG, int K2>
__device__ __forceinline__ float2 mirror_conj(const float2 (&u)[RowTile<LX>::G][RowTile<LX>::T],
                                              int t, int lane_row0)
{
    using C = RowTile<LX>;
    constexpr int gsrc = (C::G - GG) % C::G;
    const int tl = (GG == 0) ? ((C::T - t) % C::T) : (C::T - 1 - t);
    float2 a = shfl2(u[gsrc][C::T - 1 - K2], lane_row0 + tl);
    if (GG == 0 && t == 0) a = u[0][(C::T - K2) % C::T];
    return cconj(a);
}

// xf: R2C along x of the n0*n1 data rows (half-length complex FFT).  FULL:
// periodic x (n2 == 2 LX, every pair holds data); otherwise n2 == LX and the
// pairs LX/2 .. LX-1 are zero padding (not read).  Lane offsets are 32-bit
// from the component bases (fast_x_addressable).
template <int LX, int PRED, bool FULL>
__global__ void __launch_bounds__(RowTile<LX>::NT)
xf_fast(DemagGeom g, const float2 *__restrict__ tw, const float2 *__restrict__ rtw,
        const float *__restrict__ m, const float *__restrict__ msat_map,
        const float *__restrict__ tau0, float h, const float *__restrict__ norm_map,
        float *__restrict__ mpred, float2 *__restrict__ xb, unsigned int *ho_dev,
        unsigned int *ho_host)
{
    using C = RowTile<LX>;
    if (ho_dev != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        // the previous launch's result word to the host, cleared for the next
        const unsigned int e = __hip_atomic_load(ho_dev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ho_host, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(ho_dev, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    constexpr int NQ = FULL ? C::E : C::E / 2; // cell pairs 2 (t + T q), q < NQ
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S, *srtw = stw + LX;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int r = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const int row = tile * C::R + r;
    const bool ok = row < g.nrows;
    const int crow = ok ? row : (int)g.nrows - 1; // rows past the end load the last row
    for (int q = threadIdx.x; q < LX; q += C::NT) stw[q] = tw[q];
    for (int q = threadIdx.x; q <= LX; q += C::NT) srtw[q] = rtw[q];
    const size_t N = g.N;
    const unsigned cb = (unsigned)(crow * g.n2 + 2 * t) * 4u;
    float2 v[3][C::E];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const unsigned o = cb + (unsigned)(q * 2 * C::T * 4);
        float a0[2], a1[2], a2[2];
        {
            const float2 x0 = ld_off<float2>(m, o), x1 = ld_off<float2>(m + N, o),
                         x2 = ld_off<float2>(m + 2 * N, o);
            a0[0] = x0.x; a0[1] = x0.y;
            a1[0] = x1.x; a1[1] = x1.y;
            a2[0] = x2.x; a2[1] = x2.y;
        }
        if (PRED) {
            const float2 p0 = ld_off<float2>(tau0, o), p1 = ld_off<float2>(tau0 + N, o),
                         p2 = ld_off<float2>(tau0 + 2 * N, o);
            const float2 nm = norm_map ? ld_off<float2>(norm_map, o) : float2{1.0f, 1.0f};
            const float pt0[2] = {p0.x, p0.y}, pt1[2] = {p1.x, p1.y}, pt2[2] = {p2.x, p2.y};
            const float nms[2] = {nm.x, nm.y};
#pragma unroll
            for (int s = 0; s < 2; s++) {
                a0[s] = a0[s] + h * pt0[s];
                a1[s] = a1[s] + h * pt1[s];
                a2[s] = a2[s] + h * pt2[s];
                const float len = sqrtf(a0[s] * a0[s] + a1[s] * a1[s] + a2[s] * a2[s]);
                float scale = (nms[s] == 0.0f) ? 0.0f : 1.0f;
                scale = (len != 0.0f) ? scale / len : 0.0f;
                a0[s] *= scale;
                a1[s] *= scale;
                a2[s] *= scale;
            }
            if (ok) {
                st_off<float2>(mpred, o, float2{a0[0], a0[1]});
                st_off<float2>(mpred + N, o, float2{a1[0], a1[1]});
                st_off<float2>(mpred + 2 * N, o, float2{a2[0], a2[1]});
            }
        }
        if (msat_map) {
            const float2 ms = ld_off<float2>(msat_map, o);
            a0[0] *= ms.x; a1[0] *= ms.x; a2[0] *= ms.x;
            a0[1] *= ms.y; a1[1] *= ms.y; a2[1] *= ms.y;
        }
        v[0][q] = float2{a0[0], a0[1]};
        v[1][q] = float2{a1[0], a1[1]};
        v[2][q] = float2{a2[0], a2[1]};
    }
    __syncthreads();
    float2 u[3][C::G][C::T];
#pragma unroll
    for (int c = 0; c < 3; c++) row_forward<LX, !FULL>(v[c], u[c], r, t, sbuf, stw);

    const int lane_row0 = (threadIdx.x & 63) - t;
    const unsigned ob = (unsigned)((size_t)crow * g.sx * sizeof(float2));
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float2 *dst = xb + (size_t)c * g.nrows * g.sx;
        static_for<C::G * C::T>([&](auto I) {
            constexpr int gg = decltype(I)::value / C::T, k2 = decltype(I)::value % C::T;
            const float2 b = mirror_conj<LX, gg, k2>(u[c], t, lane_row0);
            const float2 a = u[c][gg][k2];
            const int k = t * C::G + gg + C::E * k2;
            const float2 Ev = float2{0.5f * (a.x + b.x), 0.5f * (a.y + b.y)};
            const float2 Ov = float2{0.5f * (a.y - b.y), -0.5f * (a.x - b.x)};
            const float2 X = cadd(Ev, cmul(Ov, srtw[k]));
            if (ok) st_off<float2>(dst, ob + (unsigned)k * 8u, X);
            if (ok && k == 0) {
                // Nyquist bin: E(0) - O(0) = Re Z0 - Im Z0
                st_off<float2>(dst, ob + (unsigned)LX * 8u, float2{a.x - a.y, 0.0f});
            }
        });
    }
}

// Two neighbouring cells (k0 even, k0 + 1 < n2) of one row at once, for
// uniform material parameters (MAPS false) or the Msat, Aex, alpha and Ku
// maps, the B_ext and anisU masks, the thermal field and H_dex (MAPS true),
// and no spin-transfer term: the same per-cell arithmetic as cell_epilogue
// (exchange6 order dim 0, dim 2, dim 1; SumUpdater order), with 8-byte loads
// and stores of the cell pair.

__device__ __forceinline__ int nb_lo(int pos, int len, int wrap)
{
    int lo = pos - 1;
    lo = (lo < 0 && wrap) ? len + lo : lo;
    return max(lo, 0);
}
__device__ __forceinline__ int nb_hi(int pos, int len, int wrap)
{
    int hi = pos + 1;
    hi = (hi == len && wrap) ? hi - len : hi;
    return min(hi, len - 1);
}

// Aex/Msat of the cell pair at byte offset b (div_zero(mask_unity(aex),
// mask_unity(msat)) of exchange6_at), and of one cell
__device__ __forceinline__ float2 lex_pair(const FieldParams &fp, unsigned b)
{
    const float2 a = fp.aex_map ? ld_off<float2>(fp.aex_map, b) : float2{1.0f, 1.0f};
    const float2 s = fp.msat_map ? ld_off<float2>(fp.msat_map, b) : float2{1.0f, 1.0f};
    return float2{div_zero(a.x, s.x), div_zero(a.y, s.y)};
}
__device__ __forceinline__ float lex_one(const FieldParams &fp, unsigned b)
{
    const float a = fp.aex_map ? ld_off<float>(fp.aex_map, b) : 1.0f;
    const float s = fp.msat_map ? ld_off<float>(fp.msat_map, b) : 1.0f;
    return div_zero(a, s);
}

template <int MODE, bool MAPS = false>
__device__ __forceinline__ void pair_epilogue(const DemagGeom &g, const FieldParams &fp,
                                              const EpilogueIO &io, unsigned I0, int i, int j,
                                              int k0, const float hd0[3], const float hd1[3],
                                              float &errmax)
{
    const size_t N = g.N;
    const unsigned b0 = I0 * 4u; // byte offset of cell I0 in a component array
    if (MODE == EPI_B) {
#pragma unroll
        for (int c = 0; c < 3; c++) st_off<float2>(io.hout + c * N, b0, float2{hd0[c], hd1[c]});
        return;
    }
    const float *__restrict__ m = io.m;
    float m0[3], m1[3], H0[3], H1[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float2 a = ld_off<float2>(m + c * N, b0);
        m0[c] = a.x;
        m1[c] = a.y;
    }
    // neighbour byte offsets: dim 0 (planes), dim 2 (along the row), dim 1
    // (rows); negative steps wrap in the unsigned sum and land on valid cells
    const int s0 = g.n1 * g.n2, s1 = g.n2;
    const unsigned o0l = b0 + 4u * (unsigned)((nb_lo(i, g.n0, fp.pbc[0]) - i) * s0);
    const unsigned o0h = b0 + 4u * (unsigned)((nb_hi(i, g.n0, fp.pbc[0]) - i) * s0);
    const unsigned o1l = b0 + 4u * (unsigned)((nb_lo(j, g.n1, fp.pbc[1]) - j) * s1);
    const unsigned o1h = b0 + 4u * (unsigned)((nb_hi(j, g.n1, fp.pbc[1]) - j) * s1);
    const unsigned o2l = b0 + 4u * (unsigned)(nb_lo(k0, g.n2, fp.pbc[2]) - k0); // left of cell 0
    const unsigned o2h = b0 + 4u * (unsigned)(nb_hi(k0 + 1, g.n2, fp.pbc[2]) - k0); // right of 1
    const float p0 = fp.ex_pre * fp.ex_c[0], p1 = fp.ex_pre * fp.ex_c[1],
                p2 = fp.ex_pre * fp.ex_c[2];
    if constexpr (MAPS) {
        // exchange weights: harmonic mean of the pair's and each neighbour's
        // Aex/Msat (w<dim><lo|hi><cell>)
        const float2 lc = lex_pair(fp, b0);
        const float2 lzl = lex_pair(fp, o0l), lzh = lex_pair(fp, o0h);
        const float2 lyl = lex_pair(fp, o1l), lyh = lex_pair(fp, o1h);
        const float lxl = lex_one(fp, o2l), lxh = lex_one(fp, o2h);
        const float w0l0 = avg_geom_zero(lc.x, lzl.x), w0h0 = avg_geom_zero(lc.x, lzh.x);
        const float w0l1 = avg_geom_zero(lc.y, lzl.y), w0h1 = avg_geom_zero(lc.y, lzh.y);
        const float w2l0 = avg_geom_zero(lc.x, lxl), w2h0 = avg_geom_zero(lc.x, lc.y);
        const float w2l1 = avg_geom_zero(lc.y, lc.x), w2h1 = avg_geom_zero(lc.y, lxh);
        const float w1l0 = avg_geom_zero(lc.x, lyl.x), w1h0 = avg_geom_zero(lc.x, lyh.x);
        const float w1l1 = avg_geom_zero(lc.y, lyl.y), w1h1 = avg_geom_zero(lc.y, lyh.y);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float *mc = m + c * N;
            const float2 zl = ld_off<float2>(mc, o0l), zh = ld_off<float2>(mc, o0h);
            const float2 yl = ld_off<float2>(mc, o1l), yh = ld_off<float2>(mc, o1h);
            const float xl = ld_off<float>(mc, o2l), xh = ld_off<float>(mc, o2h);
            float e0 = p0 * (w0l0 * (zl.x - m0[c]) + w0h0 * (zh.x - m0[c]));
            float e1 = p0 * (w0l1 * (zl.y - m1[c]) + w0h1 * (zh.y - m1[c]));
            e0 += p2 * (w2l0 * (xl - m0[c]) + w2h0 * (m1[c] - m0[c]));
            e1 += p2 * (w2l1 * (m0[c] - m1[c]) + w2h1 * (xh - m1[c]));
            e0 += p1 * (w1l0 * (yl.x - m0[c]) + w1h0 * (yh.x - m0[c]));
            e1 += p1 * (w1l1 * (yl.y - m1[c]) + w1h1 * (yh.y - m1[c]));
            // B_ext mask: multiplier * map (the SumUpdater's vecMadd)
            float hx0 = fp.hext[c], hx1 = fp.hext[c];
            if (fp.bext_map) {
                const float2 bm = ld_off<float2>(fp.bext_map + c * N, b0);
                hx0 = fp.hext[c] * bm.x;
                hx1 = fp.hext[c] * bm.y;
            }
            H0[c] = hd0[c] + e0 + hx0;
            H1[c] = hd1[c] + e1 + hx1;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float *mc = m + c * N;
            const float2 zl = ld_off<float2>(mc, o0l), zh = ld_off<float2>(mc, o0h);
            const float2 yl = ld_off<float2>(mc, o1l), yh = ld_off<float2>(mc, o1h);
            const float xl = ld_off<float>(mc, o2l), xh = ld_off<float>(mc, o2h);
            float e0 = p0 * ((zl.x - m0[c]) + (zh.x - m0[c]));
            float e1 = p0 * ((zl.y - m1[c]) + (zh.y - m1[c]));
            e0 += p2 * ((xl - m0[c]) + (m1[c] - m0[c]));
            e1 += p2 * ((m0[c] - m1[c]) + (xh - m1[c]));
            e0 += p1 * ((yl.x - m0[c]) + (yh.x - m0[c]));
            e1 += p1 * ((yl.y - m1[c]) + (yh.y - m1[c]));
            H0[c] = hd0[c] + e0 + fp.hext[c];
            H1[c] = hd1[c] + e1 + fp.hext[c];
        }
    }
    // Msat of the pair (a zero Msat reads as 1 in the anisotropy constant and
    // masks the renormalised cell)
    float2 ms = float2{1.0f, 1.0f};
    if constexpr (MAPS) {
        if (fp.msat_map) ms = ld_off<float2>(fp.msat_map, b0);
    }
    if (fp.anis) {
        float k = fp.ku2, k1 = fp.ku2;
        if constexpr (MAPS) {
            k = fp.ku2 / (ms.x == 0.0f ? 1.0f : ms.x);
            k1 = fp.ku2 / (ms.y == 0.0f ? 1.0f : ms.y);
            if (fp.ku_map) {
                const float2 q = ld_off<float2>(fp.ku_map, b0);
                k *= q.x;
                k1 *= q.y;
            }
        }
        // the anisU mask gives each cell's axis as multiplier * map
        float u0[3] = {fp.u[0], fp.u[1], fp.u[2]}, u1[3] = {fp.u[0], fp.u[1], fp.u[2]};
        if constexpr (MAPS) {
            if (fp.u_map) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float2 um = ld_off<float2>(fp.u_map + c * N, b0);
                    u0[c] = fp.u[c] * um.x;
                    u1[c] = fp.u[c] * um.y;
                }
            }
        }
        const float mu0 = m0[0] * u0[0] + m0[1] * u0[1] + m0[2] * u0[2];
        const float mu1 = m1[0] * u1[0] + m1[1] * u1[1] + m1[2] * u1[2];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            H0[c] += k * mu0 * u0[c];
            H1[c] += k1 * mu1 * u1[c];
        }
    }
    // H_therm, the SumUpdater's last term, then H_dex from its own
    // convolution (demagexch with an Msat mask)
    if constexpr (MAPS) {
        if (fp.htherm) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float2 t = ld_off<float2>(fp.htherm + c * N, b0);
                H0[c] += t.x;
                H1[c] += t.y;
            }
        }
        if (fp.hdex) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float2 t = ld_off<float2>(fp.hdex + c * N, b0);
                H0[c] += t.x;
                H1[c] += t.y;
            }
        }
    }
    if (io.hout) {
#pragma unroll
        for (int c = 0; c < 3; c++) st_off<float2>(io.hout + c * N, b0, float2{H0[c], H1[c]});
    }
    if (MODE == EPI_HEFF && io.tout == nullptr) return;
    float t0[3], t1[3];
    float2 alpha = float2{fp.alpha, fp.alpha};
    if constexpr (MAPS) {
        if (fp.alpha_map) {
            const float2 a = ld_off<float2>(fp.alpha_map, b0);
            alpha = float2{fp.alpha * a.x, fp.alpha * a.y};
        }
    }
    llg_torque(m0, H0, alpha.x, t0);
    llg_torque(m1, H1, alpha.y, t1);
    if (io.tout) {
#pragma unroll
        for (int c = 0; c < 3; c++) st_off<float2>(io.tout + c * N, b0, float2{t0[c], t1[c]});
    }
    if (MODE == EPI_CORR) {
        const float hh = 0.5f * io.h, mh = -0.5f * io.h;
        float a0[3], a1[3];
        float d = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float2 p = ld_off<float2>(io.tau0 + c * N, b0);
            a0[c] = m0[c] + (hh * t0[c] + mh * p.x);
            a1[c] = m1[c] + (hh * t1[c] + mh * p.y);
            d = fmaxf(d, fmaxf(fabsf(t0[c] - p.x), fabsf(t1[c] - p.y)));
        }
        errmax = fmaxf(errmax, (d != d) ? __builtin_inff() : d);
        const float l0 = sqrtf(a0[0] * a0[0] + a0[1] * a0[1] + a0[2] * a0[2]);
        const float l1 = sqrtf(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
        const float s0 = (ms.x == 0.0f) ? 0.0f : 1.0f, s1 = (ms.y == 0.0f) ? 0.0f : 1.0f;
        const float q0 = (l0 != 0.0f) ? s0 / l0 : 0.0f;
        const float q1 = (l1 != 0.0f) ? s1 / l1 : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++)
            st_off<float2>(io.mnew + c * N, b0, float2{a0[c] * q0, a1[c] * q1});
    }
}

// xi: C2R along x (imaginary parts of DC and Nyquist ignored), keep k < n2,
// then the fused field / torque / RK epilogue, one of three variants:
// XI_UNIFORM the pair epilogue for uniform parameters, XI_MAPS the pair
// epilogue with maps and masks (Msat, Aex, alpha, Ku, B_ext, anisU), the
// thermal field and H_dex, XI_GENERAL the per-cell epilogue (cell_epilogue:
// all of these plus the Zhang-Li torque).  Each is
// its own kernel: the general epilogue's register demand would otherwise set
// the occupancy of the pair variants.
enum XiVariant { XI_UNIFORM = 0, XI_MAPS = 1, XI_GENERAL = 2 };

// Minimum waves per SIMD the xi kernel is built for.  The map and general
// epilogues of the power-of-two lengths up to 256 take all 256 VGPRs (one
// wave per SIMD) when left alone; bounded to two waves they spill a few
// dozen registers and still run faster (256-point rows at 256x256x64: Msat
// map 146 -> 103 us, Zhang-Li 136 -> 127 us).  Longer rows, and the other
// lengths, spill hundreds at two waves.
template <int LX, int VAR> constexpr int xi_min_waves()
{
    return (VAR != XI_UNIFORM && LX <= 256 && (LX & (LX - 1)) == 0) ? 2 : 1;
}

template <int LX, int MODE, bool FULL, int VAR>
__global__ void __launch_bounds__(RowTile<LX>::NT)
    __attribute__((amdgpu_waves_per_eu(xi_min_waves<LX, VAR>())))
xi_fast(DemagGeom g, const float2 *__restrict__ tw, const float2 *__restrict__ rtw,
        const float2 *__restrict__ xb, FieldParams fp, EpilogueIO io)
{
    using C = RowTile<LX>;
    constexpr int NQ = FULL ? C::E : C::E / 2; // cell pairs 2 (t + T q) < n2
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S, *srtw = stw + LX;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int r = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const int row = tile * C::R + r;
    const bool ok = row < g.nrows;
    const int crow = ok ? row : (int)g.nrows - 1;
    for (int q = threadIdx.x; q < LX; q += C::NT) stw[q] = tw[q];
    for (int q = threadIdx.x; q <= LX; q += C::NT) srtw[q] = rtw[q];
    __syncthreads();
    const int lane_row0 = (threadIdx.x & 63) - t;
    const unsigned ob = (unsigned)((size_t)crow * g.sx * sizeof(float2));
    float2 v[3][C::E];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float2 *src = xb + (size_t)c * g.nrows * g.sx;
        float2 u[C::G][C::T];
#pragma unroll
        for (int gg = 0; gg < C::G; gg++)
#pragma unroll
            for (int k2 = 0; k2 < C::T; k2++) {
                const int k = t * C::G + gg + C::E * k2;
                float2 x = ld_off<float2>(src, ob + (unsigned)k * 8u);
                if (k == 0) x.y = 0.0f;
                u[gg][k2] = x;
            }
        const float2 nyq = float2{ld_off<float2>(src, ob + (unsigned)LX * 8u).x, 0.0f};
        float2 z[C::G][C::T];
        static_for<C::G * C::T>([&](auto I) {
            constexpr int gg = decltype(I)::value / C::T, k2 = decltype(I)::value % C::T;
            float2 b = mirror_conj<LX, gg, k2>(u, t, lane_row0);
            const int k = t * C::G + gg + C::E * k2;
            if (k == 0) b = nyq; // conj(X[LX]) with the imaginary part dropped
            const float2 a = u[gg][k2];
            const float2 Ev = cadd(a, b);
            const float2 Ov = cmulc(csub(a, b), srtw[k]);
            z[gg][k2] = float2{Ev.x - Ov.y, Ev.y + Ov.x};
        });
        row_inverse<LX>(z, v[c], r, t, sbuf, stw);
    }

    float errmax = 0.0f;
    // cell pairs (n2 is even on this path: n2 = LX padded to 2 LX, or periodic
    // with n2 = 2 LX), with 8-byte I/O in the pair variants
    if (ok) {
        const int i = row / g.n1, j = row - i * g.n1;
        // the pairs' loads are left to the scheduler: a scheduling barrier per
        // pair (one pair's loads in flight at a time) measured 3 % slower
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int k0 = 2 * (t + C::T * q);
            float hd0[3], hd1[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                hd0[c] = fp.demag_scale * v[c][q].x;
                hd1[c] = fp.demag_scale * v[c][q].y;
            }
            if (VAR == XI_GENERAL && MODE != EPI_B) {
                // both cells' outputs stored as float2 pairs, as the
                // two-cell variants store them
                const long I0 = (long)row * g.n2 + k0;
                CellOut o0, o1;
                cell_epilogue<MODE>(g, fp, io, I0, i, j, k0, hd0, errmax, &o0);
                cell_epilogue<MODE>(g, fp, io, I0 + 1, i, j, k0 + 1, hd1, errmax, &o1);
                const size_t N = g.N;
                const unsigned b0 = (unsigned)I0 * 4u;
                if (io.hout) {
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        st_off<float2>(io.hout + c * N, b0, float2{o0.h[c], o1.h[c]});
                }
                if (io.tout) {
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        st_off<float2>(io.tout + c * N, b0, float2{o0.t[c], o1.t[c]});
                }
                if (MODE == EPI_CORR) {
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        st_off<float2>(io.mnew + c * N, b0, float2{o0.mn[c], o1.mn[c]});
                }
            } else {
                pair_epilogue<MODE, VAR == XI_MAPS>(g, fp, io, (unsigned)(row * g.n2 + k0), i, j,
                                                    k0, hd0, hd1, errmax);
            }
        }
    }
    if (MODE == EPI_CORR) block_max_to(errmax, io.err);
}

// ------------------------------------------------------- launch helpers

template <typename K> static void allow_lds_fast(K kernel, size_t bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}


// Launchers of one transform length (defined in fast_launch.hpp, instantiated
// by length in the files that include it).
template <int L>
int xf_len(const DemagPlanImpl *P, int pred, const float *m, const float *msat_map,
           const float *tau0, float h, const float *norm_map, float *mpred, hipStream_t st);
template <int L> int yf_len(const DemagPlanImpl *P, hipStream_t st);
template <int L> int yi_len(const DemagPlanImpl *P, hipStream_t st);
template <int L> int z_len(const DemagPlanImpl *P, hipStream_t st);
template <int L>
int z_len_geom(const DemagGeom &g, const float2 *tw, const float *K, float2 *buf, hipStream_t st);
template <int L, int VAR>
int xi_len(const DemagPlanImpl *P, int mode, const FieldParams &fp, const EpilogueIO &io,
           hipStream_t st);

// the lengths of each pass (fast_pow2_length, fast_odd_short, fast_odd_long)
#define MMX_POW2_LENGTHS(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)
#define MMX_R3_SHORT(X) X(24) X(48) X(96) X(192)
#define MMX_R5_SHORT(X) X(40) X(80)
#define MMX_R7_SHORT(X) X(56) X(112)
#define MMX_R3_LONG(X) X(384)
#define MMX_R5_LONG(X) X(160) X(320)
#define MMX_R7_LONG(X) X(224) X(448)
#define MMX_SHORT_LENGTHS(X) MMX_POW2_LENGTHS(X) MMX_R3_SHORT(X) MMX_R5_SHORT(X) MMX_R7_SHORT(X)
#define MMX_ALL_LENGTHS(X) MMX_SHORT_LENGTHS(X) MMX_R3_LONG(X) MMX_R5_LONG(X) MMX_R7_LONG(X)

} // namespace mmx
