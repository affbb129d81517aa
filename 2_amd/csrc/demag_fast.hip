// demag_fast.hip -- the demag passes for power-of-two transform lengths.
//
// Same five passes as demag.hip (xf, yf, z, yi, xi; films: xf, y2d, xi) with
// the FFT held in registers (four-step split, reg_fft.hpp): each pass reads its
// input once and writes its output once, with one LDS transpose per transform
// direction.  Strided passes work on tiles of W spectrum columns (W >= 16, so
// each row segment is at least 128 bytes on 128-byte aligned rows); blocks are
// mapped to tiles XCD by XCD (xcd_tile).
#include "epilogue.hpp"
#include "reg_fft.hpp"

namespace mmx {

// Work split of a strided (y or z) pass: T threads per line, E registers, W
// lines (spectrum columns) per workgroup, LDS row pitch S.
template <int L> struct ColTile {
    static constexpr int T = Split<L>::T;
    static constexpr int E = Split<L>::E;
    static constexpr int G = E / T;
    static constexpr int W = (L == 2048) ? 4 : (L == 1024 ? 8 : 256 / T);
    static constexpr int S = T * W + 8;
    static constexpr int NT = T * W;                     // threads
    static constexpr int LDS = (E * S + L) * 8;           // transpose + twiddles
};

// Row pass (x): R rows per workgroup, T threads per row.
template <int LX> struct RowTile {
    static constexpr int T = Split<LX>::T;
    static constexpr int E = Split<LX>::E;
    static constexpr int G = E / T;
    static constexpr int R = (LX == 2048) ? 4 : (LX == 1024 ? 8 : 256 / T);
    static constexpr int S = T * R + 8;
    static constexpr int NT = T * R;
    static constexpr int LDS = (E * S + 2 * LX + 1) * 8;
};

struct Kern6f {
    float xx, yy, zz, yz, xz, xy;
};

__device__ __forceinline__ Kern6f kern6_at(const DemagGeom &g, const float *__restrict__ K, int kz,
                                           int ky, int kx)
{
    const bool mz = kz > g.p0 / 2, my = ky > g.p1 / 2;
    const int kzp = mz ? g.p0 - kz : kz;
    const int kyp = my ? g.p1 - ky : ky;
    const float sz = mz ? -1.0f : 1.0f, sy = my ? -1.0f : 1.0f;
    const long plane = (long)g.ko0 * g.ko1 * g.h2;
    const long o = ((long)kzp * g.ko1 + kyp) * g.h2 + kx;
    Kern6f r;
    r.xx = K[o];
    r.yy = K[plane + o];
    r.zz = K[2 * plane + o];
    r.yz = sy * K[3 * plane + o];
    r.xz = sz * K[4 * plane + o];
    r.xy = sz * sy * K[5 * plane + o];
    return r;
}

__device__ __forceinline__ void kern6_apply(const Kern6f &k, float2 &a0, float2 &a1, float2 &a2)
{
    const float2 s0 = a0, s1 = a1, s2 = a2;
    a0 = float2{k.xx * s0.x + k.xy * s1.x + k.xz * s2.x, k.xx * s0.y + k.xy * s1.y + k.xz * s2.y};
    a1 = float2{k.xy * s0.x + k.yy * s1.x + k.yz * s2.x, k.xy * s0.y + k.yy * s1.y + k.yz * s2.y};
    a2 = float2{k.xz * s0.x + k.yz * s1.x + k.zz * s2.x, k.xz * s0.y + k.yz * s1.y + k.zz * s2.y};
}

// ---------------------------------------------------- column transforms
//
// Thread (w, t) of a column tile: w = tid % W (column), t = tid / W.
// Transpose buffer index of Y[t][k1] of column w: k1*S + t*W + w.

template <int L>
__device__ __forceinline__ void col_forward(float2 (&v)[ColTile<L>::E],
                                            float2 (&u)[ColTile<L>::G][ColTile<L>::T], int w,
                                            int t, float2 *sbuf, const float2 *stw)
{
    using C = ColTile<L>;
    fs_first<L, -1>(v, t, stw);
#pragma unroll
    for (int k1 = 0; k1 < C::E; k1++) sbuf[k1 * C::S + t * C::W + w] = v[k1];
    __syncthreads();
#pragma unroll
    for (int gg = 0; gg < C::G; gg++) {
#pragma unroll
        for (int t2 = 0; t2 < C::T; t2++) u[gg][t2] = sbuf[(t * C::G + gg) * C::S + t2 * C::W + w];
        fft_reg<C::T, -1>(u[gg]);
    }
    __syncthreads();
}

template <int L>
__device__ __forceinline__ void col_inverse(float2 (&u)[ColTile<L>::G][ColTile<L>::T],
                                            float2 (&v)[ColTile<L>::E], int w, int t,
                                            float2 *sbuf, const float2 *stw)
{
    using C = ColTile<L>;
#pragma unroll
    for (int gg = 0; gg < C::G; gg++) {
        fft_reg<C::T, +1>(u[gg]);
        const int k1 = t * C::G + gg;
#pragma unroll
        for (int t2 = 0; t2 < C::T; t2++) {
            const float2 x = (t2 == 0) ? u[gg][0] : cmulc(u[gg][t2], stw[t2 * k1]);
            sbuf[k1 * C::S + t2 * C::W + w] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < C::E; k1++) v[k1] = sbuf[k1 * C::S + t * C::W + w];
    fft_reg<C::E, +1>(v);
    __syncthreads();
}

// yf: FFT along y of the n0 nonzero planes (rows j >= n1 are zeros)
template <int L>
__global__ void __launch_bounds__(ColTile<L>::NT)
yf_fast(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float2 *__restrict__ xb,
        float2 *__restrict__ yb)
{
    using C = ColTile<L>;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, rest = tile / ntx, i = rest % g.n0, c = rest / g.n0;
    const int kx0 = tx * C::W;
    const int w = threadIdx.x % C::W, t = threadIdx.x / C::W;
    for (int q = threadIdx.x; q < L; q += C::NT) stw[q] = tw[q];
    const bool ok = kx0 + w < g.h2;
    const float2 *src = xb + ((long)c * g.nrows + (long)i * g.n1) * g.sx + kx0 + w;
    float2 v[C::E];
#pragma unroll
    for (int q = 0; q < C::E; q++) {
        const int j = t + C::T * q;
        v[q] = (ok && j < g.n1) ? src[(long)j * g.sx] : float2{0.0f, 0.0f};
    }
    __syncthreads();
    float2 u[C::G][C::T];
    col_forward<L>(v, u, w, t, sbuf, stw);
    if (!ok) return;
    float2 *dst = yb + ((long)c * g.n0 + i) * (long)g.p1 * g.sx + kx0 + w;
#pragma unroll
    for (int gg = 0; gg < C::G; gg++)
#pragma unroll
        for (int k2 = 0; k2 < C::T; k2++) dst[(long)(t * C::G + gg + C::E * k2) * g.sx] = u[gg][k2];
}

// yi: inverse FFT along y, keep j < n1
template <int L>
__global__ void __launch_bounds__(ColTile<L>::NT)
yi_fast(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float2 *__restrict__ yb,
        float2 *__restrict__ xb)
{
    using C = ColTile<L>;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, rest = tile / ntx, i = rest % g.n0, c = rest / g.n0;
    const int kx0 = tx * C::W;
    const int w = threadIdx.x % C::W, t = threadIdx.x / C::W;
    for (int q = threadIdx.x; q < L; q += C::NT) stw[q] = tw[q];
    const bool ok = kx0 + w < g.h2;
    const float2 *src = yb + ((long)c * g.n0 + i) * (long)g.p1 * g.sx + kx0 + w;
    float2 u[C::G][C::T];
#pragma unroll
    for (int gg = 0; gg < C::G; gg++)
#pragma unroll
        for (int k2 = 0; k2 < C::T; k2++)
            u[gg][k2] = ok ? src[(long)(t * C::G + gg + C::E * k2) * g.sx] : float2{0.0f, 0.0f};
    __syncthreads();
    float2 v[C::E];
    col_inverse<L>(u, v, w, t, sbuf, stw);
    if (!ok) return;
    float2 *dst = xb + ((long)c * g.nrows + (long)i * g.n1) * g.sx + kx0 + w;
#pragma unroll
    for (int q = 0; q < C::E; q++) {
        const int j = t + C::T * q;
        if (j < g.n1) dst[(long)j * g.sx] = v[q];
    }
}

// z: FFT along z (planes i >= n0 are zeros), kernel multiply, inverse, keep i < n0
template <int L>
__global__ void __launch_bounds__(ColTile<L>::NT)
z_fast(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float *__restrict__ K,
       float2 *__restrict__ yb)
{
    using C = ColTile<L>;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, ky = tile / ntx;
    const int kx0 = tx * C::W;
    const int w = threadIdx.x % C::W, t = threadIdx.x / C::W;
    for (int q = threadIdx.x; q < L; q += C::NT) stw[q] = tw[q];
    const bool ok = kx0 + w < g.h2;
    const long cstride = (long)g.n0 * g.p1 * g.sx, istride = (long)g.p1 * g.sx;
    float2 *base = yb + (long)ky * g.sx + kx0 + w;
    float2 u[3][C::G][C::T];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float2 v[C::E];
#pragma unroll
        for (int q = 0; q < C::E; q++) {
            const int i = t + C::T * q;
            v[q] = (ok && i < g.n0) ? base[c * cstride + i * istride] : float2{0.0f, 0.0f};
        }
        col_forward<L>(v, u[c], w, t, sbuf, stw);
    }
    if (ok) {
#pragma unroll
        for (int gg = 0; gg < C::G; gg++)
#pragma unroll
            for (int k2 = 0; k2 < C::T; k2++) {
                const Kern6f k = kern6_at(g, K, t * C::G + gg + C::E * k2, ky, kx0 + w);
                kern6_apply(k, u[0][gg][k2], u[1][gg][k2], u[2][gg][k2]);
            }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float2 v[C::E];
        col_inverse<L>(u[c], v, w, t, sbuf, stw);
        if (ok) {
#pragma unroll
            for (int q = 0; q < C::E; q++) {
                const int i = t + C::T * q;
                if (i < g.n0) base[c * cstride + i * istride] = v[q];
            }
        }
    }
}

// films (p0 == 1): y transform, kernel multiply and inverse y in one kernel
template <int L>
__global__ void __launch_bounds__(ColTile<L>::NT)
y2d_fast(DemagGeom g, int ntx, const float2 *__restrict__ tw, const float *__restrict__ K,
         float2 *__restrict__ xb)
{
    using C = ColTile<L>;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int kx0 = tile * C::W;
    const int w = threadIdx.x % C::W, t = threadIdx.x / C::W;
    for (int q = threadIdx.x; q < L; q += C::NT) stw[q] = tw[q];
    const bool ok = kx0 + w < g.h2;
    const long cstride = (long)g.nrows * g.sx;
    float2 *base = xb + kx0 + w;
    float2 u[3][C::G][C::T];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float2 v[C::E];
#pragma unroll
        for (int q = 0; q < C::E; q++) {
            const int j = t + C::T * q;
            v[q] = (ok && j < g.n1) ? base[c * cstride + (long)j * g.sx] : float2{0.0f, 0.0f};
        }
        col_forward<L>(v, u[c], w, t, sbuf, stw);
    }
    if (ok) {
#pragma unroll
        for (int gg = 0; gg < C::G; gg++)
#pragma unroll
            for (int k2 = 0; k2 < C::T; k2++) {
                const Kern6f k = kern6_at(g, K, 0, t * C::G + gg + C::E * k2, kx0 + w);
                kern6_apply(k, u[0][gg][k2], u[1][gg][k2], u[2][gg][k2]);
            }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float2 v[C::E];
        col_inverse<L>(u[c], v, w, t, sbuf, stw);
        if (ok) {
#pragma unroll
            for (int q = 0; q < C::E; q++) {
                const int j = t + C::T * q;
                if (j < g.n1) base[c * cstride + (long)j * g.sx] = v[q];
            }
        }
    }
}

// ---------------------------------------------------------- row transforms
//
// Thread (r, t): r = tid / T (row of the tile), t = tid % T.  Transpose
// buffer index of Y[t][k1] of row r: k1*S + r*T + t.

template <int LX>
__device__ __forceinline__ void row_forward(float2 (&v)[RowTile<LX>::E],
                                            float2 (&u)[RowTile<LX>::G][RowTile<LX>::T], int r,
                                            int t, float2 *sbuf, const float2 *stw)
{
    using C = RowTile<LX>;
    fs_first<LX, -1>(v, t, stw);
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

// xf: R2C along x of the n0*n1 data rows (half-length complex FFT)
template <int LX, int PRED>
__global__ void __launch_bounds__(RowTile<LX>::NT)
xf_fast(DemagGeom g, const float2 *__restrict__ tw, const float2 *__restrict__ rtw,
        const float *__restrict__ m, const float *__restrict__ msat_map,
        const float *__restrict__ tau0, float h, const float *__restrict__ norm_map,
        float *__restrict__ mpred, float2 *__restrict__ xb)
{
    using C = RowTile<LX>;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S, *srtw = stw + LX;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int r = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const long row = (long)tile * C::R + r;
    const bool ok = row < g.nrows;
    for (int q = threadIdx.x; q < LX; q += C::NT) stw[q] = tw[q];
    for (int q = threadIdx.x; q <= LX; q += C::NT) srtw[q] = rtw[q];
    const long N = g.N;
    float2 v[3][C::E];
#pragma unroll
    for (int q = 0; q < C::E; q++) {
        float a[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
        const int n = t + C::T * q;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int k = 2 * n + s;
            if (ok && k < g.n2) {
                const long I = row * g.n2 + k;
                float a0 = m[I], a1 = m[N + I], a2 = m[2 * N + I];
                if (PRED) {
                    a0 = a0 + h * tau0[I];
                    a1 = a1 + h * tau0[N + I];
                    a2 = a2 + h * tau0[2 * N + I];
                    const float norm = norm_map ? norm_map[I] : 1.0f;
                    const float len = sqrtf(a0 * a0 + a1 * a1 + a2 * a2);
                    float scale = (norm == 0.0f) ? 0.0f : 1.0f;
                    scale = (len != 0.0f) ? scale / len : 0.0f;
                    a0 *= scale;
                    a1 *= scale;
                    a2 *= scale;
                    mpred[I] = a0;
                    mpred[N + I] = a1;
                    mpred[2 * N + I] = a2;
                }
                if (msat_map) {
                    const float ms = msat_map[I];
                    a0 *= ms;
                    a1 *= ms;
                    a2 *= ms;
                }
                a[s][0] = a0;
                a[s][1] = a1;
                a[s][2] = a2;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) v[c][q] = float2{a[0][c], a[1][c]};
    }
    __syncthreads();
    float2 u[3][C::G][C::T];
#pragma unroll
    for (int c = 0; c < 3; c++) row_forward<LX>(v[c], u[c], r, t, sbuf, stw);

    const int lane_row0 = (threadIdx.x & 63) - t;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float2 *dst = xb + ((long)c * g.nrows + row) * g.sx;
        static_for<C::G * C::T>([&](auto I) {
            constexpr int gg = decltype(I)::value / C::T, k2 = decltype(I)::value % C::T;
            const float2 b = mirror_conj<LX, gg, k2>(u[c], t, lane_row0);
            const float2 a = u[c][gg][k2];
            const int k = t * C::G + gg + C::E * k2;
            const float2 Ev = float2{0.5f * (a.x + b.x), 0.5f * (a.y + b.y)};
            const float2 Ov = float2{0.5f * (a.y - b.y), -0.5f * (a.x - b.x)};
            const float2 X = cadd(Ev, cmul(Ov, srtw[k]));
            if (ok) dst[k] = X;
            if (ok && k == 0) {
                // Nyquist bin: E(0) - O(0) = Re Z0 - Im Z0
                dst[LX] = float2{a.x - a.y, 0.0f};
            }
        });
    }
}

// xi: C2R along x (imaginary parts of DC and Nyquist ignored), keep k < n2,
// then the fused field / torque / RK epilogue
template <int LX, int MODE>
__global__ void __launch_bounds__(RowTile<LX>::NT)
xi_fast(DemagGeom g, const float2 *__restrict__ tw, const float2 *__restrict__ rtw,
        const float2 *__restrict__ xb, FieldParams fp, EpilogueIO io)
{
    using C = RowTile<LX>;
    extern __shared__ float2 sm[];
    float2 *sbuf = sm, *stw = sm + C::E * C::S, *srtw = stw + LX;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int r = threadIdx.x / C::T, t = threadIdx.x % C::T;
    const long row = (long)tile * C::R + r;
    const bool ok = row < g.nrows;
    for (int q = threadIdx.x; q < LX; q += C::NT) stw[q] = tw[q];
    for (int q = threadIdx.x; q <= LX; q += C::NT) srtw[q] = rtw[q];
    __syncthreads();
    const int lane_row0 = (threadIdx.x & 63) - t;
    float2 v[3][C::E];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float2 *src = xb + ((long)c * g.nrows + (ok ? row : 0)) * g.sx;
        float2 u[C::G][C::T];
#pragma unroll
        for (int gg = 0; gg < C::G; gg++)
#pragma unroll
            for (int k2 = 0; k2 < C::T; k2++) {
                const int k = t * C::G + gg + C::E * k2;
                float2 x = ok ? src[k] : float2{0.0f, 0.0f};
                if (k == 0) x.y = 0.0f;
                u[gg][k2] = x;
            }
        const float2 nyq = ok ? float2{src[LX].x, 0.0f} : float2{0.0f, 0.0f};
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
    if (ok) {
        const int i = (int)(row / g.n1), j = (int)(row - (long)i * g.n1);
#pragma unroll
        for (int q = 0; q < C::E; q++) {
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int k = 2 * (t + C::T * q) + s;
                if (k < g.n2) {
                    const long I = row * g.n2 + k;
                    float hd[3];
#pragma unroll
                    for (int c = 0; c < 3; c++) hd[c] = fp.demag_scale * (s ? v[c][q].y : v[c][q].x);
                    cell_epilogue<MODE>(g, fp, io, I, i, j, k, hd, errmax);
                }
            }
        }
    }
    if (MODE == EPI_CORR) block_max_to(errmax, io.err);
}

// ------------------------------------------------------------ host side

template <typename K> static void allow_lds_fast(K kernel, size_t bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

#define MMX_FAST_LENGTHS(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)

int fast_yf(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    switch (g.p1) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = ColTile<L>;                                                                \
        const int ntx = div_up(g.h2, C::W);                                                  \
        allow_lds_fast(yf_fast<L>, C::LDS);                                                  \
        yf_fast<L><<<dim3(ntx * g.n0 * 3), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw,      \
                                                                      P->xbuf, P->ybuf);     \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_yf: length %d", g.p1); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int fast_yi(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    switch (g.p1) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = ColTile<L>;                                                                \
        const int ntx = div_up(g.h2, C::W);                                                  \
        allow_lds_fast(yi_fast<L>, C::LDS);                                                  \
        yi_fast<L><<<dim3(ntx * g.n0 * 3), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw,      \
                                                                      P->ybuf, P->xbuf);     \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_yi: length %d", g.p1); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int fast_z(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    switch (g.p0) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = ColTile<L>;                                                                \
        const int ntx = div_up(g.h2, C::W);                                                  \
        allow_lds_fast(z_fast<L>, C::LDS);                                                   \
        z_fast<L><<<dim3(ntx * g.p1), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fz.tw, P->kern,  \
                                                                 P->ybuf);                   \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_z: length %d", g.p0); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int fast_y2d(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    switch (g.p1) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = ColTile<L>;                                                                \
        const int ntx = div_up(g.h2, C::W);                                                  \
        allow_lds_fast(y2d_fast<L>, C::LDS);                                                 \
        y2d_fast<L><<<dim3(ntx), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw, P->kern,       \
                                                            P->xbuf);                        \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_y2d: length %d", g.p1); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int fast_xf(const DemagPlanImpl *P, int pred, const float *m, const float *msat_map,
            const float *tau0, float h, const float *norm_map, float *mpred, hipStream_t st)
{
    const DemagGeom &g = P->g;
    switch (g.Lx) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = RowTile<L>;                                                                \
        const int nb = div_up(g.nrows, C::R);                                                \
        if (pred) {                                                                          \
            allow_lds_fast(xf_fast<L, 1>, C::LDS);                                           \
            xf_fast<L, 1><<<dim3(nb), dim3(C::NT), C::LDS, st>>>(                            \
                g, P->fx.tw, P->fx.rtw, m, msat_map, tau0, h, norm_map, mpred, P->xbuf);     \
        } else {                                                                             \
            allow_lds_fast(xf_fast<L, 0>, C::LDS);                                           \
            xf_fast<L, 0><<<dim3(nb), dim3(C::NT), C::LDS, st>>>(                            \
                g, P->fx.tw, P->fx.rtw, m, msat_map, nullptr, 0.0f, nullptr, nullptr,        \
                P->xbuf);                                                                    \
        }                                                                                    \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_xf: length %d", g.Lx); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

template <int L, int MODE>
static void launch_xi(const DemagPlanImpl *P, const FieldParams &fp, const EpilogueIO &io,
                      hipStream_t st)
{
    using C = RowTile<L>;
    const int nb = div_up(P->g.nrows, C::R);
    allow_lds_fast(xi_fast<L, MODE>, C::LDS);
    xi_fast<L, MODE><<<dim3(nb), dim3(C::NT), C::LDS, st>>>(P->g, P->fx.tw, P->fx.rtw, P->xbuf,
                                                            fp, io);
}

int fast_xi(const DemagPlanImpl *P, int mode, const FieldParams &fp, const EpilogueIO &io,
            hipStream_t st)
{
    const DemagGeom &g = P->g;
    switch (g.Lx) {
#define CASE(L)                                                                              \
    case L:                                                                                  \
        switch (mode) {                                                                      \
        case EPI_B: launch_xi<L, EPI_B>(P, fp, io, st); break;                               \
        case EPI_HEFF: launch_xi<L, EPI_HEFF>(P, fp, io, st); break;                         \
        case EPI_TAU0: launch_xi<L, EPI_TAU0>(P, fp, io, st); break;                         \
        default: launch_xi<L, EPI_CORR>(P, fp, io, st); break;                               \
        }                                                                                    \
        break;
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_xi: length %d", g.Lx); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

} // namespace mmx
