// demag_tiles.hpp -- one workgroup's share of each generic (any-length,
// LDS Stockham) demag pass, given its tile coordinates and its LDS.  Called by
// the single-simulation kernels (demag.hip) and by the batched kernels that
// run many simulations per launch (batch.hip).  Pass structure and the
// reference functions replaced: see demag.hip.
#pragma once

#include "epilogue.hpp"

namespace mmx {

// Each pass is a tile function (one workgroup's work, given its tile
// coordinates and the LDS).

template <int PRED>
__device__ __forceinline__ void
xf_tile(const DemagGeom &g, const LineFFT<float> &fx, int RB, long row0,
        const float *__restrict__ m, const float *__restrict__ msat_map,
        const float *__restrict__ tau0, float h, const float *__restrict__ norm_map,
        float *__restrict__ mpred, float2 *__restrict__ xb, float2 *sm)
{
    const int Lx = g.Lx, ls = Lx + 1, nl = 3 * RB;
    float2 *A = sm, *B = sm + nl * ls;
    const long N = g.N;

    for (int t = threadIdx.x; t < RB * Lx; t += blockDim.x) {
        const int r = t / Lx, nn = t - r * Lx;
        const long row = row0 + r;
        float v[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
        if (row < g.nrows) {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int k = 2 * nn + q;
                if (k < g.n2) {
                    const long I = row * g.n2 + k;
                    float a0 = m[I], a1 = m[N + I], a2 = m[2 * N + I];
                    if (PRED) {
                        // predictor m = y0 + dt*gamma*tau0 (maddKern), then the
                        // normUpdater (normalize.cu:14-43)
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
                    v[q][0] = a0;
                    v[q][1] = a1;
                    v[q][2] = a2;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) A[(c * RB + r) * ls + nn] = float2{v[0][c], v[1][c]};
    }
    __syncthreads();
    const float2 *R = lds_fft<float, -1>(A, B, nl, ls, 1, fx);

    // split the half-length spectrum into the R2C output X[0..Lx]
    for (int t = threadIdx.x; t < nl * (Lx + 1); t += blockDim.x) {
        const int l = t / (Lx + 1), k = t - l * (Lx + 1);
        const int c = l / RB, r = l - c * RB;
        const long row = row0 + r;
        if (row >= g.nrows) continue;
        const float2 a = R[l * ls + (k == Lx ? 0 : k)];
        const float2 b = cconj(R[l * ls + (k == 0 ? 0 : Lx - k)]);
        const float2 E = float2{0.5f * (a.x + b.x), 0.5f * (a.y + b.y)};
        const float2 O = float2{0.5f * (a.y - b.y), -0.5f * (a.x - b.x)};
        xb[((long)c * g.nrows + row) * g.sx + k] = cadd(E, cmul(O, fx.rtw[k]));
    }
}

__device__ __forceinline__ void
yf_tile(const DemagGeom &g, const LineFFT<float> &fy, int W, int kx0, int i, int c,
        const float2 *__restrict__ xb, float2 *__restrict__ yb, float2 *sm)
{
    const int p1 = g.p1, ls = p1 + 1;
    float2 *A = sm, *B = sm + W * ls;
    const float2 *src = xb + ((long)c * g.nrows + (long)i * g.n1) * g.sx;
    for (int t = threadIdx.x; t < p1 * W; t += blockDim.x) {
        const int w = t % W, j = t / W, kx = kx0 + w;
        float2 v = float2{0.0f, 0.0f};
        if (j < g.n1 && kx < g.h2) v = src[(long)j * g.sx + kx];
        A[w * ls + j] = v;
    }
    __syncthreads();
    const float2 *R = lds_fft<float, -1>(A, B, W, ls, 1, fy);
    float2 *dst = yb + ((long)c * g.n0 + i) * (long)p1 * g.sx;
    for (int t = threadIdx.x; t < p1 * W; t += blockDim.x) {
        const int w = t % W, ky = t / W, kx = kx0 + w;
        if (kx < g.h2) dst[(long)ky * g.sx + kx] = R[w * ls + ky];
    }
}

__device__ __forceinline__ void
z_tile(const DemagGeom &g, const LineFFT<float> &fz, int W, int kx0, int ky,
       const float *__restrict__ K, float2 *__restrict__ yb, float2 *sm)
{
    const int p0 = g.p0, ls = p0 + 1, nl = 3 * W;
    float2 *A = sm, *B = sm + nl * ls;
    const long cstride = (long)g.n0 * g.p1 * g.sx; // component
    const long istride = (long)g.p1 * g.sx;        // plane
    float2 *base = yb + (long)ky * g.sx + kx0;
    for (int t = threadIdx.x; t < 3 * p0 * W; t += blockDim.x) {
        const int w = t % W, rest = t / W, i = rest % p0, c = rest / p0;
        float2 v = float2{0.0f, 0.0f};
        if (i < g.n0 && kx0 + w < g.h2) v = base[c * cstride + i * istride + w];
        A[(c * W + w) * ls + i] = v;
    }
    __syncthreads();
    float2 *R = lds_fft<float, -1>(A, B, nl, ls, 1, fz);
    float2 *Q = (R == A) ? B : A;
    for (int t = threadIdx.x; t < p0 * W; t += blockDim.x) {
        const int w = t % W, kz = t / W, kx = kx0 + w;
        if (kx >= g.h2) continue;
        const Kern6 k = kern_at(g, K, kz, ky, kx);
        float2 &a0 = R[(0 * W + w) * ls + kz];
        float2 &a1 = R[(1 * W + w) * ls + kz];
        float2 &a2 = R[(2 * W + w) * ls + kz];
        kern_apply(k, a0, a1, a2);
    }
    __syncthreads();
    const float2 *S = lds_fft<float, +1>(R, Q, nl, ls, 1, fz);
    for (int t = threadIdx.x; t < 3 * g.n0 * W; t += blockDim.x) {
        const int w = t % W, rest = t / W, i = rest % g.n0, c = rest / g.n0;
        if (kx0 + w < g.h2) base[c * cstride + i * istride + w] = S[(c * W + w) * ls + i];
    }
}

__device__ __forceinline__ void
y2d_tile(const DemagGeom &g, const LineFFT<float> &fy, int W, int kx0,
         const float *__restrict__ K, float2 *__restrict__ xb, float2 *sm)
{
    const int p1 = g.p1, ls = p1 + 1, nl = 3 * W;
    float2 *A = sm, *B = sm + nl * ls;
    const long cstride = (long)g.nrows * g.sx;
    for (int t = threadIdx.x; t < 3 * p1 * W; t += blockDim.x) {
        const int w = t % W, rest = t / W, j = rest % p1, c = rest / p1, kx = kx0 + w;
        float2 v = float2{0.0f, 0.0f};
        if (j < g.n1 && kx < g.h2) v = xb[c * cstride + (long)j * g.sx + kx];
        A[(c * W + w) * ls + j] = v;
    }
    __syncthreads();
    float2 *R = lds_fft<float, -1>(A, B, nl, ls, 1, fy);
    float2 *Q = (R == A) ? B : A;
    for (int t = threadIdx.x; t < p1 * W; t += blockDim.x) {
        const int w = t % W, ky = t / W, kx = kx0 + w;
        if (kx >= g.h2) continue;
        const Kern6 k = kern_at(g, K, 0, ky, kx);
        kern_apply(k, R[(0 * W + w) * ls + ky], R[(1 * W + w) * ls + ky], R[(2 * W + w) * ls + ky]);
    }
    __syncthreads();
    const float2 *S = lds_fft<float, +1>(R, Q, nl, ls, 1, fy);
    for (int t = threadIdx.x; t < 3 * g.n1 * W; t += blockDim.x) {
        const int w = t % W, rest = t / W, j = rest % g.n1, c = rest / g.n1, kx = kx0 + w;
        if (kx < g.h2) xb[c * cstride + (long)j * g.sx + kx] = S[(c * W + w) * ls + j];
    }
}

__device__ __forceinline__ void
yi_tile(const DemagGeom &g, const LineFFT<float> &fy, int W, int kx0, int i, int c,
        const float2 *__restrict__ yb, float2 *__restrict__ xb, float2 *sm)
{
    const int p1 = g.p1, ls = p1 + 1;
    float2 *A = sm, *B = sm + W * ls;
    const float2 *src = yb + ((long)c * g.n0 + i) * (long)p1 * g.sx;
    for (int t = threadIdx.x; t < p1 * W; t += blockDim.x) {
        const int w = t % W, ky = t / W, kx = kx0 + w;
        float2 v = float2{0.0f, 0.0f};
        if (kx < g.h2) v = src[(long)ky * g.sx + kx];
        A[w * ls + ky] = v;
    }
    __syncthreads();
    const float2 *R = lds_fft<float, +1>(A, B, W, ls, 1, fy);
    float2 *dst = xb + ((long)c * g.nrows + (long)i * g.n1) * g.sx;
    for (int t = threadIdx.x; t < g.n1 * W; t += blockDim.x) {
        const int w = t % W, j = t / W, kx = kx0 + w;
        if (kx < g.h2) dst[(long)j * g.sx + kx] = R[w * ls + j];
    }
}

template <int MODE>
__device__ __forceinline__ void
xi_tile(const DemagGeom &g, const LineFFT<float> &fx, int RB, long row0,
        const float2 *__restrict__ xb, const FieldParams &fp, const EpilogueIO &io, float2 *sm)
{
    const int Lx = g.Lx, ls = Lx + 1, nl = 3 * RB;
    float2 *A = sm, *B = sm + nl * ls;

    // C2R pre-processing (imaginary parts of DC and Nyquist ignored)
    for (int t = threadIdx.x; t < nl * Lx; t += blockDim.x) {
        const int l = t / Lx, k = t - l * Lx;
        const int c = l / RB, r = l - c * RB;
        const long row = row0 + r;
        float2 z = float2{0.0f, 0.0f};
        if (row < g.nrows) {
            const float2 *src = xb + ((long)c * g.nrows + row) * g.sx;
            float2 a = src[k], bb = src[Lx - k];
            if (k == 0) {
                a.y = 0.0f;
                bb.y = 0.0f;
            }
            const float2 b = cconj(bb);
            const float2 E = cadd(a, b);
            const float2 O = cmulc(csub(a, b), fx.rtw[k]);
            z = float2{E.x - O.y, E.y + O.x};
        }
        A[l * ls + k] = z;
    }
    __syncthreads();
    const float2 *R = lds_fft<float, +1>(A, B, nl, ls, 1, fx);

    float errmax = 0.0f;
    for (int t = threadIdx.x; t < RB * g.n2; t += blockDim.x) {
        const int r = t / g.n2, k = t - r * g.n2;
        const long row = row0 + r;
        if (row >= g.nrows) continue;
        const int i = (int)(row / g.n1), j = (int)(row - (long)i * g.n1);
        const long I = row * g.n2 + k;
        float hd[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float2 z = R[(c * RB + r) * ls + (k >> 1)];
            hd[c] = fp.demag_scale * ((k & 1) ? z.y : z.x);
        }
        cell_epilogue<MODE>(g, fp, io, I, i, j, k, hd, errmax);
    }
    if (MODE == EPI_CORR) block_max_to(errmax, io.err);
}

// Z transform + kernel multiply + inverse of one spectrum column (ky, kx) for
// short z (P0 <= 8, N0 = P0/2 nonzero planes in and kept out): direct DFTs in
// registers, one thread per column, no LDS.  Same transform pair as z_tile
// (forward exp(-2 pi i k n / P0), unnormalised inverse).
template <int P0, int N0>
__device__ __forceinline__ void z_column_small(const DemagGeom &g, const float *__restrict__ K,
                                               float2 *__restrict__ yb, int ky, int kx)
{
    const long cstride = (long)g.n0 * g.p1 * g.sx;
    const long istride = (long)g.p1 * g.sx;
    float2 *base = yb + (long)ky * g.sx + kx;
    float2 a[3][N0];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int i = 0; i < N0; i++) a[c][i] = base[c * cstride + i * istride];
    float cs[P0], sn[P0];
#pragma unroll
    for (int q = 0; q < P0; q++) sincospif(2.0f * (float)q / (float)P0, &sn[q], &cs[q]);
    float2 X[3][P0];
#pragma unroll
    for (int kz = 0; kz < P0; kz++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float2 acc = float2{0.0f, 0.0f};
#pragma unroll
            for (int i = 0; i < N0; i++) {
                const int q = (kz * i) % P0;
                acc.x += a[c][i].x * cs[q] + a[c][i].y * sn[q];
                acc.y += a[c][i].y * cs[q] - a[c][i].x * sn[q];
            }
            X[c][kz] = acc;
        }
#pragma unroll
    for (int kz = 0; kz < P0; kz++) {
        const Kern6 k = kern_at(g, K, kz, ky, kx);
        kern_apply(k, X[0][kz], X[1][kz], X[2][kz]);
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int i = 0; i < N0; i++) {
            float2 acc = float2{0.0f, 0.0f};
#pragma unroll
            for (int kz = 0; kz < P0; kz++) {
                const int q = (kz * i) % P0;
                acc.x += X[c][kz].x * cs[q] - X[c][kz].y * sn[q];
                acc.y += X[c][kz].y * cs[q] + X[c][kz].x * sn[q];
            }
            base[c * cstride + i * istride] = acc;
        }
}

// True when z_column_small covers the geometry.
__host__ __device__ inline bool z_small_applies(const DemagGeom &g)
{
    return (g.p0 == 4 || g.p0 == 6 || g.p0 == 8) && 2 * g.n0 == g.p0;
}

// Dispatch on the z length (the caller checked z_small_applies).
__device__ __forceinline__ void z_column_small_any(const DemagGeom &g, const float *__restrict__ K,
                                                   float2 *__restrict__ yb, int ky, int kx)
{
    switch (g.p0) {
    case 4: z_column_small<4, 2>(g, K, yb, ky, kx); break;
    case 6: z_column_small<6, 3>(g, K, yb, ky, kx); break;
    default: z_column_small<8, 4>(g, K, yb, ky, kx); break;
    }
}

} // namespace mmx
