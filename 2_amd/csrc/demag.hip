// demag.hip -- zero-padding-aware demag convolution passes for gfx950.
//
// One H_eff evaluation on a padded grid (p0,p1,p2) with data (n0,n1,n2):
//   XF  R2C along x of the n0*n1 nonzero rows (half-length complex FFT),
//       input optionally formed on the fly (RK predictor + normalisation);
//   YF  FFT along y of the n0 nonzero planes, zero rows j >= n1 never loaded;
//   Z   FFT along z (zero planes i >= n0 never loaded), 3x3 symmetric kernel
//       multiply, inverse FFT along z, only i < n0 kept;
//   YI  inverse FFT along y, only j < n1 kept;
//   XI  C2R along x, only k < n2 kept, fused with exchange, anisotropy,
//       Zeeman, the LLG/Zhang-Li torque and the RK stage update.
// For p0 == 1 (films) YF/Z/YI collapse into one kernel.
// Replaces MaxwellPlan.update (maxwell_convolution.go:272-309): zeroing,
// CopyPad3D, 3 cuFFT R2C, 9 cmadd passes, 3 cuFFT C2R and CopyUnPad3D.
#include "epilogue.hpp"
#include "reg_fft.hpp"

namespace mmx {

// ------------------------------------------------------------------ XF pass

template <int PRED>
__global__ void __launch_bounds__(kBlock)
xf_kernel(DemagGeom g, LineFFT<float> fx, int RB, const float *__restrict__ m,
          const float *__restrict__ msat_map, const float *__restrict__ tau0, float h,
          const float *__restrict__ norm_map, float *__restrict__ mpred,
          float2 *__restrict__ xb)
{
    extern __shared__ float2 sm[];
    const int Lx = g.Lx, ls = Lx + 1, nl = 3 * RB;
    float2 *A = sm, *B = sm + nl * ls;
    const long row0 = (long)blockIdx.x * RB;
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

// ------------------------------------------------------------------ YF pass

__global__ void __launch_bounds__(kBlock)
yf_kernel(DemagGeom g, LineFFT<float> fy, int W, const float2 *__restrict__ xb,
          float2 *__restrict__ yb)
{
    extern __shared__ float2 sm[];
    const int p1 = g.p1, ls = p1 + 1;
    float2 *A = sm, *B = sm + W * ls;
    const int kx0 = blockIdx.x * W, i = blockIdx.y, c = blockIdx.z;
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

// -------------------------------------------------- Z pass (+ kernel multiply)

__global__ void __launch_bounds__(kBlock)
z_kernel(DemagGeom g, LineFFT<float> fz, int W, const float *__restrict__ K,
         float2 *__restrict__ yb)
{
    extern __shared__ float2 sm[];
    const int p0 = g.p0, ls = p0 + 1, nl = 3 * W;
    float2 *A = sm, *B = sm + nl * ls;
    const int kx0 = blockIdx.x * W, ky = blockIdx.y;
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

// ---------------------------------------------- film pass: y, multiply, y^-1

__global__ void __launch_bounds__(kBlock)
y2d_kernel(DemagGeom g, LineFFT<float> fy, int W, const float *__restrict__ K,
           float2 *__restrict__ xb)
{
    extern __shared__ float2 sm[];
    const int p1 = g.p1, ls = p1 + 1, nl = 3 * W;
    float2 *A = sm, *B = sm + nl * ls;
    const int kx0 = blockIdx.x * W;
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

// ------------------------------------------------------------------ YI pass

__global__ void __launch_bounds__(kBlock)
yi_kernel(DemagGeom g, LineFFT<float> fy, int W, const float2 *__restrict__ yb,
          float2 *__restrict__ xb)
{
    extern __shared__ float2 sm[];
    const int p1 = g.p1, ls = p1 + 1;
    float2 *A = sm, *B = sm + W * ls;
    const int kx0 = blockIdx.x * W, i = blockIdx.y, c = blockIdx.z;
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

// ------------------------------------------------- XI pass (+ fused epilogue)

template <int MODE>
__global__ void __launch_bounds__(kBlock)
xi_kernel(DemagGeom g, LineFFT<float> fx, int RB, const float2 *__restrict__ xb,
          FieldParams fp, EpilogueIO io)
{
    extern __shared__ float2 sm[];
    const int Lx = g.Lx, ls = Lx + 1, nl = 3 * RB;
    float2 *A = sm, *B = sm + nl * ls;
    const long row0 = (long)blockIdx.x * RB;

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

// ------------------------------------------------------------ host launchers

static inline void mark(DemagPlanImpl *P, hipStream_t st)
{
    if (P->events) (void)hipEventRecord(P->events[P->nevent++], st);
}

static size_t lds_bytes_x(const DemagPlanImpl *P)
{
    return 2u * 3u * P->tiles.rows_x * (P->g.Lx + 1) * sizeof(float2);
}
static size_t lds_bytes_y(const DemagPlanImpl *P, int ncomp)
{
    return 2u * ncomp * P->tiles.w_y * (P->g.p1 + 1) * sizeof(float2);
}
static size_t lds_bytes_z(const DemagPlanImpl *P)
{
    return 2u * 3u * P->tiles.w_z * (P->g.p0 + 1) * sizeof(float2);
}

template <typename K> static void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

bool fast_length(int L) { return fast_length_impl(L); }

static int y_and_z_passes(DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    const int ntx_y = div_up(g.h2, P->tiles.w_y);
    if (g.p0 == 1 && P->fast_y) {
        MMX_TRY(fast_y2d(P, st));
        mark(P, st);
        return MMX_OK;
    }
    if (g.p0 == 1) {
        const size_t b = lds_bytes_y(P, 3);
        allow_lds(y2d_kernel, b);
        y2d_kernel<<<dim3(ntx_y), dim3(kBlock), b, st>>>(g, P->fy, P->tiles.w_y, P->kern, P->xbuf);
        MMX_CHECK_LAUNCH();
        mark(P, st);
        return MMX_OK;
    }
    if (P->fast_y) {
        MMX_TRY(fast_yf(P, st));
        mark(P, st);
    } else {
        const size_t b = lds_bytes_y(P, 1);
        allow_lds(yf_kernel, b);
        yf_kernel<<<dim3(ntx_y, g.n0, 3), dim3(kBlock), b, st>>>(g, P->fy, P->tiles.w_y, P->xbuf,
                                                                   P->ybuf);
        MMX_CHECK_LAUNCH();
        mark(P, st);
    }
    if (P->fast_z) {
        MMX_TRY(fast_z(P, st));
        mark(P, st);
    } else {
        const int ntx_z = div_up(g.h2, P->tiles.w_z);
        const size_t b = lds_bytes_z(P);
        allow_lds(z_kernel, b);
        z_kernel<<<dim3(ntx_z, g.p1), dim3(kBlock), b, st>>>(g, P->fz, P->tiles.w_z, P->kern,
                                                              P->ybuf);
        MMX_CHECK_LAUNCH();
        mark(P, st);
    }
    if (P->fast_y) {
        MMX_TRY(fast_yi(P, st));
        mark(P, st);
    } else {
        const size_t b = lds_bytes_y(P, 1);
        allow_lds(yi_kernel, b);
        yi_kernel<<<dim3(ntx_y, g.n0, 3), dim3(kBlock), b, st>>>(g, P->fy, P->tiles.w_y, P->ybuf,
                                                                   P->xbuf);
        MMX_CHECK_LAUNCH();
        mark(P, st);
    }
    return MMX_OK;
}

int demag_forward(DemagPlanImpl *P, const float *m, const float *msat_map, hipStream_t st)
{
    const DemagGeom &g = P->g;
    const size_t b = lds_bytes_x(P);
    allow_lds(xf_kernel<0>, b);
    const int nb = div_up(g.nrows, P->tiles.rows_x);
    mark(P, st);
    if (P->fast_x) {
        MMX_TRY(fast_xf(P, 0, m, msat_map, nullptr, 0.0f, nullptr, nullptr, st));
    } else {
        xf_kernel<0><<<dim3(nb), dim3(kBlock), b, st>>>(g, P->fx, P->tiles.rows_x, m, msat_map,
                                                         nullptr, 0.0f, nullptr, nullptr, P->xbuf);
        MMX_CHECK_LAUNCH();
    }
    mark(P, st);
    return y_and_z_passes(P, st);
}

int demag_forward_pred(DemagPlanImpl *P, const float *m0, const float *tau0, float h,
                       const float *norm_map, float *mpred, hipStream_t st)
{
    const DemagGeom &g = P->g;
    const size_t b = lds_bytes_x(P);
    allow_lds(xf_kernel<1>, b);
    const int nb = div_up(g.nrows, P->tiles.rows_x);
    // the demag source is M = m * Msat_map, norm map = Msat map
    mark(P, st);
    if (P->fast_x) {
        MMX_TRY(fast_xf(P, 1, m0, norm_map, tau0, h, norm_map, mpred, st));
    } else {
        xf_kernel<1><<<dim3(nb), dim3(kBlock), b, st>>>(g, P->fx, P->tiles.rows_x, m0, norm_map,
                                                         tau0, h, norm_map, mpred, P->xbuf);
        MMX_CHECK_LAUNCH();
    }
    mark(P, st);
    return y_and_z_passes(P, st);
}

int demag_inverse(DemagPlanImpl *P, int mode, const FieldParams &fp, const EpilogueIO &io,
                  hipStream_t st)
{
    const DemagGeom &g = P->g;
    const size_t b = lds_bytes_x(P);
    const int nb = div_up(g.nrows, P->tiles.rows_x);
    const int RB = P->tiles.rows_x;
    if (P->fast_x && fast_xi_applies(fp)) {
        MMX_TRY(fast_xi(P, mode, fp, io, st));
        mark(P, st);
        return MMX_OK;
    }
    switch (mode) {
    case EPI_B:
        allow_lds(xi_kernel<EPI_B>, b);
        xi_kernel<EPI_B><<<dim3(nb), dim3(kBlock), b, st>>>(g, P->fx, RB, P->xbuf, fp, io);
        break;
    case EPI_HEFF:
        allow_lds(xi_kernel<EPI_HEFF>, b);
        xi_kernel<EPI_HEFF><<<dim3(nb), dim3(kBlock), b, st>>>(g, P->fx, RB, P->xbuf, fp, io);
        break;
    case EPI_TAU0:
        allow_lds(xi_kernel<EPI_TAU0>, b);
        xi_kernel<EPI_TAU0><<<dim3(nb), dim3(kBlock), b, st>>>(g, P->fx, RB, P->xbuf, fp, io);
        break;
    case EPI_CORR:
        allow_lds(xi_kernel<EPI_CORR>, b);
        xi_kernel<EPI_CORR><<<dim3(nb), dim3(kBlock), b, st>>>(g, P->fx, RB, P->xbuf, fp, io);
        break;
    default:
        set_error("demag_inverse: bad mode %d", mode);
        return MMX_ERR_ARG;
    }
    MMX_CHECK_LAUNCH();
    mark(P, st);
    return MMX_OK;
}

} // namespace mmx
