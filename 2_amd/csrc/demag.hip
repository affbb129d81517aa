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
#include <atomic>
#include <cstdlib>

#include "demag_tiles.hpp"
#include "reg_fft.hpp"

namespace mmx {

// ------------------------------------------------------------------ XF pass

template <int PRED>
__global__ void __launch_bounds__(kBlock)
xf_kernel(DemagGeom g, LineFFT<float> fx, int RB, const float *__restrict__ m,
          const float *__restrict__ msat_map, const float *__restrict__ tau0, float h,
          const float *__restrict__ norm_map, float *__restrict__ mpred,
          float2 *__restrict__ xb, unsigned int *ho_dev, unsigned int *ho_host)
{
    extern __shared__ float2 sm[];
    if (ho_dev != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        // the previous launch's result word to the host, cleared for the next
        const unsigned int e = __hip_atomic_load(ho_dev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ho_host, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(ho_dev, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    xf_tile<PRED>(g, fx, RB, (long)blockIdx.x * RB, m, msat_map, tau0, h, norm_map, mpred, xb,
                  sm);
}

// ------------------------------------------------------------------ YF pass

__global__ void __launch_bounds__(kBlock)
yf_kernel(DemagGeom g, LineFFT<float> fy, int W, const float2 *__restrict__ xb,
          float2 *__restrict__ yb)
{
    extern __shared__ float2 sm[];
    yf_tile(g, fy, W, blockIdx.x * W, blockIdx.y, blockIdx.z, xb, yb, sm);
}

// -------------------------------------------------- Z pass (+ kernel multiply)

__global__ void __launch_bounds__(kBlock)
z_kernel(DemagGeom g, LineFFT<float> fz, int W, const float *__restrict__ K,
         float2 *__restrict__ yb)
{
    extern __shared__ float2 sm[];
    z_tile(g, fz, W, blockIdx.x * W, blockIdx.y, K, yb, sm);
}

// short z (z_small_applies): one thread per spectrum column
__global__ void __launch_bounds__(kBlock)
z_small_kernel(DemagGeom g, const float *__restrict__ K, float2 *__restrict__ yb)
{
    const long col = (long)blockIdx.x * kBlock + threadIdx.x;
    if (col >= (long)g.p1 * g.h2) return;
    const int ky = (int)(col / g.h2), kx = (int)(col - (long)ky * g.h2);
    z_column_small_any(g, K, yb, ky, kx);
}

// ---------------------------------------------- film pass: y, multiply, y^-1

__global__ void __launch_bounds__(kBlock)
y2d_kernel(DemagGeom g, LineFFT<float> fy, int W, const float *__restrict__ K,
           float2 *__restrict__ xb)
{
    extern __shared__ float2 sm[];
    y2d_tile(g, fy, W, blockIdx.x * W, K, xb, sm);
}

// ------------------------------------------------------------------ YI pass

__global__ void __launch_bounds__(kBlock)
yi_kernel(DemagGeom g, LineFFT<float> fy, int W, const float2 *__restrict__ yb,
          float2 *__restrict__ xb)
{
    extern __shared__ float2 sm[];
    yi_tile(g, fy, W, blockIdx.x * W, blockIdx.y, blockIdx.z, yb, xb, sm);
}

// ------------------------------------------------- XI pass (+ fused epilogue)

template <int MODE>
__global__ void __launch_bounds__(kBlock)
xi_kernel(DemagGeom g, LineFFT<float> fx, int RB, const float2 *__restrict__ xb,
          FieldParams fp, EpilogueIO io)
{
    extern __shared__ float2 sm[];
    xi_tile<MODE>(g, fx, RB, (long)blockIdx.x * RB, xb, fp, io, sm);
}

// ------------------------------------------------------------ host launchers

// Pass ids of the timed evaluation: 3-D xf 0, yf 1, z 2, yi 3, xi 4; films
// xf 0, y2d 1, xi 2.  mark(P, st, pass) ends a launch of `pass` (-1: start).
enum : int { PASS_START = -1, PASS_XF = 0, PASS_YF = 1, PASS_Y2D = 1, PASS_Z = 2, PASS_YI = 3 };

static inline void mark(DemagPlanImpl *P, hipStream_t st, int pass)
{
    if (P->events && P->nevent < kMaxMarks) {
        P->event_pass[P->nevent] = pass;
        (void)hipEventRecord(P->events[P->nevent++], st);
    }
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
bool fast_ylength(int L) { return fast_ylength_impl(L); }

static int y_and_z_passes(DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    const int ntx_y = div_up(g.h2, P->tiles.w_y);
    // films: the y transform on the z pass's kernel (three component groups
    // per column tile) wherever the length has a register FFT
    if (g.p0 == 1 && fast_y2d_as_z_applies(g)) {
        MMX_TRY(fast_y2d_as_z(P, st));
        mark(P, st, PASS_Y2D);
        return MMX_OK;
    }
    if (g.p0 == 1) {
        const size_t b = lds_bytes_y(P, 3);
        allow_lds(y2d_kernel, b);
        y2d_kernel<<<dim3(ntx_y), dim3(kBlock), b, st>>>(g, P->fy, P->tiles.w_y, P->kern, P->xbuf);
        MMX_CHECK_LAUNCH();
        mark(P, st, PASS_Y2D);
        return MMX_OK;
    }
    if (P->fast_y) {
        MMX_TRY(fast_yf(P, st));
        mark(P, st, PASS_YF);
    } else {
        const size_t b = lds_bytes_y(P, 1);
        allow_lds(yf_kernel, b);
        yf_kernel<<<dim3(ntx_y, g.n0, 3), dim3(kBlock), b, st>>>(g, P->fy, P->tiles.w_y, P->xbuf,
                                                                   P->ybuf);
        MMX_CHECK_LAUNCH();
        mark(P, st, PASS_YF);
    }
    if (P->fast_z) {
        MMX_TRY(fast_z(P, st));
        mark(P, st, PASS_Z);
    } else if (z_small_applies(g)) {
        z_small_kernel<<<dim3((unsigned)div_up((long)g.p1 * g.h2, kBlock)), dim3(kBlock), 0, st>>>(
            g, P->kern, P->ybuf);
        MMX_CHECK_LAUNCH();
        mark(P, st, PASS_Z);
    } else {
        const int ntx_z = div_up(g.h2, P->tiles.w_z);
        const size_t b = lds_bytes_z(P);
        allow_lds(z_kernel, b);
        z_kernel<<<dim3(ntx_z, g.p1), dim3(kBlock), b, st>>>(g, P->fz, P->tiles.w_z, P->kern,
                                                              P->ybuf);
        MMX_CHECK_LAUNCH();
        mark(P, st, PASS_Z);
    }
    if (P->fast_y) {
        MMX_TRY(fast_yi(P, st));
        mark(P, st, PASS_YI);
    } else {
        const size_t b = lds_bytes_y(P, 1);
        allow_lds(yi_kernel, b);
        yi_kernel<<<dim3(ntx_y, g.n0, 3), dim3(kBlock), b, st>>>(g, P->fy, P->tiles.w_y, P->ybuf,
                                                                   P->xbuf);
        MMX_CHECK_LAUNCH();
        mark(P, st, PASS_YI);
    }
    return MMX_OK;
}

static std::atomic<int> g_fast_x_min{kFastXMinBlocks};

int fast_x_min_blocks() { return g_fast_x_min.load(std::memory_order_relaxed); }

int set_fast_x_min_blocks(int n)
{
    return g_fast_x_min.exchange(n, std::memory_order_relaxed);
}

bool fast_xf_used(const DemagPlanImpl *P)
{
    return P->fast_x && fast_x_blocks(P->g) >= fast_x_min_blocks();
}

int demag_forward(DemagPlanImpl *P, const float *m, const float *msat_map, hipStream_t st)
{
    const DemagGeom &g = P->g;
    const size_t b = lds_bytes_x(P);
    allow_lds(xf_kernel<0>, b);
    const int nb = div_up(g.nrows, P->tiles.rows_x);
    mark(P, st, PASS_START);
    if (fast_xf_used(P)) {
        MMX_TRY(fast_xf(P, 0, m, msat_map, nullptr, 0.0f, nullptr, nullptr, st));
    } else {
        xf_kernel<0><<<dim3(nb), dim3(kBlock), b, st>>>(g, P->fx, P->tiles.rows_x, m, msat_map,
                                                         nullptr, 0.0f, nullptr, nullptr, P->xbuf,
                                                         P->handoff_dev, P->handoff_host);
        MMX_CHECK_LAUNCH();
    }
    if (P->after_x) MMX_HIP(hipEventRecord(P->after_x, st));
    mark(P, st, PASS_XF);
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
    mark(P, st, PASS_START);
    if (fast_xf_used(P)) {
        MMX_TRY(fast_xf(P, 1, m0, norm_map, tau0, h, norm_map, mpred, st));
    } else {
        xf_kernel<1><<<dim3(nb), dim3(kBlock), b, st>>>(g, P->fx, P->tiles.rows_x, m0, norm_map,
                                                         tau0, h, norm_map, mpred, P->xbuf, nullptr,
                                                         nullptr);
        MMX_CHECK_LAUNCH();
    }
    mark(P, st, PASS_XF);
    return y_and_z_passes(P, st);
}

int demag_inverse(DemagPlanImpl *P, int mode, const FieldParams &fp, const EpilogueIO &io,
                  hipStream_t st)
{
    const DemagGeom &g = P->g;
    const size_t b = lds_bytes_x(P);
    const int nb = div_up(g.nrows, P->tiles.rows_x);
    const int RB = P->tiles.rows_x;
    const int pass_xi = (g.p0 == 1) ? 2 : 4;
    if (P->fast_x && fast_x_blocks(g) >= fast_x_min_blocks()) {
        MMX_TRY(fast_xi(P, mode, fp, io, st));
        mark(P, st, pass_xi);
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
    mark(P, st, pass_xi);
    return MMX_OK;
}

} // namespace mmx
