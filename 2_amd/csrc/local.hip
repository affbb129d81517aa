// local.hip -- the libmumax2 entry points on the hot path, as single-launch
// gfx950 kernels: normalize, torque, exchange6, uniaxial anisotropy,
// Zhang-Li, the madd family and the Reductor.  Each C entry point names the
// reference function it replaces (include/mmx.h).
#include <mutex>
#include <vector>

#include "epilogue.hpp"

namespace mmx {

static inline int grid1d(long n) { return div_up(n, kBlock); }

// ------------------------------------------------------------ pointwise ops

__global__ void __launch_bounds__(kBlock) normalize_kernel(float *mx, float *my, float *mz,
                                                           const float *__restrict__ norm_map,
                                                           int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float norm = norm_map ? norm_map[i] : 1.0f;
    const float x = mx[i], y = my[i], z = mz[i];
    const float len = sqrtf(x * x + y * y + z * z);
    float scale = (norm == 0.0f) ? 0.0f : 1.0f;
    scale = (len != 0.0f) ? scale / len : 0.0f;
    mx[i] = x * scale;
    my[i] = y * scale;
    mz[i] = z * scale;
}

__global__ void __launch_bounds__(kBlock)
torque_kernel(float *__restrict__ tx, float *__restrict__ ty, float *__restrict__ tz,
              const float *__restrict__ mx, const float *__restrict__ my,
              const float *__restrict__ mz, const float *__restrict__ hx,
              const float *__restrict__ hy, const float *__restrict__ hz,
              const float *__restrict__ alpha_map, float alpha_mul, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float alpha = alpha_map ? alpha_mul * alpha_map[i] : alpha_mul;
    const float mc[3] = {mx[i], my[i], mz[i]};
    const float H[3] = {hx[i], hy[i], hz[i]};
    float t[3];
    llg_torque(mc, H, alpha, t);
    tx[i] = t[0];
    ty[i] = t[1];
    tz[i] = t[2];
}

// exchange6 over a SoA vector stored as 3 separate pointers
__global__ void __launch_bounds__(kBlock)
exchange6_kernel(float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
                 const float *__restrict__ mx, const float *__restrict__ my,
                 const float *__restrict__ mz, const float *__restrict__ msat_map,
                 const float *__restrict__ aex_map, float pre, int N0, int N1, int N2, int w0,
                 int w1, int w2, float c0, float c1, float c2)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long N = (long)N0 * N1 * N2;
    if (I >= N) return;
    const int k = (int)(I % N2), j = (int)((I / N2) % N1), i = (int)(I / ((long)N1 * N2));
    const float *mm[3] = {mx, my, mz};
    const bool maps = (msat_map != nullptr) || (aex_map != nullptr);
    const float lex0 = maps ? div_zero(mask_unity(aex_map, I), mask_unity(msat_map, I)) : 1.0f;
    const float m0[3] = {mx[I], my[I], mz[I]};
    float H[3] = {0.0f, 0.0f, 0.0f};
    const int order[3] = {0, 2, 1};
    const int wrap[3] = {w0, w1, w2};
    const float cc[3] = {c0, c1, c2};
#pragma unroll
    for (int o = 0; o < 3; o++) {
        const int dim = order[o];
        const int pos = dim == 0 ? i : (dim == 1 ? j : k);
        const int len = dim == 0 ? N0 : (dim == 1 ? N1 : N2);
        const long stride = dim == 0 ? (long)N1 * N2 : (dim == 1 ? (long)N2 : 1L);
        int lo = pos - 1;
        lo = (lo < 0 && wrap[dim]) ? len + lo : lo;
        lo = max(lo, 0);
        int hi = pos + 1;
        hi = (hi == len && wrap[dim]) ? hi - len : hi;
        hi = min(hi, len - 1);
        const long A = I + (long)(lo - pos) * stride, B = I + (long)(hi - pos) * stride;
        float l1 = 1.0f, l2 = 1.0f;
        if (maps) {
            l1 = avg_geom_zero(lex0, div_zero(mask_unity(aex_map, A), mask_unity(msat_map, A)));
            l2 = avg_geom_zero(lex0, div_zero(mask_unity(aex_map, B), mask_unity(msat_map, B)));
        }
        const float pc = pre * cc[dim];
#pragma unroll
        for (int c = 0; c < 3; c++) H[c] += pc * (l1 * (mm[c][A] - m0[c]) + l2 * (mm[c][B] - m0[c]));
    }
    hx[I] = H[0];
    hy[I] = H[1];
    hz[I] = H[2];
}

// exchange6 without maps, its stencil staged in LDS: a workgroup takes a
// kEJ x kEK tile of (j, k) in one i-plane and stages m there with a one-cell
// halo, each LDS position holding the cell the neighbour rule names (clamped
// at open edges, wrapped at periodic ones), so the j and k neighbours come
// from LDS and only i - 1 and i + 1 from global memory.  Per cell the
// arithmetic is that of exchange6_kernel with unit weights.
constexpr int kEJ = 4, kEK = 64;
static_assert(kEJ * kEK == kBlock, "one thread per tile cell");

__device__ inline int nb_index(int x, int len, int wrap)
{
    if (x < 0) x = wrap ? x + len : 0;
    else if (x >= len) x = wrap ? x - len : len - 1;
    return min(max(x, 0), len - 1);   // positions past a partial tile: any cell
}

__global__ void __launch_bounds__(kBlock)
exchange6_lds_kernel(float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
                     const float *__restrict__ mx, const float *__restrict__ my,
                     const float *__restrict__ mz, float pre, int N0, int N1, int N2, int w0,
                     int w1, int w2, float c0, float c1, float c2, int nj, int nk)
{
    __shared__ float tile[3][kEJ + 2][kEK + 2];
    const int b = blockIdx.x;
    const int k0 = (b % nk) * kEK, j0 = ((b / nk) % nj) * kEJ, i = b / (nk * nj);
    const long plane = (long)N1 * N2;
    const float *mm[3] = {mx, my, mz};
    for (int q = threadIdx.x; q < (kEJ + 2) * (kEK + 2); q += kBlock) {
        const int jj = q / (kEK + 2), kk = q % (kEK + 2);
        const long src = (long)i * plane + (long)nb_index(j0 + jj - 1, N1, w1) * N2 +
                         nb_index(k0 + kk - 1, N2, w2);
#pragma unroll
        for (int c = 0; c < 3; c++) tile[c][jj][kk] = mm[c][src];
    }
    __syncthreads();
    const int tj = threadIdx.x / kEK, tk = threadIdx.x % kEK;
    const int j = j0 + tj, k = k0 + tk;
    if (j >= N1 || k >= N2) return;
    const long I = (long)i * plane + (long)j * N2 + k;
    const long A0 = (long)nb_index(i - 1, N0, w0) * plane + (long)j * N2 + k;
    const long B0 = (long)nb_index(i + 1, N0, w0) * plane + (long)j * N2 + k;
    const float l1 = 1.0f, l2 = 1.0f;
    float H[3] = {0.0f, 0.0f, 0.0f};
    float m0[3], a[3], bb[3];
    // dimension order 0, 2, 1 as in exchange6_kernel
#pragma unroll
    for (int c = 0; c < 3; c++) {
        m0[c] = tile[c][tj + 1][tk + 1];
        a[c] = mm[c][A0];
        bb[c] = mm[c][B0];
    }
    const float p0 = pre * c0, p2 = pre * c2, p1 = pre * c1;
#pragma unroll
    for (int c = 0; c < 3; c++) H[c] += p0 * (l1 * (a[c] - m0[c]) + l2 * (bb[c] - m0[c]));
#pragma unroll
    for (int c = 0; c < 3; c++)
        H[c] += p2 * (l1 * (tile[c][tj + 1][tk] - m0[c]) + l2 * (tile[c][tj + 1][tk + 2] - m0[c]));
#pragma unroll
    for (int c = 0; c < 3; c++)
        H[c] += p1 * (l1 * (tile[c][tj][tk + 1] - m0[c]) + l2 * (tile[c][tj + 2][tk + 1] - m0[c]));
    hx[I] = H[0];
    hy[I] = H[1];
    hz[I] = H[2];
}

__global__ void __launch_bounds__(kBlock)
uniaxial_kernel(float *__restrict__ hx, float *__restrict__ hy, float *__restrict__ hz,
                const float *__restrict__ mx, const float *__restrict__ my,
                const float *__restrict__ mz, const float *__restrict__ ku_map,
                const float *__restrict__ msat_map, float ku2, const float *__restrict__ umx,
                float ux, const float *__restrict__ umy, float uy, const float *__restrict__ umz,
                float uz, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float ms = 1.0f;
    if (msat_map) {
        ms = msat_map[i];
        if (ms == 0.0f) ms = 1.0f;
    }
    const float k = ku_map ? (ku2 / ms) * ku_map[i] : ku2 / ms;
    const float u0 = umx ? ux * umx[i] : ux;
    const float u1 = umy ? uy * umy[i] : uy;
    const float u2 = umz ? uz * umz[i] : uz;
    const float mu = mx[i] * u0 + my[i] * u1 + mz[i] * u2;
    hx[i] = k * mu * u0;
    hy[i] = k * mu * u1;
    hz[i] = k * mu * u2;
}

struct ZhangLiArgs {
    int n[3];
    int pbc[3];
    float step[3];
    float jmul[3];
    float pre, xi, alpha;
};

__global__ void __launch_bounds__(kBlock)
zhangli_kernel(float *__restrict__ sx, float *__restrict__ sy, float *__restrict__ sz,
               const float *__restrict__ mx, const float *__restrict__ my,
               const float *__restrict__ mz, const float *__restrict__ jx,
               const float *__restrict__ jy, const float *__restrict__ jz,
               const float *__restrict__ msat_map, const float *__restrict__ pol_map,
               const float *__restrict__ xi_map, const float *__restrict__ alpha_map,
               ZhangLiArgs a)
{
    const long N = (long)a.n[0] * a.n[1] * a.n[2];
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int k = (int)(I % a.n[2]), j = (int)((I / a.n[2]) % a.n[1]),
              i = (int)(I / ((long)a.n[1] * a.n[2]));
    float stt[3];
    const float *mp[3] = {mx, my, mz};
    const float *jp[3] = {jx, jy, jz};
    zhangli_at(a.n, a.pbc, a.step, mp, I, i, j, k, jp, a.jmul, msat_map, pol_map, xi_map,
               alpha_map, a.pre, a.xi, a.alpha, stt);
    sx[I] = stt[0];
    sy[I] = stt[1];
    sz[I] = stt[2];
}

__global__ void __launch_bounds__(kBlock)
vec_madd_kernel(float *__restrict__ dx, float *__restrict__ dy, float *__restrict__ dz,
                const float *__restrict__ ax, const float *__restrict__ ay,
                const float *__restrict__ az, const float *__restrict__ bx,
                const float *__restrict__ by, const float *__restrict__ bz, float mx, float my,
                float mz, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (bx) {
        dx[i] = ax[i] + mx * bx[i];
        dy[i] = ay[i] + my * by[i];
        dz[i] = az[i] + mz * bz[i];
    } else {
        dx[i] = ax[i] + mx;
        dy[i] = ay[i] + my;
        dz[i] = az[i] + mz;
    }
}

__global__ void __launch_bounds__(kBlock) madd_kernel(float *dst, const float *a,
                                                      const float *__restrict__ b, float mul,
                                                      int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    dst[i] = b ? a[i] + mul * b[i] : a[i] + mul;
}

__global__ void __launch_bounds__(kBlock) madd2_kernel(float *a, const float *__restrict__ b,
                                                       float mb, const float *__restrict__ c,
                                                       float mc, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float B = b ? b[i] : 1.0f;
    const float C = c ? c[i] : 1.0f;
    a[i] += mb * B + mc * C;
}

__global__ void __launch_bounds__(kBlock) mul_kernel(float *dst, const float *a, const float *b,
                                                     int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    dst[i] = a[i] * b[i];
}

// ---------------------------------------------------------------- reductor

enum RedOp : int { RED_SUM = 0, RED_DOT, RED_MAX, RED_MIN, RED_MAXABS, RED_MAXDIFF, RED_MAXNORM };
static_assert(RED_SUM == MMX_RED_SUM && RED_DOT == MMX_RED_DOT && RED_MAX == MMX_RED_MAX &&
                  RED_MIN == MMX_RED_MIN && RED_MAXABS == MMX_RED_MAXABS &&
                  RED_MAXDIFF == MMX_RED_MAXDIFF && RED_MAXNORM == MMX_RED_MAXNORM,
              "reduction codes of mmx.h");

// One partial per workgroup: FP64 sums (sum, dot) or float extrema.
template <int OP>
__global__ void __launch_bounds__(kBlock) reduce_kernel(const float *__restrict__ a,
                                                        const float *__restrict__ b,
                                                        const float *__restrict__ c, long n,
                                                        double *__restrict__ part)
{
    __shared__ double sh[kBlock / 64];
    double acc;
    if (OP == RED_SUM || OP == RED_DOT || OP == RED_MAXABS || OP == RED_MAXDIFF || OP == RED_MAXNORM)
        acc = 0.0;
    else if (OP == RED_MAX)
        acc = -__builtin_inf();
    else
        acc = __builtin_inf();
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        if (OP == RED_SUM) acc += (double)a[i];
        if (OP == RED_DOT) acc += (double)a[i] * (double)b[i];
        if (OP == RED_MAX) acc = fmax(acc, (double)a[i]);
        if (OP == RED_MIN) acc = fmin(acc, (double)a[i]);
        if (OP == RED_MAXABS) acc = fmax(acc, (double)fabsf(a[i]));
        if (OP == RED_MAXDIFF) acc = fmax(acc, (double)fabsf(a[i] - b[i]));
        if (OP == RED_MAXNORM) {
            const float x = a[i], y = b[i], z = c[i];
            acc = fmax(acc, (double)(x * x + y * y + z * z));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double v = __shfl_xor(acc, o, 64);
        if (OP == RED_SUM || OP == RED_DOT) acc += v;
        else if (OP == RED_MIN) acc = fmin(acc, v);
        else acc = fmax(acc, v);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            if (OP == RED_SUM || OP == RED_DOT) r += sh[w];
            else if (OP == RED_MIN) r = fmin(r, sh[w]);
            else r = fmax(r, sh[w]);
        }
        part[blockIdx.x] = r;
    }
}

// The reductor's partial count: at most this many workgroups, each reducing
// a grid-strided share of the input to one FP64 partial.
constexpr int kRedMaxBlocks = 2048;

static int red_blocks(long n)
{
    return (int)std::min<long>(kRedMaxBlocks, std::max<long>(1, div_up(n, 4L * kBlock)));
}

static void launch_reduce(int op, const float *a, const float *b, const float *c, long n, int nb,
                          hipStream_t st, double *part)
{
    switch (op) {
    case RED_SUM: reduce_kernel<RED_SUM><<<nb, kBlock, 0, st>>>(a, b, c, n, part); break;
    case RED_DOT: reduce_kernel<RED_DOT><<<nb, kBlock, 0, st>>>(a, b, c, n, part); break;
    case RED_MAX: reduce_kernel<RED_MAX><<<nb, kBlock, 0, st>>>(a, b, c, n, part); break;
    case RED_MIN: reduce_kernel<RED_MIN><<<nb, kBlock, 0, st>>>(a, b, c, n, part); break;
    case RED_MAXABS: reduce_kernel<RED_MAXABS><<<nb, kBlock, 0, st>>>(a, b, c, n, part); break;
    case RED_MAXDIFF: reduce_kernel<RED_MAXDIFF><<<nb, kBlock, 0, st>>>(a, b, c, n, part); break;
    default: reduce_kernel<RED_MAXNORM><<<nb, kBlock, 0, st>>>(a, b, c, n, part); break;
    }
}

// Host finish of the partials: compensated sum (sum, dot) or extremum.
static double finish_partials(int op, const double *hp, int nb)
{
    double res = hp[0];
    if (op == RED_SUM || op == RED_DOT) {
        double s = 0.0, comp = 0.0;
        for (int i = 0; i < nb; i++) {
            const double y = hp[i] - comp;
            const double t = s + y;
            comp = (t - s) - y;
            s = t;
        }
        res = s;
    } else {
        for (int i = 1; i < nb; i++)
            res = (op == RED_MIN) ? std::min(res, hp[i]) : std::max(res, hp[i]);
    }
    return res;
}

// ------------------------------------------------ allocation-free reductions
// mmx_reduce_noalloc keeps its partials in this file-scope device array: one
// copy per GPU, resident until the process ends (16 KiB).  On the host the
// name `scratch` is a stand-in, not a device address: each device's address is
// looked up once with hipGetSymbolAddress (noalloc_setup) and only that
// address is handed to kernels and copies.
static __device__ __align__(256) double scratch[kRedMaxBlocks];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static double *g_noalloc_parts[kMaxDevices];

// Looks up this device's copy of `scratch`; `dev` must be the current device.
int noalloc_setup(int dev)
{
    MMX_REQUIRE(dev >= 0 && dev < kMaxDevices, MMX_ERR_ARG,
                "allocation-free reductions: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_parts[dev]) return MMX_OK;
    void *p = nullptr;
    MMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MMX_REQUIRE(p != nullptr, MMX_ERR_HIP, "allocation-free reductions: no device address");
    g_noalloc_parts[dev] = static_cast<double *>(p);
    return MMX_OK;
}

static int noalloc_parts(double **out)
{
    int dev = -1;
    MMX_HIP(hipGetDevice(&dev));
    MMX_TRY(noalloc_setup(dev)); // first call on a device that mmx_init did not set up
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    *out = g_noalloc_parts[dev];
    return MMX_OK;
}

} // namespace mmx

struct mmx_reductor {
    long n = 0;
    int blocks = 0;
    double *dpart = nullptr;
    std::vector<double> hpart;
};

namespace mmx {

static int reduce_run(mmx_reductor *r, int op, const float *a, const float *b, const float *c,
                      long n, hipStream_t st, double *out)
{
    MMX_REQUIRE(r && a, MMX_ERR_ARG, "reductor: NULL argument");
    MMX_REQUIRE(n > 0 && n <= r->n, MMX_ERR_ARG, "reductor: size %ld outside (0, %ld]", n, r->n);
    const int nb = std::min(r->blocks, red_blocks(n));
    launch_reduce(op, a, b, c, n, nb, st, r->dpart);
    MMX_CHECK_LAUNCH();
    MMX_HIP(hipMemcpyAsync(r->hpart.data(), r->dpart, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    MMX_HIP(hipStreamSynchronize(st));
    *out = finish_partials(op, r->hpart.data(), nb);
    return MMX_OK;
}

} // namespace mmx

using namespace mmx;

extern "C" {

int mmx_normalize(float *mx, float *my, float *mz, const float *norm_map, int N, void *stream)
{
    MMX_REQUIRE(mx && my && mz && N >= 0, MMX_ERR_ARG, "mmx_normalize: bad arguments");
    if (N == 0) return MMX_OK;
    normalize_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(mx, my, mz, norm_map, N);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_torque(float *tx, float *ty, float *tz, const float *mx, const float *my, const float *mz,
               const float *hx, const float *hy, const float *hz, const float *alpha_map,
               float alpha_mul, int N, void *stream)
{
    MMX_REQUIRE(tx && ty && tz && mx && my && mz && hx && hy && hz && N >= 0, MMX_ERR_ARG,
                "mmx_torque: NULL array");
    if (N == 0) return MMX_OK;
    torque_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(tx, ty, tz, mx, my, mz, hx, hy, hz,
                                                                alpha_map, alpha_mul, N);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_exchange6(float *hx, float *hy, float *hz, const float *mx, const float *my,
                  const float *mz, const float *msat_map, const float *aex_map, float pre,
                  int N0, int N1, int N2, int periodic0, int periodic1, int periodic2,
                  float cell0, float cell1, float cell2, void *stream)
{
    MMX_REQUIRE(hx && hy && hz && mx && my && mz, MMX_ERR_ARG, "mmx_exchange6: NULL array");
    MMX_REQUIRE(N0 > 0 && N1 > 0 && N2 > 0, MMX_ERR_ARG, "mmx_exchange6: bad size");
    MMX_REQUIRE(cell0 > 0 && cell1 > 0 && cell2 > 0, MMX_ERR_ARG, "mmx_exchange6: bad cell size");
    const long N = (long)N0 * N1 * N2;
    // 1/cell^2 in double, then rounded (exchange6.cu:152-154)
    const float c0 = (float)(1.0 / ((double)cell0 * (double)cell0));
    const float c1 = (float)(1.0 / ((double)cell1 * (double)cell1));
    const float c2 = (float)(1.0 / ((double)cell2 * (double)cell2));
    const long nj = div_up(N1, kEJ), nk = div_up(N2, kEK);
    // the LDS tile is 64 cells wide in x: narrow grids (N2 < 32) leave most
    // of it idle and take the per-cell kernel; grid size below 2^32 threads
    if (!msat_map && !aex_map && N2 >= kEK / 2 && (long)N0 * nj * nk * kBlock < (1L << 32)) {
        exchange6_lds_kernel<<<dim3((unsigned)(N0 * nj * nk)), kBlock, 0, as_stream(stream)>>>(
            hx, hy, hz, mx, my, mz, pre, N0, N1, N2, periodic0, periodic1, periodic2, c0, c1, c2,
            (int)nj, (int)nk);
    } else {
        exchange6_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(
            hx, hy, hz, mx, my, mz, msat_map, aex_map, pre, N0, N1, N2, periodic0, periodic1,
            periodic2, c0, c1, c2);
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_uniaxial_anisotropy(float *hx, float *hy, float *hz, const float *mx, const float *my,
                            const float *mz, const float *ku_map, const float *msat_map,
                            float ku2_mu0msat_mul, const float *u_map_x, float u_mul_x,
                            const float *u_map_y, float u_mul_y, const float *u_map_z,
                            float u_mul_z, int N, void *stream)
{
    MMX_REQUIRE(hx && hy && hz && mx && my && mz && N >= 0, MMX_ERR_ARG,
                "mmx_uniaxial_anisotropy: NULL array");
    if (N == 0) return MMX_OK;
    uniaxial_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(
        hx, hy, hz, mx, my, mz, ku_map, msat_map, ku2_mu0msat_mul, u_map_x, u_mul_x, u_map_y,
        u_mul_y, u_map_z, u_mul_z, N);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_zhangli(float *sttx, float *stty, float *sttz, const float *mx, const float *my,
                const float *mz, const float *jx_map, const float *jy_map, const float *jz_map,
                const float *msat_map, const float *pol_map, const float *xi_map,
                const float *alpha_map, float jmul_x, float jmul_y, float jmul_z, float pred,
                float xi_mul, float alpha_mul, int N0, int N1, int N2, float cell0, float cell1,
                float cell2, int periodic0, int periodic1, int periodic2, void *stream)
{
    MMX_REQUIRE(sttx && stty && sttz && mx && my && mz, MMX_ERR_ARG, "mmx_zhangli: NULL array");
    MMX_REQUIRE(N0 > 0 && N1 > 0 && N2 > 0, MMX_ERR_ARG, "mmx_zhangli: bad size");
    const long N = (long)N0 * N1 * N2;
    hipStream_t st = as_stream(stream);
    ZhangLiArgs a{};
    a.n[0] = N0; a.n[1] = N1; a.n[2] = N2;
    a.pbc[0] = periodic0; a.pbc[1] = periodic1; a.pbc[2] = periodic2;
    a.step[0] = 1.0f / (12.0f * cell0);
    a.step[1] = 1.0f / (12.0f * cell1);
    a.step[2] = 1.0f / (12.0f * cell2);
    a.jmul[0] = jmul_x; a.jmul[1] = jmul_y; a.jmul[2] = jmul_z;
    a.pre = pred; a.xi = xi_mul; a.alpha = alpha_mul;
    zhangli_kernel<<<grid1d(N), kBlock, 0, st>>>(sttx, stty, sttz, mx, my, mz, jx_map, jy_map,
                                                 jz_map, msat_map, pol_map, xi_map, alpha_map, a);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_vec_madd(float *dstx, float *dsty, float *dstz, const float *ax, const float *ay,
                 const float *az, const float *bx, const float *by, const float *bz, float mulx,
                 float muly, float mulz, int N, void *stream)
{
    MMX_REQUIRE(dstx && dsty && dstz && ax && ay && az && N >= 0, MMX_ERR_ARG,
                "mmx_vec_madd: NULL array");
    MMX_REQUIRE((bx && by && bz) || (!bx && !by && !bz), MMX_ERR_ARG,
                "mmx_vec_madd: b must be all NULL or all set");
    if (N == 0) return MMX_OK;
    vec_madd_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(dstx, dsty, dstz, ax, ay, az, bx,
                                                                  by, bz, mulx, muly, mulz, N);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_madd(float *dst, const float *a, const float *b, float mul, int N, void *stream)
{
    MMX_REQUIRE(dst && a && N >= 0, MMX_ERR_ARG, "mmx_madd: NULL array");
    if (N == 0) return MMX_OK;
    madd_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(dst, a, b, mul, N);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_madd1(float *a, const float *b, float mulb, int N, void *stream)
{
    MMX_REQUIRE(a && N >= 0, MMX_ERR_ARG, "mmx_madd1: NULL array");
    if (N == 0) return MMX_OK;
    madd2_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(a, b, mulb, nullptr, 0.0f, N);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_madd2(float *a, const float *b, float mulb, const float *c, float mulc, int N,
              void *stream)
{
    MMX_REQUIRE(a && N >= 0, MMX_ERR_ARG, "mmx_madd2: NULL array");
    if (N == 0) return MMX_OK;
    madd2_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(a, b, mulb, c, mulc, N);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_mul(float *dst, const float *a, const float *b, int N, void *stream)
{
    MMX_REQUIRE(dst && a && b && N >= 0, MMX_ERR_ARG, "mmx_mul: NULL array");
    if (N == 0) return MMX_OK;
    mul_kernel<<<grid1d(N), kBlock, 0, as_stream(stream)>>>(dst, a, b, N);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int mmx_reductor_create(mmx_reductor **r, long n)
{
    MMX_REQUIRE(r && n > 0, MMX_ERR_ARG, "mmx_reductor_create: bad arguments");
    mmx_reductor *x = new mmx_reductor();
    x->n = n;
    x->blocks = red_blocks(n);
    x->hpart.resize(x->blocks);
    if (hipMalloc(&x->dpart, sizeof(double) * x->blocks) != hipSuccess) {
        delete x;
        set_error("mmx_reductor_create: out of device memory");
        return MMX_ERR_NOMEM;
    }
    *r = x;
    return MMX_OK;
}

int mmx_reductor_destroy(mmx_reductor *r)
{
    if (!r) return MMX_OK;
    (void)hipFree(r->dpart);
    delete r;
    return MMX_OK;
}

int mmx_reductor_sum(mmx_reductor *r, const float *a, long n, void *stream, double *out)
{
    return reduce_run(r, RED_SUM, a, nullptr, nullptr, n, as_stream(stream), out);
}

int mmx_reductor_dot(mmx_reductor *r, const float *a, const float *b, long n, void *stream,
                     double *out)
{
    MMX_REQUIRE(b, MMX_ERR_ARG, "mmx_reductor_dot: NULL b");
    return reduce_run(r, RED_DOT, a, b, nullptr, n, as_stream(stream), out);
}

static int reduce_f(mmx_reductor *r, int op, const float *a, const float *b, const float *c,
                    long n, void *stream, float *out)
{
    double d = 0;
    MMX_TRY(reduce_run(r, op, a, b, c, n, as_stream(stream), &d));
    *out = (float)d;
    return MMX_OK;
}

int mmx_reductor_max(mmx_reductor *r, const float *a, long n, void *stream, float *out)
{
    return reduce_f(r, RED_MAX, a, nullptr, nullptr, n, stream, out);
}
int mmx_reductor_min(mmx_reductor *r, const float *a, long n, void *stream, float *out)
{
    return reduce_f(r, RED_MIN, a, nullptr, nullptr, n, stream, out);
}
int mmx_reductor_maxabs(mmx_reductor *r, const float *a, long n, void *stream, float *out)
{
    return reduce_f(r, RED_MAXABS, a, nullptr, nullptr, n, stream, out);
}
int mmx_reductor_maxdiff(mmx_reductor *r, const float *a, const float *b, long n, void *stream,
                         float *out)
{
    MMX_REQUIRE(b, MMX_ERR_ARG, "mmx_reductor_maxdiff: NULL b");
    return reduce_f(r, RED_MAXDIFF, a, b, nullptr, n, stream, out);
}
int mmx_reductor_maxnorm(mmx_reductor *r, const float *ax, const float *ay, const float *az,
                         long n, void *stream, float *out)
{
    MMX_REQUIRE(ay && az, MMX_ERR_ARG, "mmx_reductor_maxnorm: NULL component");
    double d = 0;
    MMX_TRY(reduce_run(r, RED_MAXNORM, ax, ay, az, n, as_stream(stream), &d));
    *out = (float)sqrt(d);
    return MMX_OK;
}

int mmx_reduce_noalloc(int op, const float *a, const float *b, const float *c, long n,
                       void *stream, double *out)
{
    MMX_REQUIRE(op >= MMX_RED_SUM && op <= MMX_RED_MAXNORM, MMX_ERR_ARG,
                "mmx_reduce_noalloc: unknown operation %d", op);
    MMX_REQUIRE(a && out, MMX_ERR_ARG, "mmx_reduce_noalloc: NULL argument");
    MMX_REQUIRE(b || !(op == MMX_RED_DOT || op == MMX_RED_MAXDIFF || op == MMX_RED_MAXNORM),
                MMX_ERR_ARG, "mmx_reduce_noalloc: operation %d needs b", op);
    MMX_REQUIRE(c || op != MMX_RED_MAXNORM, MMX_ERR_ARG, "mmx_reduce_noalloc: maxnorm needs c");
    MMX_REQUIRE(n > 0 && n <= MMX_NOALLOC_MAX_N, MMX_ERR_SIZE,
                "mmx_reduce_noalloc: size %ld outside (0, %ld] (no workspace is allocated)", n,
                (long)MMX_NOALLOC_MAX_N);
    double *parts = nullptr;
    MMX_TRY(noalloc_parts(&parts));
    const int nb = red_blocks(n); // <= kRedMaxBlocks: every partial read is written below
    hipStream_t st = as_stream(stream);
    launch_reduce(op, a, b, c, n, nb, st, parts);
    MMX_CHECK_LAUNCH();
    double hp[kRedMaxBlocks];
    MMX_HIP(hipMemcpyAsync(hp, parts, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    MMX_HIP(hipStreamSynchronize(st));
    const double res = finish_partials(op, hp, nb);
    // the float results of the reductor's max/min/maxabs/maxdiff/maxnorm calls
    if (op == MMX_RED_SUM || op == MMX_RED_DOT) *out = res;
    else if (op == MMX_RED_MAXNORM) *out = (double)(float)sqrt(res);
    else *out = (double)(float)res;
    return MMX_OK;
}

} // extern "C"
