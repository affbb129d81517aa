// surrogate.hip -- learned correction of a 2x-coarsened demag field.
//
// No reference counterpart (the reference has no learned component).  For a
// film of n0 <= 5 layers on an n1 x n2 in-plane grid (internal z,y,x order):
//
//   coarse m  = means of 2x2 in-plane blocks of m, layers kept (coarsening the
//               few layers of a film loses most of the field: 92 % error on a
//               random state against 23 % for in-plane coarsening);
//   coarse B  = exact demag of the coarse grid (its own demag plan);
//   net input = per in-plane pixel, channel c*n0+i = m_c of layer i,
//               3*n0 + c*n0+i = the upsampled coarse B_c / (mu0 Ms), and
//               6*n0, 6*n0+1 = the position in the coarse cell (+-1 for odd /
//               even y, x: the residual operator has period 2), in bf16;
//   net       = 4 conv 3x3 layers 32 -> 32 -> 32 -> 32 -> 16 (ReLU after the
//               first three), BF16 MFMA with FP32 accumulation (gfx950's
//               v_mfma_f32_16x16x32_bf16, 32 channels or pixels per
//               instruction; 16-channel inputs on v_mfma_f32_16x16x16_bf16),
//               as implicit GEMMs over 16 x 8 pixel tiles;
//   B_fine   ~= upsampled coarse B + mu0 Ms * out[c*n0+i].
//
// Training (on the GPU, by these kernels): relative squared error of the
// residual against the exact fine field (sum d^2 / sum target^2 per sample), backward-data as the same convolution kernel with flipped
// and transposed weights and a ReLU-mask epilogue, weight gradients as MFMA
// products over the pixel tiles with per-tile partial sums reduced in a fixed
// order (deterministic), Adam on FP32 master weights.  The CPU oracle
// (oracle/surrogate.py) rounds to bf16 at the same points.
//
// In the RK12 solver (sim.cpp) the corrector's field evaluation uses the
// surrogate; the first stage (at the accepted m) stays exact, and every
// `check_every` steps the exact field at the accepted m judges the surrogate's
// there: above the tolerance the solver falls back to the exact path.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "demag_tiles.hpp"
#include "sim_impl.hpp"

namespace mmx {


typedef unsigned short bf16_t;
typedef short sh4 __attribute__((ext_vector_type(4)));
typedef short sh8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float fl4 __attribute__((ext_vector_type(4)));

// D += A B over k = 32 on gfx950's double-K bf16 MFMA: lane (n = lane % 16,
// g = lane / 16) passes 8 values of row n of A and 8 of column n of B, and
// holds D[4 g + e][n], e < 4.  Both operands index their k by the same
// (g, element), so the sum covers each k once whatever order the unit uses.
__device__ __forceinline__ fl4 mfma_k32(sh8 a, sh8 b, fl4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                   __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

constexpr int kSurCh[5] = {32, 32, 32, 32, 16};

__device__ __forceinline__ bf16_t f2bf(float x)
{
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned int)h << 16); }

struct SurGeom {
    int n0, n1, n2;       // fine grid
    int c0, c1, c2;       // coarse grid
    int f0;               // z coarsening factor (1 or 2)
    long N, Nc;
    float inv_mu0ms;      // 1 / (mu0 Ms)
    float msat;
};

// --------------------------------------------------------- field kernels

// Block means of m (times the Msat map) on the coarse grid; with PRED the fine
// m is first the RK predictor normalize(m0 + h tau0) (maddKern + normUpdater,
// as in the exact predictor), written to mpred.
template <int PRED>
__global__ void __launch_bounds__(kBlock)
sur_coarsen(SurGeom s, const float *__restrict__ m0, const float *__restrict__ tau0, float h,
            const float *__restrict__ msat_map, float *__restrict__ mpred, float *__restrict__ mc)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= s.Nc) return;
    const int ck = (int)(t % s.c2), cj = (int)((t / s.c2) % s.c1), ci = (int)(t / ((long)s.c2 * s.c1));
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < s.f0; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                const long I = ((long)(ci * s.f0 + a) * s.n1 + (cj * 2 + b)) * s.n2 + (ck * 2 + c);
                float v0 = m0[I], v1 = m0[s.N + I], v2 = m0[2 * s.N + I];
                if (PRED) {
                    v0 = v0 + h * tau0[I];
                    v1 = v1 + h * tau0[s.N + I];
                    v2 = v2 + h * tau0[2 * s.N + I];
                    const float norm = msat_map ? msat_map[I] : 1.0f;
                    const float len = sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
                    float scale = (norm == 0.0f) ? 0.0f : 1.0f;
                    scale = (len != 0.0f) ? scale / len : 0.0f;
                    v0 *= scale;
                    v1 *= scale;
                    v2 *= scale;
                    mpred[I] = v0;
                    mpred[s.N + I] = v1;
                    mpred[2 * s.N + I] = v2;
                }
                const float ms = msat_map ? msat_map[I] : 1.0f;
                acc[0] += v0 * ms;
                acc[1] += v1 * ms;
                acc[2] += v2 * ms;
            }
    const float inv = 1.0f / (float)(s.f0 * 4);
    mc[t] = acc[0] * inv;
    mc[s.Nc + t] = acc[1] * inv;
    mc[2 * s.Nc + t] = acc[2] * inv;
}

__device__ __forceinline__ long coarse_of(const SurGeom &s, int i, int j, int k)
{
    return ((long)(i / s.f0) * s.c1 + j / 2) * s.c2 + k / 2;
}

// net input channel ch of pixel (y, x), from m and the coarse B (tesla)
__device__ __forceinline__ float input_channel(const SurGeom &s, const float *__restrict__ m,
                                               const float *__restrict__ bc, int y, int x, int ch)
{
    const int nm = 3 * s.n0;
    if (ch < nm) {
        const int c = ch / s.n0, i = ch - c * s.n0;
        return m[c * s.N + ((long)i * s.n1 + y) * s.n2 + x];
    }
    if (ch < 2 * nm) {
        const int q = ch - nm, c = q / s.n0, i = q - c * s.n0;
        return bc[c * s.Nc + coarse_of(s, i, y, x)] * s.inv_mu0ms;
    }
    if (ch == 2 * nm) return (y & 1) ? 1.0f : -1.0f;
    if (ch == 2 * nm + 1) return (x & 1) ? 1.0f : -1.0f;
    return 0.0f;
}

// packed bf16 input [P][32] (used by training and the tests; inference packs
// inside the first convolution)
__global__ void __launch_bounds__(kBlock)
sur_pack(SurGeom s, const float *__restrict__ m, const float *__restrict__ bc,
         bf16_t *__restrict__ x0)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x; // pixel * 32 + channel
    const long P = (long)s.n1 * s.n2;
    if (t >= P * 32) return;
    const int ch = (int)(t & 31);
    const long p = t >> 5;
    const int y = (int)(p / s.n2), x = (int)(p - (long)y * s.n2);
    x0[t] = f2bf(input_channel(s, m, bc, y, x, ch));
}

// ------------------------------------------------------------ convolution

enum { CE_RELU_BF16 = 0, CE_LINEAR_F32 = 1, CE_MASK_BF16 = 2 };

// out[p][co] = epilogue(sum_tap sum_ci in[p + off(tap)][ci] * wt[tap][co][ci])
// over a 16 x 8 pixel tile per workgroup of 4 waves (2 rows each).  The MFMA
// computes D = A B with A = weights (rows co, k = ci) and B = the input
// (k = ci, columns = 16 pixels of a row), so each lane ends up holding 4
// consecutive output channels of one pixel.  PACK: the input is formed from
// (m, coarse B) instead of being read from `in`.
template <int CIN, int COUT, int EPI, int PACK>
__global__ void __launch_bounds__(kBlock)
sur_conv(const bf16_t *__restrict__ in, const bf16_t *__restrict__ wt,
         const float *__restrict__ bias, const bf16_t *__restrict__ mask, void *__restrict__ out,
         int H, int W, SurGeom s, const float *__restrict__ m, const float *__restrict__ bc)
{
    constexpr int CP = CIN + 8; // LDS pitch per pixel (bf16 elements)
    constexpr int NQ = CIN / 4;
    __shared__ __align__(16) bf16_t tile[10 * 18 * CP];
    __shared__ __align__(16) bf16_t wsh[9 * COUT * CP];
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 8;
    for (int t = threadIdx.x; t < 10 * 18 * NQ; t += kBlock) {
        const int q = t % NQ, pc = t / NQ, cx = pc % 18, cy = pc / 18;
        const int gy = y0 + cy - 1, gx = x0 + cx - 1;
        bf16_t v[4] = {0, 0, 0, 0};
        if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
            if (PACK) {
#pragma unroll
                for (int e = 0; e < 4; e++) v[e] = f2bf(input_channel(s, m, bc, gy, gx, 4 * q + e));
            } else {
                const uint2 u = *reinterpret_cast<const uint2 *>(in + ((long)gy * W + gx) * CIN + 4 * q);
                v[0] = (bf16_t)(u.x & 0xffffu);
                v[1] = (bf16_t)(u.x >> 16);
                v[2] = (bf16_t)(u.y & 0xffffu);
                v[3] = (bf16_t)(u.y >> 16);
            }
        }
        bf16_t *d = tile + (cy * 18 + cx) * CP + 4 * q;
        *reinterpret_cast<uint2 *>(d) = uint2{(unsigned)v[0] | ((unsigned)v[1] << 16),
                                              (unsigned)v[2] | ((unsigned)v[3] << 16)};
    }
    for (int t = threadIdx.x; t < 9 * COUT * NQ; t += kBlock) {
        const int q = t % NQ, r = t / NQ;
        *reinterpret_cast<uint2 *>(wsh + r * CP + 4 * q) =
            *reinterpret_cast<const uint2 *>(wt + (long)r * CIN + 4 * q);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kq = (lane >> 4) * 4, kq8 = (lane >> 4) * 8;
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int row = wave * 2 + r;
        fl4 acc[COUT / 16];
#pragma unroll
        for (int mt = 0; mt < COUT / 16; mt++) acc[mt] = fl4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
            const bf16_t *bp = tile + ((row + dy) * 18 + n + dx) * CP;
            if constexpr (CIN % 32 == 0) { // 32 input channels per MFMA
#pragma unroll
                for (int kb = 0; kb < CIN; kb += 32) {
                    const sh8 b = *reinterpret_cast<const sh8 *>(bp + kb + kq8);
#pragma unroll
                    for (int mt = 0; mt < COUT / 16; mt++) {
                        const sh8 a = *reinterpret_cast<const sh8 *>(
                            wsh + (tap * COUT + mt * 16 + n) * CP + kb + kq8);
                        acc[mt] = mfma_k32(a, b, acc[mt]);
                    }
                }
            } else { // 16 channels (the last layer's backward-data input)
#pragma unroll
                for (int kb = 0; kb < CIN; kb += 16) {
                    const sh4 b = *reinterpret_cast<const sh4 *>(bp + kb + kq);
#pragma unroll
                    for (int mt = 0; mt < COUT / 16; mt++) {
                        const sh4 a = *reinterpret_cast<const sh4 *>(
                            wsh + (tap * COUT + mt * 16 + n) * CP + kb + kq);
                        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[mt], 0, 0, 0);
                    }
                }
            }
        }
        const long pix = (long)(y0 + row) * W + x0 + n;
#pragma unroll
        for (int mt = 0; mt < COUT / 16; mt++) {
            const int co = mt * 16 + kq;
            float v[4] = {acc[mt][0], acc[mt][1], acc[mt][2], acc[mt][3]};
            if (EPI == CE_LINEAR_F32) {
#pragma unroll
                for (int e = 0; e < 4; e++) v[e] += bias[co + e];
                *reinterpret_cast<float4 *>(reinterpret_cast<float *>(out) + pix * COUT + co) =
                    float4{v[0], v[1], v[2], v[3]};
            } else {
                if (EPI == CE_RELU_BF16) {
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = fmaxf(v[e] + bias[co + e], 0.0f);
                } else {
                    const uint2 mk = *reinterpret_cast<const uint2 *>(mask + pix * COUT + co);
                    const bf16_t mb[4] = {(bf16_t)(mk.x & 0xffffu), (bf16_t)(mk.x >> 16),
                                          (bf16_t)(mk.y & 0xffffu), (bf16_t)(mk.y >> 16)};
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = bf2f(mb[e]) > 0.0f ? v[e] : 0.0f;
                }
                const unsigned lo = (unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16);
                const unsigned hi = (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16);
                *reinterpret_cast<uint2 *>(reinterpret_cast<bf16_t *>(out) + pix * COUT + co) =
                    uint2{lo, hi};
            }
        }
    }
}

// Weight and bias gradients of one layer over a 16 x 16 pixel tile:
// part[tile][(tap*CIN + ci)*COUT + co] = sum_p dz[p][co] x[p + off(tap)][ci],
// part[tile][9*CIN*COUT + co] = sum_p dz[p][co].  MFMA with A = dz^T (rows co,
// k = pixels) and B = the shifted input (k = pixels, columns ci).
template <int CIN, int COUT>
__global__ void __launch_bounds__(kBlock)
sur_conv_wgrad(const bf16_t *__restrict__ x, const bf16_t *__restrict__ dz, float *__restrict__ part,
               int H, int W)
{
    constexpr int XP = CIN + 8, DP = COUT + 8;
    constexpr int NQX = CIN / 4, NQD = COUT / 4;
    __shared__ __align__(16) bf16_t xt[18 * 18 * XP];
    __shared__ __align__(16) bf16_t dzt[256 * DP];
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    const long tileid = (long)blockIdx.y * gridDim.x + blockIdx.x;
    for (int t = threadIdx.x; t < 18 * 18 * NQX; t += kBlock) {
        const int q = t % NQX, pc = t / NQX, cx = pc % 18, cy = pc / 18;
        const int gy = y0 + cy - 1, gx = x0 + cx - 1;
        uint2 u = uint2{0u, 0u};
        if (gy >= 0 && gy < H && gx >= 0 && gx < W)
            u = *reinterpret_cast<const uint2 *>(x + ((long)gy * W + gx) * CIN + 4 * q);
        *reinterpret_cast<uint2 *>(xt + (cy * 18 + cx) * XP + 4 * q) = u;
    }
    for (int t = threadIdx.x; t < 256 * NQD; t += kBlock) {
        const int q = t % NQD, p = t / NQD;
        const int gy = y0 + p / 16, gx = x0 + p % 16;
        *reinterpret_cast<uint2 *>(dzt + p * DP + 4 * q) =
            *reinterpret_cast<const uint2 *>(dz + ((long)gy * W + gx) * COUT + 4 * q);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kq = (lane >> 4) * 4, kq8 = (lane >> 4) * 8;
    constexpr int MT = COUT / 16, NT = CIN / 16, ITEMS = 9 * MT * NT;
    float *dst = part + tileid * (long)(9 * CIN * COUT + COUT);
    for (int it = wave; it < ITEMS; it += 4) {
        const int tap = it / (MT * NT), mt = (it / NT) % MT, nt = it % NT;
        const int dy = tap / 3, dx = tap % 3;
        fl4 acc = fl4{0.0f, 0.0f, 0.0f, 0.0f};
        // k = the tile's 256 pixels, 32 per MFMA
#pragma unroll 4
        for (int ks = 0; ks < 8; ks++) {
            sh8 a, b;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int p = ks * 32 + kq8 + e, py = p >> 4, px = p & 15;
                a[e] = (short)dzt[p * DP + mt * 16 + n];
                b[e] = (short)xt[((py + dy) * 18 + px + dx) * XP + nt * 16 + n];
            }
            acc = mfma_k32(a, b, acc);
        }
        // D[m = co][n = ci]: lane holds co = mt*16 + kq + e, ci = nt*16 + n
#pragma unroll
        for (int e = 0; e < 4; e++)
            dst[(tap * CIN + nt * 16 + n) * COUT + mt * 16 + kq + e] = acc[e];
    }
    // bias gradient: wave w sums pixels 64 w .. 64 w + 63 of channel `lane`,
    // then the four partial sums in wave order
    __shared__ float bsum[4][COUT];
    if (lane < COUT) {
        float sum = 0.0f;
#pragma unroll 8
        for (int p = 64 * wave; p < 64 * wave + 64; p++) sum += bf2f(dzt[p * DP + lane]);
        bsum[wave][lane] = sum;
    }
    __syncthreads();
    if (threadIdx.x < COUT)
        dst[9 * CIN * COUT + threadIdx.x] =
            ((bsum[0][threadIdx.x] + bsum[1][threadIdx.x]) + bsum[2][threadIdx.x]) +
            bsum[3][threadIdx.x];
}

// grad[e] = sum over tiles of part[tile][e] in a fixed order, with the
// tiles split over the 16 waves of a workgroup: wave g sums tiles
// [g per, (g + 1) per) for 64 consecutive elements, then wave 0 adds the 16
// partial sums in wave order (deterministic; one serial chain per element
// over all tiles was latency-bound: 98 us per layer at 256 tiles).
constexpr int kRedGroups = 16;
__global__ void __launch_bounds__(64 * kRedGroups)
sur_reduce_tiles(const float *__restrict__ part, int ntile, int len, float *__restrict__ grad)
{
    __shared__ float ps[kRedGroups][64];
    const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + lane;
    const int per = (ntile + kRedGroups - 1) / kRedGroups;
    const int t0 = grp * per, t1 = min(ntile, t0 + per);
    float s = 0.0f;
    if (e < len) {
#pragma unroll 4
        for (int t = t0; t < t1; t++) s += part[(long)t * len + e];
    }
    ps[grp][lane] = s;
    __syncthreads();
    if (grp == 0 && e < len) {
        float r = 0.0f;
        for (int g = 0; g < kRedGroups; g++) r += ps[g][lane];
        grad[e] = r;
    }
}

// out[0] = sum of the n values of part in a fixed order: thread i of one
// workgroup sums i, i + 1024, ..., then a tree in LDS (deterministic).
__global__ void __launch_bounds__(1024)
sur_reduce_one(const float *__restrict__ part, int n, float *__restrict__ out)
{
    __shared__ float ps[1024];
    float s = 0.0f;
    for (int t = threadIdx.x; t < n; t += 1024) s += part[t];
    ps[threadIdx.x] = s;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (threadIdx.x < h) ps[threadIdx.x] += ps[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = ps[0];
}

__device__ __forceinline__ float target_at(const SurGeom &s, const float *__restrict__ bfine,
                                           const float *__restrict__ bc,
                                           const float *__restrict__ tgt_direct, long p, int ch)
{
    const int nused = 3 * s.n0;
    if (tgt_direct) return tgt_direct[p * nused + ch];
    const int c = ch / s.n0, i = ch - c * s.n0;
    const int y = (int)(p / s.n2), x = (int)(p - (long)y * s.n2);
    const long I = ((long)i * s.n1 + y) * s.n2 + x;
    return (bfine[c * s.N + I] - bc[c * s.Nc + coarse_of(s, i, y, x)]) * s.inv_mu0ms;
}

// per-block partial sums of target^2
__global__ void __launch_bounds__(kBlock)
sur_target_sq(SurGeom s, const float *__restrict__ bfine, const float *__restrict__ bc,
              const float *__restrict__ tgt_direct, float *__restrict__ part)
{
    __shared__ float red[kBlock / 64];
    const long P = (long)s.n1 * s.n2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x; // pixel * 16 + channel
    float v2 = 0.0f;
    if (t < P * 16 && (int)(t & 15) < 3 * s.n0) {
        const float v = target_at(s, bfine, bc, tgt_direct, t >> 4, (int)(t & 15));
        v2 = v * v;
    }
    v2 = wave_sum(v2);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v2;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s2 = 0.0f;
        for (int w = 0; w < kBlock / 64; w++) s2 += red[w];
        part[blockIdx.x] = s2;
    }
}

// d loss / d out = 2 (out - target) / sum target^2 (bf16) and per-block
// partial sums of the squared error; *t2 = sum target^2 (device)
__global__ void __launch_bounds__(kBlock)
sur_loss(SurGeom s, const float *__restrict__ out, const float *__restrict__ bfine,
         const float *__restrict__ bc, const float *__restrict__ tgt_direct,
         const float *__restrict__ t2, bf16_t *__restrict__ gout, float *__restrict__ part)
{
    __shared__ float red[kBlock / 64];
    const long P = (long)s.n1 * s.n2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x; // pixel * 16 + channel
    const int nused = 3 * s.n0;
    float e2 = 0.0f;
    const float gscale = 2.0f / fmaxf(*t2, 1e-30f);
    if (t < P * 16) {
        const int ch = (int)(t & 15);
        const long p = t >> 4;
        float g = 0.0f;
        if (ch < nused) {
            const float d = out[t] - target_at(s, bfine, bc, tgt_direct, p, ch);
            e2 = d * d;
            g = gscale * d;
        }
        gout[t] = f2bf(g);
    }
    e2 = wave_sum(e2);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s2 = 0.0f;
        for (int w = 0; w < kBlock / 64; w++) s2 += red[w];
        part[blockIdx.x] = s2;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void __launch_bounds__(kBlock)
sur_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
         float *__restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float a = b1 * m1[e] + (1.0f - b1) * gi;
    const float b = b2 * m2[e] + (1.0f - b2) * gi * gi;
    m1[e] = a;
    m2[e] = b;
    w[e] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// bf16 copies of a layer's weights: forward [tap][co][ci], backward-data
// [tap][ci][co] with the tap flipped (the transposed convolution)
__global__ void __launch_bounds__(kBlock)
sur_prep_weights(const float *__restrict__ W, int cin, int cout, bf16_t *__restrict__ wt,
                 bf16_t *__restrict__ wbt)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x; // canonical (tap, ci, co)
    if (e >= 9 * cin * cout) return;
    const int co = e % cout, ci = (e / cout) % cin, tap = e / (cout * cin);
    const bf16_t v = f2bf(W[e]);
    wt[((long)tap * cout + co) * cin + ci] = v;
    wbt[((long)(8 - tap) * cin + ci) * cout + co] = v;
}

// H_demag (A/m) = upsampled coarse B / mu0 + Ms * out, into the field epilogue;
// optionally also out_scale * H_demag into hd_out (1: A/m, mu0: tesla)
template <int MODE>
__global__ void __launch_bounds__(kBlock)
sur_epilogue(DemagGeom g, SurGeom s, FieldParams fp, EpilogueIO io, const float *__restrict__ out,
             const float *__restrict__ bc, float *__restrict__ hd_out, float out_scale)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float errmax = 0.0f;
    if (I < s.N) {
        const int k = (int)(I % s.n2), j = (int)((I / s.n2) % s.n1), i = (int)(I / ((long)s.n2 * s.n1));
        const long p = (long)j * s.n2 + k;
        const long ic = coarse_of(s, i, j, k);
        float hd[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
            hd[c] = bc[c * s.Nc + ic] * (1.0f / (float)kMu0) + s.msat * out[p * 16 + c * s.n0 + i];
        if (hd_out) {
#pragma unroll
            for (int c = 0; c < 3; c++) hd_out[c * s.N + I] = out_scale * hd[c];
        }
        if (MODE != EPI_B) cell_epilogue<MODE>(g, fp, io, I, i, j, k, hd, errmax);
    }
    if (MODE == EPI_CORR) block_max_to(errmax, io.err);
}

// sum (a - b)^2 and sum c^2 per block (FP64 partials)
__global__ void __launch_bounds__(kBlock)
sur_diff_norm(const float *__restrict__ a, const float *__restrict__ b, const float *__restrict__ c,
              long n, double *__restrict__ part)
{
    __shared__ double red[2][kBlock / 64];
    double d2 = 0.0, b2 = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double d = (double)a[t] - (double)b[t];
        d2 += d * d;
        b2 += (double)c[t] * (double)c[t];
    }
    for (int o = 32; o > 0; o >>= 1) {
        d2 += __shfl_down(d2, o);
        b2 += __shfl_down(b2, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = d2;
        red[1][threadIdx.x >> 6] = b2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int w = 0; w < kBlock / 64; w++) {
            x += red[0][w];
            y += red[1][w];
        }
        part[2 * blockIdx.x] = x;
        part[2 * blockIdx.x + 1] = y;
    }
}

} // namespace mmx

// ------------------------------------------------------------ host object

struct mmx_surrogate {
    mmx_sim *sim = nullptr;
    mmx::SurGeom s{};
    int H = 0, W = 0;
    long P = 0;
    mmx_demag_plan *cplan = nullptr;
    float *mc = nullptr, *bc = nullptr, *bfine = nullptr, *hd = nullptr;
    mmx::bf16_t *x0 = nullptr, *act[3] = {}, *gout = nullptr, *dz[3] = {};
    float *out = nullptr, *tgt = nullptr, *part = nullptr, *loss_part = nullptr;
    double *norm_part = nullptr;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    mmx::bf16_t *wt[4] = {}, *wbt[4] = {};
    int np = 0, off_w[4] = {}, off_b[4] = {};
    long adam_t = 0;
};

namespace mmx {
namespace {

template <typename T> int dalloc(T **p, size_t count)
{
    if (hipMalloc(p, sizeof(T) * count) != hipSuccess) {
        set_error("surrogate: cannot allocate %zu bytes", sizeof(T) * count);
        return MMX_ERR_NOMEM;
    }
    return MMX_OK;
}

void sur_free(mmx_surrogate *S)
{
    void *ptrs[] = {S->mc, S->bc, S->bfine, S->hd, S->x0, S->act[0], S->act[1], S->act[2],
                    S->gout, S->dz[0], S->dz[1], S->dz[2], S->out, S->tgt, S->part, S->loss_part,
                    S->norm_part, S->params, S->grads, S->adam_m, S->adam_v, S->wt[0], S->wt[1],
                    S->wt[2], S->wt[3], S->wbt[0], S->wbt[1], S->wbt[2], S->wbt[3]};
    for (void *p : ptrs) (void)hipFree(p);
    if (S->cplan) plan_destroy(S->cplan);
}

int prep_weights(mmx_surrogate *S, hipStream_t st)
{
    for (int l = 0; l < 4; l++) {
        const int n = 9 * kSurCh[l] * kSurCh[l + 1];
        sur_prep_weights<<<dim3(div_up(n, kBlock)), dim3(kBlock), 0, st>>>(
            S->params + S->off_w[l], kSurCh[l], kSurCh[l + 1], S->wt[l], S->wbt[l]);
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

// the net on a packed input (PACK = 0) or formed from (m, coarse B)
int net_forward(mmx_surrogate *S, bool pack, const float *m, hipStream_t st)
{
    const dim3 grid(S->W / 16, S->H / 8), blk(kBlock);
    const float *b[4] = {S->params + S->off_b[0], S->params + S->off_b[1], S->params + S->off_b[2],
                         S->params + S->off_b[3]};
    if (pack)
        sur_conv<32, 32, CE_RELU_BF16, 1><<<grid, blk, 0, st>>>(nullptr, S->wt[0], b[0], nullptr,
                                                                S->act[0], S->H, S->W, S->s, m, S->bc);
    else
        sur_conv<32, 32, CE_RELU_BF16, 0><<<grid, blk, 0, st>>>(S->x0, S->wt[0], b[0], nullptr,
                                                                S->act[0], S->H, S->W, S->s, nullptr, nullptr);
    sur_conv<32, 32, CE_RELU_BF16, 0><<<grid, blk, 0, st>>>(S->act[0], S->wt[1], b[1], nullptr,
                                                            S->act[1], S->H, S->W, S->s, nullptr, nullptr);
    sur_conv<32, 32, CE_RELU_BF16, 0><<<grid, blk, 0, st>>>(S->act[1], S->wt[2], b[2], nullptr,
                                                            S->act[2], S->H, S->W, S->s, nullptr, nullptr);
    sur_conv<32, 16, CE_LINEAR_F32, 0><<<grid, blk, 0, st>>>(S->act[2], S->wt[3], b[3], nullptr,
                                                             S->out, S->H, S->W, S->s, nullptr, nullptr);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

// loss (given the target: exact fine B + coarse B, or a direct target),
// backward, gradients into S->grads; *loss = mean squared error
int net_backward(mmx_surrogate *S, const float *tgt_direct, double *loss, hipStream_t st)
{
    const long n16 = S->P * 16;
    const int nlb = (int)div_up(n16, kBlock);
    float *t2 = S->loss_part + nlb; // sum target^2 (one float after the partials)
    sur_target_sq<<<dim3(nlb), dim3(kBlock), 0, st>>>(S->s, S->bfine, S->bc, tgt_direct, S->loss_part);
    sur_reduce_one<<<dim3(1), dim3(1024), 0, st>>>(S->loss_part, nlb, t2);
    sur_loss<<<dim3(nlb), dim3(kBlock), 0, st>>>(S->s, S->out, S->bfine, S->bc, tgt_direct, t2,
                                                 S->gout, S->loss_part);
    const dim3 grid(S->W / 16, S->H / 8), blk(kBlock);
    const dim3 wgrid(S->W / 16, S->H / 16);
    const int ntile = (S->W / 16) * (S->H / 16);
    const bf16_t *xin[4] = {S->x0, S->act[0], S->act[1], S->act[2]};
    // layer 3 (32 -> 16): weight gradient, then d act[2]
    sur_conv_wgrad<32, 16><<<wgrid, blk, 0, st>>>(xin[3], S->gout, S->part, S->H, S->W);
    int len = 9 * 32 * 16 + 16;
    sur_reduce_tiles<<<dim3(div_up(len, 64)), dim3(64 * kRedGroups), 0, st>>>(
        S->part, ntile, len, S->grads + S->off_w[3]);
    sur_conv<16, 32, CE_MASK_BF16, 0><<<grid, blk, 0, st>>>(S->gout, S->wbt[3], nullptr, S->act[2],
                                                            S->dz[2], S->H, S->W, S->s, nullptr, nullptr);
    for (int l = 2; l >= 0; l--) {
        sur_conv_wgrad<32, 32><<<wgrid, blk, 0, st>>>(xin[l], S->dz[l], S->part, S->H, S->W);
        len = 9 * 32 * 32 + 32;
        sur_reduce_tiles<<<dim3(div_up(len, 64)), dim3(64 * kRedGroups), 0, st>>>(
            S->part, ntile, len, S->grads + S->off_w[l]);
        if (l > 0)
            sur_conv<32, 32, CE_MASK_BF16, 0><<<grid, blk, 0, st>>>(
                S->dz[l], S->wbt[l], nullptr, S->act[l - 1], S->dz[l - 1], S->H, S->W, S->s, nullptr,
                nullptr);
    }
    MMX_CHECK_LAUNCH();
    if (loss) {
        std::vector<float> lp(nlb + 1);
        MMX_HIP(hipMemcpyAsync(lp.data(), S->loss_part, sizeof(float) * (nlb + 1),
                               hipMemcpyDeviceToHost, st));
        MMX_HIP(hipStreamSynchronize(st));
        double sum = 0.0;
        for (int b = 0; b < nlb; b++) sum += lp[b];
        *loss = lp[nlb] > 0 ? sum / (double)lp[nlb] : 0.0;
    }
    return MMX_OK;
}

int adam_step(mmx_surrogate *S, float lr, hipStream_t st)
{
    S->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - (float)std::pow(0.9, (double)S->adam_t);
    const float c2 = 1.0f - (float)std::pow(0.999, (double)S->adam_t);
    sur_adam<<<dim3(div_up(S->np, kBlock)), dim3(kBlock), 0, st>>>(S->params, S->grads, S->adam_m,
                                                                    S->adam_v, S->np, lr, b1, b2,
                                                                    eps, c1, c2);
    MMX_CHECK_LAUNCH();
    return prep_weights(S, st);
}

// coarse m (optionally from the RK predictor) and coarse B
int coarse_field(mmx_surrogate *S, const float *m, const float *tau0, float h, float *mpred,
                 hipStream_t st)
{
    const int nb = (int)div_up(S->s.Nc, kBlock);
    const float *map = S->sim->msat_map;
    if (tau0)
        sur_coarsen<1><<<dim3(nb), dim3(kBlock), 0, st>>>(S->s, m, tau0, h, map, mpred, S->mc);
    else
        sur_coarsen<0><<<dim3(nb), dim3(kBlock), 0, st>>>(S->s, m, nullptr, 0.0f, map, nullptr, S->mc);
    MMX_CHECK_LAUNCH();
    const long Nc = S->s.Nc;
    return mmx_demag_update_b(S->cplan, S->bc, S->bc + Nc, S->bc + 2 * Nc, S->mc, S->mc + Nc,
                              S->mc + 2 * Nc, nullptr, (float)S->sim->msat, st);
}

} // namespace

// Surrogate field evaluation at m (or at the predictor of m0 when tau0 is
// given): H_demag estimate into the epilogue of MODE.  Used by sim.cpp.
int surrogate_eval(mmx_surrogate *S, int mode, const float *m0, const float *tau0, float h,
                   float *mpred, const FieldParams &fp, const EpilogueIO &io, float *hd_out,
                   hipStream_t st)
{
    const float *m = tau0 ? mpred : m0;
    MMX_TRY(coarse_field(S, m0, tau0, h, mpred, st));
    MMX_TRY(net_forward(S, true, m, st));
    const DemagGeom &g = S->sim->plan->impl.g;
    const int nb = (int)div_up(S->s.N, kBlock);
    switch (mode) {
    case EPI_B: sur_epilogue<EPI_B><<<dim3(nb), dim3(kBlock), 0, st>>>(g, S->s, fp, io, S->out, S->bc, hd_out, 1.0f); break;
    case EPI_HEFF: sur_epilogue<EPI_HEFF><<<dim3(nb), dim3(kBlock), 0, st>>>(g, S->s, fp, io, S->out, S->bc, hd_out, 1.0f); break;
    case EPI_TAU0: sur_epilogue<EPI_TAU0><<<dim3(nb), dim3(kBlock), 0, st>>>(g, S->s, fp, io, S->out, S->bc, hd_out, 1.0f); break;
    default: sur_epilogue<EPI_CORR><<<dim3(nb), dim3(kBlock), 0, st>>>(g, S->s, fp, io, S->out, S->bc, hd_out, 1.0f); break;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

// ||a - b|| / ||c|| over n floats (L2; synchronises)
int surrogate_rel_l2(mmx_surrogate *S, const float *a, const float *b, const float *c, long n,
                     hipStream_t st, double *rel)
{
    const int nb = 512;
    sur_diff_norm<<<dim3(nb), dim3(kBlock), 0, st>>>(a, b, c, n, S->norm_part);
    MMX_CHECK_LAUNCH();
    std::vector<double> h(2 * nb);
    MMX_HIP(hipMemcpyAsync(h.data(), S->norm_part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, st));
    MMX_HIP(hipStreamSynchronize(st));
    double d2 = 0.0, b2 = 0.0;
    for (int i = 0; i < nb; i++) {
        d2 += h[2 * i];
        b2 += h[2 * i + 1];
    }
    *rel = b2 > 0 ? std::sqrt(d2 / b2) : 0.0;
    return MMX_OK;
}

float *surrogate_hd_buffer(mmx_surrogate *S) { return S->hd; }

} // namespace mmx

using namespace mmx;

extern "C" {

int mmx_surrogate_create(mmx_surrogate **sur, mmx_sim *sim, unsigned long long seed)
{
    MMX_REQUIRE(sur && sim, MMX_ERR_ARG, "mmx_surrogate_create: NULL argument");
    MMX_TRY(sim_prepare(sim));
    const int n0 = sim->n[0], n1 = sim->n[1], n2 = sim->n[2];
    MMX_REQUIRE(n0 >= 1 && n0 <= 5 && n1 % 16 == 0 && n2 % 16 == 0, MMX_ERR_SIZE,
                "surrogate: needs 1 <= n0 <= 5 and n1, n2 multiples of 16 (got %d %d %d)", n0, n1, n2);
    MMX_REQUIRE(!sim->pbc[0] && !sim->pbc[1] && !sim->pbc[2], MMX_ERR_ARG,
                "surrogate: periodic grids are not supported");
    mmx_surrogate *S = new mmx_surrogate();
    S->sim = sim;
    SurGeom &s = S->s;
    s.n0 = n0; s.n1 = n1; s.n2 = n2;
    s.f0 = 1; // the layers are kept
    s.c0 = n0 / s.f0; s.c1 = n1 / 2; s.c2 = n2 / 2;
    s.N = (long)n0 * n1 * n2;
    s.Nc = (long)s.c0 * s.c1 * s.c2;
    s.msat = (float)sim->msat;
    s.inv_mu0ms = (float)(1.0 / (kMu0 * sim->msat));
    S->H = n1;
    S->W = n2;
    S->P = (long)n1 * n2;
    int off = 0;
    for (int l = 0; l < 4; l++) {
        S->off_w[l] = off;
        off += 9 * kSurCh[l] * kSurCh[l + 1];
        S->off_b[l] = off;
        off += kSurCh[l + 1];
    }
    S->np = off;
    const double ccell[3] = {sim->cell[0] * s.f0, sim->cell[1] * 2, sim->cell[2] * 2};
    const int zero[3] = {0, 0, 0};
    int rc = plan_create(&S->cplan, s.c0, s.c1, s.c2, ccell, zero, sim->kernel_type, sim->accuracy);
    const long P = S->P;
    const int ntile = (n2 / 16) * (n1 / 16);
    if (rc == MMX_OK) rc = dalloc(&S->mc, 3 * s.Nc);
    if (rc == MMX_OK) rc = dalloc(&S->bc, 3 * s.Nc);
    if (rc == MMX_OK) rc = dalloc(&S->bfine, 3 * s.N);
    if (rc == MMX_OK) rc = dalloc(&S->hd, 3 * s.N);
    if (rc == MMX_OK) rc = dalloc(&S->x0, P * 32);
    for (int l = 0; l < 3 && rc == MMX_OK; l++) rc = dalloc(&S->act[l], P * 32);
    for (int l = 0; l < 3 && rc == MMX_OK; l++) rc = dalloc(&S->dz[l], P * 32);
    if (rc == MMX_OK) rc = dalloc(&S->gout, P * 16);
    if (rc == MMX_OK) rc = dalloc(&S->out, P * 16);
    if (rc == MMX_OK) rc = dalloc(&S->tgt, P * 3 * n0);
    if (rc == MMX_OK) rc = dalloc(&S->part, (size_t)ntile * (9 * 32 * 32 + 32));
    if (rc == MMX_OK) rc = dalloc(&S->loss_part, (size_t)div_up(P * 16, kBlock) + 1);
    if (rc == MMX_OK) rc = dalloc(&S->norm_part, 2 * 512);
    if (rc == MMX_OK) rc = dalloc(&S->params, S->np);
    if (rc == MMX_OK) rc = dalloc(&S->grads, S->np);
    if (rc == MMX_OK) rc = dalloc(&S->adam_m, S->np);
    if (rc == MMX_OK) rc = dalloc(&S->adam_v, S->np);
    for (int l = 0; l < 4 && rc == MMX_OK; l++) rc = dalloc(&S->wt[l], 9 * kSurCh[l] * kSurCh[l + 1]);
    for (int l = 0; l < 4 && rc == MMX_OK; l++) rc = dalloc(&S->wbt[l], 9 * kSurCh[l] * kSurCh[l + 1]);
    if (rc == MMX_OK) {
        // He initialisation (first layer over its 6 n0 + 2 used inputs), a
        // small last layer, zero biases
        std::vector<float> w(S->np, 0.0f);
        std::mt19937_64 gen(seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        for (int l = 0; l < 4; l++) {
            const int fan = 9 * (l == 0 ? 6 * n0 + 2 : kSurCh[l]);
            const double sd = std::sqrt(2.0 / fan) * (l == 3 ? 0.1 : 1.0);
            for (int e = 0; e < 9 * kSurCh[l] * kSurCh[l + 1]; e++) w[S->off_w[l] + e] = (float)(sd * nd(gen));
        }
        if (hipMemcpy(S->params, w.data(), sizeof(float) * S->np, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(S->adam_m, 0, sizeof(float) * S->np) != hipSuccess ||
            hipMemset(S->adam_v, 0, sizeof(float) * S->np) != hipSuccess ||
            hipMemset(S->grads, 0, sizeof(float) * S->np) != hipSuccess) {
            set_error("surrogate: initialisation copy failed");
            rc = MMX_ERR_HIP;
        }
    }
    if (rc == MMX_OK) rc = prep_weights(S, sim->st);
    if (rc == MMX_OK && hipStreamSynchronize(sim->st) != hipSuccess) rc = MMX_ERR_HIP;
    if (rc != MMX_OK) {
        sur_free(S);
        delete S;
        return rc;
    }
    *sur = S;
    return MMX_OK;
}

int mmx_surrogate_destroy(mmx_surrogate *sur)
{
    if (!sur) return MMX_OK;
    if (sur->sim && sur->sim->st) (void)hipStreamSynchronize(sur->sim->st);
    sur_free(sur);
    delete sur;
    return MMX_OK;
}

int mmx_surrogate_num_params(const mmx_surrogate *sur, int *count)
{
    MMX_REQUIRE(sur && count, MMX_ERR_ARG, "mmx_surrogate_num_params: NULL argument");
    *count = sur->np;
    return MMX_OK;
}

int mmx_surrogate_get_params(mmx_surrogate *sur, float *host, int count)
{
    MMX_REQUIRE(sur && host && count == sur->np, MMX_ERR_ARG, "mmx_surrogate_get_params: bad arguments");
    MMX_HIP(hipStreamSynchronize(sur->sim->st));
    MMX_HIP(hipMemcpy(host, sur->params, sizeof(float) * count, hipMemcpyDeviceToHost));
    return MMX_OK;
}

int mmx_surrogate_set_params(mmx_surrogate *sur, const float *host, int count)
{
    MMX_REQUIRE(sur && host && count == sur->np, MMX_ERR_ARG, "mmx_surrogate_set_params: bad arguments");
    MMX_HIP(hipStreamSynchronize(sur->sim->st));
    MMX_HIP(hipMemcpy(sur->params, host, sizeof(float) * count, hipMemcpyHostToDevice));
    sur->adam_t = 0;
    MMX_HIP(hipMemset(sur->adam_m, 0, sizeof(float) * sur->np));
    MMX_HIP(hipMemset(sur->adam_v, 0, sizeof(float) * sur->np));
    MMX_TRY(prep_weights(sur, sur->sim->st));
    MMX_HIP(hipStreamSynchronize(sur->sim->st));
    return MMX_OK;
}

int mmx_surrogate_get_grads(mmx_surrogate *sur, float *host, int count)
{
    MMX_REQUIRE(sur && host && count == sur->np, MMX_ERR_ARG, "mmx_surrogate_get_grads: bad arguments");
    MMX_HIP(hipStreamSynchronize(sur->sim->st));
    MMX_HIP(hipMemcpy(host, sur->grads, sizeof(float) * count, hipMemcpyDeviceToHost));
    return MMX_OK;
}

int mmx_surrogate_net(mmx_surrogate *sur, const float *x_host, const float *tgt_host, float *out_host,
                      double *loss)
{
    MMX_REQUIRE(sur && x_host && out_host, MMX_ERR_ARG, "mmx_surrogate_net: NULL argument");
    mmx_surrogate *S = sur;
    hipStream_t st = S->sim->st;
    const long P = S->P;
    // the input rounded to bf16 on the host side of the copy (same rounding)
    std::vector<unsigned short> xb(P * 32);
    for (long e = 0; e < P * 32; e++) {
        unsigned int u;
        std::memcpy(&u, &x_host[e], 4);
        if ((u & 0x7f800000u) == 0x7f800000u) xb[e] = (unsigned short)(u >> 16);
        else xb[e] = (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
    MMX_HIP(hipMemcpyAsync(S->x0, xb.data(), sizeof(unsigned short) * P * 32, hipMemcpyHostToDevice, st));
    MMX_TRY(net_forward(S, false, nullptr, st));
    if (tgt_host) {
        MMX_HIP(hipMemcpyAsync(S->tgt, tgt_host, sizeof(float) * P * 3 * S->s.n0, hipMemcpyHostToDevice, st));
        MMX_TRY(net_backward(S, S->tgt, loss, st));
    }
    MMX_HIP(hipMemcpyAsync(out_host, S->out, sizeof(float) * P * 16, hipMemcpyDeviceToHost, st));
    MMX_HIP(hipStreamSynchronize(st));
    return MMX_OK;
}

int mmx_surrogate_train_step(mmx_surrogate *sur, const float *m_dev, float lr, double *loss)
{
    MMX_REQUIRE(sur && m_dev && lr >= 0, MMX_ERR_ARG, "mmx_surrogate_train_step: bad arguments");
    mmx_surrogate *S = sur;
    mmx_sim *x = S->sim;
    MMX_REQUIRE(!x->msat_map, MMX_ERR_ARG, "surrogate: an Msat map is not supported");
    hipStream_t st = x->st;
    const long N = S->s.N;
    // target: the exact fine field; input: the coarse field
    MMX_TRY(mmx_demag_update_b(x->plan, S->bfine, S->bfine + N, S->bfine + 2 * N, m_dev, m_dev + N,
                               m_dev + 2 * N, nullptr, (float)x->msat, st));
    MMX_TRY(coarse_field(S, m_dev, nullptr, 0.0f, nullptr, st));
    const int np32 = (int)div_up(S->P * 32, kBlock);
    sur_pack<<<dim3(np32), dim3(kBlock), 0, st>>>(S->s, m_dev, S->bc, S->x0);
    MMX_CHECK_LAUNCH();
    MMX_TRY(net_forward(S, false, nullptr, st));
    MMX_TRY(net_backward(S, nullptr, loss, st));
    if (lr > 0) MMX_TRY(adam_step(S, lr, st));
    return MMX_OK;
}

int mmx_surrogate_predict_b(mmx_surrogate *sur, const float *m_dev, float *b_dev, int coarse_only)
{
    MMX_REQUIRE(sur && m_dev && b_dev, MMX_ERR_ARG, "mmx_surrogate_predict_b: NULL argument");
    mmx_surrogate *S = sur;
    hipStream_t st = S->sim->st;
    MMX_TRY(coarse_field(S, m_dev, nullptr, 0.0f, nullptr, st));
    if (coarse_only) {
        MMX_HIP(hipMemsetAsync(S->out, 0, sizeof(float) * S->P * 16, st));
    } else {
        MMX_TRY(net_forward(S, true, m_dev, st));
    }
    const DemagGeom &g = S->sim->plan->impl.g;
    const FieldParams fp = sim_field_params(S->sim);
    EpilogueIO io{};
    sur_epilogue<EPI_B><<<dim3((unsigned)div_up(S->s.N, kBlock)), dim3(kBlock), 0, st>>>(
        g, S->s, fp, io, S->out, S->bc, b_dev, (float)kMu0);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

} // extern "C"
