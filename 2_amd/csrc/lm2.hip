// lm2.hip -- libmumax2's C interface on this path, argument for argument.
//
// The Go package mumax/gpu calls libmumax2 through cgo with every array as a
// float** (one device pointer per used GPU, gpu/array.go:122-131), a stream
// array (CUstream*, here hipStream_t*), sizes per GPU, and no return value:
// an error prints and aborts the process (libmumax2/gpu_safe.h:36-41).  The
// functions below keep exactly that calling convention, so each cgo wrapper in
// libmumax2.go changes only the symbol it calls.  One GPU is used
// (mmx_setUsedGPUs accepts one device), so every array is its element [0].
// A map argument is uniform (1) when the float** or its element is NULL.
//
// Most entry points forward to the mmx_* kernels of local.hip; the FFT-side
// helpers (kernelMulMicromag3D, cmadd, copyPad3D/copyUnPad3D, zeroArray) and
// the float32 block partials of the Reductor are defined here.
#include <cstdio>
#include <cstdlib>

#include "mmx_internal.hpp"
#include "../../include/mmx_libmumax2.h"

namespace mmx {

namespace {

[[noreturn]] void lm2_abort(const char *fn)
{
    std::fprintf(stderr, "mmx %s: %s\n", fn, mmx_last_error());
    std::fflush(stderr);
    std::abort();
}

inline void lm2_ok(int rc, const char *fn)
{
    if (rc != MMX_OK) lm2_abort(fn);
}

// device 0's part of a multi-GPU array (NULL for a NULL array)
inline float *part0(float **a) { return a ? a[0] : nullptr; }

inline hipStream_t stream0(hipStream_t *s) { return s ? s[0] : nullptr; }

inline void need(bool cond, const char *fn, const char *what)
{
    if (!cond) {
        set_error("%s", what);
        lm2_abort(fn);
    }
}

inline void launched(const char *fn)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch: %s", hipGetErrorString(e));
        lm2_abort(fn);
    }
}

inline unsigned blocks_for(long n) { return (unsigned)std::max<long>(1, div_up(n, kBlock)); }

// ------------------------------------------------------------- FFT side

// In-place product of the 3 transformed components (interleaved complex,
// n complex values each) with the symmetric real kernel (n values per
// component; NULL component = 0).
__global__ void __launch_bounds__(kBlock)
kmul3_kernel(float2 *__restrict__ fx, float2 *__restrict__ fy, float2 *__restrict__ fz,
             const float *__restrict__ kxx, const float *__restrict__ kyy,
             const float *__restrict__ kzz, const float *__restrict__ kyz,
             const float *__restrict__ kxz, const float *__restrict__ kxy, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 a = fx[i], b = fy[i], c = fz[i];
    const float xx = kxx ? kxx[i] : 0.0f, yy = kyy ? kyy[i] : 0.0f, zz = kzz ? kzz[i] : 0.0f;
    const float yz = kyz ? kyz[i] : 0.0f, xz = kxz ? kxz[i] : 0.0f, xy = kxy ? kxy[i] : 0.0f;
    fx[i] = float2{a.x * xx + b.x * xy + c.x * xz, a.y * xx + b.y * xy + c.y * xz};
    fy[i] = float2{a.x * xy + b.x * yy + c.x * yz, a.y * xy + b.y * yy + c.y * yz};
    fz[i] = float2{a.x * xz + b.x * yz + c.x * zz, a.y * xz + b.y * yz + c.y * zz};
}

// dst += (re + i im) * k * src, interleaved complex dst/src, real k
__global__ void __launch_bounds__(kBlock)
cmadd_kernel(float2 *__restrict__ dst, float re, float im, const float *__restrict__ k,
             const float2 *__restrict__ src, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 s = src[i];
    const float kk = k[i];
    float2 d = dst[i];
    d.x += kk * (re * s.x - im * s.y);
    d.y += kk * (im * s.x + re * s.y);
    dst[i] = d;
}

// Copies the (s0,s1,s2) block of src into the corner of the (d0,d1,d2) dst
// (pad) or the corner of src into all of dst (unpad): cells (i,j,k) below
// both extents.  One thread per cell of the copied block, grid-strided.
__global__ void __launch_bounds__(kBlock)
block_copy_kernel(float *__restrict__ dst, int d1, int d2, const float *__restrict__ src, int s1,
                  int s2, int b0, int b1, int b2)
{
    const long n = (long)b0 * b1 * b2;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % b2);
        const long r = t / b2;
        const int j = (int)(r % b1), i = (int)(r / b1);
        dst[((long)i * d1 + j) * d2 + k] = src[((long)i * s1 + j) * s2 + k];
    }
}

__global__ void __launch_bounds__(kBlock) fill_zero_kernel(float *__restrict__ a, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        a[t] = 0.0f;
}

// ------------------------------------------------------- block partials
//
// The Reductor (gpu/reductor.go:19-212) launches `blocks` partial reductions
// and finishes them on the host.  Here each of the `blocks` workgroups reduces
// a grid-strided share of the input (sums in FP64, rounded once to float;
// extrema exactly) and writes output[block].

enum PartOp : int { P_SUM, P_SDOT, P_MAX, P_MIN, P_MAXABS, P_MAXDIFF, P_MAXSUM, P_NORM3SQ, P_NORM3SQDIFF };

template <int OP>
__global__ void __launch_bounds__(kBlock)
partial_kernel(const float *__restrict__ a, const float *__restrict__ b,
               const float *__restrict__ c, const float *__restrict__ d,
               const float *__restrict__ e, const float *__restrict__ f, float *__restrict__ out,
               int n)
{
    constexpr bool SUM = OP == P_SUM || OP == P_SDOT;
    __shared__ double sh[kBlock / 64];
    double acc = SUM ? 0.0 : (OP == P_MIN ? (double)__builtin_inff() : -(double)__builtin_inff());
    if (OP == P_MAXABS || OP == P_MAXDIFF || OP == P_MAXSUM || OP == P_NORM3SQ || OP == P_NORM3SQDIFF)
        acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        if (OP == P_SUM) acc += (double)a[i];
        if (OP == P_SDOT) acc += (double)a[i] * (double)b[i];
        if (OP == P_MAX) acc = fmax(acc, (double)a[i]);
        if (OP == P_MIN) acc = fmin(acc, (double)a[i]);
        if (OP == P_MAXABS) acc = fmax(acc, (double)fabsf(a[i]));
        if (OP == P_MAXDIFF) acc = fmax(acc, (double)fabsf(a[i] - b[i]));
        if (OP == P_MAXSUM) acc = fmax(acc, (double)fabsf(a[i] + b[i]));
        if (OP == P_NORM3SQ) {
            const float x = a[i], y = b[i], z = c[i];
            acc = fmax(acc, (double)(x * x + y * y + z * z));
        }
        if (OP == P_NORM3SQDIFF) {
            const float x = a[i] - d[i], y = b[i] - e[i], z = c[i] - f[i];
            acc = fmax(acc, (double)(x * x + y * y + z * z));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double v = __shfl_xor(acc, o, 64);
        acc = SUM ? acc + v : (OP == P_MIN ? fmin(acc, v) : fmax(acc, v));
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int w = 1; w < kBlock / 64; w++)
            r = SUM ? r + sh[w] : (OP == P_MIN ? fmin(r, sh[w]) : fmax(r, sh[w]));
        out[blockIdx.x] = (float)r;
    }
}

template <int OP>
void run_partial(const char *fn, float **a, float **b, float **c, float **d, float **e, float **f,
                 float **out, int blocks, int n, hipStream_t *streams)
{
    need(a && a[0] && out && out[0], fn, "NULL input or output array");
    need(blocks > 0 && n >= 0, fn, "bad block count or size");
    if (OP == P_SDOT || OP == P_MAXDIFF || OP == P_MAXSUM || OP == P_NORM3SQ || OP == P_NORM3SQDIFF)
        need(b && b[0], fn, "NULL second input");
    if (OP == P_NORM3SQ || OP == P_NORM3SQDIFF) need(c && c[0], fn, "NULL third input");
    if (OP == P_NORM3SQDIFF) need(d && d[0] && e && e[0] && f && f[0], fn, "NULL input");
    partial_kernel<OP><<<dim3(blocks), dim3(kBlock), 0, stream0(streams)>>>(
        part0(a), part0(b), part0(c), part0(d), part0(e), part0(f), part0(out), n);
    launched(fn);
}

} // namespace

} // namespace mmx

using namespace mmx;

extern "C" {

void mmx_setUsedGPUs(int *gpuIds, int Ngpu)
{
    need(gpuIds && Ngpu == 1, "setUsedGPUs", "one GPU per process (Ngpu must be 1)");
    lm2_ok(mmx_init(gpuIds[0]), "setUsedGPUs");
}

void mmx_normalizeAsync(float **mx, float **my, float **mz, float **norm_map, hipStream_t *stream,
                        int Npart)
{
    need(mx && my && mz, "normalizeAsync", "NULL array");
    lm2_ok(mmx_normalize(part0(mx), part0(my), part0(mz), part0(norm_map), Npart, stream0(stream)),
           "normalizeAsync");
}

void mmx_torqueAsync(float **tx, float **ty, float **tz, float **mx, float **my, float **mz,
                     float **hx, float **hy, float **hz, float **alpha_map, float alpha_mul,
                     hipStream_t *stream, int Npart)
{
    need(tx && ty && tz && mx && my && mz && hx && hy && hz, "torqueAsync", "NULL array");
    lm2_ok(mmx_torque(part0(tx), part0(ty), part0(tz), part0(mx), part0(my), part0(mz), part0(hx),
                      part0(hy), part0(hz), part0(alpha_map), alpha_mul, Npart, stream0(stream)),
           "torqueAsync");
}

void mmx_exchange6Async(float **hx, float **hy, float **hz, float **mx, float **my, float **mz,
                        float **msat, float **aex, float Aex2_mu0MsatMul, int N0, int N1Part,
                        int N2, int periodic0, int periodic1, int periodic2, float cellSizeX,
                        float cellSizeY, float cellSizeZ, hipStream_t *streams)
{
    need(hx && hy && hz && mx && my && mz, "exchange6Async", "NULL array");
    lm2_ok(mmx_exchange6(part0(hx), part0(hy), part0(hz), part0(mx), part0(my), part0(mz),
                         part0(msat), part0(aex), Aex2_mu0MsatMul, N0, N1Part, N2, periodic0,
                         periodic1, periodic2, cellSizeX, cellSizeY, cellSizeZ, stream0(streams)),
           "exchange6Async");
}

void mmx_uniaxialAnisotropyAsync(float **hx, float **hy, float **hz, float **mx, float **my,
                                 float **mz, float **Ku1_map, float **MSat_map,
                                 float Ku2_Mu0Msat_mul, float **anisU_mapx, float anisU_mulx,
                                 float **anisU_mapy, float anisU_muly, float **anisU_mapz,
                                 float anisU_mulz, hipStream_t *stream, int Npart)
{
    need(hx && hy && hz && mx && my && mz, "uniaxialAnisotropyAsync", "NULL array");
    lm2_ok(mmx_uniaxial_anisotropy(part0(hx), part0(hy), part0(hz), part0(mx), part0(my), part0(mz),
                                   part0(Ku1_map), part0(MSat_map), Ku2_Mu0Msat_mul,
                                   part0(anisU_mapx), anisU_mulx, part0(anisU_mapy), anisU_muly,
                                   part0(anisU_mapz), anisU_mulz, Npart, stream0(stream)),
           "uniaxialAnisotropyAsync");
}

void mmx_zhangli_async(float **sttx, float **stty, float **sttz, float **mx, float **my,
                       float **mz, float **jx, float **jy, float **jz, float **msat, float **pol,
                       float **ee, float **alpha, const float jMul_x, const float jMul_y,
                       const float jMul_z, const float pred, const float eeMul,
                       const float alphaMul, const int sx, const int sy, const int sz,
                       const float csx, const float csy, const float csz, const int pbc_x,
                       const int pbc_y, const int pbc_z, hipStream_t *stream)
{
    need(sttx && stty && sttz && mx && my && mz, "zhangli_async", "NULL array");
    lm2_ok(mmx_zhangli(part0(sttx), part0(stty), part0(sttz), part0(mx), part0(my), part0(mz),
                       part0(jx), part0(jy), part0(jz), part0(msat), part0(pol), part0(ee),
                       part0(alpha), jMul_x, jMul_y, jMul_z, pred, eeMul, alphaMul, sx, sy, sz,
                       csx, csy, csz, pbc_x, pbc_y, pbc_z, stream0(stream)),
           "zhangli_async");
}

void mmx_maddAsync(float **dst, float **a, float **b, float mulB, hipStream_t *stream, int Npart)
{
    need(dst && a, "maddAsync", "NULL array");
    lm2_ok(mmx_madd(part0(dst), part0(a), part0(b), mulB, Npart, stream0(stream)), "maddAsync");
}

void mmx_madd1Async(float **a, float **b, float mulB, hipStream_t *stream, int Npart)
{
    need(a != nullptr, "madd1Async", "NULL array");
    lm2_ok(mmx_madd1(part0(a), part0(b), mulB, Npart, stream0(stream)), "madd1Async");
}

void mmx_madd2Async(float **a, float **b, float mulB, float **c, float mulC, hipStream_t *stream,
                    int Npart)
{
    need(a != nullptr, "madd2Async", "NULL array");
    lm2_ok(mmx_madd2(part0(a), part0(b), mulB, part0(c), mulC, Npart, stream0(stream)),
           "madd2Async");
}

void mmx_vecMaddAsync(float **dstx, float **dsty, float **dstz, float **ax, float **ay,
                      float **az, float **bx, float **by, float **bz, float mulBx, float mulBy,
                      float mulBz, hipStream_t *stream, int Npart)
{
    need(dstx && dsty && dstz && ax && ay && az, "vecMaddAsync", "NULL array");
    lm2_ok(mmx_vec_madd(part0(dstx), part0(dsty), part0(dstz), part0(ax), part0(ay), part0(az),
                        part0(bx), part0(by), part0(bz), mulBx, mulBy, mulBz, Npart,
                        stream0(stream)),
           "vecMaddAsync");
}

void mmx_mulAsync(float **dst, float **a, float **b, hipStream_t *stream, int Npart)
{
    need(dst && a && b, "mulAsync", "NULL array");
    lm2_ok(mmx_mul(part0(dst), part0(a), part0(b), Npart, stream0(stream)), "mulAsync");
}

void mmx_cmaddAsync(float **dst, float a, float b, float **kern, float **src, hipStream_t *stream,
                    int NComplexPart)
{
    need(dst && dst[0] && kern && kern[0] && src && src[0], "cmaddAsync", "NULL array");
    need(NComplexPart >= 0, "cmaddAsync", "negative size");
    if (NComplexPart == 0) return;
    cmadd_kernel<<<dim3(blocks_for(NComplexPart)), dim3(kBlock), 0, stream0(stream)>>>(
        reinterpret_cast<float2 *>(dst[0]), a, b, kern[0],
        reinterpret_cast<const float2 *>(src[0]), NComplexPart);
    launched("cmaddAsync");
}

void mmx_kernelMulMicromag3DAsync(float **fftMx, float **fftMy, float **fftMz, float **fftKxx,
                                  float **fftKyy, float **fftKzz, float **fftKyz, float **fftKxz,
                                  float **fftKxy, hipStream_t *stream, int partLen3D)
{
    need(fftMx && fftMx[0] && fftMy && fftMy[0] && fftMz && fftMz[0], "kernelMulMicromag3DAsync",
         "NULL array");
    need(partLen3D > 0 && partLen3D % 2 == 0, "kernelMulMicromag3DAsync",
         "partLen3D must be a positive even number of floats");
    const int n = partLen3D / 2;
    kmul3_kernel<<<dim3(blocks_for(n)), dim3(kBlock), 0, stream0(stream)>>>(
        reinterpret_cast<float2 *>(fftMx[0]), reinterpret_cast<float2 *>(fftMy[0]),
        reinterpret_cast<float2 *>(fftMz[0]), part0(fftKxx), part0(fftKyy), part0(fftKzz),
        part0(fftKyz), part0(fftKxz), part0(fftKxy), n);
    launched("kernelMulMicromag3DAsync");
}

void mmx_copyPad3DAsync(float **dst, int D0, int D1, int D2, float **src, int S0, int S1, int S2,
                        int Ncomp, hipStream_t *streams)
{
    need(dst && dst[0] && src && src[0], "copyPad3DAsync", "NULL array");
    need(S0 > 0 && S1 > 0 && S2 > 0 && S0 <= D0 && S1 <= D1 && S2 <= D2 && Ncomp > 0,
         "copyPad3DAsync", "source block must fit the destination");
    const long sblock = (long)S0 * S1 * S2, dblock = (long)D0 * D1 * D2;
    for (int c = 0; c < Ncomp; c++) {
        block_copy_kernel<<<dim3(std::min<long>(65535, blocks_for(sblock))), dim3(kBlock), 0,
                            stream0(streams)>>>(dst[0] + c * dblock, D1, D2, src[0] + c * sblock,
                                                S1, S2, S0, S1, S2);
        launched("copyPad3DAsync");
    }
}

void mmx_copyUnPad3DAsync(float **dst, int D0, int D1, int D2, float **src, int S0, int S1,
                          int S2, int Ncomp, hipStream_t *streams)
{
    need(dst && dst[0] && src && src[0], "copyUnPad3DAsync", "NULL array");
    need(D0 > 0 && D1 > 0 && D2 > 0 && D0 <= S0 && D1 <= S1 && D2 <= S2 && Ncomp > 0,
         "copyUnPad3DAsync", "destination block must fit the source");
    const long sblock = (long)S0 * S1 * S2, dblock = (long)D0 * D1 * D2;
    for (int c = 0; c < Ncomp; c++) {
        block_copy_kernel<<<dim3(std::min<long>(65535, blocks_for(dblock))), dim3(kBlock), 0,
                            stream0(streams)>>>(dst[0] + c * dblock, D1, D2, src[0] + c * sblock,
                                                S1, S2, D0, D1, D2);
        launched("copyUnPad3DAsync");
    }
}

void mmx_zeroArrayAsync(float **A, int N, hipStream_t *streams)
{
    need(A && A[0] && N >= 0, "zeroArrayAsync", "NULL array or negative size");
    if (N == 0) return;
    fill_zero_kernel<<<dim3(std::min<long>(65535, blocks_for(N))), dim3(kBlock), 0,
                       stream0(streams)>>>(A[0], N);
    launched("zeroArrayAsync");
}

void mmx_partialSumAsync(float **input, float **output, int blocksPerGPU,
                         int threadsPerBlockPerGPU, int NPerGPU, hipStream_t *streams)
{
    (void)threadsPerBlockPerGPU; // the workgroup size is this library's own
    run_partial<P_SUM>("partialSumAsync", input, nullptr, nullptr, nullptr, nullptr, nullptr,
                       output, blocksPerGPU, NPerGPU, streams);
}

void mmx_partialSDotAsync(float **input1, float **input2, float **output, int blocks,
                          int threadsPerBlock, int N, hipStream_t *stream)
{
    (void)threadsPerBlock;
    run_partial<P_SDOT>("partialSDotAsync", input1, input2, nullptr, nullptr, nullptr, nullptr,
                        output, blocks, N, stream);
}

void mmx_partialMaxAsync(float **input, float **output, int blocksPerGPU,
                         int threadsPerBlockPerGPU, int NPerGPU, hipStream_t *streams)
{
    (void)threadsPerBlockPerGPU;
    run_partial<P_MAX>("partialMaxAsync", input, nullptr, nullptr, nullptr, nullptr, nullptr,
                       output, blocksPerGPU, NPerGPU, streams);
}

void mmx_partialMinAsync(float **input, float **output, int blocksPerGPU,
                         int threadsPerBlockPerGPU, int NPerGPU, hipStream_t *streams)
{
    (void)threadsPerBlockPerGPU;
    run_partial<P_MIN>("partialMinAsync", input, nullptr, nullptr, nullptr, nullptr, nullptr,
                       output, blocksPerGPU, NPerGPU, streams);
}

void mmx_partialMaxAbsAsync(float **input, float **output, int blocksPerGPU,
                            int threadsPerBlockPerGPU, int NPerGPU, hipStream_t *streams)
{
    (void)threadsPerBlockPerGPU;
    run_partial<P_MAXABS>("partialMaxAbsAsync", input, nullptr, nullptr, nullptr, nullptr, nullptr,
                          output, blocksPerGPU, NPerGPU, streams);
}

void mmx_partialMaxDiffAsync(float **a, float **b, float **output, int blocksPerGPU,
                             int threadsPerBlockPerGPU, int NPerGPU, hipStream_t *streams)
{
    (void)threadsPerBlockPerGPU;
    run_partial<P_MAXDIFF>("partialMaxDiffAsync", a, b, nullptr, nullptr, nullptr, nullptr, output,
                           blocksPerGPU, NPerGPU, streams);
}

void mmx_partialMaxSumAsync(float **a, float **b, float **output, int blocksPerGPU,
                            int threadsPerBlockPerGPU, int NPerGPU, hipStream_t *streams)
{
    (void)threadsPerBlockPerGPU;
    run_partial<P_MAXSUM>("partialMaxSumAsync", a, b, nullptr, nullptr, nullptr, nullptr, output,
                          blocksPerGPU, NPerGPU, streams);
}

void mmx_partialMaxNorm3SqAsync(float **x, float **y, float **z, float **output, int blocksPerGPU,
                                int threadsPerBlockPerGPU, int NPerGPU, hipStream_t *streams)
{
    (void)threadsPerBlockPerGPU;
    run_partial<P_NORM3SQ>("partialMaxNorm3SqAsync", x, y, z, nullptr, nullptr, nullptr, output,
                           blocksPerGPU, NPerGPU, streams);
}

void mmx_partialMaxNorm3SqDiffAsync(float **x1, float **y1, float **z1, float **x2, float **y2,
                                    float **z2, float **output, int blocksPerGPU,
                                    int threadsPerBlockPerGPU, int NPerGPU, hipStream_t *streams)
{
    (void)threadsPerBlockPerGPU;
    run_partial<P_NORM3SQDIFF>("partialMaxNorm3SqDiffAsync", x1, y1, z1, x2, y2, z2, output,
                               blocksPerGPU, NPerGPU, streams);
}

} // extern "C"
