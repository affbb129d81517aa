// implicit.hip -- the fixed-point iterations of the implicit solvers
// (solver/bdf-euler, solver/am12) in one pass per iteration.
//
// One launch forms the next iterate and the maxima the iterators test:
//   MODE 0  yb = y0 + h dy                 (gpu.Madd, libmumax2.go:167-176)
//   MODE 1  yb = y0 + h (dy + dy0)          (gpu.AddMadd, add.cu:36-43; h = 0.5 h_step)
//   MODE 2  no new iterate: compares a = m with b = y0
// and, for each component c of the 3-component field, with b the new iterate
// (MODE 0, 1) or y0 (MODE 2):
//   out[3c + 0] = max |m - b|   (Reductor.MaxDiff, reduce.cu:595-695)
//   out[3c + 1] = max |b|       (Reductor.MaxAbs)
//   out[3c + 2] = max |m + b|   (Reductor.MaxSum, reduce.cu:698-790)
// Maxima are exact in any order; NaN counts as +inf (an iterator that meets
// one stops).  The 9 words are cleared by the caller and hold float bits.
#include <algorithm>
#include <cstring>

#include "mmx_internal.hpp"

namespace mmx {

namespace {

constexpr int kImpBlock = 256;

__device__ __forceinline__ float nan_to_inf(float v) { return (v != v) ? __builtin_inff() : v; }

template <int MODE>
__global__ void __launch_bounds__(kImpBlock)
implicit_kernel(const float *__restrict__ y0, const float *__restrict__ dy,
                const float *__restrict__ dy0, float h, const float *__restrict__ m,
                float *__restrict__ yb, long N, unsigned int *__restrict__ out)
{
    __shared__ float part[9][kImpBlock / 64];
    float mx[9];
#pragma unroll
    for (int k = 0; k < 9; k++) mx[k] = 0.0f;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const long I = c * N + i;
            float b;
            if (MODE == 0) {
                b = y0[I] + h * dy[I];
                yb[I] = b;
            } else if (MODE == 1) {
                b = y0[I] + h * (dy[I] + dy0[I]);
                yb[I] = b;
            } else {
                b = y0[I];
            }
            const float a = m[I];
            mx[3 * c + 0] = fmaxf(mx[3 * c + 0], nan_to_inf(fabsf(a - b)));
            mx[3 * c + 1] = fmaxf(mx[3 * c + 1], nan_to_inf(fabsf(b)));
            mx[3 * c + 2] = fmaxf(mx[3 * c + 2], nan_to_inf(fabsf(a + b)));
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        float v = mx[k];
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
        if (lane == 0) part[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        float v = part[threadIdx.x][0];
        for (int w = 1; w < kImpBlock / 64; w++) v = fmaxf(v, part[threadIdx.x][w]);
        // non-negative floats order like their bit patterns
        atomicMax(out + threadIdx.x, __float_as_uint(v));
    }
}

} // namespace

int implicit_pass(int mode, const float *y0, const float *dy, const float *dy0, float h,
                  const float *m, float *yb, long N, unsigned int *dout, unsigned int *hout,
                  hipStream_t st, double res[9])
{
    MMX_REQUIRE(mode >= 0 && mode <= 2 && N > 0 && y0 && m && dout && hout, MMX_ERR_ARG,
                "implicit_pass: bad arguments");
    MMX_REQUIRE(mode == 2 || (dy && yb), MMX_ERR_ARG, "implicit_pass: missing derivative");
    MMX_REQUIRE(mode != 1 || dy0, MMX_ERR_ARG, "implicit_pass: missing first derivative");
    MMX_HIP(hipMemsetAsync(dout, 0, 9 * sizeof(unsigned int), st));
    const long nb = std::min<long>(div_up(N, kImpBlock), 2048);
    switch (mode) {
    case 0:
        implicit_kernel<0><<<dim3((unsigned)nb), dim3(kImpBlock), 0, st>>>(y0, dy, nullptr, h, m,
                                                                           yb, N, dout);
        break;
    case 1:
        implicit_kernel<1><<<dim3((unsigned)nb), dim3(kImpBlock), 0, st>>>(y0, dy, dy0, h, m, yb,
                                                                           N, dout);
        break;
    default:
        implicit_kernel<2><<<dim3((unsigned)nb), dim3(kImpBlock), 0, st>>>(y0, nullptr, nullptr,
                                                                           0.0f, m, nullptr, N,
                                                                           dout);
        break;
    }
    MMX_CHECK_LAUNCH();
    MMX_HIP(hipMemcpyAsync(hout, dout, 9 * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    MMX_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 9; k++) {
        float f;
        memcpy(&f, hout + k, sizeof(float));
        res[k] = (double)f;
    }
    return MMX_OK;
}

} // namespace mmx
