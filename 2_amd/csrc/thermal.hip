// thermal.hip -- the fluctuating field of temperature/brown
// (modules/mod_temperature.go:71-134, libmumax2/temperature.cu:11-70).
//
// H_therm_c(i) = n_c(i) sqrt(alpha_i T_i a / (b Ms_i)), 0 where the Msat map
// is 0, with a = float32(alpha 2 kB T) and b = float32(mu0 V gamma dt Msat)
// from the multipliers (the reference's alphaKB2tempMul and
// mu0VgammaDtMsatMul) and n standard normal.
//
// The normals are our own: the reference draws them from cuRAND's default
// generator; here they come from the counter-based Philox4x32-10 generator
// (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC 2011): one call per cell and draw, counter (cell lo, cell hi, draw lo,
// draw hi), key (seed lo, seed hi); its four 32-bit words become uniforms
// (x + 1/2) 2^-32 in (0, 1) and Box-Muller pairs in FP64,
// n0 = r cos(2 pi u1), n1 = r sin(2 pi u1) with r = sqrt(-2 ln u0), and
// n2 = r' cos(2 pi u3) with r' = sqrt(-2 ln u2).  The field is therefore a
// function of (seed, draw, cell) only: the same on any launch shape.
#include <algorithm>

#include "mmx_internal.hpp"

namespace mmx {

namespace {

constexpr int kThermBlock = 256;

struct U4 {
    unsigned int x, y, z, w;
};

__host__ __device__ inline unsigned int mulhilo(unsigned int a, unsigned int b, unsigned int *hi)
{
    const unsigned long long p = (unsigned long long)a * b;
    *hi = (unsigned int)(p >> 32);
    return (unsigned int)p;
}

// Philox4x32 with 10 rounds (multipliers and key increments of the paper)
__host__ __device__ inline U4 philox4x32_10(U4 c, unsigned int k0, unsigned int k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        unsigned int hi0, hi1;
        const unsigned int lo0 = mulhilo(0xD2511F53u, c.x, &hi0);
        const unsigned int lo1 = mulhilo(0xCD9E8D57u, c.z, &hi1);
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

__device__ inline double uniform01(unsigned int x) { return ((double)x + 0.5) * 0x1.0p-32; }

__global__ void __launch_bounds__(kThermBlock)
therm_kernel(unsigned int k0, unsigned int k1, unsigned int d0, unsigned int d1, long N,
             const float *__restrict__ alpha_map, const float *__restrict__ temp_map,
             const float *__restrict__ msat_map, float a, float b, float *__restrict__ out)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        const U4 r = philox4x32_10(U4{(unsigned int)i, (unsigned int)((unsigned long long)i >> 32),
                                      d0, d1},
                                   k0, k1);
        const double ra = sqrt(-2.0 * log(uniform01(r.x)));
        const double rb = sqrt(-2.0 * log(uniform01(r.z)));
        const double ta = 2.0 * M_PI * uniform01(r.y);
        const double tb = 2.0 * M_PI * uniform01(r.w);
        const float n[3] = {(float)(ra * cos(ta)), (float)(ra * sin(ta)), (float)(rb * cos(tb))};
        const float am = alpha_map ? alpha_map[i] : 1.0f;
        const float tm = temp_map ? temp_map[i] : 1.0f;
        const float mm = msat_map ? msat_map[i] : 1.0f;
        const float s = (mm != 0.0f) ? sqrtf((am * tm * a) / (b * mm)) : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) out[c * N + i] = (mm != 0.0f) ? n[c] * s : 0.0f;
    }
}

} // namespace

int therm_draw(unsigned long long seed, unsigned long long draw, long N, const float *alpha_map,
               const float *temp_map, const float *msat_map, float a, float b, float *out,
               hipStream_t st)
{
    MMX_REQUIRE(N > 0 && out, MMX_ERR_ARG, "therm_draw: bad arguments");
    const long nb = std::min<long>(div_up(N, kThermBlock), 4096);
    therm_kernel<<<dim3((unsigned)nb), dim3(kThermBlock), 0, st>>>(
        (unsigned int)seed, (unsigned int)(seed >> 32), (unsigned int)draw,
        (unsigned int)(draw >> 32), N, alpha_map, temp_map, msat_map, a, b, out);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

} // namespace mmx

extern "C" int mmx_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2],
                                 unsigned int out[4])
{
    MMX_REQUIRE(ctr && key && out, MMX_ERR_ARG, "mmx_philox4x32_10: NULL argument");
    const mmx::U4 r = mmx::philox4x32_10(mmx::U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
    return MMX_OK;
}
