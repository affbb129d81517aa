// kernel_build.hip -- one-time demag kernel construction in FP64 on the GPU.
//
// Compiled with -ffp-contract=off: the Kernel_Arne integration below performs
// the same IEEE double operations in the same order as the reference's Go code
// (modules/kernel_magnetostatic.go:30-181), so the float32 kernel it stores is
// the reference's kernel bit for bit (checked against the FP64 CPU oracle).
// The kernel is then transformed with an FP64 FFT (the reference uses a float32
// cuFFT here, maxwell_convolution.go:386-396), the real part of the octant
// kz <= p0/2, ky <= p1/2, kx <= p2/2 is kept and scaled by 1/(p0 p1 p2).
#include <cmath>
#include <cstring>
#include <vector>

#include "demag_plan.hpp"

namespace mmx {

// tensor component of (source s, destination d) in the stored order
// XX,YY,ZZ,YZ,XZ,XY; -1 for the lower triangle (common/index.go:44-47)
__device__ __constant__ int kStoredIdx[3][3] = {{0, 5, 4}, {-1, 1, 3}, {-1, -1, 2}};

struct ArneArgs {
    int p[3];
    double cell[3];
    int pbc[3];
    int x1, x2, y1, y2, z1, z2;
    double L;
    double accuracy;
    int s;
    int accumulate; // periodic images: several (x,y,z) map to one element
};

__device__ inline int wrap_idx(int number, int max)
{
    while (number < 0) number += max;
    while (number >= max) number -= max;
    return number;
}

__device__ inline double delta_cells(int d)
{
    if (d < 0) d = -d;
    if (d > 0) d -= 1;
    return (double)d;
}

__global__ void __launch_bounds__(kBlock) arne_kernel(ArneArgs a, float *__restrict__ K6)
{
    const int X = 0, Y = 1, Z = 2;
    const long nx = a.x2 - a.x1 + 1, ny = a.y2 - a.y1 + 1, nz = a.z2 - a.z1 + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny * nz) return;
    const int z = a.z1 + (int)(t % nz);
    const int y = a.y1 + (int)((t / nz) % ny);
    const int x = a.x1 + (int)(t / (nz * ny));
    const double *cell = a.cell;
    const int s = a.s;
    const int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
    double R[3], R2[3], pole[3];
    const int xw = wrap_idx(x, a.p[X]);
    const int yw = wrap_idx(y, a.p[Y]);
    const int zw = wrap_idx(z, a.p[Z]);
    R[X] = (double)x * cell[X];
    R[Y] = (double)y * cell[Y];
    R[Z] = (double)z * cell[Z];

    const double dx = delta_cells(x) * cell[X];
    const double dy = delta_cells(y) * cell[Y];
    const double dz = delta_cells(z) * cell[Z];
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d == 0) d = a.L;
    const double maxSize = d / a.accuracy;
    int nv = (int)(fmax(cell[v] / maxSize, 1.) + 0.5);
    int nw = (int)(fmax(cell[w] / maxSize, 1.) + 0.5);
    const int nx_ = (int)(fmax(cell[X] / maxSize, 1.) + 0.5);
    const int ny_ = (int)(fmax(cell[Y] / maxSize, 1.) + 0.5);
    const int nz_ = (int)(fmax(cell[Z] / maxSize, 1.) + 0.5);
    nv *= 2;
    nw *= 2;
    const double scale = 1 / (double)(nv * nw * nx_ * ny_ * nz_);
    const double surface = cell[v] * cell[w];
    const double charge = surface * scale;
    const double pu1 = cell[u] / 2.;
    const double pu2 = -pu1;

    double Bf[3] = {0, 0, 0};
    for (int i = 0; i < nv; i++) {
        const double pv = -(cell[v] / 2.) + cell[v] / (double)(2 * nv) + (double)i * (cell[v] / (double)nv);
        pole[v] = pv;
        for (int j = 0; j < nw; j++) {
            const double pw = -(cell[w] / 2.) + cell[w] / (double)(2 * nw) + (double)j * (cell[w] / (double)nw);
            pole[w] = pw;
            for (int al = 0; al < nx_; al++) {
                const double rx = R[X] - cell[X] / 2. + cell[X] / (double)(2 * nx_) + (cell[X] / (double)nx_) * (double)al;
                for (int be = 0; be < ny_; be++) {
                    const double ry = R[Y] - cell[Y] / 2. + cell[Y] / (double)(2 * ny_) + (cell[Y] / (double)ny_) * (double)be;
                    for (int ga = 0; ga < nz_; ga++) {
                        const double rz = R[Z] - cell[Z] / 2. + cell[Z] / (double)(2 * nz_) + (cell[Z] / (double)nz_) * (double)ga;
                        pole[u] = pu1;
                        R2[X] = rx - pole[X];
                        R2[Y] = ry - pole[Y];
                        R2[Z] = rz - pole[Z];
                        double r = sqrt(R2[X] * R2[X] + R2[Y] * R2[Y] + R2[Z] * R2[Z]);
                        double qr = charge / (4. * kPi * r * r * r);
                        const double bx = R2[X] * qr;
                        const double by = R2[Y] * qr;
                        const double bz = R2[Z] * qr;
                        pole[u] = pu2;
                        R2[X] = rx - pole[X];
                        R2[Y] = ry - pole[Y];
                        R2[Z] = rz - pole[Z];
                        r = sqrt(R2[X] * R2[X] + R2[Y] * R2[Y] + R2[Z] * R2[Z]);
                        qr = -charge / (4. * kPi * r * r * r);
                        Bf[X] += (bx + R2[X] * qr);
                        Bf[Y] += (by + R2[Y] * qr);
                        Bf[Z] += (bz + R2[Z] * qr);
                    }
                }
            }
        }
    }
    const long P = (long)a.p[0] * a.p[1] * a.p[2];
    const long off = ((long)xw * a.p[Y] + yw) * a.p[Z] + zw;
    for (int dd = 0; dd < 3; dd++) {
        const int idx = kStoredIdx[s][dd];
        if (idx < 0) continue;
        if (a.accumulate)
            atomicAdd(&K6[idx * P + off], (float)Bf[dd]);
        else
            K6[idx * P + off] = (float)Bf[dd];
    }
}

// max |x| of a float array into *dst (bits of a non-negative float)
__global__ void __launch_bounds__(kBlock) maxabs_kernel(const float *__restrict__ a, long n,
                                                        unsigned int *dst)
{
    float m = 0.0f;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(a[t]));
    block_max_to(m, dst);
}

__global__ void __launch_bounds__(kBlock) real_to_complex_f64(const float *__restrict__ a, long n,
                                                              double2 *__restrict__ z)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        z[t] = double2{(double)a[t], 0.0};
}

// In-place FP64 FFT of every line along one axis; the element t of a line is
// at base + t*inner.  One line per workgroup (one-time use, not hot).
__global__ void __launch_bounds__(kBlock) axis_fft_f64(double2 *__restrict__ data, long inner,
                                                       LineFFT<double> f)
{
    extern __shared__ double2 smd[];
    const int L = f.L;
    double2 *A = smd, *B = smd + L;
    const long line = blockIdx.x;
    const long outer = line / inner, in = line - outer * inner;
    double2 *base = data + outer * inner * L + in;
    for (int t = threadIdx.x; t < L; t += blockDim.x) A[t] = base[(long)t * inner];
    __syncthreads();
    const double2 *R = lds_fft<double, -1>(A, B, 1, L, 1, f);
    for (int t = threadIdx.x; t < L; t += blockDim.x) base[(long)t * inner] = R[t];
}

__global__ void __launch_bounds__(kBlock) zero_strided(float *__restrict__ out, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        out[6 * t] = 0.0f;
}

// Octant extraction with the 1/Np scale; realness statistics over it.
__global__ void __launch_bounds__(kBlock) extract_octant(const double2 *__restrict__ full, int p0,
                                                         int p1, int p2, int ko0, int ko1,
                                                         int h2, double inv,
                                                         float *__restrict__ out,
                                                         unsigned int *maxbad,
                                                         unsigned int *maxgood)
{
    const long n = (long)ko0 * ko1 * h2;
    float bad = 0.0f, good = 0.0f;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int kx = (int)(t % h2);
        const int ky = (int)((t / h2) % ko1);
        const int kz = (int)(t / ((long)h2 * ko1));
        const double2 v = full[((long)kz * p1 + ky) * p2 + kx];
        out[6 * t] = (float)(v.x * inv); // component slot of `out` set by the caller
        bad = fmaxf(bad, (float)fabs(v.y));
        good = fmaxf(good, (float)fabs(v.x));
    }
    block_max_to(bad, maxbad);
    __syncthreads();
    block_max_to(good, maxgood);
}

static int make_line_fft_f64(int L, LineFFT<double> &f, std::vector<double2> &host)
{
    f.L = L;
    if (!factor_radices(L, f.radix, &f.nstages)) {
        set_error("FFT length %d has a prime factor > 7", L);
        return MMX_ERR_SIZE;
    }
    host.resize(L);
    for (int t = 0; t < L; t++) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)L;
        host[t] = double2{(double)cosl(a), (double)(-sinl(a))};
    }
    return MMX_OK;
}

int build_demag_kernel(DemagPlanImpl *P, bool keep_real)
{
    const DemagGeom &g = P->g;
    const int p[3] = {g.p0, g.p1, g.p2};
    const long Pn = (long)p[0] * p[1] * p[2];
    MMX_REQUIRE(p[0] > 0 && p[1] > 1 && p[2] > 1, MMX_ERR_SIZE,
                "Kernel_Arne needs padded sizes p1, p2 > 1 (got %d %d %d)", p[0], p[1], p[2]);
    MMX_REQUIRE(p[1] % 2 == 0 && p[2] % 2 == 0 && (p[0] == 1 || p[0] % 2 == 0), MMX_ERR_SIZE,
                "Kernel_Arne needs even padded sizes (got %d %d %d)", p[0], p[1], p[2]);
    MMX_REQUIRE(p[0] <= 4096 && p[1] <= 4096 && p[2] <= 4096, MMX_ERR_SIZE,
                "kernel build: padded sizes above 4096 are not supported (%d %d %d)", p[0], p[1],
                p[2]);
    MMX_REQUIRE(P->cell_kernel_type == MMX_KERNEL_ARNE, MMX_ERR_ARG,
                "kernel type %d not available", P->cell_kernel_type);

    float *K6 = nullptr;
    double2 *full = nullptr;
    unsigned int *stats = nullptr;
    void *twd = nullptr;
    int rc = MMX_OK;
    auto fail = [&](int code) {
        rc = code;
        return code;
    };
    hipError_t e = hipMalloc(&K6, sizeof(float) * 6 * Pn);
    if (e != hipSuccess) {
        set_error("kernel build: cannot allocate %ld bytes", (long)(sizeof(float) * 6 * Pn));
        return MMX_ERR_NOMEM;
    }
    e = hipMalloc(&full, sizeof(double2) * Pn);
    if (e == hipSuccess) e = hipMalloc(&stats, sizeof(unsigned int) * 16);
    if (e != hipSuccess) {
        (void)hipFree(K6);
        (void)hipFree(full);
        set_error("kernel build: cannot allocate the FP64 work array");
        return MMX_ERR_NOMEM;
    }
    do {
        if (hipMemset(K6, 0, sizeof(float) * 6 * Pn) != hipSuccess) { fail(MMX_ERR_HIP); break; }
        if (hipMemset(stats, 0, sizeof(unsigned int) * 16) != hipSuccess) { fail(MMX_ERR_HIP); break; }

        // ranges of Kernel_Arne (kernel_magnetostatic.go:50-65)
        ArneArgs a{};
        for (int i = 0; i < 3; i++) {
            a.p[i] = p[i];
            a.cell[i] = P->cell[i];
            a.pbc[i] = P->pbc[i];
        }
        a.x1 = -(p[0] - 1) / 2; a.x2 = p[0] / 2 - 1;
        a.y1 = -(p[1] - 1) / 2; a.y2 = p[1] / 2 - 1;
        a.z1 = -(p[2] - 1) / 2; a.z2 = p[2] / 2 - 1;
        if (p[0] == 1 && P->pbc[0] == 0) a.x2 = 0;
        a.x1 *= (P->pbc[0] + 1); a.x2 *= (P->pbc[0] + 1);
        a.y1 *= (P->pbc[1] + 1); a.y2 *= (P->pbc[1] + 1);
        a.z1 *= (P->pbc[2] + 1); a.z2 *= (P->pbc[2] + 1);
        double L = P->cell[0];
        if (P->cell[1] < L) L = P->cell[1];
        if (P->cell[2] < L) L = P->cell[2];
        a.L = L;
        a.accuracy = (double)P->accuracy;
        a.accumulate = (P->pbc[0] || P->pbc[1] || P->pbc[2]) ? 1 : 0;
        const long npts = (long)(a.x2 - a.x1 + 1) * (a.y2 - a.y1 + 1) * (a.z2 - a.z1 + 1);
        for (int s = 0; s < 3; s++) {
            a.s = s;
            arne_kernel<<<dim3(div_up(npts, kBlock)), dim3(kBlock)>>>(a, K6);
            if (hipGetLastError() != hipSuccess) { set_error("arne_kernel launch failed"); fail(MMX_ERR_HIP); break; }
        }
        if (rc) break;

        // IsZero test per component (maxwell_convolution.go:412-418)
        const int gb = 1024;
        for (int c = 0; c < 6; c++)
            maxabs_kernel<<<dim3(gb), dim3(kBlock)>>>(K6 + c * Pn, Pn, stats + c);
        if (hipGetLastError() != hipSuccess) { set_error("maxabs launch failed"); fail(MMX_ERR_HIP); break; }
        unsigned int hs[16];
        if (hipMemcpy(hs, stats, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess) { fail(MMX_ERR_HIP); break; }
        float cmax[6], kmax = 0.0f;
        for (int c = 0; c < 6; c++) {
            std::memcpy(&cmax[c], &hs[c], 4);
            kmax = fmaxf(kmax, cmax[c]);
        }

        // FP64 twiddles for the three lengths
        LineFFT<double> f[3];
        std::vector<double2> tw[3];
        size_t off[4] = {0, 0, 0, 0};
        for (int ax = 0; ax < 3; ax++) {
            if (make_line_fft_f64(p[ax], f[ax], tw[ax])) { fail(MMX_ERR_SIZE); break; }
            off[ax + 1] = off[ax] + tw[ax].size();
        }
        if (rc) break;
        if (hipMalloc(&twd, sizeof(double2) * off[3]) != hipSuccess) { fail(MMX_ERR_NOMEM); break; }
        for (int ax = 0; ax < 3; ax++) {
            if (hipMemcpy((double2 *)twd + off[ax], tw[ax].data(), sizeof(double2) * tw[ax].size(),
                          hipMemcpyHostToDevice) != hipSuccess) { fail(MMX_ERR_HIP); break; }
            f[ax].tw = (const double2 *)twd + off[ax];
        }
        if (rc) break;
        for (int ax = 0; ax < 3; ax++) {
            const size_t b = 2 * sizeof(double2) * p[ax];
            if (b > 65536)
                (void)hipFuncSetAttribute(reinterpret_cast<const void *>(axis_fft_f64),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);
        }

        const long octant = (long)g.ko0 * g.ko1 * g.h2;
        const double inv = 1.0 / (double)Pn;
        for (int c = 0; c < 6; c++) {
            // interleaved storage: component c of point o at P->kern[6*o + c]
            // (6*octant floats); extract_octant / zero_strided write out[6*t]
            float *dst = P->kern + c;
            if (kmax == 0.0f || cmax[c] / kmax < 1e-5f) {
                zero_strided<<<dim3(1024), dim3(kBlock)>>>(dst, octant);
                if (hipGetLastError() != hipSuccess) { set_error("kernel zero launch failed"); fail(MMX_ERR_HIP); break; }
                continue;
            }
            real_to_complex_f64<<<dim3(4096), dim3(kBlock)>>>(K6 + c * Pn, Pn, full);
            long inner = 1;
            for (int ax = 2; ax >= 0; ax--) {
                if (p[ax] > 1) {
                    const size_t b = 2 * sizeof(double2) * p[ax];
                    axis_fft_f64<<<dim3((unsigned)(Pn / p[ax])), dim3(kBlock), b>>>(full, inner, f[ax]);
                }
                inner *= p[ax];
            }
            extract_octant<<<dim3(1024), dim3(kBlock)>>>(full, p[0], p[1], p[2], g.ko0, g.ko1, g.h2,
                                                         inv, dst, stats + 8, stats + 9);
            if (hipGetLastError() != hipSuccess) { set_error("kernel FFT launch failed"); fail(MMX_ERR_HIP); break; }
        }
        if (rc) break;
        if (hipMemcpy(hs, stats, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess) { fail(MMX_ERR_HIP); break; }
        float bad, good;
        std::memcpy(&bad, &hs[8], 4);
        std::memcpy(&good, &hs[9], 4);
        P->bad_over_good = good > 0 ? (double)bad / (double)good : 0.0;
        if (P->bad_over_good > 1e-5) {
            set_error("FFT kernel max bad/good part = %g", P->bad_over_good);
            fail(MMX_ERR_STATE);
            break;
        }
    } while (0);
    if (rc == MMX_OK && keep_real) {
        P->real_kernel = K6;
        K6 = nullptr;
    }
    (void)hipFree(K6);
    (void)hipFree(full);
    (void)hipFree(stats);
    (void)hipFree(twd);
    if (rc == MMX_OK && hipDeviceSynchronize() != hipSuccess) {
        set_error("kernel build: device error");
        return MMX_ERR_HIP;
    }
    return rc;
}

} // namespace mmx
