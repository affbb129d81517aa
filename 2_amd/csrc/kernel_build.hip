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
    int x1, x2, y1, y2, z1, z2;
    double L;
    double accuracy;
    int s;
};

__device__ inline double delta_cells(int d)
{
    if (d < 0) d = -d;
    if (d > 0) d -= 1;
    return (double)d;
}

// First displacement in [lo, hi] whose wrapped index is w (w in [0, p)), or
// hi + 1 when there is none; the others follow at +p.  kernel.go:19-28 wraps
// x into [0, p) by adding or subtracting p.
__device__ inline int first_image(int w, int p, int lo)
{
    int x = w;
    while (x - p >= lo) x -= p;
    return x;
}

// B (FP64) at displacement (x, y, z) of a unit s-dipole: the surface-charge /
// volume integration of kernel_magnetostatic.go:94-165, same operations in the
// same order.
__device__ inline void arne_point(const ArneArgs &a, int x, int y, int z, double Bf[3])
{
    const int X = 0, Y = 1, Z = 2;
    const double *cell = a.cell;
    const int s = a.s;
    const int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
    double R[3], R2[3], pole[3];
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

    Bf[0] = Bf[1] = Bf[2] = 0.0;
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
}

// One thread per padded-grid element (xw, yw, zw): the element is the float32
// sum, starting from 0, of float32(B) over every displacement (x, y, z) in the
// loop ranges that wraps onto it, added in the reference's loop order (x, then
// y, then z ascending; kernel_magnetostatic.go:91-101,168-172).  Without
// periodic images an element has at most one displacement; with them the
// element is the same float32 sum the reference forms, with no atomics, so the
// kernel is deterministic and bit-identical to it.  Every element is written.
__global__ void __launch_bounds__(kBlock) arne_kernel(ArneArgs a, float *__restrict__ K6)
{
    const int X = 0, Y = 1, Z = 2;
    const long P = (long)a.p[X] * a.p[Y] * a.p[Z];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    const int zw = (int)(t % a.p[Z]);
    const int yw = (int)((t / a.p[Z]) % a.p[Y]);
    const int xw = (int)(t / ((long)a.p[Z] * a.p[Y]));
    const int xs = first_image(xw, a.p[X], a.x1);
    const int ys = first_image(yw, a.p[Y], a.y1);
    const int zs = first_image(zw, a.p[Z], a.z1);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int x = xs; x <= a.x2; x += a.p[X])
        for (int y = ys; y <= a.y2; y += a.p[Y])
            for (int z = zs; z <= a.z2; z += a.p[Z]) {
                double Bf[3];
                arne_point(a, x, y, z, Bf);
                for (int dd = 0; dd < 3; dd++) acc[dd] += (float)Bf[dd];
            }
    for (int dd = 0; dd < 3; dd++) {
        const int idx = kStoredIdx[a.s][dd];
        if (idx >= 0) K6[idx * P + t] = acc[dd];
    }
}

// ------------------------------------------------------------ Newell, FP64
//
// No reference counterpart (the reference's dipole kernel is Kernel_Arne).
// Newell, Williams & Dunlop, J. Geophys. Res. 98 (1993) 9551: the cell-pair
// averaged tensor is a 27-point sixth-order difference of f (diagonal) and g
// (off-diagonal).  That difference cancels ~log10(64 |R|^6 / V^2) digits, so
// from NEWELL_RSWITCH largest cell sizes on the tensor is the point-dipole
// tensor averaged over the offset of two points drawn uniformly from the two
// cells (triangular density 1 - |t| per axis), with a 5-point Gauss rule of
// that density per axis (error below 1e-10 relative there).  K = -N.

constexpr double kNewellRSwitch = 8.0;
constexpr int kNewellNQ = 5;

struct NewellArgs {
    double d[3];
    double rs2;        // (kNewellRSwitch * largest cell size)^2
    double t[kNewellNQ], w[kNewellNQ];
    int e[3];          // octant extent per axis
};

__device__ inline double nw_f(double x, double y, double z)
{
    x = fabs(x); y = fabs(y); z = fabs(z);
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double R = sqrt(x2 + y2 + z2);
    double t = 0.0;
    if (y > 0 && x2 + z2 > 0) t += 0.5 * y * (z2 - x2) * asinh(y / sqrt(x2 + z2));
    if (z > 0 && x2 + y2 > 0) t += 0.5 * z * (y2 - x2) * asinh(z / sqrt(x2 + y2));
    if (x * y * z > 0) t -= x * y * z * atan(y * z / (x * R));
    t += (2.0 * x2 - y2 - z2) * R / 6.0;
    return t;
}

__device__ inline double nw_g(double x, double y, double z)
{
    const double sgn = ((x < 0) != (y < 0)) ? -1.0 : 1.0;
    x = fabs(x); y = fabs(y); z = fabs(z);
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double R = sqrt(x2 + y2 + z2);
    double t = 0.0;
    if (z > 0 && x2 + y2 > 0) t += x * y * z * asinh(z / sqrt(x2 + y2));
    if (x > 0 && y2 + z2 > 0) t += y / 6.0 * (3.0 * z2 - y2) * asinh(x / sqrt(y2 + z2));
    if (y > 0 && x2 + z2 > 0) t += x / 6.0 * (3.0 * z2 - x2) * asinh(y / sqrt(x2 + z2));
    if (z > 0) t -= z2 * z / 6.0 * atan(x * y / (z * R));
    if (y > 0) t -= z * y2 / 2.0 * atan(x * z / (y * R));
    if (x > 0) t -= z * x2 / 2.0 * atan(y * z / (x * R));
    t -= x * y * R / 3.0;
    return sgn * t;
}

// stored component c = axis pair (a, b), XX,YY,ZZ,YZ,XZ,XY
__device__ __constant__ int kNwA[6] = {0, 1, 2, 1, 0, 0};
__device__ __constant__ int kNwB[6] = {0, 1, 2, 2, 2, 1};

// One thread per octant displacement R = (i,j,k) * d, i < e0, j < e1, k < e2;
// writes the 6 components of K(R) to oct[6*t + c].
__global__ void __launch_bounds__(kBlock) newell_octant(NewellArgs a, double *__restrict__ oct)
{
    const long ne = (long)a.e[0] * a.e[1] * a.e[2];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne) return;
    const int k = (int)(t % a.e[2]);
    const int j = (int)((t / a.e[2]) % a.e[1]);
    const int i = (int)(t / ((long)a.e[2] * a.e[1]));
    const double R[3] = {i * a.d[0], j * a.d[1], k * a.d[2]};
    const double V = a.d[0] * a.d[1] * a.d[2];
    double out[6];
    if (R[0] * R[0] + R[1] * R[1] + R[2] * R[2] >= a.rs2) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int qi = 0; qi < kNewellNQ; qi++)
            for (int qj = 0; qj < kNewellNQ; qj++)
                for (int qk = 0; qk < kNewellNQ; qk++) {
                    const double x[3] = {R[0] + a.t[qi] * a.d[0], R[1] + a.t[qj] * a.d[1],
                                         R[2] + a.t[qk] * a.d[2]};
                    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
                    const double W = a.w[qi] * a.w[qj] * a.w[qk] / (r2 * r2 * sqrt(r2));
                    acc[0] += W * (3.0 * x[0] * x[0] - r2);
                    acc[1] += W * (3.0 * x[1] * x[1] - r2);
                    acc[2] += W * (3.0 * x[2] * x[2] - r2);
                    acc[3] += W * (3.0 * x[1] * x[2]);
                    acc[4] += W * (3.0 * x[0] * x[2]);
                    acc[5] += W * (3.0 * x[0] * x[1]);
                }
        for (int c = 0; c < 6; c++) out[c] = V / (4.0 * kPi) * acc[c];
    } else {
        const double w3[3] = {-1.0, 2.0, -1.0};
        for (int c = 0; c < 6; c++) {
            const int ax = kNwA[c], bx = kNwB[c];
            int b2 = bx, c2 = 3 - ax - bx;
            if (ax == bx) { b2 = (ax + 1) % 3; c2 = (ax + 2) % 3; }
            double s = 0.0;
            for (int u = 0; u < 3; u++)
                for (int v = 0; v < 3; v++)
                    for (int w = 0; w < 3; w++) {
                        const double xa = R[ax] + (u - 1) * a.d[ax];
                        const double xb = R[b2] + (v - 1) * a.d[b2];
                        const double xc = R[c2] + (w - 1) * a.d[c2];
                        const double fv = (ax == bx) ? nw_f(xa, xb, xc) : nw_g(xa, xb, xc);
                        s += w3[u] * w3[v] * w3[w] * fv;
                    }
            out[c] = -s / (4.0 * kPi * V);
        }
    }
    // K_ab is odd in R_a and in R_b: exactly 0 on those planes
    if (i == 0 || k == 0) out[4] = 0.0;          // XZ: axes 0, 2
    if (j == 0 || k == 0) out[3] = 0.0;          // YZ: axes 1, 2
    if (i == 0 || j == 0) out[5] = 0.0;          // XY: axes 0, 1
    for (int c = 0; c < 6; c++) oct[6 * t + c] = out[c];
}

// Component c of the padded real-space kernel from the octant: displacement
// of padded index q is q (q < p/2) or q - p (q > p/2); the p/2 planes are 0
// (the displacement range Kernel_Arne fills).  Off-diagonal components are
// odd in both of their axes.  Optionally also the float32 copy (real_f32).
__global__ void __launch_bounds__(kBlock) newell_expand(const double *__restrict__ oct, int e0, int e1,
                                                        int e2, int p0, int p1, int p2, int c,
                                                        double2 *__restrict__ full,
                                                        float *__restrict__ real_f32)
{
    const long P = (long)p0 * p1 * p2;
    const int ax = kNwA[c], bx = kNwB[c];
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < P; t += (long)gridDim.x * blockDim.x) {
        const int q[3] = {(int)(t / ((long)p1 * p2)), (int)((t / p2) % p1), (int)(t % p2)};
        const int p[3] = {p0, p1, p2};
        int disp[3];
        bool zero = false;
        for (int x = 0; x < 3; x++) {
            if (p[x] == 1) { disp[x] = 0; continue; }
            if (2 * q[x] == p[x]) zero = true;
            disp[x] = (2 * q[x] < p[x]) ? q[x] : q[x] - p[x];
        }
        double v = 0.0;
        if (!zero) {
            const long o = ((long)abs(disp[0]) * e1 + abs(disp[1])) * e2 + abs(disp[2]);
            v = oct[6 * o + c];
            if (ax != bx) {
                if (disp[ax] < 0) v = -v;
                if (disp[bx] < 0) v = -v;
            }
        }
        full[t] = double2{v, 0.0};
        if (real_f32) real_f32[t] = (float)v;
    }
}

// max |oct[6 t + c]| per component c into dst[c] (float bits)
__global__ void __launch_bounds__(kBlock) maxabs_oct6(const double *__restrict__ oct, long ne,
                                                      unsigned int *dst)
{
    float m[6] = {0, 0, 0, 0, 0, 0};
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ne; t += (long)gridDim.x * blockDim.x)
        for (int c = 0; c < 6; c++) m[c] = fmaxf(m[c], (float)fabs(oct[6 * t + c]));
    for (int c = 0; c < 6; c++) {
        block_max_to(m[c], dst + c);
        __syncthreads();
    }
}

// 5-point Gauss rule on [-1,1] for the density 1 - |t|: nodes 0, +-sqrt(u1),
// +-sqrt(u2) with u1, u2 the roots of the quadratic orthogonal to 1 and u
// under u dmu(u), mu the distribution of u = t^2 (moments E[t^2j] =
// 2/((2j+1)(2j+2))); weights from the moment equations.
static void newell_rule5(double t[5], double w[5])
{
    long double mu[5];
    for (int j = 0; j < 5; j++) mu[j] = 2.0L / ((2.0L * j + 1.0L) * (2.0L * j + 2.0L));
    // u^2 + c1 u + c0:  mu3 + c1 mu2 + c0 mu1 = 0,  mu4 + c1 mu3 + c0 mu2 = 0
    const long double det = mu[2] * mu[2] - mu[1] * mu[3];
    const long double c1 = (-mu[3] * mu[2] + mu[1] * mu[4]) / det;
    const long double c0 = (-mu[2] * mu[4] + mu[3] * mu[3]) / det;
    const long double disc = sqrtl(c1 * c1 - 4.0L * c0);
    const long double u1 = (-c1 + disc) / 2.0L, u2 = (-c1 - disc) / 2.0L;
    // 2 (W1 u1 + W2 u2) = mu1,  2 (W1 u1^2 + W2 u2^2) = mu2
    const long double A = (mu[2] / 2.0L - u2 * mu[1] / 2.0L) / (u1 - u2);
    const long double B = mu[1] / 2.0L - A;
    const long double W1 = A / u1, W2 = B / u2;
    t[0] = (double)-sqrtl(u1); w[0] = (double)W1;
    t[1] = (double)-sqrtl(u2); w[1] = (double)W2;
    t[2] = 0.0;                w[2] = (double)(1.0L - 2.0L * (W1 + W2));
    t[3] = (double)sqrtl(u2);  w[3] = (double)W2;
    t[4] = (double)sqrtl(u1);  w[4] = (double)W1;
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

// The octant storage takes K at kz > p0/2 or ky > p1/2 from its mirror
// point, sign-flipped for the components odd in that axis (XZ and XY in axis
// 0, YZ and XY in axis 1).  For a kernel given by the caller this maximum
// deviation of the full FP64 spectrum from that rule is checked.
__global__ void __launch_bounds__(kBlock) octant_mirror_dev(const double2 *__restrict__ full,
                                                            int p0, int p1, int p2, int h2,
                                                            int odd0, int odd1,
                                                            unsigned int *maxdev)
{
    const long n = (long)p0 * p1 * h2;
    float dev = 0.0f;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int kx = (int)(t % h2);
        const int ky = (int)((t / h2) % p1);
        const int kz = (int)(t / ((long)h2 * p1));
        const bool mz = kz > p0 / 2, my = ky > p1 / 2;
        if (!mz && !my) continue;
        const int kzp = mz ? p0 - kz : kz, kyp = my ? p1 - ky : ky;
        const double sign = ((mz && odd0) ? -1.0 : 1.0) * ((my && odd1) ? -1.0 : 1.0);
        const double a = full[((long)kz * p1 + ky) * p2 + kx].x;
        const double b = full[((long)kzp * p1 + kyp) * p2 + kx].x;
        dev = fmaxf(dev, (float)fabs(a - sign * b));
    }
    block_max_to(dev, maxdev);
}

// ------------------------------------------------------------ demagexch
//
// Exch6NgbrKernel (modules/kernel_exchange.go:102-124): the 6-neighbour
// Laplacian on the padded grid, the same for the three diagonal components,
// its seven float32 values assigned one after another at wrapped positions
// (kernel.go:19-27): a later value replaces an earlier one at the same
// position (on a length-1 axis the centre ends up 1/h_x^2).  dexKernUpdater
// (mod_demagexchange.go:134-149): dex = float32(Ms K_d + cex K_ex) with
// cex = 2 Aex / (mu0 Ms), from the float32 Arne kernel.

struct ExchPoints {
    long off[7];
    float val[7];
    int n;
};

static ExchPoints exch6_points(const int p[3], const double cell[3])
{
    const double h[3] = {cell[0] * cell[0], cell[1] * cell[1], cell[2] * cell[2]};
    const int d[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    const float v[7] = {(float)(-2 / h[0] - 2 / h[1] - 2 / h[2]), (float)(1 / h[0]), (float)(1 / h[0]),
                        (float)(1 / h[1]), (float)(1 / h[1]), (float)(1 / h[2]), (float)(1 / h[2])};
    ExchPoints e{};
    for (int t = 0; t < 7; t++) {
        long off = 0;
        for (int a = 0; a < 3; a++) off = off * p[a] + ((d[t][a] % p[a]) + p[a]) % p[a];
        int k = 0;
        while (k < e.n && e.off[k] != off) k++;
        if (k == e.n) e.n++;
        e.off[k] = off;
        e.val[k] = v[t];
    }
    return e;
}

__global__ void __launch_bounds__(kBlock) dex_combine(float *__restrict__ K6, long Pn, double ms,
                                                      double cex, ExchPoints e)
{
    const long n = 6 * Pn;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long c = i / Pn, o = i - c * Pn;
        double v = ms * (double)K6[i];
        if (c < 3)
            for (int k = 0; k < e.n; k++)
                if (e.off[k] == o) v += cex * (double)e.val[k];
        K6[i] = (float)v;
    }
}

__global__ void __launch_bounds__(kBlock) fold_kernel(const float *__restrict__ kd,
                                                      const float *__restrict__ kx, double inv_ms,
                                                      float *__restrict__ out, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)((double)kd[i] + (double)kx[i] * inv_ms);
}

static int make_line_fft_f64(int L, LineFFT<double> &f, std::vector<double2> &host)
{
    f.L = L;
    if (!factor_radices(L, f.radix, &f.nstages)) {
        set_error("FFT length %d has more than %d prime factors", L, kMaxStages);
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
    const bool newell = P->cell_kernel_type == MMX_KERNEL_NEWELL;
    MMX_REQUIRE(P->cell_kernel_type == MMX_KERNEL_ARNE || newell, MMX_ERR_ARG,
                "kernel type %d not available", P->cell_kernel_type);
    MMX_REQUIRE(p[0] > 0 && p[1] > 1 && p[2] > 1, MMX_ERR_SIZE,
                "demag kernel needs padded sizes p1, p2 > 1 (got %d %d %d)", p[0], p[1], p[2]);
    MMX_REQUIRE(p[1] % 2 == 0 && p[2] % 2 == 0 && (p[0] == 1 || p[0] % 2 == 0), MMX_ERR_SIZE,
                "demag kernel needs even padded sizes (got %d %d %d)", p[0], p[1], p[2]);
    MMX_REQUIRE(p[0] <= 4096 && p[1] <= 4096 && p[2] <= 4096, MMX_ERR_SIZE,
                "kernel build: padded sizes above 4096 are not supported (%d %d %d)", p[0], p[1],
                p[2]);
    MMX_REQUIRE(!newell || (P->pbc[0] == 0 && P->pbc[1] == 0 && P->pbc[2] == 0), MMX_ERR_ARG,
                "the Newell kernel supports non-periodic grids only");

    float *K6 = nullptr;     // Arne: float32 real-space kernel (6 x Pn); Newell: kept copy
    double *oct = nullptr;   // Newell: FP64 octant table (6 per displacement)
    double2 *full = nullptr;
    unsigned int *stats = nullptr;
    void *twd = nullptr;
    int rc = MMX_OK;
    auto fail = [&](int code) {
        rc = code;
        return code;
    };
    int e[3];
    for (int a = 0; a < 3; a++) e[a] = (p[a] == 1) ? 1 : p[a] / 2;
    const long ne = (long)e[0] * e[1] * e[2];
    hipError_t err = hipSuccess;
    if (!newell || keep_real) err = hipMalloc(&K6, sizeof(float) * 6 * Pn);
    if (err == hipSuccess && newell) err = hipMalloc(&oct, sizeof(double) * 6 * ne);
    if (err == hipSuccess) err = hipMalloc(&full, sizeof(double2) * Pn);
    if (err == hipSuccess) err = hipMalloc(&stats, sizeof(unsigned int) * 16);
    if (err != hipSuccess) {
        (void)hipFree(K6);
        (void)hipFree(oct);
        (void)hipFree(full);
        (void)hipFree(stats);
        set_error("kernel build: cannot allocate the FP64 work arrays (%ld points)", Pn);
        return MMX_ERR_NOMEM;
    }
    do {
        if (K6 && hipMemset(K6, 0, sizeof(float) * 6 * Pn) != hipSuccess) { fail(MMX_ERR_HIP); break; }
        if (hipMemset(stats, 0, sizeof(unsigned int) * 16) != hipSuccess) { fail(MMX_ERR_HIP); break; }

        const int gb = 1024;
        if (newell) {
            NewellArgs a{};
            double dm = 0.0;
            for (int x = 0; x < 3; x++) {
                a.d[x] = P->cell[x];
                a.e[x] = e[x];
                dm = fmax(dm, P->cell[x]);
            }
            a.rs2 = kNewellRSwitch * kNewellRSwitch * dm * dm;
            newell_rule5(a.t, a.w);
            newell_octant<<<dim3(div_up(ne, kBlock)), dim3(kBlock)>>>(a, oct);
            maxabs_oct6<<<dim3(gb), dim3(kBlock)>>>(oct, ne, stats);
            if (hipGetLastError() != hipSuccess) { set_error("newell_octant launch failed"); fail(MMX_ERR_HIP); break; }
        } else if (P->host_kernel) {
            // ConvPlan.loadKernel (convplan.go:85-115): the caller's kernel
            for (int c = 0; c < 6; c++) {
                if (!P->host_kernel[c]) continue;
                if (hipMemcpy(K6 + c * Pn, P->host_kernel[c], sizeof(float) * Pn,
                              hipMemcpyHostToDevice) != hipSuccess) { fail(MMX_ERR_HIP); break; }
            }
            if (rc) break;
            for (int c = 0; c < 6; c++)
                maxabs_kernel<<<dim3(gb), dim3(kBlock)>>>(K6 + c * Pn, Pn, stats + c);
            if (hipGetLastError() != hipSuccess) { set_error("maxabs launch failed"); fail(MMX_ERR_HIP); break; }
        } else {
            // ranges of Kernel_Arne (kernel_magnetostatic.go:50-65)
            ArneArgs a{};
            for (int i = 0; i < 3; i++) {
                a.p[i] = p[i];
                a.cell[i] = P->cell[i];
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
            for (int s = 0; s < 3; s++) {
                a.s = s;
                arne_kernel<<<dim3(div_up(Pn, kBlock)), dim3(kBlock)>>>(a, K6);
                if (hipGetLastError() != hipSuccess) { set_error("arne_kernel launch failed"); fail(MMX_ERR_HIP); break; }
            }
            if (rc) break;
            if (P->dex) {
                dex_combine<<<dim3(4096), dim3(kBlock)>>>(K6, Pn, P->dex_ms, P->dex_cex,
                                                          exch6_points(p, P->cell));
                if (hipGetLastError() != hipSuccess) { set_error("dex_combine launch failed"); fail(MMX_ERR_HIP); break; }
            }
            // IsZero test per component (maxwell_convolution.go:412-418)
            for (int c = 0; c < 6; c++)
                maxabs_kernel<<<dim3(gb), dim3(kBlock)>>>(K6 + c * Pn, Pn, stats + c);
            if (hipGetLastError() != hipSuccess) { set_error("maxabs launch failed"); fail(MMX_ERR_HIP); break; }
        }
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
            // (ConvPlan.loadKernel, used by demagexch, transforms every
            // component: convplan.go:85-115)
            const bool unused = P->host_kernel ? P->host_kernel[c] == nullptr
                                               : (!P->dex && (kmax == 0.0f || cmax[c] / kmax < 1e-5f));
            if (unused) {
                zero_strided<<<dim3(1024), dim3(kBlock)>>>(dst, octant);
                if (hipGetLastError() != hipSuccess) { set_error("kernel zero launch failed"); fail(MMX_ERR_HIP); break; }
                continue;
            }
            if (newell)
                newell_expand<<<dim3(4096), dim3(kBlock)>>>(oct, e[0], e[1], e[2], p[0], p[1], p[2], c,
                                                            full, K6 ? K6 + c * Pn : nullptr);
            else
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
            if (P->host_kernel) {
                // components XX,YY,ZZ,YZ,XZ,XY: odd in axis 0 for XZ, XY; in axis 1 for YZ, XY
                const int odd0 = (c == 4 || c == 5), odd1 = (c == 3 || c == 5);
                octant_mirror_dev<<<dim3(1024), dim3(kBlock)>>>(full, p[0], p[1], p[2], g.h2, odd0,
                                                                odd1, stats + 10);
            }
            if (hipGetLastError() != hipSuccess) { set_error("kernel FFT launch failed"); fail(MMX_ERR_HIP); break; }
        }
        if (rc) break;
        if (hipMemcpy(hs, stats, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess) { fail(MMX_ERR_HIP); break; }
        float bad, good;
        std::memcpy(&bad, &hs[8], 4);
        std::memcpy(&good, &hs[9], 4);
        P->bad_over_good = good > 0 ? (double)bad / (double)good : 0.0;
        if (P->host_kernel) {
            // ConvPlan only warns about imaginary parts (convplan.go:143-152);
            // the octant storage needs the mirror symmetry, so that is enforced
            float mdev;
            std::memcpy(&mdev, &hs[10], 4);
            if (good > 0 && (double)mdev / (double)good > 1e-5) {
                set_error("convolution kernel: transformed kernel deviates from the mirror "
                          "symmetry of a demag/exchange kernel by %g of its maximum",
                          (double)mdev / (double)good);
                fail(MMX_ERR_ARG);
                break;
            }
        } else if (P->bad_over_good > 1e-5) {
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
    (void)hipFree(oct);
    (void)hipFree(full);
    (void)hipFree(stats);
    (void)hipFree(twd);
    if (rc == MMX_OK && hipDeviceSynchronize() != hipSuccess) {
        set_error("kernel build: device error");
        return MMX_ERR_HIP;
    }
    return rc;
}

int build_dex_kernel(const DemagPlanImpl *P, int accuracy, double ms, double aex, float *out)
{
    MMX_REQUIRE(P && out && accuracy >= 0 && ms != 0.0, MMX_ERR_ARG, "build_dex_kernel: bad arguments");
    DemagPlanImpl Q = *P;   // geometry, cells, periodicity; nothing of P is written
    Q.kern = out;
    Q.real_kernel = nullptr;
    Q.accuracy = accuracy;
    Q.cell_kernel_type = MMX_KERNEL_ARNE;
    Q.dex = true;
    Q.dex_ms = ms;
    Q.dex_cex = (2.0 * aex) / (kMu0 * ms);
    return build_demag_kernel(&Q, false);
}

int fold_kernels(const float *kd, const float *kx, double ms, float *out, long n)
{
    MMX_REQUIRE(kd && kx && out && n > 0 && ms != 0.0, MMX_ERR_ARG, "fold_kernels: bad arguments");
    fold_kernel<<<dim3(1024), dim3(kBlock)>>>(kd, kx, 1.0 / ms, out, n);
    MMX_CHECK_LAUNCH();
    MMX_HIP(hipDeviceSynchronize());
    return MMX_OK;
}

} // namespace mmx
