/*
 * mmx_oracle.c -- FP64 CPU restatement of the mumax/2 hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see mmx_oracle.h).  Compiled with
 * -ffp-contract=off so that no multiply-add is fused behind our back: the
 * Kernel_Arne restatement is then bit-for-bit the Go code's arithmetic
 * (Go's amd64 backend does not fuse either), and the GPU's FP64 kernel
 * builder, compiled the same way, must match it exactly.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   orc_kernel_arne      closed forms: cube self-term -1/3 per diagonal, trace
 *                        -1, thin-film N_zz -> -1 (tests/dipole.py:29), symmetry
 *                        K_sd = K_ds, convergence of the cube average with
 *                        accuracy; uniform-cube field -Ms/3 (tests/demag3D.py:27-30).
 *   orc_fft_line         numpy.fft on random data (all radices), Parseval.
 *   orc_demag_kernel_fft realness check (maxwell_convolution.go:608-610),
 *                        against numpy.fft.rfftn of the same kernel.
 *   orc_demag_conv       orc_demag_brute (independent O(N^2) real-space sum)
 *                        on tiny grids; linearity; dipole.py / demag3D.py.
 *   orc_exchange6        uniform m -> 0; spin wave closed form; Neumann edge;
 *                        PBC wrap; harmonic-mean weights by hand on 2 cells.
 *   orc_uniaxial         closed form (m.u)u.
 *   orc_torque           tau . m = 0; tau = 0 for m || H; closed form on
 *                        hand-made vectors with the handedness flip.
 *   orc_normalize        |m| = 1, zero-mask -> 0.
 *   orc_zhangli          uniform m -> 0; linear texture -> analytic
 *                        5-point derivative; SP5 <m>(t) trace
 *                        (certification/stdproblem5/m.txt).
 *   orc_rk12_step        macrospin vs the analytic LLG solution; controller
 *                        rules on scalar sequences; SP4 trace
 *                        (tests/stdsolution4.txt).
 *   orc_average etc.     numpy on random data.
 */
#include "mmx_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------ sizes */

/* engine/zeropad.go:12-22: double every dimension that is >1 and not periodic. */
void orc_pad_size(const int n[3], const int pbc[3], int p[3])
{
    for (int i = 0; i < 3; i++)
        p[i] = (n[i] > 1 && pbc[i] == 0) ? 2 * n[i] : n[i];
}

/* modules/kernel.go:19-28 */
static int wrap_index(int number, int max)
{
    while (number < 0) number += max;
    while (number >= max) number -= max;
    return number;
}

/* modules/kernel_magnetostatic.go:185-193: nearest distance in cells. */
static double delta_cells(int d)
{
    if (d < 0) d = -d;
    if (d > 0) d -= 1;
    return (double)d;
}

/* Tensor component index of (source s, destination d), common/index.go:44-47. */
static const int full_tensor_idx[3][3] = {{0, 5, 4}, {8, 1, 3}, {7, 6, 2}};

/* ------------------------------------------------------------ Kernel_Arne */

/* modules/kernel_magnetostatic.go:30-181.  Same loop nesting, same
 * floating-point expression order, FP64 accumulation and a float32 "+=" store.
 * Parallelised over the x loop only: every destination element is written by
 * exactly one (s, x, y, z) iteration when pbc == 0, so the result does not
 * depend on the thread count.  With pbc != 0 the wrapped images are summed in
 * the reference's order by running that case serially. */
int orc_kernel_arne(const int p[3], const double cell[3], const int pbc[3],
                    int accuracy_, float *K9)
{
    const int X = 0, Y = 1, Z = 2;
    if (!(p[0] > 0 && p[1] > 1 && p[2] > 1)) return -1;
    if (!(cell[0] > 0 && cell[1] > 0 && cell[2] > 0)) return -1;
    if (accuracy_ < 0) return -1; /* 0: maxSize = d/0 = inf, one point per axis */
    if (p[1] % 2 || p[2] % 2) return -1;
    if (p[0] > 1 && p[0] % 2) return -1;

    const long P = (long)p[0] * p[1] * p[2];
    const double accuracy = (double)accuracy_;

    int x1 = -(p[X] - 1) / 2, x2 = p[X] / 2 - 1;
    int y1 = -(p[Y] - 1) / 2, y2 = p[Y] / 2 - 1;
    int z1 = -(p[Z] - 1) / 2, z2 = p[Z] / 2 - 1;
    if (p[X] == 1 && pbc[X] == 0) x2 = 0;
    x1 *= (pbc[X] + 1); x2 *= (pbc[X] + 1);
    y1 *= (pbc[Y] + 1); y2 *= (pbc[Y] + 1);
    z1 *= (pbc[Z] + 1); z2 *= (pbc[Z] + 1);

    double L = cell[X];
    if (cell[Y] < L) L = cell[Y];
    if (cell[Z] < L) L = cell[Z];

    const int serial = (pbc[0] || pbc[1] || pbc[2]);

    for (int s = 0; s < 3; s++) {
        const int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
#pragma omp parallel for schedule(dynamic, 1) if (!serial)
        for (int x = x1; x <= x2; x++) {
            double R[3], R2[3], pole[3];
            const int xw = wrap_index(x, p[X]);
            R[X] = (double)x * cell[X];
            for (int y = y1; y <= y2; y++) {
                const int yw = wrap_index(y, p[Y]);
                R[Y] = (double)y * cell[Y];
                for (int z = z1; z <= z2; z++) {
                    const int zw = wrap_index(z, p[Z]);
                    R[Z] = (double)z * cell[Z];

                    double dx = delta_cells(x) * cell[X];
                    double dy = delta_cells(y) * cell[Y];
                    double dz = delta_cells(z) * cell[Z];
                    double d = sqrt(dx * dx + dy * dy + dz * dz);
                    if (d == 0) d = L;
                    double maxSize = d / accuracy;
                    int nv = (int)(fmax(cell[v] / maxSize, 1.) + 0.5);
                    int nw = (int)(fmax(cell[w] / maxSize, 1.) + 0.5);
                    int nx = (int)(fmax(cell[X] / maxSize, 1.) + 0.5);
                    int ny = (int)(fmax(cell[Y] / maxSize, 1.) + 0.5);
                    int nz = (int)(fmax(cell[Z] / maxSize, 1.) + 0.5);
                    nv *= 2;
                    nw *= 2;

                    double scale = 1 / (double)(nv * nw * nx * ny * nz);
                    double surface = cell[v] * cell[w];
                    double charge = surface * scale;
                    double pu1 = cell[u] / 2.;
                    double pu2 = -pu1;

                    double B[3] = {0, 0, 0};
                    for (int i = 0; i < nv; i++) {
                        double pv = -(cell[v] / 2.) + cell[v] / (double)(2 * nv) +
                                    (double)i * (cell[v] / (double)nv);
                        pole[v] = pv;
                        for (int j = 0; j < nw; j++) {
                            double pw = -(cell[w] / 2.) + cell[w] / (double)(2 * nw) +
                                        (double)j * (cell[w] / (double)nw);
                            pole[w] = pw;
                            for (int a = 0; a < nx; a++) {
                                double rx = R[X] - cell[X] / 2. + cell[X] / (double)(2 * nx) +
                                            (cell[X] / (double)nx) * (double)a;
                                for (int b = 0; b < ny; b++) {
                                    double ry = R[Y] - cell[Y] / 2. + cell[Y] / (double)(2 * ny) +
                                                (cell[Y] / (double)ny) * (double)b;
                                    for (int c = 0; c < nz; c++) {
                                        double rz = R[Z] - cell[Z] / 2. + cell[Z] / (double)(2 * nz) +
                                                    (cell[Z] / (double)nz) * (double)c;
                                        pole[u] = pu1;
                                        R2[X] = rx - pole[X];
                                        R2[Y] = ry - pole[Y];
                                        R2[Z] = rz - pole[Z];
                                        double r = sqrt(R2[X] * R2[X] + R2[Y] * R2[Y] + R2[Z] * R2[Z]);
                                        double qr = charge / (4. * ORC_PI * r * r * r);
                                        double bx = R2[X] * qr;
                                        double by = R2[Y] * qr;
                                        double bz = R2[Z] * qr;

                                        pole[u] = pu2;
                                        R2[X] = rx - pole[X];
                                        R2[Y] = ry - pole[Y];
                                        R2[Z] = rz - pole[Z];
                                        r = sqrt(R2[X] * R2[X] + R2[Y] * R2[Y] + R2[Z] * R2[Z]);
                                        qr = -charge / (4. * ORC_PI * r * r * r);
                                        B[X] += (bx + R2[X] * qr);
                                        B[Y] += (by + R2[Y] * qr);
                                        B[Z] += (bz + R2[Z] * qr);
                                    }
                                }
                            }
                        }
                    }
                    const long off = ((long)xw * p[Y] + yw) * p[Z] + zw;
                    for (int dd = 0; dd < 3; dd++)
                        K9[(long)full_tensor_idx[s][dd] * P + off] += (float)B[dd];
                }
            }
        }
    }
    return 0;
}

/* -------------------------------------------------------------------- FFT */

/* Mixed-radix decimation-in-time DFT, written from the textbook recursion
 * X[u + m q1] = sum_q W_n^{q (u + m q1)} Y_q[u], Y_q = DFT_m(x[q::p]).
 * w[t] = exp(sign 2 pi i t / n0); the recursion level of length n reads
 * w[(t mod n) * (n0 / n)]. */
static void dft_rec(const cplx *in, long istride, cplx *out, int n,
                    const int *fac, const cplx *w, int wstep, cplx *tmp)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int pr = fac[0], m = n / pr;
    for (int q = 0; q < pr; q++)
        dft_rec(in + q * istride, istride * pr, out + (long)q * m, m, fac + 1,
                w, wstep * pr, tmp + pr);
    for (int u = 0; u < m; u++) {
        for (int q = 0; q < pr; q++) tmp[q] = out[(long)q * m + u];
        for (int q1 = 0; q1 < pr; q1++) {
            const long k = u + (long)m * q1;
            cplx acc = tmp[0];
            for (int q = 1; q < pr; q++)
                acc += tmp[q] * w[((q * k) % n) * wstep];
            out[k] = acc;
        }
    }
}

static int factorize(int n, int *fac)
{
    int nf = 0;
    const int pref[] = {4, 2, 3, 5, 7};
    for (int i = 0; i < 5; i++)
        while (n % pref[i] == 0 && n > 1) {
            fac[nf++] = pref[i];
            n /= pref[i];
        }
    for (int d = 11; n > 1; d += 2)
        while (n % d == 0) {
            fac[nf++] = d;
            n /= d;
        }
    return nf;
}

typedef struct {
    int n, sign;
    int fac[64];
    cplx *w;
} fft_plan;

static void fft_plan_init(fft_plan *pl, int n, int sign)
{
    pl->n = n;
    pl->sign = sign;
    factorize(n, pl->fac);
    pl->w = (cplx *)malloc(sizeof(cplx) * (n > 0 ? n : 1));
    for (int t = 0; t < n; t++) {
        double a = 2.0 * ORC_PI * (double)t / (double)n;
        pl->w[t] = cos(a) + (double)sign * sin(a) * _Complex_I;
    }
}

static void fft_plan_free(fft_plan *pl) { free(pl->w); }

/* out-of-place on a strided line, via scratch buffers */
static void fft_exec(const fft_plan *pl, cplx *line, cplx *scratch, cplx *tmp)
{
    if (pl->n <= 1) return;
    dft_rec(line, 1, scratch, pl->n, pl->fac, pl->w, 1, tmp);
    memcpy(line, scratch, sizeof(cplx) * pl->n);
}

int orc_fft_line(double *z, int n, int sign)
{
    if (n <= 0) return -1;
    fft_plan pl;
    fft_plan_init(&pl, n, sign);
    cplx *scratch = (cplx *)malloc(sizeof(cplx) * n);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4096);
    fft_exec(&pl, (cplx *)z, scratch, tmp);
    free(scratch);
    free(tmp);
    fft_plan_free(&pl);
    return 0;
}

/* Transform along one axis of a complex (a0,a1,a2) array; lines whose index
 * along the axis is >= nin are treated as zero on input, only the first nout
 * outputs are kept.  Works in place (nin, nout <= length). */
static void fft_axis(cplx *data, const int dims[3], int axis, int nin, int nout,
                     int sign)
{
    const int L = dims[axis];
    long stride = 1;
    for (int a = axis + 1; a < 3; a++) stride *= dims[a];
    long nlines = 1;
    for (int a = 0; a < 3; a++)
        if (a != axis) nlines *= dims[a];
    fft_plan pl;
    fft_plan_init(&pl, L, sign);
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * L);
        cplx *scr = (cplx *)malloc(sizeof(cplx) * L);
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4096);
#pragma omp for schedule(static)
        for (long ln = 0; ln < nlines; ln++) {
            /* decompose line index into the two other coordinates */
            long inner = 1;
            for (int a = axis + 1; a < 3; a++) inner *= dims[a];
            long outer_i = ln / inner, inner_i = ln % inner;
            cplx *base = data + outer_i * inner * L + inner_i;
            for (int t = 0; t < L; t++) line[t] = (t < nin) ? base[t * stride] : 0;
            fft_exec(&pl, line, scr, tmp);
            for (int t = 0; t < nout; t++) base[t * stride] = line[t];
        }
        free(line);
        free(scr);
        free(tmp);
    }
    fft_plan_free(&pl);
}

/* ------------------------------------------------- kernel FFT (LoadKernel) */

int orc_demag_kernel_fft(const int p[3], const float *K9, const double *K9d,
                         double *khat6, double *bad_over_good)
{
    const long P = (long)p[0] * p[1] * p[2];
    const int h2 = p[2] / 2 + 1;
    const long H = (long)p[0] * p[1] * h2;
    double kmax = 0;
    for (long t = 0; t < 9 * P; t++) {
        double a = K9d ? fabs(K9d[t]) : fabs((double)K9[t]);
        if (a > kmax) kmax = a;
    }
    cplx *buf = (cplx *)malloc(sizeof(cplx) * P);
    if (!buf) return -2;
    double maxbad = 0, maxgood = 0;
    const double inv = 1.0 / (double)P;
    const int dims[3] = {p[0], p[1], p[2]};
    for (int c = 0; c < 6; c++) {
        double cmax = 0;
        for (long t = 0; t < P; t++) {
            double a = K9d ? fabs(K9d[c * P + t]) : fabs((double)K9[c * P + t]);
            if (a > cmax) cmax = a;
        }
        double *dst = khat6 + c * H;
        if (kmax == 0 || cmax / kmax < 1e-5) {
            memset(dst, 0, sizeof(double) * H);
            continue;
        }
        for (long t = 0; t < P; t++)
            buf[t] = K9d ? K9d[c * P + t] : (double)K9[c * P + t];
        fft_axis(buf, dims, 2, p[2], p[2], -1);
        fft_axis(buf, dims, 1, p[1], p[1], -1);
        fft_axis(buf, dims, 0, p[0], p[0], -1);
        for (long ij = 0; ij < (long)p[0] * p[1]; ij++)
            for (int k = 0; k < h2; k++) {
                cplx v = buf[ij * p[2] + k];
                if (fabs(cimag(v)) > maxbad) maxbad = fabs(cimag(v));
                if (fabs(creal(v)) > maxgood) maxgood = fabs(creal(v));
                dst[ij * h2 + k] = creal(v) * inv;
            }
    }
    free(buf);
    if (bad_over_good) *bad_over_good = maxgood > 0 ? maxbad / maxgood : 0;
    return 0;
}

/* The same transform for a kernel whose components are generated on demand
 * (fill writes component c as interleaved complex (re, 0) for all p0*p1*p2
 * points); cmax6 holds max|K_c| for the IsZero test.  Keeps one complex
 * work array instead of the 9-component real-space kernel. */
int orc_demag_kernel_fft_cb(const int p[3], const double cmax6[6],
                            void (*fill)(void *ctx, int c, double *re_im), void *ctx,
                            double *khat6, double *bad_over_good)
{
    const long P = (long)p[0] * p[1] * p[2];
    const int h2 = p[2] / 2 + 1;
    const long H = (long)p[0] * p[1] * h2;
    double kmax = 0;
    for (int c = 0; c < 6; c++)
        if (cmax6[c] > kmax) kmax = cmax6[c];
    cplx *buf = (cplx *)malloc(sizeof(cplx) * P);
    if (!buf) return -2;
    double maxbad = 0, maxgood = 0;
    const double inv = 1.0 / (double)P;
    const int dims[3] = {p[0], p[1], p[2]};
    for (int c = 0; c < 6; c++) {
        double *dst = khat6 + c * H;
        if (kmax == 0 || cmax6[c] / kmax < 1e-5) {
            memset(dst, 0, sizeof(double) * H);
            continue;
        }
        fill(ctx, c, (double *)buf);
        fft_axis(buf, dims, 2, p[2], p[2], -1);
        fft_axis(buf, dims, 1, p[1], p[1], -1);
        fft_axis(buf, dims, 0, p[0], p[0], -1);
#pragma omp parallel for reduction(max : maxbad, maxgood)
        for (long ij = 0; ij < (long)p[0] * p[1]; ij++)
            for (int k = 0; k < h2; k++) {
                cplx v = buf[ij * p[2] + k];
                if (fabs(cimag(v)) > maxbad) maxbad = fabs(cimag(v));
                if (fabs(creal(v)) > maxgood) maxgood = fabs(creal(v));
                dst[ij * h2 + k] = creal(v) * inv;
            }
    }
    free(buf);
    if (bad_over_good) *bad_over_good = maxgood > 0 ? maxbad / maxgood : 0;
    return 0;
}

/* stored component for (s,d): XX,YY,ZZ,YZ,XZ,XY (upper triangle, the lower
 * one aliases it: maxwell_convolution.go:372-384) */
static int symm_idx(int s, int d)
{
    static const int t[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
    return t[s][d];
}

/* ------------------------------------------------ demag FFT convolution */

int orc_demag_conv(const int n[3], const int p[3], const double *khat6,
                   const double *m, const double *msat_map, double *out)
{
    const long N = (long)n[0] * n[1] * n[2];
    const int h2 = p[2] / 2 + 1;
    const long H = (long)p[0] * p[1] * h2;
    cplx *S[3];
    for (int c = 0; c < 3; c++) {
        S[c] = (cplx *)calloc(H, sizeof(cplx));
        if (!S[c]) return -2;
    }
    /* forward: x (R2C on the nonzero rows), then y, then z */
    for (int c = 0; c < 3; c++) {
        fft_plan pl;
        fft_plan_init(&pl, p[2], -1);
#pragma omp parallel
        {
            cplx *line = (cplx *)malloc(sizeof(cplx) * p[2]);
            cplx *scr = (cplx *)malloc(sizeof(cplx) * p[2]);
            cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4096);
#pragma omp for schedule(static)
            for (long ij = 0; ij < (long)n[0] * n[1]; ij++) {
                long i = ij / n[1], j = ij % n[1];
                for (int k = 0; k < p[2]; k++) {
                    if (k < n[2]) {
                        long ci = ij * n[2] + k;
                        double ms = msat_map ? msat_map[ci] : 1.0;
                        line[k] = m[c * N + ci] * ms;
                    } else
                        line[k] = 0;
                }
                fft_exec(&pl, line, scr, tmp);
                cplx *dst = S[c] + (i * p[1] + j) * h2;
                for (int k = 0; k < h2; k++) dst[k] = line[k];
            }
            free(line);
            free(scr);
            free(tmp);
        }
        fft_plan_free(&pl);
        const int dims[3] = {p[0], p[1], h2};
        fft_axis(S[c], dims, 1, n[1], p[1], -1);
        fft_axis(S[c], dims, 0, n[0], p[0], -1);
    }
    /* kernel multiply: F_d = sum_s K_sd S_s */
#pragma omp parallel for schedule(static)
    for (long t = 0; t < H; t++) {
        cplx s0 = S[0][t], s1 = S[1][t], s2 = S[2][t];
        cplx f[3];
        for (int d = 0; d < 3; d++)
            f[d] = khat6[symm_idx(0, d) * H + t] * s0 + khat6[symm_idx(1, d) * H + t] * s1 +
                   khat6[symm_idx(2, d) * H + t] * s2;
        S[0][t] = f[0];
        S[1][t] = f[1];
        S[2][t] = f[2];
    }
    /* inverse: z, y (keep the first n outputs), then C2R along x */
    for (int c = 0; c < 3; c++) {
        const int dims[3] = {p[0], p[1], h2};
        fft_axis(S[c], dims, 0, p[0], n[0], +1);
        fft_axis(S[c], dims, 1, p[1], n[1], +1);
        fft_plan pl;
        fft_plan_init(&pl, p[2], +1);
#pragma omp parallel
        {
            cplx *line = (cplx *)malloc(sizeof(cplx) * p[2]);
            cplx *scr = (cplx *)malloc(sizeof(cplx) * p[2]);
            cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4096);
#pragma omp for schedule(static)
            for (long ij = 0; ij < (long)n[0] * n[1]; ij++) {
                long i = ij / n[1], j = ij % n[1];
                const cplx *src = S[c] + (i * p[1] + j) * h2;
                for (int k = 0; k < p[2]; k++) {
                    if (k < h2)
                        line[k] = src[k];
                    else
                        line[k] = conj(src[p[2] - k]);
                }
                fft_exec(&pl, line, scr, tmp);
                for (int k = 0; k < n[2]; k++) out[c * N + ij * n[2] + k] = creal(line[k]);
            }
            free(line);
            free(scr);
            free(tmp);
        }
        fft_plan_free(&pl);
    }
    for (int c = 0; c < 3; c++) free(S[c]);
    return 0;
}

int orc_demag_brute(const int n[3], const int p[3], const float *K9,
                    const double *m, const double *msat_map, double *out)
{
    const long N = (long)n[0] * n[1] * n[2];
    const long P = (long)p[0] * p[1] * p[2];
    double kmax = 0;
    for (long t = 0; t < 9 * P; t++)
        if (fabs((double)K9[t]) > kmax) kmax = fabs((double)K9[t]);
    int zero[6];
    for (int c = 0; c < 6; c++) {
        double cmax = 0;
        for (long t = 0; t < P; t++)
            if (fabs((double)K9[c * P + t]) > cmax) cmax = fabs((double)K9[c * P + t]);
        zero[c] = (kmax == 0 || cmax / kmax < 1e-5);
    }
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < N; ci++) {
        int i = (int)(ci / ((long)n[1] * n[2])), j = (int)((ci / n[2]) % n[1]), k = (int)(ci % n[2]);
        double acc[3] = {0, 0, 0};
        for (long cj = 0; cj < N; cj++) {
            int i2 = (int)(cj / ((long)n[1] * n[2])), j2 = (int)((cj / n[2]) % n[1]), k2 = (int)(cj % n[2]);
            int di = wrap_index(i - i2, p[0]), dj = wrap_index(j - j2, p[1]), dk = wrap_index(k - k2, p[2]);
            long off = ((long)di * p[1] + dj) * p[2] + dk;
            double ms = msat_map ? msat_map[cj] : 1.0;
            for (int d = 0; d < 3; d++)
                for (int s = 0; s < 3; s++) {
                    int c = symm_idx(s, d);
                    if (zero[c]) continue;
                    acc[d] += (double)K9[c * P + off] * m[s * N + cj] * ms;
                }
        }
        for (int d = 0; d < 3; d++) out[d * N + ci] = acc[d];
    }
    return 0;
}

/* ------------------------------------------------------ local field terms */

static double mask_unity(const double *msk, long i) { return msk ? msk[i] : 1.0; }
static double div_zero(double a, double b) { return (b == 0.0) ? 0.0 : a / b; }
static double avg_geom_zero(double a, double b)
{
    double s = a + b;
    return (s == 0.0) ? 0.0 : 2.0 * a * b / s;
}

void orc_exchange6(const int n[3], const int pbc[3], const double cell[3],
                   double pre, const double *msat_map, const double *aex_map,
                   const double *m, double *h)
{
    const int N0 = n[0], N1 = n[1], N2 = n[2];
    const long N = (long)N0 * N1 * N2;
    const double c0 = 1.0 / (cell[0] * cell[0]);
    const double c1 = 1.0 / (cell[1] * cell[1]);
    const double c2 = 1.0 / (cell[2] * cell[2]);
    const double *mx = m, *my = m + N, *mz = m + 2 * N;
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < N; ci++) {
        const int i = (int)(ci / ((long)N1 * N2)), j = (int)((ci / N2) % N1), k = (int)(ci % N2);
        const double lex0 = div_zero(mask_unity(aex_map, ci), mask_unity(msat_map, ci));
        double H[3] = {0, 0, 0};
        /* order of the reference: dim 0, dim 2, dim 1 (exchange6.cu:48-133) */
        const int order[3] = {0, 2, 1};
        for (int o = 0; o < 3; o++) {
            const int dim = order[o];
            const int pos = (dim == 0) ? i : (dim == 1) ? j : k;
            const int len = (dim == 0) ? N0 : (dim == 1) ? N1 : N2;
            const double cc = (dim == 0) ? c0 : (dim == 1) ? c1 : c2;
            int lo = pos - 1;
            lo = (lo < 0 && pbc[dim]) ? len + lo : lo;
            if (lo < 0) lo = 0;
            int hi = pos + 1;
            hi = (hi == len && pbc[dim]) ? hi - len : hi;
            if (hi > len - 1) hi = len - 1;
            long A, B;
            if (dim == 0) {
                A = ((long)lo * N1 + j) * N2 + k;
                B = ((long)hi * N1 + j) * N2 + k;
            } else if (dim == 1) {
                A = ((long)i * N1 + lo) * N2 + k;
                B = ((long)i * N1 + hi) * N2 + k;
            } else {
                A = ((long)i * N1 + j) * N2 + lo;
                B = ((long)i * N1 + j) * N2 + hi;
            }
            const double l1 = avg_geom_zero(lex0, div_zero(mask_unity(aex_map, A), mask_unity(msat_map, A)));
            const double l2 = avg_geom_zero(lex0, div_zero(mask_unity(aex_map, B), mask_unity(msat_map, B)));
            H[0] += pre * cc * (l1 * (mx[A] - mx[ci]) + l2 * (mx[B] - mx[ci]));
            H[1] += pre * cc * (l1 * (my[A] - my[ci]) + l2 * (my[B] - my[ci]));
            H[2] += pre * cc * (l1 * (mz[A] - mz[ci]) + l2 * (mz[B] - mz[ci]));
        }
        h[ci] = H[0];
        h[N + ci] = H[1];
        h[2 * N + ci] = H[2];
    }
}

void orc_uniaxial(long N, double ku2, const double *ku_map,
                  const double *msat_map, const double u[3],
                  const double *u_map, const double *m, double *h)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) {
        double ms = 1.0;
        if (msat_map) {
            ms = msat_map[i];
            if (ms == 0.0) ms = 1.0; /* uniaxialanisotropy.cu:30-40 */
        }
        double k = ku_map ? (ku2 / ms) * ku_map[i] : ku2 / ms;
        double uc[3];
        for (int c = 0; c < 3; c++) /* mul * map, uniaxialanisotropy.cu:51-79 */
            uc[c] = u_map ? u[c] * u_map[c * N + i] : u[c];
        double mu = m[i] * uc[0] + m[N + i] * uc[1] + m[2 * N + i] * uc[2];
        h[i] = k * mu * uc[0];
        h[N + i] = k * mu * uc[1];
        h[2 * N + i] = k * mu * uc[2];
    }
}

void orc_torque(long N, const double *m, const double *h,
                const double *alpha_map, double alpha_mul, double *tau)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) {
        double alpha = alpha_map ? alpha_mul * alpha_map[i] : alpha_mul;
        double Mx = m[i], My = m[N + i], Mz = m[2 * N + i];
        double Hx = h[i], Hy = h[N + i], Hz = h[2 * N + i];
        double ax = My * Hz - Hy * Mz;
        double ay = -Mx * Hz + Hx * Mz;
        double az = Mx * Hy - Hx * My;
        double bx = -My * az + ay * Mz;
        double by = +Mx * az - ax * Mz;
        double bz = -Mx * ay + ax * My;
        double gilb = 1.0 / (1.0 + alpha * alpha);
        tau[i] = gilb * (ax + bx * alpha);
        tau[N + i] = gilb * (ay + by * alpha);
        tau[2 * N + i] = gilb * (az + bz * alpha);
    }
}

void orc_normalize(long N, double *m, const double *norm_map)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) {
        double norm = norm_map ? norm_map[i] : 1.0;
        double x = m[i], y = m[N + i], z = m[2 * N + i];
        double len = sqrt(x * x + y * y + z * z);
        double scale = (norm == 0.0) ? 0.0 : 1.0;
        scale = (len != 0.0) ? scale / len : 0.0;
        m[i] = x * scale;
        m[N + i] = y * scale;
        m[2 * N + i] = z * scale;
    }
}

/* negated ("LHR") cross product, libmumax2/common_func.h:196-201 */
static void crossf(const double a[3], const double b[3], double r[3])
{
    r[0] = -a[1] * b[2] + a[2] * b[1];
    r[1] = -a[2] * b[0] + a[0] * b[2];
    r[2] = -a[0] * b[1] + a[1] * b[0];
}

void orc_zhangli(const int n[3], const int pbc[3], const double cell[3],
                 const double *m, const double *jmask, const double jmul[3],
                 const double *msat_map, const double *pol_map,
                 const double *xi_map, const double *alpha_map, double pre,
                 double xi_mul, double alpha_mul, double *stt)
{
    const int S0 = n[0], S1 = n[1], S2 = n[2];
    const long N = (long)S0 * S1 * S2;
    const double ms0 = 1.0 / (12.0 * cell[0]);
    const double ms1 = 1.0 / (12.0 * cell[1]);
    const double ms2 = 1.0 / (12.0 * cell[2]);
#pragma omp parallel for schedule(static)
    for (long x0 = 0; x0 < N; x0++) {
        const int i = (int)(x0 / ((long)S1 * S2)), j = (int)((x0 / S2) % S1), k = (int)(x0 % S2);
        double m_sat = msat_map ? msat_map[x0] : 1.0;
        double j0[3];
        for (int c = 0; c < 3; c++) j0[c] = jmask ? jmask[c * N + x0] * jmul[c] : jmul[c];
        double njn = sqrt(j0[0] * j0[0] + j0[1] * j0[1] + j0[2] * j0[2]);
        if (m_sat == 0.0 || njn == 0.0) {
            stt[x0] = stt[N + x0] = stt[2 * N + x0] = 0.0;
            continue;
        }
        m_sat = 1.0 / m_sat;
        double PP = pol_map ? pol_map[x0] : 1.0;
        double pred = njn * PP * pre;
        double ee = xi_map ? xi_map[x0] * xi_mul : xi_mul;
        double alpha = alpha_map ? alpha_map[x0] * alpha_mul : alpha_mul;
        pred = pred / ((1.0 + ee * ee) * (1.0 + alpha * alpha));
        double pret = pred * (ee - alpha);
        pred = pred * (1.0 + ee * alpha);

        /* 5-point first derivative; out-of-range neighbours -> centre cell
         * unless periodic (zhang-li_torque.cu:184-249) */
        int nb[3][4];
        const int pos[3] = {i, j, k}, len[3] = {S0, S1, S2};
        for (int d = 0; d < 3; d++) {
            int c = pos[d];
            int b2 = c - 2, b1 = c - 1, f1 = c + 1, f2 = c + 2;
            if (!pbc[d]) {
                if (b2 < 0) b2 = c;
                if (b1 < 0) b1 = c;
                if (f1 >= len[d]) f1 = c;
                if (f2 >= len[d]) f2 = c;
            }
            /* periodic wrap: the reference adds/subtracts len once, which
             * equals a modulo for len >= 2 */
            b2 = ((b2 % len[d]) + len[d]) % len[d];
            b1 = ((b1 % len[d]) + len[d]) % len[d];
            f1 = f1 % len[d];
            f2 = f2 % len[d];
            nb[d][0] = b2; nb[d][1] = b1; nb[d][2] = f1; nb[d][3] = f2;
        }
        const double step[3] = {ms0, ms1, ms2};
        double dm[3][3]; /* dm[comp][dim] */
        for (int c = 0; c < 3; c++) {
            const double *mc = m + c * N;
            for (int d = 0; d < 3; d++) {
                long idx[4];
                for (int t = 0; t < 4; t++) {
                    int q = nb[d][t];
                    if (d == 0) idx[t] = ((long)q * S1 + j) * S2 + k;
                    else if (d == 1) idx[t] = ((long)i * S1 + q) * S2 + k;
                    else idx[t] = ((long)i * S1 + j) * S2 + q;
                }
                dm[c][d] = step[d] * (mc[idx[0]] - 8.0 * mc[idx[1]] + 8.0 * mc[idx[2]] - mc[idx[3]]);
            }
        }
        double nj0[3];
        double len_j = sqrt(j0[0] * j0[0] + j0[1] * j0[1] + j0[2] * j0[2]);
        double inv = (len_j != 0.0) ? 1.0 / len_j : 0.0;
        for (int c = 0; c < 3; c++) nj0[c] = j0[c] * inv;
        double dmdj[3];
        for (int c = 0; c < 3; c++)
            dmdj[c] = dm[c][0] * nj0[0] + dm[c][1] * nj0[1] + dm[c][2] * nj0[2];
        double mv[3] = {m[x0], m[N + x0], m[2 * N + x0]};
        double dmdjxm[3], mxdmxm[3];
        crossf(dmdj, mv, dmdjxm);
        crossf(mv, dmdjxm, mxdmxm);
        for (int c = 0; c < 3; c++)
            stt[c * N + x0] = m_sat * ((pred * mxdmxm[c]) + (pret * dmdjxm[c]));
    }
}

/* ---------------------------------------------------------- composite path */

int orc_heff(const orc_problem *P, const double *m, double *h,
             double *h_demag, double *h_ex, double *h_anis)
{
    const long N = (long)P->n[0] * P->n[1] * P->n[2];
    double *tmp = (double *)malloc(sizeof(double) * 3 * N);
    if (!tmp) return -2;
    for (long t = 0; t < 3 * N; t++) h[t] = 0;
    if (P->demag) {
        int rc = orc_demag_conv(P->n, P->p, P->khat6, m, P->msat_map, tmp);
        if (rc) { free(tmp); return rc; }
        /* B = mu0 Ms conv, added to H_eff with weight 1/mu0 */
        for (long t = 0; t < 3 * N; t++) {
            double b = ORC_MU0 * P->msat * tmp[t];
            h[t] += b / ORC_MU0;
            if (h_demag) h_demag[t] = b / ORC_MU0;
        }
    }
    {
        double pre = 2.0 * P->aex / (ORC_MU0 * P->msat);
        orc_exchange6(P->n, P->pbc, P->cell, pre, P->msat_map, P->aex_map, m, tmp);
        for (long t = 0; t < 3 * N; t++) h[t] += tmp[t];
        if (h_ex) memcpy(h_ex, tmp, sizeof(double) * 3 * N);
    }
    if (P->anis) {
        double ku2 = 2.0 * P->ku / (ORC_MU0 * P->msat);
        orc_uniaxial(N, ku2, P->ku_map, P->msat_map, P->anis_u, P->u_map, m, tmp);
        for (long t = 0; t < 3 * N; t++) h[t] += tmp[t];
        if (h_anis) memcpy(h_anis, tmp, sizeof(double) * 3 * N);
    }
    for (int c = 0; c < 3; c++) { /* B_ext mask: vecMadd, mul * map (add.cu:66-101) */
        double add = P->bext[c] / ORC_MU0;
        for (long t = 0; t < N; t++)
            h[c * N + t] += P->bext_map ? add * P->bext_map[c * N + t] : add;
    }
    free(tmp);
    return 0;
}

int orc_torque_full(const orc_problem *P, const double *m, const double *h,
                    double *tau)
{
    const long N = (long)P->n[0] * P->n[1] * P->n[2];
    orc_torque(N, m, h, P->alpha_map, P->alpha, tau);
    if (P->zhangli) {
        double *stt = (double *)malloc(sizeof(double) * 3 * N);
        if (!stt) return -2;
        /* pred = P muB / (e Ms), mod_zhang-li.go:62-68 */
        double pre = P->pol * ORC_MUB / (ORC_QE * P->msat);
        orc_zhangli(P->n, P->pbc, P->cell, m, P->jmask, P->jmul, P->msat_map,
                    P->pol_map, P->xi_map, P->alpha_map, pre, P->xi, P->alpha, stt);
        for (long t = 0; t < 3 * N; t++) tau[t] += stt[t] / P->gamma;
        free(stt);
    }
    return 0;
}

/* -------------------------------------------------------------- reductions */

void orc_average(long N, long ncell, const double *a, double *avg)
{
    for (int c = 0; c < 3; c++) {
        double s = 0, comp = 0;
        for (long i = 0; i < N; i++) { /* Kahan */
            double y = a[c * N + i] - comp;
            double t = s + y;
            comp = (t - s) - y;
            s = t;
        }
        avg[c] = s / (double)ncell;
    }
}

double orc_maxnorm(long N, const double *a)
{
    double mx = 0;
    for (long i = 0; i < N; i++) {
        double v = a[i] * a[i] + a[N + i] * a[N + i] + a[2 * N + i] * a[2 * N + i];
        if (v > mx) mx = v;
    }
    return sqrt(mx);
}

double orc_maxdiff(long n, const double *a, const double *b)
{
    double mx = 0;
    for (long i = 0; i < n; i++) {
        double v = fabs(a[i] - b[i]);
        if (v > mx) mx = v;
    }
    return mx;
}

/* ------------------------------------------------------------------ RK12 */

void orc_rk12_init(orc_rk12 *S)
{
    memset(S, 0, sizeof(*S));
    /* LoadRK12 defaults (solve_rk12.go:36-42) then LoadMicromag overrides
     * (mod_micromag.go:31-33) */
    S->mindt = 1e-15;
    S->maxdt = 1e38;
    S->dt = 1e-15;
    S->maxerr = 1.0 / 2000.0;
}

static int eval_torque(const orc_problem *P, orc_rk12 *S, double *m, double *h,
                       double *tau)
{
    const long N = (long)P->n[0] * P->n[1] * P->n[2];
    if (!S->m_normalized) {
        orc_normalize(N, m, P->msat_map);
        S->m_normalized = 1;
    }
    int rc = orc_heff(P, m, h, NULL, NULL, NULL);
    if (rc) return rc;
    S->evals++;
    return orc_torque_full(P, m, h, tau);
}

/* engine/solve_rk12.go:91-214 */
int orc_rk12_step(const orc_problem *P, orc_rk12 *S, double *m)
{
    const long N = (long)P->n[0] * P->n[1] * P->n[2];
    const long n3 = 3 * N;
    double *y0 = (double *)malloc(sizeof(double) * n3);
    double *dy0 = (double *)malloc(sizeof(double) * n3);
    double *tau = (double *)malloc(sizeof(double) * n3);
    double *h = (double *)malloc(sizeof(double) * n3);
    if (!y0 || !dy0 || !tau || !h) return -2;
    int rc = eval_torque(P, S, m, h, tau);
    if (rc) goto done;
    const double gamma = P->gamma;
    const double t0 = S->t;
    memcpy(dy0, tau, sizeof(double) * n3);
    memcpy(y0, m, sizeof(double) * n3);
    const int maxTry = 10;
    const double headRoom = 0.8;
    int try_ = 0;
    for (;;) {
        double dt = S->dt;
        if (try_ > 0) {
            memcpy(m, y0, sizeof(double) * n3);
            memcpy(tau, dy0, sizeof(double) * n3);
        }
        double hh = dt * gamma;
        for (long t = 0; t < n3; t++) m[t] = m[t] + hh * tau[t];
        S->m_normalized = 0;
        S->t = t0 + dt;
        rc = eval_torque(P, S, m, h, tau);
        if (rc) goto done;
        int badStep = 0;
        double minFactor = 2.0;
        for (long t = 0; t < n3; t++) m[t] += 0.5 * hh * tau[t] + (-0.5 * hh) * dy0[t];
        double err = orc_maxdiff(n3, tau, dy0) * hh;
        S->err = err;
        if (err > S->maxerr) {
            S->badsteps++;
            badStep = 1;
        }
        if ((!badStep || try_ == maxTry - 1) && err > S->peakerr) S->peakerr = err;
        double factor = badStep ? pow(S->maxerr / err, 1. / 3.) : sqrt(S->maxerr / err);
        factor *= headRoom;
        if (factor > 1.5) factor = 1.5;
        if (factor < 0.1) factor = 0.1;
        if (factor < minFactor) minFactor = factor;
        S->m_normalized = 0;
        double newDt = dt * minFactor;
        if (newDt < S->mindt) newDt = S->mindt;
        if (newDt > S->maxdt) newDt = S->maxdt;
        S->dt = newDt;
        if (!badStep || newDt == S->mindt) break;
        if (try_ > maxTry) {
            rc = -1;
            goto done;
        }
        try_++;
    }
    S->step++;
done:
    free(y0);
    free(dy0);
    free(tau);
    free(h);
    return rc;
}

int orc_rk12_try(const orc_problem *P, orc_rk12 *S, double *m, double dt,
                 double *err)
{
    const long N = (long)P->n[0] * P->n[1] * P->n[2];
    const long n3 = 3 * N;
    double *dy0 = (double *)malloc(sizeof(double) * n3);
    double *tau = (double *)malloc(sizeof(double) * n3);
    double *h = (double *)malloc(sizeof(double) * n3);
    if (!dy0 || !tau || !h) return -2;
    int rc = eval_torque(P, S, m, h, tau);
    if (!rc) {
        memcpy(dy0, tau, sizeof(double) * n3);
        double hh = dt * P->gamma;
        for (long t = 0; t < n3; t++) m[t] = m[t] + hh * tau[t];
        S->m_normalized = 0;
        S->t += dt;
        rc = eval_torque(P, S, m, h, tau);
        if (!rc) {
            for (long t = 0; t < n3; t++) m[t] += 0.5 * hh * tau[t] + (-0.5 * hh) * dy0[t];
            if (err) *err = orc_maxdiff(n3, tau, dy0) * hh;
            S->m_normalized = 0;
            S->step++;
        }
    }
    free(dy0);
    free(tau);
    free(h);
    return rc;
}
