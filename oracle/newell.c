/* newell.c -- FP64 Newell demag tensor on the CPU (test infrastructure).
 *
 * The reference has no Newell kernel (its dipole kernel is Kernel_Arne,
 * modules/kernel_magnetostatic.go:30-181); this is the oracle of the build's
 * FP64 Newell construction.  The formulas are Newell, Williams & Dunlop,
 * "A generalization of the demagnetizing tensor for nonuniform magnetization",
 * J. Geophys. Res. 98 (1993) 9551: the cell-averaged tensor
 *
 *   N_aa(R) = 1/(4 pi V) sum_{i,j,k in -1,0,1} w_i w_j w_k f(R_a+i d_a, R_b+j d_b, R_c+k d_c)
 *   N_ab(R) = same with g(R_a.., R_b.., R_c..),   w_0 = 2, w_{+-1} = -1,
 *
 * with f, g as below.  The 27-point combination is a sixth-order difference
 * of functions that grow like |R|^3 while N falls like V/|R|^3, so it loses
 * about log10(64 |R|^6 / V^2) digits: in FP64 it is exact to ~1e-9 up to
 * 8 cells and useless beyond ~100.  Far away the tensor is instead the
 * point-dipole tensor averaged over the offset s = u' - u of two points drawn
 * uniformly from the two cells (each component of s has the triangular
 * density (1 - |t|) on [-1, 1] in units of the cell size), evaluated with an
 * n-point Gauss rule of that weight per axis.
 *
 * Sign convention of the reference's kernel: K = -N (B = mu0 Ms K * m; the
 * self term of a cube is -1/3 per diagonal component).
 *
 * What pins it (tests/test_oracle_newell.py): the same formulas in __float128
 * (method 2, immune to the cancellation at these distances) at near and far
 * points; the trace identity (N_xx + N_yy + N_zz = 1 at R = 0 and 0 elsewhere,
 * which fails for a dropped or mis-permuted f term); the cube self term 1/3;
 * the point-dipole limit; the quadrature rule's moments; the convergence of
 * Kernel_Arne to it with the integration accuracy.
 */
#include "mmx_oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ f and g */

/* Newell's f(x, y, z), even in each argument (evaluated at |x|,|y|,|z|). */
static double nw_f(double x, double y, double z)
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

/* Newell's g(x, y, z), odd in x and y, even in z. */
static double nw_g(double x, double y, double z)
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

static __float128 nw_fq(__float128 x, __float128 y, __float128 z)
{
    x = fabsq(x); y = fabsq(y); z = fabsq(z);
    const __float128 x2 = x * x, y2 = y * y, z2 = z * z;
    const __float128 R = sqrtq(x2 + y2 + z2);
    __float128 t = 0;
    if (y > 0 && x2 + z2 > 0) t += y * (z2 - x2) * asinhq(y / sqrtq(x2 + z2)) / 2;
    if (z > 0 && x2 + y2 > 0) t += z * (y2 - x2) * asinhq(z / sqrtq(x2 + y2)) / 2;
    if (x * y * z > 0) t -= x * y * z * atanq(y * z / (x * R));
    t += (2 * x2 - y2 - z2) * R / 6;
    return t;
}

static __float128 nw_gq(__float128 x, __float128 y, __float128 z)
{
    const int neg = (x < 0) != (y < 0);
    x = fabsq(x); y = fabsq(y); z = fabsq(z);
    const __float128 x2 = x * x, y2 = y * y, z2 = z * z;
    const __float128 R = sqrtq(x2 + y2 + z2);
    __float128 t = 0;
    if (z > 0 && x2 + y2 > 0) t += x * y * z * asinhq(z / sqrtq(x2 + y2));
    if (x > 0 && y2 + z2 > 0) t += y * (3 * z2 - y2) * asinhq(x / sqrtq(y2 + z2)) / 6;
    if (y > 0 && x2 + z2 > 0) t += x * (3 * z2 - x2) * asinhq(y / sqrtq(x2 + z2)) / 6;
    if (z > 0) t -= z2 * z * atanq(x * y / (z * R)) / 6;
    if (y > 0) t -= z * y2 * atanq(x * z / (y * R)) / 2;
    if (x > 0) t -= z * x2 * atanq(y * z / (x * R)) / 2;
    t -= x * y * R / 3;
    return neg ? -t : t;
}

/* stored order XX,YY,ZZ,YZ,XZ,XY: axis pairs (a,b), third axis c */
static const int kPairA[6] = {0, 1, 2, 1, 0, 0};
static const int kPairB[6] = {0, 1, 2, 2, 2, 1};

static void newell_analytic(const double R[3], const double d[3], double out6[6])
{
    static const double w[3] = {-1.0, 2.0, -1.0};
    const double V = d[0] * d[1] * d[2];
    for (int c = 0; c < 6; c++) {
        const int a = kPairA[c], b = kPairB[c];
        int cc = 3 - a - b;
        int bb = b;
        if (a == b) { bb = (a + 1) % 3; cc = (a + 2) % 3; }
        double s = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    const double xa = R[a] + (i - 1) * d[a];
                    const double xb = R[bb] + (j - 1) * d[bb];
                    const double xc = R[cc] + (k - 1) * d[cc];
                    const double v = (a == b) ? nw_f(xa, xb, xc) : nw_g(xa, xb, xc);
                    s += w[i] * w[j] * w[k] * v;
                }
        out6[c] = -s / (4.0 * ORC_PI * V);
    }
}

static void newell_analytic_q(const double R[3], const double d[3], double out6[6])
{
    const __float128 w[3] = {-1, 2, -1};
    const __float128 V = (__float128)d[0] * d[1] * d[2];
    for (int c = 0; c < 6; c++) {
        const int a = kPairA[c], b = kPairB[c];
        int cc = 3 - a - b;
        int bb = b;
        if (a == b) { bb = (a + 1) % 3; cc = (a + 2) % 3; }
        __float128 s = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    const __float128 xa = (__float128)R[a] + (i - 1) * (__float128)d[a];
                    const __float128 xb = (__float128)R[bb] + (j - 1) * (__float128)d[bb];
                    const __float128 xc = (__float128)R[cc] + (k - 1) * (__float128)d[cc];
                    const __float128 v = (a == b) ? nw_fq(xa, xb, xc) : nw_gq(xa, xb, xc);
                    s += w[i] * w[j] * w[k] * v;
                }
        out6[c] = (double)(-s / (4 * M_PIq * V));
    }
}

/* ---------------------------------------------- Gauss rule for 1 - |t| */

/* n-point Gauss rule on [-1,1] for the density 1 - |t| (total weight 1),
 * n = 4 or 5.  The even moments are E[t^2j] = 2/((2j+1)(2j+2)).  Nodes come in
 * pairs +-sqrt(u): for n = 4 the two u are the roots of the quadratic
 * orthogonal to 1 and u under the distribution of u = t^2; for n = 5 they are
 * orthogonal under u dmu(u), plus the node 0.  Weights from the moment
 * equations. */
int orc_newell_rule(int n, double *t, double *w)
{
    long double mu[6];
    for (int j = 0; j < 6; j++) mu[j] = 2.0L / ((2.0L * j + 1.0L) * (2.0L * j + 2.0L));
    if (n != 4 && n != 5) return -1;
    const int o = (n == 5) ? 1 : 0; /* moments of u dmu are mu[j+1] */
    /* u^2 + c1 u + c0 orthogonal to 1 and u: */
    /*   mu[o+2] + c1 mu[o+1] + c0 mu[o] = 0, mu[o+3] + c1 mu[o+2] + c0 mu[o+1] = 0 */
    const long double a11 = mu[o + 1], a12 = mu[o], b1 = -mu[o + 2];
    const long double a21 = mu[o + 2], a22 = mu[o + 1], b2 = -mu[o + 3];
    const long double det = a11 * a22 - a12 * a21;
    const long double c1 = (b1 * a22 - a12 * b2) / det;
    const long double c0 = (a11 * b2 - b1 * a21) / det;
    const long double disc = sqrtl(c1 * c1 - 4.0L * c0);
    const long double u1 = (-c1 + disc) / 2.0L, u2 = (-c1 - disc) / 2.0L;
    /* 2 (W1 u1 + W2 u2) = mu1, 2 (W1 u1^2 + W2 u2^2) = mu2 */
    const long double A = (mu[2] / 2.0L - u2 * mu[1] / 2.0L) / (u1 - u2); /* W1 u1 */
    const long double B = mu[1] / 2.0L - A;                                /* W2 u2 */
    const long double W1 = A / u1, W2 = B / u2;
    const long double r1 = sqrtl(u1), r2 = sqrtl(u2);
    t[0] = (double)-r1; w[0] = (double)W1;
    t[1] = (double)-r2; w[1] = (double)W2;
    if (n == 5) {
        t[2] = 0.0; w[2] = (double)(1.0L - 2.0L * (W1 + W2));
        t[3] = (double)r2; w[3] = (double)W2;
        t[4] = (double)r1; w[4] = (double)W1;
    } else {
        t[2] = (double)r2; w[2] = (double)W2;
        t[3] = (double)r1; w[3] = (double)W1;
    }
    return 0;
}

static void newell_far(const double R[3], const double d[3], int nq, double out6[6])
{
    double t[5], w[5];
    orc_newell_rule(nq, t, w);
    const double V = d[0] * d[1] * d[2];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < nq; i++)
        for (int j = 0; j < nq; j++)
            for (int k = 0; k < nq; k++) {
                const double x[3] = {R[0] + t[i] * d[0], R[1] + t[j] * d[1], R[2] + t[k] * d[2]};
                const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
                const double r = sqrt(r2);
                const double ir5 = 1.0 / (r2 * r2 * r);
                const double W = w[i] * w[j] * w[k];
                for (int c = 0; c < 6; c++) {
                    const int a = kPairA[c], b = kPairB[c];
                    const double v = 3.0 * x[a] * x[b] - (a == b ? r2 : 0.0);
                    acc[c] += W * v * ir5;
                }
            }
    /* N = -(V/4pi) <d_a d_b (1/r)>,  K = -N */
    for (int c = 0; c < 6; c++) out6[c] = V / (4.0 * ORC_PI) * acc[c];
}

/* Distance (in units of the largest cell size) from which the far-field rule
 * replaces the analytic formulas, and the rule's order. */
#define NEWELL_RSWITCH 8.0
#define NEWELL_NQ 5

void orc_newell_point(const double R[3], const double d[3], int method, double out6[6])
{
    switch (method) {
    case 1: newell_analytic(R, d, out6); return;
    case 2: newell_analytic_q(R, d, out6); return;
    case 3: newell_far(R, d, NEWELL_NQ, out6); return;
    case 4: newell_far(R, d, 4, out6); return;
    default: break;
    }
    double dm = d[0];
    if (d[1] > dm) dm = d[1];
    if (d[2] > dm) dm = d[2];
    const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
    if (r2 >= NEWELL_RSWITCH * NEWELL_RSWITCH * dm * dm)
        newell_far(R, d, NEWELL_NQ, out6);
    else
        newell_analytic(R, d, out6);
    /* K_ab is odd in R_a and in R_b: exactly 0 on those planes */
    for (int c = 3; c < 6; c++)
        if (R[kPairA[c]] == 0.0 || R[kPairB[c]] == 0.0) out6[c] = 0.0;
}

/* octant table of K at displacements (i,j,k) * cell, e = octant extent */
static double *newell_octant_table(const int p[3], const double cell[3], int e[3], long *ne_out)
{
    for (int a = 0; a < 3; a++) e[a] = (p[a] == 1) ? 1 : p[a] / 2;
    const long ne = (long)e[0] * e[1] * e[2];
    double *oct = (double *)malloc(sizeof(double) * 6 * ne);
    if (!oct) return NULL;
#pragma omp parallel for schedule(dynamic, 64)
    for (long t = 0; t < ne; t++) {
        const int k = (int)(t % e[2]), j = (int)((t / e[2]) % e[1]), i = (int)(t / ((long)e[2] * e[1]));
        const double R[3] = {i * cell[0], j * cell[1], k * cell[2]};
        orc_newell_point(R, cell, 0, oct + 6 * t);
    }
    *ne_out = ne;
    return oct;
}

/* value of component c at padded index t from the octant table (0 on the
 * p/2 planes; off-diagonal components odd in both of their axes) */
static double newell_at(const double *oct, const int p[3], const int e[3], int c, long t)
{
    const int z = (int)(t % p[2]), y = (int)((t / p[2]) % p[1]), x = (int)(t / ((long)p[2] * p[1]));
    const int q[3] = {x, y, z};
    int disp[3];
    for (int a = 0; a < 3; a++) {
        if (p[a] == 1) { disp[a] = 0; continue; }
        if (2 * q[a] == p[a]) return 0.0;
        disp[a] = (2 * q[a] < p[a]) ? q[a] : q[a] - p[a];
    }
    const long o = ((long)abs(disp[0]) * e[1] + abs(disp[1])) * e[2] + abs(disp[2]);
    double v = oct[6 * o + c];
    const int a = kPairA[c], b = kPairB[c];
    if (a != b) {
        if (disp[a] < 0) v = -v;
        if (disp[b] < 0) v = -v;
    }
    return v;
}

struct newell_ctx {
    const double *oct;
    int p[3], e[3];
};

static void newell_fill(void *vctx, int c, double *re_im)
{
    const struct newell_ctx *x = (const struct newell_ctx *)vctx;
    const long P = (long)x->p[0] * x->p[1] * x->p[2];
#pragma omp parallel for schedule(static)
    for (long t = 0; t < P; t++) {
        re_im[2 * t] = newell_at(x->oct, x->p, x->e, c, t);
        re_im[2 * t + 1] = 0.0;
    }
}

int orc_newell_khat(const int p[3], const double cell[3], const int pbc[3], double *khat6,
                    double *bad_over_good)
{
    if (pbc[0] || pbc[1] || pbc[2]) return -1;
    struct newell_ctx x;
    long ne = 0;
    double *oct = newell_octant_table(p, cell, x.e, &ne);
    if (!oct) return -2;
    double cmax[6] = {0, 0, 0, 0, 0, 0};
    for (long t = 0; t < ne; t++)
        for (int c = 0; c < 6; c++)
            if (fabs(oct[6 * t + c]) > cmax[c]) cmax[c] = fabs(oct[6 * t + c]);
    x.oct = oct;
    for (int a = 0; a < 3; a++) x.p[a] = p[a];
    const int rc = orc_demag_kernel_fft_cb(p, cmax, newell_fill, &x, khat6, bad_over_good);
    free(oct);
    return rc;
}

/* The padded real-space kernel with the Newell tensor, laid out like
 * orc_kernel_arne's output but in double: 9 x p0*p1*p2, tensor order
 * XX,YY,ZZ,YZ,XZ,XY,ZY,ZX,YX.  Displacements |R_a| <= p_a/2 - 1 (the range
 * Kernel_Arne fills, kernel_magnetostatic.go:51-57; the p_a/2 planes stay 0).
 * Non-periodic only.  Returns 0, or -1 for a periodic axis. */
int orc_kernel_newell(const int p[3], const double cell[3], const int pbc[3], double *K9d)
{
    if (pbc[0] || pbc[1] || pbc[2]) return -1;
    const long P = (long)p[0] * p[1] * p[2];
    int e[3];
    long ne = 0;
    double *oct = newell_octant_table(p, cell, e, &ne);
    if (!oct) return -2;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < P; t++) {
        for (int c = 0; c < 6; c++) K9d[c * P + t] = newell_at(oct, p, e, c, t);
        K9d[6 * P + t] = K9d[3 * P + t]; /* ZY = YZ */
        K9d[7 * P + t] = K9d[4 * P + t]; /* ZX = XZ */
        K9d[8 * P + t] = K9d[5 * P + t]; /* YX = XY */
    }
    free(oct);
    return 0;
}
