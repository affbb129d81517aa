/*
 * mmx_oracle.h -- FP64 CPU oracle of the mumax/2 effective-field + LLG/RK12 path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (2_amd/, include/) links,
 * loads or calls this library.  It is the checker: tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py use it to judge the
 * HIP path, never as a substitute for it.
 *
 * Conventions (same as the reference's libmumax2):
 *   - arrays are in the reference's *internal* order (z,y,x): index
 *     i*N1*N2 + j*N2 + k with i = user z (slowest), k = user x (fastest)
 *     (src/mumax/gpu/doc.go:22-35);
 *   - vector fields are SoA: component c of a 3-vector array of N cells sits at
 *     base + c*N; component 0 is the user z component;
 *   - a NULL map means "uniform": the map is 1 and the scalar multiplier
 *     carries the value (src/libmumax2/common_func.h:298-306).
 * Every function is FP64 throughout; the only float32 object is the real-space
 * demag kernel, which the reference stores as float32 (kernel_magnetostatic.go:172).
 *
 * Parity status of each function (what pins it) is listed in the header of
 * mmx_oracle.c next to its definition.
 */
#ifndef MMX_ORACLE_H
#define MMX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants, src/mumax/common/constants.go:16-26. */
#define ORC_PI 3.14159265358979323846264338327950288
#define ORC_MU0 (4.0 * ORC_PI * 1e-7)
#define ORC_GAMMA0 2.211e5
#define ORC_MUB 9.2740091523e-24
#define ORC_QE 1.60217646e-19

/* engine/zeropad.go:12-22 */
void orc_pad_size(const int n[3], const int pbc[3], int p[3]);

/* modules/kernel_magnetostatic.go:30-181.  K9 holds 9 float32 components of
 * p0*p1*p2 values each, in the reference's tensor order XX,YY,ZZ,YZ,XZ,XY,ZY,ZX,YX
 * (common/index.go:23-31), K[FullTensorIdx[s][d]] = field d of source s.
 * Must be zeroed by the caller (the reference accumulates with +=). */
int orc_kernel_arne(const int p[3], const double cell[3], const int pbc[3],
                    int accuracy, float *K9);

/* FP64 Newell tensor (newell.c; no reference counterpart).  K = -N for
 * displacement R (internal axes) and cell d, stored order XX,YY,ZZ,YZ,XZ,XY.
 * method 0: analytic formulas below 8 largest cell sizes, 5-point-per-axis
 * far-field rule beyond (the production method); 1: analytic FP64 everywhere;
 * 2: analytic in __float128; 3: far-field rule (5 points); 4: far-field rule
 * (4 points). */
void orc_newell_point(const double R[3], const double d[3], int method, double out6[6]);
/* n-point (n = 4, 5) Gauss rule for the density 1 - |t| on [-1, 1]. */
int orc_newell_rule(int n, double *t, double *w);
/* Padded real-space Newell kernel, 9 x p0*p1*p2 doubles (orc_kernel_arne's
 * layout); non-periodic only (-1 otherwise). */
int orc_kernel_newell(const int p[3], const double cell[3], const int pbc[3], double *K9d);
/* Its transform (as orc_demag_kernel_fft), without the real-space kernel. */
int orc_newell_khat(const int p[3], const double cell[3], const int pbc[3], double *khat6,
                    double *bad_over_good);
/* orc_demag_kernel_fft for components generated on demand by fill(ctx, c,
 * interleaved complex out[p0*p1*p2]); cmax6 = max|K_c| for the IsZero test. */
int orc_demag_kernel_fft_cb(const int p[3], const double cmax6[6],
                            void (*fill)(void *ctx, int c, double *re_im), void *ctx,
                            double *khat6, double *bad_over_good);

/* In-place complex FFT of one contiguous line, interleaved (re,im), length n,
 * sign -1 = forward (exp(-2 pi i jk/n)), +1 = backward, unnormalised. */
int orc_fft_line(double *z, int n, int sign);

/* modules/maxwell_convolution.go:313-436 + 576-612: FFT each of the 6 stored
 * components (XX,YY,ZZ,YZ,XZ,XY) of the padded real-space kernel, keep the
 * real part of the kx < p2/2+1 half spectrum, multiply by 1/(p0 p1 p2).
 * A component whose max|K_c| < 1e-5 max|K| is set to zero (IsZero, :416-418).
 * khat6: 6 x p0*p1*(p2/2+1) doubles.  If Kd (double) is non-NULL it is used
 * instead of K9 (float).  *bad_over_good receives max|Im|/max|Re| (the
 * reference panics above 1e-5).  */
int orc_demag_kernel_fft(const int p[3], const float *K9, const double *K9d,
                         double *khat6, double *bad_over_good);

/* MaxwellPlan.update for the dipole kernel (maxwell_convolution.go:222-309):
 * out_d = sum_s IFFT(khat_sd * FFT(pad(M_s))) cropped to n, with
 * M_s = m_s * msat_map (NULL: 1).  No mu0*Ms factor: the caller applies it. */
int orc_demag_conv(const int n[3], const int p[3], const double *khat6,
                   const double *m, const double *msat_map, double *out);

/* The same convolution by brute-force real-space summation on the padded
 * (circular) grid with the float32 kernel K9 (upper triangle, zero components
 * skipped as above).  O(N^2): tiny grids only.  */
int orc_demag_brute(const int n[3], const int p[3], const float *K9,
                    const double *m, const double *msat_map, double *out);

/* libmumax2/exchange6.cu:12-163 (+ common_func.h:298-317).  pre = 2A/(mu0 Ms). */
void orc_exchange6(const int n[3], const int pbc[3], const double cell[3],
                   double pre, const double *msat_map, const double *aex_map,
                   const double *m, double *h);

/* libmumax2/uniaxialanisotropy.cu:12-121; ku2 = 2 Ku/(mu0 Ms); u_map: the
 * anisU mask (3 components, u_c = u[c] * u_map_c) or NULL. */
void orc_uniaxial(long N, double ku2, const double *ku_map,
                  const double *msat_map, const double u[3],
                  const double *u_map, const double *m, double *h);

/* libmumax2/torque.cu:13-78: tau/gamma, internal-coordinate formula. */
void orc_torque(long N, const double *m, const double *h,
                const double *alpha_map, double alpha_mul, double *tau);

/* libmumax2/normalize.cu:14-60 */
void orc_normalize(long N, double *m, const double *norm_map);

/* libmumax2/zhang-li_torque.cu:123-300 (single device, so the y-neighbour
 * "device slices" are the array itself).  jmask: 3 component maps or NULL. */
void orc_zhangli(const int n[3], const int pbc[3], const double cell[3],
                 const double *m, const double *jmask, const double jmul[3],
                 const double *msat_map, const double *pol_map,
                 const double *xi_map, const double *alpha_map, double pred,
                 double xi_mul, double alpha_mul, double *stt);

/* Problem description for the composite field and the solver. */
typedef struct {
    int n[3];
    int pbc[3];
    double cell[3];
    double msat;            const double *msat_map;
    double aex;             const double *aex_map;
    double alpha;           const double *alpha_map;
    double gamma;
    double bext[3];         /* uniform applied field B_ext in T */
    int demag;  int p[3];   const double *khat6;
    int anis;   double ku;  const double *ku_map;  double anis_u[3];
    int zhangli; double jmul[3]; const double *jmask; double pol; double xi;
    /* MASK quantities (NULL = uniform): B_ext (3 components,
     * mod_zeeman.go:26), anisU (3, mod_uniaxialanisotropy.go:30),
     * polarisation and xi (zhang-li_torque.cu:167,171) */
    const double *bext_map; const double *u_map;
    const double *pol_map;  const double *xi_map;
} orc_problem;

/* SumUpdater for H_eff (engine/update_sum.go:29-75, modules/hfield.go:19-29):
 * H_eff = B_demag/mu0 + H_ex + H_anis + B_ext/mu0.  Terms returned separately
 * when the pointers are non-NULL. */
int orc_heff(const orc_problem *P, const double *m, double *h,
             double *h_demag, double *h_ex, double *h_anis);

/* torque quantity (modules/mod_llg.go:59-70, + zhang-li added with weight
 * 1/gamma, modules/mod_zhang-li.go:50-70, engine/update_addterm.go:26-40). */
int orc_torque_full(const orc_problem *P, const double *m, const double *h,
                    double *tau);

/* engine/solve_rk12.go:20-214 solver state. */
typedef struct {
    double t, dt, mindt, maxdt, maxerr;
    double err, peakerr;
    long step, badsteps;
    int m_normalized;       /* lazy normalisation flag (normUpdater)      */
    int evals;              /* H_eff evaluations so far                     */
} orc_rk12;

void orc_rk12_init(orc_rk12 *S);
/* One accepted step with the reference's retry loop.  Returns 0, or -1 when
 * the solver cannot converge (the reference panics). */
int orc_rk12_step(const orc_problem *P, orc_rk12 *S, double *m);
/* One predictor/corrector try with a given dt and no step control
 * (for same-dt replays).  *err receives the reference's error estimate. */
int orc_rk12_try(const orc_problem *P, orc_rk12 *S, double *m, double dt,
                 double *err);

/* <m> and maxtorque reductions (engine/update_reduce.go:47-108,
 * gpu/reductor.go:69-188), compensated FP64 sums. */
void orc_average(long N, long ncell, const double *a, double *avg);
double orc_maxnorm(long N, const double *a);
double orc_maxdiff(long n, const double *a, const double *b);

#ifdef __cplusplus
}
#endif
#endif
