/*
 * mmx.h -- C ABI of the MI355X (gfx950) micromagnetic solver core.
 *
 * Drop-in for the hot path of mumax/2: the libmumax2 kernels that the Go
 * package mumax/gpu binds through cgo, the cuFFT-based demag convolution
 * behind MaxwellPlan/FFTPlanX, the Reductor, and the engine-side pieces that
 * drive them (SumUpdater for H_eff, the torque updaters, solver/rk12).
 *
 * Conventions (same as libmumax2: the headers in src/libmumax2/ and src/mumax/gpu/libmumax2.go):
 *  - arrays are float32 device pointers in the reference's internal (z,y,x)
 *    order: index i*N1*N2 + j*N2 + k, i = user z (slowest), k = user x;
 *    component 0 of a vector is the user z component (src/mumax/gpu/doc.go:22-35);
 *  - a vector field is passed as three component pointers (SoA);
 *  - a NULL map pointer means "uniform": the map is 1 and the scalar
 *    multiplier carries the value (src/libmumax2/common_func.h:298-306);
 *  - kernels run asynchronously on `stream` (a hipStream_t, NULL = the null
 *    stream); the caller synchronises, as the Go wrappers do with Stream.Sync();
 *  - the reference is multi-device in form (float** per device); one GPU is
 *    one device, so these entry points take the single device's pointers;
 *  - every entry point returns an int status, 0 = MMX_OK.  The reference
 *    aborts on errors (src/libmumax2/gpu_safe.h:36-41); a caller wanting that
 *    behaviour aborts on a nonzero status (mmx_last_error() says why).
 * Only the plan/reductor/sim *_create calls allocate device memory.
 */
#ifndef MMX_H
#define MMX_H

#ifdef __cplusplus
extern "C" {
#endif

#define MMX_OK 0
#define MMX_ERR_ARG 1      /* bad argument: sizes, required pointer NULL      */
#define MMX_ERR_SIZE 2     /* unsupported size (FFT length with prime > 7)    */
#define MMX_ERR_HIP 3      /* HIP runtime error                               */
#define MMX_ERR_NOMEM 4    /* device allocation failed                        */
#define MMX_ERR_CONVERGE 5 /* rk12 cannot converge after maxTry bad steps     */
#define MMX_ERR_NAN 6      /* NaN/Inf in the error estimate                   */
#define MMX_ERR_NAME 7     /* unknown quantity or module name                 */
#define MMX_ERR_STATE 8    /* call not valid in the current state             */

/* Text of the last error on this thread (never NULL). */
const char *mmx_last_error(void);

/* Library version string. */
const char *mmx_version(void);

/* Selects the GPU used by later calls; replaces setUsedGPUs
 * (src/libmumax2/multigpu.cu:11-22) for the one-device case. */
int mmx_init(int device);

/* ------------------------------------------------------------------------
 * Field and torque kernels
 * ---------------------------------------------------------------------- */

/* m <- m/|m|, 0 where norm_map == 0 or |m| == 0.
 * Replaces normalizeAsync (src/libmumax2/normalize.h, normalize.cu:45-60). */
int mmx_normalize(float *mx, float *my, float *mz, const float *norm_map,
                  int N, void *stream);

/* tau/gamma = (m x H - alpha m x (m x H)) / (1 + alpha^2), internal coords.
 * Replaces torqueAsync (src/libmumax2/torque.h:35, torque.cu:59-78). */
int mmx_torque(float *tx, float *ty, float *tz, const float *mx,
               const float *my, const float *mz, const float *hx,
               const float *hy, const float *hz, const float *alpha_map,
               float alpha_mul, int N, void *stream);

/* 6-neighbour exchange field, harmonic-mean Aex/Msat weights, clamped
 * (Neumann) or wrapped edges.  pre = 2 Aex_mul / (mu0 Msat_mul).
 * Replaces exchange6Async (src/libmumax2/exchange6.h:20, exchange6.cu:145-163). */
int mmx_exchange6(float *hx, float *hy, float *hz, const float *mx,
                  const float *my, const float *mz, const float *msat_map,
                  const float *aex_map, float pre, int N0, int N1, int N2,
                  int periodic0, int periodic1, int periodic2, float cell0,
                  float cell1, float cell2, void *stream);

/* H = (2 Ku/(mu0 Ms)) Ku_map/Msat_map (m.u) u.
 * Replaces uniaxialAnisotropyAsync (src/libmumax2/uniaxialanisotropy.h,
 * uniaxialanisotropy.cu:234-265). */
int mmx_uniaxial_anisotropy(float *hx, float *hy, float *hz, const float *mx,
                            const float *my, const float *mz,
                            const float *ku_map, const float *msat_map,
                            float ku2_mu0msat_mul, const float *u_map_x,
                            float u_mul_x, const float *u_map_y, float u_mul_y,
                            const float *u_map_z, float u_mul_z, int N,
                            void *stream);

/* Zhang-Li spin-transfer torque (5-point derivative along the current),
 * pred = P muB/(e Ms).  One launch for the whole volume.
 * Replaces zhangli_async (src/libmumax2/zhang-li_torque.h:21-35,
 * zhang-li_torque.cu:307-429). */
int mmx_zhangli(float *sttx, float *stty, float *sttz, const float *mx,
                const float *my, const float *mz, const float *jx_map,
                const float *jy_map, const float *jz_map,
                const float *msat_map, const float *pol_map,
                const float *xi_map, const float *alpha_map, float jmul_x,
                float jmul_y, float jmul_z, float pred, float xi_mul,
                float alpha_mul, int N0, int N1, int N2, float cell0,
                float cell1, float cell2, int periodic0, int periodic1,
                int periodic2, void *stream);

/* ------------------------------------------------------------------------
 * Linear algebra used by the sum updaters and the RK stages
 * ---------------------------------------------------------------------- */

/* dst_c = a_c + mul_c * b_c  (b == NULL: dst_c = a_c + mul_c).
 * Replaces vecMaddAsync (src/libmumax2/add.h, add.cu:119-150). */
int mmx_vec_madd(float *dstx, float *dsty, float *dstz, const float *ax,
                 const float *ay, const float *az, const float *bx,
                 const float *by, const float *bz, float mulx, float muly,
                 float mulz, int N, void *stream);

/* dst = a + mul*b  (b == NULL: dst = a + mul).
 * Replaces maddAsync (src/libmumax2/add.h, add.cu:102-118). */
int mmx_madd(float *dst, const float *a, const float *b, float mul, int N,
             void *stream);

/* a += mulb*b  (b == NULL: b = 1).  Replaces madd1Async (add.cu:169-177). */
int mmx_madd1(float *a, const float *b, float mulb, int N, void *stream);

/* a += mulb*b + mulc*c (NULL: 1).  Replaces madd2Async (add.cu:191-199). */
int mmx_madd2(float *a, const float *b, float mulb, const float *c,
              float mulc, int N, void *stream);

/* dst = a*b.  Replaces mulAsync (src/libmumax2/mul.h). */
int mmx_mul(float *dst, const float *a, const float *b, int N, void *stream);

/* ------------------------------------------------------------------------
 * Reductor (src/mumax/gpu/reductor.go:19-212 + libmumax2/reduce.cu)
 * The reference returns float32 results of float32 partial sums; here sums
 * are accumulated in FP64 and returned as double.  All calls synchronise.
 * ---------------------------------------------------------------------- */
typedef struct mmx_reductor mmx_reductor;

/* Reductor.Init: buffers for arrays of up to n floats. */
int mmx_reductor_create(mmx_reductor **r, long n);
int mmx_reductor_destroy(mmx_reductor *r);
/* Reductor.Sum / Dot */
int mmx_reductor_sum(mmx_reductor *r, const float *a, long n, void *stream,
                     double *out);
int mmx_reductor_dot(mmx_reductor *r, const float *a, const float *b, long n,
                     void *stream, double *out);
/* Reductor.Max / Min / MaxAbs / MaxDiff */
int mmx_reductor_max(mmx_reductor *r, const float *a, long n, void *stream,
                     float *out);
int mmx_reductor_min(mmx_reductor *r, const float *a, long n, void *stream,
                     float *out);
int mmx_reductor_maxabs(mmx_reductor *r, const float *a, long n, void *stream,
                        float *out);
int mmx_reductor_maxdiff(mmx_reductor *r, const float *a, const float *b,
                         long n, void *stream, float *out);
/* Reductor.MaxNorm: sqrt(max(ax^2+ay^2+az^2)) */
int mmx_reductor_maxnorm(mmx_reductor *r, const float *ax, const float *ay,
                         const float *az, long n, void *stream, float *out);

/* Allocation-free reduction.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * result as the mmx_reductor_* call named by `op`, for n in (0,
 * MMX_NOALLOC_MAX_N] floats (the reference's int sizes); larger n fails with
 * MMX_ERR_SIZE.  Operands: a (sum, max, min, maxabs), a and b (dot, maxdiff),
 * a, b, c as the x, y, z components (maxnorm).  The result is returned as a
 * double: the FP64 sum for sum/dot, the float result for the others.  The
 * partials go to a 16 KiB device array built into the library (one per GPU,
 * held until the process ends), so no call allocates or frees; calls that
 * overlap on one GPU (several host threads or streams) race on it.  The call
 * synchronises `stream`. */
#define MMX_RED_SUM 0
#define MMX_RED_DOT 1
#define MMX_RED_MAX 2
#define MMX_RED_MIN 3
#define MMX_RED_MAXABS 4
#define MMX_RED_MAXDIFF 5
#define MMX_RED_MAXNORM 6
#define MMX_NOALLOC_MAX_N 2147483647L
int mmx_reduce_noalloc(int op, const float *a, const float *b, const float *c,
                       long n, void *stream, double *out);

/* ------------------------------------------------------------------------
 * Demag convolution plan: MaxwellPlan (src/mumax/modules/maxwell_convolution.go)
 * with its FFTPlanX (src/mumax/gpu/fftplanX.go) and kernel construction
 * (Kernel_Arne, src/mumax/modules/kernel_magnetostatic.go:30-181).
 * The plan owns its spectrum buffers and the transformed kernel.
 * ---------------------------------------------------------------------- */
typedef struct mmx_demag_plan mmx_demag_plan;

#define MMX_KERNEL_ARNE 0   /* the reference's kernel (accuracy as given)  */
#define MMX_KERNEL_NEWELL 1 /* Newell tensor, FP64 (no reference counterpart) */

/* MaxwellPlan.init + loadDipoleKernel + LoadKernel.  Builds the kernel in
 * FP64 on the GPU, transforms it in FP64, stores 6 real symmetric components
 * (octant) in float32.  Sizes in internal order.  MMX_KERNEL_ARNE is the
 * reference's Kernel_Arne at the given accuracy; MMX_KERNEL_NEWELL is the
 * analytic cell-averaged tensor (Newell, Williams & Dunlop 1993) in FP64 with
 * a quadrature far field (accuracy ignored; non-periodic grids only). */
int mmx_demag_plan_create(mmx_demag_plan **plan, int n0, int n1, int n2,
                          double cell0, double cell1, double cell2, int pbc0,
                          int pbc1, int pbc2, int kernel_type, int accuracy);
int mmx_demag_plan_destroy(mmx_demag_plan *plan);

/* MaxwellPlan.UpdateB for the demag source: B = mu0 Ms (K * M) with
 * M = m * msat_map (NULL: 1), Ms = msat_mul.  Output in tesla. */
int mmx_demag_update_b(mmx_demag_plan *plan, float *bx, float *by, float *bz,
                       const float *mx, const float *my, const float *mz,
                       const float *msat_map, float msat_mul, void *stream);

/* Tuning of the demag x passes (no counterpart in the reference, whose FFT is
 * cuFFT's choice): the register-FFT x passes run when they span at least
 * `nblocks` workgroups, smaller grids take the generic LDS passes (faster
 * there: fewer, longer serial chains otherwise).  Default 256; 0 runs the
 * register passes wherever their lengths allow.  Process-wide; *previous (if
 * not NULL) receives the value it replaces. */
int mmx_set_fast_x_min_blocks(int nblocks, int *previous);

/* Plan geometry: padded size p[3]; octant kernel size ko[3] = (p0/2+1,
 * p1/2+1, p2/2+1); realness ratio max|Im|/max|Re| of the FFT'd kernel
 * (the reference panics above 1e-5, maxwell_convolution.go:608-610). */
int mmx_demag_plan_info(const mmx_demag_plan *plan, int p[3], int ko[3],
                        double *bad_over_good);
/* Copy the transformed kernel (6 x ko0*ko1*ko2 float32, components
 * XX,YY,ZZ,YZ,XZ,XY, already scaled by 1/(p0 p1 p2)) to host memory. */
int mmx_demag_plan_kernel(const mmx_demag_plan *plan, float *host);
/* Copy the real-space kernel (6 x p0*p1*p2 float32, same component order,
 * as Kernel_Arne stores it) to host memory; only valid until the plan's
 * build scratch is released (flag keep_real_space of the sim, or plans
 * created with MMX_KEEP_REAL_KERNEL in kernel_type). */
#define MMX_KEEP_REAL_KERNEL 0x100
int mmx_demag_plan_real_kernel(const mmx_demag_plan *plan, float *host);

/* ------------------------------------------------------------------------
 * Convolution plan: ConvPlan (src/mumax/gpu/convplan.go:22-231), the
 * convolution that the demagexch module runs with its combined kernel
 * (src/mumax/modules/mod_demagexchange.go:26-205).
 * ---------------------------------------------------------------------- */
typedef struct mmx_convplan mmx_convplan;

/* ConvPlan.Init + loadKernel (convplan.go:33-115).  dataSize: the non-zero
 * input block (internal order); logicSize: the kernel's size, per axis the
 * data size (periodic axis, or length 1) or twice it (zero padding), as
 * PadSize gives it (engine/zeropad.go:12-22); kernel: 6 host float32 arrays of
 * logicSize (components XX,YY,ZZ,YZ,XZ,XY, common/index.go:57-63; NULL =
 * unused), not normalised (the plan applies 1/prod(logicSize)).  The kernel is
 * transformed in FP64 and stored as its octant (6 real values per point), which
 * needs the mirror symmetry of a demag/exchange kernel: even in every axis
 * except, for an off-diagonal component, its two axes; a kernel without it is
 * refused (MMX_ERR_ARG).  The imaginary-part ratio is recorded
 * (mmx_convplan_info), as ConvPlan warns about it (convplan.go:143-152).
 * Allocates the plan's device buffers. */
int mmx_convplan_create(mmx_convplan **cp, const int dataSize[3],
                        const int logicSize[3], const float *const kernel[6]);
int mmx_convplan_destroy(mmx_convplan *cp);
/* ConvPlan.Convolve (convplan.go:163-176): out = kernel (*) in for the 3
 * components (component c at base + c*N, N = cells of dataSize; in != out),
 * asynchronous on `stream`. */
int mmx_convplan_convolve(mmx_convplan *cp, const float *in, float *out,
                          void *stream);
/* max |Im| / max |Re| of the transformed kernel */
int mmx_convplan_info(const mmx_convplan *cp, double *bad_over_good);

/* ------------------------------------------------------------------------
 * Simulation engine for this path: the Engine/Quant/module/solver pieces of
 * src/mumax/engine and src/mumax/modules that drive the kernels above
 * (micromagnetism = demag + exchange6 + zeeman + llg + solver/rk12 +
 * maxtorque, mod_micromag.go:22-34; optional anisotropy/uniaxial, zhang-li).
 * Quantity names are the reference's; values are in internal (z,y,x) order
 * (the reference's API layer swaps user XYZ, engine/api.go:43-45).
 * ---------------------------------------------------------------------- */
typedef struct mmx_sim mmx_sim;

/* setgridsize + setcellsize (+ periodic), load('micromagnetism'). */
int mmx_sim_create(mmx_sim **sim, int n0, int n1, int n2, double cell0,
                   double cell1, double cell2, int pbc0, int pbc1, int pbc2);
int mmx_sim_destroy(mmx_sim *sim);
/* Engine.LoadModule: "anisotropy/uniaxial", "zhang-li", and the solvers
 * "solver/rk12" (default, adaptive), "solver/heun" and "solver/euler" (fixed
 * dt; solve_heun.go:27-79, solve_euler.go:22-50), "solver/bdf-euler" (fixed
 * dt, fixed-point backward Euler; solve_euler_backward.go:29-91) and
 * "temperature/brown" (Brown's fluctuating field H_therm added to H_eff,
 * mod_temperature.go:22-134; see mmx_philox4x32_10 for the noise),
 * "solver/am12" (adaptive implicit-Euler predictor / trapezoidal corrector;
 * solve_am01.go:45-288; loading it sets mindt = 1e-38, maxdt = 1e38 as the
 * module does), which replace each other, and "demagexch" (H_dex =
 * conv(m, Ms K_d + 2 Aex / (mu0 Ms) K_ex) added to H_eff, K_d Kernel_Arne at
 * demag_acc, K_ex the 6-neighbour Laplacian; mod_demagexchange.go:26-205; the
 * H_eff evaluations fold it into the demag convolution, so it costs no pass;
 * an Msat mask, the surrogate and batches are refused with it) ("micromagnetism", "demag",
 * "exchange6", "zeeman", "llg", "maxtorque" are already loaded).  Names are
 * case-insensitive, as the reference's quantity names are. */
int mmx_sim_load(mmx_sim *sim, const char *module);
/* Demag kernel choice (default MMX_KERNEL_ARNE, accuracy 6); a change
 * rebuilds the plan at the next update. */
int mmx_sim_set_kernel(mmx_sim *sim, int kernel_type, int accuracy);
/* setv/getv on VALUE/MASK multipliers: Msat, Aex, alpha, gamma, B_ext[3],
 * Ku, anisU[3], j[3], polarisation, xi, t, dt, mindt, maxdt, m_maxerror,
 * m_maxIterations and m_maxIterError (bdf-euler; default 3, 1e-5),
 * m_maxAbsError and m_maxRelError (am12; default 1e-4, 1e-3), Temp,
 * Therm_seed and cutoff_dt (temperature/brown), demag_acc (demagexch; a
 * positive integer, default 8);
 * read-only: step, badsteps (am12: of the last step), bdf_iterations (of the
 * last step), m_error, m_peakerror, <m>[3], <H_therm>[3], maxtorque,
 * evals (H_eff evaluations so far), and the energies of micromag/energy
 * (modules/mod_energy.go:24-100): E_zeeman, E_ex, E_demag, E_anis and their
 * sum E, in joule, as FP64 sums with compensated partials. */
int mmx_sim_setv(mmx_sim *sim, const char *name, const double *v, int ncomp);
int mmx_sim_getv(mmx_sim *sim, const char *name, double *v, int ncomp);
/* setmask: spatial maps (host float32, N per component) for Msat, Aex,
 * alpha, Ku, j (3 comps), Temp; NULL host pointer removes the map. */
int mmx_sim_setmask(mmx_sim *sim, const char *name, const float *host,
                    int ncomp);
/* setarray/getarray (host float32, SoA, internal order): m (set/get),
 * H_eff, torque, B, H_therm, H_dex (get; updated first, like Quant.Update). */
int mmx_sim_setarray(mmx_sim *sim, const char *name, const float *host,
                     int ncomp);
int mmx_sim_getarray(mmx_sim *sim, const char *name, float *host, int ncomp);
/* Device pointer of a quantity's storage (component c at ptr + c*N).  The
 * pointer is valid until the next step, try or batch run (m and its work
 * buffer swap on an accepted step) or the sim is destroyed; taking the "m"
 * pointer marks m for normalisation at the next update (the caller may write
 * through it). */
int mmx_sim_device_array(mmx_sim *sim, const char *name, float **dev);
/* The stream all of the sim's work is queued on (hipStream_t). */
int mmx_sim_stream(mmx_sim *sim, void **stream);
/* Quant.Update of H_eff (one effective-field evaluation), torque, B or
 * H_dex (demagexch; computed into the storage B's update uses).
 * Asynchronous: queues the kernels and returns. */
int mmx_sim_update(mmx_sim *sim, const char *name);
/* Engine.Step / Steps / Run with the loaded solver.  A solver that cannot
 * converge returns MMX_ERR_CONVERGE where the reference panics (am12 also
 * after 1000 step cuts within one step, where the reference retries
 * without bound). */
int mmx_sim_step(mmx_sim *sim);
int mmx_sim_steps(mmx_sim *sim, int nsteps);
int mmx_sim_run(mmx_sim *sim, double duration);
/* Times the passes of `reps` H_eff evaluations with HIP events on the sim's
 * stream (no reference counterpart; the benchmark's per-kernel clock).
 * ms[i] = average ms of pass i; *npass = 5 for 3-D grids (xf, yf, z, yi, xi),
 * 3 for films (xf, y2d, xi). */
int mmx_sim_time_passes(mmx_sim *sim, int reps, double *ms, int *npass);
/* Same-dt replay: one predictor/corrector try with the given dt, no step
 * control (no reference counterpart; used to compare m(t) step by step). */
int mmx_sim_try_step(mmx_sim *sim, double dt, double *err);

/* CRC-64/ISO as Go's hash/crc64 with crc64.ISO computes it, continuing
 * from crc (0 to start): the checksum of the #dump002 output format
 * (src/mumax/dump/writer.go:12,30-37,60-67).  Host memory, no GPU. */
unsigned long long mmx_crc64_iso(const void *data, unsigned long long n,
                                 unsigned long long crc);

/* ------------------------------------------------------------------------
 * Learned demag correction (no reference counterpart).  For a film of
 * 1 <= n0 <= 5 layers with n1, n2 multiples of 16: the exact demag field of
 * the grid coarsened 2x in the film plane, upsampled, plus a 4-layer 3x3
 * convolutional net (width
 * 32, BF16 matrix cores, FP32 accumulation) that predicts the fine-grid
 * residual from m and the coarse field.  Trained on the GPU against the exact
 * field.  Used by a simulation (mmx_sim_set_surrogate) for the RK12
 * corrector's field; the first stage stays exact and judges the surrogate
 * every check_every steps.
 * ---------------------------------------------------------------------- */
typedef struct mmx_surrogate mmx_surrogate;

/* For the grid, cell and Msat of `sim` (which must outlive it); weights
 * He-initialised from `seed`. */
int mmx_surrogate_create(mmx_surrogate **sur, mmx_sim *sim, unsigned long long seed);
int mmx_surrogate_destroy(mmx_surrogate *sur);
/* Parameters: per layer l = 0..3 the weights [3][3][Cin][Cout] then the
 * biases [Cout], channels 32, 32, 32, 32, 16; FP32. */
int mmx_surrogate_num_params(const mmx_surrogate *sur, int *count);
int mmx_surrogate_get_params(mmx_surrogate *sur, float *host, int count);
int mmx_surrogate_set_params(mmx_surrogate *sur, const float *host, int count);
/* Gradients of the last mmx_surrogate_train_step / mmx_surrogate_net. */
int mmx_surrogate_get_grads(mmx_surrogate *sur, float *host, int count);
/* One training step on the magnetisation m_dev (device, 3*N floats): exact
 * fine field as the target, forward, relative squared error of the residual
 * (sum d^2 / sum target^2), backward, Adam with learning rate lr (lr = 0:
 * gradients only).  *loss (may be NULL) = the loss before the update. */
int mmx_surrogate_train_step(mmx_surrogate *sur, const float *m_dev, float lr, double *loss);
/* The net alone on a host input x[n1*n2][32] (rounded to bf16): out[n1*n2][16]
 * (FP32); with a target tgt[n1*n2][3*n0] also the loss and the gradients. */
int mmx_surrogate_net(mmx_surrogate *sur, const float *x_host, const float *tgt_host,
                      float *out_host, double *loss);
/* The surrogate's B (tesla, device 3*N floats) at m_dev; coarse_only = 1
 * leaves the learned correction out (the method without the net). */
int mmx_surrogate_predict_b(mmx_surrogate *sur, const float *m_dev, float *b_dev, int coarse_only);
/* Use (sur != NULL) or stop using the surrogate in mmx_sim_step; tol: relative
 * L2 residual of the demag field above which the solver falls back to the
 * exact field (getv: surrogate_on, surrogate_residual, surrogate_checks,
 * surrogate_fallbacks, surrogate_evals). */
int mmx_sim_set_surrogate(mmx_sim *sim, mmx_surrogate *sur, double tol, int check_every);

/* ------------------------------------------------------------------------
 * Batches: many independent simulations of uneven sizes stepped together
 * (no reference counterpart; the reference runs one simulation per process,
 * Engine.Run engine.go:489-524).  One launch per demag pass covers every
 * simulation, the RK12 controller (solve_rk12.go:124-214) runs on the device
 * after each try, and tries are replayed from a hipGraph.  The batch borrows
 * the simulations: they must outlive it, and their settings are re-read at
 * every run/steps call; afterwards each simulation's state (m, t, dt, step,
 * ...) is the one its own mmx_sim_run would have produced, up to float32
 * rounding (the batch uses the any-length FFT passes).
 * ---------------------------------------------------------------------- */
typedef struct mmx_batch mmx_batch;

int mmx_batch_create(mmx_batch **batch, mmx_sim *const *sims, int nsim);
int mmx_batch_destroy(mmx_batch *batch);
/* Engine.Run / Steps per simulation: each one advances until t >= t_start +
 * duration (or has taken nsteps accepted steps). */
int mmx_batch_run(mmx_batch *batch, double duration);
int mmx_batch_steps(mmx_batch *batch, int nsteps);
/* Tries per graph replay (default 8); 0 = plain launches, no graph. */
int mmx_batch_set_graph(mmx_batch *batch, int tries_per_graph);
/* One persistent launch per `tries_per_launch` tries (0: off, the graph or
 * plain launches above): a cooperative launch whose workgroups run every pass
 * of every simulation and the step controller, separated by grid-wide
 * barriers, and stop early when no simulation is active. */
int mmx_batch_set_persistent(mmx_batch *batch, int tries_per_launch);
/* Autotuner: for each of the 5 passes (x forward, y forward or film y,
 * z, y inverse, x inverse) times the pass over the batch at LDS budgets of
 * 8..128 KB per workgroup (every simulation's tile shape follows from the
 * budget) and keeps the fastest; lds_kb[5] receives the choices (may be
 * NULL).  Does not change m or the solver state. */
int mmx_batch_tune(mmx_batch *batch, int reps, int *lds_kb);
/* Sets the per-pass LDS budgets (KB, 1..160) directly. */
int mmx_batch_set_lds(mmx_batch *batch, const int *lds_kb);
/* Workgroups per pass launch and dynamic LDS bytes per workgroup (5 each). */
int mmx_batch_info(const mmx_batch *batch, int *nblocks, int *lds_bytes);

/* ------------------------------------------------------------------------
 * Counter-based random numbers of temperature/brown.  No counterpart in the
 * reference, which draws its noise from cuRAND's default generator
 * (mod_temperature.go:60-106); here H_therm is a function of (Therm_seed,
 * draw, cell) through Philox4x32-10 (Salmon et al., SC 2011), and this is
 * that block function on the host: 4 counter words and 2 key words in, 4
 * words out.  H_therm of draw d (the d-th field drawn since Therm_seed was
 * set) in cell i uses counter (i lo, i hi, d lo, d hi) and key (seed lo,
 * seed hi); words 0/1 and 2/3 give Box-Muller pairs (FP64) for components
 * 0, 1 and 2. */
int mmx_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]);

#ifdef __cplusplus
}
#endif
#endif
