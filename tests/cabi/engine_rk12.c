/* engine_rk12.c -- the Go engine's field and solver sequence written against
 * the C ABI alone, as the cgo wrappers of mumax/gpu would drive it: the
 * caller allocates every device buffer (array.go:182-192), passes component
 * pointers and the stream, NULL for uniform maps, and checks each status.
 *
 *   H_eff = B/mu0 + H_ex + B_ext/mu0   (SumUpdater, update_sum.go:29-97;
 *                                        B from MaxwellPlan.update,
 *                                        maxwell_convolution.go:272-309)
 *   tau0  = torque(m, H_eff(m))        (mod_llg.go:59-70)
 *   m1    = normalize(m + h tau0)      (solve_rk12.go:97-123)
 *   tau1  = torque(m1, H_eff(m1))
 *   err   = MaxDiff(tau1, tau0) h      (solve_rk12.go:124-158)
 *   m     = normalize(m1 + h/2 (tau1 - tau0))
 *
 * with h = float32(dt gamma).  The same tries also run through the fused
 * simulation (mmx_sim_try_step) of the same library, for comparison.
 *
 * Input file (little-endian): int32 n[3]; float64 cell[3], Msat, Aex, alpha,
 * B_ext[3] (internal order), dt; int32 ntries; float32 m[3N] (SoA).
 * Output file: per try float64 err_engine, err_sim; then float32 m[3N] of the
 * engine sequence and float32 m[3N] of the fused simulation. */
#include <stdio.h>
#include <stdlib.h>

#include <hip/hip_runtime_api.h>

#include "mmx.h"

#define MU0 (4.0 * 3.14159265358979323846 * 1e-7)

static void die(const char *what, int rc)
{
    fprintf(stderr, "engine_rk12: %s failed (%d): %s\n", what, rc, mmx_last_error());
    exit(1);
}

#define CHECK(call)                             \
    do {                                        \
        int rc_ = (call);                       \
        if (rc_ != 0) die(#call, rc_);          \
    } while (0)

#define HCHECK(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "engine_rk12: %s: %s\n", #call, hipGetErrorString(e_)); \
            exit(1);                                                              \
        }                                                                         \
    } while (0)

static float *dev3(long N)
{
    void *p = NULL;
    HCHECK(hipMalloc(&p, sizeof(float) * 3 * (size_t)N));
    return (float *)p;
}

struct engine {
    int n[3];
    long N;
    float cellf[3];
    double msat, aex, alpha, bext[3];
    mmx_demag_plan *plan;
    mmx_reductor *red;
    hipStream_t st;
    float *b, *hex, *h;
};

/* H_eff at m into e->h */
static void heff(struct engine *e, const float *m)
{
    const long N = e->N;
    CHECK(mmx_demag_update_b(e->plan, e->b, e->b + N, e->b + 2 * N, m, m + N, m + 2 * N, NULL,
                             (float)e->msat, e->st));
    CHECK(mmx_exchange6(e->hex, e->hex + N, e->hex + 2 * N, m, m + N, m + 2 * N, NULL, NULL,
                        (float)((2 * e->aex) / (MU0 * e->msat)), e->n[0], e->n[1], e->n[2], 0, 0, 0,
                        e->cellf[0], e->cellf[1], e->cellf[2], e->st));
    const float inv_mu0 = (float)(1.0 / MU0);
    CHECK(mmx_vec_madd(e->h, e->h + N, e->h + 2 * N, e->hex, e->hex + N, e->hex + 2 * N, e->b,
                       e->b + N, e->b + 2 * N, inv_mu0, inv_mu0, inv_mu0, (int)N, e->st));
    CHECK(mmx_vec_madd(e->h, e->h + N, e->h + 2 * N, e->h, e->h + N, e->h + 2 * N, NULL, NULL,
                       NULL, (float)(e->bext[0] / MU0), (float)(e->bext[1] / MU0),
                       (float)(e->bext[2] / MU0), (int)N, e->st));
}

static void torque(struct engine *e, float *tau, const float *m)
{
    const long N = e->N;
    heff(e, m);
    CHECK(mmx_torque(tau, tau + N, tau + 2 * N, m, m + N, m + 2 * N, e->h, e->h + N, e->h + 2 * N,
                     NULL, (float)e->alpha, (int)N, e->st));
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: engine_rk12 input output\n");
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        perror(argv[1]);
        return 1;
    }
    struct engine e = {0};
    double cell[3], dt;
    int ntries;
    int ok = fread(e.n, sizeof(int), 3, f) == 3 && fread(cell, sizeof(double), 3, f) == 3 &&
             fread(&e.msat, sizeof(double), 1, f) == 1 && fread(&e.aex, sizeof(double), 1, f) == 1 &&
             fread(&e.alpha, sizeof(double), 1, f) == 1 && fread(e.bext, sizeof(double), 3, f) == 3 &&
             fread(&dt, sizeof(double), 1, f) == 1 && fread(&ntries, sizeof(int), 1, f) == 1;
    if (!ok || e.n[0] < 1 || e.n[1] < 1 || e.n[2] < 1 || ntries < 1 || ntries > 1000) {
        fprintf(stderr, "engine_rk12: bad header\n");
        return 1;
    }
    e.N = (long)e.n[0] * e.n[1] * e.n[2];
    const long N = e.N;
    float *m0 = (float *)malloc(sizeof(float) * 3 * N);
    float *out = (float *)malloc(sizeof(float) * 3 * N);
    double *errs = (double *)malloc(sizeof(double) * 2 * ntries);
    if (!m0 || !out || !errs || fread(m0, sizeof(float), 3 * N, f) != (size_t)(3 * N)) {
        fprintf(stderr, "engine_rk12: bad input\n");
        return 1;
    }
    fclose(f);
    for (int d = 0; d < 3; d++) e.cellf[d] = (float)cell[d];

    CHECK(mmx_init(0));
    HCHECK(hipStreamCreateWithFlags(&e.st, hipStreamNonBlocking));
    CHECK(mmx_demag_plan_create(&e.plan, e.n[0], e.n[1], e.n[2], cell[0], cell[1], cell[2], 0, 0, 0,
                                MMX_KERNEL_ARNE, 6));
    CHECK(mmx_reductor_create(&e.red, 3 * N));
    float *m = dev3(N), *m1 = dev3(N), *tau0 = dev3(N), *tau1 = dev3(N);
    e.b = dev3(N);
    e.hex = dev3(N);
    e.h = dev3(N);
    HCHECK(hipMemcpy(m, m0, sizeof(float) * 3 * N, hipMemcpyHostToDevice));

    /* the fused simulation on the same input */
    mmx_sim *sim = NULL;
    CHECK(mmx_sim_create(&sim, e.n[0], e.n[1], e.n[2], cell[0], cell[1], cell[2], 0, 0, 0));
    CHECK(mmx_sim_setv(sim, "Msat", &e.msat, 1));
    CHECK(mmx_sim_setv(sim, "Aex", &e.aex, 1));
    CHECK(mmx_sim_setv(sim, "alpha", &e.alpha, 1));
    CHECK(mmx_sim_setv(sim, "B_ext", e.bext, 3));
    CHECK(mmx_sim_setarray(sim, "m", m0, 3));
    double gamma = 0;
    CHECK(mmx_sim_getv(sim, "gamma", &gamma, 1));

    CHECK(mmx_normalize(m, m + N, m + 2 * N, NULL, (int)N, e.st));
    for (int k = 0; k < ntries; k++) {
        const double dtk = dt * (1 + 0.1 * k);
        const float h = (float)(dtk * gamma);
        torque(&e, tau0, m);
        CHECK(mmx_madd(m1, m, tau0, h, (int)(3 * N), e.st));
        CHECK(mmx_normalize(m1, m1 + N, m1 + 2 * N, NULL, (int)N, e.st));
        torque(&e, tau1, m1);
        float maxdiff = 0;
        CHECK(mmx_reductor_maxdiff(e.red, tau1, tau0, 3 * N, e.st, &maxdiff));
        errs[2 * k] = (double)(maxdiff * h);
        CHECK(mmx_madd2(m1, tau1, 0.5f * h, tau0, -0.5f * h, (int)(3 * N), e.st));
        CHECK(mmx_normalize(m1, m1 + N, m1 + 2 * N, NULL, (int)N, e.st));
        float *t = m;
        m = m1;
        m1 = t;
        CHECK(mmx_sim_try_step(sim, dtk, &errs[2 * k + 1]));
    }
    HCHECK(hipStreamSynchronize(e.st));

    FILE *g = fopen(argv[2], "wb");
    if (!g) {
        perror(argv[2]);
        return 1;
    }
    fwrite(errs, sizeof(double), 2 * ntries, g);
    HCHECK(hipMemcpy(out, m, sizeof(float) * 3 * N, hipMemcpyDeviceToHost));
    fwrite(out, sizeof(float), 3 * N, g);
    CHECK(mmx_sim_getarray(sim, "m", out, 3));
    fwrite(out, sizeof(float), 3 * N, g);
    fclose(g);

    CHECK(mmx_sim_destroy(sim));
    CHECK(mmx_reductor_destroy(e.red));
    CHECK(mmx_demag_plan_destroy(e.plan));
    HCHECK(hipFree(m));
    HCHECK(hipFree(m1));
    HCHECK(hipFree(tau0));
    HCHECK(hipFree(tau1));
    HCHECK(hipFree(e.b));
    HCHECK(hipFree(e.hex));
    HCHECK(hipFree(e.h));
    HCHECK(hipStreamDestroy(e.st));
    printf("engine_rk12: %d tries on %d x %d x %d done\n", ntries, e.n[0], e.n[1], e.n[2]);
    return 0;
}
