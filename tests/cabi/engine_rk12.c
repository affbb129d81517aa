/* engine_rk12.c -- the Go engine's field and solver sequence written against
 * the C ABI alone, as the cgo wrappers of mumax/gpu drive libmumax2: every
 * array is a float** of per-device pointers (one GPU: one element), streams a
 * hipStream_t*, the argument lists of src/libmumax2's headers
 * (include/mmx_libmumax2.h); the caller allocates every device buffer
 * (array.go:182-192), NULL is a uniform map, and the Reductor's partials are
 * finished on the host (reductor.go:19-212).
 *
 *   H_eff = B/mu0 + H_ex + B_ext/mu0   (SumUpdater, update_sum.go:29-97;
 *                                        B from MaxwellPlan.update,
 *                                        maxwell_convolution.go:272-309,
 *                                        replaced by mmx_demag_update_b)
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
#include "mmx_libmumax2.h"

#define MU0 (4.0 * 3.14159265358979323846 * 1e-7)

/* a one-device float** (the Go Array's pointer slice) */
#define DEV(p) ((float *[1]){(float *)(p)})

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
    hipStream_t st[1];
    float *b, *hex, *h;
    float *part;      /* the Reductor's device partials */
    float *hpart;
    int blocks;
};

/* H_eff at m into e->h */
static void heff(struct engine *e, const float *m)
{
    const long N = e->N;
    CHECK(mmx_demag_update_b(e->plan, e->b, e->b + N, e->b + 2 * N, m, m + N, m + 2 * N, NULL,
                             (float)e->msat, e->st[0]));
    mmx_exchange6Async(DEV(e->hex), DEV(e->hex + N), DEV(e->hex + 2 * N), DEV(m), DEV(m + N),
                       DEV(m + 2 * N), NULL, NULL, (float)((2 * e->aex) / (MU0 * e->msat)),
                       e->n[0], e->n[1], e->n[2], 0, 0, 0, e->cellf[0], e->cellf[1], e->cellf[2],
                       e->st);
    const float inv_mu0 = (float)(1.0 / MU0);
    mmx_vecMaddAsync(DEV(e->h), DEV(e->h + N), DEV(e->h + 2 * N), DEV(e->hex), DEV(e->hex + N),
                     DEV(e->hex + 2 * N), DEV(e->b), DEV(e->b + N), DEV(e->b + 2 * N), inv_mu0,
                     inv_mu0, inv_mu0, e->st, (int)N);
    mmx_vecMaddAsync(DEV(e->h), DEV(e->h + N), DEV(e->h + 2 * N), DEV(e->h), DEV(e->h + N),
                     DEV(e->h + 2 * N), NULL, NULL, NULL, (float)(e->bext[0] / MU0),
                     (float)(e->bext[1] / MU0), (float)(e->bext[2] / MU0), e->st, (int)N);
}

static void torque(struct engine *e, float *tau, const float *m)
{
    const long N = e->N;
    heff(e, m);
    mmx_torqueAsync(DEV(tau), DEV(tau + N), DEV(tau + 2 * N), DEV(m), DEV(m + N), DEV(m + 2 * N),
                    DEV(e->h), DEV(e->h + N), DEV(e->h + 2 * N), NULL, (float)e->alpha, e->st,
                    (int)N);
}

/* Reductor.MaxDiff: block partials on the device, the maximum on the host */
static float max_diff(struct engine *e, const float *a, const float *b, long n)
{
    mmx_partialMaxDiffAsync(DEV(a), DEV(b), DEV(e->part), e->blocks, 512, (int)n, e->st);
    HCHECK(hipMemcpyAsync(e->hpart, e->part, sizeof(float) * e->blocks, hipMemcpyDeviceToHost,
                          e->st[0]));
    HCHECK(hipStreamSynchronize(e->st[0]));
    float r = e->hpart[0];
    for (int i = 1; i < e->blocks; i++)
        if (e->hpart[i] > r) r = e->hpart[i];
    return r;
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

    int gpus[1] = {0};
    mmx_setUsedGPUs(gpus, 1);
    HCHECK(hipStreamCreateWithFlags(&e.st[0], hipStreamNonBlocking));
    CHECK(mmx_demag_plan_create(&e.plan, e.n[0], e.n[1], e.n[2], cell[0], cell[1], cell[2], 0, 0, 0,
                                MMX_KERNEL_ARNE, 6));
    e.blocks = (int)((3 * N + 1023) / 1024);
    if (e.blocks > 2048) e.blocks = 2048;
    HCHECK(hipMalloc((void **)&e.part, sizeof(float) * e.blocks));
    e.hpart = (float *)malloc(sizeof(float) * e.blocks);
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

    mmx_normalizeAsync(DEV(m), DEV(m + N), DEV(m + 2 * N), NULL, e.st, (int)N);
    for (int k = 0; k < ntries; k++) {
        const double dtk = dt * (1 + 0.1 * k);
        const float h = (float)(dtk * gamma);
        torque(&e, tau0, m);
        mmx_maddAsync(DEV(m1), DEV(m), DEV(tau0), h, e.st, (int)(3 * N));
        mmx_normalizeAsync(DEV(m1), DEV(m1 + N), DEV(m1 + 2 * N), NULL, e.st, (int)N);
        torque(&e, tau1, m1);
        const float maxdiff = max_diff(&e, tau1, tau0, 3 * N);
        errs[2 * k] = (double)(maxdiff * h);
        mmx_madd2Async(DEV(m1), DEV(tau1), 0.5f * h, DEV(tau0), -0.5f * h, e.st, (int)(3 * N));
        mmx_normalizeAsync(DEV(m1), DEV(m1 + N), DEV(m1 + 2 * N), NULL, e.st, (int)N);
        float *t = m;
        m = m1;
        m1 = t;
        CHECK(mmx_sim_try_step(sim, dtk, &errs[2 * k + 1]));
    }
    HCHECK(hipStreamSynchronize(e.st[0]));

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
    HCHECK(hipFree(e.part));
    free(e.hpart);
    CHECK(mmx_demag_plan_destroy(e.plan));
    HCHECK(hipFree(m));
    HCHECK(hipFree(m1));
    HCHECK(hipFree(tau0));
    HCHECK(hipFree(tau1));
    HCHECK(hipFree(e.b));
    HCHECK(hipFree(e.hex));
    HCHECK(hipFree(e.h));
    HCHECK(hipStreamDestroy(e.st[0]));
    printf("engine_rk12: %d tries on %d x %d x %d done\n", ntries, e.n[0], e.n[1], e.n[2]);
    return 0;
}
