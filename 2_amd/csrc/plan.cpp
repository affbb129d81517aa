// plan.cpp -- demag plan construction and the plan-level C entry points.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "demag_plan.hpp"

namespace mmx {

static thread_local char g_err[512] = "no error";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

bool factor_radices(int L, int *radix, int *nstages)
{
    int n = 0;
    int rem = L;
    // powers of two: as many radix-8 stages as possible, then one 4 or 2
    while (rem % 8 == 0 && n < kMaxStages) { radix[n++] = 8; rem /= 8; }
    if (rem % 4 == 0 && n < kMaxStages) { radix[n++] = 4; rem /= 4; }
    if (rem % 2 == 0 && n < kMaxStages) { radix[n++] = 2; rem /= 2; }
    const int odd[3] = {3, 5, 7};
    for (int q = 0; q < 3; q++)
        while (rem % odd[q] == 0 && n < kMaxStages) { radix[n++] = odd[q]; rem /= odd[q]; }
    // any other prime: a stage of the direct-DFT kind (stockham_stage_any)
    for (int p = 11; rem > 1 && n < kMaxStages; p += 2) {
        if ((long)p * p > rem) p = rem; // rem itself is prime
        while (rem % p == 0 && n < kMaxStages) { radix[n++] = p; rem /= p; }
    }
    *nstages = n;
    return rem == 1;
}

// Host twiddles of one length (double, rounded once to float).
static void twiddles_f32(int L, std::vector<float2> &tw, std::vector<float2> &rtw)
{
    tw.resize(L);
    for (int t = 0; t < L; t++) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * t / (long double)L;
        tw[t] = float2{(float)cosl(a), (float)(-sinl(a))};
    }
    rtw.resize(L + 1);
    for (int k = 0; k <= L; k++) {
        const long double a = 3.14159265358979323846264338327950288L * k / (long double)L;
        rtw[k] = float2{(float)cosl(a), (float)(-sinl(a))};
    }
}

static int pick_width(int L, int ncomp, int maxw, size_t budget)
{
    int w = maxw;
    while (w > 1 && 2ull * ncomp * w * (L + 1) * sizeof(float2) > budget) w /= 2;
    return w;
}

static int plan_init(DemagPlanImpl *P, int n0, int n1, int n2, const double cell[3],
                     const int pbc[3], int kernel_type, int accuracy, bool keep_real)
{
    DemagGeom &g = P->g;
    MMX_REQUIRE(n0 > 0 && n1 > 0 && n2 > 0, MMX_ERR_ARG, "demag plan: bad size %d %d %d", n0, n1, n2);
    MMX_REQUIRE(cell[0] > 0 && cell[1] > 0 && cell[2] > 0, MMX_ERR_ARG, "demag plan: bad cell size");
    MMX_REQUIRE(accuracy > 0, MMX_ERR_ARG, "demag plan: accuracy must be > 0");
    g.n0 = n0; g.n1 = n1; g.n2 = n2;
    const int n[3] = {n0, n1, n2};
    int p[3];
    for (int i = 0; i < 3; i++) p[i] = (n[i] > 1 && pbc[i] == 0) ? 2 * n[i] : n[i];
    g.p0 = p[0]; g.p1 = p[1]; g.p2 = p[2];
    MMX_REQUIRE(g.p2 % 2 == 0 && g.p2 >= 2, MMX_ERR_SIZE,
                "demag plan: padded x size %d must be even (user x size > 1)", g.p2);
    MMX_REQUIRE(g.p1 >= 2, MMX_ERR_SIZE, "demag plan: user y size must be > 1");
    g.h2 = g.p2 / 2 + 1;
    g.Lx = g.p2 / 2;
    g.sx = (g.h2 + 15) / 16 * 16; // 128-byte aligned spectrum rows
    g.ko0 = g.p0 / 2 + 1;
    g.ko1 = g.p1 / 2 + 1;
    g.N = (long)n0 * n1 * n2;
    g.nrows = (long)n0 * n1;
    for (int i = 0; i < 3; i++) {
        P->cell[i] = cell[i];
        P->pbc[i] = pbc[i];
    }
    P->cell_kernel_type = kernel_type;
    P->accuracy = accuracy;
    P->fast_x = fast_length(g.Lx) && fast_x_addressable(g);
    // 3-D y passes take the longer odd lengths too; films run y2d
    P->fast_y = (g.p0 == 1) ? fast_length(g.p1) : fast_ylength(g.p1);
    P->fast_z = g.p0 > 1 && fast_length(g.p0) && fast_z_addressable(g);

    // FFT descriptors
    LineFFT<float> *fs[3] = {&P->fx, &P->fy, &P->fz};
    const int lens[3] = {g.Lx, g.p1, g.p0};
    std::vector<float2> tw[3], rtw[3];
    size_t off[7] = {0};
    for (int a = 0; a < 3; a++) {
        fs[a]->L = lens[a];
        MMX_REQUIRE(factor_radices(lens[a], fs[a]->radix, &fs[a]->nstages), MMX_ERR_SIZE,
                    "demag plan: FFT length %d has more than %d prime factors", lens[a],
                    kMaxStages);
        twiddles_f32(lens[a], tw[a], rtw[a]);
    }
    // layout: tw x | rtw x | tw y | tw z
    off[1] = tw[0].size();
    off[2] = off[1] + rtw[0].size();
    off[3] = off[2] + tw[1].size();
    off[4] = off[3] + tw[2].size();
    MMX_HIP(hipMalloc(&P->tw_mem, sizeof(float2) * off[4]));
    float2 *twd = (float2 *)P->tw_mem;
    MMX_HIP(hipMemcpy(twd + off[0], tw[0].data(), sizeof(float2) * tw[0].size(), hipMemcpyHostToDevice));
    MMX_HIP(hipMemcpy(twd + off[1], rtw[0].data(), sizeof(float2) * rtw[0].size(), hipMemcpyHostToDevice));
    MMX_HIP(hipMemcpy(twd + off[2], tw[1].data(), sizeof(float2) * tw[1].size(), hipMemcpyHostToDevice));
    MMX_HIP(hipMemcpy(twd + off[3], tw[2].data(), sizeof(float2) * tw[2].size(), hipMemcpyHostToDevice));
    P->fx.tw = twd + off[0];
    P->fx.rtw = twd + off[1];
    P->fy.tw = twd + off[2];
    P->fz.tw = twd + off[3];

    // tiles: keep a workgroup's LDS (two ping-pong buffers) at or under 80 KB
    const size_t budget = 80 * 1024;
    int rb = 16;
    while (rb > 1 && 2ull * 3 * rb * (g.Lx + 1) * sizeof(float2) > budget) rb--;
    // small grids: spread the rows over more workgroups
    while (rb > 1 && g.nrows / rb < 1024) rb /= 2;
    P->tiles.rows_x = rb;
    P->tiles.w_y = pick_width(g.p1, g.p0 == 1 ? 3 : 1, 16, budget);
    P->tiles.w_z = pick_width(g.p0, 3, 16, budget);
    // a workgroup can hold at most 160 KB of LDS on gfx950
    const size_t lds_max = 160 * 1024;
    const size_t need_x = 2ull * 3 * rb * (g.Lx + 1) * sizeof(float2);
    const size_t need_y = 2ull * (g.p0 == 1 ? 3 : 1) * P->tiles.w_y * (g.p1 + 1) * sizeof(float2);
    const size_t need_z = 2ull * 3 * P->tiles.w_z * (g.p0 + 1) * sizeof(float2);
    MMX_REQUIRE(need_x <= lds_max && need_y <= lds_max && need_z <= lds_max, MMX_ERR_SIZE,
                "demag plan: padded size %d x %d x %d needs more LDS per line than a workgroup has",
                g.p0, g.p1, g.p2);

    // spectrum buffers
    const size_t xbytes = sizeof(float2) * 3 * (size_t)g.nrows * g.sx;
    hipError_t e = hipMalloc(&P->xbuf, xbytes);
    MMX_REQUIRE(e == hipSuccess, MMX_ERR_NOMEM, "demag plan: cannot allocate %zu bytes", xbytes);
    if (g.p0 > 1) {
        const size_t ybytes = sizeof(float2) * 3 * (size_t)g.n0 * g.p1 * g.sx;
        e = hipMalloc(&P->ybuf, ybytes);
        MMX_REQUIRE(e == hipSuccess, MMX_ERR_NOMEM, "demag plan: cannot allocate %zu bytes", ybytes);
    }
    const size_t kbytes = sizeof(float) * 6 * (size_t)g.ko0 * g.ko1 * g.h2;
    e = hipMalloc(&P->kern, kbytes);
    MMX_REQUIRE(e == hipSuccess, MMX_ERR_NOMEM, "demag plan: cannot allocate %zu bytes", kbytes);
    return build_demag_kernel(P, keep_real);
}

static void plan_free(DemagPlanImpl *P)
{
    (void)hipFree(P->tw_mem);
    (void)hipFree(P->xbuf);
    (void)hipFree(P->ybuf);
    (void)hipFree(P->kern);
    (void)hipFree(P->real_kernel);
    P->tw_mem = nullptr;
    P->xbuf = P->ybuf = nullptr;
    P->kern = P->real_kernel = nullptr;
}

int plan_create(mmx_demag_plan **plan, int n0, int n1, int n2, const double cell[3],
                const int pbc[3], int kernel_type, int accuracy,
                const float *const *host_kernel)
{
    MMX_REQUIRE(plan, MMX_ERR_ARG, "demag plan: NULL output pointer");
    const bool keep = (kernel_type & MMX_KEEP_REAL_KERNEL) != 0;
    mmx_demag_plan *pl = new mmx_demag_plan();
    pl->impl.host_kernel = host_kernel;
    int rc = plan_init(&pl->impl, n0, n1, n2, cell, pbc, kernel_type & 0xff, accuracy, keep);
    pl->impl.host_kernel = nullptr; // the caller's arrays are not kept
    if (rc != MMX_OK) {
        plan_free(&pl->impl);
        delete pl;
        return rc;
    }
    *plan = pl;
    return MMX_OK;
}

void plan_destroy(mmx_demag_plan *plan)
{
    if (!plan) return;
    plan_free(&plan->impl);
    delete plan;
}

} // namespace mmx

using namespace mmx;

extern "C" {

const char *mmx_last_error(void) { return g_err; }

const char *mmx_version(void) { return "mmx 0.1 (gfx950)"; }

int mmx_init(int device)
{
    int n = 0;
    MMX_HIP(hipGetDeviceCount(&n));
    MMX_REQUIRE(device >= 0 && device < n, MMX_ERR_ARG, "mmx_init: device %d of %d", device, n);
    MMX_HIP(hipSetDevice(device));
    return noalloc_setup(device);
}

int mmx_demag_plan_create(mmx_demag_plan **plan, int n0, int n1, int n2, double cell0,
                          double cell1, double cell2, int pbc0, int pbc1, int pbc2,
                          int kernel_type, int accuracy)
{
    const double cell[3] = {cell0, cell1, cell2};
    const int pbc[3] = {pbc0, pbc1, pbc2};
    return plan_create(plan, n0, n1, n2, cell, pbc, kernel_type, accuracy, nullptr);
}

int mmx_demag_plan_destroy(mmx_demag_plan *plan)
{
    plan_destroy(plan);
    return MMX_OK;
}

int mmx_demag_update_b(mmx_demag_plan *plan, float *bx, float *by, float *bz, const float *mx,
                       const float *my, const float *mz, const float *msat_map, float msat_mul,
                       void *stream)
{
    MMX_REQUIRE(plan && bx && by && bz && mx && my && mz, MMX_ERR_ARG,
                "mmx_demag_update_b: NULL argument");
    DemagPlanImpl *P = &plan->impl;
    const long N = P->g.N;
    // the passes address a vector as base + c*N (the Go Array layout,
    // gpu/array.go:122-131)
    MMX_REQUIRE(my == mx + N && mz == mx + 2 * N && by == bx + N && bz == bx + 2 * N, MMX_ERR_ARG,
                "mmx_demag_update_b: components must be contiguous (c*N apart)");
    hipStream_t st = as_stream(stream);
    MMX_TRY(demag_forward(P, mx, msat_map, st));
    FieldParams fp{};
    fp.demag_scale = (float)(kMu0 * (double)msat_mul);
    EpilogueIO io{};
    io.hout = bx;
    return demag_inverse(P, EPI_B, fp, io, st);
}

int mmx_set_fast_x_min_blocks(int nblocks, int *previous)
{
    MMX_REQUIRE(nblocks >= 0, MMX_ERR_ARG, "mmx_set_fast_x_min_blocks: negative count %d", nblocks);
    const int old = set_fast_x_min_blocks(nblocks);
    if (previous) *previous = old;
    return MMX_OK;
}

int mmx_demag_plan_info(const mmx_demag_plan *plan, int p[3], int ko[3], double *bad_over_good)
{
    MMX_REQUIRE(plan, MMX_ERR_ARG, "mmx_demag_plan_info: NULL plan");
    const DemagGeom &g = plan->impl.g;
    if (p) { p[0] = g.p0; p[1] = g.p1; p[2] = g.p2; }
    if (ko) { ko[0] = g.ko0; ko[1] = g.ko1; ko[2] = g.h2; }
    if (bad_over_good) *bad_over_good = plan->impl.bad_over_good;
    return MMX_OK;
}

int mmx_demag_plan_kernel(const mmx_demag_plan *plan, float *host)
{
    MMX_REQUIRE(plan && host, MMX_ERR_ARG, "mmx_demag_plan_kernel: NULL argument");
    const DemagGeom &g = plan->impl.g;
    const size_t n = (size_t)g.ko0 * g.ko1 * g.h2;
    std::vector<float> aos(6 * n);
    MMX_HIP(hipMemcpy(aos.data(), plan->impl.kern, sizeof(float) * 6 * n, hipMemcpyDeviceToHost));
    for (size_t o = 0; o < n; o++)
        for (int c = 0; c < 6; c++) host[c * n + o] = aos[6 * o + c];
    return MMX_OK;
}

int mmx_demag_plan_real_kernel(const mmx_demag_plan *plan, float *host)
{
    MMX_REQUIRE(plan && host, MMX_ERR_ARG, "mmx_demag_plan_real_kernel: NULL argument");
    MMX_REQUIRE(plan->impl.real_kernel, MMX_ERR_STATE,
                "mmx_demag_plan_real_kernel: plan was created without MMX_KEEP_REAL_KERNEL");
    const DemagGeom &g = plan->impl.g;
    MMX_HIP(hipMemcpy(host, plan->impl.real_kernel,
                      sizeof(float) * 6 * (size_t)g.p0 * g.p1 * g.p2, hipMemcpyDeviceToHost));
    return MMX_OK;
}

int mmx_convplan_create(mmx_convplan **cp, const int dataSize[3], const int logicSize[3],
                        const float *const kernel[6])
{
    MMX_REQUIRE(cp && dataSize && logicSize && kernel, MMX_ERR_ARG,
                "mmx_convplan_create: NULL argument");
    int pbc[3];
    for (int a = 0; a < 3; a++) {
        const int n = dataSize[a], l = logicSize[a];
        MMX_REQUIRE(n > 0 && (l == n || (n > 1 && l == 2 * n)), MMX_ERR_SIZE,
                    "mmx_convplan_create: axis %d: kernel size %d for data size %d (must be the "
                    "data size or twice it)", a, l, n);
        pbc[a] = (l == n && n > 1) ? 1 : 0;
    }
    bool any = false;
    for (int c = 0; c < 6; c++) any = any || kernel[c] != nullptr;
    MMX_REQUIRE(any, MMX_ERR_ARG, "mmx_convplan_create: every kernel component is NULL");
    // the cell size only matters for a built kernel; a unit cell is passed
    const double cell[3] = {1.0, 1.0, 1.0};
    mmx_demag_plan *pl = nullptr;
    MMX_TRY(plan_create(&pl, dataSize[0], dataSize[1], dataSize[2], cell, pbc, MMX_KERNEL_ARNE, 6,
                        kernel));
    mmx_convplan *c = new mmx_convplan();
    c->plan = pl;
    *cp = c;
    return MMX_OK;
}

int mmx_convplan_destroy(mmx_convplan *cp)
{
    if (!cp) return MMX_OK;
    plan_destroy(cp->plan);
    delete cp;
    return MMX_OK;
}

int mmx_convplan_convolve(mmx_convplan *cp, const float *in, float *out, void *stream)
{
    MMX_REQUIRE(cp && cp->plan && in && out, MMX_ERR_ARG, "mmx_convplan_convolve: NULL argument");
    MMX_REQUIRE(in != out, MMX_ERR_ARG, "mmx_convplan_convolve: in and out must differ");
    DemagPlanImpl *P = &cp->plan->impl;
    const hipStream_t st = as_stream(stream);
    MMX_TRY(demag_forward(P, in, nullptr, st));
    FieldParams fp{};
    fp.demag_scale = 1.0f;
    EpilogueIO io{};
    io.hout = out;
    return demag_inverse(P, EPI_B, fp, io, st);
}

int mmx_convplan_info(const mmx_convplan *cp, double *bad_over_good)
{
    MMX_REQUIRE(cp && cp->plan && bad_over_good, MMX_ERR_ARG, "mmx_convplan_info: NULL argument");
    *bad_over_good = cp->plan->impl.bad_over_good;
    return MMX_OK;
}

} // extern "C"
