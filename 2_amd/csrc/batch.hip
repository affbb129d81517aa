// batch.hip -- many independent simulations of uneven sizes stepped together.
//
// No reference counterpart: the reference runs one simulation per process
// (Engine.Run, engine.go:489-524).  Here a batch borrows existing mmx_sim
// objects and advances all of them with RK12 (solve_rk12.go:91-214) using
//
//  * one launch per demag pass for all simulations: a block -> simulation
//    table per pass; each block runs the same tile function as the
//    single-simulation generic kernels (demag_tiles.hpp), with that
//    simulation's geometry, FFT plan, buffers and field parameters, read from
//    a device table;
//  * step control on the device: a one-thread-per-simulation kernel applies
//    the RK12 controller after every predictor/corrector try and flips the
//    simulation's m double buffer on acceptance, so no host round trip is
//    needed between tries;
//  * a hipGraph holding K tries ([stage0 passes, predictor/corrector passes,
//    controller] x K), replayed until every simulation has reached its stop
//    time or step count (no memset is recorded: the controller clears the
//    error words itself);
//  * an autotuner that picks, per pass, the LDS budget from which every
//    simulation's tile shape (rows or columns per workgroup) follows, by
//    timing the pass over the whole batch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "demag_tiles.hpp"
#include "sim_impl.hpp"

namespace mmx {

enum BatchPass { BP_XF = 0, BP_Y = 1, BP_Z = 2, BP_YI = 3, BP_XI = 4, BP_N = 5 };

// Per-simulation descriptor, read-only during a launch.
struct BatchSim {
    DemagGeom g;
    LineFFT<float> fx, fy, fz;
    FieldParams fp;
    float2 *xbuf, *ybuf;
    const float *kern;
    float *mbuf[2];
    float *mpred, *tau0;
    const float *msat_map;
    int zsmall;       // z pass as one thread per column (z_column_small)
    int tile[BP_N];   // rows (XF, XI) or columns (Y, Z, YI) per workgroup
    int ntx[BP_N];    // column tiles per line set (Y, Z, YI)
    int first[BP_N];  // first block of this simulation in each pass
};

// Per-simulation RK12 state, owned by the device during a run.
struct BatchState {
    double t, dt, t0, t_stop, mindt, maxdt, maxerr, err, peakerr, gamma;
    long long step, step_stop, badsteps, evals;
    int cur, tries, need_stage0, active, error;
    float h;
    unsigned int err_bits;
    int pad_;
};

// ------------------------------------------------------------- kernels
//
// Each pass is a device function of one virtual block `vb` (a tile of one
// simulation, from the pass's block table), called by the per-pass kernels
// with vb = blockIdx.x and by the persistent kernel for vb = blockIdx.x,
// blockIdx.x + gridDim.x, ...

template <int STAGE>
__device__ __forceinline__ void bxf_block(const BatchSim *__restrict__ sims,
                                          const BatchState *__restrict__ st,
                                          const int *__restrict__ map, int vb, float2 *sm)
{
    const int s = map[vb];
    const BatchSim &S = sims[s];
    const BatchState &Z = st[s];
    if (!Z.active || (STAGE == 0 && !Z.need_stage0)) return;
    const long row0 = (long)(vb - S.first[BP_XF]) * S.tile[BP_XF];
    const float *m = S.mbuf[Z.cur];
    if (STAGE == 0)
        xf_tile<0>(S.g, S.fx, S.tile[BP_XF], row0, m, S.msat_map, nullptr, 0.0f, nullptr, nullptr,
                   S.xbuf, sm);
    else
        xf_tile<1>(S.g, S.fx, S.tile[BP_XF], row0, m, S.msat_map, S.tau0, Z.h, S.msat_map,
                   S.mpred, S.xbuf, sm);
}

template <int STAGE>
__device__ __forceinline__ void by_block(const BatchSim *__restrict__ sims,
                                         const BatchState *__restrict__ st,
                                         const int *__restrict__ map, int vb, float2 *sm)
{
    const int s = map[vb];
    const BatchSim &S = sims[s];
    const BatchState &Z = st[s];
    if (!Z.active || (STAGE == 0 && !Z.need_stage0)) return;
    const int b = vb - S.first[BP_Y], W = S.tile[BP_Y], ntx = S.ntx[BP_Y];
    if (S.g.p0 == 1) {
        y2d_tile(S.g, S.fy, W, b * W, S.kern, S.xbuf, sm);
    } else {
        const int kt = b % ntx, i = (b / ntx) % S.g.n0, c = b / (ntx * S.g.n0);
        yf_tile(S.g, S.fy, W, kt * W, i, c, S.xbuf, S.ybuf, sm);
    }
}

template <int STAGE>
__device__ __forceinline__ void bz_block(const BatchSim *__restrict__ sims,
                                         const BatchState *__restrict__ st,
                                         const int *__restrict__ map, int vb, float2 *sm)
{
    const int s = map[vb];
    const BatchSim &S = sims[s];
    const BatchState &Z = st[s];
    if (!Z.active || (STAGE == 0 && !Z.need_stage0)) return;
    const int b = vb - S.first[BP_Z], W = S.tile[BP_Z], ntx = S.ntx[BP_Z];
    if (S.zsmall) {
        const long col = (long)b * kBlock + threadIdx.x;
        if (col >= (long)S.g.p1 * S.g.h2) return;
        const int ky = (int)(col / S.g.h2), kx = (int)(col - (long)ky * S.g.h2);
        z_column_small_any(S.g, S.kern, S.ybuf, ky, kx);
        return;
    }
    z_tile(S.g, S.fz, W, (b % ntx) * W, b / ntx, S.kern, S.ybuf, sm);
}

template <int STAGE>
__device__ __forceinline__ void byi_block(const BatchSim *__restrict__ sims,
                                          const BatchState *__restrict__ st,
                                          const int *__restrict__ map, int vb, float2 *sm)
{
    const int s = map[vb];
    const BatchSim &S = sims[s];
    const BatchState &Z = st[s];
    if (!Z.active || (STAGE == 0 && !Z.need_stage0)) return;
    const int b = vb - S.first[BP_YI], W = S.tile[BP_YI], ntx = S.ntx[BP_YI];
    const int kt = b % ntx, i = (b / ntx) % S.g.n0, c = b / (ntx * S.g.n0);
    yi_tile(S.g, S.fy, W, kt * W, i, c, S.ybuf, S.xbuf, sm);
}

// STAGE 0: tau0 = torque(m) (solve_rk12.go:97-115); STAGE 1: corrector,
// normalised m into the other buffer, max |tau1 - tau0| into err_bits.
template <int STAGE>
__device__ __forceinline__ void bxi_block(const BatchSim *__restrict__ sims,
                                          BatchState *__restrict__ st,
                                          const int *__restrict__ map, int vb, float2 *sm)
{
    const int s = map[vb];
    const BatchSim &S = sims[s];
    BatchState &Z = st[s];
    if (!Z.active || (STAGE == 0 && !Z.need_stage0)) return;
    const long row0 = (long)(vb - S.first[BP_XI]) * S.tile[BP_XI];
    EpilogueIO io{};
    if (STAGE == 0) {
        io.m = S.mbuf[Z.cur];
        io.tout = S.tau0;
        xi_tile<EPI_TAU0>(S.g, S.fx, S.tile[BP_XI], row0, S.xbuf, S.fp, io, sm);
    } else {
        io.m = S.mpred;
        io.tau0 = S.tau0;
        io.mnew = S.mbuf[Z.cur ^ 1];
        io.h = Z.h;
        io.err = &Z.err_bits;
        xi_tile<EPI_CORR>(S.g, S.fx, S.tile[BP_XI], row0, S.xbuf, S.fp, io, sm);
    }
}

#define MMX_BATCH_KERNEL(NAME, BODY, STATE_T)                                                \
    template <int STAGE>                                                                     \
    __global__ void __launch_bounds__(kBlock)                                                \
    NAME(const BatchSim *__restrict__ sims, STATE_T *__restrict__ st,                        \
         const int *__restrict__ map)                                                        \
    {                                                                                        \
        extern __shared__ float2 sm[];                                                       \
        BODY<STAGE>(sims, st, map, blockIdx.x, sm);                                          \
    }
MMX_BATCH_KERNEL(bxf_kernel, bxf_block, const BatchState)
MMX_BATCH_KERNEL(by_kernel, by_block, const BatchState)
MMX_BATCH_KERNEL(bz_kernel, bz_block, const BatchState)
MMX_BATCH_KERNEL(byi_kernel, byi_block, const BatchState)
MMX_BATCH_KERNEL(bxi_kernel, bxi_block, BatchState)
#undef MMX_BATCH_KERNEL

// RK12Solver.Step's controller (solve_rk12.go:124-214) after one try, for one
// simulation; mirrors the host loop of mmx_sim_step.
__device__ __forceinline__ void control_one(BatchState &Z)
{
    if (!Z.active) return;
    const int maxTry = 10;
    const double headRoom = 0.8;
    const float maxdiff = __uint_as_float(Z.err_bits);
    Z.err_bits = 0u;
    Z.evals += Z.need_stage0 ? 2 : 1;
    if (!isfinite(maxdiff)) {
        Z.error = MMX_ERR_NAN;
        Z.active = 0;
        return;
    }
    const float stepDiff = maxdiff * Z.h;
    const double err = (double)stepDiff;
    const double dt = Z.dt;
    Z.t = Z.t0 + dt;
    Z.err = err;
    bool bad = false;
    if (err > Z.maxerr) {
        Z.badsteps++;
        bad = true;
    }
    if ((!bad || Z.tries == maxTry - 1) && err > Z.peakerr) Z.peakerr = err;
    double factor = bad ? pow(Z.maxerr / err, 1.0 / 3.0) : sqrt(Z.maxerr / err);
    factor *= headRoom;
    if (factor > 1.5) factor = 1.5;
    if (factor < 0.1) factor = 0.1;
    const double minFactor = factor < 2.0 ? factor : 2.0;
    double newDt = dt * minFactor;
    if (newDt < Z.mindt) newDt = Z.mindt;
    if (newDt > Z.maxdt) newDt = Z.maxdt;
    Z.dt = newDt;
    if (!bad || newDt == Z.mindt) {
        Z.cur ^= 1;
        Z.step++;
        Z.tries = 0;
        Z.need_stage0 = 1;
        Z.t0 = Z.t;
        if (Z.t >= Z.t_stop || Z.step >= Z.step_stop) Z.active = 0;
    } else {
        if (Z.tries > maxTry) {
            Z.error = MMX_ERR_CONVERGE;
            Z.active = 0;
            return;
        }
        Z.tries++;
        Z.need_stage0 = 0;
    }
    Z.h = (float)(Z.dt * Z.gamma);
}

__global__ void __launch_bounds__(64) bcontrol_kernel(BatchState *__restrict__ st, int nsim)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsim) control_one(st[s]);
}

// ---------------------------------------------------- the persistent launch
//
// One cooperative launch (every workgroup resident) runs up to `tries` tries
// of the whole batch: the ten passes of a try and the controller one after
// another, separated by grid-wide barriers, each pass's virtual blocks dealt
// round-robin to the workgroups.  It stops early once no simulation is
// active.  A barrier wait is bounded: past kSpinLimit polls the workgroup
// records the time-out in `fail` and leaves; a timed-out barrier leaves the
// barrier words inconsistent, so the host resets them before the next launch.

struct PersistArgs {
    const int *map[BP_N];
    int nblk[BP_N];
    int nsim;
    int tries;
    unsigned int *bar;   // [0]: arrivals, [1]: generation
    unsigned int *fail;  // nonzero: a barrier wait timed out
    int *nactive;        // active simulations after the last controller run
};

constexpr long kSpinLimit = 1L << 25;

// Grid-wide barrier with agent-scope release/acquire (the workgroups of the
// 8 XCDs do not share an L2).  Returns false when the wait timed out.
__device__ bool grid_barrier(const PersistArgs &a)
{
    __shared__ int ok;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int gen =
            __hip_atomic_load(&a.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned int arrived =
            __hip_atomic_fetch_add(&a.bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        ok = 1;
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(&a.bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&a.bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long n = 0;
            while (__hip_atomic_load(&a.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) ==
                   gen) {
                __builtin_amdgcn_s_sleep(2);
                if (++n > kSpinLimit) {
                    __hip_atomic_store(a.fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
    }
    __syncthreads();
    __threadfence();
    return ok != 0;
}

template <int STAGE>
__device__ bool persist_passes(const BatchSim *__restrict__ sims, BatchState *__restrict__ st,
                               const PersistArgs &a, float2 *sm)
{
#define MMX_PASS(P, BODY)                                                                    \
    if (a.nblk[P] > 0) {                                                                     \
        for (int vb = blockIdx.x; vb < a.nblk[P]; vb += gridDim.x) {                         \
            BODY<STAGE>(sims, st, a.map[P], vb, sm);                                         \
            __syncthreads(); /* the next tile reuses the LDS */                              \
        }                                                                                    \
        if (!grid_barrier(a)) return false;                                                  \
    }
    MMX_PASS(BP_XF, bxf_block)
    MMX_PASS(BP_Y, by_block)
    MMX_PASS(BP_Z, bz_block)
    MMX_PASS(BP_YI, byi_block)
    MMX_PASS(BP_XI, bxi_block)
#undef MMX_PASS
    return true;
}

__global__ void __launch_bounds__(kBlock)
bpersist_kernel(const BatchSim *__restrict__ sims, BatchState *__restrict__ st, PersistArgs a)
{
    extern __shared__ float2 sm[];
    for (int k = 0; k < a.tries; k++) {
        if (!persist_passes<0>(sims, st, a, sm)) return;
        if (!persist_passes<1>(sims, st, a, sm)) return;
        if (blockIdx.x == 0) {
            for (int s = threadIdx.x; s < a.nsim; s += blockDim.x) control_one(st[s]);
            __syncthreads();
            if (threadIdx.x == 0) {
                int n = 0;
                for (int s = 0; s < a.nsim; s++) n += st[s].active;
                __hip_atomic_store(a.nactive, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (!grid_barrier(a)) return;
        if (__hip_atomic_load(a.nactive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    }
}

// ------------------------------------------------------------ host side

struct BatchImpl {
    std::vector<mmx_sim *> sims;
    std::vector<BatchSim> hs;
    std::vector<BatchState> hst;
    BatchSim *dsims = nullptr;
    BatchState *dst = nullptr;
    int *dmap[BP_N] = {};
    int nblk[BP_N] = {};
    size_t lds[BP_N] = {};
    size_t cap[BP_N] = {};      // LDS budget per pass (tuner's choice)
    hipStream_t st = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_tries = 0;        // tries recorded in the graph (0: none yet)
    int tries_per_graph = 8;    // 0: eager launches
    int persist_tries = 0;      // > 0: one persistent launch per this many tries
    unsigned int *dpersist = nullptr; // barrier words, time-out flag, active count
    bool tables_dirty = true;
};

} // namespace mmx

struct mmx_batch {
    mmx::BatchImpl impl;
};

namespace mmx {
namespace {

// LDS bytes a pass needs for tile parameter v of simulation g
size_t pass_lds(int pass, const DemagGeom &g, int v)
{
    switch (pass) {
    case BP_XF:
    case BP_XI: return 2ull * 3 * v * (g.Lx + 1) * sizeof(float2);
    case BP_Y: return 2ull * (g.p0 == 1 ? 3 : 1) * v * (g.p1 + 1) * sizeof(float2);
    case BP_Z: return 2ull * 3 * v * (g.p0 + 1) * sizeof(float2);
    default: return 2ull * v * (g.p1 + 1) * sizeof(float2);
    }
}

bool pass_applies(int pass, const DemagGeom &g)
{
    return !(g.p0 == 1 && (pass == BP_Z || pass == BP_YI));
}

// largest tile parameter (<= 16) whose LDS fits the budget, at least 1
int pick_tile(int pass, const DemagGeom &g, size_t budget)
{
    int v = 16;
    while (v > 1 && pass_lds(pass, g, v) > budget) v--;
    return v;
}

int pass_blocks(int pass, const DemagGeom &g, int v, int *ntx)
{
    *ntx = 0;
    switch (pass) {
    case BP_XF:
    case BP_XI: return (int)div_up(g.nrows, v);
    case BP_Y:
        *ntx = (int)div_up(g.h2, v);
        return g.p0 == 1 ? *ntx : *ntx * g.n0 * 3;
    case BP_Z:
        *ntx = (int)div_up(g.h2, v);
        return *ntx * g.p1;
    default:
        *ntx = (int)div_up(g.h2, v);
        return *ntx * g.n0 * 3;
    }
}

template <typename K> void allow_lds_batch(K kernel, size_t bytes)
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

void set_attributes(const BatchImpl &B)
{
    allow_lds_batch(bxf_kernel<0>, B.lds[BP_XF]);
    allow_lds_batch(bxf_kernel<1>, B.lds[BP_XF]);
    allow_lds_batch(by_kernel<0>, B.lds[BP_Y]);
    allow_lds_batch(by_kernel<1>, B.lds[BP_Y]);
    allow_lds_batch(bz_kernel<0>, B.lds[BP_Z]);
    allow_lds_batch(bz_kernel<1>, B.lds[BP_Z]);
    allow_lds_batch(byi_kernel<0>, B.lds[BP_YI]);
    allow_lds_batch(byi_kernel<1>, B.lds[BP_YI]);
    allow_lds_batch(bxi_kernel<0>, B.lds[BP_XI]);
    allow_lds_batch(bxi_kernel<1>, B.lds[BP_XI]);
}

void drop_graph(BatchImpl &B)
{
    if (B.gexec) (void)hipGraphExecDestroy(B.gexec);
    if (B.graph) (void)hipGraphDestroy(B.graph);
    B.gexec = nullptr;
    B.graph = nullptr;
    B.graph_tries = 0;
}

// Tile shapes and block tables from the per-pass LDS budgets.
int build_tables(BatchImpl &B)
{
    const int ns = (int)B.sims.size();
    std::vector<int> maps[BP_N];
    for (int p = 0; p < BP_N; p++) {
        B.lds[p] = 0;
        for (int s = 0; s < ns; s++) {
            BatchSim &S = B.hs[s];
            S.first[p] = (int)maps[p].size();
            S.tile[p] = 1;
            S.ntx[p] = 0;
            if (!pass_applies(p, S.g)) continue;
            if (p == BP_Z && S.zsmall) {
                const int nb = (int)div_up((long)S.g.p1 * S.g.h2, kBlock);
                maps[p].insert(maps[p].end(), (size_t)nb, s);
                continue;
            }
            const int v = pick_tile(p, S.g, B.cap[p]);
            S.tile[p] = v;
            const int nb = pass_blocks(p, S.g, v, &S.ntx[p]);
            maps[p].insert(maps[p].end(), (size_t)nb, s);
            B.lds[p] = std::max(B.lds[p], pass_lds(p, S.g, v));
        }
        MMX_REQUIRE(B.lds[p] <= 160 * 1024, MMX_ERR_SIZE,
                    "batch: pass %d needs %zu bytes of LDS per workgroup", p, B.lds[p]);
        B.nblk[p] = (int)maps[p].size();
        (void)hipFree(B.dmap[p]);
        B.dmap[p] = nullptr;
        if (B.nblk[p] > 0) {
            MMX_HIP(hipMalloc(&B.dmap[p], sizeof(int) * B.nblk[p]));
            MMX_HIP(hipMemcpy(B.dmap[p], maps[p].data(), sizeof(int) * B.nblk[p],
                              hipMemcpyHostToDevice));
        }
    }
    set_attributes(B);
    drop_graph(B);
    B.tables_dirty = false;
    return MMX_OK;
}

// Descriptor and state of every simulation from its current host settings.
int load_sims(BatchImpl &B, double duration, long long nsteps)
{
    const int ns = (int)B.sims.size();
    for (int s = 0; s < ns; s++) {
        mmx_sim *x = B.sims[s];
        MMX_REQUIRE(x->solver == 0, MMX_ERR_ARG,
                    "batch: simulation %d has another solver than solver/rk12 loaded", s);
        MMX_REQUIRE(!x->thermal, MMX_ERR_ARG,
                    "batch: simulation %d has temperature/brown loaded", s);
        MMX_REQUIRE(!x->dex, MMX_ERR_ARG, "batch: simulation %d has demagexch loaded", s);
        x->spec_valid = false; // the batch advances m, t, dt itself
        MMX_TRY(sim_prepare(x));
        MMX_HIP(hipStreamSynchronize(x->st));
        const DemagPlanImpl &P = x->plan->impl;
        BatchSim &S = B.hs[s];
        S.g = P.g;
        S.fx = P.fx;
        S.fy = P.fy;
        S.fz = P.fz;
        S.fp = sim_field_params(x);
        S.xbuf = P.xbuf;
        S.ybuf = P.ybuf;
        S.kern = P.kern;
        S.mbuf[0] = x->m;
        S.mbuf[1] = x->mnew;
        S.mpred = x->mpred;
        S.tau0 = x->tau0;
        S.msat_map = x->msat_map;
        S.zsmall = z_small_applies(P.g) ? 1 : 0;
        BatchState &Z = B.hst[s];
        std::memset(&Z, 0, sizeof(Z));
        Z.t = Z.t0 = x->t;
        Z.dt = x->dt;
        Z.mindt = x->mindt;
        Z.maxdt = x->maxdt;
        Z.maxerr = x->maxerr;
        Z.err = x->err;
        Z.peakerr = x->peakerr;
        Z.gamma = x->gamma;
        Z.step = x->step;
        Z.badsteps = x->badsteps;
        Z.evals = x->evals;
        Z.t_stop = duration >= 0 ? x->t + duration : 1e300;
        Z.step_stop = nsteps >= 0 ? x->step + nsteps : (1LL << 62);
        Z.cur = 0;
        Z.need_stage0 = 1;
        Z.active = (Z.t < Z.t_stop && Z.step < Z.step_stop) ? 1 : 0;
        Z.h = (float)(Z.dt * Z.gamma);
    }
    if (B.tables_dirty) MMX_TRY(build_tables(B));
    MMX_HIP(hipMemcpyAsync(B.dsims, B.hs.data(), sizeof(BatchSim) * ns, hipMemcpyHostToDevice,
                           B.st));
    MMX_HIP(hipMemcpyAsync(B.dst, B.hst.data(), sizeof(BatchState) * ns, hipMemcpyHostToDevice,
                           B.st));
    return MMX_OK;
}

template <typename K>
void launch(K kernel, int nblk, size_t lds, hipStream_t st, const BatchSim *sims, BatchState *bst,
            const int *map)
{
    if (nblk > 0) kernel<<<dim3(nblk), dim3(kBlock), lds, st>>>(sims, bst, map);
}

// one evaluation of stage STAGE over the batch: the five passes
template <int STAGE> int eval_passes(BatchImpl &B, int only_pass = -1)
{
    auto want = [&](int p) { return only_pass < 0 || only_pass == p; };
    if (want(BP_XF)) launch(bxf_kernel<STAGE>, B.nblk[BP_XF], B.lds[BP_XF], B.st, B.dsims, B.dst, B.dmap[BP_XF]);
    if (want(BP_Y)) launch(by_kernel<STAGE>, B.nblk[BP_Y], B.lds[BP_Y], B.st, B.dsims, B.dst, B.dmap[BP_Y]);
    if (want(BP_Z)) launch(bz_kernel<STAGE>, B.nblk[BP_Z], B.lds[BP_Z], B.st, B.dsims, B.dst, B.dmap[BP_Z]);
    if (want(BP_YI)) launch(byi_kernel<STAGE>, B.nblk[BP_YI], B.lds[BP_YI], B.st, B.dsims, B.dst, B.dmap[BP_YI]);
    if (want(BP_XI)) launch(bxi_kernel<STAGE>, B.nblk[BP_XI], B.lds[BP_XI], B.st, B.dsims, B.dst, B.dmap[BP_XI]);
    if (hipGetLastError() != hipSuccess) {
        set_error("batch pass launch failed");
        return MMX_ERR_HIP;
    }
    return MMX_OK;
}

int queue_try(BatchImpl &B)
{
    MMX_TRY(eval_passes<0>(B));
    MMX_TRY(eval_passes<1>(B));
    const int ns = (int)B.sims.size();
    bcontrol_kernel<<<dim3(div_up(ns, 64)), dim3(64), 0, B.st>>>(B.dst, ns);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int ensure_graph(BatchImpl &B)
{
    if (B.gexec && B.graph_tries == B.tries_per_graph) return MMX_OK;
    drop_graph(B);
    MMX_HIP(hipStreamBeginCapture(B.st, hipStreamCaptureModeThreadLocal));
    int rc = MMX_OK;
    for (int k = 0; k < B.tries_per_graph && rc == MMX_OK; k++) rc = queue_try(B);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(B.st, &g);
    if (rc != MMX_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) {
        set_error("batch: graph capture failed (%s)", hipGetErrorString(e));
        return MMX_ERR_HIP;
    }
    B.graph = g;
    MMX_HIP(hipGraphInstantiate(&B.gexec, B.graph, nullptr, nullptr, 0));
    B.graph_tries = B.tries_per_graph;
    return MMX_OK;
}

// Up to B.persist_tries tries of the whole batch in one cooperative launch
// (bpersist_kernel); the grid is what can be resident at once, at most the
// largest pass's block count.
int persist_launch(BatchImpl &B)
{
    size_t lds = 0;
    int maxblk = 1;
    PersistArgs a{};
    for (int p = 0; p < BP_N; p++) {
        lds = std::max(lds, B.lds[p]);
        maxblk = std::max(maxblk, B.nblk[p]);
        a.map[p] = B.dmap[p];
        a.nblk[p] = B.nblk[p];
    }
    a.nsim = (int)B.sims.size();
    a.tries = B.persist_tries;
    a.bar = B.dpersist;
    a.fail = B.dpersist + 2;
    a.nactive = reinterpret_cast<int *>(B.dpersist + 3);
    allow_lds_batch(bpersist_kernel, lds);
    int dev = 0, coop = 0, ncu = 0, per_cu = 0;
    MMX_HIP(hipGetDevice(&dev));
    MMX_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    MMX_REQUIRE(coop, MMX_ERR_STATE, "batch: the device has no cooperative launch");
    MMX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    MMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bpersist_kernel, kBlock, lds));
    MMX_REQUIRE(per_cu > 0, MMX_ERR_SIZE, "batch: the persistent kernel does not fit a CU (%zu B LDS)",
                lds);
    const int grid = std::max(1, std::min(per_cu * ncu, maxblk));
    // barrier words, time-out flag and active count start at zero (a timed-out
    // launch may leave them inconsistent)
    MMX_HIP(hipMemsetAsync(B.dpersist, 0, 4 * sizeof(unsigned int), B.st));
    const BatchSim *sims = B.dsims;
    BatchState *st = B.dst;
    void *args[] = {&sims, &st, &a};
    MMX_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(bpersist_kernel),
                                       dim3(grid), dim3(kBlock), args, (unsigned int)lds, B.st));
    unsigned int h[4];
    MMX_HIP(hipMemcpyAsync(h, B.dpersist, sizeof(h), hipMemcpyDeviceToHost, B.st));
    MMX_HIP(hipStreamSynchronize(B.st));
    MMX_REQUIRE(h[2] == 0, MMX_ERR_HIP, "batch: a grid barrier of the persistent launch timed out");
    return MMX_OK;
}

// state back to the host; optionally into the simulations
int fetch_state(BatchImpl &B, bool into_sims, int *nactive)
{
    const int ns = (int)B.sims.size();
    MMX_HIP(hipMemcpyAsync(B.hst.data(), B.dst, sizeof(BatchState) * ns, hipMemcpyDeviceToHost,
                           B.st));
    MMX_HIP(hipStreamSynchronize(B.st));
    int act = 0, err_sim = -1, err_code = 0;
    for (int s = 0; s < ns; s++) {
        act += B.hst[s].active;
        if (B.hst[s].error && err_sim < 0) {
            err_sim = s;
            err_code = B.hst[s].error;
        }
    }
    if (nactive) *nactive = act;
    if (into_sims) {
        for (int s = 0; s < ns; s++) {
            mmx_sim *x = B.sims[s];
            const BatchState &Z = B.hst[s];
            x->t = Z.t;
            x->dt = Z.dt;
            x->err = Z.err;
            x->peakerr = Z.peakerr;
            x->step = (long)Z.step;
            x->badsteps = (long)Z.badsteps;
            x->evals = (long)Z.evals;
            if (Z.cur == 1) std::swap(x->m, x->mnew);
            x->m_normalized = true;
        }
    }
    if (err_sim >= 0) {
        if (err_code == MMX_ERR_NAN)
            set_error("batch: NaN or Inf in the torque of simulation %d (t = %g)", err_sim,
                      B.hst[err_sim].t);
        else
            set_error("batch: simulation %d cannot converge after 10 bad steps", err_sim);
        return err_code;
    }
    return MMX_OK;
}

int batch_advance(BatchImpl &B, double duration, long long nsteps)
{
    MMX_TRY(load_sims(B, duration, nsteps));
    const int ns = (int)B.sims.size();
    int nactive = 0;
    for (int s = 0; s < ns; s++) nactive += B.hst[s].active;
    const long long guard = (nsteps >= 0 ? (nsteps + 1) * 11 : (1LL << 40));
    long long tries = 0;
    int rc = MMX_OK;
    while (nactive > 0 && rc == MMX_OK) {
        if (B.persist_tries > 0) {
            rc = persist_launch(B);
            tries += B.persist_tries;
        } else if (B.tries_per_graph > 0) {
            rc = ensure_graph(B);
            if (rc == MMX_OK && hipGraphLaunch(B.gexec, B.st) != hipSuccess) {
                set_error("batch: graph launch failed");
                rc = MMX_ERR_HIP;
            }
            tries += B.tries_per_graph;
        } else {
            rc = queue_try(B);
            tries++;
        }
        if (rc == MMX_OK) rc = fetch_state(B, false, &nactive);
        if (rc == MMX_OK && tries > guard) {
            set_error("batch: step budget exceeded");
            rc = MMX_ERR_STATE;
        }
    }
    const int rc2 = fetch_state(B, true, nullptr);
    return rc != MMX_OK ? rc : rc2;
}

} // namespace
} // namespace mmx

using namespace mmx;

extern "C" {

int mmx_batch_create(mmx_batch **batch, mmx_sim *const *sims, int nsim)
{
    MMX_REQUIRE(batch && sims && nsim > 0, MMX_ERR_ARG, "mmx_batch_create: bad arguments");
    for (int s = 0; s < nsim; s++)
        MMX_REQUIRE(sims[s], MMX_ERR_ARG, "mmx_batch_create: simulation %d is NULL", s);
    mmx_batch *b = new mmx_batch();
    BatchImpl &B = b->impl;
    B.sims.assign(sims, sims + nsim);
    B.hs.assign(nsim, BatchSim{});
    B.hst.assign(nsim, BatchState{});
    for (int p = 0; p < BP_N; p++) B.cap[p] = 32 * 1024;
    int rc = MMX_OK;
    if (hipMalloc(&B.dsims, sizeof(BatchSim) * nsim) != hipSuccess ||
        hipMalloc(&B.dst, sizeof(BatchState) * nsim) != hipSuccess ||
        hipMalloc(&B.dpersist, 4 * sizeof(unsigned int)) != hipSuccess ||
        hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking) != hipSuccess) {
        set_error("mmx_batch_create: cannot allocate the batch tables");
        rc = MMX_ERR_NOMEM;
    }
    if (rc == MMX_OK) rc = load_sims(B, 0.0, 0);
    if (rc != MMX_OK) {
        mmx_batch_destroy(b);
        return rc;
    }
    *batch = b;
    return MMX_OK;
}

int mmx_batch_destroy(mmx_batch *batch)
{
    if (!batch) return MMX_OK;
    BatchImpl &B = batch->impl;
    if (B.st) (void)hipStreamSynchronize(B.st);
    drop_graph(B);
    for (int p = 0; p < BP_N; p++) (void)hipFree(B.dmap[p]);
    (void)hipFree(B.dsims);
    (void)hipFree(B.dst);
    (void)hipFree(B.dpersist);
    if (B.st) (void)hipStreamDestroy(B.st);
    delete batch;
    return MMX_OK;
}

int mmx_batch_set_graph(mmx_batch *batch, int tries_per_graph)
{
    MMX_REQUIRE(batch && tries_per_graph >= 0 && tries_per_graph <= 256, MMX_ERR_ARG,
                "mmx_batch_set_graph: tries per graph must be 0..256");
    batch->impl.tries_per_graph = tries_per_graph;
    drop_graph(batch->impl);
    return MMX_OK;
}

int mmx_batch_set_persistent(mmx_batch *batch, int tries_per_launch)
{
    MMX_REQUIRE(batch && tries_per_launch >= 0 && tries_per_launch <= 4096, MMX_ERR_ARG,
                "mmx_batch_set_persistent: tries per launch must be 0..4096");
    batch->impl.persist_tries = tries_per_launch;
    return MMX_OK;
}

int mmx_batch_run(mmx_batch *batch, double duration)
{
    MMX_REQUIRE(batch && duration >= 0, MMX_ERR_ARG, "mmx_batch_run: bad arguments");
    return batch_advance(batch->impl, duration, -1);
}

int mmx_batch_steps(mmx_batch *batch, int nsteps)
{
    MMX_REQUIRE(batch && nsteps >= 0, MMX_ERR_ARG, "mmx_batch_steps: bad arguments");
    return batch_advance(batch->impl, -1.0, nsteps);
}

int mmx_batch_tune(mmx_batch *batch, int reps, int *lds_kb)
{
    MMX_REQUIRE(batch && reps > 0, MMX_ERR_ARG, "mmx_batch_tune: bad arguments");
    BatchImpl &B = batch->impl;
    // state: every simulation active, tau0 recomputed (stage 0 only: m and the
    // solver state are not changed)
    MMX_TRY(load_sims(B, 1e300, -1));
    static const size_t caps[] = {8 * 1024, 16 * 1024, 32 * 1024, 64 * 1024, 128 * 1024};
    const int ncap = (int)(sizeof(caps) / sizeof(caps[0]));
    hipEvent_t e0, e1;
    MMX_HIP(hipEventCreate(&e0));
    MMX_HIP(hipEventCreate(&e1));
    int rc = MMX_OK;
    for (int p = 0; p < BP_N && rc == MMX_OK; p++) {
        size_t best_cap = B.cap[p];
        float best = 1e30f;
        for (int c = 0; c < ncap && rc == MMX_OK; c++) {
            B.cap[p] = caps[c];
            rc = build_tables(B);
            if (rc != MMX_OK) break;
            if (hipMemcpyAsync(B.dsims, B.hs.data(), sizeof(BatchSim) * B.sims.size(),
                               hipMemcpyHostToDevice, B.st) != hipSuccess) {
                rc = MMX_ERR_HIP;
                break;
            }
            rc = eval_passes<0>(B); // inputs of pass p in place
            float ms = 0.0f;
            for (int r = 0; r < reps && rc == MMX_OK; r++) {
                (void)hipEventRecord(e0, B.st);
                rc = eval_passes<0>(B, p);
                (void)hipEventRecord(e1, B.st);
                if (hipEventSynchronize(e1) != hipSuccess) rc = MMX_ERR_HIP;
                float t = 0.0f;
                (void)hipEventElapsedTime(&t, e0, e1);
                ms += t;
            }
            if (rc == MMX_OK && ms < best) {
                best = ms;
                best_cap = caps[c];
            }
        }
        B.cap[p] = best_cap;
        if (lds_kb) lds_kb[p] = (int)(best_cap / 1024);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == MMX_OK) rc = build_tables(B);
    if (rc == MMX_OK)
        rc = hipMemcpy(B.dsims, B.hs.data(), sizeof(BatchSim) * B.sims.size(),
                       hipMemcpyHostToDevice) == hipSuccess ? MMX_OK : MMX_ERR_HIP;
    return rc;
}

int mmx_batch_set_lds(mmx_batch *batch, const int *lds_kb)
{
    MMX_REQUIRE(batch && lds_kb, MMX_ERR_ARG, "mmx_batch_set_lds: NULL argument");
    for (int p = 0; p < BP_N; p++)
        MMX_REQUIRE(lds_kb[p] >= 1 && lds_kb[p] <= 160, MMX_ERR_ARG,
                    "mmx_batch_set_lds: budget %d KB out of range", lds_kb[p]);
    for (int p = 0; p < BP_N; p++) batch->impl.cap[p] = (size_t)lds_kb[p] * 1024;
    batch->impl.tables_dirty = true;
    return MMX_OK;
}

int mmx_batch_info(const mmx_batch *batch, int *nblocks, int *lds_bytes)
{
    MMX_REQUIRE(batch, MMX_ERR_ARG, "mmx_batch_info: NULL batch");
    for (int p = 0; p < BP_N; p++) {
        if (nblocks) nblocks[p] = batch->impl.nblk[p];
        if (lds_bytes) lds_bytes[p] = (int)batch->impl.lds[p];
    }
    return MMX_OK;
}

} // extern "C"
