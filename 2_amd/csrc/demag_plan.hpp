// demag_plan.hpp -- the demag convolution plan and the fused field epilogue.
#pragma once

#include "mmx_internal.hpp"

namespace mmx {

// Geometry of one zero-padded convolution, internal (z,y,x) order.
struct DemagGeom {
    int n0, n1, n2;    // data
    int p0, p1, p2;    // padded (engine/zeropad.go:12-22)
    int h2;            // p2/2 + 1 complex values per row after the R2C pass
    int Lx;            // p2/2: the R2C pass is a half-length complex FFT
    int sx;            // row stride (complex) of the spectrum buffers
    int ko0, ko1;      // octant kernel extents p0/2+1, p1/2+1
    long N;            // n0 n1 n2
    long nrows;        // n0 n1
};

// What the last (C2R) pass does with the demag field of each cell.
enum EpilogueMode : int {
    EPI_B = 0,      // out = demag_scale * conv (B in tesla, or H)
    EPI_HEFF = 1,   // H_eff = demag + exchange + anisotropy + Zeeman
    EPI_TAU0 = 2,   // first RK stage: tau0 = torque(m, H_eff) (+ Zhang-Li / gamma)
    EPI_CORR = 3,   // RK corrector: tau1, m_new = norm(m + h/2 (tau1 - tau0)), max|tau1-tau0|
};

// Everything the fused epilogue needs besides the demag spectrum.
struct FieldParams {
    float demag_scale;          // Ms (H) or mu0 Ms (B)
    float ex_pre;               // 2 Aex / (mu0 Ms)
    float ex_c[3];              // 1/cell^2, FP64 then rounded (exchange6.cu:152-154)
    int pbc[3];
    const float *msat_map;
    const float *aex_map;
    int anis;
    float ku2;                  // 2 Ku / (mu0 Ms)
    const float *ku_map;
    float u[3];
    float hext[3];              // B_ext / mu0
    float alpha;
    const float *alpha_map;
    int zl;                     // Zhang-Li on
    float zl_pre;               // P muB / (e Ms)
    float zl_xi;
    float zl_j[3];              // current density multipliers
    const float *zl_jmask;      // 3 components or NULL
    float zl_step[3];           // 1/(12 cell)
    float inv_gamma;            // weight of zzt in the torque sum
    const float *htherm = nullptr; // temperature/brown field (3 N) or NULL
    // MASK quantities (NULL = uniform): B_ext (3 N, hext_c * map_c), anisU
    // (3 N, u_c * map_c), polarisation and xi (N each, Zhang-Li prefactor)
    const float *bext_map = nullptr;
    const float *u_map = nullptr;
    // demagexch with an Msat mask: H_dex = conv(m, dex) computed by its own
    // convolution (3 N), added like H_therm
    const float *hdex = nullptr;
    const float *pol_map = nullptr;
    const float *xi_map = nullptr;
};

// Per-launch state of the epilogue (pointers that change between calls).
struct EpilogueIO {
    const float *m;       // magnetisation the field is evaluated at
    float *hout;          // optional H_eff / B output
    float *tout;          // optional torque output (tau/gamma)
    const float *tau0;    // EPI_CORR: first-stage torque
    float *mnew;          // EPI_CORR: corrected, normalised m
    float h;              // EPI_CORR: dt*gamma (float32, solve_rk12.go:153)
    unsigned int *err;    // EPI_CORR: max |tau1 - tau0| (float bits)
};

// Tunables of the pass kernels.
struct DemagTiles {
    int rows_x;   // rows per workgroup in the x passes
    int w_y;      // spectrum columns per workgroup in the y passes
    int w_z;      // spectrum columns per workgroup in the z pass
};

struct DemagPlanImpl {
    DemagGeom g{};
    DemagTiles tiles{};
    int cell_kernel_type = MMX_KERNEL_ARNE;
    int accuracy = 6;
    double cell[3] = {0, 0, 0};
    int pbc[3] = {0, 0, 0};
    double bad_over_good = 0;
    // FFT descriptors (float) for x (half length), y, z
    LineFFT<float> fx, fy, fz;
    void *tw_mem = nullptr;      // twiddle tables (device)
    float2 *xbuf = nullptr;      // 3 x nrows x sx
    float2 *ybuf = nullptr;      // 3 x n0 x p1 x sx
    float *kern = nullptr;       // 6 x ko0 x ko1 x h2 (octant, real, 1/Np)
    float *real_kernel = nullptr;// optional: 6 x p0 p1 p2 real-space kernel
    bool fast_x = false, fast_y = false, fast_z = false; // register-FFT passes
    hipEvent_t *events = nullptr; // when set: recorded before/after every launch
    int *event_pass = nullptr;    // pass id of the launch that each event ends
    int nevent = 0;
    // Set by the caller around one demag_forward: the x pass's first thread
    // moves the word *handoff_dev (left by the previous launch on the stream)
    // into the pinned host word *handoff_host and clears it, and after_x is
    // recorded once the x pass has been queued.
    unsigned int *handoff_dev = nullptr, *handoff_host = nullptr;
    hipEvent_t after_x = nullptr;
    // ConvPlan: the real-space kernel comes from the caller (6 host float32
    // arrays of the padded size, NULL = zero component), every component is
    // transformed and the realness ratio is recorded, not enforced
    const float *const *host_kernel = nullptr;
    // demagexch kernel build: real-space kernel Ms K_d + dex_cex K_ex
    bool dex = false;
    double dex_ms = 0, dex_cex = 0;
};

// Pass names in launch order (3-D grids; films run xf, y2d, xi).
constexpr int kMaxPasses = 5;
// Events recorded by one timed evaluation: the start and one per launch.
constexpr int kMaxMarks = kMaxPasses + 1;

// Kernel construction (kernel_build.hip)
int build_demag_kernel(DemagPlanImpl *P, bool keep_real);
// The demagexch octant (Kernel_Arne at `accuracy` combined with the exchange
// kernel, transformed like the demag kernel) into out (6 x octant floats),
// and out = kd + kx / ms over n floats; both on the null stream, synchronised.
int build_dex_kernel(const DemagPlanImpl *P, int accuracy, double ms, double aex, float *out);
int fold_kernels(const float *kd, const float *kx, double ms, float *out, long n);

// Launch the five (3-D) or three (2-D) passes; the last one runs `mode`.
int demag_forward(DemagPlanImpl *P, const float *m, const float *msat_map,
                  hipStream_t st);
int demag_forward_pred(DemagPlanImpl *P, const float *m0, const float *tau0,
                       float h, const float *norm_map, float *mpred,
                       hipStream_t st);
int demag_inverse(DemagPlanImpl *P, int mode, const FieldParams &fp,
                  const EpilogueIO &io, hipStream_t st);

// Power-of-two passes (demag_fast.hip)
bool fast_length(int L);   // x passes, film y, z
bool fast_ylength(int L);  // 3-D y passes
int fast_xf(const DemagPlanImpl *P, int pred, const float *m, const float *msat_map,
            const float *tau0, float h, const float *norm_map, float *mpred, hipStream_t st);
int fast_yf(const DemagPlanImpl *P, hipStream_t st);
int fast_z(const DemagPlanImpl *P, hipStream_t st);
bool fast_z_addressable(const DemagGeom &g);
bool fast_x_addressable(const DemagGeom &g);
int fast_yi(const DemagPlanImpl *P, hipStream_t st);
int fast_x_blocks(const DemagGeom &g);
bool fast_y2d_as_z_applies(const DemagGeom &g);
int fast_y2d_as_z(const DemagPlanImpl *P, hipStream_t st);
// The register-FFT x passes run when they have at least this many
// workgroups (one per CU); below it a handful of workgroups each carry a long
// serial chain and the generic LDS kernels, spread over more workgroups, are
// faster (128x32x1 film: xf 8.3 against 12.5 us, xi 9.8 against 17.6 us;
// 4x256x256, 64 workgroups: xf 13.1 against 16.3, xi 14.6 against 20.2; at
// 16x256x256, 256 workgroups, the register passes win: xf 17.2 against 30.5).
constexpr int kFastXMinBlocks = 256;
// the value in force (mmx_set_fast_x_min_blocks)
int fast_x_min_blocks();
int set_fast_x_min_blocks(int n); // returns the previous value
// Whether demag_forward runs the register-FFT x pass for this plan.
bool fast_xf_used(const DemagPlanImpl *P);
int fast_xi(const DemagPlanImpl *P, int mode, const FieldParams &fp, const EpilogueIO &io,
            hipStream_t st);
// The fast C2R pass's two-cell epilogue covers uniform parameters
// (fast_xi_applies) and the Msat, Aex, alpha and Ku maps, the B_ext and
// anisU masks, the thermal field and H_dex (fast_xi_maps_applies: no
// spin-transfer term); with the Zhang-Li torque fast_xi runs the general
// per-cell epilogue (xi_gen.hip).
bool fast_xi_applies(const FieldParams &fp);
bool fast_xi_maps_applies(const FieldParams &fp);

} // namespace mmx

namespace mmx {
// Plan construction (plan.cpp); host_kernel: see DemagPlanImpl::host_kernel.
int plan_create(mmx_demag_plan **plan, int n0, int n1, int n2, const double cell[3],
                const int pbc[3], int kernel_type, int accuracy,
                const float *const *host_kernel = nullptr);
void plan_destroy(mmx_demag_plan *plan);
} // namespace mmx

struct mmx_demag_plan {
    mmx::DemagPlanImpl impl;
};

struct mmx_convplan {
    mmx_demag_plan *plan = nullptr;
};
