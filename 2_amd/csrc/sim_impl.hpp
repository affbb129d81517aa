// sim_impl.hpp -- the simulation object behind mmx_sim_* (private to the
// library): parameters, solver state, device buffers.  Shared by sim.cpp and
// by the batched stepping in batch.hip.
#pragma once

#include "demag_plan.hpp"

struct mmx_sim {
    int n[3] = {0, 0, 0};
    int pbc[3] = {0, 0, 0};
    double cell[3] = {0, 0, 0};
    long N = 0;

    // VALUE / MASK multipliers (mod_micromag.go:22-34 and the modules)
    double msat = 0, aex = 0, alpha = 0, gamma = mmx::kGamma0;
    double bext[3] = {0, 0, 0};
    bool anis = false;
    double ku = 0, anisU[3] = {0, 0, 0};
    bool zl = false;
    double j[3] = {0, 0, 0}, pol = 1.0, xi = 0.05; // mod_zhang-li.go:31-36
    float *msat_map = nullptr, *aex_map = nullptr, *alpha_map = nullptr, *ku_map = nullptr,
          *j_map = nullptr;
    // MASK quantities B_ext, anisU (3 components), polarisation, xi
    float *bext_map = nullptr, *u_map = nullptr, *pol_map = nullptr, *xi_map = nullptr;

    // solver/rk12 (solve_rk12.go:33-72 + mod_micromag.go:31-33)
    double t = 0, dt = 1e-15, mindt = 1e-15, maxdt = 1e38, maxerr = 1.0 / 2000.0;
    double err = 0, peakerr = 0;
    long step = 0, badsteps = 0, evals = 0;
    bool m_normalized = false;

    // 0 solver/rk12, 1 solver/heun, 2 solver/euler, 3 solver/bdf-euler, 4 solver/am12
    int solver = 0;
    // implicit solvers (solve_euler_backward.go:104-146, solve_am01.go:328-398)
    int bdf_maxiter = 3;
    double bdf_maxitererr = 1e-5, am_abserr = 1e-4, am_relerr = 1e-3;
    long bdf_iters = 0;   // bdf_iterations of the last step
    double am_alpha = 0;  // the predictor's last convergence ratio

    // temperature/brown (mod_temperature.go:22-134): Temp multiplier and map,
    // Therm_seed, cutoff_dt; the field and the state of its updater
    bool thermal = false;
    // demagexch (mod_demagexchange.go): H_dex = conv(m, Ms K_d + 2 Aex/(mu0 Ms) K_ex)
    // with K_d the Arne kernel at demag_acc; the H_eff evaluations convolve with
    // kern_fold = demag octant + kern_dex / Ms (one convolution for both fields)
    bool dex = false;
    int dex_acc = 8;
    float *kern_dex = nullptr, *kern_fold = nullptr;
    float *hdex = nullptr; // H_dex of the current stage when an Msat mask is set (3 N)
    bool dex_valid = false;
    double dex_ms = 0, dex_aex = 0;
    int dex_built_acc = 0;
    double temp = 0, therm_seed = 0, cutoff_dt = 0;
    float *temp_map = nullptr, *htherm = nullptr;
    unsigned long long therm_draws = 0; // draws since the seed was set
    long long therm_seed_cache = -10000000000LL;
    double therm_last_time = 0;
    bool therm_valid = false;           // drawn for (therm_step, therm_dt, therm_key)
    long therm_step = -1;
    double therm_dt = 0, therm_key[4] = {0, 0, 0, 0};
    int kernel_type = MMX_KERNEL_ARNE, accuracy = 6;
    mmx_demag_plan *plan = nullptr;
    mmx_reductor *red = nullptr;
    float *m = nullptr, *mnew = nullptr, *mpred = nullptr, *tau0 = nullptr, *heff = nullptr,
          *torq = nullptr, *bf = nullptr;
    // solver/rk12: the next step's first stage, computed at the corrected m
    // while the host runs the step controller (valid until the state changes)
    float *tau0_spec = nullptr;
    bool spec_valid = false;
    hipEvent_t ev_err = nullptr;
    unsigned int *derr = nullptr; // [0]: RK error; [4 .. 12]: implicit-solver maxima
    bool derr_clear = false;      // derr[0] is known to be zero (cleared by a hand-off)
    unsigned int *herr = nullptr; // pinned, same layout
    hipStream_t st = nullptr;

    // learned inner-stage field (borrowed; mmx_sim_set_surrogate)
    mmx_surrogate *sur = nullptr;
    bool sur_on = false;
    double sur_tol = 1e-3;
    int sur_check = 8;
    long sur_checks = 0, sur_fallbacks = 0, sur_evals = 0;
    double sur_residual = 0.0;
};

namespace mmx {
// Msat/gamma set, demag plan built, m normalised (queued on s->st).
int sim_prepare(mmx_sim *s);
// Surrogate field evaluation (surrogate.hip): at m0, or at the RK predictor
// normalize(m0 + h tau0) written to mpred when tau0 is given; the H_demag
// estimate goes into the epilogue of `mode` (and to hd_out if not NULL).
int surrogate_eval(mmx_surrogate *S, int mode, const float *m0, const float *tau0, float h,
                   float *mpred, const FieldParams &fp, const EpilogueIO &io, float *hd_out,
                   hipStream_t st);
// ||a - b|| / ||c|| over n floats (synchronises).
int surrogate_rel_l2(mmx_surrogate *S, const float *a, const float *b, const float *c, long n,
                     hipStream_t st, double *rel);
// the surrogate's H_demag buffer (3N floats)
float *surrogate_hd_buffer(mmx_surrogate *S);
// The epilogue parameters of the simulation's current settings (H_eff mode).
FieldParams sim_field_params(const mmx_sim *s);
// Fills out (3 N) with the temperature/brown field of draw `draw` of `seed`
// (thermal.hip); a = float32(alpha 2 kB T), b = float32(mu0 V gamma dt Msat).
int therm_draw(unsigned long long seed, unsigned long long draw, long N, const float *alpha_map,
               const float *temp_map, const float *msat_map, float a, float b, float *out,
               hipStream_t st);
// One fixed-point pass of the implicit solvers (implicit.hip): the next
// iterate and, per component c, res[3c] = max|m - b|, res[3c+1] = max|b|,
// res[3c+2] = max|m + b| (synchronises).
int implicit_pass(int mode, const float *y0, const float *dy, const float *dy0, float h,
                  const float *m, float *yb, long N, unsigned int *dout, unsigned int *hout,
                  hipStream_t st, double res[9]);
} // namespace mmx
