/*
 * fivo2d.h -- C ABI of the MI355X (gfx950) 2+1D finite-volume hydro core.
 *
 * The library replaces the bodies of FiVoHydro.jl's 2+1D sweep and stepper
 * functions -- the functions `run_sim_2d!` (main2D.jl:371-561) calls with the
 * `(U, g, tau, model, work)` convention -- with hand-written FP64 HIP kernels.
 * Every entry point below names the reference function it replaces.
 *
 * Conventions
 *  - Host state arrays use the reference's layout: U[a + nv*i] (Julia
 *    column-major Matrix(nv, Ntot)), flat cell index i = ix*Nytot + iy,
 *    0-based, y fastest (src2d/grid2d.jl:9-16,85). The device keeps a
 *    structure-of-arrays copy; the layout swap happens inside the copy calls.
 *  - The caller owns every host pointer. The library copies in/out inside the
 *    call and keeps no host pointer. The context owns all device memory.
 *  - Status codes: 0 = ok; >0 = numerical outcome (e.g. FIVO_STEP_FAILED when a
 *    step still fails after dt halving, mirroring `ok=false` in main2D.jl:466-473);
 *    <0 = API or HIP error, text via fivo_last_error().
 *  - A context is single-threaded and owns one HIP stream. Calls return once
 *    their results are visible to the host.
 */
#ifndef FIVO2D_H
#define FIVO2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FIVO_ABI_VERSION 1

/* ---- status codes ---- */
#define FIVO_OK              0
#define FIVO_STEP_FAILED     1   /* step failed after max_halve halvings */
#define FIVO_ERR_ARG        -1   /* invalid argument / refused knob */
#define FIVO_ERR_HIP        -2   /* HIP runtime error */
#define FIVO_ERR_STATE      -3   /* call sequence error (e.g. no state set) */

/* ---- equation of state kinds (src/eos.jl) ---- */
#define FIVO_EOS_CONFORMAL_HQ 0  /* ConformalHQEOS, src/eos.jl:39-84 */
#define FIVO_EOS_LATTICE_HRG  1  /* LatticeHRGEOS,  src/eos.jl:318-519 */

/* ---- boundary conditions (src2d/bc2d.jl:14) ---- */
#define FIVO_BC_OUTFLOW  0
#define FIVO_BC_PERIODIC 1

/* ---- integrators (src2d/timestepper2d.jl:74,148) ---- */
#define FIVO_SSPRK2 2
#define FIVO_SSPRK3 3

/* ---- Terms switches (src/terms.jl:82-106), bit k = field k of `Terms` ---- */
#define FIVO_T_NU_GRADALPHA    (1u << 0)
#define FIVO_T_FM_GRADT        (1u << 1)
#define FIVO_T_FM_INERTIAL     (1u << 2)
#define FIVO_T_FM_NU_GRADU     (1u << 3)
#define FIVO_T_FM_EXPANSION    (1u << 4)
#define FIVO_T_FM_DLNH         (1u << 5)
#define FIVO_T_M2_NU_GRADIENT  (1u << 6)
#define FIVO_T_M2_BG_GRADU     (1u << 7)
#define FIVO_T_M2_BG_DLNT      (1u << 8)
#define FIVO_T_M2_BG_DALPHA    (1u << 9)
#define FIVO_T_M2_EXPANSION    (1u << 10)
#define FIVO_T_M2_PI_SIGMA     (1u << 11)
#define FIVO_T_M2_PIQ_SIGMA    (1u << 12)
#define FIVO_T_M2_VORTICITY    (1u << 13)
#define FIVO_T_M2_PROJECTOR    (1u << 14)
#define FIVO_T_M2_ACCEL_NU     (1u << 15)
#define FIVO_T_M2_NU_GRADTH    (1u << 16)
#define FIVO_T_SHEAR_VORTICITY (1u << 17)
/* `Terms()` defaults: everything on except the two vorticity couplings */
#define FIVO_TERMS_DEFAULT (0x3FFFFu & ~(FIVO_T_M2_VORTICITY | FIVO_T_SHEAR_VORTICITY))

/* Grid2D (src2d/grid2d.jl:19-61): box [-xmax,xmax] x [-ymax,ymax], nghost = 3. */
typedef struct fivo_grid_desc {
    int32_t nx, ny, nghost;
    int32_t _pad;
    double xmax, ymax;
} fivo_grid_desc;

/* EOS parameters: ConformalHQEOS (src/eos.jl:39-46) and LatticeHRGEOS
 * (src/eos.jl:318-347). Unused fields of the other kind are ignored. */
typedef struct fivo_eos_desc {
    int32_t kind;
    int32_t hq_times_fmGeV3;
    double g_eff, m_hq, g_hq;
    double a1, a2, a3, a4, b1, b2, b3, b4, c, d;
    double canon_factor;
} fivo_eos_desc;

/* IdealDiffVisc2DModel (src2d/primitives2d.jl:41-210) as a flat POD, with the
 * transport models QGPViscosity / SimpleBulkViscosity (src/primitives.jl:47-108)
 * folded in. Field meaning and defaults follow build_model_2d (main2D.jl:85-123). */
typedef struct fivo_model_desc {
    fivo_eos_desc eos;
    int32_t with_charge, enable_shear, enable_bulk, enable_diff;
    int32_t consistent_fm, consistent_m2;
    int32_t advect_nu, relax_advect_nu, advect_Pi, relax_advect_Pi;
    int32_t advect_pi, relax_advect_pi, relax_advect_m2;
    int32_t vacuum_ramp_relax, shear_projected_deriv, diff_projected_deriv;
    int32_t shear_constraint_project;   /* 1 = :project, 0 = :monitor */
    uint32_t terms;                     /* FIVO_T_* bitmask */
    /* shear */
    double eta_over_s, tauShear_coeff, deltaShear_factor, taupi_pi_factor, pi_clip_factor;
    /* bulk */
    double zeta_over_s, tauPi_coeff, deltaPi_factor, lambda_Pi_pi_factor,
           lambda_pi_Pi_factor, Pi_clip_factor;
    /* charge diffusion */
    double kappa_coeff, transport_mass, tauN_coeff, deltaN_factor, nu_clip_factor,
           vacuum_n_lo, vacuum_n_hi;
    /* domain and vacuum cut */
    double r_domain, T_vac_cut, E_vac_cut;   /* E_vac_cut <= 0: derived from T_vac_cut */
} fivo_model_desc;

/* Host views of the Work2D arrays (src2d/work2d.jl:15-100). Each pointer is
 * nullable and, when set, points at Ntot doubles (ok: Ntot bytes). */
typedef struct fivo_work_view {
    double *yT, *phi, *mu, *alpha, *ux, *uy, *P, *n, *e;
    uint8_t *ok;
    double *x0_yT, *x0_phi, *x0_ux, *x0_uy;
    double *ux_prev, *uy_prev, *alpha_prev, *T_prev, *nux_prev, *nuy_prev;
    double *shear_res;
} fivo_work_view;

/* Per-call counters of the primitive pass (rhs2d.jl:89-112). nretry and nstaged,
 * not in the reference: cells whose warm-started solve failed and were retried
 * from the cold seed (cons_to_prim_2d!, src2d/primrec2d.jl:465-486), and solves
 * closed by the staged phi fallback (src2d/primrec2d.jl:723-738). */
typedef struct fivo_counts {
    int64_t nprimfail, nvacuum;
    double amax;
    int64_t nretry, nstaged;
} fivo_counts;

/* run_sim_2d! options (main2D.jl:371-378) */
typedef struct fivo_run_opts {
    double tau0, taufinal, cfl, cfltau, dump_dt;
    int32_t integrator, bc;
    int32_t max_halve, max_mood;
    int64_t max_steps;
    int32_t reset_history;
    int32_t _pad;
    /* called at tau0, every dump_dt and at the end, like `on_dump(τ, U, work)`;
     * the callback may call fivo_get_state / fivo_get_work on the context */
    void (*on_dump)(void* user, double tau, void* ctx);
    void* user;
} fivo_run_opts;

/* run_sim_2d! result NamedTuple (main2D.jl:558-560) */
typedef struct fivo_run_result {
    int32_t ok;
    int32_t _pad;
    double tau;
    int64_t nsteps, nprimfail, nvacuum;
    double max_shear_res, maxu, dQ, Q0, Q1, minPtot;
    /* extra accounting, not in the reference NamedTuple */
    int64_t nrhs, nhalvings, nmood;
    int64_t newton_iters;
    int64_t nlearned;            /* cells started from the learned seed */
    int64_t nlearned_fallback;   /* of those, cells that fell back to the cold seed */
    int64_t nqueued;             /* cells handed from the warm to the cooperative primitive kernel */
    int64_t nretry;              /* warm-started solves retried from the cold seed (primrec2d.jl:465-486) */
    int64_t nstaged;             /* solves closed by the staged phi fallback (primrec2d.jl:723-738) */
} fivo_run_result;

typedef struct fivo_ctx fivo_ctx;

/* ---- lifetime ---------------------------------------------------------- */
/* build_model_2d (main2D.jl:85-123) + make_grid2d + allocate_state + make_work.
 * Refuses the knobs reject_unwired_knobs_2d refuses (src2d/primitives2d.jl:246-272). */
int fivo_create(const fivo_grid_desc* g, const fivo_model_desc* m, fivo_ctx** out);
void fivo_destroy(fivo_ctx* ctx);
const char* fivo_last_error(const fivo_ctx* ctx);   /* ctx may be NULL */
int fivo_nvars(const fivo_ctx* ctx);
int64_t fivo_ntot(const fivo_ctx* ctx);
double fivo_e_vac_cut(const fivo_ctx* ctx);         /* derived E_vac_cut (main2D.jl:100-105) */

/* ---- state and work ---------------------------------------------------- */
int fivo_set_state(fivo_ctx* ctx, const double* U, int nv, int64_t ntot);
int fivo_get_state(fivo_ctx* ctx, double* U, int nv, int64_t ntot);
int fivo_set_work(fivo_ctx* ctx, const fivo_work_view* w);
int fivo_get_work(fivo_ctx* ctx, fivo_work_view* w);
/* NaN-fill the warm starts and histories (main2D.jl:407-415) */
int fivo_reset_history(fivo_ctx* ctx);

/* ---- initial conditions ------------------------------------------------ */
/* set_cell! on every cell (main2D.jl:142-154) from host arrays of Ntot values
 * (ux/uy/Pi/pixx/pixy/piyy/nux/nuy nullable = 0), then finalize_ic!
 * (main2D.jl:164-171). nbad = cells whose prim_to_cons failed. */
int fivo_init_from_prims(fivo_ctx* ctx, double tau0, const double* T, const double* alpha,
                         const double* ux, const double* uy, const double* Pi,
                         const double* pixx, const double* pixy, const double* piyy,
                         const double* nux, const double* nuy, int bc, int64_t* nbad);
/* The project's seeded hot-spot event on the device: entropy-like density
 * s = sum_k A_k exp(-|x - x_k|^2 / 2 sigma^2) over the cell centres (spots: x_k,
 * y_k, A_k per spot, 3*nspots doubles), T = Tpeak cbrt(s / max s) floored at
 * T_MIN, alpha = 2 log(s / max s) floored at alpha_floor, u = 0, no dissipative
 * fields; then set_cell! on every cell and finalize_ic! as fivo_init_from_prims.
 * The reference builds events on the host from a grid file
 * (initialize_from_grid_csv!, main2D.jl:238-319); this builds them on the device. */
int fivo_init_hotspots(fivo_ctx* ctx, double tau0, int32_t nspots, const double* spots, double sigma,
                       double Tpeak, double alpha_floor, int bc, int64_t* nbad);
/* finalize_ic! alone (main2D.jl:164-171) on the current state */
int fivo_finalize_ic(fivo_ctx* ctx, double tau0, int bc);

/* ---- sweeps -------------------------------------------------------------- */
/* apply_bc_2d! (src2d/bc2d.jl:14-65) */
int fivo_apply_bc(fivo_ctx* ctx, int bc);
/* rhs_2d! (src2d/rhs2d.jl:169-249). force_first_order: nullable Ntot bytes.
 * dU: nullable host Ntot*nv doubles (reference layout). */
int fivo_rhs(fivo_ctx* ctx, double tau, int bc, const uint8_t* force_first_order,
             double* dU, fivo_counts* counts);
/* update_primitives_2d! (src2d/rhs2d.jl:57-65) */
int fivo_update_primitives(fivo_ctx* ctx, double tau, int bc, fivo_counts* counts);
/* sanitize_stage_2d! (src2d/floors2d.jl:218-227) */
int fivo_sanitize(fivo_ctx* ctx, double tau, int bc);
/* state_ok_2d (src2d/timestepper2d.jl:56-66); nbad = inadmissible interior cells */
int fivo_state_ok(fivo_ctx* ctx, double tau, int* ok, int64_t* nbad);
/* compute_dt_2d (src2d/timestepper2d.jl:37-47) from the last RHS's amax */
int fivo_compute_dt(fivo_ctx* ctx, double tau, double cfl, double cfltau, double* dt);
/* step_ssprk2_2d! / step_ssprk3_2d! (src2d/timestepper2d.jl:74-188) */
int fivo_step(fivo_ctx* ctx, int integrator, double tau, double dtau, int bc,
              int max_halve, int max_mood, int* ok, double* dtau_used);
/* relax_dissipative_2d! (src2d/dissipation2d.jl:293-363) */
int fivo_relax(fivo_ctx* ctx, double tau, double dtau);
/* project_shear! (main2D.jl:332-358) */
int fivo_project_shear(fivo_ctx* ctx, double* max_rel, double* mean_rel);
/* run_sim_2d! (main2D.jl:371-561), device resident */
int fivo_run(fivo_ctx* ctx, const fivo_run_opts* o, fivo_run_result* r);

/* ---- readout: a background for a transport code ------------------------- */
/* One time slice of export_background2d (tools/export_background2d.jl:79-94),
 * from the primitives of the last primitive pass (call fivo_update_primitives at
 * the slice's tau first, as the reference does). For every interior cell
 * (i, j) = (ix - nghost, iy - nghost):
 *     T = exp(yT),  v^x = u^x / u^tau,  v^y = u^y / u^tau,  n,  u^tau = sqrt(1 + u.u)
 * (the VELOCITY a transport code reads, not the four-velocity), stored at element
 * (i, j, k) of column-major (Nx, Ny, nt) tables: offset i + Nx (j + Ny k), Julia's
 * T_grid[i, j, k]. device != 0: the pointers are device memory and the tables stay
 * in HBM across slices; 0: host memory. n may be NULL. maxima (nullable) receives
 * max |u| and max |v| over the slice (NaN if any cell's is NaN), for the
 * exporter's |v| < 1 check. */
int fivo_background_slice(fivo_ctx* ctx, int32_t k, int32_t nt, double* T, double* vx, double* vy,
                          double* n, int32_t device, double* maxima);

/* ---- timing hooks for benchmarks --------------------------------------- */
/* Total device time (ms) of the kernels by category since fivo_enable_timing,
 * measured with HIP events on the context stream: 0 = primitive pass (warm
 * kernel), 1 = faces+update, 2 = boundary fill, 3 = relaxation, 4 = everything,
 * 5 = primitive pass (cooperative kernel), 6 = learned seeds, 7 = background
 * slice, 8 = stage end (ghost fill, pending test and scalar copies of an RK stage
 * in one launch). Returns the count of launches. */
int64_t fivo_kernel_time(const fivo_ctx* ctx, int category, double* total_ms);
int fivo_enable_timing(fivo_ctx* ctx, int on);
/* Histograms of the primitive passes of the last fivo_run (cumulative over the run),
 * 24 counts: [0, 8) every cell of every warm pass -- 0 finished at its first
 * residual (converged warm start, vacuum), 1-6 converged after that many warm
 * Newton iterations, 7 queued for the cooperative solve; [8, 16) the cooperative
 * solves that resumed a warm solve and [16, 24) those started from a seed (cold or
 * learned), by total Newton iterations in bins of ten (the last bin: 70 or more).
 * No counterpart in the reference: a diagnostic of where the pass's time goes. */
int fivo_get_histograms(const fivo_ctx* ctx, int64_t* counts);

/* ---- probes of the relaxation's source terms ---------------------------------- */
/* Evaluate the device code of a source term on n states given row by row (host
 * arrays, row-major), with the term switches `terms` (FIVO_T_* bits). These exist so
 * that tests can run the reference's fixed-state gates (1-D limit, rotation
 * covariance, tracelessness, Bjorken limit) on the HIP path; no solver call uses them.
 *
 * fivo_probe_fm_source replaces nothing: it evaluates consistent_fm_source_2d
 * (src2d/hq_consistent_firstmoment2d.jl:218-282) with its 22 arguments in the
 * reference's order (ux, uy, ut, T, dxT, dyT, dtT, nux, nuy, n, dn_dT, tau_n, D_s, h,
 * h', theta, ax, ay, dxux, dxuy, dyux, dyuy) -> (s^x, s^y).
 * fivo_probe_m2_source evaluates consistent_m2_source_2d
 * (src2d/hq_consistent_m2_2d.jl:247-427) with its 44 arguments in the reference's
 * order (ux, uy, ut, tau, T, dxT, dyT, dtT, nux, nuy, pixx_Q, pixy_Q, piyy_Q, pieta_Q,
 * Pi_Q, dt_alpha, dx_alpha, dy_alpha, theta_nu, a_nu^x, a_nu^y, dt_nux, dt_nuy,
 * dx_nux, dx_nuy, dy_nux, dy_nuy, theta, ax, ay, dt_ux, dt_uy, dxux, dxuy, dyux, dyuy,
 * n, tau_n, D_s, h, h', tau_M, eta_M, m) -> the rates (pixx_Q, pixy_Q, piyy_Q, Pi_Q).
 * No counterpart in the reference (its gates call the Julia functions directly). */
int fivo_probe_fm_source(uint32_t terms, int64_t n, const double* in, double* out);
int fivo_probe_m2_source(uint32_t terms, int64_t n, const double* in, double* out);

/* ---- learned Newton warm start --------------------------------------------- */
/* A 6 -> 64 -> 64 -> 6 GELU MLP (BF16 MFMA, FP32 accumulate) mapping features of a
 * cell's conserved state (log E, log(D + 1e-30), |S|/E, Pi/E, s.pi.s/E,
 * s.nu/(|nu| + D), s = S/|S|) to its primitives (yT, phi, asinh(u.s), u.s_perp,
 * log e, log P). With a context's learned seed set, cells without a warm start
 * start Newton from the network's output instead of seed_from_conserved_2d
 * (src2d/primrec2d.jl:323-347); Newton still converges to the reference
 * tolerance, and a cell whose solve fails falls back to the cold seed. No
 * counterpart in the reference. Parameters: W1[64][6], b1[64], W2[64][64],
 * b2[64], W3[6][64], b3[6] (FIVO_MLP_NPARAM floats); normalisation: feature
 * mean[6], std[6], target mean[6], std[6]. */
#define FIVO_MLP_NPARAM 4998
#define FIVO_MLP_NNORM 24
typedef struct fivo_mlp fivo_mlp;
/* an untrained network (Glorot-uniform from `seed`) with room for `capacity` pairs */
int fivo_mlp_create(int64_t capacity, uint64_t seed, fivo_mlp** out);
void fivo_mlp_destroy(fivo_mlp* m);
const char* fivo_mlp_last_error(const fivo_mlp* m);
/* append the (features, targets) pairs of the converged interior cells of the
 * context's last primitive pass at tau; total = pairs held */
int fivo_mlp_collect(fivo_mlp* m, fivo_ctx* ctx, double tau, int64_t* total);
int fivo_mlp_size(const fivo_mlp* m, int64_t* npairs);
/* mean / std of the held pairs; normalises them in place (collection ends) */
int fivo_mlp_fit_normalization(fivo_mlp* m);
/* Adam on minibatches of the held pairs; loss[epoch] = mean over pairs of the
 * summed squared error of the normalised outputs during that epoch */
int fivo_mlp_train(fivo_mlp* m, int32_t epochs, int32_t batch, float lr, float* loss);
int fivo_mlp_get_params(const fivo_mlp* m, float* params, float* norm);   /* either nullable */
int fivo_mlp_set_params(fivo_mlp* m, const float* params, const float* norm);
/* mean squared error of each normalised output over the held pairs */
int fivo_mlp_eval(fivo_mlp* m, float* mse);
/* the network on n normalised feature rows (n x 6) -> normalised outputs (n x 6) */
int fivo_mlp_forward(fivo_mlp* m, int64_t n, const float* x, float* y);
/* loss and its gradient w.r.t. the parameters on n normalised (x, target) rows */
int fivo_mlp_grad(fivo_mlp* m, int64_t n, const float* x, const float* t, float* loss, float* grad);
/* use the network for the context's cold starts (m = NULL: the cold seed);
 * the network must outlive its use */
int fivo_set_learned_seed(fivo_ctx* ctx, const fivo_mlp* m);

/* ---- ensembles of events ------------------------------------------------- */
/* The reference runs fluctuating events one at a time, each with its own
 * run_sim_2d! (main2D.jl:371-561). An ensemble runs many such events at once:
 * each phase of the SSPRK3 or SSPRK2 step (with its MOOD stage retries) is one
 * launch over every event, the step is a captured hipGraph, and every event
 * keeps its own dt, halvings, retries and stopping time, so each event's result
 * is the one fivo_run gives for it alone. */
typedef struct fivo_batch fivo_batch;
/* nev events on their own grids, one model (build_model_2d, main2D.jl:85-123) */
int fivo_batch_create(const fivo_grid_desc* grids, int32_t nev, const fivo_model_desc* m, fivo_batch** out);
void fivo_batch_destroy(fivo_batch* b);
const char* fivo_batch_last_error(const fivo_batch* b);  /* b may be NULL */
/* the one-event context of event ev: set its state and initial conditions, read
 * it back, or run it alone with the one-event API */
fivo_ctx* fivo_batch_event(fivo_batch* b, int32_t ev);
/* time the fused face kernel for every candidate tile shape on the current
 * states at tau (after a primitive pass there) and keep the fastest; ms gets
 * one time per candidate (FIVO_BATCH_NTILES), tx/ty the chosen tile. The Newton
 * warm starts that pass moves are restored, so a later run is the same with or
 * without tuning. */
#define FIVO_BATCH_NTILES 4
int fivo_batch_autotune(fivo_batch* b, double tau, int32_t* tx, int32_t* ty, float* ms);
/* use candidate tile `tile` (0..FIVO_BATCH_NTILES-1) without timing */
int fivo_batch_set_tile(fivo_batch* b, int32_t tile);
/* run_sim_2d! for every event (SSPRK3, or SSPRK2 with up to 4 MOOD retries per
 * stage; no dump callback): res[ev] per event.
 * tau0: nullable per-event start times (else o->tau0 for all), e.g. to continue
 * a run whose events stopped at different times. Returns FIVO_STEP_FAILED when
 * some event failed, like fivo_run for it alone. */
int fivo_batch_run(fivo_batch* b, const fivo_run_opts* o, const double* tau0, fivo_run_result* res);
/* fivo_set_learned_seed for every event of the ensemble */
int fivo_batch_set_learned_seed(fivo_batch* b, const fivo_mlp* m);

/* ---- freeze-out surface and Cooper-Frye yield (freezeout2d.jl) ------------
 * A tracker bound to one context records, after each primitive pass the caller
 * makes, the last downward crossing of T_fo in every interior cell: the
 * surface tau = tau_fo(x, y) and the flow (u^x, u^y) there (freeze_surface's
 * record!, freezeout2d.jl:97-119). Cooper-Frye then integrates
 *   dN/dy = f_yield / (2 pi hbar c)^3 sum dSigma_mu p^mu f(E_s / T_fo),
 *   dSigma_mu p^mu = tau_fo (m_T cosh eta - d_x tau_fo p_x - d_y tau_fo p_y),
 *   E_s = u^tau m_T cosh eta - u^x p_x - u^y p_y,
 * over the cells 1..nx-2 x 1..ny-2 (0-based) with a frozen tau_fo and finite
 * gradients (central differences, one-sided where one neighbour has none) on
 * a p_T x phi x eta product grid, with Boltzmann and Bose statistics
 * (pion_yield, freezeout2d.jl:144-192). Surface arrays are nx*ny doubles in
 * the grid's interior order, index a*ny + b (x index a, y index b, 0-based). */
typedef struct fivo_freezeout fivo_freezeout;
typedef struct fivo_cf_opts {
    int32_t n_eta, n_pt, n_phi;   /* grid sizes, each in 2..256 (40, 40, 24) */
    int32_t _pad;
    double eta_max;               /* eta in [-eta_max, eta_max] (6.0) */
    double pt_lo, pt_hi;          /* GeV (0.1, 5.0) */
    double mass;                  /* GeV (0.138, the pion) */
    double yield_factor;          /* 2 for pi+ + pi-, 2.5 for feed-down (5.0) */
} fivo_cf_opts;
/* a tracker on ctx at T_fo (GeV): every cell not yet hot, no surface
 * (freezeout2d.jl:87-91); the tracker must not outlive ctx */
int fivo_freezeout_create(fivo_ctx* ctx, double T_fo, fivo_freezeout** out);
void fivo_freezeout_destroy(fivo_freezeout* f);
/* record! at tau from the context's current primitives (T = exp(yT), u^x,
 * u^y), as left by the last fivo_rhs, fivo_update_primitives or fivo_run:
 * T > T_fo marks the cell hot and clears its tau_fo (a cell that had frozen
 * counts as re-heated); a hot cell at or below T_fo without a tau_fo freezes
 * at the linear crossing between the previous sample and this one, with the
 * flow of this sample (freezeout2d.jl:97-119) */
int fivo_freezeout_record(fivo_freezeout* f, double tau);
/* the surface: tau_fo (NaN where not frozen), u^x, u^y there (nullable each);
 * counts[3] = cells ever hot, cells frozen, re-heating events (nullable)
 * (freezeout2d.jl:127-132) */
int fivo_freezeout_surface(fivo_freezeout* f, double* tau_fo, double* ux, double* uy, int64_t* counts);
/* replace the surface (tau_fo NaN = not frozen); the hot flags and the
 * re-heating count are left as they are */
int fivo_freezeout_set_surface(fivo_freezeout* f, const double* tau_fo, const double* ux, const double* uy);
/* pion_yield (freezeout2d.jl:144-192) on the current surface:
 * out[3] = dN/dy Boltzmann, dN/dy Bose, the share of |dSigma.p| f_Boltzmann
 * carried by dSigma.p < 0 */
int fivo_freezeout_pion_yield(fivo_freezeout* f, const fivo_cf_opts* o, double* out);

/* fivo_freezeout_pion_yield on a surface already in device memory, allocating
 * nothing. No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). tau_fo, ux, uy: device arrays of nx*ny doubles in
 * the tracker's order (NaN tau_fo = not frozen); dx, dy: cell sizes (fm);
 * out: a device array of 3 doubles, written on `stream` (a hipStream_t, NULL =
 * the default stream) without synchronising -- read it after the stream. The
 * scratch (cell list and per-workgroup partials, 4.03 MiB) is one static device
 * array per GPU, held until the process ends: nx*ny above
 * FIVO_CF_NOALLOC_MAX_CELLS is refused with FIVO_ERR_ARG, and calls that
 * overlap on one GPU race on it. Errors: fivo_last_error(NULL). */
#define FIVO_CF_NOALLOC_MAX_CELLS (1 << 20)
int fivo_pion_yield_noalloc(const double* tau_fo, const double* ux, const double* uy, int32_t nx, int32_t ny,
                            double dx, double dy, double T_fo, const fivo_cf_opts* o, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FIVO2D_H */
