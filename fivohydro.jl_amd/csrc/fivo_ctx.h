// =============================================================================
// fivo_ctx.h -- device context of the 2+1D hot path: HBM layout, launch
// parameters and the kernel launchers shared by the C ABI (fivo_api.hip) and the
// kernels (fivo_kernels.hip).
//
// HBM layout (structure of arrays): every cell field is a contiguous array of
// `pitch` doubles (pitch = Ntot rounded up to 64), cell index c = ix*Nytot + iy
// (y fastest, the reference's flat order). A state buffer is nv such arrays back
// to back: field a of buffer U is U + a*pitch. A y-row of a field is contiguous,
// so a wavefront of 64 lanes along y issues one 512-byte load per field.
// =============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "fivo_device.h"
#include "fivo_mlp.h"
#include "../../include/fivo2d.h"

namespace fivo {

// Model constants needed on the device (flat, passed by value as a kernel argument)
struct Params {
    int Nx, Ny, ng, Nxtot, Nytot, Ntot;
    int pitch, nv;
    int iNux, iNuy, iPi, iPixx, iPixy, iPiyy, iPieta;   // -1 when the sector is absent
    int iPQ0;                                            // first charm second-moment row, -1 if absent
    int hasNu, hasPi, hasShear, hasM2;
    int has_rdom;
    double dx, dy, xmin, ymin;
    double Evac, rdom2;
    fd::EosP eos;
    // relaxation (src2d/dissipation2d.jl, src/primitives.jl:30-108)
    int enable_shear, enable_bulk, enable_diff, consistent_fm, consistent_m2;
    int relax_advect_pi, relax_advect_Pi, relax_advect_nu, relax_advect_m2;
    // passive-scalar transport of the dissipative fields through the face flux, with
    // the grad.v source (advect_*, src2d/fluxes2d.jl:55-67, src2d/rhs2d.jl:200-224)
    int advect_nu, advect_Pi, advect_pi;
    int shear_projected_deriv, diff_projected_deriv, vacuum_ramp_relax;
    uint32_t terms;
    double eta_over_s, Cs, deltaShear_factor, taupi_pi_factor, pi_clip_factor;
    double zeta_over_s, Czeta, deltaPi_factor, lambda_Pi_pi_factor, lambda_pi_Pi_factor, Pi_clip_factor;
    double kappa_coeff, transport_mass, tauN_coeff, deltaN_factor, nu_clip_factor, vacuum_n_lo, vacuum_n_hi;
};

// Small device-resident scalars read back once per stage. Each on a 128-byte line
// of its own: device-wide atomics on one line are serialised (~12 ns apiece on
// MI355X), so counters sharing a line would also wait for each other's updates.
#define FIVO_SCALAR(name) unsigned long long name; unsigned long long name##_line_[15]
struct DevScalars {
    // cleared before each primitive pass (reset_scalars): amax_bits .. nwarm
    FIVO_SCALAR(amax_bits);     // max |signal speed| of the last RHS (positive double bits)
    FIVO_SCALAR(nbad);          // inadmissible cells of the last stage
    FIVO_SCALAR(nprimfail);     // last primitive pass
    FIVO_SCALAR(nvacuum);
    FIVO_SCALAR(npending);      // near-bound cells awaiting the inversion test
    FIVO_SCALAR(ncold);         // cells of the current primitive pass queued for the cooperative solve
    FIVO_SCALAR(nresq);         // of a pass's queue, resumed solves (at the queue's far end)
    FIVO_SCALAR(nwarm);         // cells of the current primitive pass whose warm start iterates
    // cumulative
    FIVO_SCALAR(newton_iters);
    FIVO_SCALAR(shear_max_bits);
    FIVO_SCALAR(nlearned);      // cumulative: cells started from the learned seed
    FIVO_SCALAR(nfallback);     // cumulative: of those, cells that fell back to the cold seed
    FIVO_SCALAR(nqueued);       // cumulative: cells queued for the cooperative kernel
    FIVO_SCALAR(bg_umax_bits);  // background slice: max |u|, max |v| (double bits; NaN sticks)
    FIVO_SCALAR(bg_vmax_bits);
    FIVO_SCALAR(nretry);        // cumulative: warm-started solves that failed and were retried from the cold seed
    FIVO_SCALAR(nstaged);       // cumulative: solves closed by the staged phi fallback (bisection of the charge row)
    // cumulative histograms of the primitive passes (fivo_get_histograms), each bin on
    // its own line: [0, 8) every cell of a warm pass -- 0 finished at its first
    // residual, 1-6 converged after that many warm iterations, 7 queued for the
    // cooperative solve; [8, 16) the cooperative solves that resumed a warm solve and
    // [16, 24) those started from a seed, by total Newton iterations in bins of ten
    unsigned long long hist[24][16];
};
#undef FIVO_SCALAR


// Per-cell work arrays on the device (Work2D, src2d/work2d.jl:15-100)
// a cell queued by the warm primitive kernel: resume its Newton at iteration
// `it` from the iterate kept in the warm-start arrays, or (it = 0) solve it
// from the cold seed (it >= 1000: only the staged fallback is left); ev is the
// event of an ensemble (0 for one event)
struct PrimJob { int cell, it, ev; };

struct DevWork {
    double *yT, *phi, *mu, *alpha, *ux, *uy, *P, *n, *e, *vx, *vy;
    uint8_t *ok, *bad, *force, *ftmp;
    double *x0_yT, *x0_phi, *x0_ux, *x0_uy;
    double *ux_prev, *uy_prev, *alpha_prev, *T_prev, *nux_prev, *nuy_prev;
    double *shear_res;
    double *gradAx, *gradAy;
};

// timing categories: warm primitive kernel, fused faces, ghost fill, relaxation,
// everything, cooperative primitive kernel, learned seeds, background slice, stage end
enum { TK_PRIM = 0, TK_FACES = 1, TK_BC = 2, TK_RELAX = 3, TK_ALL = 4, TK_COOP = 5, TK_SEED = 6, TK_BG = 7,
       TK_TAIL = 8, TK_N = 9 };

struct Ctx {
    fivo_grid_desc gd;
    fivo_model_desc md;
    Params p;
    int dev = 0;
    hipStream_t stream = nullptr;
    void* pool = nullptr;           // one device allocation carved into the arrays below
    size_t pool_bytes = 0;
    double* U[4] = {nullptr, nullptr, nullptr, nullptr};   // U[cur] = state, others stage outputs
    int cur = 0;
    double* dU = nullptr;           // RHS output for fivo_rhs
    double* snap = nullptr;         // pre-relaxation snapshot (dissipation2d.jl:299-303)
    double* partials = nullptr;     // reduction partials
    PrimJob* jobs = nullptr;        // cells of a primitive pass queued for the cooperative solve
    int* wlist = nullptr;           // cells of a primitive pass whose warm Newton iterates
    uint8_t* dirty = nullptr;       // face tiles touching a queued cell (one byte per tile of the smallest tile)
    int ntiles_max = 0;
    hipStream_t stream2 = nullptr;  // the face tiles that run beside the cooperative kernel
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int npartials = 0;
    DevWork w;
    DevScalars* sc = nullptr;       // device
    DevScalars* sc_host = nullptr;  // pinned mirror
    // a step whose stages are all launched before any is checked: the scalars after
    // each stage (slots 0-2) and after the refresh pass (slot 3), pinned; the Newton
    // warm starts after stages 1 and 2 (2 x 4 pitch doubles, the rollback of a failed
    // stage); the shear projection's partials, pinned, read at the next synchronisation
    DevScalars* sc_stage_host = nullptr;
    DevScalars* sc_slots = nullptr;          // device copies of the stage scalars (k_stage_end), 4 slots
    unsigned int* tail_done = nullptr;       // k_stage_end's finished-block count (zero between launches)
    double* ws = nullptr;
    double* partials_host = nullptr;
    double amax_host = 0.0;         // last RHS amax, as read back
    std::string err;
    // tile shape of the fused face kernel (chosen by grid size or autotuned)
    int tile = 0;
    // learned warm start for cells without one (null: the reference's cold seed)
    int learned = 0;
    MlpDev mlp{};
    // timing
    bool timing = false;
    double t_ms[TK_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t t_cnt[TK_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> evpool;
    std::vector<int> ev_cat;
    int ev_used = 0;
    // run accounting
    int64_t nrhs = 0, nhalvings = 0, nmood = 0;
    int64_t last_primfail = 0, last_vacuum = 0;
    // an accepted all-stage attempt left the per-pass device scalars zero and the
    // new state's ghosts filled, so the refresh pass that follows needs neither
    bool pass_clean = false;
};

// ---- launchers (fivo_kernels.hip) ------------------------------------------
struct FaceLaunch {
    const double* Ub;      // base state (RK input = stage input)
    const double* Uold;    // RK weight state (may equal Ub)
    double* Uout;          // stage output (mode 1) -- never aliases Ub
    double* dUout;         // RHS output (mode 0)
    int mode;              // 0: write dU; 1: combine + sanitize + admissibility
    int combine;           // 0: Uin + D k; 1: wold Uold + wnew (Uin + D k); 2: (Uold + 2 (Uin + D k))/3
    double tau, tau_state, D, wold, wnew;
    const uint8_t* force;  // MOOD first-order mask or null
};
hipError_t launch_bc(Ctx& c, double* U, int bc);
hipError_t launch_prim(Ctx& c, const double* U, double tau);        // the three below in order
hipError_t launch_prim_warm(Ctx& c, const double* U, double tau);
hipError_t launch_prim_seed(Ctx& c, const double* U, double tau);   // no-op without a learned seed
hipError_t launch_prim_coop(Ctx& c, const double* U, double tau);
// pass 0: every tile; 1: the tiles no queued cell touches; 2: the rest (see FaceK)
hipError_t launch_faces(Ctx& c, const FaceLaunch& f, int pass = 0, hipStream_t st = nullptr);
hipError_t launch_mark_tiles(Ctx& c);   // marks from the queue of the last warm primitive pass
hipError_t launch_expand(Ctx& c, const uint8_t* in, uint8_t* out);
hipError_t launch_sanitize(Ctx& c, double* U, double tau, int use_P);
hipError_t launch_admissible(Ctx& c, const double* U, double tau, int use_work, uint8_t* bad_out);
hipError_t launch_admissible_pending(Ctx& c, const double* U, double tau);
// ghost fill + pending test of a stage output, its scalars to sc_slots[slot], the
// warm starts to ws_dst (nullable), per-pass scalars cleared (clear): one launch
hipError_t launch_stage_end(Ctx& c, double* U, double tau, int bc, int slot, double* ws_dst, bool clear);
hipError_t launch_init_prims(Ctx& c, double* U, double tau0, const double* const* prims, unsigned long long* nbad);
hipError_t launch_fill(Ctx& c, double* p, double v, int64_t n);
hipError_t launch_hotspot_sum(Ctx& c, double* s_out, int nspots, const double* spots, double sigma, double* part);
hipError_t launch_hotspot_prims(Ctx& c, const double* s_in, double* T, double* alpha, double smax, double Tpeak,
                                double alpha_floor);
hipError_t launch_history(Ctx& c);
hipError_t launch_reduce_diag(Ctx& c, const double* U, int what);
hipError_t launch_project_shear(Ctx& c, double* U, int correct);
hipError_t launch_relax(Ctx& c, double* U, double tau, double dtau);
// the relaxation's source terms at arbitrary states (which: 0 first moment, 1 second moment)
hipError_t launch_probe(int which, uint32_t terms, int64_t n, const double* in, double* out, hipStream_t s);
// the text fivo_last_error(NULL) returns, for errors without a context (fivo_api.hip)
void set_last_error(const std::string& msg);
hipError_t launch_layout(Ctx& c, const double* aos_dev, double* soa, int to_soa);
// one slice of a transport background into column-major (Nx, Ny, *) tables at
// element offset `off` (n may be null); maxima into sc->bg_{u,v}max_bits
hipError_t launch_background(Ctx& c, int64_t off, double* T, double* vx, double* vy, double* n);

// ---- ensembles (fivo_batch.hip, fivo_batch_kernels.hip) -----------------------
// An ensemble is a set of events, each a one-event context (its own grid, state
// buffers, work arrays and scalars) sharing one model. The ensemble kernels run
// one phase of the time step for every event in one launch: each grid kind maps
// a block to (event, local block) through a block -> event table.
enum { KB_CELLS = 0,      // 256 cells of Ntot per block
       KB_INTERIOR = 1,   // 256 interior cells per block
       KB_GHOST = 2,      // 256 ghost cells per block
       KB_FACES = 3,      // face tiles + ghost-face blocks
       KB_COOP = 4,       // 64-lane blocks of the cooperative primitive kernel
       KB_N = 5 };
// per event, device resident, fixed during a run
struct EvView {
    Params p;
    DevWork w;
    double* U[4];
    DevScalars* sc;
    double* partials;     // 2 per interior block
    PrimJob* jobs;        // queued cells of the event's primitive pass: sc->ncold seeded from the
                          // front, sc->nresq resumed from the far end (jobs[Ntot - 1] down)
    int* wlist;           // cells of the event's warm pass whose Newton iterates (count: sc->nwarm)
    int tiles_y, nmain;   // face tiling of the ensemble's tile shape
    int b0[KB_N];         // first block of the event in each grid kind
    int nb[KB_N];         // blocks of the event in each grid kind
};
// per event run control: the state of run_sim_2d!'s loop (main2D.jl:371-561)
enum { EV_RUNNING = 0, EV_DONE = 1, EV_FAILED = 2 };
struct EvCtl {
    double tau, D, taufinal, amax, max_shear;
    double shear_step;  // shear residual maximum of the current replay's projection
    int cur;        // index of the accepted state among U[0..3]
    int state;      // EV_*
    int fresh;      // next attempt starts a new step (history snapshot, new dt)
    int stage;      // stages of the current attempt passed so far
    int nhalve;     // halvings of the current step
    int have_hist;
    int accepted;   // the current replay accepted a step
    int dohist;
    int mood_attempt;   // MOOD retry of the current stage being run (0: the first attempt)
    long long nsteps, max_steps, nprimfail, nvacuum, nhalvings, nrhs, nmood;
};
struct BatchScalars {
    unsigned long long nrunning;   // events still running after the last replay
    unsigned long long pad[7];
};
struct BatchRun {                  // kernel arguments shared by the ensemble kernels
    Params p0;                     // the model part of every event's Params
    const EvView* ev;
    EvCtl* ctl;
    BatchScalars* bs;
    const int* map[KB_N];
    int nev;
    int bc;
    int max_halve;
    int dissip;
    int learned;
    int integ;                     // FIVO_SSPRK2 or FIVO_SSPRK3
    int max_mood;                  // MOOD retries per stage (SSPRK2; 0 with SSPRK3)
    MlpDev mlp;
    double cfl, cfltau;
};
// phases of one replay of the step graph; a MOOD retry of stage s is phase
// s + PH_RETRY * attempt (attempt >= 1)
enum { PH_INIT = 0, PH_STAGE1 = 1, PH_STAGE2 = 2, PH_STAGE3 = 3, PH_REFRESH = 4, PH_RETRY = 8 };
// host side of an ensemble
struct Batch {
    std::vector<Ctx*> evs;          // one-event contexts: state, work arrays, scalars
    int nev = 0;
    Params p0;                      // model part shared by the events (grid of event 0)
    int dev = 0;
    hipStream_t stream = nullptr;
    EvView* d_ev = nullptr;
    EvCtl* d_ctl = nullptr;
    EvCtl* h_ctl = nullptr;         // pinned mirror
    BatchScalars* bs = nullptr;
    BatchScalars* bs_host = nullptr;
    int* d_map[KB_N] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int nblocks[KB_N] = {0, 0, 0, 0, 0};
    std::vector<EvView> h_ev;
    int tile = -1;                  // face tile index chosen by the autotuner
    int coop_lanes = 8;             // lanes per queued cell in kb_coop (8 or 16)
    float tile_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_tile = -1, graph_dissip = -1;
    int shear_project = 0;          // project_shear! corrects pi^eta_eta
    BatchRun run{};
    std::string err;
};
hipError_t launch_batch_bc(Batch& b, int phase, int role);
hipError_t launch_batch_prim(Batch& b, int phase);
hipError_t launch_batch_faces(Batch& b, int phase, int tile);
hipError_t launch_batch_pending(Batch& b, int phase);
hipError_t launch_batch_expand(Batch& b, int phase, int round);   // one round of expand_bad_2d!
hipError_t launch_batch_history(Batch& b);
hipError_t launch_batch_relax(Batch& b);
hipError_t launch_batch_project_shear(Batch& b);
hipError_t launch_batch_ctl(Batch& b, int what);
MlpDev mlp_dev(const fivo_mlp* m);
int batch_face_blocks(const Params& p, int tile, int* tiles_y, int* nmain);
constexpr int BATCH_NTILES = 4;    // face tile candidates of the autotuner
void batch_tile_shape(int tile, int* tx, int* ty);

}  // namespace fivo
