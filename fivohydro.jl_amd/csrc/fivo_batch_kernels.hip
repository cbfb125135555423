// =============================================================================
// fivo_batch_kernels.hip -- the ensemble kernels: one launch runs one phase of
// run_sim_2d!'s step (main2D.jl:371-561) for every event of an ensemble.
//
// A block finds its event through the block -> event table of its grid kind
// and runs the same body as the one-event kernel (fivo_cells.h) on that event's
// arrays. The per-event control state (EvCtl) says whether the event takes
// part in a phase (running, stage passed, step accepted), which buffers are the
// stage input and output, and the stage time; the control kernels at the end of
// this file advance it on the device, so a whole step of every event is a fixed
// sequence of launches that is captured once in a hipGraph and replayed.
//
// Event tables and control state are read through the constant address space:
// they do not change while a kernel runs, and the loads become scalar loads.
// =============================================================================
#include "fivo_cells.h"

namespace fivo {
using namespace fd;

typedef const __attribute__((address_space(4))) EvView* CEv;
typedef const __attribute__((address_space(4))) EvCtl* CCtl;
typedef const __attribute__((address_space(4))) int* CInt;

// an event's Params: the shared model part with the event's grid
__device__ __forceinline__ Params ev_params(const BatchRun& R, CEv E) {
    Params p = R.p0;
    p.Nx = E->p.Nx; p.Ny = E->p.Ny; p.Nxtot = E->p.Nxtot; p.Nytot = E->p.Nytot; p.Ntot = E->p.Ntot;
    p.pitch = E->p.pitch; p.dx = E->p.dx; p.dy = E->p.dy; p.xmin = E->p.xmin; p.ymin = E->p.ymin;
    return p;
}
__device__ __forceinline__ DevWork ev_work(CEv E) {
    DevWork w;
    w.yT = E->w.yT; w.phi = E->w.phi; w.mu = E->w.mu; w.alpha = E->w.alpha; w.ux = E->w.ux; w.uy = E->w.uy;
    w.P = E->w.P; w.n = E->w.n; w.e = E->w.e; w.vx = E->w.vx; w.vy = E->w.vy;
    w.ok = E->w.ok; w.bad = E->w.bad; w.force = E->w.force; w.ftmp = E->w.ftmp;
    w.x0_yT = E->w.x0_yT; w.x0_phi = E->w.x0_phi; w.x0_ux = E->w.x0_ux; w.x0_uy = E->w.x0_uy;
    w.ux_prev = E->w.ux_prev; w.uy_prev = E->w.uy_prev; w.alpha_prev = E->w.alpha_prev; w.T_prev = E->w.T_prev;
    w.nux_prev = E->w.nux_prev; w.nuy_prev = E->w.nuy_prev; w.shear_res = E->w.shear_res;
    w.gradAx = E->w.gradAx; w.gradAy = E->w.gradAy;
    return w;
}

__device__ __forceinline__ int block_event(const BatchRun& R, int kind) {
    return __builtin_amdgcn_readfirstlane(((CInt)R.map[kind])[blockIdx.x]);
}
// a phase is a stage and, for a MOOD retry, its attempt (phase = stage + PH_RETRY attempt)
__device__ __forceinline__ int ph_stage(int phase) { return phase & (PH_RETRY - 1); }
__device__ __forceinline__ int ph_attempt(int phase) { return phase / PH_RETRY; }
__device__ __forceinline__ bool ev_active(CCtl c, int phase) {
    const int ps = ph_stage(phase);
    if (ps == PH_REFRESH) return c->accepted != 0;
    return c->state == EV_RUNNING && c->stage == (ps == PH_INIT ? 0 : ps - 1) && c->mood_attempt == ph_attempt(phase);
}
// stage input role (offset from the accepted buffer) and the time of its primitive pass
__device__ __forceinline__ int phase_in_role(int phase) {
    const int ps = ph_stage(phase);
    return ps == PH_STAGE2 ? 1 : (ps == PH_STAGE3 ? 2 : 0);
}
__device__ __forceinline__ double phase_tau(CCtl c, int phase) {
    const int ps = ph_stage(phase);
    return ps == PH_STAGE2 ? c->tau + c->D : (ps == PH_STAGE3 ? c->tau + 0.5 * c->D : c->tau);
}

// ---- ghost fill of buffer `role` -------------------------------------------------------
__global__ __launch_bounds__(256) void kb_bc(BatchRun R, int phase, int role) {
    const int e = block_event(R, KB_GHOST);
    CCtl c = (CCtl)(R.ctl + e);
    if (!ev_active(c, phase)) return;
    CEv E = (CEv)(R.ev + e);
    const Params p = ev_params(R, E);
    const int t = (blockIdx.x - E->b0[KB_GHOST]) * 256 + threadIdx.x;
    if (t < bc_nghost(p)) bc_cell(E->U[(c->cur + role) & 3], p, R.bc == FIVO_BC_PERIODIC, t);
}

// ---- primitive pass: warm part, one lane per cell --------------------------------------
// first launch: every cell's warm start is tested; converged and vacuum cells are
// finished, the cells whose Newton iterates are listed (per event)
template <int EK, bool NCH>
__global__ __launch_bounds__(256) void kb_prim_check(BatchRun R, int phase) {
    const int e = block_event(R, KB_CELLS);
    CCtl c = (CCtl)(R.ctl + e);
    if (!ev_active(c, phase)) return;
    CEv E = (CEv)(R.ev + e);
    const Params p = ev_params(R, E);
    const DevWork w = ev_work(E);
    const double* U = E->U[(c->cur + phase_in_role(phase)) & 3];
    const double tau = phase_tau(c, phase);
    DevScalars* sc = E->sc;
    const int i = (blockIdx.x - E->b0[KB_CELLS]) * 256 + threadIdx.x;
    unsigned long long its = 0, nfail = 0, nvac = 0, nq = 0, nre = 0;
    bool iterate = false;
    if (i < p.Ntot)
        prim_cell<EK, NCH, true>(U, p, tau, w, i, e, &sc->ncold, E->jobs, its, nfail, nvac, nq, nre, &iterate,
                                 nullptr, &sc->nresq, p.Ntot);
    wave_append(iterate, i, &sc->nwarm, E->wlist);
    prim_counts(sc, its, nfail, nvac, nq);
    rare_counts(sc, nre, 0);
}
// second launch: the event's listed cells, densely packed into wavefronts (the
// list is at most the event's cell count, so one cell per thread; blocks past it exit)
template <int EK, bool NCH>
__global__ __launch_bounds__(256) void kb_prim(BatchRun R, int phase) {
    const int e = block_event(R, KB_CELLS);
    CCtl c = (CCtl)(R.ctl + e);
    if (!ev_active(c, phase)) return;
    CEv E = (CEv)(R.ev + e);
    DevScalars* sc = E->sc;
    const int n = (int)sc->nwarm;
    const int k0 = (blockIdx.x - E->b0[KB_CELLS]) * 256;
    if (k0 >= n) return;
    const Params p = ev_params(R, E);
    const DevWork w = ev_work(E);
    const double* U = E->U[(c->cur + phase_in_role(phase)) & 3];
    const double tau = phase_tau(c, phase);
    const int k = k0 + threadIdx.x;
    unsigned long long its = 0, nfail = 0, nvac = 0, nq = 0, nre = 0;
    if (k < n)
        prim_cell<EK, NCH>(U, p, tau, w, E->wlist[k], e, &sc->ncold, E->jobs, its, nfail, nvac, nq, nre, nullptr,
                           nullptr, &sc->nresq, p.Ntot);
    prim_counts(sc, its, nfail, nvac, nq);
    rare_counts(sc, nre, 0);
}

// ---- primitive pass: the event's queued cells, G lanes per cell -----------------------------
// G = 8 (a large ensemble): two wavefronts per SIMD (registers capped at 256, a few
// spilled), since its queue holds tens of thousands of cells and occupancy beats
// per-wave latency. G = 16 (a small one, a short queue): one wavefront per SIMD,
// so the registers past 256 live in AGPRs instead of scratch (1.4 % on a 512^2
// event's batched + learned step).
template <int EK, bool NCH, int G>
__global__ __launch_bounds__(64, G == 8 ? 2 : 1) void kb_coop(BatchRun R, int phase) {
    const int e = block_event(R, KB_COOP);
    CCtl c = (CCtl)(R.ctl + e);
    if (!ev_active(c, phase)) return;
    CEv E = (CEv)(R.ev + e);
    // seeded solves at the queue's front, resumed ones at its far end (prim_cell)
    const int nf = (int)E->sc->ncold, n = nf + (int)E->sc->nresq;
    const int lb = blockIdx.x - E->b0[KB_COOP];
    const int ngroups = E->nb[KB_COOP] * (64 / G);
    int t = (lb * 64 + (int)threadIdx.x) / G;
    if (__builtin_amdgcn_readfirstlane(lb * (64 / G)) >= n) return;
    const Params p = ev_params(R, E);
    const DevWork w = ev_work(E);
    const double* U = E->U[(c->cur + phase_in_role(phase)) & 3];
    const double tau = phase_tau(c, phase);
    const PrimJob* jobs = E->jobs;
    unsigned long long its = 0, nfail = 0, nvac = 0, nfb = 0, nre = 0, nst = 0;
    for (; t < n; t += ngroups) {
        int itsum;
        coop_job<EK, NCH, G>(U, p, tau, w, t < nf ? jobs[t] : jobs[p.Ntot - 1 - (t - nf)], itsum, nfail, nvac, nfb,
                             nre, nst);
        if (glane<G>() == 0) its += itsum;
    }
    prim_counts(E->sc, its, nfail, nvac);
    if (glane<G>() == 0 && nfb) atomicAdd(&E->sc->nfallback, nfb);
    const bool lead = glane<G>() == 0;
    rare_counts(E->sc, lead ? nre : 0, lead ? nst : 0);
}

// ---- learned seeds of the event's queued cold cells, 16 per wavefront ---------------------
template <int EK, bool NCH>
__global__ __launch_bounds__(64) void kb_seed(BatchRun R, int phase) {
    const int e = block_event(R, KB_COOP);
    CCtl c = (CCtl)(R.ctl + e);
    if (!ev_active(c, phase)) return;
    CEv E = (CEv)(R.ev + e);
    const int n = (int)E->sc->ncold;
    const int lb = blockIdx.x - E->b0[KB_COOP];
    if (lb * 16 >= n) return;
    const Params p = ev_params(R, E);
    const DevWork w = ev_work(E);
    const double* U = E->U[(c->cur + phase_in_role(phase)) & 3];
    const double tau = phase_tau(c, phase);
    unsigned long long used = 0;
    for (int tile = lb; tile * 16 < n; tile += E->nb[KB_COOP])
        used += seed_tile<EK, NCH>(U, p, tau, w, E->jobs, n, tile, R.mlp);
    used = wave_sum(used);
    if (threadIdx.x == 0 && used) atomicAdd(&E->sc->nlearned, used);
}

// ---- the fused face / stage-update kernel --------------------------------------------
// PH_INIT: the RHS at tau0 (its amax gives the first dt; output to the scratch
// buffer); stage s of SSPRK3 (step_ssprk3_2d!, src2d/timestepper2d.jl:148-188):
//   1: U1 = U + D k(U, tau)            2: U2 = 3/4 U + 1/4 (U1 + D k(U1, tau + D))
//   3: U  = 1/3 U + 2/3 (U2 + D k(U2, tau + D/2)), each sanitised at tau + D
template <int EK, bool NCH, int TX, int TY>
__global__ __launch_bounds__((face_threads<TX, TY>())) void kb_faces(BatchRun R, int phase) {
    extern __shared__ double sF[];
    const int e = block_event(R, KB_FACES);
    CCtl c = (CCtl)(R.ctl + e);
    if (!ev_active(c, phase)) return;
    CEv E = (CEv)(R.ev + e);
    FaceK K;
    K.p = ev_params(R, E);
    K.w = ev_work(E);
    K.sc = E->sc;
    K.tiles_y = E->tiles_y;
    K.nmain = E->nmain;
    K.remap = 0;
    K.pass = 0;           // every tile: inside a captured step the kernels run in stream order
    K.dirty = nullptr;
    const int cur = c->cur;
    const int s = ph_stage(phase) == PH_INIT ? 1 : ph_stage(phase), attempt = ph_attempt(phase);
    FaceLaunch& f = K.f;
    f.Ub = E->U[(cur + s - 1) & 3];
    f.Uold = E->U[cur & 3];
    f.Uout = E->U[(cur + s) & 3];
    f.dUout = nullptr;
    f.mode = 1;
    f.combine = s - 1;
    f.tau = phase_tau(c, phase);
    f.tau_state = c->tau + c->D;
    f.D = c->D;
    // step_ssprk3_2d!: 3/4, 1/4 in stage 2; step_ssprk2_2d!: 1/2, 1/2 (timestepper2d.jl:74-188)
    f.wold = s == 2 ? (R.integ == FIVO_SSPRK2 ? 0.5 : 0.75) : (s == 1 ? 1.0 : 0.0);
    f.wnew = s == 2 ? (R.integ == FIVO_SSPRK2 ? 0.5 : 0.25) : 0.0;
    // a retry takes first-order faces on the mask expanded `attempt` times (kb_expand)
    f.force = attempt == 0 ? nullptr : ((attempt & 1) ? K.w.force : K.w.ftmp);
    faces_body<EK, NCH, TX, TY, 1>(K, blockIdx.x - E->b0[KB_FACES], sF);
}

// ---- expand_bad_2d! (src2d/floors2d.jl:369-383): one round of the MOOD mask's growth -------
// round 0 reads the stage's bad mask; the rounds alternate between force and ftmp,
// so after `attempt` rounds the mask is in force (odd) or ftmp (even)
__global__ __launch_bounds__(256) void kb_expand(BatchRun R, int phase, int round) {
    const int e = block_event(R, KB_CELLS);
    CCtl c = (CCtl)(R.ctl + e);
    if (!ev_active(c, phase)) return;
    CEv E = (CEv)(R.ev + e);
    const Params p = ev_params(R, E);
    const int t = (blockIdx.x - E->b0[KB_CELLS]) * 256 + threadIdx.x;
    if (t >= p.Ntot) return;
    const uint8_t* in = round == 0 ? E->w.bad : (((round - 1) & 1) == 0 ? E->w.force : E->w.ftmp);
    uint8_t* out = (round & 1) == 0 ? E->w.force : E->w.ftmp;
    const int ix = t / p.Nytot, iy = t - ix * p.Nytot;
    uint8_t v = in[t];
    if (!v && ix >= p.ng && ix < p.ng + p.Nx && iy >= p.ng && iy < p.ng + p.Ny)
        v = in[t - 1] | in[t + 1] | in[t - p.Nytot] | in[t + p.Nytot];
    out[t] = v ? 1 : 0;
}

// ---- near-bound cells of a stage output: cold inversion ---------------------------------
template <int EK, bool NCH>
__global__ __launch_bounds__(256) void kb_pending(BatchRun R, int phase) {
    const int e = block_event(R, KB_INTERIOR);
    CCtl c = (CCtl)(R.ctl + e);
    if (!ev_active(c, phase)) return;
    CEv E = (CEv)(R.ev + e);
    if (E->sc->npending == 0) return;
    const Params p = ev_params(R, E);
    const int s = ph_stage(phase) == PH_INIT ? 1 : ph_stage(phase);
    pending_cell<EK, NCH>(E->U[(c->cur + s) & 3], p, c->tau + c->D, (uint8_t*)E->w.bad, E->sc,
                          (blockIdx.x - E->b0[KB_INTERIOR]) * 256 + threadIdx.x);
}

// ---- history snapshot before a step (main2D.jl:458-463) ----------------------------------
__global__ __launch_bounds__(256) void kb_history(BatchRun R) {
    const int e = block_event(R, KB_CELLS);
    CCtl c = (CCtl)(R.ctl + e);
    if (!(c->state == EV_RUNNING && c->dohist)) return;
    CEv E = (CEv)(R.ev + e);
    const int t = (blockIdx.x - E->b0[KB_CELLS]) * 256 + threadIdx.x;
    if (t < E->p.Ntot) history_cell(ev_work(E), t);
}

// ---- project_shear! per event, block partials ----------------------------------------
__global__ __launch_bounds__(256) void kb_project_shear(BatchRun R, int correct) {
    const int e = block_event(R, KB_INTERIOR);
    CCtl c = (CCtl)(R.ctl + e);
    if (!c->accepted) return;
    CEv E = (CEv)(R.ev + e);
    const Params p = ev_params(R, E);
    const int lb = blockIdx.x - E->b0[KB_INTERIOR];
    project_shear_block(E->U[c->cur & 3], p, ev_work(E), correct, lb * 256 + threadIdx.x, E->partials + 2 * lb);
}

// ---- control: one thread per event -------------------------------------------------------
// what: 2 run start, 0 after the initial RHS, 1 step begin, 10+s after stage s,
// 20 after relaxation, 21 step end
__device__ __forceinline__ void reset_pass(DevScalars* sc) {
    sc->amax_bits = 0; sc->nbad = 0; sc->nprimfail = 0; sc->nvacuum = 0; sc->npending = 0; sc->ncold = 0; sc->nresq = 0; sc->nwarm = 0;
}
__global__ __launch_bounds__(64) void kb_ctl(BatchRun R, int what) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= R.nev) return;
    EvCtl& c = R.ctl[e];
    const EvView& E = R.ev[e];
    DevScalars* sc = E.sc;
    if (what == 0) {
        // initial RHS done: its amax sets the first dt (run_sim_2d!, main2D.jl:384-400)
        c.amax = __longlong_as_double((long long)sc->amax_bits);
        c.nrhs += 1;
        c.stage = 0;
        reset_pass(sc);
        return;
    }
    if (what == 2) {   // run start: the initial RHS follows
        c.stage = 0;
        reset_pass(sc);
        return;
    }
    if (what == 1) {
        c.accepted = 0;
        c.dohist = 0;
        if (c.state != EV_RUNNING) return;
        if (c.fresh) {
            if (!(c.tau < c.taufinal && c.nsteps < c.max_steps)) { c.state = EV_DONE; return; }
            // compute_dt_2d (src2d/timestepper2d.jl:37-47)
            const double amax = c.amax > 1e-6 ? c.amax : 1e-6;
            const double dcfl = R.cfl * (E.p.dx < E.p.dy ? E.p.dx : E.p.dy) / amax;
            const double dtau = R.cfltau * c.tau;
            double D = dcfl < dtau ? dcfl : dtau;
            const double rem = c.taufinal - c.tau;
            D = D < rem ? D : rem;
            if (D <= 0) { c.state = EV_DONE; return; }
            c.D = D;
            c.nhalve = 0;
            if (c.have_hist || c.nsteps > 0) { c.dohist = 1; c.have_hist = 1; }
        }
        c.stage = 0;
        c.mood_attempt = 0;
        reset_pass(sc);
        return;
    }
    if (what % 100 >= 11 && what % 100 <= 13) {
        // stage check; what = 10 + stage + 100 attempt
        const int s = what % 100 - 10, attempt = what / 100, ns = R.integ == FIVO_SSPRK2 ? 2 : 3;
        if (!(c.state == EV_RUNNING && c.stage == s - 1 && c.mood_attempt == attempt)) return;
        c.nrhs += 1;
        if (sc->nbad == 0) {
            c.stage = s;
            c.mood_attempt = 0;
            if (s == ns) c.amax = __longlong_as_double((long long)sc->amax_bits);
            // without relaxation the step's primitive counts are those of its last pass
            if (s == ns && !R.dissip) { c.nprimfail += (long long)sc->nprimfail; c.nvacuum += (long long)sc->nvacuum; }
        } else if (attempt < R.max_mood) {
            // _stage_2d! (timestepper2d.jl:113-140): redo the stage with first-order
            // faces on the bad mask expanded 1 + attempt times
            c.mood_attempt = attempt + 1;
            c.nmood += 1;
        } else {
            c.stage = -1;   // attempt failed: no later stage of this replay runs
            c.mood_attempt = 0;
        }
        if (c.stage == ns || c.stage < 0) {
            // step outcome: accept, or halve dt and retry next replay
            if (c.stage == ns) {
                c.cur = (c.cur + ns) & 3;
                c.tau += c.D;
                c.nsteps += 1;
                c.accepted = 1;
                c.fresh = 1;
            } else {
                c.nhalvings += 1;
                c.nhalve += 1;
                c.D *= 0.5;
                c.fresh = 0;
                if (c.nhalve > R.max_halve) c.state = EV_FAILED;
            }
        }
        reset_pass(sc);
        return;
    }
    if (what == 20) {   // relaxation wrote U[(cur+1)%4]
        if (c.accepted && R.dissip) {
            c.cur = (c.cur + 1) & 3;
            c.nprimfail += (long long)sc->nprimfail;
            c.nvacuum += (long long)sc->nvacuum;
        }
        return;
    }
    if (what == 21) {   // step end: shear residual of the accepted step
        if (c.accepted && E.p.enable_shear) {
            const double m = c.shear_step;
            c.max_shear = m > c.max_shear ? m : c.max_shear;
        }
        if (c.state == EV_RUNNING && !(c.tau < c.taufinal && c.nsteps < c.max_steps) && c.fresh) c.state = EV_DONE;
        if (c.state == EV_RUNNING) atomicAdd(&R.bs->nrunning, 1ull);
        return;
    }
}
__global__ void kb_zero_running(BatchRun R) { R.bs->nrunning = 0; }

// the maximum of an event's per-block shear residuals after the projection, one
// wavefront per event (the maximum of a set: the same value in any order; NaN
// partials are skipped, like the sequential `a > m ? a : m`)
__global__ __launch_bounds__(64) void kb_shear_max(BatchRun R) {
    const int e = blockIdx.x;
    EvCtl& c = R.ctl[e];
    const EvView& E = R.ev[e];
    double m = 0.0;
    if (c.accepted) {
        const int nbi = E.nb[KB_INTERIOR];
        for (int b = threadIdx.x; b < nbi; b += 64) m = E.partials[2 * b] > m ? E.partials[2 * b] : m;
    }
    m = wave_max_pos(m);
    if (threadIdx.x == 0) c.shear_step = m;
}

// ---- launchers -------------------------------------------------------------------------------
#define EOS_DISPATCH_B(b, BODY)                                                      \
    do {                                                                             \
        if ((b).p0.eos.kind == 1) { constexpr int EK = 1; constexpr bool NCH = true; BODY; } \
        else if ((b).p0.eos.has_charge) { constexpr int EK = 2; constexpr bool NCH = true; BODY; } \
        else { constexpr int EK = 0; constexpr bool NCH = false; BODY; }             \
    } while (0)

void batch_tile_shape(int tile, int* tx, int* ty) {
    static const int T[BATCH_NTILES][2] = {{7, 8}, {8, 28}, {12, 20}, {15, 16}};
    *tx = T[tile][0]; *ty = T[tile][1];
}
int batch_face_blocks(const Params& p, int tile, int* tiles_y, int* nmain) {
    int nb = 0;
    switch (tile) {
        case 0: face_grid<7, 8>(p, tiles_y, nmain, &nb); break;
        case 1: face_grid<8, 28>(p, tiles_y, nmain, &nb); break;
        case 2: face_grid<12, 20>(p, tiles_y, nmain, &nb); break;
        default: face_grid<15, 16>(p, tiles_y, nmain, &nb); break;
    }
    return nb;
}

hipError_t launch_batch_bc(Batch& b, int phase, int role) {
    hipLaunchKernelGGL(kb_bc, dim3(b.nblocks[KB_GHOST]), dim3(256), 0, b.stream, b.run, phase, role);
    return hipGetLastError();
}
hipError_t launch_batch_prim(Batch& b, int phase) {
    EOS_DISPATCH_B(b, hipLaunchKernelGGL((kb_prim_check<EK, NCH>), dim3(b.nblocks[KB_CELLS]), dim3(256), 0,
                                         b.stream, b.run, phase));
    EOS_DISPATCH_B(b, hipLaunchKernelGGL((kb_prim<EK, NCH>), dim3(b.nblocks[KB_CELLS]), dim3(256), 0, b.stream,
                                         b.run, phase));
    if (b.run.learned)
        EOS_DISPATCH_B(b, hipLaunchKernelGGL((kb_seed<EK, NCH>), dim3(b.nblocks[KB_COOP]), dim3(64), 0, b.stream,
                                             b.run, phase));
    if (b.coop_lanes == 16)
        EOS_DISPATCH_B(b, hipLaunchKernelGGL((kb_coop<EK, NCH, 16>), dim3(b.nblocks[KB_COOP]), dim3(64), 0,
                                             b.stream, b.run, phase));
    else
        EOS_DISPATCH_B(b, hipLaunchKernelGGL((kb_coop<EK, NCH, 8>), dim3(b.nblocks[KB_COOP]), dim3(64), 0,
                                             b.stream, b.run, phase));
    return hipGetLastError();
}
template <int EK, bool NCH, int TX, int TY>
static hipError_t batch_faces_tx(Batch& b, int phase) {
    constexpr int NF = (TX + 1) * TY + TX * (TY + 1);
    const int nvf = b.p0.iPQ0 >= 0 ? b.p0.iPQ0 : b.p0.nv;
    const size_t lds = sizeof(double) * (size_t)nvf * NF;
    static size_t lds_set = 0;   // the attribute is set once, before any graph capture
    if (lds > 65536 && lds > lds_set) {
        hipError_t e = hipFuncSetAttribute((const void*)kb_faces<EK, NCH, TX, TY>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        lds_set = lds;
    }
    hipLaunchKernelGGL((kb_faces<EK, NCH, TX, TY>), dim3(b.nblocks[KB_FACES]), dim3(face_threads<TX, TY>()), lds, b.stream, b.run,
                       phase);
    return hipGetLastError();
}
hipError_t launch_batch_faces(Batch& b, int phase, int tile) {
    hipError_t e = hipSuccess;
    switch (tile) {
        case 0: EOS_DISPATCH_B(b, e = (batch_faces_tx<EK, NCH, 7, 8>(b, phase))); break;
        case 1: EOS_DISPATCH_B(b, e = (batch_faces_tx<EK, NCH, 8, 28>(b, phase))); break;
        case 2: EOS_DISPATCH_B(b, e = (batch_faces_tx<EK, NCH, 12, 20>(b, phase))); break;
        default: EOS_DISPATCH_B(b, e = (batch_faces_tx<EK, NCH, 15, 16>(b, phase))); break;
    }
    return e;
}
hipError_t launch_batch_expand(Batch& b, int phase, int round) {
    hipLaunchKernelGGL(kb_expand, dim3(b.nblocks[KB_CELLS]), dim3(256), 0, b.stream, b.run, phase, round);
    return hipGetLastError();
}
hipError_t launch_batch_pending(Batch& b, int phase) {
    EOS_DISPATCH_B(b, hipLaunchKernelGGL((kb_pending<EK, NCH>), dim3(b.nblocks[KB_INTERIOR]), dim3(256), 0,
                                         b.stream, b.run, phase));
    return hipGetLastError();
}
hipError_t launch_batch_history(Batch& b) {
    hipLaunchKernelGGL(kb_history, dim3(b.nblocks[KB_CELLS]), dim3(256), 0, b.stream, b.run);
    return hipGetLastError();
}
hipError_t launch_batch_project_shear(Batch& b) {
    hipLaunchKernelGGL(kb_project_shear, dim3(b.nblocks[KB_INTERIOR]), dim3(256), 0, b.stream, b.run,
                       b.shear_project);
    return hipGetLastError();
}
// what >= 0: kb_ctl's phases; -1 clears the running count; -2 the shear residual maxima
hipError_t launch_batch_ctl(Batch& b, int what) {
    if (what == -1) hipLaunchKernelGGL(kb_zero_running, dim3(1), dim3(1), 0, b.stream, b.run);
    else if (what == -2) hipLaunchKernelGGL(kb_shear_max, dim3(b.nev), dim3(64), 0, b.stream, b.run);
    else hipLaunchKernelGGL(kb_ctl, dim3((b.nev + 63) / 64), dim3(64), 0, b.stream, b.run, what);
    return hipGetLastError();
}

}  // namespace fivo
