// =============================================================================
// fivo_batch.hip -- host side of an ensemble of events (C ABI fivo_batch_*).
//
// Each event is a one-event context (fivo_create), so its state, work arrays
// and scalars live where the one-event API expects them. The ensemble adds the
// per-event views and control state on the device, the block -> event tables
// of every grid kind, and the step graph: one replay of the captured graph
// advances every running event by one attempted SSPRK3 step (accept, or halve
// dt for the next replay), relaxes the accepted ones and decides which events
// are done. The host only replays the graph and polls one counter.
// =============================================================================
#include <cmath>
#include <cstring>
#include <limits>
#include "fivo_ctx.h"

using namespace fivo;

namespace {
thread_local std::string g_berr;

int bfail(Batch* b, int code, const std::string& msg) {
    if (b) b->err = msg; else g_berr = msg;
    return code;
}
#define BCHK(b, expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return bfail(b, FIVO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

// per-event block counts of every grid kind, then the block -> event tables
int build_maps(Batch& b) {
    std::vector<int> maps[KB_N];
    for (int k = 0; k < KB_N; ++k) b.nblocks[k] = 0;
    for (int e = 0; e < b.nev; ++e) {
        EvView& v = b.h_ev[e];
        const Params& p = v.p;
        v.nb[KB_CELLS] = nblk(p.Ntot, 256);
        v.nb[KB_INTERIOR] = nblk((int64_t)p.Nx * p.Ny, 256);
        v.nb[KB_GHOST] = nblk(2 * (int64_t)p.ng * p.Nytot + 2 * (int64_t)p.ng * p.Nx, 256);
        v.nb[KB_FACES] = batch_face_blocks(p, b.tile, &v.tiles_y, &v.nmain);
        // 64 / coop_lanes cells in flight per 64-lane block; one group per queued
        // cell of a warm pass (the queue holds ~0.5 % of the cells), so no group
        // solves two cells one after the other; blocks past the queue exit at once
        v.nb[KB_COOP] = std::max(4, std::min(1024, p.Ntot / 64));
        for (int k = 0; k < KB_N; ++k) {
            v.b0[k] = b.nblocks[k];
            b.nblocks[k] += v.nb[k];
            maps[k].insert(maps[k].end(), v.nb[k], e);
        }
    }
    for (int k = 0; k < KB_N; ++k) {
        if (b.d_map[k]) BCHK(&b, hipFree(b.d_map[k]));
        BCHK(&b, hipMalloc(&b.d_map[k], sizeof(int) * maps[k].size()));
        BCHK(&b, hipMemcpy(b.d_map[k], maps[k].data(), sizeof(int) * maps[k].size(), hipMemcpyHostToDevice));
        b.run.map[k] = b.d_map[k];
    }
    BCHK(&b, hipMemcpy(b.d_ev, b.h_ev.data(), sizeof(EvView) * b.nev, hipMemcpyHostToDevice));
    return FIVO_OK;
}

int sync_events(Batch& b) {
    for (Ctx* c : b.evs) BCHK(&b, hipStreamSynchronize(c->stream));
    return FIVO_OK;
}

// sum of D over the interior, max |u|, min P + Pi (the run's diagnostics), per event
int event_diag(Batch& b, Ctx& c, int what, double* q, double* maxu, double* minp) {
    BCHK(&b, launch_reduce_diag(c, c.U[c.cur], what));
    std::vector<double> h(3 * (size_t)c.npartials);
    BCHK(&b, hipMemcpyAsync(h.data(), c.partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
    BCHK(&b, hipStreamSynchronize(c.stream));
    double s = 0.0, mu = 0.0, mp = std::numeric_limits<double>::infinity();
    for (int k = 0; k < c.npartials; ++k) {
        s += h[3 * k];
        mu = h[3 * k + 1] > mu ? h[3 * k + 1] : mu;
        mp = h[3 * k + 2] < mp ? h[3 * k + 2] : mp;
    }
    *q = s;
    if (maxu) *maxu = mu;
    if (minp) *minp = mp;
    return FIVO_OK;
}

// the launches of one replay: every running event attempts one step
int record_step(Batch& b) {
    BCHK(&b, launch_batch_ctl(b, 1));
    BCHK(&b, launch_batch_history(b));
    const int ns = b.run.integ == FIVO_SSPRK2 ? 2 : 3;
    for (int s = 1; s <= ns; ++s) {
        if (s == 1) BCHK(&b, launch_batch_bc(b, PH_STAGE1, 0));
        BCHK(&b, launch_batch_prim(b, s));
        BCHK(&b, launch_batch_faces(b, s, b.tile));
        BCHK(&b, launch_batch_bc(b, s, s));
        BCHK(&b, launch_batch_pending(b, s));
        BCHK(&b, launch_batch_ctl(b, 10 + s));
        // MOOD retries (SSPRK2): every launch of retry a runs only the events whose
        // check asked for it; the mask is the stage's bad mask expanded a times
        for (int a = 1; a <= b.run.max_mood; ++a) {
            const int ph = s + PH_RETRY * a;
            for (int r = 0; r < a; ++r) BCHK(&b, launch_batch_expand(b, ph, r));
            BCHK(&b, launch_batch_prim(b, ph));
            BCHK(&b, launch_batch_faces(b, ph, b.tile));
            BCHK(&b, launch_batch_bc(b, ph, s));
            BCHK(&b, launch_batch_pending(b, ph));
            BCHK(&b, launch_batch_ctl(b, 100 * a + 10 + s));
        }
    }
    if (b.run.dissip) {
        BCHK(&b, launch_batch_bc(b, PH_REFRESH, 0));
        BCHK(&b, launch_batch_prim(b, PH_REFRESH));
        BCHK(&b, launch_batch_relax(b));
        BCHK(&b, launch_batch_ctl(b, 20));
    }
    if (b.p0.enable_shear) {
        BCHK(&b, launch_batch_project_shear(b));
        BCHK(&b, launch_batch_ctl(b, -2));
    }
    BCHK(&b, launch_batch_ctl(b, -1));
    BCHK(&b, launch_batch_ctl(b, 21));
    return FIVO_OK;
}

int capture_step(Batch& b) {
    if (b.gexec) { BCHK(&b, hipGraphExecDestroy(b.gexec)); b.gexec = nullptr; }
    if (b.graph) { BCHK(&b, hipGraphDestroy(b.graph)); b.graph = nullptr; }
    BCHK(&b, hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    int rc = record_step(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b.stream, &g);
    if (rc != FIVO_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return bfail(&b, FIVO_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b.graph = g;
    BCHK(&b, hipGraphInstantiate(&b.gexec, b.graph, nullptr, nullptr, 0));
    return FIVO_OK;
}

// the primitive pass at tau of every event (the autotuner's input state)
int prim_at(Batch& b, double tau) {
    for (int e = 0; e < b.nev; ++e) {
        EvCtl& c = b.h_ctl[e];
        std::memset(&c, 0, sizeof(EvCtl));
        c.tau = tau; c.D = 0.0; c.cur = b.evs[e]->cur; c.state = EV_RUNNING; c.fresh = 1;
    }
    BCHK(&b, hipMemcpyAsync(b.d_ctl, b.h_ctl, sizeof(EvCtl) * b.nev, hipMemcpyHostToDevice, b.stream));
    BCHK(&b, launch_batch_ctl(b, 2));
    BCHK(&b, launch_batch_bc(b, PH_INIT, 0));
    BCHK(&b, launch_batch_prim(b, PH_INIT));
    return FIVO_OK;
}
}  // namespace

extern "C" {

const char* fivo_batch_last_error(const fivo_batch* bp) {
    return bp ? reinterpret_cast<const Batch*>(bp)->err.c_str() : g_berr.c_str();
}

int fivo_batch_create(const fivo_grid_desc* grids, int32_t nev, const fivo_model_desc* m, fivo_batch** out) {
    if (!grids || !m || !out || nev < 1) return bfail(nullptr, FIVO_ERR_ARG, "batch_create: bad argument");
    *out = nullptr;
    Batch* b = new Batch();
    auto bail = [&](int code, std::string msg) {
        fivo_batch_destroy(reinterpret_cast<fivo_batch*>(b));
        return bfail(nullptr, code, msg);
    };
    b->nev = nev;
    for (int e = 0; e < nev; ++e) {
        fivo_ctx* c = nullptr;
        int rc = fivo_create(&grids[e], m, &c);
        if (rc != FIVO_OK) return bail(rc, std::string("event ") + std::to_string(e) + ": " + fivo_last_error(nullptr));
        b->evs.push_back(reinterpret_cast<Ctx*>(c));
    }
    b->p0 = b->evs[0]->p;
    b->shear_project = m->shear_constraint_project ? 1 : 0;
    hipError_t e;
    if ((e = hipGetDevice(&b->dev)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&b->d_ev, sizeof(EvView) * nev)) != hipSuccess ||
        (e = hipMalloc(&b->d_ctl, sizeof(EvCtl) * nev)) != hipSuccess ||
        (e = hipHostMalloc(&b->h_ctl, sizeof(EvCtl) * nev, hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc(&b->bs, sizeof(BatchScalars))) != hipSuccess ||
        (e = hipHostMalloc(&b->bs_host, sizeof(BatchScalars), hipHostMallocDefault)) != hipSuccess)
        return bail(FIVO_ERR_HIP, hipGetErrorString(e));
    b->h_ev.resize(nev);
    int64_t cells = 0;
    for (int k = 0; k < nev; ++k) {
        Ctx& c = *b->evs[k];
        EvView& v = b->h_ev[k];
        std::memset(&v, 0, sizeof(EvView));
        v.p = c.p; v.w = c.w; v.sc = c.sc; v.partials = c.partials; v.jobs = c.jobs; v.wlist = c.wlist;
        for (int r = 0; r < 4; ++r) v.U[r] = c.U[r];
        cells += (int64_t)c.p.Nx * c.p.Ny;
    }
    // until the autotuner runs: the one-event choice for the mean event size
    b->tile = (cells / nev >= 300 * 300) ? 2 : 0;
    // Up to ~1M cells the queue of a pass (~0.5 % of the cells) leaves most SIMDs
    // idle, so the pass is latency-bound and 16 lanes per cell (the Jacobian
    // evaluated ahead) win; past that the queue fills the GPU and 8 lanes, twice
    // the cells in flight, win (measured: 64 events of 128-384^2, 32 of 384^2).
    b->coop_lanes = cells <= ((int64_t)1 << 20) ? 16 : 8;
    b->run.p0 = b->p0;
    b->run.ev = b->d_ev; b->run.ctl = b->d_ctl; b->run.bs = b->bs; b->run.nev = nev;
    int rc = build_maps(*b);
    if (rc != FIVO_OK) return bail(rc, b->err);
    *out = reinterpret_cast<fivo_batch*>(b);
    return FIVO_OK;
}

void fivo_batch_destroy(fivo_batch* bp) {
    if (!bp) return;
    Batch* b = reinterpret_cast<Batch*>(bp);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (int k = 0; k < KB_N; ++k) if (b->d_map[k]) (void)hipFree(b->d_map[k]);
    if (b->d_ev) (void)hipFree(b->d_ev);
    if (b->d_ctl) (void)hipFree(b->d_ctl);
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    if (b->bs) (void)hipFree(b->bs);
    if (b->bs_host) (void)hipHostFree(b->bs_host);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (Ctx* c : b->evs) fivo_destroy(reinterpret_cast<fivo_ctx*>(c));
    delete b;
}

fivo_ctx* fivo_batch_event(fivo_batch* bp, int32_t ev) {
    Batch* b = reinterpret_cast<Batch*>(bp);
    if (!b || ev < 0 || ev >= b->nev) return nullptr;
    return reinterpret_cast<fivo_ctx*>(b->evs[ev]);
}

int fivo_batch_set_tile(fivo_batch* bp, int32_t tile) {
    Batch& b = *reinterpret_cast<Batch*>(bp);
    if (tile < 0 || tile >= BATCH_NTILES) return bfail(&b, FIVO_ERR_ARG, "batch_set_tile: no such tile");
    b.tile = tile;
    return build_maps(b);
}

int fivo_batch_autotune(fivo_batch* bp, double tau, int32_t* tx, int32_t* ty, float* ms) {
    Batch& b = *reinterpret_cast<Batch*>(bp);
    int rc;
    if ((rc = sync_events(b)) != FIVO_OK) return rc;
    // the timing needs real primitives, but a primitive pass moves the Newton warm
    // starts: keep them in each event's snapshot buffer (nv >= 4 rows) and put them
    // back afterwards, so tuning leaves the solver state as it found it
    auto x0_copy = [&](bool save) -> int {
        for (Ctx* c : b.evs) {
            double* x0[4] = {c->w.x0_yT, c->w.x0_phi, c->w.x0_ux, c->w.x0_uy};
            for (int k = 0; k < 4; ++k) {
                double* keep = c->snap + (size_t)k * c->p.pitch;
                BCHK(&b, hipMemcpyAsync(save ? keep : x0[k], save ? x0[k] : keep, sizeof(double) * c->p.Ntot,
                                        hipMemcpyDeviceToDevice, b.stream));
            }
        }
        return FIVO_OK;
    };
    if ((rc = x0_copy(true)) != FIVO_OK) return rc;
    if ((rc = prim_at(b, tau)) != FIVO_OK) return rc;
    hipEvent_t t0, t1;
    BCHK(&b, hipEventCreate(&t0));
    BCHK(&b, hipEventCreate(&t1));
    int best = -1;
    float best_ms = 0.0f;
    for (int t = 0; t < BATCH_NTILES; ++t) {
        b.tile = t;
        if ((rc = build_maps(b)) != FIVO_OK) return rc;
        BCHK(&b, launch_batch_faces(b, PH_INIT, t));     // warm-up launch
        BCHK(&b, hipEventRecord(t0, b.stream));
        const int reps = 3;
        for (int r = 0; r < reps; ++r) BCHK(&b, launch_batch_faces(b, PH_INIT, t));
        BCHK(&b, hipEventRecord(t1, b.stream));
        BCHK(&b, hipEventSynchronize(t1));
        float m = 0.0f;
        BCHK(&b, hipEventElapsedTime(&m, t0, t1));
        b.tile_ms[t] = m / reps;
        if (ms) ms[t] = b.tile_ms[t];
        if (best < 0 || b.tile_ms[t] < best_ms) { best = t; best_ms = b.tile_ms[t]; }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    b.tile = best;
    if ((rc = build_maps(b)) != FIVO_OK) return rc;
    if ((rc = x0_copy(false)) != FIVO_OK) return rc;
    int X, Y;
    batch_tile_shape(best, &X, &Y);
    if (tx) *tx = X;
    if (ty) *ty = Y;
    BCHK(&b, hipStreamSynchronize(b.stream));
    return FIVO_OK;
}

int fivo_batch_run(fivo_batch* bp, const fivo_run_opts* o, const double* tau0, fivo_run_result* res) {
    Batch& b = *reinterpret_cast<Batch*>(bp);
    if (!o || !res) return bfail(&b, FIVO_ERR_ARG, "batch_run: null argument");
    if (o->integrator != FIVO_SSPRK3 && o->integrator != FIVO_SSPRK2)
        return bfail(&b, FIVO_ERR_ARG, "batch_run: integrator must be SSPRK2 or SSPRK3");
    if (o->max_mood < 0 || o->max_mood > 4) return bfail(&b, FIVO_ERR_ARG, "batch_run: max_mood must be 0..4");
    if (o->on_dump) return bfail(&b, FIVO_ERR_ARG, "batch_run: no dump callback for ensembles");
    int rc;
    if ((rc = sync_events(b)) != FIVO_OK) return rc;
    const fivo_model_desc& md = b.evs[0]->md;
    const bool dissip = md.enable_shear || md.enable_bulk || md.enable_diff;
    std::vector<double> Q0(b.nev);
    for (int e = 0; e < b.nev; ++e) {
        Ctx& c = *b.evs[e];
        EvCtl& k = b.h_ctl[e];
        std::memset(&k, 0, sizeof(EvCtl));
        BCHK(&b, hipMemsetAsync(c.sc, 0, sizeof(DevScalars), c.stream));
        int have_hist = 0;
        if (o->reset_history) {
            if ((rc = fivo_reset_history(reinterpret_cast<fivo_ctx*>(&c))) < 0) return bfail(&b, rc, c.err);
        } else {
            double a0;
            BCHK(&b, hipMemcpyAsync(&a0, c.w.alpha_prev, sizeof(double), hipMemcpyDeviceToHost, c.stream));
            BCHK(&b, hipStreamSynchronize(c.stream));
            have_hist = std::isfinite(a0) ? 1 : 0;
        }
        if ((rc = event_diag(b, c, 0, &Q0[e], nullptr, nullptr)) != FIVO_OK) return rc;
        k.tau = tau0 ? tau0[e] : o->tau0; k.D = 0.0; k.taufinal = o->taufinal; k.cur = c.cur; k.state = EV_RUNNING; k.fresh = 1;
        k.have_hist = have_hist; k.max_steps = o->max_steps;
    }
    if ((rc = sync_events(b)) != FIVO_OK) return rc;
    b.run.bc = o->bc; b.run.max_halve = o->max_halve; b.run.dissip = dissip ? 1 : 0;
    // step_ssprk3_2d! halves only; step_ssprk2_2d! retries a stage first (MOOD)
    b.run.integ = o->integrator;
    b.run.max_mood = o->integrator == FIVO_SSPRK2 ? o->max_mood : 0;
    b.run.cfl = o->cfl; b.run.cfltau = o->cfltau;
    BCHK(&b, hipMemcpyAsync(b.d_ctl, b.h_ctl, sizeof(EvCtl) * b.nev, hipMemcpyHostToDevice, b.stream));
    // the initial RHS at tau0 (run_sim_2d!, main2D.jl:384-400): amax for the first dt
    BCHK(&b, launch_batch_ctl(b, 2));
    BCHK(&b, launch_batch_bc(b, PH_INIT, 0));
    BCHK(&b, launch_batch_prim(b, PH_INIT));
    BCHK(&b, launch_batch_faces(b, PH_INIT, b.tile));
    BCHK(&b, launch_batch_ctl(b, 0));
    if ((rc = capture_step(b)) != FIVO_OK) return rc;
    // replay until no event runs; each replay ends by counting the running events
    const int64_t max_replays = (o->max_steps > 0 ? o->max_steps : 1) * (int64_t)(o->max_halve + 2) + 16;
    int64_t replays = 0;
    for (;;) {
        for (int r = 0; r < 2; ++r) BCHK(&b, hipGraphLaunch(b.gexec, b.stream));
        replays += 2;
        BCHK(&b, hipMemcpyAsync(b.bs_host, b.bs, sizeof(BatchScalars), hipMemcpyDeviceToHost, b.stream));
        BCHK(&b, hipStreamSynchronize(b.stream));
        if (b.bs_host->nrunning == 0) break;
        if (replays > max_replays) return bfail(&b, FIVO_ERR_STATE, "batch_run: events still running");
    }
    BCHK(&b, hipMemcpy(b.h_ctl, b.d_ctl, sizeof(EvCtl) * b.nev, hipMemcpyDeviceToHost));
    int status = FIVO_OK;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int e = 0; e < b.nev; ++e) {
        Ctx& c = *b.evs[e];
        const EvCtl& k = b.h_ctl[e];
        fivo_run_result& r = res[e];
        std::memset(&r, 0, sizeof(r));
        c.cur = k.cur;
        c.amax_host = k.amax;
        BCHK(&b, hipMemcpy(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost));
        r.tau = k.tau; r.nsteps = k.nsteps; r.nprimfail = k.nprimfail; r.nvacuum = k.nvacuum;
        r.max_shear_res = k.max_shear; r.Q0 = Q0[e];
        r.nrhs = k.nrhs; r.nhalvings = k.nhalvings; r.nmood = k.nmood;
        r.newton_iters = (int64_t)c.sc_host->newton_iters;
        r.nlearned = (int64_t)c.sc_host->nlearned;
        r.nlearned_fallback = (int64_t)c.sc_host->nfallback;
        r.nqueued = (int64_t)c.sc_host->nqueued;
        r.nretry = (int64_t)c.sc_host->nretry;
        r.nstaged = (int64_t)c.sc_host->nstaged;
        if (k.state == EV_FAILED) {
            r.ok = 0;
            r.maxu = r.dQ = r.Q1 = r.minPtot = nan;
            status = FIVO_STEP_FAILED;
            continue;
        }
        double Q1, maxu, minp;
        if ((rc = event_diag(b, c, 1, &Q1, &maxu, &minp)) != FIVO_OK) return rc;
        r.ok = 1; r.maxu = maxu; r.Q1 = Q1; r.minPtot = minp;
        r.dQ = std::fabs(Q0[e]) > 0 ? std::fabs(Q1 - Q0[e]) / std::fabs(Q0[e]) : 0.0;
    }
    return status;
}

int fivo_batch_set_learned_seed(fivo_batch* bp, const fivo_mlp* m) {
    Batch& b = *reinterpret_cast<Batch*>(bp);
    b.run.learned = m ? 1 : 0;
    if (m) b.run.mlp = mlp_dev(m);
    for (Ctx* c : b.evs) fivo_set_learned_seed(reinterpret_cast<fivo_ctx*>(c), m);
    return FIVO_OK;
}

}  // extern "C"
