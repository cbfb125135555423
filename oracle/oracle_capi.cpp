// =============================================================================
// oracle/oracle_capi.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
// C entry points of the CPU oracle, prefixed `orc_`, with the same signatures as
// the library's `fivo_*` ABI so tests can drive both through one harness. Also
// exports single-cell functions (EOS, Bessel, recovery) for unit checks.
// =============================================================================
#include <cstring>
#include <omp.h>
#include "oracle_solver.hpp"

using namespace orc;

static const double NaNv = std::numeric_limits<double>::quiet_NaN();

static void alloc_work(Ctx& c) {
    const size_t N = c.g.Ntot, nv = c.m.L.nv;
    Work& w = c.w;
    for (auto* v : {&w.yT, &w.phi, &w.mu, &w.alpha, &w.ux, &w.uy, &w.P, &w.n, &w.e, &w.vxC, &w.vyC,
                    &w.theta, &w.slope_tmp, &w.shear_res})
        v->assign(N, 0.0);
    for (auto* v : {&w.x0_yT, &w.x0_phi, &w.x0_ux, &w.x0_uy, &w.ux_prev, &w.uy_prev, &w.alpha_prev,
                    &w.T_prev, &w.nux_prev, &w.nuy_prev})
        v->assign(N, NaNv);
    w.ok.assign(N, 0); w.bad.assign(N, 0); w.bad_tmp.assign(N, 0);
    for (int d = 0; d < 2; ++d) {
        w.ULp[d].assign(4 * N, 0.0); w.URp[d].assign(4 * N, 0.0); w.sig[d].assign(4 * N, 0.0);
        w.ULc[d].assign(nv * N, 0.0); w.URc[d].assign(nv * N, 0.0); w.Fh[d].assign(nv * N, 0.0);
        w.alphaF[d].assign(N, 0.0); w.gradA[d].assign(N, 0.0);
    }
    for (auto* v : {&w.S, &w.k, &w.U1, &w.U2, &w.diss_snap}) v->assign(nv * N, 0.0);
}

static void reset_history(Ctx& c) {
    Work& w = c.w;
    for (auto* v : {&w.x0_yT, &w.x0_phi, &w.x0_ux, &w.x0_uy, &w.ux_prev, &w.uy_prev, &w.alpha_prev,
                    &w.T_prev, &w.nux_prev, &w.nuy_prev})
        std::fill(v->begin(), v->end(), NaNv);
}

extern "C" {

int orc_create(const fivo_grid_desc* gd, const fivo_model_desc* md, Ctx** out) {
    *out = nullptr;
    if (!gd || !md || gd->nx <= 8 || gd->ny <= 8 || gd->nghost != 3) return FIVO_ERR_ARG;
    Ctx* c = new Ctx();
    c->g.make(gd->nx, gd->ny, gd->xmax, gd->ymax, gd->nghost);
    c->m.d = *md;
    c->m.eos = EOS(md->eos);
    c->m.L.make(md->with_charge != 0, md->enable_bulk != 0, md->enable_shear != 0, md->consistent_m2 != 0);
    // derived vacuum cut (main2D.jl:100-105)
    double Evac = md->E_vac_cut;
    if (Evac <= 0.0) Evac = eos_Pne(md->T_vac_cut, 0.0, c->m.eos).e;
    c->m.Evac = Evac;
    c->U.assign((size_t)c->m.L.nv * c->g.Ntot, 0.0);
    alloc_work(*c);
    *out = c;
    return FIVO_OK;
}
void orc_destroy(Ctx* c) { delete c; }
int orc_nvars(const Ctx* c) { return c->m.L.nv; }
int64_t orc_ntot(const Ctx* c) { return c->g.Ntot; }
double orc_e_vac_cut(const Ctx* c) { return c->m.Evac; }
int orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); return omp_get_max_threads(); }

int orc_set_state(Ctx* c, const double* U, int nv, int64_t ntot) {
    if (nv != c->m.L.nv || ntot != c->g.Ntot) return FIVO_ERR_ARG;
    std::memcpy(c->U.data(), U, sizeof(double) * nv * ntot);
    return FIVO_OK;
}
int orc_get_state(Ctx* c, double* U, int nv, int64_t ntot) {
    if (nv != c->m.L.nv || ntot != c->g.Ntot) return FIVO_ERR_ARG;
    std::memcpy(U, c->U.data(), sizeof(double) * nv * ntot);
    return FIVO_OK;
}

static void cp_out(double* dst, const std::vector<double>& src) {
    if (dst) std::memcpy(dst, src.data(), sizeof(double) * src.size());
}
static void cp_in(std::vector<double>& dst, const double* src) {
    if (src) std::memcpy(dst.data(), src, sizeof(double) * dst.size());
}
int orc_get_work(Ctx* c, fivo_work_view* v) {
    Work& w = c->w;
    cp_out(v->yT, w.yT); cp_out(v->phi, w.phi); cp_out(v->mu, w.mu); cp_out(v->alpha, w.alpha);
    cp_out(v->ux, w.ux); cp_out(v->uy, w.uy); cp_out(v->P, w.P); cp_out(v->n, w.n); cp_out(v->e, w.e);
    if (v->ok) std::memcpy(v->ok, w.ok.data(), w.ok.size());
    cp_out(v->x0_yT, w.x0_yT); cp_out(v->x0_phi, w.x0_phi); cp_out(v->x0_ux, w.x0_ux); cp_out(v->x0_uy, w.x0_uy);
    cp_out(v->ux_prev, w.ux_prev); cp_out(v->uy_prev, w.uy_prev); cp_out(v->alpha_prev, w.alpha_prev);
    cp_out(v->T_prev, w.T_prev); cp_out(v->nux_prev, w.nux_prev); cp_out(v->nuy_prev, w.nuy_prev);
    cp_out(v->shear_res, w.shear_res);
    return FIVO_OK;
}
int orc_set_work(Ctx* c, const fivo_work_view* v) {
    Work& w = c->w;
    cp_in(w.yT, v->yT); cp_in(w.phi, v->phi); cp_in(w.mu, v->mu); cp_in(w.alpha, v->alpha);
    cp_in(w.ux, v->ux); cp_in(w.uy, v->uy); cp_in(w.P, v->P); cp_in(w.n, v->n); cp_in(w.e, v->e);
    if (v->ok) std::memcpy(w.ok.data(), v->ok, w.ok.size());
    cp_in(w.x0_yT, v->x0_yT); cp_in(w.x0_phi, v->x0_phi); cp_in(w.x0_ux, v->x0_ux); cp_in(w.x0_uy, v->x0_uy);
    cp_in(w.ux_prev, v->ux_prev); cp_in(w.uy_prev, v->uy_prev); cp_in(w.alpha_prev, v->alpha_prev);
    cp_in(w.T_prev, v->T_prev); cp_in(w.nux_prev, v->nux_prev); cp_in(w.nuy_prev, v->nuy_prev);
    cp_in(w.shear_res, v->shear_res);
    return FIVO_OK;
}
int orc_reset_history(Ctx* c) { reset_history(*c); return FIVO_OK; }
// diagnostics: enable per-cell Newton iteration recording; copy the histogram out
void orc_debug_iters(Ctx* c, int enable, int64_t* hist256) {
    if (enable) { c->w.dbg_iters.assign(c->g.Ntot, -1); c->w.dbg_class.assign(c->g.Ntot, 0); }
    else { c->w.dbg_iters.clear(); c->w.dbg_class.clear(); }
    if (hist256) for (int k = 0; k < 256; ++k) hist256[k] = k < (int)c->w.dbg_hist.size() ? c->w.dbg_hist[k] : 0;
}
void orc_debug_hist2(Ctx* c, int64_t* h) {
    for (int k = 0; k < 2 * 8 * 256; ++k) h[k] = k < (int)c->w.dbg_hist2.size() ? c->w.dbg_hist2[k] : 0;
}
// expand_bad_2d! on a host mask of Ntot bytes, in place (unit checks of the MOOD dilation)
void orc_debug_expand(Ctx* c, uint8_t* bad, int radius) {
    std::copy(bad, bad + c->g.Ntot, c->w.bad.begin());
    expand_bad(*c, radius);
    std::copy(c->w.bad.begin(), c->w.bad.end(), bad);
}
// per-cell iterations and class of the last primitive pass (recording enabled)
void orc_debug_cells(Ctx* c, int32_t* iters, int32_t* cls) {
    const size_t n = c->w.dbg_iters.size();
    for (size_t i = 0; i < n; ++i) { iters[i] = c->w.dbg_iters[i]; cls[i] = c->w.dbg_class[i]; }
}

// finalize_ic! (main2D.jl:164-171)
int orc_finalize_ic(Ctx* c, double tau0, int bc) {
    enforce_floors(*c, c->U, E_FLOOR);
    enforce_S_energy_constraint(*c, c->U, CHI_SRE, nullptr);
    enforce_nu_charge_constraint(*c, c->U, tau0, 0.999);
    apply_bc(*c, c->U, bc);
    return FIVO_OK;
}

// set_cell! over every cell (main2D.jl:142-154), then finalize_ic!
int orc_init_from_prims(Ctx* c, double tau0, const double* T, const double* alpha, const double* ux,
                        const double* uy, const double* Pi, const double* pixx, const double* pixy,
                        const double* piyy, const double* nux, const double* nuy, int bc, int64_t* nbad) {
    const Layout& L = c->m.L;
    const int nv = L.nv;
    int64_t bad = 0;
    for (int i = 0; i < c->g.Ntot; ++i) {
        double vx = ux ? ux[i] : 0.0, vy = uy ? uy[i] : 0.0;
        Diss d{nux ? nux[i] : 0.0, nuy ? nuy[i] : 0.0, Pi ? Pi[i] : 0.0,
               pixx ? pixx[i] : 0.0, pixy ? pixy[i] : 0.0, piyy ? piyy[i] : 0.0, 0.0};
        double ut = std::sqrt(1 + vx * vx + vy * vy);
        d.pieta = shear_project_pieta(vx, vy, ut, d.pixx, d.pixy, d.piyy);
        // prim_to_cons_2d! (src2d/primrec2d.jl:753-796)
        double Tc = jmax(T[i], T_MIN), mu = alpha[i] * T[i];
        PNE r = eos_Pne_2d(Tc, mu, c->m.eos);
        if (!(isf(r.P) && isf(r.n) && isf(r.e)) || r.e <= 0.0) { bad += 1; continue; }
        double* Uc = &c->U[nv * (size_t)i];
        double weff = r.e + r.P + d.Pi, Ptot = r.P + d.Pi;
        Pi4 p = shear_tau_row(vx, vy, ut, d.pixx, d.pixy, d.piyy);
        Uc[L.iD] = tau0 * (r.n * ut + nu_tau(vx, vy, ut, d.nux, d.nuy));
        Uc[L.iSx] = weff * ut * vx + p.tx;
        Uc[L.iSy] = weff * ut * vy + p.ty;
        Uc[L.iE] = weff * ut * ut - Ptot + p.tt;
        if (L.hasNu) { Uc[L.iNux] = d.nux; Uc[L.iNuy] = d.nuy; }
        if (L.hasPi) Uc[L.iPi] = d.Pi;
        if (L.hasShear) { Uc[L.iPixx] = d.pixx; Uc[L.iPixy] = d.pixy; Uc[L.iPiyy] = d.piyy; Uc[L.iPieta] = d.pieta; }
    }
    if (nbad) *nbad = bad;
    return orc_finalize_ic(c, tau0, bc);
}

int orc_apply_bc(Ctx* c, int bc) { apply_bc(*c, c->U, bc); return FIVO_OK; }

int orc_rhs(Ctx* c, double tau, int bc, const uint8_t* force, double* dU, fivo_counts* cnt) {
    rhs(*c, c->w.k, c->U, tau, bc, force);
    if (dU) std::memcpy(dU, c->w.k.data(), sizeof(double) * c->w.k.size());
    if (cnt) {
        cnt->nprimfail = c->w.primfail; cnt->nvacuum = c->w.vacuum; cnt->amax = c->w.amax;
        cnt->nretry = c->w.nretry; cnt->nstaged = c->w.nstaged;
    }
    return FIVO_OK;
}
// update_primitives_2d! (src2d/rhs2d.jl:57-65)
int orc_update_primitives(Ctx* c, double tau, int bc, fivo_counts* cnt) {
    apply_bc(*c, c->U, bc);
    primitive_pass(*c, c->U, tau);
    if (cnt) {
        cnt->nprimfail = c->w.primfail; cnt->nvacuum = c->w.vacuum; cnt->amax = c->w.amax;
        cnt->nretry = c->w.nretry; cnt->nstaged = c->w.nstaged;
    }
    return FIVO_OK;
}
// export_background2d's slice loop (tools/export_background2d.jl:79-94): T = exp(yT),
// v^i = u^i / u^tau, n of every interior cell at (i, j, k) of column-major (Nx, Ny, nt)
// tables; max |u| and max |v| with Julia's max (a NaN propagates). Host tables only.
int orc_background_slice(Ctx* c, int32_t k, int32_t nt, double* T, double* vx, double* vy, double* n,
                         int32_t /*device*/, double* maxima) {
    if (!T || !vx || !vy || nt < 1 || k < 0 || k >= nt) return FIVO_ERR_ARG;
    const Grid& g = c->g;
    const Work& w = c->w;
    const auto jmax = [](double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NaNv : (a > b ? a : b); };
    double maxu = 0.0, maxv = 0.0;
    for (int i = 0; i < g.Nx; ++i)
        for (int j = 0; j < g.Ny; ++j) {
            const int l = g.lin(i + g.ng, j + g.ng);
            const double ux = w.ux[l], uy = w.uy[l], ut = std::sqrt(1 + ux * ux + uy * uy);
            const size_t o = (size_t)i + (size_t)g.Nx * ((size_t)j + (size_t)g.Ny * k);
            T[o] = std::exp(w.yT[l]);
            vx[o] = ux / ut;
            vy[o] = uy / ut;
            if (n) n[o] = w.n[l];
            maxu = jmax(maxu, std::hypot(ux, uy));
            maxv = jmax(maxv, std::hypot(vx[o], vy[o]));
        }
    if (maxima) { maxima[0] = maxu; maxima[1] = maxv; }
    return FIVO_OK;
}
int orc_sanitize(Ctx* c, double tau, int bc) { sanitize_stage(*c, c->U, tau, bc); return FIVO_OK; }
int orc_state_ok(Ctx* c, double tau, int* ok, int64_t* nbad) {
    *ok = state_ok(*c, c->U, tau, nbad) ? 1 : 0;
    return FIVO_OK;
}
int orc_compute_dt(Ctx* c, double tau, double cfl, double cfltau, double* dt) {
    *dt = compute_dt(*c, tau, cfl, cfltau);
    return FIVO_OK;
}
int orc_step(Ctx* c, int integ, double tau, double dtau, int bc, int max_halve, int max_mood, int* ok,
             double* used) {
    bool r = integ == FIVO_SSPRK3 ? step_ssprk3(*c, tau, dtau, bc, max_halve, used)
                                  : step_ssprk2(*c, tau, dtau, bc, max_halve, max_mood, used);
    *ok = r ? 1 : 0;
    return r ? FIVO_OK : FIVO_STEP_FAILED;
}
int orc_relax(Ctx* c, double tau, double dtau) { relax_dissipative(*c, tau, dtau); return FIVO_OK; }
int orc_project_shear(Ctx* c, double* mx, double* mean) { project_shear(*c, mx, mean); return FIVO_OK; }

// run_sim_2d! (main2D.jl:371-561)
int orc_run(Ctx* c, const fivo_run_opts* o, fivo_run_result* res) {
    const Grid& g = c->g;
    const Layout& L = c->m.L;
    Work& w = c->w;
    const int nv = L.nv, ng = g.ng;
    const fivo_model_desc& d = c->m.d;
    std::memset(res, 0, sizeof(*res));
    c->nrhs = 0; c->nhalvings = 0; c->nmood = 0; w.newton_iters = 0; w.nretry_run = 0; w.nstaged_run = 0;
    if (o->reset_history) reset_history(*c);
    rhs(*c, w.k, c->U, o->tau0, o->bc, nullptr);
    double Qin = 0.0;
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) Qin += c->U[L.iD + nv * (size_t)g.lin(ix, iy)];
    double tau = o->tau0;
    int64_t nsteps = 0, nprimfail = 0, nvacuum = 0;
    double max_shear_res = 0.0, next_dump = o->tau0;
    if (o->on_dump) { o->on_dump(o->user, tau, c); next_dump = o->tau0 + o->dump_dt; }
    const bool dissip = d.enable_shear || d.enable_bulk || d.enable_diff;
    while (tau < o->taufinal && nsteps < o->max_steps) {
        double D = compute_dt(*c, tau, o->cfl, o->cfltau);
        D = jmin(D, o->taufinal - tau);
        if (D <= 0) break;
        if (isf(w.alpha_prev[0]) || nsteps > 0) {
            w.alpha_prev = w.alpha;
            for (size_t j = 0; j < w.T_prev.size(); ++j) w.T_prev[j] = std::exp(w.yT[j]);
        }
        double used;
        bool ok = o->integrator == FIVO_SSPRK3
                      ? step_ssprk3(*c, tau, D, o->bc, o->max_halve, &used)
                      : step_ssprk2(*c, tau, D, o->bc, o->max_halve, o->max_mood, &used);
        if (!ok) {
            res->ok = 0; res->tau = tau; res->nsteps = nsteps; res->nprimfail = nprimfail;
            res->nvacuum = nvacuum; res->max_shear_res = max_shear_res;
            res->maxu = res->dQ = res->Q1 = res->minPtot = NaNv; res->Q0 = Qin;
            res->nrhs = c->nrhs; res->nhalvings = c->nhalvings; res->nmood = c->nmood;
            res->newton_iters = w.newton_iters;
            res->nretry = w.nretry_run; res->nstaged = w.nstaged_run;
            return FIVO_STEP_FAILED;
        }
        tau += used;
        nsteps += 1;
        if (dissip) {
            apply_bc(*c, c->U, o->bc);
            primitive_pass(*c, c->U, tau);
            relax_dissipative(*c, tau, used);
        }
        nprimfail += w.primfail;
        nvacuum += w.vacuum;
        if (L.hasShear) {
            double mx, mean;
            project_shear(*c, &mx, &mean);
            max_shear_res = jmax(max_shear_res, mx);
        }
        if (o->on_dump && tau >= next_dump - 1e-12) { o->on_dump(o->user, tau, c); next_dump += o->dump_dt; }
    }
    if (o->on_dump) o->on_dump(o->user, tau, c);
    double minPtot = std::numeric_limits<double>::infinity(), maxu = 0.0, Qout = 0.0;
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            int i = g.lin(ix, iy);
            Qout += c->U[L.iD + nv * (size_t)i];
            double u = std::hypot(w.ux[i], w.uy[i]);
            if (u > maxu) maxu = u;
            if (w.P[i] > 0) {
                double pt = w.P[i] + (L.hasPi ? c->U[L.iPi + nv * (size_t)i] : 0.0);
                if (pt < minPtot) minPtot = pt;
            }
        }
    res->ok = 1; res->tau = tau; res->nsteps = nsteps; res->nprimfail = nprimfail; res->nvacuum = nvacuum;
    res->max_shear_res = max_shear_res; res->maxu = maxu;
    res->dQ = std::fabs(Qin) > 0 ? std::fabs(Qout - Qin) / std::fabs(Qin) : 0.0;
    res->Q0 = Qin; res->Q1 = Qout; res->minPtot = minPtot;
    res->nrhs = c->nrhs; res->nhalvings = c->nhalvings; res->nmood = c->nmood;
    res->newton_iters = w.newton_iters;
    res->nretry = w.nretry_run; res->nstaged = w.nstaged_run;
    return FIVO_OK;
}

// ---- single-cell exports for unit checks ---------------------------------------
void orc_eos_Pne(const fivo_eos_desc* e, double T, double mu, double* out3) {
    PNE r = eos_Pne(T, mu, EOS(*e));
    out3[0] = r.P; out3[1] = r.n; out3[2] = r.e;
}
void orc_eos_Pne_2d(const fivo_eos_desc* e, double T, double mu, double* out3) {
    PNE r = eos_Pne_2d(T, mu, EOS(*e));
    out3[0] = r.P; out3[1] = r.n; out3[2] = r.e;
}
double orc_eos_cs2(const fivo_eos_desc* e, double T) { return eos_cs2(T, 0.0, EOS(*e)); }
void orc_light(const fivo_eos_desc* e, double T, double* out3) {
    EOS q(*e);
    out3[0] = light_P(T, q); out3[1] = light_dP_dT(T, q); out3[2] = light_d2P_dT2(T, q);
}
double orc_besselkx(int nu, double x) { return safe_besselkx(nu, x); }
double orc_besselk2x_fast(double x) { return safe_besselk2x_2d(x); }
// prim_to_cons_2d! for one state: cons4 = (tau*J^tau, Sx, Sy, E)
int orc_prim_to_cons(const fivo_eos_desc* e, double T, double mu, double ux, double uy, const double* diss7,
                     double tau, double* cons4) {
    EOS q(*e);
    double Tc = jmax(T, T_MIN), ut = std::sqrt(1 + ux * ux + uy * uy);
    PNE r = eos_Pne_2d(Tc, mu, q);
    if (!(isf(r.P) && isf(r.n) && isf(r.e)) || r.e <= 0.0) return 0;
    Diss d{diss7[0], diss7[1], diss7[2], diss7[3], diss7[4], diss7[5], diss7[6]};
    double weff = r.e + r.P + d.Pi, Ptot = r.P + d.Pi;
    Pi4 p = shear_tau_row(ux, uy, ut, d.pixx, d.pixy, d.piyy);
    cons4[0] = tau * (r.n * ut + nu_tau(ux, uy, ut, d.nux, d.nuy));
    cons4[1] = weff * ut * ux + p.tx;
    cons4[2] = weff * ut * uy + p.ty;
    cons4[3] = weff * ut * ut - Ptot + p.tt;
    return 1;
}
// cons_to_prim_2d! for one state. out = (T, mu, ux, uy, n, e, P, ok, reason, iters, resnorm)
void orc_cons_to_prim(const fivo_eos_desc* e, const double* cons4, const double* diss7, double tau,
                      const double* warm4, double* out) {
    EOS q(*e);
    Cons cs{cons4[0], cons4[1], cons4[2], cons4[3]};
    Diss d{diss7[0], diss7[1], diss7[2], diss7[3], diss7[4], diss7[5], diss7[6]};
    PrimStat st;
    int its = 0;
    PrimOut r = cons_to_prim(cs, d, tau, q, warm4[0], warm4[1], warm4[2], warm4[3], st, &its);
    out[0] = r.T; out[1] = r.mu; out[2] = r.ux; out[3] = r.uy; out[4] = r.n; out[5] = r.e; out[6] = r.P;
    out[7] = r.ok ? 1.0 : 0.0; out[8] = st.reason; out[9] = its; out[10] = st.resnorm;
}

}  // extern "C"
