// =============================================================================
// oracle/oracle_solver.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
// CPU restatement of the 2+1D sweeps (src2d/bc2d.jl, rhs2d.jl,
// reconstruction2d.jl, fluxes2d.jl, floors2d.jl, timestepper2d.jl). OpenMP over
// x-columns with a dynamic schedule, like the reference's `:greedy` threads.
// =============================================================================
#include "oracle_solver.hpp"
#include <algorithm>
#include <cstring>

namespace orc {

static const double NaN = std::numeric_limits<double>::quiet_NaN();

// dissipatives_at (src2d/state_layout2d.jl:89-100)
static inline Diss diss_at(const Layout& L, const double* Ui) {
    Diss d;
    d.nux = L.hasNu ? Ui[L.iNux] : 0.0;
    d.nuy = L.hasNu ? Ui[L.iNuy] : 0.0;
    d.Pi = L.hasPi ? Ui[L.iPi] : 0.0;
    d.pixx = L.hasShear ? Ui[L.iPixx] : 0.0;
    d.pixy = L.hasShear ? Ui[L.iPixy] : 0.0;
    d.piyy = L.hasShear ? Ui[L.iPiyy] : 0.0;
    d.pieta = L.hasShear ? Ui[L.iPieta] : 0.0;
    return d;
}

// apply_bc_2d! (src2d/bc2d.jl:14-65)
void apply_bc(Ctx& c, std::vector<double>& U, int bc) {
    const Grid& g = c.g;
    const int ng = g.ng, nv = c.m.L.nv;
    const int ix0 = ng, ix1 = ng + g.Nx - 1, iy0 = ng, iy1 = ng + g.Ny - 1;
    auto cp = [&](int dst, int src) { for (int a = 0; a < nv; ++a) U[a + nv * dst] = U[a + nv * src]; };
    if (bc == FIVO_BC_PERIODIC) {
        for (int k = 1; k <= ng; ++k)
            for (int iy = 0; iy < g.Nytot; ++iy) {
                cp(g.lin(ng - k, iy), g.lin(ix1 + 1 - k, iy));
                cp(g.lin(ix1 + k, iy), g.lin(ix0 - 1 + k, iy));
            }
        for (int k = 1; k <= ng; ++k)
            for (int ix = 0; ix < g.Nxtot; ++ix) {
                cp(g.lin(ix, ng - k), g.lin(ix, iy1 + 1 - k));
                cp(g.lin(ix, iy1 + k), g.lin(ix, iy0 - 1 + k));
            }
        return;
    }
    for (int k = 1; k <= ng; ++k)
        for (int iy = 0; iy < g.Nytot; ++iy) {
            cp(g.lin(ng - k, iy), g.lin(ix0, iy));
            cp(g.lin(ix1 + k, iy), g.lin(ix1, iy));
        }
    for (int k = 1; k <= ng; ++k)
        for (int ix = 0; ix < g.Nxtot; ++ix) {
            cp(g.lin(ix, ng - k), g.lin(ix, iy0));
            cp(g.lin(ix, iy1 + k), g.lin(ix, iy1));
        }
}

// _primitive_pass! (src2d/rhs2d.jl:67-160)
void primitive_pass(Ctx& c, std::vector<double>& U, double tau) {
    const Grid& g = c.g;
    const Layout& L = c.m.L;
    const EOS& q = c.m.eos;
    Work& w = c.w;
    const int nv = L.nv;
    int64_t nfail = 0, nvac = 0, nit = 0, nre = 0, nst = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nfail, nvac, nit, nre, nst)
    for (int ix = 0; ix < g.Nxtot; ++ix) {
        for (int iy = 0; iy < g.Nytot; ++iy) {
            int i = g.lin(ix, iy);
            const double* Ui = &U[nv * (size_t)i];
            Cons cs{safe_div(Ui[L.iD], tau), Ui[L.iSx], Ui[L.iSy], Ui[L.iE]};
            Diss d = diss_at(L, Ui);
            PrimStat st;
            int its = 0, it1 = 0, natt = 0;
            const bool warm0 = isf(w.x0_yT[i]) && isf(w.x0_ux[i]) && isf(w.x0_uy[i]);
            PrimOut r = cons_to_prim(cs, d, tau, q, w.x0_yT[i], w.x0_phi[i], w.x0_ux[i], w.x0_uy[i],
                                     st, &its, &it1, &natt);
            nit += its;
            nre += natt == 2 ? 1 : 0;
            nst += st.staged;
            if (!w.dbg_iters.empty()) {
                w.dbg_iters[i] = (st.reason == PRR_VACUUM) ? -1 : its;
                // class: 1 warm start, 2 fresh-seed retry, 4 ok; first attempt's iterations above bit 8
                w.dbg_class[i] = (warm0 ? 1 : 0) | (natt == 2 ? 2 : 0) | (r.ok ? 4 : 0) | (it1 << 8);
            }
            double T = r.T, mu = r.mu, ux = r.ux, uy = r.uy, n = r.n, e = r.e, P = r.P;
            bool ok = r.ok;
            if (!ok) {
                if (st.reason == PRR_VACUUM) nvac += 1; else nfail += 1;
                T = T_MIN; mu = 0.0; ux = 0.0; uy = 0.0;
                PNE f = eos_Pne_2d(T, mu, q);
                P = f.P; n = f.n; e = f.e;
                e = jmax(e, 0.0); n = 0.0;
            }
            double Tm = jmax(T, T_MIN);
            w.yT[i] = std::log(Tm);
            w.phi[i] = jclamp((mu - q.mass()) / Tm, -PHI_CAP, PHI_CAP);
            w.mu[i] = mu;
            w.alpha[i] = mu / Tm;
            w.ux[i] = ux; w.uy[i] = uy;
            w.P[i] = P; w.n[i] = n; w.e[i] = e;
            w.ok[i] = ok ? 1 : 0;
            double ut = std::sqrt(1 + ux * ux + uy * uy), iut = safe_inv(ut);
            w.vxC[i] = ux * iut; w.vyC[i] = uy * iut;
            if (ok) {
                w.x0_yT[i] = w.yT[i]; w.x0_phi[i] = w.phi[i]; w.x0_ux[i] = ux; w.x0_uy[i] = uy;
            } else {
                w.x0_yT[i] = NaN; w.x0_phi[i] = NaN; w.x0_ux[i] = NaN; w.x0_uy[i] = NaN;
            }
            if (L.hasShear) w.shear_res[i] = shear_rel_residual(ux, uy, ut, d.pixx, d.pixy, d.piyy, d.pieta);
        }
    }
    w.primfail = nfail; w.vacuum = nvac; w.newton_iters += nit;
    w.nretry = nre; w.nstaged = nst; w.nretry_run += nre; w.nstaged_run += nst;
    if (!w.dbg_iters.empty()) {
        if (w.dbg_hist.size() < 256) w.dbg_hist.assign(256, 0);
        if (w.dbg_hist2.size() < 2 * 8 * 256) w.dbg_hist2.assign(2 * 8 * 256, 0);
        for (size_t i = 0; i < w.dbg_iters.size(); ++i) {
            const int v = w.dbg_iters[i];
            if (v < 0) continue;
            w.dbg_hist[v < 255 ? v : 255] += 1;
            const int cl = w.dbg_class[i] & 7, v1 = w.dbg_class[i] >> 8;
            w.dbg_hist2[cl * 256 + (v < 255 ? v : 255)] += 1;
            w.dbg_hist2[8 * 256 + cl * 256 + (v1 < 255 ? v1 : 255)] += 1;
        }
    }
}

// reconstruct_muscl_2d! (src2d/reconstruction2d.jl:27-103); dir 0 = x, 1 = y
static void reconstruct(Ctx& c, int dir) {
    const Grid& g = c.g;
    Work& w = c.w;
    const int st = dir == 0 ? g.Nytot : 1, ng = g.ng;
    std::vector<double>& sig = w.sig[dir];
    std::vector<double>& UL = w.ULp[dir];
    std::vector<double>& UR = w.URp[dir];
    const double* f[4] = {w.yT.data(), w.phi.data(), w.ux.data(), w.uy.data()};
    std::fill(sig.begin(), sig.end(), 0.0);
    int ixlo = dir == 0 ? 1 : 0, ixhi = dir == 0 ? g.Nxtot - 2 : g.Nxtot - 1;
    int iylo = dir == 0 ? 0 : 1, iyhi = dir == 0 ? g.Nytot - 1 : g.Nytot - 2;
#pragma omp parallel for schedule(static)
    for (int ix = ixlo; ix <= ixhi; ++ix)
        for (int iy = iylo; iy <= iyhi; ++iy) {
            int i = g.lin(ix, iy), im = i - st, ip = i + st;
            for (int v = 0; v < 4; ++v) sig[v + 4 * (size_t)i] = mc_limiter(f[v][i] - f[v][im], f[v][ip] - f[v][i]);
        }
    int fxhi = dir == 0 ? g.Nxtot - 2 : g.Nxtot - 1, fyhi = dir == 0 ? g.Nytot - 1 : g.Nytot - 2;
    const int kmax = dir == 0 ? g.Nxtot : g.Nytot;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix <= fxhi; ++ix)
        for (int iy = 0; iy <= fyhi; ++iy) {
            int i = g.lin(ix, iy), ip = i + st;
            int k = (dir == 0 ? ix : iy) + 1;   // 1-based index along dir of the left cell
            if (k <= ng || k >= kmax - ng) {
                for (int v = 0; v < 4; ++v) { UL[v + 4 * (size_t)i] = f[v][i]; UR[v + 4 * (size_t)i] = f[v][ip]; }
                continue;
            }
            int im = i - st, ipp = ip + st;
            for (int v = 0; v < 4; ++v) {
                double l = f[v][i] + 0.5 * sig[v + 4 * (size_t)i];
                double r = f[v][ip] - 0.5 * sig[v + 4 * (size_t)ip];
                l = jclamp(l, jmin3(f[v][im], f[v][i], f[v][ip]), jmax3(f[v][im], f[v][i], f[v][ip]));
                r = jclamp(r, jmin3(f[v][i], f[v][ip], f[v][ipp]), jmax3(f[v][i], f[v][ip], f[v][ipp]));
                UL[v + 4 * (size_t)i] = l; UR[v + 4 * (size_t)i] = r;
            }
        }
}

// prim_to_cons_2d! (src2d/primrec2d.jl:753-796); writes column Uc[0..nv) when ok
static inline bool prim_to_cons(double* Uc, double T, double mu, double ux, double uy, const Diss& d,
                                double tau, const EOS& q, const Layout& L, Prim& pr) {
    double Tc = jmax(T, T_MIN);
    double ut = std::sqrt(1 + ux * ux + uy * uy);
    PNE r = eos_Pne_2d(Tc, mu, q);
    if (!(isf(r.P) && isf(r.n) && isf(r.e)) || r.e <= 0.0) {
        pr = {Tc, mu, 0.0, 0.0, 0.0, 0.0, 0.0, d.nux, d.nuy, d.Pi, d.pixx, d.pixy, d.piyy, d.pieta, false};
        return false;
    }
    double weff = r.e + r.P + d.Pi, Ptot = r.P + d.Pi;
    Pi4 p = shear_tau_row(ux, uy, ut, d.pixx, d.pixy, d.piyy);
    double Jt = r.n * ut + nu_tau(ux, uy, ut, d.nux, d.nuy);
    Uc[L.iD] = tau * Jt;
    Uc[L.iSx] = weff * ut * ux + p.tx;
    Uc[L.iSy] = weff * ut * uy + p.ty;
    Uc[L.iE] = weff * ut * ut - Ptot + p.tt;
    if (L.hasNu) { Uc[L.iNux] = d.nux; Uc[L.iNuy] = d.nuy; }
    if (L.hasPi) Uc[L.iPi] = d.Pi;
    if (L.hasShear) { Uc[L.iPixx] = d.pixx; Uc[L.iPixy] = d.pixy; Uc[L.iPiyy] = d.piyy; Uc[L.iPieta] = d.pieta; }
    pr = {Tc, mu, ux, uy, r.n, r.e, r.P, d.nux, d.nuy, d.Pi, d.pixx, d.pixy, d.piyy, d.pieta, true};
    return true;
}

// _floor_prim_2d! (src2d/rhs2d.jl:259-270)
static inline Prim floor_prim(double* Uc, const Layout& L, double Emin) {
    for (int a = 0; a < L.nv; ++a) Uc[a] = 0.0;
    Uc[L.iE] = Emin;
    return {T_MIN, 0.0, 0.0, 0.0, 0.0, Emin, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, true};
}

// flux_cell_2d! (src2d/fluxes2d.jl:24-69)
static inline void flux_cell(double* F, const Prim& p, double tau, int dir, const Model& m) {
    const Layout& L = m.L;
    for (int a = 0; a < L.nv; ++a) F[a] = 0.0;
    double ux = p.ux, uy = p.uy;
    double ut = std::sqrt(1 + ux * ux + uy * uy), iut = safe_inv(ut);
    double vx = ux * iut, vy = uy * iut;
    double Ptot = p.P + p.Pi, weff = (p.e + p.P) + p.Pi;
    Pi4 s = shear_tau_row(ux, uy, ut, p.pixx, p.pixy, p.piyy);
    double Jx = p.n * ux + p.nux, Jy = p.n * uy + p.nuy, v;
    if (dir == 0) {
        F[L.iD] = tau * Jx;
        F[L.iSx] = weff * ux * ux + Ptot + p.pixx;
        F[L.iSy] = weff * ux * uy + p.pixy;
        F[L.iE] = weff * ut * ux + s.tx;
        v = vx;
    } else {
        F[L.iD] = tau * Jy;
        F[L.iSx] = weff * ux * uy + p.pixy;
        F[L.iSy] = weff * uy * uy + Ptot + p.piyy;
        F[L.iE] = weff * ut * uy + s.ty;
        v = vy;
    }
    if (L.hasNu && m.d.advect_nu) { F[L.iNux] = p.nux * v; F[L.iNuy] = p.nuy * v; }
    if (L.hasPi && m.d.advect_Pi) F[L.iPi] = p.Pi * v;
    if (L.hasShear && m.d.advect_pi) {
        F[L.iPixx] = p.pixx * v; F[L.iPixy] = p.pixy * v; F[L.iPiyy] = p.piyy * v; F[L.iPieta] = p.pieta * v;
    }
}

// wavespeeds_2d (src2d/fluxes2d.jl:87-114)
static inline void wavespeeds(const Prim& p, int dir, double pinn, const EOS& q, double* lm, double* lp) {
    double cs2 = jclamp(eos_cs2(p.T, p.mu, q), 0.0, 0.999);
    double wv = p.e + p.P;
    if (wv > TINY) cs2 = jmin(cs2 + (std::fabs(p.Pi) + std::fabs(pinn)) / wv, 0.999);
    double ut = std::sqrt(1 + p.ux * p.ux + p.uy * p.uy), iut = safe_inv(ut);
    double vx = p.ux * iut, vy = p.uy * iut;
    double v2 = vx * vx + vy * vy;
    double vn = dir == 0 ? vx : vy;
    double den = 1 - v2 * cs2;
    if (den <= TINY) { *lm = -1.0; *lp = 1.0; return; }
    double cs = safe_sqrt(cs2);
    double rad = (1 - v2) * (den - vn * vn * (1 - cs2));
    double disc = cs * safe_sqrt(jmax(rad, 0.0));
    double l1 = (vn * (1 - cs2) - disc) / den;
    double l2 = (vn * (1 - cs2) + disc) / den;
    *lm = jclamp(l1, -1.0, 1.0);
    *lp = jclamp(l2, -1.0, 1.0);
}

// hlle_flux_2d! (src2d/fluxes2d.jl:121-157); returns (sL, sR)
static inline void hlle(double* Fh, const double* ULc, const double* URc, const Prim& pL, const Prim& pR,
                        double tau, int dir, const Model& m, double* FL, double* FR, double* sLo, double* sRo) {
    flux_cell(FL, pL, tau, dir, m);
    flux_cell(FR, pR, tau, dir, m);
    double pinnL = dir == 0 ? pL.pixx : pL.piyy, pinnR = dir == 0 ? pR.pixx : pR.piyy;
    double lmL, lpL, lmR, lpR;
    wavespeeds(pL, dir, pinnL, m.eos, &lmL, &lpL);
    wavespeeds(pR, dir, pinnR, m.eos, &lmR, &lpR);
    double sL = jmin(lmL, lmR), sR = jmax(lpL, lpR);
    const int nv = m.L.nv;
    if (sL >= 0) { for (int a = 0; a < nv; ++a) Fh[a] = FL[a]; }
    else if (sR <= 0) { for (int a = 0; a < nv; ++a) Fh[a] = FR[a]; }
    else {
        double inv = 1 / (sR - sL + TINY);
        for (int a = 0; a < nv; ++a) Fh[a] = (sR * FL[a] - sL * FR[a] + sR * sL * (URc[a] - ULc[a])) * inv;
    }
    *sLo = sL; *sRo = sR;
}

// _faces_2d! (src2d/rhs2d.jl:282-371)
static double faces(Ctx& c, const std::vector<double>& U, double tau, int dir, double Emin,
                    const uint8_t* force) {
    const Grid& g = c.g;
    const Model& m = c.m;
    const Layout& L = m.L;
    Work& w = c.w;
    const int nv = L.nv, st = dir == 0 ? g.Nytot : 1;
    const int fxhi = dir == 0 ? g.Nxtot - 2 : g.Nxtot - 1, fyhi = dir == 0 ? g.Nytot - 1 : g.Nytot - 2;
    double* Fh = w.Fh[dir].data();
    double* ULc = w.ULc[dir].data();
    double* URc = w.URc[dir].data();
    const double* ULp = w.ULp[dir].data();
    const double* URp = w.URp[dir].data();
    const double mass = m.eos.mass();
    double amax = 0.0;
#pragma omp parallel reduction(max : amax)
    {
        std::vector<double> FL(nv), FR(nv);
        double amax_t = 0.0;
#pragma omp for schedule(dynamic, 1)
        for (int ix = 0; ix <= fxhi; ++ix)
            for (int iy = 0; iy <= fyhi; ++iy) {
                int i = g.lin(ix, iy), ip = i + st;
                Diss dL = diss_at(L, &U[nv * (size_t)i]), dR = diss_at(L, &U[nv * (size_t)ip]);
                bool forced = force && (force[i] || force[ip]);
                bool good = false;
                Prim pL, pR;
                double* uL = ULc + nv * (size_t)i;
                double* uR = URc + nv * (size_t)i;
                if (!forced) {
                    const double* l = ULp + 4 * (size_t)i;
                    const double* r = URp + 4 * (size_t)i;
                    double TL = std::exp(l[0]), muL = mass + TL * l[1];
                    double TR = std::exp(r[0]), muR = mass + TR * r[1];
                    bool okL = prim_to_cons(uL, TL, muL, l[2], l[3], dL, tau, m.eos, L, pL);
                    bool okR = prim_to_cons(uR, TR, muR, r[2], r[3], dR, tau, m.eos, L, pR);
                    good = okL && okR && uL[L.iD] >= 0.0 && uR[L.iD] >= 0.0 && uL[L.iE] >= Emin && uR[L.iE] >= Emin;
                }
                if (!good) {
                    double TLc = std::exp(w.yT[i]), muLc = w.mu[i];
                    double TRc = std::exp(w.yT[ip]), muRc = w.mu[ip];
                    bool okL = prim_to_cons(uL, TLc, muLc, w.ux[i], w.uy[i], dL, tau, m.eos, L, pL);
                    bool okR = prim_to_cons(uR, TRc, muRc, w.ux[ip], w.uy[ip], dR, tau, m.eos, L, pR);
                    if (!(okL && okR)) {
                        if (!okL) pL = floor_prim(uL, L, Emin);
                        if (!okR) pR = floor_prim(uR, L, Emin);
                    }
                }
                double sL, sR;
                hlle(Fh + nv * (size_t)i, uL, uR, pL, pR, tau, dir, m, FL.data(), FR.data(), &sL, &sR);
                double a = jmax(std::fabs(sL), std::fabs(sR));
                if (a > amax_t) amax_t = a;
            }
        amax = std::max(amax, amax_t);
    }
    return amax;
}

// rhs_2d! (src2d/rhs2d.jl:169-249)
void rhs(Ctx& c, std::vector<double>& dU, std::vector<double>& U, double tau, int bc,
         const uint8_t* force) {
    const Grid& g = c.g;
    const Layout& L = c.m.L;
    Work& w = c.w;
    const int nv = L.nv, ng = g.ng, Nyt = g.Nytot;
    c.nrhs += 1;
    apply_bc(c, U, bc);
    primitive_pass(c, U, tau);
    reconstruct(c, 0);
    reconstruct(c, 1);
    std::fill(w.Fh[0].begin(), w.Fh[0].end(), 0.0);
    std::fill(w.Fh[1].begin(), w.Fh[1].end(), 0.0);
    double ax = faces(c, U, tau, 0, E_FLOOR, force);
    double ay = faces(c, U, tau, 1, E_FLOOR, force);
    w.amax = std::max(ax, ay);
    // sources (src2d/rhs2d.jl:200-231, src2d/fluxes2d.jl:166-177)
    const fivo_model_desc& md = c.m.d;
    const bool adv_nu = L.hasNu && md.advect_nu, adv_Pi = L.hasPi && md.advect_Pi,
               adv_pi = L.hasShear && md.advect_pi;
    const bool need_div = adv_nu || adv_Pi || adv_pi;
    std::fill(w.theta.begin(), w.theta.end(), 0.0);
    if (need_div) {
        // compute_divv_2d! (src2d/rhs2d.jl:29-41)
        const double invdx = 1 / g.dx, invdy = 1 / g.dy;
#pragma omp parallel for schedule(static)
        for (int ix = 1; ix < g.Nxtot - 1; ++ix)
            for (int iy = 1; iy < g.Nytot - 1; ++iy) {
                const int i = g.lin(ix, iy);
                w.theta[i] = 0.5 * (w.vxC[i + Nyt] - w.vxC[i - Nyt]) * invdx + 0.5 * (w.vyC[i + 1] - w.vyC[i - 1]) * invdy;
            }
    }
    std::fill(w.S.begin(), w.S.end(), 0.0);
    const double itau = safe_inv(tau);
#pragma omp parallel for schedule(static)
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            int i = g.lin(ix, iy);
            const double* Ui = &U[nv * (size_t)i];
            Diss d = diss_at(L, Ui);
            double* S = &w.S[nv * (size_t)i];
            S[L.iSx] = -Ui[L.iSx] * itau;
            S[L.iSy] = -Ui[L.iSy] * itau;
            S[L.iE] = -(Ui[L.iE] + (w.P[i] + d.Pi) + d.pieta) * itau;
            if (need_div) {
                const double divv = w.theta[i];
                if (adv_nu) { S[L.iNux] += Ui[L.iNux] * divv; S[L.iNuy] += Ui[L.iNuy] * divv; }
                if (adv_Pi) S[L.iPi] += Ui[L.iPi] * divv;
                if (adv_pi) {
                    S[L.iPixx] += Ui[L.iPixx] * divv; S[L.iPixy] += Ui[L.iPixy] * divv;
                    S[L.iPiyy] += Ui[L.iPiyy] * divv; S[L.iPieta] += Ui[L.iPieta] * divv;
                }
            }
        }
    std::fill(dU.begin(), dU.end(), 0.0);
    const double idx = 1 / g.dx, idy = 1 / g.dy;
    const double* Fx = w.Fh[0].data();
    const double* Fy = w.Fh[1].data();
#pragma omp parallel for schedule(static)
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            int i = g.lin(ix, iy), ixm = i - Nyt;
            for (int a = 0; a < nv; ++a)
                dU[a + nv * (size_t)i] = -(Fx[a + nv * (size_t)i] - Fx[a + nv * (size_t)ixm]) * idx -
                                         (Fy[a + nv * (size_t)i] - Fy[a + nv * (size_t)(i - 1)]) * idy +
                                         w.S[a + nv * (size_t)i];
        }
}

// enforce_floors_2d! (src2d/floors2d.jl:24-104)
void enforce_floors(Ctx& c, std::vector<double>& U, double Emin) {
    const Grid& g = c.g;
    const Layout& L = c.m.L;
    const int nv = L.nv, ng = g.ng;
    const double Evac = jmax(Emin, c.m.Evac);
    const double rdom = c.m.d.r_domain;
    const double rdom2 = isf(rdom) ? rdom * rdom : std::numeric_limits<double>::infinity();
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            double* Ui = &U[nv * (size_t)g.lin(ix, iy)];
            auto blank = [&]() {
                Ui[L.iSx] = 0.0; Ui[L.iSy] = 0.0; Ui[L.iD] = 0.0;
                if (L.hasNu) { Ui[L.iNux] = 0.0; Ui[L.iNuy] = 0.0; }
                if (L.hasPi) Ui[L.iPi] = 0.0;
                if (L.hasShear) { Ui[L.iPixx] = 0.0; Ui[L.iPixy] = 0.0; Ui[L.iPiyy] = 0.0; Ui[L.iPieta] = 0.0; }
            };
            if (isf(rdom)) {
                double x = g.xC[ix], y = g.yC[iy];
                if (x * x + y * y > rdom2) { Ui[L.iE] = Emin; blank(); continue; }
            }
            double Ei = Ui[L.iE];
            if (!isf(Ei)) { Ui[L.iE] = Emin; blank(); continue; }
            if (Ei < Evac) { Ui[L.iE] = Emin; blank(); continue; }
            if (Ei <= 1e6 * Emin) blank();
            double Di = Ui[L.iD];
            if (!isf(Di) || Di < 0.0) Ui[L.iD] = 0.0;
            if (!isf(Ui[L.iSx])) Ui[L.iSx] = 0.0;
            if (!isf(Ui[L.iSy])) Ui[L.iSy] = 0.0;
            if (L.hasNu) {
                if (!isf(Ui[L.iNux])) Ui[L.iNux] = 0.0;
                if (!isf(Ui[L.iNuy])) Ui[L.iNuy] = 0.0;
            }
            if (L.hasPi && !isf(Ui[L.iPi])) Ui[L.iPi] = 0.0;
            if (L.hasShear) {
                if (!isf(Ui[L.iPixx])) Ui[L.iPixx] = 0.0;
                if (!isf(Ui[L.iPixy])) Ui[L.iPixy] = 0.0;
                if (!isf(Ui[L.iPiyy])) Ui[L.iPiyy] = 0.0;
                if (!isf(Ui[L.iPieta])) Ui[L.iPieta] = 0.0;
            }
        }
}

// Peff of enforce_S_energy_constraint_2d! / cell_admissible_2d (src2d/floors2d.jl:136-149, 320-330)
static inline double peff_of(const Layout& L, const double* Ui, const double* P, int i) {
    double Peff = 0.0;
    if (P) { double p = P[i]; if (isf(p)) Peff = jmax(p, 0.0); }
    if (L.hasPi) { double v = Ui[L.iPi]; if (isf(v)) Peff += std::fabs(v); }
    if (L.hasShear) {
        const int ks[4] = {L.iPixx, L.iPixy, L.iPiyy, L.iPieta};
        for (int k : ks) { double v = Ui[k]; if (isf(v)) Peff += std::fabs(v); }
    }
    return Peff;
}

// enforce_S_energy_constraint_2d! (src2d/floors2d.jl:118-162)
void enforce_S_energy_constraint(Ctx& c, std::vector<double>& U, double chi, const double* P) {
    const Grid& g = c.g;
    const Layout& L = c.m.L;
    const int nv = L.nv, ng = g.ng;
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            int i = g.lin(ix, iy);
            double* Ui = &U[nv * (size_t)i];
            double Ei = Ui[L.iE], Sx = Ui[L.iSx], Sy = Ui[L.iSy];
            if (!(isf(Ei) && isf(Sx) && isf(Sy))) continue;
            if (Ei <= 0) continue;
            double Peff = peff_of(L, Ui, P, i);
            double Smax = chi * (Ei + Peff);
            double aS = std::hypot(Sx, Sy);
            if (aS > Smax) {
                double sc = Smax / (aS + TINY);
                Ui[L.iSx] = Sx * sc; Ui[L.iSy] = Sy * sc;
            }
        }
}

// enforce_nu_charge_constraint_2d! (src2d/floors2d.jl:183-213)
void enforce_nu_charge_constraint(Ctx& c, std::vector<double>& U, double tau, double chi) {
    const Grid& g = c.g;
    const Layout& L = c.m.L;
    if (!L.hasNu) return;
    const int nv = L.nv, ng = g.ng;
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            double* Ui = &U[nv * (size_t)g.lin(ix, iy)];
            double Dt = Ui[L.iD] / tau, nux = Ui[L.iNux], nuy = Ui[L.iNuy];
            if (!(isf(Dt) && isf(nux) && isf(nuy))) continue;
            if (Dt <= 0.0) { Ui[L.iNux] = 0.0; Ui[L.iNuy] = 0.0; continue; }
            double cap = chi * Dt, an = std::hypot(nux, nuy);
            if (an > cap) {
                double sc = cap / (an + TINY);
                Ui[L.iNux] = nux * sc; Ui[L.iNuy] = nuy * sc;
            }
        }
}

// sanitize_stage_2d! (src2d/floors2d.jl:218-227)
void sanitize_stage(Ctx& c, std::vector<double>& U, double tau, int bc) {
    enforce_floors(c, U, E_FLOOR);
    enforce_S_energy_constraint(c, U, CHI_SRE, c.w.P.data());
    enforce_nu_charge_constraint(c, U, tau, 0.999);
    apply_bc(c, U, bc);
}

// cell_admissible_2d (src2d/floors2d.jl:281-337)
bool cell_admissible(Ctx& c, const std::vector<double>& U, int i, double tau, bool use_work) {
    const Layout& L = c.m.L;
    const Work& w = c.w;
    const double* Ui = &U[L.nv * (size_t)i];
    const double Emin = E_FLOOR, chi = CHI_SRE, v_eps = 1e-12;
    double E = Ui[L.iE], Sx = Ui[L.iSx], Sy = Ui[L.iSy], Dt = Ui[L.iD];
    if (!(isf(E) && isf(Sx) && isf(Sy) && isf(Dt))) return false;
    if (!(E >= Emin)) return false;
    if (!(Dt >= 0.0)) return false;
    if (L.hasNu && !(isf(Ui[L.iNux]) && isf(Ui[L.iNuy]))) return false;
    if (L.hasPi && !isf(Ui[L.iPi])) return false;
    if (L.hasShear && !(isf(Ui[L.iPixx]) && isf(Ui[L.iPixy]) && isf(Ui[L.iPiyy]) && isf(Ui[L.iPieta])))
        return false;
    if (E <= 1e6 * Emin) return true;
    if (use_work && w.ok[i]) {
        double vx = w.vxC[i], vy = w.vyC[i];
        if (!(isf(vx) && isf(vy))) return false;
        if (!((vx * vx + vy * vy) < (1 - v_eps) * (1 - v_eps))) return false;
    }
    double Peff = peff_of(L, Ui, use_work ? w.P.data() : nullptr, i);
    if (std::hypot(Sx, Sy) <= SR_MARGIN_2D * chi * (E + Peff)) return true;
    Diss d = diss_at(L, Ui);
    PrimStat st;
    Cons cs{safe_div(Dt, tau), Sx, Sy, E};
    PrimOut r = cons_to_prim(cs, d, tau, c.m.eos, NaN, NaN, NaN, NaN, st);
    return r.ok || st.reason == PRR_VACUUM;
}

// state_ok_2d (src2d/timestepper2d.jl:56-66); counts every bad cell instead of
// stopping at the first (same boolean)
bool state_ok(Ctx& c, const std::vector<double>& U, double tau, int64_t* nbad) {
    const Grid& g = c.g;
    const int ng = g.ng;
    int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cnt)
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy)
            if (!cell_admissible(c, U, g.lin(ix, iy), tau, true)) cnt += 1;
    if (nbad) *nbad = cnt;
    return cnt == 0;
}

// expand_bad_2d! (src2d/floors2d.jl:369-383): `radius` rounds of growing the mask
// by the four neighbours of every marked cell, interior cells only
void expand_bad(Ctx& c, int radius) {
    const Grid& g = c.g;
    Work& w = c.w;
    const int ng = g.ng;
    for (int r = 0; r < radius; ++r) {
        w.bad_tmp = w.bad;
        for (int ix = ng; ix < ng + g.Nx; ++ix)
            for (int iy = ng; iy < ng + g.Ny; ++iy) {
                int i = g.lin(ix, iy);
                if (w.bad_tmp[i]) continue;
                if (w.bad[i - 1] || w.bad[i + 1] || w.bad[i - g.Nytot] || w.bad[i + g.Nytot]) w.bad_tmp[i] = 1;
            }
        w.bad = w.bad_tmp;
    }
}

// mark_bad_2d! + expand_bad_2d! (src2d/floors2d.jl:350-383)
static void mark_and_expand(Ctx& c, const std::vector<double>& U, double tau, int radius) {
    const Grid& g = c.g;
    Work& w = c.w;
    const int ng = g.ng;
    std::fill(w.bad.begin(), w.bad.end(), 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            int i = g.lin(ix, iy);
            if (!cell_admissible(c, U, i, tau, true)) w.bad[i] = 1;
        }
    expand_bad(c, radius);
}

// compute_dt_2d (src2d/timestepper2d.jl:37-47)
double compute_dt(Ctx& c, double tau, double cfl, double cfltau) {
    double amax = jmax(c.w.amax, 1e-6);
    double dt_cfl = cfl * std::min(c.g.dx, c.g.dy) / amax;
    double dt_tau = cfltau * tau;
    return jmin(dt_cfl, dt_tau);
}

// _stage_2d! (src2d/timestepper2d.jl:113-140)
static bool stage(Ctx& c, std::vector<double>& Uout, const std::vector<double>& Uin,
                  const std::vector<double>& Uold, std::vector<double>& Ubase, double tau, double tau_state,
                  double D, double wold, double wnew, int bc, int max_mood) {
    Work& w = c.w;
    const uint8_t* mask = nullptr;
    const size_t N = Uout.size();
    for (int attempt = 0; attempt <= max_mood; ++attempt) {
        rhs(c, w.k, Ubase, tau, bc, mask);
        if (wnew == 0.0) {
            for (size_t j = 0; j < N; ++j) Uout[j] = Uin[j] + D * w.k[j];
        } else {
            for (size_t j = 0; j < N; ++j) Uout[j] = wold * Uold[j] + wnew * (Uin[j] + D * w.k[j]);
        }
        sanitize_stage(c, Uout, tau_state, bc);
        if (state_ok(c, Uout, tau_state, nullptr)) return true;
        if (attempt == max_mood) return false;
        c.nmood += 1;
        mark_and_expand(c, Uout, tau_state, 1 + attempt);
        mask = w.bad.data();
    }
    return false;
}

// step_ssprk2_2d! (src2d/timestepper2d.jl:74-103)
bool step_ssprk2(Ctx& c, double tau, double dtau, int bc, int max_halve, int max_mood, double* used) {
    Work& w = c.w;
    double D = dtau;
    for (int h = 0; h <= max_halve; ++h) {
        w.U2 = c.U;
        // stage 1: Uout = U1, Uin = U, Uold = U2, Ubase = U
        bool ok1 = stage(c, w.U1, c.U, w.U2, c.U, tau, tau + D, D, 1.0, 0.0, bc, max_mood);
        if (ok1) {
            bool ok2 = stage(c, c.U, w.U1, w.U2, w.U1, tau + D, tau + D, D, 0.5, 0.5, bc, max_mood);
            if (ok2) { *used = D; return true; }
        }
        c.U = w.U2;
        D *= 0.5;
        c.nhalvings += 1;
    }
    *used = D;
    return false;
}

// step_ssprk3_2d! (src2d/timestepper2d.jl:148-188)
bool step_ssprk3(Ctx& c, double tau, double dtau, int bc, int max_halve, double* used) {
    Work& w = c.w;
    double D = dtau;
    const size_t N = c.U.size();
    for (int h = 0; h <= max_halve; ++h) {
        w.U2 = c.U;
        rhs(c, w.k, c.U, tau, bc, nullptr);
        for (size_t j = 0; j < N; ++j) w.U1[j] = c.U[j] + D * w.k[j];
        sanitize_stage(c, w.U1, tau + D, bc);
        if (!state_ok(c, w.U1, tau + D, nullptr)) { c.U = w.U2; D *= 0.5; c.nhalvings += 1; continue; }
        rhs(c, w.k, w.U1, tau + D, bc, nullptr);
        for (size_t j = 0; j < N; ++j) w.U1[j] = 0.75 * w.U2[j] + 0.25 * (w.U1[j] + D * w.k[j]);
        sanitize_stage(c, w.U1, tau + D, bc);
        if (!state_ok(c, w.U1, tau + D, nullptr)) { c.U = w.U2; D *= 0.5; c.nhalvings += 1; continue; }
        rhs(c, w.k, w.U1, tau + 0.5 * D, bc, nullptr);
        for (size_t j = 0; j < N; ++j) c.U[j] = (w.U2[j] + 2.0 * (w.U1[j] + D * w.k[j])) / 3.0;
        sanitize_stage(c, c.U, tau + D, bc);
        if (state_ok(c, c.U, tau + D, nullptr)) { *used = D; return true; }
        c.U = w.U2;
        D *= 0.5;
        c.nhalvings += 1;
    }
    *used = D;
    return false;
}

// project_shear! (main2D.jl:332-358)
void project_shear(Ctx& c, double* maxr, double* meanr) {
    const Grid& g = c.g;
    const Layout& L = c.m.L;
    *maxr = 0.0; *meanr = 0.0;
    if (!L.hasShear) return;
    const bool correct = c.m.d.shear_constraint_project != 0;
    const int nv = L.nv, ng = g.ng;
    double mx = 0.0, sum = 0.0;
    int64_t cnt = 0;
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            int i = g.lin(ix, iy);
            double* Ui = &c.U[nv * (size_t)i];
            double ux = c.w.ux[i], uy = c.w.uy[i], ut = std::sqrt(1 + ux * ux + uy * uy);
            double rel = shear_rel_residual(ux, uy, ut, Ui[L.iPixx], Ui[L.iPixy], Ui[L.iPiyy], Ui[L.iPieta]);
            mx = jmax(mx, rel); sum += rel; cnt += 1;
            if (correct) Ui[L.iPieta] = shear_project_pieta(ux, uy, ut, Ui[L.iPixx], Ui[L.iPixy], Ui[L.iPiyy]);
        }
    *maxr = mx;
    *meanr = cnt > 0 ? sum / cnt : 0.0;
}

}  // namespace orc
