// =============================================================================
// oracle/oracle_primrec.hpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
// CPU restatement of the 2+1D primitive recovery (src2d/primrec2d.jl).
// =============================================================================
#pragma once
#include "oracle_physics.hpp"

namespace orc {

struct Diss { double nux, nuy, Pi, pixx, pixy, piyy, pieta; };

// solve_gauss_n! (src2d/primrec2d.jl:33-87), n <= 4
static inline bool solve_gauss_n(double* delta, const double J[4][4], const double* F, int n) {
    double A[4][5];
    for (int r = 0; r < n; ++r) {
        for (int c = 0; c < n; ++c) A[r][c] = J[r][c];
        A[r][n] = -F[r];
    }
    double smax = 0.0;
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) smax = jmax(smax, std::fabs(A[r][c]));
    if (!isf(smax) || smax == 0.0) return false;
    double pivtol = 1e-15 * smax;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double amax = std::fabs(A[k][k]);
        for (int r = k + 1; r < n; ++r)
            if (std::fabs(A[r][k]) > amax) { amax = std::fabs(A[r][k]); p = r; }
        if (amax < pivtol) return false;
        if (p != k)
            for (int c = k; c <= n; ++c) { double t = A[k][c]; A[k][c] = A[p][c]; A[p][c] = t; }
        double invp = 1.0 / A[k][k];
        for (int r = k + 1; r < n; ++r) {
            double f = A[r][k] * invp;
            if (f == 0.0) continue;
            for (int c = k; c <= n; ++c) A[r][c] -= f * A[k][c];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = A[r][n];
        for (int c = r + 1; c < n; ++c) s -= A[r][c] * delta[c];
        delta[r] = s / A[r][r];
    }
    for (int r = 0; r < n; ++r) if (!isf(delta[r])) return false;
    return true;
}

// _invert_e_for_T (src2d/primrec2d.jl:300-314)
static inline double invert_e_for_T(double e_target, double phi, const EOS& q) {
    double lo = std::log(T_MIN), hi = std::log(T_SOLVE_MAX);
    if (e_target <= 0.0) return lo;
    for (int it = 0; it < 60; ++it) {
        double mid = 0.5 * (lo + hi);
        double Tm = std::exp(mid), P, em;
        eos_Pe_2d(Tm, q.mass() + Tm * phi, q, &P, &em);
        if (!isf(em) || em > e_target) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// seed_from_conserved_2d (src2d/primrec2d.jl:323-347)
static inline void seed_from_conserved(double E, double Sx, double Sy, double phi, const EOS& q,
                                       double* yT_o, double* ux_o, double* uy_o) {
    double M = std::sqrt(Sx * Sx + Sy * Sy);
    if (!isf(E) || E <= 0.0) { *yT_o = std::log(jmax(0.15, T_MIN)); *ux_o = 0.0; *uy_o = 0.0; return; }
    double P = E / 3, yT = std::log(jmax(0.15, T_MIN)), v = 0.0;
    for (int it = 0; it < 3; ++it) {
        double denom = E + P;
        v = denom > TINY ? jmin(M / denom, 0.999999) : 0.0;
        double e = denom * (1 - v * v) - P;
        if (e <= 0.0) e = jmax(E, TINY);
        yT = invert_e_for_T(e, phi, q);
        double T = std::exp(yT), Pn, en;
        eos_Pe_2d(T, q.mass() + T * phi, q, &Pn, &en);
        if (!isf(Pn)) break;
        P = Pn;
    }
    double gam = 1 / std::sqrt(jmax(1 - v * v, 1e-16));
    double us = gam * v;
    double ux = M > TINY ? us * Sx / M : 0.0;
    double uy = M > TINY ? us * Sy / M : 0.0;
    *yT_o = yT;
    *ux_o = jclamp(ux, -U_CAP_2D, U_CAP_2D);
    *uy_o = jclamp(uy, -U_CAP_2D, U_CAP_2D);
}

struct Cons { double D, Sx, Sy, E; };

// evalF_2d! (src2d/primrec2d.jl:360-390)
static inline void evalF(double* F, const double* x, const Cons& c, const Diss& d, double tau,
                         const EOS& q, bool ncharge, EOSCache* ec) {
    double yT = x[0], phi = x[1], ux = x[2], uy = x[3];
    const double inf = std::numeric_limits<double>::infinity();
    if (!(isf(yT) && isf(phi) && isf(ux) && isf(uy))) { F[0] = F[1] = F[2] = F[3] = inf; return; }
    double T = std::exp(yT);
    double mu = q.mass() + T * phi;
    double ut = std::sqrt(1 + ux * ux + uy * uy);
    PNE r = eos_Pne_2d(T, mu, q, ec);
    if (!(isf(r.P) && isf(r.n) && isf(r.e))) { F[0] = F[1] = F[2] = F[3] = inf; return; }
    double weff = r.e + r.P + d.Pi, Ptot = r.P + d.Pi;
    Pi4 p = shear_tau_row(ux, uy, ut, d.pixx, d.pixy, d.piyy);
    F[0] = ncharge ? (r.n * ut + nu_tau(ux, uy, ut, d.nux, d.nuy) - c.D) : 0.0;
    F[1] = weff * ut * ux + p.tx - c.Sx;
    F[2] = weff * ut * uy + p.ty - c.Sy;
    F[3] = weff * ut * ut - Ptot + p.tt - c.E;
}

struct PrimOut { double T, mu, ux, uy, n, e, P; bool ok; };
struct PrimStat { uint8_t reason; int iters; double resnorm; double rows[4]; int staged = 0; };   // staged: phi by bisection

// _solve_phi_bisect (src2d/primrec2d.jl:416-434)
static inline bool solve_phi_bisect(double D, double T, double ux, double uy, double nux, double nuy,
                                    const EOS& q, EOSCache* c, double* phi_o) {
    double ut = std::sqrt(1 + ux * ux + uy * uy);
    double target = D - nu_tau(ux, uy, ut, nux, nuy);
    double m = q.mass();
    if (!isf(target) || target <= 0.0) { *phi_o = -PHI_CAP; return false; }
    double lo = -PHI_CAP, hi = PHI_CAP;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        double nm = eos_Pne_2d(T, m + T * mid, q, c).n;
        if (isf(nm) && nm * ut > target) hi = mid; else lo = mid;
    }
    double phi = 0.5 * (lo + hi);
    double nf = eos_Pne_2d(T, m + T * phi, q, c).n;
    *phi_o = phi;
    return isf(nf);
}

// _cons_to_prim_attempt_2d! (src2d/primrec2d.jl:511-742)
static inline PrimOut cons_to_prim_attempt(const Cons& c, const Diss& d, double tau, const EOS& q,
                                           double yT0, double phi0, double ux0, double uy0,
                                           PrimStat& st, int maxit = 80, double tol = 1e-12) {
    const double yT_lo = std::log(T_MIN), yT_hi = std::log(T_SOLVE_MAX);
    st.reason = PRR_UNSET; st.iters = 0; st.resnorm = std::numeric_limits<double>::quiet_NaN();
    if (!isf(c.E) || c.E <= E_VAC || !isf(c.D) || c.D < D_VAC) {
        st.reason = PRR_VACUUM; st.resnorm = 0.0;
        double P = eos_Pne(T_MIN, 0.0, q).P;
        return {T_MIN, 0.0, 0.0, 0.0, 0.0, jmax(c.E, 0.0), P, false};
    }
    bool ncharge = q.has_charge();
    EOSCache ec; ec.reset();
    int n_unk = ncharge ? 4 : 3;
    double x[4];
    x[1] = (ncharge && isf(phi0)) ? jclamp(phi0, -PHI_CAP, PHI_CAP) : 0.0;
    if (isf(yT0) && isf(ux0) && isf(uy0)) {
        x[0] = jclamp(yT0, yT_lo, yT_hi);
        x[2] = jclamp(ux0, -U_CAP_2D, U_CAP_2D);
        x[3] = jclamp(uy0, -U_CAP_2D, U_CAP_2D);
    } else {
        double sy, sx, syy;
        seed_from_conserved(c.E, c.Sx, c.Sy, x[1], q, &sy, &sx, &syy);
        x[0] = jclamp(sy, yT_lo, yT_hi); x[2] = sx; x[3] = syy;
    }
    const int slot[4] = {0, ncharge ? 1 : 2, ncharge ? 2 : 3, 3};
    double F[4], Fp[4], Fm[4], J[4][4], delta[4];
    evalF(F, x, c, d, tau, q, ncharge, &ec);
    if (!(isf(F[0]) && isf(F[1]) && isf(F[2]) && isf(F[3]))) {
        st.reason = PRR_NONFINITE_INITIAL;
        double T = jclamp(std::exp(x[0]), T_MIN, T_SOLVE_MAX), mu = q.mass();
        PNE r = eos_Pne_2d(T, mu, q);
        return {T, mu, 0.0, 0.0, 0.0, jmax(r.e, 0.0), r.P, false};
    }
    double sD = jmax(std::fabs(c.D), 1e-12);
    double sS = jmax(jmax(jmax(std::fabs(c.Sx), std::fabs(c.Sy)), 1e-10 * std::fabs(c.E)), 1e-12);
    double sE = jmax(std::fabs(c.E), 1e-12);
    const double rs[4] = {1 / sD, 1 / sS, 1 / sS, 1 / sE};
    bool converged = false;
    double last_nrm = std::numeric_limits<double>::quiet_NaN();
    int r_lo = ncharge ? 0 : 1;
    for (int it = 1; it <= maxit; ++it) {
        double nrm = 0.0;
        for (int r = r_lo; r < 4; ++r) nrm = jmax(nrm, std::fabs(F[r] * rs[r]));
        last_nrm = nrm;
        if (nrm < tol) { converged = true; st.reason = PRR_CONVERGED; st.iters = it - 1; break; }
        bool ok_jac = true;
        for (int k = 0; k < n_unk; ++k) {
            int s = slot[k];
            double h = 1e-6 * jmax(1.0, std::fabs(x[s]));
            double xs = x[s];
            x[s] = xs + h; evalF(Fp, x, c, d, tau, q, ncharge, &ec);
            x[s] = xs - h; evalF(Fm, x, c, d, tau, q, ncharge, &ec);
            x[s] = xs;
            bool fin = true;
            for (int r = 0; r < 4; ++r) fin = fin && isf(Fp[r]) && isf(Fm[r]);
            if (!fin) { ok_jac = false; break; }
            double inv2h = 1 / (2 * h);
            for (int r = 0; r < n_unk; ++r) {
                int rr = ncharge ? r : r + 1;
                J[r][k] = ((Fp[rr] - Fm[rr]) * inv2h) * rs[rr];
            }
        }
        if (!ok_jac) { st.reason = PRR_NONFINITE_JACOBIAN; break; }
        double Fact[4];
        for (int r = 0; r < n_unk; ++r) {
            int rr = ncharge ? r : r + 1;
            delta[r] = 0.0;
            Fact[r] = F[rr] * rs[rr];
        }
        if (!solve_gauss_n(delta, J, Fact, n_unk)) { st.reason = PRR_LINSOLVE_FAILED; break; }
        double alpha = 1.0;
        for (int k = 0; k < n_unk; ++k) {
            double ad = std::fabs(delta[k]);
            if (ad > 1.0) alpha = jmin(alpha, 1.0 / ad);
        }
        double xs[4] = {x[0], x[1], x[2], x[3]};
        bool accepted = false;
        for (int ls = 0; ls < 12; ++ls) {
            x[0] = xs[0]; x[1] = xs[1]; x[2] = xs[2]; x[3] = xs[3];
            for (int k = 0; k < n_unk; ++k) x[slot[k]] += alpha * delta[k];
            x[0] = jclamp(x[0], yT_lo, yT_hi);
            if (ncharge) x[1] = jclamp(x[1], -PHI_CAP, PHI_CAP);
            x[2] = jclamp(x[2], -U_CAP_2D, U_CAP_2D);
            x[3] = jclamp(x[3], -U_CAP_2D, U_CAP_2D);
            evalF(F, x, c, d, tau, q, ncharge, &ec);
            if (isf(F[0]) && isf(F[1]) && isf(F[2]) && isf(F[3])) {
                double trial = 0.0;
                for (int r = r_lo; r < 4; ++r) trial = jmax(trial, std::fabs(F[r] * rs[r]));
                if (trial < nrm * (1 - 1e-4 * alpha)) { accepted = true; break; }
            }
            alpha *= 0.5;
        }
        if (!accepted) { st.reason = PRR_LINESEARCH_FAILED; break; }
        if (it == maxit) st.iters = maxit;
    }
    st.resnorm = last_nrm;
    double T = jclamp(std::exp(x[0]), T_MIN, T_SOLVE_MAX);
    double mu = ncharge ? (q.mass() + T * x[1]) : q.mass();
    double ux = x[2], uy = x[3];
    PNE r = eos_Pne_2d(T, mu, q, &ec);
    if (!converged) st.reason = (st.reason == PRR_UNSET) ? PRR_MAXIT : st.reason;
    evalF(F, x, c, d, tau, q, ncharge, &ec);
    double resnorm = 0.0;
    for (int rr = r_lo; rr < 4; ++rr) {
        st.rows[rr] = std::fabs(F[rr] * rs[rr]);
        resnorm = jmax(resnorm, st.rows[rr]);
    }
    st.resnorm = resnorm;
    if (!(isf(resnorm) && resnorm < tol)) {
        double hyd = jmax(jmax(st.rows[1], st.rows[2]), st.rows[3]);
        if (ncharge && isf(hyd) && hyd < tol) {
            double phib;
            bool okb = solve_phi_bisect(c.D, T, ux, uy, d.nux, d.nuy, q, &ec, &phib);
            x[1] = jclamp(phib, -PHI_CAP, PHI_CAP);
            mu = q.mass() + T * x[1];
            r = eos_Pne_2d(T, mu, q, &ec);
            st.reason = okb ? PRR_CONVERGED : PRR_RESIDUAL_TOO_LARGE;
            st.staged = 1;
            return {T, mu, ux, uy, jmax(r.n, 0.0), jmax(r.e, 0.0), r.P, true};
        }
        st.reason = PRR_RESIDUAL_TOO_LARGE;
        return {T, mu, ux, uy, r.n, jmax(r.e, 0.0), r.P, false};
    }
    st.reason = PRR_CONVERGED;
    return {T, mu, ux, uy, r.n, jmax(r.e, 0.0), r.P, true};
}

// cons_to_prim_2d! (src2d/primrec2d.jl:465-486): one fresh-seed retry
static inline PrimOut cons_to_prim(const Cons& c, const Diss& d, double tau, const EOS& q,
                                   double yT0, double phi0, double ux0, double uy0, PrimStat& st,
                                   int* iters_total = nullptr, int* first_iters = nullptr,
                                   int* attempts = nullptr) {
    PrimOut r = cons_to_prim_attempt(c, d, tau, q, yT0, phi0, ux0, uy0, st);
    if (iters_total) *iters_total += st.iters;
    if (first_iters) *first_iters = st.iters;
    if (attempts) *attempts = 1;
    if (r.ok) return r;
    if (st.reason == PRR_VACUUM) return r;
    if (!(isf(yT0) && isf(ux0) && isf(uy0))) return r;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    r = cons_to_prim_attempt(c, d, tau, q, nan, nan, nan, nan, st);
    if (iters_total) *iters_total += st.iters;
    if (attempts) *attempts = 2;
    return r;
}

}  // namespace orc
