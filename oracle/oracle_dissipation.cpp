// =============================================================================
// oracle/oracle_dissipation.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
// CPU restatement of the operator-split relaxation (src2d/dissipation2d.jl),
// the transport coefficients (src/primitives.jl:30-108, src2d/transport2d.jl) and
// the charm moment closures (src2d/hq_consistent_firstmoment2d.jl,
// src2d/hq_consistent_m2_2d.jl).
// =============================================================================
#include <cstring>
#include "oracle_solver.hpp"

namespace orc {

static const double NaNd = std::numeric_limits<double>::quiet_NaN();
static inline double sq(double x) { return x * x; }
static inline bool term(const Model& m, uint32_t bit) { return (m.d.terms & bit) != 0; }

// eos_entropy (src/eos.jl:306-310)
static inline double eos_entropy(double T, double mu, double n, double e, double P) {
    T = jmax(T, T_MIN);
    double s = (e + P - mu * n) / T;
    return jmax(s, 0.0);
}
// shear_coeffs_2d (src2d/dissipation2d.jl:108-116) with QGPViscosity (src/primitives.jl:52-62)
static inline void shear_coeffs(const Model& m, double T, double mu, double n, double e, double P,
                                double* eta, double* taupi, double* deltapi) {
    double s = eos_entropy(T, mu, n, e, P);
    double et = jmax(m.d.eta_over_s, 0.0) * jmax(s, 0.0) * INVFMGEV;
    double Ts = posden(T) * posden(s);
    double tp = safe_div(et, (Ts * posden(m.d.tauShear_coeff)));
    *eta = et; *taupi = tp; *deltapi = m.d.deltaShear_factor * tp;
}
// bulk_coeffs_2d (src2d/dissipation2d.jl:118-124) with SimpleBulkViscosity (src/primitives.jl:95-108)
static inline void bulk_coeffs(const Model& m, double T, double mu, double n, double e, double P,
                               double* zeta, double* tauPi) {
    double s = eos_entropy(T, mu, n, e, P);
    double cs2 = eos_cs2(T, mu, m.eos);
    double peak = m.d.zeta_over_s / (1.0 + sq((T - 0.175) / 0.024));
    double z = peak * INVFMGEV * jmax(s, 0.0);
    double sp = posden(s);
    double D = jmax(std::fabs(1.0 / 3.0 - cs2), 1e-8);
    *zeta = z;
    *tauPi = safe_div(z, (posden(T) * sp * posden(m.d.tauPi_coeff))) * (1 / (D * D)) + 0.1;
}

struct Kin { double th, ax, ay, at, dxux, dxuy, dyux, dyuy; };

// kinematics_2d (src2d/dissipation2d.jl:141-167)
static inline Kin kinematics(const Ctx& c, int i, double D, double tau) {
    const Work& w = c.w;
    const Grid& g = c.g;
    double invdx = 1 / (2 * g.dx), invdy = 1 / (2 * g.dy);
    int ixm = i - g.Nytot, ixp = i + g.Nytot;
    Kin k;
    k.dxux = (w.ux[ixp] - w.ux[ixm]) * invdx;
    k.dxuy = (w.uy[ixp] - w.uy[ixm]) * invdx;
    k.dyux = (w.ux[i + 1] - w.ux[i - 1]) * invdy;
    k.dyuy = (w.uy[i + 1] - w.uy[i - 1]) * invdy;
    double ux = w.ux[i], uy = w.uy[i], ut = std::sqrt(1 + ux * ux + uy * uy);
    double uxp = w.ux_prev[i], uyp = w.uy_prev[i];
    bool have = isf(uxp) && isf(uyp);
    double dtux = have ? (ux - uxp) / D : 0.0;
    double dtuy = have ? (uy - uyp) / D : 0.0;
    double dtut = (ux * dtux + uy * dtuy) * safe_inv(ut);
    k.th = dtut + k.dxux + k.dyuy + ut * safe_inv(tau);
    k.ax = ut * dtux + ux * k.dxux + uy * k.dyux;
    k.ay = ut * dtuy + ux * k.dxuy + uy * k.dyuy;
    k.at = (ux * k.ax + uy * k.ay) * safe_inv(ut);
    return k;
}

// ns_shear_target_2d (src2d/dissipation2d.jl:181-193): (-2 eta sigma^xx, ^xy, ^yy, ^eta_eta)
static inline void ns_shear(double ux, double uy, double ut, double tau, const Kin& k, double eta, double* o) {
    double third = k.th / 3;
    double sxx = k.dxux + ux * k.ax - third * (1 + ux * ux);
    double syy = k.dyuy + uy * k.ay - third * (1 + uy * uy);
    double sxy = 0.5 * (k.dxuy + k.dyux) + 0.5 * (ux * k.ay + uy * k.ax) - third * (ux * uy);
    double seta = ut * safe_inv(tau) - third;
    o[0] = -2 * eta * sxx; o[1] = -2 * eta * sxy; o[2] = -2 * eta * syy; o[3] = -2 * eta * seta;
}

// pi_sigma_contractions_2d (src2d/hq_consistent_m2_2d.jl:201-210)
static inline void pi_sigma(double ux, double uy, double pxx, double pxy, double pyy, double peta, double sxx,
                            double sxy, double syy, double seta, double* ps, double* cxx, double* cxy, double* cyy) {
    double ux2 = ux * ux, uy2 = uy * uy, d = 1 + ux2 + uy2;
    *ps = (2 * pxy * sxy + pyy * syy + 2 * pxy * sxy * ux2 + 2 * pyy * syy * ux2 + pyy * syy * (ux2 * ux2) -
           2 * pxy * sxx * ux * uy - 2 * pyy * sxy * ux * uy - 2 * pxy * syy * ux * uy -
           2 * pyy * sxy * pow3(ux) * uy - 2 * pxy * syy * pow3(ux) * uy + 2 * pxy * sxy * uy2 +
           pyy * sxx * ux2 * uy2 + 4 * pxy * sxy * ux2 * uy2 - 2 * pxy * sxx * ux * pow3(uy) +
           pxx * sxx * sq(1 + uy2) + peta * seta * sq(d) +
           pxx * ux * uy * (syy * ux * uy - 2 * sxy * (1 + uy2))) / sq(d);
    *cxx = pxx * sxx + pxy * sxy - ((pxx * ux + pxy * uy) * (sxx * ux + sxy * uy)) / d;
    *cyy = pxy * sxy + pyy * syy - ((pxy * ux + pyy * uy) * (sxy * ux + syy * uy)) / d;
    *cxy = (pyy * (sxy + sxy * ux2 - sxx * ux * uy) + pxy * (sxx + syy + syy * ux2 - 2 * sxy * ux * uy + sxx * uy2) +
            pxx * (sxy - syy * ux * uy + sxy * uy2)) / (2 * d);
}

// vorticity_coupling_2d (src2d/hq_consistent_m2_2d.jl:442-468)
static inline void vorticity(double ux, double uy, double ut, double pxx, double pxy, double pyy, double ax,
                             double ay, double dtux, double dtuy, double dxux, double dxuy, double dyux,
                             double dyuy, double* w) {
    double iu = safe_inv(ut);
    double at = (ux * ax + uy * ay) * iu;
    double dxut = (ux * dxux + uy * dxuy) * iu, dyut = (ux * dyux + uy * dyuy) * iu;
    double Atx = -dtux + ut * ax, Axt = dxut + ux * at;
    double Aty = -dtuy + ut * ay, Ayt = dyut + uy * at;
    double Axy = dxuy + ux * ay, Ayx = dyux + uy * ax;
    double wtx = 0.5 * (Atx - Axt), wty = 0.5 * (Aty - Ayt), wxy = 0.5 * (Axy - Ayx);
    Pi4 P = shear_tau_row(ux, uy, ut, pxx, pxy, pyy);
    double Mxx = -P.tx * wtx - pxy * wxy;
    double Mxy = -P.tx * wty + pxx * wxy;
    double Myx = -P.ty * wtx - pyy * wxy;
    double Myy = -P.ty * wty + pxy * wxy;
    w[0] = 2 * Mxx; w[1] = Mxy + Myx; w[2] = 2 * Myy;
}

// _diff_tauN_impl_2d (src2d/transport2d.jl:28-123)
static inline double diff_tauN(double Tm, double alpha, const EOS& q, double DsT, double tauD) {
    double m = q.mass(), z = m / Tm;
    if (z <= 0.0) return 0.0;
    if (q.kind == FIVO_EOS_LATTICE_HRG) {
        double b1 = safe_besselkx(1, z), b2 = safe_besselkx(2, z);
        double b3 = b1 + 4.0 / z * b2;
        double b4 = b2 + 6.0 / z * b3;
        double b5 = b3 + 8.0 / z * b4;
        double ex = std::exp(jclamp(alpha - z, -700.0, 700.0));
        double m5 = (m * m) * (m * m) * m;
        double tauq = q.g_hq * (DsT / (96.0 * PI2 * pow3(Tm))) * m5 * ex * (2.0 * b1 - 3.0 * b3 + b5);
        double norm = q.g_hq * (Tm / (2 * PI2)) * (m * m) * ex * b2 * FMGEV3;   // src2d/transport2d.jl:19-26
        if (std::fabs(norm) <= TINY) return 0.0;
        double tn = tauq / norm * (FMGEV * FMGEV);
        if (!isf(tn)) return 0.0;
        return jmax(tn, 0.0) * tauD;
    }
    if (z > 50.0) {
        double K2a = safe_besselkx(2, z), K3a = safe_besselkx(3, z);
        double tg = std::fabs(K2a) > TINY ? DsT * z * (K3a / K2a) / Tm : DsT * z / Tm;
        return jmin((tg / FMGEV) * tauD, 1e20);
    }
    double K1x = safe_besselkx(1, z), K2x = safe_besselkx(2, z), K3x = safe_besselkx(3, z), K5x = safe_besselkx(5, z);
    double num = 2 * K1x - 3 * K3x + K5x;
    double den = jmax(std::fabs(K2x), TINY);
    double ratio = num / den * jsign(K2x == 0 ? 1.0 : K2x);
    double z3 = pow3(z) / Tm;
    if (!isf(z3) || z3 > 1e50) z3 = 1e50;
    double tg = (DsT / 48) * z3 * ratio;
    if (!isf(tg) || std::fabs(tg) > 1e50) return 1e50 * jsign(tg);
    return (tg / FMGEV) * tauD;
}

// vacuum_weight_2d (src2d/transport2d.jl:143-149)
static inline double vacuum_weight(double n, const Model& m) {
    double lo = m.d.vacuum_n_lo, hi = m.d.vacuum_n_hi;
    if (hi <= lo) return 1.0;
    if (n <= lo) return 0.0;
    if (n >= hi) return 1.0;
    return (n - lo) / (hi - lo);
}

// tauM_charm_2d / etaM_charm_2d (src2d/transport2d.jl:172-186)
static inline double tauM_charm(double T, double taun, double m) {
    double Tm = jmax(T, T_MIN), z = m / Tm;
    double K2 = safe_besselkx(2, z), K3 = safe_besselkx(3, z), K4 = safe_besselkx(4, z);
    if (std::fabs(K3) <= TINY) return 0.0;
    double v = taun * K4 * K2 / (2 * K3 * K3);
    return (isf(v) && v > 0) ? v : 0.0;
}
static inline double etaM_charm(double T, double taun) { return jmax(T, T_MIN) * taun / 2; }

// hq_h_hprime_2d, hq_dn_dT_2d (src2d/hq_consistent_firstmoment2d.jl:145-185)
static inline void hq_h_hprime(double T, const EOS& q, double* h, double* hp) {
    double Tm = jmax(T, T_MIN), m = q.mass(), z = m / Tm;
    *h = 0.0; *hp = 0.0;
    if (z <= 0.0) return;
    double K1 = safe_besselkx(1, z), K2 = safe_besselkx(2, z), K3 = safe_besselkx(3, z), K4 = safe_besselkx(4, z);
    if (std::fabs(K2) <= TINY) return;
    double hh = m * K3 / K2;
    double hpp = (m * z / Tm) * ((K2 + K4) * K2 - (K1 + K3) * K3) / (2.0 * K2 * K2);
    *h = isf(hh) ? hh : 0.0;
    *hp = isf(hpp) ? hpp : 0.0;
}
static inline double hq_dn_dT(double T, double n, const EOS& q) {
    double Tm = jmax(T, T_MIN), m = q.mass(), z = m / Tm;
    if (z <= 0.0) return 0.0;
    double K1 = safe_besselkx(1, z), K2 = safe_besselkx(2, z);
    if (std::fabs(K2) <= TINY) return 0.0;
    double v = (n / Tm) * (3.0 + z * K1 / K2);
    return isf(v) ? v : 0.0;
}

// consistent_fm_source_2d (src2d/hq_consistent_firstmoment2d.jl:218-282)
static inline void consistent_fm_source(const Model& md, double ux, double uy, double ut, double T, double dxT,
                                        double dyT, double dtT, double nux, double nuy, double n, double dn_dT,
                                        double taun, double Ds, double hp, const Kin& k, double* sx, double* sy) {
    double Tm = jmax(T, T_MIN);
    double DT = ut * dtT + ux * dxT + uy * dyT;
    double gTx = dxT + ux * DT, gTy = dyT + uy * DT;
    double nut = nu_tau(ux, uy, ut, nux, nuy);
    double dtux = (k.ax - ux * k.dxux - uy * k.dyux) * safe_inv(ut);
    double dtuy = (k.ay - ux * k.dxuy - uy * k.dyuy) * safe_inv(ut);
    double ngux = nut * dtux + nux * k.dxux + nuy * k.dyux;
    double nguy = nut * dtuy + nux * k.dxuy + nuy * k.dyuy;
    double expansion = term(md, FIVO_T_FM_EXPANSION) ? taun * k.th : 0.0;
    double dlnh = term(md, FIVO_T_FM_DLNH) ? (Ds / Tm) * hp * DT : 0.0;
    double br = expansion + dlnh;
    double pref = (Ds / Tm) * (n + Tm * dn_dT);
    *sx = (term(md, FIVO_T_FM_GRADT) ? pref * gTx : 0.0) + (term(md, FIVO_T_FM_INERTIAL) ? taun * n * k.ax : 0.0) +
          (term(md, FIVO_T_FM_NU_GRADU) ? taun * ngux : 0.0) + nux * br;
    *sy = (term(md, FIVO_T_FM_GRADT) ? pref * gTy : 0.0) + (term(md, FIVO_T_FM_INERTIAL) ? taun * n * k.ay : 0.0) +
          (term(md, FIVO_T_FM_NU_GRADU) ? taun * nguy : 0.0) + nuy * br;
}

// sigma_nu_2d (src2d/hq_consistent_m2_2d.jl:171-186): the current's shear
static inline void sigma_nu(double uxv, double uyv, double uT, double tau, double nux, double nuy, double dtnx,
                            double dxnx, double dynx, double dtny, double dxny, double dyny, double dtux, double dxux,
                            double dyux, double dtuy, double dxuy, double dyuy, double* o) {
    const double ux2 = uxv * uxv, uy2 = uyv * uyv, ux3 = pow3(uxv), uy3 = pow3(uyv);
    const double ux4 = ux2 * ux2, uy4 = uy2 * uy2, ux5 = ux4 * uxv, uy5 = uy4 * uyv;
    const double D = 1 + ux2 + uy2;
    const double D32 = D * std::sqrt(D);   // (1 + u^2)^(3/2)
    o[0] = (tau * D * (2 * dtnx * ux3 - dtny * uyv + 2 * dxnx * uT - dyny * uT - ux2 * (dtny * uyv - 2 * dxnx * uT + dyny * uT) + uxv * (2 * dtnx + 3 * dtnx * uy2 + 3 * dynx * uyv * uT)) +
            nuy * ((-1 - dtux * tau * uxv - ux2 + 2 * dtux * tau * ux3) * uy3 - 2 * tau * uxv * sq(1 + ux2) * (dtuy * uxv + dxuy * uT) - tau * (-1 + 2 * ux2) * uy2 * (dtuy + dtuy * ux2 - dyux * uxv * uT) + (1 + ux2) * uyv * (-1 + 2 * dtux * tau * ux3 + dyuy * tau * uT + ux2 * (-1 + 2 * dxux * tau * uT - 2 * dyuy * tau * uT))) +
            nux * (ux5 * (-1 + 2 * dtuy * tau * uyv) + tau * uyv * (1 + uy2) * (dtux * uyv + dyux * uT) + 2 * tau * ux4 * (-dtux - dtux * uy2 + dxuy * uyv * uT) - 2 * tau * ux2 * (dtux + 2 * dtux * uy2 + dtux * uy4 + (-dxuy + dyux) * uyv * uT + dyux * uy3 * uT) - uxv * (1 + dtuy * tau * uy3 + 2 * dxux * tau * uT + uy2 * (1 + 2 * dxux * tau * uT + dyuy * tau * uT)) + ux3 * (2 * dtuy * tau * uyv + 2 * dtuy * tau * uy3 - 2 * (1 + dxux * tau * uT) + uy2 * (-1 - 2 * dxux * tau * uT + 2 * dyuy * tau * uT)))) /
           (3 * tau * D32);
    o[1] = (tau * D * (3 * dtny * ux3 + 3 * dynx * (1 + uy2) * uT + 3 * (dtnx * uyv * (1 + uy2) + dxny * uT) + ux2 * (dtnx * uyv + 3 * dxny * uT) + uxv * (3 * dtny + dtny * uy2 + (dxnx + dyny) * uyv * uT)) -
            nuy * (4 * dtuy * tau * ux5 * uyv + 3 * dxuy * tau * uyv * uT + 4 * tau * ux4 * uyv * (-(dtux * uyv) + dxuy * uT) - tau * ux2 * uyv * (-7 * dxuy * uT + 2 * dtux * uyv * (3 + 2 * uy2) + dyux * uT * (3 + 4 * uy2)) + ux3 * (10 * dtuy * tau * uyv + 4 * dtuy * tau * uy3 + 3 * dyuy * tau * uT + uy2 * (2 - 4 * dxux * tau * uT + 4 * dyuy * tau * uT)) + uxv * (6 * dtuy * tau * uyv + 4 * dtuy * tau * uy3 + 2 * uy4 + 3 * dyuy * tau * uT + uy2 * (2 - 3 * dxux * tau * uT + 4 * dyuy * tau * uT))) +
            nux * (2 * ux4 * uyv * (-1 + 2 * dtuy * tau * uyv) - 3 * dxux * tau * uyv * (1 + uy2) * uT - 4 * tau * ux3 * uyv * (dtux + dtux * uy2 - dxuy * uyv * uT) + ux2 * uyv * (-2 + 6 * dtuy * tau * uyv + 4 * dtuy * tau * uy3 - 4 * dxux * tau * uT + 3 * dyuy * tau * uT + uy2 * (-2 - 4 * dxux * tau * uT + 4 * dyuy * tau * uT)) - tau * uxv * (dyux * uT * (3 + 7 * uy2 + 4 * uy4) + uyv * (6 * dtux + 10 * dtux * uy2 + 4 * dtux * uy4 - 3 * dxuy * uyv * uT)))) /
           (6 * tau * D32);
    o[2] = (tau * D * (-(dtnx * uxv) + 2 * dtny * uyv + 3 * dtny * ux2 * uyv - dtnx * uxv * uy2 + 2 * dtny * uy3 + 3 * dxny * uxv * uyv * uT - dxnx * (1 + uy2) * uT + 2 * dyny * (1 + uy2) * uT) +
            nux * (ux3 * (-1 - dtuy * tau * uyv - uy2 + 2 * dtuy * tau * uy3) - 2 * tau * uyv * sq(1 + uy2) * (dtux * uyv + dyux * uT) - tau * ux2 * (-1 + 2 * uy2) * (dtux + dtux * uy2 - dxuy * uyv * uT) + uxv * (1 + uy2) * (-1 + 2 * dtuy * tau * uy3 + dxux * tau * uT + uy2 * (-1 - 2 * dxux * tau * uT + 2 * dyuy * tau * uT))) +
            nuy * ((-1 + 2 * dtux * tau * uxv) * uy5 + tau * uxv * (1 + ux2) * (dtuy * uxv + dxuy * uT) + 2 * tau * uy4 * (-dtuy - dtuy * ux2 + dyux * uxv * uT) - 2 * tau * uy2 * (dtuy + 2 * dtuy * ux2 + dtuy * ux4 + (dxuy - dyux) * uxv * uT + dxuy * ux3 * uT) + uy3 * (2 * dtux * tau * uxv + 2 * dtux * tau * ux3 + ux2 * (-1 + 2 * dxux * tau * uT - 2 * dyuy * tau * uT) - 2 * (1 + dyuy * tau * uT)) - uyv * (1 + dtux * tau * ux3 + 2 * dyuy * tau * uT + ux2 * (1 + dxux * tau * uT + 2 * dyuy * tau * uT)))) /
           (3 * tau * D32);
}

// rank1_traceless_2d (src2d/hq_consistent_m2_2d.jl:225-236)
static inline void rank1_traceless(double Ax, double Ay, double Bx, double By, double ux, double uy, double ut,
                                   double* q) {
    double u2 = ut * ut;
    double tr = (Ay * (By + By * (ux * ux) - Bx * ux * uy) + Ax * (Bx - By * ux * uy + Bx * (uy * uy))) / u2;
    q[0] = Ax * Bx - (1 + ux * ux) * tr / 3;
    q[2] = Ay * By - (1 + uy * uy) * tr / 3;
    q[1] = (Ay * Bx + Ax * By) / 2 - (ux * uy) * tr / 3;
}

struct M2In {
    double ux, uy, ut, tau, T, dxT, dyT, dtT, nx, ny, peta;
    double dta, dxa, dya, thnu, anx, any, dtnx, dtny, dxnx, dxny, dynx, dyny;
    Kin k;
    double dtux, dtuy, n, taun, Ds, h, hp, tauM, etaM, m;
};

// consistent_m2_source_2d (src2d/hq_consistent_m2_2d.jl:247-427); returns the four rates
static inline void m2_source(const Model& md, const M2In& a, double pxx, double pxy, double pyy, double PiQ,
                             double* out) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    double Tm = jmax(a.T, T_MIN);
    if (!(a.tauM > 0.0)) return;
    const double ux = a.ux, uy = a.uy, ut = a.ut;
    double etaQ = a.etaM, zetaQ = (5.0 / 3.0) * etaQ;
    // hq_m2_coeffs_2d (src2d/hq_consistent_m2_2d.jl:125-146)
    double z = a.m / Tm;
    double K2 = safe_besselkx(2, z), K3 = safe_besselkx(3, z), K4 = safe_besselkx(4, z);
    double K5 = K3 + 8 * K4 / z;
    double dlnC = (std::fabs(K4) > TINY && std::fabs(K3) > TINY) ? z * ((K3 + K5) / (2 * K4) - (K2 + K4) / (2 * K3)) : 0.0;
    if (!isf(dlnC)) dlnC = 0.0;
    double etabar = a.taun * a.n * Tm / 2;
    double lam_a = a.Ds * ((a.m * a.m) + 6 * Tm * a.h) / (2 * Tm);
    double Aco = (a.m * a.m) + 5 * Tm * a.h;
    double ABr = (std::fabs(K3) > TINY) ? 5 + z * K2 / K3 : 5.0;
    if (!isf(ABr)) ABr = 5.0;
    double s4[4];
    ns_shear(ux, uy, ut, a.tau, a.k, -0.5, s4);
    const double sxx = s4[0], sxy = s4[1], syy = s4[2], seta = s4[3];
    double anudot = a.k.ax * a.nx + a.k.ay * a.ny - ((a.k.ax * ux + a.k.ay * uy) * (a.nx * ux + a.ny * uy)) / (ut * ut);
    double sn[3];
    sigma_nu(ux, uy, ut, a.tau, a.nx, a.ny, a.dtnx, a.dxnx, a.dynx, a.dtny, a.dxny, a.dyny, a.dtux, a.k.dxux,
             a.k.dyux, a.dtuy, a.k.dxuy, a.k.dyuy, sn);
    double DT = ut * a.dtT + ux * a.dxT + uy * a.dyT;
    double DlnT = DT / Tm;
    double DlnC = dlnC * DlnT;
    double Da = ut * a.dta + ux * a.dxa + uy * a.dya;
    double geo = term(md, FIVO_T_M2_EXPANSION) ? a.tauM * ((5.0 / 3.0) * a.k.th + DlnC) : 0.0;
    double ps, cxx, cxy, cyy;
    pi_sigma(ux, uy, pxx, pxy, pyy, a.peta, sxx, sxy, syy, seta, &ps, &cxx, &cxy, &cyy);
    double gTx = (a.h + Tm * a.hp) * (a.dxT + ux * DT);
    double gTy = (a.h + Tm * a.hp) * (a.dyT + uy * DT);
    double qa[3], qg[3];
    rank1_traceless(a.k.ax, a.k.ay, a.nx, a.ny, ux, uy, ut, qa);
    rank1_traceless(a.nx, a.ny, gTx, gTy, ux, uy, ut, qg);
    bool ia = term(md, FIVO_T_M2_ACCEL_NU), ig = term(md, FIVO_T_M2_NU_GRADTH);
    double q_xx = (ia ? 2 * lam_a * qa[0] : 0.0) + (ig ? (a.Ds / Tm) * qg[0] : 0.0);
    double q_yy = (ia ? 2 * lam_a * qa[2] : 0.0) + (ig ? (a.Ds / Tm) * qg[2] : 0.0);
    double q_xy = (ia ? 2 * lam_a * qa[1] : 0.0) + (ig ? (a.Ds / Tm) * qg[1] : 0.0);
    bool i1 = term(md, FIVO_T_M2_NU_GRADIENT), i2 = term(md, FIVO_T_M2_BG_GRADU), i3 = term(md, FIVO_T_M2_PI_SIGMA),
         i4 = term(md, FIVO_T_M2_PIQ_SIGMA);
    const double tM = a.tauM;
    double Sxx = pxx + (i1 ? 2 * etaQ * sn[0] : 0.0) + geo * pxx + (i3 ? 2 * tM * (cxx - (1 + ux * ux) * ps / 3) : 0.0) +
                 (i4 ? 2 * tM * PiQ * sxx : 0.0) + (i2 ? 2 * etabar * sxx : 0.0) + q_xx;
    double Syy = pyy + (i1 ? 2 * etaQ * sn[2] : 0.0) + geo * pyy + (i3 ? 2 * tM * (cyy - (1 + uy * uy) * ps / 3) : 0.0) +
                 (i4 ? 2 * tM * PiQ * syy : 0.0) + (i2 ? 2 * etabar * syy : 0.0) + q_yy;
    double Sxy = pxy + (i1 ? 2 * etaQ * sn[1] : 0.0) + geo * pxy + (i3 ? 2 * tM * (cxy - (ux * uy) * ps / 3) : 0.0) +
                 (i4 ? 2 * tM * PiQ * sxy : 0.0) + (i2 ? 2 * etabar * sxy : 0.0) + q_xy;
    if (term(md, FIVO_T_M2_VORTICITY)) {
        double w[3];
        vorticity(ux, uy, ut, pxx, pxy, pyy, a.k.ax, a.k.ay, a.dtux, a.dtuy, a.k.dxux, a.k.dxuy, a.k.dyux, a.k.dyuy, w);
        Sxx += tM * w[0]; Sxy += tM * w[1]; Syy += tM * w[2];
    }
    if (term(md, FIVO_T_M2_PROJECTOR)) {
        double at = (ux * a.k.ax + uy * a.k.ay) / ut;
        Pi4 P = shear_tau_row(ux, uy, ut, pxx, pxy, pyy);
        double cx = -P.tx * at + pxx * a.k.ax + pxy * a.k.ay;
        double cy = -P.ty * at + pxy * a.k.ax + pyy * a.k.ay;
        Sxx -= tM * (2 * ux * cx);
        Sxy -= tM * (ux * cy + uy * cx);
        Syy -= tM * (2 * uy * cy);
    }
    double SB = PiQ + (i1 ? zetaQ * a.thnu : 0.0) + geo * PiQ + (i3 ? (2.0 / 3.0) * tM * ps : 0.0) +
                etabar * ((term(md, FIVO_T_M2_BG_DLNT) ? ABr * DlnT : 0.0) + (i2 ? (5.0 / 3.0) * a.k.th : 0.0)) +
                (ia ? (a.Ds * Aco / (3 * Tm)) * anudot : 0.0) +
                (ig ? (5.0 / 6.0) * (a.Ds / Tm) * (a.nx * gTx + a.ny * gTy - ((a.nx * ux + a.ny * uy) * (gTx * ux + gTy * uy)) / (ut * ut)) : 0.0) +
                (term(md, FIVO_T_M2_BG_DALPHA) ? etabar * Da : 0.0);
    double den = tM * ut;
    out[0] = -Sxx / den; out[1] = -Sxy / den; out[2] = -Syy / den; out[3] = -SB / den;
}

// compute_alpha_grad_fv_2d! (src2d/dissipation2d.jl:225-260)
static void alpha_grad(Ctx& c) {
    const Grid& g = c.g;
    Work& w = c.w;
    const int ng = g.ng;
    const double* a = w.alpha.data();
    for (int dir = 0; dir < 2; ++dir) {
        const int st = dir == 0 ? g.Nytot : 1;
        std::vector<double>& face = w.alphaF[dir];
        std::vector<double>& grad = w.gradA[dir];
        std::vector<double>& sl = w.slope_tmp;
        const double invd = dir == 0 ? 1 / g.dx : 1 / g.dy;
        std::fill(sl.begin(), sl.end(), 0.0);
        std::fill(face.begin(), face.end(), 0.0);
        std::fill(grad.begin(), grad.end(), 0.0);
        int lo = dir == 0 ? 1 : 0, hi = dir == 0 ? g.Nxtot - 2 : g.Nxtot - 1;
        int ylo = dir == 0 ? 0 : 1, yhi = dir == 0 ? g.Nytot - 1 : g.Nytot - 2;
        for (int ix = lo; ix <= hi; ++ix)
            for (int iy = ylo; iy <= yhi; ++iy) {
                int i = g.lin(ix, iy);
                sl[i] = mc_limiter(a[i] - a[i - st], a[i + st] - a[i]);
            }
        int fhi = dir == 0 ? g.Nxtot - 2 : g.Nxtot - 1, gyhi = dir == 0 ? g.Nytot - 1 : g.Nytot - 2;
        const int kmax = dir == 0 ? g.Nxtot : g.Nytot;
        for (int ix = 0; ix <= fhi; ++ix)
            for (int iy = 0; iy <= gyhi; ++iy) {
                int i = g.lin(ix, iy), ip = i + st;
                int k = (dir == 0 ? ix : iy) + 1;
                if (k <= ng || k >= kmax - ng) face[i] = 0.5 * (a[i] + a[ip]);
                else face[i] = 0.5 * ((a[i] + 0.5 * sl[i]) + (a[ip] - 0.5 * sl[ip]));
            }
        for (int ix = ng; ix < ng + g.Nx; ++ix)
            for (int iy = ng; iy < ng + g.Ny; ++iy) {
                int i = g.lin(ix, iy);
                grad[i] = (face[i] - face[i - st]) * invd;
            }
    }
}

// relax_dissipative_2d! (src2d/dissipation2d.jl:293-363) and the per-cell updates
void relax_dissipative(Ctx& c, double tau, double D) {
    const Model& md = c.m;
    const fivo_model_desc& d = md.d;
    const Layout& L = md.L;
    if (!(d.enable_shear || d.enable_bulk || d.enable_diff)) return;
    const Grid& g = c.g;
    Work& w = c.w;
    const int nv = L.nv, ng = g.ng, Nyt = g.Nytot;
    w.diss_snap = c.U;
    const std::vector<double>& Q = w.diss_snap;
    if (d.enable_diff) alpha_grad(c);
    const double invdx = 1 / g.dx, invdy = 1 / g.dy;
    auto upw = [&](int a, int i, double ux, double uy) {
        double dqx = ux >= 0 ? (Q[a + nv * (size_t)i] - Q[a + nv * (size_t)(i - Nyt)]) * invdx
                             : (Q[a + nv * (size_t)(i + Nyt)] - Q[a + nv * (size_t)i]) * invdx;
        double dqy = uy >= 0 ? (Q[a + nv * (size_t)i] - Q[a + nv * (size_t)(i - 1)]) * invdy
                             : (Q[a + nv * (size_t)(i + 1)] - Q[a + nv * (size_t)i]) * invdy;
        return ux * dqx + uy * dqy;
    };
#pragma omp parallel for schedule(dynamic, 1)
    for (int ix = ng; ix < ng + g.Nx; ++ix)
        for (int iy = ng; iy < ng + g.Ny; ++iy) {
            const int i = g.lin(ix, iy);
            if (!w.ok[i]) continue;
            double* Ui = &c.U[nv * (size_t)i];
            const double ux = w.ux[i], uy = w.uy[i], ut = std::sqrt(1 + ux * ux + uy * uy);
            const Kin k = kinematics(c, i, D, tau);
            const double T = std::exp(w.yT[i]), mu = w.mu[i], n = w.n[i], e = w.e[i], P = w.P[i];
            // ---- bulk (src2d/dissipation2d.jl:377-416)
            if (d.enable_bulk && L.hasPi) {
                double Pin = Ui[L.iPi];
                double zeta, tPi;
                bulk_coeffs(md, T, mu, n, e, P, &zeta, &tPi);
                double PiNS = -zeta * k.th;
                double A = safe_div(tPi * ut, D);
                double adv = d.relax_advect_Pi ? -tPi * upw(L.iPi, i, ux, uy) : 0.0;
                double dPi = d.deltaPi_factor * tPi;
                if (d.lambda_Pi_pi_factor != 0.0 && d.enable_shear && L.hasShear) {
                    double s4[4], ps, a1, a2, a3;
                    ns_shear(ux, uy, ut, tau, k, -0.5, s4);
                    pi_sigma(ux, uy, Ui[L.iPixx], Ui[L.iPixy], Ui[L.iPiyy], Ui[L.iPieta], s4[0], s4[1], s4[2], s4[3],
                             &ps, &a1, &a2, &a3);
                    adv -= d.lambda_Pi_pi_factor * tPi * ps;
                }
                Pin = dPi == 0.0 ? (A * Pin + PiNS + adv) / (A + 1)
                                 : (A * Pin + PiNS + adv) / jmax(A + 1 + dPi * k.th, 0.5 * (A + 1));
                Pin = jmax(Pin, -0.99 * jmax(P, 0.0));
                if (d.Pi_clip_factor > 0) {
                    double cap = d.Pi_clip_factor * std::fabs(P);
                    Pin = jclamp(Pin, -cap, cap);
                }
                Ui[L.iPi] = Pin;
            }
            // ---- shear (src2d/dissipation2d.jl:432-534)
            if (d.enable_shear && L.hasShear) {
                double eta, tpi, dpi;
                shear_coeffs(md, T, mu, n, e, P, &eta, &tpi, &dpi);
                double pixx = Ui[L.iPixx], pixy = Ui[L.iPixy], piyy = Ui[L.iPiyy], pieta = Ui[L.iPieta];
                double ns[4];
                ns_shear(ux, uy, ut, tau, k, eta, ns);
                Pi4 Pt = shear_tau_row(ux, uy, ut, pixx, pixy, piyy);
                double cx = -Pt.tx * k.at + pixx * k.ax + pixy * k.ay;
                double cy = -Pt.ty * k.at + pixy * k.ax + piyy * k.ay;
                double A = safe_div(tpi * ut, D);
                double den = jmax(A + 1 + dpi * k.th, 0.5 * (A + 1));
                double pd = d.shear_projected_deriv ? tpi : 0.0;
                bool ra = d.relax_advect_pi != 0;
                double axx = ra ? -tpi * upw(L.iPixx, i, ux, uy) : 0.0;
                double axy = ra ? -tpi * upw(L.iPixy, i, ux, uy) : 0.0;
                double ayy = ra ? -tpi * upw(L.iPiyy, i, ux, uy) : 0.0;
                if (term(md, FIVO_T_SHEAR_VORTICITY)) {
                    double iut = safe_inv(ut);
                    double dtux = (k.ax - ux * k.dxux - uy * k.dyux) * iut;
                    double dtuy = (k.ay - ux * k.dxuy - uy * k.dyuy) * iut;
                    double wv[3];
                    vorticity(ux, uy, ut, pixx, pixy, piyy, k.ax, k.ay, dtux, dtuy, k.dxux, k.dxuy, k.dyux, k.dyuy, wv);
                    axx -= tpi * wv[0]; axy -= tpi * wv[1]; ayy -= tpi * wv[2];
                }
                if (d.taupi_pi_factor != 0.0 || d.lambda_pi_Pi_factor != 0.0) {
                    double s4[4];
                    ns_shear(ux, uy, ut, tau, k, -0.5, s4);
                    if (d.taupi_pi_factor != 0.0) {
                        double tpp = d.taupi_pi_factor * tpi, ps, cxx, cxy, cyy;
                        pi_sigma(ux, uy, pixx, pixy, piyy, pieta, s4[0], s4[1], s4[2], s4[3], &ps, &cxx, &cxy, &cyy);
                        axx -= tpp * (cxx - (1 + ux * ux) * ps / 3);
                        axy -= tpp * (cxy - (ux * uy) * ps / 3);
                        ayy -= tpp * (cyy - (1 + uy * uy) * ps / 3);
                    }
                    if (d.lambda_pi_Pi_factor != 0.0 && d.enable_bulk && L.hasPi) {
                        double lP = d.lambda_pi_Pi_factor * tpi * Ui[L.iPi];
                        axx -= lP * s4[0]; axy -= lP * s4[1]; ayy -= lP * s4[2];
                    }
                }
                pixx = (A * pixx + ns[0] + pd * (2 * ux * cx) + axx) / den;
                pixy = (A * pixy + ns[1] + pd * (ux * cy + uy * cx) + axy) / den;
                piyy = (A * piyy + ns[2] + pd * (2 * uy * cy) + ayy) / den;
                if (d.pi_clip_factor > 0) {
                    double cap = d.pi_clip_factor * std::fabs(P);
                    pixx = jclamp(pixx, -cap, cap); pixy = jclamp(pixy, -cap, cap); piyy = jclamp(piyy, -cap, cap);
                }
                Ui[L.iPixx] = pixx; Ui[L.iPixy] = pixy; Ui[L.iPiyy] = piyy;
                Ui[L.iPieta] = shear_project_pieta(ux, uy, ut, pixx, pixy, piyy);
            }
            // ---- charge (src2d/dissipation2d.jl:552-705)
            if (d.enable_diff && L.hasNu) {
                double wv = vacuum_weight(n, md);
                if (wv <= 0.0) { Ui[L.iNux] = 0.0; Ui[L.iNuy] = 0.0; continue; }
                // diff_coeffs_2d (src2d/transport2d.jl:131-140)
                double Tm = posden(T), al = mu / Tm;
                double kappa = safe_div(d.kappa_coeff, Tm) * jmax(n, 0.0) / FMGEV;
                double taun = diff_tauN(Tm, al, md.eos, d.kappa_coeff, d.tauN_coeff);
                double dN = d.deltaN_factor * taun;
                double nux = Ui[L.iNux], nuy = Ui[L.iNuy];
                double dxa = w.gradA[0][i], dya = w.gradA[1][i];
                double ap = w.alpha_prev[i];
                double dta = isf(ap) ? (w.alpha[i] - ap) / D : 0.0;
                double nsx = 0.0, nsy = 0.0;
                if (term(md, FIVO_T_NU_GRADALPHA)) {
                    double Da = ut * dta + ux * dxa + uy * dya;
                    nsx = -kappa * (dxa + ux * Da);
                    nsy = -kappa * (dya + uy * Da);
                }
                if (d.consistent_fm) {
                    double Tp = w.T_prev[i];
                    double dtT = isf(Tp) ? (T - Tp) / D : 0.0;
                    double dxT = (std::exp(w.yT[i + Nyt]) - std::exp(w.yT[i - Nyt])) / (2 * g.dx);
                    double dyT = (std::exp(w.yT[i + 1]) - std::exp(w.yT[i - 1])) / (2 * g.dy);
                    double h, hp;
                    hq_h_hprime(T, md.eos, &h, &hp);
                    double dndT = hq_dn_dT(T, n, md.eos);
                    double Ds = safe_div(d.kappa_coeff, T) / FMGEV;
                    double csx, csy;
                    consistent_fm_source(md, ux, uy, ut, T, dxT, dyT, dtT, nux, nuy, n, dndT, taun, Ds, hp, k, &csx, &csy);
                    nsx -= csx; nsy -= csy;
                }
                nsx *= wv; nsy *= wv;
                if (d.vacuum_ramp_relax) { taun = wv * taun; dN = wv * dN; }
                double nut = nu_tau(ux, uy, ut, nux, nuy);
                double nua = -nut * k.at + nux * k.ax + nuy * k.ay;
                double A = safe_div(taun * ut, D);
                double den = jmax(A + 1 + dN * k.th, 0.5 * (A + 1));
                double pdn = d.diff_projected_deriv ? taun : 0.0;
                double anx = d.relax_advect_nu ? -taun * upw(L.iNux, i, ux, uy) : 0.0;
                double any_ = d.relax_advect_nu ? -taun * upw(L.iNuy, i, ux, uy) : 0.0;
                nux = (A * nux + nsx + pdn * ux * nua + anx) / den;
                nuy = (A * nuy + nsy + pdn * uy * nua + any_) / den;
                if (d.nu_clip_factor > 0) {
                    double cap = d.nu_clip_factor * jmax(n, 0.0) * ut, mag = std::hypot(nux, nuy);
                    if (mag > cap && mag > TINY) { double sc = cap / mag; nux *= sc; nuy *= sc; }
                }
                Ui[L.iNux] = nux; Ui[L.iNuy] = nuy;
                // ---- charm second moment (src2d/dissipation2d.jl:722-888)
                if (d.consistent_m2 && L.hasM2) {
                    double pQxx = Ui[L.iPQxx], pQxy = Ui[L.iPQxy], pQyy = Ui[L.iPQyy], pQeta = Ui[L.iPQeta];
                    double PiQv = Ui[L.iPiQ];
                    double mq = d.transport_mass > 0.0 ? d.transport_mass : md.eos.mass();
                    double tMq = tauM_charm(T, taun, mq);
                    bool have_hist = isf(w.alpha_prev[i]) && isf(w.T_prev[i]);
                    if (tMq > 0.0 && have_hist) {
                        M2In a;
                        a.ux = ux; a.uy = uy; a.ut = ut; a.tau = tau; a.T = T; a.k = k;
                        a.dtnx = isf(w.nux_prev[i]) ? (nux - w.nux_prev[i]) / D : 0.0;
                        a.dtny = isf(w.nuy_prev[i]) ? (nuy - w.nuy_prev[i]) / D : 0.0;
                        const int ixm = i - Nyt, ixp = i + Nyt;
                        a.dxnx = (Q[L.iNux + nv * (size_t)ixp] - Q[L.iNux + nv * (size_t)ixm]) * 0.5 * invdx;
                        a.dxny = (Q[L.iNuy + nv * (size_t)ixp] - Q[L.iNuy + nv * (size_t)ixm]) * 0.5 * invdx;
                        a.dynx = (Q[L.iNux + nv * (size_t)(i + 1)] - Q[L.iNux + nv * (size_t)(i - 1)]) * 0.5 * invdy;
                        a.dyny = (Q[L.iNuy + nv * (size_t)(i + 1)] - Q[L.iNuy + nv * (size_t)(i - 1)]) * 0.5 * invdy;
                        a.dtux = isf(w.ux_prev[i]) ? (ux - w.ux_prev[i]) / D : 0.0;
                        a.dtuy = isf(w.uy_prev[i]) ? (uy - w.uy_prev[i]) / D : 0.0;
                        a.anx = ut * a.dtnx + ux * a.dxnx + uy * a.dynx;
                        a.any = ut * a.dtny + ux * a.dxny + uy * a.dyny;
                        double nut2 = nu_tau(ux, uy, ut, nux, nuy);
                        a.thnu = (ux * a.dtnx + uy * a.dtny + nux * a.dtux + nuy * a.dtuy) * safe_inv(ut) -
                                 nut2 * (ux * a.dtux + uy * a.dtuy) * safe_inv(ut * ut) + a.dxnx + a.dyny +
                                 nut2 * safe_inv(tau);
                        double Tp2 = w.T_prev[i];
                        a.dtT = isf(Tp2) ? (T - Tp2) / D : 0.0;
                        a.dxT = (std::exp(w.yT[ixp]) - std::exp(w.yT[ixm])) * 0.5 * invdx;
                        a.dyT = (std::exp(w.yT[i + 1]) - std::exp(w.yT[i - 1])) * 0.5 * invdy;
                        hq_h_hprime(T, md.eos, &a.h, &a.hp);
                        a.Ds = safe_div(d.kappa_coeff, T) / FMGEV;
                        a.nx = nux; a.ny = nuy; a.peta = pQeta;
                        a.dta = dta; a.dxa = dxa; a.dya = dya;
                        a.n = n; a.taun = taun; a.tauM = tMq; a.etaM = etaM_charm(T, taun); a.m = mq;
                        double s0[4], sp[4];
                        m2_source(md, a, pQxx, pQxy, pQyy, PiQv, s0);
                        double A2 = safe_div(tMq * ut, D);
                        double dp = 1e-7 * jmax(jmax(jmax(jmax(1.0, std::fabs(pQxx)), std::fabs(pQxy)), std::fabs(pQyy)),
                                                std::fabs(PiQv));
                        double cc[4];
                        m2_source(md, a, pQxx + dp, pQxy, pQyy, PiQv, sp); cc[0] = -(sp[0] - s0[0]) / dp * tMq * ut;
                        m2_source(md, a, pQxx, pQxy + dp, pQyy, PiQv, sp); cc[1] = -(sp[1] - s0[1]) / dp * tMq * ut;
                        m2_source(md, a, pQxx, pQxy, pQyy + dp, PiQv, sp); cc[2] = -(sp[2] - s0[2]) / dp * tMq * ut;
                        m2_source(md, a, pQxx, pQxy, pQyy, PiQv + dp, sp); cc[3] = -(sp[3] - s0[3]) / dp * tMq * ut;
                        auto dn = [&](double cv) { return jmax(A2 + cv, 0.5 * (A2 + 1)); };
                        bool ram = d.relax_advect_m2 != 0;
                        double amxx = ram ? -tMq * upw(L.iPQxx, i, ux, uy) : 0.0;
                        double amxy = ram ? -tMq * upw(L.iPQxy, i, ux, uy) : 0.0;
                        double amyy = ram ? -tMq * upw(L.iPQyy, i, ux, uy) : 0.0;
                        double amB = ram ? -tMq * upw(L.iPiQ, i, ux, uy) : 0.0;
                        double Rxx = -tMq * ut * s0[0] - cc[0] * pQxx - amxx;
                        double Rxy = -tMq * ut * s0[1] - cc[1] * pQxy - amxy;
                        double Ryy = -tMq * ut * s0[2] - cc[2] * pQyy - amyy;
                        double RB = -tMq * ut * s0[3] - cc[3] * PiQv - amB;
                        pQxx = (A2 * pQxx - Rxx) / dn(cc[0]);
                        pQxy = (A2 * pQxy - Rxy) / dn(cc[1]);
                        pQyy = (A2 * pQyy - Ryy) / dn(cc[2]);
                        PiQv = (A2 * PiQv - RB) / dn(cc[3]);
                        pQeta = shear_project_pieta(ux, uy, ut, pQxx, pQxy, pQyy);
                        Ui[L.iPQxx] = pQxx; Ui[L.iPQxy] = pQxy; Ui[L.iPQyy] = pQyy; Ui[L.iPQeta] = pQeta;
                        Ui[L.iPiQ] = PiQv;
                    }
                }
            }
        }
    w.ux_prev = w.ux;
    w.uy_prev = w.uy;
    if (d.consistent_m2 && L.hasNu) {
        for (int j = 0; j < g.Ntot; ++j) {
            w.nux_prev[j] = c.U[L.iNux + nv * (size_t)j];
            w.nuy_prev[j] = c.U[L.iNuy + nv * (size_t)j];
        }
    }
    (void)NaNd;
}

}  // namespace orc

// single-value probes of the charge-sector coefficients (test infrastructure)
extern "C" {
double orc_diff_tauN(const fivo_eos_desc* e, double T, double alpha, double DsT, double tauN_coeff) {
    return orc::diff_tauN(orc::jmax(T, orc::T_MIN), alpha, orc::EOS(*e), DsT, tauN_coeff);
}
void orc_hq_h_hprime(const fivo_eos_desc* e, double T, double* out2) {
    orc::hq_h_hprime(T, orc::EOS(*e), &out2[0], &out2[1]);
}

// consistent_fm_source_2d (src2d/hq_consistent_firstmoment2d.jl:218-282) on n
// states, 22 arguments per row in the reference's order -> (s^x, s^y)
int orc_probe_fm_source(uint32_t terms, int64_t n, const double* in, double* out) {
    orc::Model md{};
    md.d.terms = terms;
    for (int64_t r = 0; r < n; ++r) {
        const double* a = in + 22 * r;
        orc::Kin k;
        k.th = a[15]; k.ax = a[16]; k.ay = a[17]; k.dxux = a[18]; k.dxuy = a[19]; k.dyux = a[20]; k.dyuy = a[21];
        k.at = (a[0] * k.ax + a[1] * k.ay) * orc::safe_inv(a[2]);
        orc::consistent_fm_source(md, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11],
                                  a[12], a[14], k, &out[2 * r], &out[2 * r + 1]);
    }
    return 0;
}

// consistent_m2_source_2d (src2d/hq_consistent_m2_2d.jl:247-427) on n states, 44
// arguments per row in the reference's order -> (dpxx, dpxy, dpyy, dPiQ)
int orc_probe_m2_source(uint32_t terms, int64_t n, const double* in, double* out) {
    orc::Model md{};
    md.d.terms = terms;
    for (int64_t r = 0; r < n; ++r) {
        const double* v = in + 44 * r;
        orc::M2In a;
        a.ux = v[0]; a.uy = v[1]; a.ut = v[2]; a.tau = v[3]; a.T = v[4]; a.dxT = v[5]; a.dyT = v[6]; a.dtT = v[7];
        a.nx = v[8]; a.ny = v[9]; a.peta = v[13];
        a.dta = v[15]; a.dxa = v[16]; a.dya = v[17]; a.thnu = v[18]; a.anx = v[19]; a.any = v[20];
        a.dtnx = v[21]; a.dtny = v[22]; a.dxnx = v[23]; a.dxny = v[24]; a.dynx = v[25]; a.dyny = v[26];
        a.k.th = v[27]; a.k.ax = v[28]; a.k.ay = v[29];
        a.dtux = v[30]; a.dtuy = v[31];
        a.k.dxux = v[32]; a.k.dxuy = v[33]; a.k.dyux = v[34]; a.k.dyuy = v[35];
        a.k.at = (a.ux * a.k.ax + a.uy * a.k.ay) * orc::safe_inv(a.ut);
        a.n = v[36]; a.taun = v[37]; a.Ds = v[38]; a.h = v[39]; a.hp = v[40]; a.tauM = v[41]; a.etaM = v[42];
        a.m = v[43];
        orc::m2_source(md, a, v[10], v[11], v[12], v[14], out + 4 * r);
    }
    return 0;
}

// project_shear_traceless_2d's pi^eta_eta (src2d/shear2d.jl:131-143) on n states
// (ux, uy, ut, pixx, pixy, piyy) -> pi^eta_eta = pi^tautau - pixx - piyy
void orc_probe_pieta(int64_t n, const double* in, double* out) {
    for (int64_t r = 0; r < n; ++r) {
        const double* a = in + 6 * r;
        orc::Pi4 P = orc::shear_tau_row(a[0], a[1], a[2], a[3], a[4], a[5]);
        out[r] = P.tt - a[3] - a[5];
    }
}
}
