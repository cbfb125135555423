// =============================================================================
// fivo_relax.hip -- operator-split Israel-Stewart relaxation of the dissipative
// fields, one kernel per accepted step (src2d/dissipation2d.jl:293-888).
//
// k_relax reads the state after the primitive refresh (it plays the role of the
// reference's pre-relaxation snapshot Q, dissipation2d.jl:299-303) and writes
// the next state buffer, so no snapshot copy is made: interior cells with a
// converged recovery are relaxed, every other cell is copied through. The
// limited grad(alpha) (dissipation2d.jl:225-260) is evaluated inline from a
// 5-point stencil. The velocity and current histories are advanced in place,
// each thread touching only its own cell.
//
// FP contraction is off in this file: the charm second moment measures its
// implicit coefficients by finite differences with a 1e-7 relative step
// (dissipation2d.jl:842-877), which amplifies rounding differences ~1e7, so the
// arithmetic is kept to the same roundings as the CPU restatement.
// =============================================================================
#include "fivo_ctx.h"

#pragma clang fp contract(off)

namespace fivo {
using namespace fd;

#define P_ (p.pitch)
constexpr double FMGEV = 1.0 / 0.1973269804;
constexpr double INVFMGEV = 0.1973269804;
constexpr double PI2 = 3.141592653589793 * 3.141592653589793;

__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ bool term(const Params& p, uint32_t bit) { return (p.terms & bit) != 0; }

struct Kin { double th, ax, ay, at, dxux, dxuy, dyux, dyuy; };

// eos_entropy (src/eos.jl:306-310)
__device__ __forceinline__ double entropy(double T, double mu, double n, double e, double P) {
    T = jmax(T, T_MIN);
    return jmax((e + P - mu * n) / T, 0.0);
}

// ns_shear_target_2d (src2d/dissipation2d.jl:181-193)
__device__ __forceinline__ void ns_shear(double ux, double uy, double ut, double tau, const Kin& k, double eta,
                                         double o[4]) {
    const double third = k.th / 3;
    const double sxx = k.dxux + ux * k.ax - third * (1 + ux * ux);
    const double syy = k.dyuy + uy * k.ay - third * (1 + uy * uy);
    const double sxy = 0.5 * (k.dxuy + k.dyux) + 0.5 * (ux * k.ay + uy * k.ax) - third * (ux * uy);
    const double seta = ut * safe_inv(tau) - third;
    o[0] = -2 * eta * sxx; o[1] = -2 * eta * sxy; o[2] = -2 * eta * syy; o[3] = -2 * eta * seta;
}

// pi_sigma_contractions_2d (src2d/hq_consistent_m2_2d.jl:201-210)
__device__ __forceinline__ void pi_sigma(double ux, double uy, double pxx, double pxy, double pyy, double peta,
                                         double sxx, double sxy, double syy, double seta, double* ps, double* cxx,
                                         double* cxy, double* cyy) {
    const double ux2 = ux * ux, uy2 = uy * uy, d = 1 + ux2 + uy2;
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
__device__ __forceinline__ void vorticity(double ux, double uy, double ut, double pxx, double pxy, double pyy,
                                          double ax, double ay, double dtux, double dtuy, const Kin& k, double w[3]) {
    const double iu = safe_inv(ut);
    const double at = (ux * ax + uy * ay) * iu;
    const double dxut = (ux * k.dxux + uy * k.dxuy) * iu, dyut = (ux * k.dyux + uy * k.dyuy) * iu;
    const double wtx = 0.5 * ((-dtux + ut * ax) - (dxut + ux * at));
    const double wty = 0.5 * ((-dtuy + ut * ay) - (dyut + uy * at));
    const double wxy = 0.5 * ((k.dxuy + ux * ay) - (k.dyux + uy * ax));
    double tt, tx, ty;
    shear_tau_row(ux, uy, ut, pxx, pxy, pyy, &tt, &tx, &ty);
    const double Mxx = -tx * wtx - pxy * wxy, Mxy = -tx * wty + pxx * wxy;
    const double Myx = -ty * wtx - pyy * wxy, Myy = -ty * wty + pxy * wxy;
    w[0] = 2 * Mxx; w[1] = Mxy + Myx; w[2] = 2 * Myy;
}

// _diff_tauN_impl_2d (src2d/transport2d.jl:28-123)
template <int EK>
__device__ __forceinline__ double diff_tauN(double Tm, double alpha, const EosP& q, double g_hq, double DsT,
                                            double tauD) {
    const double m = q.m, z = m / Tm;
    if (z <= 0.0) return 0.0;
    if (EK == 1) {
        const double b1 = safe_besselkx(1, z), b2 = safe_besselkx(2, z);
        const double b3 = b1 + 4.0 / z * b2;
        const double b4 = b2 + 6.0 / z * b3;
        const double b5 = b3 + 8.0 / z * b4;
        const double ex = exp(jclamp(alpha - z, -700.0, 700.0));
        const double m5 = (m * m) * (m * m) * m;
        const double tauq = g_hq * (DsT / (96.0 * PI2 * pow3(Tm))) * m5 * ex * (2.0 * b1 - 3.0 * b3 + b5);
        const double norm = g_hq * (Tm / (2 * PI2)) * (m * m) * ex * b2 * q.fmGeV3;
        if (fabs(norm) <= TINY) return 0.0;
        const double tn = tauq / norm * (FMGEV * FMGEV);
        if (!isf(tn)) return 0.0;
        return jmax(tn, 0.0) * tauD;
    }
    if (z > 50.0) {
        const double K2a = safe_besselkx(2, z), K3a = safe_besselkx(3, z);
        const double tg = fabs(K2a) > TINY ? DsT * z * (K3a / K2a) / Tm : DsT * z / Tm;
        return jmin((tg / FMGEV) * tauD, 1e20);
    }
    const double K1x = safe_besselkx(1, z), K2x = safe_besselkx(2, z), K3x = safe_besselkx(3, z),
                 K5x = safe_besselkx(5, z);
    const double num = 2 * K1x - 3 * K3x + K5x, den = jmax(fabs(K2x), TINY);
    const double ratio = num / den * jsign(K2x == 0 ? 1.0 : K2x);
    double z3 = pow3(z) / Tm;
    if (!isf(z3) || z3 > 1e50) z3 = 1e50;
    const double tg = (DsT / 48) * z3 * ratio;
    if (!isf(tg) || fabs(tg) > 1e50) return 1e50 * jsign(tg);
    return (tg / FMGEV) * tauD;
}

// hq_h_hprime_2d (src2d/hq_consistent_firstmoment2d.jl:145-162)
__device__ __forceinline__ void hq_h_hprime(double T, double m, double* h, double* hp) {
    const double Tm = jmax(T, T_MIN), z = m / Tm;
    *h = 0.0; *hp = 0.0;
    if (z <= 0.0) return;
    const double K1 = safe_besselkx(1, z), K2 = safe_besselkx(2, z), K3 = safe_besselkx(3, z),
                 K4 = safe_besselkx(4, z);
    if (fabs(K2) <= TINY) return;
    const double hh = m * K3 / K2;
    const double hpp = (m * z / Tm) * ((K2 + K4) * K2 - (K1 + K3) * K3) / (2.0 * K2 * K2);
    *h = isf(hh) ? hh : 0.0;
    *hp = isf(hpp) ? hpp : 0.0;
}

// sigma_nu_2d (src2d/hq_consistent_m2_2d.jl:171-186), transverse block
__device__ __forceinline__ void sigma_nu(double uxv, double uyv, double uT, double tau, double nux, double nuy,
                                         double dtnx, double dxnx, double dynx, double dtny, double dxny, double dyny,
                                         double dtux, double dxux, double dyux, double dtuy, double dxuy, double dyuy,
                                         double o[3]) {
    const double ux2 = uxv * uxv, uy2 = uyv * uyv, ux3 = pow3(uxv), uy3 = pow3(uyv);
    const double ux4 = ux2 * ux2, uy4 = uy2 * uy2, ux5 = ux4 * uxv, uy5 = uy4 * uyv;
    const double D = 1 + ux2 + uy2;
    const double D32 = D * sqrt(D);
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
__device__ __forceinline__ void rank1_traceless(double Ax, double Ay, double Bx, double By, double ux, double uy,
                                                double ut, double q[3]) {
    const double tr = (Ay * (By + By * (ux * ux) - Bx * ux * uy) + Ax * (Bx - By * ux * uy + Bx * (uy * uy))) / (ut * ut);
    q[0] = Ax * Bx - (1 + ux * ux) * tr / 3;
    q[2] = Ay * By - (1 + uy * uy) * tr / 3;
    q[1] = (Ay * Bx + Ax * By) / 2 - (ux * uy) * tr / 3;
}

// consistent_fm_source_2d (src2d/hq_consistent_firstmoment2d.jl:218-282): the
// first moment's sources (s^x, s^y) in the row's source-on-the-left convention;
// the relaxation subtracts them from the Navier-Stokes target.
__device__ __forceinline__ void fm_source(const Params& p, double ux, double uy, double ut, double T, double dxT,
                                          double dyT, double dtT, double nux, double nuy, double n, double dndT,
                                          double taun, double Ds, double hp, const Kin& k, double* sx, double* sy) {
    const double Tm2 = jmax(T, T_MIN);
    const double DT = ut * dtT + ux * dxT + uy * dyT;
    const double gTx = dxT + ux * DT, gTy = dyT + uy * DT;
    const double nut = nu_tau(ux, uy, ut, nux, nuy);
    const double dtux = (k.ax - ux * k.dxux - uy * k.dyux) * safe_inv(ut);
    const double dtuy = (k.ay - ux * k.dxuy - uy * k.dyuy) * safe_inv(ut);
    const double ngux = nut * dtux + nux * k.dxux + nuy * k.dyux;
    const double nguy = nut * dtuy + nux * k.dxuy + nuy * k.dyuy;
    const double expansion = term(p, FIVO_T_FM_EXPANSION) ? taun * k.th : 0.0;
    const double dlnh = term(p, FIVO_T_FM_DLNH) ? (Ds / Tm2) * hp * DT : 0.0;
    const double br = expansion + dlnh;
    const double pref = (Ds / Tm2) * (n + Tm2 * dndT);
    *sx = (term(p, FIVO_T_FM_GRADT) ? pref * gTx : 0.0) + (term(p, FIVO_T_FM_INERTIAL) ? taun * n * k.ax : 0.0) +
          (term(p, FIVO_T_FM_NU_GRADU) ? taun * ngux : 0.0) + nux * br;
    *sy = (term(p, FIVO_T_FM_GRADT) ? pref * gTy : 0.0) + (term(p, FIVO_T_FM_INERTIAL) ? taun * n * k.ay : 0.0) +
          (term(p, FIVO_T_FM_NU_GRADU) ? taun * nguy : 0.0) + nuy * br;
}

// everything the charm second-moment source takes besides the four fields
struct M2In {
    double ux, uy, ut, tau, T, dxT, dyT, dtT, nx, ny, peta;
    double dta, dxa, dya, thnu, dtnx, dtny, dxnx, dxny, dynx, dyny;
    Kin k;
    double dtux, dtuy, n, taun, Ds, h, hp, tauM, etaM;
    // hq_m2_coeffs_2d (src2d/hq_consistent_m2_2d.jl:125-146), field independent
    double dlnC, etabar, lam_a, Aco, ABr;
};

// hq_m2_coeffs_2d (src2d/hq_consistent_m2_2d.jl:125-146): the field-independent
// coefficients, from a.taun, a.n, a.Ds, a.h and the scaled K2..K4 at z = m / Tm
__device__ __forceinline__ void m2_coeffs(M2In& a, double Tm, double z, double K2, double K3, double K4, double m) {
    const double K5 = K3 + 8 * K4 / z;
    const double dlnC = (fabs(K4) > TINY && fabs(K3) > TINY) ? z * ((K3 + K5) / (2 * K4) - (K2 + K4) / (2 * K3)) : 0.0;
    a.dlnC = isf(dlnC) ? dlnC : 0.0;
    a.etabar = a.taun * a.n * Tm / 2;
    a.lam_a = a.Ds * ((m * m) + 6 * Tm * a.h) / (2 * Tm);
    a.Aco = (m * m) + 5 * Tm * a.h;
    const double ABr = (fabs(K3) > TINY) ? 5 + z * K2 / K3 : 5.0;
    a.ABr = isf(ABr) ? ABr : 5.0;
}

// consistent_m2_source_2d (src2d/hq_consistent_m2_2d.jl:247-427): the four rates.
// Out of line: the update evaluates it five times. INST separates the copy the
// relaxation calls (0) from the probe's (1): the same source, but a second caller
// of one out-of-line copy costs the relaxation kernel 224 bytes of scratch per lane.
template <int INST>
__device__ __noinline__ void m2_source(const Params& p, const M2In& a, double pxx, double pxy, double pyy, double PiQ,
                                       double out[4]) {
    const double Tm = jmax(a.T, T_MIN);
    const double ux = a.ux, uy = a.uy, ut = a.ut;
    const double etaQ = a.etaM, zetaQ = (5.0 / 3.0) * etaQ;
    double s4[4];
    ns_shear(ux, uy, ut, a.tau, a.k, -0.5, s4);
    const double sxx = s4[0], sxy = s4[1], syy = s4[2], seta = s4[3];
    const double anudot = a.k.ax * a.nx + a.k.ay * a.ny - ((a.k.ax * ux + a.k.ay * uy) * (a.nx * ux + a.ny * uy)) / (ut * ut);
    double sn[3];
    sigma_nu(ux, uy, ut, a.tau, a.nx, a.ny, a.dtnx, a.dxnx, a.dynx, a.dtny, a.dxny, a.dyny, a.dtux, a.k.dxux, a.k.dyux,
             a.dtuy, a.k.dxuy, a.k.dyuy, sn);
    const double DT = ut * a.dtT + ux * a.dxT + uy * a.dyT;
    const double DlnT = DT / Tm;
    const double DlnC = a.dlnC * DlnT;
    const double Da = ut * a.dta + ux * a.dxa + uy * a.dya;
    const double geo = term(p, FIVO_T_M2_EXPANSION) ? a.tauM * ((5.0 / 3.0) * a.k.th + DlnC) : 0.0;
    double ps, cxx, cxy, cyy;
    pi_sigma(ux, uy, pxx, pxy, pyy, a.peta, sxx, sxy, syy, seta, &ps, &cxx, &cxy, &cyy);
    const double gTx = (a.h + Tm * a.hp) * (a.dxT + ux * DT);
    const double gTy = (a.h + Tm * a.hp) * (a.dyT + uy * DT);
    double qa[3], qg[3];
    rank1_traceless(a.k.ax, a.k.ay, a.nx, a.ny, ux, uy, ut, qa);
    rank1_traceless(a.nx, a.ny, gTx, gTy, ux, uy, ut, qg);
    const bool ia = term(p, FIVO_T_M2_ACCEL_NU), ig = term(p, FIVO_T_M2_NU_GRADTH);
    const double q_xx = (ia ? 2 * a.lam_a * qa[0] : 0.0) + (ig ? (a.Ds / Tm) * qg[0] : 0.0);
    const double q_yy = (ia ? 2 * a.lam_a * qa[2] : 0.0) + (ig ? (a.Ds / Tm) * qg[2] : 0.0);
    const double q_xy = (ia ? 2 * a.lam_a * qa[1] : 0.0) + (ig ? (a.Ds / Tm) * qg[1] : 0.0);
    const bool i1 = term(p, FIVO_T_M2_NU_GRADIENT), i2 = term(p, FIVO_T_M2_BG_GRADU),
               i3 = term(p, FIVO_T_M2_PI_SIGMA), i4 = term(p, FIVO_T_M2_PIQ_SIGMA);
    const double tM = a.tauM;
    double Sxx = pxx + (i1 ? 2 * etaQ * sn[0] : 0.0) + geo * pxx + (i3 ? 2 * tM * (cxx - (1 + ux * ux) * ps / 3) : 0.0) +
                 (i4 ? 2 * tM * PiQ * sxx : 0.0) + (i2 ? 2 * a.etabar * sxx : 0.0) + q_xx;
    double Syy = pyy + (i1 ? 2 * etaQ * sn[2] : 0.0) + geo * pyy + (i3 ? 2 * tM * (cyy - (1 + uy * uy) * ps / 3) : 0.0) +
                 (i4 ? 2 * tM * PiQ * syy : 0.0) + (i2 ? 2 * a.etabar * syy : 0.0) + q_yy;
    double Sxy = pxy + (i1 ? 2 * etaQ * sn[1] : 0.0) + geo * pxy + (i3 ? 2 * tM * (cxy - (ux * uy) * ps / 3) : 0.0) +
                 (i4 ? 2 * tM * PiQ * sxy : 0.0) + (i2 ? 2 * a.etabar * sxy : 0.0) + q_xy;
    if (term(p, FIVO_T_M2_VORTICITY)) {
        double w[3];
        vorticity(ux, uy, ut, pxx, pxy, pyy, a.k.ax, a.k.ay, a.dtux, a.dtuy, a.k, w);
        Sxx += tM * w[0]; Sxy += tM * w[1]; Syy += tM * w[2];
    }
    if (term(p, FIVO_T_M2_PROJECTOR)) {
        const double at = (ux * a.k.ax + uy * a.k.ay) / ut;
        double tt, tx, ty;
        shear_tau_row(ux, uy, ut, pxx, pxy, pyy, &tt, &tx, &ty);
        const double cx = -tx * at + pxx * a.k.ax + pxy * a.k.ay;
        const double cy = -ty * at + pxy * a.k.ax + pyy * a.k.ay;
        Sxx -= tM * (2 * ux * cx);
        Sxy -= tM * (ux * cy + uy * cx);
        Syy -= tM * (2 * uy * cy);
    }
    const double SB = PiQ + (i1 ? zetaQ * a.thnu : 0.0) + geo * PiQ + (i3 ? (2.0 / 3.0) * tM * ps : 0.0) +
                      a.etabar * ((term(p, FIVO_T_M2_BG_DLNT) ? a.ABr * DlnT : 0.0) + (i2 ? (5.0 / 3.0) * a.k.th : 0.0)) +
                      (ia ? (a.Ds * a.Aco / (3 * Tm)) * anudot : 0.0) +
                      (ig ? (5.0 / 6.0) * (a.Ds / Tm) * (a.nx * gTx + a.ny * gTy - ((a.nx * ux + a.ny * uy) * (gTx * ux + gTy * uy)) / (ut * ut)) : 0.0) +
                      (term(p, FIVO_T_M2_BG_DALPHA) ? a.etabar * Da : 0.0);
    const double den = tM * ut;
    out[0] = -Sxx / den; out[1] = -Sxy / den; out[2] = -Syy / den; out[3] = -SB / den;
}

// the limited, face-reconstructed d(alpha)/dx_dir at interior cell i (src2d/dissipation2d.jl:225-260)
__device__ __forceinline__ double alpha_grad_dir(const double* __restrict__ a, int i, int st, int k0, int kmax, int ng,
                                                 double invd) {
    // face f(j) between j and j+st; k1 = 1-based index along dir of the face's left cell
    auto face = [&](int j, int k1) {
        const double aj = a[j], ajp = a[j + st];
        if (k1 <= ng || k1 >= kmax - ng) return 0.5 * (aj + ajp);
        const double slj = mc_limiter(aj - a[j - st], ajp - aj);
        const double slp = mc_limiter(ajp - aj, a[j + 2 * st] - ajp);
        return 0.5 * ((aj + 0.5 * slj) + (ajp - 0.5 * slp));
    };
    return (face(i, k0 + 1) - face(i - st, k0)) * invd;
}

template <int EK>
__device__ __forceinline__ void relax_cell(const double* __restrict__ Q, double* __restrict__ U, const Params& p,
                                           const DevWork& w, double tau, double D, int i) {
    const int ix = i / p.Nytot, iy = i - ix * p.Nytot;
    const bool interior = ix >= p.ng && ix < p.ng + p.Nx && iy >= p.ng && iy < p.ng + p.Ny;
    const double ux = w.ux[i], uy = w.uy[i];
    const int nv = p.nv;
    if (!interior || !w.ok[i]) {
        for (int a = 0; a < nv; ++a) U[a * P_ + i] = Q[a * P_ + i];
    } else {
        const int Nyt = p.Nytot;
        const double ut = sqrt(1 + ux * ux + uy * uy);
        const double invdx = 1 / p.dx, invdy = 1 / p.dy;
        // kinematics_2d (src2d/dissipation2d.jl:141-167)
        Kin k;
        {
            const double i2dx = 1 / (2 * p.dx), i2dy = 1 / (2 * p.dy);
            k.dxux = (w.ux[i + Nyt] - w.ux[i - Nyt]) * i2dx;
            k.dxuy = (w.uy[i + Nyt] - w.uy[i - Nyt]) * i2dx;
            k.dyux = (w.ux[i + 1] - w.ux[i - 1]) * i2dy;
            k.dyuy = (w.uy[i + 1] - w.uy[i - 1]) * i2dy;
            const double uxp = w.ux_prev[i], uyp = w.uy_prev[i];
            const bool have = isf(uxp) && isf(uyp);
            const double dtux = have ? (ux - uxp) / D : 0.0, dtuy = have ? (uy - uyp) / D : 0.0;
            const double dtut = (ux * dtux + uy * dtuy) * safe_inv(ut);
            k.th = dtut + k.dxux + k.dyuy + ut * safe_inv(tau);
            k.ax = ut * dtux + ux * k.dxux + uy * k.dyux;
            k.ay = ut * dtuy + ux * k.dxuy + uy * k.dyuy;
            k.at = (ux * k.ax + uy * k.ay) * safe_inv(ut);
        }
        // upwind transverse derivative of stored row a (src2d/dissipation2d.jl:309-315)
        auto upw = [&](int a) {
            const double* q = Q + a * P_;
            const double c = q[i];
            const double dqx = ux >= 0 ? (c - q[i - Nyt]) * invdx : (q[i + Nyt] - c) * invdx;
            const double dqy = uy >= 0 ? (c - q[i - 1]) * invdy : (q[i + 1] - c) * invdy;
            return ux * dqx + uy * dqy;
        };
        const double T = exp(w.yT[i]), mu = w.mu[i], n = w.n[i], e = w.e[i], P = w.P[i];
        double Pin = p.hasPi ? Q[p.iPi * P_ + i] : 0.0;
        double pixx = 0, pixy = 0, piyy = 0, pieta = 0;
        if (p.hasShear) {
            pixx = Q[p.iPixx * P_ + i]; pixy = Q[p.iPixy * P_ + i];
            piyy = Q[p.iPiyy * P_ + i]; pieta = Q[p.iPieta * P_ + i];
        }
        double cs2 = 0.0, s = 0.0;
        if (p.enable_bulk || p.enable_shear) s = entropy(T, mu, n, e, P);
        // ---- bulk (src2d/dissipation2d.jl:377-416)
        if (p.enable_bulk && p.hasPi) {
            cs2 = eos_cs2<EK>(T, p.eos);
            const double peak = p.zeta_over_s / (1.0 + sq((T - 0.175) / 0.024));
            const double zeta = peak * INVFMGEV * jmax(s, 0.0);
            const double Dl = jmax(fabs(1.0 / 3.0 - cs2), 1e-8);
            const double tPi = safe_div(zeta, (posden(T) * posden(s) * posden(p.Czeta))) * (1 / (Dl * Dl)) + 0.1;
            const double PiNS = -zeta * k.th;
            const double A = safe_div(tPi * ut, D);
            double adv = p.relax_advect_Pi ? -tPi * upw(p.iPi) : 0.0;
            const double dPi = p.deltaPi_factor * tPi;
            if (p.lambda_Pi_pi_factor != 0.0 && p.enable_shear && p.hasShear) {
                double s4[4], ps, a1, a2, a3;
                ns_shear(ux, uy, ut, tau, k, -0.5, s4);
                pi_sigma(ux, uy, pixx, pixy, piyy, pieta, s4[0], s4[1], s4[2], s4[3], &ps, &a1, &a2, &a3);
                adv -= p.lambda_Pi_pi_factor * tPi * ps;
            }
            Pin = dPi == 0.0 ? (A * Pin + PiNS + adv) / (A + 1)
                             : (A * Pin + PiNS + adv) / jmax(A + 1 + dPi * k.th, 0.5 * (A + 1));
            Pin = jmax(Pin, -0.99 * jmax(P, 0.0));
            if (p.Pi_clip_factor > 0) {
                const double cap = p.Pi_clip_factor * fabs(P);
                Pin = jclamp(Pin, -cap, cap);
            }
        }
        if (p.hasPi) U[p.iPi * P_ + i] = Pin;
        // ---- shear (src2d/dissipation2d.jl:432-534)
        if (p.enable_shear && p.hasShear) {
            const double et = jmax(p.eta_over_s, 0.0) * jmax(s, 0.0) * INVFMGEV;
            const double tpi = safe_div(et, ((posden(T) * posden(s)) * posden(p.Cs)));
            const double dpi = p.deltaShear_factor * tpi;
            double ns[4];
            ns_shear(ux, uy, ut, tau, k, et, ns);
            double tt, tx, ty;
            shear_tau_row(ux, uy, ut, pixx, pixy, piyy, &tt, &tx, &ty);
            const double cx = -tx * k.at + pixx * k.ax + pixy * k.ay;
            const double cy = -ty * k.at + pixy * k.ax + piyy * k.ay;
            const double A = safe_div(tpi * ut, D);
            const double den = jmax(A + 1 + dpi * k.th, 0.5 * (A + 1));
            const double pd = p.shear_projected_deriv ? tpi : 0.0;
            const bool ra = p.relax_advect_pi != 0;
            double axx = ra ? -tpi * upw(p.iPixx) : 0.0;
            double axy = ra ? -tpi * upw(p.iPixy) : 0.0;
            double ayy = ra ? -tpi * upw(p.iPiyy) : 0.0;
            if (term(p, FIVO_T_SHEAR_VORTICITY)) {
                const double iut = safe_inv(ut);
                const double dtux = (k.ax - ux * k.dxux - uy * k.dyux) * iut;
                const double dtuy = (k.ay - ux * k.dxuy - uy * k.dyuy) * iut;
                double wv[3];
                vorticity(ux, uy, ut, pixx, pixy, piyy, k.ax, k.ay, dtux, dtuy, k, wv);
                axx -= tpi * wv[0]; axy -= tpi * wv[1]; ayy -= tpi * wv[2];
            }
            if (p.taupi_pi_factor != 0.0 || p.lambda_pi_Pi_factor != 0.0) {
                double s4[4];
                ns_shear(ux, uy, ut, tau, k, -0.5, s4);
                if (p.taupi_pi_factor != 0.0) {
                    const double tpp = p.taupi_pi_factor * tpi;
                    double ps, cxx, cxy, cyy;
                    pi_sigma(ux, uy, pixx, pixy, piyy, pieta, s4[0], s4[1], s4[2], s4[3], &ps, &cxx, &cxy, &cyy);
                    axx -= tpp * (cxx - (1 + ux * ux) * ps / 3);
                    axy -= tpp * (cxy - (ux * uy) * ps / 3);
                    ayy -= tpp * (cyy - (1 + uy * uy) * ps / 3);
                }
                if (p.lambda_pi_Pi_factor != 0.0 && p.enable_bulk && p.hasPi) {
                    const double lP = p.lambda_pi_Pi_factor * tpi * Pin;
                    axx -= lP * s4[0]; axy -= lP * s4[1]; ayy -= lP * s4[2];
                }
            }
            double nxx = (A * pixx + ns[0] + pd * (2 * ux * cx) + axx) / den;
            double nxy = (A * pixy + ns[1] + pd * (ux * cy + uy * cx) + axy) / den;
            double nyy = (A * piyy + ns[2] + pd * (2 * uy * cy) + ayy) / den;
            if (p.pi_clip_factor > 0) {
                const double cap = p.pi_clip_factor * fabs(P);
                nxx = jclamp(nxx, -cap, cap); nxy = jclamp(nxy, -cap, cap); nyy = jclamp(nyy, -cap, cap);
            }
            double t2, x2, y2;
            shear_tau_row(ux, uy, ut, nxx, nxy, nyy, &t2, &x2, &y2);
            pixx = nxx; pixy = nxy; piyy = nyy; pieta = t2 - nxx - nyy;
        }
        if (p.hasShear) {
            U[p.iPixx * P_ + i] = pixx; U[p.iPixy * P_ + i] = pixy;
            U[p.iPiyy * P_ + i] = piyy; U[p.iPieta * P_ + i] = pieta;
        }
        // ---- charge (src2d/dissipation2d.jl:552-705), then the charm second moment
        double nux = p.hasNu ? Q[p.iNux * P_ + i] : 0.0, nuy = p.hasNu ? Q[p.iNuy * P_ + i] : 0.0;
        double pq[5] = {0, 0, 0, 0, 0};
        if (p.hasM2)
            for (int k5 = 0; k5 < 5; ++k5) pq[k5] = Q[(p.iPQ0 + k5) * P_ + i];
        if (p.enable_diff && p.hasNu) {
            double wv;
            {
                const double lo = p.vacuum_n_lo, hi = p.vacuum_n_hi;
                wv = hi <= lo ? 1.0 : (n <= lo ? 0.0 : (n >= hi ? 1.0 : (n - lo) / (hi - lo)));
            }
            if (wv <= 0.0) {
                nux = 0.0; nuy = 0.0;
            } else {
                const double Tm = posden(T), al = mu / Tm;
                const double kappa = safe_div(p.kappa_coeff, Tm) * jmax(n, 0.0) / FMGEV;
                double taun = diff_tauN<EK>(Tm, al, p.eos, p.eos.ghq, p.kappa_coeff, p.tauN_coeff);
                double dN = p.deltaN_factor * taun;
                const double dxa = alpha_grad_dir(w.alpha, i, Nyt, ix, p.Nxtot, p.ng, invdx);
                const double dya = alpha_grad_dir(w.alpha, i, 1, iy, p.Nytot, p.ng, invdy);
                const double ap = w.alpha_prev[i];
                const double dta = isf(ap) ? (w.alpha[i] - ap) / D : 0.0;
                double nsx = 0.0, nsy = 0.0;
                if (term(p, FIVO_T_NU_GRADALPHA)) {
                    const double Da = ut * dta + ux * dxa + uy * dya;
                    nsx = -kappa * (dxa + ux * Da);
                    nsy = -kappa * (dya + uy * Da);
                }
                double h = 0.0, hp = 0.0;
                const bool need_h = p.consistent_fm || (p.consistent_m2 && p.hasM2);
                if (need_h) hq_h_hprime(T, p.eos.m, &h, &hp);
                if (p.consistent_fm) {
                    // consistent_fm_source_2d (src2d/hq_consistent_firstmoment2d.jl:218-282)
                    const double Tp = w.T_prev[i];
                    const double dtT = isf(Tp) ? (T - Tp) / D : 0.0;
                    const double dxT = (exp(w.yT[i + Nyt]) - exp(w.yT[i - Nyt])) / (2 * p.dx);
                    const double dyT = (exp(w.yT[i + 1]) - exp(w.yT[i - 1])) / (2 * p.dy);
                    double dndT = 0.0;
                    {
                        const double z = p.eos.m / jmax(T, T_MIN);
                        if (z > 0.0) {
                            const double K1 = safe_besselkx(1, z), K2 = safe_besselkx(2, z);
                            if (fabs(K2) > TINY) {
                                const double v = (n / jmax(T, T_MIN)) * (3.0 + z * K1 / K2);
                                dndT = isf(v) ? v : 0.0;
                            }
                        }
                    }
                    const double Ds = safe_div(p.kappa_coeff, T) / FMGEV;
                    double csx, csy;
                    fm_source(p, ux, uy, ut, T, dxT, dyT, dtT, nux, nuy, n, dndT, taun, Ds, hp, k, &csx, &csy);
                    nsx -= csx; nsy -= csy;
                }
                nsx *= wv; nsy *= wv;
                if (p.vacuum_ramp_relax) { taun = wv * taun; dN = wv * dN; }
                const double nut = nu_tau(ux, uy, ut, nux, nuy);
                const double nua = -nut * k.at + nux * k.ax + nuy * k.ay;
                const double A = safe_div(taun * ut, D);
                const double den = jmax(A + 1 + dN * k.th, 0.5 * (A + 1));
                const double pdn = p.diff_projected_deriv ? taun : 0.0;
                const double anx = p.relax_advect_nu ? -taun * upw(p.iNux) : 0.0;
                const double any_ = p.relax_advect_nu ? -taun * upw(p.iNuy) : 0.0;
                double nx = (A * nux + nsx + pdn * ux * nua + anx) / den;
                double ny = (A * nuy + nsy + pdn * uy * nua + any_) / den;
                if (p.nu_clip_factor > 0) {
                    const double cap = p.nu_clip_factor * jmax(n, 0.0) * ut, mag = hypot(nx, ny);
                    if (mag > cap && mag > TINY) { const double sc = cap / mag; nx *= sc; ny *= sc; }
                }
                nux = nx; nuy = ny;
                // ---- charm second moment (src2d/dissipation2d.jl:722-888)
                if (p.consistent_m2 && p.hasM2) {
                    const double mq = p.transport_mass > 0.0 ? p.transport_mass : p.eos.m;
                    const double Tm3 = jmax(T, T_MIN), z = mq / Tm3;
                    const double K2 = safe_besselkx(2, z), K3 = safe_besselkx(3, z), K4 = safe_besselkx(4, z);
                    double tMq = 0.0;
                    if (fabs(K3) > TINY) {   // tauM_charm_2d (src2d/transport2d.jl:172-180)
                        const double v = taun * K4 * K2 / (2 * K3 * K3);
                        tMq = (isf(v) && v > 0) ? v : 0.0;
                    }
                    const bool have_hist = isf(w.alpha_prev[i]) && isf(w.T_prev[i]);
                    if (tMq > 0.0 && have_hist) {
                        M2In a;
                        a.ux = ux; a.uy = uy; a.ut = ut; a.tau = tau; a.T = T; a.k = k;
                        a.dtnx = isf(w.nux_prev[i]) ? (nux - w.nux_prev[i]) / D : 0.0;
                        a.dtny = isf(w.nuy_prev[i]) ? (nuy - w.nuy_prev[i]) / D : 0.0;
                        const double* Qx = Q + p.iNux * P_;
                        const double* Qy = Q + p.iNuy * P_;
                        a.dxnx = (Qx[i + Nyt] - Qx[i - Nyt]) * 0.5 * invdx;
                        a.dxny = (Qy[i + Nyt] - Qy[i - Nyt]) * 0.5 * invdx;
                        a.dynx = (Qx[i + 1] - Qx[i - 1]) * 0.5 * invdy;
                        a.dyny = (Qy[i + 1] - Qy[i - 1]) * 0.5 * invdy;
                        a.dtux = isf(w.ux_prev[i]) ? (ux - w.ux_prev[i]) / D : 0.0;
                        a.dtuy = isf(w.uy_prev[i]) ? (uy - w.uy_prev[i]) / D : 0.0;
                        const double nut2 = nu_tau(ux, uy, ut, nux, nuy);
                        a.thnu = (ux * a.dtnx + uy * a.dtny + nux * a.dtux + nuy * a.dtuy) * safe_inv(ut) -
                                 nut2 * (ux * a.dtux + uy * a.dtuy) * safe_inv(ut * ut) + a.dxnx + a.dyny +
                                 nut2 * safe_inv(tau);
                        const double Tp2 = w.T_prev[i];
                        a.dtT = isf(Tp2) ? (T - Tp2) / D : 0.0;
                        a.dxT = (exp(w.yT[i + Nyt]) - exp(w.yT[i - Nyt])) * 0.5 * invdx;
                        a.dyT = (exp(w.yT[i + 1]) - exp(w.yT[i - 1])) * 0.5 * invdy;
                        a.h = h; a.hp = hp;
                        a.Ds = safe_div(p.kappa_coeff, T) / FMGEV;
                        a.nx = nux; a.ny = nuy; a.peta = pq[3];
                        a.dta = dta; a.dxa = dxa; a.dya = dya;
                        a.n = n; a.taun = taun; a.tauM = tMq; a.etaM = jmax(T, T_MIN) * taun / 2;
                        m2_coeffs(a, Tm3, z, K2, K3, K4, mq);
                        double s0[4], sp[4], cc[4];
                        m2_source<0>(p, a, pq[0], pq[1], pq[2], pq[4], s0);
                        const double A2 = safe_div(tMq * ut, D);
                        const double dp = 1e-7 * jmax(jmax(jmax(jmax(1.0, fabs(pq[0])), fabs(pq[1])), fabs(pq[2])), fabs(pq[4]));
                        m2_source<0>(p, a, pq[0] + dp, pq[1], pq[2], pq[4], sp); cc[0] = -(sp[0] - s0[0]) / dp * tMq * ut;
                        m2_source<0>(p, a, pq[0], pq[1] + dp, pq[2], pq[4], sp); cc[1] = -(sp[1] - s0[1]) / dp * tMq * ut;
                        m2_source<0>(p, a, pq[0], pq[1], pq[2] + dp, pq[4], sp); cc[2] = -(sp[2] - s0[2]) / dp * tMq * ut;
                        m2_source<0>(p, a, pq[0], pq[1], pq[2], pq[4] + dp, sp); cc[3] = -(sp[3] - s0[3]) / dp * tMq * ut;
                        const bool ram = p.relax_advect_m2 != 0;
                        const int row[4] = {p.iPQ0, p.iPQ0 + 1, p.iPQ0 + 2, p.iPQ0 + 4};
                        const int fld[4] = {0, 1, 2, 4};
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            const double am = ram ? -tMq * upw(row[c]) : 0.0;
                            const double R = -tMq * ut * s0[c] - cc[c] * pq[fld[c]] - am;
                            pq[fld[c]] = (A2 * pq[fld[c]] - R) / jmax(A2 + cc[c], 0.5 * (A2 + 1));
                        }
                        double t2, x2, y2;
                        shear_tau_row(ux, uy, ut, pq[0], pq[1], pq[2], &t2, &x2, &y2);
                        pq[3] = t2 - pq[0] - pq[2];
                    }
                }
            }
        }
        if (p.hasNu) { U[p.iNux * P_ + i] = nux; U[p.iNuy * P_ + i] = nuy; }
        if (p.hasM2)
            for (int k5 = 0; k5 < 5; ++k5) U[(p.iPQ0 + k5) * P_ + i] = pq[k5];
        for (int a = 0; a < 4; ++a) U[a * P_ + i] = Q[a * P_ + i];
    }
    // histories for the next step (src2d/dissipation2d.jl:346-359)
    w.ux_prev[i] = ux;
    w.uy_prev[i] = uy;
    if (p.consistent_m2 && p.hasNu) {
        w.nux_prev[i] = U[p.iNux * P_ + i];
        w.nuy_prev[i] = U[p.iNuy * P_ + i];
    }
}

// Two wavefronts per SIMD: the cell body needs ~290 registers, so at one wave per
// SIMD nothing hides its latency; capping at 256 spills a few values to scratch
// (measured 4-5 % faster per step on ensembles, ~1 % on one 512^2 event).
template <int EK>
__global__ __launch_bounds__(256, 2) void k_relax(const double* __restrict__ Q, double* __restrict__ U, Params p,
                                               DevWork w, double tau, double D) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p.Ntot) relax_cell<EK>(Q, U, p, w, tau, D, i);
}

// the ensemble form: every event whose step was accepted in this replay
typedef const __attribute__((address_space(4))) EvView* CEvR;
typedef const __attribute__((address_space(4))) EvCtl* CCtlR;
template <int EK>
__global__ __launch_bounds__(256, 2) void kb_relax(BatchRun R) {
    const int e = __builtin_amdgcn_readfirstlane(((const __attribute__((address_space(4))) int*)R.map[KB_CELLS])[blockIdx.x]);
    CCtlR c = (CCtlR)(R.ctl + e);
    if (!c->accepted) return;
    CEvR E = (CEvR)(R.ev + e);
    Params p = R.p0;
    p.Nx = E->p.Nx; p.Ny = E->p.Ny; p.Nxtot = E->p.Nxtot; p.Nytot = E->p.Nytot; p.Ntot = E->p.Ntot;
    p.pitch = E->p.pitch; p.dx = E->p.dx; p.dy = E->p.dy; p.xmin = E->p.xmin; p.ymin = E->p.ymin;
    DevWork w;
    w.yT = E->w.yT; w.mu = E->w.mu; w.alpha = E->w.alpha; w.ux = E->w.ux; w.uy = E->w.uy;
    w.P = E->w.P; w.n = E->w.n; w.e = E->w.e; w.ok = E->w.ok;
    w.ux_prev = E->w.ux_prev; w.uy_prev = E->w.uy_prev; w.alpha_prev = E->w.alpha_prev; w.T_prev = E->w.T_prev;
    w.nux_prev = E->w.nux_prev; w.nuy_prev = E->w.nuy_prev; w.gradAx = E->w.gradAx; w.gradAy = E->w.gradAy;
    w.phi = E->w.phi; w.vx = E->w.vx; w.vy = E->w.vy; w.bad = E->w.bad; w.force = E->w.force; w.ftmp = E->w.ftmp;
    w.x0_yT = E->w.x0_yT; w.x0_phi = E->w.x0_phi; w.x0_ux = E->w.x0_ux; w.x0_uy = E->w.x0_uy;
    w.shear_res = E->w.shear_res;
    const int i = (blockIdx.x - E->b0[KB_CELLS]) * 256 + threadIdx.x;
    const int cur = c->cur;
    if (i < p.Ntot) relax_cell<EK>(E->U[cur & 3], E->U[(cur + 1) & 3], p, w, c->tau, c->D, i);
}

hipError_t launch_batch_relax(Batch& b) {
    if (b.p0.eos.kind == 1)
        hipLaunchKernelGGL((kb_relax<1>), dim3(b.nblocks[KB_CELLS]), dim3(256), 0, b.stream, b.run);
    else if (b.p0.eos.has_charge)
        hipLaunchKernelGGL((kb_relax<2>), dim3(b.nblocks[KB_CELLS]), dim3(256), 0, b.stream, b.run);
    else
        hipLaunchKernelGGL((kb_relax<0>), dim3(b.nblocks[KB_CELLS]), dim3(256), 0, b.stream, b.run);
    return hipGetLastError();
}

// ---- probes: the relaxation's source terms at arbitrary states ----------------------
// One thread per input row, arguments in the reference functions' own order, so a
// test can hand the device code the states the reference's gates use.
constexpr int PROBE_FM_IN = 22, PROBE_M2_IN = 44;

// consistent_fm_source_2d(ux, uy, ut, T, dxT, dyT, dtT, nux, nuy, n, dn_dT, taun, Ds,
// h, hp, theta, ax, ay, dxux, dxuy, dyux, dyuy) -> (s^x, s^y)
__global__ void k_probe_fm(Params p, int64_t n, const double* __restrict__ in, double* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double* a = in + r * PROBE_FM_IN;
    Kin k;
    k.th = a[15]; k.ax = a[16]; k.ay = a[17]; k.dxux = a[18]; k.dxuy = a[19]; k.dyux = a[20]; k.dyuy = a[21];
    k.at = (a[0] * k.ax + a[1] * k.ay) * safe_inv(a[2]);
    fm_source(p, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11], a[12], a[14], k,
              out + 2 * r, out + 2 * r + 1);
}

// consistent_m2_source_2d(ux, uy, ut, tau, T, dxT, dyT, dtT, nux, nuy, pxx, pxy, pyy,
// peta, PiQ, dta, dxa, dya, theta_nu, anux, anuy, dtnux, dtnuy, dxnux, dxnuy, dynux,
// dynuy, theta, ax, ay, dtux, dtuy, dxux, dxuy, dyux, dyuy, n, taun, Ds, h, hp, tauM,
// etaM, m) -> (dpxx, dpxy, dpyy, dPiQ); the coefficients come from m2_coeffs as in
// the relaxation
__global__ void k_probe_m2(Params p, int64_t n, const double* __restrict__ in, double* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double* v = in + r * PROBE_M2_IN;
    M2In a;
    a.ux = v[0]; a.uy = v[1]; a.ut = v[2]; a.tau = v[3]; a.T = v[4]; a.dxT = v[5]; a.dyT = v[6]; a.dtT = v[7];
    a.nx = v[8]; a.ny = v[9]; a.peta = v[13];
    a.dta = v[15]; a.dxa = v[16]; a.dya = v[17]; a.thnu = v[18];
    a.dtnx = v[21]; a.dtny = v[22]; a.dxnx = v[23]; a.dxny = v[24]; a.dynx = v[25]; a.dyny = v[26];
    a.k.th = v[27]; a.k.ax = v[28]; a.k.ay = v[29];
    a.dtux = v[30]; a.dtuy = v[31];
    a.k.dxux = v[32]; a.k.dxuy = v[33]; a.k.dyux = v[34]; a.k.dyuy = v[35];
    a.k.at = (a.ux * a.k.ax + a.uy * a.k.ay) * safe_inv(a.ut);
    a.n = v[36]; a.taun = v[37]; a.Ds = v[38]; a.h = v[39]; a.hp = v[40]; a.tauM = v[41]; a.etaM = v[42];
    const double m = v[43];
    double* o = out + 4 * r;
    o[0] = o[1] = o[2] = o[3] = 0.0;
    if (!(a.tauM > 0.0)) return;
    const double Tm = jmax(a.T, T_MIN), z = m / Tm;
    m2_coeffs(a, Tm, z, safe_besselkx(2, z), safe_besselkx(3, z), safe_besselkx(4, z), m);
    m2_source<1>(p, a, v[10], v[11], v[12], v[14], o);
}

hipError_t launch_probe(int which, uint32_t terms, int64_t n, const double* in, double* out, hipStream_t s) {
    Params p{};
    p.terms = terms;
    const int nb = (int)((n + 63) / 64);
    if (which == 0)
        hipLaunchKernelGGL(k_probe_fm, dim3(nb), dim3(64), 0, s, p, n, in, out);
    else
        hipLaunchKernelGGL(k_probe_m2, dim3(nb), dim3(64), 0, s, p, n, in, out);
    return hipGetLastError();
}

hipError_t launch_relax(Ctx& c, double* U, double tau, double dtau) {
    // read the current state, write the next buffer, make it current
    const int nxt = (c.cur + 1) % 4;
    const int nb = (c.p.Ntot + 255) / 256;
    if (c.p.eos.kind == 1)
        hipLaunchKernelGGL((k_relax<1>), dim3(nb), dim3(256), 0, c.stream, U, c.U[nxt], c.p, c.w, tau, dtau);
    else if (c.p.eos.has_charge)
        hipLaunchKernelGGL((k_relax<2>), dim3(nb), dim3(256), 0, c.stream, U, c.U[nxt], c.p, c.w, tau, dtau);
    else
        hipLaunchKernelGGL((k_relax<0>), dim3(nb), dim3(256), 0, c.stream, U, c.U[nxt], c.p, c.w, tau, dtau);
    c.cur = nxt;
    return hipGetLastError();
}

}  // namespace fivo
