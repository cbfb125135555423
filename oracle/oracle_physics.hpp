// =============================================================================
// oracle/oracle_physics.hpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// CPU FP64 restatement of FiVoHydro.jl's 2+1D per-cell physics, written to be
// read side by side with the reference. It is the parity checker for the HIP
// library and the CPU baseline of the benchmark; nothing in the shipped library
// includes, links or calls it. Every function cites the reference file:line it
// restates. Julia semantics are reproduced where they differ from C++:
//   * `min`/`max` propagate NaN (Base.min/Base.max)            -> jmin/jmax
//   * `clamp(x,a,b) = x < a ? a : (x > b ? b : x)` passes NaN   (src/utils.jl:6)
//   * `sign(±0.0) = ±0.0`, so the MC limiter returns 0 on flat data (src/utils.jl:14-22)
//   * `muladd`/`evalpoly` fuse into one rounding                 -> std::fma
// Compiled with -ffp-contract=off so no other operation is fused.
// =============================================================================
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>
#include "../include/fivo2d.h"

namespace orc {

// ---- src/constants.jl:10-42 --------------------------------------------------
static const double HBARC    = 0.1973269804;
static const double FMGEV    = 1.0 / 0.1973269804;
static const double INVFMGEV = 0.1973269804;
static const double FMGEV3   = (1.0 / 0.1973269804) * (1.0 / 0.1973269804) * (1.0 / 0.1973269804);
static const double T_MIN       = 1e-20;
static const double T_SOLVE_MAX = 10.0;
static const double PHI_CAP     = 10000.0;
static const double E_VAC       = 1e-20;
static const double D_VAC       = 0.0;
static const double E_FLOOR     = 1e-20;
static const double CHI_SRE     = 0.99999999;
static const double TINY        = 1e-300;
static const double EPS_POS     = 1e-50;
static const double U_CAP_2D    = 1.0e4;             // src2d/primrec2d.jl:265
static const double SR_MARGIN_2D = 0.98;             // src2d/floors2d.jl:268
static const double BESSELKX_ASYM_X = 1.0e4;         // src/eos.jl:25
static const double PI_ = 3.141592653589793;
static const double PI2 = 3.141592653589793 * 3.141592653589793;   // src/eos.jl:7

// PrimRecReason (src/constants.jl:49-60)
enum PRR : uint8_t {
    PRR_UNSET = 0, PRR_VACUUM = 1, PRR_CONVERGED = 2, PRR_NONFINITE_INITIAL = 3,
    PRR_NONFINITE_JACOBIAN = 4, PRR_LINSOLVE_FAILED = 5, PRR_LINESEARCH_FAILED = 6,
    PRR_EOS_INVALID = 7, PRR_RESIDUAL_TOO_LARGE = 8, PRR_MAXIT = 9
};

// ---- Julia Base semantics ----------------------------------------------------
static inline bool isf(double x) { return std::isfinite(x); }
static inline double jmax(double x, double y) {
    bool pick_y = (y > x) || (!std::signbit(y) && std::signbit(x));
    return pick_y ? (std::isnan(x) ? x : y) : (std::isnan(y) ? y : x);
}
static inline double jmin(double x, double y) {
    bool pick_y = (y < x) || (std::signbit(y) && !std::signbit(x));
    return pick_y ? (std::isnan(x) ? x : y) : (std::isnan(y) ? y : x);
}
static inline double jmax3(double a, double b, double c) { return jmax(jmax(a, b), c); }
static inline double jmin3(double a, double b, double c) { return jmin(jmin(a, b), c); }
static inline double jclamp(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }
static inline double jsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }
static inline double posden(double x) { return x > EPS_POS ? x : EPS_POS; }  // src/constants.jl:39
static inline double safe_inv(double x) { return 1.0 / posden(x); }
static inline double safe_div(double a, double b) { return a / posden(b); }
static inline double safe_sqrt(double x) { return x <= 0 ? 0.0 : std::sqrt(x); }  // src/utils.jl:5
static inline double pow4(double x) { double x2 = x * x; return x2 * x2; }
static inline double pow3(double x) { return x * x * x; }

// mc_limiter (src/utils.jl:14-22)
static inline double mc_limiter(double a, double b) {
    if (jsign(a) != jsign(b)) return 0.0;
    double s = jsign(a), aa = std::fabs(a), bb = std::fabs(b);
    return s * jmin(2 * aa, jmin(2 * bb, 0.5 * (aa + bb)));
}

// ---- Bessel K ----------------------------------------------------------------
// e^x K0(x), e^x K1(x): the rational minimax approximations published with Boost
// (J. Maddock, coefficients by P. Holoborodko), the same table the reference
// vendors (src2d/bessel2d.jl:37-103). Horner evaluation with fused multiply-add.
namespace bk {
static const double K0_P1[] = {-1.372509002685546267e-1, 2.574916117833312855e-1,
    1.395474602146869316e-2, 5.445476986653926759e-4, 7.125159422136622118e-6};
static const double K0_Q1[] = {1.000000000000000000e+00, -5.458333438017788530e-02,
    1.291052816975251298e-03, -1.367653946978586591e-05};
static const double K0_P2[] = {1.159315156584124484e-01, 2.789828789146031732e-01,
    2.524892993216121934e-02, 8.460350907213637784e-04, 1.491471924309617534e-05,
    1.627106892422088488e-07, 1.208266102392756055e-09, 6.611686391749704310e-12};
static const double K0_P3[] = {2.533141373155002416e-1, 3.628342133984595192e0,
    1.868441889406606057e1, 4.306243981063412784e1, 4.424116209627428189e1,
    1.562095339356220468e1, -1.810138978229410898e0, -1.414237994269995877e0,
    -9.369168119754924625e-2};
static const double K0_Q3[] = {1.000000000000000000e0, 1.494194694879908328e1,
    8.265296455388554217e1, 2.162779506621866970e2, 2.845145155184222157e2,
    1.851714491916334995e2, 5.486540717439723515e1, 6.118075837628957015e0,
    1.586261269326235053e-1};
static const double K0_Y = 1.137250900268554688;
static const double K1_Y  = 8.69547128677368164e-2;
static const double K1_Y2 = 1.45034217834472656;
static const double K1_P1[] = {-3.62137953440350228e-3, 7.11842087490330300e-3,
    1.00302560256614306e-5, 1.77231085381040811e-6};
static const double K1_Q1[] = {1.00000000000000000e0, -4.80414794429043831e-2,
    9.85972641934416525e-4, -8.91196859397070326e-6};
static const double K1_P2[] = {-3.07965757829206184e-1, -7.80929703673074907e-02,
    -2.70619343754051620e-3, -2.49549522229072008e-5};
static const double K1_Q2[] = {1.00000000000000000e0, -2.36316836412163098e-2,
    2.64524577525962719e-4, -1.49749618004162787e-6};
static const double K1_P3[] = {-1.97028041029226295e-1, -2.32408961548087617e0,
    -7.98269784507699938e0, -2.39968410774221632e0, 3.28314043780858713e1,
    5.67713761158496058e1, 3.30907788466509823e1, 6.62582288933739787e0,
    3.08851840645286691e-1};
static const double K1_Q3[] = {1.00000000000000000e0, 1.41811409298826118e1,
    7.35979466317556420e1, 1.77821793937080859e2, 2.11014501598705982e2,
    1.19425262951064454e2, 2.88448064302447607e1, 2.27912927104139732e0,
    2.50358186953478678e-2};
template <int N>
static inline double horner(double x, const double (&p)[N]) {
    double ex = p[N - 1];
    for (int i = N - 2; i >= 0; --i) ex = std::fma(x, ex, p[i]);
    return ex;
}
}  // namespace bk

// src2d/bessel2d.jl:79-89
static inline double besselk0x(double x) {
    using namespace bk;
    if (x <= 1.0) {
        double a = x * x / 4;
        double s = std::fma(horner(a, K0_P1), 1.0 / horner(a, K0_Q1), K0_Y);
        a = std::fma(s, a, 1.0);
        return std::fma(-a, std::log(x), horner(x * x, K0_P2)) * std::exp(x);
    }
    double ix = 1.0 / x;
    return std::fma(horner(ix, K0_P3), 1.0 / horner(ix, K0_Q3), 1.0) / std::sqrt(x);
}
// src2d/bessel2d.jl:91-103
static inline double besselk1x(double x) {
    using namespace bk;
    if (x <= 1.0) {
        double z = x * x, a = z / 4;
        double pq = std::fma(horner(a, K1_P1), 1.0 / horner(a, K1_Q1), K1_Y);
        pq = std::fma(pq * a, a, (a / 2 + 1));
        a = pq * x / 2;
        pq = std::fma(horner(z, K1_P2) / horner(z, K1_Q2), x, 1.0 / x);
        return std::fma(a, std::log(x), pq) * std::exp(x);
    }
    double ix = 1.0 / x;
    return std::fma(horner(ix, K1_P3), 1.0 / horner(ix, K1_Q3), K1_Y2) / std::sqrt(x);
}
// Large-x series of safe_besselkx (src/eos.jl:27-34)
static inline double besselkx_asym(int nu, double x) {
    double mu = 4.0 * double(nu) * double(nu);
    double ix = 1.0 / (8.0 * x);
    double s = 1.0 + (mu - 1.0) * ix * (1.0 + (mu - 9.0) * ix / 2.0 * (1.0 + (mu - 25.0) * ix / 3.0));
    return std::sqrt(PI_ / (2.0 * x)) * s;
}
// e^x K_nu(x), nu = 0..5, for 0 < x < 1e4: K0, K1 above and the upward
// recurrence K_{n+1} = K_{n-1} + (2n/x) K_n (Abramowitz & Stegun 9.6.26), which is
// stable for K. Stands in for SpecialFunctions.besselkx (AMOS), which the
// reference calls at src2d/transport2d.jl:23-105 and src/eos.jl:29.
static inline double besselkx_rec(int nu, double x) {
    double km = besselk0x(x);
    if (nu == 0) return km;
    double k = besselk1x(x);
    for (int n = 1; n < nu; ++n) {
        double kp = km + (2.0 * n / x) * k;
        km = k; k = kp;
    }
    return k;
}
// safe_besselkx (src/eos.jl:27-34)
static inline double safe_besselkx(int nu, double x) {
    if (!(isf(x) && x > 0.0)) return 0.0;
    if (x < BESSELKX_ASYM_X) return besselkx_rec(nu, x);
    return besselkx_asym(nu, x);
}
// safe_besselk2x_2d (src2d/bessel2d.jl:111-115)
static inline double safe_besselk2x_2d(double x) {
    if (!(isf(x) && x > 0.0)) return 0.0;
    if (x < BESSELKX_ASYM_X) return besselk0x(x) + (2 / x) * besselk1x(x);
    return besselkx_asym(2, x);
}

// ---- EOS (src/eos.jl) ----------------------------------------------------------
struct EOS {
    int kind;
    bool hq_fm;
    double g_eff, m_hq, g_hq;
    double a1, a2, a3, a4, b1, b2, b3, b4, c, d, canon;
    explicit EOS(const fivo_eos_desc& e)
        : kind(e.kind), hq_fm(e.hq_times_fmGeV3 != 0), g_eff(e.g_eff), m_hq(e.m_hq),
          g_hq(e.g_hq), a1(e.a1), a2(e.a2), a3(e.a3), a4(e.a4), b1(e.b1), b2(e.b2),
          b3(e.b3), b4(e.b4), c(e.c), d(e.d), canon(e.canon_factor) {}
    EOS() = default;
    double mass() const { return m_hq; }
    // eos_has_charge_2d (src2d/primrec2d.jl:273-276)
    bool has_charge() const { return kind == FIVO_EOS_LATTICE_HRG ? true : (g_hq > 0.0); }
};

static const double LK0 = 0.0005624486560000001, LK1 = 0.003652264, LK2 = 0.023716,
                    LK3 = 0.154, LK4 = 5.208957878352717;

// light_P (src/eos.jl:357-377)
static inline double light_P(double T, const EOS& q) {
    T = jmax(T, T_MIN);
    double A = 0.01 * (q.d * q.d), B = 0.1 * (q.c * q.c);
    double E = -(A / (T * T)) - (B / T);
    E = jclamp(E, -700.0, 700.0);
    double expE = std::exp(E);
    double N = LK0 * q.a4 + LK1 * q.a3 * T + LK2 * q.a2 * (T * T) + LK3 * q.a1 * pow3(T) + LK4 * pow4(T);
    double D = LK0 * q.b4 + LK1 * q.b3 * T + LK2 * q.b2 * (T * T) + LK3 * q.b1 * pow3(T) + pow4(T);
    return FMGEV3 * expE * pow4(T) * (N / D);
}
// light_dP_dT (src/eos.jl:379-412)
static inline double light_dP_dT(double T, const EOS& q) {
    T = jmax(T, T_MIN);
    double A = 0.01 * (q.d * q.d), B = 0.1 * (q.c * q.c);
    double E = -(A / (T * T)) - (B / T);
    E = jclamp(E, -700.0, 700.0);
    double expE = std::exp(E);
    double Ep = (2 * A / pow3(T)) + (B / (T * T));
    double N = LK0 * q.a4 + LK1 * q.a3 * T + LK2 * q.a2 * (T * T) + LK3 * q.a1 * pow3(T) + LK4 * pow4(T);
    double D = LK0 * q.b4 + LK1 * q.b3 * T + LK2 * q.b2 * (T * T) + LK3 * q.b1 * pow3(T) + pow4(T);
    double Np = LK1 * q.a3 + 2 * LK2 * q.a2 * T + 3 * LK3 * q.a1 * (T * T) + 4 * LK4 * pow3(T);
    double Dp = LK1 * q.b3 + 2 * LK2 * q.b2 * T + 3 * LK3 * q.b1 * (T * T) + 4 * pow3(T);
    double G = pow4(T) * N, H = D;
    double Gp = 4 * pow3(T) * N + pow4(T) * Np, Hp = Dp;
    double F = G / H;
    double Fp = (Gp * H - G * Hp) / (H * H);
    return FMGEV3 * expE * (Fp + F * Ep);
}
// light_d2P_dT2 (src/eos.jl:414-460)
static inline double light_d2P_dT2(double T, const EOS& q) {
    T = jmax(T, T_MIN);
    double A = 0.01 * (q.d * q.d), B = 0.1 * (q.c * q.c);
    double E = -(A / (T * T)) - (B / T);
    E = jclamp(E, -700.0, 700.0);
    double expE = std::exp(E);
    double Ep = (2 * A / pow3(T)) + (B / (T * T));
    double Epp = (-6 * A / pow4(T)) - (2 * B / pow3(T));
    double N = LK0 * q.a4 + LK1 * q.a3 * T + LK2 * q.a2 * (T * T) + LK3 * q.a1 * pow3(T) + LK4 * pow4(T);
    double D = LK0 * q.b4 + LK1 * q.b3 * T + LK2 * q.b2 * (T * T) + LK3 * q.b1 * pow3(T) + pow4(T);
    double Np = LK1 * q.a3 + 2 * LK2 * q.a2 * T + 3 * LK3 * q.a1 * (T * T) + 4 * LK4 * pow3(T);
    double Dp = LK1 * q.b3 + 2 * LK2 * q.b2 * T + 3 * LK3 * q.b1 * (T * T) + 4 * pow3(T);
    double Npp = 2 * LK2 * q.a2 + 6 * LK3 * q.a1 * T + 12 * LK4 * (T * T);
    double Dpp = 2 * LK2 * q.b2 + 6 * LK3 * q.b1 * T + 12 * (T * T);
    double G = pow4(T) * N, H = D;
    double Gp = 4 * pow3(T) * N + pow4(T) * Np, Hp = Dp;
    double Gpp = 12 * (T * T) * N + 8 * pow3(T) * Np + pow4(T) * Npp, Hpp = Dpp;
    double F = G / H;
    double Fp = (Gp * H - G * Hp) / (H * H);
    double Q = (Gp * H - G * Hp);
    double Qp = (Gpp * H - G * Hpp);
    double Fpp = (Qp * H - 2 * Q * Hp) / (H * H * H);
    return FMGEV3 * expE * (Fpp + 2 * Fp * Ep + F * (Epp + Ep * Ep));
}

struct PNE { double P, n, e; };

// eos_Pne for ConformalHQEOS (src/eos.jl:51-84) and LatticeHRGEOS (src/eos.jl:481-519)
static inline PNE eos_Pne(double T, double mu, const EOS& q) {
    T = jmax(T, T_MIN);
    if (q.kind == FIVO_EOS_CONFORMAL_HQ) {
        double Plight = (PI2 / 90) * q.g_eff * pow4(T) * FMGEV3;
        double elight = 3.0 * Plight;
        double m = q.m_hq, g = q.g_hq, x = m / T;
        if (!isf(x) || x <= 0.0) return {Plight, 0.0, elight};
        double A = g * (m * m) / (2 * PI2);
        double K2x = safe_besselkx(2, x), K1x = safe_besselkx(1, x);
        double Efac = std::exp((mu - m) / T);
        double n_hq = A * T * Efac * K2x;
        double P_hq = A * (T * T) * Efac * K2x;
        double e_hq = A * (T * T) * Efac * (3.0 * K2x + x * K1x);
        if (q.hq_fm) { n_hq *= FMGEV3; P_hq *= FMGEV3; e_hq *= FMGEV3; }
        return {Plight + P_hq, n_hq, elight + e_hq};
    }
    double Pl = light_P(T, q);
    double el = -Pl + T * light_dP_dT(T, q);    // light_e (src/eos.jl:463-467)
    double m = q.m_hq, g = q.g_hq, x = jmin(m / T, 100000.0);
    if (!isf(x) || x <= 0.0) return {Pl, 0.0, el};
    double A = q.canon * g * (m * m) / (2 * PI2);
    double K2x = safe_besselkx(2, x);
    double z = jclamp((mu - m) / T, -700.0, 700.0);
    double n_hq = A * T * std::exp(z) * K2x;
    double P_hq = 0.0, e_hq = 0.0;
    if (q.hq_fm) { n_hq *= FMGEV3; P_hq *= FMGEV3; e_hq *= FMGEV3; }
    return {Pl + P_hq, n_hq, el + e_hq};
}

// eos_cs2 (src/eos.jl:49 and :471-476)
static inline double eos_cs2(double T, double /*mu*/, const EOS& q) {
    if (q.kind == FIVO_EOS_CONFORMAL_HQ) return 1.0 / 3.0;
    T = jmax(T, T_MIN);
    double Pp = light_dP_dT(T, q), Ppp = light_d2P_dT2(T, q);
    return (Ppp == 0.0) ? (1.0 / 3.0) : (Pp / (T * Ppp));
}

// EOSCache2D: 4-slot memo of the T-only part, bit-neutral (src2d/primrec2d.jl:115-231)
struct EOSCache {
    double T[4], Pl[4], el[4], K2[4];
    int next;
    void reset() { for (int j = 0; j < 4; ++j) T[j] = std::numeric_limits<double>::quiet_NaN(); next = 0; }
};

// _lhrg_assemble_2d (src2d/primrec2d.jl:147-166)
static inline PNE lhrg_assemble(double T, double mu, const EOS& q, double x, double Pl, double el,
                                double K2x) {
    if (!isf(x) || x <= 0.0) return {Pl, 0.0, el};
    double m = q.m_hq, g = q.g_hq;
    double A = q.canon * g * (m * m) / (2 * PI2);
    double z = jclamp((mu - m) / T, -700.0, 700.0);
    double n_hq = A * T * std::exp(z) * K2x;
    double P_hq = 0.0, e_hq = 0.0;
    if (q.hq_fm) { n_hq *= FMGEV3; P_hq *= FMGEV3; e_hq *= FMGEV3; }
    return {Pl + P_hq, n_hq, el + e_hq};
}
static inline double lhrg_x(double T, const EOS& q) { return jmin(q.m_hq / T, 100000.0); }
static inline double lhrg_K2(double x) { return (!isf(x) || x <= 0.0) ? 0.0 : safe_besselk2x_2d(x); }

// eos_Pne_2d (src2d/primrec2d.jl:178-206); cache may be null
static inline PNE eos_Pne_2d(double T, double mu, const EOS& q, EOSCache* c = nullptr) {
    if (q.kind != FIVO_EOS_LATTICE_HRG) return eos_Pne(T, mu, q);
    T = jmax(T, T_MIN);
    double x = lhrg_x(T, q);
    if (!c) {
        double Pl = light_P(T, q);
        double el = -Pl + T * light_dP_dT(T, q);
        return lhrg_assemble(T, mu, q, x, Pl, el, lhrg_K2(x));
    }
    int k = -1;
    for (int j = 0; j < 4; ++j) if (c->T[j] == T) { k = j; break; }
    if (k < 0) {
        k = c->next;
        c->next = (k == 3) ? 0 : k + 1;
        double Pl = light_P(T, q);
        c->T[k] = T; c->Pl[k] = Pl;
        c->el[k] = -Pl + T * light_dP_dT(T, q);
        c->K2[k] = lhrg_K2(x);
    }
    return lhrg_assemble(T, mu, q, x, c->Pl[k], c->el[k], c->K2[k]);
}
// eos_Pe_2d (src2d/primrec2d.jl:218-234): P and e only (no Bessel for LHRG)
static inline void eos_Pe_2d(double T, double mu, const EOS& q, double* P, double* e) {
    if (q.kind != FIVO_EOS_LATTICE_HRG) { PNE r = eos_Pne(T, mu, q); *P = r.P; *e = r.e; return; }
    T = jmax(T, T_MIN);
    double Pl = light_P(T, q);
    double el = -Pl + T * light_dP_dT(T, q);
    *P = Pl + 0.0; *e = el + 0.0;
}

// ---- shear algebra (src2d/shear2d.jl) ----------------------------------------
struct Pi4 { double tt, tx, ty; };
// shear_tensor_contravariant_2d (src2d/shear2d.jl:61-75): τ-row from orthogonality
static inline Pi4 shear_tau_row(double ux, double uy, double ut, double pixx, double pixy,
                                double piyy) {
    double iu = safe_inv(ut);
    double tx = (ux * pixx + uy * pixy) * iu;
    double ty = (ux * pixy + uy * piyy) * iu;
    double tt = (ux * tx + uy * ty) * iu;
    return {tt, tx, ty};
}
// shear_constraint_residual_2d (src2d/shear2d.jl:88-101) -> relative residual
static inline double shear_rel_residual(double ux, double uy, double ut, double pixx, double pixy,
                                        double piyy, double pieta) {
    Pi4 p = shear_tau_row(ux, uy, ut, pixx, pixy, piyy);
    double R = -p.tt + pixx + piyy + pieta;
    double scale = jmax(jmax(jmax(jmax(std::fabs(p.tt), std::fabs(pixx)), std::fabs(piyy)),
                             std::fabs(pieta)), std::fabs(pixy));
    return scale > TINY ? std::fabs(R) / scale : 0.0;
}
// project_shear_traceless_2d (src2d/shear2d.jl:131-142)
static inline double shear_project_pieta(double ux, double uy, double ut, double pixx, double pixy,
                                         double piyy) {
    Pi4 p = shear_tau_row(ux, uy, ut, pixx, pixy, piyy);
    return p.tt - pixx - piyy;
}
// nu_tau_2d (src2d/shear2d.jl:180-181)
static inline double nu_tau(double ux, double uy, double ut, double nux, double nuy) {
    return (ux * nux + uy * nuy) * safe_inv(ut);
}

}  // namespace orc
