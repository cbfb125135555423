// =============================================================================
// fivo_device.h -- per-cell FP64 physics of the 2+1D hot path, as device code.
//
// Restates, for one thread per cell or face, the reference's per-cell functions:
// the EOS (src/eos.jl:39-84, 318-519; src2d/primrec2d.jl:115-234), the scaled
// Bessel K (src2d/bessel2d.jl:79-115, src/eos.jl:27-34), the primitive recovery
// (src2d/primrec2d.jl:33-742), the forward map, wave speeds and HLLE flux
// (src2d/primrec2d.jl:753-796, src2d/fluxes2d.jl:24-177) and the MC limiter
// (src/utils.jl:14-22). Julia semantics are kept where they change results:
// NaN-propagating min/max, the NaN-passing clamp, sign(0) = 0.
//
// Design for gfx950: everything is FP64 VALU; the light-sector EOS evaluates
// P, dP/dT and d2P/dT2 from one exp; the Newton keeps the T-only part of the
// EOS of the current iterate in registers (the reference's bit-neutral memo,
// src2d/primrec2d.jl:115-206), so 6 of 8 Jacobian columns cost no exp of the
// light sector; the 4x4 solve is fully unrolled so nothing spills to scratch.
// =============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fd {

constexpr double T_MIN = 1e-20;
constexpr double T_SOLVE_MAX = 10.0;
constexpr double PHI_CAP = 10000.0;
constexpr double E_VAC = 1e-20;
constexpr double E_FLOOR = 1e-20;
constexpr double CHI_SRE = 0.99999999;
constexpr double TINY = 1e-300;
constexpr double EPS_POS = 1e-50;
constexpr double U_CAP = 1.0e4;
constexpr double SR_MARGIN = 0.98;
constexpr double ASYM_X = 1.0e4;
constexpr double PI_ = 3.141592653589793;

enum : int { PRR_UNSET = 0, PRR_VACUUM = 1, PRR_CONVERGED = 2, PRR_NONFINITE_INITIAL = 3,
             PRR_NONFINITE_JACOBIAN = 4, PRR_LINSOLVE_FAILED = 5, PRR_LINESEARCH_FAILED = 6,
             PRR_RESIDUAL_TOO_LARGE = 8, PRR_MAXIT = 9 };

// ---- Julia Base semantics ----------------------------------------------------
__device__ __forceinline__ bool isf(double x) { return __builtin_isfinite(x); }
__device__ __forceinline__ bool isnan_(double x) { return x != x; }
__device__ __forceinline__ bool sgn(double x) { return __double_as_longlong(x) < 0; }
__device__ __forceinline__ double jmax(double x, double y) {
    bool py = (y > x) || (!sgn(y) && sgn(x));
    return py ? (isnan_(x) ? x : y) : (isnan_(y) ? y : x);
}
__device__ __forceinline__ double jmin(double x, double y) {
    bool py = (y < x) || (sgn(y) && !sgn(x));
    return py ? (isnan_(x) ? x : y) : (isnan_(y) ? y : x);
}
__device__ __forceinline__ double jclamp(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }
__device__ __forceinline__ double jsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }
__device__ __forceinline__ double posden(double x) { return x > EPS_POS ? x : EPS_POS; }
__device__ __forceinline__ double safe_inv(double x) { return 1.0 / posden(x); }
__device__ __forceinline__ double safe_div(double a, double b) { return a / posden(b); }
__device__ __forceinline__ double safe_sqrt(double x) { return x <= 0 ? 0.0 : sqrt(x); }
__device__ __forceinline__ double pow4(double x) { double x2 = x * x; return x2 * x2; }
__device__ __forceinline__ double pow3(double x) { return x * x * x; }
// max of N values none of which is NaN, as a tree of v_max: exact and independent of
// the order, so it equals any sequential `max` loop over the same values
template <int N>
__device__ __forceinline__ double max_tree(const double* v) {
    if constexpr (N == 1) return v[0];
    else return fmax(max_tree<N / 2>(v), max_tree<N - N / 2>(v + N / 2));
}
// Julia's `m = 0.0; for r: m = max(m, |x_r|)` over non-negative |x_r|: NaN when
// any is NaN (a sum carries it beside the tree), the tree's max otherwise
template <int N>
__device__ __forceinline__ double max_abs_nan(const double* v) {
    double a[N], s = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) { a[r] = fabs(v[r]); s += a[r]; }
    const double m = max_tree<N>(a);
    return isnan_(s) ? s : m;
}
__device__ __forceinline__ double dinf() { return __longlong_as_double(0x7ff0000000000000LL); }
// Julia's max(x, c) and clamp(x, a, b) where x is known not to be NaN (Newton
// iterates and trial points are finite, their temperatures exp of a finite yT):
// one v_max / v_min each, the same values
__device__ __forceinline__ double max_nn(double x, double c) { return fmax(x, c); }
__device__ __forceinline__ double clamp_nn(double x, double a, double b) { return fmin(fmax(x, a), b); }

// mc_limiter (src/utils.jl:14-22)
__device__ __forceinline__ double mc_limiter(double a, double b) {
    double sa = jsign(a), sb = jsign(b);
    if (sa != sb) return 0.0;
    double aa = fabs(a), bb = fabs(b);
    return sa * jmin(2 * aa, jmin(2 * bb, 0.5 * (aa + bb)));
}

// ---- Bessel K ----------------------------------------------------------------
// e^x K0(x), e^x K1(x): Boost's rational minimax approximations (coefficients by
// P. Holoborodko), the table the reference uses (src2d/bessel2d.jl:37-103).
// evalpoly: p0 + x*(p1 + x*(... + x*pn)) with one rounding per step (fma)
__device__ __forceinline__ double hp(double, double c0) { return c0; }
template <typename... R>
__device__ __forceinline__ double hp(double x, double c0, R... r) { return fma(x, hp(x, r...), c0); }
#define FD_H(x, ...) hp(x, __VA_ARGS__)
__device__ __forceinline__ double k0x_small_P1(double a) { return FD_H(a, -1.372509002685546267e-1, 2.574916117833312855e-1, 1.395474602146869316e-2, 5.445476986653926759e-4, 7.125159422136622118e-6); }
__device__ __forceinline__ double k0x_small_Q1(double a) { return FD_H(a, 1.0, -5.458333438017788530e-02, 1.291052816975251298e-03, -1.367653946978586591e-05); }
__device__ __forceinline__ double k0x_small_P2(double a) { return FD_H(a, 1.159315156584124484e-01, 2.789828789146031732e-01, 2.524892993216121934e-02, 8.460350907213637784e-04, 1.491471924309617534e-05, 1.627106892422088488e-07, 1.208266102392756055e-09, 6.611686391749704310e-12); }
__device__ __forceinline__ double k0x_big_P3(double a) { return FD_H(a, 2.533141373155002416e-1, 3.628342133984595192e0, 1.868441889406606057e1, 4.306243981063412784e1, 4.424116209627428189e1, 1.562095339356220468e1, -1.810138978229410898e0, -1.414237994269995877e0, -9.369168119754924625e-2); }
__device__ __forceinline__ double k0x_big_Q3(double a) { return FD_H(a, 1.0, 1.494194694879908328e1, 8.265296455388554217e1, 2.162779506621866970e2, 2.845145155184222157e2, 1.851714491916334995e2, 5.486540717439723515e1, 6.118075837628957015e0, 1.586261269326235053e-1); }
__device__ __forceinline__ double k1x_small_P1(double a) { return FD_H(a, -3.62137953440350228e-3, 7.11842087490330300e-3, 1.00302560256614306e-5, 1.77231085381040811e-6); }
__device__ __forceinline__ double k1x_small_Q1(double a) { return FD_H(a, 1.0, -4.80414794429043831e-2, 9.85972641934416525e-4, -8.91196859397070326e-6); }
__device__ __forceinline__ double k1x_small_P2(double a) { return FD_H(a, -3.07965757829206184e-1, -7.80929703673074907e-02, -2.70619343754051620e-3, -2.49549522229072008e-5); }
__device__ __forceinline__ double k1x_small_Q2(double a) { return FD_H(a, 1.0, -2.36316836412163098e-2, 2.64524577525962719e-4, -1.49749618004162787e-6); }
__device__ __forceinline__ double k1x_big_P3(double a) { return FD_H(a, -1.97028041029226295e-1, -2.32408961548087617e0, -7.98269784507699938e0, -2.39968410774221632e0, 3.28314043780858713e1, 5.67713761158496058e1, 3.30907788466509823e1, 6.62582288933739787e0, 3.08851840645286691e-1); }
__device__ __forceinline__ double k1x_big_Q3(double a) { return FD_H(a, 1.0, 1.41811409298826118e1, 7.35979466317556420e1, 1.77821793937080859e2, 2.11014501598705982e2, 1.19425262951064454e2, 2.88448064302447607e1, 2.27912927104139732e0, 2.50358186953478678e-2); }
#undef FD_H

// e^x K0 and e^x K1 together: the large-x branch shares 1/x and sqrt(x),
// the small-x branch shares log(x) and exp(x). Out of line (it is large and
// called from many places); the pair comes back in registers.
struct K01 { double k0, k1; };
__device__ __noinline__ K01 besselk01x(double x) {
    K01 r;
    if (x <= 1.0) {
        double lx = log(x), ex = exp(x), z = x * x, a = z / 4;
        double s = fma(k0x_small_P1(a), 1.0 / k0x_small_Q1(a), 1.137250900268554688);
        double a0 = fma(s, a, 1.0);
        r.k0 = fma(-a0, lx, k0x_small_P2(z)) * ex;
        double pq = fma(k1x_small_P1(a), 1.0 / k1x_small_Q1(a), 8.69547128677368164e-2);
        pq = fma(pq * a, a, (a / 2 + 1));
        double a1 = pq * x / 2;
        double pq2 = fma(k1x_small_P2(z) / k1x_small_Q2(z), x, 1.0 / x);
        r.k1 = fma(a1, lx, pq2) * ex;
    } else {
        double ix = 1.0 / x, sx = sqrt(x);
        r.k0 = fma(k0x_big_P3(ix), 1.0 / k0x_big_Q3(ix), 1.0) / sx;
        r.k1 = fma(k1x_big_P3(ix), 1.0 / k1x_big_Q3(ix), 1.45034217834472656) / sx;
    }
    return r;
}
// large-x series of safe_besselkx (src/eos.jl:27-34)
__device__ __forceinline__ double besselkx_asym(int nu, double x) {
    double mu = 4.0 * double(nu) * double(nu);
    double ix = 1.0 / (8.0 * x);
    double s = 1.0 + (mu - 1.0) * ix * (1.0 + (mu - 9.0) * ix / 2.0 * (1.0 + (mu - 25.0) * ix / 3.0));
    return sqrt(PI_ / (2.0 * x)) * s;
}
// safe_besselkx(nu, x), nu = 0..5: K0, K1 and the upward recurrence
// K_{n+1} = K_{n-1} + (2n/x) K_n (Abramowitz & Stegun 9.6.26), stable for K.
__device__ __forceinline__ double safe_besselkx(int nu, double x) {
    if (!(isf(x) && x > 0.0)) return 0.0;
    if (x >= ASYM_X) return besselkx_asym(nu, x);
    const K01 b = besselk01x(x);
    double km = b.k0, k = b.k1;
    if (nu == 0) return km;
    for (int n = 1; n < nu; ++n) { double kp = km + (2.0 * n / x) * k; km = k; k = kp; }
    return k;
}
// safe_besselkx(1, x) and safe_besselkx(2, x) from one K0/K1 evaluation
__device__ __forceinline__ void besselk12x(double x, double* k1, double* k2) {
    if (!(isf(x) && x > 0.0)) { *k1 = 0.0; *k2 = 0.0; return; }
    if (x >= ASYM_X) { *k1 = besselkx_asym(1, x); *k2 = besselkx_asym(2, x); return; }
    const K01 b = besselk01x(x);
    *k1 = b.k1;
    *k2 = b.k0 + (2.0 * 1 / x) * b.k1;
}
// safe_besselk2x_2d (src2d/bessel2d.jl:111-115)
__device__ __forceinline__ double besselk2x_fast(double x) {
    if (!(isf(x) && x > 0.0)) return 0.0;
    if (x >= ASYM_X) return besselkx_asym(2, x);
    const K01 b = besselk01x(x);
    return b.k0 + (2 / x) * b.k1;
}

// ---- EOS constants (host-prepared, same arithmetic as the reference) ----------
struct EosP {
    int kind;            // 0 conformal, 1 lattice-HRG
    int has_charge;      // eos_has_charge_2d
    int hq_fm;
    int conf_charm;      // conformal with g_hq > 0: charm sector evaluated
    double m;            // hq_mass
    double ghq;          // g_hq (charm degeneracy)
    double aSB;          // (pi^2/90) g_eff
    double Aconf;        // g m^2 / (2 pi^2)
    double Alhrg;        // canon g m^2 / (2 pi^2)
    double fmGeV3;
    // lattice light sector: A = 0.01 d^2, B = 0.1 c^2 and the products of the
    // fit coefficients with the k-constants in src/eos.jl:357-460
    double LA, LB, nA4, nA3, nA2, nA1, nA0, dB4, dB3, dB2, dB1;
    double nP1, nP2, nP3, nP4, dP1, dP2, dP3;   // first-derivative coefficients
    double nQ2, nQ3, nQ4, dQ2, dQ3;             // second-derivative coefficients
};

// light_P, light_dP_dT, light_d2P_dT2 (src/eos.jl:357-460) from one exp. T >= T_MIN.
template <int ORDER>
__device__ __forceinline__ void light_eval(double T, const EosP& q, double* P, double* dP, double* d2P) {
    double T2 = T * T, T3 = pow3(T), T4 = pow4(T);
    double E = -(q.LA / T2) - (q.LB / T);
    E = jclamp(E, -700.0, 700.0);
    double expE = exp(E);
    double N = q.nA0 + q.nA1 * T + q.nA2 * T2 + q.nA3 * T3 + q.nA4 * T4;
    double D = q.dB4 + q.dB3 * T + q.dB2 * T2 + q.dB1 * T3 + T4;
    double ND = N / D;
    *P = q.fmGeV3 * expE * T4 * ND;
    if (ORDER < 1) return;
    double Ep = (2 * q.LA / T3) + (q.LB / T2);
    double Np = q.nP1 + q.nP2 * T + q.nP3 * T2 + q.nP4 * T3;
    double Dp = q.dP1 + q.dP2 * T + q.dP3 * T2 + 4 * T3;
    double G = T4 * N, H = D;
    double Gp = 4 * T3 * N + T4 * Np, Hp = Dp;
    double F = G / H;
    double Q = (Gp * H - G * Hp);
    double Fp = Q / (H * H);
    *dP = q.fmGeV3 * expE * (Fp + F * Ep);
    if (ORDER < 2) return;
    double Epp = (-6 * q.LA / T4) - (2 * q.LB / T3);
    double Npp = q.nQ2 + q.nQ3 * T + q.nQ4 * T2;
    double Dpp = q.dQ2 + q.dQ3 * T + 12 * T2;
    double Gpp = 12 * T2 * N + 8 * T3 * Np + T4 * Npp, Hpp = Dpp;
    double Qp = (Gpp * H - G * Hpp);
    double Fpp = (Qp * H - 2 * Q * Hp) / (H * H * H);
    *d2P = q.fmGeV3 * expE * (Fpp + 2 * Fp * Ep + F * (Epp + Ep * Ep));
}

// T-only part of the EOS at T (>= T_MIN): what the reference memoises per T
struct TPart { double T, Pl, el, K2, K1, x; };

// EOS kinds as template parameter EK: 0 = conformal without the charm sector
// (g_hq = 0), 2 = conformal with it, 1 = lattice-HRG.
template <int EK>
__device__ __forceinline__ TPart eos_tpart(double T, const EosP& q) {
    constexpr bool CH = EK != 0;
    TPart t;
    t.T = T;
    if (EK != 1) {
        t.Pl = q.aSB * pow4(T) * q.fmGeV3;
        t.el = 3.0 * t.Pl;
        t.x = q.m / T;
        t.K2 = 0.0; t.K1 = 0.0;
        if (CH) besselk12x(t.x, &t.K1, &t.K2);
    } else {
        double P, dP, d2P;
        light_eval<1>(T, q, &P, &dP, &d2P);
        t.Pl = P;
        t.el = -P + T * dP;
        t.x = jmin(q.m / T, 100000.0);
        t.K2 = (!isf(t.x) || t.x <= 0.0) ? 0.0 : besselk2x_fast(t.x);
        t.K1 = 0.0;
    }
    return t;
}
// assemble (P, n, e) from the T-part and mu: eos_Pne (src/eos.jl:51-84) for the
// conformal EOS, _lhrg_assemble_2d (src2d/primrec2d.jl:147-166) for lattice-HRG.
template <int EK>
__device__ __forceinline__ void eos_assemble(const TPart& t, double mu, const EosP& q, double* P, double* n,
                                             double* e) {
    constexpr bool CH = EK != 0;
    const double T = t.T;
    if (EK != 1) {
        if (!CH || !isf(t.x) || t.x <= 0.0) { *P = t.Pl; *n = 0.0; *e = t.el; return; }
        double Efac = exp((mu - q.m) / T);
        double nh = q.Aconf * T * Efac * t.K2;
        double Ph = q.Aconf * (T * T) * Efac * t.K2;
        double eh = q.Aconf * (T * T) * Efac * (3.0 * t.K2 + t.x * t.K1);
        if (q.hq_fm) { nh *= q.fmGeV3; Ph *= q.fmGeV3; eh *= q.fmGeV3; }
        *P = t.Pl + Ph; *n = nh; *e = t.el + eh;
    } else {
        if (!isf(t.x) || t.x <= 0.0) { *P = t.Pl; *n = 0.0; *e = t.el; return; }
        double z = jclamp((mu - q.m) / T, -700.0, 700.0);
        double nh = q.Alhrg * T * exp(z) * t.K2;
        if (q.hq_fm) nh *= q.fmGeV3;
        *P = t.Pl + 0.0; *n = nh; *e = t.el + 0.0;
    }
}
template <int EK>
__device__ __forceinline__ void eos_Pne_2d(double T, double mu, const EosP& q, double* P, double* n, double* e) {
    TPart t = eos_tpart<EK>(jmax(T, T_MIN), q);
    eos_assemble<EK>(t, mu, q, P, n, e);
}
// eos_Pe_2d (src2d/primrec2d.jl:218-234): P and e for the Newton seed
template <int EK>
__device__ __forceinline__ void eos_Pe_2d(double T, double mu, const EosP& q, double* P, double* e) {
    T = jmax(T, T_MIN);
    if (EK != 1) {
        double n;
        TPart t = eos_tpart<EK>(T, q);
        eos_assemble<EK>(t, mu, q, P, &n, e);
    } else {
        double Pl, dP, d2P;
        light_eval<1>(T, q, &Pl, &dP, &d2P);
        *P = Pl + 0.0;
        *e = (-Pl + T * dP) + 0.0;
    }
}
// eos_cs2 (src/eos.jl:49, 471-476)
template <int EK>
__device__ __forceinline__ double eos_cs2(double T, const EosP& q) {
    if (EK != 1) return 1.0 / 3.0;
    T = jmax(T, T_MIN);
    double P, dP, d2P;
    light_eval<2>(T, q, &P, &dP, &d2P);
    return (d2P == 0.0) ? (1.0 / 3.0) : (dP / (T * d2P));
}

// ---- shear algebra (src2d/shear2d.jl:61-181) ---------------------------------
__device__ __forceinline__ void shear_tau_row(double ux, double uy, double ut, double pixx, double pixy,
                                              double piyy, double* tt, double* tx, double* ty) {
    double iu = safe_inv(ut);
    *tx = (ux * pixx + uy * pixy) * iu;
    *ty = (ux * pixy + uy * piyy) * iu;
    *tt = (ux * *tx + uy * *ty) * iu;
}
__device__ __forceinline__ double nu_tau(double ux, double uy, double ut, double nux, double nuy) {
    return (ux * nux + uy * nuy) * safe_inv(ut);
}

// dissipative dofs of one cell in physical units (dissipatives_at)
struct Diss { double nux, nuy, Pi, pixx, pixy, piyy, pieta; };

// ---- primitive recovery (src2d/primrec2d.jl:300-742) --------------------------
struct PrimRes { double T, mu, ux, uy, n, e, P; int ok, reason, iters; int staged = 0; };   // staged: phi by bisection

// Group-cooperative form: G consecutive lanes of a wavefront solve one cell
// (G = 8 or 16; G = 0 is the one-lane form). Every lane of a group holds the
// same iterate; the lanes split the independent evaluations (finite-difference
// columns, line-search trials, bisection midpoints) and pick results with
// shuffles, so the iterate sequence is the sequential one, bit for bit. With
// G = 16 the upper eight lanes also evaluate the next iteration's Jacobian
// columns at the first line-search trial while the lower eight test the trials.
template <int G> __device__ __forceinline__ int glane() { return threadIdx.x & (G - 1); }
template <int G> __device__ __forceinline__ double gshfl(double v, int src) { return __shfl(v, src, G); }
template <int G> __device__ __forceinline__ unsigned gbits(bool pred) {   // pred of the group's lanes, bit l = lane l
    const unsigned long long b = __ballot(pred);
    return (unsigned)((b >> ((threadIdx.x & 63) & ~(G - 1))) & ((1ull << G) - 1));
}

// `nsteps` steps of  mid = (lo+hi)/2; pred(mid) ? hi = mid : lo = mid.
// COOP: L = log2(G) steps per round (3, 4 or 5 for 8, 16 or 32 lanes); lane l
// takes the midpoint of node l of the L-level tree below [lo, hi] (node n has
// children 2n+1 for pred true, 2n+2 otherwise), reached by the same halvings the
// sequential loop would make on its way there, so every midpoint is bit for bit
// the one the sequential bisection evaluates; the predicates' ballot then replays
// the sequential choices.
template <int G, class Pred>
__device__ __forceinline__ void bisect(double& lo, double& hi, int nsteps, Pred pred) {
    if constexpr (G == 0) {
        for (int it = 0; it < nsteps; ++it) {
            double mid = 0.5 * (lo + hi);
            if (pred(mid)) hi = mid; else lo = mid;
        }
        return;
    }
    constexpr int LG = G >= 32 ? 5 : (G >= 16 ? 4 : 3);
    const int l = glane<G>();
    for (int done = 0; done < nsteps;) {
        const int L = nsteps - done < LG ? nsteps - done : LG;
        const int nn = (1 << L) - 1;
        bool pl = false;
        if (l < nn) {
            // descend to node l: the bits of l + 1 below its leading one, 0 = pred true
            const int m = l + 1;
            const int d = 31 - __builtin_clz((unsigned)m);
            double a = lo, z = hi;
            for (int j = d - 1; j >= 0; --j) {
                const double mid = 0.5 * (a + z);
                if ((m >> j) & 1) a = mid; else z = mid;
            }
            pl = pred(0.5 * (a + z));
        }
        const unsigned b = gbits<G>(pl);
        int node = 0;
        for (int lev = 0; lev < L; ++lev) {
            const double mid = 0.5 * (lo + hi);
            const bool p = (b >> node) & 1u;
            if (p) hi = mid; else lo = mid;
            node = 2 * node + (p ? 1 : 2);
        }
        done += L;
    }
}

// _invert_e_for_T (src2d/primrec2d.jl:300-314): 60 bisection steps on log T
template <int EK, int G>
__device__ __forceinline__ double invert_e_for_T(double e_target, double phi, const EosP& q) {
    double lo = log(T_MIN), hi = log(T_SOLVE_MAX);
    if (e_target <= 0.0) return lo;
    bisect<G>(lo, hi, 60, [&](double mid) {
        double Tm = exp(mid), P, em;
        eos_Pe_2d<EK>(Tm, q.m + Tm * phi, q, &P, &em);
        return !isf(em) || em > e_target;
    });
    return 0.5 * (lo + hi);
}
// seed_from_conserved_2d (src2d/primrec2d.jl:323-347)
template <int EK, int G>
__device__ __forceinline__ void seed_from_conserved(double E, double Sx, double Sy, double phi, const EosP& q,
                                                 double* yT_o, double* ux_o, double* uy_o) {
    double M = sqrt(Sx * Sx + Sy * Sy);
    if (!isf(E) || E <= 0.0) { *yT_o = log(0.15); *ux_o = 0.0; *uy_o = 0.0; return; }
    double P = E / 3, yT = log(0.15), v = 0.0;
    for (int it = 0; it < 3; ++it) {
        double denom = E + P;
        v = denom > TINY ? jmin(M / denom, 0.999999) : 0.0;
        double e = denom * (1 - v * v) - P;
        if (e <= 0.0) e = jmax(E, TINY);
        yT = invert_e_for_T<EK, G>(e, phi, q);
        double T = exp(yT), Pn, en;
        eos_Pe_2d<EK>(T, q.m + T * phi, q, &Pn, &en);
        if (!isf(Pn)) break;
        P = Pn;
    }
    double gam = 1 / sqrt(jmax(1 - v * v, 1e-16));
    double us = gam * v;
    double ux = M > TINY ? us * Sx / M : 0.0;
    double uy = M > TINY ? us * Sy / M : 0.0;
    *yT_o = yT;
    *ux_o = jclamp(ux, -U_CAP, U_CAP);
    *uy_o = jclamp(uy, -U_CAP, U_CAP);
}

// conserved inputs of one cell and the Newton row scales
struct CellIn {
    double D, Sx, Sy, E;
    Diss d;
    double rs[4];
};

// evalF_2d! (src2d/primrec2d.jl:360-390) at x with T = exp(x0) and the EOS
// T-part of max(T, T_MIN) given. Writes F (all Inf when the state or the EOS is
// non-finite, like the reference) and returns whether every component is finite.
template <int EK, bool NCH>
__device__ __forceinline__ bool evalF(double F[4], double x0, double x1, double x2, double x3, double T,
                                      const TPart& tp, const CellIn& c, const EosP& q) {
    if (!(isf(x0) && isf(x1) && isf(x2) && isf(x3))) { F[0] = F[1] = F[2] = F[3] = dinf(); return false; }
    double mu = q.m + T * x1;
    double ut = sqrt(1 + x2 * x2 + x3 * x3);
    double P, n, e;
    eos_assemble<EK>(tp, mu, q, &P, &n, &e);
    if (!(isf(P) && isf(n) && isf(e))) { F[0] = F[1] = F[2] = F[3] = dinf(); return false; }
    const Diss& d = c.d;
    double weff = e + P + d.Pi, Ptot = P + d.Pi;
    double tt, tx, ty;
    shear_tau_row(x2, x3, ut, d.pixx, d.pixy, d.piyy, &tt, &tx, &ty);
    F[0] = NCH ? (n * ut + nu_tau(x2, x3, ut, d.nux, d.nuy) - c.D) : 0.0;
    F[1] = weff * ut * x2 + tx - c.Sx;
    F[2] = weff * ut * x3 + ty - c.Sy;
    F[3] = weff * ut * ut - Ptot + tt - c.E;
    return isf(F[0]) && isf(F[1]) && isf(F[2]) && isf(F[3]);
}

// solve_gauss_n! (src2d/primrec2d.jl:33-87), n = N, fully unrolled on registers
template <int N>
__device__ __forceinline__ bool solve_gauss(double J[4][4], const double F[4], double delta[4]) {
    double A[N][N + 1];
#pragma unroll
    for (int r = 0; r < N; ++r) {
#pragma unroll
        for (int c = 0; c < N; ++c) A[r][c] = J[r][c];
        A[r][N] = -F[r];
    }
    // smax = max |A_rc| (Julia's loop of `max`, which fails the check below on any
    // NaN or Inf): a finiteness test and a tree of v_max, exact and order-free for
    // finite values, in place of a 16-deep chain of NaN-propagating maxima
    double aa[N * N];
    bool fin_in = true;
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) { aa[r * N + c] = fabs(A[r][c]); fin_in = fin_in && isf(A[r][c]); }
    const double smax = max_tree<N * N>(aa);
    if (!fin_in || smax == 0.0) return false;
    const double pivtol = 1e-15 * smax;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        // partial pivoting, the first row of the largest |A_rk| (entries stay finite:
        // |f| <= 1 below); a pairwise tree for four candidates, the sequential scan
        // otherwise; select-based swap keeps the matrix in registers
        double amax = fabs(A[k][k]);
        int p = k;
        if constexpr (N == 4) {
            if (k == 0) {
                const double a0 = fabs(A[0][0]), a1 = fabs(A[1][0]), a2 = fabs(A[2][0]), a3 = fabs(A[3][0]);
                const bool s1 = a1 > a0, s3 = a3 > a2;
                const double m01 = s1 ? a1 : a0, m23 = s3 ? a3 : a2;
                const bool hi = m23 > m01;
                amax = hi ? m23 : m01;
                p = hi ? (s3 ? 3 : 2) : (s1 ? 1 : 0);
            } else {
#pragma unroll
                for (int r = k + 1; r < N; ++r) if (fabs(A[r][k]) > amax) { amax = fabs(A[r][k]); p = r; }
            }
        } else {
#pragma unroll
            for (int r = k + 1; r < N; ++r) if (fabs(A[r][k]) > amax) { amax = fabs(A[r][k]); p = r; }
        }
        if (amax < pivtol) return false;
#pragma unroll
        for (int r = k + 1; r < N; ++r) {
            bool sw = (p == r);
#pragma unroll
            for (int c = k; c <= N; ++c) {
                double a = A[k][c], b = A[r][c];
                A[k][c] = sw ? b : a;
                A[r][c] = sw ? a : b;
            }
        }
        double invp = 1.0 / A[k][k];
#pragma unroll
        for (int r = k + 1; r < N; ++r) {
            double f = A[r][k] * invp;
            if (f != 0.0) {
#pragma unroll
                for (int c = k; c <= N; ++c) A[r][c] -= f * A[k][c];
            }
        }
    }
#pragma unroll
    for (int r = N - 1; r >= 0; --r) {
        double s = A[r][N];
#pragma unroll
        for (int c = r + 1; c < N; ++c) s -= A[r][c] * delta[c];
        delta[r] = s / A[r][r];
    }
    bool fin = true;
#pragma unroll
    for (int r = 0; r < N; ++r) fin = fin && isf(delta[r]);
    return fin;
}

// _solve_phi_bisect (src2d/primrec2d.jl:416-434) at the (clamped) final T
template <int EK, int G>
__device__ __forceinline__ bool solve_phi_bisect(double D, double ux, double uy, double nux, double nuy,
                                              const TPart& tp, const EosP& q, double* phi_o) {
    double ut = sqrt(1 + ux * ux + uy * uy);
    double target = D - nu_tau(ux, uy, ut, nux, nuy);
    if (!isf(target) || target <= 0.0) { *phi_o = -PHI_CAP; return false; }
    double lo = -PHI_CAP, hi = PHI_CAP, P, nm, e;
    bisect<G>(lo, hi, 200, [&](double mid) {
        double Pb, nb, eb;
        eos_assemble<EK>(tp, q.m + tp.T * mid, q, &Pb, &nb, &eb);
        return isf(nb) && nb * ut > target;
    });
    double phi = 0.5 * (lo + hi);
    eos_assemble<EK>(tp, q.m + tp.T * phi, q, &P, &nm, &e);
    *phi_o = phi;
    return isf(nm);
}

// A lane's memo of EOS T-parts keyed by the exact log-temperature they were taken
// at: the same yT gives the same T and T-part, so a hit changes no result (the
// reference's memo, EOSCache2D in src2d/primrec2d.jl:115-206, is bit-neutral the
// same way). A cooperative group's lanes evaluate the finite-difference columns and
// line-search trials at points whose yT repeats from one iteration to the next once
// T and u have converged: in a charge-free cell's phi walk (the cold retry, ~56
// iterations) yT stays on one double or moves between two neighbouring ones, so
// two slots let those iterations skip every exp and T-part. Groups of 16 and 32
// lanes keep it (their registers past 256 live in AGPRs); 8-lane groups, which
// run two wavefronts per SIMD, do not (it would spill to scratch).
struct TCache {
    double y[2], T[2];
    TPart tp[2];
};
__device__ __forceinline__ int tcache_slot(const TCache& c, double y0) {
    return y0 == c.y[0] ? 0 : (y0 == c.y[1] ? 1 : -1);
}
__device__ __forceinline__ void tcache_put(TCache& c, double y0, double T, const TPart& tp) {
    c.y[1] = c.y[0]; c.T[1] = c.T[0]; c.tp[1] = c.tp[0];
    c.y[0] = y0; c.T[0] = T; c.tp[0] = tp;
}

// The Newton iterate of _cons_to_prim_attempt_2d!. `it` is the reference's loop
// variable: an iterate can stop before iteration `it` and resume from (x, it),
// since T, the T-part and F are pure functions of x.
struct NewtonState {
    double x[4];
    double Tr;      // exp(x[0])
    TPart tp;       // EOS T-part at max(Tr, T_MIN)
    double F[4];    // residual at x
    int it, iters, reason;
    bool converged;
    // G = 16: the Jacobian at x, evaluated ahead by the line search that chose x
    // (lane 8 + 2k holds column k); valid while jnext, non-finite if jbad
    double jcol[4];
    bool jnext, jbad;
    int jbase;      // lane of column 0
    TCache tc;      // cooperative forms: this lane's T-parts by yT
};
constexpr int NEWTON_MAXIT = 80;
__device__ __forceinline__ bool newton_paused(const NewtonState& s) {
    return !s.converged && s.reason == PRR_UNSET && s.it <= NEWTON_MAXIT;
}

template <bool NCH>
__device__ __forceinline__ CellIn make_cell(double D, double Sx, double Sy, double E, const Diss& d) {
    CellIn c;
    c.D = D; c.Sx = Sx; c.Sy = Sy; c.E = E; c.d = d;
    const double sD = jmax(fabs(D), 1e-12);
    const double sS = jmax(jmax(jmax(fabs(Sx), fabs(Sy)), 1e-10 * fabs(E)), 1e-12);
    const double sE = jmax(fabs(E), 1e-12);
    c.rs[0] = 1 / sD; c.rs[1] = 1 / sS; c.rs[2] = 1 / sS; c.rs[3] = 1 / sE;
    return c;
}

// set the iterate to x and evaluate T, the T-part and F there
template <int EK, bool NCH>
__device__ __forceinline__ bool newton_set(NewtonState& s, const CellIn& c, const EosP& q) {
    s.Tr = exp(s.x[0]);
    s.tp = eos_tpart<EK>(jmax(s.Tr, T_MIN), q);
    s.tc.y[0] = s.tc.y[1] = __longlong_as_double(0x7ff8000000000000LL);   // empty: NaN matches nothing
    return evalF<EK, NCH>(s.F, s.x[0], s.x[1], s.x[2], s.x[3], s.Tr, s.tp, c, q);
}

// head of _cons_to_prim_attempt_2d! (src2d/primrec2d.jl:511-590): vacuum test,
// warm start or cold seed, first residual. Returns true with the final R when
// the attempt ends here (vacuum, non-finite start). SEED = false: the caller
// guarantees a finite warm start or a vacuum state.
template <int EK, bool NCH, bool SEED, int G>
__device__ __forceinline__ bool newton_begin(const CellIn& c, double yT0, double phi0, double ux0, double uy0,
                                             const EosP& q, NewtonState& s, PrimRes& R) {
    const double yT_lo = log(T_MIN), yT_hi = log(T_SOLVE_MAX);
    if (!isf(c.E) || c.E <= E_VAC || !isf(c.D) || c.D < 0.0) {
        double P, n, e;
        eos_Pne_2d<EK>(T_MIN, 0.0, q, &P, &n, &e);   // eos_Pne at T_MIN: identical branch
        R = {T_MIN, 0.0, 0.0, 0.0, 0.0, jmax(c.E, 0.0), P, 0, PRR_VACUUM, 0};
        return true;
    }
    s.x[1] = (NCH && isf(phi0)) ? jclamp(phi0, -PHI_CAP, PHI_CAP) : 0.0;
    if (isf(yT0) && isf(ux0) && isf(uy0)) {
        s.x[0] = jclamp(yT0, yT_lo, yT_hi);
        s.x[2] = jclamp(ux0, -U_CAP, U_CAP);
        s.x[3] = jclamp(uy0, -U_CAP, U_CAP);
    } else if (SEED) {
        double sy, sx, syy;
        seed_from_conserved<EK, G>(c.E, c.Sx, c.Sy, s.x[1], q, &sy, &sx, &syy);
        s.x[0] = jclamp(sy, yT_lo, yT_hi); s.x[2] = sx; s.x[3] = syy;
    }
    s.it = 1; s.iters = 0; s.reason = PRR_UNSET; s.converged = false;
    if (!newton_set<EK, NCH>(s, c, q)) {
        double T = jclamp(s.Tr, T_MIN, T_SOLVE_MAX), P, n, e;
        eos_Pne_2d<EK>(T, q.m, q, &P, &n, &e);
        R = {T, q.m, 0.0, 0.0, 0.0, jmax(e, 0.0), P, 0, PRR_NONFINITE_INITIAL, 0};
        return true;
    }
    return false;
}

// finite-difference Jacobian, scaled rows (src2d/primrec2d.jl:605-632)
// Cooperative form: lane l < 8 evaluates column l/2 at x +- h e_k (fd_lane).
template <bool NCH>
__device__ __forceinline__ void fd_point(const double x[4], int l, double y[4], double& h, bool& act, int& si) {
    constexpr int NU = NCH ? 4 : 3;
    const int k = l >> 1;
    act = k < NU;
    si = NCH ? k : (k == 0 ? 0 : k + 1);
    const double xs = si == 0 ? x[0] : si == 1 ? x[1] : si == 2 ? x[2] : x[3];
    h = 1e-6 * max_nn(fabs(xs), 1.0);
    const double xv = (l & 1) ? xs - h : xs + h;
    y[0] = si == 0 ? xv : x[0];
    y[1] = si == 1 ? xv : x[1];
    y[2] = si == 2 ? xv : x[2];
    y[3] = si == 3 ? xv : x[3];
}
template <int G>
__device__ __forceinline__ void fd_column(const double Fl[4], double h, const CellIn& c, double col[4]) {
    // even lane 2k forms column k from its partner's F(x - h e_k)
    const double inv2h = 1 / (2 * h);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double Fm = __shfl_xor(Fl[r], 1, G);
        col[r] = ((Fl[r] - Fm) * inv2h) * c.rs[r];
    }
}

template <int EK, bool NCH, int G>
__device__ __forceinline__ bool newton_jacobian(NewtonState& s, const CellIn& c, const EosP& q,
                                                double J[4][4]) {
    constexpr int NU = NCH ? 4 : 3;
    if constexpr (G == 0) {
#pragma unroll
        for (int k = 0; k < NU; ++k) {
            const int si = NCH ? k : (k == 0 ? 0 : k + 1);
            double xs = s.x[si];
            double h = 1e-6 * max_nn(fabs(xs), 1.0);
            double Fp[4], Fm[4];
            double xp[4] = {s.x[0], s.x[1], s.x[2], s.x[3]}, xm[4] = {s.x[0], s.x[1], s.x[2], s.x[3]};
            xp[si] = xs + h; xm[si] = xs - h;
            bool a, b;
            if (si == 0) {
                const double Tp = exp(xp[0]), Tm = exp(xm[0]);
                TPart tpp = eos_tpart<EK>(max_nn(Tp, T_MIN), q);
                TPart tpm = eos_tpart<EK>(max_nn(Tm, T_MIN), q);
                a = evalF<EK, NCH>(Fp, xp[0], xp[1], xp[2], xp[3], Tp, tpp, c, q);
                b = evalF<EK, NCH>(Fm, xm[0], xm[1], xm[2], xm[3], Tm, tpm, c, q);
            } else {
                a = evalF<EK, NCH>(Fp, xp[0], xp[1], xp[2], xp[3], s.Tr, s.tp, c, q);
                b = evalF<EK, NCH>(Fm, xm[0], xm[1], xm[2], xm[3], s.Tr, s.tp, c, q);
            }
            if (!(a && b)) return false;
            double inv2h = 1 / (2 * h);
#pragma unroll
            for (int r = 0; r < NU; ++r) {
                const int rr = NCH ? r : r + 1;
                J[r][k] = ((Fp[rr] - Fm[rr]) * inv2h) * c.rs[rr];
            }
        }
        return true;
    }
    // lane l evaluates column k = l/2 at x + h e_k (l even) or x - h e_k (l odd)
    double y[4], h;
    bool act;
    int si;
    fd_point<NCH>(s.x, glane<G>(), y, h, act, si);
    double Tl = s.Tr;
    TPart tl = s.tp;
    if (act && si == 0) {
        const int slot = G >= 16 ? tcache_slot(s.tc, y[0]) : -1;
        if (slot >= 0) {
            Tl = slot == 0 ? s.tc.T[0] : s.tc.T[1];
            tl = slot == 0 ? s.tc.tp[0] : s.tc.tp[1];
        } else {
            Tl = exp(y[0]);
            tl = eos_tpart<EK>(max_nn(Tl, T_MIN), q);
            if (G >= 16) tcache_put(s.tc, y[0], Tl, tl);
        }
    }
    double Fl[4] = {0.0, 0.0, 0.0, 0.0};
    bool okl = true;
    if (act) okl = evalF<EK, NCH>(Fl, y[0], y[1], y[2], y[3], Tl, tl, c, q);
    if (gbits<G>(!okl) != 0u) return false;
    double col[4];
    fd_column<G>(Fl, h, c, col);
#pragma unroll
    for (int kk = 0; kk < NU; ++kk)
#pragma unroll
        for (int r = 0; r < NU; ++r) {
            const int rr = NCH ? r : r + 1;
            J[r][kk] = gshfl<G>(col[rr], 2 * kk);
        }
    return true;
}

// the line search of src2d/primrec2d.jl:648-691: trial ls at alpha 2^-ls,
// accepted at the first trial with a finite residual below (1 - 1e-4 a) nrm.
// On return x is the accepted trial, or the last one when none is.
template <int EK, bool NCH, int G>
__device__ __forceinline__ bool newton_linesearch(NewtonState& s, const CellIn& c, const EosP& q,
                                                  const double delta[4], double alpha0, double nrm) {
    constexpr int NU = NCH ? 4 : 3;
    constexpr int rlo = NCH ? 0 : 1;
    const double yT_lo = log(T_MIN), yT_hi = log(T_SOLVE_MAX);
    const double xs0 = s.x[0], xs1 = s.x[1], xs2 = s.x[2], xs3 = s.x[3];
    auto trial_point = [&](double alpha, double t[4]) {
        t[0] = xs0; t[1] = xs1; t[2] = xs2; t[3] = xs3;
#pragma unroll
        for (int k = 0; k < NU; ++k) {
            const int si = NCH ? k : (k == 0 ? 0 : k + 1);
            t[si] += alpha * delta[k];
        }
        // finite: x and delta are, and |alpha delta_k| <= 1
        t[0] = clamp_nn(t[0], yT_lo, yT_hi);
        if (NCH) t[1] = clamp_nn(t[1], -PHI_CAP, PHI_CAP);
        t[2] = clamp_nn(t[2], -U_CAP, U_CAP);
        t[3] = clamp_nn(t[3], -U_CAP, U_CAP);
    };
    if constexpr (G == 0) {
        double alpha = alpha0;
        for (int ls = 0; ls < 12; ++ls) {
            double t[4];
            trial_point(alpha, t);
            s.x[0] = t[0]; s.x[1] = t[1]; s.x[2] = t[2]; s.x[3] = t[3];
            s.Tr = exp(t[0]);
            const double Tn = max_nn(s.Tr, T_MIN);
            if (Tn != s.tp.T) s.tp = eos_tpart<EK>(Tn, q);
            if (evalF<EK, NCH>(s.F, t[0], t[1], t[2], t[3], s.Tr, s.tp, c, q)) {
                double tr[4 - rlo];       // evalF returned finite rows
#pragma unroll
                for (int r = rlo; r < 4; ++r) tr[r - rlo] = fabs(s.F[r] * c.rs[r]);
                const double trial = max_tree<4 - rlo>(tr);
                if (trial < nrm * (1 - 1e-4 * alpha)) return true;
            }
            alpha *= 0.5;
        }
        return false;
    }
    // trials 0..7 on lanes 0..7, then 8..11 on lanes 0..3. G = 16: in the first
    // round lanes 8..15 evaluate the Jacobian's columns at trial 0 (fd_point at
    // that point, T and T-part as newton_jacobian takes them there), kept for the
    // next iteration when trial 0 is the accepted one.
    const int l = glane<G>();
    for (int round = 0; round < 2; ++round) {
        const bool ahead = G >= 16 && round == 0 && l >= 8;
        const int at = ahead ? (l - 8) >> 3 : 0;   // the trial whose Jacobian lane l evaluates
        const int ls = ahead ? at : 8 * round + l;
        const bool act = !ahead && ls < 12;
        // alpha0 halved ls times: one multiplication by 2^-ls, exact while the product
        // is a normal number (alpha0 >= 2^-1000); the halving loop otherwise
        double alpha;
        if (alpha0 >= 0x1p-1000) {
            alpha = alpha0 * __longlong_as_double((long long)(1023 - ls) << 52);
        } else {
            alpha = alpha0;
            for (int j = 0; j < ls; ++j) alpha *= 0.5;
        }
        double t[4];
        trial_point(alpha, t);
        double y[4] = {t[0], t[1], t[2], t[3]}, h = 1.0;
        bool fdact = false;
        int si = 1;
        if (ahead) fd_point<NCH>(t, (l - 8) & 7, y, h, fdact, si);
        const bool fd0 = fdact && si == 0;
        const int slot = G >= 16 ? tcache_slot(s.tc, y[0]) : -1;
        // exp only where a lane evaluates and its memo misses (an idle lane's T is
        // never picked: the source lane below is an evaluating one)
        double Tl = s.Tr;
        if (slot == 0) Tl = s.tc.T[0];
        else if (slot == 1) Tl = s.tc.T[1];
        else if (act || fdact) Tl = exp(y[0]);
        const double Tn = max_nn(Tl, T_MIN);
        TPart tl = s.tp;
        double Fl[4] = {0.0, 0.0, 0.0, 0.0};
        bool accl = false, okl = true;
        if (act || fdact) {
            if (slot >= 0) {
                if (fd0 || Tn != s.tp.T) tl = slot == 0 ? s.tc.tp[0] : s.tc.tp[1];
            } else {
                if (fd0 || Tn != s.tp.T) tl = eos_tpart<EK>(Tn, q);
                if (G >= 16) tcache_put(s.tc, y[0], Tl, tl);    // tl is the T-part at Tn either way
            }
            okl = evalF<EK, NCH>(Fl, y[0], y[1], y[2], y[3], Tl, tl, c, q);
            if (act && okl) {
                double tr[4 - rlo];       // okl: finite rows
#pragma unroll
                for (int r = rlo; r < 4; ++r) tr[r - rlo] = fabs(Fl[r] * c.rs[r]);
                accl = max_tree<4 - rlo>(tr) < nrm * (1 - 1e-4 * alpha);
            }
        }
        const unsigned b = gbits<G>(accl);
        const bool last = round == 1;
        if (b != 0u || last) {
            const int src = b != 0u ? __builtin_ctz(b) : 3;   // trial 11 is lane 3 of round 1
#pragma unroll
            for (int r = 0; r < 4; ++r) { s.x[r] = gshfl<G>(t[r], src); s.F[r] = gshfl<G>(Fl[r], src); }
            s.Tr = gshfl<G>(Tl, src);
            s.tp.T = gshfl<G>(tl.T, src); s.tp.Pl = gshfl<G>(tl.Pl, src); s.tp.el = gshfl<G>(tl.el, src);
            s.tp.K2 = gshfl<G>(tl.K2, src); s.tp.K1 = gshfl<G>(tl.K1, src); s.tp.x = gshfl<G>(tl.x, src);
            if constexpr (G >= 16) {
                fd_column<G>(Fl, h, c, s.jcol);
                s.jbad = gbits<G>(fdact && !okl && at == src) != 0u;
                s.jnext = round == 0 && b != 0u && src < (G - 8) / 8;
                s.jbase = 8 + 8 * src;
            }
            return b != 0u;
        }
    }
    return false;
}

constexpr double NEWTON_TOL = 1e-12;
// max_r |F_r| s_r over the rows of the system (src2d/primrec2d.jl:594-601)
template <bool NCH>
__device__ __forceinline__ double newton_scaled_norm(const NewtonState& s, const CellIn& c) {
    constexpr int rlo = NCH ? 0 : 1;
    double sr[4 - rlo];
#pragma unroll
    for (int r = rlo; r < 4; ++r) sr[r - rlo] = s.F[r] * c.rs[r];
    return max_abs_nan<4 - rlo>(sr);
}

// Newton iterations it .. itcap of _cons_to_prim_attempt_2d! (src2d/primrec2d.jl:591-692)
template <int EK, bool NCH, int G>
__device__ __forceinline__ void newton_iterate(NewtonState& s, const CellIn& c, const EosP& q, int itcap) {
    constexpr int NU = NCH ? 4 : 3;
    constexpr int rlo = NCH ? 0 : 1;
    s.jnext = false;
    for (; s.it <= NEWTON_MAXIT; ++s.it) {
        if (s.it > itcap) return;                    // paused before iteration s.it
        const double nrm = newton_scaled_norm<NCH>(s, c);
        if (nrm < NEWTON_TOL) { s.converged = true; s.reason = PRR_CONVERGED; s.iters = s.it - 1; return; }
        double J[4][4];
        bool ahead = false;
        if constexpr (G >= 16) {
            if (s.jnext) {                  // evaluated by the line search that chose x
                if (s.jbad) { s.reason = PRR_NONFINITE_JACOBIAN; return; }
#pragma unroll
                for (int kk = 0; kk < NU; ++kk)
#pragma unroll
                    for (int r = 0; r < NU; ++r) {
                        const int rr = NCH ? r : r + 1;
                        J[r][kk] = gshfl<G>(s.jcol[rr], s.jbase + 2 * kk);
                    }
                ahead = true;
            }
        }
        if (!ahead && !newton_jacobian<EK, NCH, G>(s, c, q, J)) { s.reason = PRR_NONFINITE_JACOBIAN; return; }
        double Fact[4], delta[4];
#pragma unroll
        for (int r = 0; r < NU; ++r) {
            const int rr = NCH ? r : r + 1;
            Fact[r] = s.F[rr] * c.rs[rr];
            delta[r] = 0.0;
        }
        if (!solve_gauss<NU>(J, Fact, delta)) { s.reason = PRR_LINSOLVE_FAILED; return; }
        // alpha = min(1, min over |delta_k| > 1 of 1/|delta_k|) = 1/max|delta_k| when that
        // exceeds 1 (delta is finite, and 1/x rounds monotonically)
        double ad[NU];
#pragma unroll
        for (int k = 0; k < NU; ++k) ad[k] = fabs(delta[k]);
        const double admax = max_tree<NU>(ad);
        const double alpha = admax > 1.0 ? 1.0 / admax : 1.0;
        if (!newton_linesearch<EK, NCH, G>(s, c, q, delta, alpha, nrm)) {
            s.reason = PRR_LINESEARCH_FAILED;
            return;
        }
        if (s.it == NEWTON_MAXIT) s.iters = NEWTON_MAXIT;
    }
}

// tail of _cons_to_prim_attempt_2d! (src2d/primrec2d.jl:694-742): final state,
// residual acceptance, staged phi fallback. BISECT = false: returns with
// reason PRR_UNSET where the staged fallback would start, leaving it to a caller
// that resumes the tail with BISECT = true from the same iterate.
template <int EK, bool NCH, int G, bool BISECT = true>
__device__ __forceinline__ PrimRes newton_finish(NewtonState& s, const CellIn& c, const EosP& q) {
    constexpr int rlo = NCH ? 0 : 1;
    constexpr double tol = 1e-12;
    PrimRes R;
    double T = jclamp(s.Tr, T_MIN, T_SOLVE_MAX);
    double mu = NCH ? (q.m + T * s.x[1]) : q.m;
    double ux = s.x[2], uy = s.x[3];
    TPart tf = (jmax(T, T_MIN) == s.tp.T) ? s.tp : eos_tpart<EK>(jmax(T, T_MIN), q);
    double P, n, e;
    eos_assemble<EK>(tf, mu, q, &P, &n, &e);
    // residual at the final iterate (tp belongs to x[0])
    double rows[4] = {0, 0, 0, 0};
    double resnorm = 0.0;
    evalF<EK, NCH>(s.F, s.x[0], s.x[1], s.x[2], s.x[3], s.Tr, s.tp, c, q);
#pragma unroll
    for (int r = rlo; r < 4; ++r) { rows[r] = fabs(s.F[r] * c.rs[r]); resnorm = jmax(resnorm, rows[r]); }
    const int iters = s.iters;
    if (!(isf(resnorm) && resnorm < tol)) {
        double hyd = jmax(jmax(rows[1], rows[2]), rows[3]);
        if (NCH && isf(hyd) && hyd < tol) {
            if (!BISECT) { R.reason = PRR_UNSET; R.ok = 0; R.iters = iters; return R; }
            double phib;
            bool okb = solve_phi_bisect<EK, G>(c.D, ux, uy, c.d.nux, c.d.nuy, tf, q, &phib);
            double ph = jclamp(phib, -PHI_CAP, PHI_CAP);
            mu = q.m + T * ph;
            eos_assemble<EK>(tf, mu, q, &P, &n, &e);
            R = {T, mu, ux, uy, jmax(n, 0.0), jmax(e, 0.0), P, 1, okb ? PRR_CONVERGED : PRR_RESIDUAL_TOO_LARGE, iters};
            R.staged = 1;
            return R;
        }
        R = {T, mu, ux, uy, n, jmax(e, 0.0), P, 0, PRR_RESIDUAL_TOO_LARGE, iters};
        return R;
    }
    R = {T, mu, ux, uy, n, jmax(e, 0.0), P, 1, PRR_CONVERGED, iters};
    return R;
}

// _cons_to_prim_attempt_2d! (src2d/primrec2d.jl:511-742). NCH: eos_has_charge_2d.
template <int EK, bool NCH, bool SEED = true, int G = 0>
__device__ __forceinline__ PrimRes cons_to_prim_attempt(const CellIn& c, double yT0, double phi0, double ux0,
                                                        double uy0, const EosP& q) {
    NewtonState s;
    PrimRes R;
    if (newton_begin<EK, NCH, SEED, G>(c, yT0, phi0, ux0, uy0, q, s, R)) return R;
    newton_iterate<EK, NCH, G>(s, c, q, NEWTON_MAXIT);
    return newton_finish<EK, NCH, G>(s, c, q);
}

// cons_to_prim_2d! (src2d/primrec2d.jl:465-486): one fresh-seed retry. Written
// as a two-pass loop so the Newton body is instantiated once.
template <int EK, bool NCH>
__device__ __forceinline__ PrimRes cons_to_prim(const CellIn& c, double yT0, double phi0, double ux0, double uy0,
                                                const EosP& q) {
    PrimRes r;
    int its = 0;
    const bool warm = isf(yT0) && isf(ux0) && isf(uy0);
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        r = cons_to_prim_attempt<EK, NCH, true>(c, yT0, phi0, ux0, uy0, q);
        its += r.iters;
        if (r.ok || r.reason == PRR_VACUUM || !warm) break;
        yT0 = nan; phi0 = nan; ux0 = nan; uy0 = nan;
    }
    r.iters = its;
    return r;
}

}  // namespace fd
