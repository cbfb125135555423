// =============================================================================
// fivo_api.hip -- the C ABI (include/fivo2d.h): context, state/work transfer and
// the device-resident driver loop that replaces run_sim_2d! (main2D.jl:371-561)
// and the steppers (src2d/timestepper2d.jl:74-188).
//
// Control flow stays on the host; every array stays in HBM. One small pinned
// read-back per RK stage carries the stage's inadmissible-cell count and the
// RHS's max signal speed (the only data-dependent decisions: accept, MOOD
// retry, dt halving, next dt).
// =============================================================================
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>
#include "fivo_ctx.h"

using namespace fivo;

namespace {

thread_local std::string g_err;

int fail(Ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg; else g_err = msg;
    return code;
}
#define HIPCHK(c, x)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail((c), FIVO_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr double kPI2 = 3.141592653589793 * 3.141592653589793;
constexpr double kFMGEV3 = (1.0 / 0.1973269804) * (1.0 / 0.1973269804) * (1.0 / 0.1973269804);
constexpr double kLK0 = 0.0005624486560000001, kLK1 = 0.003652264, kLK2 = 0.023716, kLK3 = 0.154,
                 kLK4 = 5.208957878352717;

// EOS constants with the same arithmetic as the reference (src/eos.jl)
fd::EosP make_eos(const fivo_eos_desc& e) {
    fd::EosP q;
    std::memset(&q, 0, sizeof(q));
    q.kind = e.kind;
    q.hq_fm = e.hq_times_fmGeV3 != 0;
    q.m = e.m_hq;
    q.ghq = e.g_hq;
    q.fmGeV3 = kFMGEV3;
    q.aSB = (kPI2 / 90) * e.g_eff;
    q.Aconf = e.g_hq * (e.m_hq * e.m_hq) / (2 * kPI2);
    q.Alhrg = e.canon_factor * e.g_hq * (e.m_hq * e.m_hq) / (2 * kPI2);
    q.conf_charm = (e.kind == FIVO_EOS_CONFORMAL_HQ && e.g_hq > 0.0) ? 1 : 0;
    q.has_charge = e.kind == FIVO_EOS_LATTICE_HRG ? 1 : (e.g_hq > 0.0 ? 1 : 0);
    q.LA = 0.01 * (e.d * e.d);
    q.LB = 0.1 * (e.c * e.c);
    q.nA0 = kLK0 * e.a4; q.nA1 = kLK1 * e.a3; q.nA2 = kLK2 * e.a2; q.nA3 = kLK3 * e.a1; q.nA4 = kLK4;
    q.dB4 = kLK0 * e.b4; q.dB3 = kLK1 * e.b3; q.dB2 = kLK2 * e.b2; q.dB1 = kLK3 * e.b1;
    q.nP1 = kLK1 * e.a3; q.nP2 = 2 * kLK2 * e.a2; q.nP3 = 3 * kLK3 * e.a1; q.nP4 = 4 * kLK4;
    q.dP1 = kLK1 * e.b3; q.dP2 = 2 * kLK2 * e.b2; q.dP3 = 3 * kLK3 * e.b1;
    q.nQ2 = 2 * kLK2 * e.a2; q.nQ3 = 6 * kLK3 * e.a1; q.nQ4 = 12 * kLK4;
    q.dQ2 = 2 * kLK2 * e.b2; q.dQ3 = 6 * kLK3 * e.b1;
    return q;
}

// e(T_vac_cut, mu = 0) on the host, for the derived E_vac_cut (main2D.jl:100-105):
// eos_Pne's energy density. Lattice-HRG: light sector only (src/eos.jl:481-519).
double host_eos_e(const fivo_eos_desc& e, double T);

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Kernel timing with HIP events on the context stream. Event pairs are recorded
// around each launch and resolved at the next stream synchronisation, so timing
// adds no synchronisation of its own.
struct Timer {
    Ctx& c;
    int cat, i0 = -1;
    Timer(Ctx& c_, int cat_) : c(c_), cat(cat_) {
        if (!c.timing) return;
        i0 = c.ev_used;
        if (c.ev_used + 2 > (int)c.evpool.size()) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a); (void)hipEventCreate(&b);
            c.evpool.push_back(a); c.evpool.push_back(b);
        }
        c.ev_used += 2;
        (void)hipEventRecord(c.evpool[i0], c.stream);
    }
    ~Timer() {
        if (i0 < 0) return;
        (void)hipEventRecord(c.evpool[i0 + 1], c.stream);
        c.ev_cat.push_back(cat);
    }
};
// resolve recorded event pairs; call after a stream synchronisation
void resolve_timing(Ctx& c) {
    if (!c.timing) return;
    for (size_t k = 0; k < c.ev_cat.size(); ++k) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c.evpool[2 * k], c.evpool[2 * k + 1]);
        c.t_ms[c.ev_cat[k]] += ms; c.t_cnt[c.ev_cat[k]] += 1;
        c.t_ms[TK_ALL] += ms; c.t_cnt[TK_ALL] += 1;
    }
    c.ev_cat.clear();
    c.ev_used = 0;
}

// zero the device scalars first..last (fields of DevScalars, with the lines between)
hipError_t clear_scalars(Ctx& c, unsigned long long* first, unsigned long long* last) {
    return hipMemsetAsync(first, 0, (size_t)((char*)(last + 1) - (char*)first), c.stream);
}
int reset_scalars(Ctx& c) {
    // amax, nbad, nprimfail, nvacuum, npending, ncold, nresq, nwarm -> 0 (newton_iters kept)
    HIPCHK(&c, clear_scalars(c, &c.sc->amax_bits, &c.sc->nwarm));
    return FIVO_OK;
}
int read_scalars(Ctx& c) {
    HIPCHK(&c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    resolve_timing(c);
    c.amax_host = __builtin_bit_cast(double, (unsigned long long)c.sc_host->amax_bits);
    c.last_primfail = (int64_t)c.sc_host->nprimfail;
    c.last_vacuum = (int64_t)c.sc_host->nvacuum;
    return FIVO_OK;
}

int do_bc(Ctx& c, double* U, int bc) {
    Timer t(c, TK_BC);
    HIPCHK(&c, launch_bc(c, U, bc));
    return FIVO_OK;
}
int do_prim(Ctx& c, const double* U, double tau) {
    {
        Timer t(c, TK_PRIM);
        HIPCHK(&c, launch_prim_warm(c, U, tau));
    }
    if (c.learned) {
        Timer t(c, TK_SEED);
        HIPCHK(&c, launch_prim_seed(c, U, tau));
    }
    Timer t(c, TK_COOP);
    HIPCHK(&c, launch_prim_coop(c, U, tau));
    return FIVO_OK;
}
int do_faces(Ctx& c, const FaceLaunch& f) {
    Timer t(c, TK_FACES);
    HIPCHK(&c, launch_faces(c, f));
    return FIVO_OK;
}
// The primitive pass and the fused faces of one RHS. The face tiles that read no
// cell queued for the cooperative kernel do not wait for it: they run on a second
// stream beside it, the rest after it. Per-kernel timing serialises the two.
int do_prim_faces(Ctx& c, const double* U, double tau, const FaceLaunch& f) {
    if (c.timing) {
        const int rc = do_prim(c, U, tau);
        return rc < 0 ? rc : do_faces(c, f);
    }
    HIPCHK(&c, launch_prim_warm(c, U, tau));
    HIPCHK(&c, launch_mark_tiles(c));
    HIPCHK(&c, hipEventRecord(c.ev_fork, c.stream));
    HIPCHK(&c, hipStreamWaitEvent(c.stream2, c.ev_fork, 0));
    HIPCHK(&c, launch_faces(c, f, 1, c.stream2));
    HIPCHK(&c, hipEventRecord(c.ev_join, c.stream2));
    HIPCHK(&c, launch_prim_seed(c, U, tau));
    HIPCHK(&c, launch_prim_coop(c, U, tau));
    HIPCHK(&c, hipStreamWaitEvent(c.stream, c.ev_join, 0));
    HIPCHK(&c, launch_faces(c, f, 2, c.stream));
    return FIVO_OK;
}

// rhs_2d! (src2d/rhs2d.jl:169-249) into c.dU
int rhs_dev(Ctx& c, double tau, int bc, const uint8_t* force_dev) {
    double* U = c.U[c.cur];
    int rc;
    if ((rc = reset_scalars(c)) < 0) return rc;
    if ((rc = do_bc(c, U, bc)) < 0) return rc;
    HIPCHK(&c, hipMemsetAsync(c.dU, 0, sizeof(double) * (size_t)c.p.nv * c.p.pitch, c.stream));
    FaceLaunch f{};
    f.Ub = U; f.Uold = U; f.Uout = nullptr; f.dUout = c.dU; f.mode = 0; f.tau = tau; f.force = force_dev;
    if ((rc = do_prim_faces(c, U, tau, f)) < 0) return rc;
    c.nrhs += 1;
    return read_scalars(c);
}

// _stage_2d! (src2d/timestepper2d.jl:113-140) on device buffers: 1 = admissible
int stage(Ctx& c, double* Ub, const double* Uold, double* Uout, double tau, double tau_state, double D,
          int combine, double wold, double wnew, int bc, int max_mood, bool bc_first) {
    const uint8_t* mask = nullptr;
    int rc;
    for (int attempt = 0; attempt <= max_mood; ++attempt) {
        if ((rc = reset_scalars(c)) < 0) return rc;
        if (bc_first && attempt == 0 && (rc = do_bc(c, Ub, bc)) < 0) return rc;
        FaceLaunch f{};
        f.Ub = Ub; f.Uold = Uold; f.Uout = Uout; f.dUout = nullptr; f.mode = 1; f.combine = combine;
        f.tau = tau; f.tau_state = tau_state; f.D = D; f.wold = wold; f.wnew = wnew; f.force = mask;
        if ((rc = do_prim_faces(c, Ub, tau, f)) < 0) return rc;
        if ((rc = do_bc(c, Uout, bc)) < 0) return rc;
        c.nrhs += 1;
        if ((rc = read_scalars(c)) < 0) return rc;
        if (c.sc_host->npending) {
            HIPCHK(&c, launch_admissible_pending(c, Uout, tau_state));
            if ((rc = read_scalars(c)) < 0) return rc;
        }
        if (c.sc_host->nbad == 0) return 1;
        if (attempt == max_mood) return 0;
        c.nmood += 1;
        // mark_bad_2d! is the kernel's bad mask; expand_bad_2d! radius 1 + attempt
        const uint8_t* src = c.w.bad;
        for (int r = 0; r < 1 + attempt; ++r) {
            uint8_t* dst = (r % 2 == 0) ? c.w.force : c.w.ftmp;
            HIPCHK(&c, launch_expand(c, src, dst));
            src = dst;
        }
        mask = src;
    }
    return 0;
}

// One RK stage's launches up to its face update, without reading anything back
// (the first stage of an attempt also clears the per-pass scalars and fills the
// ghosts of its input; k_stage_end does both for the stages after it).
int stage_launch(Ctx& c, double* Ub, const double* Uold, double* Uout, double tau, double tau_state, double D,
                 int combine, double wold, double wnew, int bc, bool first) {
    int rc;
    if (first && (rc = reset_scalars(c)) < 0) return rc;
    if (first && (rc = do_bc(c, Ub, bc)) < 0) return rc;
    FaceLaunch f{};
    f.Ub = Ub; f.Uold = Uold; f.Uout = Uout; f.dUout = nullptr; f.mode = 1; f.combine = combine;
    f.tau = tau; f.tau_state = tau_state; f.D = D; f.wold = wold; f.wnew = wnew; f.force = nullptr;
    if ((rc = do_prim_faces(c, Ub, tau, f)) < 0) return rc;
    c.nrhs += 1;
    return FIVO_OK;
}

// read_scalars' host state from a copy of the device scalars
void take_scalars(Ctx& c, const DevScalars& s) {
    *c.sc_host = s;
    c.amax_host = __builtin_bit_cast(double, (unsigned long long)s.amax_bits);
    c.last_primfail = (int64_t)s.nprimfail;
    c.last_vacuum = (int64_t)s.nvacuum;
}

// Every stage of one SSPRK attempt (no MOOD retries) launched back to back, each
// ending in k_stage_end (ghost fill, pending test, the device scalars to a slot,
// the warm starts to a rollback slot), the slots read back in one copy, one
// synchronisation for the attempt instead of one per stage. *failed = the first stage whose output is
// inadmissible (1-based), 0 if none. After a failure the later stages' changes that
// the retry would see are rolled back -- the Newton warm starts (copied after each
// stage but the last) and the device counters (from the failed stage's slot) -- so
// the halved retry starts from exactly what the stage-by-stage loop leaves.
int step_all_stages(Ctx& c, int integ, double tau, double D, int bc, int* failed) {
    const int ns = integ == FIVO_SSPRK3 ? 3 : 2;
    double* U = c.U[c.cur];
    double* S1 = c.U[(c.cur + 1) % 4];
    double* S2 = c.U[(c.cur + 2) % 4];
    double* S3 = c.U[(c.cur + 3) % 4];
    const int64_t nrhs0 = c.nrhs;
    const size_t wsn = 4 * (size_t)c.p.pitch;   // x0_yT, x0_phi, x0_ux, x0_uy: consecutive pitch blocks
    for (int k = 0; k < ns; ++k) {
        int rc;
        double* out = k == 0 ? S1 : (k == 1 ? S2 : S3);
        if (k == 0) rc = stage_launch(c, U, U, S1, tau, tau + D, D, 0, 1.0, 0.0, bc, true);
        else if (integ == FIVO_SSPRK3 && k == 1) rc = stage_launch(c, S1, U, S2, tau + D, tau + D, D, 1, 0.75, 0.25, bc, false);
        else if (integ == FIVO_SSPRK3) rc = stage_launch(c, S2, U, S3, tau + 0.5 * D, tau + D, D, 2, 0.0, 0.0, bc, false);
        else rc = stage_launch(c, S1, U, S2, tau + D, tau + D, D, 1, 0.5, 0.5, bc, false);
        if (rc < 0) return rc;
        Timer t(c, TK_TAIL);
        HIPCHK(&c, launch_stage_end(c, out, tau + D, bc, k, k + 1 < ns ? c.ws + k * wsn : nullptr, true));
    }
    HIPCHK(&c, hipMemcpyAsync(c.sc_stage_host, c.sc_slots, ns * sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    resolve_timing(c);
    *failed = 0;
    for (int k = 0; k < ns; ++k)
        if (c.sc_stage_host[k].nbad != 0) { *failed = k + 1; break; }
    const int last = *failed ? *failed - 1 : ns - 1;
    if (*failed && *failed < ns) {
        HIPCHK(&c, hipMemcpyAsync(c.w.x0_yT, c.ws + last * wsn, sizeof(double) * wsn, hipMemcpyDeviceToDevice,
                                  c.stream));
        HIPCHK(&c, hipMemcpyAsync(c.sc, &c.sc_stage_host[last], sizeof(DevScalars), hipMemcpyHostToDevice, c.stream));
        c.nrhs = nrhs0 + *failed;
    }
    take_scalars(c, c.sc_stage_host[last]);
    c.pass_clean = *failed == 0;
    return FIVO_OK;
}

// step_ssprk2_2d! / step_ssprk3_2d! (src2d/timestepper2d.jl:74-188). Without MOOD
// retries every attempt but the last launches its stages back to back
// (step_all_stages); the last attempt, and every attempt with MOOD, checks each
// stage before the next, so a step that fails leaves the state of the reference's
// loop behind.
int step_dev(Ctx& c, int integ, double tau, double dtau, int bc, int max_halve, int max_mood, int* ok,
             double* used) {
    double D = dtau;
    int rc;
    const bool batch = integ == FIVO_SSPRK3 || max_mood == 0;
    c.pass_clean = false;
    for (int h = 0; h <= max_halve; ++h) {
        if (batch && h < max_halve) {
            int failed = 0;
            if ((rc = step_all_stages(c, integ, tau, D, bc, &failed)) < 0) return rc;
            if (!failed) {
                c.cur = (c.cur + (integ == FIVO_SSPRK3 ? 3 : 2)) % 4;
                *ok = 1; *used = D;
                return FIVO_OK;
            }
            D *= 0.5;
            c.nhalvings += 1;
            continue;
        }
        double* U = c.U[c.cur];
        double* S1 = c.U[(c.cur + 1) % 4];
        double* S2 = c.U[(c.cur + 2) % 4];
        double* S3 = c.U[(c.cur + 3) % 4];
        if (integ == FIVO_SSPRK3) {
            rc = stage(c, U, U, S1, tau, tau + D, D, 0, 1.0, 0.0, bc, 0, true);
            if (rc < 0) return rc;
            if (rc == 1) {
                rc = stage(c, S1, U, S2, tau + D, tau + D, D, 1, 0.75, 0.25, bc, 0, false);
                if (rc < 0) return rc;
                if (rc == 1) {
                    rc = stage(c, S2, U, S3, tau + 0.5 * D, tau + D, D, 2, 0.0, 0.0, bc, 0, false);
                    if (rc < 0) return rc;
                    if (rc == 1) { c.cur = (c.cur + 3) % 4; *ok = 1; *used = D; return FIVO_OK; }
                }
            }
        } else {
            rc = stage(c, U, U, S1, tau, tau + D, D, 0, 1.0, 0.0, bc, max_mood, true);
            if (rc < 0) return rc;
            if (rc == 1) {
                rc = stage(c, S1, U, S2, tau + D, tau + D, D, 1, 0.5, 0.5, bc, max_mood, false);
                if (rc < 0) return rc;
                if (rc == 1) { c.cur = (c.cur + 2) % 4; *ok = 1; *used = D; return FIVO_OK; }
            }
        }
        D *= 0.5;
        c.nhalvings += 1;
    }
    *ok = 0;
    *used = D;
    return FIVO_STEP_FAILED;
}

int relax_dev(Ctx& c, double tau, double dtau) {
    if (!(c.md.enable_shear || c.md.enable_bulk || c.md.enable_diff)) return FIVO_OK;
    Timer t(c, TK_RELAX);
    HIPCHK(&c, launch_relax(c, c.U[c.cur], tau, dtau));
    return FIVO_OK;
}

int reset_history_dev(Ctx& c) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    double* arrs[] = {c.w.x0_yT, c.w.x0_phi, c.w.x0_ux, c.w.x0_uy, c.w.ux_prev, c.w.uy_prev,
                      c.w.alpha_prev, c.w.T_prev, c.w.nux_prev, c.w.nuy_prev};
    for (double* a : arrs) HIPCHK(&c, launch_fill(c, a, nan, c.p.pitch));
    return FIVO_OK;
}

// host-side sums of the per-block diagnostics partials, in block order
int diag(Ctx& c, int what, double* q, double* maxu, double* minp) {
    HIPCHK(&c, launch_reduce_diag(c, c.U[c.cur], what));
    std::vector<double> h(3 * (size_t)c.npartials);
    HIPCHK(&c, hipMemcpyAsync(h.data(), c.partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    double s = 0.0, mu = 0.0, mp = std::numeric_limits<double>::infinity();
    for (int b = 0; b < c.npartials; ++b) {
        s += h[3 * b];
        mu = h[3 * b + 1] > mu ? h[3 * b + 1] : mu;
        mp = h[3 * b + 2] < mp ? h[3 * b + 2] : mp;
    }
    *q = s;
    if (maxu) *maxu = mu;
    if (minp) *minp = mp;
    return FIVO_OK;
}

int project_shear_dev(Ctx& c, double* mx, double* mean) {
    *mx = 0.0; *mean = 0.0;
    if (!c.md.enable_shear) return FIVO_OK;
    HIPCHK(&c, launch_project_shear(c, c.U[c.cur], c.md.shear_constraint_project ? 1 : 0));
    std::vector<double> h(2 * (size_t)c.npartials);
    HIPCHK(&c, hipMemcpyAsync(h.data(), c.partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    double m = 0.0, s = 0.0;
    for (int b = 0; b < c.npartials; ++b) { m = h[2 * b] > m ? h[2 * b] : m; s += h[2 * b + 1]; }
    *mx = m;
    *mean = s / (double)((int64_t)c.p.Nx * c.p.Ny);
    return FIVO_OK;
}

}  // namespace

// ---- host EOS energy density for the derived vacuum cut ------------------------------
namespace {
double host_besselkx(int nu, double x);
double host_eos_e(const fivo_eos_desc& e, double T) {
    T = T > 1e-20 ? T : 1e-20;
    if (e.kind == FIVO_EOS_CONFORMAL_HQ) {
        double Pl = (kPI2 / 90) * e.g_eff * ((T * T) * (T * T)) * kFMGEV3;
        double el = 3.0 * Pl, m = e.m_hq, x = m / T;
        if (!std::isfinite(x) || x <= 0.0) return el;
        double A = e.g_hq * (m * m) / (2 * kPI2);
        double K2 = host_besselkx(2, x), K1 = host_besselkx(1, x);
        double Ef = std::exp((0.0 - m) / T);
        double eh = A * (T * T) * Ef * (3.0 * K2 + x * K1);
        if (e.hq_times_fmGeV3) eh *= kFMGEV3;
        return el + eh;
    }
    // light_e = -P + T dP/dT (src/eos.jl:463-467)
    double A = 0.01 * (e.d * e.d), B = 0.1 * (e.c * e.c);
    double Ex = -(A / (T * T)) - (B / T);
    Ex = Ex < -700.0 ? -700.0 : (Ex > 700.0 ? 700.0 : Ex);
    double expE = std::exp(Ex), T2 = T * T, T3 = T * T * T, T4 = T2 * T2;
    double N = kLK0 * e.a4 + kLK1 * e.a3 * T + kLK2 * e.a2 * T2 + kLK3 * e.a1 * T3 + kLK4 * T4;
    double D = kLK0 * e.b4 + kLK1 * e.b3 * T + kLK2 * e.b2 * T2 + kLK3 * e.b1 * T3 + T4;
    double P = kFMGEV3 * expE * T4 * (N / D);
    double Ep = (2 * A / T3) + (B / T2);
    double Np = kLK1 * e.a3 + 2 * kLK2 * e.a2 * T + 3 * kLK3 * e.a1 * T2 + 4 * kLK4 * T3;
    double Dp = kLK1 * e.b3 + 2 * kLK2 * e.b2 * T + 3 * kLK3 * e.b1 * T2 + 4 * T3;
    double G = T4 * N, H = D, Gp = 4 * T3 * N + T4 * Np, Hp = Dp;
    double F = G / H, Fp = (Gp * H - G * Hp) / (H * H);
    double dP = kFMGEV3 * expE * (Fp + F * Ep);
    return (-P + T * dP) + 0.0;
}
// scaled Bessel K on the host (A&S 9.8.5-9.8.8 polynomial forms are not used:
// the same Boost rational table as the device, for one call at create time)
double host_k01(double x, int which) {
    auto ev = [](double z, std::initializer_list<double> c) {
        double r = 0; const double* p = c.begin(); int n = (int)c.size();
        r = p[n - 1];
        for (int i = n - 2; i >= 0; --i) r = std::fma(z, r, p[i]);
        return r;
    };
    if (x <= 1.0) {
        double z = x * x, a = z / 4;
        if (which == 0) {
            double s = std::fma(ev(a, {-1.372509002685546267e-1, 2.574916117833312855e-1, 1.395474602146869316e-2,
                                      5.445476986653926759e-4, 7.125159422136622118e-6}),
                                1.0 / ev(a, {1.0, -5.458333438017788530e-02, 1.291052816975251298e-03,
                                             -1.367653946978586591e-05}), 1.137250900268554688);
            double a0 = std::fma(s, a, 1.0);
            return std::fma(-a0, std::log(x), ev(z, {1.159315156584124484e-01, 2.789828789146031732e-01,
                                                    2.524892993216121934e-02, 8.460350907213637784e-04,
                                                    1.491471924309617534e-05, 1.627106892422088488e-07,
                                                    1.208266102392756055e-09, 6.611686391749704310e-12})) * std::exp(x);
        }
        double pq = std::fma(ev(a, {-3.62137953440350228e-3, 7.11842087490330300e-3, 1.00302560256614306e-5,
                                   1.77231085381040811e-6}),
                             1.0 / ev(a, {1.0, -4.80414794429043831e-2, 9.85972641934416525e-4,
                                          -8.91196859397070326e-6}), 8.69547128677368164e-2);
        pq = std::fma(pq * a, a, (a / 2 + 1));
        double a1 = pq * x / 2;
        double pq2 = std::fma(ev(z, {-3.07965757829206184e-1, -7.80929703673074907e-02, -2.70619343754051620e-3,
                                     -2.49549522229072008e-5}) /
                                  ev(z, {1.0, -2.36316836412163098e-2, 2.64524577525962719e-4, -1.49749618004162787e-6}),
                              x, 1.0 / x);
        return std::fma(a1, std::log(x), pq2) * std::exp(x);
    }
    double ix = 1.0 / x;
    if (which == 0)
        return std::fma(ev(ix, {2.533141373155002416e-1, 3.628342133984595192e0, 1.868441889406606057e1,
                                4.306243981063412784e1, 4.424116209627428189e1, 1.562095339356220468e1,
                                -1.810138978229410898e0, -1.414237994269995877e0, -9.369168119754924625e-2}),
                        1.0 / ev(ix, {1.0, 1.494194694879908328e1, 8.265296455388554217e1, 2.162779506621866970e2,
                                      2.845145155184222157e2, 1.851714491916334995e2, 5.486540717439723515e1,
                                      6.118075837628957015e0, 1.586261269326235053e-1}), 1.0) / std::sqrt(x);
    return std::fma(ev(ix, {-1.97028041029226295e-1, -2.32408961548087617e0, -7.98269784507699938e0,
                            -2.39968410774221632e0, 3.28314043780858713e1, 5.67713761158496058e1,
                            3.30907788466509823e1, 6.62582288933739787e0, 3.08851840645286691e-1}),
                    1.0 / ev(ix, {1.0, 1.41811409298826118e1, 7.35979466317556420e1, 1.77821793937080859e2,
                                  2.11014501598705982e2, 1.19425262951064454e2, 2.88448064302447607e1,
                                  2.27912927104139732e0, 2.50358186953478678e-2}), 1.45034217834472656) / std::sqrt(x);
}
double host_besselkx(int nu, double x) {
    if (!(std::isfinite(x) && x > 0.0)) return 0.0;
    if (x >= 1.0e4) {
        double mu = 4.0 * double(nu) * double(nu), ix = 1.0 / (8.0 * x);
        double s = 1.0 + (mu - 1.0) * ix * (1.0 + (mu - 9.0) * ix / 2.0 * (1.0 + (mu - 25.0) * ix / 3.0));
        return std::sqrt(3.141592653589793 / (2.0 * x)) * s;
    }
    double km = host_k01(x, 0);
    if (nu == 0) return km;
    double k = host_k01(x, 1);
    for (int n = 1; n < nu; ++n) { double kp = km + (2.0 * n / x) * k; km = k; k = kp; }
    return k;
}
}  // namespace

// =============================================================================
namespace fivo {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace fivo

extern "C" {

const char* fivo_last_error(const fivo_ctx* ctx) {
    return ctx ? reinterpret_cast<const Ctx*>(ctx)->err.c_str() : g_err.c_str();
}

int fivo_create(const fivo_grid_desc* g, const fivo_model_desc* m, fivo_ctx** out) {
    if (!out) return fail(nullptr, FIVO_ERR_ARG, "null out");
    *out = nullptr;
    if (!g || !m) return fail(nullptr, FIVO_ERR_ARG, "null descriptor");
    if (g->nx <= 8 || g->ny <= 8 || g->nghost != 3 || !(g->xmax > 0) || !(g->ymax > 0))
        return fail(nullptr, FIVO_ERR_ARG, "grid: need nx, ny > 8, nghost = 3, xmax, ymax > 0");
    if (m->eos.kind != FIVO_EOS_CONFORMAL_HQ && m->eos.kind != FIVO_EOS_LATTICE_HRG)
        return fail(nullptr, FIVO_ERR_ARG, "unknown EOS kind");
    // build_model_2d (main2D.jl:85-123): one transport per dissipative field
    if ((m->advect_nu && m->relax_advect_nu) || (m->advect_Pi && m->relax_advect_Pi) ||
        (m->advect_pi && m->relax_advect_pi))
        return fail(nullptr, FIVO_ERR_ARG, "Pick ONE for transport: advect_* OR relax_advect_* (not both)");
    if (m->consistent_fm && m->tauN_coeff != 1.0)
        return fail(nullptr, FIVO_ERR_ARG, "tauN_coeff cannot be combined with consistent_fm = true");
    if (m->consistent_m2 && !m->enable_diff)
        return fail(nullptr, FIVO_ERR_ARG, "consistent_m2 requires enable_diff");
    if (m->enable_diff && !m->with_charge)
        return fail(nullptr, FIVO_ERR_ARG, "enable_diff requires with_charge");
    {
        // check_terms (src/terms.jl:272-283): a switch changed from its default in
        // a disabled sector would act on nothing
        static const char* const names[18] = {
            "nu_gradalpha", "fm_gradT", "fm_inertial", "fm_nu_gradu", "fm_expansion", "fm_dlnh",
            "m2_nu_gradient", "m2_bg_gradu", "m2_bg_DlnT", "m2_bg_Dalpha", "m2_expansion", "m2_pi_sigma",
            "m2_PiQ_sigma", "m2_vorticity", "m2_projector", "m2_accel_nu", "m2_nu_gradTh", "shear_vorticity"};
        const unsigned changed = (m->terms ^ FIVO_TERMS_DEFAULT) & 0x3FFFFu;
        if (m->terms & ~0x3FFFFu) return fail(nullptr, FIVO_ERR_ARG, "terms: unknown bits set");
        for (int k = 0; k < 18; ++k) {
            if (!((changed >> k) & 1u)) continue;
            const bool on = k == 0 ? m->enable_diff != 0
                          : k <= 5 ? m->consistent_fm != 0
                          : k <= 16 ? m->consistent_m2 != 0
                                    : m->enable_shear != 0;
            if (on) continue;
            const char* sector = k == 0 ? "enable_diff" : k <= 5 ? "consistent_fm" : k <= 16 ? "consistent_m2"
                                                                                        : "enable_shear";
            return fail(nullptr, FIVO_ERR_ARG,
                        std::string("terms: `") + names[k] + "` changes a term of the `" + sector +
                            "` sector, but `" + sector + " = false`, so the switch would act on nothing");
        }
    }
    Ctx* c = new Ctx();
    c->gd = *g;
    c->md = *m;
    Params& p = c->p;
    std::memset(&p, 0, sizeof(p));
    p.Nx = g->nx; p.Ny = g->ny; p.ng = g->nghost;
    p.Nxtot = p.Nx + 2 * p.ng; p.Nytot = p.Ny + 2 * p.ng; p.Ntot = p.Nxtot * p.Nytot;
    p.pitch = (p.Ntot + 63) / 64 * 64;
    p.dx = (g->xmax - (-g->xmax)) / p.Nx; p.dy = (g->ymax - (-g->ymax)) / p.Ny;
    p.xmin = -g->xmax; p.ymin = -g->ymax;
    int nv = 4;
    p.iNux = p.iNuy = p.iPi = p.iPixx = p.iPixy = p.iPiyy = p.iPieta = p.iPQ0 = -1;
    p.hasNu = m->with_charge != 0; p.hasPi = m->enable_bulk != 0; p.hasShear = m->enable_shear != 0;
    p.hasM2 = m->consistent_m2 != 0;
    if (p.hasNu) { p.iNux = nv++; p.iNuy = nv++; }
    if (p.hasPi) p.iPi = nv++;
    if (p.hasShear) { p.iPixx = nv++; p.iPixy = nv++; p.iPiyy = nv++; p.iPieta = nv++; }
    if (p.hasM2) { p.iPQ0 = nv; nv += 5; }
    p.nv = nv;
    p.eos = make_eos(m->eos);
    double Evac = m->E_vac_cut;
    if (Evac <= 0.0) Evac = host_eos_e(m->eos, m->T_vac_cut);
    p.Evac = Evac;
    p.has_rdom = std::isfinite(m->r_domain) ? 1 : 0;
    p.rdom2 = p.has_rdom ? m->r_domain * m->r_domain : 0.0;
    p.enable_shear = m->enable_shear; p.enable_bulk = m->enable_bulk; p.enable_diff = m->enable_diff;
    p.consistent_fm = m->consistent_fm; p.consistent_m2 = m->consistent_m2;
    p.relax_advect_pi = m->relax_advect_pi; p.relax_advect_Pi = m->relax_advect_Pi;
    p.relax_advect_nu = m->relax_advect_nu; p.relax_advect_m2 = m->relax_advect_m2;
    p.advect_nu = p.hasNu && m->advect_nu; p.advect_Pi = p.hasPi && m->advect_Pi;
    p.advect_pi = p.hasShear && m->advect_pi;
    p.shear_projected_deriv = m->shear_projected_deriv; p.diff_projected_deriv = m->diff_projected_deriv;
    p.vacuum_ramp_relax = m->vacuum_ramp_relax; p.terms = m->terms;
    p.eta_over_s = m->eta_over_s; p.Cs = m->tauShear_coeff; p.deltaShear_factor = m->deltaShear_factor;
    p.taupi_pi_factor = m->taupi_pi_factor; p.pi_clip_factor = m->pi_clip_factor;
    p.zeta_over_s = m->zeta_over_s; p.Czeta = m->tauPi_coeff; p.deltaPi_factor = m->deltaPi_factor;
    p.lambda_Pi_pi_factor = m->lambda_Pi_pi_factor; p.lambda_pi_Pi_factor = m->lambda_pi_Pi_factor;
    p.Pi_clip_factor = m->Pi_clip_factor;
    p.kappa_coeff = m->kappa_coeff; p.transport_mass = m->transport_mass; p.tauN_coeff = m->tauN_coeff;
    p.deltaN_factor = m->deltaN_factor; p.nu_clip_factor = m->nu_clip_factor;
    p.vacuum_n_lo = m->vacuum_n_lo; p.vacuum_n_hi = m->vacuum_n_hi;
    // tile: small grids take the small tile so the launch still covers the chip
    c->tile = ((int64_t)p.Nx * p.Ny >= 300 * 300) ? 0 : 1;
    c->npartials = nblk((int64_t)p.Nx * p.Ny, 256);

    hipError_t e;
    if ((e = hipGetDevice(&c->dev)) != hipSuccess) { delete c; return fail(nullptr, FIVO_ERR_HIP, hipGetErrorString(e)); }
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete c; return fail(nullptr, FIVO_ERR_HIP, hipGetErrorString(e));
    }
    const size_t P = p.pitch;
    const size_t ndbl = 4 * nv * P + nv * P + nv * P + 8 * P + 24 * P + 3 * (size_t)c->npartials + 64;
    c->ntiles_max = ((p.Nx + 6) / 7) * ((p.Ny + 7) / 8);   // the smallest face tile, 7 x 8
    const size_t nbyte = ndbl * sizeof(double) + 4 * P + sizeof(PrimJob) * P + sizeof(int) * P + (size_t)c->ntiles_max + 256;
    c->pool_bytes = nbyte;
    if ((e = hipMalloc(&c->pool, nbyte)) != hipSuccess) {
        (void)hipStreamDestroy(c->stream); delete c;
        return fail(nullptr, FIVO_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    double* d = static_cast<double*>(c->pool);
    for (int k = 0; k < 4; ++k) { c->U[k] = d; d += nv * P; }
    c->dU = d; d += nv * P;
    c->snap = d; d += nv * P;
    c->ws = d; d += 8 * P;
    double** wd[] = {&c->w.yT, &c->w.phi, &c->w.mu, &c->w.alpha, &c->w.ux, &c->w.uy, &c->w.P, &c->w.n, &c->w.e,
                     &c->w.vx, &c->w.vy, &c->w.x0_yT, &c->w.x0_phi, &c->w.x0_ux, &c->w.x0_uy, &c->w.ux_prev,
                     &c->w.uy_prev, &c->w.alpha_prev, &c->w.T_prev, &c->w.nux_prev, &c->w.nuy_prev,
                     &c->w.shear_res, &c->w.gradAx, &c->w.gradAy};
    for (double** a : wd) { *a = d; d += P; }
    c->partials = d; d += 3 * (size_t)c->npartials;
    d += 64;
    uint8_t* b = reinterpret_cast<uint8_t*>(d);
    c->w.ok = b; b += P;
    c->w.bad = b; b += P;
    c->w.force = b; b += P;
    c->w.ftmp = b; b += P;
    c->jobs = reinterpret_cast<PrimJob*>(b); b += sizeof(PrimJob) * P;
    c->wlist = reinterpret_cast<int*>(b); b += sizeof(int) * P;
    c->dirty = b; b += c->ntiles_max;
    b += (4 - (reinterpret_cast<uintptr_t>(b) & 3)) & 3;   // within the pool's 256 spare bytes
    c->tail_done = reinterpret_cast<unsigned int*>(b); b += sizeof(unsigned int);
    if ((e = hipMalloc(&c->sc, sizeof(DevScalars))) != hipSuccess ||
        (e = hipHostMalloc(&c->sc_host, sizeof(DevScalars), hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc(&c->sc_stage_host, 4 * sizeof(DevScalars), hipHostMallocDefault)) != hipSuccess ||
        (e = hipMalloc(&c->sc_slots, 4 * sizeof(DevScalars))) != hipSuccess ||
        (e = hipHostMalloc(&c->partials_host, sizeof(double) * 2 * (size_t)c->npartials, hipHostMallocDefault)) !=
            hipSuccess ||
        (e = hipEventCreate(&c->ev[0])) != hipSuccess || (e = hipEventCreate(&c->ev[1])) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming)) != hipSuccess) {
        fivo_destroy(reinterpret_cast<fivo_ctx*>(c));
        return fail(nullptr, FIVO_ERR_HIP, hipGetErrorString(e));
    }
    (void)hipMemsetAsync(c->pool, 0, nbyte, c->stream);
    (void)hipMemsetAsync(c->sc, 0, sizeof(DevScalars), c->stream);
    int rc = reset_history_dev(*c);
    if (rc == FIVO_OK && (e = hipStreamSynchronize(c->stream)) != hipSuccess) rc = FIVO_ERR_HIP;
    if (rc != FIVO_OK) {
        std::string msg = c->err;
        fivo_destroy(reinterpret_cast<fivo_ctx*>(c));
        return fail(nullptr, FIVO_ERR_HIP, msg);
    }
    *out = reinterpret_cast<fivo_ctx*>(c);
    return FIVO_OK;
}

void fivo_destroy(fivo_ctx* ctx) {
    if (!ctx) return;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->pool) (void)hipFree(c->pool);
    if (c->sc) (void)hipFree(c->sc);
    if (c->sc_host) (void)hipHostFree(c->sc_host);
    if (c->sc_stage_host) (void)hipHostFree(c->sc_stage_host);
    if (c->sc_slots) (void)hipFree(c->sc_slots);
    if (c->partials_host) (void)hipHostFree(c->partials_host);
    for (auto& ev : c->ev) if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : c->evpool) (void)hipEventDestroy(ev);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) { (void)hipStreamSynchronize(c->stream2); (void)hipStreamDestroy(c->stream2); }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int fivo_nvars(const fivo_ctx* ctx) { return reinterpret_cast<const Ctx*>(ctx)->p.nv; }
int64_t fivo_ntot(const fivo_ctx* ctx) { return reinterpret_cast<const Ctx*>(ctx)->p.Ntot; }
double fivo_e_vac_cut(const fivo_ctx* ctx) { return reinterpret_cast<const Ctx*>(ctx)->p.Evac; }

int fivo_set_state(fivo_ctx* ctx, const double* U, int nv, int64_t ntot) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (!U || nv != c.p.nv || ntot != c.p.Ntot) return fail(&c, FIVO_ERR_ARG, "set_state: shape mismatch");
    HIPCHK(&c, hipMemcpyAsync(c.dU, U, sizeof(double) * nv * ntot, hipMemcpyHostToDevice, c.stream));
    HIPCHK(&c, launch_layout(c, c.dU, c.U[c.cur], 1));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}
int fivo_get_state(fivo_ctx* ctx, double* U, int nv, int64_t ntot) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (!U || nv != c.p.nv || ntot != c.p.Ntot) return fail(&c, FIVO_ERR_ARG, "get_state: shape mismatch");
    HIPCHK(&c, launch_layout(c, c.U[c.cur], c.dU, 0));
    HIPCHK(&c, hipMemcpyAsync(U, c.dU, sizeof(double) * nv * ntot, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

static int work_xfer(Ctx& c, const fivo_work_view* v, bool to_dev) {
    const size_t nb = sizeof(double) * c.p.Ntot;
    struct { double* h; double* d; } f[] = {
        {v->yT, c.w.yT}, {v->phi, c.w.phi}, {v->mu, c.w.mu}, {v->alpha, c.w.alpha}, {v->ux, c.w.ux},
        {v->uy, c.w.uy}, {v->P, c.w.P}, {v->n, c.w.n}, {v->e, c.w.e}, {v->x0_yT, c.w.x0_yT},
        {v->x0_phi, c.w.x0_phi}, {v->x0_ux, c.w.x0_ux}, {v->x0_uy, c.w.x0_uy}, {v->ux_prev, c.w.ux_prev},
        {v->uy_prev, c.w.uy_prev}, {v->alpha_prev, c.w.alpha_prev}, {v->T_prev, c.w.T_prev},
        {v->nux_prev, c.w.nux_prev}, {v->nuy_prev, c.w.nuy_prev}, {v->shear_res, c.w.shear_res}};
    for (auto& x : f) {
        if (!x.h) continue;
        if (to_dev) HIPCHK(&c, hipMemcpyAsync(x.d, x.h, nb, hipMemcpyHostToDevice, c.stream));
        else HIPCHK(&c, hipMemcpyAsync(x.h, x.d, nb, hipMemcpyDeviceToHost, c.stream));
    }
    if (v->ok) {
        if (to_dev) HIPCHK(&c, hipMemcpyAsync(c.w.ok, v->ok, c.p.Ntot, hipMemcpyHostToDevice, c.stream));
        else HIPCHK(&c, hipMemcpyAsync(v->ok, c.w.ok, c.p.Ntot, hipMemcpyDeviceToHost, c.stream));
    }
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}
int fivo_set_work(fivo_ctx* ctx, const fivo_work_view* w) {
    if (!w) return FIVO_ERR_ARG;
    return work_xfer(*reinterpret_cast<Ctx*>(ctx), w, true);
}
int fivo_get_work(fivo_ctx* ctx, fivo_work_view* w) {
    if (!w) return FIVO_ERR_ARG;
    return work_xfer(*reinterpret_cast<Ctx*>(ctx), w, false);
}
int fivo_reset_history(fivo_ctx* ctx) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    int rc = reset_history_dev(c);
    if (rc) return rc;
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

int fivo_finalize_ic(fivo_ctx* ctx, double tau0, int bc) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    HIPCHK(&c, launch_sanitize(c, c.U[c.cur], tau0, 0));
    HIPCHK(&c, launch_bc(c, c.U[c.cur], bc));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

int fivo_init_from_prims(fivo_ctx* ctx, double tau0, const double* T, const double* alpha, const double* ux,
                         const double* uy, const double* Pi, const double* pixx, const double* pixy,
                         const double* piyy, const double* nux, const double* nuy, int bc, int64_t* nbad) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (!T || !alpha) return fail(&c, FIVO_ERR_ARG, "init_from_prims: T and alpha are required");
    const double* h[10] = {T, alpha, ux, uy, Pi, pixx, pixy, piyy, nux, nuy};
    const double* dptr[10];
    const size_t nb = sizeof(double) * c.p.Ntot;
    // stage the host fields in the dU scratch (nv*pitch >= 4*pitch) and the snapshot buffer
    double* stage_buf[10];
    for (int k = 0; k < 10; ++k) stage_buf[k] = (k < 4 ? c.dU + k * (size_t)c.p.pitch : c.snap + (k - 4) * (size_t)c.p.pitch);
    if (c.p.nv < 6) {
        // nv*pitch of the snapshot may hold fewer than 6 fields: use the stage buffers
        for (int k = 4; k < 10; ++k) stage_buf[k] = c.U[(c.cur + 1 + (k - 4) / 4) % 4] + ((k - 4) % 4) * (size_t)c.p.pitch;
    }
    for (int k = 0; k < 10; ++k) {
        dptr[k] = nullptr;
        if (!h[k]) continue;
        HIPCHK(&c, hipMemcpyAsync(stage_buf[k], h[k], nb, hipMemcpyHostToDevice, c.stream));
        dptr[k] = stage_buf[k];
    }
    HIPCHK(&c, hipMemsetAsync(c.U[c.cur], 0, sizeof(double) * (size_t)c.p.nv * c.p.pitch, c.stream));
    HIPCHK(&c, hipMemsetAsync(c.sc, 0, sizeof(DevScalars), c.stream));
    HIPCHK(&c, launch_init_prims(c, c.U[c.cur], tau0, dptr, nullptr));
    HIPCHK(&c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    if (nbad) *nbad = (int64_t)c.sc_host->nbad;
    return fivo_finalize_ic(ctx, tau0, bc);
}

int fivo_init_hotspots(fivo_ctx* ctx, double tau0, int32_t nspots, const double* spots, double sigma, double Tpeak,
                       double alpha_floor, int bc, int64_t* nbad) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (!spots || nspots < 1 || !(sigma > 0) || !(Tpeak > 0)) return fail(&c, FIVO_ERR_ARG, "init_hotspots: bad argument");
    const size_t P = c.p.pitch;
    if ((size_t)3 * nspots > (size_t)c.p.nv * P) return fail(&c, FIVO_ERR_ARG, "init_hotspots: too many spots for the grid");
    // dU scratch: s, T, alpha; snapshot buffer: the spot table
    double* s = c.dU;
    double* T = c.dU + P;
    double* al = c.dU + 2 * P;
    double* sp = c.snap;
    const int nblk_ = (int)((c.p.Ntot + 255) / 256);
    if (nblk_ > 3 * c.npartials + 64) return fail(&c, FIVO_ERR_ARG, "init_hotspots: reduction buffer too small");
    HIPCHK(&c, hipMemcpyAsync(sp, spots, sizeof(double) * 3 * nspots, hipMemcpyHostToDevice, c.stream));
    HIPCHK(&c, launch_hotspot_sum(c, s, nspots, sp, sigma, c.partials));
    std::vector<double> part(nblk_);
    HIPCHK(&c, hipMemcpyAsync(part.data(), c.partials, sizeof(double) * nblk_, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    double smax = 0.0;
    for (double v : part) smax = v > smax ? v : smax;
    if (!(smax > 0)) return fail(&c, FIVO_ERR_ARG, "init_hotspots: no spot reaches the grid");
    HIPCHK(&c, launch_hotspot_prims(c, s, T, al, smax, Tpeak, alpha_floor));
    const double* dptr[10] = {T, al, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    HIPCHK(&c, hipMemsetAsync(c.U[c.cur], 0, sizeof(double) * (size_t)c.p.nv * P, c.stream));
    HIPCHK(&c, hipMemsetAsync(c.sc, 0, sizeof(DevScalars), c.stream));
    HIPCHK(&c, launch_init_prims(c, c.U[c.cur], tau0, dptr, nullptr));
    HIPCHK(&c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    if (nbad) *nbad = (int64_t)c.sc_host->nbad;
    return fivo_finalize_ic(ctx, tau0, bc);
}

int fivo_apply_bc(fivo_ctx* ctx, int bc) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    HIPCHK(&c, launch_bc(c, c.U[c.cur], bc));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

int fivo_rhs(fivo_ctx* ctx, double tau, int bc, const uint8_t* force, double* dU, fivo_counts* cnt) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    const uint8_t* fdev = nullptr;
    if (force) {
        HIPCHK(&c, hipMemcpyAsync(c.w.force, force, c.p.Ntot, hipMemcpyHostToDevice, c.stream));
        fdev = c.w.force;
    }
    // this call's retries and staged fallbacks
    HIPCHK(&c, clear_scalars(c, &c.sc->nretry, &c.sc->nstaged));
    int rc = rhs_dev(c, tau, bc, fdev);
    if (rc < 0) return rc;
    if (dU) {
        // SoA dU -> reference layout, via the snapshot buffer as staging
        HIPCHK(&c, launch_layout(c, c.dU, c.snap, 0));
        HIPCHK(&c, hipMemcpyAsync(dU, c.snap, sizeof(double) * c.p.nv * c.p.Ntot, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(&c, hipStreamSynchronize(c.stream));
    }
    if (cnt) {
        cnt->nprimfail = c.last_primfail; cnt->nvacuum = c.last_vacuum; cnt->amax = c.amax_host;
        cnt->nretry = (int64_t)c.sc_host->nretry; cnt->nstaged = (int64_t)c.sc_host->nstaged;
    }
    return FIVO_OK;
}

int fivo_update_primitives(fivo_ctx* ctx, double tau, int bc, fivo_counts* cnt) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    int rc;
    if ((rc = reset_scalars(c)) < 0) return rc;
    HIPCHK(&c, clear_scalars(c, &c.sc->nretry, &c.sc->nstaged));
    // keep amax: update_primitives_2d! does not touch the signal speeds
    if ((rc = do_bc(c, c.U[c.cur], bc)) < 0) return rc;
    if ((rc = do_prim(c, c.U[c.cur], tau)) < 0) return rc;
    double amax = c.amax_host;
    if ((rc = read_scalars(c)) < 0) return rc;
    c.amax_host = amax;
    if (cnt) {
        cnt->nprimfail = c.last_primfail; cnt->nvacuum = c.last_vacuum; cnt->amax = c.amax_host;
        cnt->nretry = (int64_t)c.sc_host->nretry; cnt->nstaged = (int64_t)c.sc_host->nstaged;
    }
    return FIVO_OK;
}

int fivo_background_slice(fivo_ctx* ctx, int32_t k, int32_t nt, double* T, double* vx, double* vy, double* n,
                          int32_t device, double* maxima) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (!T || !vx || !vy || nt < 1 || k < 0 || k >= nt)
        return fail(&c, FIVO_ERR_ARG, "background_slice: need T, vx, vy and 0 <= k < nt");
    const int64_t cells = (int64_t)c.p.Nx * c.p.Ny;
    HIPCHK(&c, clear_scalars(c, &c.sc->bg_umax_bits, &c.sc->bg_vmax_bits));
    if (device) {
        Timer t(c, TK_BG);
        HIPCHK(&c, launch_background(c, cells * k, T, vx, vy, n));
    } else {
        // host tables: the slice is staged in the snapshot buffer (nv >= 4 rows of Ntot)
        double* st = c.snap;
        {
            Timer t(c, TK_BG);
            HIPCHK(&c, launch_background(c, 0, st, st + cells, st + 2 * cells, n ? st + 3 * cells : nullptr));
        }
        double* dst[4] = {T, vx, vy, n};
        for (int f = 0; f < 4; ++f)
            if (dst[f])
                HIPCHK(&c, hipMemcpyAsync(dst[f] + cells * k, st + cells * f, sizeof(double) * cells,
                                          hipMemcpyDeviceToHost, c.stream));
    }
    HIPCHK(&c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    resolve_timing(c);
    if (maxima) {
        const unsigned long long b[2] = {c.sc_host->bg_umax_bits, c.sc_host->bg_vmax_bits};
        for (int j = 0; j < 2; ++j) std::memcpy(&maxima[j], &b[j], sizeof(double));
    }
    return FIVO_OK;
}

int fivo_sanitize(fivo_ctx* ctx, double tau, int bc) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    HIPCHK(&c, launch_sanitize(c, c.U[c.cur], tau, 1));
    HIPCHK(&c, launch_bc(c, c.U[c.cur], bc));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

int fivo_state_ok(fivo_ctx* ctx, double tau, int* ok, int64_t* nbad) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    HIPCHK(&c, hipMemsetAsync(&c.sc->nbad, 0, sizeof(unsigned long long), c.stream));
    HIPCHK(&c, launch_admissible(c, c.U[c.cur], tau, 1, nullptr));
    HIPCHK(&c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    if (ok) *ok = c.sc_host->nbad == 0 ? 1 : 0;
    if (nbad) *nbad = (int64_t)c.sc_host->nbad;
    return FIVO_OK;
}

int fivo_compute_dt(fivo_ctx* ctx, double tau, double cfl, double cfltau, double* dt) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    double amax = c.amax_host > 1e-6 ? c.amax_host : 1e-6;   // amax is finite and >= 0
    double dcfl = cfl * (c.p.dx < c.p.dy ? c.p.dx : c.p.dy) / amax;
    double dtau = cfltau * tau;
    *dt = dcfl < dtau ? dcfl : dtau;
    return FIVO_OK;
}

int fivo_step(fivo_ctx* ctx, int integ, double tau, double dtau, int bc, int max_halve, int max_mood, int* ok,
              double* used) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    return step_dev(c, integ, tau, dtau, bc, max_halve, max_mood, ok, used);
}

int fivo_relax(fivo_ctx* ctx, double tau, double dtau) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    int rc = relax_dev(c, tau, dtau);
    if (rc) return rc;
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

int fivo_project_shear(fivo_ctx* ctx, double* mx, double* mean) {
    return project_shear_dev(*reinterpret_cast<Ctx*>(ctx), mx, mean);
}

int fivo_run(fivo_ctx* ctx, const fivo_run_opts* o, fivo_run_result* r) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (!o || !r) return fail(&c, FIVO_ERR_ARG, "run: null argument");
    std::memset(r, 0, sizeof(*r));
    c.nrhs = 0; c.nhalvings = 0; c.nmood = 0;
    int rc;
    HIPCHK(&c, hipMemsetAsync(c.sc, 0, sizeof(DevScalars), c.stream));
    bool have_hist = false;
    if (o->reset_history) {
        if ((rc = reset_history_dev(c)) < 0) return rc;
    } else {
        double a0;
        HIPCHK(&c, hipMemcpyAsync(&a0, c.w.alpha_prev, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(&c, hipStreamSynchronize(c.stream));
        have_hist = std::isfinite(a0);
    }
    if ((rc = rhs_dev(c, o->tau0, o->bc, nullptr)) < 0) return rc;
    double Qin;
    if ((rc = diag(c, 0, &Qin, nullptr, nullptr)) < 0) return rc;
    double tau = o->tau0, max_shear = 0.0, next_dump = o->tau0;
    int64_t nsteps = 0, nprimfail = 0, nvacuum = 0;
    if (o->on_dump) { o->on_dump(o->user, tau, ctx); next_dump = o->tau0 + o->dump_dt; }
    const bool dissip = c.md.enable_shear || c.md.enable_bulk || c.md.enable_diff;
    // The refresh pass's counts and the shear projection's residual of a step are
    // copied to pinned memory without a synchronisation and taken in at the next one
    // (the next step's, a dump's or the run's end): the step's only wait is its stages'.
    bool defer_refresh = false, defer_shear = false;
    auto take_deferred = [&]() {
        if (defer_refresh) {
            nprimfail += (int64_t)c.sc_stage_host[3].nprimfail;
            nvacuum += (int64_t)c.sc_stage_host[3].nvacuum;
            defer_refresh = false;
        }
        if (defer_shear) {
            double m = 0.0;
            for (int b = 0; b < c.npartials; ++b) m = c.partials_host[2 * b] > m ? c.partials_host[2 * b] : m;
            max_shear = m > max_shear ? m : max_shear;
            defer_shear = false;
        }
    };
    while (tau < o->taufinal && nsteps < o->max_steps) {
        double D;
        fivo_compute_dt(ctx, tau, o->cfl, o->cfltau, &D);
        double rem = o->taufinal - tau;
        D = D < rem ? D : rem;
        if (D <= 0) break;
        if (have_hist || nsteps > 0) { HIPCHK(&c, launch_history(c)); have_hist = true; }
        int ok = 0;
        double used = 0;
        rc = step_dev(c, o->integrator, tau, D, o->bc, o->max_halve, o->max_mood, &ok, &used);
        if (rc < 0) return rc;
        take_deferred();   // step_dev synchronised after the deferred copies
        if (!ok) {
            r->ok = 0; r->tau = tau; r->nsteps = nsteps; r->nprimfail = nprimfail; r->nvacuum = nvacuum;
            r->max_shear_res = max_shear;
            r->maxu = r->dQ = r->Q1 = r->minPtot = std::numeric_limits<double>::quiet_NaN();
            r->Q0 = Qin; r->nrhs = c.nrhs; r->nhalvings = c.nhalvings; r->nmood = c.nmood;
            r->newton_iters = (int64_t)c.sc_host->newton_iters;
            r->nlearned = (int64_t)c.sc_host->nlearned;
            r->nlearned_fallback = (int64_t)c.sc_host->nfallback;
            r->nqueued = (int64_t)c.sc_host->nqueued;
            r->nretry = (int64_t)c.sc_host->nretry;
            r->nstaged = (int64_t)c.sc_host->nstaged;
            return FIVO_STEP_FAILED;
        }
        tau += used;
        nsteps += 1;
        if (dissip) {
            // update_primitives_2d! + relax_dissipative_2d!; amax stays the last RHS's
            // after an accepted all-stage attempt the per-pass scalars are clear and
            // the last stage-end kernel has filled the new state's ghosts
            if (!c.pass_clean && (rc = reset_scalars(c)) < 0) return rc;
            if (!c.pass_clean && (rc = do_bc(c, c.U[c.cur], o->bc)) < 0) return rc;
            c.pass_clean = false;
            if ((rc = do_prim(c, c.U[c.cur], tau)) < 0) return rc;
            if ((rc = relax_dev(c, tau, used)) < 0) return rc;
            HIPCHK(&c, hipMemcpyAsync(&c.sc_stage_host[3], c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost,
                                      c.stream));
            defer_refresh = true;
        } else {
            nprimfail += c.last_primfail;
            nvacuum += c.last_vacuum;
        }
        if (c.md.enable_shear) {
            HIPCHK(&c, launch_project_shear(c, c.U[c.cur], c.md.shear_constraint_project ? 1 : 0));
            HIPCHK(&c, hipMemcpyAsync(c.partials_host, c.partials, sizeof(double) * 2 * (size_t)c.npartials,
                                      hipMemcpyDeviceToHost, c.stream));
            defer_shear = true;
        }
        if (o->on_dump && tau >= next_dump - 1e-12) {
            HIPCHK(&c, hipStreamSynchronize(c.stream));
            take_deferred();
            o->on_dump(o->user, tau, ctx);
            next_dump += o->dump_dt;
        }
    }
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    resolve_timing(c);
    take_deferred();
    if (o->on_dump) o->on_dump(o->user, tau, ctx);
    double Qout, maxu, minp;
    if ((rc = diag(c, 1, &Qout, &maxu, &minp)) < 0) return rc;
    HIPCHK(&c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(&c, hipStreamSynchronize(c.stream));
    r->ok = 1; r->tau = tau; r->nsteps = nsteps; r->nprimfail = nprimfail; r->nvacuum = nvacuum;
    r->max_shear_res = max_shear; r->maxu = maxu;
    r->dQ = std::fabs(Qin) > 0 ? std::fabs(Qout - Qin) / std::fabs(Qin) : 0.0;
    r->Q0 = Qin; r->Q1 = Qout; r->minPtot = minp;
    r->nrhs = c.nrhs; r->nhalvings = c.nhalvings; r->nmood = c.nmood;
    r->newton_iters = (int64_t)c.sc_host->newton_iters;
    r->nlearned = (int64_t)c.sc_host->nlearned;
    r->nlearned_fallback = (int64_t)c.sc_host->nfallback;
    r->nqueued = (int64_t)c.sc_host->nqueued;
    r->nretry = (int64_t)c.sc_host->nretry;
    r->nstaged = (int64_t)c.sc_host->nstaged;
    return FIVO_OK;
}

int fivo_set_learned_seed(fivo_ctx* ctx, const fivo_mlp* m) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    c.learned = m ? 1 : 0;
    if (m) c.mlp = mlp_dev(m);
    return FIVO_OK;
}

int fivo_enable_timing(fivo_ctx* ctx, int on) {
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    c.timing = on != 0;
    for (int k = 0; k < TK_N; ++k) { c.t_ms[k] = 0; c.t_cnt[k] = 0; }
    return FIVO_OK;
}
static int probe(int which, uint32_t terms, int64_t n, const double* in, double* out, const char* who) {
    const int nin = which == 0 ? 22 : 44, nout = which == 0 ? 2 : 4;
    if (n < 0 || (n > 0 && (!in || !out))) return fail(nullptr, FIVO_ERR_ARG, std::string(who) + ": bad arguments");
    if (terms & ~0x3FFFFu) return fail(nullptr, FIVO_ERR_ARG, std::string(who) + ": terms: unknown bits set");
    if (n == 0) return FIVO_OK;
    double *din = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&din, sizeof(double) * nin * n);
    if (e == hipSuccess) e = hipMalloc(&dout, sizeof(double) * nout * n);
    if (e == hipSuccess) e = hipMemcpy(din, in, sizeof(double) * nin * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe(which, terms, n, din, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * nout * n, hipMemcpyDeviceToHost);
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return fail(nullptr, FIVO_ERR_HIP, std::string(who) + ": " + hipGetErrorString(e));
    return FIVO_OK;
}

int fivo_probe_fm_source(uint32_t terms, int64_t n, const double* in, double* out) {
    return probe(0, terms, n, in, out, "fivo_probe_fm_source");
}

int fivo_probe_m2_source(uint32_t terms, int64_t n, const double* in, double* out) {
    return probe(1, terms, n, in, out, "fivo_probe_m2_source");
}

int fivo_get_histograms(const fivo_ctx* ctx, int64_t* counts) {
    const Ctx& c = *reinterpret_cast<const Ctx*>(ctx);
    if (!counts) return FIVO_ERR_ARG;
    for (int k = 0; k < 24; ++k) counts[k] = (int64_t)c.sc_host->hist[k][0];
    return FIVO_OK;
}

int64_t fivo_kernel_time(const fivo_ctx* ctx, int cat, double* total_ms) {
    const Ctx& c = *reinterpret_cast<const Ctx*>(ctx);
    if (cat < 0 || cat >= TK_N) return -1;
    if (total_ms) *total_ms = c.t_ms[cat];
    return c.t_cnt[cat];
}

}  // extern "C"
