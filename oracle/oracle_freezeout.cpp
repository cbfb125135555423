// =============================================================================
// oracle/oracle_freezeout.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
// CPU restatement of the freeze-out surface and its Cooper-Frye pion yield
// (freezeout2d.jl:68-192), with the C entry points orc_freezeout_* of the same
// signatures as the library's fivo_freezeout_*.
//
//   record      freeze_surface's record! (freezeout2d.jl:97-119)
//   grad1       freezeout2d.jl:131-141
//   pion_yield  freezeout2d.jl:144-192: the loops in the reference's order
//               (cell, p_T, phi, eta), one partial per cell, the cells summed in
//               increasing order (the reference adds per-row partials
//               atomically, in no fixed order).
// Pinned by tests/test_oracle_freezeout.py: exact crossings of linear cooling
// histories, the re-heating count, closed forms of the integral (a static
// surface: Bessel K1 and its Bose series; a tilted surface with transverse flow:
// K0, K1, I0, I1), one-sided gradients at a rim, and invariants of the
// negative-dSigma share.
// =============================================================================
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>
#include "oracle_solver.hpp"

using namespace orc;

namespace {

// hbar c = HBARC (GeV fm, freezeout2d.jl:59), from oracle_physics.hpp
const double NaNv = std::numeric_limits<double>::quiet_NaN();

struct FreezeOut {
    Ctx* c = nullptr;
    double T_fo = 0.156, tau_prev = NaNv;
    int nx = 0, ny = 0;
    double dx = 0.0, dy = 0.0;
    std::vector<double> tfo, uxf, uyf, Tprev;
    std::vector<uint8_t> above;
    int64_t refroze = 0;
};

double grad1(const std::vector<double>& t, int k, int s, double h) {
    const double lo = t[k - s], hi = t[k + s], c = t[k];
    if (!std::isnan(lo) && !std::isnan(hi)) return (hi - lo) / (2.0 * h);
    if (!std::isnan(hi)) return (hi - c) / h;
    if (!std::isnan(lo)) return (c - lo) / h;
    return NaNv;
}

}  // namespace

extern "C" {

struct orc_freezeout;

int orc_freezeout_create(Ctx* c, double T_fo, orc_freezeout** out) {
    if (!out) return FIVO_ERR_ARG;
    *out = nullptr;
    if (!c) return FIVO_ERR_ARG;
    if (!(T_fo > 0.0)) { c->err = "freezeout_create: T_fo must be > 0"; return FIVO_ERR_ARG; }
    if (c->g.Nx < 3 || c->g.Ny < 3) { c->err = "freezeout_create: need nx, ny >= 3"; return FIVO_ERR_ARG; }
    FreezeOut* f = new FreezeOut();
    f->c = c;
    f->T_fo = T_fo;
    f->nx = c->g.Nx;
    f->ny = c->g.Ny;
    f->dx = c->g.dx;
    f->dy = c->g.dy;
    const size_t n = (size_t)f->nx * f->ny;
    f->tfo.assign(n, NaNv);
    f->uxf.assign(n, 0.0);
    f->uyf.assign(n, 0.0);
    f->Tprev.assign(n, NaNv);
    f->above.assign(n, 0);
    *out = reinterpret_cast<orc_freezeout*>(f);
    return FIVO_OK;
}

void orc_freezeout_destroy(orc_freezeout* fp) { delete reinterpret_cast<FreezeOut*>(fp); }

// record! (freezeout2d.jl:97-119)
int orc_freezeout_record(orc_freezeout* fp, double tau) {
    if (!fp) return FIVO_ERR_ARG;
    FreezeOut& f = *reinterpret_cast<FreezeOut*>(fp);
    const Grid& g = f.c->g;
    const Work& w = f.c->w;
    for (int a = 0; a < f.nx; ++a)
        for (int b = 0; b < f.ny; ++b) {
            const int k = a * f.ny + b, i = g.lin(g.ng + a, g.ng + b);
            const double T = std::exp(w.yT[i]);
            if (T > f.T_fo) {
                if (!std::isnan(f.tfo[k])) f.refroze += 1;
                f.above[k] = 1;
                f.tfo[k] = NaNv;
            } else if (f.above[k] && std::isnan(f.tfo[k])) {
                const double Tp = f.Tprev[k];
                f.tfo[k] = (std::isnan(Tp) || Tp <= f.T_fo) ? tau
                                                             : f.tau_prev + (tau - f.tau_prev) * (Tp - f.T_fo) / (Tp - T);
                f.uxf[k] = w.ux[i];
                f.uyf[k] = w.uy[i];
            }
            f.Tprev[k] = T;
        }
    f.tau_prev = tau;
    return FIVO_OK;
}

int orc_freezeout_surface(orc_freezeout* fp, double* tau_fo, double* ux, double* uy, int64_t* counts) {
    if (!fp) return FIVO_ERR_ARG;
    FreezeOut& f = *reinterpret_cast<FreezeOut*>(fp);
    const size_t n = f.tfo.size();
    if (tau_fo) std::memcpy(tau_fo, f.tfo.data(), n * sizeof(double));
    if (ux) std::memcpy(ux, f.uxf.data(), n * sizeof(double));
    if (uy) std::memcpy(uy, f.uyf.data(), n * sizeof(double));
    if (counts) {
        int64_t nhot = 0, nfroze = 0;
        for (size_t k = 0; k < n; ++k) {
            nhot += f.above[k] ? 1 : 0;
            nfroze += std::isnan(f.tfo[k]) ? 0 : 1;
        }
        counts[0] = nhot;
        counts[1] = nfroze;
        counts[2] = f.refroze;
    }
    return FIVO_OK;
}

int orc_freezeout_set_surface(orc_freezeout* fp, const double* tau_fo, const double* ux, const double* uy) {
    if (!fp) return FIVO_ERR_ARG;
    FreezeOut& f = *reinterpret_cast<FreezeOut*>(fp);
    if (!tau_fo || !ux || !uy) { f.c->err = "freezeout_set_surface: null array"; return FIVO_ERR_ARG; }
    const size_t n = f.tfo.size();
    std::memcpy(f.tfo.data(), tau_fo, n * sizeof(double));
    std::memcpy(f.uxf.data(), ux, n * sizeof(double));
    std::memcpy(f.uyf.data(), uy, n * sizeof(double));
    return FIVO_OK;
}

// pion_yield (freezeout2d.jl:144-192)
int orc_freezeout_pion_yield(orc_freezeout* fp, const fivo_cf_opts* o, double* out) {
    if (!fp) return FIVO_ERR_ARG;
    FreezeOut& f = *reinterpret_cast<FreezeOut*>(fp);
    if (!o || !out) { f.c->err = "pion_yield: null argument"; return FIVO_ERR_ARG; }
    if (o->n_eta < 2 || o->n_pt < 2 || o->n_phi < 2 || o->n_eta > 256 || o->n_pt > 256 || o->n_phi > 256) {
        f.c->err = "pion_yield: n_eta, n_pt and n_phi must lie in 2..256";
        return FIVO_ERR_ARG;
    }
    if (!(o->eta_max > 0.0) || !(o->pt_hi > o->pt_lo) || !(o->pt_lo >= 0.0) || !(o->mass >= 0.0)) {
        f.c->err = "pion_yield: need eta_max > 0, 0 <= pt_lo < pt_hi, mass >= 0 and T_fo > 0";
        return FIVO_ERR_ARG;
    }
    // the grids (freezeout2d.jl:146-149)
    const int ne = o->n_eta, np = o->n_pt, nq = o->n_phi;
    const double deta = 2.0 * o->eta_max / (ne - 1), dpt = (o->pt_hi - o->pt_lo) / (np - 1),
                 dphi = 2.0 * M_PI / nq;
    std::vector<double> ch(ne), pt(np), mT(np), cq(nq), sq(nq);
    for (int e = 0; e < ne; ++e) ch[e] = std::cosh(-o->eta_max + e * deta);
    for (int e = 0; e < np; ++e) {
        pt[e] = o->pt_lo + e * dpt;
        mT[e] = std::sqrt(pt[e] * pt[e] + o->mass * o->mass);
    }
    for (int e = 0; e < nq; ++e) {
        cq[e] = std::cos(e * dphi);
        sq[e] = std::sin(e * dphi);
    }
    const double wconst = dpt * dphi * deta * f.dx * f.dy, T_fo = f.T_fo;
    const int nx = f.nx, ny = f.ny;
    std::vector<double> cell(4 * (size_t)nx * ny, 0.0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int a = 1; a < nx - 1; ++a)
        for (int b = 1; b < ny - 1; ++b) {
            const int k = a * ny + b;
            if (std::isnan(f.tfo[k])) continue;
            const double gx = grad1(f.tfo, k, ny, f.dx), gy = grad1(f.tfo, k, 1, f.dy);
            if (std::isnan(gx) || std::isnan(gy)) continue;
            const double tau = f.tfo[k], ux = f.uxf[k], uy = f.uyf[k];
            const double ut = std::sqrt(1.0 + ux * ux + uy * uy);
            double acc = 0.0, accB = 0.0, neg = 0.0, pos = 0.0;
            for (int ip = 0; ip < np; ++ip)
                for (int iq = 0; iq < nq; ++iq) {
                    const double px = pt[ip] * cq[iq], py = pt[ip] * sq[iq];
                    const double gp = gx * px + gy * py, up = ux * px + uy * py;
                    const double wpt = pt[ip] * wconst;
                    for (int e = 0; e < ne; ++e) {
                        const double dSp = tau * (mT[ip] * ch[e] - gp);
                        const double x = (ut * mT[ip] * ch[e] - up) / T_fo;
                        const double fB = std::exp(-x);
                        const double fb = x > 1e-8 ? fB / (1.0 - fB) : 0.0;   // 1/(exp(x) - 1)
                        const double w = dSp * wpt;
                        acc += w * fB;
                        accB += w * fb;
                        if (dSp < 0.0) neg -= w * fB;
                        else pos += w * fB;
                    }
                }
            cell[4 * (size_t)k] = acc;
            cell[4 * (size_t)k + 1] = accB;
            cell[4 * (size_t)k + 2] = neg;
            cell[4 * (size_t)k + 3] = pos;
        }
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t k = 0; k < (size_t)nx * ny; ++k)
        for (int q = 0; q < 4; ++q) s[q] += cell[4 * k + q];
    const double t = 2.0 * M_PI * HBARC;
    const double pref = o->yield_factor / (t * t * t);
    out[0] = pref * s[0];
    out[1] = pref * s[1];
    out[2] = s[2] / std::fmax(s[3] + s[2], 1e-30);
    return FIVO_OK;
}

}  // extern "C"
