// =============================================================================
// oracle/oracle_solver.hpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
// CPU restatement of the 2+1D grid sweeps, stepper and driver (src2d/*.jl,
// main2D.jl). Layout follows the reference: U[a + nv*i], i = ix*Nytot + iy.
// =============================================================================
#pragma once
#include <string>
#include <vector>
#include "oracle_primrec.hpp"

namespace orc {

// Grid2D / make_grid2d (src2d/grid2d.jl:19-74)
struct Grid {
    int Nx, Ny, ng, Nxtot, Nytot, Ntot;
    double xmin, xmax, ymin, ymax, dx, dy;
    std::vector<double> xC, yC;
    void make(int nx, int ny, double xm, double ym, int nghost) {
        Nx = nx; Ny = ny; ng = nghost;
        xmin = -xm; xmax = xm; ymin = -ym; ymax = ym;
        dx = (xmax - xmin) / Nx; dy = (ymax - ymin) / Ny;
        Nxtot = Nx + 2 * ng; Nytot = Ny + 2 * ng; Ntot = Nxtot * Nytot;
        xC.resize(Nxtot); yC.resize(Nytot);
        for (int i = 1; i <= Nxtot; ++i) xC[i - 1] = xmin + (i - ng - 0.5) * dx;
        for (int j = 1; j <= Nytot; ++j) yC[j - 1] = ymin + (j - ng - 0.5) * dy;
    }
    int lin(int ix, int iy) const { return ix * Nytot + iy; }   // 0-based
};

// StateLayout2D / make_layout2d (src2d/state_layout2d.jl:61-84)
struct Layout {
    int nv = 4;
    int iD = 0, iSx = 1, iSy = 2, iE = 3;
    bool hasNu = false, hasPi = false, hasShear = false, hasM2 = false;
    int iNux = -1, iNuy = -1, iPi = -1, iPixx = -1, iPixy = -1, iPiyy = -1, iPieta = -1;
    int iPQxx = -1, iPQxy = -1, iPQyy = -1, iPQeta = -1, iPiQ = -1;
    void make(bool charge, bool bulk, bool shear, bool m2) {
        nv = 4;
        hasNu = charge; hasPi = bulk; hasShear = shear; hasM2 = m2;
        if (charge) { iNux = nv++; iNuy = nv++; }
        if (bulk) iPi = nv++;
        if (shear) { iPixx = nv++; iPixy = nv++; iPiyy = nv++; iPieta = nv++; }
        if (m2) { iPQxx = nv++; iPQxy = nv++; iPQyy = nv++; iPQeta = nv++; iPiQ = nv++; }
    }
};

// PrimIdealVisc2D (src2d/primitives2d.jl:10-26)
struct Prim {
    double T, mu, ux, uy, n, e, P, nux, nuy, Pi, pixx, pixy, piyy, pieta;
    bool ok;
};

struct Model {
    fivo_model_desc d;
    EOS eos;
    Layout L;
    double Evac;   // derived E_vac_cut
};

// Work2D (src2d/work2d.jl:15-139), the arrays this restatement needs
struct Work {
    std::vector<double> yT, phi, mu, alpha, ux, uy, P, n, e, vxC, vyC;
    std::vector<uint8_t> ok, bad, bad_tmp;
    std::vector<double> x0_yT, x0_phi, x0_ux, x0_uy;
    std::vector<double> ux_prev, uy_prev, alpha_prev, T_prev, nux_prev, nuy_prev;
    std::vector<double> ULp[2], URp[2], sig[2];   // 4 x Ntot, per direction
    std::vector<double> ULc[2], URc[2], Fh[2];    // nv x Ntot, per direction
    std::vector<double> S, k, U1, U2, diss_snap, theta;
    std::vector<double> slope_tmp, alphaF[2], gradA[2];
    std::vector<double> shear_res;
    std::vector<int> dbg_iters;      // Newton iterations of each cell in the last primitive pass
    std::vector<int64_t> dbg_hist;   // histogram of per-cell iterations over all passes
    std::vector<int64_t> dbg_hist2;  // [total|first-attempt][class][iterations]
    std::vector<int> dbg_class;      // per cell: warm start, retry, ok, first-attempt iterations
    int64_t primfail = 0, vacuum = 0, newton_iters = 0;
    int64_t nretry = 0, nstaged = 0;          // last primitive pass: fresh-seed retries, staged phi fallbacks
    int64_t nretry_run = 0, nstaged_run = 0;  // the same, summed since the run started
    double amax = 0.0;
};

struct Ctx {
    Grid g;
    Model m;
    std::vector<double> U;
    Work w;
    std::string err;
    int nthreads = 0;
    int64_t nrhs = 0, nhalvings = 0, nmood = 0;
};

// src2d/* sweeps (oracle_solver.cpp)
void apply_bc(Ctx& c, std::vector<double>& U, int bc);
void primitive_pass(Ctx& c, std::vector<double>& U, double tau);
void rhs(Ctx& c, std::vector<double>& dU, std::vector<double>& U, double tau, int bc,
         const uint8_t* force_first_order);
void enforce_floors(Ctx& c, std::vector<double>& U, double Emin);
void enforce_S_energy_constraint(Ctx& c, std::vector<double>& U, double chi, const double* P);
void enforce_nu_charge_constraint(Ctx& c, std::vector<double>& U, double tau, double chi);
void sanitize_stage(Ctx& c, std::vector<double>& U, double tau, int bc);
bool cell_admissible(Ctx& c, const std::vector<double>& U, int i, double tau, bool use_work);
bool state_ok(Ctx& c, const std::vector<double>& U, double tau, int64_t* nbad);
void expand_bad(Ctx& c, int radius);
double compute_dt(Ctx& c, double tau, double cfl, double cfltau);
bool step_ssprk2(Ctx& c, double tau, double dtau, int bc, int max_halve, int max_mood, double* used);
bool step_ssprk3(Ctx& c, double tau, double dtau, int bc, int max_halve, double* used);
void project_shear(Ctx& c, double* maxr, double* meanr);
// dissipation (oracle_dissipation.cpp)
void relax_dissipative(Ctx& c, double tau, double dtau);

}  // namespace orc
