// =============================================================================
// fivo_cells.h -- device bodies of the 2+1D RHS and time step, shared by the
// one-event kernels (fivo_kernels.hip) and the ensemble kernels
// (fivo_batch_kernels.hip). Each body works on one event's arrays, given its
// Params, DevWork and an explicit thread / block index.
//
//   bc_cell        ghost fill, outflow or periodic          src2d/bc2d.jl:14-65
//   prim_cell      warm primitive recovery of one cell      src2d/rhs2d.jl:67-160
//   coop_job       a queued cell, cooperative Newton        src2d/primrec2d.jl:465-742
//   faces_body     fused: MUSCL + forward map + HLLE for every face of a tile into
//                  LDS, then per cell the flux divergence + Bjorken source, the RK
//                  combination, the stage sanitiser and the admissibility test
//                    src2d/reconstruction2d.jl:27-103, rhs2d.jl:169-371,
//                    fluxes2d.jl:24-177, timestepper2d.jl:113-188, floors2d.jl:24-337
//   pending_cell, history_cell, project_shear_block
// =============================================================================
#pragma once
#include <algorithm>
#include "fivo_ctx.h"
#include "fivo_mlp.h"

namespace fivo {
using namespace fd;

#define P_ (p.pitch)

// ---- helpers ---------------------------------------------------------------------
__device__ __forceinline__ Diss load_diss(const double* __restrict__ U, const Params& p, int i) {
    Diss d;
    d.nux = p.hasNu ? U[p.iNux * P_ + i] : 0.0;
    d.nuy = p.hasNu ? U[p.iNuy * P_ + i] : 0.0;
    d.Pi = p.hasPi ? U[p.iPi * P_ + i] : 0.0;
    d.pixx = p.hasShear ? U[p.iPixx * P_ + i] : 0.0;
    d.pixy = p.hasShear ? U[p.iPixy * P_ + i] : 0.0;
    d.piyy = p.hasShear ? U[p.iPiyy * P_ + i] : 0.0;
    d.pieta = p.hasShear ? U[p.iPieta * P_ + i] : 0.0;
    return d;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max_pos(double v) {   // NaN ignored, like `a > amax`
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { double u = __shfl_xor(v, o, 64); v = u > v ? u : v; }
    return v;
}
__device__ __forceinline__ void atomic_max_pos(unsigned long long* a, double v) {
    if (v > 0.0) atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// XCD-aware block order: consecutive logical tiles land on the same XCD (and L2)
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int nx = 8;
    int xcd = b % nx, idx = b / nx, per = nb / nx, rem = nb % nx;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// ---- ghost fill (src2d/bc2d.jl:14-65) -----------------------------------------------
// Outflow x-then-y copies equal a copy from the clamped cell; periodic x-then-y
// wraps equal a copy from the doubly wrapped cell. One thread per ghost cell t.
__device__ __forceinline__ int bc_nghost(const Params& p) { return 2 * p.ng * p.Nytot + 2 * p.ng * p.Nx; }
__device__ __forceinline__ void bc_cell(double* __restrict__ U, const Params& p, int periodic, int t) {
    const int ng = p.ng;
    const int nA = ng * p.Nytot, nB = ng * p.Nx;
    int ix, iy;
    if (t < nA) { ix = t / p.Nytot; iy = t - ix * p.Nytot; }
    else if ((t -= nA) < nA) { ix = ng + p.Nx + t / p.Nytot; iy = t % p.Nytot; }
    else if ((t -= nA) < nB) { ix = ng + t / ng; iy = t % ng; }
    else if ((t -= nB) < nB) { ix = ng + t / ng; iy = ng + p.Ny + t % ng; }
    else return;
    int sx, sy;
    if (periodic) {
        sx = ix < ng ? ix + p.Nx : (ix >= ng + p.Nx ? ix - p.Nx : ix);
        sy = iy < ng ? iy + p.Ny : (iy >= ng + p.Ny ? iy - p.Ny : iy);
    } else {
        sx = ix < ng ? ng : (ix >= ng + p.Nx ? ng + p.Nx - 1 : ix);
        sy = iy < ng ? ng : (iy >= ng + p.Ny ? ng + p.Ny - 1 : iy);
    }
    const int dst = ix * p.Nytot + iy, src = sx * p.Nytot + sy;
    for (int a = 0; a < p.nv; ++a) U[a * P_ + dst] = U[a * P_ + src];
}

// ---- primitive pass (src2d/rhs2d.jl:67-160) ---------------------------------------
// Two kernels. k_prim gives every cell whose Newton starts warm (the production
// path) or is vacuum PRIM_WARM_ITCAP iterations, one lane per cell; almost all
// cells converge there. A cell still iterating is paused (its iterate kept in
// the warm-start arrays) and queued; a cell with no warm start, or whose warm
// solve failed (the fresh-seed retry of cons_to_prim_2d!,
// src2d/primrec2d.jl:465-486), is queued for the cold seed. k_prim_coop then
// solves the queue with a group of lanes per cell: the slow cells set the pass's
// latency, and cooperating lanes cut each iteration to about two residual
// evaluations on the critical path instead of nine (one, when the line search
// accepts a trial whose Jacobian the group's upper lanes evaluated beside it:
// the first trial with 16 lanes, the first three with 32).
constexpr int PRIM_WARM_ITCAP = 6;
constexpr int PRIM_JOB_TAIL = 1000;   // job.it = PRIM_JOB_TAIL + iters: only the staged fallback is left
__device__ __forceinline__ void prim_finalize(const Params& p, const DevWork& w, int i, const PrimRes& r0,
                                              const Diss& d, unsigned long long& nfail, unsigned long long& nvac) {
    double T = r0.T, mu = r0.mu, ux = r0.ux, uy = r0.uy, n = r0.n, e = r0.e, P = r0.P;
    if (!r0.ok) {
        if (r0.reason == PRR_VACUUM) nvac += 1; else nfail += 1;
        T = T_MIN; mu = 0.0; ux = 0.0; uy = 0.0;
        // eos_Pne_2d(T_MIN, 0): the light sector at T_MIN, no charm density
        if (p.eos.kind == 1) {
            double Pl, dP, d2P;
            light_eval<1>(T_MIN, p.eos, &Pl, &dP, &d2P);
            P = Pl + 0.0; e = jmax((-Pl + T_MIN * dP) + 0.0, 0.0);
        } else {
            // conformal: the charm sector carries exp(-m/T_MIN) = 0
            const double Pl = p.eos.aSB * pow4(T_MIN) * p.eos.fmGeV3, el = 3.0 * Pl;
            double Ph = 0.0, eh = 0.0;
            if (p.eos.conf_charm) {
                const double x = p.eos.m / T_MIN;
                const double K2 = besselkx_asym(2, x), K1 = besselkx_asym(1, x);
                const double Ef = exp((0.0 - p.eos.m) / T_MIN);
                Ph = p.eos.Aconf * (T_MIN * T_MIN) * Ef * K2;
                eh = p.eos.Aconf * (T_MIN * T_MIN) * Ef * (3.0 * K2 + x * K1);
                if (p.eos.hq_fm) { Ph *= p.eos.fmGeV3; eh *= p.eos.fmGeV3; }
            }
            P = Pl + Ph; e = jmax(el + eh, 0.0);
        }
        n = 0.0;
    }
    const double Tm = jmax(T, T_MIN);
    const double yT = log(Tm);
    const double phi = jclamp((mu - p.eos.m) / Tm, -PHI_CAP, PHI_CAP);
    w.yT[i] = yT; w.phi[i] = phi; w.mu[i] = mu; w.alpha[i] = mu / Tm;
    w.ux[i] = ux; w.uy[i] = uy; w.P[i] = P; w.n[i] = n; w.e[i] = e;
    w.ok[i] = r0.ok ? 1 : 0;
    const double ut = sqrt(1 + ux * ux + uy * uy), iut = safe_inv(ut);
    w.vx[i] = ux * iut; w.vy[i] = uy * iut;
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    w.x0_yT[i] = r0.ok ? yT : nan;
    w.x0_phi[i] = r0.ok ? phi : nan;
    w.x0_ux[i] = r0.ok ? ux : nan;
    w.x0_uy[i] = r0.ok ? uy : nan;
    if (p.hasShear) {
        double tt, tx, ty;
        shear_tau_row(ux, uy, ut, d.pixx, d.pixy, d.piyy, &tt, &tx, &ty);
        double R = -tt + d.pixx + d.piyy + d.pieta;
        double scale = jmax(jmax(jmax(jmax(fabs(tt), fabs(d.pixx)), fabs(d.piyy)), fabs(d.pieta)), fabs(d.pixy));
        w.shear_res[i] = scale > TINY ? fabs(R) / scale : 0.0;
    }
}

// the rare recovery branches (fresh-seed retries, staged phi fallbacks): one ballot
// when none happened in the wavefront
__device__ __forceinline__ void rare_counts(DevScalars* sc, unsigned long long nre, unsigned long long nst) {
    if (__ballot(nre != 0 || nst != 0) == 0) return;
    nre = wave_sum(nre); nst = wave_sum(nst);
    if ((threadIdx.x & 63) == 0) {
        if (nre) atomicAdd(&sc->nretry, nre);
        if (nst) atomicAdd(&sc->nstaged, nst);
    }
}

// The pass's counts summed over the block (up to four wavefronts) and added with
// one atomic per nonzero sum: thousands of wavefronts adding to one counter are
// serialised on its line. Every thread of the block calls it.
__device__ __forceinline__ void prim_counts(DevScalars* sc, unsigned long long its, unsigned long long nfail,
                                            unsigned long long nvac, unsigned long long nq = 0) {
    __shared__ unsigned long long part[4][4];
    its = wave_sum(its); nfail = wave_sum(nfail); nvac = wave_sum(nvac); nq = wave_sum(nq);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { part[wv][0] = its; part[wv][1] = nfail; part[wv][2] = nvac; part[wv][3] = nq; }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int nw = (blockDim.x + 63) >> 6;
        unsigned long long v = 0;
        for (int k = 0; k < nw; ++k) v += part[k][threadIdx.x];
        unsigned long long* dst[4] = {&sc->newton_iters, &sc->nprimfail, &sc->nvacuum, &sc->nqueued};
        if (v) atomicAdd(dst[threadIdx.x], v);
    }
}

// list[(*cnt)++] = val for the lanes with `take`, one atomic per wavefront, the
// lanes in order; all lanes of the wavefront call it
__device__ __forceinline__ void wave_append(bool take, int val, unsigned long long* cnt, int* list) {
    const unsigned long long m = __ballot(take);
    if (m == 0ull) return;
    const int lead = __builtin_ctzll(m);
    unsigned long long base = 0;
    if ((int)(threadIdx.x & 63) == lead) base = atomicAdd(cnt, (unsigned long long)__popcll(m));
    base = __shfl(base, lead, 64);
    if (take) list[base + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = val;
}

// Histogram bins of a block's threads (bin < 0: none), summed in LDS and added to
// sc->hist[base + bin] with one atomic per nonzero bin. Every thread of the block
// calls it.
__device__ __forceinline__ void block_hist(DevScalars* sc, int base, int bin0, int bin1 = -1) {
    __shared__ unsigned int h[16];
    if (threadIdx.x < 16) h[threadIdx.x] = 0u;
    __syncthreads();
    if (bin0 >= 0) atomicAdd(&h[bin0], 1u);
    if (bin1 >= 0) atomicAdd(&h[bin1], 1u);
    __syncthreads();
    if (threadIdx.x < 16 && h[threadIdx.x]) atomicAdd(&sc->hist[base + threadIdx.x][0], (unsigned long long)h[threadIdx.x]);
}

// the warm part of the primitive pass for cell i; queued cells go to jobs[*njobs++].
// CHECK: only the cells whose warm start is already converged (or that end at their
// first residual: vacuum, non-finite) are finished here; a cell that would iterate
// sets *iterate and is left to a second launch over the compacted list of such
// cells, which repeats this function without CHECK from the same inputs. About
// 80 % of a warm pass's cells need no iteration and the rest 2-5, so wavefronts
// of all cells would idle most of their lanes through every iteration.
template <int EK, bool NCH, bool CHECK = false>
__device__ __forceinline__ void prim_cell(const double* __restrict__ U, const Params& p, double tau,
                                          const DevWork& w, int i, int ev, unsigned long long* njobs,
                                          PrimJob* jobs, unsigned long long& its, unsigned long long& nfail,
                                          unsigned long long& nvac, unsigned long long& nq,
                                          unsigned long long& nre, bool* iterate = nullptr, int* hbin = nullptr,
                                          unsigned long long* nres = nullptr, int qcap = 0) {
    const double D = safe_div(U[i], tau), Sx = U[P_ + i], Sy = U[2 * P_ + i], E = U[3 * P_ + i];
    const CellIn c = make_cell<NCH>(D, Sx, Sy, E, load_diss(U, p, i));
    const double yT0 = w.x0_yT[i], ux0 = w.x0_ux[i], uy0 = w.x0_uy[i];
    const bool vac = !isf(E) || E <= E_VAC || !isf(D) || D < 0.0;
    const bool warm = isf(yT0) && isf(ux0) && isf(uy0);
    int job = 0;                                 // < 0: done here; 0: cold seed; > 0: resume
    int cell_iters = 0;
    if (vac || warm) {
        NewtonState s;
        PrimRes r;
        bool fin = newton_begin<EK, NCH, false, 0>(c, yT0, w.x0_phi[i], ux0, uy0, p.eos, s, r);
        bool paused = false;
        if (!fin) {
            if constexpr (CHECK) {
                // newton_iterate's test at iteration 1: converged with no iteration
                if (!(newton_scaled_norm<NCH>(s, c) < NEWTON_TOL)) { *iterate = true; return; }
                s.converged = true; s.reason = PRR_CONVERGED; s.iters = s.it - 1;
            } else {
                newton_iterate<EK, NCH, 0>(s, c, p.eos, PRIM_WARM_ITCAP);
            }
            paused = newton_paused(s);
            if (paused) {
                job = s.it;
                w.x0_yT[i] = s.x[0]; w.x0_phi[i] = s.x[1]; w.x0_ux[i] = s.x[2]; w.x0_uy[i] = s.x[3];
            } else {
                r = newton_finish<EK, NCH, 0, false>(s, c, p.eos);
                if (r.reason == PRR_UNSET) {      // staged phi fallback: done by the group
                    paused = true;
                    job = PRIM_JOB_TAIL + s.iters;
                    w.x0_yT[i] = s.x[0]; w.x0_phi[i] = s.x[1]; w.x0_ux[i] = s.x[2]; w.x0_uy[i] = s.x[3];
                }
            }
        }
        if (!paused) {
            its += r.iters;
            cell_iters = r.iters;
            if (r.ok || r.reason == PRR_VACUUM) { prim_finalize(p, w, i, r, c.d, nfail, nvac); job = -1; }
            else if (warm) nre += 1;             // the warm start failed: the cell retries from the cold seed
        }
    }
    if (hbin) *hbin = job >= 0 ? 7 : (CHECK ? 0 : min(cell_iters, 6));
    if (job >= 0) {
        // with nres: resumed solves fill the queue from its far end (jobs[qcap - 1]
        // down), seeded ones from the front, so the cooperative kernel pairs like
        // with like in a wavefront and a ~56-iteration seeded walk never shares one
        // with a resumed solve's divergent branches
        if (nres && job > 0) {
            const unsigned long long k = atomicAdd(nres, 1ull);
            jobs[qcap - 1 - (long long)k] = PrimJob{i, job, ev};
        } else {
            const unsigned long long k = atomicAdd(njobs, 1ull);
            jobs[k] = PrimJob{i, job, ev};
        }
        nq += 1;
    }
}

// A learned start converges in a few iterations when the network's guess is
// close; one that has not converged after LEARNED_ITCAP iterations is given up
// for the cold seed, so a poor guess costs a bounded detour.
constexpr int LEARNED_ITCAP = 12;

// one queued cell, solved by the G lanes of a group (all lanes return the same counts).
// Job modes: it > 0 resumes the paused warm solve; it < 0 starts from the learned
// seed left in the warm-start arrays; it == 0 starts from the cold seed. A
// failed resumed or learned solve retries from the cold seed.
template <int EK, bool NCH, int G>
__device__ __forceinline__ void coop_job(const double* __restrict__ U, const Params& p, double tau, const DevWork& w,
                                         const PrimJob& jb, int& itsum, unsigned long long& nfail,
                                         unsigned long long& nvac, unsigned long long& nfallback,
                                         unsigned long long& nre, unsigned long long& nst) {
    const int i = jb.cell;
    const double D = safe_div(U[i], tau), Sx = U[P_ + i], Sy = U[2 * P_ + i], E = U[3 * P_ + i];
    const CellIn c = make_cell<NCH>(D, Sx, Sy, E, load_diss(U, p, i));
    NewtonState s;
    PrimRes r;
    itsum = 0;
    int mode = jb.it > 0 ? 1 : (jb.it < 0 ? 2 : 0);   // 1 resume, 2 learned seed, 0 cold seed
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        bool fin = false;
        if (mode == 1) {
            s.x[0] = w.x0_yT[i]; s.x[1] = w.x0_phi[i]; s.x[2] = w.x0_ux[i]; s.x[3] = w.x0_uy[i];
            s.it = jb.it; s.iters = 0; s.reason = PRR_UNSET; s.converged = false;
            newton_set<EK, NCH>(s, c, p.eos);
        } else {
            const bool ls = mode == 2;
            fin = newton_begin<EK, NCH, true, G>(c, ls ? w.x0_yT[i] : nan, ls ? w.x0_phi[i] : nan,
                                                    ls ? w.x0_ux[i] : nan, ls ? w.x0_uy[i] : nan, p.eos, s, r);
        }
        if (!fin) {
            if (mode == 1 && jb.it >= PRIM_JOB_TAIL) s.iters = jb.it - PRIM_JOB_TAIL;   // tail only
            else newton_iterate<EK, NCH, G>(s, c, p.eos, mode == 2 ? LEARNED_ITCAP : NEWTON_MAXIT);
            if (mode == 2 && newton_paused(s)) {         // a learned start that is not close: give it up
                itsum += LEARNED_ITCAP;
                nfallback += 1;
                mode = 0;
                continue;
            }
            r = newton_finish<EK, NCH, G>(s, c, p.eos);
        }
        itsum += r.iters;
        nst += r.staged;
        if (mode == 0 || r.ok || r.reason == PRR_VACUUM) break;
        if (mode == 2) nfallback += 1; else nre += 1;   // failed learned start / failed warm start
        mode = 0;
    }
    if (glane<G>() == 0) prim_finalize(p, w, i, r, c.d, nfail, nvac);
}

// The learned seed for 16 queued cold cells (jobs [16 tile, 16 tile + 16)) on one
// wavefront: features of the cell's conserved state -> the MLP on MFMA tiles ->
// seed (yT, phi, ux, uy) into the warm-start arrays, job marked learned (it = -1).
//
// With the charge unknown (NCH) the seed is used only where the charge row pins
// phi: the row n u^tau + nu^tau = D is scaled by max(|D|, 1e-12)
// (src2d/primrec2d.jl:580-583), so below D = 1e-12 its tolerance is an absolute
// 1e-24 on n u^tau and phi is fixed only to ~1e-24 / (n u^tau) -- in a
// charge-free cell not at all: the cold solve walks phi down from 0 until
// n u^tau < 1e-24 and stops wherever that walk crosses the floor. Which phi a
// solve returns there depends on where it started, and the charge dynamics of
// the dilute edge (grad alpha, nu, the charge flux of the reconstructed phi)
// carry that into D and, through the mu n term of the entropy density in the
// transport coefficients, into T. Such cells keep the reference's cold seed, so
// they take the reference's path exactly. A seeded cell needs
// D - |nu| >= max(1e-12, 1e-3 D) (|nu^tau| < |nu|, so n u^tau is at least
// that): its phi is then fixed to ~1e-9 or better by the relative tolerance.
// Above that, phi is raised to at least phi_c, where the scaled charge residual
// is 1e-11, so the Jacobian's charge row does not vanish at the start.
constexpr double LEARNED_CHARGE_MIN = 1e-12;

template <int EK, bool NCH>
__device__ __forceinline__ int seed_tile(const double* __restrict__ U, const Params& p, double tau,
                                         const DevWork& w, PrimJob* jobs, int n, int tile, const MlpDev& M) {
    const int l = threadIdx.x & 63, col = l & 15;
    const int j = tile * 16 + col;
    bool cold = false;
    int cell = 0;
    float fn[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    MlpFrame fr;
    fr.sx = 1.0; fr.sy = 0.0;
    if (j < n && jobs[j].it == 0) {
        cell = jobs[j].cell;
        const double D = safe_div(U[cell], tau), Sx = U[P_ + cell], Sy = U[2 * P_ + cell], E = U[3 * P_ + cell];
        const Diss d = load_diss(U, p, cell);
        const bool charge_ok = !NCH || (D - sqrt(d.nux * d.nux + d.nuy * d.nuy) >= jmax(LEARNED_CHARGE_MIN, 1e-3 * D));
        if (isf(E) && E > E_VAC && isf(D) && D >= 0.0 && isf(Sx) && isf(Sy) && charge_ok) {
            fr = mlp_frame(Sx, Sy);
            double f[MLP_IN];
            mlp_features(D, Sx, Sy, E, d, fr, f);
            cold = true;
            for (int k = 0; k < MLP_IN; ++k) {
                const double z = (f[k] - (double)M.norm[k]) / (double)M.norm[6 + k];
                cold = cold && isf(z);
                fn[k] = isf(z) ? (float)jclamp(z, -50.0, 50.0) : 0.f;
            }
            if (!cold) for (int k = 0; k < MLP_IN; ++k) fn[k] = 0.f;
        }
    }
    MlpFwd F;
    mlp_forward_tile(M.tiles, M.bias, mlp_feature_operand(fn), F);
    int used = 0;
    if (l < 16 && cold) {
        double x[4];
        mlp_seed_from_outputs(F.y, M.norm, fr, x);
        if (NCH) {
            const double T = jmax(exp(x[0]), T_MIN);
            const TPart tp = eos_tpart<EK>(T, p.eos);
            double P0, n0, e0;
            eos_assemble<EK>(tp, p.eos.m, p.eos, &P0, &n0, &e0);     // phi = 0: n = n0 e^phi
            const double ut = sqrt(1 + x[2] * x[2] + x[3] * x[3]);
            const double D = safe_div(U[cell], tau);
            if (n0 > 0 && isf(n0)) {
                const double phic = log(1e-11 * jmax(fabs(D), 1e-12) / (ut * n0));
                if (isf(phic)) x[1] = jclamp(jmax(x[1], phic), -PHI_CAP, PHI_CAP);
            }
        }
        if (isf(x[0]) && isf(x[1]) && isf(x[2]) && isf(x[3])) {
            w.x0_yT[cell] = x[0]; w.x0_phi[cell] = x[1]; w.x0_ux[cell] = x[2]; w.x0_uy[cell] = x[3];
            jobs[j].it = -1;
            used = 1;
        }
    }
    return used;
}


// ---- face physics -------------------------------------------------------------------
struct FPrim { double T, ux, uy, n, e, P, cs2; };

// prim_to_cons_2d! (src2d/primrec2d.jl:753-796) for a face state; also returns
// c_s^2 at the same T (eos_cs2, src/eos.jl:471-476) from the same light-sector exp.
template <int EK>
__device__ __forceinline__ bool p2c(double T, double mu, double ux, double uy, const Diss& d, double tau,
                                    const EosP& q, double cons[4], FPrim& pr) {
    const double Tc = jmax(T, T_MIN);
    const double ut = sqrt(1 + ux * ux + uy * uy);
    double P, n, e, cs2;
    if (EK != 1) {
        TPart t = eos_tpart<EK>(Tc, q);
        eos_assemble<EK>(t, mu, q, &P, &n, &e);
        cs2 = 1.0 / 3.0;
    } else {
        double Pl, dP, d2P;
        light_eval<2>(Tc, q, &Pl, &dP, &d2P);
        TPart t;
        t.T = Tc; t.Pl = Pl; t.el = -Pl + Tc * dP; t.x = jmin(q.m / Tc, 100000.0);
        t.K2 = (!isf(t.x) || t.x <= 0.0) ? 0.0 : besselk2x_fast(t.x);
        t.K1 = 0.0;
        eos_assemble<1>(t, mu, q, &P, &n, &e);
        cs2 = (d2P == 0.0) ? (1.0 / 3.0) : (dP / (Tc * d2P));
    }
    if (!(isf(P) && isf(n) && isf(e)) || e <= 0.0) return false;
    const double weff = e + P + d.Pi, Ptot = P + d.Pi;
    double tt, tx, ty;
    shear_tau_row(ux, uy, ut, d.pixx, d.pixy, d.piyy, &tt, &tx, &ty);
    cons[0] = tau * (n * ut + nu_tau(ux, uy, ut, d.nux, d.nuy));
    cons[1] = weff * ut * ux + tx;
    cons[2] = weff * ut * uy + ty;
    cons[3] = weff * ut * ut - Ptot + tt;
    pr = {Tc, ux, uy, n, e, P, cs2};
    return true;
}

// flux_cell_2d! (src2d/fluxes2d.jl:24-69), the conserved rows (the advected
// dissipative rows are formed by the face, below)
__device__ __forceinline__ void flux4(const FPrim& s, const Diss& d, double tau, int dir, double F[4]) {
    const double ux = s.ux, uy = s.uy;
    const double ut = sqrt(1 + ux * ux + uy * uy);
    const double Ptot = s.P + d.Pi, weff = (s.e + s.P) + d.Pi;
    double tt, tx, ty;
    shear_tau_row(ux, uy, ut, d.pixx, d.pixy, d.piyy, &tt, &tx, &ty);
    if (dir == 0) {
        F[0] = tau * (s.n * ux + d.nux);
        F[1] = weff * ux * ux + Ptot + d.pixx;
        F[2] = weff * ux * uy + d.pixy;
        F[3] = weff * ut * ux + tx;
    } else {
        F[0] = tau * (s.n * uy + d.nuy);
        F[1] = weff * ux * uy + d.pixy;
        F[2] = weff * uy * uy + Ptot + d.piyy;
        F[3] = weff * ut * uy + ty;
    }
}

// wavespeeds_2d (src2d/fluxes2d.jl:87-114)
__device__ __forceinline__ void wavespeeds(const FPrim& s, double Pi, double pinn, int dir, double* lm, double* lp) {
    double cs2 = jclamp(s.cs2, 0.0, 0.999);
    const double wv = s.e + s.P;
    if (wv > TINY) cs2 = jmin(cs2 + (fabs(Pi) + fabs(pinn)) / wv, 0.999);
    const double ut = sqrt(1 + s.ux * s.ux + s.uy * s.uy), iut = safe_inv(ut);
    const double vx = s.ux * iut, vy = s.uy * iut;
    const double v2 = vx * vx + vy * vy;
    const double vn = dir == 0 ? vx : vy;
    const double den = 1 - v2 * cs2;
    if (den <= TINY) { *lm = -1.0; *lp = 1.0; return; }
    const double cs = safe_sqrt(cs2);
    const double rad = (1 - v2) * (den - vn * vn * (1 - cs2));
    const double disc = cs * safe_sqrt(jmax(rad, 0.0));
    *lm = jclamp((vn * (1 - cs2) - disc) / den, -1.0, 1.0);
    *lp = jclamp((vn * (1 - cs2) + disc) / den, -1.0, 1.0);
}

struct FaceK {
    Params p;
    FaceLaunch f;
    DevWork w;
    DevScalars* sc;
    int tiles_y, nmain;
    int remap;            // XCD-aware tile order (one grid per event)
    // 0: every tile; 1: the tiles no queued cell of the primitive pass touches (run
    // beside the cooperative kernel); 2: the marked tiles and the ghost-region faces
    // (after it), clearing the marks
    int pass;
    uint8_t* dirty;       // one mark per tile
};
// main tiles + extra blocks for the ghost-region faces of a TX x TY tiling
// Threads of a face-tile block: the tile's cells rounded up to whole wavefronts.
// A TX x TY tile has 2 TX TY + TX + TY faces; tiles with at most two faces per
// thread (e.g. 15 x 16, 12 x 20, 8 x 28, 7 x 8) finish the face phase in two
// rounds, where a 256-cell tile needs a third round for its TX + TY halo faces.
template <int TX, int TY>
constexpr int face_threads() { return (TX * TY + 63) / 64 * 64; }

// floor(a / b) for b > 0 and any sign of a
__device__ __forceinline__ int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Mark the face tiles whose work reads the primitives of cell i: a tile's faces
// read the cells [x0 - 2, x0 + TX + 1] x [y0 - 2, y0 + TY + 1] at most (MUSCL
// stencils of its halo faces), and its cell update its own cells.
template <int TX, int TY>
__device__ __forceinline__ void mark_tiles(int Nytot, int ng, int i, uint8_t* dirty, int tiles_y, int nmain) {
    const int gx = i / Nytot;
    const int ix = gx - ng, iy = i - gx * Nytot - ng;   // interior-origin coordinates
    const int tiles_x = nmain / tiles_y;
    const int alo = max(floordiv(ix - 2, TX), 0), ahi = min(floordiv(ix + 2, TX), tiles_x - 1);
    const int blo = max(floordiv(iy - 2, TY), 0), bhi = min(floordiv(iy + 2, TY), tiles_y - 1);
    for (int a = alo; a <= ahi; ++a)
        for (int b = blo; b <= bhi; ++b) dirty[a * tiles_y + b] = 1;
}

template <int TX, int TY, int NT = face_threads<TX, TY>()>
__host__ __device__ inline void face_grid(const Params& p, int* tiles_y, int* nmain, int* nblocks) {
    const int tiles_x = (p.Nx + TX - 1) / TX;
    *tiles_y = (p.Ny + TY - 1) / TY;
    *nmain = tiles_x * *tiles_y;
    const int ng = p.ng;
    const int nextra = 2 * ng * (p.Nxtot - 1) + 2 * p.Ny * (ng - 1) + 2 * ng * (p.Nytot - 1) + 2 * p.Nx * (ng - 1);
    *nblocks = *nmain + (nextra + NT - 1) / NT;
}

// One face: left cell i, right cell i + st (st = Nytot for x, 1 for y);
// k0 = 0-based index along dir of the left cell. _faces_2d! (src2d/rhs2d.jl:282-371).
// Writes the nv-row numerical flux to out[a*ostride] when WRITE; returns max |s|.
template <int EK, bool NCH, bool WRITE>
__device__ __forceinline__ double face(const FaceK& K, int i, int dir, int k0, bool forced, double* out,
                                       int ostride) {
    const Params& p = K.p;
    const DevWork& w = K.w;
    const double* __restrict__ U = K.f.Ub;
    const double tau = K.f.tau;
    const int st = dir == 0 ? p.Nytot : 1;
    const int kmax = dir == 0 ? p.Nxtot : p.Nytot;
    const int ip = i + st;
    const Diss dL = load_diss(U, p, i), dR = load_diss(U, p, ip);
    const double yTi = w.yT[i], yTp = w.yT[ip], uxi = w.ux[i], uxp = w.ux[ip], uyi = w.uy[i], uyp = w.uy[ip];
    double cL[4], cR[4];
    FPrim sL, sR;
    bool good = false, okL = false, okR = false;
    if (!forced) {
        const double phii = w.phi[i], phip = w.phi[ip];
        double ul[4], ur[4];
        if (k0 < p.ng || k0 >= kmax - p.ng - 1) {
            ul[0] = yTi; ul[1] = phii; ul[2] = uxi; ul[3] = uyi;
            ur[0] = yTp; ur[1] = phip; ur[2] = uxp; ur[3] = uyp;
        } else {
            const int im = i - st, ipp = ip + st;
            const double fm[4] = {w.yT[im], w.phi[im], w.ux[im], w.uy[im]};
            const double f0[4] = {yTi, phii, uxi, uyi};
            const double f1[4] = {yTp, phip, uxp, uyp};
            const double f2[4] = {w.yT[ipp], w.phi[ipp], w.ux[ipp], w.uy[ipp]};
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double si = mc_limiter(f0[v] - fm[v], f1[v] - f0[v]);
                const double sp = mc_limiter(f1[v] - f0[v], f2[v] - f1[v]);
                ul[v] = jclamp(f0[v] + 0.5 * si, jmin(jmin(fm[v], f0[v]), f1[v]), jmax(jmax(fm[v], f0[v]), f1[v]));
                ur[v] = jclamp(f1[v] - 0.5 * sp, jmin(jmin(f0[v], f1[v]), f2[v]), jmax(jmax(f0[v], f1[v]), f2[v]));
            }
        }
        const double TL = exp(ul[0]), muL = p.eos.m + TL * ul[1];
        const double TR = exp(ur[0]), muR = p.eos.m + TR * ur[1];
        okL = p2c<EK>(TL, muL, ul[2], ul[3], dL, tau, p.eos, cL, sL);
        okR = p2c<EK>(TR, muR, ur[2], ur[3], dR, tau, p.eos, cR, sR);
        good = okL && okR && cL[0] >= 0.0 && cR[0] >= 0.0 && cL[3] >= E_FLOOR && cR[3] >= E_FLOOR;
    }
    if (!good) {
        okL = p2c<EK>(exp(yTi), w.mu[i], uxi, uyi, dL, tau, p.eos, cL, sL);
        okR = p2c<EK>(exp(yTp), w.mu[ip], uxp, uyp, dR, tau, p.eos, cR, sR);
    }
    // _floor_prim_2d! (src2d/rhs2d.jl:259-270): vacuum side at the energy floor
    Diss zL = dL, zR = dR;
    if (!okL) {
        cL[0] = cL[1] = cL[2] = 0.0; cL[3] = E_FLOOR;
        sL = {T_MIN, 0.0, 0.0, 0.0, E_FLOOR, 0.0, eos_cs2<EK>(T_MIN, p.eos)};
        zL = {0, 0, 0, 0, 0, 0, 0};
    }
    if (!okR) {
        cR[0] = cR[1] = cR[2] = 0.0; cR[3] = E_FLOOR;
        sR = {T_MIN, 0.0, 0.0, 0.0, E_FLOOR, 0.0, eos_cs2<EK>(T_MIN, p.eos)};
        zR = {0, 0, 0, 0, 0, 0, 0};
    }
    // hlle_flux_2d! (src2d/fluxes2d.jl:121-157)
    double FL[4], FR[4];
    flux4(sL, zL, tau, dir, FL);
    flux4(sR, zR, tau, dir, FR);
    double lmL, lpL, lmR, lpR;
    wavespeeds(sL, zL.Pi, dir == 0 ? zL.pixx : zL.piyy, dir, &lmL, &lpL);
    wavespeeds(sR, zR.Pi, dir == 0 ? zR.pixx : zR.piyy, dir, &lmR, &lpR);
    const double a_l = jmin(lmL, lmR), a_r = jmax(lpL, lpR);
    if (WRITE) {
        const int mode = a_l >= 0 ? 0 : (a_r <= 0 ? 1 : 2);
        const double inv = 1 / (a_r - a_l + TINY);
#pragma unroll
        for (int a = 0; a < 4; ++a)
            out[a * ostride] = mode == 0 ? FL[a] : (mode == 1 ? FR[a] : (a_r * FL[a] - a_l * FR[a] + a_r * a_l * (cR[a] - cL[a])) * inv);
        if (p.nv > 4) {
            // dissipative rows: physical flux stored * v_n when the field is advected
            // (advect_*, src2d/fluxes2d.jl:55-67), zero otherwise, and the HLLE jump of
            // the stored values (zero on a floor side); charm second-moment rows never
            // enter the face states
            const double jf = mode == 2 ? a_r * a_l : 0.0;
            double vL = 0.0, vR = 0.0;
            if (p.advect_nu || p.advect_Pi || p.advect_pi) {
                const double iuL = safe_inv(sqrt(1 + sL.ux * sL.ux + sL.uy * sL.uy));
                const double iuR = safe_inv(sqrt(1 + sR.ux * sR.ux + sR.uy * sR.uy));
                vL = (dir == 0 ? sL.ux : sL.uy) * iuL;
                vR = (dir == 0 ? sR.ux : sR.uy) * iuR;
            }
            auto row = [&](int idx, double l, double r, int adv) {
                if (idx < 0) return;
                const double fl = adv ? l * vL : 0.0, fr = adv ? r * vR : 0.0;
                out[idx * ostride] = mode == 0 ? fl : (mode == 1 ? fr : (a_r * fl - a_l * fr + jf * (r - l)) * inv);
            };
            row(p.iNux, zL.nux, zR.nux, p.advect_nu); row(p.iNuy, zL.nuy, zR.nuy, p.advect_nu);
            row(p.iPi, zL.Pi, zR.Pi, p.advect_Pi);
            row(p.iPixx, zL.pixx, zR.pixx, p.advect_pi); row(p.iPixy, zL.pixy, zR.pixy, p.advect_pi);
            row(p.iPiyy, zL.piyy, zR.piyy, p.advect_pi); row(p.iPieta, zL.pieta, zR.pieta, p.advect_pi);
        }
    }
    const double aa = jmax(fabs(a_l), fabs(a_r));
    return aa;
}

// ---- cell update: floors, S and nu bounds, admissibility -----------------------------
struct CellState { double D, Sx, Sy, E, nux, nuy, Pi, pixx, pixy, piyy, pieta; };

__device__ __forceinline__ void floors_cell(CellState& s, const Params& p, int ix, int iy) {
    // enforce_floors_2d! (src2d/floors2d.jl:24-104)
    const double Emin = E_FLOOR;
    const double Evac = jmax(Emin, p.Evac);
    auto blank = [&]() {
        s.Sx = 0.0; s.Sy = 0.0; s.D = 0.0;
        s.nux = 0.0; s.nuy = 0.0; s.Pi = 0.0; s.pixx = 0.0; s.pixy = 0.0; s.piyy = 0.0; s.pieta = 0.0;
    };
    if (p.has_rdom) {
        const double x = p.xmin + (ix + 1 - p.ng - 0.5) * p.dx, y = p.ymin + (iy + 1 - p.ng - 0.5) * p.dy;
        if (x * x + y * y > p.rdom2) { s.E = Emin; blank(); return; }
    }
    if (!isf(s.E)) { s.E = Emin; blank(); return; }
    if (s.E < Evac) { s.E = Emin; blank(); return; }
    if (s.E <= 1e6 * Emin) blank();
    if (!isf(s.D) || s.D < 0.0) s.D = 0.0;
    if (!isf(s.Sx)) s.Sx = 0.0;
    if (!isf(s.Sy)) s.Sy = 0.0;
    if (!isf(s.nux)) s.nux = 0.0;
    if (!isf(s.nuy)) s.nuy = 0.0;
    if (!isf(s.Pi)) s.Pi = 0.0;
    if (!isf(s.pixx)) s.pixx = 0.0;
    if (!isf(s.pixy)) s.pixy = 0.0;
    if (!isf(s.piyy)) s.piyy = 0.0;
    if (!isf(s.pieta)) s.pieta = 0.0;
}
__device__ __forceinline__ double peff_cell(const CellState& s, const Params& p, double P, bool useP) {
    double Peff = 0.0;
    if (useP && isf(P)) Peff = jmax(P, 0.0);
    if (p.hasPi && isf(s.Pi)) Peff += fabs(s.Pi);
    if (p.hasShear) {
        if (isf(s.pixx)) Peff += fabs(s.pixx);
        if (isf(s.pixy)) Peff += fabs(s.pixy);
        if (isf(s.piyy)) Peff += fabs(s.piyy);
        if (isf(s.pieta)) Peff += fabs(s.pieta);
    }
    return Peff;
}
__device__ __forceinline__ void bounds_cell(CellState& s, const Params& p, double P, bool useP, double tau) {
    // enforce_S_energy_constraint_2d! (src2d/floors2d.jl:118-162)
    if (isf(s.E) && isf(s.Sx) && isf(s.Sy) && !(s.E <= 0)) {
        const double Smax = CHI_SRE * (s.E + peff_cell(s, p, P, useP));
        const double aS = hypot(s.Sx, s.Sy);
        if (aS > Smax) { const double sc = Smax / (aS + TINY); s.Sx = s.Sx * sc; s.Sy = s.Sy * sc; }
    }
    // enforce_nu_charge_constraint_2d! (src2d/floors2d.jl:183-213)
    if (p.hasNu) {
        const double Dt = s.D / tau;
        if (isf(Dt) && isf(s.nux) && isf(s.nuy)) {
            if (Dt <= 0.0) { s.nux = 0.0; s.nuy = 0.0; }
            else {
                const double cap = 0.999 * Dt, an = hypot(s.nux, s.nuy);
                if (an > cap) { const double sc = cap / (an + TINY); s.nux = s.nux * sc; s.nuy = s.nuy * sc; }
            }
        }
    }
}
// cell_admissible_2d (src2d/floors2d.jl:281-337), cheap filters only:
// returns 1 admissible, 0 inadmissible, 2 = within the margin of the causal bound,
// where the reference settles it by inverting the state (admissible_newton below).
__device__ __forceinline__ int admissible_filters(const CellState& s, const Params& p, bool use_work, int ok,
                                                  double vx, double vy, double P) {
    const double Emin = E_FLOOR;
    if (!(isf(s.E) && isf(s.Sx) && isf(s.Sy) && isf(s.D))) return 0;
    if (!(s.E >= Emin)) return 0;
    if (!(s.D >= 0.0)) return 0;
    if (p.hasNu && !(isf(s.nux) && isf(s.nuy))) return 0;
    if (p.hasPi && !isf(s.Pi)) return 0;
    if (p.hasShear && !(isf(s.pixx) && isf(s.pixy) && isf(s.piyy) && isf(s.pieta))) return 0;
    if (s.E <= 1e6 * Emin) return 1;
    if (use_work && ok) {
        if (!(isf(vx) && isf(vy))) return 0;
        if (!((vx * vx + vy * vy) < (1 - 1e-12) * (1 - 1e-12))) return 0;
    }
    if (hypot(s.Sx, s.Sy) <= SR_MARGIN * CHI_SRE * (s.E + peff_cell(s, p, P, use_work))) return 1;
    return 2;
}
template <int EK, bool NCH>
__device__ __forceinline__ bool admissible_newton(const CellState& s, const Params& p, double tau) {
    const Diss d = {s.nux, s.nuy, s.Pi, s.pixx, s.pixy, s.piyy, s.pieta};
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    PrimRes r = cons_to_prim<EK, NCH>(make_cell<NCH>(safe_div(s.D, tau), s.Sx, s.Sy, s.E, d), nan, nan, nan, nan,
                                      p.eos);
    return r.ok || r.reason == PRR_VACUUM;
}

__device__ __forceinline__ CellState load_cell(const double* __restrict__ U, const Params& p, int i) {
    CellState s;
    s.D = U[i]; s.Sx = U[P_ + i]; s.Sy = U[2 * P_ + i]; s.E = U[3 * P_ + i];
    const Diss d = load_diss(U, p, i);
    s.nux = d.nux; s.nuy = d.nuy; s.Pi = d.Pi; s.pixx = d.pixx; s.pixy = d.pixy; s.piyy = d.piyy; s.pieta = d.pieta;
    return s;
}
__device__ __forceinline__ void store_cell(double* __restrict__ U, const Params& p, int i, const CellState& s) {
    U[i] = s.D; U[P_ + i] = s.Sx; U[2 * P_ + i] = s.Sy; U[3 * P_ + i] = s.E;
    if (p.hasNu) { U[p.iNux * P_ + i] = s.nux; U[p.iNuy * P_ + i] = s.nuy; }
    if (p.hasPi) U[p.iPi * P_ + i] = s.Pi;
    if (p.hasShear) {
        U[p.iPixx * P_ + i] = s.pixx; U[p.iPixy * P_ + i] = s.pixy;
        U[p.iPiyy * P_ + i] = s.piyy; U[p.iPieta * P_ + i] = s.pieta;
    }
}

// ---- the fused face / update kernel ---------------------------------------------------
// MODE 0: write the RHS to f.dUout (rhs_2d!); MODE 1: the RK stage update
// NT threads per block: face_threads (two faces per thread), or twice that where
// only a few tiles run (the second pass beside the cooperative solve), so the face
// phase takes one round and the block's latency shrinks
template <int EK, bool NCH, int TX, int TY, int MODE, int NT = face_threads<TX, TY>()>
__device__ __forceinline__ void faces_body(const FaceK& K, const int bid, double* __restrict__ sF) {
    constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1), NF = NFX + NFY;
    const Params& p = K.p;
    const int tid = threadIdx.x;
    double amax = 0.0;
    if (bid < K.nmain) {
        const int b = K.remap ? xcd_remap(bid, K.nmain) : bid;
        if (K.pass != 0 && ((K.dirty[b] != 0) != (K.pass == 2))) return;   // the other pass's tile
        const int tx_t = b / K.tiles_y, ty_t = b - tx_t * K.tiles_y;
        const int x0 = p.ng + tx_t * TX, y0 = p.ng + ty_t * TY;
        const int nxt = min(TX, p.ng + p.Nx - x0), nyt = min(TY, p.ng + p.Ny - y0);
        const uint8_t* force = K.f.force;
        for (int f = tid; f < NF; f += NT) {
            // one call site for both directions keeps a single copy of the face code
            int i = 0, dir = 0, k0 = 0, st = 0;
            bool valid;
            if (f < NFX) {
                const int fx = f / TY, fy = f - fx * TY;
                valid = fx <= nxt && fy < nyt;
                k0 = x0 + fx - 1; i = k0 * p.Nytot + y0 + fy; st = p.Nytot;
            } else {
                const int g = f - NFX, gx = g / (TY + 1), gy = g - gx * (TY + 1);
                valid = gx < nxt && gy <= nyt;
                k0 = y0 + gy - 1; i = (x0 + gx) * p.Nytot + k0; dir = 1; st = 1;
            }
            if (valid) {
                const bool forced = force && (force[i] || force[i + st]);
                double a = face<EK, NCH, true>(K, i, dir, k0, forced, sF + f, NF);
                amax = a > amax ? a : amax;
            }
        }
        __syncthreads();
        if (K.pass == 2 && tid == 0) K.dirty[b] = 0;   // every thread read the mark before the barrier
        // ---- per cell: divergence + source, then the RK stage update ----
        const int cx = tid / TY, cy = tid - cx * TY;
        if (cx < nxt && cy < nyt) {
            const int ix = x0 + cx, iy = y0 + cy, i = ix * p.Nytot + iy;
            const int fL = cx * TY + cy, fR = fL + TY, gB = NFX + cx * (TY + 1) + cy, gT = gB + 1;
            const double idx = 1 / p.dx, idy = 1 / p.dy, itau = safe_inv(K.f.tau);
            const double* __restrict__ Ub = K.f.Ub;
            CellState b = load_cell(Ub, p, i);
            const double Pw = K.w.P[i];
            auto du = [&](int a) {
                return -(sF[a * NF + fR] - sF[a * NF + fL]) * idx - (sF[a * NF + gT] - sF[a * NF + gB]) * idy;
            };
            // geometric source (src2d/fluxes2d.jl:166-177)
            const double S1 = -b.Sx * itau, S2 = -b.Sy * itau, S3 = -(b.E + (Pw + b.Pi) + b.pieta) * itau;
            // advected dissipative fields: + q div v, div v from the cell velocities
            // (compute_divv_2d!, src2d/rhs2d.jl:29-41, 200-224)
            double divv = 0.0;
            if (p.advect_nu || p.advect_Pi || p.advect_pi)
                divv = 0.5 * (K.w.vx[i + p.Nytot] - K.w.vx[i - p.Nytot]) * idx +
                       0.5 * (K.w.vy[i + 1] - K.w.vy[i - 1]) * idy;
            const double Snx = p.advect_nu ? b.nux * divv : 0.0, Sny = p.advect_nu ? b.nuy * divv : 0.0;
            const double SPi = p.advect_Pi ? b.Pi * divv : 0.0;
            const double Sxx = p.advect_pi ? b.pixx * divv : 0.0, Sxy = p.advect_pi ? b.pixy * divv : 0.0;
            const double Syy = p.advect_pi ? b.piyy * divv : 0.0, Set = p.advect_pi ? b.pieta * divv : 0.0;
            if (MODE == 0) {
                double* dU = K.f.dUout;
                dU[i] = du(0) + 0.0;
                dU[P_ + i] = du(1) + S1;
                dU[2 * P_ + i] = du(2) + S2;
                dU[3 * P_ + i] = du(3) + S3;
                if (p.hasNu) { dU[p.iNux * P_ + i] = du(p.iNux) + Snx; dU[p.iNuy * P_ + i] = du(p.iNuy) + Sny; }
                if (p.hasPi) dU[p.iPi * P_ + i] = du(p.iPi) + SPi;
                if (p.hasShear) {
                    dU[p.iPixx * P_ + i] = du(p.iPixx) + Sxx; dU[p.iPixy * P_ + i] = du(p.iPixy) + Sxy;
                    dU[p.iPiyy * P_ + i] = du(p.iPiyy) + Syy; dU[p.iPieta * P_ + i] = du(p.iPieta) + Set;
                }
                if (p.iPQ0 >= 0)
                    for (int a = p.iPQ0; a < p.nv; ++a) dU[a * P_ + i] = 0.0;
            } else {
                const double D = K.f.D, wo = K.f.wold, wn = K.f.wnew;
                const int ct = K.f.combine;
                const double* __restrict__ Uo = K.f.Uold;
                auto comb = [&](int a, double in, double k) {
                    if (ct == 0) return in + D * k;
                    const double old = Uo[a * P_ + i];
                    if (ct == 1) return wo * old + wn * (in + D * k);
                    return (old + 2.0 * (in + D * k)) / 3.0;
                };
                CellState s;
                s.D = comb(0, b.D, du(0) + 0.0);
                s.Sx = comb(1, b.Sx, du(1) + S1);
                s.Sy = comb(2, b.Sy, du(2) + S2);
                s.E = comb(3, b.E, du(3) + S3);
                s.nux = p.hasNu ? comb(p.iNux, b.nux, du(p.iNux) + Snx) : 0.0;
                s.nuy = p.hasNu ? comb(p.iNuy, b.nuy, du(p.iNuy) + Sny) : 0.0;
                s.Pi = p.hasPi ? comb(p.iPi, b.Pi, du(p.iPi) + SPi) : 0.0;
                s.pixx = p.hasShear ? comb(p.iPixx, b.pixx, du(p.iPixx) + Sxx) : 0.0;
                s.pixy = p.hasShear ? comb(p.iPixy, b.pixy, du(p.iPixy) + Sxy) : 0.0;
                s.piyy = p.hasShear ? comb(p.iPiyy, b.piyy, du(p.iPiyy) + Syy) : 0.0;
                s.pieta = p.hasShear ? comb(p.iPieta, b.pieta, du(p.iPieta) + Set) : 0.0;
                if (p.iPQ0 >= 0) {
                    // charm second-moment rows: never in the face states, zero flux
                    for (int k = 0; k < 5; ++k) {
                        const int a = p.iPQ0 + k;
                        K.f.Uout[a * P_ + i] = comb(a, Ub[a * P_ + i], 0.0);
                    }
                }
                // sanitize_stage_2d! at tau_state (src2d/floors2d.jl:218-227)
                floors_cell(s, p, ix, iy);
                bounds_cell(s, p, Pw, true, K.f.tau_state);
                store_cell(K.f.Uout, p, i, s);
                // state_ok_2d / mark_bad_2d! (src2d/floors2d.jl:281-363); near-bound
                // cells are flagged 2 and inverted by k_admissible_pending
                const int adm = admissible_filters(s, p, true, K.w.ok[i], K.w.vx[i], K.w.vy[i], Pw);
                K.w.bad[i] = adm == 1 ? 0 : (adm == 0 ? 1 : 2);
                if (adm == 0) atomicAdd(&K.sc->nbad, 1ull);
                if (adm == 2) atomicAdd(&K.sc->npending, 1ull);
            }
        }
    } else {
        if (K.pass == 1) return;   // ghost cells may be queued: these wait for every primitive
        // faces outside every interior cell's stencil: the reference evaluates them
        // too and their signal speeds enter amax (src2d/rhs2d.jl:290-367)
        const int ng = p.ng, Nx = p.Nx, Ny = p.Ny, Nxt = p.Nxtot, Nyt = p.Nytot;
        int t = (bid - K.nmain) * NT + tid;
        const int r1 = ng * (Nxt - 1), r3 = Ny * (ng - 1), r5 = ng * (Nyt - 1), r7 = Nx * (ng - 1);
        int ix = -1, iy = -1, dir = 0;
        if (t < r1) { ix = t / ng; iy = t % ng; }
        else if ((t -= r1) < r1) { ix = t / ng; iy = ng + Ny + t % ng; }
        else if ((t -= r1) < r3) { iy = ng + t / (ng - 1); ix = t % (ng - 1); }
        else if ((t -= r3) < r3) { iy = ng + t / (ng - 1); ix = ng + Nx + t % (ng - 1); }
        else if ((t -= r3) < r5) { dir = 1; ix = t / (Nyt - 1); iy = t % (Nyt - 1); }
        else if ((t -= r5) < r5) { dir = 1; ix = ng + Nx + t / (Nyt - 1); iy = t % (Nyt - 1); }
        else if ((t -= r5) < r7) { dir = 1; ix = ng + t / (ng - 1); iy = t % (ng - 1); }
        else if ((t -= r7) < r7) { dir = 1; ix = ng + t / (ng - 1); iy = ng + Ny + t % (ng - 1); }
        if (ix >= 0) {
            const int i = ix * Nyt + iy;
            double a = face<EK, NCH, false>(K, i, dir, dir == 0 ? ix : iy, false, nullptr, 0);
            amax = a > amax ? a : amax;
        }
    }
    amax = wave_max_pos(amax);
    if ((tid & 63) == 0) atomic_max_pos(&K.sc->amax_bits, amax);
}

// the pending near-bound cells of a stage: settle each by a cold inversion
template <int EK, bool NCH>
__device__ __forceinline__ void pending_cell(const double* U, const Params& p, double tau, uint8_t* bad,
                                             DevScalars* sc, int t) {
    if (t >= p.Nx * p.Ny) return;
    const int ix = p.ng + t / p.Ny, iy = p.ng + t % p.Ny, i = ix * p.Nytot + iy;
    if (bad[i] != 2) return;
    CellState s = load_cell(U, p, i);
    const bool adm = admissible_newton<EK, NCH>(s, p, tau);
    bad[i] = adm ? 0 : 1;
    if (!adm) atomicAdd(&sc->nbad, 1ull);
}

// history snapshot before the step (main2D.jl:458-463)
__device__ __forceinline__ void history_cell(const DevWork& w, int t) {
    w.alpha_prev[t] = w.alpha[t];
    w.T_prev[t] = exp(w.yT[t]);
}

// project_shear! (main2D.jl:332-358): (max, sum) of the residual over the 256
// interior cells t of one block, into part[0..1]
__device__ __forceinline__ void project_shear_block(double* U, const Params& p, const DevWork& w, int correct,
                                                    int t, double* part) {
    __shared__ double s0[256], s1[256];
    double rel = 0.0;
    if (t < p.Nx * p.Ny) {
        const int ix = p.ng + t / p.Ny, iy = p.ng + t % p.Ny, i = ix * p.Nytot + iy;
        const double ux = w.ux[i], uy = w.uy[i], ut = sqrt(1 + ux * ux + uy * uy);
        const double pixx = U[p.iPixx * P_ + i], pixy = U[p.iPixy * P_ + i], piyy = U[p.iPiyy * P_ + i];
        const double pieta = U[p.iPieta * P_ + i];
        double tt, tx, ty;
        shear_tau_row(ux, uy, ut, pixx, pixy, piyy, &tt, &tx, &ty);
        const double R = -tt + pixx + piyy + pieta;
        const double scale = jmax(jmax(jmax(jmax(fabs(tt), fabs(pixx)), fabs(piyy)), fabs(pieta)), fabs(pixy));
        rel = scale > TINY ? fabs(R) / scale : 0.0;
        if (correct) U[p.iPieta * P_ + i] = tt - pixx - piyy;
    }
    s0[threadIdx.x] = rel; s1[threadIdx.x] = rel;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s0[threadIdx.x] = jmax(s0[threadIdx.x], s0[threadIdx.x + o]);
            s1[threadIdx.x] += s1[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[0] = s0[0]; part[1] = s1[0]; }
}

}  // namespace fivo
