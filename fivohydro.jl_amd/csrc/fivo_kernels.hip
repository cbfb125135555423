// =============================================================================
// fivo_kernels.hip -- the FP64 HIP kernels of one event's 2+1D RHS and time step
// (gfx950): thin wrappers around the bodies of fivo_cells.h, plus the kernels
// only the one-event API needs (MOOD dilation, standalone sanitiser and
// admissibility, IC forward map, layout copies, diagnostics) and the launchers.
// =============================================================================
#include "fivo_cells.h"

namespace fivo {
using namespace fd;

__global__ __launch_bounds__(256) void k_bc(double* __restrict__ U, Params p, int periodic) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < bc_nghost(p)) bc_cell(U, p, periodic, t);
}

// warm pass, first launch: every cell; the cells whose Newton iterates are listed
template <int EK, bool NCH>
__global__ __launch_bounds__(256) void k_prim_check(const double* __restrict__ U, Params p, double tau, DevWork w,
                                                    DevScalars* sc, PrimJob* jobs, int* wlist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long its = 0, nfail = 0, nvac = 0, nq = 0, nre = 0;
    bool iterate = false;
    int bin = -1;
    if (i < p.Ntot)
        prim_cell<EK, NCH, true>(U, p, tau, w, i, 0, &sc->ncold, jobs, its, nfail, nvac, nq, nre, &iterate, &bin,
                                 &sc->nresq, p.Ntot);
    wave_append(iterate, i, &sc->nwarm, wlist);
    block_hist(sc, 0, iterate ? -1 : bin);
    prim_counts(sc, its, nfail, nvac, nq);
    rare_counts(sc, nre, 0);
}
// warm pass, second launch: the listed cells, densely packed into wavefronts. The
// list (~20 % of the cells) holds fewer wavefronts than the GPU has SIMDs, so the
// kernel is scheduled for one wavefront per SIMD (98 -> 95 us per launch at 512^2)
template <int EK, bool NCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_prim(const double* __restrict__ U, Params p, double tau, DevWork w,
                                              DevScalars* sc, PrimJob* jobs, const int* __restrict__ wlist) {
    const int n = (int)sc->nwarm;
    unsigned long long its = 0, nfail = 0, nvac = 0, nq = 0, nre = 0;
    int bin0 = -1, bin1 = -1;   // the grid holds the list of a warm pass; a second cell per thread is rare
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int bin = -1;
        prim_cell<EK, NCH>(U, p, tau, w, wlist[k], 0, &sc->ncold, jobs, its, nfail, nvac, nq, nre, nullptr, &bin,
                           &sc->nresq, p.Ntot);
        if (bin0 < 0) bin0 = bin; else if (bin1 < 0) bin1 = bin; else atomicAdd(&sc->hist[bin][0], 1ull);
    }
    block_hist(sc, 0, bin0, bin1);
    prim_counts(sc, its, nfail, nvac, nq);
    rare_counts(sc, nre, 0);
}

template <int EK, bool NCH, int G>
__global__ __launch_bounds__(256) void k_prim_coop(const double* __restrict__ U, Params p, double tau, DevWork w,
                                                   DevScalars* sc, const PrimJob* __restrict__ jobs) {
    // seeded solves at the queue's front, resumed ones at its far end (prim_cell)
    const int nf = (int)sc->ncold, n = nf + (int)sc->nresq;
    const int ngroups = gridDim.x * (blockDim.x / G);
    unsigned long long its = 0, nfail = 0, nvac = 0, nfb = 0, nre = 0, nst = 0;
    int bin0 = -1, bin1 = -1;
    for (int t = (blockIdx.x * blockDim.x + threadIdx.x) / G; t < n; t += ngroups) {
        int itsum;
        const PrimJob jb = t < nf ? jobs[t] : jobs[p.Ntot - 1 - (t - nf)];
        coop_job<EK, NCH, G>(U, p, tau, w, jb, itsum, nfail, nvac, nfb, nre, nst);
        if (glane<G>() == 0) {
            its += itsum;
            const int bin = (jb.it > 0 ? 0 : 8) + min(itsum / 10, 7);   // resumed / seeded solves
            if (bin0 < 0) bin0 = bin; else if (bin1 < 0) bin1 = bin; else atomicAdd(&sc->hist[8 + bin][0], 1ull);
        }
    }
    block_hist(sc, 8, bin0, bin1);
    prim_counts(sc, its, nfail, nvac);
    if (glane<G>() == 0 && nfb) atomicAdd(&sc->nfallback, nfb);
    const bool lead = glane<G>() == 0;
    rare_counts(sc, lead ? nre : 0, lead ? nst : 0);
}

// learned seeds for the queued cold cells, one wavefront per 16 cells
template <int EK, bool NCH>
__global__ __launch_bounds__(64) void k_prim_seed(const double* __restrict__ U, Params p, double tau, DevWork w,
                                                  DevScalars* sc, PrimJob* jobs, MlpDev M) {
    const int n = (int)sc->ncold;
    unsigned long long used = 0;
    for (int tile = blockIdx.x; tile * 16 < n; tile += gridDim.x)
        used += seed_tile<EK, NCH>(U, p, tau, w, jobs, n, tile, M);
    used = wave_sum(used);
    if (threadIdx.x == 0 && used) atomicAdd(&sc->nlearned, used);
}

template <int EK, bool NCH, int TX, int TY, int MODE>
__global__ __launch_bounds__((face_threads<TX, TY>())) void k_faces(FaceK K) {
    extern __shared__ double sF[];
    faces_body<EK, NCH, TX, TY, MODE>(K, blockIdx.x, sF);
}
// the same with one face per thread (the few tiles of the second pass)
template <int EK, bool NCH, int TX, int TY, int MODE>
__global__ __launch_bounds__((2 * face_threads<TX, TY>())) void k_faces_wide(FaceK K) {
    extern __shared__ double sF[];
    faces_body<EK, NCH, TX, TY, MODE, 2 * face_threads<TX, TY>()>(K, blockIdx.x, sF);
}

template <int EK, bool NCH>
__global__ __launch_bounds__(256) void k_admissible_pending(const double* U, Params p, double tau, uint8_t* bad,
                                                            DevScalars* sc) {
    if (sc->npending == 0) return;   // no cell carries the pending mark
    pending_cell<EK, NCH>(U, p, tau, bad, sc, blockIdx.x * blockDim.x + threadIdx.x);
}

// The end of one RK stage of the plain loop in one launch: blocks [0, nbc) fill
// the ghost cells of the stage output (bc_cell), blocks [nbc, nbc + npd) settle
// its pending near-bound cells (pending_cell, striding over the interior), the
// others copy the Newton warm starts (wsn doubles from x0_yT on) to the stage's
// rollback slot when ws_dst is set. Once the scalars are final -- at once when no
// cell is pending (the first pending block), else after the last pending block --
// they are copied to the stage's slot and, with clear, the per-pass ones
// (amax_bits .. nwarm) are zeroed for the next stage.
template <int EK, bool NCH>
__global__ __launch_bounds__(256) void k_stage_end(double* __restrict__ U, Params p, int periodic, double tau,
                                                   uint8_t* bad, DevScalars* sc, DevScalars* slot,
                                                   const double* __restrict__ ws_src, double* __restrict__ ws_dst,
                                                   int64_t wsn, int nbc, int npd, int clear, unsigned int* done) {
    const int b = blockIdx.x;
    if (b < nbc) {
        bc_cell(U, p, periodic, b * 256 + threadIdx.x);
        return;
    }
    if (b >= nbc + npd) {
        if (!ws_dst) return;
        const int64_t n2 = wsn / 2;
        const double2* __restrict__ s2 = reinterpret_cast<const double2*>(ws_src);
        double2* __restrict__ d2 = reinterpret_cast<double2*>(ws_dst);
        const int64_t stride = (int64_t)(gridDim.x - nbc - npd) * 256;
        for (int64_t k = (int64_t)(b - nbc - npd) * 256 + threadIdx.x; k < n2; k += stride) d2[k] = s2[k];
        return;
    }
    // pending blocks: every block of the kernel reads npending before any writes it
    const bool pend = sc->npending != 0;
    __shared__ bool last;
    if (pend) {
        const int n = p.Nx * p.Ny;
        for (int t = (b - nbc) * 256 + threadIdx.x; t < n; t += npd * 256) pending_cell<EK, NCH>(U, p, tau, bad, sc, t);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            last = atomicAdd(done, 1u) == (unsigned int)npd - 1;
        }
        __syncthreads();
        if (!last) return;
        __threadfence();
        if (threadIdx.x == 0) *done = 0;
    } else if (b != nbc) {
        return;
    }
    constexpr int NW = (int)(sizeof(DevScalars) / sizeof(unsigned long long));
    const unsigned long long* src = reinterpret_cast<const unsigned long long*>(sc);
    unsigned long long* dst = reinterpret_cast<unsigned long long*>(slot);
    for (int k = threadIdx.x; k < NW; k += 256) dst[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (clear) {
        unsigned long long* f0 = &sc->amax_bits;
        const int nz = (int)(&sc->nwarm - f0) + 1;
        for (int k = threadIdx.x; k < nz; k += 256) f0[k] = 0;
    }
}

__global__ void k_history(DevWork w, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) history_cell(w, t);
}

__global__ __launch_bounds__(256) void k_project_shear(double* U, Params p, DevWork w, int correct, double* part) {
    project_shear_block(U, p, w, correct, blockIdx.x * blockDim.x + threadIdx.x, part + 2 * blockIdx.x);
}

// ---- MOOD mask dilation (src2d/floors2d.jl:369-383) ------------------------------------
__global__ __launch_bounds__(256) void k_expand(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, Params p) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.Ntot) return;
    const int ix = t / p.Nytot, iy = t - ix * p.Nytot;
    uint8_t v = in[t];
    if (!v && ix >= p.ng && ix < p.ng + p.Nx && iy >= p.ng && iy < p.ng + p.Ny)
        v = in[t - 1] | in[t + 1] | in[t - p.Nytot] | in[t + p.Nytot];
    out[t] = v ? 1 : 0;
}

// ---- standalone sanitiser / admissibility (fivo_sanitize, fivo_state_ok, finalize_ic) --
__global__ __launch_bounds__(256) void k_sanitize(double* U, Params p, const double* Pw, double tau) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.Nx * p.Ny) return;
    const int ix = p.ng + t / p.Ny, iy = p.ng + t % p.Ny, i = ix * p.Nytot + iy;
    CellState s = load_cell(U, p, i);
    floors_cell(s, p, ix, iy);
    bounds_cell(s, p, Pw ? Pw[i] : 0.0, Pw != nullptr, tau);
    store_cell(U, p, i, s);
}
template <int EK, bool NCH>
__global__ __launch_bounds__(256) void k_admissible(const double* U, Params p, DevWork w, int use_work, double tau,
                                                    uint8_t* bad, DevScalars* sc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.Nx * p.Ny) return;
    const int ix = p.ng + t / p.Ny, iy = p.ng + t % p.Ny, i = ix * p.Nytot + iy;
    CellState s = load_cell(U, p, i);
    int adm = admissible_filters(s, p, use_work != 0, use_work ? w.ok[i] : 0, use_work ? w.vx[i] : 0.0,
                                 use_work ? w.vy[i] : 0.0, use_work ? w.P[i] : 0.0);
    if (adm == 2) adm = admissible_newton<EK, NCH>(s, p, tau) ? 1 : 0;
    if (bad) bad[i] = adm ? 0 : 1;
    if (!adm) atomicAdd(&sc->nbad, 1ull);
}
// ---- IC forward map: set_cell! on every cell (main2D.jl:142-154) -------------------------
struct InitK { const double* a[10]; };
template <int EK>
__global__ __launch_bounds__(256) void k_init(double* U, Params p, double tau0, InitK in, DevScalars* sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.Ntot) return;
    auto g = [&](int k) { return in.a[k] ? in.a[k][i] : 0.0; };
    const double T = in.a[0][i], alpha = in.a[1][i], ux = g(2), uy = g(3);
    Diss d = {g(8), g(9), g(4), g(5), g(6), g(7), 0.0};
    const double ut = sqrt(1 + ux * ux + uy * uy);
    double tt, tx, ty;
    shear_tau_row(ux, uy, ut, d.pixx, d.pixy, d.piyy, &tt, &tx, &ty);
    d.pieta = tt - d.pixx - d.piyy;   // project_shear_traceless_2d (src2d/shear2d.jl:131-142)
    const double Tc = jmax(T, T_MIN), mu = alpha * T;
    double P, n, e;
    eos_Pne_2d<EK>(Tc, mu, p.eos, &P, &n, &e);
    if (!(isf(P) && isf(n) && isf(e)) || e <= 0.0) { atomicAdd(&sc->nbad, 1ull); return; }
    const double weff = e + P + d.Pi, Ptot = P + d.Pi;
    CellState s;
    s.D = tau0 * (n * ut + nu_tau(ux, uy, ut, d.nux, d.nuy));
    s.Sx = weff * ut * ux + tx;
    s.Sy = weff * ut * uy + ty;
    s.E = weff * ut * ut - Ptot + tt;
    s.nux = d.nux; s.nuy = d.nuy; s.Pi = d.Pi; s.pixx = d.pixx; s.pixy = d.pixy; s.piyy = d.piyy; s.pieta = d.pieta;
    store_cell(U, p, i, s);
}

// ---- hot-spot initial state on the device -------------------------------------------------
// s(x) = sum_k A_k exp(-|x - x_k|^2 / (2 sigma^2)) over every cell centre (ghosts
// included), in spot order, each product and sum rounded separately; per-block
// maxima into part. Then T = Tpeak cbrt(s / max s) (>= T_MIN) and
// alpha = 2 log(s / max s) (>= alpha_floor): the project's seeded event generator.
__global__ __launch_bounds__(256) void k_hotspot_sum(double* __restrict__ s_out, Params p, int nspots,
                                                     const double* __restrict__ spots, double inv2s2,
                                                     double* part) {
    __shared__ double sm[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (t < p.Ntot) {
        const int ix = t / p.Nytot, iy = t - ix * p.Nytot;
        const double ox = (double)(ix + 1 - p.ng) - 0.5, oy = (double)(iy + 1 - p.ng) - 0.5;
        const double px = ox * p.dx, py = oy * p.dy;
        const double X = p.xmin + px, Y = p.ymin + py;
        for (int k = 0; k < nspots; ++k) {
            const double ddx = X - spots[3 * k], ddy = Y - spots[3 * k + 1];
            const double a2 = ddx * ddx, b2 = ddy * ddy;
            const double d2 = a2 + b2;
            const double arg = -d2 * inv2s2;
            const double g = exp(arg);
            const double term = spots[3 * k + 2] * g;
            s = s + term;
        }
        s_out[t] = s;
    }
    sm[threadIdx.x] = t < p.Ntot ? s : 0.0;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = jmax(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}
__global__ __launch_bounds__(256) void k_hotspot_prims(const double* __restrict__ s_in, double* __restrict__ T,
                                                       double* __restrict__ alpha, int n, double smax, double Tpeak,
                                                       double alpha_floor) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double s = s_in[t] / smax;
    const double Tv = Tpeak * cbrt(s);
    T[t] = jmax(Tv, T_MIN);
    const double a = s > 0 ? 2.0 * log(jmax(s, 1e-300)) : alpha_floor;
    alpha[t] = jmax(a, alpha_floor);
}

// ---- small utility kernels ------------------------------------------------------------------
__global__ void k_fill(double* a, double v, int64_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}
// reference AoS (U[a + nv*i]) <-> device SoA
__global__ void k_layout(const double* __restrict__ src, double* __restrict__ dst, Params p, int to_soa) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.Ntot * p.nv) return;
    if (to_soa) { const int i = t / p.nv, a = t - i * p.nv; dst[a * P_ + i] = src[t]; }
    else { const int a = t / p.Ntot, i = t - a * p.Ntot; dst[a + p.nv * i] = src[a * P_ + i]; }
}
// diagnostics partials over the interior, per block (deterministic order):
// what 0: sum D; what 1: (sum D, max hypot(ux,uy), min over P>0 of P+Pi)
__global__ __launch_bounds__(256) void k_diag(const double* U, Params p, DevWork w, int what, double* part) {
    __shared__ double s0[256], s1[256], s2[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double q = 0.0, mu = 0.0, mp = dinf();
    if (t < p.Nx * p.Ny) {
        const int ix = p.ng + t / p.Ny, iy = p.ng + t % p.Ny, i = ix * p.Nytot + iy;
        q = U[i];
        if (what == 1) {
            mu = hypot(w.ux[i], w.uy[i]);
            if (w.P[i] > 0) mp = w.P[i] + (p.hasPi ? U[p.iPi * P_ + i] : 0.0);
        }
    }
    s0[threadIdx.x] = q; s1[threadIdx.x] = mu; s2[threadIdx.x] = mp;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s0[threadIdx.x] += s0[threadIdx.x + o];
            s1[threadIdx.x] = s1[threadIdx.x + o] > s1[threadIdx.x] ? s1[threadIdx.x + o] : s1[threadIdx.x];
            s2[threadIdx.x] = s2[threadIdx.x + o] < s2[threadIdx.x] ? s2[threadIdx.x + o] : s2[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[3 * blockIdx.x] = s0[0]; part[3 * blockIdx.x + 1] = s1[0]; part[3 * blockIdx.x + 2] = s2[0]; }
}
// ---- launchers -------------------------------------------------------------------------------
#define EOS_DISPATCH(c, BODY)                                                    \
    do {                                                                         \
        if ((c).p.eos.kind == 1) { constexpr int EK = 1; constexpr bool NCH = true; BODY; } \
        else if ((c).p.eos.has_charge) { constexpr int EK = 2; constexpr bool NCH = true; BODY; } \
        else { constexpr int EK = 0; constexpr bool NCH = false; BODY; }         \
    } while (0)

static inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

hipError_t launch_bc(Ctx& c, double* U, int bc) {
    const Params& p = c.p;
    const int n = 2 * p.ng * p.Nytot + 2 * p.ng * p.Nx;
    hipLaunchKernelGGL(k_bc, dim3(nblk(n, 256)), dim3(256), 0, c.stream, U, p, bc == FIVO_BC_PERIODIC ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_prim_warm(Ctx& c, const double* U, double tau) {
    const Params& p = c.p;
    // the queue and list counters are cleared with the other per-pass scalars; the
    // iterating cells (~20 % of a warm pass) on a grid that strides over any count
    EOS_DISPATCH(c, hipLaunchKernelGGL((k_prim_check<EK, NCH>), dim3(nblk(p.Ntot, 256)), dim3(256), 0, c.stream,
                                       U, p, tau, c.w, c.sc, c.jobs, c.wlist));
    EOS_DISPATCH(c, hipLaunchKernelGGL((k_prim<EK, NCH>), dim3(std::min(nblk(p.Ntot, 256), 1024)), dim3(256), 0,
                                       c.stream, U, p, tau, c.w, c.sc, c.jobs, c.wlist));
    return hipGetLastError();
}
hipError_t launch_prim_seed(Ctx& c, const double* U, double tau) {
    if (!c.learned) return hipSuccess;
    const Params& p = c.p;
    EOS_DISPATCH(c, hipLaunchKernelGGL((k_prim_seed<EK, NCH>), dim3(std::min(nblk(p.Ntot, 16), 4096)), dim3(64), 0,
                                       c.stream, U, p, tau, c.w, c.sc, c.jobs, c.mlp));
    return hipGetLastError();
}
hipError_t launch_prim_coop(Ctx& c, const double* U, double tau) {
    const Params& p = c.p;
    // one group of G lanes per queued cell. The queue of a warm pass holds ~0.5 %
    // of the cells: up to 2^19 cells it leaves most SIMDs idle, so the pass is
    // latency-bound and 32 lanes (the Jacobian evaluated ahead at three trials)
    // are fastest; larger grids queue enough cells to fill the GPU, where fewer
    // lanes per cell keep more cells in flight. 2048 blocks hold the queue of a
    // warm pass and stride over the queue of a cold one.
    const int G = p.Ntot <= (1 << 19) ? 32 : (p.Ntot <= (1 << 21) ? 16 : 8);
    const int nb = std::min(nblk((int64_t)p.Ntot * G, 256), 2048);
    if (G == 32)
        EOS_DISPATCH(c, hipLaunchKernelGGL((k_prim_coop<EK, NCH, 32>), dim3(nb), dim3(256), 0, c.stream, U, p, tau,
                                           c.w, c.sc, c.jobs));
    else if (G == 16)
        EOS_DISPATCH(c, hipLaunchKernelGGL((k_prim_coop<EK, NCH, 16>), dim3(nb), dim3(256), 0, c.stream, U, p, tau,
                                           c.w, c.sc, c.jobs));
    else
        EOS_DISPATCH(c, hipLaunchKernelGGL((k_prim_coop<EK, NCH, 8>), dim3(nb), dim3(256), 0, c.stream, U, p, tau,
                                           c.w, c.sc, c.jobs));
    return hipGetLastError();
}
hipError_t launch_prim(Ctx& c, const double* U, double tau) {
    hipError_t e = launch_prim_warm(c, U, tau);
    if (e == hipSuccess) e = launch_prim_seed(c, U, tau);
    if (e == hipSuccess) e = launch_prim_coop(c, U, tau);
    return e;
}

template <int EK, bool NCH, int TX, int TY>
static hipError_t faces_tx(Ctx& c, const FaceLaunch& f, int pass, hipStream_t st) {
    const Params& p = c.p;
    FaceK K;
    K.p = p; K.f = f; K.w = c.w; K.sc = c.sc; K.remap = 1; K.pass = pass; K.dirty = c.dirty;
    int nb;
    constexpr int NTW = 2 * face_threads<TX, TY>();
    // the second pass (the ~10 % of tiles near a queued cell, after the cooperative
    // solve) runs one face per thread: its blocks are few, so their latency counts
    const bool wide = pass == 2;
    if (wide) face_grid<TX, TY, NTW>(p, &K.tiles_y, &K.nmain, &nb);
    else face_grid<TX, TY>(p, &K.tiles_y, &K.nmain, &nb);
    constexpr int NF = (TX + 1) * TY + TX * (TY + 1);
    const int nvf = p.iPQ0 >= 0 ? p.iPQ0 : p.nv;   // rows with a face flux
    const size_t lds = sizeof(double) * (size_t)nvf * NF;
    static size_t lds_set = 0;
    if (lds > 65536 && lds > lds_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_faces<EK, NCH, TX, TY, 0>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_faces<EK, NCH, TX, TY, 1>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_faces_wide<EK, NCH, TX, TY, 0>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_faces_wide<EK, NCH, TX, TY, 1>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        lds_set = lds;
    }
    if (wide) {
        if (f.mode == 0) hipLaunchKernelGGL((k_faces_wide<EK, NCH, TX, TY, 0>), dim3(nb), dim3(NTW), lds, st, K);
        else hipLaunchKernelGGL((k_faces_wide<EK, NCH, TX, TY, 1>), dim3(nb), dim3(NTW), lds, st, K);
    } else {
        if (f.mode == 0) hipLaunchKernelGGL((k_faces<EK, NCH, TX, TY, 0>), dim3(nb), dim3(face_threads<TX, TY>()), lds, st, K);
        else hipLaunchKernelGGL((k_faces<EK, NCH, TX, TY, 1>), dim3(nb), dim3(face_threads<TX, TY>()), lds, st, K);
    }
    return hipGetLastError();
}

hipError_t launch_faces(Ctx& c, const FaceLaunch& f, int pass, hipStream_t st) {
    hipError_t e = hipSuccess;
    if (!st) st = c.stream;
    // 15 x 16 cells (one 256-lane block, two face rounds) from 300^2 up; 7 x 8 (one
    // wavefront) below, where larger tiles would leave CUs without blocks
    if (c.tile == 0) EOS_DISPATCH(c, e = (faces_tx<EK, NCH, 15, 16>(c, f, pass, st)));
    else EOS_DISPATCH(c, e = (faces_tx<EK, NCH, 7, 8>(c, f, pass, st)));
    return e;
}

template <int TX, int TY>
__global__ __launch_bounds__(256) void k_mark_tiles(Params p, const DevScalars* __restrict__ sc,
                                                    const PrimJob* __restrict__ jobs, uint8_t* dirty, int tiles_y,
                                                    int nmain) {
    const int nf = (int)sc->ncold, n = nf + (int)sc->nresq;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        mark_tiles<TX, TY>(p.Nytot, p.ng, (t < nf ? jobs[t] : jobs[p.Ntot - 1 - (t - nf)]).cell, dirty, tiles_y, nmain);
}
template <int TX, int TY>
static hipError_t mark_tx(Ctx& c) {
    int tiles_y, nmain, nb;
    face_grid<TX, TY>(c.p, &tiles_y, &nmain, &nb);
    hipLaunchKernelGGL((k_mark_tiles<TX, TY>), dim3(64), dim3(256), 0, c.stream, c.p, c.sc, c.jobs, c.dirty, tiles_y,
                       nmain);
    return hipGetLastError();
}
hipError_t launch_mark_tiles(Ctx& c) {
    return c.tile == 0 ? mark_tx<15, 16>(c) : mark_tx<7, 8>(c);
}

hipError_t launch_expand(Ctx& c, const uint8_t* in, uint8_t* out) {
    hipLaunchKernelGGL(k_expand, dim3(nblk(c.p.Ntot, 256)), dim3(256), 0, c.stream, in, out, c.p);
    return hipGetLastError();
}

hipError_t launch_sanitize(Ctx& c, double* U, double tau, int use_P) {
    hipLaunchKernelGGL(k_sanitize, dim3(nblk(c.p.Nx * c.p.Ny, 256)), dim3(256), 0, c.stream, U, c.p,
                       use_P ? (const double*)c.w.P : nullptr, tau);
    return hipGetLastError();
}

hipError_t launch_admissible(Ctx& c, const double* U, double tau, int use_work, uint8_t* bad_out) {
    EOS_DISPATCH(c, hipLaunchKernelGGL((k_admissible<EK, NCH>), dim3(nblk(c.p.Nx * c.p.Ny, 256)), dim3(256), 0,
                                       c.stream, U, c.p, c.w, use_work, tau, bad_out, c.sc));
    return hipGetLastError();
}

hipError_t launch_admissible_pending(Ctx& c, const double* U, double tau) {
    EOS_DISPATCH(c, hipLaunchKernelGGL((k_admissible_pending<EK, NCH>), dim3(nblk(c.p.Nx * c.p.Ny, 256)), dim3(256),
                                       0, c.stream, U, c.p, tau, c.w.bad, c.sc));
    return hipGetLastError();
}

hipError_t launch_stage_end(Ctx& c, double* U, double tau, int bc, int slot, double* ws_dst, bool clear) {
    const Params& p = c.p;
    const int nbc = nblk(2 * (int64_t)p.ng * p.Nytot + 2 * (int64_t)p.ng * p.Nx, 256);
    const int npd = std::min(nblk((int64_t)p.Nx * p.Ny, 256), 64);
    const int ncp = ws_dst ? 256 : 0;
    const int64_t wsn = 4 * (int64_t)p.pitch;   // x0_yT, x0_phi, x0_ux, x0_uy: consecutive pitch blocks
    EOS_DISPATCH(c, hipLaunchKernelGGL((k_stage_end<EK, NCH>), dim3(nbc + npd + ncp), dim3(256), 0, c.stream, U, p,
                                       bc == FIVO_BC_PERIODIC ? 1 : 0, tau, c.w.bad, c.sc, c.sc_slots + slot,
                                       c.w.x0_yT, ws_dst, wsn, nbc, npd, clear ? 1 : 0, c.tail_done));
    return hipGetLastError();
}

hipError_t launch_init_prims(Ctx& c, double* U, double tau0, const double* const* prims, unsigned long long*) {
    InitK in;
    for (int k = 0; k < 10; ++k) in.a[k] = prims[k];
    EOS_DISPATCH(c, hipLaunchKernelGGL((k_init<EK>), dim3(nblk(c.p.Ntot, 256)), dim3(256), 0, c.stream, U, c.p,
                                       tau0, in, c.sc));
    return hipGetLastError();
}

hipError_t launch_hotspot_sum(Ctx& c, double* s_out, int nspots, const double* spots, double sigma, double* part) {
    hipLaunchKernelGGL(k_hotspot_sum, dim3(nblk(c.p.Ntot, 256)), dim3(256), 0, c.stream, s_out, c.p, nspots, spots,
                       1.0 / (2 * sigma * sigma), part);
    return hipGetLastError();
}
hipError_t launch_hotspot_prims(Ctx& c, const double* s_in, double* T, double* alpha, double smax, double Tpeak,
                                double alpha_floor) {
    hipLaunchKernelGGL(k_hotspot_prims, dim3(nblk(c.p.Ntot, 256)), dim3(256), 0, c.stream, s_in, T, alpha,
                       c.p.Ntot, smax, Tpeak, alpha_floor);
    return hipGetLastError();
}

hipError_t launch_fill(Ctx& c, double* a, double v, int64_t n) {
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, c.stream, a, v, n);
    return hipGetLastError();
}

hipError_t launch_history(Ctx& c) {
    hipLaunchKernelGGL(k_history, dim3(nblk(c.p.Ntot, 256)), dim3(256), 0, c.stream, c.w, c.p.Ntot);
    return hipGetLastError();
}

hipError_t launch_reduce_diag(Ctx& c, const double* U, int what) {
    hipLaunchKernelGGL(k_diag, dim3(nblk(c.p.Nx * c.p.Ny, 256)), dim3(256), 0, c.stream, U, c.p, c.w, what,
                       c.partials);
    return hipGetLastError();
}

hipError_t launch_project_shear(Ctx& c, double* U, int correct) {
    hipLaunchKernelGGL(k_project_shear, dim3(nblk(c.p.Nx * c.p.Ny, 256)), dim3(256), 0, c.stream, U, c.p, c.w,
                       correct, c.partials);
    return hipGetLastError();
}

// ---- readout: one time slice of a transport background -------------------------------
// export_background2d (tools/export_background2d.jl:79-94): T = exp(yT) and the
// VELOCITY v^i = u^i / u^tau (u^tau = sqrt(1 + u^2)) and n of every interior cell,
// from the primitives of the last pass, into column-major (Nx, Ny, nt) tables
// (element (i, j, k) at i + Nx (j + Ny k)). The grid is y-fastest and the tables
// x-fastest: a 16 x 16 tile (one cell per thread) is read along y into LDS and
// written along x, so both sides move whole 128-byte lines. max |u| and max |v|
// go to the scalars; a NaN sticks (its bits sort above +inf), as Julia's max
// propagates it.
constexpr int BG_TILE = 16;
__device__ __forceinline__ unsigned long long bg_key(double h) {   // h >= 0 or NaN
    return h != h ? 0x7ff8000000000000ull : (unsigned long long)__double_as_longlong(h);
}
__global__ __launch_bounds__(BG_TILE * BG_TILE) void k_background(Params p, DevWork w, int64_t off,
                                                                  double* __restrict__ T, double* __restrict__ vx,
                                                                  double* __restrict__ vy, double* __restrict__ n,
                                                                  DevScalars* sc) {
    __shared__ double s[4][BG_TILE][BG_TILE + 1];
    const int tiles_y = (p.Ny + BG_TILE - 1) / BG_TILE;
    const int bx = blockIdx.x / tiles_y, by = blockIdx.x - bx * tiles_y;
    const int x0 = bx * BG_TILE, y0 = by * BG_TILE;
    const int tid = threadIdx.x;
    unsigned long long ku = 0, kv = 0;
    {
        const int xl = tid / BG_TILE, yl = tid % BG_TILE;
        const int x = x0 + xl, y = y0 + yl;
        if (x < p.Nx && y < p.Ny) {
            const int l = (x + p.ng) * p.Nytot + y + p.ng;
            const double ux = w.ux[l], uy = w.uy[l];
            // separate statements: no multiply-add fusion, as the reference rounds
            const double uxx = ux * ux;
            const double uyy = uy * uy;
            const double ut = sqrt(1.0 + uxx + uyy);
            const double ax = ux / ut, ay = uy / ut;
            s[0][xl][yl] = exp(w.yT[l]);
            s[1][xl][yl] = ax;
            s[2][xl][yl] = ay;
            s[3][xl][yl] = w.n[l];
            ku = bg_key(hypot(ux, uy));
            kv = bg_key(hypot(ax, ay));
        }
    }
    __syncthreads();
    {
        const int yl = tid / BG_TILE, xl = tid % BG_TILE;
        const int x = x0 + xl, y = y0 + yl;
        if (x < p.Nx && y < p.Ny) {
            const int64_t o = off + x + (int64_t)p.Nx * y;
            T[o] = s[0][xl][yl];
            vx[o] = s[1][xl][yl];
            vy[o] = s[2][xl][yl];
            if (n) n[o] = s[3][xl][yl];
        }
    }
    // one atomic per block at most, and none once the block's maximum is not above
    // what the scalars already hold (they only grow, so a stale read is safe):
    // thousands of atomics on one address would cost ten times the slice itself
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long a = __shfl_xor(ku, o, 64), b = __shfl_xor(kv, o, 64);
        ku = a > ku ? a : ku;
        kv = b > kv ? b : kv;
    }
    __shared__ unsigned long long red[2][BG_TILE * BG_TILE / 64];
    if ((tid & 63) == 0) { red[0][tid >> 6] = ku; red[1][tid >> 6] = kv; }
    __syncthreads();
    if (tid == 0) {
        for (int j = 1; j < BG_TILE * BG_TILE / 64; ++j) {
            ku = red[0][j] > ku ? red[0][j] : ku;
            kv = red[1][j] > kv ? red[1][j] : kv;
        }
        if (ku > __atomic_load_n(&sc->bg_umax_bits, __ATOMIC_RELAXED)) atomicMax(&sc->bg_umax_bits, ku);
        if (kv > __atomic_load_n(&sc->bg_vmax_bits, __ATOMIC_RELAXED)) atomicMax(&sc->bg_vmax_bits, kv);
    }
}

hipError_t launch_background(Ctx& c, int64_t off, double* T, double* vx, double* vy, double* n) {
    const int nb = ((c.p.Nx + BG_TILE - 1) / BG_TILE) * ((c.p.Ny + BG_TILE - 1) / BG_TILE);
    hipLaunchKernelGGL(k_background, dim3(nb), dim3(BG_TILE * BG_TILE), 0, c.stream, c.p, c.w, off, T, vx, vy, n,
                       c.sc);
    return hipGetLastError();
}

hipError_t launch_layout(Ctx& c, const double* src, double* dst, int to_soa) {
    hipLaunchKernelGGL(k_layout, dim3(nblk((int64_t)c.p.Ntot * c.p.nv, 256)), dim3(256), 0, c.stream, src, dst,
                       c.p, to_soa);
    return hipGetLastError();
}

}  // namespace fivo
