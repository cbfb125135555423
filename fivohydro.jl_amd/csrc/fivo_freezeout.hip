// =============================================================================
// fivo_freezeout.hip -- the freeze-out surface tau = tau_fo(x, y) and its
// Cooper-Frye pion yield (C ABI fivo_freezeout_*, freezeout2d.jl:68-192).
//
// Three sweeps, all over the interior cells of one context:
//   k_fo_record  one thread per cell: the last downward crossing of T_fo,
//                linearly interpolated between two samples (record!,
//                freezeout2d.jl:97-119); re-heated cells are counted.
//   k_cf_list    one workgroup: the cells that enter the integral (a frozen
//                tau_fo and finite gradients, freezeout2d.jl:160-169) in
//                increasing cell order, by a block-wide prefix sum, so that
//                the sum below runs in the same order on every call.
//   k_cf_sum     a fixed grid of CF_NB workgroups striding over the items
//                (cell, p_T, phi); each item loops over eta. FP64 throughout;
//                the bound is the FP64 rate of two exponentials-worth of work per
//                (cell, p_T, phi, eta) point, not memory (the surface is a few
//                hundred KB and stays in L2).
//   k_cf_final   one workgroup: the per-workgroup partials in a fixed order.
// Every reduction has a fixed shape, so the yield is bit-reproducible.
// =============================================================================
#include <cmath>
#include <limits>
#include <mutex>
#include <vector>
#include "fivo_ctx.h"

namespace fivo {

constexpr int CF_NB = 1024;     // workgroups of the Cooper-Frye sum
constexpr int CF_BS = 256;      // threads per workgroup of the sum and the final pass
constexpr int CF_LIST_BS = 1024;
constexpr int CF_NMAX = 256;    // largest n_eta, n_pt, n_phi
constexpr double HBARC = 0.1973269804;   // GeV fm (freezeout2d.jl:59)

struct CfScratch {
    unsigned long long refroze;  // re-heating events
    int nlist;                   // cells on the list of the last k_cf_list
    int _pad;
    double out[4];
    double part[4 * CF_NB];      // per-workgroup (Boltzmann, Bose, negative, positive)
};

struct CfGrid {
    int n_eta, n_pt, n_phi;
    double eta_lo, deta, pt_lo, dpt, dphi, mass, T_fo;
    double wconst;               // dpt * dphi * deta * dx * dy
};

struct FreezeOut {
    Ctx* c = nullptr;
    double T_fo = 0.156;
    double tau_prev = std::numeric_limits<double>::quiet_NaN();
    int nx = 0, ny = 0, ng = 0, Nytot = 0;
    double dx = 0.0, dy = 0.0;
    void* pool = nullptr;
    double *tfo = nullptr, *uxf = nullptr, *uyf = nullptr, *Tprev = nullptr;
    uint8_t* above = nullptr;
    int* list = nullptr;
    CfScratch* sc = nullptr;
};

namespace {

__device__ __forceinline__ double dnan() { return __builtin_nan(""); }

// record! (freezeout2d.jl:97-119) for one interior cell k = a*ny + b
__global__ void k_fo_record(const double* __restrict__ yT, const double* __restrict__ ux,
                            const double* __restrict__ uy, int nx, int ny, int ng, int Nytot, double T_fo,
                            double tau, double tau_prev, double* __restrict__ tfo, double* __restrict__ uxf,
                            double* __restrict__ uyf, double* __restrict__ Tprev, uint8_t* __restrict__ above,
                            unsigned long long* refroze) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nx * ny) return;
    const int a = k / ny, b = k - a * ny;
    const int i = (ng + a) * Nytot + ng + b;
    const double T = exp(yT[i]);
    const double tf = tfo[k];
    if (T > T_fo) {                                   // still (or again) hot
        if (!isnan(tf)) atomicAdd(refroze, 1ull);
        above[k] = 1;
        tfo[k] = dnan();
    } else if (above[k] && isnan(tf)) {               // the crossing, interpolated
        const double Tp = Tprev[k];
        tfo[k] = (isnan(Tp) || Tp <= T_fo) ? tau : tau_prev + (tau - tau_prev) * (Tp - T_fo) / (Tp - T);
        uxf[k] = ux[i];
        uyf[k] = uy[i];
    }
    Tprev[k] = T;
}

// grad1 (freezeout2d.jl:131-141): central where both neighbours froze, one-sided
// where one did, NaN where neither did. s = the neighbour stride.
__device__ __forceinline__ double grad1(const double* __restrict__ t, int k, int s, double h) {
    const double lo = t[k - s], hi = t[k + s], c = t[k];
    if (!isnan(lo) && !isnan(hi)) return (hi - lo) / (2.0 * h);
    if (!isnan(hi)) return (hi - c) / h;
    if (!isnan(lo)) return (c - lo) / h;
    return dnan();
}

__device__ __forceinline__ bool cf_cell_ok(const double* __restrict__ t, int k, int ny, double dx, double dy) {
    return !isnan(t[k]) && !isnan(grad1(t, k, ny, dx)) && !isnan(grad1(t, k, 1, dy));
}

// The cells 1..nx-2 x 1..ny-2 that enter the integral, in increasing k. One
// workgroup: each thread takes a contiguous run of cells, counts, an exclusive
// prefix sum over the threads places its run, and it writes the run.
__global__ __launch_bounds__(CF_LIST_BS) void k_cf_list(const double* __restrict__ tfo, int nx, int ny, double dx,
                                                        double dy, int* __restrict__ list, int* nlist) {
    __shared__ int s_wave[CF_LIST_BS / 64];
    const int ci = ny - 2, m = (nx - 2) * ci;
    const int t = threadIdx.x;
    const int per = (m + CF_LIST_BS - 1) / CF_LIST_BS;
    const int q0 = t * per, q1 = min(m, q0 + per);
    int cnt = 0;
    for (int q = q0; q < q1; ++q) {
        const int k = (1 + q / ci) * ny + 1 + q % ci;
        cnt += cf_cell_ok(tfo, k, ny, dx, dy) ? 1 : 0;
    }
    // inclusive scan within the wave, then over the wave totals
    const int lane = t & 63, wv = t >> 6;
    int inc = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) s_wave[wv] = inc;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wv; ++w) base += s_wave[w];
    int pos = base + inc - cnt;
    for (int q = q0; q < q1; ++q) {
        const int k = (1 + q / ci) * ny + 1 + q % ci;
        if (cf_cell_ok(tfo, k, ny, dx, dy)) list[pos++] = k;
    }
    if (t == CF_LIST_BS - 1) *nlist = pos;
}

__device__ __forceinline__ double wave_sum64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// pion_yield's sum (freezeout2d.jl:160-187) over items (cell, p_T, phi), each over eta
__global__ __launch_bounds__(CF_BS) void k_cf_sum(const double* __restrict__ tfo, const double* __restrict__ uxf,
                                                  const double* __restrict__ uyf, int ny, double dx, double dy,
                                                  const int* __restrict__ list, const int* __restrict__ nlist,
                                                  CfGrid G, double* __restrict__ part) {
    __shared__ double s_ch[CF_NMAX], s_c[CF_NMAX], s_s[CF_NMAX], s_pt[CF_NMAX], s_mT[CF_NMAX];
    __shared__ double s_red[4][CF_BS / 64];
    for (int e = threadIdx.x; e < G.n_eta; e += CF_BS) s_ch[e] = cosh(G.eta_lo + e * G.deta);
    for (int e = threadIdx.x; e < G.n_phi; e += CF_BS) {
        const double ph = e * G.dphi;
        s_c[e] = cos(ph);
        s_s[e] = sin(ph);
    }
    for (int e = threadIdx.x; e < G.n_pt; e += CF_BS) {
        const double pt = G.pt_lo + e * G.dpt;
        s_pt[e] = pt;
        s_mT[e] = sqrt(pt * pt + G.mass * G.mass);
    }
    __syncthreads();
    const int64_t per_cell = (int64_t)G.n_pt * G.n_phi;
    const int64_t nitem = (int64_t)(*nlist) * per_cell;
    double acc = 0.0, accB = 0.0, neg = 0.0, pos = 0.0;
    for (int64_t it = (int64_t)blockIdx.x * CF_BS + threadIdx.x; it < nitem; it += (int64_t)CF_NB * CF_BS) {
        const int j = (int)(it / per_cell);
        const int r = (int)(it - (int64_t)j * per_cell);
        const int ip = r / G.n_phi, iq = r - ip * G.n_phi;
        const int k = list[j];
        const double gx = grad1(tfo, k, ny, dx), gy = grad1(tfo, k, 1, dy);
        const double tau = tfo[k], ux = uxf[k], uy = uyf[k];
        const double ut = sqrt(1.0 + ux * ux + uy * uy);
        const double pt = s_pt[ip], mT = s_mT[ip];
        const double px = pt * s_c[iq], py = pt * s_s[iq];
        const double gp = gx * px + gy * py, up = ux * px + uy * py;
        const double wpt = pt * G.wconst;
        for (int e = 0; e < G.n_eta; ++e) {
            const double ch = s_ch[e];
            const double dSp = tau * (mT * ch - gp);
            const double x = (ut * mT * ch - up) / G.T_fo;
            const double fB = exp(-x);
            // 1/(exp(x) - 1) = fB/(1 - fB)
            const double fb = x > 1e-8 ? fB / (1.0 - fB) : 0.0;
            const double w = dSp * wpt;
            acc += w * fB;
            accB += w * fb;
            if (dSp < 0.0) neg -= w * fB;
            else pos += w * fB;
        }
    }
    const double v[4] = {wave_sum64(acc), wave_sum64(accB), wave_sum64(neg), wave_sum64(pos)};
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 4; ++q) s_red[q][wv] = v[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int w = 0; w < CF_BS / 64; ++w) s += s_red[threadIdx.x][w];
        part[4 * blockIdx.x + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(CF_BS) void k_cf_final(const double* __restrict__ part, double pref,
                                                    double* __restrict__ out) {
    __shared__ double s_red[4][CF_BS / 64];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < CF_NB; b += CF_BS)
        for (int q = 0; q < 4; ++q) v[q] += part[4 * b + q];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < 4; ++q) {
        v[q] = wave_sum64(v[q]);
        if (lane == 0) s_red[q][wv] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[4];
        for (int q = 0; q < 4; ++q) {
            s[q] = 0.0;
            for (int w = 0; w < CF_BS / 64; ++w) s[q] += s_red[q][w];
        }
        out[0] = pref * s[0];
        out[1] = pref * s[1];
        out[2] = s[2] / fmax(s[3] + s[2], 1e-30);
    }
}

int fo_fail(Ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg; else set_last_error(msg);
    return code;
}
#define FOCHK(c, x)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            return fo_fail((c), FIVO_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

// the Cooper-Frye grid of o (freezeout2d.jl:146-149) and the measure constant
int cf_grid(const fivo_cf_opts* o, double T_fo, double dx, double dy, CfGrid& G, double& pref, std::string& err) {
    if (!o) { err = "pion_yield: null options"; return FIVO_ERR_ARG; }
    if (o->n_eta < 2 || o->n_pt < 2 || o->n_phi < 2 || o->n_eta > CF_NMAX || o->n_pt > CF_NMAX ||
        o->n_phi > CF_NMAX) {
        err = "pion_yield: n_eta, n_pt and n_phi must lie in 2..256";
        return FIVO_ERR_ARG;
    }
    if (!(o->eta_max > 0.0) || !(o->pt_hi > o->pt_lo) || !(o->pt_lo >= 0.0) || !(o->mass >= 0.0) ||
        !(T_fo > 0.0)) {
        err = "pion_yield: need eta_max > 0, 0 <= pt_lo < pt_hi, mass >= 0 and T_fo > 0";
        return FIVO_ERR_ARG;
    }
    G.n_eta = o->n_eta;
    G.n_pt = o->n_pt;
    G.n_phi = o->n_phi;
    G.eta_lo = -o->eta_max;
    G.deta = 2.0 * o->eta_max / (o->n_eta - 1);
    G.pt_lo = o->pt_lo;
    G.dpt = (o->pt_hi - o->pt_lo) / (o->n_pt - 1);
    G.dphi = 2.0 * M_PI / o->n_phi;
    G.mass = o->mass;
    G.T_fo = T_fo;
    G.wconst = G.dpt * G.dphi * G.deta * dx * dy;
    const double t = 2.0 * M_PI * HBARC;
    pref = o->yield_factor / (t * t * t);
    return FIVO_OK;
}

// The three Cooper-Frye launches on stream st; scratch: list (>= (nx-2)(ny-2)
// ints), nlist, part (4 CF_NB doubles), out (3 doubles).
hipError_t launch_cooper_frye(const double* tfo, const double* uxf, const double* uyf, int nx, int ny, double dx,
                              double dy, const CfGrid& G, double pref, int* list, int* nlist, double* part,
                              double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_cf_list, dim3(1), dim3(CF_LIST_BS), 0, st, tfo, nx, ny, dx, dy, list, nlist);
    hipLaunchKernelGGL(k_cf_sum, dim3(CF_NB), dim3(CF_BS), 0, st, tfo, uxf, uyf, ny, dx, dy, list, nlist, G, part);
    hipLaunchKernelGGL(k_cf_final, dim3(1), dim3(CF_BS), 0, st, part, pref, out);
    return hipGetLastError();
}

// ---- the entry point that allocates nothing ------------------------------------
// Its scratch -- the cell list, its length and the per-workgroup partials -- is
// this file-scope device array, one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host the name is a
// stand-in, not a device address): it uses the address hipGetSymbolAddress
// returns on each device, kept in g_scratch_dev[device].
constexpr size_t NA_OFF_NLIST = 0;
constexpr size_t NA_OFF_PART = 256;
constexpr size_t NA_OFF_LIST = NA_OFF_PART + sizeof(double) * 4 * CF_NB;
constexpr size_t NA_BYTES = NA_OFF_LIST + sizeof(int) * (size_t)FIVO_CF_NOALLOC_MAX_CELLS;
static_assert(NA_OFF_LIST % 256 == 0, "scratch sections stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[NA_MAX_DEVICES] = {};

// the device address of `scratch` on the current device, looked up once per device
int scratch_on_device(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_last_error(std::string("pion_yield_noalloc: hipGetDevice: ") + hipGetErrorString(e));
        return FIVO_ERR_HIP;
    }
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        set_last_error("pion_yield_noalloc: device index beyond 63");
        return FIVO_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(std::string("pion_yield_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return FIVO_ERR_HIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_dev[dev]);
    return FIVO_OK;
}
}  // namespace

}  // namespace fivo

using namespace fivo;

extern "C" {

int fivo_freezeout_create(fivo_ctx* ctx, double T_fo, fivo_freezeout** out) {
    if (!out) return FIVO_ERR_ARG;
    *out = nullptr;
    if (!ctx) return fo_fail(nullptr, FIVO_ERR_ARG, "freezeout_create: null context");
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (!(T_fo > 0.0)) return fo_fail(&c, FIVO_ERR_ARG, "freezeout_create: T_fo must be > 0");
    if (c.gd.nx < 3 || c.gd.ny < 3) return fo_fail(&c, FIVO_ERR_ARG, "freezeout_create: need nx, ny >= 3");
    FreezeOut* f = new FreezeOut();
    f->c = &c;
    f->T_fo = T_fo;
    f->nx = c.gd.nx;
    f->ny = c.gd.ny;
    f->ng = c.gd.nghost;
    f->Nytot = c.gd.ny + 2 * c.gd.nghost;
    f->dx = 2.0 * c.gd.xmax / c.gd.nx;
    f->dy = 2.0 * c.gd.ymax / c.gd.ny;
    const size_t n = (size_t)f->nx * f->ny, nl = (size_t)(f->nx - 2) * (f->ny - 2);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bd = up(n * sizeof(double)), bu = up(n), bl = up(nl * sizeof(int)), bs = up(sizeof(CfScratch));
    hipError_t e = hipMalloc(&f->pool, 4 * bd + bu + bl + bs);
    if (e != hipSuccess) {
        delete f;
        return fo_fail(&c, FIVO_ERR_HIP, std::string("freezeout_create: hipMalloc: ") + hipGetErrorString(e));
    }
    char* p = static_cast<char*>(f->pool);
    f->tfo = reinterpret_cast<double*>(p);
    f->uxf = reinterpret_cast<double*>(p + bd);
    f->uyf = reinterpret_cast<double*>(p + 2 * bd);
    f->Tprev = reinterpret_cast<double*>(p + 3 * bd);
    f->above = reinterpret_cast<uint8_t*>(p + 4 * bd);
    f->list = reinterpret_cast<int*>(p + 4 * bd + bu);
    f->sc = reinterpret_cast<CfScratch*>(p + 4 * bd + bu + bl);
    // tau_fo = NaN, flow 0, Tprev = NaN, nothing hot (freezeout2d.jl:87-91)
    std::vector<double> nan(n, std::numeric_limits<double>::quiet_NaN());
    if ((e = hipMemcpyAsync(f->tfo, nan.data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream)) != hipSuccess ||
        (e = hipMemcpyAsync(f->Tprev, nan.data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream)) != hipSuccess ||
        (e = hipMemsetAsync(f->uxf, 0, n * sizeof(double), c.stream)) != hipSuccess ||
        (e = hipMemsetAsync(f->uyf, 0, n * sizeof(double), c.stream)) != hipSuccess ||
        (e = hipMemsetAsync(f->above, 0, n, c.stream)) != hipSuccess ||
        (e = hipMemsetAsync(f->sc, 0, sizeof(CfScratch), c.stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c.stream)) != hipSuccess) {
        (void)hipFree(f->pool);
        delete f;
        return fo_fail(&c, FIVO_ERR_HIP, std::string("freezeout_create: ") + hipGetErrorString(e));
    }
    *out = reinterpret_cast<fivo_freezeout*>(f);
    return FIVO_OK;
}

void fivo_freezeout_destroy(fivo_freezeout* fp) {
    if (!fp) return;
    FreezeOut* f = reinterpret_cast<FreezeOut*>(fp);
    if (f->c) (void)hipStreamSynchronize(f->c->stream);
    (void)hipFree(f->pool);
    delete f;
}

int fivo_freezeout_record(fivo_freezeout* fp, double tau) {
    if (!fp) return FIVO_ERR_ARG;
    FreezeOut& f = *reinterpret_cast<FreezeOut*>(fp);
    Ctx& c = *f.c;
    const int n = f.nx * f.ny;
    hipLaunchKernelGGL(k_fo_record, dim3((n + 255) / 256), dim3(256), 0, c.stream, c.w.yT, c.w.ux, c.w.uy, f.nx,
                       f.ny, f.ng, f.Nytot, f.T_fo, tau, f.tau_prev, f.tfo, f.uxf, f.uyf, f.Tprev, f.above,
                       &f.sc->refroze);
    FOCHK(&c, hipGetLastError());
    FOCHK(&c, hipStreamSynchronize(c.stream));
    f.tau_prev = tau;
    return FIVO_OK;
}

int fivo_freezeout_surface(fivo_freezeout* fp, double* tau_fo, double* ux, double* uy, int64_t* counts) {
    if (!fp) return FIVO_ERR_ARG;
    FreezeOut& f = *reinterpret_cast<FreezeOut*>(fp);
    Ctx& c = *f.c;
    const size_t n = (size_t)f.nx * f.ny;
    std::vector<double> t(n);
    std::vector<uint8_t> ab(n);
    unsigned long long nre = 0;
    FOCHK(&c, hipMemcpyAsync(t.data(), f.tfo, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FOCHK(&c, hipMemcpyAsync(ab.data(), f.above, n, hipMemcpyDeviceToHost, c.stream));
    FOCHK(&c, hipMemcpyAsync(&nre, &f.sc->refroze, sizeof(nre), hipMemcpyDeviceToHost, c.stream));
    if (ux) FOCHK(&c, hipMemcpyAsync(ux, f.uxf, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (uy) FOCHK(&c, hipMemcpyAsync(uy, f.uyf, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FOCHK(&c, hipStreamSynchronize(c.stream));
    if (tau_fo) std::copy(t.begin(), t.end(), tau_fo);
    if (counts) {
        int64_t nhot = 0, nfroze = 0;
        for (size_t k = 0; k < n; ++k) {
            nhot += ab[k] ? 1 : 0;
            nfroze += std::isnan(t[k]) ? 0 : 1;
        }
        counts[0] = nhot;
        counts[1] = nfroze;
        counts[2] = (int64_t)nre;
    }
    return FIVO_OK;
}

int fivo_freezeout_set_surface(fivo_freezeout* fp, const double* tau_fo, const double* ux, const double* uy) {
    if (!fp) return FIVO_ERR_ARG;
    FreezeOut& f = *reinterpret_cast<FreezeOut*>(fp);
    Ctx& c = *f.c;
    if (!tau_fo || !ux || !uy) return fo_fail(&c, FIVO_ERR_ARG, "freezeout_set_surface: null array");
    const size_t n = (size_t)f.nx * f.ny;
    FOCHK(&c, hipMemcpyAsync(f.tfo, tau_fo, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FOCHK(&c, hipMemcpyAsync(f.uxf, ux, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FOCHK(&c, hipMemcpyAsync(f.uyf, uy, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    FOCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

int fivo_pion_yield_noalloc(const double* tau_fo, const double* ux, const double* uy, int32_t nx, int32_t ny,
                            double dx, double dy, double T_fo, const fivo_cf_opts* o, double* out, void* stream) {
    if (!tau_fo || !ux || !uy || !out) return fo_fail(nullptr, FIVO_ERR_ARG, "pion_yield_noalloc: null array");
    if (nx < 3 || ny < 3) return fo_fail(nullptr, FIVO_ERR_ARG, "pion_yield_noalloc: need nx, ny >= 3");
    if ((int64_t)nx * ny > FIVO_CF_NOALLOC_MAX_CELLS)
        return fo_fail(nullptr, FIVO_ERR_ARG,
                       "pion_yield_noalloc: nx*ny = " + std::to_string((int64_t)nx * ny) + " exceeds the scratch limit " +
                           std::to_string(FIVO_CF_NOALLOC_MAX_CELLS) + " cells");
    if (!(dx > 0.0) || !(dy > 0.0)) return fo_fail(nullptr, FIVO_ERR_ARG, "pion_yield_noalloc: need dx, dy > 0");
    CfGrid G;
    double pref;
    std::string err;
    int rc = cf_grid(o, T_fo, dx, dy, G, pref, err);
    if (rc != FIVO_OK) return fo_fail(nullptr, rc, err);
    unsigned char* sp = nullptr;
    if ((rc = scratch_on_device(&sp)) != FIVO_OK) return rc;
    // every scratch word the kernels read is written first in the same call
    // (k_cf_list writes the list and its length, k_cf_sum every partial), so
    // nothing is cleared and the previous call's contents do not matter
    const hipError_t e = launch_cooper_frye(tau_fo, ux, uy, nx, ny, dx, dy, G, pref,
                                            reinterpret_cast<int*>(sp + NA_OFF_LIST),
                                            reinterpret_cast<int*>(sp + NA_OFF_NLIST),
                                            reinterpret_cast<double*>(sp + NA_OFF_PART), out,
                                            static_cast<hipStream_t>(stream));
    if (e != hipSuccess)
        return fo_fail(nullptr, FIVO_ERR_HIP, std::string("pion_yield_noalloc: ") + hipGetErrorString(e));
    return FIVO_OK;
}

int fivo_freezeout_pion_yield(fivo_freezeout* fp, const fivo_cf_opts* o, double* out) {
    if (!fp) return FIVO_ERR_ARG;
    FreezeOut& f = *reinterpret_cast<FreezeOut*>(fp);
    Ctx& c = *f.c;
    if (!out) return fo_fail(&c, FIVO_ERR_ARG, "pion_yield: null output");
    CfGrid G;
    double pref;
    std::string err;
    const int rc = cf_grid(o, f.T_fo, f.dx, f.dy, G, pref, err);
    if (rc != FIVO_OK) return fo_fail(&c, rc, err);
    FOCHK(&c, launch_cooper_frye(f.tfo, f.uxf, f.uyf, f.nx, f.ny, f.dx, f.dy, G, pref, f.list, &f.sc->nlist,
                                 f.sc->part, f.sc->out, c.stream));
    FOCHK(&c, hipMemcpyAsync(out, f.sc->out, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FOCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

}  // extern "C"
