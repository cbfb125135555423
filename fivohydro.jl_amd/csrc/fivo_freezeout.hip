// =============================================================================
// fivo_freezeout.hip -- the freeze-out surface tau = tau_fo(x, y) and its
// Cooper-Frye pion yield (C ABI fivo_freezeout_*, freezeout2d.jl:68-192).
//
// Three sweeps, all over the interior cells of one context:
//   k_fo_record  one thread per cell: the last downward crossing of T_fo,
//                linearly interpolated between two samples (record!,
//                freezeout2d.jl:97-119); re-heated cells are counted.
//   k_cf_count / k_cf_scan / k_cf_write
//                the cells that enter the integral (a frozen tau_fo and finite
//                gradients, freezeout2d.jl:160-169) in increasing cell order:
//                per-workgroup counts, their prefix sum, then each workgroup
//                writes its cells at its offset; the sum below therefore runs
//                in the same order on every call.
//   k_cf_sum     a fixed grid of CF_NB workgroups striding over batches of
//                (cell, p_T) pairs. exp(-E*/T) splits into a factor of eta and
//                one of phi, so a batch puts n_eta + n_phi exponentials per pair
//                into LDS and every (cell, p_T, eta, phi) point costs a product,
//                a division (the Bose factor) and a few FMAs. FP64 throughout;
//                bound by FP64 issue, not memory (the surface is a few hundred
//                KB and stays in L2).
//   k_cf_final   one workgroup: the per-workgroup partials in a fixed order.
// Every reduction has a fixed shape, so the yield is bit-reproducible.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>
#include <vector>
#include "fivo_ctx.h"

namespace fivo {

constexpr int CF_NB = 1024;     // workgroups of the Cooper-Frye sum
constexpr int CF_BS = 256;      // threads per workgroup of the sum, the list and the final pass
constexpr int CF_PER = 4;       // cells per thread of k_cf_count / k_cf_write
constexpr int CF_CHUNK = CF_BS * CF_PER;
constexpr int CF_PAIRS_MAX = 32;            // (cell, p_T) pairs per batch of k_cf_sum
constexpr size_t CF_LDS_BUDGET = 32 * 1024; // dynamic LDS of one k_cf_sum workgroup (several fit a CU)
constexpr int CF_PK = 10;                   // per-pair constants in LDS (9 used, even for alignment)
constexpr int CF_NMAX = 256;    // largest n_eta, n_pt, n_phi
constexpr double HBARC = 0.1973269804;   // GeV fm (freezeout2d.jl:59)

struct CfScratch {
    unsigned long long refroze;  // re-heating events
    int nlist;                   // cells on the list of the last k_cf_write
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
    int* bcnt = nullptr;         // per-workgroup counts, then offsets, of the list
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

// The cells 1..nx-2 x 1..ny-2 that enter the integral, in increasing k, in three
// launches: each workgroup counts the cells of its run of CF_CHUNK (q = its cell
// number in the interior-minus-rim order), one workgroup turns the counts into
// offsets, and each workgroup writes its run's cells at its offset.
__device__ __forceinline__ int cf_k(int q, int ny) {
    const int ci = ny - 2;
    return (1 + q / ci) * ny + 1 + q % ci;
}

__device__ __forceinline__ int block_exclusive_scan(int v, int* s_wave, int* total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    if (lane == 63) s_wave[wv] = inc;
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        if (w < wv) base += s_wave[w];
        tot += s_wave[w];
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

__global__ __launch_bounds__(CF_BS) void k_cf_count(const double* __restrict__ tfo, int nx, int ny, double dx,
                                                    double dy, int* __restrict__ bcnt) {
    __shared__ int s_wave[CF_BS / 64];
    const int m = (nx - 2) * (ny - 2);
    const int q0 = blockIdx.x * CF_CHUNK + threadIdx.x * CF_PER;
    int c = 0;
    for (int q = q0; q < min(m, q0 + CF_PER); ++q) c += cf_cell_ok(tfo, cf_k(q, ny), ny, dx, dy) ? 1 : 0;
    int tot;
    block_exclusive_scan(c, s_wave, &tot);
    if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

// exclusive offsets of nb counts in place, and their total
__global__ __launch_bounds__(1024) void k_cf_scan(int* __restrict__ bcnt, int nb, int* nlist) {
    __shared__ int s_wave[16];
    const int per = (nb + 1023) / 1024;
    const int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
    int c = 0;
    for (int b = b0; b < b1; ++b) c += bcnt[b];
    int tot;
    int off = block_exclusive_scan(c, s_wave, &tot);
    for (int b = b0; b < b1; ++b) {
        const int v = bcnt[b];
        bcnt[b] = off;
        off += v;
    }
    if (threadIdx.x == 0) *nlist = tot;
}

__global__ __launch_bounds__(CF_BS) void k_cf_write(const double* __restrict__ tfo, int nx, int ny, double dx,
                                                    double dy, const int* __restrict__ boff, int* __restrict__ list) {
    __shared__ int s_wave[CF_BS / 64];
    const int m = (nx - 2) * (ny - 2);
    const int q0 = blockIdx.x * CF_CHUNK + threadIdx.x * CF_PER;
    int c = 0;
    for (int q = q0; q < min(m, q0 + CF_PER); ++q) c += cf_cell_ok(tfo, cf_k(q, ny), ny, dx, dy) ? 1 : 0;
    int tot;
    int pos = boff[blockIdx.x] + block_exclusive_scan(c, s_wave, &tot);
    for (int q = q0; q < min(m, q0 + CF_PER); ++q) {
        const int k = cf_k(q, ny);
        if (cf_cell_ok(tfo, k, ny, dx, dy)) list[pos++] = k;
    }
}

__device__ __forceinline__ double wave_sum64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// 1/d for d in [1e-8, 1]: the hardware estimate and two Newton-Raphson steps
// r <- r + r (1 - d r) (each squares the relative error), no scaling for
// denormals or infinities since d stays far from both.
__device__ __forceinline__ double recip_nr(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// pion_yield's sum (freezeout2d.jl:160-187). exp(-E*/T) factorises,
//   exp(-(u^tau m_T cosh eta - u.p) / T) = exp(-u^tau m_T cosh eta / T) exp(u.p / T),
// into a factor of (cell, p_T, eta) and one of (cell, p_T, phi). A workgroup takes P
// (cell, p_T) pairs at a time: their constants and the phi factors into LDS, then
// every (pair, eta) task takes its eta factor and sums its n_phi points, so each
// point costs a product, a reciprocal (the Bose factor 1/(exp(x) - 1) =
// fB/(1 - fB)) and a few FMAs, and a pair n_eta + n_phi exponentials instead of
// n_eta n_phi. The same terms as the reference's loops, grouped differently:
//   w = tau (m_T cosh eta - grad tau_fo . p) p_T dp_T dphi deta dx dy
//     = W0(eta) - Gw(phi),
//   x = a(eta) - b(phi) > 1e-8  <=>  b < a - 1e-8,
// and the positive-dSigma sum is the total plus the negative share (w and
// dSigma.p have the same sign, so max(-w, 0) picks the negative terms).
__global__ __launch_bounds__(CF_BS) void k_cf_sum(const double* __restrict__ tfo, const double* __restrict__ uxf,
                                                  const double* __restrict__ uyf, int ny, double dx, double dy,
                                                  const int* __restrict__ list, const int* __restrict__ nlist,
                                                  CfGrid G, int P, double* __restrict__ part) {
    extern __shared__ __align__(16) double lds[];
    __shared__ double s_red[3][CF_BS / 64];
    const int ne = G.n_eta, nq = G.n_phi;
    double* s_ch = lds;                  // cosh eta (an even number of slots)
    double* s_c = s_ch + ((ne + 1) & ~1); // cos phi
    double* s_s = s_c + nq;              // sin phi
    double* s_pt = s_s + nq;             // p_T
    double* s_mT = s_pt + G.n_pt;        // m_T
    double* pK = s_mT + G.n_pt;          // per pair: tau, m_T, p_T w, u^tau, u^x, u^y, gx, gy, p_T
    // per (pair, phi): exp(b), b = u.p / T (16-byte aligned: every section above
    // has an even number of doubles), then Gw = tau p_T w grad tau_fo . p
    double2* pBb = reinterpret_cast<double2*>(pK + CF_PK * P);
    double* pG = reinterpret_cast<double*>(pBb + P * nq);
    for (int e = threadIdx.x; e < ne; e += CF_BS) s_ch[e] = cosh(G.eta_lo + e * G.deta);
    for (int e = threadIdx.x; e < nq; e += CF_BS) {
        const double ph = e * G.dphi;
        s_c[e] = cos(ph);
        s_s[e] = sin(ph);
    }
    for (int e = threadIdx.x; e < G.n_pt; e += CF_BS) {
        const double pt = G.pt_lo + e * G.dpt;
        s_pt[e] = pt;
        s_mT[e] = sqrt(pt * pt + G.mass * G.mass);
    }
    const double iT = 1.0 / G.T_fo;
    const int64_t npair = (int64_t)(*nlist) * G.n_pt;
    const int64_t nbatch = (npair + P - 1) / P;
    double acc = 0.0, accB = 0.0, neg = 0.0;
    for (int64_t bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {     // the same trip count for the whole block
        const int64_t p0 = bt * P;
        const int np = (int)min((int64_t)P, npair - p0);
        __syncthreads();                                              // the last batch's tables are read
        for (int t = threadIdx.x; t < np; t += CF_BS) {
            const int64_t pp = p0 + t;
            const int j = (int)(pp / G.n_pt), ip = (int)(pp - (int64_t)j * G.n_pt);
            const int k = list[j];
            const double ux = uxf[k], uy = uyf[k];
            double* K = pK + CF_PK * t;
            K[0] = tfo[k]; K[1] = s_mT[ip]; K[2] = s_pt[ip] * G.wconst;
            K[3] = sqrt(1.0 + ux * ux + uy * uy); K[4] = ux; K[5] = uy;
            K[6] = grad1(tfo, k, ny, dx); K[7] = grad1(tfo, k, 1, dy); K[8] = s_pt[ip];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < np * nq; t += CF_BS) {
            const int pr = t / nq, q = t - pr * nq;
            const double* K = pK + CF_PK * pr;
            const double px = K[8] * s_c[q], py = K[8] * s_s[q];
            const double b = (K[4] * px + K[5] * py) * iT;
            pBb[t] = make_double2(exp(b), b);
            pG[t] = K[0] * K[2] * (K[6] * px + K[7] * py);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < np * ne; t += CF_BS) {
            const int pr = t / ne, e = t - pr * ne;
            const double* K = pK + CF_PK * pr;
            const double mTch = K[1] * s_ch[e];
            const double a = K[3] * mTch * iT;
            const double Ea = exp(-a), a8 = a - 1e-8;
            const double W0 = K[0] * K[2] * mTch;
            const double2* Bb = pBb + pr * nq;
            const double* Gw = pG + pr * nq;
#pragma unroll 2
            for (int q = 0; q < nq; ++q) {
                const double w = W0 - Gw[q];
                const double2 v = Bb[q];
                const double fB = Ea * v.x;
                const double fb = v.y < a8 ? fB * recip_nr(1.0 - fB) : 0.0;
                acc = fma(w, fB, acc);
                accB = fma(w, fb, accB);
                neg = fma(fmax(-w, 0.0), fB, neg);
            }
        }
    }
    const double v[3] = {wave_sum64(acc), wave_sum64(accB), wave_sum64(neg)};
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 3; ++q) s_red[q][wv] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < 3; ++q)
            for (int w = 0; w < CF_BS / 64; ++w) s[q] += s_red[q][w];
        double* o = part + 4 * blockIdx.x;
        o[0] = s[0];
        o[1] = s[1];
        o[2] = s[2];
        o[3] = s[0] + s[2];      // positive-dSigma sum
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

// workgroups of k_cf_count / k_cf_write over the (nx-2)(ny-2) candidate cells
inline int cf_list_blocks(int nx, int ny) { return ((nx - 2) * (ny - 2) + CF_CHUNK - 1) / CF_CHUNK; }

// (cell, p_T) pairs per batch of k_cf_sum and the dynamic LDS they take
inline void cf_batch(const CfGrid& G, int& P, size_t& lds) {
    const size_t tab = sizeof(double) * (((G.n_eta + 1) & ~1) + 2 * G.n_phi + 2 * G.n_pt);
    const size_t pair = sizeof(double) * (CF_PK + 3 * G.n_phi);
    P = (int)std::min<size_t>(CF_PAIRS_MAX, (CF_LDS_BUDGET - tab) / pair);
    lds = tab + (size_t)P * pair;
}

// The Cooper-Frye launches on stream st; scratch: list (>= (nx-2)(ny-2) ints),
// bcnt (cf_list_blocks ints), nlist, part (4 CF_NB doubles), out (3 doubles).
hipError_t launch_cooper_frye(const double* tfo, const double* uxf, const double* uyf, int nx, int ny, double dx,
                              double dy, const CfGrid& G, double pref, int* list, int* bcnt, int* nlist, double* part,
                              double* out, hipStream_t st) {
    const int nb = cf_list_blocks(nx, ny);
    int P;
    size_t lds;
    cf_batch(G, P, lds);
    hipLaunchKernelGGL(k_cf_count, dim3(nb), dim3(CF_BS), 0, st, tfo, nx, ny, dx, dy, bcnt);
    hipLaunchKernelGGL(k_cf_scan, dim3(1), dim3(1024), 0, st, bcnt, nb, nlist);
    hipLaunchKernelGGL(k_cf_write, dim3(nb), dim3(CF_BS), 0, st, tfo, nx, ny, dx, dy, bcnt, list);
    hipLaunchKernelGGL(k_cf_sum, dim3(CF_NB), dim3(CF_BS), lds, st, tfo, uxf, uyf, ny, dx, dy, list, nlist, G, P,
                       part);
    hipLaunchKernelGGL(k_cf_final, dim3(1), dim3(CF_BS), 0, st, part, pref, out);
    return hipGetLastError();
}

// ---- the entry point that allocates nothing ------------------------------------
// Its scratch -- the cell list, its length, the list's per-workgroup counts and
// the sum's per-workgroup partials -- is
// this file-scope device array, one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host the name is a
// stand-in, not a device address): it uses the address hipGetSymbolAddress
// returns on each device, kept in g_scratch_dev[device].
constexpr size_t NA_OFF_NLIST = 0;
constexpr size_t NA_OFF_PART = 256;
constexpr size_t NA_OFF_BCNT = NA_OFF_PART + sizeof(double) * 4 * CF_NB;
constexpr size_t NA_BCNT_MAX = (FIVO_CF_NOALLOC_MAX_CELLS + CF_CHUNK - 1) / CF_CHUNK;
constexpr size_t NA_OFF_LIST = NA_OFF_BCNT + ((sizeof(int) * NA_BCNT_MAX + 255) & ~(size_t)255);
constexpr size_t NA_BYTES = NA_OFF_LIST + sizeof(int) * (size_t)FIVO_CF_NOALLOC_MAX_CELLS;
static_assert(NA_OFF_BCNT % 256 == 0 && NA_OFF_LIST % 256 == 0, "scratch sections stay 256-byte aligned");

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
    const size_t bb = up(sizeof(int) * (size_t)cf_list_blocks(f->nx, f->ny));
    hipError_t e = hipMalloc(&f->pool, 4 * bd + bu + bl + bs + bb);
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
    f->bcnt = reinterpret_cast<int*>(p + 4 * bd + bu + bl + bs);
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
    // (k_cf_count the counts, k_cf_scan their offsets and the list length,
    // k_cf_write the list, k_cf_sum every partial), so nothing is cleared and
    // the previous call's contents do not matter
    const hipError_t e = launch_cooper_frye(tau_fo, ux, uy, nx, ny, dx, dy, G, pref,
                                            reinterpret_cast<int*>(sp + NA_OFF_LIST),
                                            reinterpret_cast<int*>(sp + NA_OFF_BCNT),
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
    FOCHK(&c, launch_cooper_frye(f.tfo, f.uxf, f.uyf, f.nx, f.ny, f.dx, f.dy, G, pref, f.list, f.bcnt, &f.sc->nlist,
                                 f.sc->part, f.sc->out, c.stream));
    FOCHK(&c, hipMemcpyAsync(out, f.sc->out, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FOCHK(&c, hipStreamSynchronize(c.stream));
    return FIVO_OK;
}

}  // extern "C"
