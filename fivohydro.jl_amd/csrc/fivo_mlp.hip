// =============================================================================
// fivo_mlp.hip -- training and evaluation of the learned Newton warm start
// (C ABI fivo_mlp_*): the pair collector, normalisation, the BF16-MFMA
// forward/backward kernel, Adam, and the packing of the weights into MFMA tiles.
//
// Training data are (features, targets) pairs of converged cells that the
// solver itself produced (fivo_mlp_collect after a primitive pass). Loss: the
// mean over samples of the summed squared error of the 6 normalised outputs.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#include "fivo_ctx.h"
#include "fivo_mlp.h"

namespace fivo {
using namespace fd;

struct Mlp {
    int dev = 0;
    hipStream_t stream = nullptr;
    float* params = nullptr;   // FP32 master parameters
    float* grad = nullptr;
    float* am = nullptr;       // Adam moments
    float* av = nullptr;
    s4* tiles = nullptr;       // packed BF16 A tiles
    float* norm = nullptr;     // MLP_NNORM
    float* data = nullptr;     // capacity x 12: features then targets
    int64_t capacity = 0;
    unsigned long long* count = nullptr;
    double* acc = nullptr;     // reductions: loss, statistics (64 doubles)
    double* part = nullptr;    // per-block partials of the statistics
    int normalized = 0;
    int64_t adam_t = 0;
    std::string err;
};

namespace {

// ---- packing the parameters into MFMA A tiles -----------------------------------
__global__ void k_mlp_pack(const float* __restrict__ prm, s4* __restrict__ tiles) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;     // one thread per (tile, lane)
    if (gid >= MLP_NTILES * 64) return;
    const int tile = gid >> 6, l = gid & 63;
    const int m = l & 15, g = l >> 4;
    float v[4];
    for (int e = 0; e < 4; ++e) {
        const int k = 4 * g + e;
        float a = 0.f;
        if (tile < MLP_T_W2) {                           // W1 rows 16t.., inputs k < 6
            const int t = tile - MLP_T_W1;
            if (k < MLP_IN) a = prm[MLP_OFF_W1 + (16 * t + m) * MLP_IN + k];
        } else if (tile < MLP_T_W3) {                    // W2[16u + m][16t + k]
            const int q = tile - MLP_T_W2, u = q >> 2, t = q & 3;
            a = prm[MLP_OFF_W2 + (16 * u + m) * MLP_H + 16 * t + k];
        } else if (tile < MLP_T_W3T) {                   // W3[m][16t + k], rows m < 6
            const int t = tile - MLP_T_W3;
            if (m < MLP_OUT) a = prm[MLP_OFF_W3 + m * MLP_H + 16 * t + k];
        } else if (tile < MLP_T_W2T) {                   // W3^T: A[m = j][k = o] = W3[o][16t + m]
            const int t = tile - MLP_T_W3T;
            if (k < MLP_OUT) a = prm[MLP_OFF_W3 + k * MLP_H + 16 * t + m];
        } else {                                         // W2^T: A[m = j_in][k = j_out] = W2[16u + k][16t + m]
            const int q = tile - MLP_T_W2T, t = q >> 2, u = q & 3;
            a = prm[MLP_OFF_W2 + (16 * u + k) * MLP_H + 16 * t + m];
        }
        v[e] = a;
    }
    tiles[gid] = pack4(v[0], v[1], v[2], v[3]);
}

// ---- pair collection from a context ---------------------------------------------
constexpr double MLP_PHI_TRAIN_MAX = 200.0;
__global__ __launch_bounds__(256) void k_mlp_collect(const double* __restrict__ U, Params p, DevWork w, double tau,
                                                     float* __restrict__ data, unsigned long long* count,
                                                     long long capacity) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.Nx * p.Ny) return;
    const int ix = p.ng + t / p.Ny, iy = p.ng + t % p.Ny, i = ix * p.Nytot + iy;
    if (!w.ok[i]) return;
    const double D = safe_div(U[i], tau), Sx = U[p.pitch + i], Sy = U[2 * p.pitch + i], E = U[3 * p.pitch + i];
    if (!(isf(D) && isf(Sx) && isf(Sy) && isf(E)) || E <= E_VAC) return;
    Diss d;
    d.nux = p.hasNu ? U[p.iNux * p.pitch + i] : 0.0;
    d.nuy = p.hasNu ? U[p.iNuy * p.pitch + i] : 0.0;
    d.Pi = p.hasPi ? U[p.iPi * p.pitch + i] : 0.0;
    d.pixx = p.hasShear ? U[p.iPixx * p.pitch + i] : 0.0;
    d.pixy = p.hasShear ? U[p.iPixy * p.pitch + i] : 0.0;
    d.piyy = p.hasShear ? U[p.iPiyy * p.pitch + i] : 0.0;
    d.pieta = 0.0;
    const MlpFrame fr = mlp_frame(Sx, Sy);
    double f[MLP_IN], g[MLP_OUT];
    mlp_features(D, Sx, Sy, E, d, fr, f);
    mlp_targets(w.yT[i], w.phi[i], w.ux[i], w.uy[i], w.e[i], w.P[i], fr, g);
    for (int k = 0; k < MLP_IN; ++k) if (!isf(f[k])) return;
    for (int k = 0; k < MLP_OUT; ++k) if (!isf(g[k])) return;
    // phi pinned at the cap by the staged charge fallback carries no information
    // about the state (any phi with a negligible density solves those cells)
    if (fabs(g[1]) > MLP_PHI_TRAIN_MAX) return;
    const unsigned long long s = atomicAdd(count, 1ull);
    if ((long long)s >= capacity) return;
    float* r = data + 12 * s;
    for (int k = 0; k < 6; ++k) { r[k] = (float)f[k]; r[6 + k] = (float)g[k]; }
}

// column sums and sums of squares of the 12 columns, per block (FP64)
__global__ __launch_bounds__(256) void k_mlp_stats(const float* __restrict__ data, long long n, double* part) {
    __shared__ double sh[256];
    for (int c = 0; c < 24; ++c) {
        double v = 0.0;
        for (long long s = blockIdx.x * 256LL + threadIdx.x; s < n; s += (long long)gridDim.x * 256) {
            const double x = data[12 * s + (c % 12)];
            v += c < 12 ? x : x * x;
        }
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[24 * blockIdx.x + c] = sh[0];
        __syncthreads();
    }
}
__global__ void k_mlp_normalize(float* __restrict__ data, long long n, const float* __restrict__ norm) {
    const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    float* r = data + 12 * s;
    for (int k = 0; k < 6; ++k) {
        r[k] = (r[k] - norm[k]) / norm[6 + k];
        r[6 + k] = (r[6 + k] - norm[12 + k]) / norm[18 + k];
    }
}

// ---- forward + backward over a set of samples: gradients and loss ------------------
// One wavefront per block; the wavefront takes `tiles_per_wave` tiles of 16
// samples. Sample s of the set is data row (a * s + c) mod n (a permutation of
// the rows when gcd(a, n) = 1): minibatches without moving data.
struct TrainArgs {
    const float* data;       // normalised rows
    long long n;             // rows in the set
    long long first;         // first sample of this call
    long long count;         // samples of this call
    long long a, c;          // row permutation
    int tiles_per_wave;
    float inv_count;         // 1 / count: the loss is a mean
};
__device__ __forceinline__ float bfr(float x) { return bf16_to_f(bf16_rne(x)); }

__global__ __launch_bounds__(64) void k_mlp_grad(TrainArgs A, const float* __restrict__ prm,
                                                 const s4* __restrict__ tiles, float* __restrict__ grad,
                                                 double* loss_acc) {
    // transposes through LDS: activations as [unit][sample] BF16
    __shared__ uint16_t sH1[64][16], sH2[64][16], sZ2[64][16], sZ1[64][16], sY[16][16], sX[16][16];
    const int l = threadIdx.x;
    const int g = l >> 4, col = l & 15, r0 = 4 * g;
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    f4 aw1[4], aw2[16], aw3[4];
    float db1[4][4], db2[4][4], db3[4];
    for (int t = 0; t < 4; ++t) { aw1[t] = zero; aw3[t] = zero; for (int e = 0; e < 4; ++e) { db1[t][e] = 0; db2[t][e] = 0; } }
    for (int q = 0; q < 16; ++q) aw2[q] = zero;
    for (int e = 0; e < 4; ++e) db3[e] = 0;
    double lsum = 0.0;
    auto rd = [&](uint16_t (*M)[16], int u) {       // A or B operand, rows 16u + col, samples 4g..4g+3
        s4 v;
        for (int e = 0; e < 4; ++e) v[e] = (short)M[16 * u + col][4 * g + e];
        return v;
    };
    for (int tt = 0; tt < A.tiles_per_wave; ++tt) {
        const long long s0 = A.first + ((long long)blockIdx.x * A.tiles_per_wave + tt) * 16;
        if (s0 >= A.first + A.count) break;
        const long long s = s0 + col;
        const bool valid = s < A.first + A.count;
        float fn[MLP_IN], tn[MLP_OUT];
        for (int k = 0; k < 6; ++k) { fn[k] = 0.f; tn[k] = 0.f; }
        if (valid) {
            const long long row = (A.a * (s % A.n) + A.c) % A.n;
            const float* r = A.data + 12 * row;
            for (int k = 0; k < 6; ++k) { fn[k] = r[k]; tn[k] = r[6 + k]; }
        }
        const s4 bx = mlp_feature_operand(fn);
        MlpFwd F;
        mlp_forward_tile(tiles, prm, bx, F);
        // dY = 2 (y - t) / count on rows < 6 of valid samples (D layout: rows r0+e, sample col)
        f4 dy = zero;
        for (int e = 0; e < 4; ++e) {
            const int o = r0 + e;
            float tv = 0.f;
            for (int q = 0; q < MLP_OUT; ++q) tv = (o == q) ? tn[q] : tv;
            if (valid && o < MLP_OUT) {
                const float d = F.y[e] - tv;
                lsum += (double)d * d;
                dy[e] = 2.0f * d * A.inv_count;
                db3[e] += dy[e];
            }
        }
        // LDS: H1, H2 (bf16 activations), dY, X (bf16, as the MFMA saw them)
        for (int t = 0; t < 4; ++t)
            for (int e = 0; e < 4; ++e) {
                sH1[16 * t + r0 + e][col] = (uint16_t)F.h1b[t][e];
                sH2[16 * t + r0 + e][col] = (uint16_t)F.h2b[t][e];
            }
        for (int e = 0; e < 4; ++e) {
            sY[r0 + e][col] = bf16_rne(dy[e]);
            sX[r0 + e][col] = (uint16_t)bx[e];
        }
        __syncthreads();
        // dW3[o][j] += sum_i dY[o][i] H2[j][i]
        {
            const s4 ay = rd(sY, 0);
            for (int u = 0; u < 4; ++u) aw3[u] = mfma16(ay, rd(sH2, u), aw3[u]);
        }
        // dH2 = W3^T dY, dZ2 = dH2 * gelu'(z2)
        const s4 bdy = pack4(dy[0], dy[1], dy[2], dy[3]);
        f4 dz2[4];
        for (int t = 0; t < 4; ++t) {
            f4 h = mfma16(tiles[(MLP_T_W3T + t) * 64 + l], bdy, zero);
            for (int e = 0; e < 4; ++e) { h[e] *= gelu_d(F.z2[t][e]); db2[t][e] += h[e]; }
            dz2[t] = h;
        }
        s4 bdz2[4];
        for (int t = 0; t < 4; ++t) {
            bdz2[t] = pack4(dz2[t][0], dz2[t][1], dz2[t][2], dz2[t][3]);
            for (int e = 0; e < 4; ++e) sZ2[16 * t + r0 + e][col] = (uint16_t)bdz2[t][e];
        }
        __syncthreads();
        // dW2[j_out][j_in] += sum_i dZ2[j_out][i] H1[j_in][i]
        for (int u = 0; u < 4; ++u) {
            const s4 a = rd(sZ2, u);
            for (int t = 0; t < 4; ++t) aw2[4 * u + t] = mfma16(a, rd(sH1, t), aw2[4 * u + t]);
        }
        // dH1 = W2^T dZ2, dZ1 = dH1 * gelu'(z1)
        for (int t = 0; t < 4; ++t) {
            f4 h = zero;
            for (int u = 0; u < 4; ++u) h = mfma16(tiles[(MLP_T_W2T + 4 * t + u) * 64 + l], bdz2[u], h);
            for (int e = 0; e < 4; ++e) { h[e] *= gelu_d(F.z1[t][e]); db1[t][e] += h[e]; }
            const s4 b = pack4(h[0], h[1], h[2], h[3]);
            for (int e = 0; e < 4; ++e) sZ1[16 * t + r0 + e][col] = (uint16_t)b[e];
        }
        __syncthreads();
        // dW1[j][k] += sum_i dZ1[j][i] X[k][i]
        {
            const s4 bxT = rd(sX, 0);
            for (int t = 0; t < 4; ++t) aw1[t] = mfma16(rd(sZ1, t), bxT, aw1[t]);
        }
        __syncthreads();
    }
    // bias gradients: sum over the 16 sample columns of each lane group
    auto colsum = [&](float v) {
        for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
        return v;
    };
    for (int t = 0; t < 4; ++t)
        for (int e = 0; e < 4; ++e) {
            const float b1 = colsum(db1[t][e]), b2 = colsum(db2[t][e]);
            if (col == 0) {
                atomicAdd(&grad[MLP_OFF_B1 + 16 * t + r0 + e], b1);
                atomicAdd(&grad[MLP_OFF_B2 + 16 * t + r0 + e], b2);
            }
        }
    for (int e = 0; e < 4; ++e) {
        const float b3 = colsum(db3[e]);
        if (col == 0 && r0 + e < MLP_OUT) atomicAdd(&grad[MLP_OFF_B3 + r0 + e], b3);
    }
    // weight gradients from the D-layout accumulators (rows r0+e, column col)
    for (int t = 0; t < 4; ++t)
        for (int e = 0; e < 4; ++e)
            if (col < MLP_IN) atomicAdd(&grad[MLP_OFF_W1 + (16 * t + r0 + e) * MLP_IN + col], aw1[t][e]);
    for (int u = 0; u < 4; ++u)
        for (int t = 0; t < 4; ++t)
            for (int e = 0; e < 4; ++e)
                atomicAdd(&grad[MLP_OFF_W2 + (16 * u + r0 + e) * MLP_H + 16 * t + col], aw2[4 * u + t][e]);
    for (int u = 0; u < 4; ++u)
        for (int e = 0; e < 4; ++e)
            if (r0 + e < MLP_OUT) atomicAdd(&grad[MLP_OFF_W3 + (r0 + e) * MLP_H + 16 * u + col], aw3[u][e]);
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (l == 0) atomicAdd(loss_acc, lsum);
}

// Adam on the master parameters; clears the gradient
__global__ void k_mlp_adam(float* __restrict__ prm, float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1c, float b2c) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= MLP_NPARAM) return;
    const float g = grad[k];
    const float mk = 0.9f * m[k] + 0.1f * g;
    const float vk = 0.999f * v[k] + 0.001f * g * g;
    m[k] = mk;
    v[k] = vk;
    prm[k] -= lr * (mk / b1c) / (sqrtf(vk / b2c) + 1e-8f);
    grad[k] = 0.f;
}
__global__ void k_mlp_fill(float* a, float v, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}
__global__ void k_mlp_zero_d(double* a, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = 0.0;
}

// forward only, normalised space: y[s][o] for s < n (testing and evaluation)
__global__ __launch_bounds__(64) void k_mlp_forward(const float* __restrict__ x, long long n,
                                                    const float* __restrict__ prm, const s4* __restrict__ tiles,
                                                    float* __restrict__ y) {
    const int l = threadIdx.x, col = l & 15, r0 = 4 * (l >> 4);
    const long long s = blockIdx.x * 16LL + col;
    float fn[MLP_IN];
    for (int k = 0; k < MLP_IN; ++k) fn[k] = s < n ? x[MLP_IN * s + k] : 0.f;
    MlpFwd F;
    mlp_forward_tile(tiles, prm, mlp_feature_operand(fn), F);
    for (int e = 0; e < 4; ++e)
        if (s < n && r0 + e < MLP_OUT) y[MLP_OUT * s + r0 + e] = F.y[e];
}

// per-output summed squared error over the held (normalised) pairs
__global__ __launch_bounds__(64) void k_mlp_eval(const float* __restrict__ data, long long n,
                                                 const float* __restrict__ prm, const s4* __restrict__ tiles,
                                                 double* sse) {
    const int l = threadIdx.x, col = l & 15, r0 = 4 * (l >> 4);
    double acc[4] = {0, 0, 0, 0};
    for (long long tile = blockIdx.x; tile * 16 < n; tile += gridDim.x) {
        const long long s = tile * 16 + col;
        float fn[MLP_IN], tn[MLP_OUT];
        for (int k = 0; k < 6; ++k) { fn[k] = s < n ? data[12 * s + k] : 0.f; tn[k] = s < n ? data[12 * s + 6 + k] : 0.f; }
        MlpFwd F;
        mlp_forward_tile(tiles, prm, mlp_feature_operand(fn), F);
        for (int e = 0; e < 4; ++e) {
            const int o = r0 + e;
            float tv = 0.f;
            for (int q = 0; q < MLP_OUT; ++q) tv = (o == q) ? tn[q] : tv;
            if (s < n && o < MLP_OUT) { const double d = F.y[e] - tv; acc[e] += d * d; }
        }
    }
    for (int e = 0; e < 4; ++e) {
        double v = acc[e];
        for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
        if (col == 0 && r0 + e < MLP_OUT) atomicAdd(&sse[r0 + e], v);
    }
}

int mfail(Mlp* m, int code, const std::string& s) {
    if (m) m->err = s;
    return code;
}
#define MCHK(m, expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return mfail(m, FIVO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int pack(Mlp& m) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_NTILES * 64 + 255) / 256), dim3(256), 0, m.stream, m.params, m.tiles);
    MCHK(&m, hipGetLastError());
    return FIVO_OK;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

int run_grad(Mlp& m, const float* data, long long n, long long first, long long count, long long a, long long c) {
    TrainArgs A;
    A.data = data; A.n = n; A.first = first; A.count = count; A.a = a; A.c = c;
    A.tiles_per_wave = 8;
    A.inv_count = 1.0f / (float)count;
    const long long ntile = (count + 15) / 16;
    const int nb = (int)((ntile + A.tiles_per_wave - 1) / A.tiles_per_wave);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(64), 0, m.stream, A, m.params, m.tiles, m.grad, m.acc);
    MCHK(&m, hipGetLastError());
    return FIVO_OK;
}
}  // namespace
}  // namespace fivo

using namespace fivo;

extern "C" {

const char* fivo_mlp_last_error(const fivo_mlp* mp) {
    return mp ? reinterpret_cast<const Mlp*>(mp)->err.c_str() : "";
}

int fivo_mlp_create(int64_t capacity, uint64_t seed, fivo_mlp** out) {
    if (!out || capacity < 0) return FIVO_ERR_ARG;
    *out = nullptr;
    Mlp* m = new Mlp();
    m->capacity = capacity;
    hipError_t e;
    if ((e = hipGetDevice(&m->dev)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&m->params, sizeof(float) * MLP_NPARAM)) != hipSuccess ||
        (e = hipMalloc(&m->grad, sizeof(float) * MLP_NPARAM)) != hipSuccess ||
        (e = hipMalloc(&m->am, sizeof(float) * MLP_NPARAM)) != hipSuccess ||
        (e = hipMalloc(&m->av, sizeof(float) * MLP_NPARAM)) != hipSuccess ||
        (e = hipMalloc(&m->tiles, sizeof(s4) * MLP_NTILES * 64)) != hipSuccess ||
        (e = hipMalloc(&m->norm, sizeof(float) * MLP_NNORM)) != hipSuccess ||
        (e = hipMalloc(&m->count, sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc(&m->acc, sizeof(double) * 64)) != hipSuccess ||
        (e = hipMalloc(&m->part, sizeof(double) * 24 * 1024)) != hipSuccess ||
        (capacity > 0 && (e = hipMalloc(&m->data, sizeof(float) * 12 * (size_t)capacity)) != hipSuccess)) {
        fivo_mlp_destroy(reinterpret_cast<fivo_mlp*>(m));
        return FIVO_ERR_HIP;
    }
    // Glorot-uniform weights from a splitmix64 stream, zero biases, identity normalisation
    std::vector<float> p(MLP_NPARAM, 0.f);
    uint64_t s = seed;
    auto uni = [&](float a) { return (float)((splitmix(s) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * a; };
    const float a1 = std::sqrt(6.f / (MLP_IN + MLP_H)), a2 = std::sqrt(6.f / (MLP_H + MLP_H)),
                a3 = std::sqrt(6.f / (MLP_H + MLP_OUT));
    for (int k = 0; k < MLP_H * MLP_IN; ++k) p[MLP_OFF_W1 + k] = uni(a1);
    for (int k = 0; k < MLP_H * MLP_H; ++k) p[MLP_OFF_W2 + k] = uni(a2);
    for (int k = 0; k < MLP_OUT * MLP_H; ++k) p[MLP_OFF_W3 + k] = uni(a3);
    float nrm[MLP_NNORM];
    for (int k = 0; k < 6; ++k) { nrm[k] = 0.f; nrm[6 + k] = 1.f; nrm[12 + k] = 0.f; nrm[18 + k] = 1.f; }
    int rc = FIVO_OK;
    if ((e = hipMemcpy(m->params, p.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(m->norm, nrm, sizeof(nrm), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(m->grad, 0, sizeof(float) * MLP_NPARAM)) != hipSuccess ||
        (e = hipMemset(m->am, 0, sizeof(float) * MLP_NPARAM)) != hipSuccess ||
        (e = hipMemset(m->av, 0, sizeof(float) * MLP_NPARAM)) != hipSuccess ||
        (e = hipMemset(m->count, 0, sizeof(unsigned long long))) != hipSuccess)
        rc = FIVO_ERR_HIP;
    if (rc == FIVO_OK) rc = pack(*m);
    if (rc == FIVO_OK && hipStreamSynchronize(m->stream) != hipSuccess) rc = FIVO_ERR_HIP;
    if (rc != FIVO_OK) { fivo_mlp_destroy(reinterpret_cast<fivo_mlp*>(m)); return rc; }
    *out = reinterpret_cast<fivo_mlp*>(m);
    return FIVO_OK;
}

void fivo_mlp_destroy(fivo_mlp* mp) {
    if (!mp) return;
    Mlp* m = reinterpret_cast<Mlp*>(mp);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* q : {(void*)m->params, (void*)m->grad, (void*)m->am, (void*)m->av, (void*)m->tiles, (void*)m->norm,
                    (void*)m->count, (void*)m->acc, (void*)m->part, (void*)m->data})
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int fivo_mlp_collect(fivo_mlp* mp, fivo_ctx* ctx, double tau, int64_t* total) {
    Mlp& m = *reinterpret_cast<Mlp*>(mp);
    Ctx& c = *reinterpret_cast<Ctx*>(ctx);
    if (m.normalized) return mfail(&m, FIVO_ERR_STATE, "mlp_collect: the set is already normalised");
    MCHK(&m, hipStreamSynchronize(c.stream));
    const int n = c.p.Nx * c.p.Ny;
    hipLaunchKernelGGL(k_mlp_collect, dim3((n + 255) / 256), dim3(256), 0, m.stream, c.U[c.cur], c.p, c.w, tau,
                       m.data, m.count, (long long)m.capacity);
    MCHK(&m, hipGetLastError());
    unsigned long long k = 0;
    MCHK(&m, hipMemcpyAsync(&k, m.count, sizeof(k), hipMemcpyDeviceToHost, m.stream));
    MCHK(&m, hipStreamSynchronize(m.stream));
    if (total) *total = (int64_t)std::min<unsigned long long>(k, (unsigned long long)m.capacity);
    return FIVO_OK;
}

int fivo_mlp_fit_normalization(fivo_mlp* mp) {
    Mlp& m = *reinterpret_cast<Mlp*>(mp);
    unsigned long long k = 0;
    MCHK(&m, hipMemcpy(&k, m.count, sizeof(k), hipMemcpyDeviceToHost));
    const long long n = (long long)std::min<unsigned long long>(k, (unsigned long long)m.capacity);
    if (n < 2) return mfail(&m, FIVO_ERR_STATE, "mlp_fit_normalization: no pairs collected");
    if (m.normalized) return mfail(&m, FIVO_ERR_STATE, "mlp_fit_normalization: already normalised");
    const int nb = 1024;
    hipLaunchKernelGGL(k_mlp_stats, dim3(nb), dim3(256), 0, m.stream, m.data, n, m.part);
    MCHK(&m, hipGetLastError());
    std::vector<double> part(24 * nb);
    MCHK(&m, hipMemcpyAsync(part.data(), m.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, m.stream));
    MCHK(&m, hipStreamSynchronize(m.stream));
    double sum[24] = {0};
    for (int b = 0; b < nb; ++b) for (int c = 0; c < 24; ++c) sum[c] += part[24 * b + c];
    float nrm[MLP_NNORM];
    for (int c = 0; c < 12; ++c) {
        const double mu = sum[c] / n, var = std::max(sum[12 + c] / n - mu * mu, 0.0);
        const float sd = (float)std::max(std::sqrt(var), 1e-6);
        const int base = c < 6 ? 0 : 12, k6 = c % 6;
        nrm[base + k6] = (float)mu;
        nrm[base + 6 + k6] = sd;
    }
    MCHK(&m, hipMemcpyAsync(m.norm, nrm, sizeof(nrm), hipMemcpyHostToDevice, m.stream));
    hipLaunchKernelGGL(k_mlp_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, m.data, n, m.norm);
    MCHK(&m, hipGetLastError());
    MCHK(&m, hipStreamSynchronize(m.stream));
    m.normalized = 1;
    return FIVO_OK;
}

int fivo_mlp_train(fivo_mlp* mp, int32_t epochs, int32_t batch, float lr, float* loss) {
    Mlp& m = *reinterpret_cast<Mlp*>(mp);
    if (!m.normalized) return mfail(&m, FIVO_ERR_STATE, "mlp_train: fit the normalisation first");
    if (epochs < 1 || batch < 16) return mfail(&m, FIVO_ERR_ARG, "mlp_train: bad epochs/batch");
    unsigned long long k = 0;
    MCHK(&m, hipMemcpy(&k, m.count, sizeof(k), hipMemcpyDeviceToHost));
    const long long n = (long long)std::min<unsigned long long>(k, (unsigned long long)m.capacity);
    uint64_t rs = 0x5eed0000ull + (uint64_t)m.adam_t;
    for (int ep = 0; ep < epochs; ++ep) {
        // a multiplier coprime to n: the epoch visits every row once
        long long a = (long long)(splitmix(rs) % (unsigned long long)n) | 1;
        while (std::gcd(a, n) != 1) a += 2;
        const long long c = (long long)(splitmix(rs) % (unsigned long long)n);
        hipLaunchKernelGGL(k_mlp_zero_d, dim3(1), dim3(64), 0, m.stream, m.acc, 64);
        for (long long first = 0; first < n; first += batch) {
            const long long cnt = std::min<long long>(batch, n - first);
            int rc = run_grad(m, m.data, n, first, cnt, a, c);
            if (rc != FIVO_OK) return rc;
            m.adam_t += 1;
            const float b1c = 1.f - std::pow(0.9f, (float)m.adam_t), b2c = 1.f - std::pow(0.999f, (float)m.adam_t);
            hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m.stream, m.params, m.grad,
                               m.am, m.av, lr, b1c, b2c);
            MCHK(&m, hipGetLastError());
            if ((rc = pack(m)) != FIVO_OK) return rc;
        }
        double ls = 0;
        MCHK(&m, hipMemcpyAsync(&ls, m.acc, sizeof(double), hipMemcpyDeviceToHost, m.stream));
        MCHK(&m, hipStreamSynchronize(m.stream));
        if (loss) loss[ep] = (float)(ls / n);
    }
    return FIVO_OK;
}

int fivo_mlp_size(const fivo_mlp* mp, int64_t* npairs) {
    const Mlp& m = *reinterpret_cast<const Mlp*>(mp);
    unsigned long long k = 0;
    if (hipMemcpy(&k, m.count, sizeof(k), hipMemcpyDeviceToHost) != hipSuccess) return FIVO_ERR_HIP;
    *npairs = (int64_t)std::min<unsigned long long>(k, (unsigned long long)m.capacity);
    return FIVO_OK;
}

int fivo_mlp_get_params(const fivo_mlp* mp, float* params, float* norm) {
    const Mlp& m = *reinterpret_cast<const Mlp*>(mp);
    if (params && hipMemcpy(params, m.params, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost) != hipSuccess)
        return FIVO_ERR_HIP;
    if (norm && hipMemcpy(norm, m.norm, sizeof(float) * MLP_NNORM, hipMemcpyDeviceToHost) != hipSuccess)
        return FIVO_ERR_HIP;
    return FIVO_OK;
}

int fivo_mlp_set_params(fivo_mlp* mp, const float* params, const float* norm) {
    Mlp& m = *reinterpret_cast<Mlp*>(mp);
    if (params) MCHK(&m, hipMemcpy(m.params, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    if (norm) MCHK(&m, hipMemcpy(m.norm, norm, sizeof(float) * MLP_NNORM, hipMemcpyHostToDevice));
    int rc = pack(m);
    if (rc != FIVO_OK) return rc;
    MCHK(&m, hipStreamSynchronize(m.stream));
    return FIVO_OK;
}

int fivo_mlp_forward(fivo_mlp* mp, int64_t n, const float* x, float* y) {
    Mlp& m = *reinterpret_cast<Mlp*>(mp);
    if (n < 1 || !x || !y) return mfail(&m, FIVO_ERR_ARG, "mlp_forward: bad argument");
    float *dx = nullptr, *dy = nullptr;
    MCHK(&m, hipMalloc(&dx, sizeof(float) * MLP_IN * n));
    MCHK(&m, hipMalloc(&dy, sizeof(float) * MLP_OUT * n));
    MCHK(&m, hipMemcpyAsync(dx, x, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice, m.stream));
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 15) / 16)), dim3(64), 0, m.stream, dx, (long long)n,
                       m.params, m.tiles, dy);
    MCHK(&m, hipGetLastError());
    MCHK(&m, hipMemcpyAsync(y, dy, sizeof(float) * MLP_OUT * n, hipMemcpyDeviceToHost, m.stream));
    MCHK(&m, hipStreamSynchronize(m.stream));
    (void)hipFree(dx);
    (void)hipFree(dy);
    return FIVO_OK;
}

int fivo_mlp_eval(fivo_mlp* mp, float* mse) {
    Mlp& m = *reinterpret_cast<Mlp*>(mp);
    if (!m.normalized) return mfail(&m, FIVO_ERR_STATE, "mlp_eval: fit the normalisation first");
    unsigned long long k = 0;
    MCHK(&m, hipMemcpy(&k, m.count, sizeof(k), hipMemcpyDeviceToHost));
    const long long n = (long long)std::min<unsigned long long>(k, (unsigned long long)m.capacity);
    hipLaunchKernelGGL(k_mlp_zero_d, dim3(1), dim3(64), 0, m.stream, m.acc, 64);
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)std::min<long long>((n + 15) / 16, 65536)), dim3(64), 0, m.stream,
                       m.data, n, m.params, m.tiles, m.acc + 8);
    MCHK(&m, hipGetLastError());
    double h[MLP_OUT];
    MCHK(&m, hipMemcpyAsync(h, m.acc + 8, sizeof(h), hipMemcpyDeviceToHost, m.stream));
    MCHK(&m, hipStreamSynchronize(m.stream));
    for (int o = 0; o < MLP_OUT; ++o) mse[o] = (float)(h[o] / n);
    return FIVO_OK;
}

int fivo_mlp_grad(fivo_mlp* mp, int64_t n, const float* x, const float* t, float* loss, float* grad) {
    Mlp& m = *reinterpret_cast<Mlp*>(mp);
    if (n < 1 || !x || !t || !grad) return mfail(&m, FIVO_ERR_ARG, "mlp_grad: bad argument");
    std::vector<float> rows(12 * (size_t)n);
    for (int64_t s = 0; s < n; ++s)
        for (int k = 0; k < 6; ++k) { rows[12 * s + k] = x[6 * s + k]; rows[12 * s + 6 + k] = t[6 * s + k]; }
    float* d = nullptr;
    MCHK(&m, hipMalloc(&d, sizeof(float) * rows.size()));
    MCHK(&m, hipMemcpyAsync(d, rows.data(), sizeof(float) * rows.size(), hipMemcpyHostToDevice, m.stream));
    hipLaunchKernelGGL(k_mlp_fill, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m.stream, m.grad, 0.f, MLP_NPARAM);
    hipLaunchKernelGGL(k_mlp_zero_d, dim3(1), dim3(64), 0, m.stream, m.acc, 64);
    int rc = run_grad(m, d, n, 0, n, 1, 0);
    if (rc != FIVO_OK) return rc;
    double ls = 0;
    MCHK(&m, hipMemcpyAsync(grad, m.grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m.stream));
    MCHK(&m, hipMemcpyAsync(&ls, m.acc, sizeof(double), hipMemcpyDeviceToHost, m.stream));
    hipLaunchKernelGGL(k_mlp_fill, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m.stream, m.grad, 0.f, MLP_NPARAM);
    MCHK(&m, hipStreamSynchronize(m.stream));
    (void)hipFree(d);
    if (loss) *loss = (float)(ls / n);
    return FIVO_OK;
}

}  // extern "C"

namespace fivo {
MlpDev mlp_dev(const fivo_mlp* mp) {
    const Mlp& m = *reinterpret_cast<const Mlp*>(mp);
    MlpDev d;
    d.tiles = m.tiles;
    d.bias = m.params;
    d.norm = m.norm;
    return d;
}
}  // namespace fivo
