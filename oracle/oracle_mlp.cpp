// =============================================================================
// oracle/oracle_mlp.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
// CPU reference of the learned warm start's network (6 -> 64 -> 64 -> 6, GELU)
// and of its loss gradient, with the operands of every matrix product rounded
// to BF16 (round to nearest even) exactly where the GPU's MFMA tiles take BF16
// inputs; accumulation in double. The GPU library's forward and gradient are
// checked against these (tests/test_gpu_mlp.py).
//
// Loss: L = (1/n) sum_i sum_o (y_o(x_i) - t_o,i)^2 over normalised outputs.
// Parameter vector: W1[64][6], b1[64], W2[64][64], b2[64], W3[6][64], b3[6].
// =============================================================================
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
constexpr int NI = 6, NH = 64, NO = 6;
constexpr int W1 = 0, B1 = 384, W2 = 448, B2 = 4544, W3 = 4608, B3 = 4992, NP = 4998;

int g_exact = 0;      // 1: no BF16 rounding (pins the backward pass against finite differences)
float bf(float x) {   // round to the nearest BF16, ties to even
    if (g_exact) return x;
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float y;
    std::memcpy(&y, &u, 4);
    return y;
}
float gelu(float z) { return 0.5f * z * (1.0f + std::erf(z * 0.70710678118654752f)); }
float gelu_d(float z) {
    return 0.5f * (1.0f + std::erf(z * 0.70710678118654752f)) + z * 0.39894228040143268f * std::exp(-0.5f * z * z);
}

struct Fwd {
    float xb[NI], z1[NH], h1b[NH], z2[NH], h2b[NH], y[NO];
};
void forward(const float* p, const float* x, Fwd& f) {
    for (int k = 0; k < NI; ++k) f.xb[k] = bf(x[k]);
    for (int j = 0; j < NH; ++j) {
        double s = 0;
        for (int k = 0; k < NI; ++k) s += (double)bf(p[W1 + j * NI + k]) * f.xb[k];
        f.z1[j] = (float)s + p[B1 + j];
        f.h1b[j] = bf(gelu(f.z1[j]));
    }
    for (int j = 0; j < NH; ++j) {
        double s = 0;
        for (int k = 0; k < NH; ++k) s += (double)bf(p[W2 + j * NH + k]) * f.h1b[k];
        f.z2[j] = (float)s + p[B2 + j];
        f.h2b[j] = bf(gelu(f.z2[j]));
    }
    for (int o = 0; o < NO; ++o) {
        double s = 0;
        for (int k = 0; k < NH; ++k) s += (double)bf(p[W3 + o * NH + k]) * f.h2b[k];
        f.y[o] = (float)s + p[B3 + o];
    }
}
}  // namespace

extern "C" {

void orc_mlp_set_exact(int exact) { g_exact = exact; }

void orc_mlp_forward(const float* p, int64_t n, const float* x, float* y) {
    Fwd f;
    for (int64_t i = 0; i < n; ++i) {
        forward(p, x + NI * i, f);
        for (int o = 0; o < NO; ++o) y[NO * i + o] = f.y[o];
    }
}

void orc_mlp_grad(const float* p, int64_t n, const float* x, const float* t, double* loss, double* g) {
    std::memset(g, 0, sizeof(double) * NP);
    double L = 0;
    Fwd f;
    const float inv = 1.0f / (float)n;
    for (int64_t i = 0; i < n; ++i) {
        forward(p, x + NI * i, f);
        float dy[NO], dyb[NO], dz2[NH], dz2b[NH], dz1b[NH];
        for (int o = 0; o < NO; ++o) {
            const float d = f.y[o] - t[NO * i + o];
            L += (double)d * d;
            dy[o] = 2.0f * d * inv;
            dyb[o] = bf(dy[o]);
            g[B3 + o] += dy[o];
            for (int j = 0; j < NH; ++j) g[W3 + o * NH + j] += (double)dyb[o] * f.h2b[j];
        }
        for (int j = 0; j < NH; ++j) {
            double s = 0;
            for (int o = 0; o < NO; ++o) s += (double)bf(p[W3 + o * NH + j]) * dyb[o];
            dz2[j] = (float)s * gelu_d(f.z2[j]);
            dz2b[j] = bf(dz2[j]);
            g[B2 + j] += dz2[j];
        }
        for (int jo = 0; jo < NH; ++jo)
            for (int ji = 0; ji < NH; ++ji) g[W2 + jo * NH + ji] += (double)dz2b[jo] * f.h1b[ji];
        for (int ji = 0; ji < NH; ++ji) {
            double s = 0;
            for (int jo = 0; jo < NH; ++jo) s += (double)bf(p[W2 + jo * NH + ji]) * dz2b[jo];
            const float dz1 = (float)s * gelu_d(f.z1[ji]);
            dz1b[ji] = bf(dz1);
            g[B1 + ji] += dz1;
        }
        for (int j = 0; j < NH; ++j)
            for (int k = 0; k < NI; ++k) g[W1 + j * NI + k] += (double)dz1b[j] * f.xb[k];
    }
    *loss = L / (double)n;
}

}  // extern "C"
