// =============================================================================
// fivo_mlp.h -- the learned Newton warm start: a 6 -> 64 -> 64 -> 6 GELU MLP on
// BF16 MFMA tiles (v_mfma_f32_16x16x16_bf16, FP32 accumulate).
//
// The network maps features of a cell's conserved state to its primitive
// variables; its output replaces the cold seed of the primitive recovery
// (seed_from_conserved_2d, src2d/primrec2d.jl:323-347) for cells without a warm
// start, and Newton still polishes every cell to the reference tolerance, with
// the cold seed as the fallback when it does not converge. It has no
// counterpart in the reference.
//
// Samples are the columns of every product (Z = W X): with the MFMA lane layout
//   A[m][k]: lane m + 16 (k/4), element k%4     (16 x 16 bf16)
//   B[k][n]: lane n + 16 (k/4), element k%4
//   D[m][n]: lane n + 16 (m/4), element m%4     (16 x 16 fp32)
// a layer's output tile D (rows = units, columns = samples) is, after the
// activation, the next layer's B operand in the same registers.
// =============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fivo_device.h"

namespace fivo {

constexpr int MLP_IN = 6, MLP_H = 64, MLP_OUT = 6;
// parameter vector: W1[64][6], b1[64], W2[64][64], b2[64], W3[6][64], b3[6]
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = 384, MLP_OFF_W2 = 448, MLP_OFF_B2 = 4544, MLP_OFF_W3 = 4608,
              MLP_OFF_B3 = 4992, MLP_NPARAM = 4998;
// normalisation block: feature mean[6], feature std[6], target mean[6], target std[6]
constexpr int MLP_NNORM = 24;
// packed A tiles (64 lanes x 4 bf16 each): forward W1 (4), W2 (16), W3 (4);
// backward W3^T (4), W2^T (16)
constexpr int MLP_T_W1 = 0, MLP_T_W2 = 4, MLP_T_W3 = 20, MLP_T_W3T = 24, MLP_T_W2T = 28, MLP_NTILES = 44;

typedef __attribute__((ext_vector_type(4))) float f4;
typedef __attribute__((ext_vector_type(4))) short s4;

// what an inference needs, device pointers
struct MlpDev {
    const s4* tiles;     // [MLP_NTILES][64]
    const float* bias;   // the parameter vector (FP32); biases are read from it
    const float* norm;   // MLP_NNORM
};

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__host__ __device__ inline uint16_t bf16_rne(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

// GELU (exact, erf form) and its derivative
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ f4 mfma16(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ __forceinline__ s4 pack4(float a, float b, float c, float d) {
    s4 r;
    r[0] = (short)bf16_rne(a); r[1] = (short)bf16_rne(b); r[2] = (short)bf16_rne(c); r[3] = (short)bf16_rne(d);
    return r;
}

// ---- features and targets (double precision, per cell) ------------------------
// features: log E, log(D + 1e-30), |S|/E, Pi/E, (s.pi.s)/E, (s.nu)/(|nu| + D + 1e-30),
// with s = S/|S| (x axis when S = 0); targets: yT, phi, asinh(u.s), u.s_perp, log e, log P
struct MlpFrame { double sx, sy; };
__device__ __forceinline__ MlpFrame mlp_frame(double Sx, double Sy) {
    const double M = sqrt(Sx * Sx + Sy * Sy);
    MlpFrame f;
    if (M > 1e-300) { f.sx = Sx / M; f.sy = Sy / M; } else { f.sx = 1.0; f.sy = 0.0; }
    return f;
}
__device__ __forceinline__ void mlp_features(double D, double Sx, double Sy, double E, const fd::Diss& d,
                                             const MlpFrame& fr, double f[MLP_IN]) {
    const double M = sqrt(Sx * Sx + Sy * Sy);
    f[0] = log(E);
    f[1] = log(fd::jmax(D, 0.0) + 1e-30);
    f[2] = M / E;
    f[3] = d.Pi / E;
    const double pss = fr.sx * fr.sx * d.pixx + 2.0 * fr.sx * fr.sy * d.pixy + fr.sy * fr.sy * d.piyy;
    f[4] = pss / E;
    const double nus = fr.sx * d.nux + fr.sy * d.nuy;
    f[5] = nus / (sqrt(d.nux * d.nux + d.nuy * d.nuy) + fd::jmax(D, 0.0) + 1e-30);
}
__device__ __forceinline__ void mlp_targets(double yT, double phi, double ux, double uy, double e, double P,
                                            const MlpFrame& fr, double t[MLP_OUT]) {
    const double us = fr.sx * ux + fr.sy * uy, up = -fr.sy * ux + fr.sx * uy;
    t[0] = yT;
    t[1] = phi;
    t[2] = asinh(us);
    t[3] = up;
    t[4] = log(fd::jmax(e, 1e-300));
    t[5] = log(fd::jmax(P, 1e-300));
}

// ---- one 16-sample forward pass on a wavefront ----------------------------------
// bx: this lane's B operand of the normalised features (rows k = 4 (lane/16) + e).
// z1, z2: pre-activations (kept for the backward pass); h1b, h2b: BF16 activations
// as B operands; y: the 16 x 16 output tile (rows >= MLP_OUT are zero padding).
struct MlpFwd { f4 z1[4], z2[4], y; s4 h1b[4], h2b[4]; };
__device__ __forceinline__ void mlp_forward_tile(const s4* __restrict__ tiles, const float* __restrict__ prm,
                                                 s4 bx, MlpFwd& F) {
    const int l = threadIdx.x & 63;
    const int r0 = 4 * (l >> 4);           // first unit row this lane holds in a D tile
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f4 z = mfma16(tiles[(MLP_T_W1 + t) * 64 + l], bx, zero);
        float h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) { z[e] += prm[MLP_OFF_B1 + 16 * t + r0 + e]; h[e] = gelu(z[e]); }
        F.z1[t] = z;
        F.h1b[t] = pack4(h[0], h[1], h[2], h[3]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        f4 z = zero;
#pragma unroll
        for (int t = 0; t < 4; ++t) z = mfma16(tiles[(MLP_T_W2 + 4 * u + t) * 64 + l], F.h1b[t], z);
        float h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) { z[e] += prm[MLP_OFF_B2 + 16 * u + r0 + e]; h[e] = gelu(z[e]); }
        F.z2[u] = z;
        F.h2b[u] = pack4(h[0], h[1], h[2], h[3]);
    }
    f4 y = zero;
#pragma unroll
    for (int t = 0; t < 4; ++t) y = mfma16(tiles[(MLP_T_W3 + t) * 64 + l], F.h2b[t], y);
#pragma unroll
    for (int e = 0; e < 4; ++e) y[e] += (r0 + e < MLP_OUT) ? prm[MLP_OFF_B3 + r0 + e] : 0.f;
    F.y = y;
}

// B operand of normalised features for the sample of this lane (lane % 16):
// rows 4 (lane/16) .. +3 of the padded 16-row feature column
__device__ __forceinline__ s4 mlp_feature_operand(const float fn[MLP_IN]) {
    const int g = (threadIdx.x & 63) >> 4;
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = 4 * g + e;
        float x = 0.f;
#pragma unroll
        for (int q = 0; q < MLP_IN; ++q) x = (k == q) ? fn[q] : x;
        v[e] = x;
    }
    return pack4(v[0], v[1], v[2], v[3]);
}

// ---- the learned seed of a queued cell -----------------------------------------
// Lanes 0..15 of a wavefront end with the seed (yT, phi, ux, uy) of their sample;
// the network's outputs 0..3 of a sample are in lane sample (group 0) of y.
__device__ __forceinline__ void mlp_seed_from_outputs(const f4& y, const float* __restrict__ norm,
                                                      const MlpFrame& fr, double x[4]) {
    double o[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = (double)y[e] * (double)norm[18 + e] + (double)norm[12 + e];
    const double yT = fd::jclamp(o[0], log(fd::T_MIN), log(fd::T_SOLVE_MAX));
    const double phi = fd::jclamp(o[1], -fd::PHI_CAP, fd::PHI_CAP);
    const double us = sinh(fd::jclamp(o[2], -10.0, 10.0)), up = o[3];
    x[0] = yT;
    x[1] = phi;
    x[2] = fd::jclamp(us * fr.sx - up * fr.sy, -fd::U_CAP, fd::U_CAP);
    x[3] = fd::jclamp(us * fr.sy + up * fr.sx, -fd::U_CAP, fd::U_CAP);
}

}  // namespace fivo
