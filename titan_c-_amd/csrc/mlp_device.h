// mlp_device.h -- device side of the charge-exchange surrogate 8-64-64-64-5 (SiLU): the BF16 MFMA
// forward pass of one wavefront (32 cells per pass, FP32 accumulation), the input features of a
// cell, the training-box test and the source assembly.  Shared by the standalone surrogate
// kernels and the fused step kernel, and written without contractible a*b+c patterns and with
// explicit reciprocals, so every translation unit (whatever its -ffp-contract and division
// flags) produces the same bits.
//
// Inference maps one wavefront onto v_mfma_f32_32x32x16_bf16 with the network transposed:
// D[32 features x 32 cells] = W[32 x 16] . H^T[16 x 32 cells].  Lane l supplies W row l%32,
// k = 8*(l/32)..+7, and cell l%32's features k = 8*(l/32)..+7; it receives features
// 8*(i/4) + 4*(l/32) + i%4 (i = 0..15) of cell l%32.  The rows of every hidden layer are stored
// permuted so that accumulator registers 8*(kb%2)..+7 of tile kb/2 ARE the next layer's B operand
// for k-block kb: activations never leave the registers between layers.
#pragma once
#include <hip/hip_runtime.h>

#include "mlp.h"
#include "cx_device.h"

namespace titan {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// parameter offsets (row-major W[out][in], then b[out])
constexpr int oW1 = 0, oB1 = oW1 + 64 * 8, oW2 = oB1 + 64, oB2 = oW2 + 64 * 64, oW3 = oB2 + 64, oB3 = oW3 + 64 * 64,
              oW4 = oB3 + 64, oB4 = oW4 + 5 * 64;
// packed operand bases (in lane-vectors) and bias bases
constexpr int vL1 = 0, vL2 = 128, vL3 = 640, vL4 = 1152;
constexpr int bL1 = 0, bL2 = 64, bL3 = 128, bL4 = 192;
// normalisation block offsets
constexpr int kNormMean = 0, kNormStd = 8, kNormScale = 16, kNormLo = 21, kNormHi = 29;

// logical hidden feature held by (tile t, accumulator register i, lane half h)
__host__ __device__ inline int hidden_logical(int t, int i, int h) { return 16 * (2 * t + i / 8) + 8 * h + (i % 8); }
// physical row of (register i, lane half h) inside a 32-row tile
__host__ __device__ inline int tile_row(int i, int h) { return (i / 4) * 8 + h * 4 + (i % 4); }

__device__ __forceinline__ float silu_f(float x) { return x * __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

__device__ __forceinline__ void act_to_b(const f32x16 &a, const f32x16 &b, bf16x8 B[4]) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
        B[0][e] = (__bf16)silu_f(a[e]);
        B[1][e] = (__bf16)silu_f(a[8 + e]);
        B[2][e] = (__bf16)silu_f(b[e]);
        B[3][e] = (__bf16)silu_f(b[8 + e]);
    }
}

__device__ __forceinline__ f32x16 bias_init(const float *bias, int base, int t, int h) {
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = bias[base + (t * 2 + h) * 16 + i];
    return acc;
}

// One 32-cell pass of the network (all 64 lanes active).  W: packed lane-vectors, bias: packed
// biases, in: this lane's B operand of layer 1.  Returns the final accumulator (5 outputs:
// registers 0..3 of half 0 and register 0 of half 1).
__device__ __forceinline__ f32x16 mlp_pass(const bf16x8 *W, const float *bias, bf16x8 in, int lane) {
    const int h = lane >> 5;
    f32x16 a0 = bias_init(bias, bL1, 0, h), a1 = bias_init(bias, bL1, 1, h);
    a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[vL1 + 0 * 64 + lane], in, a0, 0, 0, 0);
    a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[vL1 + 1 * 64 + lane], in, a1, 0, 0, 0);
    bf16x8 B[4];
    act_to_b(a0, a1, B);
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
        const int vb = layer == 0 ? vL2 : vL3, bb = layer == 0 ? bL2 : bL3;
        f32x16 c0 = bias_init(bias, bb, 0, h), c1 = bias_init(bias, bb, 1, h);
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
            c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[vb + (0 * 4 + kb) * 64 + lane], B[kb], c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[vb + (1 * 4 + kb) * 64 + lane], B[kb], c1, 0, 0, 0);
        }
        act_to_b(c0, c1, B);
    }
    f32x16 o = bias_init(bias, bL4, 0, h);
#pragma unroll
    for (int kb = 0; kb < 4; kb++) o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[vL4 + kb * 64 + lane], B[kb], o, 0, 0, 0);
    return o;
}

// Block-cooperative copy of the packed operands (and optionally the normalisation block) into LDS.
__device__ __forceinline__ void load_packed(const uint16_t *__restrict__ wpack, const float *__restrict__ bpack,
                                            bf16x8 *sW, float *sB) {
    const uint4 *src = reinterpret_cast<const uint4 *>(wpack);
    uint4 *dst = reinterpret_cast<uint4 *>(sW);
    for (int i = threadIdx.x; i < kMlpPackVec; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < kMlpPackBias; i += blockDim.x) sB[i] = bpack[i];
}

// Raw input features of a cell: ln n_p, ln c_p^2 (= p/rho), ln n_H, ln c_H^2, V_p - V_H, |V_p - V_H|.
// s = (rho, Vx, Vy, Vz, p, ...), hn = (n_H, c_H^2, V_H).  The logarithms are FP32 (the features
// are rounded to BF16 for the network); the velocity differences and their norm are FP64.
__device__ __forceinline__ void mlp_features(const double *s, const double *hn, float xr[8]) {
    const double d0 = s[1] - hn[2], d1 = s[2] - hn[3], d2 = s[3] - hn[4];
    // hardware log2, scaled after the difference (no multiply-add for a contracting unit to fuse)
    const float ln2 = 0.69314718055994531f;
    const float l2r = __builtin_amdgcn_logf((float)s[0]);
    xr[0] = l2r * ln2;
    xr[1] = (__builtin_amdgcn_logf((float)s[4]) - l2r) * ln2;
    xr[2] = __builtin_amdgcn_logf((float)hn[0]) * ln2;
    xr[3] = __builtin_amdgcn_logf((float)hn[1]) * ln2;
    xr[4] = (float)d0;
    xr[5] = (float)d1;
    xr[6] = (float)d2;
    xr[7] = (float)sqrt(__builtin_fma(d0, d0, __builtin_fma(d1, d1, d2 * d2)));
}

__device__ __forceinline__ bool mlp_in_box(const float xr[8], const float *norm) {
    bool in = true;
#pragma unroll
    for (int e = 0; e < 8; e++) in = in && (xr[e] >= norm[kNormLo + e]) && (xr[e] <= norm[kNormHi + e]);
    return in;
}

__device__ __forceinline__ __bf16 mlp_normalise(float v, const float *norm, int e) {
    return (__bf16)((v - norm[kNormMean + e]) * __builtin_amdgcn_rcpf(norm[kNormStd + e]));
}

// FP32 pieces of the closed-form estimates, each built from hardware instructions only
// (v_log_f32, v_exp_f32, v_sqrt_f32, v_rcp_f32) so that every translation unit, whatever its
// contraction and FP32 division / square-root flags, forms the same bits.
__device__ __forceinline__ float est_exp(float x) {
#pragma clang fp contract(off)
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}
__device__ __forceinline__ float est_ln(float x) {
#pragma clang fp contract(off)
    return __builtin_amdgcn_logf(x) * 0.69314718055994531f;
}
// erf by Abramowitz & Stegun 7.1.26 (absolute error below 1.5e-7), x >= 0
__device__ __forceinline__ float est_erf(float x) {
#pragma clang fp contract(off)
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * x);
    const float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f +
                                                                                   t * 1.061405429f))));
    return 1.0f - poly * est_exp(-(x * x));
}
// the Lindsay-Stebbings cross-section of sigma_ls (cx_device.h) in FP32
__device__ __forceinline__ float est_sigma_ls(float g) {
#pragma clang fp contract(off)
    const float e_per_g2 = (float)(0.5 * 1.67262192369e-24 * (1.03804e6 * 1.03804e6) / 1.602176634e-9);
    const float nl16 = (float)(1e-16 * 0.06 * 4.2113 * 1.495978707e13);
    float E = e_per_g2 * (g * g);
    if (E < 1e-12f) E = 1e-12f;
    const float a = 4.15f - 0.531f * est_ln(E);
    const float b = 1.0f - est_exp(-67.3f * __builtin_amdgcn_rcpf(E));
    // b^4.5 (b in [0, 1]: log2(0) = -inf gives 0)
    return ((a * a) * __builtin_amdgcn_exp2f(4.5f * __builtin_amdgcn_logf(b))) * nl16;
}

// Closed-form estimates, per n_p n_H, of the five sources the network corrects (the scaled residual
// form): over the relative-velocity Maxwellian of drift U and width c (c^2 = c_p^2 + c_H^2), the
// sigma = const moments <g> and <g^3> and the momentum moment U u_m (u_m^2 = U^2 + 64 c^2/(9 pi)),
// each times sigma at its speed (<g>, u_m, <g^3>^(1/3)), estimate J0, J1, J2; they give the rate
// B0, the momentum source -B1 e and the plasma-frame energy source (U B1 (1 + kappa) - kappa B2)/2
// (kappa = (c_p^2 - c_H^2)/c^2).  b[k] = the estimate of output k, m[k] > 0 its scale: the
// network's output k is (source_k / (n_p n_H) - b[k]) / m[k].  The training labels take the same
// formulas in FP64 (surrogate.py); these FP32 estimates stay within ~1e-6 of them, far inside the
// network's own error, and cost a few hardware transcendentals instead of a dozen FP64 library
// calls per cell.
__device__ __forceinline__ void mlp_baseline(float U, float c, float kappa, const float e[3], float b[5], float m[5]) {
#pragma clang fp contract(off)
    const float rpi = 0.56418958354775628695f;  // 1/sqrt(pi)
    const float yc = __builtin_amdgcn_rcpf(c);
    const float w = U * yc;
    float g1, g3;
    if (w < 1e-3f) {  // the series of the closed forms below at small drift
        g1 = c * (2.0f * rpi) * (1.0f + w * w * (1.0f / 3.0f));
        g3 = c * c * c * (4.0f * rpi) * (1.0f + w * w);
    } else {
        const float ew = est_exp(-(w * w)) * rpi, erw = est_erf(w), yw = __builtin_amdgcn_rcpf(w);
        g1 = c * (ew + (w + 0.5f * yw) * erw);
        g3 = c * c * c * ((w * w + 2.5f) * ew + (w * w * w + 3.0f * w + 0.75f * yw) * erw);
    }
    const float um = __builtin_amdgcn_sqrtf(U * U + (float)(64.0 / (9.0 * 3.14159265358979323846)) * (c * c));
    const float cg3 = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(g3) * (1.0f / 3.0f));  // <g^3>^(1/3)
    const float B0 = est_sigma_ls(g1) * g1;
    const float B1 = est_sigma_ls(um) * U * um;
    const float B2 = est_sigma_ls(cg3) * g3;
    b[0] = B0; m[0] = B0;
#pragma unroll
    for (int k = 0; k < 3; k++) { b[1 + k] = -B1 * e[k]; m[1 + k] = B1; }
    b[4] = 0.5f * (U * B1 * (1.0f + kappa) - kappa * B2);
    m[4] = 0.5f * (U * B1 * (1.0f + kappa) + (fabsf(kappa) + 0.01f) * B2);
#pragma unroll
    for (int k = 0; k < 5; k++) m[k] = m[k] > 1e-30f ? m[k] : 1e-30f;
}

// Source of one cell from the network's 5 outputs y (normalised): rate, S_m, and the plasma-frame
// energy source per n_p n_H, each the closed-form estimate plus the output times its scale;
// S_E(lab) = S_E' + V_p . S_m.  s = (rho, V, p, ...), hn = (n_H, c_H^2, V_H).
__device__ __forceinline__ void mlp_source(const float y[5], const float *norm, const double *s, const double *hn,
                                           double o[5]) {
#pragma clang fp contract(off)
    const double nn = s[0] * hn[0];
    // the relative frame in FP32 from FP64 differences (reciprocals and roots by hardware instruction)
    const float d0 = (float)(s[1] - hn[2]), d1 = (float)(s[2] - hn[3]), d2 = (float)(s[3] - hn[4]);
    const float U = __builtin_amdgcn_sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
    const float cp2 = (float)s[4] * __builtin_amdgcn_rcpf((float)s[0]), c2 = cp2 + (float)hn[1];
    const float yU = __builtin_amdgcn_rcpf(U);
    const float e[3] = {U > 0.0f ? d0 * yU : 1.0f, U > 0.0f ? d1 * yU : 0.0f, U > 0.0f ? d2 * yU : 0.0f};
    float b[5], m[5];
    mlp_baseline(U, __builtin_amdgcn_sqrtf(c2), (cp2 - (float)hn[1]) * __builtin_amdgcn_rcpf(c2), e, b, m);
    double yy[5];
#pragma unroll
    for (int k = 0; k < 5; k++)
        yy[k] = __builtin_fma((double)y[k] * (double)norm[kNormScale + k], (double)m[k], (double)b[k]);
    o[0] = nn * yy[0];
    o[1] = nn * yy[1];
    o[2] = nn * yy[2];
    o[3] = nn * yy[3];
    o[4] = __builtin_fma(s[3], o[3], __builtin_fma(s[2], o[2], __builtin_fma(s[1], o[1], nn * yy[4])));
}

}  // namespace titan
