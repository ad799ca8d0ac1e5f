// k_train.hip -- training of the charge-exchange surrogate (8-64-64-64-5, SiLU) on the matrix cores in
// FP32: one fused kernel per mini-batch does the forward pass, the backward pass and the weight
// gradient of each wavefront's 32 samples on v_mfma_f32_32x32x2_f32, and one kernel sums the
// wavefronts' gradients in a fixed order, takes the Adam step and repacks the weights for the
// next batch.  Mini-batches are drawn by a keyed permutation on the device, so an epoch needs no
// host round trip.
//
// Layout (32 samples per wavefront): a layer's pre-activations Z [features x samples] live in the
// MFMA accumulator layout -- lane l holds sample l%32 and, in register r of tile t, feature
// 32t + 8(r/4) + 4(l/32) + r%4.  The next layer reads them back as its B operand without moving
// them: its k-step (t, r) pairs features 32t + 8(r/4) + r%4 (lanes 0..31) and that + 4 (lanes
// 32..63), and the weights are packed in that k order (the forward pack fw).  The backward data
// products W^T dZ use the same trick with the transposed packs (bw).  The weight gradient
// dZ A^T sums over samples, which the accumulator layout holds across lanes, so dZ and A pass
// through LDS transposed ([feature][sample]) first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "../../include/titan_gpu.h"
#include "mlp.h"
#include "mlp_device.h"

namespace titan {

#define HIPCHK_T(expr)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            std::fprintf(stderr, "titan_gpu(train): %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return TITAN_E_HIP;                                                                \
        }                                                                                      \
    } while (0)

// forward packs: one float per lane per (tile, k-step) of A = W_l (rows = outputs)
constexpr int fwL1 = 0, fwL2 = fwL1 + 2 * 4 * 64, fwL3 = fwL2 + 2 * 32 * 64, fwL4 = fwL3 + 2 * 32 * 64,
              kFwFloats = fwL4 + 32 * 64;
// backward packs of A = W_l^T (rows = inputs) for l = 4, 3, 2
constexpr int bw4 = 0, bw3 = bw4 + 2 * 4 * 64, bw2 = bw3 + 2 * 32 * 64, kBwFloats = bw2 + 2 * 32 * 64;
// a wavefront's gradient partial: the kMlpParams gradients, its squared-error sum, padding
constexpr int kPartLoss = (int)kMlpParams, kPartStride = 145 * 64;
static_assert(kPartStride > kPartLoss, "partial row holds the loss");

// k-step and lane half of hidden feature f as a B operand read from the accumulator layout
__host__ __device__ inline void feat_kstep(int f, int &kk, int &h) {
    const int t = f / 32, rem = f % 32, r = 4 * (rem / 8) + rem % 4;
    h = (rem % 8) / 4;
    kk = 16 * t + r;
}

// Where parameter p goes in the forward (fi) and backward (bi) packs; -1 = not packed (biases, W1^T).
__host__ __device__ inline void pack_slots(int p, int &fi, int &bi) {
    fi = -1;
    bi = -1;
    int kk, h;
    if (p < oB1) {
        const int o = p / 8, i = p % 8;
        fi = fwL1 + ((o / 32) * 4 + i / 2) * 64 + (i % 2) * 32 + o % 32;
    } else if ((p >= oW2 && p < oB2) || (p >= oW3 && p < oB3)) {
        const bool l2 = p < oB2;
        const int q = p - (l2 ? oW2 : oW3), o = q / 64, i = q % 64;
        feat_kstep(i, kk, h);
        fi = (l2 ? fwL2 : fwL3) + ((o / 32) * 32 + kk) * 64 + h * 32 + o % 32;
        feat_kstep(o, kk, h);
        bi = (l2 ? bw2 : bw3) + ((i / 32) * 32 + kk) * 64 + h * 32 + i % 32;
    } else if (p >= oW4 && p < oB4) {
        const int q = p - oW4, o = q / 64, i = q % 64;
        feat_kstep(i, kk, h);
        fi = fwL4 + kk * 64 + h * 32 + o;
        bi = bw4 + ((i / 32) * 4 + o % 4) * 64 + (o / 4) * 32 + i % 32;  // output row o < 5: k-step o%4, half o/4
    }
}

// One workgroup: clears both packs (the output layer's padding rows stay zero), then scatters P.
__global__ void k_pack_train(const float *__restrict__ P, float *__restrict__ fw, float *__restrict__ bw) {
    for (int q = threadIdx.x; q < kFwFloats; q += blockDim.x) fw[q] = 0.0f;
    for (int q = threadIdx.x; q < kBwFloats; q += blockDim.x) bw[q] = 0.0f;
    __syncthreads();
    for (int q = threadIdx.x; q < kMlpParams; q += blockDim.x) {
        int fi, bi;
        pack_slots(q, fi, bi);
        if (fi >= 0) fw[fi] = P[q];
        if (bi >= 0) bw[bi] = P[q];
    }
}

__device__ __forceinline__ uint64_t mix64t(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// A permutation of [0, n): a 4-round Feistel network on `bits` (even) bits keyed by `key`, walked
// until it lands below n (its orbit through i < n returns below n; the walk is capped, after which
// i maps to itself).
__device__ __forceinline__ uint32_t perm_index(uint32_t i, uint32_t n, int bits, uint64_t key) {
    const int half = bits / 2;
    const uint32_t mask = (1u << half) - 1u;
    uint32_t x = i;
    for (int walk = 0; walk < 256; walk++) {
        uint32_t L = x >> half, R = x & mask;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t F = (uint32_t)mix64t(key ^ (((uint64_t)R << 8) | (uint64_t)r)) & mask;
            const uint32_t t = R;
            R = L ^ F;
            L = t;
        }
        x = (L << half) | R;
        if (x < n) return x;
    }
    return i;
}

// SiLU and its derivative from the hardware exp2 and reciprocal (about 1 ulp each): training
// tolerates that, and the IEEE quotient and expf would triple the kernel's vector instructions.
__device__ __forceinline__ float sigm_t(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_t(float z) { return z * sigm_t(z); }
__device__ __forceinline__ float dsilu_t(float z) {
    const float s = sigm_t(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x16 mfma2(float a, float b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int acc_row(int r, int h) { return 8 * (r / 4) + 4 * h + r % 4; }

// Z = W A_prev + b for a 64-output hidden layer whose input is a 64-feature accumulator-layout
// activation (act = SiLU of zp).
__device__ __forceinline__ void hidden_fwd(const float *__restrict__ fw, int base, const float *__restrict__ bias,
                                           const f32x16 zp[2], f32x16 z[2], int lane, int h) {
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) z[t][r] = bias[32 * t + acc_row(r, h)];
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const float b = silu_t(zp[ti][r]);
            const int kk = 16 * ti + r;
#pragma unroll
            for (int t = 0; t < 2; t++) z[t] = mfma2(fw[base + (t * 32 + kk) * 64 + lane], b, z[t]);
        }
}

// dZ_prev = (W^T dZ) * silu'(zp) with W^T from a backward pack (64 x 64 layer).
__device__ __forceinline__ void hidden_bwd(const float *__restrict__ bw, int base, const f32x16 dz[2],
                                           const f32x16 zp[2], f32x16 dzp[2], int lane) {
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) dzp[t][r] = 0.0f;
#pragma unroll
    for (int to = 0; to < 2; to++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int kk = 16 * to + r;
#pragma unroll
            for (int t = 0; t < 2; t++) dzp[t] = mfma2(bw[base + (t * 32 + kk) * 64 + lane], dz[to][r], dzp[t]);
        }
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) dzp[t][r] *= dsilu_t(zp[t][r]);
}

// Accumulator-layout [rows x 32 samples] tiles -> LDS [row][sample].
__device__ __forceinline__ void to_lds(float (*dst)[33], const f32x16 *v, int tiles, int col, int h, bool silu) {
    for (int t = 0; t < tiles; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) dst[32 * t + acc_row(r, h)][col] = silu ? silu_t(v[t][r]) : v[t][r];
}

// Partial weight and bias gradient of one layer over the wavefront's 32 samples: dW = dZ A^T from
// sT[0] = dZ [nout rows][32] and sT[1] = A [nin rows][32]; db = the row sums of dZ.
__device__ __forceinline__ void layer_wgrad(float (*sT)[64][33], int nout, int nin, int woff, int boff,
                                            float *__restrict__ part, int lane, int col, int h) {
    const int tm = nout > 32 ? 2 : 1, tn = nin > 32 ? 2 : 1;
    for (int a = 0; a < tm; a++)
        for (int b = 0; b < tn; b++) {
            f32x16 acc;
#pragma unroll
            for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                const int s = 2 * kk + h;
                acc = mfma2(sT[0][32 * a + col][s], sT[1][32 * b + col][s], acc);
            }
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int o = 32 * a + acc_row(r, h), i = 32 * b + col;
                if (o < nout && i < nin) part[woff + o * nin + i] = acc[r];
            }
        }
    if (lane < nout) {
        float sacc = 0.0f;
        for (int s = 0; s < 32; s++) sacc += sT[0][lane][s];
        part[boff + lane] = sacc;
    }
}

// A workgroup of kTrainWaves wavefronts; wavefront w of workgroup g takes 32 samples,
// base + 32*(kTrainWaves*g + w) + (0..31) of the batch, each at index perm_index(.) of x/y
// (bits > 0) or at its own position (bits = 0).  The workgroup first copies both weight packs
// into LDS, where every wavefront reads its MFMA A operands.  Row kTrainWaves*g + w of part
// (kPartStride floats) receives the summed gradient over the wavefront's samples of the loss
// sum_s |y_s - t_s|^2 * inv_n5, and at kPartLoss the sum of its squared errors.
constexpr int kTrainWaves = 4;
__global__ void __launch_bounds__(64 * kTrainWaves) k_train_wave(int64_t nb, int64_t base, uint32_t n_all, int bits,
                                                                 uint64_t key, const float *__restrict__ x,
                                                                 const float *__restrict__ yt,
                                                                 const float *__restrict__ P,
                                                                 const float *__restrict__ gfw,
                                                                 const float *__restrict__ gbw, float inv_n5,
                                                                 float *__restrict__ part) {
    __shared__ float4 sW4[(kFwFloats + kBwFloats) / 4];
    __shared__ float sTw[kTrainWaves][2][64][33];
    static_assert(kFwFloats % 4 == 0 && kBwFloats % 4 == 0, "packs copy as float4");
    for (int i = threadIdx.x; i < kFwFloats / 4; i += blockDim.x) sW4[i] = reinterpret_cast<const float4 *>(gfw)[i];
    for (int i = threadIdx.x; i < kBwFloats / 4; i += blockDim.x)
        sW4[kFwFloats / 4 + i] = reinterpret_cast<const float4 *>(gbw)[i];
    const float *fw = reinterpret_cast<const float *>(sW4), *bw = fw + kFwFloats;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
    float(*sT)[64][33] = sTw[wave];
    const int64_t wg = (int64_t)kTrainWaves * blockIdx.x + wave;
    const int64_t s_in_batch = 32 * wg + col;
    const bool valid = s_in_batch < nb;
    uint32_t gi = 0;
    if (valid) gi = bits > 0 ? perm_index((uint32_t)(base + s_in_batch), n_all, bits, key) : (uint32_t)(base + s_in_batch);
    float *pw = part + wg * kPartStride;
    __syncthreads();
    // ---- forward
    f32x16 z1[2], z2[2], z3[2], y;
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) z1[t][r] = P[oB1 + 32 * t + acc_row(r, h)];
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
        const float b = valid ? x[8 * (int64_t)gi + 2 * kk + h] : 0.0f;
#pragma unroll
        for (int t = 0; t < 2; t++) z1[t] = mfma2(fw[fwL1 + (t * 4 + kk) * 64 + lane], b, z1[t]);
    }
    hidden_fwd(fw, fwL2, P + oB2, z1, z2, lane, h);
    hidden_fwd(fw, fwL3, P + oB3, z2, z3, lane, h);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = acc_row(r, h);
        y[r] = row < 5 ? P[oB4 + row] : 0.0f;
    }
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int r = 0; r < 16; r++) y = mfma2(fw[fwL4 + (16 * ti + r) * 64 + lane], silu_t(z3[ti][r]), y);
    // ---- output error
    f32x16 dy;
    float l2 = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = acc_row(r, h);
        dy[r] = 0.0f;
        if (row < 5 && valid) {
            const float e = y[r] - yt[5 * (int64_t)gi + row];
            dy[r] = 2.0f * e * inv_n5;
            l2 += e * e;
        }
    }
    // ---- output layer: dW4, db4 and dZ3
    to_lds(sT[0], &dy, 1, col, h, false);
    to_lds(sT[1], z3, 2, col, h, true);
    __syncthreads();
    layer_wgrad(sT, 5, 64, oW4, oB4, pw, lane, col, h);
    f32x16 dz3[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
#pragma unroll
        for (int r = 0; r < 16; r++) dz3[t][r] = 0.0f;
#pragma unroll
        for (int r = 0; r < 4; r++) dz3[t] = mfma2(bw[bw4 + (t * 4 + r) * 64 + lane], dy[r], dz3[t]);
#pragma unroll
        for (int r = 0; r < 16; r++) dz3[t][r] *= dsilu_t(z3[t][r]);
    }
    // ---- layer 3
    __syncthreads();
    to_lds(sT[0], dz3, 2, col, h, false);
    to_lds(sT[1], z2, 2, col, h, true);
    __syncthreads();
    layer_wgrad(sT, 64, 64, oW3, oB3, pw, lane, col, h);
    f32x16 dz2[2];
    hidden_bwd(bw, bw3, dz3, z2, dz2, lane);
    // ---- layer 2
    __syncthreads();
    to_lds(sT[0], dz2, 2, col, h, false);
    to_lds(sT[1], z1, 2, col, h, true);
    __syncthreads();
    layer_wgrad(sT, 64, 64, oW2, oB2, pw, lane, col, h);
    f32x16 dz1[2];
    hidden_bwd(bw, bw2, dz2, z1, dz1, lane);
    // ---- layer 1 (inputs from x)
    __syncthreads();
    to_lds(sT[0], dz1, 2, col, h, false);
    for (int i = h; i < 32; i += 2) sT[1][i][col] = (i < 8 && valid) ? x[8 * (int64_t)gi + i] : 0.0f;
    __syncthreads();
    layer_wgrad(sT, 64, 8, oW1, oB1, pw, lane, col, h);
    // ---- squared error of the wavefront
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l2 += __shfl_xor(l2, off, 64);
    if (lane == 0) pw[kPartLoss] = l2;
}

// Sums the wavefronts' gradient partials: column p of part, over rows w = q (mod 4) for q = 0..3 in
// row order, then the four sums in q order (a fixed order, so a batch's gradient is reproducible).
// Then, per parameter: optionally stores the gradient in grad (accum != 0: adds to it), takes the
// Adam step on P (adam != 0) and refreshes the training packs; column kPartLoss adds the batch
// loss to *loss_acc.  256 threads per block, 64 columns per block.
__global__ void __launch_bounds__(256) k_reduce_adam(int nw, const float *__restrict__ part, float inv_n5,
                                                     float *__restrict__ grad, int accum, int adam,
                                                     float *__restrict__ P, float *__restrict__ m,
                                                     float *__restrict__ v, float lr, float b1, float b2, float eps,
                                                     float c1, float c2, float *__restrict__ fw,
                                                     float *__restrict__ bw, double *__restrict__ loss_acc) {
    __shared__ float red[4][64];
    const int q = threadIdx.x >> 6, c = threadIdx.x & 63;
    const int p = blockIdx.x * 64 + c;
    float g = 0.0f;
#pragma unroll 8
    for (int w = q; w < nw; w += 4) g += part[(int64_t)w * kPartStride + p];
    red[q][c] = g;
    __syncthreads();
    if (q != 0) return;
    g = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
    if (p == kPartLoss) {
        loss_acc[0] += (double)g * (double)inv_n5;
        return;
    }
    if (p >= kMlpParams) return;
    if (grad) grad[p] = accum ? grad[p] + g : g;
    if (adam) {
        const float mi = b1 * m[p] + (1.0f - b1) * g;
        const float vi = b2 * v[p] + (1.0f - b2) * g * g;
        m[p] = mi;
        v[p] = vi;
        const float np = P[p] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        P[p] = np;
        int fi, bi;
        pack_slots(p, fi, bi);
        if (fi >= 0) fw[fi] = np;
        if (bi >= 0) bw[bi] = np;
    }
}

// Device staging buffer of a call, freed on every return path.
struct StageBuf {
    float *p = nullptr;
    ~StageBuf() {
        if (p) (void)hipFree(p);
    }
};

static int ensure_train(MlpState &s, int64_t max_waves) {
    if (!s.d_fw) {
        HIPCHK_T(hipMalloc((void **)&s.d_fw, sizeof(float) * kFwFloats));
        HIPCHK_T(hipMalloc((void **)&s.d_bw, sizeof(float) * kBwFloats));
        HIPCHK_T(hipMalloc((void **)&s.d_lacc, sizeof(double)));
    }
    if (s.part_waves < max_waves) {
        if (s.d_part) hipFree(s.d_part);
        s.d_part = nullptr;
        s.part_waves = 0;
        HIPCHK_T(hipMalloc((void **)&s.d_part, sizeof(float) * kPartStride * max_waves));
        s.part_waves = max_waves;
    }
    return TITAN_OK;
}

static void launch_pack(MlpState &s, hipStream_t st) {
    // one block: the kernel zeroes every slot before any is written
    hipLaunchKernelGGL(k_pack_train, dim3(1), dim3(1024), 0, st, s.d_params, s.d_fw, s.d_bw);
}

int mlp_loss_grad_host(MlpState &s, hipStream_t st, int64_t n, const float *x, const float *y, float *loss,
                       float *grad) {
    if (!s.d_params) return TITAN_E_NOT_TRAINED;
    if (n <= 0 || n > INT32_MAX) return TITAN_E_ARG;
    const int64_t chunk = 32 * 4096;  // samples per gradient pass (bounds the partial buffer)
    const int64_t max_w = kTrainWaves * ((std::min(n, chunk) + 32 * kTrainWaves - 1) / (32 * kTrainWaves));
    int rc = ensure_train(s, max_w);
    if (rc != TITAN_OK) return rc;
    StageBuf bx, by;
    HIPCHK_T(hipMalloc((void **)&bx.p, sizeof(float) * 8 * n));
    HIPCHK_T(hipMalloc((void **)&by.p, sizeof(float) * 5 * n));
    float *dx = bx.p, *dy = by.p;
    HIPCHK_T(hipMemcpyAsync(dx, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice, st));
    HIPCHK_T(hipMemcpyAsync(dy, y, sizeof(float) * 5 * n, hipMemcpyHostToDevice, st));
    HIPCHK_T(hipMemsetAsync(s.d_lacc, 0, sizeof(double), st));
    launch_pack(s, st);
    const float inv_n5 = 1.0f / (float)(5 * n);
    for (int64_t b0 = 0; b0 < n; b0 += chunk) {
        const int64_t nb = std::min(chunk, n - b0), nw = (nb + 31) / 32;
        const unsigned ng = (unsigned)((nw + kTrainWaves - 1) / kTrainWaves);
        hipLaunchKernelGGL(k_train_wave, dim3(ng), dim3(64 * kTrainWaves), 0, st, nb, b0, (uint32_t)n, 0, 0ull, dx, dy,
                           s.d_params, s.d_fw, s.d_bw, inv_n5, s.d_part);
        hipLaunchKernelGGL(k_reduce_adam, dim3(kPartStride / 64), dim3(256), 0, st, (int)nw, s.d_part, inv_n5,
                           s.d_grad, b0 > 0 ? 1 : 0, 0, s.d_params, s.d_m, s.d_v, 0.0f,
                           0.0f, 0.0f, 0.0f, 1.0f, 1.0f, s.d_fw, s.d_bw, s.d_lacc);
    }
    HIPCHK_T(hipGetLastError());
    HIPCHK_T(hipMemcpyAsync(grad, s.d_grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st));
    double lsum = 0.0;
    HIPCHK_T(hipMemcpyAsync(&lsum, s.d_lacc, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK_T(hipStreamSynchronize(st));
    *loss = (float)lsum;
    return TITAN_OK;
}

int mlp_train_host(MlpState &s, hipStream_t st, int64_t n, const float *x, const float *y, int epochs, int batch,
                   float lr, uint64_t seed, float *loss_hist) {
    if (!s.d_params) return TITAN_E_NOT_TRAINED;
    if (n <= 0 || n > (int64_t)1 << 30) return TITAN_E_ARG;
    if (batch > n) batch = (int)n;
    const int64_t nw = (batch + 31) / 32;
    const unsigned ng = (unsigned)((nw + kTrainWaves - 1) / kTrainWaves);
    int rc = ensure_train(s, (int64_t)ng * kTrainWaves);
    if (rc != TITAN_OK) return rc;
    StageBuf bx, by;
    HIPCHK_T(hipMalloc((void **)&bx.p, sizeof(float) * 8 * n));
    HIPCHK_T(hipMalloc((void **)&by.p, sizeof(float) * 5 * n));
    float *dx = bx.p, *dy = by.p;
    HIPCHK_T(hipMemcpyAsync(dx, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice, st));
    HIPCHK_T(hipMemcpyAsync(dy, y, sizeof(float) * 5 * n, hipMemcpyHostToDevice, st));
    launch_pack(s, st);
    int bits = 2;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)n) bits += 2;
    const float inv_n5 = 1.0f / (float)(5 * (int64_t)batch);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int64_t nbatches = n / batch;
    for (int ep = 0; ep < epochs; ep++) {
        const uint64_t key = (seed + 0x9E3779B97F4A7C15ull * (uint64_t)(ep + 1)) * 0xD1B54A32D192ED03ull;
        HIPCHK_T(hipMemsetAsync(s.d_lacc, 0, sizeof(double), st));
        for (int64_t bi = 0; bi < nbatches; bi++) {
            hipLaunchKernelGGL(k_train_wave, dim3(ng), dim3(64 * kTrainWaves), 0, st, (int64_t)batch, bi * batch,
                               (uint32_t)n, bits, key, dx, dy, s.d_params, s.d_fw, s.d_bw, inv_n5, s.d_part);
            s.adam_t++;
            const float c1 = 1.0f - std::pow(b1, (float)s.adam_t), c2 = 1.0f - std::pow(b2, (float)s.adam_t);
            hipLaunchKernelGGL(k_reduce_adam, dim3(kPartStride / 64), dim3(256), 0, st, (int)nw, s.d_part, inv_n5,
                               (float *)nullptr, 0, 1, s.d_params, s.d_m, s.d_v, lr, b1, b2,
                               eps, c1, c2, s.d_fw, s.d_bw, s.d_lacc);
        }
        HIPCHK_T(hipGetLastError());
        if (loss_hist) {
            double l = 0.0;
            HIPCHK_T(hipMemcpyAsync(&l, s.d_lacc, sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK_T(hipStreamSynchronize(st));
            loss_hist[ep] = (float)(l / (double)std::max<int64_t>(1, nbatches));
        }
    }
    HIPCHK_T(hipStreamSynchronize(st));
    return mlp_repack(s, st);
}

}  // namespace titan
