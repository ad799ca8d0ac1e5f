// k_mlp.hip -- charge-exchange surrogate: BF16 MFMA inference (FP32 accumulation) and FP32
// training kernels (forward, backward, Adam).  The wavefront forward pass lives in mlp_device.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/titan_gpu.h"
#include "cx_device.h"
#include "mlp.h"
#include "mlp_device.h"

namespace titan {

#define HIPCHK_M(expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            std::fprintf(stderr, "titan_gpu(mlp): %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return TITAN_E_HIP;                                                               \
        }                                                                                     \
    } while (0)

// Gathers a pass's outputs: after the pass for cells 32p..32p+31, lane l < 32 holds outputs 0..3
// of cell 32p+l (registers 0..3) and lane l >= 32 holds output 4 of cell 32p+l-32 (register 0).
__device__ __forceinline__ void stash_outputs(const f32x16 &o, int lane, int pass, float (*sOut)[5]) {
    const int cell = 32 * pass + (lane & 31);
    if (lane < 32) {
        sOut[cell][0] = o[0]; sOut[cell][1] = o[1]; sOut[cell][2] = o[2]; sOut[cell][3] = o[3];
    } else {
        sOut[cell][4] = o[0];
    }
}

// Forward pass of normalised inputs (tests and training diagnostics): x [n*8] -> y [n*5].
__global__ void __launch_bounds__(256) k_mlp_forward(int64_t n, const float *__restrict__ x,
                                                     const uint16_t *__restrict__ wpack,
                                                     const float *__restrict__ bpack, float *__restrict__ y) {
    __shared__ bf16x8 sW[kMlpPackVec];
    __shared__ float sB[kMlpPackBias];
    __shared__ float sOut[4][64][5];
    load_packed(wpack, bpack, sW, sB);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * 64;
    if (base >= n) return;  // whole wavefront exits together
    for (int pass = 0; pass < 2; pass++) {
        const int64_t cell = base + 32 * pass + (lane & 31);
        bf16x8 in;
#pragma unroll
        for (int e = 0; e < 8; e++) in[e] = (__bf16)0.0f;
        if (lane < 32 && cell < n) {
#pragma unroll
            for (int e = 0; e < 8; e++) in[e] = (__bf16)x[8 * cell + e];
        }
        const f32x16 o = mlp_pass(sW, sB, in, lane);
        stash_outputs(o, lane, pass, sOut[wave]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const int64_t cell = base + lane;
    if (cell < n) {
#pragma unroll
        for (int k = 0; k < 5; k++) y[5 * cell + k] = sOut[wave][lane][k];
    }
}

// Surrogate source for every cell.  norm = in_mean[8] in_std[8] out_scale[5] lo[8] hi[8].
__global__ void __launch_bounds__(256) k_mlp_cells(int64_t nc, const double *__restrict__ prim,
                                                   const double *__restrict__ neutrals,
                                                   const uint16_t *__restrict__ wpack,
                                                   const float *__restrict__ bpack, const float *__restrict__ norm,
                                                   double *__restrict__ src, int32_t *__restrict__ list,
                                                   unsigned long long *count, int64_t cap,
                                                   unsigned long long *fallback_total) {
    __shared__ bf16x8 sW[kMlpPackVec];
    __shared__ float sB[kMlpPackBias];
    __shared__ float sN[kMlpNorm];
    __shared__ float sOut[4][64][5];
    load_packed(wpack, bpack, sW, sB);
    for (int i = threadIdx.x; i < kMlpNorm; i += blockDim.x) sN[i] = norm[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * 64;
    if (base >= nc) return;
    // this lane's own cell (pass lane/32) -- its raw inputs are kept for the output stage
    const int64_t mycell = base + lane;
    double st[8] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0}, hn[5] = {0.0, 1.0, 0.0, 0.0, 0.0};
    bool inrange = false;
    float xr[8];
    if (mycell < nc) {
#pragma unroll
        for (int k = 0; k < 8; k++) st[k] = prim[8 * mycell + k];
#pragma unroll
        for (int k = 0; k < 5; k++) hn[k] = neutrals[5 * mycell + k];
        mlp_features(st, hn, xr);
        inrange = mlp_in_box(xr, sN);
    } else {
#pragma unroll
        for (int e = 0; e < 8; e++) xr[e] = 0.0f;
    }
    for (int pass = 0; pass < 2; pass++) {
        // lanes l and l+32 both handle cell base + 32*pass + l%32; the owner of those raw inputs
        // is lane 32*pass + l%32, fetched by a cross-lane read
        const int src_lane = 32 * pass + (lane & 31);
        bf16x8 in;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const float v = __shfl(xr[e], src_lane, 64);
            in[e] = (lane < 32) ? mlp_normalise(v, sN, e) : (__bf16)0.0f;
        }
        const f32x16 o = mlp_pass(sW, sB, in, lane);
        stash_outputs(o, lane, pass, sOut[wave]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    if (mycell >= nc) return;
    float y[5];
#pragma unroll
    for (int k = 0; k < 5; k++) y[k] = sOut[wave][lane][k];
    double o[5];
    mlp_source(y, sN, st, hn, o);
#pragma unroll
    for (int k = 0; k < 5; k++) src[5 * mycell + k] = o[k];
    if (!inrange) {
        const unsigned long long k = atomicAdd(count, 1ull);
        atomicAdd(fallback_total, 1ull);
        if ((int64_t)k < cap) list[k] = (int32_t)mycell;
    }
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

// ---- training kernels (FP32) ---------------------------------------------------------------
// layer forward: thread per (sample, output); act: 1 = SiLU
__global__ void k_fwd_layer(int64_t n, int nin, int nout, const float *__restrict__ W, const float *__restrict__ b,
                            const float *__restrict__ ain, float *__restrict__ z, float *__restrict__ aout, int act) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * nout) return;
    const int64_t s = i / nout;
    const int o = (int)(i % nout);
    float acc = b[o];
    for (int k = 0; k < nin; k++) acc += W[o * nin + k] * ain[s * nin + k];
    z[i] = acc;
    aout[i] = act ? acc / (1.0f + expf(-acc)) : acc;
}

// dY = 2 (Y - T) / (n*5); per-block partial loss
__global__ void k_out_grad(int64_t n, const float *__restrict__ yp, const float *__restrict__ yt,
                           float *__restrict__ dy, float *__restrict__ loss_part) {
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < n * 5) {
        const float e = yp[i] - yt[i];
        dy[i] = 2.0f * e / (float)(n * 5);
        l = e * e;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// backward through a layer's weights and the previous activation:
// dz_prev[s][k] = (sum_o dz[s][o] W[o][k]) * silu'(z_prev[s][k])
__global__ void k_bwd_data(int64_t n, int nin, int nout, const float *__restrict__ W, const float *__restrict__ dz,
                           const float *__restrict__ zprev, float *__restrict__ dzprev) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * nin) return;
    const int64_t s = i / nin;
    const int k = (int)(i % nin);
    float acc = 0.0f;
    for (int o = 0; o < nout; o++) acc += dz[s * nout + o] * W[o * nin + k];
    const float zz = zprev[i];
    const float sg = 1.0f / (1.0f + expf(-zz));
    dzprev[i] = acc * (sg * (1.0f + zz * (1.0f - sg)));
}

// partial weight gradient over a chunk of 256 samples: part[chunk][o*nin + k] and bias part
__global__ void k_wgrad_part(int64_t n, int nin, int nout, const float *__restrict__ dz, const float *__restrict__ ain,
                             float *__restrict__ part) {
    __shared__ float sdz[64][64];
    __shared__ float sa[64][65];
    const int64_t s0 = (int64_t)blockIdx.x * 256;
    const int np = nout * nin + nout;
    float acc[17];
#pragma unroll
    for (int j = 0; j < 17; j++) acc[j] = 0.0f;
    for (int sub = 0; sub < 4; sub++) {
        const int64_t sb = s0 + 64 * sub;
        __syncthreads();
        for (int idx = threadIdx.x; idx < 64 * 64; idx += blockDim.x) {
            const int r = idx / 64, cidx = idx % 64;
            const int64_t s = sb + r;
            sdz[r][cidx] = (s < n && cidx < nout) ? dz[s * nout + cidx] : 0.0f;
            sa[r][cidx] = (s < n && cidx < nin) ? ain[s * nin + cidx] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 17; j++) {
            const int p = threadIdx.x + 256 * j;
            if (p >= np) break;
            float a = 0.0f;
            if (p < nout * nin) {
                const int o = p / nin, k = p % nin;
                for (int r = 0; r < 64; r++) a += sdz[r][o] * sa[r][k];
            } else {
                const int o = p - nout * nin;
                for (int r = 0; r < 64; r++) a += sdz[r][o];
            }
            acc[j] += a;
        }
    }
#pragma unroll
    for (int j = 0; j < 17; j++) {
        const int p = threadIdx.x + 256 * j;
        if (p < np) part[(int64_t)blockIdx.x * np + p] = acc[j];
    }
}

// grad[off + p] (weights) / grad[boff + o] (bias) = sum over chunks, fixed order
__global__ void k_wgrad_reduce(int64_t chunks, int nin, int nout, const float *__restrict__ part, float *__restrict__ grad,
                               int woff, int boff) {
    const int np = nout * nin + nout;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    float a = 0.0f;
    for (int64_t c = 0; c < chunks; c++) a += part[c * np + p];
    if (p < nout * nin) grad[woff + p] = a;
    else grad[boff + (p - nout * nin)] = a;
}

__global__ void k_adam(int np, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- host side -----------------------------------------------------------------------------
static uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Builds the MFMA operand packs from FP32 parameters (logical layout).
static void pack_params(const float *P, std::vector<uint16_t> &wv, std::vector<float> &bv) {
    wv.assign((size_t)kMlpPackVec * 8, 0);
    bv.assign(kMlpPackBias, 0.0f);
    // logical row held at physical row p of a hidden layer: inverse of (t,i,h) -> p
    int logical_of_phys[64];
    for (int t = 0; t < 2; t++)
        for (int i = 0; i < 16; i++)
            for (int h = 0; h < 2; h++) logical_of_phys[32 * t + tile_row(i, h)] = hidden_logical(t, i, h);
    struct L { int vbase, bbase, tiles, kbs, nin, nout, woff, boff; bool hidden; };
    const L layers[4] = {{vL1, bL1, 2, 1, 8, 64, oW1, oB1, true}, {vL2, bL2, 2, 4, 64, 64, oW2, oB2, true},
                         {vL3, bL3, 2, 4, 64, 64, oW3, oB3, true}, {vL4, bL4, 1, 4, 64, 5, oW4, oB4, false}};
    for (const L &ly : layers) {
        for (int t = 0; t < ly.tiles; t++)
            for (int kb = 0; kb < ly.kbs; kb++)
                for (int lane = 0; lane < 64; lane++) {
                    const int p = 32 * t + (lane & 31);
                    const int o = ly.hidden ? logical_of_phys[p] : p;  // output layer: row p = output p
                    for (int e = 0; e < 8; e++) {
                        const int k = 16 * kb + 8 * (lane >> 5) + e;
                        float val = 0.0f;
                        if (o < ly.nout && k < ly.nin) val = P[ly.woff + o * ly.nin + k];
                        wv[(size_t)(ly.vbase + (t * ly.kbs + kb) * 64 + lane) * 8 + e] = f2bf(val);
                    }
                }
        for (int t = 0; t < ly.tiles; t++)
            for (int h = 0; h < 2; h++)
                for (int i = 0; i < 16; i++) {
                    const int p = 32 * t + tile_row(i, h);
                    const int o = ly.hidden ? logical_of_phys[p] : p;
                    bv[ly.bbase + (t * 2 + h) * 16 + i] = (o < ly.nout) ? P[ly.boff + o] : 0.0f;
                }
    }
}

void mlp_free(MlpState &s) {
    void *ptrs[] = {s.d_params, s.d_grad, s.d_m, s.d_v, s.d_wpack, s.d_bpack, s.d_norm, s.d_list, s.d_count, s.d_ws,
                    s.d_stash};
    for (void *p : ptrs)
        if (p) hipFree(p);
    s = MlpState();
}

static int ensure_alloc(MlpState &s) {
    if (!s.d_params) {
        HIPCHK_M(hipMalloc((void **)&s.d_params, sizeof(float) * kMlpParams));
        HIPCHK_M(hipMalloc((void **)&s.d_grad, sizeof(float) * kMlpParams));
        HIPCHK_M(hipMalloc((void **)&s.d_m, sizeof(float) * kMlpParams));
        HIPCHK_M(hipMalloc((void **)&s.d_v, sizeof(float) * kMlpParams));
        HIPCHK_M(hipMalloc((void **)&s.d_wpack, sizeof(uint16_t) * kMlpPackVec * 8));
        HIPCHK_M(hipMalloc((void **)&s.d_bpack, sizeof(float) * kMlpPackBias));
        HIPCHK_M(hipMalloc((void **)&s.d_norm, sizeof(float) * kMlpNorm));
        HIPCHK_M(hipMalloc((void **)&s.d_count, sizeof(unsigned long long)));
        HIPCHK_M(hipMemset(s.d_m, 0, sizeof(float) * kMlpParams));
        HIPCHK_M(hipMemset(s.d_v, 0, sizeof(float) * kMlpParams));
    }
    return TITAN_OK;
}

static int repack(MlpState &s, hipStream_t st) {
    std::vector<float> P(kMlpParams);
    HIPCHK_M(hipMemcpyAsync(P.data(), s.d_params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st));
    HIPCHK_M(hipStreamSynchronize(st));
    std::vector<uint16_t> wv;
    std::vector<float> bv;
    pack_params(P.data(), wv, bv);
    HIPCHK_M(hipMemcpyAsync(s.d_wpack, wv.data(), wv.size() * 2, hipMemcpyHostToDevice, st));
    HIPCHK_M(hipMemcpyAsync(s.d_bpack, bv.data(), bv.size() * 4, hipMemcpyHostToDevice, st));
    HIPCHK_M(hipStreamSynchronize(st));
    return TITAN_OK;
}

int mlp_set(MlpState &s, hipStream_t st, const float *params, const float *in_norm16, const float *out_scale5,
            const float *box16) {
    int rc = ensure_alloc(s);
    if (rc != TITAN_OK) return rc;
    for (int i = 0; i < 8; i++)
        if (!(in_norm16[8 + i] > 0.0f)) return TITAN_E_ARG;
    std::memcpy(s.h_norm, in_norm16, 16 * sizeof(float));
    std::memcpy(s.h_norm + 16, out_scale5, 5 * sizeof(float));
    std::memcpy(s.h_norm + 21, box16, 16 * sizeof(float));
    HIPCHK_M(hipMemcpyAsync(s.d_params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st));
    HIPCHK_M(hipMemcpyAsync(s.d_norm, s.h_norm, sizeof(float) * kMlpNorm, hipMemcpyHostToDevice, st));
    HIPCHK_M(hipMemsetAsync(s.d_m, 0, sizeof(float) * kMlpParams, st));
    HIPCHK_M(hipMemsetAsync(s.d_v, 0, sizeof(float) * kMlpParams, st));
    s.adam_t = 0;
    rc = repack(s, st);
    if (rc != TITAN_OK) return rc;
    s.ready = true;
    return TITAN_OK;
}

int mlp_get(MlpState &s, hipStream_t st, float *params) {
    if (!s.d_params) return TITAN_E_NOT_TRAINED;
    HIPCHK_M(hipMemcpyAsync(params, s.d_params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st));
    HIPCHK_M(hipStreamSynchronize(st));
    return TITAN_OK;
}

int mlp_forward_host(MlpState &s, hipStream_t st, int64_t n, const float *x, float *y) {
    if (!s.ready) return TITAN_E_NOT_TRAINED;
    if (n == 0) return TITAN_OK;
    float *dx = nullptr, *dy = nullptr;
    HIPCHK_M(hipMalloc((void **)&dx, sizeof(float) * 8 * n));
    HIPCHK_M(hipMalloc((void **)&dy, sizeof(float) * 5 * n));
    HIPCHK_M(hipMemcpyAsync(dx, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice, st));
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(256), 0, st, n, dx, s.d_wpack, s.d_bpack, dy);
    HIPCHK_M(hipGetLastError());
    HIPCHK_M(hipMemcpyAsync(y, dy, sizeof(float) * 5 * n, hipMemcpyDeviceToHost, st));
    HIPCHK_M(hipStreamSynchronize(st));
    hipFree(dx);
    hipFree(dy);
    return TITAN_OK;
}

// workspace: activations of a batch (z and a for 4 layers), dz buffers, partials, loss parts
static int ensure_ws(MlpState &s, int64_t batch) {
    const int64_t chunks = (batch + 255) / 256;
    const size_t need = (size_t)batch * (8 + 2 * (64 * 3 + 5) + 64 * 3 + 5 + 5) + (size_t)chunks * (64 * 64 + 64) +
                        (size_t)((batch * 5 + 255) / 256) + 16;
    if (s.ws_floats < need) {
        if (s.d_ws) hipFree(s.d_ws);
        s.d_ws = nullptr;
        HIPCHK_M(hipMalloc((void **)&s.d_ws, need * sizeof(float)));
        s.ws_floats = need;
    }
    return TITAN_OK;
}

// Forward + backward of one batch already on the device; fills s.d_grad; returns mean loss.
static int batch_grad(MlpState &s, hipStream_t st, int64_t n, const float *dx, const float *dyt, float *loss_out) {
    const int64_t chunks = (n + 255) / 256;
    float *w = s.d_ws;
    float *z[4], *a[4], *dz[4];
    const int dims[5] = {8, 64, 64, 64, 5};
    for (int l = 0; l < 4; l++) { z[l] = w; w += n * dims[l + 1]; a[l] = w; w += n * dims[l + 1]; }
    for (int l = 0; l < 4; l++) { dz[l] = w; w += n * dims[l + 1]; }
    float *part = w; w += chunks * (64 * 64 + 64);
    float *lpart = w;
    const int woff[4] = {oW1, oW2, oW3, oW4}, boff[4] = {oB1, oB2, oB3, oB4};
    const float *ain = dx;
    for (int l = 0; l < 4; l++) {
        const int64_t tot = n * dims[l + 1];
        hipLaunchKernelGGL(k_fwd_layer, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, dims[l], dims[l + 1],
                           s.d_params + woff[l], s.d_params + boff[l], ain, z[l], a[l], l < 3 ? 1 : 0);
        ain = a[l];
    }
    const unsigned lblocks = (unsigned)((n * 5 + 255) / 256);
    hipLaunchKernelGGL(k_out_grad, dim3(lblocks), dim3(256), 0, st, n, a[3], dyt, dz[3], lpart);
    for (int l = 3; l >= 0; l--) {
        const float *ain_l = l == 0 ? dx : a[l - 1];
        hipLaunchKernelGGL(k_wgrad_part, dim3((unsigned)chunks), dim3(256), 0, st, n, dims[l], dims[l + 1], dz[l], ain_l,
                           part);
        const int np = dims[l] * dims[l + 1] + dims[l + 1];
        hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, chunks, dims[l],
                           dims[l + 1], part, s.d_grad, woff[l], boff[l]);
        if (l > 0) {
            const int64_t tot = n * dims[l];
            hipLaunchKernelGGL(k_bwd_data, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, dims[l],
                               dims[l + 1], s.d_params + woff[l], dz[l], z[l - 1], dz[l - 1]);
        }
    }
    HIPCHK_M(hipGetLastError());
    if (loss_out) {
        std::vector<float> lp(lblocks);
        HIPCHK_M(hipMemcpyAsync(lp.data(), lpart, sizeof(float) * lblocks, hipMemcpyDeviceToHost, st));
        HIPCHK_M(hipStreamSynchronize(st));
        double tot = 0.0;
        for (float v : lp) tot += v;
        *loss_out = (float)(tot / (double)(n * 5));
    }
    return TITAN_OK;
}

int mlp_loss_grad_host(MlpState &s, hipStream_t st, int64_t n, const float *x, const float *y, float *loss,
                       float *grad) {
    if (!s.d_params) return TITAN_E_NOT_TRAINED;
    int rc = ensure_ws(s, n);
    if (rc != TITAN_OK) return rc;
    float *dx = nullptr, *dy = nullptr;
    HIPCHK_M(hipMalloc((void **)&dx, sizeof(float) * 8 * n));
    HIPCHK_M(hipMalloc((void **)&dy, sizeof(float) * 5 * n));
    HIPCHK_M(hipMemcpyAsync(dx, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice, st));
    HIPCHK_M(hipMemcpyAsync(dy, y, sizeof(float) * 5 * n, hipMemcpyHostToDevice, st));
    rc = batch_grad(s, st, n, dx, dy, loss);
    if (rc == TITAN_OK) {
        HIPCHK_M(hipMemcpyAsync(grad, s.d_grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st));
        HIPCHK_M(hipStreamSynchronize(st));
    }
    hipFree(dx);
    hipFree(dy);
    return rc;
}

// gathers a shuffled batch: xb[i] = x[perm[i]]
__global__ void k_gather_batch(int64_t nb, const int32_t *__restrict__ perm, const float *__restrict__ x,
                               const float *__restrict__ y, float *__restrict__ xb, float *__restrict__ yb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int64_t j = perm[i];
#pragma unroll
    for (int k = 0; k < 8; k++) xb[8 * i + k] = x[8 * j + k];
#pragma unroll
    for (int k = 0; k < 5; k++) yb[5 * i + k] = y[5 * j + k];
}

int mlp_train_host(MlpState &s, hipStream_t st, int64_t n, const float *x, const float *y, int epochs, int batch,
                   float lr, uint64_t seed, float *loss_hist) {
    if (!s.d_params) return TITAN_E_NOT_TRAINED;
    if (batch > n) batch = (int)n;
    int rc = ensure_ws(s, batch);
    if (rc != TITAN_OK) return rc;
    float *dx = nullptr, *dy = nullptr, *xb = nullptr, *yb = nullptr;
    int32_t *dperm = nullptr;
    HIPCHK_M(hipMalloc((void **)&dx, sizeof(float) * 8 * n));
    HIPCHK_M(hipMalloc((void **)&dy, sizeof(float) * 5 * n));
    HIPCHK_M(hipMalloc((void **)&xb, sizeof(float) * 8 * batch));
    HIPCHK_M(hipMalloc((void **)&yb, sizeof(float) * 5 * batch));
    HIPCHK_M(hipMalloc((void **)&dperm, sizeof(int32_t) * n));
    HIPCHK_M(hipMemcpyAsync(dx, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice, st));
    HIPCHK_M(hipMemcpyAsync(dy, y, sizeof(float) * 5 * n, hipMemcpyHostToDevice, st));
    std::vector<int32_t> perm(n);
    uint64_t state = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {
        state ^= state << 13; state ^= state >> 7; state ^= state << 17;
        return state;
    };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs && rc == TITAN_OK; ep++) {
        for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
        for (int64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
        HIPCHK_M(hipMemcpyAsync(dperm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
        const int64_t nbatches = n / batch;
        double lsum = 0.0;
        for (int64_t bi = 0; bi < nbatches; bi++) {
            hipLaunchKernelGGL(k_gather_batch, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, st, (int64_t)batch,
                               dperm + bi * batch, dx, dy, xb, yb);
            float l = 0.0f;
            const bool want = loss_hist && (bi % 16 == 0);
            rc = batch_grad(s, st, batch, xb, yb, want ? &l : nullptr);
            if (rc != TITAN_OK) break;
            if (want) lsum += l;
            s.adam_t++;
            const float c1 = 1.0f - std::pow(b1, (float)s.adam_t), c2 = 1.0f - std::pow(b2, (float)s.adam_t);
            hipLaunchKernelGGL(k_adam, dim3((unsigned)((kMlpParams + 255) / 256)), dim3(256), 0, st, (int)kMlpParams,
                               s.d_params, s.d_grad, s.d_m, s.d_v, lr, b1, b2, eps, c1, c2);
        }
        if (loss_hist) loss_hist[ep] = (float)(lsum / (double)std::max<int64_t>(1, (nbatches + 15) / 16));
    }
    HIPCHK_M(hipStreamSynchronize(st));
    hipFree(dx); hipFree(dy); hipFree(xb); hipFree(yb); hipFree(dperm);
    if (rc != TITAN_OK) return rc;
    rc = repack(s, st);
    if (rc == TITAN_OK) s.ready = true;
    return rc;
}

static int ensure_list(MlpState &s, int64_t nc) {
    if (s.list_cap < nc) {
        if (s.d_list) hipFree(s.d_list);
        if (s.d_stash) hipFree(s.d_stash);
        s.d_list = nullptr;
        s.d_stash = nullptr;
        s.list_cap = 0;
        HIPCHK_M(hipMalloc((void **)&s.d_list, sizeof(int32_t) * nc));
        HIPCHK_M(hipMalloc((void **)&s.d_stash, sizeof(double) * 11 * nc));
        s.list_cap = nc;
    }
    return TITAN_OK;
}

int mlp_fused_args(MlpState &s, int64_t nc, const double *neutrals, SurArgs *a) {
    if (!s.ready) return TITAN_E_NOT_TRAINED;
    const int rc = ensure_list(s, nc);
    if (rc != TITAN_OK) return rc;
    a->wpack = s.d_wpack;
    a->bpack = s.d_bpack;
    a->norm = s.d_norm;
    a->neutrals = neutrals;
    a->list = s.d_list;
    a->stash = s.d_stash;
    a->cap = s.list_cap;
    return TITAN_OK;
}

int mlp_cell_sources(MlpState &s, hipStream_t st, int64_t nc, const double *prim, const double *neutrals, int nodes,
                     double *src, unsigned long long *fallback_total) {
    if (!s.ready) return TITAN_E_NOT_TRAINED;
    const int rc = ensure_list(s, nc);
    if (rc != TITAN_OK) return rc;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, s.d_count);
    const unsigned grid = (unsigned)((nc + 255) / 256);
    hipLaunchKernelGGL(k_mlp_cells, dim3(grid), dim3(256), 0, st, nc, prim, neutrals, s.d_wpack, s.d_bpack, s.d_norm, src,
                       s.d_list, s.d_count, s.list_cap, fallback_total);
    launch_cx_list(st, s.d_list, s.d_count, s.list_cap, prim, neutrals, nodes, src);
    HIPCHK_M(hipGetLastError());
    return TITAN_OK;
}

}  // namespace titan
