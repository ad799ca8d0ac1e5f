// k_mlp.hip -- charge-exchange surrogate: BF16 MFMA inference (FP32 accumulation), parameter
// state and packing.  The wavefront forward pass lives in mlp_device.h, training in k_train.hip.
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
    void *ptrs[] = {s.d_params, s.d_grad, s.d_m, s.d_v,    s.d_wpack, s.d_bpack, s.d_norm,  s.d_list,
                    s.d_count,  s.d_stash, s.d_fw, s.d_bw, s.d_part,  s.d_lacc,  s.d_fixpart};
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

int mlp_repack(MlpState &s, hipStream_t st) {
    std::vector<float> P(kMlpParams);
    HIPCHK_M(hipMemcpyAsync(P.data(), s.d_params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st));
    HIPCHK_M(hipStreamSynchronize(st));
    std::vector<uint16_t> wv;
    std::vector<float> bv;
    pack_params(P.data(), wv, bv);
    HIPCHK_M(hipMemcpyAsync(s.d_wpack, wv.data(), wv.size() * 2, hipMemcpyHostToDevice, st));
    HIPCHK_M(hipMemcpyAsync(s.d_bpack, bv.data(), bv.size() * 4, hipMemcpyHostToDevice, st));
    HIPCHK_M(hipStreamSynchronize(st));
    s.ready = true;
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
    return mlp_repack(s, st);
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
    if (!s.d_fixpart) HIPCHK_M(hipMalloc((void **)&s.d_fixpart, sizeof(double) * 3 * kFixPar * kFixSlices));
    a->wpack = s.d_wpack;
    a->bpack = s.d_bpack;
    a->norm = s.d_norm;
    a->neutrals = neutrals;
    a->list = s.d_list;
    a->stash = s.d_stash;
    a->cap = s.list_cap;
    a->fixpart = s.d_fixpart;
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
