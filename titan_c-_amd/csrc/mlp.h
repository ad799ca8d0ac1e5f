// mlp.h -- charge-exchange surrogate network 8-64-64-64-5 (SiLU): host-side state and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace titan {

constexpr int kMlpIn = 8, kMlpH = 64, kMlpOut = 5;
constexpr int64_t kMlpParams = (int64_t)kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH * kMlpH + kMlpH +
                               kMlpOut * kMlpH + kMlpOut;  // 9221
// packed BF16 MFMA operands: 1408 lane-vectors of 8 bf16, and 224 permuted FP32 biases
constexpr int kMlpPackVec = 1408;
constexpr int kMlpPackBias = 224;
// normalisation block: in_mean[8], in_std[8], out_scale[5], box_lo[8], box_hi[8]
constexpr int kMlpNorm = 37;

// Inputs of the surrogate fused into the step kernel (device pointers, device cell order).
struct SurArgs {
    const uint16_t *wpack;   // packed BF16 operands
    const float *bpack;      // packed biases
    const float *norm;       // normalisation block
    const double *neutrals;  // [nc*5]: n_H, c_H^2, V_H
    int32_t *list;           // cells outside the training box this step (slot from S->n_fallback)
    double *stash;           // [cap*11]: flux sum Pf[9], V(now1), V(now2) of each listed cell
    int64_t cap;
    double *fixpart;         // [kFixPar*kFixSlices*3]: quadrature partial sums of the first listed cells
};

// The exact source of the first kFixPar listed cells is spread over kFixSlices workgroups each.
constexpr int kFixPar = 256, kFixSlices = 16;

struct MlpState {
    bool ready = false;
    float *d_params = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t *d_wpack = nullptr;
    float *d_bpack = nullptr;
    float *d_norm = nullptr;
    float h_norm[kMlpNorm] = {};
    int32_t *d_list = nullptr;
    int64_t list_cap = 0;
    double *d_stash = nullptr;  // fused-step fallback stash, list_cap * 11 doubles
    double *d_fixpart = nullptr;  // kFixPar * kFixSlices * 3 quadrature partial sums
    unsigned long long *d_count = nullptr;
    // training: FP32 MFMA operand packs of W (forward) and W^T (backward), per-wavefront gradient
    // partials (with their squared-error sums), the loss accumulator
    float *d_fw = nullptr, *d_bw = nullptr, *d_part = nullptr;
    double *d_lacc = nullptr;
    int64_t part_waves = 0;
    long long adam_t = 0;
};

void mlp_free(MlpState &s);
int mlp_set(MlpState &s, hipStream_t st, const float *params, const float *in_norm16, const float *out_scale5,
            const float *box16);
int mlp_get(MlpState &s, hipStream_t st, float *params);
// Rebuilds the BF16 inference packs from the FP32 parameters and marks the network ready.
int mlp_repack(MlpState &s, hipStream_t st);
int mlp_forward_host(MlpState &s, hipStream_t st, int64_t n, const float *x, float *y);
int mlp_loss_grad_host(MlpState &s, hipStream_t st, int64_t n, const float *x, const float *y, float *loss,
                       float *grad);
int mlp_train_host(MlpState &s, hipStream_t st, int64_t n, const float *x, const float *y, int epochs, int batch,
                   float lr, uint64_t seed, float *loss_hist);
// Surrogate source of every cell (device arrays in device numbering); out-of-range cells get the
// exact source.  fallback_total (device) accumulates the number of such cells.
int mlp_cell_sources(MlpState &s, hipStream_t st, int64_t nc, const double *prim, const double *neutrals, int nodes,
                     double *src, unsigned long long *fallback_total);

// Device pointers of the surrogate fused into the step kernel; allocates the fallback list and
// stash (capacity nc) on first use.
int mlp_fused_args(MlpState &s, int64_t nc, const double *neutrals, SurArgs *a);

void launch_cx_list(hipStream_t s, const int32_t *list, const unsigned long long *count, int64_t cap,
                    const double *prim, const double *neutrals, int nodes, double *src);

}  // namespace titan
