// k_step32.hip -- mixed-precision face flux: FP32 Riemann arithmetic (contracted, fast division
// and square root), FP64 states, geometry, area scaling and CFL step.
#define TITAN_FLUX_ONLY 1
#include "step_kernels.h"
#include "tile_kernel.h"
#include "launch.h"

namespace titan {

static inline unsigned grid_of32(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_face_flux_f32(hipStream_t s, const DevMesh &M, const DevParams &P, const double *prim,
                          const double *scal, const DevGeom &G, const double *X1, const double *X2, StepScalars *S,
                          double *flux, double *kflux, uint8_t *branch, long long step, double dt_override) {
    hipLaunchKernelGGL(k_face_flux<float>, dim3(grid_of32(M.n_faces, 256)), dim3(256), 0, s, M, P, prim, scal, G, X1,
                       X2, S, flux, kflux, branch, step, dt_override);
}

void launch_step_tile_f32(hipStream_t s, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                          const double *prim1, double *prim2, const double *X1, double *X2, const double *disp,
                          const double *cc1, double *cc2, const double *src, const SurArgs *sur, StepScalars *S,
                          long long step) {
    launch_tile_variants<float, false>(s, TM, M, P, prim1, prim2, nullptr, nullptr, X1, X2, disp, cc1, cc2, src, sur, S,
                                       step);
}

template <int K>
__global__ void __launch_bounds__(256) k_chlld_batch32(int64_t n, const int8_t *__restrict__ n_state,
                                                       const int8_t *__restrict__ n_disc,
                                                       const uint8_t *__restrict__ flags,
                                                       const double *__restrict__ normal,
                                                       const double *__restrict__ w, const double *__restrict__ qL,
                                                       const double *__restrict__ qR,
                                                       const double *__restrict__ kL,
                                                       const double *__restrict__ kR, double *__restrict__ q,
                                                       double *__restrict__ k, double *__restrict__ ds,
                                                       uint8_t *__restrict__ branch, int *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a[8], b[8], ka[K > 0 ? K : 1], kb[K > 0 ? K : 1];
#pragma unroll
    for (int j = 0; j < 8; j++) { a[j] = qL[8 * i + j]; b[j] = qR[8 * i + j]; }
#pragma unroll
    for (int j = 0; j < K; j++) { ka[j] = kL[K * i + j]; kb[j] = kR[K * i + j]; }
    RiemannOut<double, K> o;
    chlld_mixed<K>(n_state[i], n_disc ? n_disc[i] : 0, flags ? flags[i] : 0, normal[3 * i], normal[3 * i + 1],
                   normal[3 * i + 2], w ? w[i] : 0.0, a, b, ka, kb, o);
    if (o.ok < 0)  // ill-conditioned in FP32: the FP64 solver
        chlld_fp64_face<K>(n_state[i], n_disc ? n_disc[i] : 0, flags ? flags[i] : 0, normal[3 * i], normal[3 * i + 1],
                           normal[3 * i + 2], w ? w[i] : 0.0, a, b, ka, kb, &o);
#pragma unroll
    for (int j = 0; j < 8; j++) q[8 * i + j] = o.q[j];
#pragma unroll
    for (int j = 0; j < K; j++) k[K * i + j] = o.ok ? o.k[j] : 0.0;
    if (ds) { ds[3 * i] = o.dsl; ds[3 * i + 1] = o.dsc; ds[3 * i + 2] = o.dsr; }
    if (branch) branch[i] = (uint8_t)o.branch;
    if (!o.ok) atomicExch(err, 1);
}

void launch_chlld_batch_f32(hipStream_t s, int64_t n, int K, const int8_t *n_state, const int8_t *n_disc,
                            const uint8_t *flags, const double *normal, const double *w, const double *qL,
                            const double *qR, const double *kL, const double *kR, double *q, double *k, double *ds,
                            uint8_t *branch, int *err) {
    const dim3 g(grid_of32(n, 256)), b(256);
    switch (K) {
    case 0: hipLaunchKernelGGL((k_chlld_batch32<0>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    case 1: hipLaunchKernelGGL((k_chlld_batch32<1>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    case 2: hipLaunchKernelGGL((k_chlld_batch32<2>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    case 3: hipLaunchKernelGGL((k_chlld_batch32<3>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    default: hipLaunchKernelGGL((k_chlld_batch32<4>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    }
}

}  // namespace titan
