// k_step64.hip -- FP64 step kernels (compiled with -ffp-contract=off: reference operation order,
// no fused multiply-add, IEEE division and square root).
#include "step_kernels.h"
#include "tile_kernel.h"
#include "launch.h"

namespace titan {

static inline unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_begin_step(hipStream_t s, StepScalars *S, long long n, const DevParams &P, int restart) {
    hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(64), 0, s, S, n, P, restart);
}

void launch_motion(hipStream_t s, int64_t n3, const double *x_old, const double *disp, const StepScalars *S,
                   double *x_new) {
    if (n3 <= 0) return;
    hipLaunchKernelGGL(k_motion, dim3(grid_of(n3, 256)), dim3(256), 0, s, n3, x_old, disp, S, x_new);
}

void launch_measure(hipStream_t s, const DevMesh &M, const double *X, const DevGeom &G, int gcl) {
    // order matters: cell centres -> faces (orientation uses the centre of cells[0]) -> volumes
    hipLaunchKernelGGL(k_cell_center, dim3(grid_of(M.n_cells, 256)), dim3(256), 0, s, M, X, G);
    hipLaunchKernelGGL(k_face_geom, dim3(grid_of(M.n_faces, 256)), dim3(256), 0, s, M, X, G, gcl);
    hipLaunchKernelGGL(k_cell_volume, dim3(grid_of(M.n_cells, 256)), dim3(256), 0, s, M, X, G, gcl);
}

void launch_face_flux(hipStream_t s, int precision, const DevMesh &M, const DevParams &P, const double *prim,
                      const double *scal, const DevGeom &G, const double *X1, const double *X2, StepScalars *S,
                      double *flux, double *kflux, uint8_t *branch, long long step, double dt_override) {
    if (precision == 1) {
        launch_face_flux_f32(s, M, P, prim, scal, G, X1, X2, S, flux, kflux, branch, step, dt_override);
        return;
    }
    hipLaunchKernelGGL(k_face_flux<double>, dim3(grid_of(M.n_faces, 256)), dim3(256), 0, s, M, P, prim, scal, G, X1,
                       X2, S, flux, kflux, branch, step, dt_override);
}

void launch_step_tile(hipStream_t s, int precision, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                      const double *prim1, double *prim2, const double *sc1, double *sc2, const double *X1, double *X2,
                      const double *disp, const double *cc1, double *cc2, const double *src, const SurArgs *sur,
                      StepScalars *S, long long step) {
    if (P.n_scal) {
        if (precision == 1) launch_step_tile_sc32(s, TM, M, P, prim1, prim2, sc1, sc2, X1, X2, disp, cc1, cc2, src, sur, S, step);
        else launch_step_tile_sc64(s, TM, M, P, prim1, prim2, sc1, sc2, X1, X2, disp, cc1, cc2, src, sur, S, step);
        return;
    }
    if (precision == 1) {
        launch_step_tile_f32(s, TM, M, P, prim1, prim2, X1, X2, disp, cc1, cc2, src, sur, S, step);
        return;
    }
    launch_tile_variants<double, false>(s, TM, M, P, prim1, prim2, nullptr, nullptr, X1, X2, disp, cc1, cc2, src, sur, S,
                                        step);
}

size_t step_tile_lds_bytes(const TileMesh &TM, bool moving, bool scal) {
    return tile_lds_bytes(TM.max_nodes, TM.tile_cells, TM.max_faces, moving, scal);
}

// ---- standalone Riemann batch ----------------------------------------------------------------
template <typename T, int K>
__global__ void __launch_bounds__(256) k_chlld_batch(int64_t n, const int8_t *__restrict__ n_state,
                                                     const int8_t *__restrict__ n_disc,
                                                     const uint8_t *__restrict__ flags,
                                                     const double *__restrict__ normal,
                                                     const double *__restrict__ w, const double *__restrict__ qL,
                                                     const double *__restrict__ qR, const double *__restrict__ kL,
                                                     const double *__restrict__ kR, double *__restrict__ q,
                                                     double *__restrict__ k, double *__restrict__ ds,
                                                     uint8_t *__restrict__ branch, int *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T a[8], b[8], ka[K > 0 ? K : 1], kb[K > 0 ? K : 1];
#pragma unroll
    for (int j = 0; j < 8; j++) { a[j] = (T)qL[8 * i + j]; b[j] = (T)qR[8 * i + j]; }
#pragma unroll
    for (int j = 0; j < K; j++) { ka[j] = (T)kL[K * i + j]; kb[j] = (T)kR[K * i + j]; }
    RiemannOut<T, K> o;
    chlld<T, K>(n_state[i], n_disc ? n_disc[i] : 0, flags ? flags[i] : 0, (T)normal[3 * i], (T)normal[3 * i + 1],
                (T)normal[3 * i + 2], w ? (T)w[i] : T(0), a, b, ka, kb, o);
#pragma unroll
    for (int j = 0; j < 8; j++) q[8 * i + j] = (double)o.q[j];
#pragma unroll
    for (int j = 0; j < K; j++) k[K * i + j] = o.ok ? (double)o.k[j] : 0.0;
    if (ds) { ds[3 * i] = (double)o.dsl; ds[3 * i + 1] = (double)o.dsc; ds[3 * i + 2] = (double)o.dsr; }
    if (branch) branch[i] = (uint8_t)o.branch;
    if (!o.ok) atomicExch(err, 1);
}

void launch_chlld_batch(hipStream_t s, int precision, int64_t n, int K, const int8_t *n_state, const int8_t *n_disc,
                        const uint8_t *flags, const double *normal, const double *w, const double *qL,
                        const double *qR, const double *kL, const double *kR, double *q, double *k, double *ds,
                        uint8_t *branch, int *err) {
    if (n <= 0) return;
    if (precision == 1) {
        launch_chlld_batch_f32(s, n, K, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err);
        return;
    }
    const dim3 g(grid_of(n, 256)), b(256);
    switch (K) {
    case 0: hipLaunchKernelGGL((k_chlld_batch<double, 0>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    case 1: hipLaunchKernelGGL((k_chlld_batch<double, 1>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    case 2: hipLaunchKernelGGL((k_chlld_batch<double, 2>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    case 3: hipLaunchKernelGGL((k_chlld_batch<double, 3>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    default: hipLaunchKernelGGL((k_chlld_batch<double, 4>), g, b, 0, s, n, n_state, n_disc, flags, normal, w, qL, qR, kL, kR, q, k, ds, branch, err); break;
    }
}

}  // namespace titan
