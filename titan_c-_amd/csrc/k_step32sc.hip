// k_step32sc.hip -- the mixed-precision fused step with the passive scalars Q and rho_He carried
// (built as k_step32.hip is); its own translation unit so that the step variants compile in
// parallel.
#define TITAN_FLUX_ONLY 1
#include "step_kernels.h"
#include "tile_kernel.h"
#include "launch.h"

namespace titan {

void launch_step_tile_sc32(hipStream_t s, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                           const double *prim1, double *prim2, const double *sc1, double *sc2, const double *X1,
                           double *X2, const double *disp, const double *cc1, double *cc2, const double *src,
                           const SurArgs *sur, StepScalars *S, long long step) {
    launch_tile_variants<float, true>(s, TM, M, P, prim1, prim2, sc1, sc2, X1, X2, disp, cc1, cc2, src, sur, S, step);
}

}  // namespace titan
