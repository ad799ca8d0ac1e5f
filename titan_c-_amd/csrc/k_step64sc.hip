// k_step64sc.hip -- the FP64 fused step with the passive scalars Q and rho_He carried (built as
// k_step64.hip is: -ffp-contract=off, IEEE division and square root); its own translation unit
// so that the two halves of the step variants compile in parallel.  The geometry and bookkeeping
// kernels are k_step64.hip's (TITAN_FLUX_ONLY leaves them out here).
#define TITAN_FLUX_ONLY 1
#include "step_kernels.h"
#include "tile_kernel.h"
#include "launch.h"

namespace titan {

void launch_step_tile_sc64(hipStream_t s, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                           const double *prim1, double *prim2, const double *sc1, double *sc2, const double *X1,
                           double *X2, const double *disp, const double *cc1, double *cc2, const double *src,
                           const SurArgs *sur, StepScalars *S, long long step) {
    launch_tile_variants<double, true>(s, TM, M, P, prim1, prim2, sc1, sc2, X1, X2, disp, cc1, cc2, src, sur, S, step);
}

}  // namespace titan
