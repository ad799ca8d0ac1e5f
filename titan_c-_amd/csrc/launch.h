// launch.h -- host-callable launchers of the HIP kernels (defined in the .hip translation units).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "device_mesh.h"
#include "mlp.h"

namespace titan {

void launch_begin_step(hipStream_t s, StepScalars *S, long long n, const DevParams &P, int restart);
void launch_motion(hipStream_t s, int64_t n3, const double *x_old, const double *disp, const StepScalars *S,
                   double *x_new);
void launch_measure(hipStream_t s, const DevMesh &M, const double *X, const DevGeom &G, int gcl);
// every face: flux [Nf*9]; with P.n_scal, scal [Nc*2] (nullable otherwise) gives kflux [Nf*2]
void launch_face_flux(hipStream_t s, int precision, const DevMesh &M, const DevParams &P, const double *prim,
                      const double *scal, const DevGeom &G, const double *X1, const double *X2, StepScalars *S,
                      double *flux, double *kflux, uint8_t *branch, long long step, double dt_override);
void launch_face_flux_f32(hipStream_t s, const DevMesh &M, const DevParams &P, const double *prim,
                          const double *scal, const DevGeom &G, const double *X1, const double *X2, StepScalars *S,
                          double *flux, double *kflux, uint8_t *branch, long long step, double dt_override);

// fused step over tiles (tile_kernel.h); precision = TITAN_PREC_*.  sur (nullable): surrogate
// source fused into the update (then src is ignored and launch_cx_fixup must follow).  With
// P.n_scal the passive scalars go sc1 -> sc2 (nullable otherwise).
void launch_step_tile(hipStream_t s, int precision, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                      const double *prim1, double *prim2, const double *sc1, double *sc2, const double *X1, double *X2,
                      const double *disp, const double *cc1, double *cc2, const double *src, const SurArgs *sur,
                      StepScalars *S, long long step);
void launch_step_tile_sc64(hipStream_t s, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                           const double *prim1, double *prim2, const double *sc1, double *sc2, const double *X1,
                           double *X2, const double *disp, const double *cc1, double *cc2, const double *src,
                           const SurArgs *sur, StepScalars *S, long long step);
void launch_step_tile_sc32(hipStream_t s, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                           const double *prim1, double *prim2, const double *sc1, double *sc2, const double *X1,
                           double *X2, const double *disp, const double *cc1, double *cc2, const double *src,
                           const SurArgs *sur, StepScalars *S, long long step);
void launch_step_tile_f32(hipStream_t s, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                          const double *prim1, double *prim2, const double *X1, double *X2, const double *disp,
                          const double *cc1, double *cc2, const double *src, const SurArgs *sur, StepScalars *S,
                          long long step);
size_t step_tile_lds_bytes(const TileMesh &TM, bool moving, bool scal);

// fitted surfaces (k_fit.hip): surface faces and nodes of layer X1 -> surface nodes of X2
// (time > 0: explicit step length, else S->dt); then every ray's other nodes of X
void launch_fit_surface(hipStream_t s, const DevMesh &M, const DevParams &P, const FitDev &F, const double *prim,
                        const double *X1, double *X2, const double *cell_center, StepScalars *S, double time);
void launch_fit_rays(hipStream_t s, const FitDev &F, double *X);
// k_mk.hip: Monte-Carlo neutral densities
void launch_mk_pack(hipStream_t s, const DevMesh &M, const double *X, MkCell *pack);
void launch_mk_gather(hipStream_t s, int64_t n, const int32_t *faces, const DevGeom &G, double *na);
void launch_mk_particles(hipStream_t s, const DevMesh &M, const MkDev &K, const double *X, const double *prim,
                         const DevGeom &G);
// zone borders: running sums of n distributions at offsets off (into cdf [n*nb]); zeroing n of them
// Cell::MK_normir_Moments + Cell::MK_calc_Sm of n cells' pickup-ion tallies, in place (k_mk.hip);
// tt: nw*nw doubles of scratch for the cell-independent angle integrals
void launch_mk_pui_finish(hipStream_t s, int64_t n, int32_t nw, double wr, double a2, double kn, const double *vol,
                          const double *trig, double *tt, double *sm, double *sp);
void launch_mkz_cdf(hipStream_t s, int64_t n, const int64_t *off, const double *dist, int64_t nb, double *cdf);
void launch_mkz_clear(hipStream_t s, int64_t n, const int64_t *off, double *dist, int64_t nb);

// chlld over n independent states; err (device int) set to 1 on a fatal branch
void launch_chlld_batch(hipStream_t s, int precision, int64_t n, int K, const int8_t *n_state,
                        const int8_t *n_disc, const uint8_t *flags, const double *normal, const double *w,
                        const double *qL, const double *qR, const double *kL, const double *kR, double *q,
                        double *k, double *ds, uint8_t *branch, int *err);
void launch_chlld_batch_f32(hipStream_t s, int64_t n, int K, const int8_t *n_state, const int8_t *n_disc,
                            const uint8_t *flags, const double *normal, const double *w, const double *qL,
                            const double *qR, const double *kL, const double *kR, double *q, double *k,
                            double *ds, uint8_t *branch, int *err);

// exact charge-exchange source: in10 [n*10] (n_p, c_p^2, V_p, n_H, c_H^2, V_H) -> out5 [n*5]
void launch_cx_exact(hipStream_t s, int64_t n, const double *in10, int nodes, double *out5);
// per-cell exact source from the plasma state and the neutral moments [Nc*5]
void launch_cx_cells(hipStream_t s, int64_t nc, const double *prim, const double *neutrals, int nodes,
                     double *src);
// Calc_sourse_MF_Bera plasma source per cell from the state, the passive scalars scal [Nc*2]
// (nullable: Q = rho_He = 0) with the cell codes' zones, and hfluid [Nc * n_h * 5]
void launch_bera_cells(hipStream_t s, int64_t nc, const double *prim, const double *scal, const uint32_t *cell_code,
                       const double *hfluid, int n_h, double a2, double ddp, double *src);
// the same with the hydrogen fluids' sources hout [nc][n_h][5] (n_h <= kBeraMaxH); arise: nullable
void launch_bera_cells_h(hipStream_t s, int64_t nc, const double *prim, const double *scal, const uint32_t *cell_code,
                         const double *hfluid, int n_h, double a2, double ddp, double n_h_lism, const int8_t *arise,
                         double *src, double *hout);
// exact source + update of the cells the surrogate step listed as out of its training box
void launch_cx_fixup(hipStream_t s, const DevMesh &M, const double *prim1, const double *neutrals, int nodes,
                     const int32_t *list, const double *stash, int64_t cap, double *part, StepScalars *S,
                     double *prim2);

}  // namespace titan
