// titan_fit.h -- host side of the fitted surfaces (TS, HP, BS) and their rays.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "../../include/titan_gpu.h"
#include "device_mesh.h"
#include "titan_mesh.h"

namespace titan {

struct FitState;

// Surface lists from the face flags (bit3 TS, bit2 HP, bit4 BS) of a mesh in reference numbering.
int fit_create(FitState **out, int64_t nn, int64_t nf, const int32_t *face_nodes, const int32_t *face_cells,
               const int32_t *cell_faces, const uint8_t *face_flags, const DeviceOrder &o);
void fit_destroy(FitState *F);
int fit_set_rays(FitState *F, const titan_rays *r, const DeviceOrder &o);
int fit_set_rays_ext(FitState *F, const titan_rays_ext *r, const DeviceOrder &o);
// Setka::Smooth_angle_HP from node layer 0 x0 (reference numbering): circle centres and the
// azimuthal neighbours added to each listed HP node's smoothing set.
int fit_smooth_angle(FitState *F, int64_t n, const int32_t *nodes, const int32_t *nb6, const double *x0);
int fit_set_hp_grid(FitState *F, const titan_hp_grid *g, const DeviceOrder &o);
// fitCircleRobust (Setka_move.cpp:2544-2620) of n points [n*2] -> centre_radius [3].
int circle_fit(int64_t n, const double *pts, double *centre_radius);
// kind 1: Setka::Smooth_head_TS3, 2: Smooth_head_HP3, 3: Find_Yzel_Sosed_for_BS, from the node
// coordinates x0 [nn*3] (reference numbering) of node layer 0.
int fit_smooth(FitState *F, int kind, const double *x0);
// Checks the parameters against the set-up (rays, smoothing data), uploads what changed and
// fills the device view.
int fit_prepare(FitState *F, const titan_params &p, hipStream_t s, FitDev *dev);
// Averaged velocity of the surface nodes of the last fitting call -> v [nn*3] (reference numbering).
int fit_node_velocity(FitState *F, hipStream_t s, int64_t nn, double *v);
bool fit_has_rays(const FitState *F);

// Smooth_head_TS3's sphere fit (Setka_move.cpp:2257-2355) of n points [n*3].
int sphere_fit(int64_t n, const double *pts, double centre[3]);

}  // namespace titan
