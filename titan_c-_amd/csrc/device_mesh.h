// device_mesh.h -- device-resident mesh, geometry and step scalars shared by host and kernels.
//
// Layout in HBM (all arrays in the device numbering: cells along a Morton curve of their
// initial centres, faces ordered by their lower device cell, nodes by first use):
//   cell_nodes [Nc*8] i32, cell_faces [Nc*6] i32 (a cell's faces in ascending REFERENCE face id,
//   the summation order of Setka_physics.cpp:1702-1733), cell_code [Nc] u32,
//   face_nodes [Nf*4] i32, face_cells [Nf*2] i32, face_code [Nf] u32, face_bidx [Nf] i32,
//   bstate [Nb*8] f64, node xyz [2][Nn*3] f64, prim [2][Nc*8] f64 (cell-major, 64 B per cell),
//   passive scalars (when carried) scal [2][Nc*2] f64 = (Q, rho_He) per cell and bscal [Nb*2],
//   geometry per layer: cell centre [Nc*3], volume [Nc], face centre/normal [Nf*3], area [Nf].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace titan {

// cell_code bits: 0..5 = sign of face j (1: the cell is cells[0] of its j-th face),
// 6 = is_TVD, 7 = in the inner area (Cell::is_inner), 8..11 = zone (Type_cell::Zone_1..4)
__host__ __device__ inline uint32_t cell_code_make(uint32_t fsign, uint32_t tvd, uint32_t zone) {
    return (fsign & 63u) | ((tvd & 1u) << 6) | ((zone & 15u) << 8);
}
// face_code bits: 0..3 = boundary code, 4..7 = Riemann method, 8..15 = chlld flags
__host__ __device__ inline uint32_t face_code_make(uint32_t bc, uint32_t method, uint32_t flags) {
    return (bc & 15u) | ((method & 15u) << 4) | ((flags & 255u) << 8);
}

// The area a step of Setka::Go covers (Setka_physics.cpp:1526-1537): 0 every cell (this build's
// continuing steps), 1 the inner area (is_inner_area = true), 2 the outer area.
__host__ __device__ inline bool cell_in_area(uint32_t code, int area) {
    return area == 0 || (((code >> 7) & 1u) != 0u) == (area == 1);
}
// The area's face list (Setka_physics.cpp:168-194): a face belongs to it when one of its cells
// does; a boundary face goes with its one cell.
__host__ __device__ inline bool face_in_area(uint32_t codeA, uint32_t codeB, bool boundary, int area) {
    if (area == 0) return true;
    return cell_in_area(codeA, area) || (!boundary && cell_in_area(codeB, area));
}

struct DevMesh {
    int64_t n_nodes, n_cells, n_faces;
    const int32_t *cell_nodes;
    const int32_t *cell_faces;
    const uint32_t *cell_code;
    const int32_t *face_nodes;
    const int32_t *face_cells;
    const uint32_t *face_code;
    const int32_t *face_bidx;
    const double *bstate;
    const double *bscal;       // [Nb*2] Q, rho_He of the fixed boundary states (passive scalars)
    const int32_t *face_ref;   // device face -> reference face id (error reports)
    const int32_t *cell_ref;   // device cell -> reference cell id
    const int32_t *face_tvd;   // [Nf*2] second neighbours (AA, BB) of the faces that reconstruct to
                               // second order, (-1, -1) on every other face
};

struct DevGeom {
    double *cell_center;
    double *cell_volume;
    double *face_center;
    double *face_normal;
    double *face_area;
};

struct DevParams {
    double gamma, kfl, v_inf;
    int32_t move_mesh, n_disc, geom_mode, tvd;  // tvd: some face reconstructs to second order
    double motion_omega, motion_phase;
    int32_t area;                    // cell_in_area: the cells this step updates
    int32_t n_scal;                  // passive scalars carried by the step: 0, or 2 (Q, rho_He)
};

// The passive scalars of the plasma (Phys_param.cpp:232-233): Q, the heliopause marker, and the
// helium density rho_He, advected with the flow (the konvect part of Setka::Culc_Gran_Potok,
// Setka_physics.cpp:2349-2361,2485-2493).
constexpr int kScal = 2;

// Fitted surfaces and their rays (Setka::Culc_Velocity_surface, Setka_move.cpp:111-1357, and
// Luch::dvigenie of the A_Luch rays, Luch.cpp:6-152).  Surface kinds: 1 TS, 2 HP, 3 BS; arrays
// indexed by kind have slot 0 unused.
// The B, C, D, E, H, G rays of Luch::dvigenie (Luch.cpp:153-739).  Rays sit in `order` family by
// family (fam_off[f] .. fam_off[f+1] for family f = 1..6), the given order within one.
struct RayExtDev {
    int64_t n;
    const int8_t *type;          // [n] 1 B, 2 C/C2, 3 D, 4 E, 5 H, 6 G
    const int64_t *off;          // [n+1] into nodes
    const int32_t *nodes;        // device nodes, inner to outer
    const int32_t *opor;         // [n*4] device support nodes, -1 unused
    const double *trig;          // [n*4] cos(the), sin(the), cos(phi), sin(phi)
    const int32_t *order;        // rays by family
    int64_t fam_off[8];
    int32_t serial;              // bit f: family f has rays that touch a node another of them writes
    int32_t N4, N5, MF;
    double L6, L7, dd1, dd2, dd3, dd4, dd5i, dd6i, dd7, dd8;  // dd5i, dd6i truncated to integers
    double ba1, ba2, ba3, ba4, ba5, md1, md2, md3, ea1, ea2, ea3;
};

struct FitDev {
    int32_t n_sf, n_sn;
    const int32_t *sf_face;      // [n_sf] device face of each surface face
    const uint8_t *sf_kind;      // [n_sf]
    double *sf_vel;              // [n_sf*3] velocity_K * ds * n of each surface face
    const int32_t *sn_node;      // [n_sn] device node of each surface node
    const uint8_t *sn_kind;      // [n_sn]
    const int32_t *sn_face_off;  // [n_sn+1] its surface faces (indices into sf_*), ascending reference id
    const int32_t *sn_face;
    const int32_t *sn_nb_off;    // [n_sn+1] its smoothing neighbours (device nodes), ascending reference id
    const int32_t *sn_nb;
    const double *sn_centre;     // [n_sn*3] centre of the sphere smoothing (xc, yc, zc)
    double *sn_vel;              // [n_sn*3] averaged node velocity of the last call
    int64_t n_rays;
    int32_t npr;                 // nodes per ray, inner to outer sphere
    const int32_t *ray_nodes;    // [n_rays*npr] device nodes
    const double *ray_trig;      // [n_rays*4] cos(the), sin(the), cos(phi), sin(phi)
    int64_t n_nodes;
    const int2 *node_ray;        // [n_nodes] (ray, index on the ray) of each device node; ray -1: none
    double *ray_R;               // [n_rays*3] TS, HP, BS distances from the origin (func_R)
    double R0, R1, R5, da1, da2, da3, da4, da5;
    int32_t M0, M1, M11, M2, M3, M4;
    int32_t move[4], sglag[4];
    double velocity[4];          // velocity_TS, velocity_HP, velocity_BS
    double k_smooth[4];          // sglag_TS_k_sphere, sglag_HP_k_sphere, sglag_BS_k
    double phi_lim;              // HP sphere smoothing where the polar angle is <= phi_a + pi/180
    RayExtDev ext;               // the other ray families (ext.n = 0: none)
    // HP smoothing on the ray grids (Setka_move.cpp:700-1130), with sglag[2]
    const double *sn_cc;         // [n_sn*2] Smooth_angle_HP circle centre (xcc, ycc); NaN: none; null: no table
    const int32_t *sn_grid;      // [n_sn*8] device nodes AA1, AA11, AA3, AA33, AA2, AA22, AA4, AA44; -1: none
    double k_angle, k_hp, hp_angle, hp_along;  // sglag_HP_k_angle, sglag_HP_k, sglag_HP_angle, sglag_HP_along
    // corrections after the move (Setka_move.cpp:1280-1357)
    int32_t n_tail, tail_len, N4;
    const int32_t *tail;         // [n_tail*tail_len] device HP nodes of each azimuth's D rays
    int32_t n_bs;
    const int32_t *bs_ref, *bs_off, *bs_nodes;  // per azimuth: the reference BS node, its CSR list
    const int32_t *sn_ref;       // [n_sn] reference node id of each surface node (error reports)
};

// Monte-Carlo neutrals (Setka::MK_go, Setka_MK.cpp:1225-2316): particles are numbered face by face
// over the injection faces (ascending reference face id); first[q] is the id of face q's first
// particle, first[n_in] the particle count.
struct alignas(16) MkCell {  // 240 bytes
    double x[24];      // the cell's 8 node positions (cell_nodes order)
    uint32_t loc[4];   // face j (device cell_faces order): its 4 nodes as 4-bit indices into x, in
                       // bits 16 (j & 1) of loc[j >> 1], the face's own node order
    int32_t nbr[6];    // the cell across face j, or -1 - boundary code
    int32_t pad[2];
};

struct MkDev {
    int64_t n_in, n_particles;
    const int32_t *face;        // [n_in] device face of each injection face
    const int64_t *first;       // [n_in+1]
    const double *mu;           // [n_in] weight of each of the face's particles
    uint64_t seed;
    double v_inf, a2, kn;
    int32_t max_steps, pad;
    const MkCell *pack;         // [Nc] flight records of the cells (k_mk_pack)
    double *tally;              // [Nc*5] mu * time of sorts outside 1..4, then of sorts 1..4
    double *out;                // [Nc*5] the tallies per unit volume, reference cell order
    unsigned long long *stats;  // [9] particles, charge exchanges, left, passes through the Sun's
                                // sphere, lost, births dropped, segments, flights with a capped
                                // exchange, flights ended on the zone's border
    // the Sun's sphere (Setka_MK.cpp:2138-2160): its radius R0 (0: none) and the inner boundary
    // faces -- device face ids ascending, and per face the inner face across each of its 4 edges
    // (edge k joins its nodes k and k+1 mod 4)
    double R0;
    int32_t n_inner, pad2;
    const int32_t *inner_face;  // [n_inner]
    const int32_t *inner_nb;    // [n_inner*4]
    // Monte-Carlo zones (Setka::MK_go(zone_MK), Setka_MK.cpp:1225-2316): zone 0 flies over the
    // whole mesh; zone 1..7 flies inside that zone's cells, starts particles on the faces its list
    // holds and ends them there.  An injection with side >= 0 draws its velocity from the face's
    // distribution on that side (in_cdf: its bins' running sums); a flight that ends on an interior
    // listed face deposits its weight into the face's distribution on the side it leaves by.
    int32_t zone, r;            // zone; bins 3r x 6r x 6r per distribution
    const int8_t *cell_mk;      // [Nc] MK zone of each cell
    const uint8_t *face_mask;   // [Nf] bit z-1: on zone z's list
    const int32_t *dslot;       // [Nf] distribution slot of a listed interior face, -1 otherwise
    double *dist;               // [n_slots][4][2][108 r^3] weight per bin
    const int8_t *in_sort;      // [n_in] 1..4
    const int8_t *in_side;      // [n_in] -1: the boundary Maxwellian; 0 / 1: the distribution's side
    const int64_t *in_cdf;      // [n_in] offset of the injection's running sums in cdf, -1: none
    const double *cdf;          // running sums of the incoming distributions' bins
    const double *scal;         // [Nc*kScal] Q, rho_He of the flown layer (null: rho_He = 0), for the
                                // thermal-proton split of Plasma_components_1
    // pickup-ion source tallies of MK_source_S (Cell::MK_Add_pui_source): per segment, bin iw of
    // the relative speed u = |V - V_p| on (0, pui_wr) in pui_nw bins; pui_sm[c*nw + iw] += mu t,
    // pui_sp[(c*nw + iw)*2 + k-1] += mu t nu_ex with k = pui_arise[(zone-1)*4 + sort-1] (0: none)
    int32_t pui_nw;
    int8_t pui_arise[16];
    double pui_wr;
    double *pui_sm, *pui_sp;
};

// Failures of a step, ranked in the order the reference meets them -- the surface fit before the
// face loop (Setka_physics.cpp:1616-1627), the face loop before the cell loop -- so that the
// reported one does not depend on which workgroup ran first; within a rank the smallest reference
// id wins (one 64-bit atomicMin of rank << 48 | id).
enum ErrRank : unsigned {
    kErrFitContact = 0,     // NaN / subnormal heliopause contact speed (TITAN_E_SURFACE)
    kErrFitRiemann = 1,     // chlld's fatal branch in the surface fit (TITAN_E_RIEMANN)
    kErrFitQuad = 2,        // HP grid parabola through points with a repeated abscissa (TITAN_E_SURFACE)
    kErrFaceRiemann = 3,    // chlld's fatal branch or an unknown boundary code in the face loop
    kErrCellNonfinite = 4,  // NaN density after the update (TITAN_E_NONFINITE_STATE)
    kErrCellFallback = 5,   // surrogate fallback list full (TITAN_E_ARG)
};

// Device-side scalars of the step loop (one instance per context).
struct StepScalars {
    double dt;                       // dt of the current step (lagged, Setka_physics.cpp:1606)
    double t;                        // simulated time at the start of the current step
    double motion_fac;               // dt * omega * cos(omega t + phase)
    double pad0;
    unsigned long long dtmin_bits[2];  // running min of the CFL step, slot = step & 1
    unsigned long long n_rho_floor, n_p_floor, n_fallback;
    unsigned long long n_fp64_faces;  // faces the mixed solver handed to the FP64 solver
    unsigned long long err_key;      // first failure: rank << 48 | reference id (ErrRank), ~0 = none
};

}  // namespace titan

namespace titan {

// Tiles of the fused step kernel: TILE consecutive device cells, the nodes they use (local
// indices < 4096), and every face touching them (faces on a tile border appear in both tiles).
// A face record is 16 bytes: x = 4 local node ids (12 bits each) | face code << 48; y = cells[0]
// (device id, low 32 bits) | cells[1] << 32 (device id, or -(boundary-state index + 2)).
constexpr int kTileThreads = 256;  // workgroup size of the fused step kernel
constexpr int kXcds = 8;           // accelerator dies of an MI355X (workgroups dealt round robin)

struct TileMesh {
    int64_t n_tiles;
    int32_t tile_cells, max_nodes, max_faces;
    int32_t vol_cached;         // vol[] holds each cell's volume at the step's old node layer
    const int32_t *cell_off;    // [n_tiles+1] first device cell of each tile (tiles are runs of cells)
    const int32_t *node_off;    // [n_tiles+1] into nodes
    const int32_t *nodes;       // device node ids of each tile, ascending
    const int32_t *own;         // [n_tiles*2] device node range whose X2 this tile writes
    const int32_t *face_off;    // [n_tiles+1] into faces
    const ulonglong2 *faces;    // face records (see above)
    const int32_t *face_ref;    // reference face id of each record (error reports only)
    const uint16_t *cell_lnodes;  // [Nc*8] tile-local node index of each cell node
    const uint16_t *cell_lfaces;  // [Nc*6] tile-local face slot of each cell face (reference order)
    const int2 *face_tvd;       // per record: DevMesh::face_tvd of its face
    double *vol;                // [Nc] cell volumes: read at the old layer (if vol_cached), written at
                                // the new one -- the next step's old layer (same nodes, same formula)
};

__host__ __device__ inline unsigned long long tile_face_pack(const int ln[4], uint32_t code) {
    return (unsigned long long)(ln[0] & 0xfff) | ((unsigned long long)(ln[1] & 0xfff) << 12) |
           ((unsigned long long)(ln[2] & 0xfff) << 24) | ((unsigned long long)(ln[3] & 0xfff) << 36) |
           ((unsigned long long)(code & 0xffff) << 48);
}

}  // namespace titan
