/*
 * titan_oracle.h -- CPU restatement of the TITAN ideal-MHD step (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the parity checker for the HIP implementation.  Only the test suite,
 * the smoke entry and the CPU-baseline leg of the benchmark may load it.  The product
 * (libtitan_gpu.so) never links or calls it.
 *
 * Every function restates the arithmetic of the reference (SergeyMSU/TITAN_C-) in the
 * same operation order, so that FP64 results can be compared bit-for-bit:
 *   or_chlld            Phys_param::chlld            TITAN/source/Phys_param.cpp:882-1878
 *   or_get_bazis        get_bazis                    TITAN/source/Help_funk.cpp:219-232
 *   or_measure          Setka::Calculating_measure   TITAN/source/Setka.cpp:2465-2502
 *                         Cell::Culc_center          TITAN/source/Cell.cpp:909-930
 *                         Gran::Culc_measure         TITAN/source/Gran.cpp:112-218
 *                         Cell::Culc_volume (m=1)    TITAN/source/Cell.cpp:958-1000
 *   or_face_velocity    Gran::culc_velosity          TITAN/source/Gran.cpp:571-589
 *   or_face_states      Setka::Snos_on_Gran (first order, TVD, boundary)
 *                                                    TITAN/source/Setka_gran_snos.cpp:3-1071
 *   or_init_tvd         Setka::Init_TVD              TITAN/source/Setka_physics.cpp:648-744
 *   or_minmod, or_linear  minmod / linear           TITAN/source/Help_funk.cpp:26-70
 *   or_face_flux        Setka::Culc_Gran_Potok       TITAN/source/Setka_physics.cpp:2280-2571
 *   or_update           cell loop of Setka::Go       TITAN/source/Setka_physics.cpp:1671-2027
 *   or_go               Setka::Go                    TITAN/source/Setka_physics.cpp:1484-2278
 *   or_new_connect      Setka::New_connect           TITAN/source/Setka.cpp:2066-2249
 *   or_cx_rate_ref      Culc_nu_maxwel               MK_source/MK_source/MK_source/MK_source.cpp:38-103
 *   or_bera_source      Setka::Calc_sourse_MF_Bera   TITAN/source/Setka_physics.cpp:746-1258
 *                       (plasma part; Plasma_components_1, Phys_param.cpp:628-665)
 *   or_cx_source, or_sigma_ls, or_mlp_*  the learned-source path (no reference counterpart:
 *                       Maxwellian CX moments with the Lindsay-Stebbings cross-section; MLP)
 *   or_culc_velocity_surface  Setka::Culc_Velocity_surface  TITAN/source/Setka_move.cpp:111-1357
 *   or_dvigenie         Luch::dvigenie (A_Luch)      TITAN/source/Luch.cpp:6-152
 *   or_mk_setup, or_mk_go  Setka::MK_go / MK_fly_immit  TITAN/source/Setka_MK.cpp:1225-2850
 *                       (Gran::Get_Random_pozition, Luch_crossing, rayTriangleIntersect,
 *                       Gran.cpp:220-285,451-561; Cell::MK_Add_particle, Cell.cpp:1042-1083)
 *   or_set_mk_zone      Setka::Set_MK_Zone           TITAN/source/Setka_MK.cpp:142-683
 *   or_mk_setup_zone, or_mk_go (zone > 0)  MK_prepare + MK_go(zone_MK)  Setka_MK.cpp:686-1752
 *   or_mk_deposit, or_mk_sample_batch  AMR_f::Add_particle, Get_random_velosity  AMR_f.cpp:4-316,
 *                       AMR_cell.cpp:273-327, Gran::Read_AMR (Gran.cpp:16-99)
 *
 * What pins each function (tests/test_oracle_*.py), so that a dropped term, a wrong sign or
 * index or a transposed operand fails a check that is not a retyping of the formula:
 *   or_chlld            consistency with the reference's physical flux (Get_Potok,
 *                       Phys_param.cpp:826-874) for equal states, rotational invariance, L/R
 *                       mirror symmetry, HLLD -> HLLC at Bn = 0, monotone regions and flux
 *                       continuity across the fan, the fatal branch, raspad_testing inputs.
 *                       Last-bit parity with the compiled reference is UNPINNED: TITAN/source
 *                       needs Eigen3/Boost/CGAL (absent), so Eigen's summation order inside
 *                       get_bazis/normalized() is assumed to be left-to-right.
 *   or_get_bazis        orthonormality and handedness.
 *   or_measure(_gcl)    exact values on axis-aligned boxes, Test_geometr invariants, translation
 *                       invariance, GCL closure and swept-volume identity.
 *   or_face_velocities  rigid translation gives w = d.n/dt.
 *   or_motion, or_det_cos  rigid translation; det_cos against libm cos.
 *   or_face_states      a spherically symmetric wind is carried onto every face exactly (both
 *                       spherical frames), boundary rules on known answers.  TVD: a linear field
 *                       is reconstructed exactly (plain case and the one-sided lines at Zone_2/3
 *                       and Zone_3/4 transitions), an extremum keeps the cell value, the
 *                       positivity fallback, and the wind again through the all-Zone_1 case.
 *                       Zone_1/Zone_2 transitions (Setka_gran_snos.cpp:401-660): along radial
 *                       lines of a zone-banded shell the wind's V and r^(2 gamma) p come through
 *                       exactly where the case interpolates them, and zig-zag densities against
 *                       closed-form interpolation / extrapolation / extremum values fix which
 *                       line each side uses and that density is interpolated unscaled.
 *   or_init_tvd         straight-line neighbours on a regular cube; the side neighbour (cosine
 *                       -1/sqrt 5) next to a boundary.
 *   or_minmod, or_linear  known answers, including the 1e-8 sign threshold.
 *   or_update, or_go    mass conservation, quiescent state, floors at their thresholds, frozen B,
 *                       volume ratio and sources, lagged dt schedule, Brio-Wu profile.
 *   the _k forms        (passive scalars Q, rho_He) a scalar equal to 2 rho (or rho/4) -- a power of
 *                       two, so every operation scales exactly -- keeps that ratio bit for bit through
 *                       reconstruction (all zone cases, Zone_1 and TVD), every chlld region and the
 *                       update, which fixes each scalar formula to be rho's; total Q conserved on a
 *                       closed box; the rho_He clips at their thresholds; determ_zone's four outcomes
 *                       and the al switch of the closed-form source against its al = 1 / al = 2 forms.
 *   or_new_connect      hand-worked two-cell mesh and face counts.
 *   or_cx_rate_ref      the reference's own MK_source program (compiled from its source into
 *                       oracle/_ref) at sampled table points: tests/golden/mk_source_int1.json.
 *   or_bera_source      exact zero without atoms and at thermal + mechanical equilibrium (which
 *                       also fixes the proton temperature of Plasma_components_1), exact
 *                       additivity over fluids, the cold-beam transfer rates, rotation and
 *                       Galilean covariance, and -- with sigma = 1 -- 6-D Monte Carlo Maxwellian
 *                       averages of the momentum and energy exchange within 1.5-3 % (a sign flip
 *                       of the thermal drag or swapped temperatures in U_m / U_E leave that band).
 *   or_cx_source        sigma = const closed form, Cartesian quadrature, 6-D Monte Carlo,
 *                       bilinearity and Galilean invariance; or_sigma_ls: published values.
 *   or_mlp_*            closed forms and finite differences; or_mlp_forward_bf16 against
 *                       or_mlp_forward on BF16-exact weights and inputs with the activations'
 *                       rounding bounded by 2^-9 relative.
 *   or_culc_velocity_surface, or_dvigenie  closed-form wave speeds of uniform states on the
 *                       TS/HP/BS faces, node means, the sphere and neighbour smoothing terms,
 *                       straight rays and the end slopes of the ray polynomials, the surface
 *                       neighbour counts of a partial surface, a fitted free stream kept exactly.
 *   or_mk_setup, or_mk_go  free flight holds the LISM density in every cell of a convex domain
 *                       fed from all faces; detailed balance with protons in the same Maxwellian
 *                       (Kn 1 and 0.3); exchange counts monotone in Kn; the particle-count rule;
 *                       inflow against outflow through a closed surface; the exchange sampler's
 *                       mean and mean square against quadrature of its weighted Maxwellian;
 *                       planted mistakes (or_mk.mistake) fail these checks.  Per particle the flight
 *                       is UNPINNED against reference output: its random numbers come from files
 *                       it does not ship.
 *   or_set_mk_zone      every interior face between two MK zones is on both zones' lists and
 *                       every listed face touches its zone (a dropped rule, the mistake bit, fails).
 *   or_mk_deposit, or_mk_sample_batch  draws land in each bin in proportion to its deposited
 *                       weight (chi-square); within a bin Vn has density ~ Vn (closed-form mean)
 *                       and Vt is uniform.
 *   or_mk_setup_zone, or_mk_go (zone)  a zone without outer boundary deposits exactly the weight
 *                       it injects; the 19-pass zone sequence of Setka.cpp:200-226 reproduces the
 *                       one-zone flight's mass per zone (12% free flight, 25% with exchange, r = 4).
 */
#ifndef TITAN_ORACLE_H
#define TITAN_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Planted mistakes (test-only switch, or_set_mistake): each bit makes one plausible restatement
 * error, so a test can show that its pin catches it.  0 restores the restatement. */
enum {
    OR_MISTAKE_POWELL_MOM_8PI = 1 << 0,    /* Powell momentum term B/(8 pi) instead of B/(4 pi)   */
    OR_MISTAKE_POWELL_IND_SIGN = 1 << 1,   /* Powell induction term +v div B                      */
    OR_MISTAKE_POWELL_ENERGY = 1 << 2,     /* Powell energy term dropped                          */
    OR_MISTAKE_PDIVB_HALF = 1 << 3,        /* PdivB without the factor 1/2 of the face mean       */
    OR_MISTAKE_ZONE2_8PI = 1 << 4,         /* Zone_2 decay: pressure compensation over 8 pi       */
    OR_MISTAKE_HLLD_VZZ_SIGN = 1 << 5,     /* HLLD double-star v_t: sign(Bn) term negated          */
    OR_MISTAKE_HLLD_BZZ_WEIGHTS = 1 << 6,  /* HLLD double-star b_t: sqrt(rho*) weights on the wrong side */
    OR_MISTAKE_HLLD_TB_BN2 = 1 << 7,       /* HLLD star b_t: Bn^2 dropped from the left numerator  */
    OR_MISTAKE_SCAL_ZONE1_PLAIN = 1 << 8,  /* passive scalars from a Zone_1 cell not rescaled by r^2   */
    OR_MISTAKE_SCAL_TVD_R2_KEPT = 1 << 9,  /* TVD 1112 / 1122: the scalars' r^2 lines kept (not overwritten) */
    OR_MISTAKE_HE_CLIP_RHO = 1 << 10       /* rho_He above the new density clipped to it, not to a tenth */
};
void or_set_mistake(int bits);

/* boundary codes of a face (reference Type_Gran, Setka_physics.cpp:212-236) */
enum { OR_BC_INTERIOR = 0, OR_BC_INNER_HARD = 1, OR_BC_OUTER_HARD = 2, OR_BC_OUTER_SOFT = 3,
       OR_BC_ZERO_GRADIENT = 4 };

/* Mesh in reference numbering (faces in New_connect creation order). */
typedef struct {
    int64_t n_nodes, n_cells, n_faces;
    const int32_t *cell_nodes;   /* [n_cells*8] hexahedron nodes                      */
    const int32_t *cell_faces;   /* [n_cells*6] faces of a cell, ascending face id    */
    const int32_t *face_nodes;   /* [n_faces*4]                                       */
    const int32_t *face_cells;   /* [n_faces*2] cells[0], cells[1] (-1 on boundary)   */
    const int8_t  *face_bc;      /* [n_faces] OR_BC_*                                 */
    const int8_t  *face_method;  /* [n_faces] 0 Lax, 1 HLL, 2 HLLC, 3 HLLD            */
    const uint8_t *face_flags;   /* [n_faces] bit0 left_ydar (TS_hard), bit1 contact,
                                    bit2 heliopause face (Gran::type2 == HP)            */
    const double  *face_bstate;  /* [n_faces*8] fixed boundary state (rho,V,p,B)      */
    const int8_t  *cell_zone;    /* [n_cells] 1..4 (Type_cell::Zone_1..Zone_4)        */
    const uint8_t *cell_is_tvd;  /* [n_cells] 0 => B frozen and no TVD reconstruction  */
    const int32_t *face_tvd;     /* [n_faces*2] TVD neighbours AA (behind cells[0]) and BB
                                    (behind cells[1]), -1 = none (or_init_tvd); nullable */
    const uint8_t *cell_inner;   /* [n_cells] Cell::is_inner (Setka::Init_boundary_grans); nullable */
    const double  *face_bscal;   /* [n_faces*2] Q, rho_He of the fixed boundary state; needed by the
                                    *_k functions on hard faces                                  */
} or_mesh;

typedef struct {
    double gamma;        /* 5/3                                   */
    double kfl;          /* CFL number KFL                        */
    double v_inf;        /* Velosity_inf (outer-soft suction fix) */
    int32_t move_mesh;   /* move_setka                            */
    int32_t n_disc;      /* wave-speed estimate passed to chlld   */
    /* prescribed node motion: x += dt * (omega*cos(omega*t+phase)) * D_i */
    double motion_omega, motion_phase;
    int32_t tvd;         /* TVD: second-order limited face states (Setka_gran_snos.cpp:186-736) */
    int32_t snos_on_hp;  /* Snos_on_HP: TVD also on heliopause faces (Phys_param.h:130)        */
    double par_a_2;      /* charge-exchange cross-section slope a_2 (Phys_param.h:99)           */
    double par_n_H_LISM; /* n_H of the LISM (Phys_param.h:100)                                  */
    double par_Kn;       /* Knudsen number (Phys_param.h:101)                                   */
    int32_t area;        /* 0 every cell; 1 / 2 the inner / outer area of Setka::Go
                            (Setka_physics.cpp:1526-1537): only the area's faces are evaluated
                            (and give dt), only its cells updated, the others keep their state */
} or_params;

typedef struct {
    double *cell_center;  /* [n_cells*3] */
    double *cell_volume;  /* [n_cells]   */
    double *face_center;  /* [n_faces*3] */
    double *face_normal;  /* [n_faces*3] */
    double *face_area;    /* [n_faces]   */
} or_geom;

/* ---- Riemann solver ---- */
/* Returns 0, or -1 on the reference's fatal "wrong choice" branches. branch = n_state*16 + region
 * (+8 when HLLD's |Bn|>epsb, i.e. ibn=1); dsl/dsc/dsr as in the reference. */
int or_chlld(int n_state, int n_disc, int flags, double al, double be, double ge, double w,
             const double *qL, const double *qR, int K, const double *kL, const double *kR,
             double *q, double *k, double *dsl, double *dsc, double *dsr, int *branch);
int or_chlld_batch(int64_t n, const int8_t *n_state, const int8_t *n_disc, const uint8_t *flags,
                   const double *normal, const double *w, const double *qL, const double *qR,
                   int K, const double *kL, const double *kR, double *q, double *k,
                   double *ds, uint8_t *branch);
void or_get_bazis(const double *n, double *t, double *m);
/* physical flux F(U).n in the lab frame (Phys_param::Get_Potok, Phys_param.cpp:826-874);
 * output order of the reference: [mass, mom x3, B x3, energy] */
void or_get_potok(double gamma, double rho, double p, double u, double v, double w, double bx,
                  double by, double bz, double n1, double n2, double n3, double *potok);

/* ---- geometry ---- */
void or_measure(const or_mesh *m, const double *coords, or_geom *g);
void or_measure_gcl(const or_mesh *m, const double *coords, or_geom *g);
double or_face_velocity(const or_mesh *m, int64_t f, const double *x_now1, const double *x_now2,
                        const double *normal_now1, double dt);
void or_face_velocities(const or_mesh *m, const double *x1, const double *x2, const double *normal1,
                        const double *area1, double dt, int gcl, double *w);
double or_det_cos(double x);
void or_motion(int64_t n_nodes, const double *x_old, const double *disp, double dt, double t,
               double omega, double phase, double *x_new);

/* ---- TVD neighbours (Setka::Init_TVD, Setka_physics.cpp:648-744) from the layer-0 geometry:
 * tvd[2f] = AA, tvd[2f+1] = BB (-1 = none).  Help_funk.cpp:26-70 limiter pieces exported for tests. */
void or_init_tvd(const or_mesh *m, const or_geom *g0, int32_t *tvd);
double or_minmod(double x, double y);
double or_linear(double x1, double t1, double x2, double t2, double x3, double t3, double y);

/* ---- face / cell passes ---- */
int or_face_states(const or_mesh *m, const or_params *p, int64_t f, const double *prim,
                   const or_geom *g, double *qL, double *qR);
/* The _k forms carry the passive scalars Q, rho_He (scal [n_cells*2], kflux [n_faces*2], scal_new). */
int or_face_states_k(const or_mesh *m, const or_params *p, int64_t f, const double *prim, const double *scal,
                     const or_geom *g, double *qL, double *qR, double *kL, double *kR);
int or_face_flux_k(const or_mesh *m, const or_params *p, const double *prim, const double *scal, const or_geom *g,
                   const double *x_now1, const double *x_now2, double dt, double *flux, double *kflux,
                   double *dt_face, uint8_t *branch);
int or_update_k(const or_mesh *m, const or_params *p, const double *prim_old, const double *scal_old,
                const double *vol_old, const double *vol_new, const double *flux, const double *kflux,
                const double *src, double dt, double *prim_new, double *scal_new, int64_t *n_rho_floor,
                int64_t *n_p_floor);
int or_determ_zone(double Q, double rho, int cell_type);
void or_bera_source_k(const double *s, double Q, double rho_He, int cell_type, int n_h, const double *h, double a2,
                      double n_h_lism, double kn, double *out);
/* or_bera_source_k with the hydrogen fluids' sources too: hout [n_h*5] (rho, m_x, m_y, m_z, E) per
 * fluid, n_h <= 16 (Setka_physics.cpp:1005-1094, 1241-1250); arise [4*n_h] the fluid (1-based) the
 * atoms made from fluid k join in zone z at [(z-1)*n_h + k], NULL for fluid z (no pickup ions). */
void or_bera_source_h(const double *s, double Q, double rho_He, int cell_type, int n_h, const double *h, double a2,
                      double n_h_lism, double kn, const int8_t *arise, double *out, double *hout);
/* flux[f*9 + 0..7] = chlld flux * area, flux[f*9+8] = PdivB; dt_face[f] (1e7 on non-interior) */
int or_face_flux(const or_mesh *m, const or_params *p, const double *prim, const or_geom *g,
                 const double *x_now1, const double *x_now2, double dt, double *flux,
                 double *dt_face, uint8_t *branch);
/* src: optional [n_cells*5] sources (mass-rate, m_x, m_y, m_z, E) per unit volume and time */
int or_update(const or_mesh *m, const or_params *p, const double *prim_old, const double *vol_old,
              const double *vol_new, const double *flux, const double *src, double dt,
              double *prim_new, int64_t *n_rho_floor, int64_t *n_p_floor);

/* ---- fitted surfaces: Setka::Culc_Velocity_surface (Setka_move.cpp:111-1357) and Luch::dvigenie
 * of the A_Luch / A2_Luch rays (Luch.cpp:6-152).  Surface faces: face_flags bit3 TS, bit2 HP,
 * bit4 BS; kinds 1 TS, 2 HP, 3 BS (index 0 unused).  The per-node lists are set-up data (the
 * surface neighbours and sphere centres of Smooth_head_TS3 / Smooth_head_HP3 /
 * Find_Yzel_Sosed_for_BS), in reference ids. */
typedef struct {
    int32_t move[4], sglag[4];
    double velocity[4], k_smooth[4];
    double phi_lim;              /* phi_a + pi/180 */
    int64_t n_sn;                /* surface nodes, ascending id */
    const int32_t *sn_node;
    const uint8_t *sn_kind;
    const int32_t *sn_face_off;  /* [n_sn+1] faces that move the node, ascending face id */
    const int32_t *sn_face;
    const int32_t *sn_nb_off;    /* [n_sn+1] smoothing neighbours, ascending node id */
    const int32_t *sn_nb;
    const double *sn_centre;     /* [n_sn*3] */
    int64_t n_rays;
    int32_t npr;
    const int32_t *ray_nodes;    /* [n_rays*npr] */
    const double *ray_angles;    /* [n_rays*2] the, phi */
    double R0, R1, R5, da1, da2, da3, da4, da5;
    int32_t M0, M1, M11, M2, M3, M4;
    const struct or_rays_ext_s *ext; /* the other ray families (nullable), moved after the A rays */
    /* HP smoothing on the ray grids, with sglag[2] (Setka_move.cpp:700-1130): */
    const double *sn_cc;         /* [n_sn*2] circle centre (xcc, ycc) of Smooth_angle_HP, NaN: none (nullable) */
    const int32_t *sn_grid;      /* [n_sn*8] AA1, AA11, AA3, AA33, AA2, AA22, AA4, AA44, -1 row: none (nullable) */
    double k_angle, k_hp, hp_angle, hp_along; /* sglag_HP_k_angle, sglag_HP_k, sglag_HP_angle, sglag_HP_along */
    /* corrections after the move (Setka_move.cpp:1280-1357) */
    int32_t n_tail, tail_len, N4;
    const int32_t *tail;         /* [n_tail*tail_len] the HP nodes (Yzels_opor[1]) of each azimuth's D rays */
    int32_t n_bs;
    const int32_t *bs_ref;       /* [n_bs] the last A ray's BS node of each azimuth */
    const int32_t *bs_off;       /* [n_bs+1] */
    const int32_t *bs_nodes;     /* the BS nodes of its B, E and D rays */
} or_fit;

/* The other ray families of Luch::dvigenie (Luch.cpp:153-739): B_Luch (around the flank), C_Luch /
 * C2_Luch (tailward, bending to the x axis), D_Luch, E_Luch (in planes x = const of the tail),
 * H_Luch, G_Luch (cubic bridges between support nodes).  Each ray: its type, nodes (Luch::Yzels,
 * inner to outer), support nodes (Luch::Yzels_opor[0..3], -1 where the type has fewer) and angles.
 * or_dvigenie_ext moves the families in the reference's order (Setka.cpp:1850-1880: B, C, D, E,
 * then H, G), one ray after another; B and D also set the x of support nodes of theirs.  Geo_param
 * values as named; per-ray overrides of them are not taken. */
enum { OR_RAY_B = 1, OR_RAY_C = 2, OR_RAY_D = 3, OR_RAY_E = 4, OR_RAY_H = 5, OR_RAY_G = 6 };
typedef struct or_rays_ext_s {
    int64_t n;
    const int8_t *type;
    const int64_t *off;        /* [n+1] into nodes */
    const int32_t *nodes;
    const int32_t *opor;       /* [n*4] */
    const double *angles;      /* [n*2] the, phi */
    double R0, R1, R5, L6, L7;
    int32_t M0, M1, M11, M2, M3, M4, MF, N4, N5;
    double dd1, dd2, dd3, dd4, dd5, dd6, dd7, dd8;
    double ba1, ba2, ba3, ba4, ba5, md1, md2, md3, ea1, ea2, ea3;
} or_rays_ext;
void or_dvigenie_ext(const or_rays_ext *r, double *x);
/* Surface velocities from layer `now` (x, prim, geometry g of that layer) moved over `time` into
 * x_next (surface nodes only); node_velocity [n_sn*3] (nullable). Returns -1 on a chlld failure,
 * -8 on a NaN/subnormal HP contact speed, -9 on three grid points with a repeated abscissa
 * (solveQuadraticEquation, Help_funk.cpp:290-305, stops the reference there). */
int or_culc_velocity_surface(const or_mesh *m, const or_params *p, const or_fit *fit, const double *x,
                             const double *prim, const or_geom *g, double time, double *x_next,
                             double *node_velocity);
/* The HP grid terms of one node A [3] (test access to the step's own routines): V5 [5*3] = the angle
 * term (circle centre (cy, cz), neighbours ring [n_ring*3], n_ring <= 64, weight c_angle_term), then
 * the four macros1 terms of its neighbours P8 [8*3] (AA1, AA11, AA3, AA33, AA2, AA22, AA4, AA44).
 * Returns -1 on a repeated abscissa. */
int or_hp_terms(const double *A, const double *P8, double cy, double cz, const double *ring, int32_t n_ring,
                double c_angle_term, double c_along, double c_angle, double time, double *V5);
/* fitCircleRobust (Setka_move.cpp:2544-2620) of n points [n*2]: centre [2], returns the radius */
double or_circle_fit(int64_t n, const double *pts, double *centre);
/* Every ray's nodes other than index M0 and its TS/HP/BS nodes, in x. */
void or_dvigenie(const or_fit *fit, double *x);

/* ---- Monte-Carlo neutral densities: the flight loop of Setka::MK_go / MK_fly_immit
 * (Setka_MK.cpp:1225-2316) for the LISM hydrogen that enters through the outer boundary (sort 4).
 * Injection faces: face ids, particles per face and weight per particle (or_mk_setup restates
 * the counts of Setka_MK.cpp:1233,1380-1390).  tally [n_cells*5] += mu * time per cell: the sum
 * over sorts, then sorts 1..4 (Cell::MK_Add_particle, Cell.cpp:1042-1083).  A particle that comes
 * within 1.01 R0 of the centre after crossing a face is carried along its velocity to where it
 * leaves that sphere and goes on in the cell the ray from the centre through that point enters
 * (Setka_MK.cpp:2138-2160).  stats [8]: particles, charge exchanges, left through the outer
 * boundary, passes through the Sun's sphere, lost (no exit found, no cell after a pass, or the
 * step cap), births dropped (the birth sampler's cap), flight segments (one per tally: a cell
 * crossed or a flight to an exchange), flights in which an exchange sampler hit its cap.  The
 * random numbers come from a counter-based
 * SplitMix64 stream per particle (the reference's Sensor objects read rnd_my.txt, which it does
 * not ship). */
typedef struct {
    int64_t n_in;
    const int32_t *face;
    const int32_t *count;
    const double *mu;
    uint64_t seed;
    double v_inf, a2, kn;
    int32_t max_steps;
    int32_t mistake;  /* 0; for tests that show the pins catch a plausible error: bit 0 drops the
                         |V.n| weight of the birth velocity, bit 1 the g sigma(g) weight of the
                         exchanged velocity, bit 2 absorbs the particles at the inner sphere,
                         bit 3 bins the pickup-ion tallies by |V| instead of |V - V_p| */
    int32_t nthreads; /* OpenMP threads (<= 1: one) */
    double R0;        /* radius of the inner (Sun's) sphere the particles pass through (Geo_param::R0;
                         0: no sphere) */
    /* Monte-Carlo zones (Setka::MK_go(zone_MK)): zone 0 is one zone holding the whole mesh; zone
     * 1..7 flies inside the cells of that MK zone, injects at the faces its list holds and ends a
     * flight on them, depositing into the face's outgoing distribution (interior faces) */
    int32_t zone, pad_;
    const int8_t *cell_mk;     /* [n_cells] MK zone of each cell (or_set_mk_zone); zone > 0 */
    const uint8_t *face_mask;  /* [n_faces] bit z-1: the face is on zone z's list; zone > 0 */
    const int8_t *in_sort;     /* [n_in] sort of each injection (1..4); NULL: 4 */
    const int8_t *in_side;     /* [n_in] -1: the boundary Maxwellian (Velosity_initial); 0 / 1: the
                                  face's distribution on that side; NULL: -1 */
    struct or_mk_dist *dist;   /* zone > 0: the distributions read and deposited into */
    const double *scal;        /* [n_cells*2] Q, rho_He per cell, for the thermal-proton split of
                                  Plasma_components_1 at the cell's type (Setka_MK.cpp:1944-1966);
                                  NULL: rho_He = 0 */
    /* the pickup-ion source tallies of MK_source_S (Cell::MK_Add_pui_source, Cell.cpp:1087-1133),
     * one per flight segment at the relative speed u = |V - V_p|, bin iw = int(u / pui_wr * pui_nw)
     * clamped to the grid: pui_sm [n_cells][nw] += mu time; pui_sp [n_cells][nw][2] (row k-1 of the
     * column-major (2, nw) matrix) += mu time nu_ex, k = pui_arise[(zone-1)*4 + sort-1] (0: none),
     * zone = determ_zone of the cell (Q routes it when carried, else the cell's type).  pui_nw = 0:
     * no tallies. */
    int32_t pui_nw, pad3_;
    double pui_wr;
    const int8_t *pui_arise;   /* [4*4] */
    double *pui_sm, *pui_sp;
} or_mk;
/* Velocity distributions on the zone borders (AMR_f at its unrefined grid of Gran::Read_AMR,
 * Gran.cpp:16-99): (0, 20) x (-20, 20)^2 in the face's frame (Vn, get_bazis(Vn)), 3r x 6r x 6r
 * bins, per listed interior face (slot), sort 1..4 and side (0: Vn = the face normal, what leaves
 * cells[0]; 1: Vn = -normal).  w holds the weight deposited per bin (AMR_f::Add_particle,
 * AMR_f.cpp:240-316); the reference's normalised f (Normir_velocity_volume) times the bin's volume,
 * its centre speed and the face area -- what Culk_SpotokV and Get_random_velosity read -- is the
 * same number. */
typedef struct or_mk_dist {
    int32_t n_slots, r;
    const int32_t *slot;  /* [n_faces] slot of each listed interior face, -1 otherwise */
    double *w;            /* [n_slots][4][2][108 r^3]: bins (i, j, k) of the 3r x 6r x 6r grid, k fastest */
} or_mk_dist;
/* Setka::Set_MK_Zone (Setka_MK.cpp:142-683) on the centres of g: cell_mk [n_cells] = 1 for Zone_1
 * cells, 2 / 3, 4 / 5, 6 / 7 for Zone_2, Zone_3, Zone_4 cells with centre x > 0 / x <= 0; face_mask
 * [n_faces] bit z-1 when the rules of zone z (Setka_MK.cpp:356-641) list the face.  mistake bit 0
 * (tests only): zone 2 drops its rule for faces between Zone_2 cells across x = 0.  Returns 0; -1 a
 * cell zone outside 1..4 (Error 86743207564); -2 a zone whose rules list a face twice
 * (ERROR 6435856408). */
int or_set_mk_zone(const or_mesh *m, const or_geom *g, int mistake, int8_t *cell_mk, uint8_t *face_mask);
/* MK_prepare + the counts of MK_go(zone) (Setka_MK.cpp:686-1100,1225-1400): the flux into the zone
 * through every listed face (interior faces: the weight of each sort's distribution on the side
 * entering the zone; boundary faces: Get_Spotok_inf for sort 4), mu_expect = total / (n_per_face x
 * listed faces), then per sort 1..4 and listed face (ascending id) the injections: face, sort,
 * side, count and weight, sized 4 n_faces by the caller.  Returns their number, or -3 when a listed
 * face's entering cell is not in the zone (Error 9767653421). */
int64_t or_mk_setup_zone(const or_mesh *m, const or_geom *g, double v_inf, int32_t n_per_face, int32_t zone,
                         const int8_t *cell_mk, const uint8_t *face_mask, const or_mk_dist *dist, int32_t *face,
                         int8_t *sort, int8_t *side, int32_t *count, double *mu);
/* AMR_f::Add_particle into distribution (slot, sort, side) for a face of normal n (test access):
 * the bin of V in the frame (Vn, t, m), coordinates clipped 1e-7 inside the box. */
void or_mk_deposit(or_mk_dist *d, int32_t slot, int sort, int side, const double *n, const double *V, double mu);
/* n draws of AMR_f::Get_random_velosity from distribution (slot, sort, side) of a face of normal n
 * (test access), one SplitMix64 stream per draw from seed: V [n*3]. */
void or_mk_sample_batch(const or_mk_dist *d, int32_t slot, int sort, int side, const double *n, int64_t n_draws,
                        uint64_t seed, double *V);
/* Fluxes of the LISM Maxwellian into the domain through the outer boundary faces (bc 2, 3):
 * flux[f] = Get_Spotok_inf(n) * area (Setka_MK.cpp:2316-2324); count and mu by the rule of
 * Setka_MK.cpp:1233,1380-1390 with N_per_gran = n_per_face.  Returns the number of faces used;
 * face/count/mu sized n_faces by the caller. */
int64_t or_mk_setup(const or_mesh *m, const or_geom *g, double v_inf, int32_t n_per_face, int32_t *face,
                    int32_t *count, double *mu);
/* n draws of the exchange sampler alone (M_K_Change_Velosity, units of cp) for fixed U, V, one
 * SplitMix64 stream per draw from seed; capped = draws that hit the 10000-try cap. */
void or_mk_change_velosity_batch(int64_t n, uint64_t seed, double a2, const double *U, const double *V, double cp,
                                 int mistake, double *W, int64_t *capped);
/* With mk->zone > 0, stats has a ninth entry: flights that ended on the zone's border. */
/* Cell::MK_normir_Moments (Cell.cpp:1225-1241) and Cell::MK_calc_Sm (Cell.cpp:1135-1169) of the
 * pickup-ion tallies of n cells: sm_out[c][ij] = the loss rate at w_ij = (ij + 0.5) wr / nw from the
 * volume-normalised sm tally, sum_j sum_{k=1..40} ff_j d sigma(d) sin(th_k) dth 2 pi / (4 pi) / kn
 * with d = |w - Vh_j| at angle th_k = k pi / 40; sp_out = sp / (V 4 pi / 3 (w2^3 - w1^3)). */
void or_mk_pui_finish(int64_t n, int32_t nw, double wr, double a2, double kn, const double *vol, const double *sm,
                      const double *sp, double *sm_out, double *sp_out);
/* Phys_param::Plasma_components_1 (Phys_param.cpp:628-665): th = (rho_Th, p_Th, T_Th) of a plasma
 * cell (rho, p, rho_He) of type zone 1..4. */
void or_plasma_components_1(double rho, double p, double rho_He, int zone, double th[3]);
int or_mk_go(const or_mesh *m, const or_mk *mk, const double *coords, const double *prim, const or_geom *g,
             double *tally, int64_t *stats);

/* Full explicit loop of Setka::Go.  coords[2][n_nodes*3], prim[2][n_cells*8] hold both time
 * layers and are updated in place; disp [n_nodes*3] is the motion field (may be NULL when not
 * moving); with fit (nullable) and move_mesh 2 the nodes move by or_culc_velocity_surface and
 * or_dvigenie instead.  gcl selects the exactly-conservative geometry.  On return *final_layer says which
 * layer holds the newest state; dt_hist (optional, [steps]) receives the dt used at each step. */
int or_go(const or_mesh *m, const or_params *p, double *coords, double *prim, const double *disp,
          int64_t steps, int gcl, int nthreads, int *final_layer, double *dt_hist,
          double *t_sim, int n_h, const double *h_fluids, const or_fit *fit);
/* or_go with the passive scalars: scal [2][n_cells*2] (layer 0 in), the newest in layer *final_layer */
int or_go_k(const or_mesh *m, const or_params *p, double *coords, double *prim, double *scal, const double *disp,
            int64_t steps, int gcl, int nthreads, int *final_layer, double *dt_hist, double *t_sim, int n_h,
            const double *h_fluids, const or_fit *fit);
/* ---- Setka::Calc_sourse_MF_Bera (Setka_physics.cpp:766-1258), plasma part: state s[8], n_h
 * hydrogen fluids h[5k..5k+4] = (rho_Hk, p_Hk, V_Hk); out[5] = (rho, m_x, m_y, m_z, E) sources. */
void or_bera_source(const double *s, int n_h, const double *h, double a2, double n_h_lism, double kn,
                    double *out);

/* ---- mesh indexing (New_connect) ---- */
/* Builds faces of n_cells hexahedra; arrays sized 6*n_cells by the caller; returns n_faces. */
int64_t or_new_connect(int64_t n_cells, const int32_t *cell_nodes, int64_t n_nodes,
                       int32_t *face_nodes, int32_t *face_cells, int32_t *cell_faces);

/* ---- charge exchange ---- */
/* I(U,c) of MK_source.cpp with sigma = (1 - a2 ln g)^2, N1 x N2 midpoint rule. */
double or_cx_rate_ref(double U, double c, double a2, int N1, int N2);
/* Lindsay-Stebbings (2005) cross-section in code units (g in units of 10.3804 km/s). */
double or_sigma_ls(double g);
/* Exact FP64 Maxwellian charge-exchange source for one cell (per unit volume and time).
 * in: n_p, c_p^2 (=p/rho), V_p[3], n_H, c_H^2, V_H[3].  out[5]: rate, S_m (3), S_E (lab frame) */
void or_cx_source(double n_p, double cp2, const double *vp, double n_h, double ch2,
                  const double *vh, int N1, int N2, double *out);

/* ---- MLP (8 -> 64 -> 64 -> 64 -> 5, SiLU) in FP64 ---- */
/* params layout: W1[64*8] b1[64] W2[64*64] b2[64] W3[64*64] b3[64] W4[5*64] b4[5] (row = out) */
int64_t or_mlp_n_params(void);
void or_mlp_forward(const double *params, int64_t n, const double *x, double *y);
/* the same network with the matrix-core inference's roundings: BF16 weights and layer inputs
 * (features and activations), FP32 biases */
void or_mlp_forward_bf16(const double *params, int64_t n, const double *x, double *y);
/* mean-squared-error loss over n samples and its gradient (same layout as params) */
double or_mlp_loss_grad(const double *params, int64_t n, const double *x, const double *y_target,
                        double *grad);

#ifdef __cplusplus
}
#endif
#endif
