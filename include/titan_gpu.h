/*
 * titan_gpu.h -- C ABI of the MI355X (gfx950) implementation of TITAN's explicit ideal-MHD step.
 *
 * The reference (SergeyMSU/TITAN_C-) has no library boundary: its step sits behind C++ member
 * functions of one executable.  Each entry point below names the reference seam it replaces.
 * Conventions:
 *   - return value 0 = success, negative = TITAN_E_* status (the reference prints and calls
 *     exit(-1) at the corresponding sites; this library returns instead);
 *   - all arrays are host arrays owned by the caller, in REFERENCE numbering (cells in the
 *     caller's order, faces in New_connect order); no pointer is retained after a call;
 *   - one context per host thread; every call is synchronous on the context's HIP stream;
 *   - the state of a cell is 8 FP64 primitives (rho, Vx, Vy, Vz, p, Bx, By, Bz), B in the
 *     reference's Gaussian code units (magnetic pressure B^2/8pi), and, when the context carries
 *     them (titan_params.passive_scalars), the passive scalars Q and rho_He (titan_set_scalars).
 */
#ifndef TITAN_GPU_H
#define TITAN_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TITAN_OK 0
#define TITAN_E_ARG (-1)            /* invalid argument / size                                 */
#define TITAN_E_HIP (-2)            /* HIP runtime error                                       */
#define TITAN_E_RIEMANN (-3)        /* impossible wave ordering (Phys_param.cpp:1194-1206,1834-1839) */
#define TITAN_E_NONFINITE_STATE (-4)/* NaN density (Setka_physics.cpp:1962-1978)               */
#define TITAN_E_TOPOLOGY (-5)       /* face/cell topology violates Setka::Test_geometr          */
#define TITAN_E_NOT_TRAINED (-6)    /* surrogate requested before parameters were set          */
#define TITAN_E_IO (-7)             /* file cannot be opened, written, or is truncated / damaged
                                       (Setka_physics.cpp:4110-4113,4213-4216 print and exit) */
#define TITAN_E_SURFACE (-8)        /* NaN or subnormal contact speed of a heliopause face while
                                       fitting the surfaces (Setka_move.cpp:336-350 print and exit),
                                       or an HP grid parabola through two points of equal abscissa
                                       (solveQuadraticEquation, Help_funk.cpp:290-300, exits)   */

/* face boundary codes (Type_Gran, Setka_physics.cpp:212-236; ZERO_GRADIENT is an addition) */
#define TITAN_BC_INTERIOR 0
#define TITAN_BC_INNER_HARD 1       /* fixed state (plasma_condition 2)                    */
#define TITAN_BC_OUTER_HARD 2       /* fixed state (plasma_condition 2)                    */
#define TITAN_BC_OUTER_SOFT 3       /* cell copy + suction fix (plasma_condition 1)        */
#define TITAN_BC_ZERO_GRADIENT 4    /* pure cell copy                                      */

/* geometry modes */
#define TITAN_GEOM_REF 0            /* Gran::Culc_measure / Cell::Culc_volume formulas     */
#define TITAN_GEOM_GCL 1            /* exactly conservative (bilinear faces, swept volumes) */

/* arithmetic of the face flux */
#define TITAN_PREC_FP64 0
#define TITAN_PREC_MIXED 1          /* FP32 Riemann arithmetic; FP64 states, accumulation, geometry */

/* per-cell charge-exchange source */
#define TITAN_SRC_NONE 0
#define TITAN_SRC_EXACT 1           /* FP64 Maxwellian quadrature (Lindsay-Stebbings sigma) */
#define TITAN_SRC_MLP 2             /* BF16-MFMA surrogate, exact fallback out of range     */
#define TITAN_SRC_BERA 3            /* the reference's closed-form source (Calc_sourse_MF_Bera) with
                                       the hydrogen fluids of titan_set_h_fluids                */

typedef struct titan_ctx titan_ctx;

/* Mesh in reference numbering (TITAN/source/Gran.h:252-282, Cell.h:17-36, Yzel.h:352). */
typedef struct {
    int64_t n_nodes, n_cells, n_faces;
    const int32_t *cell_nodes;   /* [n_cells*8] hexahedron (Cell::yzels)                       */
    const int32_t *face_nodes;   /* [n_faces*4] (Gran::yzels)                                  */
    const int32_t *face_cells;   /* [n_faces*2] Gran::cells[0], cells[1] (-1 on the boundary)  */
    const int32_t *cell_faces;   /* [n_cells*6] Cell::grans, ascending face id                 */
    const int8_t *face_bc;       /* [n_faces] TITAN_BC_*                                       */
    const int8_t *face_method;   /* [n_faces] 0 Lax, 1 HLL, 2 HLLC, 3 HLLD (Gran::Get_method)  */
    const uint8_t *face_flags;   /* [n_faces] bit0 left_ydar (TS_hard), bit1 contact, bit2
                                    heliopause face (Gran::type2 == HP), bit3 termination-shock
                                    face (type2 == TS), bit4 bow-shock face (type2 == BS);
                                    nullable                                                   */
    const double *face_bstate;   /* [n_faces*8] fixed state of hard boundary faces; nullable   */
    const int8_t *cell_zone;     /* [n_cells] 1..4 = Type_cell::Zone_1..Zone_4                 */
    const uint8_t *cell_is_tvd;  /* [n_cells] Cell::is_TVD (0 => B frozen)                     */
    const double *node_xyz;      /* [n_nodes*3] initial coordinates (Yzel::coord[0])           */
    const double *node_disp;     /* [n_nodes*3] motion amplitude vectors D_i; nullable         */
    const double *face_bscalar;  /* [n_faces*2] Q, rho_He of the fixed state of hard boundary faces
                                    (Gran::parameters, read by Setka_gran_snos.cpp:51-57); required
                                    when passive_scalars is on and a hard face exists, else nullable */
} titan_mesh;

/* Parameters the hot path reads (Phys_param.h:95-122, set_parameters.txt names in brackets). */
typedef struct {
    double gamma;                /* [gamma]                                                    */
    double kfl;                  /* [KFL]                                                      */
    double v_inf;                /* [Velosity_inf]                                             */
    int32_t move_mesh;           /* [move_setka]: 0 fixed; 1 prescribed motion
                                    x += dt*omega*cos(omega t+phase)*D_i (node_disp); 2 fitted
                                    surfaces: Setka::Culc_Velocity_surface + Luch::dvigenie of
                                    every ray before each step (Setka_physics.cpp:1616-1627;
                                    titan_set_rays)                                            */
    int32_t n_disc;              /* wave-speed estimate of chlld (0 on the step path)          */
    int32_t geom_mode;           /* TITAN_GEOM_*                                               */
    int32_t precision;           /* TITAN_PREC_*                                               */
    int32_t source_mode;         /* TITAN_SRC_*                                                */
    int32_t cx_nodes;            /* quadrature nodes per direction of the exact source (e.g. 100) */
    double motion_omega, motion_phase;
    int32_t tvd;                 /* [TVD]: second-order limited face states (Setka::Snos_on_Gran, 
                                    Setka_gran_snos.cpp:186-736) on faces whose cells both have
                                    is_TVD; neighbours chosen once from the initial nodes
                                    (Setka::Init_TVD, Setka_physics.cpp:648-744)                */
    int32_t snos_on_hp;          /* [Snos_on_HP]: TVD also on faces flagged heliopause (face_flags
                                    bit2; Setka_gran_snos.cpp:188-189)                          */
    double par_a_2;              /* [par_a_2] sigma = (1 - a_2 ln U)^2 of TITAN_SRC_BERA (0.102046) */
    double par_n_H_LISM;         /* [par_n_H_LISM] (3.0); TITAN_SRC_BERA scales by n_H_LISM/Kn   */
    double par_Kn;               /* [par_Kn] (40.0906)                                          */
    /* fitted surfaces (Phys_param.h:137-162; defaults of Phys_param.cpp:500-520 in brackets) */
    int32_t move_TS, move_HP, move_BS;    /* [move_TS], [move_HP], [move_BS] (1, 1, 1)         */
    int32_t sglag_TS, sglag_HP, sglag_BS; /* smoothing [sglag_TS] (1), [sglag_HP] (1), [sglag_BS]
                                             (0).  sglag_HP applies the sphere term near the
                                             nose (Setka_move.cpp:566-622) only                */
    double velocity_TS, velocity_HP, velocity_BS; /* (0.1, 0.1, 0.3)                            */
    double sglag_TS_k_sphere, sglag_HP_k_sphere, sglag_BS_k; /* (0.01, 0.01, 0.05)              */
    int32_t passive_scalars;     /* 1: the step carries the plasma's passive scalars Q (heliopause
                                    marker) and rho_He (Phys_param.cpp:232-233, always present in the
                                    reference): reconstructed as rho is (Setka_gran_snos.cpp:364-395,
                                    715-725, 848-852), konvect fluxes of chlld (Phys_param.cpp:
                                    1079-1105,1216-1218,1356-1418,1701-1820) on every face
                                    (Setka_physics.cpp:2349-2361,2485-2493), advected and clipped in
                                    the cell loop (:1781-1817,2004-2016); Q also routes the closed-form
                                    source's zone (determ_zone, :1427-1482) and rho_He its proton split
                                    (Plasma_components_1, Phys_param.cpp:628-665).  0: neither is
                                    carried (read as 0 by TITAN_SRC_BERA).                          */
} titan_params;

/* Rays of node motion: Luch objects of type A_Luch / A2_Luch (Luch.h:9-25), each a straight line
 * of nodes from the inner sphere R0 to R5 at polar angle `the` from +x and azimuth `phi`
 * (x = r cos the, y = r sin the cos phi, z = r sin the sin phi).  Node j of a ray: j < M0 on
 * [R0, R1); j = M0 at R1 (never moved); M1 nodes up to the termination shock at index
 * M0+M1+1 (Yzels_opor[1]); M11 evenly spaced nodes past it and M2-M11 more up to the heliopause
 * at M0+M1+M2+2 (Yzels_opor[2]); M3 nodes up to the bow shock at M0+M1+M2+M3+3 (Yzels_opor[3]);
 * M4 nodes up to R5 at the last index.  nodes_per_ray = M0+M1+M2+M3+M4+5. */
typedef struct {
    int64_t n_rays;
    int32_t nodes_per_ray;
    const int32_t *ray_nodes;    /* [n_rays*nodes_per_ray] reference node ids, inner to outer   */
    const double *ray_angles;    /* [n_rays*2] (the, phi) = Luch::parameters["the"], ["phi"]     */
    double R0, R1, R5;           /* Geo_param::R0, R1, R5                                       */
    int32_t M0, M1, M11, M2, M3, M4; /* Geo_param::M0 ... M4                                    */
    double da1, da2, da3, da4, da5;  /* Geo_param::da1 ... da5 (per-ray overrides not taken)     */
    double phi_a;                /* polar angle from +x of A_Luch[0][1]->Yzels[0]: the HP sphere
                                    smoothing and Smooth_head_HP3 act where the angle is at most
                                    phi_a + pi/180 (Setka_move.cpp:566-580, 2048-2064)           */
} titan_rays;

/* The other ray families of Luch::dvigenie (Luch.cpp:153-739), moved after the A rays by every
 * fitted step and by titan_dvigenie, family by family in the reference's order (Setka.cpp:1850-1880
 * sorts the rays A, B, C, D, E; H and G follow): B_Luch, C_Luch and C2_Luch, D_Luch, E_Luch, H_Luch,
 * G_Luch.  Each ray: its nodes (Luch::Yzels, reference ids, inner to outer), support nodes
 * (Luch::Yzels_opor[0..3], -1 where its type has fewer) and angles (Luch::parameters "the", "phi";
 * D, E, H and G read only phi).  R0, R1, R5 and M0..M4 are those of titan_set_rays, which must come
 * first.  Nodes each type needs (indices on the ray): B M0+M1+M2+M3+M4+5, C M0+M1+M11+N4+N5+2,
 * D M2-M11+M3+M4+4, E M3+M4+3, H 5, G M2-M11-MF+2; support slots read: B 1-3, C 1, D 0-2, E 0-2,
 * H 0-3, G 0-2.  B sets the x of its support nodes 2, 3 and D of its 1, 2.  Rays of one family run
 * in parallel unless one reads or writes a node another of its family writes; such a family runs
 * ray after ray in the given order, as the reference does.  Per-ray overrides of the Geo_param
 * values (Luch::parameters "dd3", "ba1", ...) are not taken. */
typedef struct {
    int64_t n_rays;
    const int8_t *ray_type;      /* TITAN_RAY_B ... TITAN_RAY_G                                  */
    const int64_t *ray_off;      /* [n_rays+1] offsets into ray_nodes                            */
    const int32_t *ray_nodes;
    const int32_t *ray_opor;     /* [n_rays*4]                                                   */
    const double *ray_angles;    /* [n_rays*2] (the, phi)                                        */
    double L6, L7;               /* Geo_param::L6, L7: x of the tail's first and last C-ray node */
    int32_t N4, N5, MF;          /* Geo_param::N4, N5, MF                                        */
    double dd1, dd2, dd3, dd4, dd5, dd6, dd7, dd8; /* Geo_param (Hermite slopes; the C rays take
                                    dd5 and dd6 truncated to int, as Luch.cpp:166-167 does)      */
    double ba1, ba2, ba3, ba4, ba5, md1, md2, md3, ea1, ea2, ea3; /* Geo_param spacing laws      */
} titan_rays_ext;
enum { TITAN_RAY_B = 1, TITAN_RAY_C = 2, TITAN_RAY_D = 3, TITAN_RAY_E = 4, TITAN_RAY_H = 5, TITAN_RAY_G = 6 };

typedef struct {
    int64_t steps_done;
    double dt_last;              /* dt used by the last step                                   */
    double dt_next;              /* min over interior faces computed in the last step          */
    double t_sim;                /* simulated time                                             */
    int64_t n_rho_floor, n_p_floor; /* floor activations (Setka_physics.cpp:1784,1981), last step */
    int64_t n_fallback;          /* cells that used the exact source in MLP mode, last step    */
    double ms_total;             /* wall time of the titan_go call (HIP events)                */
    int64_t n_fp64_faces;        /* TITAN_PREC_MIXED: face solves handed to the FP64 solver (ill- 
                                    conditioned HLLD Alfven denominators), last step            */
} titan_step_stats;

/* ---- context ------------------------------------------------------------------------------- */
/* Replaces Setka::Setka + Calculating_measure + Init_physics for the hot-path subset. */
int titan_create(titan_ctx **ctx, const titan_mesh *mesh, const titan_params *params);
int titan_destroy(titan_ctx *ctx);
/* Last error of the context; index = reference face/cell id of the first failure or -1. */
int titan_last_error(titan_ctx *ctx, int64_t *index);

/* ---- state ---------------------------------------------------------------------------------- */
/* Cell::parameters[layer] for the 8 plasma keys, [n_cells*8] cell-major. */
int titan_set_state(titan_ctx *ctx, const double *prim, int layer);
int titan_get_state(titan_ctx *ctx, double *prim, int layer);
/* Cell::parameters[layer]["Q"], ["rho_He"]: qhe [n_cells*2] = (Q, rho_He) per cell.  TITAN_E_ARG when
 * the context does not carry the passive scalars.  Both are 0 until set. */
int titan_set_scalars(titan_ctx *ctx, const double *qhe, int layer);
int titan_get_scalars(titan_ctx *ctx, double *qhe, int layer);
/* Yzel::coord[layer], [n_nodes*3]. */
int titan_set_nodes(titan_ctx *ctx, const double *xyz, int layer);
int titan_get_nodes(titan_ctx *ctx, double *xyz, int layer);
/* Neutral-H moments per cell for the charge-exchange source: [n_cells*5] = n_H, c_H^2, V_H. */
int titan_set_neutrals(titan_ctx *ctx, const double *nh5);
/* Hydrogen fluids per cell for TITAN_SRC_BERA (the keys rho_Hk, p_Hk, Vx_Hk, Vy_Hk, Vz_Hk of
 * Cell::parameters, k = 1..n_h, read by Setka_physics.cpp:820-826): h [n_cells][n_h][5] =
 * (rho_Hk, p_Hk, Vx_Hk, Vy_Hk, Vz_Hk), 1 <= n_h <= 9. */
int titan_set_h_fluids(titan_ctx *ctx, int n_h, const double *h);

/* ---- geometry ------------------------------------------------------------------------------- */
/* Setka::Calculating_measure(layer) (Setka.cpp:2465-2502) on the device, in the context's
 * geometry mode.  titan_get_geometry copies any non-NULL output in reference numbering. */
int titan_calculating_measure(titan_ctx *ctx, int layer);
int titan_get_geometry(titan_ctx *ctx, int layer, double *cell_center, double *cell_volume,
                       double *face_center, double *face_normal, double *face_area);

/* ---- face loop -------------------------------------------------------------------------------- */
/* One pass of the face loop of Setka::Go (Setka_physics.cpp:1634-1662) calling
 * Setka::Culc_Gran_Potok (Setka_physics.cpp:2280-2571) for every face, with states of layer
 * `now`, geometry of layer `now` and, when the mesh moves, face velocities from layers now and
 * 1-now with time step dt.  Outputs (nullable): face_flux [n_faces*9] = 8 fluxes * area
 * (Prho, PVx..PVz, Pp(energy), PBx..PBz) and PdivB; face_kflux [n_faces*2] = PQ, Prho_He (only
 * when the context carries the passive scalars, else TITAN_E_ARG); dt_min = min over interior
 * faces of the CFL time step; branch [n_faces] = n_state<<4 | ibn<<3 | region. */
int titan_gran_potok(titan_ctx *ctx, int now, double dt, double *face_flux, double *face_kflux,
                     double *dt_min, uint8_t *branch);

/* ---- step --------------------------------------------------------------------------------- */
/* Setka::Go(is_inner_area, steps) (Setka_physics.cpp:1484-2278) over the whole mesh: the first
 * call copies layer 0 to layer 1 and starts the lagged dt at 1e-5; later calls continue. */
int titan_go(titan_ctx *ctx, int64_t steps, titan_step_stats *stats);
/* Setka::Go(is_inner_area, steps) (Setka_physics.cpp:1484-2278) as the reference runs it: the
 * steps update only the cells of one area -- the inner area (cells whose nodes are all inner,
 * Setka::Init_boundary_grans, Setka_physics.cpp:113-212) or the outer area (every other cell) --
 * from the fluxes of the area's faces (faces with a cell in the area), whose interior faces alone
 * give the next dt; cells outside the area keep their state.  Every call starts the dt schedule
 * again at 1e-5 (Setka_physics.cpp:1582).  The mesh moves as a whole (move_mesh).  Requires
 * titan_set_inner_area.  titan_go above is the same loop over every cell with the dt schedule
 * continuing across calls. */
int titan_go_area(titan_ctx *ctx, int is_inner_area, int64_t steps, titan_step_stats *stats);
/* Cell::is_inner per cell [n_cells] (nonzero = inner), the split of Setka::Init_boundary_grans:
 * a cell is inner when all its nodes are among the first M0 + 8 nodes of an A/A2/B/C/C2 ray. */
int titan_set_inner_area(titan_ctx *ctx, const uint8_t *cell_is_inner);
/* Which layer holds the newest state after titan_go. */
int titan_current_layer(titan_ctx *ctx);
/* The second neighbours of every face chosen by Setka::Init_TVD (Setka_physics.cpp:648-744) from
 * the initial nodes, in reference numbering: tvd [n_faces*2] = (AA behind cells[0], BB behind
 * cells[1]), -1 where none qualifies (Gran::Sosed_1 / Sosed_2 = nullptr).  Independent of the
 * params' tvd switch; the step uses a face's pair only if the switch is on and both of the
 * face's cells have is_TVD. */
int titan_tvd_cells(titan_ctx *ctx, int32_t *tvd);

/* ---- Riemann solver ---------------------------------------------------------------------- */
/* Phys_param::chlld (Phys_param.cpp:882-1878; Phys_param.h:240-251) over n independent face
 * states on the device.  normal [n*3] unit, w [n] (nullable = 0), qL/qR [n*8], K passive
 * scalars kL/kR [n*K] (K <= 4), outputs q [n*8], k [n*K], ds [n*3] = (dsl, dsc, dsr),
 * branch [n].  precision = TITAN_PREC_*.  Returns TITAN_E_RIEMANN if any state hit the
 * reference's fatal branch (its q is then 0). */
int titan_chlld_batch(int64_t n, const int8_t *n_state, const int8_t *n_disc, const uint8_t *flags,
                      const double *normal, const double *w, const double *qL, const double *qR,
                      int K, const double *kL, const double *kR, double *q, double *k, double *ds,
                      uint8_t *branch, int precision);

/* The same batch solve on DEVICE arrays the caller owns (every array argument is a device
 * pointer; nullable ones as above), enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and waited for before returning.  It allocates nothing: its only working memory is one 256-byte
 * array per GPU, resident until the process ends, holding the fatal-branch and bad-code flags,
 * cleared at the start of every call.  Calls that overlap on one GPU (other threads or unordered
 * streams) race on those flags.  Returns TITAN_E_ARG for n > TITAN_NOALLOC_MAX_STATES or a null
 * required array without launching anything, TITAN_E_ARG after the solve when some n_state or
 * n_disc lies outside 0..3 (its outputs are then meaningless), TITAN_E_RIEMANN as titan_chlld_batch
 * does.  No counterpart in the reference: an addition for callers that cannot allocate device
 * memory (code under HIP graph capture, per-step loops, harnesses that forbid hipMalloc). */
#define TITAN_NOALLOC_MAX_STATES ((int64_t)1 << 31)
int titan_chlld_batch_noalloc(int64_t n, const int8_t *n_state, const int8_t *n_disc, const uint8_t *flags,
                              const double *normal, const double *w, const double *qL, const double *qR,
                              int K, const double *kL, const double *kR, double *q, double *k, double *ds,
                              uint8_t *branch, int precision, void *stream);

/* ---- charge-exchange source ---------------------------------------------------------------- */
/* The plasma source of Setka::Calc_sourse_MF_Bera (Setka_physics.cpp:766-1258) for the states of
 * `layer`: out [n_cells*5] = mass, momentum (3), energy (lab frame) per unit volume and time.
 * mode TITAN_SRC_BERA: the reference's closed form with the fluids of titan_set_h_fluids (mass
 * source 0; states without PUI; Q and rho_He of the layer when carried, else 0); TITAN_SRC_EXACT /
 * TITAN_SRC_MLP: the Maxwellian
 * charge-exchange integral with the neutrals of titan_set_neutrals (out[0] = exchange rate). */
int titan_sources(titan_ctx *ctx, int layer, int mode, double *out);
/* Calc_sourse_MF_Bera's hydrogen-fluid outputs (Setka_physics.cpp:1005-1094, 1241-1250) beside its
 * plasma ones, from the fluids of titan_set_h_fluids (n_h <= TITAN_MAX_H_FLUIDS) at layer `layer`:
 * out_h [n_cells][n_h][5] = (rho, m_x, m_y, m_z, E) source of each fluid -- what its exchanges with
 * the protons take from it, plus the atoms those exchanges make, which join fluid
 * arise[(zone-1)*n_h + k] (1-based, zone = determ_zone of the cell; NULL: fluid `zone`, the
 * hydrogen_arise tables of the run without pickup ions, Phys_param.cpp:149-171) with the protons'
 * momentum and energy; all times n_H_LISM/Kn, then over n_H_LISM.  out [n_cells*5] (nullable):
 * the plasma sources, as titan_sources(TITAN_SRC_BERA).  Pickup-ion fluids are not carried. */
#define TITAN_MAX_H_FLUIDS 16
int titan_sources_h(titan_ctx *ctx, int layer, const int8_t *arise, double *out, double *out_h);
/* Standalone exact source for n states: in [n*10] = n_p, c_p^2, V_p(3), n_H, c_H^2, V_H(3). */
int titan_cx_exact_batch(int64_t n, const double *in10, int nodes, double *out5);

/* ---- Monte-Carlo neutral densities -------------------------------------------------------- */
/* Setka::MK_go with MK_fly_immit (Setka_MK.cpp:1225-2316) for the interstellar hydrogen (sort 4)
 * that enters through the outer boundary faces (bc 2, 3), on the nodes and states of `layer`:
 * particles per face and weights by the rule of Setka_MK.cpp:1233,1380-1390 (N_per_gran =
 * n_per_face, inflow Get_Spotok_inf), births by Gran::Get_Random_pozition and
 * Setka::Velosity_initial, straight flights cell to cell (Setka::Time_to_vilet), charge exchange
 * by optical depth with nu = rho_Th Velosity_1 sigma / par_Kn and the proton velocity of
 * Setka::M_K_Change_Velosity (a new particle takes the sort of the cell's zone); rho_Th and p_Th
 * are Phys_param::Plasma_components_1 at the cell's type, with the layer's rho_He when the passive
 * scalars are carried (else 0), as the flight calls it (Setka_MK.cpp:1944-1966).  density
 * [n_cells*5] = MK_n_H, MK_n_H1..MK_n_H4 of Cell::MK_Add_particle (Cell.cpp:1042-1083) per unit
 * volume.  A particle that comes within 1.01 R0 of the centre after crossing a face (R0: the
 * radius of the inner boundary faces' nodes) is carried along its velocity to where it leaves that
 * sphere and goes on in the cell the ray from the centre through that point enters
 * (Setka_MK.cpp:2138-2160).  stats [8] (nullable): particles, charge exchanges, left through the
 * outer boundary, passes through the Sun's sphere, lost (no exit found after 11 nudges to the cell
 * centre, no cell after a pass, or max_steps cells crossed), births dropped (the birth sampler gave
 * up after 10000 tries), flight segments (one per tally: a cell crossed, or a flight to an exchange),
 * flights in which an exchange sampler gave up after 10000 tries.  Random numbers:
 * a counter-based SplitMix64 stream per particle from `seed` (the reference reads rnd_my.txt,
 * which it does not ship); particles are numbered face by face in ascending reference face id.
 * The tallies are FP64 atomic sums, so repeated calls agree to rounding, not bit for bit.  The
 * first call allocates the flight records and tallies (320 bytes per cell), which the context
 * keeps until titan_destroy; calls on one context are serialised on its stream. */
int titan_mk_go(titan_ctx *ctx, int layer, int32_t n_per_face, uint64_t seed, int32_t max_steps,
                double *density, int64_t *stats);

/* The pickup-ion source tallies of MK_source_S for the following titan_mk_go / titan_mk_go_zone calls
 * (Cell::Init_S, Cell::MK_Add_pui_source, Cell.cpp:77-87, 1087-1133, called at Setka_MK.cpp:2013,
 * 2098): each flight segment adds mu t to S-[cell][iw] and mu t nu_ex to S+[cell][iw][k-1], where
 * iw = int(u / w_r * n_w) clamped to 0..n_w-1 bins the relative speed u = |V - V_p| of the atom and
 * the cell's plasma, nu_ex is the segment's exchange rate, and k = arise[(zone-1)*4 + sort-1] is
 * the pickup-ion fluid the exchange feeds (0: none) at the zone of Setka::determ_zone (Q routes it
 * when the passive scalars are carried, else the cell's type stands in).  arise [16] (nullable):
 * values 0..2; NULL takes proton_arise_1..4 of the run with pickup ions, parent 0
 * (Phys_param.cpp:93-126).  After the flights the library applies Cell::MK_normir_Moments and
 * Cell::MK_calc_Sm (Cell.cpp:1135-1169, 1225-1241): S- becomes the loss rate at each bin's centre
 * speed, S+ a density per unit cell volume and unit velocity-space volume.  n_w = 0 turns the
 * tallies off; n_w <= 4096.  The arrays take 24 n_w bytes per cell on the device. */
int titan_set_mk_source_s(titan_ctx *ctx, int32_t n_w, double w_r, const int8_t *arise);
/* The last Monte-Carlo call's S- [n_cells*n_w] and S+ [n_cells*n_w*2] (per cell the column-major
 * (2, n_w) matrix of Cell::pui_Sp), reference cell order.  TITAN_E_ARG before such a call. */
int titan_mk_get_s(titan_ctx *ctx, double *sm, double *sp);
/* Cell::MK_normir_Moments + Cell::MK_calc_Sm on n cells' raw tallies (host arrays, same layout):
 * vol [n], par_a_2, par_Kn. */
int titan_mk_s_finish_batch(int64_t n, int32_t n_w, double w_r, double a2, double kn, const double *vol,
                            const double *sm, const double *sp, double *sm_out, double *sp_out);
/* One cell's file of Cell::write_S_ToFile / Cell::read_S_FromFile (Cell.cpp:89-125, 374-420;
 * "data_SpSm/func_cells_SpSm_<number>.bin"): int32 2, n_w, n_w, then S- [n_w] and S+ [n_w*2].
 * The reader returns TITAN_E_ARG for other sizes (the reference exits), TITAN_E_IO for a missing
 * or truncated file. */
int titan_mk_s_write(const char *path, int32_t n_w, const double *sm, const double *sp);
int titan_mk_s_read(const char *path, int32_t n_w, double *sm, double *sp);
/* Setka::Set_MK_Zone (Setka_MK.cpp:142-683) from the centres of node layer 0: cell_mk [n_cells]
 * (nullable) = the MK zone 1..7 of each cell (Zone_1 -> 1; Zone_2, Zone_3, Zone_4 -> 2 / 3, 4 / 5,
 * 6 / 7 by centre x > 0 / x <= 0), face_mask [n_faces] (nullable) bit z-1 = the face is on zone z's
 * list (MK_Grans[z-1], by the rules of Setka_MK.cpp:356-641 on the face flags TS/HP/BS, boundary
 * codes, cell zones and centres).  Then the velocity distributions of the listed interior faces,
 * zeroed: per face, sort 1..4 and side, the AMR_f grid of Gran::Read_AMR -- (0, 20) x (-20, 20)^2 in
 * the face's frame (Vn, get_bazis(Vn)), 3 x 6 x 6 bins -- refined uniformly r times in each direction
 * (r = 1..4; the reference refines adaptively, AMR_f::Refine): 108 r^3 FP64 bins per face, sort
 * and side, 6,912 r^3 bytes per face.  TITAN_E_TOPOLOGY (titan_last_error: the cell or face)
 * for a cell zone outside 1..4 (Error 86743207564) or a face a zone lists twice (ERROR 6435856408).
 * Replaces any earlier zones and distributions. */
int titan_set_mk_zone(titan_ctx *ctx, int32_t r, int8_t *cell_mk, uint8_t *face_mask);
/* One pass of the zone scheme: MK_prepare(zone) + MK_go(zone) + MK_delete (Setka_MK.cpp:686-1752;
 * Setka.cpp:221-226 runs them over the zones 6,4,2,1,3,5,7,5,3,1,2,4,6,7,4,5,2,3,1).  The flux into
 * the zone through each listed face -- an interior face: each sort's distribution on the side
 * entering the zone (the weight it holds, Culk_SpotokV); a boundary face: Get_Spotok_inf of sort 4 --
 * gives mu_expect = total / (n_per_face x listed faces) and, sort by sort and face by face (ascending
 * reference id), the particle counts and weights of Setka_MK.cpp:1380-1390.  A particle from an
 * interior face draws its velocity from the distribution (AMR_f::Get_random_velosity: a bin with
 * probability its weight, Vn ~ Vn across it, Vt, Vm uniform); from a boundary face as titan_mk_go.
 * Flights stay in the zone's cells and end on a listed face, whose distribution on the side the
 * particle leaves by gains its weight in the bin of its velocity (AMR_f::Add_particle); the zone's
 * outgoing distributions are cleared first.  density [n_cells*5]: the zone's cells receive MK_n_H,
 * MK_n_H1..4 per unit volume; other cells are left as they are.  stats [9] (nullable): those of
 * titan_mk_go, then the flights that ended on the zone's border.  TITAN_E_ARG without
 * titan_set_mk_zone; TITAN_E_TOPOLOGY when a listed face's entering cell lies outside the zone
 * (Error 9767653421).  Each pass allocates the running sums of its incoming distributions
 * (864 r^3 bytes per listed interior face) for the call. */
int titan_mk_go_zone(titan_ctx *ctx, int layer, int32_t zone, int32_t n_per_face, uint64_t seed, int32_t max_steps,
                     double *density, int64_t *stats);
/* The distribution of a listed interior face (reference face id, sort 1..4, side 0: Vn along the
 * normal, 1: against it) to or from the caller: w [108 r^3] weight per bin, (i, j, k) with k fastest
 * (the reference saves them as data_AMR/func_grans_AMR_* files between passes). */
int titan_mk_dist_get(titan_ctx *ctx, int64_t face, int32_t sort, int32_t side, double *w);
int titan_mk_dist_set(titan_ctx *ctx, int64_t face, int32_t sort, int32_t side, const double *w);

/* ---- surrogate network (8-64-64-64-5 SiLU) ------------------------------------------------- */
/* Parameter layout (FP32): W1[64x8] b1[64] W2[64x64] b2[64] W3[64x64] b3[64] W4[5x64] b4[5],
 * rows = outputs; plus input normalisation mean/std [8+8], output scale [5], input box
 * lo/hi [8+8] outside which the exact source is used. */
int64_t titan_mlp_n_params(void);
int titan_mlp_set(titan_ctx *ctx, const float *params, const float *in_norm16, const float *out_scale5,
                  const float *box16);
int titan_mlp_get(titan_ctx *ctx, float *params);
/* Forward pass of n samples of already-normalised inputs x [n*8] -> y [n*5] on the BF16-MFMA
 * kernel (FP32 accumulation). */
int titan_mlp_forward(titan_ctx *ctx, int64_t n, const float *x, float *y);
/* Loss (MSE) and gradient [n_params] of n samples on the device in FP32. */
int titan_mlp_loss_grad(titan_ctx *ctx, int64_t n, const float *x, const float *y, float *loss,
                        float *grad);
/* Adam training on the device: n samples x [n*8], y [n*5] (normalised), `epochs` passes of
 * mini-batches of `batch`, learning rate lr; loss_hist [epochs] (nullable). */
int titan_mlp_train(titan_ctx *ctx, int64_t n, const float *x, const float *y, int epochs, int batch,
                    float lr, uint64_t seed, float *loss_hist);

/* ---- checkpoint files (host) ------------------------------------------------------------------ */
/* The file of Setka::Save_cell_parameters / Setka::Download_cell_parameters
 * (Setka_physics.cpp:4107-4165, 4210-4285), little-endian: per cell u64 count and count x
 * (u64 key length, key bytes, f64 value); u8 1, u64 n_nodes, n_nodes x 3 f64 (coord[0]); 1000 zero
 * bytes.  titan_checkpoint_write writes n_keys keys for every cell, values [n_cells][n_keys], and
 * the node block when n_nodes >= 0 (flag 0 otherwise).  titan_checkpoint_read fills values
 * [n_cells][n_keys] with the named keys of each cell (NaN where a cell lacks a key; key order in
 * the file is free, other keys are skipped) and, if the file has a node block, xyz [n_nodes*3]
 * (n_nodes < 0: only report the count; xyz must then be NULL); *nodes_read = node count, 0 when
 * the file has no node block.  A node count different from n_nodes >= 0 is TITAN_E_ARG. */
int titan_checkpoint_write(const char *path, int64_t n_cells, int n_keys, const char *const *keys,
                           const double *values, int64_t n_nodes, const double *xyz);
int titan_checkpoint_read(const char *path, int64_t n_cells, int n_keys, const char *const *keys,
                          double *values, int64_t n_nodes, double *xyz, int64_t *nodes_read);
/* The file of Setka::Save_cell_MK_parameters / Setka::Download_cell_MK_parameters
 * (Setka_physics.cpp:4165-4209, 4287-4336), little-endian: u64 key count, each key once (u64
 * length, bytes), then per cell (All_Cell order) one f64 per key in that order, then 1000 zero
 * bytes.  titan_mk_parameters_write writes values [n_cells][n_keys] (e.g. MK_n_H, MK_n_H1..4 of
 * titan_mk_go).  titan_mk_parameters_read copies the file's value of every named key it holds into
 * values [n_cells][n_keys], except for the cells whose MK zone (cell_mk, nullable) is zone_except,
 * which keep theirs (the zone the caller is about to compute); keys the file lacks are left as
 * they are; *keys_found (nullable) = how many of the named keys the file holds.  TITAN_E_IO for a
 * missing or truncated file; TITAN_E_ARG when the values and the 1000-byte trailer do not fill
 * the file exactly (a file of another cell count). */
int titan_mk_parameters_write(const char *path, int64_t n_cells, int n_keys, const char *const *keys,
                              const double *values);
int titan_mk_parameters_read(const char *path, int64_t n_cells, int n_keys, const char *const *keys, double *values,
                             const int8_t *cell_mk, int32_t zone_except, int32_t *keys_found);
/* Setka::Save_cell_parameters on a context: the 8 plasma keys (rho, Vx, Vy, Vz, p, Bx, By, Bz) of
 * the newest layer, and Q, rho_He when the context carries them, then n_extra caller keys with
 * extra_values [n_cells][n_extra] (e.g. the hydrogen fluids kept on the host), and the newest node
 * layer. */
int titan_save_cell_parameters(titan_ctx *ctx, const char *path, int n_extra, const char *const *extra_keys,
                               const double *extra_values);
/* Setka::Download_cell_parameters on a context: each of the 8 plasma keys (and Q, rho_He when
 * carried) a cell has in the file replaces that value in both state layers (keys a cell lacks keep
 * their values); a node block
 * replaces both node layers, and the next titan_go re-measures the geometry, as Setka::Go does
 * on entry (Setka_physics.cpp:1495-1510).  The TVD neighbours stay as chosen; titan_init_tvd
 * chooses them again from the new nodes. */
int titan_download_cell_parameters(titan_ctx *ctx, const char *path);
/* Setka::Init_TVD (Setka_physics.cpp:648-744) again, from the newest node layer. */
int titan_init_tvd(titan_ctx *ctx);

/* ---- fitted surfaces ------------------------------------------------------------------------- */
/* The mesh's rays (the A_Luch / A2_Luch part of Setka::All_Luch): node lists, angles and the
 * Geo_param counts and radii of Luch::dvigenie.  Each node may lie on one ray only. */
int titan_set_rays(titan_ctx *ctx, const titan_rays *rays);
/* HP smoothing on the reference's ray grids A_Luch[i][j], B_Luch[i][j], E_Luch[i][j], D_Luch[i][j]
 * (azimuth i, place j along the grid), given as node lists built from them; applied by every fitted
 * step and titan_culc_velocity_surface with sglag_HP.  Each listed HP node gets, after the sphere
 * term, the angle term when titan_smooth_angle_hp gave it a circle centre (Setka_move.cpp:703-745:
 * towards the circle about (x, xcc, ycc) through the mean distance of its smoothing neighbours,
 * velocity_HP * sglag_HP_k_angle, no x part) and, for grid nodes, the four macros1 terms
 * (Setka_move.cpp:4-99, 880-1130): parabolas in (x, r_yz) through AA3, AA1, AA11 and AA33, AA3, AA1
 * (weight sglag_HP_along) and in (polar angle about x, r_yz) through AA2, AA4, AA44 and AA4, AA2, AA22
 * (weight sglag_HP_angle), each times velocity_HP * sglag_HP_k.  After the move, with move_HP, the
 * tail correction (Setka_move.cpp:1281-1323) on each azimuth's D-ray HP nodes: node N4-5 onto the
 * line in (x, r_yz) through nodes N4-4 and N4-6, nodes N4-3 .. onto the r_yz of node N4-4; with
 * move_BS the BS correction (:1325-1357): each azimuth's listed BS nodes onto the r_yz of its
 * reference node.  Replaces any earlier grid; n_grid = n_tail = n_bs = 0 switches it off. */
typedef struct {
    int64_t n_grid;
    const int32_t *grid_node;    /* [n_grid] HP nodes: Yzels_opor[2] of A rays (j >= 2) and B rays,
                                    Yzels_opor[1] of E rays and D rays (j < N4 - 1)              */
    const int32_t *grid_nb;      /* [n_grid*8] AA1, AA11 (one and two back along the grid), AA3,
                                    AA33 (one and two ahead), AA2, AA22 (azimuth i-1, i-2), AA4,
                                    AA44 (i+1, i+2): Setka_move.cpp:906-1128                     */
    double sglag_HP_k_angle, sglag_HP_k, sglag_HP_angle, sglag_HP_along; /* (0.001, 0.03, 1.8, 1.0),
                                    Phys_param.cpp:513-519                                       */
    int32_t n_tail, tail_len, N4; /* tail correction: n_tail azimuths of tail_len D rays (N4 >= 6) */
    const int32_t *tail;         /* [n_tail*tail_len] D_Luch[i][k]->Yzels_opor[1]                 */
    int32_t n_bs;                /* BS correction: n_bs azimuths                                  */
    const int32_t *bs_ref;       /* [n_bs] A_Luch[i][last]->Yzels_opor[3]                         */
    const int32_t *bs_off;       /* [n_bs+1] into bs_nodes                                        */
    const int32_t *bs_nodes;     /* B_Luch[i][*]->Yzels_opor[3], E_Luch[i][*] and D_Luch[i][*]->
                                    Yzels_opor[2]                                                */
} titan_hp_grid;

/* The B, C, D, E, H and G rays (Luch.cpp:153-739; Setka.cpp:1850-1900 builds them): replaces any
 * earlier set.  TITAN_E_ARG on a ray too short for its type, a support slot or node out of range,
 * N5 < 2 with C rays, or no titan_set_rays before; a later titan_set_rays drops them. */
int titan_set_rays_ext(titan_ctx *ctx, const titan_rays_ext *rays);
/* Setka::Smooth_angle_HP (Setka_move.cpp:2663-2760) from node layer 0: for each of n HP nodes, the
 * circle fit (fitCircleRobust, :2544-2620) of its (y, z) and those of its 6 azimuthal neighbours nb6
 * (azimuth i-1, i-2, i+1, i+2, i+3, i-3) gives its centre (xcc, ycc), and the neighbours join its
 * smoothing set (the set titan_smooth_head_hp3 later clears and refills near the nose, as the
 * reference's Yzel_sosed_sglag2 is).  TITAN_E_ARG for a node that is no HP surface node. */
int titan_smooth_angle_hp(titan_ctx *ctx, int64_t n, const int32_t *nodes, const int32_t *nb6);
/* The HP grid smoothing and the corrections after the move (titan_hp_grid).  TITAN_E_ARG for a
 * grid node that is no HP surface node or appears twice, a tail node that is no HP surface node, a
 * BS node that is no BS surface node, N4 < 6 or tail_len < N4 - 3 with tails. */
int titan_set_hp_grid(titan_ctx *ctx, const titan_hp_grid *grid);
/* fitCircleRobust (Setka_move.cpp:2544-2620) of n >= 3 points [n*2]: centre_radius = (cx, cy, R). */
int titan_circle_fit(int64_t n, const double *points, double *centre_radius);
/* Set-up of the surface smoothing from node layer 0 (Yzel::coord[0]): Setka::Smooth_head_TS3
 * (Setka_move.cpp:2205-2363) gives every TS node with at least 3 points its neighbours (the other
 * nodes of its TS faces) and the centre of the sphere fitted to it and them; Smooth_head_HP3
 * (:2046-2203) the same for HP nodes within phi_a + pi/180 of +x (rays required); Setka::
 * Find_Yzel_Sosed_for_BS (:2365-2397) the neighbours of BS nodes.  titan_go with move_mesh 2,
 * and titan_culc_velocity_surface, return TITAN_E_ARG while a switched-on smoothing lacks it. */
int titan_smooth_head_ts3(titan_ctx *ctx);
int titan_smooth_head_hp3(titan_ctx *ctx);
int titan_find_yzel_sosed_for_bs(titan_ctx *ctx);
/* Setka::Culc_Velocity_surface(now, time) (Setka_move.cpp:111-1357, the parts the A_Luch rays
 * carry): from the state and nodes of layer `now`, chlld wave speeds (S_L at TS, S_M at HP, S_R
 * at BS faces, w = 0) times velocity_K times the face normal, averaged per node with the
 * smoothing terms, move the TS/HP/BS nodes into layer 1-now.  node_velocity [n_nodes*3]
 * (nullable) receives the averaged velocity of every surface node (0 elsewhere).  A node sums its
 * faces in ascending face id (the reference's order varies between runs).  With the HP grid of
 * titan_set_hp_grid and the circle centres of titan_smooth_angle_hp, the HP angle and along-grid
 * terms and the tail and BS corrections apply as in the fitted step. */
int titan_culc_velocity_surface(titan_ctx *ctx, int now, double time, double *node_velocity);
/* Luch::dvigenie(layer) of every ray (Luch.cpp:6-152; Setka_physics.cpp:1620-1624). */
int titan_dvigenie(titan_ctx *ctx, int layer);
/* The sphere fit of Smooth_head_TS3 on the host: n >= 3 points [n*3] -> centre [3]. */
int titan_sphere_fit(int64_t n, const double *points, double *centre);

/* ---- mesh construction helpers (host) -------------------------------------------------------- */
/* Setka::New_connect (Setka.cpp:2066-2249): faces of n_cells hexahedra; face_nodes [6*n_cells*4],
 * face_cells [6*n_cells*2], cell_faces [n_cells*6] sized by the caller; returns n_faces (< 0 on
 * error). */
int64_t titan_mesh_connect(int64_t n_cells, const int32_t *cell_nodes, int64_t n_nodes,
                           int32_t *face_nodes, int32_t *face_cells, int32_t *cell_faces);
/* Synthetic meshes: jittered n^3 cube on [0,1]^3 ((n+1)^3 nodes), and a spherical shell of
 * n_layers radial layers x 6*n_edge^2 cube-sphere quads (radii [n_layers+1]). */
int titan_mesh_cube(int32_t n, double jitter, uint64_t seed, double *xyz, int32_t *cell_nodes);
int titan_mesh_shell(int32_t n_edge, int32_t n_layers, const double *radii, double *xyz,
                     int32_t *cell_nodes);

/* ---- timing --------------------------------------------------------------------------------- */
/* Mean device time (ms) per launch of the named kernel family over the last titan_go call,
 * measured with HIP events on the context stream: which = 0 step kernel (flux+update),
 * 1 geometry, 2 motion, 3 source (the exact-source kernel, or in TITAN_SRC_MLP mode the
 * out-of-box fallback kernel that follows each step).  Returns -1 if not measured. */
double titan_kernel_ms(titan_ctx *ctx, int which);
int titan_set_timing(titan_ctx *ctx, int enable);
/* Layout of the fused step: info8 = {cells per tile, max nodes per tile, max faces per tile,
 * LDS bytes per workgroup, device LDS limit, tiles, face records (border faces counted twice),
 * faces that reconstruct to second order}. */
int titan_tile_info(titan_ctx *ctx, int64_t *info8);

#ifdef __cplusplus
}
#endif
#endif
