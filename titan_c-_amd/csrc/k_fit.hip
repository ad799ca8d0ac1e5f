// k_fit.hip -- fitted-surface node motion: Setka::Culc_Velocity_surface (Setka_move.cpp:111-1357)
// and Luch::dvigenie for every ray family (Luch.cpp:6-739), the production caller of
// the moving-mesh step (Setka_physics.cpp:1616-1627).
//
// The kernels of a step, all FP64 and in the reference's operation order:
//   k_fit_faces   one thread per TS/HP/BS face: the face's states at the current layer, chlld with
//                 w = 0, and velocity_K * (S_L | S_M | S_R) * n;
//   k_fit_nodes   one thread per surface node: the average of its faces' vectors and of the sphere
//                 (TS, HP) or neighbour-mean (BS) smoothing term, then the move along the radius
//                 (HP behind the Sun: along the cylinder radius);
//   k_fit_ray_radii, k_fit_rays  each ray's moved TS/HP/BS radii, then one thread per device node:
//                 the A_Luch distribution between them;
//   k_fit_rays_ext  the B, C, D, E, H and G rays (Luch.cpp:153-739), one launch per family in that
//                 order, one thread per ray.
// The reference adds the face vectors into each node under a mutex from an OpenMP loop, so its
// summation order varies from run to run; here a node sums them in ascending reference face id,
// which makes the motion repeat bit for bit.
#define TITAN_FLUX_ONLY 1
#include "step_kernels.h"
#include "launch.h"

namespace titan {

namespace {
inline unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
}  // namespace

__global__ void __launch_bounds__(256) k_fit_faces(DevMesh M, DevParams P, FitDev F,
                                                   const double *__restrict__ prim, const double *__restrict__ X,
                                                   const double *__restrict__ cell_center, StepScalars *S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F.n_sf) return;
    const int64_t f = F.sf_face[i];
    const int kind = F.sf_kind[i];
    double Gc[3], n[3], area;
    face_geom_one(M, X, cell_center, f, P.geom_mode, Gc, n, area);
    const uint32_t code = M.face_code[f];
    double qL[8], qR[8];
    const bool okbc = face_states_at(M, P, f, code & 15u, prim, cell_center, Gc, n, qL, qR);
    // chlld(Get_method(), normal, w = 0, ..., null_bn = false, n_disc = 0), no face flags
    // (Setka_move.cpp:171-175, 328-333, 422-426)
    RiemannOut<double, 0> o;
    chlld<double, 0>((int)((code >> 4) & 15u), 0, 0, n[0], n[1], n[2], 0.0, qL, qR, nullptr, nullptr, o);
    if (!o.ok || !okbc) report_error(S, kErrFitRiemann, M.face_ref[f]);
    const double ds = kind == 1 ? o.dsl : (kind == 2 ? o.dsc : o.dsr);
    // Setka_move.cpp:336-350: a NaN or subnormal contact speed stops the reference
    if (kind == 2 && (isnan(ds) || (ds != 0.0 && fabs(ds) < 2.2250738585072014e-308)))
        report_error(S, kErrFitContact, M.face_ref[f]);
    const double s = F.velocity[kind] * ds;
    F.sf_vel[3 * i] = s * n[0];
    F.sf_vel[3 * i + 1] = s * n[1];
    F.sf_vel[3 * i + 2] = s * n[2];
}

namespace {
// solveQuadraticEquation (Help_funk.cpp:290-305): a x^2 + b x + c through three points, evaluated
// at t; false where two abscissae coincide
__device__ inline bool quad3_at(double x1, double y1, double x2, double y2, double x3, double y3, double t,
                                double &r) {
    const double denom = (x1 - x2) * (x1 - x3) * (x2 - x3);
    if (denom == 0) return false;
    const double a = (x3 * (y2 - y1) + x2 * (y1 - y3) + x1 * (y3 - y2)) / denom;
    const double b = (x3 * x3 * (y1 - y2) + x2 * x2 * (y3 - y1) + x1 * x1 * (y2 - y3)) / denom;
    const double c = (x2 * x3 * (x2 - x3) * y1 + x3 * x1 * (x3 - x1) * y2 + x1 * x2 * (x1 - x2) * y3) / denom;
    r = a * (t * t) + b * t + c;
    return true;
}
__device__ inline double ryz0(const double *p) { return sqrt(0.0 * 0.0 + p[1] * p[1] + p[2] * p[2]); }
}  // namespace

// time > 0: the step length of an explicit call; otherwise the context's current dt.
__global__ void __launch_bounds__(256) k_fit_nodes(FitDev F, const double *__restrict__ X1, double *__restrict__ X2,
                                                   StepScalars *__restrict__ S, double time) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F.n_sn) return;
    const double dt = time > 0.0 ? time : S->dt;
    const int kind = F.sn_kind[i];
    const int64_t g = F.sn_node[i];
    const double A0 = X1[3 * g], A1 = X1[3 * g + 1], A2 = X1[3 * g + 2];
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    int num = 0;
    if (F.move[kind]) {
        for (int j = F.sn_face_off[i]; j < F.sn_face_off[i + 1]; j++) {
            const double *s = F.sf_vel + 3 * (int64_t)F.sn_face[j];
            v0 += s[0]; v1 += s[1]; v2 += s[2];
            num++;
        }
    }
    const int nb0 = F.sn_nb_off[i], nb1 = F.sn_nb_off[i + 1];
    bool smooth = F.sglag[kind] != 0;
    if (smooth && kind == 2) {
        // Setka_move.cpp:577-580: only near the nose, polar angle from +x at the current layer
        const double phi = polar_angle(A0, sqrt(0.0 * 0.0 + A1 * A1 + A2 * A2));
        smooth = !(phi > F.phi_lim);
    }
    if (smooth && kind == 3) smooth = (nb1 - nb0) > 5;  // Setka_move.cpp:1190
    // a TS/HP node the set-up gave no neighbours (fewer than 3 points) gets no sphere term; the
    // reference would divide by the empty neighbour count there
    if (smooth && kind != 3 && nb1 == nb0) smooth = false;
    if (smooth) {
        const double c = F.velocity[kind] * F.k_smooth[kind];
        double V0, V1, V2;
        if (kind == 3) {
            // mean of the node and its BS neighbours (Setka_move.cpp:1180-1209)
            double B0 = A0, B1 = A1, B2 = A2;
            for (int j = nb0; j < nb1; j++) {
                const double *p = X1 + 3 * (int64_t)F.sn_nb[j];
                B0 += p[0]; B1 += p[1]; B2 += p[2];
            }
            const double cnt = (double)(nb1 - nb0) + 1.0;
            B0 /= cnt; B1 /= cnt; B2 /= cnt;
            V0 = c * (B0 - A0) / dt; V1 = c * (B1 - A1) / dt; V2 = c * (B2 - A2) / dt;
        } else {
            // onto the sphere about (xc, yc, zc) through the neighbours' mean distance
            // (Setka_move.cpp:466-506 TS, 583-622 HP)
            const double C0 = F.sn_centre[3 * i], C1 = F.sn_centre[3 * i + 1], C2 = F.sn_centre[3 * i + 2];
            const double e0 = A0 - C0, e1 = A1 - C1, e2 = A2 - C2;
            const double rr = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            double r = 0.0;
            for (int j = nb0; j < nb1; j++) {
                const double *p = X1 + 3 * (int64_t)F.sn_nb[j];
                const double d0 = p[0] - C0, d1 = p[1] - C1, d2 = p[2] - C2;
                r += sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            }
            r /= (double)(nb1 - nb0);
            const double k = r / rr;
            const double P0 = k * e0 + C0, P1 = k * e1 + C1, P2 = k * e2 + C2;
            V0 = c * (P0 - A0) / dt; V1 = c * (P1 - A1) / dt; V2 = c * (P2 - A2) / dt;
        }
        v0 += V0; v1 += V1; v2 += V2;
        num++;
    }
    if (kind == 2 && F.sglag[2]) {
        if (F.sn_cc && !isnan(F.sn_cc[2 * i]) && nb1 > nb0) {
            // Setka_move.cpp:703-745: towards the circle about (x, xcc, ycc) through the mean distance
            // of the neighbours' (y, z) from its centre; no x component
            const double cy = F.sn_cc[2 * i], cz = F.sn_cc[2 * i + 1];
            const double d1 = A1 - cy, d2 = A2 - cz;
            const double rr = sqrt(d1 * d1 + d2 * d2);
            double r = 0.0;
            for (int j = nb0; j < nb1; j++) {
                const double *p = X1 + 3 * (int64_t)F.sn_nb[j];
                const double e1 = cy - p[1], e2 = cz - p[2];
                r += sqrt(e1 * e1 + e2 * e2);
            }
            r /= (double)(nb1 - nb0);
            const double k = r / rr;
            const double c = F.velocity[2] * F.k_angle;
            const double B1 = k * (A1 - cy) + cy, B2 = k * (A2 - cz) + cz;
            v0 += 0.0;
            v1 += c * (B1 - A1) / dt;
            v2 += c * (B2 - A2) / dt;
            num++;
        }
        if (F.sn_grid && F.sn_grid[8 * i] >= 0) {
            // macros1 (Setka_move.cpp:4-99): parabolas through the grid neighbours in (x, r_yz) and in
            // (polar angle about x, r_yz), each pulling r_yz towards its value at the node
            const int32_t *G = F.sn_grid + 8 * i;
            const double *a1 = X1 + 3 * (int64_t)G[0], *a11 = X1 + 3 * (int64_t)G[1], *a3 = X1 + 3 * (int64_t)G[2],
                         *a33 = X1 + 3 * (int64_t)G[3], *a2 = X1 + 3 * (int64_t)G[4], *a22 = X1 + 3 * (int64_t)G[5],
                         *a4 = X1 + 3 * (int64_t)G[6], *a44 = X1 + 3 * (int64_t)G[7];
            const double Ap[3] = {A0, A1, A2};
            const double p1 = ryz0(Ap);
            double rad[4];
            bool ok = quad3_at(a3[0], ryz0(a3), a1[0], ryz0(a1), a11[0], ryz0(a11), A0, rad[0]);
            ok &= quad3_at(a33[0], ryz0(a33), a3[0], ryz0(a3), a1[0], ryz0(a1), A0, rad[1]);
            const double t0 = polar_angle(A1, A2);
            const double t2 = polar_angle(a2[1], a2[2]), t4 = polar_angle(a4[1], a4[2]);
            const double t22 = polar_angle(a22[1], a22[2]), t44 = polar_angle(a44[1], a44[2]);
            ok &= quad3_at(t2, ryz0(a2), t4, ryz0(a4), t44, ryz0(a44), t0, rad[2]);
            ok &= quad3_at(t4, ryz0(a4), t2, ryz0(a2), t22, ryz0(a22), t0, rad[3]);
            if (!ok) report_error(S, kErrFitQuad, F.sn_ref[i]);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double c = F.velocity[2] * (q < 2 ? F.hp_along : F.hp_angle) * F.k_hp;
                double B1 = A1, B2 = A2;
                B1 *= rad[q] / p1;
                B2 *= rad[q] / p1;
                v0 += c * (A0 - A0) / dt;
                v1 += c * (B1 - A1) / dt;
                v2 += c * (B2 - A2) / dt;
                num++;
            }
        }
    }
    if (num > 0) {
        v0 /= num; v1 /= num; v2 /= num;
    }
    F.sn_vel[3 * i] = v0;
    F.sn_vel[3 * i + 1] = v1;
    F.sn_vel[3 * i + 2] = v2;
    if (F.move[kind]) {
        // Setka_move.cpp:1228-1278: along the unit radius (HP with x < 0: the radius in the yz plane)
        double B0 = A0, B1 = A1, B2 = A2;
        if (kind == 2 && !(A0 >= 0.0)) B0 = 0.0;
        const double z = B0 * B0 + B1 * B1 + B2 * B2;
        if (z > 0.0) {
            const double sz = sqrt(z);
            B0 /= sz; B1 /= sz; B2 /= sz;
        }
        const double s = (v0 * B0 + v1 * B1 + v2 * B2) * dt;
        X2[3 * g] = A0 + s * B0;
        X2[3 * g + 1] = A1 + s * B1;
        X2[3 * g + 2] = A2 + s * B2;
    }
}

// After the move (Setka_move.cpp:1281-1357): threads [0, n_tail) take an azimuth's D-ray HP nodes
// -- node N4-5 onto the line in (x, r_yz) through nodes N4-4 and N4-6, nodes N4-3 .. onto the r_yz
// of node N4-4 -- and the next n_bs threads an azimuth's BS nodes, onto the r_yz of its reference
// node.  The two sets of nodes are disjoint (HP, BS).
__global__ void __launch_bounds__(64) k_fit_fix(FitDev F, double *__restrict__ X) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < F.n_tail) {
        if (!F.move[2]) return;
        const int32_t *L = F.tail + (int64_t)t * F.tail_len;
        const int N4 = F.N4, NN = F.tail_len - 1;
        auto Q = [&](int k) { return X + 3 * (int64_t)L[k]; };
        const double h1 = ryz0(Q(N4 - 4)), xx1 = Q(N4 - 4)[0];
        const double h2 = ryz0(Q(N4 - 6)), xx2 = Q(N4 - 6)[0];
        double *yz = Q(N4 - 5);
        const double hh = ryz0(yz), xx = yz[0];
        const double d = (h1 - h2) / (xx1 - xx2);  // linear2 (Help_funk.cpp:65-70)
        const double hhh = d * (xx - xx2) + h2;
        yz[1] = yz[1] * hhh / hh;
        yz[2] = yz[2] * hhh / hh;
        for (int k = N4 - 3; k <= NN; k++) {
            double *q = Q(k);
            const double hq = ryz0(q);
            q[1] = q[1] * h1 / hq;
            q[2] = q[2] * h1 / hq;
        }
        return;
    }
    const int a = t - F.n_tail;
    if (a >= F.n_bs || !F.move[3]) return;
    const double H = ryz0(X + 3 * (int64_t)F.bs_ref[a]);
    for (int j = F.bs_off[a]; j < F.bs_off[a + 1]; j++) {
        double *q = X + 3 * (int64_t)F.bs_nodes[j];
        const double h2 = ryz0(q);
        q[1] *= H / h2;
        q[2] *= H / h2;
    }
}

// func_R (Yzel.cpp:27-30) of each ray's TS, HP and BS node: the radii Luch::dvigenie reads.
__global__ void __launch_bounds__(256) k_fit_ray_radii(FitDev F, const double *__restrict__ X) {
    const int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= F.n_rays) return;
    const int iTS = F.M0 + F.M1 + 1, iHP = iTS + F.M2 + 1, iBS = iHP + F.M3 + 1;
    const int32_t *rn = F.ray_nodes + ray * F.npr;
    const int idx[3] = {iTS, iHP, iBS};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double *p = X + 3 * (int64_t)rn[idx[k]];
        F.ray_R[3 * ray + k] = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    }
}

// Luch::dvigenie, A_Luch / A2_Luch (Luch.cpp:8-151), one thread per device node (coalesced
// writes).  Node j of a ray: the inner segment up to R1 (j < M0), then to the TS, a uniform run of
// M11 past it, the run to the HP, to the BS and to R5.  The ray's node M0 (at R1) and its TS, HP
// and BS nodes are not written, nor are nodes on no ray.
__global__ void __launch_bounds__(256) k_fit_rays(FitDev F, double *__restrict__ X) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= F.n_nodes) return;
    const int2 nr = F.node_ray[g];
    if (nr.x < 0) return;
    const int64_t ray = nr.x;
    const int j = nr.y;
    const int iTS = F.M0 + F.M1 + 1, iHP = iTS + F.M2 + 1, iBS = iHP + F.M3 + 1;
    if (j == F.M0 || j == iTS || j == iHP || j == iBS) return;
    const double R2 = F.ray_R[3 * ray], R3 = F.ray_R[3 * ray + 1], R4 = F.ray_R[3 * ray + 2];
    double r;
    if (j < F.M0) {
        r = F.R0 + j * (F.R1 - F.R0) / F.M0;
    } else if (j < iTS) {
        r = F.R1 + (j - F.M0) * (R2 - F.R1) / (F.M1 + 1);
    } else if (j <= iTS + F.M11) {
        const double dr = (R2 - F.R1) / (F.M1 + 1);
        r = R2 + dr * (j - iTS);
    } else if (j < iHP) {
        const double dr = (R2 - F.R1) / (F.M1 + 1);
        const double a = F.da1, b = F.da2;
        const double s = (j - iTS - F.M11 + 0.0) / (F.M2 - F.M11 + 1.0);
        const double u = 1.0 - 3.0 * s + 2.0 * (s * s);
        const double w = 1.0 - 2.0 * s;
        const double ss = s * (a * (u * u) +
                               s * (7.0 + b * (w * w) * (-1.0 + s) - 2.0 * s * (9.0 + 2.0 * s * (-5.0 + 2.0 * s))));
        r = R2 + dr * F.M11 + (R3 - R2 - dr * F.M11) * ss;
    } else if (j < iBS) {
        const double a = F.da3, b = F.da4;
        const double s = (j - iHP + 0.0) / (F.M3 + 1.0);
        const double ss = s * (a - (2.0 * a + b - 3.0) * s + (a + b - 2.0) * s * s);
        r = R3 + (R4 - R3) * ss;
    } else if (j < iBS + 1 + F.M4) {
        const double a = F.da5, b = 2.0;
        const double s = (j - iBS + 0.0) / (F.M4 + 1.0);
        const double ss = a * s - (3.0 * a + b - 4.0) * s * s * s + (2.0 * a + b - 3.0) * s * s * s * s;
        r = R4 + (F.R5 - R4) * ss;
    } else {
        r = F.R5;
    }
    const double *tr = F.ray_trig + 4 * ray;
    X[3 * g] = r * tr[0];
    X[3 * g + 1] = r * tr[1] * tr[2];
    X[3 * g + 2] = r * tr[1] * tr[3];
}

namespace {
// the cubic from (x0, y0) with slope (t1, tt1) to (x1, y1) with slope (t2, tt2), at u
__device__ inline void hermite(double x0, double y0, double x1, double y1, double t1, double tt1, double t2,
                               double tt2, double u, double &x, double &y) {
    const double a1 = x0, b1 = t1, c1 = -2 * t1 - t2 - 3.0 * x0 + 3.0 * x1, d1 = t1 + t2 + 2.0 * x0 - 2.0 * x1;
    const double a2 = y0, b2 = tt1, c2 = -2 * tt1 - tt2 - 3.0 * y0 + 3.0 * y1, d2 = tt1 + tt2 + 2.0 * y0 - 2.0 * y1;
    x = a1 + b1 * u + c1 * u * u + d1 * u * u * u;
    y = a2 + b2 * u + c2 * u * u + d2 * u * u * u;
}
// spacing with end slopes a, b: s (a - (2a + b - 3) s + (a + b - 2) s^2)
__device__ inline double spacing1(double s, double a, double b) {
    return s * (a - (2.0 * a + b - 3.0) * s + (a + b - 2.0) * s * s);
}
// spacing with slope a at 0, b at 1 and no curvature at 0: a s - (3a + b - 4) s^3 + (2a + b - 3) s^4
__device__ inline double spacing2(double s, double a, double b) {
    return a * s - (3.0 * a + b - 4.0) * s * s * s + (2.0 * a + b - 3.0) * s * s * s * s;
}
__device__ inline double ryz(const double *p) { return sqrt(0.0 * 0.0 + p[1] * p[1] + p[2] * p[2]); }

// One ray of family B..G, in the reference's operation order.  pol(): a node at polar radius r
// (x = r cos the, y = r sin the cos phi, z = r sin the sin phi); cyl(): at x, distance r from the axis.
__device__ void move_ext_ray(const FitDev &F, int64_t k, double *__restrict__ X) {
    const RayExtDev &E = F.ext;
    const int32_t *Y = E.nodes + E.off[k];
    const int32_t *O = E.opor + 4 * k;
    const double ct = E.trig[4 * k], st = E.trig[4 * k + 1], cp = E.trig[4 * k + 2], sp = E.trig[4 * k + 3];
    auto P = [&](int i) { return X + 3 * (int64_t)Y[i]; };
    auto Q = [&](int i) { return X + 3 * (int64_t)O[i]; };
    auto pol = [&](int i, double r) {
        double *p = P(i);
        p[0] = r * ct;
        p[1] = r * st * cp;
        p[2] = r * st * sp;
    };
    auto cyl = [&](int i, double x, double r) {
        double *p = P(i);
        p[0] = x;
        p[1] = r * cp;
        p[2] = r * sp;
    };
    const int M0 = F.M0, M1 = F.M1, M11 = F.M11, M2 = F.M2, M3 = F.M3, M4 = F.M4;
    const double R0 = F.R0, R1 = F.R1, R5 = F.R5;
    const int t = E.type[k];
    int num = 0;
    double r = 0.0, xx, yy;
    if (t == 1 || t == 2) {  // B (Luch.cpp:254-421), C / C2 (:153-252): the inner, common part
        const double *o1 = Q(1);
        const double R2 = sqrt(o1[0] * o1[0] + o1[1] * o1[1] + o1[2] * o1[2]);
        double R3 = 0.0, R4 = 0.0;
        if (t == 1) {
            R3 = sqrt(Q(2)[1] * Q(2)[1] + Q(2)[2] * Q(2)[2]);
            R4 = sqrt(Q(3)[1] * Q(3)[1] + Q(3)[2] * Q(3)[2]);
        }
        for (int j = 0; j < M0; j++) {
            r = R0 + j * (R1 - R0) / M0;
            pol(j, r);
        }
        num = M0;
        if (t == 1) pol(num, R1);
        num += 1;
        for (int j = 0; j < M1; j++) {
            r = R1 + (j + 1) * (R2 - R1) / (M1 + 1);
            pol(num + j, r);
        }
        num += M1 + 1;
        const double dr = (R2 - R1) / (M1 + 1);
        for (int j = 0; j < M11; j++) {
            r = R2 + dr * (j + 1);
            pol(num + j, r);
        }
        num += M11;
        const double x0 = r * ct, y0 = r * st;
        if (t == 1) {
            const double x3 = r * ct;
            Q(2)[0] = x3;
            Q(3)[0] = x3;
            const double t1 = ct * E.dd3, tt1 = st * E.dd3, t2 = 0.0, tt2 = 1.0 * E.dd4;
            for (int j = 0; j < M2 - M11; j++) {
                const double s = 1.0 * (j + 1) / (M2 - M11 + 1);
                hermite(x0, y0, x3, R3, t1, tt1, t2, tt2, spacing1(s, E.ba1, E.ba2), xx, yy);
                cyl(num + j, xx, yy);
            }
            num += (M2 - M11) + 1;
            for (int j = 0; j < M3; j++) {
                const double s = (j + 1.0) / (M3 + 1.0);
                r = R3 + spacing1(s, E.ba3, E.ba4) * (R4 - R3);
                cyl(num + j, x3, r);
            }
            num += M3 + 1;
            for (int j = 0; j < M4; j++) {
                const double s = (j + 1.0) / (M4 + 1.0);
                r = R4 + (R5 - R4) * spacing2(s, E.ba5, 2.0);
                cyl(num + j, x3, r);
            }
            num += M4;
            cyl(num, x3, R5);
        } else {
            const double t1 = (ct - 1.0) / 2.0 * E.dd5i, tt1 = st * E.dd5i, t2 = -1.0 * E.dd6i, tt2 = 0.0;
            double y = 0.0, z = 0.0;
            for (int j = 0; j < E.N4; j++) {
                const double s = 1.0 * (j + 1) / (E.N4 + 1);
                hermite(x0, y0, E.L6, y0, t1, tt1, t2, tt2, s, xx, yy);
                z = yy * sp;
                y = yy * cp;
                double *p = P(num + j);
                p[0] = xx;
                p[1] = y;
                p[2] = z;
            }
            num += E.N4;
            for (int j = 0; j < E.N5; j++) {
                double *p = P(num + j);
                p[0] = E.L6 + (j) * (E.L7 - E.L6) / (E.N5 - 1);
                p[1] = y;
                p[2] = z;
            }
        }
    } else if (t == 3) {  // D (Luch.cpp:423-523)
        const double x0 = Q(0)[0], y0 = sqrt(Q(0)[1] * Q(0)[1] + Q(0)[2] * Q(0)[2]);
        Q(1)[0] = x0;
        const double y1 = sqrt(Q(1)[1] * Q(1)[1] + Q(1)[2] * Q(1)[2]);
        Q(2)[0] = x0;
        const double y2 = sqrt(Q(2)[1] * Q(2)[1] + Q(2)[2] * Q(2)[2]);
        num = 1;
        for (int j = 0; j < M2 - M11; j++) {
            const double s = 1.0 * (j + 1) / (M2 - M11 + 1);
            cyl(num + j, x0, y0 + (y1 - y0) * spacing1(s, 0.1, 0.3));
        }
        num += (M2 - M11) + 1;
        for (int j = 0; j < M3; j++) {
            const double s = 1.0 * (j + 1) / (M3 + 1);
            cyl(num + j, x0, y1 + (y2 - y1) * spacing1(s, E.md1, E.md2));
        }
        num += M3 + 1;
        for (int j = 0; j < M4; j++) {
            const double s = (j + 1.0) / (M4 + 1.0);
            cyl(num + j, x0, y2 + (R5 - y2) * spacing2(s, E.md3, 2.0));
        }
        num += M4;
        cyl(num, x0, R5);
    } else if (t == 4) {  // E (Luch.cpp:525-602)
        const double y0 = sqrt(Q(1)[1] * Q(1)[1] + Q(1)[2] * Q(1)[2]);
        const double y1 = sqrt(Q(2)[1] * Q(2)[1] + Q(2)[2] * Q(2)[2]);
        const double xe = Q(0)[0];
        P(0)[0] = xe;
        num = 1;
        for (int j = 0; j < M3; j++) {
            const double s = 1.0 * (j + 1) / (M3 + 1);
            cyl(num + j, xe, y0 + (y1 - y0) * spacing1(s, E.ea1, E.ea2));
        }
        num += M3;
        P(num)[0] = xe;
        num += 1;
        for (int j = 0; j < M4; j++) {
            const double s = (j + 1.0) / (M4 + 1.0);
            cyl(num + j, xe, y1 + (R5 - y1) * spacing2(s, E.ea3, 2.0));
        }
        num += M4;
        cyl(num, xe, R5);
    } else if (t == 5) {  // H (Luch.cpp:604-661)
        const double *z3 = Q(0), *z1 = Q(1), *z4 = Q(2), *z2 = Q(3);
        const double t1 = (z1[0] - z3[0]) * E.dd7, tt1 = (ryz(z1) - ryz(z3)) * E.dd7;
        const double t2 = (z4[0] - z2[0]) * E.dd8, tt2 = (ryz(z4) - ryz(z2)) * E.dd8;
        const double x0 = z1[0], y0 = ryz(z1), x1 = z2[0], y1 = ryz(z2);
        for (int j = 0; j < 4; j++) {
            const double s = 1.0 * (j + 1) / (5);
            hermite(x0, y0, x1, y1, t1, tt1, t2, tt2, s, xx, yy);
            cyl(1 + j, xx, yy);
        }
    } else if (t == 6) {  // G (Luch.cpp:663-736)
        const double *z1 = Q(1), *z2 = Q(2), *z3 = Q(0);
        const double t1 = (z1[0] - z3[0]) * E.dd1, tt1 = (ryz(z1) - ryz(z3)) * E.dd1;
        const double t2 = 0.0, tt2 = 1.0 * E.dd2;
        const double x0 = z1[0], y0 = ryz(z1), x1 = z2[0], y1 = ryz(z2);
        const int n = M2 - M11 - E.MF + 1;
        for (int j = 0; j < n; j++) {
            const double s = 1.0 * (j + 1.0) / (n + 1.0);
            hermite(x0, y0, x1, y1, t1, tt1, t2, tt2, s, xx, yy);
            cyl(1 + j, xx, yy);
        }
    }
}
}  // namespace

// The rays of family `fam`: one thread per ray, or one thread taking them in order when rays of the
// family touch nodes another of them writes.
__global__ void __launch_bounds__(64) k_fit_rays_ext(FitDev F, double *__restrict__ X, int fam) {
    const int64_t b = F.ext.fam_off[fam], e = F.ext.fam_off[fam + 1];
    if ((F.ext.serial >> fam) & 1) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (int64_t i = b; i < e; i++) move_ext_ray(F, F.ext.order[i], X);
        return;
    }
    const int64_t i = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < e) move_ext_ray(F, F.ext.order[i], X);
}

void launch_fit_surface(hipStream_t s, const DevMesh &M, const DevParams &P, const FitDev &F, const double *prim,
                        const double *X1, double *X2, const double *cell_center, StepScalars *S, double time) {
    if (F.n_sf > 0)
        hipLaunchKernelGGL(k_fit_faces, dim3(grid_of(F.n_sf, 256)), dim3(256), 0, s, M, P, F, prim, X1, cell_center,
                           S);
    if (F.n_sn > 0) hipLaunchKernelGGL(k_fit_nodes, dim3(grid_of(F.n_sn, 256)), dim3(256), 0, s, F, X1, X2, S, time);
    if (F.n_tail + F.n_bs > 0)
        hipLaunchKernelGGL(k_fit_fix, dim3(grid_of(F.n_tail + F.n_bs, 64)), dim3(64), 0, s, F, X2);
}

void launch_fit_rays(hipStream_t s, const FitDev &F, double *X) {
    if (F.n_rays <= 0) return;
    hipLaunchKernelGGL(k_fit_ray_radii, dim3(grid_of(F.n_rays, 256)), dim3(256), 0, s, F, X);
    hipLaunchKernelGGL(k_fit_rays, dim3(grid_of(F.n_nodes, 256)), dim3(256), 0, s, F, X);
    // B, C, D, E, H, G after the A rays (Setka.cpp:1850-1880)
    for (int fam = 1; fam <= 6; fam++) {
        const int64_t cnt = F.ext.fam_off[fam + 1] - F.ext.fam_off[fam];
        if (cnt <= 0) continue;
        const unsigned g = ((F.ext.serial >> fam) & 1) ? 1u : grid_of(cnt, 64);
        hipLaunchKernelGGL(k_fit_rays_ext, dim3(g), dim3(64), 0, s, F, X, fam);
    }
}

}  // namespace titan
