// step_kernels.h -- HIP kernels of one explicit step (included by k_step64.hip and k_step32.hip).
//
// Kernel map (reference loop -> kernel):
//   Setka::Calculating_measure (Setka.cpp:2465-2502)  -> k_cell_center, k_face_geom, k_cell_volume
//   Culc_Velocity_surface + Luch::dvigenie (prescribed) -> k_motion
//   face loop, Setka_physics.cpp:1634-1662           -> k_face_flux<T> (+ wave-shuffle dt min)
//   cell loop, Setka_physics.cpp:1671-2027           -> cell_update_state, scalar_update (called by
//                                                       the fused tile kernel, tile_kernel.h)
#pragma once
#include <hip/hip_runtime.h>
#include "device_mesh.h"
#include "riemann.h"
#include "riemann_mixed.h"
#include "tvd.h"
#include "fdiv.h"
#include "fsqrt.h"

namespace titan {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double det_cos(double x) {
    // Cody-Waite reduction by pi/2 (3 parts) + Taylor polynomials on [-pi/4, pi/4]; the same
    // operations as the host's prescribed-motion law so both move nodes identically.
    const double two_over_pi = 0.63661977236758134308;
    const double p1 = 1.5707963267341256e+00;
    const double p2 = 6.0771005065061922e-11;
    const double p3 = 2.0222662487959506e-21;
    double kf = rint(x * two_over_pi);
    double r = ((x - kf * p1) - kf * p2) - kf * p3;
    long k = (long)kf;
    double r2 = r * r;
    double s = r * (1.0 + r2 * (-1.0 / 6.0 + r2 * (1.0 / 120.0 + r2 * (-1.0 / 5040.0 + r2 * (1.0 / 362880.0 +
               r2 * (-1.0 / 39916800.0 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0))))))));
    double c = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24.0 + r2 * (-1.0 / 720.0 + r2 * (1.0 / 40320.0 + r2 * (-1.0 / 3628800.0 +
               r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0 + r2 * (1.0 / 20922789888000.0))))))));
    switch (((k % 4) + 4) % 4) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
    }
}

__device__ __forceinline__ double tri_area(const double *a, const double *b, const double *c) {
    // triangleArea3D, Help_funk.cpp:193-217
    double v1x = b[0] - a[0], v1y = b[1] - a[1], v1z = b[2] - a[2];
    double v2x = c[0] - a[0], v2y = c[1] - a[1], v2z = c[2] - a[2];
    double cx = v1y * v2z - v1z * v2y;
    double cy = v1z * v2x - v1x * v2z;
    double cz = v1x * v2y - v1y * v2x;
    return 0.5 * sqrt_n(cx * cx + cy * cy + cz * cz);
}

__device__ __forceinline__ double tet_volume(const double *A, const double *B, const double *C, const double *D) {
    // tetrahedronVolume, Help_funk.cpp:253-272
    double abx = B[0] - A[0], aby = B[1] - A[1], abz = B[2] - A[2];
    double acx = C[0] - A[0], acy = C[1] - A[1], acz = C[2] - A[2];
    double adx = D[0] - A[0], ady = D[1] - A[1], adz = D[2] - A[2];
    double cx = aby * acz - abz * acy;
    double cy = abz * acx - abx * acz;
    double cz = abx * acy - aby * acx;
    double d = cx * adx + cy * ady + cz * adz;
    return div_by(fabs(d), 6.0, 1.0 / 6.0);
}

// Centre, unit normal (away from cells[0], whose centre cc holds) and area of face f from the
// nodes X: Gran::Culc_measure (Gran.cpp:112-218), or the exactly conservative variant (gcl != 0).
__device__ __forceinline__ void face_geom_one(const DevMesh &M, const double *__restrict__ X,
                                              const double *__restrict__ cell_center, int64_t f, int gcl,
                                              double c[3], double n[3], double &A) {
    double x[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double *p = X + 3 * (int64_t)M.face_nodes[4 * f + i];
        x[i][0] = p[0]; x[i][1] = p[1]; x[i][2] = p[2];
    }
    double xc = 0.0, yc = 0.0, zc = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) { xc += x[i][0]; yc += x[i][1]; zc += x[i][2]; }
    xc /= 4; yc /= 4; zc /= 4;
    double n1, n2, n3;
    if (!gcl) {
        double S = 0.0;
        S += tri_area(x[0], x[1], x[2]);
        S += tri_area(x[0], x[3], x[2]);
        A = S;
        double x1 = (x[0][0] + x[1][0]) / 2.0, y1 = (x[0][1] + x[1][1]) / 2.0, z1 = (x[0][2] + x[1][2]) / 2.0;
        double x2 = (x[1][0] + x[2][0]) / 2.0, y2 = (x[1][1] + x[2][1]) / 2.0, z2 = (x[1][2] + x[2][2]) / 2.0;
        double x3 = (x[2][0] + x[3][0]) / 2.0, y3 = (x[2][1] + x[3][1]) / 2.0, z3 = (x[2][2] + x[3][2]) / 2.0;
        double ax = x2 - x1, ay = y2 - y1, az = z2 - z1;
        double bx = x3 - x1, by = y3 - y1, bz = z3 - z1;
        n1 = ay * bz - az * by;
        n2 = az * bx - ax * bz;
        n3 = ax * by - ay * bx;
        double nn = sqrt(n1 * n1 + n2 * n2 + n3 * n3);
        n1 /= nn; n2 /= nn; n3 /= nn;
    } else {
        double ax = x[2][0] - x[0][0], ay = x[2][1] - x[0][1], az = x[2][2] - x[0][2];
        double bx = x[3][0] - x[1][0], by = x[3][1] - x[1][1], bz = x[3][2] - x[1][2];
        double s1 = 0.5 * (ay * bz - az * by);
        double s2 = 0.5 * (az * bx - ax * bz);
        double s3 = 0.5 * (ax * by - ay * bx);
        A = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
        n1 = s1 / A; n2 = s2 / A; n3 = s3 / A;
    }
    const double *cc = cell_center + 3 * (int64_t)M.face_cells[2 * f];
    if (spf(n1, n2, n3, cc[0] - xc, cc[1] - yc, cc[2] - zc) > 0.0) { n1 = -n1; n2 = -n2; n3 = -n3; }
    c[0] = xc; c[1] = yc; c[2] = zc;
    n[0] = n1; n[1] = n2; n[2] = n3;
}

#ifndef TITAN_FLUX_ONLY
// ---- step bookkeeping (one thread) --------------------------------------------------------
// step n (1-based, counted over the context's life): dt = 1e-5 for n == 1 and for the first step
// of a call that restarts the schedule as each Setka::Go call does (restart), else the min over
// interior faces of step n-1 (initial value 1e8, Setka_physics.cpp:1582,1605-1607).
__global__ void k_begin_step(StepScalars *S, long long n, DevParams P, int restart) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double dt_prev = S->dt;
    if (n == 1) {
        S->dt = 0.00001;
        S->t = S->t;  // caller initialises t
    } else {
        S->t = S->t + dt_prev;
        S->dt = restart ? 0.00001 : __longlong_as_double((long long)S->dtmin_bits[(n - 1) & 1]);
    }
    S->dtmin_bits[n & 1] = (unsigned long long)__double_as_longlong(100000000.0);
    S->motion_fac = S->dt * (P.motion_omega * det_cos(P.motion_omega * S->t + P.motion_phase));
    S->n_rho_floor = 0;
    S->n_p_floor = 0;
    S->n_fallback = 0;
    S->n_fp64_faces = 0;
}

// ---- geometry -------------------------------------------------------------------------------
__global__ void k_motion(int64_t n3, const double *__restrict__ x_old, const double *__restrict__ disp,
                         const StepScalars *__restrict__ S, double *__restrict__ x_new) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const double fac = S->motion_fac;
    x_new[i] = x_old[i] + fac * disp[i];
}

__global__ void k_cell_center(DevMesh M, const double *__restrict__ X, DevGeom G) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M.n_cells) return;
    double xc = 0.0, yc = 0.0, zc = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const double *x = X + 3 * (int64_t)M.cell_nodes[8 * c + i];
        xc += x[0]; yc += x[1]; zc += x[2];
    }
    G.cell_center[3 * c] = xc / 8;
    G.cell_center[3 * c + 1] = yc / 8;
    G.cell_center[3 * c + 2] = zc / 8;
}


// Gran::Culc_measure (Gran.cpp:112-218), or the exactly conservative variant (gcl != 0).
__global__ void k_face_geom(DevMesh M, const double *__restrict__ X, DevGeom G, int gcl) {
    int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= M.n_faces) return;
    double c[3], n[3], A;
    face_geom_one(M, X, G.cell_center, f, gcl, c, n, A);
    G.face_center[3 * f] = c[0];
    G.face_center[3 * f + 1] = c[1];
    G.face_center[3 * f + 2] = c[2];
    G.face_normal[3 * f] = n[0];
    G.face_normal[3 * f + 1] = n[1];
    G.face_normal[3 * f + 2] = n[2];
    G.face_area[f] = A;
}


// Cell::Culc_volume method 1 (Cell.cpp:958-1000), or 1/3 sum s_f xbar.n A (gcl).
__global__ void k_cell_volume(DevMesh M, const double *__restrict__ X, DevGeom G, int gcl) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M.n_cells) return;
    const double cc[3] = {G.cell_center[3 * c], G.cell_center[3 * c + 1], G.cell_center[3 * c + 2]};
    const uint32_t code = M.cell_code[c];
    double V = 0.0;
    for (int j = 0; j < 6; j++) {
        const int64_t f = M.cell_faces[6 * c + j];
        if (!gcl) {
            const double *x0 = X + 3 * (int64_t)M.face_nodes[4 * f + 0];
            const double *x1 = X + 3 * (int64_t)M.face_nodes[4 * f + 1];
            const double *x2 = X + 3 * (int64_t)M.face_nodes[4 * f + 2];
            const double *x3 = X + 3 * (int64_t)M.face_nodes[4 * f + 3];
            V += tet_volume(cc, x0, x1, x2);
            V += tet_volume(cc, x0, x3, x2);
        } else {
            const double s = ((code >> j) & 1u) ? 1.0 : -1.0;
            const double *n = G.face_normal + 3 * f;
            const double *fc = G.face_center + 3 * f;
            V += s * (spf(fc[0], fc[1], fc[2], n[0], n[1], n[2]) * G.face_area[f]);
        }
    }
    G.cell_volume[c] = gcl ? V / 3.0 : V;
}

#endif  // TITAN_FLUX_ONLY

// ---- face velocity ----------------------------------------------------------------------------
// Gran::culc_velosity (Gran.cpp:571-589)
__device__ __forceinline__ double face_velocity_ref(const DevMesh &M, int64_t f, const double *X1,
                                                    const double *X2, const double n[3], double dt) {
    double ddot = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int64_t v = M.face_nodes[4 * f + i];
        const double d0 = X2[3 * v] - X1[3 * v];
        const double d1 = X2[3 * v + 1] - X1[3 * v + 1];
        const double d2 = X2[3 * v + 2] - X1[3 * v + 2];
        ddot = ddot + ((d0 * n[0] + d1 * n[1]) + d2 * n[2]);
    }
    return ddot / 4 / dt;
}

// Exactly conservative face velocity: swept volume of the face (hexahedron old face -> new face,
// bilinear sides, volume 1/3 sum xbar.S) / (dt * area_old).
__device__ __forceinline__ double face_velocity_gcl(const DevMesh &M, int64_t f, const double *X1,
                                                    const double *X2, const double n[3], double area,
                                                    double dt) {
    double P[8][3];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int64_t v = M.face_nodes[4 * f + i];
#pragma unroll
        for (int d = 0; d < 3; d++) { P[i][d] = X1[3 * v + d]; P[4 + i][d] = X2[3 * v + d]; }
    }
    const int lf[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 7, 3}};
    double V = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double *x0 = P[lf[j][0]], *x1 = P[lf[j][1]], *x2 = P[lf[j][2]], *x3 = P[lf[j][3]];
        double ax = x2[0] - x0[0], ay = x2[1] - x0[1], az = x2[2] - x0[2];
        double bx = x3[0] - x1[0], by = x3[1] - x1[1], bz = x3[2] - x1[2];
        double s1 = 0.5 * (ay * bz - az * by);
        double s2 = 0.5 * (az * bx - ax * bz);
        double s3 = 0.5 * (ax * by - ay * bx);
        double xc = (((x0[0] + x1[0]) + x2[0]) + x3[0]) / 4;
        double yc = (((x0[1] + x1[1]) + x2[1]) + x3[1]) / 4;
        double zc = (((x0[2] + x1[2]) + x2[2]) + x3[2]) / 4;
        V += spf(xc, yc, zc, s1, s2, s3);
    }
    V = V / 3.0;
    double ax = P[2][0] - P[0][0], ay = P[2][1] - P[0][1], az = P[2][2] - P[0][2];
    double bx = P[3][0] - P[1][0], by = P[3][1] - P[1][1], bz = P[3][2] - P[1][2];
    double s1 = ay * bz - az * by, s2 = az * bx - ax * bz, s3 = ax * by - ay * bx;
    const double sg = (spf(s1, s2, s3, n[0], n[1], n[2]) >= 0.0) ? 1.0 : -1.0;
    return (sg * V) / (dt * area);
}

// ---- face states: Setka::Snos_on_Gran (Setka_gran_snos.cpp:3-180, 737-1071) -------------------
__device__ __forceinline__ double polar_angle(double x, double y) {
    if (fabs(x) + fabs(y) < 0.000001) return 0.0;
    if (x < 0) return atan(y / x) + 1.0 * kPi;
    else if (x > 0 && y >= 0) return atan(y / x);
    else if (x > 0 && y < 0) return atan(y / x) + 2.0 * kPi;
    else if (y > 0 && x >= 0 && x <= 0) return kPi / 2.0;
    else if (y < 0 && x >= 0 && x <= 0) return 3.0 * kPi / 2.0;
    return 0.0;
}

__device__ __forceinline__ void spherical_skorost(double z, double x, double y, double Vz, double Vx, double Vy,
                                                  double &Vr, double &Vphi, double &Vtheta) {
    const double r_1 = sqrt(x * x + y * y + z * z);
    const double the_1 = acos(z / r_1);
    const double phi_1 = polar_angle(x, y);
    Vr = Vx * sin(the_1) * cos(phi_1) + Vy * sin(the_1) * sin(phi_1) + Vz * cos(the_1);
    Vtheta = Vx * cos(the_1) * cos(phi_1) + Vy * cos(the_1) * sin(phi_1) - Vz * sin(the_1);
    Vphi = -Vx * sin(phi_1) + Vy * cos(phi_1);
}

__device__ __forceinline__ void dekard_skorost(double z, double x, double y, double Vr, double Vphi, double Vtheta,
                                               double &Vz, double &Vx, double &Vy) {
    const double r_2 = sqrt(x * x + y * y + z * z);
    const double the_2 = acos(z / r_2);
    const double phi_2 = polar_angle(x, y);
    Vx = Vr * sin(the_2) * cos(phi_2) + Vtheta * cos(the_2) * cos(phi_2) - Vphi * sin(phi_2);
    Vy = Vr * sin(the_2) * sin(phi_2) + Vtheta * cos(the_2) * sin(phi_2) + Vphi * cos(phi_2);
    Vz = Vr * cos(the_2) - Vtheta * sin(the_2);
}

// Zone_1 spherical rescaling of a cell state to the face centre G.
__device__ __noinline__ void zone1_state(const double s[8], const double Ac[3], const double G[3],
                                         double gamma, double out[8]) {
    const double r1 = sqrt((Ac[0] * Ac[0] + Ac[1] * Ac[1]) + Ac[2] * Ac[2]);
    const double rr = sqrt((G[0] * G[0] + G[1] * G[1]) + G[2] * G[2]);
    const double the1 = acos(G[2] / rr);
    const bool equatorial = (the1 > kPi / 9 && the1 < 8 * kPi / 9);
    for (int ik = 0; ik < 2; ik++) {
        const double *V = s + (ik == 0 ? 1 : 5);
        double *O = out + (ik == 0 ? 1 : 5);
        double VA0, VA1, VA2;
        if (equatorial) spherical_skorost(Ac[2], Ac[0], Ac[1], V[2], V[0], V[1], VA0, VA1, VA2);
        else spherical_skorost(Ac[0], Ac[1], Ac[2], V[0], V[1], V[2], VA0, VA1, VA2);
        if (ik == 1) {
            VA0 = VA0 * ((r1 / rr) * (r1 / rr));
            VA1 = VA1 * ((r1) / (rr));
            VA2 = VA2 * ((r1) / (rr));
        }
        if (equatorial) dekard_skorost(G[2], G[0], G[1], VA0, VA1, VA2, O[2], O[0], O[1]);
        else dekard_skorost(G[0], G[1], G[2], VA0, VA1, VA2, O[0], O[1], O[2]);
    }
    out[0] = s[0] * (r1 * r1) / (rr * rr);
    out[4] = s[4] * pow(r1, 2.0 * gamma) / pow(rr, 2.0 * gamma);
}

__device__ __forceinline__ void load_state(const double *__restrict__ prim, int64_t c, double s[8]) {
    const double4 *p = reinterpret_cast<const double4 *>(prim + 8 * c);
    const double4 a = p[0], b = p[1];
    s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
    s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
}

// Returns false on an unknown boundary code.  cell_center: centres of the layer of prim; Gc: the
// face centre at that layer.  K = kScal: the passive scalars of the same layer (scal [Nc*2]) give
// kL, kR -- the cell values, the boundary state's, or rho's reconstruction (Zone_1 x r1^2/rr^2,
// Setka_gran_snos.cpp:848-852; TVD as rho).
template <int K = 0>
__device__ __forceinline__ bool face_states_at(const DevMesh &M, const DevParams &P, int64_t f, uint32_t bc,
                                               const double *__restrict__ prim,
                                               const double *__restrict__ cell_center, const double Gc[3],
                                               const double n[3], double qL[8], double qR[8],
                                               const double *__restrict__ scal = nullptr, double *kL = nullptr,
                                               double *kR = nullptr) {
    const int64_t A = M.face_cells[2 * f];
    double sA[8];
    load_state(prim, A, sA);
    double kA[K > 0 ? K : 1];
#pragma unroll
    for (int j = 0; j < K; j++) kA[j] = scal[K * A + j];
    if (bc != 0) {
#pragma unroll
        for (int j = 0; j < K; j++) {
            kL[j] = kA[j];
            kR[j] = (bc == 1 || bc == 2) ? M.bscal[K * (int64_t)M.face_bidx[f] + j] : kA[j];
        }
        if (bc == 3) {  // Outer_Soft: plasma_condition 1
#pragma unroll
            for (int i = 0; i < 8; i++) qL[i] = sA[i];
            if (n[0] < -0.95) {
                if (qL[1] > P.v_inf / 7.0) qL[1] = P.v_inf / 5.0;
            } else {
                if (qL[1] * n[0] + qL[2] * n[1] + qL[3] * n[2] < 0.0) {
                    qL[1] = 0.1 * n[0];
                    qL[2] = 0.1 * n[1];
                    qL[3] = 0.1 * n[2];
                }
            }
#pragma unroll
            for (int i = 0; i < 8; i++) qR[i] = qL[i];
        } else if (bc == 1 || bc == 2) {  // hard faces: plasma_condition 2
            const double *bs = M.bstate + 8 * (int64_t)M.face_bidx[f];
#pragma unroll
            for (int i = 0; i < 8; i++) { qL[i] = sA[i]; qR[i] = bs[i]; }
        } else if (bc == 4) {
#pragma unroll
            for (int i = 0; i < 8; i++) { qL[i] = sA[i]; qR[i] = sA[i]; }
        } else {
            return false;
        }
        return true;
    }
    const int64_t B = M.face_cells[2 * f + 1];
    double sB[8];
    load_state(prim, B, sB);
    double kB[K > 0 ? K : 1];
#pragma unroll
    for (int j = 0; j < K; j++) kB[j] = scal[K * B + j];
    const uint32_t zA = (M.cell_code[A] >> 8) & 15u, zB = (M.cell_code[B] >> 8) & 15u;
    if (P.tvd) {
        const int64_t AA = M.face_tvd[2 * f], BB = M.face_tvd[2 * f + 1];
        if (AA >= 0) {  // second-order face (both neighbours present)
            double sAA[8], sBB[8], kAA[K > 0 ? K : 1], kBB[K > 0 ? K : 1];
            load_state(prim, AA, sAA);
            load_state(prim, BB, sBB);
#pragma unroll
            for (int j = 0; j < K; j++) { kAA[j] = scal[K * AA + j]; kBB[j] = scal[K * BB + j]; }
            const double *cc = cell_center;
            const double cA[3] = {cc[3 * A], cc[3 * A + 1], cc[3 * A + 2]};
            const double cB[3] = {cc[3 * B], cc[3 * B + 1], cc[3 * B + 2]};
            const double cAA[3] = {cc[3 * AA], cc[3 * AA + 1], cc[3 * AA + 2]};
            const double cBB[3] = {cc[3 * BB], cc[3 * BB + 1], cc[3 * BB + 2]};
            tvd_states<K>((M.cell_code[AA] >> 8) & 15u, zA, zB, (M.cell_code[BB] >> 8) & 15u, sAA, sA, sB, sBB, cAA,
                          cA, cB, cBB, Gc, P.gamma, qL, qR, kAA, kA, kB, kBB, kL, kR);
            return true;
        }
    }
    if (zA == 1u || zB == 1u) {
        const double rr = sqrt((Gc[0] * Gc[0] + Gc[1] * Gc[1]) + Gc[2] * Gc[2]);
        if (zA == 1u) {
            const double Ac[3] = {cell_center[3 * A], cell_center[3 * A + 1], cell_center[3 * A + 2]};
            zone1_state(sA, Ac, Gc, P.gamma, qL);
            const double r1 = sqrt((Ac[0] * Ac[0] + Ac[1] * Ac[1]) + Ac[2] * Ac[2]);
#pragma unroll
            for (int j = 0; j < K; j++) kL[j] = kA[j] * (r1 * r1) / (rr * rr);
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) qL[i] = sA[i];
#pragma unroll
            for (int j = 0; j < K; j++) kL[j] = kA[j];
        }
        if (zB == 1u) {
            const double Bc[3] = {cell_center[3 * B], cell_center[3 * B + 1], cell_center[3 * B + 2]};
            zone1_state(sB, Bc, Gc, P.gamma, qR);
            const double r1 = sqrt((Bc[0] * Bc[0] + Bc[1] * Bc[1]) + Bc[2] * Bc[2]);
#pragma unroll
            for (int j = 0; j < K; j++) kR[j] = kB[j] * (r1 * r1) / (rr * rr);
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) qR[i] = sB[i];
#pragma unroll
            for (int j = 0; j < K; j++) kR[j] = kB[j];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) { qL[i] = sA[i]; qR[i] = sB[i]; }
#pragma unroll
        for (int j = 0; j < K; j++) { kL[j] = kA[j]; kR[j] = kB[j]; }
    }
    return true;
}

template <int K = 0>
__device__ __forceinline__ bool face_states(const DevMesh &M, const DevParams &P, int64_t f, uint32_t bc,
                                            const double *__restrict__ prim, const DevGeom &G, const double n[3],
                                            double qL[8], double qR[8], const double *__restrict__ scal = nullptr,
                                            double *kL = nullptr, double *kR = nullptr) {
    const double Gc[3] = {G.face_center[3 * f], G.face_center[3 * f + 1], G.face_center[3 * f + 2]};
    return face_states_at<K>(M, P, f, bc, prim, G.cell_center, Gc, n, qL, qR, scal, kL, kR);
}

__device__ __forceinline__ void report_error(StepScalars *S, ErrRank rank, long long index) {
    atomicMin(&S->err_key, ((unsigned long long)rank << 48) | (unsigned long long)index);
}

// Min over a block of positive doubles: 64-lane wavefront shuffle, then LDS across waves, then
// one atomicMin on the IEEE bit pattern (order-preserving for positive values).
__device__ __forceinline__ void block_min_to(double v, unsigned long long *slot) {
    __shared__ double wmin[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wmin[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 1; i < nw; i++) m = (wmin[i] < m) ? wmin[i] : m;
        atomicMin(slot, (unsigned long long)__double_as_longlong(m));
    }
}

// block_min_to, and the sum over the block of a per-thread count added to *count_slot with one
// atomic per block (none when the sum is 0).
__device__ __forceinline__ void block_min_count_to(double v, unsigned long long cnt, unsigned long long *slot,
                                                   unsigned long long *count_slot) {
    __shared__ double wmin[16];
    __shared__ unsigned long long wcnt[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
        cnt += __shfl_xor(cnt, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { wmin[wid] = v; wcnt[wid] = cnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        unsigned long long c = wcnt[0];
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 1; i < nw; i++) { m = (wmin[i] < m) ? wmin[i] : m; c += wcnt[i]; }
        atomicMin(slot, (unsigned long long)__double_as_longlong(m));
        if (c) atomicAdd(count_slot, c);
    }
}

// ---- face flux: Setka::Culc_Gran_Potok for every face -------------------------------------------
// flux[f*9 + 0..7] = chlld flux * area, flux[f*9 + 8] = PdivB; with K = kScal passive scalars,
// kflux[f*2 + j] = their konvect flux * area.  T = double: parity arithmetic; T = float:
// chlld_mixed (FP32 wave structure, FP64 upwind flux), states, geometry, area scaling in FP64.
template <typename T, int K>
__device__ __forceinline__ double face_flux_one(const DevMesh &M, const DevParams &P, const double *__restrict__ prim,
                                                const double *__restrict__ scal, const DevGeom &G,
                                                const double *__restrict__ X1, const double *__restrict__ X2,
                                                StepScalars *S, double *__restrict__ flux, double *__restrict__ kflux,
                                                uint8_t *__restrict__ branch, double dt, int64_t f) {
    double cand = 10000000.0;
    const uint32_t code = M.face_code[f];
    const uint32_t bc = code & 15u, method = (code >> 4) & 15u, flags = (code >> 8) & 255u;
    const double n[3] = {G.face_normal[3 * f], G.face_normal[3 * f + 1], G.face_normal[3 * f + 2]};
    const double area = G.face_area[f];
    double w = 0.0;
    if (P.move_mesh) {
        w = P.geom_mode ? face_velocity_gcl(M, f, X1, X2, n, area, dt) : face_velocity_ref(M, f, X1, X2, n, dt);
    }
    double qL[8], qR[8], kL[K > 0 ? K : 1], kR[K > 0 ? K : 1];
    const bool okbc = face_states<K>(M, P, f, bc, prim, G, n, qL, qR, scal, kL, kR);
    RiemannOut<double, K> o;
    if (sizeof(T) == sizeof(double)) {
        chlld<double, K>((int)method, P.n_disc, (int)flags, n[0], n[1], n[2], w, qL, qR, kL, kR, o);
    } else {
        chlld_mixed<K>((int)method, P.n_disc, (int)flags, n[0], n[1], n[2], w, qL, qR, kL, kR, o);
        if (o.ok < 0) {
            chlld_fp64_face<K>((int)method, P.n_disc, (int)flags, n[0], n[1], n[2], w, qL, qR, kL, kR, &o);
            atomicAdd(&S->n_fp64_faces, 1ull);
        }
    }
    const bool ok = o.ok;
    if (!ok || !okbc) report_error(S, kErrFaceRiemann, M.face_ref[f]);
    if (bc == 0 && ok) {
        const int64_t A = M.face_cells[2 * f], B = M.face_cells[2 * f + 1];
        const double d0 = G.cell_center[3 * A] - G.cell_center[3 * B];
        const double d1 = G.cell_center[3 * A + 1] - G.cell_center[3 * B + 1];
        const double d2 = G.cell_center[3 * A + 2] - G.cell_center[3 * B + 2];
        const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2) / 2.0;
        const double dnt = P.kfl * dist / (smax(fabs(o.dsl), fabs(o.dsr)) + fabs(w));
        if (dnt < cand) cand = dnt;
    }
    double *fo = flux + 9 * f;
#pragma unroll
    for (int i = 0; i < 8; i++) fo[i] = o.q[i] * area;
    fo[8] = 0.5 * spf(n[0], n[1], n[2], qL[5] + qR[5], qL[6] + qR[6], qL[7] + qR[7]) * area;
#pragma unroll
    for (int j = 0; j < K; j++) kflux[K * f + j] = o.k[j] * area;
    if (branch) branch[f] = (uint8_t)o.branch;
    return cand;
}

template <typename T>
__global__ void __launch_bounds__(256) k_face_flux(DevMesh M, DevParams P, const double *__restrict__ prim,
                                                   const double *__restrict__ scal, DevGeom G,
                                                   const double *__restrict__ X1, const double *__restrict__ X2,
                                                   StepScalars *S, double *__restrict__ flux,
                                                   double *__restrict__ kflux, uint8_t *__restrict__ branch,
                                                   long long step, double dt_override) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double cand = 10000000.0;
    if (f < M.n_faces) {
        const double dt = (dt_override > 0.0) ? dt_override : S->dt;
        if (P.n_scal) cand = face_flux_one<T, kScal>(M, P, prim, scal, G, X1, X2, S, flux, kflux, branch, dt, f);
        else cand = face_flux_one<T, 0>(M, P, prim, scal, G, X1, X2, S, flux, kflux, branch, dt, f);
    }
    block_min_to(cand, &S->dtmin_bits[step & 1]);
}

// ---- cell update: cell loop of Setka::Go (Setka_physics.cpp:1695-2027), FP64 -------------------
// Pf = signed flux sum of the cell (8 fluxes x area, PdivB); src (nullable): this cell's 5 source
// values -- rate, S_m (3), S_E per unit volume and time.
__device__ __forceinline__ void cell_update_state(const DevMesh &M, uint32_t code, int64_t c, const double s[8],
                                                  double Volume, double Volume2, const double Pf[9], const double *src,
                                                  double time, StepScalars *S, double *__restrict__ prim_new) {
    // every quotient below is the reference's, correctly rounded; those by V^{n+1}, rho', 4 pi,
    // gamma - 1 and 8 pi go through the correctly rounded reciprocal (fdiv.h)
    const double cpi4 = 4.0 * kPi, ycpi4 = 1.0 / (4.0 * kPi);
    const double g1 = 5.0 / 3.0 - 1.0, yg1 = 1.0 / (5.0 / 3.0 - 1.0);
    const double c8pi = 25.13274122871834590768, y8pi = 1.0 / 25.13274122871834590768;
    const double rho = s[0], vx = s[1], vy = s[2], vz = s[3], pp = s[4], bx = s[5], by = s[6], bz = s[7];
    double Smx = 0.0, Smy = 0.0, Smz = 0.0, SE = 0.0;
    if (src) { Smx = src[1]; Smy = src[2]; Smz = src[3]; SE = src[4]; }
    const double dsk = spf(vx, vy, vz, bx, by, bz);
    const double yV = 1.0 / Volume2;
    auto perV = [&](double a) { return div_by(a, Volume2, yV); };
    double rho3 = perV(rho * Volume) - perV(time * Pf[0]);
    bool rho_null = false;
    if (rho3 < 1e-7) { rho_null = true; rho3 = 0.1; atomicAdd(&S->n_rho_floor, 1ull); }
    double u3, v3, w3, bx3, by3, bz3, p3;
    if (!rho_null) {
        const double yr = 1.0 / rho3;
        u3 = div_by(perV(rho * vx * Volume) - perV(time * (Pf[1] + div_by(bx, cpi4, ycpi4) * Pf[8])) + time * Smx, rho3, yr);
        v3 = div_by(perV(rho * vy * Volume) - perV(time * (Pf[2] + div_by(by, cpi4, ycpi4) * Pf[8])) + time * Smy, rho3, yr);
        w3 = div_by(perV(rho * vz * Volume) - perV(time * (Pf[3] + div_by(bz, cpi4, ycpi4) * Pf[8])) + time * Smz, rho3, yr);
        bx3 = perV(bx * Volume) - perV(time * (Pf[5] + vx * Pf[8]));
        by3 = perV(by * Volume) - perV(time * (Pf[6] + vy * Pf[8]));
        bz3 = perV(bz * Volume) - perV(time * (Pf[7] + vz * Pf[8]));
        p3 = ((perV((div_by(pp, g1, yg1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) +
                     div_by(bx * bx + by * by + bz * bz, c8pi, y8pi)) * Volume) -
               perV(time * (Pf[4] + div_by(dsk, cpi4, ycpi4) * Pf[8])) + time * SE) -
              0.5 * rho3 * (u3 * u3 + v3 * v3 + w3 * w3) - div_by(bx3 * bx3 + by3 * by3 + bz3 * bz3, c8pi, y8pi)) *
             g1;
        if (((code >> 8) & 15u) == 2u) {  // Zone_2 magnetic decay, tau = 1
            const double tau = 1.0;
            const double bx4 = bx3 - time * bx3 / tau;
            const double by4 = by3 - time * by3 / tau;
            const double bz4 = bz3 - time * bz3 / tau;
            const double p4 = p3 + (time / tau - (time / tau) * (time / tau)) * (g1) * (bx3 * bx3 + by3 * by3 + bz3 * bz3) /
                                       (4.0 * kPi);
            bx3 = bx4; by3 = by4; bz3 = bz4; p3 = p4;
        }
    } else {
        u3 = vx; v3 = vy; w3 = vz; bx3 = bx; by3 = by; bz3 = bz;
        p3 = rho3 / 2.0;
    }
    if (isnan(rho3)) report_error(S, kErrCellNonfinite, M.cell_ref[c]);
    if (p3 < 0.00000001) { p3 = 0.000001; atomicAdd(&S->n_p_floor, 1ull); }
    if (!((code >> 6) & 1u)) { bx3 = bx; by3 = by; bz3 = bz; }
    double4 *o = reinterpret_cast<double4 *>(prim_new + 8 * c);
    o[0] = make_double4(rho3, u3, v3, w3);
    o[1] = make_double4(p3, bx3, by3, bz3);
}

__device__ __forceinline__ void cell_update_one(const DevMesh &M, uint32_t code, int64_t c,
                                                const double *__restrict__ prim_old, double Volume, double Volume2,
                                                const double Pf[9], const double *src, double time,
                                                StepScalars *S, double *__restrict__ prim_new) {
    double s[8];
    load_state(prim_old, c, s);
    cell_update_state(M, code, c, s, Volume, Volume2, Pf, src, time, S, prim_new);
}

// The passive scalars' part of the cell loop (Setka_physics.cpp:1781-1817, 2004-2016): Q and rho_He
// advected like rho (Q3 = Q V/V' - dt P_Q/V'), rho_He clipped at 0 and, when above the new
// density (after the density floor), set to a tenth of it.  k: the cell's (Q, rho_He); Pk: their
// signed flux sums; rho, Prho: the cell's density and its flux sum, from which the new density is
// formed as the update forms it.  Independent of the sources.
__device__ __forceinline__ void scalar_update(const double k[kScal], double rho, double Prho, const double Pk[kScal],
                                              double Volume, double Volume2, double time, double out[kScal]) {
    const double yV = 1.0 / Volume2;
    auto perV = [&](double a) { return div_by(a, Volume2, yV); };
    double rho3 = perV(rho * Volume) - perV(time * Prho);
    if (rho3 < 1e-7) rho3 = 0.1;
    const double Q3 = perV(k[0] * Volume) - perV(time * Pk[0]);
    double He3 = perV(k[1] * Volume) - perV(time * Pk[1]);
    if (He3 < 0.0) He3 = 0.0;
    if (He3 > rho3) He3 = rho3 * 0.1;
    out[0] = Q3;
    out[1] = He3;
}

}  // namespace titan
