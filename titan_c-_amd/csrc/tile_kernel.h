// tile_kernel.h -- the fused explicit step: one workgroup per tile of TILE cells does node motion,
// face geometry and velocity, face states, Riemann fluxes (kept in LDS), and the cell update with
// the flux sum in the reference's order.  Included by k_step64.hip (RT = double: chlld in FP64,
// parity) and k_step32.hip (RT = float: chlld_mixed, FP32 wave structure around an FP64 upwind
// flux; face states, geometry and the update stay FP64).
//
// Work per step (Setka::Go, Setka_physics.cpp:1589-2173):
//   phase 1  tile nodes -> LDS (layer now1), new positions x + dt*omega*cos(omega t + phase)*D
//            -> LDS (layer now2) and HBM for the nodes the tile owns;
//   phase 2  own cells: centres at both layers (Cell::Culc_center), Zone_1 spherical frames;
//   phase 3  tile faces: Gran::Culc_measure at now1, Gran::culc_velosity, Snos_on_Gran,
//            Phys_param::chlld, dt candidate (Setka_physics.cpp:2428-2438) -> LDS flux slots;
//   phase 4  own cells: volumes at both layers (Cell::Culc_volume), signed flux sum in ascending
//            reference face order, update (Setka_physics.cpp:1695-2027) -> prim at now2;
//   then a wavefront-shuffle min of the dt candidates and one atomic per workgroup.
// Faces on a tile border are solved by both neighbouring tiles (same inputs, same result); no
// flux ever leaves the workgroup and no float atomics are used, so runs repeat bit for bit.
#pragma once
#include "step_kernels.h"
#include "mlp_device.h"
#include "sphere.h"

namespace titan {


// Face geometry at one node layer (Gran::Culc_measure, Gran.cpp:112-218; or the exactly
// conservative variant); orientation from the centre of cells[0].
__device__ __forceinline__ void face_geometry(const double x[4][3], const double ca[3], int gcl, double n[3],
                                              double &A, double G[3]) {
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
        const double x1 = (x[0][0] + x[1][0]) / 2.0, y1 = (x[0][1] + x[1][1]) / 2.0, z1 = (x[0][2] + x[1][2]) / 2.0;
        const double x2 = (x[1][0] + x[2][0]) / 2.0, y2 = (x[1][1] + x[2][1]) / 2.0, z2 = (x[1][2] + x[2][2]) / 2.0;
        const double x3 = (x[2][0] + x[3][0]) / 2.0, y3 = (x[2][1] + x[3][1]) / 2.0, z3 = (x[2][2] + x[3][2]) / 2.0;
        const double ax = x2 - x1, ay = y2 - y1, az = z2 - z1;
        const double bx = x3 - x1, by = y3 - y1, bz = z3 - z1;
        n1 = ay * bz - az * by;
        n2 = az * bx - ax * bz;
        n3 = ax * by - ay * bx;
        const double nn = sqrt_n(n1 * n1 + n2 * n2 + n3 * n3), yn = 1.0 / nn;
        n1 = div_by(n1, nn, yn); n2 = div_by(n2, nn, yn); n3 = div_by(n3, nn, yn);
    } else {
        const double ax = x[2][0] - x[0][0], ay = x[2][1] - x[0][1], az = x[2][2] - x[0][2];
        const double bx = x[3][0] - x[1][0], by = x[3][1] - x[1][1], bz = x[3][2] - x[1][2];
        const double s1 = 0.5 * (ay * bz - az * by);
        const double s2 = 0.5 * (az * bx - ax * bz);
        const double s3 = 0.5 * (ax * by - ay * bx);
        A = sqrt_n(s1 * s1 + s2 * s2 + s3 * s3);
        const double yA = 1.0 / A;
        n1 = div_by(s1, A, yA); n2 = div_by(s2, A, yA); n3 = div_by(s3, A, yA);
    }
    if (spf(n1, n2, n3, ca[0] - xc, ca[1] - yc, ca[2] - zc) > 0.0) { n1 = -n1; n2 = -n2; n3 = -n3; }
    n[0] = n1; n[1] = n2; n[2] = n3;
    G[0] = xc; G[1] = yc; G[2] = zc;
}

// Swept volume of a moving face / (dt A): the exactly conservative face velocity.
__device__ __forceinline__ double swept_velocity(const double x1[4][3], const double x2[4][3], const double n[3],
                                                 double area, double dt) {
    double P[8][3];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int d = 0; d < 3; d++) { P[i][d] = x1[i][d]; P[4 + i][d] = x2[i][d]; }
    const int lf[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 7, 3}};
    double V = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double *a0 = P[lf[j][0]], *a1 = P[lf[j][1]], *a2 = P[lf[j][2]], *a3 = P[lf[j][3]];
        const double ax = a2[0] - a0[0], ay = a2[1] - a0[1], az = a2[2] - a0[2];
        const double bx = a3[0] - a1[0], by = a3[1] - a1[1], bz = a3[2] - a1[2];
        const double s1 = 0.5 * (ay * bz - az * by);
        const double s2 = 0.5 * (az * bx - ax * bz);
        const double s3 = 0.5 * (ax * by - ay * bx);
        const double xc = (((a0[0] + a1[0]) + a2[0]) + a3[0]) / 4;
        const double yc = (((a0[1] + a1[1]) + a2[1]) + a3[1]) / 4;
        const double zc = (((a0[2] + a1[2]) + a2[2]) + a3[2]) / 4;
        V += spf(xc, yc, zc, s1, s2, s3);
    }
    V = V / 3.0;
    const double ax = P[2][0] - P[0][0], ay = P[2][1] - P[0][1], az = P[2][2] - P[0][2];
    const double bx = P[3][0] - P[1][0], by = P[3][1] - P[1][1], bz = P[3][2] - P[1][2];
    const double s1 = ay * bz - az * by, s2 = az * bx - ax * bz, s3 = ax * by - ay * bx;
    const double sg = (spf(s1, s2, s3, n[0], n[1], n[2]) >= 0.0) ? 1.0 : -1.0;
    return (sg * V) / (dt * area);
}

struct TileLds {
    double *x1, *x2;   // [max_nodes*3]
    double *c1, *c2;   // [T*3]
    double *z1;        // [T*10] Zone_1 cell data (r1, r1^2g, trig e, trig p)
    double *flux;      // [max_faces*9]
    double *kflux;     // [max_faces*kScal] passive scalar fluxes (only when carried)
    uint16_t *fnodes;  // [max_faces*4] local nodes of each face slot
};

__host__ __device__ inline size_t tile_flux_doubles(int max_faces) { return 9 * (size_t)max_faces; }

// LDS carve-up; the second node layer and second centres exist only on a moving mesh, the passive
// scalar fluxes only when the step carries them.
__device__ __forceinline__ TileLds tile_lds(double *base, const TileMesh &TM, bool moving, bool scal) {
    TileLds L;
    double *p = base;
    L.x1 = p; p += 3 * TM.max_nodes;
    L.c1 = p; p += 3 * TM.tile_cells;
    L.z1 = p; p += 10 * TM.tile_cells;
    p += (p - base) & 1;  // 16-byte alignment of the flux slots (they also hold 16-byte vectors)
    L.flux = p; p += tile_flux_doubles(TM.max_faces);
    L.kflux = p;
    if (scal) p += kScal * (size_t)TM.max_faces;
    if (moving) {
        L.x2 = p; p += 3 * TM.max_nodes;
        L.c2 = p; p += 3 * TM.tile_cells;
    } else {
        L.x2 = L.x1;
        L.c2 = L.c1;
    }
    L.fnodes = reinterpret_cast<uint16_t *>(p);
    return L;
}

inline size_t tile_lds_bytes(int max_nodes, int tile_cells, int max_faces, bool moving, bool scal) {
    const size_t nd = 3 * (size_t)max_nodes + 13 * (size_t)tile_cells + 1 + tile_flux_doubles(max_faces) +
                      (scal ? kScal * (size_t)max_faces : 0) +
                      (moving ? 3 * (size_t)max_nodes + 3 * (size_t)tile_cells : 0);
    return sizeof(double) * nd + sizeof(uint16_t) * 4 * (size_t)max_faces;
}

template <typename Z>
__device__ __forceinline__ Zone1Cell<Z> load_zone1(const double *z) {
    Zone1Cell<Z> c;
    c.r1 = (Z)z[0]; c.pr1 = (Z)z[1];
    c.e = Trig<Z>{(Z)z[2], (Z)z[3], (Z)z[4], (Z)z[5]};
    c.p = Trig<Z>{(Z)z[6], (Z)z[7], (Z)z[8], (Z)z[9]};
    return c;
}

// The volume terms of one face of a cell (Cell::Culc_volume method 1: the two tetrahedra of the
// face's triangles with the cell centre; GCL: the face's xbar.n A, n out of this cell).  t[0..1]
// (REF) or t[0] (GCL) receive them; the cell's volume is their sum in face order (then / 3 for GCL).
__device__ __forceinline__ void volume_terms(const double *X, const double cc[3], const uint16_t *fnodes,
                                             uint16_t lfj, int gcl, double t[2]) {
    const uint16_t *fn = fnodes + 4 * lfj;
    const double *x0 = X + 3 * fn[0], *x1 = X + 3 * fn[1], *x2 = X + 3 * fn[2], *x3 = X + 3 * fn[3];
    if (!gcl) {
        t[0] = tet_volume(cc, x0, x1, x2);
        t[1] = tet_volume(cc, x0, x3, x2);
    } else {
        const double ax = x2[0] - x0[0], ay = x2[1] - x0[1], az = x2[2] - x0[2];
        const double bx = x3[0] - x1[0], by = x3[1] - x1[1], bz = x3[2] - x1[2];
        const double s1 = 0.5 * (ay * bz - az * by);
        const double s2 = 0.5 * (az * bx - ax * bz);
        const double s3 = 0.5 * (ax * by - ay * bx);
        const double A = sqrt_n(s1 * s1 + s2 * s2 + s3 * s3);
        double n1 = s1 / A, n2 = s2 / A, n3 = s3 / A;
        double xc = 0.0, yc = 0.0, zc = 0.0;
        xc += x0[0]; yc += x0[1]; zc += x0[2];
        xc += x1[0]; yc += x1[1]; zc += x1[2];
        xc += x2[0]; yc += x2[1]; zc += x2[2];
        xc += x3[0]; yc += x3[1]; zc += x3[2];
        xc /= 4; yc /= 4; zc /= 4;
        if (spf(n1, n2, n3, cc[0] - xc, cc[1] - yc, cc[2] - zc) > 0.0) { n1 = -n1; n2 = -n2; n3 = -n3; }
        t[0] = spf(xc, yc, zc, n1, n2, n3) * A;
        t[1] = 0.0;
    }
}

// Volume of an own cell at one layer from LDS nodes (Cell::Culc_volume method 1 or GCL).
__device__ __forceinline__ double tile_volume(const double *X, const double cc[3], const uint16_t *fnodes,
                                              const uint16_t *lf, uint32_t code, int gcl) {
    double V = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const uint16_t *fn = fnodes + 4 * lf[j];
        const double *x0 = X + 3 * fn[0], *x1 = X + 3 * fn[1], *x2 = X + 3 * fn[2], *x3 = X + 3 * fn[3];
        if (!gcl) {
            V += tet_volume(cc, x0, x1, x2);
            V += tet_volume(cc, x0, x3, x2);
        } else {
            const double ax = x2[0] - x0[0], ay = x2[1] - x0[1], az = x2[2] - x0[2];
            const double bx = x3[0] - x1[0], by = x3[1] - x1[1], bz = x3[2] - x1[2];
            const double s1 = 0.5 * (ay * bz - az * by);
            const double s2 = 0.5 * (az * bx - ax * bz);
            const double s3 = 0.5 * (ax * by - ay * bx);
            const double A = sqrt_n(s1 * s1 + s2 * s2 + s3 * s3);
            double n1 = s1 / A, n2 = s2 / A, n3 = s3 / A;
            double xc = 0.0, yc = 0.0, zc = 0.0;
            xc += x0[0]; yc += x0[1]; zc += x0[2];
            xc += x1[0]; yc += x1[1]; zc += x1[2];
            xc += x2[0]; yc += x2[1]; zc += x2[2];
            xc += x3[0]; yc += x3[1]; zc += x3[2];
            xc /= 4; yc /= 4; zc /= 4;
            // the orientation bit of the face for this cell: n points out of cells[0]; here the
            // face normal must point out of THIS cell, i.e. flip if n . (c - G) > 0
            if (spf(n1, n2, n3, cc[0] - xc, cc[1] - yc, cc[2] - zc) > 0.0) { n1 = -n1; n2 = -n2; n3 = -n3; }
            V += spf(xc, yc, zc, n1, n2, n3) * A;
        }
    }
    (void)code;
    return gcl ? V / 3.0 : V;
}

// Phase 3 for one face slot f of the tile: geometry at now1, face velocity, states (Snos_on_Gran),
// Riemann flux x area -> LDS slot, CFL candidate.  RT = double: chlld (parity); RT = float:
// chlld_mixed -- a face it reports ill-conditioned gets NaN in its PdivB slot and false is
// returned, for the FP64 pass over marked slots.
// The face's record and the states of its cells (sB only for an interior face) come from the
// caller, which issues those loads early (face_fetch).
template <bool SC>
__device__ __forceinline__ void face_fetch(const ulonglong2 &rec, const double *__restrict__ prim1,
                                           const double *__restrict__ sc1, double sA[8], double sB[8],
                                           double kA[kScal], double kB[kScal]) {
    const int64_t a = (int32_t)(rec.y & 0xffffffffu), b = (int32_t)(rec.y >> 32);
    load_state(prim1, a, sA);
    if constexpr (SC) {
        const double2 v = *reinterpret_cast<const double2 *>(sc1 + kScal * a);
        kA[0] = v.x; kA[1] = v.y;
    }
    if (((rec.x >> 48) & 15u) == 0u) {
        load_state(prim1, b, sB);
        if constexpr (SC) {
            const double2 v = *reinterpret_cast<const double2 *>(sc1 + kScal * b);
            kB[0] = v.x; kB[1] = v.y;
        }
    }
}

template <typename RT, bool TVD, bool SC>
__device__ __forceinline__ bool face_slot(int f, const ulonglong2 rec, const double sA[8], const double sB[8],
                                          const double kA[kScal], const double kB[kScal], const TileMesh &TM,
                                          const DevMesh &M, const DevParams &P, const TileLds &L, int fb, int64_t c0,
                                          int ncl, const double *__restrict__ prim1, const double *__restrict__ sc1,
                                          const double *__restrict__ cc1g, bool moving, int gcl, double dt,
                                          StepScalars *S, double &cand) {
    constexpr int K = SC ? kScal : 0;
    const int ln[4] = {(int)(rec.x & 0xfff), (int)((rec.x >> 12) & 0xfff), (int)((rec.x >> 24) & 0xfff),
                       (int)((rec.x >> 36) & 0xfff)};
    const uint32_t code = (uint32_t)(rec.x >> 48);
    const uint32_t bc = code & 15u, method = (code >> 4) & 15u, flags = (code >> 8) & 255u;
    const int64_t a = (int32_t)(rec.y & 0xffffffffu);
    const int64_t b = (int32_t)(rec.y >> 32);
#pragma unroll
    for (int i = 0; i < 4; i++) L.fnodes[4 * f + i] = (uint16_t)ln[i];
    // a face off the stepped area's face list is not evaluated (its cells are both frozen; its
    // node slots above still serve their volumes)
    if (P.area != 0 && !face_in_area(M.cell_code[a], b >= 0 ? M.cell_code[b] : 0u, bc != 0, P.area)) {
        L.flux[9 * f + 8] = 0.0;  // not a marked slot for the mixed solver's FP64 pass
        return true;
    }
    double kL[K > 0 ? K : 1], kR[K > 0 ? K : 1];
    double x1[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++) { x1[i][0] = L.x1[3 * ln[i]]; x1[i][1] = L.x1[3 * ln[i] + 1]; x1[i][2] = L.x1[3 * ln[i] + 2]; }
    const bool aloc = (a >= c0 && a < c0 + ncl);
    double ca[3];
    if (aloc) { ca[0] = L.c1[3 * (a - c0)]; ca[1] = L.c1[3 * (a - c0) + 1]; ca[2] = L.c1[3 * (a - c0) + 2]; }
    else { ca[0] = cc1g[3 * a]; ca[1] = cc1g[3 * a + 1]; ca[2] = cc1g[3 * a + 2]; }
    double n[3], area, G[3];
    face_geometry(x1, ca, gcl, n, area, G);
    double w = 0.0;
    if (moving) {
        double x2[4][3];
#pragma unroll
        for (int i = 0; i < 4; i++) { x2[i][0] = L.x2[3 * ln[i]]; x2[i][1] = L.x2[3 * ln[i] + 1]; x2[i][2] = L.x2[3 * ln[i] + 2]; }
        if (gcl) {
            w = swept_velocity(x1, x2, n, area, dt);
        } else {
            double ddot = 0.0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const double d0 = x2[i][0] - x1[i][0], d1 = x2[i][1] - x1[i][1], d2 = x2[i][2] - x1[i][2];
                ddot = ddot + ((d0 * n[0] + d1 * n[1]) + d2 * n[2]);
            }
            w = div_by(ddot / 4, dt, 1.0 / dt);  // 1/dt is loop-invariant: formed once per thread
        }
    }
    // face states (Snos_on_Gran)
    double qL[8], qR[8];
    bool okbc = true;
    double cb[3] = {0.0, 0.0, 0.0};
    if (bc != 0) {
        double l[8], r[8];
#pragma unroll
        for (int i = 0; i < 8; i++) l[i] = sA[i];
        if (bc == 3) {
            if (n[0] < -0.95) {
                if (l[1] > P.v_inf / 7.0) l[1] = P.v_inf / 5.0;
            } else if (l[1] * n[0] + l[2] * n[1] + l[3] * n[2] < 0.0) {
                l[1] = 0.1 * n[0]; l[2] = 0.1 * n[1]; l[3] = 0.1 * n[2];
            }
#pragma unroll
            for (int i = 0; i < 8; i++) r[i] = l[i];
        } else if (bc == 1 || bc == 2) {
            const double *bs = M.bstate + 8 * (int64_t)(-b - 2);
#pragma unroll
            for (int i = 0; i < 8; i++) r[i] = bs[i];
        } else {
            okbc = (bc == 4);
#pragma unroll
            for (int i = 0; i < 8; i++) r[i] = l[i];
        }
#pragma unroll
        for (int i = 0; i < 8; i++) { qL[i] = l[i]; qR[i] = r[i]; }
#pragma unroll
        for (int j = 0; j < K; j++) {
            kL[j] = kA[j];
            kR[j] = (bc == 1 || bc == 2) ? M.bscal[kScal * (int64_t)(-b - 2) + j] : kA[j];
        }
    } else {
        const bool bloc = (b >= c0 && b < c0 + ncl);
        if (bloc) { cb[0] = L.c1[3 * (b - c0)]; cb[1] = L.c1[3 * (b - c0) + 1]; cb[2] = L.c1[3 * (b - c0) + 2]; }
        else { cb[0] = cc1g[3 * b]; cb[1] = cc1g[3 * b + 1]; cb[2] = cc1g[3 * b + 2]; }
        const uint32_t zoneA = (M.cell_code[a] >> 8) & 15u, zoneB = (M.cell_code[b] >> 8) & 15u;
        const bool zA = zoneA == 1u, zB = zoneB == 1u;
        double tA[8], tB[8];
#pragma unroll
        for (int i = 0; i < 8; i++) { tA[i] = sA[i]; tB[i] = sB[i]; }
        int2 tv = {-1, -1};
        if constexpr (TVD) tv = TM.face_tvd[fb + f];
        if (TVD && tv.x >= 0) {  // second order: the line AA - A -|- B - BB (both neighbours present)
            double sAA[8], sBB[8], kAA[K > 0 ? K : 1], kBB[K > 0 ? K : 1];
            load_state(prim1, tv.x, sAA);
            load_state(prim1, tv.y, sBB);
#pragma unroll
            for (int j = 0; j < K; j++) { kAA[j] = sc1[K * (int64_t)tv.x + j]; kBB[j] = sc1[K * (int64_t)tv.y + j]; }
            const double cAA[3] = {cc1g[3 * (int64_t)tv.x], cc1g[3 * (int64_t)tv.x + 1], cc1g[3 * (int64_t)tv.x + 2]};
            const double cBB[3] = {cc1g[3 * (int64_t)tv.y], cc1g[3 * (int64_t)tv.y + 1], cc1g[3 * (int64_t)tv.y + 2]};
            tvd_states<K>((M.cell_code[tv.x] >> 8) & 15u, zoneA, zoneB, (M.cell_code[tv.y] >> 8) & 15u, sAA, tA, tB,
                          sBB, cAA, ca, cb, cBB, G, P.gamma, qL, qR, kAA, kA, kB, kBB, kL, kR);
        } else if (zA || zB) {
            const Zone1Face<double> zf = zone1_face<double>(G, P.gamma);
            if (zA) {
                const Zone1Cell<double> zc = aloc ? load_zone1<double>(L.z1 + 10 * (a - c0)) : zone1_cell<double>(ca, P.gamma);
                zone1_apply<double>(tA, zc, zf, qL);
#pragma unroll
                for (int j = 0; j < K; j++) kL[j] = div_by(kA[j] * (zc.r1 * zc.r1), zf.rr2, zf.yrr2);
            } else {
#pragma unroll
                for (int i = 0; i < 8; i++) qL[i] = tA[i];
#pragma unroll
                for (int j = 0; j < K; j++) kL[j] = kA[j];
            }
            if (zB) {
                const Zone1Cell<double> zc = bloc ? load_zone1<double>(L.z1 + 10 * (b - c0)) : zone1_cell<double>(cb, P.gamma);
                zone1_apply<double>(tB, zc, zf, qR);
#pragma unroll
                for (int j = 0; j < K; j++) kR[j] = div_by(kB[j] * (zc.r1 * zc.r1), zf.rr2, zf.yrr2);
            } else {
#pragma unroll
                for (int i = 0; i < 8; i++) qR[i] = tB[i];
#pragma unroll
                for (int j = 0; j < K; j++) kR[j] = kB[j];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) { qL[i] = tA[i]; qR[i] = tB[i]; }
#pragma unroll
            for (int j = 0; j < K; j++) { kL[j] = kA[j]; kR[j] = kB[j]; }
        }
    }
    // what the face needs after the solve (the PdivB flux and the CFL distance) is formed before
    // it, so that the states' field components and the two centres are not live across the solve
    const double pdivb = 0.5 * spf(n[0], n[1], n[2], qL[5] + qR[5], qL[6] + qR[6], qL[7] + qR[7]) * area;
    double dist = 0.0;
    if (bc == 0) {
        const double d0 = ca[0] - cb[0], d1 = ca[1] - cb[1], d2 = ca[2] - cb[2];
        dist = sqrt_n(d0 * d0 + d1 * d1 + d2 * d2) / 2.0;
    }
    RiemannOut<double, K> o;
    if constexpr (sizeof(RT) == sizeof(double)) {
        chlld<double, K>((int)method, P.n_disc, (int)flags, n[0], n[1], n[2], w, qL, qR, kL, kR, o);
    } else {
        chlld_mixed<K>((int)method, P.n_disc, (int)flags, n[0], n[1], n[2], w, qL, qR, kL, kR, o);
        if (o.ok < 0) {  // ill-conditioned in FP32: marked for the FP64 pass
            L.flux[9 * f + 8] = __builtin_nan("");
            return false;
        }
    }
    if (!o.ok || !okbc) report_error(S, kErrFaceRiemann, TM.face_ref[fb + f]);
    if (bc == 0 && o.ok) {
        const double dnt = P.kfl * dist / (smax(fabs((double)o.dsl), fabs((double)o.dsr)) + fabs(w));
        if (dnt < cand) cand = dnt;
    }
    double *fl = L.flux + 9 * f;
#pragma unroll
    for (int i = 0; i < 8; i++) fl[i] = o.q[i] * area;
    fl[8] = pdivb;
#pragma unroll
    for (int j = 0; j < K; j++) L.kflux[K * f + j] = o.k[j] * area;
    return true;
}

// Phase 4 of the surrogate step.  Lane l < 32 of wavefront w owns cell 32w + l of the tile (tile
// cells <= 32 * waves); lanes l and l + 32 form that cell's new-layer volume terms together (faces
// 0..2 and 3..5, summed by lane l in face order) and run its network pass together.  The network's
// operands are read from global memory (23 KB, cache-resident across the tiles): staging them in
// LDS once per workgroup measured slower, its copy and two barriers costing more than the cache
// reads they save.  The phase needs no barrier, and each wavefront proceeds on its own.
template <typename RT, bool SC>
__device__ __forceinline__ void surrogate_update(const TileMesh &TM, const DevMesh &M, const DevParams &P,
                                                 const TileLds &L, int64_t c0, int ncl,
                                                 const double *__restrict__ prim1, double *__restrict__ prim2,
                                                 const double *__restrict__ sc1, double *__restrict__ sc2,
                                                 const SurArgs &sur, double dt, int gcl, bool moving,
                                                 StepScalars *S) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cl = 32 * wave + (lane & 31);
    const bool cell = cl < ncl;  // this lane pair's cell exists
    const bool mine = lane < 32 && cell;
    const int64_t cg = c0 + (cell ? cl : 0);
    double st[8] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0}, hn[5] = {1.0, 1.0, 0.0, 0.0, 0.0};
    uint32_t code = 0;
    uint16_t lf[6] = {0, 0, 0, 0, 0, 0};
    float xr[8];
#pragma unroll
    for (int e = 0; e < 8; e++) xr[e] = 0.0f;
    bool inbox = false;
    if (cell) {
#pragma unroll
        for (int j = 0; j < 6; j++) lf[j] = TM.cell_lfaces[6 * cg + j];
    }
    if (mine) {  // loaded here: issued before the face phase's barrier, they would stay live across the
                 // mixed solver's FP64 re-solve call and spill
        load_state(prim1, cg, st);
#pragma unroll
        for (int k = 0; k < 5; k++) hn[k] = sur.neutrals[5 * cg + k];
        code = M.cell_code[cg];
        mlp_features(st, hn, xr);
        inbox = mlp_in_box(xr, sur.norm);
    }
    // flux sum and volumes, while the flux slots still hold the fluxes
    double Pf[9];
#pragma unroll
    for (int i = 0; i < 9; i++) Pf[i] = 0.0;
    double Pk[kScal] = {0.0, 0.0};
    double Volume = 0.0, Volume2 = 0.0;
    if (mine) {
#pragma unroll
        for (int j = 0; j < 6; j++) {
            const bool pos = (code >> j) & 1u;
            const double *fl = L.flux + 9 * lf[j];
#pragma unroll
            for (int i = 0; i < 9; i++) {
                const double v = fl[i];
                Pf[i] += pos ? v : -v;
            }
            if constexpr (SC) {
#pragma unroll
                for (int i = 0; i < kScal; i++) {
                    const double v = L.kflux[kScal * lf[j] + i];
                    Pk[i] += pos ? v : -v;
                }
            }
        }
        if (TM.vol_cached) {
            Volume = TM.vol[cg];
        } else {
            const double cA[3] = {L.c1[3 * cl], L.c1[3 * cl + 1], L.c1[3 * cl + 2]};
            Volume = tile_volume(L.x1, cA, L.fnodes, lf, code, gcl);
        }
        Volume2 = Volume;
    }
    if (moving) {
        double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (cell) {
            const double cB[3] = {L.c2[3 * cl], L.c2[3 * cl + 1], L.c2[3 * cl + 2]};
            const int j0 = lane < 32 ? 0 : 3;
#pragma unroll
            for (int j = 0; j < 3; j++) volume_terms(L.x2, cB, L.fnodes, lf[j0 + j], gcl, t + 2 * j);
        }
        double u[6];
#pragma unroll
        for (int k = 0; k < 6; k++) u[k] = __shfl(t[k], (lane & 31) + 32, 64);
        if (mine) {
            double V = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) { V += t[2 * j]; if (!gcl) V += t[2 * j + 1]; }
#pragma unroll
            for (int j = 0; j < 3; j++) { V += u[2 * j]; if (!gcl) V += u[2 * j + 1]; }
            Volume2 = gcl ? V / 3.0 : V;
        }
    }
    if (32 * wave >= ncl) return;  // wavefront-uniform: no cell of this wavefront
    bf16x8 in;
#pragma unroll
    for (int e = 0; e < 8; e++) in[e] = (lane < 32) ? mlp_normalise(xr[e], sur.norm, e) : (__bf16)0.0f;
    const f32x16 o = mlp_pass(reinterpret_cast<const bf16x8 *>(sur.wpack), sur.bpack, in, lane);
    const float y4 = __shfl(o[0], (lane & 31) + 32, 64);  // output 4 of this cell sits in lane l + 32
    if (!mine) return;
    if (moving || !TM.vol_cached) TM.vol[cg] = Volume2;
    if constexpr (SC) {  // source-independent: every cell's scalars, the fallback cells' included
        const double2 kv = *reinterpret_cast<const double2 *>(sc1 + kScal * cg);
        double k3[kScal] = {kv.x, kv.y};
        if (cell_in_area(code, P.area)) {
            const double k[kScal] = {kv.x, kv.y};
            scalar_update(k, st[0], Pf[0], Pk, Volume, Volume2, dt, k3);
        }
        *reinterpret_cast<double2 *>(sc2 + kScal * cg) = make_double2(k3[0], k3[1]);
    }
    if (!cell_in_area(code, P.area)) {  // frozen: the state carries over to the new layer
        double4 *o = reinterpret_cast<double4 *>(prim2 + 8 * cg);
        o[0] = make_double4(st[0], st[1], st[2], st[3]);
        o[1] = make_double4(st[4], st[5], st[6], st[7]);
    } else if (inbox) {
        const float y[5] = {o[0], o[1], o[2], o[3], y4};
        double s5[5];
        mlp_source(y, sur.norm, st, hn, s5);
        cell_update_state(M, code, cg, st, Volume, Volume2, Pf, s5, dt, S, prim2);
    } else {
        const unsigned long long k = atomicAdd(&S->n_fallback, 1ull);
        if ((int64_t)k < sur.cap) {
            sur.list[k] = (int32_t)cg;
            double *w = sur.stash + 11 * (int64_t)k;
#pragma unroll
            for (int i = 0; i < 9; i++) w[i] = Pf[i];
            w[9] = Volume;
            w[10] = Volume2;
        } else {
            report_error(S, kErrCellFallback, M.cell_ref[cg]);
        }
    }
}

// SUR = true: the charge-exchange source comes from the surrogate network, evaluated in phase 4
// on the matrix cores (one 32-cell MFMA pass per wavefront, its operands read from global memory
// and cache-resident across the tiles).  Cells outside the training box are not updated here: their flux
// sums and volumes go to sur.stash at a slot of S->n_fallback, and k_cx_fixup redoes their update
// with the exact source.  SUR = false: src (nullable) holds a precomputed source per cell.
//
// SC = true: the passive scalars Q and rho_He (sc1 -> sc2, [Nc*2]) are carried: reconstructed and
// solved with the face states, summed and updated with the cell (scalar_update).
template <typename RT, bool SUR, bool TVD, bool SC>
__global__ void __launch_bounds__(kTileThreads, 2)
    k_step_tile(TileMesh TM, DevMesh M, DevParams P, const double *__restrict__ prim1, double *__restrict__ prim2,
                const double *__restrict__ sc1, double *__restrict__ sc2, const double *__restrict__ X1g,
                double *__restrict__ X2g, const double *__restrict__ disp, const double *__restrict__ cc1g,
                double *__restrict__ cc2g, const double *__restrict__ src, SurArgs sur, StepScalars *S,
                long long step) {
    extern __shared__ __align__(16) double smem[];
    const TileLds L = tile_lds(smem, TM, P.move_mesh != 0, SC);
    // Workgroups go to the 8 XCDs round robin (block b on XCD b % 8), each XCD with its own L2.
    // Tile t(b) = (b % 8) * per + b / 8 gives every XCD a contiguous run of tiles, so neighbouring
    // tiles -- which share nodes and border-cell states -- are read through the same L2.
    const int per = (int)((TM.n_tiles + kXcds - 1) / kXcds);
    const int t = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
    if (t >= TM.n_tiles) return;  // padding blocks of the last run (whole workgroup, before any barrier)
    const int64_t c0 = TM.cell_off[t];
    const int ncl = TM.cell_off[t + 1] - (int)c0;
    const int nb = TM.node_off[t], nn = TM.node_off[t + 1] - nb;
    const int fb = TM.face_off[t], nf = TM.face_off[t + 1] - fb;
    const int own_lo = TM.own[2 * t], own_hi = TM.own[2 * t + 1];
    const bool moving = P.move_mesh != 0;
    const bool given = P.move_mesh == 2;
    const int gcl = P.geom_mode;
    const double dt = S->dt;
    const double fac = S->motion_fac;
    const int tid = threadIdx.x;
    // Loads that do not depend on earlier phases are issued here, so their latency overlaps the
    // phases before their use: this thread's first face record, its first two tile nodes, and
    // the node list and code of its own cell.
    ulonglong2 rec0 = {0ull, 0ull};
    if (tid < nf) rec0 = TM.faces[fb + tid];
    const int g_pre0 = tid < nn ? TM.nodes[nb + tid] : 0;
    const int g_pre1 = tid + kTileThreads < nn ? TM.nodes[nb + tid + kTileThreads] : 0;
    uint4 ln_pre = {0u, 0u, 0u, 0u};
    uint32_t code_pre = 0u;
    if (tid < ncl) {
        ln_pre = *reinterpret_cast<const uint4 *>(TM.cell_lnodes + 8 * (c0 + tid));
        code_pre = M.cell_code[c0 + tid];
    }

    // phase 1: nodes of the tile, both layers
    for (int i = tid; i < nn; i += blockDim.x) {
        const int64_t g = i == tid ? g_pre0 : (i == tid + kTileThreads ? g_pre1 : TM.nodes[nb + i]);
        const double a0 = X1g[3 * g], a1 = X1g[3 * g + 1], a2 = X1g[3 * g + 2];
        L.x1[3 * i] = a0; L.x1[3 * i + 1] = a1; L.x1[3 * i + 2] = a2;
        if (given) {  // new layer written before the step (fitted surfaces and their rays)
            L.x2[3 * i] = X2g[3 * g]; L.x2[3 * i + 1] = X2g[3 * g + 1]; L.x2[3 * i + 2] = X2g[3 * g + 2];
        } else if (moving) {
            const double b0 = a0 + fac * disp[3 * g], b1 = a1 + fac * disp[3 * g + 1], b2 = a2 + fac * disp[3 * g + 2];
            L.x2[3 * i] = b0; L.x2[3 * i + 1] = b1; L.x2[3 * i + 2] = b2;
            if (g >= own_lo && g < own_hi) { X2g[3 * g] = b0; X2g[3 * g + 1] = b1; X2g[3 * g + 2] = b2; }
        }
    }
    __syncthreads();
    const double *LX2 = moving ? L.x2 : L.x1;

    // phase 2: own cells -- centres at both layers, Zone_1 frames at now1
    for (int c = tid; c < ncl; c += blockDim.x) {
        uint16_t ln[8];
        uint32_t ccode;
        if (c == tid) {
            const uint32_t w4[4] = {ln_pre.x, ln_pre.y, ln_pre.z, ln_pre.w};
#pragma unroll
            for (int i = 0; i < 4; i++) { ln[2 * i] = (uint16_t)(w4[i] & 0xffffu); ln[2 * i + 1] = (uint16_t)(w4[i] >> 16); }
            ccode = code_pre;
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) ln[i] = TM.cell_lnodes[8 * (c0 + c) + i];
            ccode = M.cell_code[c0 + c];
        }
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const double *x = L.x1 + 3 * ln[i];
            a0 += x[0]; a1 += x[1]; a2 += x[2];
            const double *y = LX2 + 3 * ln[i];
            b0 += y[0]; b1 += y[1]; b2 += y[2];
        }
        const double cA[3] = {a0 / 8, a1 / 8, a2 / 8};
        L.c1[3 * c] = cA[0]; L.c1[3 * c + 1] = cA[1]; L.c1[3 * c + 2] = cA[2];
        if (moving) {
            L.c2[3 * c] = b0 / 8; L.c2[3 * c + 1] = b1 / 8; L.c2[3 * c + 2] = b2 / 8;
            cc2g[3 * (c0 + c)] = b0 / 8; cc2g[3 * (c0 + c) + 1] = b1 / 8; cc2g[3 * (c0 + c) + 2] = b2 / 8;
        }
        if (((ccode >> 8) & 15u) == 1u) {
            const Zone1Cell<double> z = zone1_cell<double>(cA, P.gamma);
            double *o = L.z1 + 10 * c;
            o[0] = z.r1; o[1] = z.pr1;
            o[2] = z.e.st; o[3] = z.e.ct; o[4] = z.e.sp; o[5] = z.e.cp;
            o[6] = z.p.st; o[7] = z.p.ct; o[8] = z.p.sp; o[9] = z.p.cp;
        }
    }
    // the states of this thread's first face, in flight across the barrier
    double sA0[8], sB0[8], kA0[kScal], kB0[kScal];
    if (tid < nf) face_fetch<SC>(rec0, prim1, sc1, sA0, sB0, kA0, kB0);
    __syncthreads();

    // phase 3: faces of the tile (at most two per thread: tiles are cut at 2 x kTileThreads faces)
    double cand = 10000000.0;
    bool pending = false;
    if (tid < nf)  // the first face, with its prefetched record and states
        pending = !face_slot<RT, TVD, SC>(tid, rec0, sA0, sB0, kA0, kB0, TM, M, P, L, fb, c0, ncl, prim1, sc1, cc1g,
                                          moving, gcl, dt, S, cand);
    for (int f = tid + kTileThreads; f < nf; f += kTileThreads) {
        const ulonglong2 rec = TM.faces[fb + f];
        double sA[8], sB[8], kA[kScal], kB[kScal];
        face_fetch<SC>(rec, prim1, sc1, sA, sB, kA, kB);
        pending |= !face_slot<RT, TVD, SC>(f, rec, sA, sB, kA, kB, TM, M, P, L, fb, c0, ncl, prim1, sc1, cc1g, moving,
                                           gcl, dt, S, cand);
    }
    // the update phase's loads of this thread's cell, in flight across the barrier
    uint32_t code4 = 0u;
    uint32_t lf_pre[3] = {0u, 0u, 0u};
    double s4[8];
    double v4 = 0.0;
    if (!SUR && tid < ncl) {
        code4 = M.cell_code[c0 + tid];
        if (TM.vol_cached) v4 = TM.vol[c0 + tid];
        const uint32_t *lfw = reinterpret_cast<const uint32_t *>(TM.cell_lfaces + 6 * (c0 + tid));
        lf_pre[0] = lfw[0]; lf_pre[1] = lfw[1]; lf_pre[2] = lfw[2];
        load_state(prim1, c0 + tid, s4);
    }
    uint32_t hw1 = 0u, hw2 = 0u;  // the volume helpers' face slots (faces 3..5 of cell tid - half)
    if (!SUR && moving && tid >= kTileThreads / 2 && tid - kTileThreads / 2 < ncl) {
        const uint32_t *lfw = reinterpret_cast<const uint32_t *>(TM.cell_lfaces + 6 * (c0 + tid - kTileThreads / 2));
        hw1 = lfw[1]; hw2 = lfw[2];
    }
    unsigned long long n_resolved = 0;  // this thread's faces re-solved in FP64 (counted per block)
    if constexpr (sizeof(RT) == sizeof(double)) {
        __syncthreads();
    } else if (__syncthreads_or(pending)) {
        // phase 3b: the slots the mixed solver marked ill-conditioned, solved in FP64
        for (int f = tid; f < nf; f += blockDim.x) {
            if (__builtin_isnan(L.flux[9 * f + 8])) {
                const ulonglong2 rec = TM.faces[fb + f];
                double sA[8], sB[8], kA[kScal], kB[kScal];
                face_fetch<SC>(rec, prim1, sc1, sA, sB, kA, kB);
                face_slot<double, TVD, SC>(f, rec, sA, sB, kA, kB, TM, M, P, L, fb, c0, ncl, prim1, sc1, cc1g, moving,
                                           gcl, dt, S, cand);
                n_resolved++;
            }
        }
        __syncthreads();
    }

    if constexpr (SUR) {
        surrogate_update<RT, SC>(TM, M, P, L, c0, ncl, prim1, prim2, sc1, sc2, sur, dt, gcl, moving, S);
        block_min_count_to(cand, n_resolved, &S->dtmin_bits[step & 1], &S->n_fp64_faces);
        return;
    }

    // phase 4a (moving mesh): the new-layer volume of each own cell is summed over its faces in
    // order; the terms of faces 3..5 come from the upper half of the workgroup, otherwise idle in
    // this phase, through LDS (the Zone_1 frames' space, dead after the face phase), while the
    // cell's own thread forms those of faces 0..2.  The sum is taken in the same order, so the
    // volume is the same bits as one thread's.
    const int half = kTileThreads / 2;
    const bool split = moving && ncl <= half;
    double v2part = 0.0;
    if (split) {
        double *vt = L.z1;  // [cell][6]
        if (tid >= half && tid - half < ncl) {
            const int c = tid - half;
            const uint16_t lfh[3] = {(uint16_t)(hw1 >> 16), (uint16_t)(hw2 & 0xffffu), (uint16_t)(hw2 >> 16)};
            const double cB[3] = {L.c2[3 * c], L.c2[3 * c + 1], L.c2[3 * c + 2]};
#pragma unroll
            for (int j = 0; j < 3; j++) volume_terms(L.x2, cB, L.fnodes, lfh[j], gcl, vt + 6 * c + 2 * j);
        }
        if (tid < ncl) {
            const double cB[3] = {L.c2[3 * tid], L.c2[3 * tid + 1], L.c2[3 * tid + 2]};
            uint16_t lf3[3] = {(uint16_t)(lf_pre[0] & 0xffffu), (uint16_t)(lf_pre[0] >> 16), (uint16_t)(lf_pre[1] & 0xffffu)};
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double t[2];
                volume_terms(L.x2, cB, L.fnodes, lf3[j], gcl, t);
                v2part += t[0];
                if (!gcl) v2part += t[1];
            }
        }
        __syncthreads();
        if (tid < ncl) {
#pragma unroll
            for (int j = 0; j < 3; j++) {
                v2part += vt[6 * tid + 2 * j];
                if (!gcl) v2part += vt[6 * tid + 2 * j + 1];
            }
            if (gcl) v2part = v2part / 3.0;
        }
    }

    // phase 4: own cells -- volumes, flux sum in reference order, update
    for (int c = tid; c < ncl; c += blockDim.x) {
        const int64_t cg = c0 + c;
        uint16_t lf[6];
        uint32_t code;
        double s[8];
        double vcache = 0.0;
        if (c == tid) {
            code = code4;
            vcache = v4;
#pragma unroll
            for (int i = 0; i < 3; i++) { lf[2 * i] = (uint16_t)(lf_pre[i] & 0xffffu); lf[2 * i + 1] = (uint16_t)(lf_pre[i] >> 16); }
#pragma unroll
            for (int i = 0; i < 8; i++) s[i] = s4[i];
        } else {
            code = M.cell_code[cg];
#pragma unroll
            for (int i = 0; i < 6; i++) lf[i] = TM.cell_lfaces[6 * cg + i];
            load_state(prim1, cg, s);
            if (TM.vol_cached) vcache = TM.vol[cg];
        }
        double Volume = vcache;
        if (!TM.vol_cached) {
            const double cA[3] = {L.c1[3 * c], L.c1[3 * c + 1], L.c1[3 * c + 2]};
            Volume = tile_volume(L.x1, cA, L.fnodes, lf, code, gcl);
        }
        double Volume2 = Volume;
        if (split) {
            Volume2 = v2part;
        } else if (moving) {
            const double cB[3] = {L.c2[3 * c], L.c2[3 * c + 1], L.c2[3 * c + 2]};
            Volume2 = tile_volume(L.x2, cB, L.fnodes, lf, code, gcl);
        }
        if (moving || !TM.vol_cached) TM.vol[cg] = Volume2;
        if (!cell_in_area(code, P.area)) {  // frozen: the state carries over to the new layer
            double4 *o = reinterpret_cast<double4 *>(prim2 + 8 * cg);
            o[0] = make_double4(s[0], s[1], s[2], s[3]);
            o[1] = make_double4(s[4], s[5], s[6], s[7]);
            if constexpr (SC) *reinterpret_cast<double2 *>(sc2 + kScal * cg) = *reinterpret_cast<const double2 *>(sc1 + kScal * cg);
            continue;
        }
        double Pf[9];
#pragma unroll
        for (int i = 0; i < 9; i++) Pf[i] = 0.0;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            const bool pos = (code >> j) & 1u;
            const double *fl = L.flux + 9 * lf[j];
#pragma unroll
            for (int i = 0; i < 9; i++) {
                const double v = fl[i];
                Pf[i] += pos ? v : -v;
            }
        }
        if constexpr (SC) {
            double Pk[kScal] = {0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const bool pos = (code >> j) & 1u;
#pragma unroll
                for (int i = 0; i < kScal; i++) {
                    const double v = L.kflux[kScal * lf[j] + i];
                    Pk[i] += pos ? v : -v;
                }
            }
            const double2 kv = *reinterpret_cast<const double2 *>(sc1 + kScal * cg);
            const double k[kScal] = {kv.x, kv.y};
            double k3[kScal];
            scalar_update(k, s[0], Pf[0], Pk, Volume, Volume2, dt, k3);
            *reinterpret_cast<double2 *>(sc2 + kScal * cg) = make_double2(k3[0], k3[1]);
        }
        cell_update_state(M, code, cg, s, Volume, Volume2, Pf, src ? src + 5 * cg : nullptr, dt, S, prim2);
    }
    block_min_count_to(cand, n_resolved, &S->dtmin_bits[step & 1], &S->n_fp64_faces);
}

// Launch of one k_step_tile instantiation; opens the kernel up to the tile's dynamic LDS size
// once (hipFuncSetAttribute is a no-op below the default limit but required above it).
template <typename RT, bool SUR, bool TVD, bool SC>
inline void launch_tile_variant(hipStream_t s, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                                const double *prim1, double *prim2, const double *sc1, double *sc2,
                                const double *X1, double *X2, const double *disp, const double *cc1, double *cc2,
                                const double *src, const SurArgs &sur, StepScalars *S, long long step) {
    const size_t lds = tile_lds_bytes(TM.max_nodes, TM.tile_cells, TM.max_faces, P.move_mesh != 0, SC);
    static size_t attr_bytes = 0;  // dynamic LDS size this instantiation has been opened up to
    if (lds > attr_bytes) {
        if (hipFuncSetAttribute((const void *)k_step_tile<RT, SUR, TVD, SC>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess)
            attr_bytes = lds;
        else
            (void)hipGetLastError();  // the launch below then reports the failure
    }
    const unsigned grid = (unsigned)(kXcds * ((TM.n_tiles + kXcds - 1) / kXcds));
    hipLaunchKernelGGL((k_step_tile<RT, SUR, TVD, SC>), dim3(grid), dim3(kTileThreads), lds, s, TM, M, P, prim1,
                       prim2, sc1, sc2, X1, X2, disp, cc1, cc2, src, sur, S, step);
}

// The four source / reconstruction variants of one precision and scalar choice.
template <typename RT, bool SC>
inline void launch_tile_variants(hipStream_t s, const TileMesh &TM, const DevMesh &M, const DevParams &P,
                                 const double *prim1, double *prim2, const double *sc1, double *sc2,
                                 const double *X1, double *X2, const double *disp, const double *cc1, double *cc2,
                                 const double *src, const SurArgs *sur, StepScalars *S, long long step) {
    const SurArgs none{};
    const SurArgs &sa = sur ? *sur : none;
    const double *sp = sur ? nullptr : src;
    if (sur && P.tvd)
        launch_tile_variant<RT, true, true, SC>(s, TM, M, P, prim1, prim2, sc1, sc2, X1, X2, disp, cc1, cc2, sp, sa, S, step);
    else if (sur)
        launch_tile_variant<RT, true, false, SC>(s, TM, M, P, prim1, prim2, sc1, sc2, X1, X2, disp, cc1, cc2, sp, sa, S, step);
    else if (P.tvd)
        launch_tile_variant<RT, false, true, SC>(s, TM, M, P, prim1, prim2, sc1, sc2, X1, X2, disp, cc1, cc2, sp, sa, S, step);
    else
        launch_tile_variant<RT, false, false, SC>(s, TM, M, P, prim1, prim2, sc1, sc2, X1, X2, disp, cc1, cc2, sp, sa, S, step);
}

}  // namespace titan
