// tvd.h -- second-order limited face states: the TVD branch of Setka::Snos_on_Gran
// (Setka_gran_snos.cpp:186-736) with the limiter of Help_funk.cpp:26-70.
//
// Along the line AA - A -|- B - BB (AA, BB = the face's second neighbours chosen by
// Setka::Init_TVD) every state component is placed at its distance from the face centre G
// (AA at -dd1 - d1, A at -d1, B at d2, BB at d2 + dd2) and evaluated at 0 on a minmod-limited
// line through three cells (`linear`) or on the plain line through two (`linear2`).  Which lines
// the left and right states use, and in which variables, depends on the zones of the four cells:
//
//   zones (AA A B BB)     variables (A side / B side)                  left          right
//   1 1 1 1               spherical, radially scaled / same            AA,A,B        A,B,BB
//   1 1 1 2               spherical, p scaled / same                   AA,A,B        A-B
//   1 1 2 2               spherical, p scaled / primitives             AA-A          B-BB
//   1 2 2 2               primitives                                   A-B           A,B,BB
//   2 2 3 3 | 3 3 4 4     primitives                                   AA-A          B-BB
//   2 2 2 3 | 3 3 3 4     primitives                                   AA,A,B        A-B
//   2 3 3 3 | 3 4 4 4     primitives                                   A-B           A,B,BB
//   otherwise             primitives                                   AA,A,B        A,B,BB
//
// "Spherical" = V and B as (r, phi, theta) components in the frame the face centre selects (as
// in the first-order Zone_1 rescaling); "radially scaled" = B_r x r^2, B_phi and B_theta x r,
// rho x r^2, p x r^(2 gamma), each divided by the same power of |G| after the interpolation;
// "p scaled" = only p x r^(2 gamma), with rho kept as it is.  (In the 1 1 1 2 and 1 1 2 2 rows the
// reference assigns rho x r^2 and then overwrites it with the plain interpolation of its final
// branch; the table gives the value it keeps.)  A reconstructed density or pressure below 1e-7
// falls back to the cell's own value (Setka_gran_snos.cpp:715-725).
//
// The passive scalars Q and rho_He belong to the reference's r^2 names with rho
// (Phys_param.cpp:284-286), so every case reconstructs them exactly as it does rho (x r^2 in the
// 1 1 1 1 row, plain elsewhere), with the same 1e-7 fallback to the cell's own value.
#pragma once
#include "sphere.h"

namespace titan {

// signum with a dead band of 1e-8 (Help_funk.cpp:26-40)
__device__ __forceinline__ int tvd_signum(double x) {
    if (x > 0.00000001) return 1;
    if (x < -0.00000001) return -1;
    return 0;
}

// Help_funk.cpp:43-53
__device__ __forceinline__ double tvd_minmod(double x, double y) {
    const int s = tvd_signum(x) + tvd_signum(y);
    if (s == 0) return 0.0;
    const double ax = fabs(x), ay = fabs(y);
    return (s / 2.0) * ((ay < ax) ? ay : ax);
}

// The abscissae of a face's lines are the same for all eight components: the differences
// x1 - x2 and x2 - x3 and their correctly rounded reciprocals are formed once per face, and the
// slopes (t1 - t2) / (x1 - x2) come out correctly rounded through div_by (fdiv.h).
struct TvdGap { double d, y; };  // a difference of abscissae and its reciprocal

__device__ __forceinline__ TvdGap tvd_gap(double x1, double x2) {
    const double d = x1 - x2;
    return {d, 1.0 / d};
}

// Help_funk.cpp:55-62, at y = 0: the minmod-limited line through (x1,t1), (x2,t2), (x3,t3)
__device__ __forceinline__ double tvd_linear(TvdGap g12, TvdGap g23, double x2, double t1, double t2, double t3) {
    const double d = tvd_minmod(div_by(t1 - t2, g12.d, g12.y), div_by(t2 - t3, g23.d, g23.y));
    return d * (0.0 - x2) + t2;
}

// Help_funk.cpp:64-69, at y = 0: the line through (x1,t1), (x2,t2)
__device__ __forceinline__ double tvd_linear2(TvdGap g12, double x2, double t1, double t2) {
    const double d = div_by(t1 - t2, g12.d, g12.y);
    return d * (0.0 - x2) + t2;
}

__device__ __forceinline__ double tvd_norm3(double a, double b, double c) { return sqrt_n((a * a + b * b) + c * c); }

// Variables of one reconstruction: 0 = primitives, 1 = spherical and radially scaled, 2 = spherical
// with p scaled.  Spherical slots hold (r, phi, theta) at 1..3 (V) and 5..7 (B).
__device__ __forceinline__ void tvd_forward(const double s[8], const double c[3], int mode, const Zone1Face<double> &F,
                                            double gamma, double t[8]) {
    if (mode == 0) {
#pragma unroll
        for (int i = 0; i < 8; i++) t[i] = s[i];
        return;
    }
    const double r = tvd_norm3(c[0], c[1], c[2]);
    const Trig<double> a = F.equatorial ? frame_trig<double>(c[2], c[0], c[1]) : frame_trig<double>(c[0], c[1], c[2]);
#pragma unroll
    for (int ik = 0; ik < 2; ik++) {
        const double *V = s + (ik == 0 ? 1 : 5);
        double *T = t + (ik == 0 ? 1 : 5);
        const double Vz = F.equatorial ? V[2] : V[0];
        const double Vx = F.equatorial ? V[0] : V[1];
        const double Vy = F.equatorial ? V[1] : V[2];
        double Vr = Vx * a.st * a.cp + Vy * a.st * a.sp + Vz * a.ct;
        double Vtheta = Vx * a.ct * a.cp + Vy * a.ct * a.sp - Vz * a.st;
        double Vphi = -Vx * a.sp + Vy * a.cp;
        if (ik == 1 && mode == 1) {
            Vr = Vr * (r * r);
            Vphi = Vphi * r;
            Vtheta = Vtheta * r;
        }
        T[0] = Vr; T[1] = Vphi; T[2] = Vtheta;
    }
    t[0] = mode == 1 ? s[0] * (r * r) : s[0];
    t[4] = s[4] * pow_2gamma<double>(r, gamma);
}

// Inverse of tvd_forward at the face centre.
__device__ __forceinline__ void tvd_back(const double t[8], int mode, const Zone1Face<double> &F, double q[8]) {
    if (mode == 0) {
#pragma unroll
        for (int i = 0; i < 8; i++) q[i] = t[i];
        return;
    }
    const Trig<double> b = F.t;
#pragma unroll
    for (int ik = 0; ik < 2; ik++) {
        const double *T = t + (ik == 0 ? 1 : 5);
        double *O = q + (ik == 0 ? 1 : 5);
        double Vr = T[0], Vphi = T[1], Vtheta = T[2];
        if (ik == 1 && mode == 1) {
            Vr = div_by(Vr, F.rr2, F.yrr2);
            Vphi = div_by(Vphi, F.rr, F.yrr);
            Vtheta = div_by(Vtheta, F.rr, F.yrr);
        }
        const double ox = Vr * b.st * b.cp + Vtheta * b.ct * b.cp - Vphi * b.sp;
        const double oy = Vr * b.st * b.sp + Vtheta * b.ct * b.sp + Vphi * b.cp;
        const double oz = Vr * b.ct - Vtheta * b.st;
        if (F.equatorial) { O[2] = oz; O[0] = ox; O[1] = oy; }
        else { O[0] = oz; O[1] = ox; O[2] = oy; }
    }
    q[0] = mode == 1 ? div_by(t[0], F.rr2, F.yrr2) : t[0];
    q[4] = div_by(t[4], F.prr, F.yprr);
}

// One row of the table above: variables of the A side and the B side; left line 0 = AA,A,B
// limited, 1 = AA-A, 2 = A-B; right line 0 = A,B,BB limited, 1 = A-B, 2 = B-BB.
struct TvdCase { int ma, mb, lm, rm; };

__device__ __forceinline__ TvdCase tvd_case(uint32_t zAA, uint32_t zA, uint32_t zB, uint32_t zBB) {
    const uint32_t key = (zAA << 12) | (zA << 8) | (zB << 4) | zBB;
    switch (key) {
        case 0x1111: return {1, 1, 0, 0};                 // Setka_gran_snos.cpp:216-400
        case 0x1112: return {2, 2, 0, 1};                 // :401-536
        case 0x1122: return {2, 0, 1, 2};                 // :537-648
        case 0x1222: return {0, 0, 2, 0};                 // :649-660
        case 0x2233: case 0x3344: return {0, 0, 1, 2};    // :661-673
        case 0x2223: case 0x3334: return {0, 0, 0, 1};    // :674-687
        case 0x2333: case 0x3444: return {0, 0, 2, 0};    // :688-701
        default: return {0, 0, 0, 0};                     // :702-713
    }
}

// Second-order states (qL from the A side, qR from the B side) of an interior face with cells A
// (cells[0]) and B and second neighbours AA and BB: zones z*, states s*, centres c* at the layer
// of the face centre G.  K passive scalars k* (nullable when K = 0) give kL, kR.
template <int K = 0>
__device__ __forceinline__ void tvd_states(uint32_t zAA, uint32_t zA, uint32_t zB, uint32_t zBB,
                                           const double sAA[8], const double sA[8], const double sB[8],
                                           const double sBB[8], const double cAA[3], const double cA[3],
                                           const double cB[3], const double cBB[3], const double G[3],
                                           double gamma, double qL[8], double qR[8],
                                           const double *kAA = nullptr, const double *kA = nullptr,
                                           const double *kB = nullptr, const double *kBB = nullptr,
                                           double *kL = nullptr, double *kR = nullptr) {
    const double d1 = tvd_norm3(cA[0] - G[0], cA[1] - G[1], cA[2] - G[2]);
    const double d2 = tvd_norm3(cB[0] - G[0], cB[1] - G[1], cB[2] - G[2]);
    const double dd1 = tvd_norm3(cA[0] - cAA[0], cA[1] - cAA[1], cA[2] - cAA[2]);
    const double dd2 = tvd_norm3(cB[0] - cBB[0], cB[1] - cBB[1], cB[2] - cBB[2]);
    const TvdCase k = tvd_case(zAA, zA, zB, zBB);
    Zone1Face<double> F = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0, true, {1.0, 0.0, 0.0, 1.0}};
    if (k.ma != 0) F = zone1_face<double>(G, gamma);
    double tAA[8], tA[8], tB[8], tBB[8];
    tvd_forward(sAA, cAA, k.ma, F, gamma, tAA);
    tvd_forward(sA, cA, k.ma, F, gamma, tA);
    tvd_forward(sB, cB, k.mb, F, gamma, tB);
    tvd_forward(sBB, cBB, k.mb, F, gamma, tBB);
    const double xAA = -dd1 - d1, xA = -d1, xB = d2, xBB = d2 + dd2;
    const TvdGap gL = tvd_gap(xAA, xA), gC = tvd_gap(xA, xB), gR = tvd_gap(xB, xBB);
    auto left = [&](double a, double b, double c) {
        return k.lm == 0 ? tvd_linear(gL, gC, xA, a, b, c) : (k.lm == 1 ? tvd_linear2(gL, xA, a, b) : tvd_linear2(gC, xB, b, c));
    };
    auto right = [&](double a, double b, double c) {
        return k.rm == 0 ? tvd_linear(gC, gR, xB, a, b, c) : (k.rm == 1 ? tvd_linear2(gC, xB, a, b) : tvd_linear2(gR, xBB, b, c));
    };
    double l[8], r[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        l[i] = left(tAA[i], tA[i], tB[i]);
        r[i] = right(tA[i], tB[i], tBB[i]);
    }
    tvd_back(l, k.ma, F, qL);
    tvd_back(r, k.mb, F, qR);
    if (qL[0] < 0.0000001) qL[0] = sA[0];
    if (qL[4] < 0.0000001) qL[4] = sA[4];
    if (qR[0] < 0.0000001) qR[0] = sB[0];
    if (qR[4] < 0.0000001) qR[4] = sB[4];
    if constexpr (K > 0) {
        // the scalars' own lines: rho's variables (x r^2 only in the radially scaled mode)
        double wAA = 1.0, wA = 1.0, wB = 1.0, wBB = 1.0;
        if (k.ma == 1) {
            const double rAA = tvd_norm3(cAA[0], cAA[1], cAA[2]), rA = tvd_norm3(cA[0], cA[1], cA[2]);
            wAA = rAA * rAA;
            wA = rA * rA;
        }
        if (k.mb == 1) {
            const double rB = tvd_norm3(cB[0], cB[1], cB[2]), rBB = tvd_norm3(cBB[0], cBB[1], cBB[2]);
            wB = rB * rB;
            wBB = rBB * rBB;
        }
#pragma unroll
        for (int j = 0; j < K; j++) {
            const double a = k.ma == 1 ? kAA[j] * wAA : kAA[j], b = k.ma == 1 ? kA[j] * wA : kA[j];
            const double c = k.mb == 1 ? kB[j] * wB : kB[j], d = k.mb == 1 ? kBB[j] * wBB : kBB[j];
            // the lines take each cell in the variables of its own side, as tvd_forward does
            const double lv = left(a, b, c), rv = right(b, c, d);
            kL[j] = k.ma == 1 ? div_by(lv, F.rr2, F.yrr2) : lv;
            kR[j] = k.mb == 1 ? div_by(rv, F.rr2, F.yrr2) : rv;
            if (kL[j] < 0.0000001) kL[j] = kA[j];
            if (kR[j] < 0.0000001) kR[j] = kB[j];
        }
    }
}
}  // namespace titan
