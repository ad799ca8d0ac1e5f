// sphere.h -- spherical frames of the reference's Zone_1 formulas (Help_funk.cpp:78-135) in
// algebraic form, and the first-order Zone_1 rescaling of a cell state to a face centre
// (Setka_gran_snos.cpp:744-873); shared by the step kernels and the TVD reconstruction.
#pragma once
#include <hip/hip_runtime.h>
#include "fdiv.h"
#include "fsqrt.h"

namespace titan {

template <typename Z> struct Trig { Z st, ct, sp, cp; };

// (theta, phi) trig of spherical_skorost / dekard_skorost (Help_funk.cpp:108-135) for the axis
// roles (z, x, y): theta = acos(z/r), phi = polar_angle(x, y) (Help_funk.cpp:78-106).  The
// reference evaluates sin/cos of those angles; they equal rho/r, z/r, y/rho, x/rho with
// rho = sqrt(x^2 + y^2) in every quadrant branch of polar_angle (phi = 0 when |x| + |y| < 1e-6),
// so they are formed directly from the coordinates (one square root and four divisions instead of
// acos, atan and four sin/cos; the values agree to a few ulp, the accuracy at which GPU and host
// libm already differ).  The V and B rotations share them.
template <typename Z>
__device__ __forceinline__ Trig<Z> frame_trig(Z z, Z x, Z y) {
    const Z r = sqrt_t<Z>(x * x + y * y + z * z);
    const Z rho = sqrt_t<Z>(x * x + y * y);
    Trig<Z> t;
    const Z yr = Z(1) / r;
    t.st = div_by_t<Z>(rho, r, yr);
    t.ct = div_by_t<Z>(z, r, yr);
    if (fabs(x) + fabs(y) < Z(0.000001)) { t.sp = Z(0); t.cp = Z(1); }
    else {
        const Z yrho = Z(1) / rho;
        t.sp = div_by_t<Z>(y, rho, yrho);
        t.cp = div_by_t<Z>(x, rho, yrho);
    }
    return t;
}

// r^(2 gamma): for gamma = 5/3, r^(10/3) = r^3 cbrt(r) (a cube root instead of a logarithm and an
// exponential; within a few ulp of pow)
template <typename Z>
__device__ __forceinline__ Z pow_2gamma(Z r, double gamma) {
    if (gamma == 5.0 / 3.0) return (r * r * r) * cbrt(r);
    return pow(r, Z(2.0) * (Z)gamma);
}


// Zone_1 cell data: r1, r1^(2 gamma), trig of the equatorial (z,x,y) = (c2,c0,c1) and polar
// (z,x,y) = (c0,c1,c2) frames.
template <typename Z> struct Zone1Cell { Z r1, pr1; Trig<Z> e, p; };

template <typename Z>
__device__ __forceinline__ Zone1Cell<Z> zone1_cell(const double c[3], double gamma) {
    const Z c0 = (Z)c[0], c1 = (Z)c[1], c2 = (Z)c[2];
    Zone1Cell<Z> z;
    z.r1 = sqrt_t<Z>((c0 * c0 + c1 * c1) + c2 * c2);
    z.pr1 = pow_2gamma<Z>(z.r1, gamma);
    z.e = frame_trig<Z>(c2, c0, c1);
    z.p = frame_trig<Z>(c0, c1, c2);
    return z;
}

// Face-centre data of a face touching a Zone_1 cell: |G|, |G|^(2 gamma), frame choice and trig.
template <typename Z> struct Zone1Face { Z rr, prr, rr2, yrr, yrr2, yprr; bool equatorial; Trig<Z> t; };

template <typename Z>
__device__ __forceinline__ Zone1Face<Z> zone1_face(const double G[3], double gamma) {
    const Z g0 = (Z)G[0], g1 = (Z)G[1], g2 = (Z)G[2];
    Zone1Face<Z> f;
    f.rr = sqrt_t<Z>((g0 * g0 + g1 * g1) + g2 * g2);
    f.prr = pow_2gamma<Z>(f.rr, gamma);
    // reciprocals of the face's radial scales, shared by the two cells of the face
    f.rr2 = f.rr * f.rr;
    f.yrr = Z(1) / f.rr;
    f.yrr2 = Z(1) / f.rr2;
    f.yprr = Z(1) / f.prr;
    // the reference's frame test pi/9 < acos(G2/|G|) < 8 pi/9 (Setka_gran_snos.cpp:771-773), as
    // |G2/|G|| < cos(pi/9); either frame gives the same rescaled state up to rounding
    const Z ct = div_by_t<Z>(g2, f.rr, f.yrr);
    const Z c9 = Z(0.93969262078590838405);
    f.equatorial = (ct < c9 && ct > -c9);
    f.t = f.equatorial ? frame_trig<Z>(g2, g0, g1) : frame_trig<Z>(g0, g1, g2);
    return f;
}

// Setka_gran_snos.cpp:744-873 with precomputed trig: rotate V and B of a cell state to the face
// centre (B radial x (r1/rr)^2, tangential x r1/rr), rho x r1^2/rr^2, p x (r1/rr)^(2 gamma).
template <typename Z>
__device__ __forceinline__ void zone1_apply(const Z s[8], const Zone1Cell<Z> &zc, const Zone1Face<Z> &zf, Z out[8]) {
    const Trig<Z> a = zf.equatorial ? zc.e : zc.p;
    const Trig<Z> b = zf.t;
#pragma unroll
    for (int ik = 0; ik < 2; ik++) {
        const Z *V = s + (ik == 0 ? 1 : 5);
        Z *O = out + (ik == 0 ? 1 : 5);
        // axis roles (z, x, y)
        const Z Vz = zf.equatorial ? V[2] : V[0];
        const Z Vx = zf.equatorial ? V[0] : V[1];
        const Z Vy = zf.equatorial ? V[1] : V[2];
        Z Vr = Vx * a.st * a.cp + Vy * a.st * a.sp + Vz * a.ct;
        Z Vtheta = Vx * a.ct * a.cp + Vy * a.ct * a.sp - Vz * a.st;
        Z Vphi = -Vx * a.sp + Vy * a.cp;
        if (ik == 1) {
            const Z q = div_by_t<Z>(zc.r1, zf.rr, zf.yrr);
            Vr = Vr * (q * q);
            Vphi = Vphi * q;
            Vtheta = Vtheta * q;
        }
        const Z ox = Vr * b.st * b.cp + Vtheta * b.ct * b.cp - Vphi * b.sp;
        const Z oy = Vr * b.st * b.sp + Vtheta * b.ct * b.sp + Vphi * b.cp;
        const Z oz = Vr * b.ct - Vtheta * b.st;
        if (zf.equatorial) { O[2] = oz; O[0] = ox; O[1] = oy; }
        else { O[0] = oz; O[1] = ox; O[2] = oy; }
    }
    out[0] = div_by_t<Z>(s[0] * (zc.r1 * zc.r1), zf.rr2, zf.yrr2);
    out[4] = div_by_t<Z>(s[4] * zc.pr1, zf.prr, zf.yprr);
}

}  // namespace titan
