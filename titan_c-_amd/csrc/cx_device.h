// cx_device.h -- exact FP64 Maxwellian charge-exchange source (device).
//
// Rate integral of MK_source/MK_source/MK_source/MK_source.cpp:38-103 (midpoint rule in
// R0 in [0, 5c], alpha in [0, pi]) extended by the first and third velocity moments, with the
// Lindsay & Stebbings (2005) cross-section.  For Maxwellian protons (V_p, c_p) and atoms
// (V_H, c_H) the relative velocity is Maxwellian about dV = V_p - V_H with c^2 = c_p^2 + c_H^2:
//   J0 = <g s(g)>, J1 = <g s(g) g.e>, J2 = <g^3 s(g)>, e = dV/|dV|, U = |dV|
//   rate = n_p n_H J0;  S_m = -n_p n_H J1 e;
//   S_E  = -(n_p n_H / 2) [ (V_p + V_H).e J1 + kappa (J2 - U J1) ],  kappa = (c_p^2 - c_H^2)/c^2.
#pragma once
#include <hip/hip_runtime.h>

namespace titan {

// sigma_LS(E) = (4.15 - 0.531 ln E)^2 (1 - exp(-67.3/E))^4.5 1e-16 cm^2, E = m_p g^2/2 in keV,
// times n_char L_char (0.06 cm^-3 x 4.2113 AU) so that rates are per unit code time.
// Without contraction in every unit that includes it (the surrogate's closed-form estimate calls
// it from units built with contraction on).
__device__ __forceinline__ double sigma_ls(double g) {
#pragma clang fp contract(off)
    const double e_per_g2 = 0.5 * 1.67262192369e-24 * (1.03804e6 * 1.03804e6) / 1.602176634e-9;
    const double nl = 0.06 * 4.2113 * 1.495978707e13;
    double E = e_per_g2 * g * g;
    if (E < 1e-12) E = 1e-12;
    const double a = 4.15 - 0.531 * log(E);
    const double b = 1.0 - exp(-67.3 / E);
    return a * a * pow(b, 4.5) * 1e-16 * nl;
}

// Relative-velocity frame of one (proton, atom) pair.
struct CxFrame {
    double U, c2, c, e0, e1, e2, dR0, dal;
};

__device__ __forceinline__ CxFrame cx_frame(double cp2, const double vp[3], double ch2, const double vh[3], int N1,
                                            int N2) {
    const double kPiL = 3.14159265358979323846;
    const double dv0 = vp[0] - vh[0], dv1 = vp[1] - vh[1], dv2 = vp[2] - vh[2];
    CxFrame f;
    f.U = sqrt((dv0 * dv0 + dv1 * dv1) + dv2 * dv2);
    f.c2 = cp2 + ch2;
    f.c = sqrt(f.c2);
    f.e0 = 1.0; f.e1 = 0.0; f.e2 = 0.0;
    if (f.U > 0.0) { f.e0 = dv0 / f.U; f.e1 = dv1 / f.U; f.e2 = dv2 / f.U; }
    f.dal = kPiL / N2;
    f.dR0 = 5.0 * f.c / N1;
    return f;
}

// Radial weight R0^2 exp(-R0^2/c^2) of radial node ii.
__device__ __forceinline__ double cx_radial(const CxFrame &f, int ii, double &R0) {
    R0 = (ii + 0.5) * f.dR0;
    return R0 * R0 * exp(-(R0 * R0) / f.c2);
}

// Integrand g sigma(g) sin(alpha) and its g_par, g^2 moments at (R0, alpha node jj).
__device__ __forceinline__ void cx_point(const CxFrame &f, double R0, int jj, const double *sc, double &gs,
                                         double &gpar, double &g2) {
    const double sa = sc[2 * jj], ca = sc[2 * jj + 1];
    gpar = f.U - R0 * ca;
    const double gper = R0 * sa;
    g2 = gpar * gpar + gper * gper;
    const double g = sqrt(g2);
    gs = g * sigma_ls(g) * sa;
}

__device__ __forceinline__ void cx_finish(const CxFrame &f, double J0, double J1, double J2, double n_p, double cp2,
                                          const double vp[3], double n_h, double ch2, const double vh[3],
                                          double out[5]) {
    const double kPiL = 3.14159265358979323846;
    const double norm = 2.0 * kPiL * f.dal * f.dR0 / (kPiL * sqrt(kPiL) * f.c2 * f.c);
    J0 *= norm; J1 *= norm; J2 *= norm;
    const double nn = n_p * n_h;
    const double kappa = (cp2 - ch2) / f.c2;
    const double sv = (vp[0] + vh[0]) * f.e0 + (vp[1] + vh[1]) * f.e1 + (vp[2] + vh[2]) * f.e2;
    out[0] = nn * J0;
    out[1] = -nn * J1 * f.e0;
    out[2] = -nn * J1 * f.e1;
    out[3] = -nn * J1 * f.e2;
    out[4] = -0.5 * nn * (sv * J1 + kappa * (J2 - f.U * J1));
}

// One state by one thread; sc[2*j] = sin, sc[2*j+1] = cos of the j-th alpha node (N2 nodes).
__device__ __forceinline__ void cx_source_one(double n_p, double cp2, const double vp[3], double n_h, double ch2,
                                              const double vh[3], int N1, int N2, const double *sc,
                                              double out[5]) {
    const CxFrame f = cx_frame(cp2, vp, ch2, vh, N1, N2);
    double J0 = 0.0, J1 = 0.0, J2 = 0.0;
    for (int ii = 0; ii < N1; ++ii) {
        double R0;
        const double wr = cx_radial(f, ii, R0);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int jj = 0; jj < N2; ++jj) {
            double gs, gpar, g2;
            cx_point(f, R0, jj, sc, gs, gpar, g2);
            s0 += gs;
            s1 += gs * gpar;
            s2 += gs * g2;
        }
        J0 += wr * s0;
        J1 += wr * s1;
        J2 += wr * s2;
    }
    cx_finish(f, J0, J1, J2, n_p, cp2, vp, n_h, ch2, vh, out);
}

// One state by a whole workgroup: thread t takes the nodes p = t, t + blockDim, ... of the
// N1 x N2 grid (p = ii*N2 + jj), then a fixed-order tree sum over the threads (deterministic).
// red: 3*blockDim doubles of LDS, blockDim a power of two.  Every thread of the block must call
// it; the result is valid in thread 0.
__device__ __forceinline__ void cx_source_block(double n_p, double cp2, const double vp[3], double n_h, double ch2,
                                                const double vh[3], int N1, int N2, const double *sc, double *red,
                                                double out[5]) {
    const CxFrame f = cx_frame(cp2, vp, ch2, vh, N1, N2);
    double J0 = 0.0, J1 = 0.0, J2 = 0.0;
    for (int p = threadIdx.x; p < N1 * N2; p += blockDim.x) {
        const int ii = p / N2, jj = p - ii * N2;
        double R0;
        const double wr = cx_radial(f, ii, R0);
        double gs, gpar, g2;
        cx_point(f, R0, jj, sc, gs, gpar, g2);
        const double w = wr * gs;
        J0 += w;
        J1 += w * gpar;
        J2 += w * g2;
    }
    const int t = threadIdx.x, nt = blockDim.x;
    red[t] = J0; red[nt + t] = J1; red[2 * nt + t] = J2;
    __syncthreads();
    for (int h = nt / 2; h > 0; h >>= 1) {
        if (t < h) {
            red[t] += red[t + h];
            red[nt + t] += red[nt + t + h];
            red[2 * nt + t] += red[2 * nt + t + h];
        }
        __syncthreads();
    }
    if (t == 0) cx_finish(f, red[0], red[nt], red[2 * nt], n_p, cp2, vp, n_h, ch2, vh, out);
}

// Fills sc[0..2*N2) cooperatively (block-wide); caller syncs.
__device__ __forceinline__ void cx_fill_angles(double *sc, int N2) {
    const double kPiL = 3.14159265358979323846;
    for (int j = threadIdx.x; j < N2; j += blockDim.x) {
        const double al = (j + 0.5) * (kPiL / N2);
        sc[2 * j] = sin(al);
        sc[2 * j + 1] = cos(al);
    }
}

}  // namespace titan
