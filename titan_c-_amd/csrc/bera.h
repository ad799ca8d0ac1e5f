// bera.h -- plasma part of Setka::Calc_sourse_MF_Bera (Setka_physics.cpp:746-1258): the momentum and
// energy the protons exchange with n_h hydrogen fluids by charge exchange, from the closed-form
// Maxwellian-averaged relative speeds U, U_m, U_E and sigma = (1 - a_2 ln U)^2.
//
// The proton split of Plasma_components_1 (Phys_param.cpp:628-665) reads rho_He and a zone: the
// zone of Setka::determ_zone (Setka_physics.cpp:1427-1482; 1 or 2 by cell type where Q/rho < 50,
// else 4 for a Zone_4 cell and 3 otherwise), passed on as zone - 1 (Setka_physics.cpp:769,783), so
// al = 2 unless zone - 1 > 2.  A state without the passive scalars reads Q = rho_He = 0 (zone 1
// or 2, al = 2).  The plasma mass source is 0 (charge exchange swaps a proton for a proton).  The
// hydrogen fluids' sources are formed on request; pickup ions are not carried.  Evaluated in the
// reference's operation order (FP64, no contraction in the translation units that include it);
// only log and sqrt differ from glibc in the last ulp.
#pragma once
#include <hip/hip_runtime.h>

namespace titan {

constexpr double kMeMp = 1.0 / 1836.15;  // MF_meDmp, Header.h:99
constexpr double kBeraPi = 3.14159265358979323846;

// |V1 - V2| as Eigen's norm(): sqrt((d0^2 + d1^2) + d2^2)
__device__ __forceinline__ double bera_dnorm(const double *V1, const double *V2) {
    const double d0 = V1[0] - V2[0], d1 = V1[1] - V2[1], d2 = V1[2] - V2[2];
    return sqrt((d0 * d0 + d1 * d1) + d2 * d2);
}

// U_bera, Um_bera, UE_bera (Setka_physics.cpp:746-762)
__device__ __forceinline__ double bera_u(double T1, double T2, const double *V1, const double *V2) {
    const double uu = bera_dnorm(V1, V2);
    return sqrt((4.0 / kBeraPi) * (T1 + T2) + uu * uu);
}
__device__ __forceinline__ double bera_um(double T1, double T2, const double *V1, const double *V2) {
    const double uu = bera_dnorm(V1, V2);
    return sqrt((16.0 / kBeraPi) * T1 + (9.0 * kBeraPi / 4.0) * T2 + 4.0 * (uu * uu));
}
__device__ __forceinline__ double bera_ue(double T1, double T2, const double *V1, const double *V2) {
    const double uu = bera_dnorm(V1, V2);
    return sqrt((4.0 / kBeraPi) * T1 + (64.0 / 9.0 / kBeraPi) * T2 + uu * uu);
}

__device__ __forceinline__ double bera_dot(const double *a, const double *b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// Setka::determ_zone (Setka_physics.cpp:1427-1482): 1..4 from Q/rho and the cell type (1..4).
__host__ __device__ inline int bera_zone(double Q, double rho, int cell_type) {
    if (Q / rho < 50.0) return cell_type == 1 ? 1 : 2;
    return cell_type == 4 ? 4 : 3;
}

constexpr int kBeraMaxH = 16;  // hydrogen fluids the fluid sources take

// s: plasma state (rho, V, p, B); rho_He, zone (bera_zone); h[5k..5k+4] = (rho_Hk, p_Hk, V_Hk);
// a2 = par_a_2; ddp = par_n_H_LISM / par_Kn.  out = (rho, m_x, m_y, m_z, E) sources per unit volume
// and time.  hout (nullable, n_h <= kBeraMaxH): the hydrogen fluids' (rho, m_x, m_y, m_z, E) --
// each fluid's losses to the protons (Setka_physics.cpp:1012-1036), then the atoms made from fluid
// k added to fluid arise[(zone-1)*n_h + k] (1-based; null: fluid `zone`, the run without pickup
// ions, Phys_param.cpp:149-171) in fluid order (:1043-1094), all times ddp and over n_h_lism
// (:1236-1250).
template <bool H = false>
__device__ __forceinline__ void bera_source(const double s[8], double rho_He, int zone, int n_h,
                                            const double *__restrict__ h, double a2, double ddp, double out[5],
                                            double *__restrict__ hout = nullptr,
                                            const int8_t *__restrict__ arise = nullptr, double n_h_lism = 1.0) {
    double gain[H ? 5 * kBeraMaxH : 1];  // the atoms made from each fluid, added after every loss
    // Plasma_components_1 (Phys_param.cpp:641-660) at zone - 1
    const double rho = s[0], pr = s[4];
    const int al = (zone - 1 <= 2) ? 2 : 1;
    double rho_p = -(kMeMp * al * rho_He + 4.0 * (-rho + rho_He)) / (4.0 * (1.0 + kMeMp));
    double T_p = 8.0 * (1.0 + kMeMp) * pr / (8.0 * rho - (7.0 - al + (-1.0 + al) * kMeMp) * rho_He);
    const double Vp[3] = {s[1], s[2], s[3]};
    // Setka_physics.cpp:802-808
    if (rho_p <= 1e-8) rho_p = 1e-8;
    if (T_p <= 1e-8) T_p = 1e-8;
    if (T_p > 1e5) T_p = 1e5;
    double mx = 0.0, my = 0.0, mz = 0.0, E = 0.0;
    for (int k = 0; k < n_h; k++) {
        // :820-826
        const double rho_H = h[5 * k];
        const double T_H = 2.0 * h[5 * k + 1] / ((rho_H < 1e-9) ? 1e-9 : rho_H);
        const double VH[3] = {h[5 * k + 2], h[5 * k + 3], h[5 * k + 4]};
        // :842-859, both orders of the pair
        const double U_pH = bera_u(T_p, T_H, Vp, VH), U_Hp = bera_u(T_H, T_p, VH, Vp);
        const double Um_pH = bera_um(T_p, T_H, Vp, VH), Um_Hp = bera_um(T_H, T_p, VH, Vp);
        const double UE_pH = bera_ue(T_p, T_H, Vp, VH), UE_Hp = bera_ue(T_H, T_p, VH, Vp);
        const double l_pH = 1.0 - a2 * log(U_pH), l_Hp = 1.0 - a2 * log(U_Hp);
        const double sig_pH = l_pH * l_pH, sig_Hp = l_Hp * l_Hp;
        // :913-927
        const double cHp = sig_Hp * rho_p * rho_H, cpH = sig_pH * rho_p * rho_H;
        const double dV[3] = {Vp[0] - VH[0], Vp[1] - VH[1], Vp[2] - VH[2]};
        const double tHp = T_p / Um_Hp, tpH = T_H / Um_pH;
        double Hm_Hp[3], Hm_pH[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            Hm_Hp[i] = cHp * (U_Hp * Vp[i] + tHp * dV[i]);
            Hm_pH[i] = cpH * (U_pH * VH[i] - tpH * dV[i]);
        }
        const double HE_Hp = cHp * ((0.5 * U_Hp * bera_dot(Vp, Vp) + tHp * bera_dot(Vp, dV)) + 3.0 / 4.0 * T_p * UE_Hp);
        const double HE_pH = cpH * ((0.5 * U_pH * bera_dot(VH, VH) - tpH * bera_dot(VH, dV)) + 3.0 / 4.0 * T_H * UE_pH);
        // :975-983
        mx += (-Hm_Hp[0] + Hm_pH[0]);
        my += (-Hm_Hp[1] + Hm_pH[1]);
        mz += (-Hm_Hp[2] + Hm_pH[2]);
        E += (-HE_Hp + HE_pH);
        if constexpr (H) {
            // :1012-1036: the fluid's losses Hrho[p+H], Hm[p+H], HE[p+H]; the atoms made from it
            // carry Hrho[H+p], Hm[H+p], HE[H+p]
            double *o = hout + 5 * k;
            o[0] = 0.0 + (-(cpH * U_pH));
            o[1] = 0.0 + (-Hm_pH[0]);
            o[2] = 0.0 + (-Hm_pH[1]);
            o[3] = 0.0 + (-Hm_pH[2]);
            o[4] = 0.0 + (-HE_pH);
            double *gk = gain + 5 * k;
            gk[0] = cHp * U_Hp;
            gk[1] = Hm_Hp[0];
            gk[2] = Hm_Hp[1];
            gk[3] = Hm_Hp[2];
            gk[4] = HE_Hp;
        }
    }
    // :1236-1240
    out[0] = 0.0 * ddp;
    out[1] = mx * ddp;
    out[2] = my * ddp;
    out[3] = mz * ddp;
    out[4] = E * ddp;
    if constexpr (H) {
        for (int k = 0; k < n_h; k++) {  // :1043-1094, rows of the arise table in order
            const int t = (arise ? arise[(zone - 1) * n_h + k] : zone) - 1;
            if (t < 0 || t >= n_h) continue;
#pragma unroll
            for (int i = 0; i < 5; i++) hout[5 * t + i] += gain[5 * k + i];
        }
        for (int k = 0; k < 5 * n_h; k++) hout[k] = hout[k] * ddp / n_h_lism;  // :1236-1250
    }
}

}  // namespace titan
