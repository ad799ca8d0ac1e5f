// riemann_mixed.h -- the mixed-precision form of Phys_param::chlld (Phys_param.cpp:882-1878).
//
// Splits every region's flux of the reference into an upwind part and wave corrections:
//   region left of S_L              F_L - w U_L
//   left star (HLL, HLLC, HLLD)     F_L - w U_L + (S_L - w)(U*_L - U_L)
//   left double star (HLLD)         ... + (S*_L - w)(U**_L - U*_L)
// (and the mirror on the right), which equals the reference's F_L + S_L(U*_L - U_L) - w U*_L etc.
// exactly.  The frame rotation, the upwind part F - wU, the state jumps (right minus left), the
// normal-field flux and the rotation back are FP64; the wave speeds, the region choice and the
// corrections are FP32, each correction formed in closed form from the jumps -- e.g.
// rho*_L - rho_L = rho_L (S_M - v_L)/(S_L - S_M) with S_M - v_L = (su_R rho_R dv - dpt)/D -- so
// its rounding error is relative to the jump across the face, not to the state.  A supersonic
// face is therefore FP64-exact, and the energy error no longer grows with the Mach number
// squared in cold supersonic flow, where p is a small remainder of the total energy.
//
// Region choices follow the reference's comparisons in the same order, on the FP32 speeds.  Where
// the HLLD Alfven denominator rho su (S - S_M) - Bn^2 cancels (field nearly along the normal),
// its 1/tt would amplify the FP32 rounding of the speeds: the solver then reports the face as
// ill-conditioned (out.ok = -1) and the caller solves it with chlld<double>.
#pragma once
#include "riemann.h"

namespace titan {

// The FP64 solve of a face the mixed solver reports as ill-conditioned.  Kept out of line: it is
// rare, and inlined it would raise the register allocation of every caller.
template <int K>
__device__ __noinline__ void chlld_fp64_face(int n_state, int n_disc, int flags, double al, double be, double ge,
                                             double w, const double *qL, const double *qR, const double *kL,
                                             const double *kR, RiemannOut<double, K> *out) {
    chlld<double, K>(n_state, n_disc, flags, al, be, ge, w, qL, qR, kL, kR, *out);
}

// Cancellation limit of the HLLD Alfven denominators tt = rho su (S - S_M) - Bn^2 (|tt| below
// this fraction of rho su (S - S_M) sends the face to the FP64 solver).
constexpr float kMixedGuard = 0.1f;

// out.ok: 1 solved, 0 the reference's fatal branch, -1 ill-conditioned in FP32: the caller solves
// the face with chlld<double> instead.
template <int K>
__device__ __forceinline__ void chlld_mixed(int n_state, int n_disc, int flags, double al, double be, double ge,
                                            double w, const double qL[8], const double qR[8], const double *kL,
                                            const double *kR, RiemannOut<double, K> &out) {
    const double spi4 = 3.5449077018110318;
    const double gamma_d = 5.0 / 3.0;
    const double g1d = gamma_d - 1.0;
    const float gamma = 5.0f / 3.0f;
    const float epsb = 1e-4f;
    const bool left_ydar = flags & 1, contact = (flags >> 1) & 1;

    // ---- FP64: frame, rotated states, upwind fluxes and jumps -------------------------------
    double n[3] = {al, be, ge}, t[3], m[3];
    get_bazis<double>(n, t, m);
    const double r1 = qL[0], p1 = qL[4], r2 = qR[0], p2 = qR[4];
    const double bx1 = qL[5] / spi4, by1 = qL[6] / spi4, bz1 = qL[7] / spi4;
    const double bx2 = qR[5] / spi4, by2 = qR[6] / spi4, bz2 = qR[7] / spi4;
    double vL[3], vR[3], bL[3], bR[3];
    vL[0] = spf(n[0], n[1], n[2], qL[1], qL[2], qL[3]);
    vL[1] = spf(t[0], t[1], t[2], qL[1], qL[2], qL[3]);
    vL[2] = spf(m[0], m[1], m[2], qL[1], qL[2], qL[3]);
    vR[0] = spf(n[0], n[1], n[2], qR[1], qR[2], qR[3]);
    vR[1] = spf(t[0], t[1], t[2], qR[1], qR[2], qR[3]);
    vR[2] = spf(m[0], m[1], m[2], qR[1], qR[2], qR[3]);
    bL[0] = spf(n[0], n[1], n[2], bx1, by1, bz1);
    bL[1] = spf(t[0], t[1], t[2], bx1, by1, bz1);
    bL[2] = spf(m[0], m[1], m[2], bx1, by1, bz1);
    bR[0] = spf(n[0], n[1], n[2], bx2, by2, bz2);
    bR[1] = spf(t[0], t[1], t[2], bx2, by2, bz2);
    bR[2] = spf(m[0], m[1], m[2], bx2, by2, bz2);
    const double b2Ld = (bL[0] * bL[0] + bL[1] * bL[1]) + bL[2] * bL[2];
    const double b2Rd = (bR[0] * bR[0] + bR[1] * bR[1]) + bR[2] * bR[2];
    const double ptLd = p1 + b2Ld / 2.0, ptRd = p2 + b2Rd / 2.0;
    const double sbv1d = (bL[0] * vL[0] + bL[1] * vL[1]) + bL[2] * vL[2];
    const double sbv2d = (bR[0] * vR[0] + bR[1] * vR[1]) + bR[2] * vR[2];
    const double e1d = p1 / g1d + r1 * ((vL[0] * vL[0] + vL[1] * vL[1]) + vL[2] * vL[2]) / 2.0 + b2Ld / 2.0;
    const double e2d = p2 / g1d + r2 * ((vR[0] * vR[0] + vR[1] * vR[1]) + vR[2] * vR[2]) / 2.0 + b2Rd / 2.0;
    auto fl = [&](int k) -> double {
        switch (k) {
        case 0: return r1 * vL[0];
        case 1: return r1 * vL[0] * vL[0] + ptLd - bL[0] * bL[0];
        case 2: return r1 * vL[0] * vL[1] - bL[0] * bL[1];
        case 3: return r1 * vL[0] * vL[2] - bL[0] * bL[2];
        case 4: return (e1d + ptLd) * vL[0] - bL[0] * sbv1d;
        case 5: return 0.0;
        case 6: return vL[0] * bL[1] - vL[1] * bL[0];
        default: return vL[0] * bL[2] - vL[2] * bL[0];
        }
    };
    auto fr = [&](int k) -> double {
        switch (k) {
        case 0: return r2 * vR[0];
        case 1: return r2 * vR[0] * vR[0] + ptRd - bR[0] * bR[0];
        case 2: return r2 * vR[0] * vR[1] - bR[0] * bR[1];
        case 3: return r2 * vR[0] * vR[2] - bR[0] * bR[2];
        case 4: return (e2d + ptRd) * vR[0] - bR[0] * sbv2d;
        case 5: return 0.0;
        case 6: return vR[0] * bR[1] - vR[1] * bR[0];
        default: return vR[0] * bR[2] - vR[2] * bR[0];
        }
    };
    auto ul = [&](int k) -> double {
        switch (k) {
        case 0: return r1;
        case 4: return e1d;
        case 1: case 2: case 3: return r1 * vL[k - 1];
        default: return bL[k - 5];
        }
    };
    auto ur = [&](int k) -> double {
        switch (k) {
        case 0: return r2;
        case 4: return e2d;
        case 1: case 2: case 3: return r2 * vR[k - 1];
        default: return bR[k - 5];
        }
    };

    // FP32 copies of the states and FP64-formed jumps
    const float fr1 = (float)r1, fr2 = (float)r2, fp1 = (float)p1, fp2 = (float)p2;
    float fvL[3], fvR[3], fbL[3], fbR[3], dv[3], db[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        fvL[i] = (float)vL[i]; fvR[i] = (float)vR[i]; fbL[i] = (float)bL[i]; fbR[i] = (float)bR[i];
        dv[i] = (float)(vR[i] - vL[i]);
        db[i] = (float)(bR[i] - bL[i]);
    }
    const float dpt = (float)(ptRd - ptLd);
    // the side energies less their kinetic part (p / (gamma - 1) + b^2 / 2): what the star jumps
    // add to the kinetic energy the mass correction carries, which is formed in FP64 (kin0 below)
    const float fei1 = (float)(e1d - r1 * (0.5 * ((vL[0] * vL[0] + vL[1] * vL[1]) + vL[2] * vL[2]))),
                fei2 = (float)(e2d - r2 * (0.5 * ((vR[0] * vR[0] + vR[1] * vR[1]) + vR[2] * vR[2])));
    const float b2L = (float)b2Ld, b2R = (float)b2Rd;
    const float ptL = (float)ptLd, ptR = (float)ptRd;
    const float wv0 = (float)w;

    // ---- FP32: wave speeds (Phys_param.cpp:978-1046) -----------------------------------------
    // b^2/rho + c (c -+ 2 b_n/sqrt(rho)) = (c -+ b_n/sqrt(rho))^2 + (b^2 - b_n^2)/rho is never negative,
    // but with the field along the normal and the sound speed at the Alfven speed it is a remainder
    // that FP32 rounding can take below zero: it is clamped at zero before the square root (in FP64,
    // where the reference takes it as it comes, such a state keeps its positive remainder).
    const float ro = (fr2 + fr1) / 2.0f;
    const float ap = (fp2 + fp1) / 2.0f;
    const float aaL = fbL[0] / sqrtf(fr1);
    const float b21 = b2L / fr1;
    const float cL = sqrtf(gamma * fp1 / fr1);
    float qp = sqrtf(fmaxf(b21 + cL * (cL + 2.0f * aaL), 0.0f));
    float qm = sqrtf(fmaxf(b21 + cL * (cL - 2.0f * aaL), 0.0f));
    const float cfL = (qp + qm) / 2.0f;
    const float aaR = fbR[0] / sqrtf(fr2);
    const float b22 = b2R / fr2;
    const float cR = sqrtf(gamma * fp2 / fr2);
    qp = sqrtf(fmaxf(b22 + cR * (cR + 2.0f * aaR), 0.0f));
    qm = sqrtf(fmaxf(b22 + cR * (cR - 2.0f * aaR), 0.0f));
    const float cfR = (qp + qm) / 2.0f;
    const float aC = (aaL + aaR) / 2.0f;
    const float b2o = (b22 + b21) / 2.0f;
    const float cC = sqrtf(gamma * ap / ro);
    qp = sqrtf(fmaxf(b2o + cC * (cC + 2.0f * aC), 0.0f));
    qm = sqrtf(fmaxf(b2o + cC * (cC - 2.0f * aC), 0.0f));
    const float cfC = (qp + qm) / 2.0f;

    // S_L, S_R relative to each side's normal velocity (su = S - v), formed from the fast speeds
    // and the velocity jump so that no large velocity cancels (Phys_param.cpp:1008-1037)
    float suL = 0.0f, suR = 0.0f;
    if (left_ydar) n_disc = 2;
    const float h = dv[0] / 2.0f;  // (v_L + v_R)/2 - v_L
    if (n_disc == 0) {
        suL = smin(-cfL, dv[0] - cfR);
        suR = smax(cfL - dv[0], cfR);
    } else if (n_disc == 1) {
        suL = smin(-cfL, h - cfC);
        suR = smax(cfR, cfC - h);
    } else if (n_disc == 2) {
        suL = 0.6f * smin(-cfL, h - cfC) + 0.4f * smin(-cfL, dv[0] - cfR);
        suR = 0.6f * smax(cfR, cfC - h) + 0.4f * smax(cfL - dv[0], cfR);
    } else if (n_disc == 3) {
        suL = smin(0.0f, dv[0]) - smax(cfL, cfR);
        suR = smax(-dv[0], 0.0f) + smax(cfL, cfR);
    }
    float SL = fvL[0] + suL, SR = fvR[0] + suR;
    // the same speeds with FP64 velocities: the face velocity of the upwind part when the
    // reference moves the face with S_L (left_ydar), and the normal-field dissipation speed
    double SLd = vL[0] + (double)suL, SRd = vR[0] + (double)suR;
    out.dsr = (double)SR;
    out.dsl = (double)SL;
    const float D = suR * fr2 - suL * fr1;
    // S_M relative to each side's normal velocity, from the jumps
    const float dSML = (suR * fr2 * dv[0] - dpt) / D;  // S_M - v_L
    const float dSMR = (suL * fr1 * dv[0] - dpt) / D;  // S_M - v_R
    const float SM = fvL[0] + dSML;
    const float gL = suL - dSML, gR = suR - dSMR;  // S_L - S_M, S_R - S_M
    out.dsc = (double)SM;
    // wv: the face velocity of the region tests (FP32); wd: the same in FP64 for the upwind part
    float wv = wv0;
    double wd = w;
    if (contact) { wv = SM; wd = vL[0] + (double)dSML; }
    if (n_state == 0) {
        const float TR0 = fabsf(fvL[0] + fvR[0]) / 2.0f + cfC;
        SR = TR0;
        SL = -TR0;
        SLd = -(fabs(vL[0] + vR[0]) / 2.0 + (double)cfC);
        SRd = -SLd;
    }

    // HLL state minus a side's state: U_hll - U_L = (S_R dU - dF)/(S_R - S_L), U_hll - U_R =
    // (S_L dU - dF)/(S_R - S_L), dU and dF formed in FP64
    auto dUF = [&](int k, float &dU, float &dF) {
        dU = (float)(ur(k) - ul(k));
        dF = (float)(fr(k) - fl(k));
    };

    double base[8];
    float corr[8];
#pragma unroll
    for (int k = 0; k < 8; k++) corr[k] = 0.0f;
    float kcorr[K > 0 ? K : 1];
#pragma unroll
    for (int i = 0; i < (K > 0 ? K : 1); i++) kcorr[i] = 0.0f;
    bool right = false;  // upwind side of the selected region
    // Star regions: corr[0] is the mass correction (S - w)(rho* - rho); the momentum and energy it
    // carries at the side's velocity, v corr[0] and |v|^2/2 corr[0], are added in FP64 and corr[1..4]
    // hold only the rest.  Formed in FP32 the full jumps would each round the tangential kinetic
    // energy of a hypersonic face separately, and the FP64 pressure recovery
    // p = (E - rho v^2/2 - B^2/8 pi)(gamma - 1) would see that rounding times the Mach number squared.
    bool split = false;
    int region = 0, ibn = 0;
    out.ok = 1;
    int wbn_side;  // -1 left normal field, +1 right, 0 average

    if (n_state <= 1) {  // Lax / HLL, Phys_param.cpp:1133-1257
        if (left_ydar) { wv = SL; wd = SLd; }
        if (SL > wv || left_ydar) region = 1;
        else if (SL <= wv && wv <= SR) region = 2;
        else if (SR < wv) { region = 3; right = true; }
        else { region = 0; out.ok = 0; }
        if (region == 2) {
            // S - w from the FP64 speeds: a face moving at nearly a wave speed (moving mesh) would
            // otherwise lose the difference to the FP32 rounding of S and w
            const float c = (float)((SLd - wd) / (SRd - SLd));
#pragma unroll
            for (int k = 0; k < 8; k++) {
                float dU, dF;
                dUF(k, dU, dF);
                corr[k] = c * (SR * dU - dF);
            }
#pragma unroll
            for (int i = 0; i < K; i++) {
                const float dU = (float)(kR[i] - kL[i]);
                const float dF = (float)(kR[i] * vR[0] - kL[i] * vL[0]);
                kcorr[i] = c * (SR * dU - dF);
            }
        }
        wbn_side = (wv >= SR) ? 1 : ((wv <= SL) ? -1 : 0);
    } else {
        // normal field of the HLL state (UZ[5]) relative to each side: S_R db/(S_R-S_L), S_L db/(S_R-S_L)
        const float dbnL = (float)SRd * db[0] / (SR - SL);
        const float bn = fbL[0] + dbnL;
        const float dbnR = dbnL - db[0];
        const float suRm = suR / gR;
        const float suLm = suL / gL;
        const float rzR = fr2 * suRm;
        const float rzL = fr1 * suLm;
        const float drL = fr1 * dSML / gL;  // rho*_L - rho_L
        const float drR = fr2 * dSMR / gR;  // rho*_R - rho_R
        // p_t* relative to each side (the reference's ptz)
        const float dptzL = suL * fr1 * (-dpt + fr2 * suR * dv[0]) / D;  // ptz - ptL
        const float dptzR = suR * fr2 * (-dpt + fr1 * suL * dv[0]) / D;  // ptz - ptR
        if (n_state == 2) {  // HLLC, Phys_param.cpp:1258-1475
            if (left_ydar) { wv = SL; wd = SLd; }
            if (SL >= wv || left_ydar) region = 1;
            else if (SL <= wv && SM >= wv) region = 2;
            else if (SM <= wv && SR > wv) { region = 3; right = true; }
            else if (SR <= wv) { region = 4; right = true; }
            else { region = 0; out.ok = 0; }
            if (region == 2 || region == 3) {
                const bool R = region == 3;
                const float rs = R ? fr2 : fr1, su = R ? suR : suL;
                const float dSM = R ? dSMR : dSML, dr = R ? drR : drL, rz = R ? rzR : rzL;
                const float dptz = R ? dptzR : dptzL, g = R ? gR : gL;
                const float *fv = R ? fvR : fvL, *fb = R ? fbR : fbL;
                const float ei = R ? fei2 : fei1, pt = R ? ptR : ptL;
                // HLL state relative to this side: dUZ[k] = U_hll - U_side
                float dUZ[8];
                const float cS = R ? SL : SR;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    float dU, dF;
                    dUF(k, dU, dF);
                    dUZ[k] = (cS * dU - dF) / (SR - SL);
                }
                // star state minus side state
                float dvt1, dvt2, dbt1, dbt2;
                if (fabsf(bn) <= epsb) {
                    dvt1 = 0.0f; dvt2 = 0.0f;
                    dbt1 = fb[1] * dSM / g;  // b (su/(S-SM) - 1)
                    dbt2 = fb[2] * dSM / g;
                } else {
                    dvt1 = bn * (-dUZ[6]) / su / rs;  // UZ5 (b - UZ6)/su/rho
                    dvt2 = bn * (-dUZ[7]) / su / rs;
                    dbt1 = dUZ[6];
                    dbt2 = dUZ[7];
                }
                // sbv_side - sbvz, sbvz = B_hll . (rho v)_hll / rho_hll
                const float UZ0 = rs + dUZ[0];
                float dvh[3];
#pragma unroll
                for (int i = 0; i < 3; i++) dvh[i] = (dUZ[1 + i] - dUZ[0] * fv[i]) / UZ0;
                const float dBh[3] = {R ? dbnR : dbnL, dUZ[6], dUZ[7]};
                float dsbv = 0.0f;  // sbv_side - sbvz
#pragma unroll
                for (int i = 0; i < 3; i++) dsbv -= fb[i] * dvh[i] + dBh[i] * (fv[i] + dvh[i]);
                // e*_side - e_side less |v|^2/2 (rho* - rho)
                const float dEi = (ei * dSM + pt * dSM + dptz * SM + bn * dsbv) / g;
                float dUs[8];
                dUs[0] = dr;
                dUs[1] = rz * dSM;  // (rho v_n)* - rho v_n less v_n dr
                dUs[2] = dvt1 * rz;
                dUs[3] = dvt2 * rz;
                dUs[4] = dEi;
                split = true;
                dUs[5] = R ? dbnR : dbnL;
                dUs[6] = dbt1;
                dUs[7] = dbt2;
                const float c = (float)((R ? SRd : SLd) - wd);  // S - w in FP64 (see HLL)
#pragma unroll
                for (int k = 0; k < 8; k++) corr[k] = c * dUs[k];
#pragma unroll
                for (int i = 0; i < K; i++) {
                    const float ks = (float)(R ? kR[i] : kL[i]);
                    kcorr[i] = c * (ks * dSM / (R ? gR : gL));
                }
            }
        } else {  // HLLD, Phys_param.cpp:1476-1870
            const float bn2 = bn * bn;
            float tvR, tvL;
            float dbzR1, dbzR2, dbzL1, dbzL2;  // B*_t - B_t
            const float aR = fr2 * suR * gR, ttR = aR - bn2;
            if (fabsf(ttR) < 0.000000001f) { tvR = 0.0f; dbzR1 = -fbR[1]; dbzR2 = -fbR[2]; }
            else {
                tvR = dSMR / ttR;
                const float c = fr2 * suR * dSMR / ttR;
                dbzR1 = fbR[1] * c; dbzR2 = fbR[2] * c;
            }
            const float aL = fr1 * suL * gL, ttL = aL - bn2;
            if (fabsf(ttL) < 0.000000001f) { tvL = 0.0f; dbzL1 = -fbL[1]; dbzL2 = -fbL[2]; }
            else {
                tvL = dSML / ttL;
                const float c = fr1 * suL * dSML / ttL;
                dbzL1 = fbL[1] * c; dbzL2 = fbL[2] * c;
            }
            const float dvzL1 = -bn * fbL[1] * tvL, dvzL2 = -bn * fbL[2] * tvL;  // v*_t - v_t
            const float dvzR1 = -bn * fbR[1] * tvR, dvzR2 = -bn * fbR[2] * tvR;
            const float vzL1 = fvL[1] + dvzL1, vzL2 = fvL[2] + dvzL2;
            const float vzR1 = fvR[1] + dvzR1, vzR2 = fvR[2] + dvzR2;
            const float bzL1 = fbL[1] + dbzL1, bzL2 = fbL[2] + dbzL2;
            const float bzR1 = fbR[1] + dbzR1, bzR2 = fbR[2] + dbzR2;
            // sbv - sbv* on each side (sbv* = bn SM + B*_t . v*_t)
            const float dsbL = -(fbL[0] * dSML + dbnL * fvL[0] + dbnL * dSML) -
                               (fbL[1] * dvzL1 + dbzL1 * fvL[1] + dbzL1 * dvzL1) -
                               (fbL[2] * dvzL2 + dbzL2 * fvL[2] + dbzL2 * dvzL2);
            const float dsbR = -(fbR[0] * dSMR + dbnR * fvR[0] + dbnR * dSMR) -
                               (fbR[1] * dvzR1 + dbzR1 * fvR[1] + dbzR1 * dvzR1) -
                               (fbR[2] * dvzR2 + dbzR2 * fvR[2] + dbzR2 * dvzR2);
            // e*_side - e_side less |v|^2/2 (rho* - rho)
            const float dEL = (fei1 * dSML + ptL * dSML + dptzL * SM + bn * dsbL) / gL;
            const float dER = (fei2 * dSMR + ptR * dSMR + dptzR * SM + bn * dsbR) / gR;
            const float rzRs = sqrtf(rzR);
            const float rzLs = sqrtf(rzL);
            const float rzss = rzRs + rzLs;
            const float rzps = rzRs * rzLs;
            float SZL = SM - fabsf(bn) / rzLs;
            float SZR = SM + fabsf(bn) / rzRs;
            float sbn;
            if (fabsf(bn) > epsb) { sbn = fabsf(bn) / bn; ibn = 1; }
            else { sbn = 0.0f; ibn = 0; SZL = SM; SZR = SM; }
            int j = -1;
            if (left_ydar) { wv = SL; wd = SLd; }
            if (SL >= wv || left_ydar) j = 1;
            else if (SL <= wv && SZL >= wv) j = 2;
            if (ibn == 1) {
                if (SZL <= wv && SM >= wv) j = 3;
                else if (SM <= wv && SZR >= wv) j = 4;
            }
            if (SZR <= wv && SR >= wv) j = 5;
            else if (SR <= wv) j = 6;
            if (j == -1) { out.ok = 0; j = 1; region = 0; }
            else region = j;
            right = j >= 4;
            // Alfven denominators that lost more than a digit to cancellation would carry the FP32
            // error of the wave speeds into the star states: such a face is solved in FP64
            if (j >= 2 && j <= 5 && (fabsf(ttL) < kMixedGuard * aL || fabsf(ttR) < kMixedGuard * aR)) {
                out.ok = -1;
                return;
            }
            if (j >= 2 && j <= 5) {
                split = true;
                const bool R = right;
                const float dSM = R ? dSMR : dSML, dr = R ? drR : drL, rz = R ? rzR : rzL;
                const float dUs[8] = {dr,
                                      rz * dSM,
                                      (R ? dvzR1 : dvzL1) * rz,
                                      (R ? dvzR2 : dvzL2) * rz,
                                      R ? dER : dEL,
                                      R ? dbnR : dbnL,
                                      R ? dbzR1 : dbzL1,
                                      R ? dbzR2 : dbzL2};
                const float c = (float)((R ? SRd : SLd) - wd);  // S - w in FP64 (see HLL)
#pragma unroll
                for (int k = 0; k < 8; k++) corr[k] = c * dUs[k];
#pragma unroll
                for (int i = 0; i < K; i++) {
                    const float ks = (float)(R ? kR[i] : kL[i]);
                    kcorr[i] = c * (ks * dSM / (R ? gR : gL));
                }
                if (j == 3 || j == 4) {
                    // double star minus star (density and normal momentum unchanged)
                    const float jv1 = dv[1] + dvzR1 - dvzL1, jv2 = dv[2] + dvzR2 - dvzL2;  // v*_tR - v*_tL
                    const float jb1 = db[1] + dbzR1 - dbzL1, jb2 = db[2] + dbzR2 - dbzL2;  // B*_tR - B*_tL
                    float d2v1, d2v2, d2b1, d2b2;  // U** - U* of this side (velocity, field)
                    if (!R) {
                        d2v1 = (rzRs * jv1 + sbn * jb1) / rzss;
                        d2v2 = (rzRs * jv2 + sbn * jb2) / rzss;
                        d2b1 = (rzLs * jb1 + sbn * rzps * jv1) / rzss;
                        d2b2 = (rzLs * jb2 + sbn * rzps * jv2) / rzss;
                    } else {
                        d2v1 = (-rzLs * jv1 + sbn * jb1) / rzss;
                        d2v2 = (-rzLs * jv2 + sbn * jb2) / rzss;
                        d2b1 = (-rzRs * jb1 + sbn * rzps * jv1) / rzss;
                        d2b2 = (-rzRs * jb2 + sbn * rzps * jv2) / rzss;
                    }
                    const float vz1 = R ? vzR1 : vzL1, vz2 = R ? vzR2 : vzL2;
                    const float bz1 = R ? bzR1 : bzL1, bz2 = R ? bzR2 : bzL2;
                    // sbv* - sbv** = -(B* . dv + dB . v* + dB . dv) over the tangential parts
                    const float dsb2 = -(bz1 * d2v1 + d2b1 * vz1 + d2b1 * d2v1) - (bz2 * d2v2 + d2b2 * vz2 + d2b2 * d2v2);
                    const float dE2 = R ? (rzRs * sbn * dsb2) : (-rzLs * sbn * dsb2);
                    // S*_side - w: S_M in FP64 (v_L + (S_M - v_L)) less the FP32 Alfven offset
                    const double SMd = vL[0] + (double)dSML;
                    const float c2 = (float)((R ? SMd + (double)(fabsf(bn) / rzRs) : SMd - (double)(fabsf(bn) / rzLs)) - wd);
                    corr[2] += c2 * (d2v1 * rz);
                    corr[3] += c2 * (d2v2 * rz);
                    corr[4] += c2 * dE2;
                    corr[6] += c2 * d2b1;
                    corr[7] += c2 * d2b2;
                }
            }
        }
        wbn_side = (wv > SR) ? 1 : ((wv < SL) ? -1 : 0);
    }

    // ---- FP64: upwind part, sum, normal-field flux, rotation back --------------------------
#pragma unroll
    for (int k = 0; k < 8; k++) base[k] = (right ? (fr(k) - wd * ur(k)) : (fl(k) - wd * ul(k))) + (double)corr[k];
    if (split) {  // momentum and kinetic energy of the mass correction at the side's velocity
        const double *vs = right ? vR : vL;
        const double kin0 = (double)corr[0];
        base[1] += vs[0] * kin0;
        base[2] += vs[1] * kin0;
        base[3] += vs[2] * kin0;
        base[4] += (0.5 * ((vs[0] * vs[0] + vs[1] * vs[1]) + vs[2] * vs[2])) * kin0;
    }
    double wbn;
    if (wbn_side > 0) wbn = wd * bR[0];
    else if (wbn_side < 0) wbn = wd * bL[0];
    else wbn = wd * (bL[0] + bR[0]) / 2.0;
    const double qv0 = base[1], qv1 = base[2], qv2 = base[3];
    // S_N from the FP64-velocity speeds
    const double qb0 = -smax(fabs(SLd), fabs(SRd)) * (bR[0] - bL[0]) - wbn, qb1 = base[6], qb2 = base[7];
    out.q[0] = base[0];
    out.q[4] = base[4];
#pragma unroll
    for (int ik = 0; ik < 3; ik++) {
        out.q[ik + 1] = n[ik] * qv0 + t[ik] * qv1 + m[ik] * qv2;
        out.q[ik + 5] = spi4 * (n[ik] * qb0 + t[ik] * qb1 + m[ik] * qb2);
    }
#pragma unroll
    for (int i = 0; i < K; i++) {
        const double fk = right ? (kR[i] * vR[0] - wd * kR[i]) : (kL[i] * vL[0] - wd * kL[i]);
        out.k[i] = fk + (double)kcorr[i];
    }
    out.branch = (n_state << 4) | (ibn << 3) | region;
    if (!out.ok) {
#pragma unroll
        for (int ik = 0; ik < 8; ik++) out.q[ik] = 0.0;
    }
}

}  // namespace titan
