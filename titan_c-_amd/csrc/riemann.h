// riemann.h -- device Riemann solvers (Lax / HLL / HLLC / HLLD) for a moving face.
//
// Restates Phys_param::chlld (TITAN/source/Phys_param.cpp:882-1878) and get_bazis
// (TITAN/source/Help_funk.cpp:219-232) for one face per lane.  Instantiated for T = double
// (compiled with -ffp-contract=off: every expression is evaluated in the reference's order,
// giving bit-identical fluxes and branch choices) and T = float (the mixed-precision step).
//
// Differences of form, not of result: the reference fills the intermediate states of every
// region and then picks one; here the region is chosen first with the same comparisons in the
// same order (later matching tests overwrite earlier ones, as the reference's chain of separate
// `if`s does) and only that region's formula is evaluated.  This keeps the HLLD working set in
// registers.
#pragma once
#include <hip/hip_runtime.h>
#include "fdiv.h"
#include "fsqrt.h"

namespace titan {

template <typename T> struct RConst;
template <> struct RConst<double> {
    static __device__ __forceinline__ double spi4() { return 3.5449077018110318; }  // sqrt(4*pi)
    static __device__ __forceinline__ double inv_spi4() { return 1.0 / 3.5449077018110318; }
};
template <> struct RConst<float> {
    static __device__ __forceinline__ float spi4() { return 3.5449077018110318f; }
    static __device__ __forceinline__ float inv_spi4() { return 1.0f / 3.5449077018110318f; }
};

// std::max / std::min semantics: the first argument wins on ties and NaN
template <typename T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }

template <typename T>
__device__ __forceinline__ T spf(T ax, T ay, T az, T bx, T by, T bz) {
    return ax * bx + ay * by + az * bz;
}

// get_bazis: t = normalize(a x n), a = x (or y when |n.x| > 0.9); m = n x t.
template <typename T>
__device__ __forceinline__ void get_bazis(const T n[3], T t[3], T m[3]) {
    T a0 = T(1), a1 = T(0), a2 = T(0);
    T nd = (n[0] * a0 + n[1] * a1) + n[2] * a2;
    if (fabs(nd) > T(0.9)) { a0 = T(0); a1 = T(1); a2 = T(0); }
    T c0 = a1 * n[2] - a2 * n[1];
    T c1 = a2 * n[0] - a0 * n[2];
    T c2 = a0 * n[1] - a1 * n[0];
    T z = (c0 * c0 + c1 * c1) + c2 * c2;
    if (z > T(0)) {  // z = n_y^2 + n_z^2 >= 0.19 or n_x^2 + n_z^2 >= 0.19 for a unit normal
        const T s = sqrt_t<T>(z), ys = T(1) / s;
        c0 = div_by_t<T>(c0, s, ys); c1 = div_by_t<T>(c1, s, ys); c2 = div_by_t<T>(c2, s, ys);
    }
    t[0] = c0; t[1] = c1; t[2] = c2;
    m[0] = n[1] * t[2] - n[2] * t[1];
    m[1] = n[2] * t[0] - n[0] * t[2];
    m[2] = n[0] * t[1] - n[1] * t[0];
}

// Result of one face solve.  branch = n_state<<4 | ibn<<3 | region; ok = 0 on the reference's
// fatal branches (NaN wave speeds or an impossible HLLD ordering).
template <typename T, int K>
struct RiemannOut {
    T q[8];
    T k[K > 0 ? K : 1];
    T dsl, dsc, dsr;
    int branch;
    int ok;
};

template <typename T, int K>
__device__ __forceinline__ void chlld(int n_state, int n_disc, int flags, T al, T be, T ge, T w,
                                      const T qL[8], const T qR[8], const T *kL, const T *kR,
                                      RiemannOut<T, K> &out) {
    // Quotients by a constant or by a denominator several of them share go through its correctly
    // rounded reciprocal (div_by_t, fdiv.h): the same correctly rounded results in fewer
    // instructions.  Every such denominator is a positive constant or positive by construction.
    const T spi4 = RConst<T>::spi4(), yspi4 = RConst<T>::inv_spi4();
    const T gamma = T(5.0 / 3.0);
    const T g1 = gamma - T(1), yg1 = T(1) / g1;
    const T epsb = T(1e-4);
    const bool left_ydar = flags & 1, contact = (flags >> 1) & 1;
    T wv = w;

    const T r1 = qL[0], u1 = qL[1], v1 = qL[2], w1 = qL[3], p1 = qL[4];
    const T bx1 = div_by_t<T>(qL[5], spi4, yspi4), by1 = div_by_t<T>(qL[6], spi4, yspi4),
            bz1 = div_by_t<T>(qL[7], spi4, yspi4);
    const T r2 = qR[0], u2 = qR[1], v2 = qR[2], w2 = qR[3], p2 = qR[4];
    const T bx2 = div_by_t<T>(qR[5], spi4, yspi4), by2 = div_by_t<T>(qR[6], spi4, yspi4),
            bz2 = div_by_t<T>(qR[7], spi4, yspi4);
    const T ro = (r2 + r1) / T(2);
    const T ap = (p2 + p1) / T(2);

    T n[3] = {al, be, ge}, t[3], m[3];
    get_bazis(n, t, m);

    T vL[3], vR[3], bL[3], bR[3];
    vL[0] = spf(n[0], n[1], n[2], u1, v1, w1);
    vL[1] = spf(t[0], t[1], t[2], u1, v1, w1);
    vL[2] = spf(m[0], m[1], m[2], u1, v1, w1);
    vR[0] = spf(n[0], n[1], n[2], u2, v2, w2);
    vR[1] = spf(t[0], t[1], t[2], u2, v2, w2);
    vR[2] = spf(m[0], m[1], m[2], u2, v2, w2);
    bL[0] = spf(n[0], n[1], n[2], bx1, by1, bz1);
    bL[1] = spf(t[0], t[1], t[2], bx1, by1, bz1);
    bL[2] = spf(m[0], m[1], m[2], bx1, by1, bz1);
    bR[0] = spf(n[0], n[1], n[2], bx2, by2, bz2);
    bR[1] = spf(t[0], t[1], t[2], bx2, by2, bz2);
    bR[2] = spf(m[0], m[1], m[2], bx2, by2, bz2);

    const T aaL = bL[0] / sqrt_t<T>(r1);
    const T b2L = (bL[0] * bL[0] + bL[1] * bL[1]) + bL[2] * bL[2];
    const T b21 = b2L / r1;
    const T cL = sqrt_t<T>(gamma * p1 / r1);
    T qp = sqrt_t<T>(b21 + cL * (cL + T(2) * aaL));
    T qm = sqrt_t<T>(b21 + cL * (cL - T(2) * aaL));
    const T cfL = (qp + qm) / T(2);
    const T ptL = p1 + b2L / T(2);

    const T aaR = bR[0] / sqrt_t<T>(r2);
    const T b2R = (bR[0] * bR[0] + bR[1] * bR[1]) + bR[2] * bR[2];
    const T b22 = b2R / r2;
    const T cR = sqrt_t<T>(gamma * p2 / r2);
    qp = sqrt_t<T>(b22 + cR * (cR + T(2) * aaR));
    qm = sqrt_t<T>(b22 + cR * (cR - T(2) * aaR));
    const T cfR = (qp + qm) / T(2);
    const T ptR = p2 + b2R / T(2);

    const T vC1 = (vL[0] + vR[0]) / T(2);

    T SL = T(0), SR = T(0);
    if (left_ydar) n_disc = 2;
    // the fast speed of the mean state enters only the n_disc = 1, 2 speed estimates and Lax's
    // (Phys_param.cpp:1008-1037,1053-1059); the step path (n_disc = 0, HLLC/HLLD) skips it
    T cfC = T(0);
    if (n_disc == 1 || n_disc == 2 || n_state == 0) {
        const T aC = (aaL + aaR) / T(2);
        const T b2o = (b22 + b21) / T(2);
        const T cC = sqrt_t<T>(gamma * ap / ro);
        qp = sqrt_t<T>(b2o + cC * (cC + T(2) * aC));
        qm = sqrt_t<T>(b2o + cC * (cC - T(2) * aC));
        cfC = (qp + qm) / T(2);
    }
    if (n_disc == 0) {
        SL = smin((vL[0] - cfL), (vR[0] - cfR));
        SR = smax((vL[0] + cfL), (vR[0] + cfR));
    } else if (n_disc == 1) {
        SL = smin((vL[0] - cfL), (vC1 - cfC));
        SR = smax((vR[0] + cfR), (vC1 + cfC));
    } else if (n_disc == 2) {
        const T SL_1 = smin((vL[0] - cfL), (vC1 - cfC));
        const T SR_1 = smax((vR[0] + cfR), (vC1 + cfC));
        const T SL_2 = smin((vL[0] - cfL), (vR[0] - cfR));
        const T SR_2 = smax((vL[0] + cfL), (vR[0] + cfR));
        const T oo = T(0.6);
        const T oo1 = T(1) - oo;
        SL = oo * SL_1 + oo1 * SL_2;
        SR = oo * SR_1 + oo1 * SR_2;
    } else if (n_disc == 3) {
        SL = smin(vL[0], vR[0]) - smax(cfL, cfR);
        SR = smax(vL[0], vR[0]) + smax(cfL, cfR);
    }
    out.dsr = SR;
    out.dsl = SL;

    const T suR = SR - vR[0];
    const T suL = SL - vL[0];
    // suR > 0 > suL, so this denominator (also that of ptz) is positive
    const T dsm = suR * r2 - suL * r1, ydsm = T(1) / dsm;
    const T SM = div_by_t<T>(suR * r2 * vR[0] - ptR + ptL - suL * r1 * vL[0], dsm, ydsm);
    out.dsc = SM;
    if (contact) wv = SM;

    if (n_state == 0) {
        const T TR0 = fabs(vL[0] + vR[0]) / T(2) + cfC;
        SR = TR0;
        SL = -TR0;
    }

    const T upt1 = (u1 * u1 + v1 * v1 + w1 * w1) / T(2);
    const T sbv1 = u1 * bx1 + v1 * by1 + w1 * bz1;
    const T upt2 = (u2 * u2 + v2 * v2 + w2 * w2) / T(2);
    const T sbv2 = u2 * bx2 + v2 * by2 + w2 * bz2;
    const T e1 = div_by_t<T>(p1, g1, yg1) + r1 * upt1 + b2L / T(2);
    const T e2 = div_by_t<T>(p2, g1, yg1) + r2 * upt2 + b2R / T(2);

    // Left/right physical fluxes and conserved states (Phys_param.cpp:1069-1117), formed where
    // they are used so that only the selected side stays live in registers.
    auto fl = [&](int k) -> T {
        switch (k) {
        case 0: return r1 * vL[0];
        case 1: return r1 * vL[0] * vL[0] + ptL - bL[0] * bL[0];
        case 2: return r1 * vL[0] * vL[1] - bL[0] * bL[1];
        case 3: return r1 * vL[0] * vL[2] - bL[0] * bL[2];
        case 4: return (e1 + ptL) * vL[0] - bL[0] * sbv1;
        case 5: return T(0);
        case 6: return vL[0] * bL[1] - vL[1] * bL[0];
        default: return vL[0] * bL[2] - vL[2] * bL[0];
        }
    };
    auto fr = [&](int k) -> T {
        switch (k) {
        case 0: return r2 * vR[0];
        case 1: return r2 * vR[0] * vR[0] + ptR - bR[0] * bR[0];
        case 2: return r2 * vR[0] * vR[1] - bR[0] * bR[1];
        case 3: return r2 * vR[0] * vR[2] - bR[0] * bR[2];
        case 4: return (e2 + ptR) * vR[0] - bR[0] * sbv2;
        case 5: return T(0);
        case 6: return vR[0] * bR[1] - vR[1] * bR[0];
        default: return vR[0] * bR[2] - vR[2] * bR[0];
        }
    };
    auto ul = [&](int k) -> T {
        switch (k) {
        case 0: return r1;
        case 4: return e1;
        case 1: case 2: case 3: return r1 * vL[k - 1];
        default: return bL[k - 5];
        }
    };
    auto ur = [&](int k) -> T {
        switch (k) {
        case 0: return r2;
        case 4: return e2;
        case 1: case 2: case 3: return r2 * vR[k - 1];
        default: return bR[k - 5];
        }
    };
    T FLk[K > 0 ? K : 1], FRk[K > 0 ? K : 1];
#pragma unroll
    for (int i = 0; i < K; i++) { FLk[i] = kL[i] * vL[0]; FRk[i] = kR[i] * vR[0]; }

    T qq[8];
    T SN, wbn;
    int region = 0, ibn = 0;
    out.ok = 1;

    if (n_state <= 1) {  // Lax / HLL, Phys_param.cpp:1133-1257
        if (left_ydar) wv = SL;
        T TL = SL, TR = SR;
        // UZ is needed only in the star region
        if (SL > wv || left_ydar) { TL = T(0); region = 1; }
        else if (SL <= wv && wv <= SR) { region = 2; }
        else if (SR < wv) { TR = T(0); region = 3; }
        else { region = 0; out.ok = 0; }
        const T a = TR * TL;
        const T b = TR - TL, yb = T(1) / b;
        const T dRL = SR - SL, ydRL = T(1) / dRL;
#pragma unroll
        for (int ik = 0; ik < 8; ik++) {
            T FW;
            if (region == 1) FW = wv * ul(ik);
            else if (region == 3) FW = wv * ur(ik);
            else FW = wv * div_by_t<T>(SR * ur(ik) - SL * ul(ik) + fl(ik) - fr(ik), dRL, ydRL);
            const T dq = ur(ik) - ul(ik);
            qq[ik] = div_by_t<T>(TR * fl(ik) - TL * fr(ik) + a * dq, b, yb) - FW;
        }
#pragma unroll
        for (int ik = 0; ik < K; ik++) {
            T FW;
            if (region == 1) FW = wv * kL[ik];
            else if (region == 3) FW = wv * kR[ik];
            else FW = wv * ((SR * kR[ik] - SL * kL[ik] + FLk[ik] - FRk[ik]) / (SR - SL));
            const T dq = kR[ik] - kL[ik];
            out.k[ik] = (TR * FLk[ik] - TL * FRk[ik] + a * dq) / b - FW;
        }
        SN = smax(fabs(SL), fabs(SR));
        if (wv >= SR) wbn = wv * bR[0];
        else if (wv <= SL) wbn = wv * bL[0];
        else wbn = wv * (bL[0] + bR[0]) / T(2);
    } else {
        // UZ (HLL average) components used by HLLC and HLLD: 0..3 and 5..7
        // HLL average state UZ (Phys_param.cpp:1119-1122); HLLD needs only its normal field UZ[5]
        T UZ[8];
        const T dRL = SR - SL, ydRL = T(1) / dRL;  // SR > SL
#pragma unroll
        for (int ik = 0; ik < 8; ik++)
            UZ[ik] = (n_state == 2 || ik == 5) ? div_by_t<T>(SR * ur(ik) - SL * ul(ik) + fl(ik) - fr(ik), dRL, ydRL) : T(0);
        const T suRm = suR / (SR - SM);
        const T suLm = suL / (SL - SM);
        const T rzR = r2 * suRm;
        const T rzL = r1 * suLm;
        // ptz, the star states and the energy of the star regions are formed only after the region
        // is chosen, and only for the side(s) it uses: a face whose velocity lies outside the fan
        // (every supersonic face) needs none of them.
        auto star_pt = [&]() -> T {
            return div_by_t<T>(suR * r2 * ptL - suL * r1 * ptR + r1 * r2 * suR * suL * (vR[0] - vL[0]), dsm, ydsm);
        };
        // A region left of the contact uses the left side's values and one right of it the right
        // side's: selecting the operands first (a negation or a select is exact) gives one code
        // path per kind of region, so a wavefront whose faces fall on both sides of the contact
        // runs it once.
        if (n_state == 2) {  // HLLC, Phys_param.cpp:1258-1475
            if (left_ydar) wv = SL;
            if (SL >= wv || left_ydar) region = 1;
            else if (SL <= wv && SM >= wv) region = 2;
            else if (SM <= wv && SR > wv) region = 3;
            else if (SR <= wv) region = 4;
            else { region = 0; out.ok = 0; }
            const bool R = region >= 3;
            if (region == 2 || region == 3) {
                const T e = R ? e2 : e1, sm = R ? suRm : suLm, pt = R ? ptR : ptL, v0 = R ? vR[0] : vL[0];
                const T sbv = R ? sbv2 : sbv1, S = R ? SR : SL, rz = R ? rzR : rzL;
                // the star velocity and field of the chosen side; (x / su) / r with the two
                // correctly rounded reciprocals shared by both components
                const T su = R ? suR : suL, rr = R ? r2 : r1;
                const T bt1 = R ? bR[1] : bL[1], bt2 = R ? bR[2] : bL[2];
                const T vt1 = R ? vR[1] : vL[1], vt2 = R ? vR[2] : vL[2];
                T vz1, vz2, bz1, bz2;
                if (fabs(UZ[5]) <= epsb) {
                    vz1 = vt1; vz2 = vt2;
                    bz1 = bt1 * sm; bz2 = bt2 * sm;
                } else {
                    const T ysu = T(1) / su, yr = T(1) / rr;
                    vz1 = vt1 + div_by_t<T>(div_by_t<T>(UZ[5] * (bt1 - UZ[6]), su, ysu), rr, yr);
                    vz2 = vt2 + div_by_t<T>(div_by_t<T>(UZ[5] * (bt2 - UZ[7]), su, ysu), rr, yr);
                    bz1 = UZ[6]; bz2 = UZ[7];
                }
                const T sbvz = (UZ[5] * UZ[1] + UZ[6] * UZ[2] + UZ[7] * UZ[3]) / UZ[0];
                const T ptz = star_pt();
                const T ez = e * sm + (ptz * SM - pt * v0 + UZ[5] * (sbv - sbvz)) / (S - SM);
                const T UZs[8] = {rz, SM * rz, vz1 * rz, vz2 * rz, ez, UZ[5], bz1, bz2};
#pragma unroll
                for (int ik = 0; ik < 8; ik++) {
                    const T f = R ? fr(ik) : fl(ik), u = R ? ur(ik) : ul(ik);
                    qq[ik] = f + S * (UZs[ik] - u) - wv * UZs[ik];
                }
#pragma unroll
                for (int ik = 0; ik < K; ik++) {
                    const T ks = R ? kR[ik] : kL[ik];
                    const T z = ks * sm;
                    out.k[ik] = (R ? FRk[ik] : FLk[ik]) + S * (z - ks) - wv * z;
                }
            } else {
#pragma unroll
                for (int ik = 0; ik < 8; ik++) qq[ik] = (R ? fr(ik) : fl(ik)) - wv * (R ? ur(ik) : ul(ik));
#pragma unroll
                for (int ik = 0; ik < K; ik++) out.k[ik] = (R ? FRk[ik] : FLk[ik]) - wv * (R ? kR[ik] : kL[ik]);
            }
        } else {  // HLLD, Phys_param.cpp:1476-1870
            const T bn = UZ[5];
            const T bn2 = bn * bn;
            const T rzRs = sqrt_t<T>(rzR);
            const T rzLs = sqrt_t<T>(rzL);
            const T rzss = rzRs + rzLs;
            const T rzps = rzRs * rzLs;
            T SZL = SM - fabs(bn) / rzLs;
            T SZR = SM + fabs(bn) / rzRs;
            T sbn;
            if (fabs(bn) > epsb) { sbn = bn > T(0) ? T(1) : T(-1); ibn = 1; }  // = |bn| / bn exactly
            else { sbn = T(0); ibn = 0; SZL = SM; SZR = SM; }
            int j = -1;
            if (left_ydar) wv = SL;
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
            const bool R = j >= 4;
            // star states of the sides the region uses: the left one for j = 2, 3, 4, the right
            // one for j = 3, 4, 5 (the double-star states combine both)
            T vzL1 = T(0), vzL2 = T(0), bzL1 = T(0), bzL2 = T(0), sbvL = T(0);
            T vzR1 = T(0), vzR2 = T(0), bzR1 = T(0), bzR2 = T(0), sbvR = T(0);
            if (j >= 2 && j <= 4) {
                T tvL, tbL;
                const T ttL = r1 * suL * (SL - SM) - bn2;
                if (fabs(ttL) < T(0.000000001)) { tvL = T(0); tbL = T(0); }
                else {
                    const T y = T(1) / ttL;
                    tvL = div_by_t<T>(SM - vL[0], ttL, y);
                    tbL = div_by_t<T>(r1 * suL * suL - bn2, ttL, y);
                }
                vzL1 = vL[1] - bn * bL[1] * tvL;
                vzL2 = vL[2] - bn * bL[2] * tvL;
                bzL1 = bL[1] * tbL; bzL2 = bL[2] * tbL;
                sbvL = bn * SM + bzL1 * vzL1 + bzL2 * vzL2;
            }
            if (j >= 3 && j <= 5) {
                T tvR, tbR;
                const T ttR = r2 * suR * (SR - SM) - bn2;
                if (fabs(ttR) < T(0.000000001)) { tvR = T(0); tbR = T(0); }
                else {
                    const T y = T(1) / ttR;
                    tvR = div_by_t<T>(SM - vR[0], ttR, y);
                    tbR = div_by_t<T>(r2 * suR * suR - bn2, ttR, y);
                }
                vzR1 = vR[1] - bn * bR[1] * tvR;
                vzR2 = vR[2] - bn * bR[2] * tvR;
                bzR1 = bR[1] * tbR; bzR2 = bR[2] * tbR;
                sbvR = bn * SM + bzR1 * vzR1 + bzR2 * vzR2;
            }
            if (j == 1 || j == 6) {
#pragma unroll
                for (int ik = 0; ik < 8; ik++) qq[ik] = (R ? fr(ik) : fl(ik)) - wv * (R ? ur(ik) : ul(ik));
#pragma unroll
                for (int ik = 0; ik < K; ik++) out.k[ik] = (R ? FRk[ik] : FLk[ik]) - wv * (R ? kR[ik] : kL[ik]);
            } else {
                const T e = R ? e2 : e1, sm = R ? suRm : suLm, pt = R ? ptR : ptL, v0 = R ? vR[0] : vL[0];
                const T sbv0 = R ? sbv2 : sbv1, sbvs = R ? sbvR : sbvL, S = R ? SR : SL, rz = R ? rzR : rzL;
                const T ptz = star_pt();
                const T ez = e * sm + (ptz * SM - pt * v0 + bn * (sbv0 - sbvs)) / (S - SM);
                const T vz1 = R ? vzR1 : vzL1, vz2 = R ? vzR2 : vzL2, bz1 = R ? bzR1 : bzL1, bz2 = R ? bzR2 : bzL2;
                const T UZs[8] = {rz, SM * rz, vz1 * rz, vz2 * rz, ez, bn, bz1, bz2};
                if (j == 2 || j == 5) {
#pragma unroll
                    for (int ik = 0; ik < 8; ik++) {
                        const T f = R ? fr(ik) : fl(ik), u = R ? ur(ik) : ul(ik);
                        qq[ik] = f + S * (UZs[ik] - u) - wv * UZs[ik];
                    }
                } else {
                    const T yz = T(1) / rzss;  // a sum of two positive square roots
                    const T vzz1 = div_by_t<T>(rzLs * vzL1 + rzRs * vzR1 + sbn * (bzR1 - bzL1), rzss, yz);
                    const T vzz2 = div_by_t<T>(rzLs * vzL2 + rzRs * vzR2 + sbn * (bzR2 - bzL2), rzss, yz);
                    const T bzz1 = div_by_t<T>(rzLs * bzR1 + rzRs * bzL1 + sbn * rzps * (vzR1 - vzL1), rzss, yz);
                    const T bzz2 = div_by_t<T>(rzLs * bzR2 + rzRs * bzL2 + sbn * rzps * (vzR2 - vzL2), rzss, yz);
                    const T sbzz = bn * SM + bzz1 * vzz1 + bzz2 * vzz2;
                    // ezzL = ezL - rzLs sbn (sbvL - sbzz), ezzR = ezR + rzRs sbn (sbvR - sbzz)
                    const T dz = (R ? rzRs : rzLs) * sbn * (sbvs - sbzz);
                    const T ezz = ez + (R ? dz : -dz);
                    const T UZZ[8] = {rz, SM * rz, vzz1 * rz, vzz2 * rz, ezz, bn, bzz1, bzz2};
                    const T SZ = R ? SZR : SZL;
#pragma unroll
                    for (int ik = 0; ik < 8; ik++) {
                        const T f = R ? fr(ik) : fl(ik), u = R ? ur(ik) : ul(ik);
                        qq[ik] = f + SZ * (UZZ[ik] - UZs[ik]) + S * (UZs[ik] - u) - wv * UZZ[ik];
                    }
                }
#pragma unroll
                for (int ik = 0; ik < K; ik++) {
                    const T ks = R ? kR[ik] : kL[ik];
                    const T z = ks * sm;
                    const T SZk = (j == 3 || j == 4) ? (R ? SZR : SZL) : T(0);
                    out.k[ik] = (j == 3 || j == 4) ? ((R ? FRk[ik] : FLk[ik]) + SZk * (z - z) + S * (z - ks) - wv * z)
                                                   : ((R ? FRk[ik] : FLk[ik]) + S * (z - ks) - wv * z);
                }
            }
        }
        SN = smax(fabs(SL), fabs(SR));
        if (wv > SR) wbn = wv * bR[0];
        else if (wv < SL) wbn = wv * bL[0];
        else wbn = wv * (bL[0] + bR[0]) / T(2);
    }

    const T qv0 = qq[1], qv1 = qq[2], qv2 = qq[3];
    const T qb0 = -SN * (bR[0] - bL[0]) - wbn, qb1 = qq[6], qb2 = qq[7];
    out.q[0] = qq[0];
    out.q[4] = qq[4];
#pragma unroll
    for (int ik = 0; ik < 3; ik++) {
        out.q[ik + 1] = n[ik] * qv0 + t[ik] * qv1 + m[ik] * qv2;
        const T b = n[ik] * qb0 + t[ik] * qb1 + m[ik] * qb2;
        out.q[ik + 5] = spi4 * b;
    }
    out.branch = (n_state << 4) | (ibn << 3) | region;
    if (!out.ok) {
#pragma unroll
        for (int ik = 0; ik < 8; ik++) out.q[ik] = T(0);
    }
}

}  // namespace titan
