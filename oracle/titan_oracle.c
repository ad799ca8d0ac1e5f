/*
 * titan_oracle.c -- CPU restatement of TITAN's explicit ideal-MHD step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the CPU baseline of the
 * benchmark.  Nothing in the product links it.  Build: gcc -O2 -ffp-contract=off -fopenmp
 * (no FMA contraction, IEEE division and square root), as in oracle/Makefile.
 *
 * The arithmetic follows the reference expression by expression (C evaluates a+b+c as (a+b)+c,
 * the reference's compiler does the same with contraction off).  Reference citations are given
 * per function.  Eigen small-vector reductions (Vector3d::dot / norm / normalized) are taken in
 * the order (x0*y0 + x1*y1) + x2*y2, which is the order of Eigen's unaligned-packet reduction of
 * a 3-vector; the reference binary cannot be built here (Eigen absent), so this is an assumption
 * that only affects the last bit.
 */
#define _GNU_SOURCE 1 /* sincos */
#include "titan_oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAX_H 16 /* hydrogen fluids or_bera_source_h takes */
#define KV(x) ((x) * (x))
#define KVV(x, y, z) ((x) * (x) + (y) * (y) + (z) * (z))
/* std::max / std::min semantics (first argument returned on ties and NaN) */
#define SMAX(a, b) (((a) < (b)) ? (b) : (a))
#define SMIN(a, b) (((b) < (a)) ? (b) : (a))

static const double OR_PI = 3.14159265358979323846;
static const double OR_EPSB = 1E-4; /* Header.h:96 */

/* Planted mistakes for the tests that show a pin catches a plausible error (0 = the restatement;
 * the OR_MISTAKE_* bits of titan_oracle.h).  Never set outside those tests. */
static int or_mistake_bits = 0;
void or_set_mistake(int bits) { or_mistake_bits = bits; }
#define MISTAKE(b) ((or_mistake_bits & (b)) != 0)

static double or_spi4(void) { return sqrt(4.0 * OR_PI); } /* Header.h:94 */

/* get_bazis, Help_funk.cpp:219-232: t = normalize(a x n), m = n x t, a = x or y. */
void or_get_bazis(const double *n, double *t, double *m)
{
    double a0 = 1.0, a1 = 0.0, a2 = 0.0;
    double nd = (n[0] * a0 + n[1] * a1) + n[2] * a2;
    if (fabs(nd) > 0.9) { a0 = 0.0; a1 = 1.0; a2 = 0.0; }
    double c0 = a1 * n[2] - a2 * n[1];
    double c1 = a2 * n[0] - a0 * n[2];
    double c2 = a0 * n[1] - a1 * n[0];
    double z = (c0 * c0 + c1 * c1) + c2 * c2;
    if (z > 0.0) {
        double s = sqrt(z);
        c0 = c0 / s; c1 = c1 / s; c2 = c2 / s;
    }
    t[0] = c0; t[1] = c1; t[2] = c2;
    m[0] = n[1] * t[2] - n[2] * t[1];
    m[1] = n[2] * t[0] - n[0] * t[2];
    m[2] = n[0] * t[1] - n[1] * t[0];
}

static double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* scalarProductFast(a..., b...) = ax*bx + ay*by + az*bz (Help_funk.cpp:246-251) */
static double spf(double ax, double ay, double az, double bx, double by, double bz)
{
    return ax * bx + ay * by + az * bz;
}

/* Phys_param::Get_Potok, Phys_param.cpp:826-874 (used by tests as a consistency oracle). */
void or_get_potok(double gamma, double rho, double p, double u, double v, double w, double bx,
                  double by, double bz, double n1, double n2, double n3, double *potok)
{
    const double spi4 = or_spi4();
    double n[3] = {n1, n2, n3}, t[3], m[3];
    or_get_bazis(n, t, m);
    double vn = spf(u, v, w, n1, n2, n3);
    double vt = spf(u, v, w, t[0], t[1], t[2]);
    double vm = spf(u, v, w, m[0], m[1], m[2]);
    double Bn = spf(bx, by, bz, n1, n2, n3) / spi4;
    double Bt = spf(bx, by, bz, t[0], t[1], t[2]) / spi4;
    double Bm = spf(bx, by, bz, m[0], m[1], m[2]) / spi4;
    double bb = KVV(bx, by, bz) / spi4 / spi4;
    double vv = KVV(u, v, w);
    double e = p / (gamma - 1) + rho * vv / 2.0 + bb / 2.0;
    double pT = p + bb / 2.0;
    potok[0] = rho * vn;
    potok[7] = (e + pT) * vn - Bn * spf(bx, by, bz, u, v, w) / spi4;
    double Pvn = rho * KV(vn) + pT - KV(Bn);
    double Pvt = rho * vn * vt - Bn * Bt;
    double Pvm = rho * vn * vm - Bn * Bm;
    double Pbt = Bt * vn - Bn * vt;
    double Pbm = Bm * vn - Bn * vm;
    Pbt *= spi4;
    Pbm *= spi4;
    potok[1] = Pvn * n[0] + Pvt * t[0] + Pvm * m[0];
    potok[2] = Pvn * n[1] + Pvt * t[1] + Pvm * m[1];
    potok[3] = Pvn * n[2] + Pvt * t[2] + Pvm * m[2];
    potok[4] = Pbt * t[0] + Pbm * m[0];
    potok[5] = Pbt * t[1] + Pbm * m[1];
    potok[6] = Pbt * t[2] + Pbm * m[2];
}

#define OR_KMAX 8

/* Phys_param::chlld, Phys_param.cpp:882-1878.  n_state 0 Lax, 1 HLL, 2 HLLC, 3 HLLD.
 * flags bit0 = left_ydar, bit1 = contact; null_bn is always false on the step path
 * (Setka_physics.cpp:2420-2424) and is not exposed. */
int or_chlld(int n_state, int n_disc, int flags, double al, double be, double ge, double w,
             const double *qL, const double *qR, int K, const double *kL, const double *kR,
             double *q, double *k, double *dsl, double *dsc, double *dsr, int *branch)
{
    const double spi4 = or_spi4();
    const double gamma = 5.0 / 3.0;
    const double g1 = gamma - 1.0;
    const int left_ydar = flags & 1, contact = (flags >> 1) & 1;
    double FL[8], FR[8], UL[8], UR[8], UZ[8];
    double FLk[OR_KMAX], FRk[OR_KMAX], ULk[OR_KMAX], URk[OR_KMAX], UZk[OR_KMAX];
    if (K > OR_KMAX) return -2;
    double wv = w;

    double r1 = qL[0], u1 = qL[1], v1 = qL[2], w1 = qL[3], p1 = qL[4];
    double bx1 = qL[5] / spi4, by1 = qL[6] / spi4, bz1 = qL[7] / spi4;
    double r2 = qR[0], u2 = qR[1], v2 = qR[2], w2 = qR[3], p2 = qR[4];
    double bx2 = qR[5] / spi4, by2 = qR[6] / spi4, bz2 = qR[7] / spi4;

    double ro = (r2 + r1) / 2.0;
    double ap = (p2 + p1) / 2.0;

    double n[3] = {al, be, ge}, t[3], m[3];
    or_get_bazis(n, t, m);

    double vL[3], vR[3], bL[3], bR[3];
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

    double aaL = bL[0] / sqrt(r1);
    double b2L = dot3(bL, bL);
    double b21 = b2L / r1;
    double cL = sqrt(gamma * p1 / r1);
    double qp = sqrt(b21 + cL * (cL + 2.0 * aaL));
    double qm = sqrt(b21 + cL * (cL - 2.0 * aaL));
    double cfL = (qp + qm) / 2.0;
    double ptL = p1 + b2L / 2.0;

    double aaR = bR[0] / sqrt(r2);
    double b2R = dot3(bR, bR);
    double b22 = b2R / r2;
    double cR = sqrt(gamma * p2 / r2);
    qp = sqrt(b22 + cR * (cR + 2.0 * aaR));
    qm = sqrt(b22 + cR * (cR - 2.0 * aaR));
    double cfR = (qp + qm) / 2.0;
    double ptR = p2 + b2R / 2.0;

    double aC = (aaL + aaR) / 2.0;
    double b2o = (b22 + b21) / 2.0;
    double cC = sqrt(gamma * ap / ro);
    qp = sqrt(b2o + cC * (cC + 2.0 * aC));
    qm = sqrt(b2o + cC * (cC - 2.0 * aC));
    double cfC = (qp + qm) / 2.0;
    double vC1 = (vL[0] + vR[0]) / 2.0;

    double SL = 0.0, SR = 0.0;
    if (left_ydar) n_disc = 2;
    if (n_disc == 0) {
        SL = SMIN((vL[0] - cfL), (vR[0] - cfR));
        SR = SMAX((vL[0] + cfL), (vR[0] + cfR));
    } else if (n_disc == 1) {
        SL = SMIN((vL[0] - cfL), (vC1 - cfC));
        SR = SMAX((vR[0] + cfR), (vC1 + cfC));
    } else if (n_disc == 2) {
        double SL_1 = SMIN((vL[0] - cfL), (vC1 - cfC));
        double SR_1 = SMAX((vR[0] + cfR), (vC1 + cfC));
        double SL_2 = SMIN((vL[0] - cfL), (vR[0] - cfR));
        double SR_2 = SMAX((vL[0] + cfL), (vR[0] + cfR));
        double oo = 0.6;
        double oo1 = 1.0 - oo;
        SL = oo * SL_1 + oo1 * SL_2;
        SR = oo * SR_1 + oo1 * SR_2;
    } else if (n_disc == 3) {
        SL = SMIN(vL[0], vR[0]) - SMAX(cfL, cfR);
        SR = SMAX(vL[0], vR[0]) + SMAX(cfL, cfR);
    }
    *dsr = SR;
    *dsl = SL;

    double suR = SR - vR[0];
    double suL = SL - vL[0];
    double SM = (suR * r2 * vR[0] - ptR + ptL - suL * r1 * vL[0]) / (suR * r2 - suL * r1);
    *dsc = SM;
    if (contact) wv = SM;

    if (n_state == 0) {
        double TR0 = fabs(vL[0] + vR[0]) / 2.0 + cfC;
        double TL0 = -TR0;
        SR = TR0;
        SL = TL0;
    }

    double upt1 = (KV(u1) + KV(v1) + KV(w1)) / 2.0;
    double sbv1 = u1 * bx1 + v1 * by1 + w1 * bz1;
    double upt2 = (KV(u2) + KV(v2) + KV(w2)) / 2.0;
    double sbv2 = u2 * bx2 + v2 * by2 + w2 * bz2;
    double e1 = p1 / g1 + r1 * upt1 + b2L / 2.0;
    double e2 = p2 / g1 + r2 * upt2 + b2R / 2.0;

    FL[0] = r1 * vL[0];
    FL[1] = r1 * vL[0] * vL[0] + ptL - KV(bL[0]);
    FL[2] = r1 * vL[0] * vL[1] - bL[0] * bL[1];
    FL[3] = r1 * vL[0] * vL[2] - bL[0] * bL[2];
    FL[4] = (e1 + ptL) * vL[0] - bL[0] * sbv1;
    FL[5] = 0.0;
    FL[6] = vL[0] * bL[1] - vL[1] * bL[0];
    FL[7] = vL[0] * bL[2] - vL[2] * bL[0];
    for (int i = 0; i < K; i++) FLk[i] = kL[i] * vL[0];

    FR[0] = r2 * vR[0];
    FR[1] = r2 * vR[0] * vR[0] + ptR - KV(bR[0]);
    FR[2] = r2 * vR[0] * vR[1] - bR[0] * bR[1];
    FR[3] = r2 * vR[0] * vR[2] - bR[0] * bR[2];
    FR[4] = (e2 + ptR) * vR[0] - bR[0] * sbv2;
    FR[5] = 0.0;
    FR[6] = vR[0] * bR[1] - vR[1] * bR[0];
    FR[7] = vR[0] * bR[2] - vR[2] * bR[0];
    for (int i = 0; i < K; i++) FRk[i] = kR[i] * vR[0];

    UL[0] = r1; UL[4] = e1; UR[0] = r2; UR[4] = e2;
    for (int i = 0; i < K; i++) { ULk[i] = kL[i]; URk[i] = kR[i]; }
    for (int i = 0; i < 3; i++) {
        UL[i + 1] = r1 * vL[i];
        UL[i + 5] = bL[i];
        UR[i + 1] = r2 * vR[i];
        UR[i + 5] = bR[i];
    }
    for (int ik = 0; ik < 8; ik++) UZ[ik] = (SR * UR[ik] - SL * UL[ik] + FL[ik] - FR[ik]) / (SR - SL);
    for (int ik = 0; ik < K; ik++) UZk[ik] = (SR * URk[ik] - SL * ULk[ik] + FLk[ik] - FRk[ik]) / (SR - SL);

    double qq[8];      /* rotated-frame flux, components 0..7 */
    double SN, wbn;
    int region = 0, ibn = 0;

    if (n_state <= 1) { /* Lax or HLL, Phys_param.cpp:1133-1257 */
        double dq[8], FW[8], dqk[OR_KMAX], FWk[OR_KMAX];
        for (int ik = 0; ik < 8; ik++) dq[ik] = UR[ik] - UL[ik];
        for (int ik = 0; ik < K; ik++) dqk[ik] = URk[ik] - ULk[ik];
        if (left_ydar) wv = SL;
        double TL = SL, TR = SR;
        if (SL > wv || left_ydar) {
            TL = 0.0;
            for (int ik = 0; ik < 8; ik++) FW[ik] = wv * UL[ik];
            for (int ik = 0; ik < K; ik++) FWk[ik] = wv * ULk[ik];
            region = 1;
        } else if (SL <= wv && wv <= SR) {
            for (int ik = 0; ik < 8; ik++) FW[ik] = wv * UZ[ik];
            for (int ik = 0; ik < K; ik++) FWk[ik] = wv * UZk[ik];
            region = 2;
        } else if (SR < wv) {
            TR = 0.0;
            for (int ik = 0; ik < 8; ik++) FW[ik] = wv * UR[ik];
            for (int ik = 0; ik < K; ik++) FWk[ik] = wv * URk[ik];
            region = 3;
        } else {
            if (branch) *branch = (n_state << 4);
            return -1;
        }
        double a = TR * TL;
        double b = TR - TL;
        for (int ik = 0; ik < 8; ik++) qq[ik] = (TR * FL[ik] - TL * FR[ik] + a * dq[ik]) / b - FW[ik];
        for (int ik = 0; ik < K; ik++) k[ik] = (TR * FLk[ik] - TL * FRk[ik] + a * dqk[ik]) / b - FWk[ik];
        SN = SMAX(fabs(SL), fabs(SR));
        if (wv >= SR) wbn = wv * bR[0];
        else if (wv <= SL) wbn = wv * bL[0];
        else wbn = wv * (bL[0] + bR[0]) / 2.0;
    } else if (n_state == 2) { /* HLLC, Phys_param.cpp:1258-1475 */
        double suRm = suR / (SR - SM);
        double suLm = suL / (SL - SM);
        double rzR = r2 * suRm;
        double rzL = r1 * suLm;
        double vzR[3], vzL[3], bzR[3], bzL[3];
        vzR[0] = SM;
        vzL[0] = SM;
        double ptz = (suR * r2 * ptL - suL * r1 * ptR + r1 * r2 * suR * suL * (vR[0] - vL[0])) /
                     (suR * r2 - suL * r1);
        bzR[0] = UZ[5];
        bzL[0] = UZ[5];
        vzR[1] = vR[1] + UZ[5] * (bR[1] - UZ[6]) / suR / r2;
        vzR[2] = vR[2] + UZ[5] * (bR[2] - UZ[7]) / suR / r2;
        vzL[1] = vL[1] + UZ[5] * (bL[1] - UZ[6]) / suL / r1;
        vzL[2] = vL[2] + UZ[5] * (bL[2] - UZ[7]) / suL / r1;
        bzR[1] = UZ[6];
        bzR[2] = UZ[7];
        bzL[1] = bzR[1];
        bzL[2] = bzR[2];
        double sbvz = (UZ[5] * UZ[1] + UZ[6] * UZ[2] + UZ[7] * UZ[3]) / UZ[0];
        double ezR = e2 * suRm + (ptz * SM - ptR * vR[0] + UZ[5] * (sbv2 - sbvz)) / (SR - SM);
        double ezL = e1 * suLm + (ptz * SM - ptL * vL[0] + UZ[5] * (sbv1 - sbvz)) / (SL - SM);
        if (fabs(UZ[5]) <= OR_EPSB) {
            vzR[1] = vR[1];
            vzR[2] = vR[2];
            vzL[1] = vL[1];
            vzL[2] = vL[2];
            bzR[1] = bR[1] * suRm;
            bzR[2] = bR[2] * suRm;
            bzL[1] = bL[1] * suLm;
            bzL[2] = bL[2] * suLm;
        }
        double UZL[8], UZR[8], UZLk[OR_KMAX], UZRk[OR_KMAX];
        for (int ik = 0; ik < K; ik++) { UZLk[ik] = kL[ik] * suLm; UZRk[ik] = kR[ik] * suRm; }
        UZL[0] = rzL; UZL[4] = ezL; UZR[0] = rzR; UZR[4] = ezR;
        for (int ik = 0; ik < 3; ik++) {
            UZL[ik + 1] = vzL[ik] * rzL;
            UZL[ik + 5] = bzL[ik];
            UZR[ik + 1] = vzR[ik] * rzR;
            UZR[ik + 5] = bzR[ik];
        }
        if (left_ydar) wv = SL;
        if (SL >= wv || left_ydar) {
            for (int ik = 0; ik < 8; ik++) qq[ik] = FL[ik] - wv * UL[ik];
            for (int ik = 0; ik < K; ik++) k[ik] = FLk[ik] - wv * ULk[ik];
            region = 1;
        } else if (SL <= wv && SM >= wv) {
            for (int ik = 0; ik < 8; ik++) qq[ik] = FL[ik] + SL * (UZL[ik] - UL[ik]) - wv * UZL[ik];
            for (int ik = 0; ik < K; ik++) k[ik] = FLk[ik] + SL * (UZLk[ik] - kL[ik]) - wv * UZLk[ik];
            region = 2;
        } else if (SM <= wv && SR > wv) {
            for (int ik = 0; ik < 8; ik++) qq[ik] = FR[ik] + SR * (UZR[ik] - UR[ik]) - wv * UZR[ik];
            for (int ik = 0; ik < K; ik++) k[ik] = FRk[ik] + SR * (UZRk[ik] - kR[ik]) - wv * UZRk[ik];
            region = 3;
        } else if (SR <= wv) {
            for (int ik = 0; ik < 8; ik++) qq[ik] = FR[ik] - wv * UR[ik];
            for (int ik = 0; ik < K; ik++) k[ik] = FRk[ik] - wv * URk[ik];
            region = 4;
        } else {
            if (branch) *branch = (n_state << 4);
            return -1;
        }
        SN = SMAX(fabs(SL), fabs(SR));
        if (wv > SR) wbn = wv * bR[0];
        else if (wv < SL) wbn = wv * bL[0];
        else wbn = wv * (bL[0] + bR[0]) / 2.0;
    } else if (n_state == 3) { /* HLLD, Phys_param.cpp:1476-1870 */
        double vzR[3], vzL[3], vzzL[3], vzzR[3], bzR[3], bzL[3], bzzL[3], bzzR[3];
        double ptz = (suR * r2 * ptL - suL * r1 * ptR + r1 * r2 * suR * suL * (vR[0] - vL[0])) /
                     (suR * r2 - suL * r1);
        vzL[0] = SM; vzR[0] = SM; vzzL[0] = SM; vzzR[0] = SM;
        double suRm = suR / (SR - SM);
        double suLm = suL / (SL - SM);
        double rzR = r2 * suRm;
        double rzL = r1 * suLm;
        double qzL[OR_KMAX], qzR[OR_KMAX];
        for (int ik = 0; ik < K; ik++) { qzL[ik] = kL[ik] * suLm; qzR[ik] = kR[ik] * suRm; }
        double bn = UZ[5];
        double bn2 = bn * bn;
        bzL[0] = bn; bzR[0] = bn; bzzL[0] = bn; bzzR[0] = bn;
        double tvR, tbR;
        double ttR = r2 * suR * (SR - SM) - bn2;
        if (fabs(ttR) < 0.000000001) { tvR = 0.0; tbR = 0.0; }
        else { tvR = (SM - vR[0]) / ttR; tbR = (r2 * suR * suR - bn2) / ttR; }
        double tvL, tbL;
        double ttL = r1 * suL * (SL - SM) - bn2;
        if (fabs(ttL) < 0.000000001) { tvL = 0.0; tbL = 0.0; }
        else { tvL = (SM - vL[0]) / ttL; tbL = (r1 * suL * suL - (MISTAKE(OR_MISTAKE_HLLD_TB_BN2) ? 0.0 : bn2)) / ttL; }
        vzL[1] = vL[1] - bn * bL[1] * tvL;
        vzL[2] = vL[2] - bn * bL[2] * tvL;
        vzR[1] = vR[1] - bn * bR[1] * tvR;
        vzR[2] = vR[2] - bn * bR[2] * tvR;
        bzL[1] = bL[1] * tbL;
        bzL[2] = bL[2] * tbL;
        bzR[1] = bR[1] * tbR;
        bzR[2] = bR[2] * tbR;
        double sbvL = bzL[0] * vzL[0] + bzL[1] * vzL[1] + bzL[2] * vzL[2];
        double sbvR = bzR[0] * vzR[0] + bzR[1] * vzR[1] + bzR[2] * vzR[2];
        double ezR = e2 * suRm + (ptz * SM - ptR * vR[0] + bn * (sbv2 - sbvR)) / (SR - SM);
        double ezL = e1 * suLm + (ptz * SM - ptL * vL[0] + bn * (sbv1 - sbvL)) / (SL - SM);
        double rzzR = rzR;
        double rzzL = rzL;
        double rzRs = sqrt(rzR);
        double rzLs = sqrt(rzL);
        double rzss = rzRs + rzLs;
        double rzps = rzRs * rzLs;
        double SZL = SM - fabs(bn) / rzLs;
        double SZR = SM + fabs(bn) / rzRs;
        double sbn;
        if (fabs(bn) > OR_EPSB) { sbn = fabs(bn) / bn; ibn = 1; }
        else { sbn = 0.0; ibn = 0; SZL = SM; SZR = SM; }
        const double sbv = MISTAKE(OR_MISTAKE_HLLD_VZZ_SIGN) ? -sbn : sbn;
        vzzL[1] = (rzLs * vzL[1] + rzRs * vzR[1] + sbv * (bzR[1] - bzL[1])) / rzss;
        vzzL[2] = (rzLs * vzL[2] + rzRs * vzR[2] + sbv * (bzR[2] - bzL[2])) / rzss;
        vzzR[1] = vzzL[1];
        vzzR[2] = vzzL[2];
        if (MISTAKE(OR_MISTAKE_HLLD_BZZ_WEIGHTS)) {
            bzzL[1] = (rzLs * bzL[1] + rzRs * bzR[1] + sbn * rzps * (vzR[1] - vzL[1])) / rzss;
            bzzL[2] = (rzLs * bzL[2] + rzRs * bzR[2] + sbn * rzps * (vzR[2] - vzL[2])) / rzss;
        } else {
            bzzL[1] = (rzLs * bzR[1] + rzRs * bzL[1] + sbn * rzps * (vzR[1] - vzL[1])) / rzss;
            bzzL[2] = (rzLs * bzR[2] + rzRs * bzL[2] + sbn * rzps * (vzR[2] - vzL[2])) / rzss;
        }
        bzzR[1] = bzzL[1];
        bzzR[2] = bzzL[2];
        double sbzz = bzzL[0] * vzzL[0] + bzzL[1] * vzzL[1] + bzzL[2] * vzzL[2];
        double ezzR = ezR + rzRs * sbn * (sbvR - sbzz);
        double ezzL = ezL - rzLs * sbn * (sbvL - sbzz);
        double UZL[8], UZR[8], UZZL[8], UZZR[8];
        UZL[0] = rzL; UZL[4] = ezL; UZR[0] = rzR; UZR[4] = ezR;
        for (int ik = 0; ik < 3; ik++) {
            UZL[ik + 1] = vzL[ik] * rzL;
            UZL[ik + 5] = bzL[ik];
            UZR[ik + 1] = vzR[ik] * rzR;
            UZR[ik + 5] = bzR[ik];
        }
        UZZL[0] = rzzL; UZZL[4] = ezzL; UZZR[0] = rzzR; UZZR[4] = ezzR;
        for (int ik = 0; ik < 3; ik++) {
            UZZL[ik + 1] = vzzL[ik] * rzzL;
            UZZL[ik + 5] = bzzL[ik];
            UZZR[ik + 1] = vzzR[ik] * rzzR;
            UZZR[ik + 5] = bzzR[ik];
        }
        /* passive scalars: UZZLk = UZLk = qzL (Phys_param.cpp:1560-1566,1656-1662) */
        int j = -1;
        if (left_ydar) wv = SL;
        if (SL >= wv || left_ydar) {
            for (int ik = 0; ik < 8; ik++) qq[ik] = FL[ik] - wv * UL[ik];
            for (int ik = 0; ik < K; ik++) k[ik] = FLk[ik] - wv * ULk[ik];
            j = 1;
        } else if (SL <= wv && SZL >= wv) {
            for (int ik = 0; ik < 8; ik++) qq[ik] = FL[ik] + SL * (UZL[ik] - UL[ik]) - wv * UZL[ik];
            for (int ik = 0; ik < K; ik++) k[ik] = FLk[ik] + SL * (qzL[ik] - ULk[ik]) - wv * qzL[ik];
            j = 2;
        }
        if (ibn == 1) {
            if (SZL <= wv && SM >= wv) {
                for (int ik = 0; ik < 8; ik++)
                    qq[ik] = FL[ik] + SZL * (UZZL[ik] - UZL[ik]) + SL * (UZL[ik] - UL[ik]) - wv * UZZL[ik];
                for (int ik = 0; ik < K; ik++)
                    k[ik] = FLk[ik] + SZL * (qzL[ik] - qzL[ik]) + SL * (qzL[ik] - ULk[ik]) - wv * qzL[ik];
                j = 3;
            } else if (SM <= wv && SZR >= wv) {
                for (int ik = 0; ik < 8; ik++)
                    qq[ik] = FR[ik] + SZR * (UZZR[ik] - UZR[ik]) + SR * (UZR[ik] - UR[ik]) - wv * UZZR[ik];
                for (int ik = 0; ik < K; ik++)
                    k[ik] = FRk[ik] + SZR * (qzR[ik] - qzR[ik]) + SR * (qzR[ik] - URk[ik]) - wv * qzR[ik];
                j = 4;
            }
        }
        if (SZR <= wv && SR >= wv) {
            for (int ik = 0; ik < 8; ik++) qq[ik] = FR[ik] + SR * (UZR[ik] - UR[ik]) - wv * UZR[ik];
            for (int ik = 0; ik < K; ik++) k[ik] = FRk[ik] + SR * (qzR[ik] - URk[ik]) - wv * qzR[ik];
            j = 5;
        } else if (SR <= wv) {
            for (int ik = 0; ik < 8; ik++) qq[ik] = FR[ik] - wv * UR[ik];
            for (int ik = 0; ik < K; ik++) k[ik] = FRk[ik] - wv * URk[ik];
            j = 6;
        }
        if (j == -1) {
            if (branch) *branch = (n_state << 4) | (ibn << 3);
            return -1;
        }
        region = j;
        SN = SMAX(fabs(SL), fabs(SR));
        if (wv > SR) wbn = wv * bR[0];
        else if (wv < SL) wbn = wv * bL[0];
        else wbn = wv * (bL[0] + bR[0]) / 2.0;
    } else {
        return -3;
    }

    /* Bn flux and rotation back to the lab frame (identical in all three solvers) */
    double qv0 = qq[1], qv1 = qq[2], qv2 = qq[3];
    double qb0 = -SN * (bR[0] - bL[0]) - wbn, qb1 = qq[6], qb2 = qq[7];
    q[0] = qq[0];
    q[4] = qq[4];
    for (int ik = 0; ik < 3; ik++) {
        q[ik + 1] = n[ik] * qv0 + t[ik] * qv1 + m[ik] * qv2;
        double b = n[ik] * qb0 + t[ik] * qb1 + m[ik] * qb2;
        q[ik + 5] = spi4 * b;
    }
    if (branch) *branch = (n_state << 4) | (ibn << 3) | region;
    return 0;
}

int or_chlld_batch(int64_t n, const int8_t *n_state, const int8_t *n_disc, const uint8_t *flags,
                   const double *normal, const double *w, const double *qL, const double *qR,
                   int K, const double *kL, const double *kR, double *q, double *k, double *ds,
                   uint8_t *branch)
{
    int status = 0;
    for (int64_t i = 0; i < n; i++) {
        double dsl, dsc, dsr;
        int br = 0;
        int rc = or_chlld(n_state[i], n_disc ? n_disc[i] : 0, flags ? flags[i] : 0, normal[3 * i],
                          normal[3 * i + 1], normal[3 * i + 2], w ? w[i] : 0.0, qL + 8 * i, qR + 8 * i, K,
                          K ? kL + K * i : NULL, K ? kR + K * i : NULL, q + 8 * i, K ? k + K * i : NULL,
                          &dsl, &dsc, &dsr, &br);
        if (ds) { ds[3 * i] = dsl; ds[3 * i + 1] = dsc; ds[3 * i + 2] = dsr; }
        if (branch) branch[i] = (uint8_t)br;
        if (rc != 0 && status == 0) status = rc;
    }
    return status;
}

/* ------------------------------------------------------------------------------------------ */
/* Geometry, reference mode.                                                                  */

static void tri_area_add(const double *a, const double *b, const double *c, double *S)
{
    /* triangleArea3D, Help_funk.cpp:193-217 */
    double v1x = b[0] - a[0], v1y = b[1] - a[1], v1z = b[2] - a[2];
    double v2x = c[0] - a[0], v2y = c[1] - a[1], v2z = c[2] - a[2];
    double cx = v1y * v2z - v1z * v2y;
    double cy = v1z * v2x - v1x * v2z;
    double cz = v1x * v2y - v1y * v2x;
    *S += 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

static double tet_volume(const double *A, const double *B, const double *C, const double *D)
{
    /* tetrahedronVolume, Help_funk.cpp:253-272 */
    double abx = B[0] - A[0], aby = B[1] - A[1], abz = B[2] - A[2];
    double acx = C[0] - A[0], acy = C[1] - A[1], acz = C[2] - A[2];
    double adx = D[0] - A[0], ady = D[1] - A[1], adz = D[2] - A[2];
    double cx = aby * acz - abz * acy;
    double cy = abz * acx - abx * acz;
    double cz = abx * acy - aby * acx;
    double d = cx * adx + cy * ady + cz * adz;
    return fabs(d) / 6.0;
}

void or_measure(const or_mesh *m, const double *X, or_geom *g)
{
    /* Cell::Culc_center (Cell.cpp:909-930) */
    for (int64_t c = 0; c < m->n_cells; c++) {
        double xc = 0.0, yc = 0.0, zc = 0.0;
        for (int i = 0; i < 8; i++) {
            const double *x = X + 3 * (int64_t)m->cell_nodes[8 * c + i];
            xc += x[0]; yc += x[1]; zc += x[2];
        }
        g->cell_center[3 * c] = xc / 8;
        g->cell_center[3 * c + 1] = yc / 8;
        g->cell_center[3 * c + 2] = zc / 8;
    }
    /* Gran::Culc_measure (Gran.cpp:112-218) */
    for (int64_t f = 0; f < m->n_faces; f++) {
        const double *x[4];
        for (int i = 0; i < 4; i++) x[i] = X + 3 * (int64_t)m->face_nodes[4 * f + i];
        double xc = 0.0, yc = 0.0, zc = 0.0;
        for (int i = 0; i < 4; i++) { xc += x[i][0]; yc += x[i][1]; zc += x[i][2]; }
        xc /= 4; yc /= 4; zc /= 4;
        g->face_center[3 * f] = xc;
        g->face_center[3 * f + 1] = yc;
        g->face_center[3 * f + 2] = zc;
        double S = 0.0;
        tri_area_add(x[0], x[1], x[2], &S);
        tri_area_add(x[0], x[3], x[2], &S);
        g->face_area[f] = S;
        double x1 = (x[0][0] + x[1][0]) / 2.0, y1 = (x[0][1] + x[1][1]) / 2.0, z1 = (x[0][2] + x[1][2]) / 2.0;
        double x2 = (x[1][0] + x[2][0]) / 2.0, y2 = (x[1][1] + x[2][1]) / 2.0, z2 = (x[1][2] + x[2][2]) / 2.0;
        double x3 = (x[2][0] + x[3][0]) / 2.0, y3 = (x[2][1] + x[3][1]) / 2.0, z3 = (x[2][2] + x[3][2]) / 2.0;
        double ax = x2 - x1, ay = y2 - y1, az = z2 - z1;
        double bx = x3 - x1, by = y3 - y1, bz = z3 - z1;
        double n1 = ay * bz - az * by;
        double n2 = az * bx - ax * bz;
        double n3 = ax * by - ay * bx;
        double nn = sqrt(KVV(n1, n2, n3));
        n1 /= nn; n2 /= nn; n3 /= nn;
        const double *cc = g->cell_center + 3 * (int64_t)m->face_cells[2 * f];
        if (spf(n1, n2, n3, cc[0] - xc, cc[1] - yc, cc[2] - zc) > 0.0) { n1 = -n1; n2 = -n2; n3 = -n3; }
        g->face_normal[3 * f] = n1;
        g->face_normal[3 * f + 1] = n2;
        g->face_normal[3 * f + 2] = n3;
    }
    /* Cell::Culc_volume, method 1 (Cell.cpp:958-1000): 2 tets per face, faces ascending */
    for (int64_t c = 0; c < m->n_cells; c++) {
        const double *cc = g->cell_center + 3 * c;
        double V = 0.0;
        for (int j = 0; j < 6; j++) {
            int64_t f = m->cell_faces[6 * c + j];
            const double *x0 = X + 3 * (int64_t)m->face_nodes[4 * f + 0];
            const double *x1 = X + 3 * (int64_t)m->face_nodes[4 * f + 1];
            const double *x2 = X + 3 * (int64_t)m->face_nodes[4 * f + 2];
            const double *x3 = X + 3 * (int64_t)m->face_nodes[4 * f + 3];
            V += tet_volume(cc, x0, x1, x2);
            V += tet_volume(cc, x0, x3, x2);
        }
        g->cell_volume[c] = V;
    }
}

/* Geometry, exactly-conservative mode (not in the reference): face vector area
 * S = 1/2 (x2-x0) x (x3-x1) (exact for a bilinear quad), unit normal S/|S| oriented
 * cells[0] -> cells[1], area |S|, and cell volume V = 1/3 sum_f s_f xbar_f . S_f, which is exact
 * for hexahedra with bilinear faces.  Cell and face centres are the reference's node means. */
void or_measure_gcl(const or_mesh *m, const double *X, or_geom *g)
{
    for (int64_t c = 0; c < m->n_cells; c++) {
        double xc = 0.0, yc = 0.0, zc = 0.0;
        for (int i = 0; i < 8; i++) {
            const double *x = X + 3 * (int64_t)m->cell_nodes[8 * c + i];
            xc += x[0]; yc += x[1]; zc += x[2];
        }
        g->cell_center[3 * c] = xc / 8;
        g->cell_center[3 * c + 1] = yc / 8;
        g->cell_center[3 * c + 2] = zc / 8;
    }
    for (int64_t f = 0; f < m->n_faces; f++) {
        const double *x[4];
        for (int i = 0; i < 4; i++) x[i] = X + 3 * (int64_t)m->face_nodes[4 * f + i];
        double xc = 0.0, yc = 0.0, zc = 0.0;
        for (int i = 0; i < 4; i++) { xc += x[i][0]; yc += x[i][1]; zc += x[i][2]; }
        xc /= 4; yc /= 4; zc /= 4;
        g->face_center[3 * f] = xc;
        g->face_center[3 * f + 1] = yc;
        g->face_center[3 * f + 2] = zc;
        double ax = x[2][0] - x[0][0], ay = x[2][1] - x[0][1], az = x[2][2] - x[0][2];
        double bx = x[3][0] - x[1][0], by = x[3][1] - x[1][1], bz = x[3][2] - x[1][2];
        double s1 = 0.5 * (ay * bz - az * by);
        double s2 = 0.5 * (az * bx - ax * bz);
        double s3 = 0.5 * (ax * by - ay * bx);
        double A = sqrt(KVV(s1, s2, s3));
        double n1 = s1 / A, n2 = s2 / A, n3 = s3 / A;
        const double *cc = g->cell_center + 3 * (int64_t)m->face_cells[2 * f];
        if (spf(n1, n2, n3, cc[0] - xc, cc[1] - yc, cc[2] - zc) > 0.0) { n1 = -n1; n2 = -n2; n3 = -n3; }
        g->face_normal[3 * f] = n1;
        g->face_normal[3 * f + 1] = n2;
        g->face_normal[3 * f + 2] = n3;
        g->face_area[f] = A;
    }
    for (int64_t c = 0; c < m->n_cells; c++) {
        double V = 0.0;
        for (int j = 0; j < 6; j++) {
            int64_t f = m->cell_faces[6 * c + j];
            double s = (m->face_cells[2 * f] == c) ? 1.0 : -1.0;
            const double *n = g->face_normal + 3 * f;
            const double *fc = g->face_center + 3 * f;
            double A = g->face_area[f];
            V += s * (spf(fc[0], fc[1], fc[2], n[0], n[1], n[2]) * A);
        }
        g->cell_volume[c] = V / 3.0;
    }
}

/* Gran::culc_velosity, Gran.cpp:571-589 */
double or_face_velocity(const or_mesh *m, int64_t f, const double *x1, const double *x2,
                        const double *normal1, double dt)
{
    const double *n = normal1 + 3 * f;
    double ddot = 0.0;
    for (int i = 0; i < 4; i++) {
        int64_t v = m->face_nodes[4 * f + i];
        double d0 = x2[3 * v] - x1[3 * v];
        double d1 = x2[3 * v + 1] - x1[3 * v + 1];
        double d2 = x2[3 * v + 2] - x1[3 * v + 2];
        ddot = ddot + ((d0 * n[0] + d1 * n[1]) + d2 * n[2]);
    }
    ddot = ddot / 4 / dt;
    return ddot;
}

/* Exactly-conservative face velocity: w = (volume swept by the face) / (dt * area_old).
 * The swept region is the hexahedron (old face, new face) with bilinear sides; its volume is
 * 1/3 sum over its 6 faces of xbar . S. */
static double hex_volume_exact(const double P[8][3])
{
    static const int lf[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 7, 3}};
    double V = 0.0;
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
    return V / 3.0;
}

static double or_face_velocity_gcl(const or_mesh *m, int64_t f, const double *x1, const double *x2,
                                   const double *normal1, const double *area1, double dt)
{
    /* hexahedron: bottom = old face (nodes 0..3), top = new face (4..7); oriented so that the
     * outward normal of the old-face side is -n (the face sweeps along +n for positive w) */
    double P[8][3];
    for (int i = 0; i < 4; i++) {
        int64_t v = m->face_nodes[4 * f + i];
        for (int d = 0; d < 3; d++) { P[i][d] = x1[3 * v + d]; P[4 + i][d] = x2[3 * v + d]; }
    }
    double V = hex_volume_exact(P);
    /* orientation: local face 0 = (0,3,2,1) of the old nodes; its vector area is -S_old(0,1,2,3).
     * The sign of V is + when the nodes move along the node-order normal of (0,1,2,3). */
    const double *n = normal1 + 3 * f;
    const double *x0 = P[0], *xa = P[1], *xb = P[2], *xd = P[3];
    double ax = xb[0] - x0[0], ay = xb[1] - x0[1], az = xb[2] - x0[2];
    double bx = xd[0] - xa[0], by = xd[1] - xa[1], bz = xd[2] - xa[2];
    double s1 = ay * bz - az * by, s2 = az * bx - ax * bz, s3 = ax * by - ay * bx;
    double sg = (spf(s1, s2, s3, n[0], n[1], n[2]) >= 0.0) ? 1.0 : -1.0;
    return (sg * V) / (dt * area1[f]);
}

/* All face velocities of a mesh between two node layers (reference or exactly conservative). */
void or_face_velocities(const or_mesh *m, const double *x1, const double *x2, const double *normal1,
                        const double *area1, double dt, int gcl, double *w)
{
    for (int64_t f = 0; f < m->n_faces; f++)
        w[f] = gcl ? or_face_velocity_gcl(m, f, x1, x2, normal1, area1, dt)
                   : or_face_velocity(m, f, x1, x2, normal1, dt);
}

/* Deterministic cosine (Cody-Waite reduction by pi/2 in three parts, Taylor polynomials on
 * [-pi/4, pi/4]).  Used only by the prescribed node motion so the CPU and the GPU move nodes
 * identically; accurate to a few ulp for |x| < 1e5. */
double or_det_cos(double x)
{
    const double two_over_pi = 0.63661977236758134308;
    const double p1 = 1.5707963267341256e+00;
    const double p2 = 6.0771005065061922e-11;
    const double p3 = 2.0222662487959506e-21;
    double kf = nearbyint(x * two_over_pi);
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

void or_motion(int64_t n_nodes, const double *x_old, const double *disp, double dt, double t,
               double omega, double phase, double *x_new)
{
    double fac = dt * (omega * or_det_cos(omega * t + phase));
    for (int64_t i = 0; i < 3 * n_nodes; i++) x_new[i] = x_old[i] + fac * disp[i];
}

/* ------------------------------------------------------------------------------------------ */
/* Face states: Setka::Snos_on_Gran, first order (TVD off) and boundary faces.                */

static double polar_angle(double x, double y)
{
    /* Help_funk.cpp:78-106 */
    if (fabs(x) + fabs(y) < 0.000001) return 0.0;
    if (x < 0) return atan(y / x) + 1.0 * OR_PI;
    else if (x > 0 && y >= 0) return atan(y / x);
    else if (x > 0 && y < 0) return atan(y / x) + 2.0 * OR_PI;
    else if (y > 0 && x >= 0 && x <= 0) return OR_PI / 2.0;
    else if (y < 0 && x >= 0 && x <= 0) return 3.0 * OR_PI / 2.0;
    return 0.0;
}

static void spherical_skorost(double z, double x, double y, double Vz, double Vx, double Vy,
                              double *Vr, double *Vphi, double *Vtheta)
{
    /* Help_funk.cpp:122-135 */
    double r_1 = sqrt(x * x + y * y + z * z);
    double the_1 = acos(z / r_1);
    double phi_1 = polar_angle(x, y);
    *Vr = Vx * sin(the_1) * cos(phi_1) + Vy * sin(the_1) * sin(phi_1) + Vz * cos(the_1);
    *Vtheta = Vx * cos(the_1) * cos(phi_1) + Vy * cos(the_1) * sin(phi_1) - Vz * sin(the_1);
    *Vphi = -Vx * sin(phi_1) + Vy * cos(phi_1);
}

static void dekard_skorost(double z, double x, double y, double Vr, double Vphi, double Vtheta,
                           double *Vz, double *Vx, double *Vy)
{
    /* Help_funk.cpp:108-120 */
    double r_2 = sqrt(x * x + y * y + z * z);
    double the_2 = acos(z / r_2);
    double phi_2 = polar_angle(x, y);
    *Vx = Vr * sin(the_2) * cos(phi_2) + Vtheta * cos(the_2) * cos(phi_2) - Vphi * sin(phi_2);
    *Vy = Vr * sin(the_2) * sin(phi_2) + Vtheta * cos(the_2) * sin(phi_2) + Vphi * cos(phi_2);
    *Vz = Vr * cos(the_2) - Vtheta * sin(the_2);
}

/* Zone_1 spherical rescaling of a cell state to the face centre (Setka_gran_snos.cpp:744-873). */
static void zone1_state(const double *s, const double *Ac, const double *G, double gamma, double *out)
{
    double r1 = sqrt((Ac[0] * Ac[0] + Ac[1] * Ac[1]) + Ac[2] * Ac[2]);
    double rr = sqrt((G[0] * G[0] + G[1] * G[1]) + G[2] * G[2]);
    double the1 = acos(G[2] / rr);
    for (int ik = 0; ik < 2; ik++) {
        const double *V = s + (ik == 0 ? 1 : 5);
        double *O = out + (ik == 0 ? 1 : 5);
        double VA0, VA1, VA2;
        if (the1 > OR_PI / 9 && the1 < 8 * OR_PI / 9) {
            spherical_skorost(Ac[2], Ac[0], Ac[1], V[2], V[0], V[1], &VA0, &VA1, &VA2);
            if (ik == 1) {
                VA0 = VA0 * (KV(r1 / rr));
                VA1 = VA1 * ((r1) / (rr));
                VA2 = VA2 * ((r1) / (rr));
            }
            dekard_skorost(G[2], G[0], G[1], VA0, VA1, VA2, &O[2], &O[0], &O[1]);
        } else {
            spherical_skorost(Ac[0], Ac[1], Ac[2], V[0], V[1], V[2], &VA0, &VA1, &VA2);
            if (ik == 1) {
                VA0 = VA0 * (KV(r1 / rr));
                VA1 = VA1 * ((r1) / (rr));
                VA2 = VA2 * ((r1) / (rr));
            }
            dekard_skorost(G[0], G[1], G[2], VA0, VA1, VA2, &O[0], &O[1], &O[2]);
        }
    }
    out[0] = s[0] * KV(r1) / KV(rr);
    out[4] = s[4] * pow(r1, 2.0 * gamma) / pow(rr, 2.0 * gamma);
}

/* ------------------------------------------------------------------------------------------ */
/* TVD: neighbour choice (Setka::Init_TVD, Setka_physics.cpp:648-744) and the limited
 * reconstruction of Setka::Snos_on_Gran (Setka_gran_snos.cpp:186-736), limiter Help_funk.cpp:26-70. */

static int or_signum(double x)
{
    if (x > 0.00000001) return 1;
    else if (x < -0.00000001) return -1;
    return 0;
}

double or_minmod(double x, double y)
{
    if (or_signum(x) + or_signum(y) == 0) return 0.0;
    double ax = fabs(x), ay = fabs(y);
    return ((or_signum(x) + or_signum(y)) / 2.0) * ((ay < ax) ? ay : ax);  /* std::min */
}

double or_linear(double x1, double t1, double x2, double t2, double x3, double t3, double y)
{
    double d = or_minmod((t1 - t2) / (x1 - x2), (t2 - t3) / (x2 - x3));
    return (d * (y - x2) + t2);
}

static double or_linear2(double x1, double t1, double x2, double t2, double y)
{
    double d = (t1 - t2) / (x1 - x2);
    return (d * (y - x2) + t2);
}

/* unit vector (Eigen normalized(): unchanged when the norm is 0) */
static void or_normalize(double *v)
{
    double n = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    if (n > 0.0) { v[0] /= n; v[1] /= n; v[2] /= n; }
}

/* The neighbour of cell C across its face at slot k, or -1 (Help_funk.cpp:275-288). */
static int64_t or_sosed(const or_mesh *m, int64_t C, int64_t gr)
{
    int64_t a = m->face_cells[2 * gr], b = m->face_cells[2 * gr + 1];
    if (b < 0) return -1;
    return a == C ? b : a;
}

/* Among C's other neighbours, the one whose centre is seen from the face centre most opposite to
 * vec (smallest cosine; first on ties). */
static int64_t or_tvd_pick(const or_mesh *m, const or_geom *g0, int64_t f, int64_t C, const double *vec)
{
    double best = 2.0;
    int64_t pick = -1;
    const double *G = g0->face_center + 3 * f;
    for (int k = 0; k < 6; k++) {
        int64_t gr = m->cell_faces[6 * C + k];
        if (gr == f) continue;
        int64_t cel = or_sosed(m, C, gr);
        if (cel < 0) continue;
        double v[3] = {g0->cell_center[3 * cel] - G[0], g0->cell_center[3 * cel + 1] - G[1],
                       g0->cell_center[3 * cel + 2] - G[2]};
        or_normalize(v);
        double cosine = (vec[0] * v[0] + vec[1] * v[1]) + vec[2] * v[2];
        if (cosine < best) { best = cosine; pick = cel; }
    }
    return pick;
}

void or_init_tvd(const or_mesh *m, const or_geom *g0, int32_t *tvd)
{
    for (int64_t f = 0; f < m->n_faces; f++) {
        const double *n = g0->face_normal + 3 * f;
        const double *G = g0->face_center + 3 * f;
        int64_t AA = or_tvd_pick(m, g0, f, m->face_cells[2 * f], n);
        tvd[2 * f] = (int32_t)AA;
        tvd[2 * f + 1] = -1;
        if (m->face_cells[2 * f + 1] < 0) continue;  /* boundary face: no second pick, no check */
        double mn[3] = {-n[0], -n[1], -n[2]};
        int64_t BB = or_tvd_pick(m, g0, f, m->face_cells[2 * f + 1], mn);
        tvd[2 * f + 1] = (int32_t)BB;
        /* drop a neighbour that is not behind its cell (Setka_physics.cpp:724-738) */
        if (AA >= 0) {
            double k3[3] = {g0->cell_center[3 * AA] - G[0], g0->cell_center[3 * AA + 1] - G[1],
                            g0->cell_center[3 * AA + 2] - G[2]};
            or_normalize(k3);
            if (-((n[0] * k3[0] + n[1] * k3[1]) + n[2] * k3[2]) < 0.1) tvd[2 * f] = -1;
        }
        if (BB >= 0) {
            double k3[3] = {g0->cell_center[3 * BB] - G[0], g0->cell_center[3 * BB + 1] - G[1],
                            g0->cell_center[3 * BB + 2] - G[2]};
            or_normalize(k3);
            if (((n[0] * k3[0] + n[1] * k3[1]) + n[2] * k3[2]) < 0.1) tvd[2 * f + 1] = -1;
        }
    }
}

static double or_norm3(const double *v) { return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); }

/* Spherical components (Vr, Vphi, Vtheta) of the V (ik = 0) or B (ik = 1) of state s at centre c in
 * the face's frame: equatorial roles (z,x,y) = (2,0,1), polar (0,1,2). */
static void or_sph(const double *s, const double *c, int ik, int equatorial, double *o)
{
    const double *V = s + (ik == 0 ? 1 : 5);
    if (equatorial) spherical_skorost(c[2], c[0], c[1], V[2], V[0], V[1], &o[0], &o[1], &o[2]);
    else spherical_skorost(c[0], c[1], c[2], V[0], V[1], V[2], &o[0], &o[1], &o[2]);
}

static void or_cart(const double *G, int equatorial, const double *Vs, int ik, double *out)
{
    double *O = out + (ik == 0 ? 1 : 5);
    if (equatorial) dekard_skorost(G[2], G[0], G[1], Vs[0], Vs[1], Vs[2], &O[2], &O[0], &O[1]);
    else dekard_skorost(G[0], G[1], G[2], Vs[0], Vs[1], Vs[2], &O[0], &O[1], &O[2]);
}

/* Second-order states of an interior face whose four cells (AA, A | B, BB) are known; returns 0 if
 * the face keeps the first-order states. */
/* The passive scalars Q, rho_He are r^2 names like rho (Phys_param.cpp:284-286), so each zone case
 * of Setka_gran_snos.cpp:216-713 reconstructs them by rho's line: values a, b, c, d of AA, A, B,
 * BB; radii r3, r1, rr, r2, r4 of AA, A, G, B, BB.  (In the cases 1112 and 1122 the r^2 assignment
 * is followed by a second if-chain whose final else overwrites it with the plain line.) */
static void or_tvd_r2(int zAA, int zA, int zB, int zBB, double a, double b, double c, double d, double d1,
                      double d2, double dd1, double dd2, double r3, double r1, double rr, double r2, double r4,
                      double *l, double *r)
{
    if (zAA == 1 && zA == 1 && zB == 1 && zBB == 1) { /* :364-374 */
        *l = or_linear(-dd1 - d1, a * KV(r3), -d1, b * KV(r1), d2, c * KV(r2), 0.0) / KV(rr);
        *r = or_linear(-d1, b * KV(r1), d2, c * KV(r2), d2 + dd2, d * KV(r4), 0.0) / KV(rr);
    } else if (zAA == 1 && zA == 1 && zB == 1 && zBB == 2) { /* :502-534, final else */
        if (MISTAKE(OR_MISTAKE_SCAL_TVD_R2_KEPT)) {
            *l = or_linear(-dd1 - d1, a * KV(r3), -d1, b * KV(r1), d2, c * KV(r2), 0.0) / KV(rr);
            *r = or_linear2(-d1, b * KV(r1), d2, c * KV(r2), 0.0) / KV(rr);
            return;
        }
        *l = or_linear(-dd1 - d1, a, -d1, b, d2, c, 0.0);
        *r = or_linear2(-d1, b, d2, c, 0.0);
    } else if (zAA == 1 && zA == 1 && zB == 2 && zBB == 2) { /* :618-646, final else */
        if (MISTAKE(OR_MISTAKE_SCAL_TVD_R2_KEPT)) {
            *l = or_linear2(-dd1 - d1, a * KV(r3), -d1, b * KV(r1), 0.0) / KV(rr);
            *r = or_linear2(d2, c, d2 + dd2, d, 0.0);
            return;
        }
        *l = or_linear2(-dd1 - d1, a, -d1, b, 0.0);
        *r = or_linear2(d2, c, d2 + dd2, d, 0.0);
    } else if (zAA == 1 && zA == 2 && zB == 2 && zBB == 2) { /* :649-660 */
        *l = or_linear2(-d1, b, d2, c, 0.0);
        *r = or_linear(-d1, b, d2, c, d2 + dd2, d, 0.0);
    } else if ((zAA == 2 && zA == 2 && zB == 3 && zBB == 3) || (zAA == 3 && zA == 3 && zB == 4 && zBB == 4)) {
        *l = or_linear2(-dd1 - d1, a, -d1, b, 0.0); /* :661-673 */
        *r = or_linear2(d2, c, d2 + dd2, d, 0.0);
    } else if ((zAA == 2 && zA == 2 && zB == 2 && zBB == 3) || (zAA == 3 && zA == 3 && zB == 3 && zBB == 4)) {
        *l = or_linear(-dd1 - d1, a, -d1, b, d2, c, 0.0); /* :674-687 */
        *r = or_linear2(-d1, b, d2, c, 0.0);
    } else if ((zAA == 2 && zA == 3 && zB == 3 && zBB == 3) || (zAA == 3 && zA == 4 && zB == 4 && zBB == 4)) {
        *l = or_linear2(-d1, b, d2, c, 0.0); /* :688-701 */
        *r = or_linear(-d1, b, d2, c, d2 + dd2, d, 0.0);
    } else { /* :702-713 */
        *l = or_linear(-dd1 - d1, a, -d1, b, d2, c, 0.0);
        *r = or_linear(-d1, b, d2, c, d2 + dd2, d, 0.0);
    }
}

static int or_tvd_states(const or_mesh *m, const or_params *p, int64_t f, const double *prim,
                         const double *scal, const or_geom *g, double *qL, double *qR, double *kL, double *kR)
{
    int64_t A = m->face_cells[2 * f], B = m->face_cells[2 * f + 1];
    if (!p->tvd || !m->face_tvd || !m->cell_is_tvd[A] || !m->cell_is_tvd[B]) return 0;
    /* heliopause faces stay first order unless Snos_on_HP (Setka_gran_snos.cpp:188-189) */
    if (m->face_flags && (m->face_flags[f] & 4) && !p->snos_on_hp) return 0;
    int64_t AA = m->face_tvd[2 * f], BB = m->face_tvd[2 * f + 1];
    if (AA < 0 || BB < 0) return 0;
    const double *sAA = prim + 8 * AA, *sA = prim + 8 * A, *sB = prim + 8 * B, *sBB = prim + 8 * BB;
    const double *G = g->face_center + 3 * f;
    const double *Ac = g->cell_center + 3 * A, *Bc = g->cell_center + 3 * B;
    const double *AAc = g->cell_center + 3 * AA, *BBc = g->cell_center + 3 * BB;
    double v[3];
    v[0] = Ac[0] - G[0]; v[1] = Ac[1] - G[1]; v[2] = Ac[2] - G[2];
    double d1 = or_norm3(v);
    v[0] = Bc[0] - G[0]; v[1] = Bc[1] - G[1]; v[2] = Bc[2] - G[2];
    double d2 = or_norm3(v);
    v[0] = Ac[0] - AAc[0]; v[1] = Ac[1] - AAc[1]; v[2] = Ac[2] - AAc[2];
    double dd1 = or_norm3(v);
    v[0] = Bc[0] - BBc[0]; v[1] = Bc[1] - BBc[1]; v[2] = Bc[2] - BBc[2];
    double dd2 = or_norm3(v);
    int zAA = m->cell_zone[AA], zA = m->cell_zone[A], zB = m->cell_zone[B], zBB = m->cell_zone[BB];
    double gam2 = 2.0 * p->gamma;
    if (zAA == 1 && zA == 1 && zB == 1 && zBB == 1) {
        /* Setka_gran_snos.cpp:216-400: spherical components; B radial x r^2, tangential x r;
         * rho x r^2, p x r^(2 gamma) */
        double r3 = or_norm3(AAc), r1 = or_norm3(Ac), rr = or_norm3(G), r2 = or_norm3(Bc), r4 = or_norm3(BBc);
        double the1 = acos(G[2] / rr);
        int eq = (the1 > OR_PI / 9 && the1 < 8 * OR_PI / 9);
        for (int ik = 0; ik < 2; ik++) {
            double VAA[3], VA[3], VB[3], VBB[3], Vl[3], Vr[3];
            or_sph(sAA, AAc, ik, eq, VAA); or_sph(sA, Ac, ik, eq, VA);
            or_sph(sB, Bc, ik, eq, VB); or_sph(sBB, BBc, ik, eq, VBB);
            for (int i = 0; i < 3; i++) {
                if (ik == 1) {
                    if (i == 0) {
                        Vl[i] = or_linear(-dd1 - d1, VAA[i] * KV(r3), -d1, VA[i] * KV(r1), d2, VB[i] * KV(r2), 0.0) / KV(rr);
                        Vr[i] = or_linear(-d1, VA[i] * KV(r1), d2, VB[i] * KV(r2), d2 + dd2, VBB[i] * KV(r4), 0.0) / KV(rr);
                    } else {
                        Vl[i] = or_linear(-dd1 - d1, VAA[i] * (r3), -d1, VA[i] * (r1), d2, VB[i] * (r2), 0.0) / (rr);
                        Vr[i] = or_linear(-d1, VA[i] * (r1), d2, VB[i] * (r2), d2 + dd2, VBB[i] * (r4), 0.0) / (rr);
                    }
                } else {
                    Vl[i] = or_linear(-dd1 - d1, VAA[i], -d1, VA[i], d2, VB[i], 0.0);
                    Vr[i] = or_linear(-d1, VA[i], d2, VB[i], d2 + dd2, VBB[i], 0.0);
                }
            }
            or_cart(G, eq, Vl, ik, qL);
            or_cart(G, eq, Vr, ik, qR);
        }
        qL[0] = or_linear(-dd1 - d1, sAA[0] * KV(r3), -d1, sA[0] * KV(r1), d2, sB[0] * KV(r2), 0.0) / KV(rr);
        qR[0] = or_linear(-d1, sA[0] * KV(r1), d2, sB[0] * KV(r2), d2 + dd2, sBB[0] * KV(r4), 0.0) / KV(rr);
        qL[4] = or_linear(-dd1 - d1, sAA[4] * pow(r3, gam2), -d1, sA[4] * pow(r1, gam2), d2, sB[4] * pow(r2, gam2), 0.0) / pow(rr, gam2);
        qR[4] = or_linear(-d1, sA[4] * pow(r1, gam2), d2, sB[4] * pow(r2, gam2), d2 + dd2, sBB[4] * pow(r4, gam2), 0.0) / pow(rr, gam2);
    } else if (zAA == 1 && zA == 1 && zB == 1 && zBB == 2) {
        /* :401-536: spherical V and B (unscaled), left limited over AA, A, B, right the A-B line;
         * rho plain (the reference's r^2 assignment is overwritten by its final else), p x r^(2 gamma) */
        double r3 = or_norm3(AAc), r1 = or_norm3(Ac), rr = or_norm3(G), r2 = or_norm3(Bc);
        double the1 = acos(G[2] / rr);
        int eq = (the1 > OR_PI / 9 && the1 < 8 * OR_PI / 9);
        for (int ik = 0; ik < 2; ik++) {
            double VAA[3], VA[3], VB[3], Vl[3], Vr[3];
            or_sph(sAA, AAc, ik, eq, VAA); or_sph(sA, Ac, ik, eq, VA); or_sph(sB, Bc, ik, eq, VB);
            for (int i = 0; i < 3; i++) {
                Vl[i] = or_linear(-dd1 - d1, VAA[i], -d1, VA[i], d2, VB[i], 0.0);
                Vr[i] = or_linear2(-d1, VA[i], d2, VB[i], 0.0);
            }
            or_cart(G, eq, Vl, ik, qL);
            or_cart(G, eq, Vr, ik, qR);
        }
        qL[0] = or_linear(-dd1 - d1, sAA[0], -d1, sA[0], d2, sB[0], 0.0);
        qR[0] = or_linear2(-d1, sA[0], d2, sB[0], 0.0);
        qL[4] = or_linear(-dd1 - d1, sAA[4] * pow(r3, gam2), -d1, sA[4] * pow(r1, gam2), d2, sB[4] * pow(r2, gam2), 0.0) / pow(rr, gam2);
        qR[4] = or_linear2(-d1, sA[4] * pow(r1, gam2), d2, sB[4] * pow(r2, gam2), 0.0) / pow(rr, gam2);
    } else if (zAA == 1 && zA == 1 && zB == 2 && zBB == 2) {
        /* :537-648: one-sided lines AA-A (left, spherical V and B) and B-BB (right, Cartesian);
         * rho plain, p left x r^(2 gamma) */
        double r3 = or_norm3(AAc), r1 = or_norm3(Ac), rr = or_norm3(G);
        double the1 = acos(G[2] / rr);
        int eq = (the1 > OR_PI / 9 && the1 < 8 * OR_PI / 9);
        for (int ik = 0; ik < 2; ik++) {
            double VAA[3], VA[3], Vl[3];
            or_sph(sAA, AAc, ik, eq, VAA); or_sph(sA, Ac, ik, eq, VA);
            for (int i = 0; i < 3; i++) Vl[i] = or_linear2(-dd1 - d1, VAA[i], -d1, VA[i], 0.0);
            or_cart(G, eq, Vl, ik, qL);
        }
        for (int k = 1; k < 8; k++) {
            if (k == 4) continue;
            qR[k] = or_linear2(d2, sB[k], d2 + dd2, sBB[k], 0.0);
        }
        qL[0] = or_linear2(-dd1 - d1, sAA[0], -d1, sA[0], 0.0);
        qR[0] = or_linear2(d2, sB[0], d2 + dd2, sBB[0], 0.0);
        qL[4] = or_linear2(-dd1 - d1, sAA[4] * pow(r3, gam2), -d1, sA[4] * pow(r1, gam2), 0.0) / pow(rr, gam2);
        qR[4] = or_linear2(d2, sB[4], d2 + dd2, sBB[4], 0.0);
    } else if (zAA == 1 && zA == 2 && zB == 2 && zBB == 2) {  /* :649-660 */
        for (int k = 0; k < 8; k++) {
            qL[k] = or_linear2(-d1, sA[k], d2, sB[k], 0.0);
            qR[k] = or_linear(-d1, sA[k], d2, sB[k], d2 + dd2, sBB[k], 0.0);
        }
    } else if ((zAA == 2 && zA == 2 && zB == 3 && zBB == 3) || (zAA == 3 && zA == 3 && zB == 4 && zBB == 4)) {  /* :661-673 */
        for (int k = 0; k < 8; k++) {
            qL[k] = or_linear2(-dd1 - d1, sAA[k], -d1, sA[k], 0.0);
            qR[k] = or_linear2(d2, sB[k], d2 + dd2, sBB[k], 0.0);
        }
    } else if ((zAA == 2 && zA == 2 && zB == 2 && zBB == 3) || (zAA == 3 && zA == 3 && zB == 3 && zBB == 4)) {  /* :674-687 */
        for (int k = 0; k < 8; k++) {
            qL[k] = or_linear(-dd1 - d1, sAA[k], -d1, sA[k], d2, sB[k], 0.0);
            qR[k] = or_linear2(-d1, sA[k], d2, sB[k], 0.0);
        }
    } else if ((zAA == 2 && zA == 3 && zB == 3 && zBB == 3) || (zAA == 3 && zA == 4 && zB == 4 && zBB == 4)) {  /* :688-701 */
        for (int k = 0; k < 8; k++) {
            qL[k] = or_linear2(-d1, sA[k], d2, sB[k], 0.0);
            qR[k] = or_linear(-d1, sA[k], d2, sB[k], d2 + dd2, sBB[k], 0.0);
        }
    } else {  /* :702-713 */
        for (int k = 0; k < 8; k++) {
            qL[k] = or_linear(-dd1 - d1, sAA[k], -d1, sA[k], d2, sB[k], 0.0);
            qR[k] = or_linear(-d1, sA[k], d2, sB[k], d2 + dd2, sBB[k], 0.0);
        }
    }
    /* positivity (Setka_gran_snos.cpp:715-725) */
    if (qL[0] < 0.0000001) qL[0] = sA[0];
    if (qL[4] < 0.0000001) qL[4] = sA[4];
    if (qR[0] < 0.0000001) qR[0] = sB[0];
    if (qR[4] < 0.0000001) qR[4] = sB[4];
    if (scal) {
        const double r3 = or_norm3(AAc), r1 = or_norm3(Ac), rr = or_norm3(G), r2 = or_norm3(Bc), r4 = or_norm3(BBc);
        for (int j = 0; j < 2; j++) {
            const double kA = scal[2 * A + j], kB = scal[2 * B + j];
            or_tvd_r2(zAA, zA, zB, zBB, scal[2 * AA + j], kA, kB, scal[2 * BB + j], d1, d2, dd1, dd2, r3, r1, rr, r2,
                      r4, &kL[j], &kR[j]);
            if (kL[j] < 0.0000001) kL[j] = kA; /* :718-723 */
            if (kR[j] < 0.0000001) kR[j] = kB;
        }
    }
    return 1;
}

int or_face_states_k(const or_mesh *m, const or_params *p, int64_t f, const double *prim, const double *scal,
                     const or_geom *g, double *qL, double *qR, double *kL, double *kR)
{
    int bc = m->face_bc[f];
    int64_t A = m->face_cells[2 * f];
    const double *sA = prim + 8 * A;
    if (scal && bc != OR_BC_INTERIOR) { /* the plasma names of the cell (and of the face's fixed
                                            state): Setka_gran_snos.cpp:20-57 */
        for (int j = 0; j < 2; j++) {
            kL[j] = scal[2 * A + j];
            kR[j] = (bc == OR_BC_INNER_HARD || bc == OR_BC_OUTER_HARD) ? m->face_bscal[2 * f + j] : kL[j];
        }
    }
    if (bc != OR_BC_INTERIOR) {
        const double *n = g->face_normal + 3 * f;
        if (bc == OR_BC_OUTER_SOFT) { /* plasma_condition 1: Setka_gran_snos.cpp:18-50 */
            for (int i = 0; i < 8; i++) qL[i] = sA[i];
            if (n[0] < -0.95) {
                if (qL[1] > p->v_inf / 7.0) qL[1] = p->v_inf / 5.0;
            } else {
                if (qL[1] * n[0] + qL[2] * n[1] + qL[3] * n[2] < 0.0) {
                    qL[1] = 0.1 * n[0];
                    qL[2] = 0.1 * n[1];
                    qL[3] = 0.1 * n[2];
                }
            }
            for (int i = 0; i < 8; i++) qR[i] = qL[i];
        } else if (bc == OR_BC_INNER_HARD || bc == OR_BC_OUTER_HARD) { /* plasma_condition 2 */
            for (int i = 0; i < 8; i++) { qL[i] = sA[i]; qR[i] = m->face_bstate[8 * f + i]; }
        } else if (bc == OR_BC_ZERO_GRADIENT) {
            for (int i = 0; i < 8; i++) { qL[i] = sA[i]; qR[i] = sA[i]; }
        } else {
            return -10;
        }
        return 0;
    }
    int64_t B = m->face_cells[2 * f + 1];
    const double *sB = prim + 8 * B;
    const double *G = g->face_center + 3 * f;
    if (or_tvd_states(m, p, f, prim, scal, g, qL, qR, kL, kR)) return 0;
    if (m->cell_zone[A] == 1) zone1_state(sA, g->cell_center + 3 * A, G, p->gamma, qL);
    else for (int i = 0; i < 8; i++) qL[i] = sA[i];
    if (m->cell_zone[B] == 1) zone1_state(sB, g->cell_center + 3 * B, G, p->gamma, qR);
    else for (int i = 0; i < 8; i++) qR[i] = sB[i];
    if (scal) { /* Q and rho_He like rho: x kv(r1) / kv(rr) from a Zone_1 cell (Setka_gran_snos.cpp:848-852,
                   1005-1009), else the cell's own (:877-880, 1037-1040) */
        const double rr = sqrt((G[0] * G[0] + G[1] * G[1]) + G[2] * G[2]);
        const double *Ac = g->cell_center + 3 * A, *Bc = g->cell_center + 3 * B;
        const double rA = sqrt((Ac[0] * Ac[0] + Ac[1] * Ac[1]) + Ac[2] * Ac[2]);
        const double rB = sqrt((Bc[0] * Bc[0] + Bc[1] * Bc[1]) + Bc[2] * Bc[2]);
        const int plain = MISTAKE(OR_MISTAKE_SCAL_ZONE1_PLAIN);
        for (int j = 0; j < 2; j++) {
            kL[j] = m->cell_zone[A] == 1 && !plain ? scal[2 * A + j] * KV(rA) / KV(rr) : scal[2 * A + j];
            kR[j] = m->cell_zone[B] == 1 && !plain ? scal[2 * B + j] * KV(rB) / KV(rr) : scal[2 * B + j];
        }
    }
    return 0;
}

int or_face_states(const or_mesh *m, const or_params *p, int64_t f, const double *prim,
                   const or_geom *g, double *qL, double *qR)
{
    return or_face_states_k(m, p, f, prim, NULL, g, qL, qR, NULL, NULL);
}

/* Setka::Culc_Gran_Potok (Setka_physics.cpp:2280-2571), all faces.  x_now1/x_now2 are used only
 * when the mesh moves (face velocity).  gcl = 0 reference geometry. */
/* Setka::Go's areas (Setka_physics.cpp:113-212, 1526-1537): the cells of area 1 are the inner
 * ones, of area 2 the others; a face is on an area's list when one of its cells is in the area. */
static int or_cell_in_area(const or_mesh *m, int64_t c, int area)
{
    if (area == 0) return 1;
    int in = m->cell_inner && m->cell_inner[c];
    return area == 1 ? in : !in;
}

static int or_face_in_area(const or_mesh *m, int64_t f, int area)
{
    if (area == 0) return 1;
    int64_t a = m->face_cells[2 * f], b = m->face_cells[2 * f + 1];
    return or_cell_in_area(m, a, area) || (b >= 0 && or_cell_in_area(m, b, area));
}

/* scal (nullable): the passive scalars [n_cells*2] of the layer; then kflux [n_faces*2] = their konvect
 * fluxes * area (Setka_physics.cpp:2349-2361, 2485-2493). */
static int face_flux_impl(const or_mesh *m, const or_params *p, const double *prim, const double *scal,
                          const or_geom *g, const double *x1, const double *x2, double dt, int gcl, double *flux,
                          double *kflux, double *dt_face, uint8_t *branch, int nthreads)
{
    int status = 0;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t f = 0; f < m->n_faces; f++) {
        double qL[8], qR[8], q[8], kL[2], kR[2], k[2] = {0.0, 0.0}, dsl, dsc, dsr;
        int br = 0;
        if (!or_face_in_area(m, f, p->area)) { /* not on the area's face list: not evaluated */
            for (int i = 0; i < 9; i++) flux[9 * f + i] = 0.0;
            if (scal) kflux[2 * f] = kflux[2 * f + 1] = 0.0;
            dt_face[f] = 10000000.0;
            if (branch) branch[f] = 0;
            continue;
        }
        double area = g->face_area[f];
        const double *n = g->face_normal + 3 * f;
        double w = 0.0;
        if (p->move_mesh) {
            w = gcl ? or_face_velocity_gcl(m, f, x1, x2, g->face_normal, g->face_area, dt)
                    : or_face_velocity(m, f, x1, x2, g->face_normal, dt);
        }
        double loc_time = 10000000.0;
        double dist = 0.0;
        int interior = (m->face_bc[f] == OR_BC_INTERIOR);
        if (interior) {
            const double *ca = g->cell_center + 3 * (int64_t)m->face_cells[2 * f];
            const double *cb = g->cell_center + 3 * (int64_t)m->face_cells[2 * f + 1];
            double d0 = ca[0] - cb[0], d1 = ca[1] - cb[1], d2 = ca[2] - cb[2];
            dist = sqrt(KVV(d0, d1, d2)) / 2.0;
        }
        int rc = or_face_states_k(m, p, f, prim, scal, g, qL, qR, kL, kR);
        if (rc == 0)
            rc = or_chlld(m->face_method[f], p->n_disc, m->face_flags ? m->face_flags[f] : 0, n[0], n[1],
                          n[2], w, qL, qR, scal ? 2 : 0, scal ? kL : NULL, scal ? kR : NULL, q, scal ? k : NULL, &dsl,
                          &dsc, &dsr, &br);
        if (rc != 0) {
#pragma omp critical(or_status)
            { if (status == 0) status = rc; }
            for (int i = 0; i < 9; i++) flux[9 * f + i] = 0.0;
            if (scal) kflux[2 * f] = kflux[2 * f + 1] = 0.0;
            dt_face[f] = loc_time;
            if (branch) branch[f] = (uint8_t)br;
            continue;
        }
        if (interior) {
            double dnt = p->kfl * dist / (SMAX(fabs(dsl), fabs(dsr)) + fabs(w));
            if (dnt < loc_time) loc_time = SMIN(loc_time, dnt);
        }
        for (int i = 0; i < 8; i++) flux[9 * f + i] = q[i] * area;
        flux[9 * f + 8] = (MISTAKE(OR_MISTAKE_PDIVB_HALF) ? 1.0 : 0.5) *
                          spf(n[0], n[1], n[2], qL[5] + qR[5], qL[6] + qR[6], qL[7] + qR[7]) * area;
        if (scal) { kflux[2 * f] = k[0] * area; kflux[2 * f + 1] = k[1] * area; }
        dt_face[f] = loc_time;
        if (branch) branch[f] = (uint8_t)br;
    }
    return status;
}

int or_face_flux(const or_mesh *m, const or_params *p, const double *prim, const or_geom *g,
                 const double *x_now1, const double *x_now2, double dt, double *flux, double *dt_face,
                 uint8_t *branch)
{
    return face_flux_impl(m, p, prim, NULL, g, x_now1, x_now2, dt, 0, flux, NULL, dt_face, branch, 1);
}

int or_face_flux_k(const or_mesh *m, const or_params *p, const double *prim, const double *scal, const or_geom *g,
                   const double *x_now1, const double *x_now2, double dt, double *flux, double *kflux,
                   double *dt_face, uint8_t *branch)
{
    return face_flux_impl(m, p, prim, scal, g, x_now1, x_now2, dt, 0, flux, kflux, dt_face, branch, 1);
}

/* Cell loop of Setka::Go (Setka_physics.cpp:1695-2027) for the plasma without PUI: flux sum in the
 * cell's ascending face order, Powell source, FP64 pressure recovery, Zone_2 magnetic decay, floors,
 * B frozen for cells with is_TVD == false; with scal_old (nullable) the passive scalars Q, rho_He
 * advected by their fluxes kflux and clipped (:1781-1817, 2004-2016) into scal_new. */
static int update_impl(const or_mesh *m, const or_params *p, const double *prim_old, const double *scal_old,
                       const double *vol_old, const double *vol_new, const double *flux, const double *kflux,
                       const double *src, double time, double *prim_new, double *scal_new, int64_t *n_rho_floor,
                       int64_t *n_p_floor, int nthreads)
{
    const double cpi4 = 4.0 * OR_PI;
    const double g1 = 5.0 / 3.0 - 1.0;
    int64_t nrf = 0, npf = 0;
    int status = 0;
    (void)nthreads;
    (void)p;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nrf, npf) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t c = 0; c < m->n_cells; c++) {
        if (!or_cell_in_area(m, c, p->area)) { /* outside the area: the state carries over */
            for (int i = 0; i < 8; i++) prim_new[8 * c + i] = prim_old[8 * c + i];
            if (scal_old) { scal_new[2 * c] = scal_old[2 * c]; scal_new[2 * c + 1] = scal_old[2 * c + 1]; }
            continue;
        }
        double Volume = vol_old[c];
        double Volume2 = vol_new[c];
        double P[9];
        for (int i = 0; i < 9; i++) P[i] = 0.0;
        for (int j = 0; j < 6; j++) {
            int64_t f = m->cell_faces[6 * c + j];
            int sign_potok = 1;
            if (m->face_cells[2 * f] != c) sign_potok = -1;
            for (int i = 0; i < 9; i++) P[i] += sign_potok * flux[9 * f + i];
        }
        const double *s = prim_old + 8 * c;
        double rho = s[0], vx = s[1], vy = s[2], vz = s[3], pp = s[4], bx = s[5], by = s[6], bz = s[7];
        double Smx = 0.0, Smy = 0.0, Smz = 0.0, SE = 0.0;
        if (src) { Smx = src[5 * c + 1]; Smy = src[5 * c + 2]; Smz = src[5 * c + 3]; SE = src[5 * c + 4]; }
        double dsk = spf(vx, vy, vz, bx, by, bz);
        double rho3 = rho * Volume / Volume2 - time * P[0] / Volume2;
        int rho_null = 0;
        if (rho3 < 1e-7) { rho_null = 1; rho3 = 0.1; nrf++; }
        double Q3 = 0.0, rho_He3 = 0.0;
        if (scal_old) {
            double PQ = 0.0, PHe = 0.0;
            for (int j = 0; j < 6; j++) {
                int64_t f = m->cell_faces[6 * c + j];
                int sign_potok = 1;
                if (m->face_cells[2 * f] != c) sign_potok = -1;
                PQ += sign_potok * kflux[2 * f];
                PHe += sign_potok * kflux[2 * f + 1];
            }
            const double Q = scal_old[2 * c], rho_He = scal_old[2 * c + 1];
            Q3 = Q * Volume / Volume2 - time * PQ / Volume2;
            rho_He3 = rho_He * Volume / Volume2 - time * PHe / Volume2;
            if (rho_He3 < 0.0) rho_He3 = 0.0;
            if (rho_He3 > rho3) rho_He3 = MISTAKE(OR_MISTAKE_HE_CLIP_RHO) ? rho3 : rho3 * 0.1;
            scal_new[2 * c] = Q3;
            scal_new[2 * c + 1] = rho_He3;
        }
        double u3, v3, w3, bx3, by3, bz3, p3;
        if (!rho_null) {
            /* planted mistakes: a Powell momentum term over 8 pi, an induction term of the wrong
               sign, an energy term dropped, the Zone_2 compensation over 8 pi */
            const double cpm = MISTAKE(OR_MISTAKE_POWELL_MOM_8PI) ? 2.0 * cpi4 : cpi4;
            const double sind = MISTAKE(OR_MISTAKE_POWELL_IND_SIGN) ? -1.0 : 1.0;
            const double sen = MISTAKE(OR_MISTAKE_POWELL_ENERGY) ? 0.0 : 1.0;
            const double c2 = MISTAKE(OR_MISTAKE_ZONE2_8PI) ? 8.0 * OR_PI : 4.0 * OR_PI;
            u3 = (rho * vx * Volume / Volume2 - time * (P[1] + (bx / cpm) * P[8]) / Volume2 + time * Smx) / rho3;
            v3 = (rho * vy * Volume / Volume2 - time * (P[2] + (by / cpm) * P[8]) / Volume2 + time * Smy) / rho3;
            w3 = (rho * vz * Volume / Volume2 - time * (P[3] + (bz / cpm) * P[8]) / Volume2 + time * Smz) / rho3;
            bx3 = bx * Volume / Volume2 - time * (P[5] + sind * vx * P[8]) / Volume2;
            by3 = by * Volume / Volume2 - time * (P[6] + sind * vy * P[8]) / Volume2;
            bz3 = bz * Volume / Volume2 - time * (P[7] + sind * vz * P[8]) / Volume2;
            p3 = (((pp / g1 + 0.5 * rho * KVV(vx, vy, vz) + KVV(bx, by, bz) / 25.13274122871834590768) * Volume / Volume2
                   - time * (P[4] + sen * (dsk / cpi4) * P[8]) / Volume2 + time * SE)
                  - 0.5 * rho3 * KVV(u3, v3, w3) - KVV(bx3, by3, bz3) / 25.13274122871834590768) * g1;
            if (m->cell_zone[c] == 2) {
                double tau = 1.0;
                double bx4 = bx3 - time * bx3 / tau;
                double by4 = by3 - time * by3 / tau;
                double bz4 = bz3 - time * bz3 / tau;
                double p4 = p3 + (time / tau - KV(time / tau)) * (g1) * KVV(bx3, by3, bz3) / c2;
                bx3 = bx4; by3 = by4; bz3 = bz4; p3 = p4;
            }
        } else {
            u3 = vx; v3 = vy; w3 = vz; bx3 = bx; by3 = by; bz3 = bz;
            p3 = rho3 / 2.0;
        }
        if (isnan(rho3)) {
#pragma omp critical(or_status2)
            { if (status == 0) status = -20; }
        }
        if (p3 < 0.00000001) { p3 = 0.000001; npf++; }
        double *o = prim_new + 8 * c;
        o[0] = rho3; o[1] = u3; o[2] = v3; o[3] = w3; o[4] = p3;
        if (m->cell_is_tvd[c]) { o[5] = bx3; o[6] = by3; o[7] = bz3; }
        else { o[5] = bx; o[6] = by; o[7] = bz; }
    }
    if (n_rho_floor) *n_rho_floor = nrf;
    if (n_p_floor) *n_p_floor = npf;
    return status;
}

int or_update(const or_mesh *m, const or_params *p, const double *prim_old, const double *vol_old,
              const double *vol_new, const double *flux, const double *src, double dt, double *prim_new,
              int64_t *n_rho_floor, int64_t *n_p_floor)
{
    return update_impl(m, p, prim_old, NULL, vol_old, vol_new, flux, NULL, src, dt, prim_new, NULL, n_rho_floor,
                       n_p_floor, 1);
}

int or_update_k(const or_mesh *m, const or_params *p, const double *prim_old, const double *scal_old,
                const double *vol_old, const double *vol_new, const double *flux, const double *kflux,
                const double *src, double dt, double *prim_new, double *scal_new, int64_t *n_rho_floor,
                int64_t *n_p_floor)
{
    return update_impl(m, p, prim_old, scal_old, vol_old, vol_new, flux, kflux, src, dt, prim_new, scal_new,
                       n_rho_floor, n_p_floor, 1);
}

static void geom_alloc(or_geom *g, const or_mesh *m)
{
    g->cell_center = (double *)malloc(sizeof(double) * 3 * m->n_cells);
    g->cell_volume = (double *)malloc(sizeof(double) * m->n_cells);
    g->face_center = (double *)malloc(sizeof(double) * 3 * m->n_faces);
    g->face_normal = (double *)malloc(sizeof(double) * 3 * m->n_faces);
    g->face_area = (double *)malloc(sizeof(double) * m->n_faces);
}
static void geom_free(or_geom *g)
{
    free(g->cell_center); free(g->cell_volume); free(g->face_center); free(g->face_normal); free(g->face_area);
}

/* Setka::Go, Setka_physics.cpp:1484-2278: lagged dt (first step 1e-5), layer swap, optional
 * motion and re-measure of the new layer, face loop, cell loop. */
/* ------------------------------------------------------------------------------------------ */
/* Setka::Calc_sourse_MF_Bera (Setka_physics.cpp:766-1258), plasma part: momentum and energy
 * sources of charge exchange between the protons and n_h hydrogen fluids, for a state without
 * pickup ions, Q or rho_He -- the keys Calc_sourse_MF_Bera then reads as 0, so the proton split of
 * Plasma_components_1 (Phys_param.cpp:628-665) has rho_He = 0 and al = 2 (zone <= 2). */
static const double OR_ME_MP = 1.0 / 1836.15; /* MF_meDmp, Header.h:99 */

/* |V1 - V2| as Eigen's norm(): square root of (d0^2 + d1^2) + d2^2 */
static double or_dnorm(const double *V1, const double *V2)
{
    double d0 = V1[0] - V2[0], d1 = V1[1] - V2[1], d2 = V1[2] - V2[2];
    return sqrt((d0 * d0 + d1 * d1) + d2 * d2);
}

/* U_bera, Um_bera, UE_bera (Setka_physics.cpp:746-762) */
static double or_u_bera(double T1, double T2, const double *V1, const double *V2)
{
    double uu = or_dnorm(V1, V2);
    return sqrt((4.0 / OR_PI) * (T1 + T2) + KV(uu));
}
static double or_um_bera(double T1, double T2, const double *V1, const double *V2)
{
    double uu = or_dnorm(V1, V2);
    return sqrt((16.0 / OR_PI) * T1 + (9.0 * OR_PI / 4.0) * T2 + 4.0 * KV(uu));
}
static double or_ue_bera(double T1, double T2, const double *V1, const double *V2)
{
    double uu = or_dnorm(V1, V2);
    return sqrt((4.0 / OR_PI) * T1 + (64.0 / 9.0 / OR_PI) * T2 + KV(uu));
}

static double or_dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* Setka::determ_zone (Setka_physics.cpp:1427-1482) */
int or_determ_zone(double Q, double rho, int cell_type)
{
    if (Q / rho < 50.0) return cell_type == 1 ? 1 : 2;
    return cell_type == 4 ? 4 : 3;
}

void or_bera_source(const double *s, int n_h, const double *h, double a2, double n_h_lism, double kn,
                    double *out)
{
    or_bera_source_k(s, 0.0, 0.0, 1, n_h, h, a2, n_h_lism, kn, out);
}

/* With the passive scalars: Calc_sourse_MF_Bera(cell, ..., zone = determ_zone) passes zone - 1 to
 * Plasma_components_1 (Setka_physics.cpp:769, 783), whose al is 2 for zone - 1 <= 2. */
void or_bera_source_k(const double *s, double Q, double rho_He, int cell_type, int n_h, const double *h, double a2,
                      double n_h_lism, double kn, double *out)
{
    or_bera_source_h(s, Q, rho_He, cell_type, n_h, h, a2, n_h_lism, kn, NULL, out, NULL);
}

/* The plasma sources and, in hout [n_h*5] (nullable), the hydrogen fluids' (rho, m_x, m_y, m_z, E):
 * each fluid loses what its exchanges with the protons take (Setka_physics.cpp:1012-1036), and the
 * atoms those exchanges make join fluid arise[(zone - 1) * n_h + k] (1-based; NULL: fluid `zone`,
 * the hydrogen_arise tables of the run without pickup ions, Phys_param.cpp:149-171) carrying the
 * protons' momentum and energy (:1043-1094); then every source times n_H_LISM / Kn and the
 * hydrogen ones divided by n_H_LISM (:1236-1250).  Pickup ions are not carried. */
void or_bera_source_h(const double *s, double Q, double rho_He, int cell_type, int n_h, const double *h, double a2,
                      double n_h_lism, double kn, const int8_t *arise, double *out, double *hout)
{
    /* proton density and temperature (Plasma_components_1, Phys_param.cpp:641-660) */
    const double rho = s[0], pr = s[4];
    const int zone = or_determ_zone(Q, rho, cell_type);
    const int zone_ = zone - 1;
    double g_rho[OR_MAX_H], g_m[OR_MAX_H][3], g_E[OR_MAX_H];
    const int al = zone_ <= 2 ? 2 : 1;
    double rho_p = -(OR_ME_MP * al * rho_He + 4.0 * (-rho + rho_He)) / (4.0 * (1.0 + OR_ME_MP));
    double T_p = 8.0 * (1.0 + OR_ME_MP) * pr / (8.0 * rho - (7.0 - al + (-1.0 + al) * OR_ME_MP) * rho_He);
    const double *Vp = s + 1;
    /* Setka_physics.cpp:802-808 */
    if (rho_p <= 1e-8) rho_p = 1e-8;
    if (T_p <= 1e-8) T_p = 1e-8;
    if (T_p > 1e5) T_p = 1e5;
    double mx = 0.0, my = 0.0, mz = 0.0, E = 0.0;
    for (int k = 0; k < n_h; k++) {
        /* Setka_physics.cpp:820-826 */
        const double rho_H = h[5 * k];
        const double T_H = 2.0 * h[5 * k + 1] / ((rho_H < 1e-9) ? 1e-9 : rho_H);  /* std::max */
        const double *VH = h + 5 * k + 2;
        /* :842-859: both orders of the pair (p, H) */
        const double U_pH = or_u_bera(T_p, T_H, Vp, VH), U_Hp = or_u_bera(T_H, T_p, VH, Vp);
        const double Um_pH = or_um_bera(T_p, T_H, Vp, VH), Um_Hp = or_um_bera(T_H, T_p, VH, Vp);
        const double UE_pH = or_ue_bera(T_p, T_H, Vp, VH), UE_Hp = or_ue_bera(T_H, T_p, VH, Vp);
        const double sig_pH = KV(1.0 - a2 * log(U_pH)), sig_Hp = KV(1.0 - a2 * log(U_Hp));
        /* :913-927 */
        const double cHp = sig_Hp * rho_p * rho_H, cpH = sig_pH * rho_p * rho_H;
        const double dV[3] = {Vp[0] - VH[0], Vp[1] - VH[1], Vp[2] - VH[2]};
        const double tHp = T_p / Um_Hp, tpH = T_H / Um_pH;
        double Hm_Hp[3], Hm_pH[3];
        for (int i = 0; i < 3; i++) {
            Hm_Hp[i] = cHp * (U_Hp * Vp[i] + tHp * dV[i]);
            Hm_pH[i] = cpH * (U_pH * VH[i] - tpH * dV[i]);
        }
        const double HE_Hp = cHp * ((0.5 * U_Hp * or_dot3(Vp, Vp) + tHp * or_dot3(Vp, dV)) + 3.0 / 4.0 * T_p * UE_Hp);
        const double HE_pH = cpH * ((0.5 * U_pH * or_dot3(VH, VH) - tpH * or_dot3(VH, dV)) + 3.0 / 4.0 * T_H * UE_pH);
        /* :975-983 */
        mx += (-Hm_Hp[0] + Hm_pH[0]);
        my += (-Hm_Hp[1] + Hm_pH[1]);
        mz += (-Hm_Hp[2] + Hm_pH[2]);
        E += (-HE_Hp + HE_pH);
        if (hout && k < OR_MAX_H) {
            /* :1012-1036, the fluid's losses: Hrho[p+H], Hm[p+H], HE[p+H] */
            hout[5 * k] = 0.0 + (-(cpH * U_pH));
            for (int i = 0; i < 3; i++) hout[5 * k + 1 + i] = 0.0 + (-Hm_pH[i]);
            hout[5 * k + 4] = 0.0 + (-HE_pH);
            /* the atoms made from this fluid: Hrho[H+p], Hm[H+p], HE[H+p] */
            g_rho[k] = cHp * U_Hp;
            for (int i = 0; i < 3; i++) g_m[k][i] = Hm_Hp[i];
            g_E[k] = HE_Hp;
        }
    }
    /* :1236-1240 */
    const double ddp = n_h_lism / kn;
    out[0] = 0.0 * ddp;
    out[1] = mx * ddp;
    out[2] = my * ddp;
    out[3] = mz * ddp;
    out[4] = E * ddp;
    if (hout) {
        /* :1043-1094, rows of the arise table in order */
        for (int k = 0; k < n_h && k < OR_MAX_H; k++) {
            int t = (arise ? arise[zone_ * n_h + k] : zone) - 1;
            if (t < 0 || t >= n_h) continue;
            hout[5 * t] += g_rho[k];
            for (int i = 0; i < 3; i++) hout[5 * t + 1 + i] += g_m[k][i];
            hout[5 * t + 4] += g_E[k];
        }
        /* :1236-1250 */
        for (int k = 0; k < 5 * n_h; k++) hout[k] *= ddp;
        for (int k = 0; k < 5 * n_h; k++) hout[k] /= n_h_lism;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Setka::Culc_Velocity_surface (Setka_move.cpp:111-1357).  Face loops (:126-182 TS, :186-360 HP,
 * :364-433 BS): chlld with w = 0 on the face states of layer `now`; each of the face's nodes gets
 * velocity_K * ds * n and one count.  Smoothing (:436-506 TS sphere, :566-622 HP sphere near the
 * nose, :1136-1210 BS neighbour mean) adds one more vector and count.  Average, then the move
 * along the radius (:1228-1278; HP with x < 0 along the radius in the yz plane). */
/* solveQuadraticEquation (Help_funk.cpp:290-305): the parabola a x^2 + b x + c through three points;
 * returns -1 where two abscissae coincide */
static int or_quad3(double x1, double y1, double x2, double y2, double x3, double y3, double *a, double *b, double *c)
{
    double denom = (x1 - x2) * (x1 - x3) * (x2 - x3);
    if (denom == 0) return -1;
    *a = (x3 * (y2 - y1) + x2 * (y1 - y3) + x1 * (y3 - y2)) / denom;
    *b = (x3 * x3 * (y1 - y2) + x2 * x2 * (y3 - y1) + x1 * x1 * (y2 - y3)) / denom;
    *c = (x2 * x3 * (x2 - x3) * y1 + x3 * x1 * (x3 - x1) * y2 + x1 * x2 * (x1 - x2) * y3) / denom;
    return 0;
}

/* 3x3 symmetric solve by elimination with partial pivoting (the LM step of the circle fit) */
static int or_solve3(double H[3][3], const double g[3], double d[3])
{
    double M[3][4];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) M[i][j] = H[i][j];
        M[i][3] = g[i];
    }
    for (int c = 0; c < 3; c++) {
        int piv = c;
        for (int r = c + 1; r < 3; r++)
            if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
        if (M[piv][c] == 0.0) return -1;
        for (int j = 0; j < 4; j++) { double t = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = t; }
        for (int r = c + 1; r < 3; r++) {
            double f = M[r][c] / M[c][c];
            for (int j = c; j < 4; j++) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 2; i >= 0; i--) {
        double t = M[i][3];
        for (int j = i + 1; j < 3; j++) t -= M[i][j] * d[j];
        d[i] = t / M[i][i];
    }
    return 0;
}

double or_circle_fit(int64_t n, const double *pts, double *centre)
{
    /* fitCircleRobust (Setka_move.cpp:2544-2620), called with max_iter 8 and tol 1e-8: Levenberg-
     * Marquardt on (cx, cy, R) from the origin and the mean distance; the damping is multiplied by 10
     * after every step (the reference compares the error with prev_error after setting one to the
     * other), and a step with a non-finite update only raises the damping */
    double P[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < n; i++) P[2] += sqrt(pts[2 * i] * pts[2 * i] + pts[2 * i + 1] * pts[2 * i + 1]);
    P[2] /= (double)n;
    double lambda = 0.001, prev = DBL_MAX;
    for (int it = 0; it < 8; it++) {
        double H[3][3] = {{0}}, g[3] = {0, 0, 0}, err = 0.0;
        for (int64_t i = 0; i < n; i++) {
            double dx = pts[2 * i] - P[0], dy = pts[2 * i + 1] - P[1];
            double dist = sqrt(dx * dx + dy * dy), res, J[3];
            if (dist < 1e-10) {
                res = -P[2];
                J[0] = 0.0; J[1] = 0.0; J[2] = -1.0;
            } else {
                res = dist - P[2];
                J[0] = -dx / dist; J[1] = -dy / dist; J[2] = -1.0;
            }
            err += res * res;
            for (int a = 0; a < 3; a++) {
                g[a] += J[a] * -res;
                for (int b = 0; b < 3; b++) H[a][b] += J[a] * J[b];
            }
        }
        err = sqrt(err / (double)n);
        if (fabs(prev - err) < 1e-8) break;
        prev = err;
        for (int a = 0; a < 3; a++) H[a][a] *= 1.0 + lambda;
        double d[3];
        if (or_solve3(H, g, d) || !isfinite(d[0]) || !isfinite(d[1]) || !isfinite(d[2])) {
            lambda *= 10.0;
            continue;
        }
        for (int a = 0; a < 3; a++) P[a] += d[a];
        lambda = err < prev ? lambda / 2.0 : lambda * 10.0;
    }
    centre[0] = P[0];
    centre[1] = P[1];
    return P[2];
}

static double or_ryz0(const double *p) { return sqrt(KVV(0.0, p[1], p[2])); }

/* The angle term (Setka_move.cpp:703-745) of an HP node A: towards the circle about (x, cy, cz)
 * through the mean distance of its neighbours' (y, z) from its centre; V[0] = 0. */
static void or_hp_angle_term(const double *A, double cy, double cz, const double *x, const int32_t *nb, int32_t nnb,
                             double c, double time, double V[3])
{
    double d1 = A[1] - cy, d2 = A[2] - cz;
    double rr = sqrt(d1 * d1 + d2 * d2);
    double r = 0.0;
    for (int32_t j = 0; j < nnb; j++) {
        const double *P = x + 3 * (int64_t)nb[j];
        double e1 = cy - P[1], e2 = cz - P[2];
        r += sqrt(e1 * e1 + e2 * e2);
    }
    r /= (double)nnb;
    double kk = r / rr;
    double B1 = kk * (A[1] - cy) + cy, B2 = kk * (A[2] - cz) + cz;
    V[0] = 0.0;
    V[1] = c * (B1 - A[1]) / time;
    V[2] = c * (B2 - A[2]) / time;
}

/* macros1 (Setka_move.cpp:4-99) of an HP node A with grid neighbours P8 = AA1, AA11, AA3, AA33, AA2,
 * AA22, AA4, AA44: parabolas in (x, r_yz) through AA3, AA1, AA11 and AA33, AA3, AA1, and in (polar
 * angle about x, r_yz) through AA2, AA4, AA44 and AA4, AA2, AA22, evaluated at the node; term q
 * scales A's (y, z) to that r_yz, weighted c_along (q = 0, 1) or c_angle (q = 2, 3).  -1 when a
 * parabola's abscissae repeat. */
static int or_hp_grid_terms(const double *A, const double *const P8[8], double c_along, double c_angle,
                            double time, double V[4][3])
{
    const double *a1 = P8[0], *a11 = P8[1], *a3 = P8[2], *a33 = P8[3], *a2 = P8[4], *a22 = P8[5], *a4 = P8[6],
                 *a44 = P8[7];
    int bad = 0;
    double p0 = A[0], p1 = or_ryz0(A), qa = 0, qb = 0, qc = 0, rad[4];
    bad |= or_quad3(a3[0], or_ryz0(a3), a1[0], or_ryz0(a1), a11[0], or_ryz0(a11), &qa, &qb, &qc);
    rad[0] = qa * KV(p0) + qb * p0 + qc;
    bad |= or_quad3(a33[0], or_ryz0(a33), a3[0], or_ryz0(a3), a1[0], or_ryz0(a1), &qa, &qb, &qc);
    rad[1] = qa * KV(p0) + qb * p0 + qc;
    p0 = polar_angle(A[1], A[2]);
    double t2 = polar_angle(a2[1], a2[2]), t4 = polar_angle(a4[1], a4[2]);
    double t22 = polar_angle(a22[1], a22[2]), t44 = polar_angle(a44[1], a44[2]);
    bad |= or_quad3(t2, or_ryz0(a2), t4, or_ryz0(a4), t44, or_ryz0(a44), &qa, &qb, &qc);
    rad[2] = qa * KV(p0) + qb * p0 + qc;
    bad |= or_quad3(t4, or_ryz0(a4), t2, or_ryz0(a2), t22, or_ryz0(a22), &qa, &qb, &qc);
    rad[3] = qa * KV(p0) + qb * p0 + qc;
    for (int q = 0; q < 4; q++) {
        double c = q < 2 ? c_along : c_angle;
        double B1 = A[1], B2 = A[2];
        B1 *= rad[q] / p1;
        B2 *= rad[q] / p1;
        V[q][0] = c * (A[0] - A[0]) / time;
        V[q][1] = c * (B1 - A[1]) / time;
        V[q][2] = c * (B2 - A[2]) / time;
    }
    return bad ? -1 : 0;
}

int or_hp_terms(const double *A, const double *P8, double cy, double cz, const double *ring, int32_t n_ring,
                double c_angle_term, double c_along, double c_angle, double time, double *V5)
{
    const double *Q8[8];
    for (int q = 0; q < 8; q++) Q8[q] = P8 + 3 * q;
    int32_t idx[64];
    if (n_ring > 64) return -2;
    for (int32_t j = 0; j < n_ring; j++) idx[j] = j;
    or_hp_angle_term(A, cy, cz, ring, idx, n_ring, c_angle_term, time, V5);
    return or_hp_grid_terms(A, Q8, c_along, c_angle, time, (double (*)[3])(V5 + 3));
}

static int fit_kind(uint8_t fl) { return ((fl >> 3) & 1) ? 1 : (((fl >> 2) & 1) ? 2 : (((fl >> 4) & 1) ? 3 : 0)); }

int or_culc_velocity_surface(const or_mesh *m, const or_params *p, const or_fit *fit, const double *x,
                             const double *prim, const or_geom *g, double time, double *x_next,
                             double *node_velocity)
{
    int status = 0;
    double *sp = (double *)calloc((size_t)(3 * m->n_faces), sizeof(double));
    for (int64_t f = 0; f < m->n_faces; f++) {
        int kind = m->face_flags && m->face_cells[2 * f + 1] >= 0 ? fit_kind(m->face_flags[f]) : 0;
        if (!kind || !fit->move[kind]) continue;
        double qL[8], qR[8], q[8], dsl, dsc, dsr;
        int br;
        if (or_face_states(m, p, f, prim, g, qL, qR)) status = -1;
        const double *n = g->face_normal + 3 * f;
        if (or_chlld(m->face_method[f], 0, 0, n[0], n[1], n[2], 0.0, qL, qR, 0, NULL, NULL, q, NULL, &dsl, &dsc,
                     &dsr, &br))
            status = -1;
        double ds = kind == 1 ? dsl : (kind == 2 ? dsc : dsr);
        if (kind == 2 && (isnan(ds) || fpclassify(ds) == FP_SUBNORMAL) && !status) status = -8;
        for (int k = 0; k < 3; k++) sp[3 * f + k] = fit->velocity[kind] * ds * n[k];
    }
    for (int64_t i = 0; i < fit->n_sn; i++) {
        int kind = fit->sn_kind[i];
        int64_t v = fit->sn_node[i];
        const double *A = x + 3 * v;
        double vel[3] = {0.0, 0.0, 0.0};
        int num = 0;
        if (fit->move[kind]) {
            for (int32_t j = fit->sn_face_off[i]; j < fit->sn_face_off[i + 1]; j++) {
                const double *s = sp + 3 * (int64_t)fit->sn_face[j];
                vel[0] += s[0];
                vel[1] += s[1];
                vel[2] += s[2];
                num++;
            }
        }
        int32_t nb0 = fit->sn_nb_off[i], nb1 = fit->sn_nb_off[i + 1];
        int smooth = fit->sglag[kind] != 0;
        if (smooth && kind == 2) {
            double phi = polar_angle(A[0], sqrt(KVV(0.0, A[1], A[2])));
            if (phi > fit->phi_lim) smooth = 0;
        }
        if (smooth && kind == 3 && nb1 - nb0 <= 5) smooth = 0;
        if (smooth && kind != 3 && nb1 == nb0) smooth = 0; /* no neighbours from the set-up */
        if (smooth) {
            double Bv[3], V[3];
            double c = fit->velocity[kind] * fit->k_smooth[kind];
            if (kind == 3) {
                for (int k = 0; k < 3; k++) Bv[k] = A[k];
                for (int32_t j = nb0; j < nb1; j++)
                    for (int k = 0; k < 3; k++) Bv[k] += x[3 * (int64_t)fit->sn_nb[j] + k];
                for (int k = 0; k < 3; k++) Bv[k] /= (nb1 - nb0 + 1.0);
            } else {
                const double *C = fit->sn_centre + 3 * i;
                double d[3] = {A[0] - C[0], A[1] - C[1], A[2] - C[2]};
                double rr = or_norm3(d);
                double r = 0.0;
                for (int32_t j = nb0; j < nb1; j++) {
                    const double *P = x + 3 * (int64_t)fit->sn_nb[j];
                    double e[3] = {P[0] - C[0], P[1] - C[1], P[2] - C[2]};
                    r += or_norm3(e);
                }
                r /= (double)(nb1 - nb0);
                double kk = r / rr;
                for (int k = 0; k < 3; k++) Bv[k] = kk * (A[k] - C[k]) + C[k];
            }
            for (int k = 0; k < 3; k++) V[k] = c * (Bv[k] - A[k]) / time;
            for (int k = 0; k < 3; k++) vel[k] += V[k];
            num++;
        }
        if (kind == 2 && fit->sglag[2] && fit->sn_cc && !isnan(fit->sn_cc[2 * i]) && nb1 > nb0) {
            double V[3];
            or_hp_angle_term(A, fit->sn_cc[2 * i], fit->sn_cc[2 * i + 1], x, fit->sn_nb + nb0, nb1 - nb0,
                             fit->velocity[2] * fit->k_angle, time, V);
            for (int k = 0; k < 3; k++) vel[k] += V[k];
            num++;
        }
        if (kind == 2 && fit->sglag[2] && fit->sn_grid && fit->sn_grid[8 * i] >= 0) {
            const double *P8[8];
            double V[4][3];
            for (int q = 0; q < 8; q++) P8[q] = x + 3 * (int64_t)fit->sn_grid[8 * i + q];
            if (or_hp_grid_terms(A, P8, fit->velocity[2] * fit->hp_along * fit->k_hp,
                                 fit->velocity[2] * fit->hp_angle * fit->k_hp, time, V))
                status = -9;
            for (int q = 0; q < 4; q++) {
                for (int k = 0; k < 3; k++) vel[k] += V[q][k];
                num++;
            }
        }
        if (num > 0)
            for (int k = 0; k < 3; k++) vel[k] /= num;
        if (node_velocity)
            for (int k = 0; k < 3; k++) node_velocity[3 * i + k] = vel[k];
        if (fit->move[kind]) {
            double Bd[3] = {A[0], A[1], A[2]};
            if (kind == 2 && !(A[0] >= 0.0)) Bd[0] = 0.0;
            double z = (Bd[0] * Bd[0] + Bd[1] * Bd[1]) + Bd[2] * Bd[2]; /* squaredNorm */
            if (z > 0.0) {
                double sz = sqrt(z);
                for (int k = 0; k < 3; k++) Bd[k] /= sz;
            }
            double sdot = ((vel[0] * Bd[0] + vel[1] * Bd[1]) + vel[2] * Bd[2]) * time;
            for (int k = 0; k < 3; k++) x_next[3 * v + k] = A[k] + sdot * Bd[k];
        }
    }
    if (fit->move[2] && fit->n_tail > 0) {
        /* Setka_move.cpp:1281-1323: along each azimuth's D rays the HP node N4-5 onto the line
         * through nodes N4-4 and N4-6 in (x, r_yz), and nodes N4-3 .. on to the r_yz of node N4-4 */
        const int N4 = fit->N4, NN = fit->tail_len - 1;
        for (int32_t a = 0; a < fit->n_tail; a++) {
            const int32_t *L = fit->tail + (int64_t)a * fit->tail_len;
#define TQ(k) (x_next + 3 * (int64_t)L[k])
            double h1 = or_ryz0(TQ(N4 - 4)), xx1 = TQ(N4 - 4)[0];
            double h2 = or_ryz0(TQ(N4 - 6)), xx2 = TQ(N4 - 6)[0];
            double *yz = TQ(N4 - 5);
            double hh = or_ryz0(yz), xx = yz[0];
            double hhh = or_linear2(xx1, h1, xx2, h2, xx);
            yz[1] = yz[1] * hhh / hh;
            yz[2] = yz[2] * hhh / hh;
            for (int k = N4 - 3; k <= NN; k++) {
                double *q = TQ(k);
                double hq = or_ryz0(q);
                q[1] = q[1] * h1 / hq;
                q[2] = q[2] * h1 / hq;
            }
#undef TQ
        }
    }
    if (fit->move[3] && fit->n_bs > 0) {
        /* Setka_move.cpp:1325-1357: each azimuth's B, E and D rays' BS nodes onto the r_yz of the
         * last A ray's BS node */
        for (int32_t a = 0; a < fit->n_bs; a++) {
            double H = or_ryz0(x_next + 3 * (int64_t)fit->bs_ref[a]);
            for (int32_t j = fit->bs_off[a]; j < fit->bs_off[a + 1]; j++) {
                double *q = x_next + 3 * (int64_t)fit->bs_nodes[j];
                double h2 = or_ryz0(q);
                q[1] *= H / h2;
                q[2] *= H / h2;
            }
        }
    }
    free(sp);
    return status;
}

/* cos and sin of a ray angle from one glibc sincos call.  glibc 2.35's sincos and cos differ in the
 * last bit for some arguments (cos(1.3646395517420515)), and whether a compiler turns cos(a) beside
 * sin(a) into sincos is its own choice; the library's host set-up takes the ray trig the same way. */
static void or_sincos(double a, double *s, double *c) { sincos(a, s, c); }

/* Luch::dvigenie for A_Luch / A2_Luch (Luch.cpp:8-151): the radii of a ray's nodes from R0, R1,
 * its TS, HP and BS nodes' distances from the origin in x (func_R, Yzel.cpp:27-30) and R5. */
void or_dvigenie(const or_fit *fit, double *x)
{
    const int32_t M0 = fit->M0, M1 = fit->M1, M11 = fit->M11, M2 = fit->M2, M3 = fit->M3, M4 = fit->M4;
    for (int64_t k = 0; k < fit->n_rays; k++) {
        const int32_t *Y = fit->ray_nodes + k * fit->npr;
        double st, ct, sp, cp;
        or_sincos(fit->ray_angles[2 * k], &st, &ct);
        or_sincos(fit->ray_angles[2 * k + 1], &sp, &cp);
        double R0 = fit->R0, R1 = fit->R1, R5 = fit->R5;
        const double *o1 = x + 3 * (int64_t)Y[M0 + M1 + 1];
        const double *o2 = x + 3 * (int64_t)Y[M0 + M1 + M2 + 2];
        const double *o3 = x + 3 * (int64_t)Y[M0 + M1 + M2 + M3 + 3];
        double R2 = sqrt(KVV(o1[0], o1[1], o1[2]));
        double R3 = sqrt(KVV(o2[0], o2[1], o2[2]));
        double R4 = sqrt(KVV(o3[0], o3[1], o3[2]));
        double r, a, b;
        int num = 0;
#define OR_PUT(idx, rad)                                               \
    do {                                                               \
        double *X_ = x + 3 * (int64_t)Y[idx];                          \
        X_[0] = (rad) * ct;                                            \
        X_[1] = (rad) * st * cp;                                       \
        X_[2] = (rad) * st * sp;                                       \
    } while (0)
        for (int j = 0; j < M0; j++) {
            r = R0 + j * (R1 - R0) / M0;
            OR_PUT(j, r);
        }
        num += M0 + 1;
        for (int j = 0; j < M1; j++) {
            r = R1 + (j + 1) * (R2 - R1) / (M1 + 1);
            OR_PUT(num + j, r);
        }
        num += M1 + 1;
        double dr = (R2 - R1) / (M1 + 1);
        for (int j = 0; j < M11; j++) {
            r = R2 + dr * (j + 1);
            OR_PUT(num + j, r);
        }
        num += M11;
        a = fit->da1;
        b = fit->da2;
        for (int j = 0; j < M2 - M11; j++) {
            double s = (j + 1.0) / (M2 - M11 + 1.0);
            double ss = s * (a * ((1.0 - 3.0 * s + 2.0 * (s * s)) * (1.0 - 3.0 * s + 2.0 * (s * s))) +
                             s * (7.0 + b * ((1.0 - 2.0 * s) * (1.0 - 2.0 * s)) * (-1.0 + s) -
                                  2.0 * s * (9.0 + 2.0 * s * (-5.0 + 2.0 * s))));
            r = R2 + dr * M11 + (R3 - R2 - dr * M11) * ss;
            OR_PUT(num + j, r);
        }
        num += (M2 - M11) + 1;
        a = fit->da3;
        b = fit->da4;
        for (int j = 0; j < M3; j++) {
            double s = (j + 1.0) / (M3 + 1.0);
            double ss = s * (a - (2.0 * a + b - 3.0) * s + (a + b - 2.0) * s * s);
            r = R3 + (R4 - R3) * ss;
            OR_PUT(num + j, r);
        }
        num += M3 + 1;
        a = fit->da5;
        b = 2.0;
        for (int j = 0; j < M4; j++) {
            double s = (j + 1.0) / (M4 + 1.0);
            double ss = a * s - (3.0 * a + b - 4.0) * s * s * s + (2.0 * a + b - 3.0) * s * s * s * s;
            r = R4 + (R5 - R4) * ss;
            OR_PUT(num + j, r);
        }
        OR_PUT(num + M4, R5);
#undef OR_PUT
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Monte-Carlo neutral densities (Setka::MK_go / MK_fly_immit, Setka_MK.cpp:1225-2316).            */
static const double OR_SQRTPI_MK = 1.77245385; /* sqrtpi_, Header.h:90 (the reference's literal) */

static uint64_t mk_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* per-particle counter stream, values in (0, 1) */
static double mk_rand(uint64_t *s)
{
    *s += 0x9E3779B97F4A7C15ull;
    return ((double)(mk_mix(*s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* Get_Spotok_inf (Setka_MK.cpp:2316-2324): inflow of the unit-temperature LISM Maxwellian through
 * a face of outward normal n, per unit area */
static double mk_spotok_inf(double v_inf, const double *n)
{
    double Ux = v_inf * n[0] + 0.0 * n[1] + 0.0 * n[2];
    return -(-exp(-(Ux * Ux)) + OR_SQRTPI_MK * Ux * erfc(Ux)) / (2.0 * OR_SQRTPI_MK);
}

/* Velosity_1 (Setka_MK.cpp:40-50): mean relative speed of a neutral at relative speed u in a
 * Maxwellian of thermal speed cp */
static double mk_velosity_1(double u, double cp)
{
    if (u < 0.00001)
        return 2.0 * cp / OR_SQRTPI_MK + 2.0 * u * u / (3.0 * cp * OR_SQRTPI_MK) -
               u * u * u * u / (15.0 * cp * cp * cp * OR_SQRTPI_MK);
    return exp(-u * u / (cp * cp)) * cp / OR_SQRTPI_MK + (u + (cp * cp) / (2.0 * u)) * erf(u / cp);
}

static double mk_sigma(double a2, double x) { double t = 1.0 - a2 * log(x); return t * t; }

/* Gran::rayTriangleIntersect (Gran.cpp:522-561), Moller-Trumbore with the reference's 1e-6 */
static int mk_ray_tri(const double *o, const double *d, const double *v0, const double *v1, const double *v2,
                      double *t)
{
    double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = (e1[0] * pv[0] + e1[1] * pv[1]) + e1[2] * pv[2];
    if (fabs(det) < 1e-6) return 0;
    double inv = 1.0 / det;
    double tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    double u = ((tv[0] * pv[0] + tv[1] * pv[1]) + tv[2] * pv[2]) * inv;
    if (u < 0.0 || u > 1.0) return 0;
    double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    double v = ((d[0] * qv[0] + d[1] * qv[1]) + d[2] * qv[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return 0;
    *t = ((e2[0] * qv[0] + e2[1] * qv[1]) + e2[2] * qv[2]) * inv;
    return *t > 1e-6;
}

/* Setka::Time_to_vilet with Gran::Luch_crossing (Setka_MK.cpp:2725-2785, Gran.cpp:451-510): the
 * earliest crossing of the cell's faces, each split into the triangles (0,1,2) and (0,2,3) */
static int mk_exit(const or_mesh *m, const double *x, int64_t c, const double *P, const double *V, double *time,
                   int64_t *face)
{
    double nv = sqrt((V[0] * V[0] + V[1] * V[1]) + V[2] * V[2]);
    double d[3] = {V[0] / nv, V[1] / nv, V[2] / nv};
    double tmin = 1e10;
    int found = 0;
    for (int j = 0; j < 6; j++) {
        int64_t f = m->cell_faces[6 * c + j];
        const int32_t *fn = m->face_nodes + 4 * f;
        double t1 = 0.0, t2 = 0.0;
        int b1 = mk_ray_tri(P, d, x + 3 * fn[0], x + 3 * fn[1], x + 3 * fn[2], &t1);
        t1 /= nv;
        int b2 = mk_ray_tri(P, d, x + 3 * fn[0], x + 3 * fn[2], x + 3 * fn[3], &t2);
        t2 /= nv;
        if (!b1 && !b2) continue;
        double tf = (b1 && b2) ? SMIN(t1, t2) : (b1 ? t1 : t2);
        if (tmin > tf) { tmin = tf; *face = f; found = 1; }
    }
    *time = tmin;
    return found;
}

/* spherical components in the frame whose polar axis is x (spherical_skorost with (z, x, y) =
 * (x0, x1, x2), Help_funk.cpp:108-121) and back (dekard_skorost, :123-135) */
static void mk_to_sph(const double *X, const double *Vc, double *r, double *th, double *ph)
{
    double rr = sqrt(X[1] * X[1] + X[2] * X[2] + X[0] * X[0]);
    double the = acos(X[0] / rr);
    double phi = polar_angle(X[1], X[2]);
    double Vz = Vc[0], Vx = Vc[1], Vy = Vc[2];
    *r = Vx * sin(the) * cos(phi) + Vy * sin(the) * sin(phi) + Vz * cos(the);
    *th = Vx * cos(the) * cos(phi) + Vy * cos(the) * sin(phi) - Vz * sin(the);
    *ph = -Vx * sin(phi) + Vy * cos(phi);
}
static void mk_from_sph(const double *X, double r, double th, double ph, double *Vc)
{
    double rr = sqrt(X[1] * X[1] + X[2] * X[2] + X[0] * X[0]);
    double the = acos(X[0] / rr);
    double phi = polar_angle(X[1], X[2]);
    Vc[1] = r * sin(the) * cos(phi) + th * cos(the) * cos(phi) - ph * sin(phi);
    Vc[2] = r * sin(the) * sin(phi) + th * cos(the) * sin(phi) + ph * cos(phi);
    Vc[0] = r * cos(the) - th * sin(the);
}

/* Setka::M_K_Change_Velosity (Setka_MK.cpp:2788-2850): the velocity of the proton a neutral of
 * velocity V exchanged with, in units of cp, from the Maxwellian of mean U weighted by g sigma(g);
 * the rejection loop is capped at 10000 tries (*capped set) */
static void mk_change_velosity(uint64_t *rs, double a2, const double *U, const double *V, double *W, double cp,
                               int *capped, int mistake)
{
    double X = sqrt((V[0] - U[0]) * (V[0] - U[0]) + (V[1] - U[1]) * (V[1] - U[1]) + (V[2] - U[2]) * (V[2] - U[2]));
    double p4 = 0.5 * OR_SQRTPI_MK * X / (1.0 + 0.5 * OR_SQRTPI_MK * X);
    double v1 = 0.0, v2 = 0.0, v3 = 0.0;
    for (int tries = 0; tries < 10000; tries++) {
        double k1 = mk_rand(rs), k2 = mk_rand(rs), k3 = mk_rand(rs), k4 = mk_rand(rs), k5 = mk_rand(rs),
               k6 = mk_rand(rs);
        double y1, y2, y3;
        if (p4 < k1) {
            double om1 = 1.0 - 2.0 * k4;
            double om2 = sqrt(1.0 - om1 * om1) * cos(2.0 * OR_PI * k5);
            double om3 = sqrt(1.0 - om1 * om1) * sin(2.0 * OR_PI * k5);
            double lo = sqrt(-log(k2 * k3));
            y1 = lo * om1; y2 = lo * om2; y3 = lo * om3;
        } else {
            y1 = sqrt(-log(k2)) * cos(OR_PI * k3);
            y2 = sqrt(-log(k4)) * cos(2.0 * OR_PI * k5);
            y3 = sqrt(-log(k4)) * sin(2.0 * OR_PI * k5);
        }
        v1 = y1 + U[0]; v2 = y2 + U[1]; v3 = y3 + U[2];
        double u1 = V[0] - v1, u2 = V[1] - v2, u3 = V[2] - v3;
        double uuu = sqrt(u1 * u1 + u2 * u2 + u3 * u3);
        double yy = sqrt(y1 * y1 + y2 * y2 + y3 * y3);
        double h = (uuu * mk_sigma(a2, uuu * cp)) / (mk_sigma(a2, X * cp) * (X + yy));
        if (!(h < k6) || (mistake & 2)) { W[0] = v1; W[1] = v2; W[2] = v3; return; }
    }
    *capped = 1;
    W[0] = v1; W[1] = v2; W[2] = v3;
}

/* Setka::Velosity_initial (Setka_MK.cpp:2671-2723): a velocity entering through a face of outward
 * normal n from the unit-temperature Maxwellian drifting at (v_inf, 0, 0) */
static void mk_velosity_initial(uint64_t *rs, double v_inf, const double *n, double *V, int *capped, int mistake)
{
    double t[3], mm[3];
    or_get_bazis(n, t, mm);
    double Vi[3] = {v_inf * n[0] + 0.0 * n[1] + 0.0 * n[2], v_inf * t[0] + 0.0 * t[1] + 0.0 * t[2],
                    v_inf * mm[0] + 0.0 * mm[1] + 0.0 * mm[2]};
    double p1 = fabs(Vi[0]) * OR_SQRTPI_MK / (1.0 + fabs(Vi[0]) * OR_SQRTPI_MK);
    double z = 0.0;
    int ok = 0;
    for (int tries = 0; tries < 10000; tries++) {
        double k3 = mk_rand(rs), k4 = mk_rand(rs), k5 = mk_rand(rs), k6 = mk_rand(rs);
        if (p1 > k3) z = cos(OR_PI * k5) * sqrt(-log(k4));
        else if (k4 <= 0.5) z = -sqrt(-log(2.0 * k4));
        else z = sqrt(-log(2.0 * (1.0 - k4)));
        int weight_rejects = fabs(z + Vi[0]) / (fabs(Vi[0]) + fabs(z)) < k6;
        if (mistake & 1) weight_rejects = 0;
        if (!(weight_rejects || z > -Vi[0])) { ok = 1; break; }
    }
    if (!ok) *capped = 1;
    double V1 = z + Vi[0];
    double k1 = mk_rand(rs), k2 = mk_rand(rs);
    double a = sqrt(-log(k2));
    double V2 = Vi[1] + a * cos(2.0 * OR_PI * k1);
    double V3 = Vi[2] + a * sin(2.0 * OR_PI * k1);
    for (int i = 0; i < 3; i++) V[i] = V1 * n[i] + V2 * t[i] + V3 * mm[i];
}

void or_mk_change_velosity_batch(int64_t n, uint64_t seed, double a2, const double *U, const double *V, double cp,
                                 int mistake, double *W, int64_t *capped)
{
    *capped = 0;
    for (int64_t i = 0; i < n; i++) {
        uint64_t rs = mk_mix(seed ^ mk_mix((uint64_t)i + 0x9E3779B97F4A7C15ull));
        int cap = 0;
        mk_change_velosity(&rs, a2, U, V, W + 3 * i, cp, &cap, mistake);
        *capped += cap;
    }
}

int64_t or_mk_setup(const or_mesh *m, const or_geom *g, double v_inf, int32_t n_per_face, int32_t *face,
                    int32_t *count, double *mu)
{
    double total = 0.0;
    int64_t nb = 0;
    for (int64_t f = 0; f < m->n_faces; f++) {
        if (m->face_bc[f] != OR_BC_OUTER_HARD && m->face_bc[f] != OR_BC_OUTER_SOFT) continue;
        total += mk_spotok_inf(v_inf, g->face_normal + 3 * f) * g->face_area[f];
        nb++;
    }
    if (nb == 0) return 0;
    double mu_expect = total / (1.0 * n_per_face * nb);
    int64_t k = 0;
    for (int64_t f = 0; f < m->n_faces; f++) {
        if (m->face_bc[f] != OR_BC_OUTER_HARD && m->face_bc[f] != OR_BC_OUTER_SOFT) continue;
        double sv = mk_spotok_inf(v_inf, g->face_normal + 3 * f) * g->face_area[f];
        if (sv < 0.0000001 * total / nb) continue;
        int32_t n = (int32_t)(sv / mu_expect) + 1;
        int32_t lo = n_per_face < 100 ? n_per_face : 100;
        if (n < lo) n = lo;
        face[k] = (int32_t)f;
        count[k] = n;
        mu[k] = sv / n;
        k++;
    }
    return k;
}

/* findSphereIntersectionTime (Setka_MK.cpp:4-38): the first non-negative time at which X + t V is at
 * radius R. */
static int mk_sphere_time(const double *X, const double *V, double R, double *time)
{
    const double a = V[0] * V[0] + V[1] * V[1] + V[2] * V[2];
    const double b = 2.0 * (X[0] * V[0] + X[1] * V[1] + X[2] * V[2]);
    const double c = (X[0] * X[0] + X[1] * X[1] + X[2] * X[2]) - R * R;
    const double D = b * b - 4 * a * c;
    if (D < 0) return 0;
    const double sqrtD = sqrt(D);
    const double t1 = (-b - sqrtD) / (2 * a);
    const double t2 = (-b + sqrtD) / (2 * a);
    if (t1 >= 0) { *time = t1; return 1; }
    if (t2 >= 0) { *time = t2; return 1; }
    return 0;
}

/* The ray from the origin through P hits triangle (a, b, c) (Moller-Trumbore, t > 0). */
static int mk_origin_ray_tri(const double *P, const double *a, const double *b, const double *c)
{
    const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double h[3] = {P[1] * e2[2] - P[2] * e2[1], P[2] * e2[0] - P[0] * e2[2], P[0] * e2[1] - P[1] * e2[0]};
    const double det = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
    if (fabs(det) < 1e-300) return 0;
    const double s[3] = {-a[0], -a[1], -a[2]};
    const double u = (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]) / det;
    if (u < 0.0 || u > 1.0) return 0;
    const double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    const double v = (P[0] * q[0] + P[1] * q[1] + P[2] * q[2]) / det;
    if (v < 0.0 || u + v > 1.0) return 0;
    return (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det > 0.0;
}

/* The cell a particle re-enters the mesh by after passing through the Sun's sphere: the cell on the
 * inner boundary face that the ray from the centre through its position crosses (every inner face
 * tried in turn).  Returns -1 if none. */
static int64_t mk_locate_inner(const or_mesh *m, const double *x, const double *P)
{
    for (int64_t f = 0; f < m->n_faces; f++) {
        if (m->face_bc[f] != OR_BC_INNER_HARD) continue;
        const int32_t *fn = m->face_nodes + 4 * f;
        const double *n0 = x + 3 * fn[0], *n1 = x + 3 * fn[1], *n2 = x + 3 * fn[2], *n3 = x + 3 * fn[3];
        if (mk_origin_ray_tri(P, n0, n1, n2) || mk_origin_ray_tri(P, n0, n2, n3)) return m->face_cells[2 * f];
    }
    return -1;
}

/* ------------------------------------------------------------------------------------------ */
/* Monte-Carlo zones: Setka::Set_MK_Zone (Setka_MK.cpp:142-683), the face distributions of the
 * zone borders (AMR_f.cpp:4-50,240-316; AMR_cell.cpp:273-327; Gran.cpp:16-99) and the counts of
 * MK_prepare / MK_go(zone) (Setka_MK.cpp:686-1100,1225-1400). */
int or_set_mk_zone(const or_mesh *m, const or_geom *g, int mistake, int8_t *cell_mk, uint8_t *face_mask)
{
    for (int64_t c = 0; c < m->n_cells; c++) {
        const int z = m->cell_zone[c];
        const double x = g->cell_center[3 * c];
        if (z == 1) cell_mk[c] = 1;
        else if (z == 2) cell_mk[c] = x > 0 ? 2 : 3;
        else if (z == 3) cell_mk[c] = x > 0 ? 4 : 5;
        else if (z == 4) cell_mk[c] = x > 0 ? 6 : 7;
        else return -1;
    }
    for (int64_t f = 0; f < m->n_faces; f++) {
        const int bc = m->face_bc[f];
        const int fl = m->face_flags ? m->face_flags[f] : 0;
        const int ts = (fl >> 3) & 1, hp = (fl >> 2) & 1, bs = (fl >> 4) & 1;
        const int us2 = !ts && !hp && !bs; /* Gran::type2 == Us */
        const int64_t a = m->face_cells[2 * f], b = m->face_cells[2 * f + 1];
        const int two = b >= 0; /* gr->cells.size() == 2 */
        const int za = m->cell_zone[a], zb = two ? m->cell_zone[b] : 0;
        const double X = g->face_center[3 * f];
        const double xa = g->cell_center[3 * a], xb = two ? g->cell_center[3 * b] : 0.0;
        const int across = two && xa * xb < 0.0;
        int n[7] = {0, 0, 0, 0, 0, 0, 0};
        if (ts) n[0]++; /* zone 1: Gran_TS */
        if (!(X < -0.000001)) { /* zone 2 */
            if (hp) n[1]++;
            if (ts) n[1]++;
            if (bc == OR_BC_OUTER_SOFT && za == 2) n[1]++;
            if (two && za == 2 && zb == 2 && across && !(mistake & 1)) n[1]++;
        }
        if (!(X > 0.000001)) { /* zone 3 */
            if (hp) n[2]++;
            if (ts) n[2]++;
            if (bc == OR_BC_OUTER_SOFT && za == 2) n[2]++;
            if (two && us2) {
                if (za == 2 && zb == 3) n[2]++;
                if (za == 2 && zb == 2 && across) n[2]++;
            }
        }
        if (hp && X > 0) n[3]++; /* zone 4 */
        if (bs) n[3]++;
        if (two && za == 3 && zb == 3 && across) n[3]++;
        if (hp && X < 0.0) n[4]++; /* zone 5 */
        if (bc == OR_BC_OUTER_SOFT && za == 3) n[4]++;
        if (two && us2) {
            if (za == 2 && zb == 3) n[4]++;
            if (za == 3 && zb == 3 && across) n[4]++;
            if (za == 3 && zb == 4 && xa < 0) n[4]++;
        }
        if (bs) n[5]++; /* zone 6 */
        if (bc == OR_BC_OUTER_HARD && X > 0) n[5]++;
        if (two && za == 4 && zb == 4 && across) n[5]++;
        if (bc != OR_BC_INTERIOR && X < 0 && za == 4) n[6]++; /* zone 7 */
        if (two && us2) {
            if (za == 3 && zb == 4 && xa < 0) n[6]++;
            if (za == 4 && zb == 4 && across) n[6]++;
        }
        uint8_t mask = 0;
        for (int z = 0; z < 7; z++) {
            if (n[z] > 1) return -2;
            if (n[z]) mask |= (uint8_t)(1u << z);
        }
        face_mask[f] = mask;
    }
    return 0;
}

static int64_t mk_dist_bins(const or_mk_dist *d) { return 108 * (int64_t)d->r * d->r * d->r; }

static double *mk_dist_at(const or_mk_dist *d, int32_t slot, int sort, int side)
{
    return d->w + (((int64_t)slot * 4 + (sort - 1)) * 2 + side) * mk_dist_bins(d);
}

/* the face frame of a distribution: Vn = the normal (side 0) or its negative (side 1), then
 * get_bazis (Gran::Read_AMR, AMR_f::Set_bazis) */
static void mk_dist_frame(const double *n, int side, double *e, double *t, double *mm)
{
    for (int i = 0; i < 3; i++) e[i] = side ? -n[i] : n[i];
    or_get_bazis(e, t, mm);
}

void or_mk_deposit(or_mk_dist *d, int32_t slot, int sort, int side, const double *n, const double *V, double mu)
{
    double e[3], t[3], mm[3];
    mk_dist_frame(n, side, e, t, mm);
    const int nx = 3 * d->r, ny = 6 * d->r, nz = 6 * d->r;
    const double xL = 0.0, xR = 20.0, yL = -20.0, yR = 20.0, zL = -20.0, zR = 20.0;
    double x = dot3(V, e), y = dot3(V, t), z = dot3(V, mm);
    if (x < xL) x = xL + 1e-7;
    if (x > xR) x = xR - 1e-7;
    if (y < yL) y = yL + 1e-7;
    if (y > yR) y = yR - 1e-7;
    if (z < zL) z = zL + 1e-7;
    if (z > zR) z = zR - 1e-7;
    const double dx = (xR - xL) / nx, dy = (yR - yL) / ny, dz = (zR - zL) / nz;
    int i = (x >= xL && x <= xR) ? (int)((x - xL) / dx) : 0;
    int j = (y >= yL && y <= yR) ? (int)((y - yL) / dy) : 0;
    int k = (z >= zL && z <= zR) ? (int)((z - zL) / dz) : 0;
    if (i >= nx) i = nx - 1;
    if (j >= ny) j = ny - 1;
    if (k >= nz) k = nz - 1;
    double *w = mk_dist_at(d, slot, sort, side);
#pragma omp atomic
    w[((int64_t)i * ny + j) * nz + k] += mu;
}

/* AMR_f::Get_random_velosity with AMR_cell::Get_random_velosity_in_cell: a bin with probability
 * its weight, then Vn with density ~ Vn across the bin, Vt and Vm uniform */
static void mk_dist_sample(uint64_t *rs, const or_mk_dist *d, const double *w, const double *n, int side, double *V)
{
    const int nx = 3 * d->r, ny = 6 * d->r, nz = 6 * d->r;
    const int64_t nb = mk_dist_bins(d);
    double total = 0.0;
    for (int64_t b = 0; b < nb; b++) total += w[b];
    const double ksi = mk_rand(rs) * total;
    double SS = 0.0;
    int64_t pick = -1, last = 0;
    for (int64_t b = 0; b < nb; b++) {
        if (w[b] > 0.0) last = b;
        if (SS + w[b] > ksi) { pick = b; break; }
        SS += w[b];
    }
    if (pick < 0) pick = last; /* rounding at the top of the sum */
    const int i = (int)(pick / ((int64_t)ny * nz)), j = (int)((pick / nz) % ny), k = (int)(pick % nz);
    const double dx = 20.0 / nx, dy = 40.0 / ny, dz = 40.0 / nz;
    const double xl = 0.0 + i * dx, xr = 0.0 + (i + 1) * dx, yl = -20.0 + j * dy, yr = -20.0 + (j + 1) * dy,
                 zl = -20.0 + k * dz, zr = -20.0 + (k + 1) * dz;
    const double cx = (xr + xl) / 2.0, cy = (yr + yl) / 2.0, cz = (zr + zl) / 2.0;
    const double hx = xr - xl, hy = yr - yl, hz = zr - zl;
    const double L = cx - hx / 2.0, R = cx + hx / 2.0;
    const double v0 = sqrt(L * L + mk_rand(rs) * (R * R - L * L));
    const double v1 = (cy - hy / 2.0) + mk_rand(rs) * hy;
    const double v2 = (cz - hz / 2.0) + mk_rand(rs) * hz;
    double e[3], t[3], mm[3];
    mk_dist_frame(n, side, e, t, mm);
    for (int q = 0; q < 3; q++) V[q] = (v0 * e[q] + v1 * t[q]) + v2 * mm[q];
}

void or_mk_sample_batch(const or_mk_dist *d, int32_t slot, int sort, int side, const double *n, int64_t n_draws,
                        uint64_t seed, double *V)
{
    const double *w = mk_dist_at(d, slot, sort, side);
    for (int64_t i = 0; i < n_draws; i++) {
        uint64_t rs = mk_mix(seed ^ mk_mix((uint64_t)i + 0x9E3779B97F4A7C15ull));
        mk_dist_sample(&rs, d, w, n, side, V + 3 * i);
    }
}

int64_t or_mk_setup_zone(const or_mesh *m, const or_geom *g, double v_inf, int32_t n_per_face, int32_t zone,
                         const int8_t *cell_mk, const uint8_t *face_mask, const or_mk_dist *dist, int32_t *face,
                         int8_t *sort, int8_t *side, int32_t *count, double *mu)
{
    const uint8_t bit = (uint8_t)(1u << (zone - 1));
    const int64_t nb = dist ? mk_dist_bins(dist) : 0;
    double S = 0.0;
    int64_t nlist = 0;
    double *ff = (double *)calloc((size_t)m->n_faces, sizeof(double));      /* flux of the face */
    double *fs = (double *)calloc((size_t)m->n_faces * 4, sizeof(double));  /* per sort */
    for (int64_t f = 0; f < m->n_faces; f++) {
        if (!(face_mask[f] & bit)) continue;
        nlist++;
        const int64_t a = m->face_cells[2 * f], b = m->face_cells[2 * f + 1];
        const int ni2 = cell_mk[a] == zone ? 1 : 0; /* the side entering the zone */
        const int64_t enter = ni2 ? a : b;
        if (enter < 0 || cell_mk[enter] != zone) { free(ff); free(fs); return -3; }
        if (m->face_bc[f] == OR_BC_INTERIOR) {
            const int32_t sl = dist ? dist->slot[f] : -1;
            for (int s = 1; s <= 4 && sl >= 0; s++) {
                const double *w = mk_dist_at(dist, sl, s, ni2);
                double v = 0.0;
                for (int64_t q = 0; q < nb; q++) v += w[q];
                fs[4 * f + s - 1] = v;
                ff[f] += v;
            }
        } else {
            /* the boundary Maxwellian of sort 4 (Get_Spotok_inf of the outward normal) */
            const double v = mk_spotok_inf(v_inf, g->face_normal + 3 * f) * g->face_area[f];
            fs[4 * f + 3] = v;
            ff[f] = v;
        }
        S += ff[f];
    }
    int64_t k = 0;
    if (nlist > 0 && S > 0.0) { /* no inflow at all (no pass has fed the zone yet): no particles */
        const double mu_expect = S / (1.0 * n_per_face * nlist);
        for (int s = 1; s <= 4; s++) {
            for (int64_t f = 0; f < m->n_faces; f++) {
                if (!(face_mask[f] & bit)) continue;
                if (ff[f] < 0.0000001 * S / nlist) continue;
                const double sv = fs[4 * f + s - 1];
                if (sv < 0.0000001 * ff[f]) continue;
                int32_t n = (int32_t)(sv / mu_expect) + 1;
                const int32_t lo = n_per_face < 100 ? n_per_face : 100;
                if (n < lo) n = lo;
                face[k] = (int32_t)f;
                sort[k] = (int8_t)s;
                side[k] = m->face_bc[f] == OR_BC_INTERIOR ? (int8_t)(cell_mk[m->face_cells[2 * f]] == zone ? 1 : 0)
                                                          : (int8_t)-1;
                count[k] = n;
                mu[k] = sv / n;
                k++;
            }
        }
    }
    free(ff);
    free(fs);
    return k;
}

/* Cell::MK_Add_pui_source (Cell.cpp:1087-1133) with parent 0 (a thermal proton), at the zone of
 * Setka::determ_zone(P.cel, 0) (Setka_MK.cpp:2010, 2088) */
static void mk_pui_add(const or_mk *mk, int64_t c, int zone, int sort, double u, double nu_ex, double mu, double t)
{
    const int nw = mk->pui_nw;
    const double xi = u / mk->pui_wr * nw;
    int iw = xi >= (double)nw ? nw - 1 : (xi > 0.0 ? (int)xi : 0);
    if (iw >= nw) iw = nw - 1;
#pragma omp atomic
    mk->pui_sm[c * nw + iw] += mu * t;
    if (zone < 1 || zone > 4 || sort < 1 || sort > 4) return;
    const int k = mk->pui_arise[(zone - 1) * 4 + (sort - 1)];
    if (k == 0) return;
#pragma omp atomic
    mk->pui_sp[(c * nw + iw) * 2 + (k - 1)] += mu * t * nu_ex;
}

void or_mk_pui_finish(int64_t n, int32_t nw, double wr, double a2, double kn, const double *vol, const double *sm,
                      const double *sp, double *sm_out, double *sp_out)
{
    const double pi = 3.14159265358979323846;
    const double dthe = pi / 40.0;
    double *ff = (double *)malloc(sizeof(double) * (size_t)nw);
    for (int64_t c = 0; c < n; c++) {
        /* MK_normir_Moments */
        for (int ij = 0; ij < nw; ij++) {
            const double w1 = ij * wr / nw, w2 = (ij + 1) * wr / nw;
            ff[ij] = sm[c * nw + ij] / vol[c];
            for (int ki = 0; ki < 2; ki++)
                sp_out[(c * nw + ij) * 2 + ki] =
                    sp[(c * nw + ij) * 2 + ki] / (vol[c] * 4.0 * pi * (1.0 / 3.0) * (w2 * w2 * w2 - w1 * w1 * w1));
        }
        /* MK_calc_Sm */
        for (int ij = 0; ij < nw; ij++) {
            const double w = ((ij + 0.5) * wr / nw);
            double acc = 0.0;
            for (int j = 0; j < nw; j++) {
                const double Vh = ((j + 0.5) * wr / nw);
                const double f = ff[j];
                if (f <= 0.0) continue;
                for (int k = 1; k <= 40; k++) {
                    const double the = dthe * k;
                    const double d = sqrt(Vh * Vh + w * w - 2.0 * w * Vh * cos(the));
                    if (d > 0.000000001) {
                        const double sg = (1.0 - a2 * log(d)) * (1.0 - a2 * log(d));
                        acc += f * d * sg * sin(the) * dthe * 2.0 * pi;
                    }
                }
            }
            acc = acc / (4.0 * pi);
            sm_out[c * nw + ij] = acc / kn;
        }
    }
    free(ff);
}

/* Phys_param::Plasma_components_1 (Phys_param.cpp:628-665): the thermal protons of a plasma cell
 * holding helium, th = (rho_Th, p_Th, T_Th); al = 2 electrons per helium ion for zone <= 2, else 1.
 * Pinned by mass (rho = rho_Th (1 + me) + rho_He (1 + al me / 4)), by p_Th = rho_Th T_Th / 2, and by
 * rho_He = 0 (rho_Th = rho / (1 + me), p_Th = p / 2). */
void or_plasma_components_1(double rho, double p, double rho_He, int zone, double th[3]) {
    const double me = OR_ME_MP;
    const int al = zone <= 2 ? 2 : 1;
    th[0] = -(me * al * rho_He + 4.0 * (-rho + rho_He)) / (4.0 * (1.0 + me));
    th[1] = p * (4.0 * rho - (4.0 + al * me) * rho_He) / (8.0 * rho + (-7.0 + al + (1.0 - al) * me) * rho_He);
    th[2] = 8.0 * (1.0 + me) * p / (8.0 * rho - (7.0 - al + (-1.0 + al) * me) * rho_He);
}

int or_mk_go(const or_mesh *m, const or_mk *mk, const double *x, const double *prim, const or_geom *g,
             double *tally, int64_t *stats)
{
    for (int i = 0; i < 8; i++) stats[i] = 0;
    /* particles are numbered face by face; first[q] = id of face q's first particle */
    int64_t *first = (int64_t *)malloc(sizeof(int64_t) * (size_t)(mk->n_in + 1));
    first[0] = 0;
    for (int64_t q = 0; q < mk->n_in; q++) first[q + 1] = first[q] + mk->count[q];
    const int64_t n_part = first[mk->n_in];
    int nt = mk->nthreads > 0 ? mk->nthreads : 1;
    int64_t s0_ = 0, s1_ = 0, s2_ = 0, s3_ = 0, s4_ = 0, s5_ = 0, s6_ = 0, s7_ = 0, s8_ = 0;
    /* one particle per iteration; its history depends only on its id (its random stream); with
     * several threads the tallies are added atomically, so their rounding varies with the order */
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) \
    reduction(+ : s0_, s1_, s2_, s3_, s4_, s5_, s6_, s7_, s8_)
    for (int64_t pid = 0; pid < n_part; pid++) {
        int64_t st_[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int64_t *stats = st_;
        int64_t lo = 0, hi = mk->n_in - 1;
        while (lo < hi) {
            int64_t mid = (lo + hi + 1) >> 1;
            if (first[mid] <= pid) lo = mid;
            else hi = mid - 1;
        }
        const int64_t q = lo;
        int64_t f = mk->face[q];
        const double *n = g->face_normal + 3 * f;
        const int32_t *fn = m->face_nodes + 4 * f;
        do {
            uint64_t rs = mk_mix(mk->seed ^ mk_mix((uint64_t)pid + 0x9E3779B97F4A7C15ull));
            int capped = 0;
            /* Gran::Get_Random_pozition (Gran.cpp:220-285) */
            double s0 = 0.0, s1 = 0.0;
            const double *A = x + 3 * fn[0], *B1 = x + 3 * fn[1], *Cc = x + 3 * fn[2], *B2 = x + 3 * fn[3];
            {
                double v1x = B1[0] - A[0], v1y = B1[1] - A[1], v1z = B1[2] - A[2];
                double v2x = Cc[0] - A[0], v2y = Cc[1] - A[1], v2z = Cc[2] - A[2];
                double cx = v1y * v2z - v1z * v2y, cy = v1z * v2x - v1x * v2z, cz = v1x * v2y - v1y * v2x;
                s0 = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
                v1x = B2[0] - A[0]; v1y = B2[1] - A[1]; v1z = B2[2] - A[2];
                cx = v1y * v2z - v1z * v2y; cy = v1z * v2x - v1x * v2z; cz = v1x * v2y - v1y * v2x;
                s1 = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            }
            double S = s0 + s1;
            double ksi = mk_rand(&rs);
            const double *T1 = (ksi <= s0 / S) ? B1 : B2;
            double r1 = mk_rand(&rs), r2 = mk_rand(&rs);
            double P[3], V[3];
            for (int i = 0; i < 3; i++)
                P[i] = (1.0 - sqrt(r1)) * A[i] + (sqrt(r1) * (1.0 - r2)) * T1[i] + (r2 * sqrt(r1)) * Cc[i];
            const int in_side = mk->in_side ? mk->in_side[q] : -1;
            if (in_side >= 0) /* from the face's distribution (AMR_f::Get_random_velosity) */
                mk_dist_sample(&rs, mk->dist, mk_dist_at(mk->dist, mk->dist->slot[f], mk->in_sort[q], in_side), n,
                               in_side, V);
            else
                mk_velosity_initial(&rs, mk->v_inf, n, V, &capped, mk->mistake);
            stats[0]++;
            if (capped) { stats[5]++; break; } /* no admissible velocity within the cap: dropped */
            int64_t c = in_side == 0 ? m->face_cells[2 * f + 1] : m->face_cells[2 * f];
            const double *cc = g->cell_center + 3 * c;
            for (int i = 0; i < 3; i++) P[i] = P[i] + (cc[i] - P[i]) / 1000.0; /* into the cell */
            double KSI = -log(1.0 - mk_rand(&rs)), I_do = 0.0, mu = mk->mu[q];
            int sort = mk->in_sort ? mk->in_sort[q] : 4;
            int done = 0;
            for (int it = 0; it < mk->max_steps && !done; it++) {
                double time = 0.0;
                int64_t ex = -1;
                int found = 0;
                for (int k1 = 0; k1 < 12 && !found; k1++) {
                    found = mk_exit(m, x, c, P, V, &time, &ex);
                    if (!found) {
                        const double *cn = g->cell_center + 3 * c;
                        for (int i = 0; i < 3; i++) P[i] = (k1 < 10) ? P[i] + (cn[i] - P[i]) / 800.0 : cn[i];
                    }
                }
                if (!found) { stats[4]++; done = 1; break; }
                stats[6]++;
                const double *s = prim + 8 * c;
                /* the thermal protons at the cell's type, as the flight calls it (Setka_MK.cpp:1944-1966) */
                double th[3];
                or_plasma_components_1(s[0], s[4], mk->scal ? mk->scal[2 * c + 1] : 0.0, m->cell_zone[c], th);
                double rho_Th = th[0], p_Th = th[1];
                if (rho_Th <= 1e-8) rho_Th = 1e-8;
                if (p_Th <= 1e-8 / 2.0) p_Th = 1e-8 / 2.0;
                double ro = rho_Th, cp = sqrt(2.0 * p_Th / rho_Th);
                double l = sqrt((time * V[0]) * (time * V[0]) + (time * V[1]) * (time * V[1]) + (time * V[2]) * (time * V[2]));
                double I = I_do;
                double u = sqrt((V[0] - s[1]) * (V[0] - s[1]) + (V[1] - s[2]) * (V[1] - s[2]) + (V[2] - s[3]) * (V[2] - s[3]));
                double vn = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
                double uz = mk_velosity_1(u, cp);
                double nu_ex = ro * uz * mk_sigma(mk->a2, uz) / mk->kn;
                double sig = vn / nu_ex;
                I += l / sig;
                if (I < KSI) {
                    I_do = I;
                    {
#pragma omp atomic
                    tally[5 * c] += time * mu;
                    }
                    if (sort >= 1 && sort <= 4) {
#pragma omp atomic
                        tally[5 * c + sort] += time * mu;
                    }
                    if (mk->pui_nw > 0)
                        mk_pui_add(mk, c, mk->scal ? or_determ_zone(mk->scal[2 * c], s[0], m->cell_zone[c]) : m->cell_zone[c],
                                   sort, (mk->mistake & 8) ? vn : u, nu_ex, mu, time);
                    for (int i = 0; i < 3; i++) P[i] += 1.000001 * time * V[i];
                    int64_t nc = m->face_cells[2 * ex] == c ? m->face_cells[2 * ex + 1] : m->face_cells[2 * ex];
                    if (nc < 0 && m->face_bc[ex] != OR_BC_INNER_HARD) { /* out through the outer boundary */
                        stats[2]++;
                        done = 1;
                    } else if (nc < 0 && (mk->mistake & 4)) { /* planted: absorbed at the inner sphere */
                        stats[2]++;
                        done = 1;
                    } else if (mk->R0 > 0.0 && sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]) < 1.01 * mk->R0) {
                        /* through the Sun's sphere to its far side (Setka_MK.cpp:2138-2160), on into
                           the cell there */
                        double t_;
                        if (mk_sphere_time(P, V, 1.01 * mk->R0, &t_)) {
                            for (int i = 0; i < 3; i++) P[i] += t_ * V[i];
                        } else {
                            const double nv = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
                            for (int i = 0; i < 3; i++) P[i] += (2.02 * mk->R0) * V[i] / nv;
                        }
                        stats[3]++;
                        c = mk_locate_inner(m, x, P);
                        if (c < 0) { stats[4]++; done = 1; }
                    } else if (nc < 0) { /* an inner face without the sphere: nowhere to go */
                        stats[4]++;
                        done = 1;
                    } else if (mk->zone > 0 && (mk->face_mask[ex] >> (mk->zone - 1)) & 1) {
                        /* the zone's border (Setka_MK.cpp:2165-2184): into the face's outgoing
                           distribution, on the side the particle leaves by */
                        const int nn = mk->cell_mk[m->face_cells[2 * ex]] == mk->zone ? 0 : 1;
                        if (sort >= 1 && sort <= 4 && mk->dist->slot[ex] >= 0)
                            or_mk_deposit(mk->dist, mk->dist->slot[ex], sort, nn, g->face_normal + 3 * ex, V, mu);
                        stats[8]++;
                        done = 1;
                    } else if (mk->zone > 0 && mk->cell_mk[nc] != mk->zone) {
                        stats[4]++; /* into another zone through a face no list holds: lost */
                        done = 1;
                    } else {
                        c = nc;
                    }
                } else {
                    double kk = (KSI - I_do) * sig;
                    double t_ex = kk / vn;
                    I_do = 0.0;
                    for (int i = 0; i < 3; i++) P[i] += t_ex * V[i];
                    {
#pragma omp atomic
                    tally[5 * c] += t_ex * mu;
                    }
                    if (sort >= 1 && sort <= 4) { /* other sorts: total only */
#pragma omp atomic
                        tally[5 * c + sort] += t_ex * mu;
                    }
                    if (mk->pui_nw > 0)
                        mk_pui_add(mk, c, mk->scal ? or_determ_zone(mk->scal[2 * c], s[0], m->cell_zone[c]) : m->cell_zone[c],
                                   sort, (mk->mistake & 8) ? vn : u, nu_ex, mu, t_ex);
                    double U[3], Vs[3], W[3];
                    double Vp[3] = {s[1], s[2], s[3]};
                    mk_to_sph(P, Vp, &U[0], &U[1], &U[2]);      /* (r, theta, phi) */
                    mk_to_sph(P, V, &Vs[0], &Vs[1], &Vs[2]);
                    for (int i = 0; i < 3; i++) { U[i] /= cp; Vs[i] /= cp; }
                    mk_change_velosity(&rs, mk->a2, U, Vs, W, cp, &capped, mk->mistake);
                    for (int i = 0; i < 3; i++) W[i] *= cp;
                    mk_from_sph(P, W[0], W[1], W[2], V);
                    /* P.sort = determ_zone(P.cel, 0) (Setka_MK.cpp:2088, 2121): Q routes it when
                     * the scalars are carried; without them the cell's type stands in (what Q gives
                     * where the marker follows the types, Q/rho >= 50 exactly past the HP) */
                    sort = mk->scal ? or_determ_zone(mk->scal[2 * c], s[0], m->cell_zone[c]) : m->cell_zone[c];
                    KSI = -log(1.0 - mk_rand(&rs));
                    stats[1]++;
                }
                if (it == mk->max_steps - 1 && !done) { stats[4]++; done = 1; }
            }
            if (capped) stats[7]++; /* some exchange sampler of the flight gave up */
        } while (0);
        s0_ += st_[0]; s1_ += st_[1]; s2_ += st_[2]; s3_ += st_[3]; s4_ += st_[4]; s5_ += st_[5]; s6_ += st_[6];
        s7_ += st_[7];
        s8_ += st_[8];
    }
    stats[0] = s0_; stats[1] = s1_; stats[2] = s2_; stats[3] = s3_; stats[4] = s4_; stats[5] = s5_;
    stats[6] = s6_; stats[7] = s7_;
    if (mk->zone > 0) stats[8] = s8_;
    free(first);
    return 0;
}

int or_go(const or_mesh *m, const or_params *p, double *coords, double *prim, const double *disp,
          int64_t steps, int gcl, int nthreads, int *final_layer, double *dt_hist, double *t_sim,
          int n_h, const double *h_fluids, const or_fit *fit)
{
    return or_go_k(m, p, coords, prim, NULL, disp, steps, gcl, nthreads, final_layer, dt_hist, t_sim, n_h, h_fluids,
                   fit);
}

int or_go_k(const or_mesh *m, const or_params *p, double *coords, double *prim, double *scal, const double *disp,
            int64_t steps, int gcl, int nthreads, int *final_layer, double *dt_hist, double *t_sim,
            int n_h, const double *h_fluids, const or_fit *fit)
{
    int64_t nn3 = 3 * m->n_nodes, nc8 = 8 * m->n_cells, nc2 = 2 * m->n_cells;
    /* charge-exchange sources of the step's old layer (Setka_physics.cpp:1741), when hydrogen
     * fluids are given */
    double *src = h_fluids ? (double *)malloc(sizeof(double) * 5 * m->n_cells) : NULL;
    double *X[2] = {coords, coords + nn3};
    double *S[2] = {prim, prim + nc8};
    double *K[2] = {scal, scal ? scal + nc2 : NULL};
    double *kflux = scal ? (double *)malloc(sizeof(double) * 2 * m->n_faces) : NULL;
    or_geom G[2];
    geom_alloc(&G[0], m);
    geom_alloc(&G[1], m);
    double *flux = (double *)malloc(sizeof(double) * 9 * m->n_faces);
    double *dtf = (double *)malloc(sizeof(double) * m->n_faces);
    memcpy(X[1], X[0], sizeof(double) * nn3);
    if (gcl) { or_measure_gcl(m, X[0], &G[0]); or_measure_gcl(m, X[1], &G[1]); }
    else { or_measure(m, X[0], &G[0]); or_measure(m, X[1], &G[1]); }
    memcpy(S[1], S[0], sizeof(double) * nc8);
    if (scal) memcpy(K[1], K[0], sizeof(double) * nc2);
    int now1 = 1, now2 = 0;
    double time = 0.00001, loc_time = 0.00001;
    double tsim = t_sim ? *t_sim : 0.0;
    int status = 0;
    for (int64_t step = 1; step <= steps; step++) {
        now2 = now1;
        now1 = (now1 + 1) % 2;
        time = loc_time;
        loc_time = 100000000.0;
        if (dt_hist) dt_hist[step - 1] = time;
        if (p->move_mesh == 2 && fit) {
            /* Setka_physics.cpp:1616-1627: surfaces, then every ray, then the new geometry */
            int rc0 = or_culc_velocity_surface(m, p, fit, X[now1], S[now1], &G[now1], time, X[now2], NULL);
            if (rc0 && !status) status = rc0;
            or_dvigenie(fit, X[now2]);
            if (fit->ext) or_dvigenie_ext(fit->ext, X[now2]);
            if (gcl) or_measure_gcl(m, X[now2], &G[now2]);
            else or_measure(m, X[now2], &G[now2]);
        } else if (p->move_mesh) {
            or_motion(m->n_nodes, X[now1], disp, time, tsim, p->motion_omega, p->motion_phase, X[now2]);
            if (gcl) or_measure_gcl(m, X[now2], &G[now2]);
            else or_measure(m, X[now2], &G[now2]);
        }
        int rc = face_flux_impl(m, p, S[now1], K[now1], &G[now1], X[now1], X[now2], time, gcl, flux, kflux, dtf, NULL,
                                nthreads);
        if (rc && !status) status = rc;
        for (int64_t f = 0; f < m->n_faces; f++)
            if (dtf[f] < loc_time) loc_time = SMIN(loc_time, dtf[f]);
        if (src) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
            for (int64_t c = 0; c < m->n_cells; c++)
                or_bera_source_k(S[now1] + 8 * c, scal ? K[now1][2 * c] : 0.0, scal ? K[now1][2 * c + 1] : 0.0,
                                 m->cell_zone[c], n_h, h_fluids + (int64_t)5 * n_h * c, p->par_a_2, p->par_n_H_LISM,
                                 p->par_Kn, src + 5 * c);
        }
        rc = update_impl(m, p, S[now1], K[now1], G[now1].cell_volume, G[now2].cell_volume, flux, kflux, src, time,
                         S[now2], K[now2], NULL, NULL, nthreads);
        if (rc && !status) status = rc;
        tsim += time;
        if (status) break;
    }
    if (final_layer) *final_layer = now2;
    if (t_sim) *t_sim = tsim;
    geom_free(&G[0]);
    geom_free(&G[1]);
    free(flux);
    free(kflux);
    free(dtf);
    free(src);
    return status;
}

/* ------------------------------------------------------------------------------------------ */
/* Luch::dvigenie of the B, C/C2, D, E, H and G rays (Luch.cpp:153-739), each restated in its
 * operation order.  P(k) is node k of the ray; r_yz the distance of a node from the x axis
 * (sqrt(kv(y) + kv(z)), Yzel::func_Ryz). */
static double or_ryz(const double *p) { return sqrt(p[1] * p[1] + p[2] * p[2]); }

/* the cubic from (x0, y0) with slope (t1, tt1) to (x1, y1) with slope (t2, tt2) at parameter u */
static void or_hermite(double x0, double y0, double x1, double y1, double t1, double tt1, double t2, double tt2,
                       double u, double *x, double *y)
{
    double a1 = x0, b1 = t1, c1 = -2 * t1 - t2 - 3.0 * x0 + 3.0 * x1, d1 = t1 + t2 + 2.0 * x0 - 2.0 * x1;
    double a2 = y0, b2 = tt1, c2 = -2 * tt1 - tt2 - 3.0 * y0 + 3.0 * y1, d2 = tt1 + tt2 + 2.0 * y0 - 2.0 * y1;
    *x = a1 + b1 * u + c1 * u * u + d1 * u * u * u;
    *y = a2 + b2 * u + c2 * u * u + d2 * u * u * u;
}

/* the two spacing laws of the rays: s (a - (2a + b - 3) s + (a + b - 2) s^2) (slopes a, b at the
 * ends) and a s - (3a + b - 4) s^3 + (2a + b - 3) s^4 */
static double or_sp1(double s, double a, double b) { return s * (a - (2.0 * a + b - 3.0) * s + (a + b - 2.0) * s * s); }
static double or_sp2(double s, double a, double b)
{
    return a * s - (3.0 * a + b - 4.0) * s * s * s + (2.0 * a + b - 3.0) * s * s * s * s;
}

void or_dvigenie_ext(const or_rays_ext *R, double *x)
{
    static const int order[6] = {OR_RAY_B, OR_RAY_C, OR_RAY_D, OR_RAY_E, OR_RAY_H, OR_RAY_G};
    for (int fam = 0; fam < 6; fam++) {
        for (int64_t k = 0; k < R->n; k++) {
            if (R->type[k] != order[fam]) continue;
            const int32_t *Y = R->nodes + R->off[k];
            const int32_t *O = R->opor + 4 * k;
            double st, ct, sp, cp;
            or_sincos(R->angles[2 * k], &st, &ct);
            or_sincos(R->angles[2 * k + 1], &sp, &cp);
#define P(i) (x + 3 * (int64_t)Y[i])
#define Q(i) (x + 3 * (int64_t)O[i])
#define PUT_R(i, rad)                                   \
    do {                                                \
        double *X_ = P(i);                              \
        X_[0] = (rad) * ct;                       \
        X_[1] = (rad) * st * cp;            \
        X_[2] = (rad) * st * sp;            \
    } while (0)
#define PUT_P(i, xx, rad)                               \
    do {                                                \
        double *X_ = P(i);                              \
        X_[0] = (xx);                                   \
        X_[1] = (rad) * cp;                       \
        X_[2] = (rad) * sp;                       \
    } while (0)
            const int M0 = R->M0, M1 = R->M1, M11 = R->M11, M2 = R->M2, M3 = R->M3, M4 = R->M4;
            const double R0 = R->R0, R1 = R->R1, R5 = R->R5;
            int num = 0;
            double r = 0.0, xx, yy;
            if (R->type[k] == OR_RAY_B) { /* Luch.cpp:254-421 */
                const double *o1 = Q(1), *o2 = Q(2), *o3 = Q(3);
                double R2 = sqrt(KVV(o1[0], o1[1], o1[2]));
                double R3 = sqrt(KV(o2[1]) + KV(o2[2]));
                double R4 = sqrt(KV(o3[1]) + KV(o3[2]));
                for (int j = 0; j < M0; j++) {
                    r = R0 + j * (R1 - R0) / M0;
                    PUT_R(j + num, r);
                }
                num += M0;
                PUT_R(num, R1);
                num += 1;
                for (int j = 0; j < M1; j++) {
                    r = R1 + (j + 1) * (R2 - R1) / (M1 + 1);
                    PUT_R(j + num, r);
                }
                num += M1 + 1;
                double dr = (R2 - R1) / (M1 + 1);
                for (int j = 0; j < M11; j++) {
                    r = R2 + dr * (j + 1);
                    PUT_R(num + j, r);
                }
                num += M11;
                double x3 = r * ct;
                Q(2)[0] = x3;
                Q(3)[0] = x3;
                double t1 = ct * R->dd3, tt1 = st * R->dd3, t2 = 0.0, tt2 = 1.0 * R->dd4;
                double x0 = r * ct, y0 = r * st;
                for (int j = 0; j < M2 - M11; j++) {
                    double s = 1.0 * (j + 1) / (M2 - M11 + 1);
                    double ss = or_sp1(s, R->ba1, R->ba2);
                    or_hermite(x0, y0, x3, R3, t1, tt1, t2, tt2, ss, &xx, &yy);
                    PUT_P(num + j, xx, yy);
                }
                num += (M2 - M11) + 1;
                for (int j = 0; j < M3; j++) {
                    double s = (j + 1.0) / (M3 + 1.0);
                    r = R3 + or_sp1(s, R->ba3, R->ba4) * (R4 - R3);
                    PUT_P(num + j, x3, r);
                }
                num += M3 + 1;
                for (int j = 0; j < M4; j++) {
                    double s = (j + 1.0) / (M4 + 1.0);
                    r = R4 + (R5 - R4) * or_sp2(s, R->ba5, 2.0);
                    PUT_P(num + j, x3, r);
                }
                num += M4;
                PUT_P(num, x3, R5);
            } else if (R->type[k] == OR_RAY_C) { /* Luch.cpp:153-252 (C_Luch and C2_Luch) */
                const double *o1 = Q(1);
                double R2 = sqrt(KVV(o1[0], o1[1], o1[2]));
                const int dd5 = (int)R->dd5, dd6 = (int)R->dd6; /* int in the reference */
                for (int j = 0; j < M0; j++) {
                    r = R0 + j * (R1 - R0) / M0;
                    PUT_R(num + j, r);
                }
                num += M0 + 1;
                for (int j = 0; j < M1; j++) {
                    r = R1 + (j + 1) * (R2 - R1) / (M1 + 1);
                    PUT_R(num + j, r);
                }
                double dr = (R2 - R1) / (M1 + 1);
                num += M1 + 1;
                for (int j = 0; j < M11; j++) {
                    r = R2 + dr * (j + 1);
                    PUT_R(num + j, r);
                }
                num += M11;
                double x0 = r * ct, y0 = r * st;
                double t1 = (ct - 1.0) / 2.0 * dd5, tt1 = st * dd5, t2 = -1.0 * dd6, tt2 = 0.0;
                double y = 0.0, z = 0.0;
                for (int j = 0; j < R->N4; j++) {
                    double s = 1.0 * (j + 1) / (R->N4 + 1);
                    or_hermite(x0, y0, R->L6, y0, t1, tt1, t2, tt2, s, &xx, &yy);
                    z = yy * sp;
                    y = yy * cp;
                    P(num + j)[0] = xx;
                    P(num + j)[1] = y;
                    P(num + j)[2] = z;
                }
                num += R->N4;
                for (int j = 0; j < R->N5; j++) {
                    P(num + j)[0] = R->L6 + (j) * (R->L7 - R->L6) / (R->N5 - 1);
                    P(num + j)[1] = y;
                    P(num + j)[2] = z;
                }
            } else if (R->type[k] == OR_RAY_D) { /* Luch.cpp:423-523 */
                double x0 = Q(0)[0], y0 = sqrt(KV(Q(0)[1]) + KV(Q(0)[2]));
                Q(1)[0] = x0;
                double y1 = sqrt(KV(Q(1)[1]) + KV(Q(1)[2]));
                Q(2)[0] = x0;
                double y2 = sqrt(KV(Q(2)[1]) + KV(Q(2)[2]));
                num += 1;
                for (int j = 0; j < M2 - M11; j++) {
                    double s = 1.0 * (j + 1) / (M2 - M11 + 1);
                    r = y0 + (y1 - y0) * or_sp1(s, 0.1, 0.3);
                    PUT_P(num + j, x0, r);
                }
                num += (M2 - M11) + 1;
                for (int j = 0; j < M3; j++) {
                    double s = 1.0 * (j + 1) / (M3 + 1);
                    r = y1 + (y2 - y1) * or_sp1(s, R->md1, R->md2);
                    PUT_P(num + j, x0, r);
                }
                num += M3 + 1;
                for (int j = 0; j < M4; j++) {
                    double s = (j + 1.0) / (M4 + 1.0);
                    r = y2 + (R5 - y2) * or_sp2(s, R->md3, 2.0);
                    PUT_P(num + j, x0, r);
                }
                num += M4;
                PUT_P(num, x0, R5);
            } else if (R->type[k] == OR_RAY_E) { /* Luch.cpp:525-602 */
                double y0 = sqrt(KV(Q(1)[1]) + KV(Q(1)[2])), y1 = sqrt(KV(Q(2)[1]) + KV(Q(2)[2]));
                double xe = Q(0)[0];
                P(num)[0] = xe;
                num += 1;
                for (int j = 0; j < M3; j++) {
                    double s = 1.0 * (j + 1) / (M3 + 1);
                    r = y0 + (y1 - y0) * or_sp1(s, R->ea1, R->ea2);
                    PUT_P(num + j, xe, r);
                }
                num += M3;
                P(num)[0] = xe;
                num += 1;
                for (int j = 0; j < M4; j++) {
                    double s = (j + 1.0) / (M4 + 1.0);
                    r = y1 + (R5 - y1) * or_sp2(s, R->ea3, 2.0);
                    PUT_P(num + j, xe, r);
                }
                num += M4;
                PUT_P(num, xe, R5);
            } else if (R->type[k] == OR_RAY_H) { /* Luch.cpp:604-661 */
                const double *z3 = Q(0), *z1 = Q(1), *z4 = Q(2), *z2 = Q(3);
                double t1 = (z1[0] - z3[0]) * R->dd7, tt1 = (or_ryz(z1) - or_ryz(z3)) * R->dd7;
                double t2 = (z4[0] - z2[0]) * R->dd8, tt2 = (or_ryz(z4) - or_ryz(z2)) * R->dd8;
                num += 1;
                for (int j = 0; j < 4; j++) {
                    double s = 1.0 * (j + 1) / (5);
                    or_hermite(z1[0], or_ryz(z1), z2[0], or_ryz(z2), t1, tt1, t2, tt2, s, &xx, &yy);
                    PUT_P(num + j, xx, yy);
                }
            } else if (R->type[k] == OR_RAY_G) { /* Luch.cpp:663-736 */
                const double *z1 = Q(1), *z2 = Q(2), *z3 = Q(0);
                double t1 = (z1[0] - z3[0]) * R->dd1, tt1 = (or_ryz(z1) - or_ryz(z3)) * R->dd1;
                double t2 = 0.0, tt2 = 1.0 * R->dd2;
                num = 1;
                for (int j = 0; j < M2 - M11 - R->MF + 1; j++) {
                    double s = 1.0 * (j + 1.0) / (M2 - M11 - R->MF + 1 + 1.0);
                    or_hermite(z1[0], or_ryz(z1), z2[0], or_ryz(z2), t1, tt1, t2, tt2, s, &xx, &yy);
                    PUT_P(num + j, xx, yy);
                }
            }
#undef P
#undef Q
#undef PUT_R
#undef PUT_P
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Setka::New_connect (Setka.cpp:2066-2249): 6 faces per hexahedron in a fixed local order,
 * duplicates dropped keeping the first created copy (its node order and cells[0] = creator),
 * the second creator appended as cells[1]; each cell's faces in ascending global id. */
static const int OR_LOCAL_FACES[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                         {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 7, 3}};

typedef struct { int32_t k[4]; int64_t id; } or_fkey;

static int fkey_cmp(const void *a, const void *b)
{
    const or_fkey *x = (const or_fkey *)a, *y = (const or_fkey *)b;
    for (int i = 0; i < 4; i++) {
        if (x->k[i] != y->k[i]) return x->k[i] < y->k[i] ? -1 : 1;
    }
    return x->id < y->id ? -1 : (x->id > y->id);
}

int64_t or_new_connect(int64_t n_cells, const int32_t *cell_nodes, int64_t n_nodes, int32_t *face_nodes,
                       int32_t *face_cells, int32_t *cell_faces)
{
    (void)n_nodes;
    int64_t nt = 6 * n_cells;
    or_fkey *keys = (or_fkey *)malloc(sizeof(or_fkey) * nt);
    for (int64_t c = 0; c < n_cells; c++) {
        for (int l = 0; l < 6; l++) {
            or_fkey *kk = &keys[6 * c + l];
            for (int i = 0; i < 4; i++) kk->k[i] = cell_nodes[8 * c + OR_LOCAL_FACES[l][i]];
            /* sort the 4 node ids (insertion sort) to compare node sets */
            for (int i = 1; i < 4; i++) {
                int32_t v = kk->k[i];
                int j = i - 1;
                while (j >= 0 && kk->k[j] > v) { kk->k[j + 1] = kk->k[j]; j--; }
                kk->k[j + 1] = v;
            }
            kk->id = 6 * c + l;
        }
    }
    qsort(keys, nt, sizeof(or_fkey), fkey_cmp);
    /* twin[i] = id of the other copy of the same face, or -1 */
    int64_t *twin = (int64_t *)malloc(sizeof(int64_t) * nt);
    for (int64_t i = 0; i < nt; i++) twin[i] = -1;
    for (int64_t i = 0; i + 1 < nt; i++) {
        if (memcmp(keys[i].k, keys[i + 1].k, sizeof(int32_t) * 4) == 0) {
            twin[keys[i].id] = keys[i + 1].id;
            twin[keys[i + 1].id] = keys[i].id;
        }
    }
    free(keys);
    int64_t *gid = (int64_t *)malloc(sizeof(int64_t) * nt);
    int64_t nf = 0;
    for (int64_t i = 0; i < nt; i++) {
        int64_t c = i / 6;
        int l = (int)(i % 6);
        if (twin[i] >= 0 && twin[i] < i) { gid[i] = gid[twin[i]]; continue; }
        gid[i] = nf;
        for (int j = 0; j < 4; j++) face_nodes[4 * nf + j] = cell_nodes[8 * c + OR_LOCAL_FACES[l][j]];
        face_cells[2 * nf] = (int32_t)c;
        face_cells[2 * nf + 1] = twin[i] >= 0 ? (int32_t)(twin[i] / 6) : -1;
        nf++;
    }
    for (int64_t c = 0; c < n_cells; c++) {
        int64_t v[6];
        for (int l = 0; l < 6; l++) v[l] = gid[6 * c + l];
        for (int i = 1; i < 6; i++) {
            int64_t x = v[i];
            int j = i - 1;
            while (j >= 0 && v[j] > x) { v[j + 1] = v[j]; j--; }
            v[j + 1] = x;
        }
        for (int l = 0; l < 6; l++) cell_faces[6 * c + l] = (int32_t)v[l];
    }
    free(twin);
    free(gid);
    return nf;
}

/* ------------------------------------------------------------------------------------------ */
/* Charge exchange.                                                                           */

/* MK_source.cpp:32-103 restated for one (U, c): midpoint rule in (R0 in [0,5c], alpha in
 * [0,pi]) of g sigma(g) R0^2 sin(alpha) exp(-(R0/c)^2), times 2 pi dal dR0 / (sqrt(pi) c)^3. */
double or_cx_rate_ref(double U, double c, double a2, int N1, int N2)
{
    double S = 0.0;
    double dal = OR_PI / N2;
    double dR0 = 5.0 * c / N1;
    for (int ii = 0; ii < N1; ++ii) {
        double R0 = (ii + 0.5) * 5.0 * c / N1;
        for (int jj = 0; jj < N2; ++jj) {
            double al = (jj + 0.5) * OR_PI / N2;
            double R = sqrt(KV(R0 * sin(al)) + KV(U - R0 * cos(al)));
            double sg = KV(1.0 - a2 * log(R));
            S += R * sg * KV(R0) * sin(al) * exp(-KV(R0 / c));
        }
    }
    S *= 2.0 * OR_PI * dal * dR0;
    S /= pow(sqrt(OR_PI) * c, 3);
    return S;
}

/* Lindsay & Stebbings (2005), J. Geophys. Res. 110, A12213, eq. (5):
 *   sigma(E) = (4.15 - 0.531 ln E)^2 (1 - exp(-67.3/E))^4.5 * 1e-16 cm^2, E in keV.
 * In code units: E = m_p g^2 / 2 with g in 10.3804 km/s; the result is multiplied by
 * n_char * L_char = 0.06 cm^-3 * 4.2113 AU so that rates come out per unit code time. */
double or_sigma_ls(double g)
{
    const double e_per_g2 = 0.5 * 1.67262192369e-24 * (1.03804e6 * 1.03804e6) / 1.602176634e-9; /* keV */
    const double nl = 0.06 * 4.2113 * 1.495978707e13;
    double E = e_per_g2 * g * g;
    if (E < 1e-12) E = 1e-12;
    double a = 4.15 - 0.531 * log(E);
    double b = 1.0 - exp(-67.3 / E);
    return a * a * pow(b, 4.5) * 1e-16 * nl;
}

/* Maxwellian charge-exchange source for the plasma.  With v_p = V_p + w_p, v_H = V_H + w_H
 * (Gaussian, thermal speeds c_p, c_H), the relative velocity g = v_p - v_H is Gaussian about
 * dV = V_p - V_H with thermal speed c = sqrt(c_p^2 + c_H^2), and E[w_p + w_H | g] =
 * kappa (g - dV) with kappa = (c_p^2 - c_H^2)/c^2.  Hence, with J_k = <|g| sigma |g|^k (g.e)^..>:
 *   rate  R   = n_p n_H J0,            J0 = < g sigma >
 *   S_m       = -n_p n_H J1 e,         J1 = < g sigma (g.e) >
 *   S_E (lab) = -(n_p n_H / 2) [ (V_p + V_H).e J1 + kappa (J2 - U J1) ],  J2 = < g^3 sigma >
 * where e = dV/|dV| and U = |dV|.  The three averages are the 2-D integrals of MK_source.cpp
 * (R0 in [0, 5c], alpha in [0, pi], midpoint rule), with the Lindsay-Stebbings sigma.
 * out = {R, S_mx, S_my, S_mz, S_E}. */
void or_cx_source(double n_p, double cp2, const double *vp, double n_h, double ch2, const double *vh,
                  int N1, int N2, double *out)
{
    double dv[3] = {vp[0] - vh[0], vp[1] - vh[1], vp[2] - vh[2]};
    double U = sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]);
    double c2 = cp2 + ch2;
    double c = sqrt(c2);
    double e[3] = {1.0, 0.0, 0.0};
    if (U > 0.0) { e[0] = dv[0] / U; e[1] = dv[1] / U; e[2] = dv[2] / U; }
    double J0 = 0.0, J1 = 0.0, J2 = 0.0;
    double dal = OR_PI / N2;
    double dR0 = 5.0 * c / N1;
    for (int ii = 0; ii < N1; ++ii) {
        double R0 = (ii + 0.5) * dR0;
        double wr = R0 * R0 * exp(-(R0 * R0) / c2);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int jj = 0; jj < N2; ++jj) {
            double al = (jj + 0.5) * dal;
            double sa = sin(al), ca = cos(al);
            double gpar = U - R0 * ca;
            double gper = R0 * sa;
            double g2 = gpar * gpar + gper * gper;
            double g = sqrt(g2);
            double gs = g * or_sigma_ls(g) * sa;
            s0 += gs;
            s1 += gs * gpar;
            s2 += gs * g2;
        }
        J0 += wr * s0;
        J1 += wr * s1;
        J2 += wr * s2;
    }
    double norm = 2.0 * OR_PI * dal * dR0 / (OR_PI * sqrt(OR_PI) * c2 * c);
    J0 *= norm; J1 *= norm; J2 *= norm;
    double nn = n_p * n_h;
    double kappa = (cp2 - ch2) / c2;
    double sv = (vp[0] + vh[0]) * e[0] + (vp[1] + vh[1]) * e[1] + (vp[2] + vh[2]) * e[2];
    out[0] = nn * J0;
    out[1] = -nn * J1 * e[0];
    out[2] = -nn * J1 * e[1];
    out[3] = -nn * J1 * e[2];
    out[4] = -0.5 * nn * (sv * J1 + kappa * (J2 - U * J1));
}

/* ------------------------------------------------------------------------------------------ */
/* FP64 MLP 8-64-64-64-5 with SiLU (the surrogate's CPU reference for gradient checks).        */

#define MLP_IN 8
#define MLP_H 64
#define MLP_OUT 5

int64_t or_mlp_n_params(void)
{
    return (int64_t)MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT;
}

static double silu(double x) { return x / (1.0 + exp(-x)); }
static double dsilu(double x)
{
    double s = 1.0 / (1.0 + exp(-x));
    return s * (1.0 + x * (1.0 - s));
}

typedef struct { const double *W[4], *b[4]; int nin[4], nout[4]; } mlp_view;

static mlp_view mlp_layout(const double *P)
{
    mlp_view v;
    int dims[5] = {MLP_IN, MLP_H, MLP_H, MLP_H, MLP_OUT};
    const double *p = P;
    for (int l = 0; l < 4; l++) {
        v.nin[l] = dims[l];
        v.nout[l] = dims[l + 1];
        v.W[l] = p; p += dims[l] * dims[l + 1];
        v.b[l] = p; p += dims[l + 1];
    }
    return v;
}

void or_mlp_forward(const double *params, int64_t n, const double *x, double *y)
{
    mlp_view v = mlp_layout(params);
    for (int64_t s = 0; s < n; s++) {
        double a[MLP_H], z[MLP_H];
        for (int i = 0; i < MLP_IN; i++) a[i] = x[MLP_IN * s + i];
        for (int l = 0; l < 4; l++) {
            for (int o = 0; o < v.nout[l]; o++) {
                double acc = v.b[l][o];
                for (int i = 0; i < v.nin[l]; i++) acc += v.W[l][o * v.nin[l] + i] * a[i];
                z[o] = acc;
            }
            for (int o = 0; o < v.nout[l]; o++) a[o] = (l < 3) ? silu(z[o]) : z[o];
        }
        for (int o = 0; o < MLP_OUT; o++) y[MLP_OUT * s + o] = a[o];
    }
}

/* Round to the nearest BF16 (ties to even) of the FP32 rounding of x: the conversion the device
 * applies to MFMA operands. */
static double bf16_round(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

/* The network as the BF16 matrix-core inference evaluates it: weights and every layer's input
 * (the normalised features, then each SiLU activation) rounded to BF16, biases rounded to FP32,
 * products and sums exact to FP64 (the device accumulates in FP32). */
void or_mlp_forward_bf16(const double *params, int64_t n, const double *x, double *y)
{
    mlp_view v = mlp_layout(params);
    for (int64_t s = 0; s < n; s++) {
        double a[MLP_H], z[MLP_H];
        for (int i = 0; i < MLP_IN; i++) a[i] = bf16_round(x[MLP_IN * s + i]);
        for (int l = 0; l < 4; l++) {
            for (int o = 0; o < v.nout[l]; o++) {
                double acc = (double)(float)v.b[l][o];
                for (int i = 0; i < v.nin[l]; i++) acc += bf16_round(v.W[l][o * v.nin[l] + i]) * a[i];
                z[o] = acc;
            }
            for (int o = 0; o < v.nout[l]; o++) a[o] = (l < 3) ? bf16_round(silu((double)(float)z[o])) : z[o];
        }
        for (int o = 0; o < MLP_OUT; o++) y[MLP_OUT * s + o] = a[o];
    }
}

double or_mlp_loss_grad(const double *params, int64_t n, const double *x, const double *yt, double *grad)
{
    mlp_view v = mlp_layout(params);
    int64_t np = or_mlp_n_params();
    for (int64_t i = 0; i < np; i++) grad[i] = 0.0;
    mlp_view gv = mlp_layout(grad);
    double loss = 0.0;
    for (int64_t s = 0; s < n; s++) {
        double A[5][MLP_H], Z[4][MLP_H];
        for (int i = 0; i < MLP_IN; i++) A[0][i] = x[MLP_IN * s + i];
        for (int l = 0; l < 4; l++) {
            for (int o = 0; o < v.nout[l]; o++) {
                double acc = v.b[l][o];
                for (int i = 0; i < v.nin[l]; i++) acc += v.W[l][o * v.nin[l] + i] * A[l][i];
                Z[l][o] = acc;
                A[l + 1][o] = (l < 3) ? silu(acc) : acc;
            }
        }
        double d[MLP_H];
        for (int o = 0; o < MLP_OUT; o++) {
            double e = A[4][o] - yt[MLP_OUT * s + o];
            loss += e * e;
            d[o] = 2.0 * e / (double)(n * MLP_OUT);
        }
        for (int l = 3; l >= 0; l--) {
            double *gW = (double *)gv.W[l], *gb = (double *)gv.b[l];
            double dn[MLP_H];
            for (int i = 0; i < v.nin[l]; i++) dn[i] = 0.0;
            for (int o = 0; o < v.nout[l]; o++) {
                gb[o] += d[o];
                for (int i = 0; i < v.nin[l]; i++) {
                    gW[o * v.nin[l] + i] += d[o] * A[l][i];
                    dn[i] += v.W[l][o * v.nin[l] + i] * d[o];
                }
            }
            if (l > 0)
                for (int i = 0; i < v.nin[l]; i++) d[i] = dn[i] * dsilu(Z[l - 1][i]);
        }
    }
    return loss / (double)(n * MLP_OUT);
}
