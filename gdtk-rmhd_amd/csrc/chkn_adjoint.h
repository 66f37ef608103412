// chkn_adjoint.h -- hand-written reverse-mode derivatives (vector-Jacobian products) of the
// pointwise physics in chkn_physics.h: MUSCL/van Albada reconstruction, the ideal-gas
// equation of state, the AUSMDV flux (face.cu:163-265, :347-371) and the conserved ->
// primitive decode (flow.cu:245-268).
//
// Each *_vjp function recomputes the forward intermediates it needs and accumulates
// adjoints into its outputs.  At the non-differentiable points of the scheme (|x| at 0,
// min/max ties, the sub/supersonic switch at |M| = 1, the K_SWITCH clip) the derivative of
// the branch the forward evaluation takes is used.  Host and device code.
#pragma once

#include "chkn_physics.h"

namespace chkn {

CHKN_HD double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// Partial derivatives of van_albada(a, b) = (ab + |ab| + eps) / (a^2 + b^2 + eps).
CHKN_HD void van_albada_grad(double a, double b, double& da, double& db)
{
    const double eps = 1.0e-12;
    const double ab = a * b;
    const double N = ab + fabs(ab) + eps;
    const double D = a * a + b * b + eps;
    const double k = 1.0 + sgn(ab);
    const double D2 = D * D;
    da = (b * k * D - N * 2.0 * a) / D2;
    db = (a * k * D - N * 2.0 * b) / D2;
}

// Adjoint of muscl(qL1, qL0, qR0, qR1) -> (qL, qR): accumulates into g[0..3].
CHKN_HD void muscl_vjp(double qL1, double qL0, double qR0, double qR1, double bqL, double bqR,
                       double g[4])
{
    const double c = 0.125;
    const double dm = qL0 - qL1, d = qR0 - qL0, dp = qR1 - qR0;
    const double sL = van_albada(dm, d), sR = van_albada(d, dp);
    double bdm = 0.0, bd = 0.0, bdp = 0.0;
    // qL = qL0 + sL*c*(3d + dm)
    g[1] += bqL;
    const double bsL = bqL * c * (3.0 * d + dm);
    bd += bqL * sL * c * 3.0;
    bdm += bqL * sL * c;
    // qR = qR0 - sR*c*(dp + 3d)
    g[2] += bqR;
    const double bsR = -bqR * c * (dp + 3.0 * d);
    bdp -= bqR * sR * c;
    bd -= bqR * sR * c * 3.0;
    double ga, gb;
    van_albada_grad(dm, d, ga, gb);
    bdm += bsL * ga;
    bd += bsL * gb;
    van_albada_grad(d, dp, ga, gb);
    bd += bsR * ga;
    bdp += bsR * gb;
    g[1] += bdm; g[0] -= bdm;  // dm = qL0 - qL1
    g[2] += bd;  g[1] -= bd;   // d = qR0 - qL0
    g[3] += bdp; g[2] -= bdp;  // dp = qR1 - qR0
}

// Adjoint of the AUSMDV flux with respect to its face states.  Inputs as ausmdv(); outputs
// accumulate adjoints of (rho, e, p, a, YB, vx, vy, vz) on each side: bL[8], bR[8].
CHKN_HD void ausmdv_vjp(double rhoL, double eL, double pL, double aL, double YBL, double vxL,
                        double vyL, double vzL, double rhoR, double eR, double pR, double aR,
                        double YBR, double vxR, double vyR, double vzR, const Frame& f,
                        const double Fb[6], double bL[8], double bR[8])
{
    // ---- forward (same expressions as ausmdv)
    double uL, vL, wL, uR, vR, wR;
    to_local(f, vxL, vyL, vzL, uL, vL, wL);
    to_local(f, vxR, vyR, vzR, uR, vR, wR);
    const double prL = pL / rhoL, prR = pR / rhoR;
    const double keL = 0.5 * (uL * uL + vL * vL + wL * wL);
    const double keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    const double hL = eL + prL + keL, hR = eR + prR + keR;
    const double S = prL + prR;
    const double wdL = 2.0 * prL / S, wdR = 2.0 * prR / S;
    const double cmax = fmax(aL, aR);
    const double machL = uL / cmax, machR = uR / cmax;
    const double upL = 0.5 * (uL + fabs(uL));
    const bool subL = fabs(machL) <= 1.0;
    const double usL = subL ? wdL * ((uL + cmax) * (uL + cmax) / (4.0 * cmax) - upL) + upL : upL;
    const double unR = 0.5 * (uR - fabs(uR));
    const bool subR = fabs(machR) <= 1.0;
    const double usR = subR ? wdR * (-(uR - cmax) * (uR - cmax) / (4.0 * cmax) - unR) + unR : unR;
    const double mfL = usL * rhoL, mfR = usR * rhoR;
    const double mdot = mfL + mfR;
    const double pdiff = pL - pR;
    const double pmin = fmin(pL, pR);
    const double pjump = 10.0 * fabs(pdiff) / pmin;
    const double blend = 0.5 * fmin(1.0, pjump);
    const double momV = mfL * uL + mfR * uR;
    const double momD = 0.5 * (mdot * (uL + uR) - fabs(mdot) * (uR - uL));
    const bool upw = (mdot >= 0.0);
    const double vt = upw ? vL : vR, wt = upw ? wL : wR;
    const double hu = upw ? hL : hR, Yu = upw ? YBL : YBR;

    // ---- reverse
    double X1, X2, X3;  // adjoint of the local-frame flux (transpose of to_global)
    to_local(f, Fb[1], Fb[2], Fb[3], X1, X2, X3);
    const double bmom = X1, bpface = X1;
    double bmdot = Fb[0] + X2 * vt + X3 * wt + Fb[4] * hu + Fb[5] * Yu;
    const double bvt = X2 * mdot, bwt = X3 * mdot, bhu = Fb[4] * mdot, bYu = Fb[5] * mdot;
    double buL = 0.0, bvL = 0.0, bwL = 0.0, buR = 0.0, bvR = 0.0, bwR = 0.0;
    double bhL = 0.0, bhR = 0.0;
    if (upw) { bvL += bvt; bwL += bwt; bhL += bhu; bL[4] += bYu; }
    else    { bvR += bvt; bwR += bwt; bhR += bhu; bR[4] += bYu; }
    // mom = (0.5 + blend) momV + (0.5 - blend) momD
    const double bblend = bmom * (momV - momD);
    const double bmomV = bmom * (0.5 + blend), bmomD = bmom * (0.5 - blend);
    // momD = 0.5 (mdot (uL + uR) - |mdot| (uR - uL))
    bmdot += bmomD * 0.5 * ((uL + uR) - sgn(mdot) * (uR - uL));
    buL += bmomD * 0.5 * (mdot + fabs(mdot));
    buR += bmomD * 0.5 * (mdot - fabs(mdot));
    // momV = mfL uL + mfR uR
    double bmfL = bmomV * uL, bmfR = bmomV * uR;
    buL += bmomV * mfL;
    buR += bmomV * mfR;
    // blend = 0.5 min(1, pjump); pjump = 10 |pdiff| / pmin
    if (pjump < 1.0) {
        const double bpjump = 0.5 * bblend;
        const double bpdiff = bpjump * 10.0 * sgn(pdiff) / pmin;
        const double bpmin = -bpjump * pjump / pmin;
        bL[2] += bpdiff;
        bR[2] -= bpdiff;
        if (pL <= pR) bL[2] += bpmin;
        else bR[2] += bpmin;
    }
    // pface = psL + psR; mdot = mfL + mfR
    const double bpsL = bpface, bpsR = bpface;
    bmfL += bmdot;
    bmfR += bmdot;
    const double busL = bmfL * rhoL, busR = bmfR * rhoR;
    bL[0] += bmfL * usL;
    bR[0] += bmfR * usR;
    double bcmax = 0.0, bwdL = 0.0, bwdR = 0.0, bmachL = 0.0, bmachR = 0.0;
    double bupL = 0.0, bunR = 0.0;
    if (subL) {
        const double q1 = (machL + 1.0) * (machL + 1.0);
        bL[2] += bpsL * q1 * (2.0 - machL) * 0.25;
        bmachL += bpsL * pL * 0.25 * (2.0 * (machL + 1.0) * (2.0 - machL) - q1);
        const double Q = (uL + cmax) * (uL + cmax) / (4.0 * cmax);
        bwdL += busL * (Q - upL);
        const double bQ = busL * wdL;
        bupL += busL * (1.0 - wdL);
        buL += bQ * 2.0 * (uL + cmax) / (4.0 * cmax);
        bcmax += bQ * (2.0 * (uL + cmax) / (4.0 * cmax) - (uL + cmax) * (uL + cmax) / (4.0 * cmax * cmax));
    } else {
        bL[2] += bpsL * upL / uL;
        bupL += bpsL * pL / uL;
        buL -= bpsL * pL * upL / (uL * uL);
        bupL += busL;
    }
    buL += bupL * 0.5 * (1.0 + sgn(uL));
    if (subR) {
        const double q1 = (machR - 1.0) * (machR - 1.0);
        bR[2] += bpsR * q1 * (2.0 + machR) * 0.25;
        bmachR += bpsR * pR * 0.25 * (2.0 * (machR - 1.0) * (2.0 + machR) + q1);
        const double Q = -(uR - cmax) * (uR - cmax) / (4.0 * cmax);
        bwdR += busR * (Q - unR);
        const double bQ = busR * wdR;
        bunR += busR * (1.0 - wdR);
        buR -= bQ * 2.0 * (uR - cmax) / (4.0 * cmax);
        bcmax += bQ * (2.0 * (uR - cmax) / (4.0 * cmax) + (uR - cmax) * (uR - cmax) / (4.0 * cmax * cmax));
    } else {
        bR[2] += bpsR * unR / uR;
        bunR += bpsR * pR / uR;
        buR -= bpsR * pR * unR / (uR * uR);
        bunR += busR;
    }
    buR += bunR * 0.5 * (1.0 - sgn(uR));
    // machL = uL / cmax, machR = uR / cmax
    buL += bmachL / cmax;
    buR += bmachR / cmax;
    bcmax -= (bmachL * machL + bmachR * machR) / cmax;
    if (aL >= aR) bL[3] += bcmax;
    else bR[3] += bcmax;
    // wd = 2 pr / S, S = prL + prR
    double bprL = bwdL * 2.0 / S, bprR = bwdR * 2.0 / S;
    const double bS = -(bwdL * wdL + bwdR * wdR) / S;
    bprL += bS;
    bprR += bS;
    // H = e + p/rho + ke
    bL[1] += bhL; bprL += bhL;
    bR[1] += bhR; bprR += bhR;
    buL += bhL * uL; bvL += bhL * vL; bwL += bhL * wL;
    buR += bhR * uR; bvR += bhR * vR; bwR += bhR * wR;
    // p/rho
    bL[2] += bprL / rhoL;
    bL[0] -= bprL * prL / rhoL;
    bR[2] += bprR / rhoR;
    bR[0] -= bprR * prR / rhoR;
    // velocities back to the global frame (transpose of to_local)
    double gx, gy, gz;
    to_global(f, buL, bvL, bwL, gx, gy, gz);
    bL[5] += gx; bL[6] += gy; bL[7] += gz;
    to_global(f, buR, bvR, bwR, gx, gy, gz);
    bR[5] += gx; bR[6] += gy; bR[7] += gz;
}

// Adjoint of the SBP-ASF flux (sbp_asf, face.cu:268-342) with each stencil state's p from the
// equation of state of its (rho, e, YB): accumulates the adjoints of the four stencil states
// (rho, e, YB, vx, vy, vz) into b[0..3][6].  The flux is a sum of products of the per-cell
// factors fac_j = (rho, u rho, v rho, w rho, e rho, u^2 rho, v^2 rho, w^2 rho, p, p) and
// adv_j = (u, ..., u, 1) in the face frame, so its adjoint runs back through the central (cen)
// and product-rule (prod) combinations to the factors, then to each state.
CHKN_HD void sbp_asf_vjp(const Prim s[4], const Frame& f, const double Fb[6], double b[4][6])
{
    constexpr double c = 1.0 / 12.0;
    double u[4], v[4], w[4], T[4], p[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        to_local(f, s[i].vx, s[i].vy, s[i].vz, u[i], v[i], w[i]);
        T[i] = eos_T(s[i].e, s[i].YB);
        p[i] = eos_p(s[i].rho, T[i]);
    }
    // Factors of product j at cell i (forward values).
    auto fac = [&](int j, int i) -> double {
        const double rho = s[i].rho;
        switch (j) {
            case 0: return rho;
            case 1: return u[i] * rho;
            case 2: return v[i] * rho;
            case 3: return w[i] * rho;
            case 4: return s[i].e * rho;
            case 5: return u[i] * u[i] * rho;
            case 6: return v[i] * v[i] * rho;
            case 7: return w[i] * w[i] * rho;
            default: return p[i];
        }
    };
    auto adv = [&](int j, int i) -> double { return j == 9 ? 1.0 : u[i]; };
    const double mass = c * (-fac(0, 0) * adv(0, 0) + 7.0 * fac(0, 1) * adv(0, 1) + 7.0 * fac(0, 2) * adv(0, 2) -
                             fac(0, 3) * adv(0, 3));
    const bool upw = mass >= 0.0;
    // Adjoints of the blended sums: mass = fc0; (mx, my, mz) rotated back by to_global (its
    // transpose is to_local); energy and the pressure terms with the alphas of sbp_asf.
    double bmx, bmy, bmz;
    to_local(f, Fb[1], Fb[2], Fb[3], bmx, bmy, bmz);
    double bcen[10], bprod[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) bcen[j] = bprod[j] = 0.0;
    bcen[0] = Fb[0] + Fb[5] * (upw ? s[1].YB : s[2].YB);
    bcen[1] = 0.5 * bmx; bprod[1] = 0.5 * bmx; bprod[9] = bmx;
    bcen[2] = 0.5 * bmy; bprod[2] = 0.5 * bmy;
    bcen[3] = 0.5 * bmz; bprod[3] = 0.5 * bmz;
    bcen[4] = 0.5 * Fb[4]; bprod[4] = 0.5 * Fb[4];
    bprod[5] = 0.5 * Fb[4]; bprod[6] = 0.5 * Fb[4]; bprod[7] = 0.5 * Fb[4];
    bprod[8] = Fb[4];
    // Adjoints of the factors: cen_j = c(-f0 a0 + 7 f1 a1 + 7 f2 a2 - f3 a3),
    // prod_j = c(-f0 a2 - f2 a0 + 8 f1 a2 + 8 f2 a1 - f1 a3 - f3 a1).
    double bfac[10][4], badv[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        const double gc = c * bcen[j], ge = c * bprod[j];
        const double f0 = fac(j, 0), f1 = fac(j, 1), f2 = fac(j, 2), f3 = fac(j, 3);
        const double a0 = adv(j, 0), a1 = adv(j, 1), a2 = adv(j, 2), a3 = adv(j, 3);
        bfac[j][0] = -gc * a0 - ge * a2;
        bfac[j][1] = 7.0 * gc * a1 + ge * (8.0 * a2 - a3);
        bfac[j][2] = 7.0 * gc * a2 + ge * (8.0 * a1 - a0);
        bfac[j][3] = -gc * a3 - ge * a1;
        if (j < 9) {  // adv_j = u of each cell (adv_9 = 1 is constant)
            badv[0] += -gc * f0 - ge * f2;
            badv[1] += 7.0 * gc * f1 + ge * (8.0 * f2 - f3);
            badv[2] += 7.0 * gc * f2 + ge * (8.0 * f1 - f0);
            badv[3] += -gc * f3 - ge * f1;
        }
    }
    // Factors -> (rho, e, p, u, v, w) of each cell -> its primitive state.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double rho = s[i].rho;
        double brho = bfac[0][i] + bfac[1][i] * u[i] + bfac[2][i] * v[i] + bfac[3][i] * w[i] + bfac[4][i] * s[i].e +
                      bfac[5][i] * u[i] * u[i] + bfac[6][i] * v[i] * v[i] + bfac[7][i] * w[i] * w[i];
        const double bu = bfac[1][i] * rho + 2.0 * bfac[5][i] * u[i] * rho + badv[i];
        const double bvv = bfac[2][i] * rho + 2.0 * bfac[6][i] * v[i] * rho;
        const double bww = bfac[3][i] * rho + 2.0 * bfac[7][i] * w[i] * rho;
        double be = bfac[4][i] * rho;
        const double bp = bfac[8][i] + bfac[9][i];
        // p = rho R T, T = (e + YB Q) / Cv
        brho += bp * kR * T[i];
        const double bT = bp * rho * kR;
        be += bT / kCv;
        double bx, by, bz;
        to_global(f, bu, bvv, bww, bx, by, bz);
        b[i][0] += brho;
        b[i][1] += be;
        b[i][2] += bT * kQ / kCv;
        b[i][3] += bx;
        b[i][4] += by;
        b[i][5] += bz;
    }
    b[upw ? 1 : 2][2] += Fb[5] * mass;
}

// Adjoint of face_flux<XORDER>: accumulates the adjoints of the four stencil states
// (rho, e, YB, vx, vy, vz each) into b[0..3][6] (L1, L0, R0, R1).
template <int XORDER>
CHKN_HD void face_flux_vjp(const Prim& L1, const Prim& L0, const Prim& R0, const Prim& R1,
                           const Frame& f, const double Fb[6], double b[4][6])
{
    double rhoL, rhoR, eL, eR, YBL, YBR, vxL, vxR, vyL, vyR, vzL, vzR;
    if (XORDER > 1) {
        muscl(L1.rho, L0.rho, R0.rho, R1.rho, rhoL, rhoR);
        muscl(L1.e, L0.e, R0.e, R1.e, eL, eR);
        muscl(L1.YB, L0.YB, R0.YB, R1.YB, YBL, YBR);
        muscl(L1.vx, L0.vx, R0.vx, R1.vx, vxL, vxR);
        muscl(L1.vy, L0.vy, R0.vy, R1.vy, vyL, vyR);
        muscl(L1.vz, L0.vz, R0.vz, R1.vz, vzL, vzR);
    } else {
        rhoL = L0.rho; eL = L0.e; YBL = L0.YB; vxL = L0.vx; vyL = L0.vy; vzL = L0.vz;
        rhoR = R0.rho; eR = R0.e; YBR = R0.YB; vxR = R0.vx; vyR = R0.vy; vzR = R0.vz;
    }
    const double TL = eos_T(eL, YBL), TR = eos_T(eR, YBR);
    const double pL = eos_p(rhoL, TL), pR = eos_p(rhoR, TR);
    const double aL = eos_a(TL), aR = eos_a(TR);
    double bL[8] = {0, 0, 0, 0, 0, 0, 0, 0}, bR[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    ausmdv_vjp(rhoL, eL, pL, aL, YBL, vxL, vyL, vzL, rhoR, eR, pR, aR, YBR, vxR, vyR, vzR, f, Fb,
               bL, bR);
    // Equation of state: T = (e + YB Q)/Cv, p = rho R T, a = sqrt(gamma R T).
    double side[2][6];  // adjoints of the reconstructed (rho, e, YB, vx, vy, vz)
    const double* bs[2] = {bL, bR};
    const double rho[2] = {rhoL, rhoR}, T[2] = {TL, TR}, a[2] = {aL, aR};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const double* bb = bs[q];
        const double bT = bb[2] * rho[q] * kR + bb[3] * kGamma * kR / (2.0 * a[q]);
        side[q][0] = bb[0] + bb[2] * kR * T[q];
        side[q][1] = bb[1] + bT / kCv;
        side[q][2] = bb[4] + bT * kQ / kCv;
        side[q][3] = bb[5];
        side[q][4] = bb[6];
        side[q][5] = bb[7];
    }
    if (XORDER > 1) {
        double g[4];
#define CHKN_MUSCL_VJP(Q, FIELD)                                                      \
        g[0] = g[1] = g[2] = g[3] = 0.0;                                              \
        muscl_vjp(L1.FIELD, L0.FIELD, R0.FIELD, R1.FIELD, side[0][Q], side[1][Q], g); \
        b[0][Q] += g[0]; b[1][Q] += g[1]; b[2][Q] += g[2]; b[3][Q] += g[3];
        CHKN_MUSCL_VJP(0, rho)
        CHKN_MUSCL_VJP(1, e)
        CHKN_MUSCL_VJP(2, YB)
        CHKN_MUSCL_VJP(3, vx)
        CHKN_MUSCL_VJP(4, vy)
        CHKN_MUSCL_VJP(5, vz)
#undef CHKN_MUSCL_VJP
    } else {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            b[1][q] += side[0][q];
            b[2][q] += side[1][q];
        }
    }
}

// Adjoint of decode(U) -> (rho, e, YB, vx, vy, vz) at the primitive state w:
// Ub[q] = sum_p Wb[p] dW_p/dU_q.
CHKN_HD void decode_vjp(const Prim& w, const double Wb[6], double Ub[6])
{
#pragma clang fp contract(off)
    const double rho = w.rho;
    const double dinv = 1.0 / rho;
    const double bx = Wb[3] - Wb[1] * w.vx;  // velocity adjoints including e = U4/rho - ke
    const double by = Wb[4] - Wb[1] * w.vy;
    const double bz = Wb[5] - Wb[1] * w.vz;
    const double ke = 0.5 * (w.vx * w.vx + w.vy * w.vy + w.vz * w.vz);
    Ub[1] = bx * dinv;
    Ub[2] = by * dinv;
    Ub[3] = bz * dinv;
    Ub[4] = Wb[1] * dinv;
    Ub[5] = Wb[2] * dinv;
    Ub[0] = Wb[0] - (bx * w.vx + by * w.vy + bz * w.vz + Wb[1] * (w.e + ke) + Wb[2] * w.YB) * dinv;
}

// Adjoint of encode(w) -> U (flow.cu:227-243) at the primitive state w: Wb[p] = sum_q Ub[q] dU_q/dW_p.
CHKN_HD void encode_vjp(const Prim& w, const double Ub[6], double Wb[6])
{
    const double rho = w.rho;
    const double ke = 0.5 * (w.vx * w.vx + w.vy * w.vy + w.vz * w.vz);
    Wb[0] = Ub[0] + Ub[1] * w.vx + Ub[2] * w.vy + Ub[3] * w.vz + Ub[4] * (w.e + ke) + Ub[5] * w.YB;
    Wb[1] = Ub[4] * rho;
    Wb[2] = Ub[5] * rho;
    Wb[3] = (Ub[1] + Ub[4] * w.vx) * rho;
    Wb[4] = (Ub[2] + Ub[4] * w.vy) * rho;
    Wb[5] = (Ub[3] + Ub[4] * w.vz) * rho;
}

// Adjoint of the A->B reaction step U3 -> U = encode(chem(decode(U3))) (block.cu:1174-1184,
// gas.cu:102-111: where T > Ti, YB' = 1 - (1 - YB) f with f = exp(-alpha dt); rho, e and the
// velocity are unchanged), from the state after the reaction wr and the adjoint Ub of U; f < 0
// for an instance without the reaction.  The reaction took place exactly where the state after
// it is above the ignition temperature (it only raises YB, hence T = (e + YB q)/Cv, and leaves
// states at or below Ti unchanged), and there YB before it is 1 - (1 - YB')/f.  Returns the
// adjoint of U3 in Ub.
CHKN_HD void chem_vjp(const Prim& wr, double f, double Ub[6])
{
    double Wb[6];
    encode_vjp(wr, Ub, Wb);
    Prim w3 = wr;
    if (f >= 0.0 && eos_T(wr.e, wr.YB) > kTi) {
        w3.YB = 1.0 - (1.0 - wr.YB) / f;
        Wb[2] *= f;
    }
    decode_vjp(w3, Wb, Ub);
}

// Adjoint of the wall ghost state reflect(s, f): the map is v -> F^T diag(-1,1,1) F v,
// which is symmetric, so its transpose applied to the ghost's velocity adjoint is the same
// reflection; rho, e and YB pass through.
CHKN_HD void reflect_vjp(const double Gb[6], const Frame& f, double Sb[6])
{
#pragma clang fp contract(off)
    Prim g;
    g.rho = Gb[0]; g.e = Gb[1]; g.YB = Gb[2]; g.vx = Gb[3]; g.vy = Gb[4]; g.vz = Gb[5];
    const Prim r = reflect(g, f);
    Sb[0] += r.rho; Sb[1] += r.e; Sb[2] += r.YB; Sb[3] += r.vx; Sb[4] += r.vy; Sb[5] += r.vz;
}

}  // namespace chkn
