// chkn_physics.h -- pointwise physics of the chicken step, for gfx950 device code and the
// host geometry pre-pass, in the build's `real` (double, or float for the FP32 flavour).
//
// Every expression keeps the operand order and association of the reference so that, with
// floating-point contraction disabled (-ffp-contract=off), the GPU reproduces the
// reference CPU build bit for bit.  References are to src/chicken/ in the reference tree.
// Literals are written real(...) where the reference uses its `number` constants; where
// the reference's expression mixes in a double literal, that part is evaluated in double
// as the reference's FP32 build evaluates it.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "chkn_real.h"

#define CHKN_HD __host__ __device__ __forceinline__

namespace chkn {

// Gas model, chosen at build time as the reference chooses it (gas.cu:20-48): ideal air, or
// with CHKN_AB_REACTING_GAS the reacting A->B gas (libchkn_ab.so).  The constants are the
// double literals of gas.cu:34-40 rounded to real; Cv divides in double (g-1.0).
#ifdef CHKN_AB_REACTING_GAS
constexpr real kGamma = real(6.0 / 5.0);
constexpr real kR = real(287.0);
constexpr real kQ = real(300000.0);
constexpr real kAlpha = real(1000.0);  // reaction rate, 1/s
constexpr real kTi = real(362.58);     // ignition temperature, K
#else
constexpr real kGamma = real(1.4);
constexpr real kR = real(287.1);
constexpr real kQ = real(0.0);
constexpr real kAlpha = real(0.0);
constexpr real kTi = real(0.0);
#endif
constexpr real kCv = real(double(kR) / (double(kGamma) - 1.0));
constexpr real kOneSixth = real(1) / real(6);   // number.cu:34 (1.0f/6.0f in the FP32 build)
constexpr real kFourSixth = real(4) / real(6);  // number.cu:35

// Primitive state carried between stages: the six values the reconstruction reads
// (face.cu:359-368).  p, T and a follow from rho and e through the equation of state.
struct Prim {
    real rho, e, YB, vx, vy, vz;
};

struct Frame {
    real nx, ny, nz, t1x, t1y, t1z, t2x, t2y, t2z;
};

// gas.cu:128-135 (update_from_rhoe): T, then p = rho*R*T, a = sqrt(g*R*T).
CHKN_HD real eos_T(real e, real YB) { return (e + YB * kQ) / kCv; }
CHKN_HD real eos_p(real rho, real T) { return rho * kR * T; }
CHKN_HD real eos_a(real T) { return sqrt(kGamma * kR * T); }

// vector3.cu:97-109 (global -> local) and :111-119 (local -> global).
CHKN_HD void to_local(const Frame& f, real x, real y, real z, real& u, real& v,
                      real& w)
{
    u = x * f.nx + y * f.ny + z * f.nz;
    v = x * f.t1x + y * f.t1y + z * f.t1z;
    w = x * f.t2x + y * f.t2y + z * f.t2z;
}
CHKN_HD void to_global(const Frame& f, real u, real v, real w, real& x, real& y,
                       real& z)
{
    x = u * f.nx + v * f.t1x + w * f.t2x;
    y = u * f.ny + v * f.t1y + w * f.t2y;
    z = u * f.nz + v * f.t1z + w * f.t2z;
}

// Slip-wall ghost state: interior state with the face-normal velocity negated in the face
// frame (bcs.cu:26-37).
CHKN_HD Prim reflect(const Prim& s, const Frame& f)
{
    real u, v, w;
    to_local(f, s.vx, s.vy, s.vz, u, v, w);
    u = -(u);
    Prim r = s;
    to_global(f, u, v, w, r.vx, r.vy, r.vz);
    return r;
}

// face.cu:72-79
CHKN_HD real van_albada(real a, real b)
{
    const real eps = real(1.0e-12);
    return (a * b + fabs(a * b) + eps) / (a * a + b * b + eps);
}

// face.cu:81-96
CHKN_HD void muscl(real qL1, real qL0, real qR0, real qR1, real& qL, real& qR)
{
    real dm = (qL0 - qL1);
    real d = (qR0 - qL0);
    real dp = (qR1 - qR0);
    real sL = van_albada(dm, d);
    real sR = van_albada(d, dp);
    qL = qL0 + sL * real(0.125) * (real(3.0) * d + dm);
    qR = qR0 - sR * real(0.125) * (dp + real(3.0) * d);
}

// AUSMDV flux of Wada & Liou (1994) on reconstructed face states, face.cu:163-265.
// Inputs are the gas states (rho, e, p, a, YB) and global-frame velocities on each side.
CHKN_HD void ausmdv(real rhoL, real eL, real pL, real aL, real YBL, real vxL,
                    real vyL, real vzL, real rhoR, real eR, real pR, real aR,
                    real YBR, real vxR, real vyR, real vzR, const Frame& f,
                    real F[6])
{
    real uL, vL, wL, uR, vR, wR;
    to_local(f, vxL, vyL, vzL, uL, vL, wL);
    to_local(f, vxR, vyR, vzR, uR, vR, wR);
    real prL = pL / rhoL;
    real keL = real(0.5) * (uL * uL + vL * vL + wL * wL);
    real hL = eL + prL + keL;
    real prR = pR / rhoR;
    real keR = real(0.5) * (uR * uR + vR * vR + wR * wR);
    real hR = eR + pR / rhoR + keR;
    real wdL = real(2.0) * prL / (prL + prR);
    real wdR = real(2.0) * prR / (prL + prR);
    real cmax = fmax(aL, aR);
    real machL = uL / cmax;
    real machR = uR / cmax;
    real psL, usL;
    real upL = real(0.5) * (uL + fabs(uL));
    if (fabs(machL) <= real(1.0)) {
        psL = pL * (machL + real(1.0)) * (machL + real(1.0)) * (real(2.0) - machL) * real(0.25);
        usL = wdL * ((uL + cmax) * (uL + cmax) / (real(4.0) * cmax) - upL) + upL;
    } else {
        psL = pL * upL / uL;
        usL = upL;
    }
    real psR, usR;
    real unR = real(0.5) * (uR - fabs(uR));
    if (fabs(machR) <= real(1.0)) {
        psR = pR * (machR - real(1.0)) * (machR - real(1.0)) * (real(2.0) + machR) * real(0.25);
        usR = wdR * (-(uR - cmax) * (uR - cmax) / (real(4.0) * cmax) - unR) + unR;
    } else {
        psR = pR * unR / uR;
        usR = unR;
    }
    real mfL = usL * rhoL;
    real mfR = usR * rhoR;
    real mdot = mfL + mfR;
    real pface = psL + psR;
    real pjump = pL - pR;
    pjump = real(10.0) * fabs(pjump) / fmin(pL, pR);
    real blend = real(0.5) * fmin(real(1.0), pjump);
    real momV = mfL * uL + mfR * uR;
    real momD = real(0.5) * (mdot * (uL + uR) - fabs(mdot) * (uR - uL));
    real mom = (real(0.5) + blend) * momV + (real(0.5) - blend) * momD;
    bool upw = (mdot >= real(0.0));
    real vt = upw ? vL : vR;
    real wt = upw ? wL : wR;
    F[0] = mdot;
    to_global(f, mom + pface, mdot * vt, mdot * wt, F[1], F[2], F[3]);
    F[4] = mdot * (upw ? hL : hR);
    F[5] = mdot * (upw ? YBL : YBR);
}

// face.cu:347-371 for flux_calc = ausmdv.  XORDER 2 reconstructs rho, e, YB, then applies
// the equation of state, then reconstructs the velocity; XORDER 1 uses the adjacent cells.
template <int XORDER>
CHKN_HD void face_flux(const Prim& L1, const Prim& L0, const Prim& R0, const Prim& R1,
                       const Frame& f, real F[6])
{
    real rhoL, rhoR, eL, eR, YBL, YBR, vxL, vxR, vyL, vyR, vzL, vzR;
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
    real TL = eos_T(eL, YBL), TR = eos_T(eR, YBR);
    real pL = eos_p(rhoL, TL), pR = eos_p(rhoR, TR);
    real aL = eos_a(TL), aR = eos_a(TR);
    ausmdv(rhoL, eL, pL, aL, YBL, vxL, vyL, vzL, rhoR, eR, pR, aR, YBR, vxR, vyR, vzR, f, F);
}

// Summation-by-parts alpha-split flux (face.cu:268-342, after Fisher et al., NASA/TM-2011-
// 217307): fourth-order central divergence-form and product-rule fluxes of the factored
// conservation terms over the four-cell stencil, blended with the alpha weights 1 (mass),
// 1/2 (momentum, internal energy), 0 (kinetic energy, pressure).  Uses each state's own p
// and e.  The alpha products are kept as written: 0*x is not 0 for all x.
CHKN_HD void sbp_asf(const Prim* s, const real* p, const Frame& f, real F[6])
{
    constexpr real one_twelfth = real(1.0) / real(12.0);
    real fac[10][4], adv[10][4];
    for (int i = 0; i < 4; ++i) {
        real un, ut1, ut2;
        to_local(f, s[i].vx, s[i].vy, s[i].vz, un, ut1, ut2);
        const real rho = s[i].rho;
        fac[0][i] = rho;                 adv[0][i] = un;
        fac[1][i] = un * rho;          adv[1][i] = un;
        fac[2][i] = ut1 * rho;          adv[2][i] = un;
        fac[3][i] = ut2 * rho;          adv[3][i] = un;
        fac[4][i] = s[i].e * rho;        adv[4][i] = un;
        fac[5][i] = un * un * rho;   adv[5][i] = un;
        fac[6][i] = ut1 * ut1 * rho;   adv[6][i] = un;
        fac[7][i] = ut2 * ut2 * rho;   adv[7][i] = un;
        fac[8][i] = p[i];                adv[8][i] = un;
        fac[9][i] = p[i];                adv[9][i] = real(1.0);
    }
    real cen[10], prod[10];
    for (int j = 0; j < 10; ++j) {
        cen[j] = one_twelfth * (-fac[j][0] * adv[j][0] + real(7.0) * fac[j][1] * adv[j][1] + real(7.0) * fac[j][2] * adv[j][2] -
                               fac[j][3] * adv[j][3]);
        prod[j] = one_twelfth * (-fac[j][0] * adv[j][2] - fac[j][2] * adv[j][0] + real(8.0) * fac[j][1] * adv[j][2] +
                               real(8.0) * fac[j][2] * adv[j][1] - fac[j][1] * adv[j][3] - fac[j][3] * adv[j][1]);
    }
    const real bl_mass = real(1.0), bl_mom = real(0.5), bl_ie = real(0.5), bl_ke = real(0.0), bl_p = real(0.0);
    const real mass = bl_mass * cen[0] + (real(1.0) - bl_mass) * prod[0];
    F[0] = mass;
    const real mx = bl_mom * cen[1] + (real(1.0) - bl_mom) * prod[1] + (bl_p * cen[9] + (real(1.0) - bl_p) * prod[9]);
    const real my = bl_mom * cen[2] + (real(1.0) - bl_mom) * prod[2];
    const real mz = bl_mom * cen[3] + (real(1.0) - bl_mom) * prod[3];
    to_global(f, mx, my, mz, F[1], F[2], F[3]);
    F[4] = bl_ie * cen[4] + (real(1.0) - bl_ie) * prod[4] +
           real(0.5) * (bl_ke * cen[5] + (real(1.0) - bl_ke) * prod[5] + bl_ke * cen[6] + (real(1.0) - bl_ke) * prod[6] +
                  bl_ke * cen[7] + (real(1.0) - bl_ke) * prod[7]) +
           bl_p * cen[8] + (real(1.0) - bl_p) * prod[8];
    F[5] = mass * ((mass >= real(0.0)) ? s[1].YB : s[2].YB);
}

// flow.cu:227-243
CHKN_HD void encode(const Prim& s, real U[6])
{
    real rho = s.rho;
    U[0] = rho;
    U[1] = rho * s.vx;
    U[2] = rho * s.vy;
    U[3] = rho * s.vz;
    real ke = real(0.5) * (s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
    U[4] = rho * (s.e + ke);
    U[5] = rho * s.YB;
}

// flow.cu:245-268 without the equation-of-state tail.  Returns 1 for a bad cell.
CHKN_HD int decode(const real U[6], Prim& s)
{
    bool nan = isnan(U[0]) || isnan(U[1]) || isnan(U[2]) || isnan(U[3]) || isnan(U[4]) || isnan(U[5]);
    if (nan) return 1;
    real rho = U[0];
    if (rho <= real(0.0)) return 1;
    real dinv = real(1.0) / rho;
    s.vx = U[1] * dinv;
    s.vy = U[2] * dinv;
    s.vz = U[3] * dinv;
    real e = U[4] * dinv;
    real ke = real(0.5) * (s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
    e -= ke;
    s.rho = rho;
    s.e = e;
    s.YB = U[5] * dinv;
    return 0;
}

// Signal-speed time-step estimate without the cfl factor, cell.cu:155-164.  The cfl
// multiply is applied after the minimum: rounding is monotone, so
// min(fl(cfl*x_i)) == fl(cfl*min(x_i)) for cfl > 0.
CHKN_HD real local_signal(const Prim& s, real a, real iL, real jL, real kL,
                            const Frame& fi, const Frame& fj, const Frame& fk)
{
    real is = iL / (fabs(s.vx * fi.nx + s.vy * fi.ny + s.vz * fi.nz) + a);
    real js = jL / (fabs(s.vx * fj.nx + s.vy * fj.ny + s.vz * fj.nz) + a);
    real ks = kL / (fabs(s.vx * fk.nx + s.vy * fk.ny + s.vz * fk.nz) + a);
    return fmin(fmin(is, js), ks);
}

}  // namespace chkn
