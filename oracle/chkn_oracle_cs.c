/* chkn_oracle_cs.c -- complex-step derivatives of chicken's face flux.
 *
 * TEST INFRASTRUCTURE ONLY: the product never links this file.
 *
 * The face flux of face.cu:347-371 (MUSCL/van Albada reconstruction, face.cu:72-96; the
 * ideal-gas equation of state, gas.cu:128-135; AUSMDV, face.cu:163-265; and the SBP-ASF flux,
 * face.cu:268-342) restated over complex numbers.  For a real input x and direction v, Im F(x + i h v) / h is the
 * directional derivative J v to rounding error (no subtractive cancellation), which is
 * what the adjoint kernels' vector-Jacobian products are checked against:
 * <Fbar, J v> = <J^T Fbar, v>.
 *
 * Complex-step rules for the non-analytic pieces: |z| = z or -z by the sign of Re z;
 * min/max and branch tests compare real parts; sqrt is the principal complex root.
 *
 * Pinning: Re F equals the real oracle flux (orc_face_flux) bit for bit when h v is tiny,
 * and Im F / h agrees with central differences of the real flux (tests/test_oracle_cs.py).
 */
#include <complex.h>
#include <math.h>
#include <string.h>

typedef double complex cx;

/* gas.cu:20-48: the gas model is chosen when the oracle is built, as the reference's is. */
#ifdef ORC_AB_REACTING_GAS
static const double GAS_G = 6.0 / 5.0;
static const double GAS_R = 287.0;
static const double GAS_Q = 300000.0;
#else
static const double GAS_G = 1.4;
static const double GAS_R = 287.1;
static const double GAS_Q = 0.0;
#endif
#define GAS_CV (GAS_R / (GAS_G - 1.0))

static cx cabs_cs(cx z) { return creal(z) < 0.0 ? -z : z; }
static cx cmax_cs(cx a, cx b) { return creal(a) >= creal(b) ? a : b; }  /* fmax: a unless b > a */
static cx cmin_cs(cx a, cx b) { return creal(a) <= creal(b) ? a : b; }

typedef struct { cx rho, e, YB, vx, vy, vz, p, T, a; } CS;

static void rhoe_cs(CS* s) /* gas.cu:128-135 */
{
    s->T = (s->e + s->YB * GAS_Q) / GAS_CV;
    s->p = s->rho * GAS_R * s->T;
    s->a = csqrt(GAS_G * GAS_R * s->T);
}

static cx van_albada_cs(cx a, cx b) /* face.cu:72-79 */
{
    const double eps = 1.0e-12;
    return (a * b + cabs_cs(a * b) + eps) / (a * a + b * b + eps);
}

static void interp_cs(cx qL1, cx qL0, cx qR0, cx qR1, cx* qL, cx* qR) /* face.cu:81-96 */
{
    cx dm = qL0 - qL1, d = qR0 - qL0, dp = qR1 - qR0;
    cx sL = van_albada_cs(dm, d), sR = van_albada_cs(d, dp);
    *qL = qL0 + sL * 0.125 * (3.0 * d + dm);
    *qR = qR0 - sR * 0.125 * (dp + 3.0 * d);
}

static void local_cs(const double* f, cx x, cx y, cx z, cx* u, cx* v, cx* w) /* vector3.cu:97-109 */
{
    *u = x * f[0] + y * f[1] + z * f[2];
    *v = x * f[3] + y * f[4] + z * f[5];
    *w = x * f[6] + y * f[7] + z * f[8];
}

static void global_cs(const double* f, cx u, cx v, cx w, cx* x, cx* y, cx* z) /* vector3.cu:111-119 */
{
    *x = u * f[0] + v * f[3] + w * f[6];
    *y = u * f[1] + v * f[4] + w * f[7];
    *z = u * f[2] + v * f[5] + w * f[8];
}

static void ausmdv_cs(const CS* L, const CS* R, const double* f, cx* F) /* face.cu:163-265 */
{
    cx uL, vL, wL, uR, vR, wR;
    local_cs(f, L->vx, L->vy, L->vz, &uL, &vL, &wL);
    local_cs(f, R->vx, R->vy, R->vz, &uR, &vR, &wR);
    cx pLrL = L->p / L->rho, pRrR = R->p / R->rho;
    cx keL = 0.5 * (uL * uL + vL * vL + wL * wL), keR = 0.5 * (uR * uR + vR * vR + wR * wR);
    cx HL = L->e + pLrL + keL, HR = R->e + R->p / R->rho + keR;
    cx alphaL = 2.0 * pLrL / (pLrL + pRrR), alphaR = 2.0 * pRrR / (pLrL + pRrR);
    cx am = cmax_cs(L->a, R->a);
    cx ML = uL / am, MR = uR / am;
    cx pLplus, uLplus, pRminus, uRminus;
    cx dvL = 0.5 * (uL + cabs_cs(uL));
    if (fabs(creal(ML)) <= 1.0) {
        pLplus = L->p * (ML + 1.0) * (ML + 1.0) * (2.0 - ML) * 0.25;
        uLplus = alphaL * ((uL + am) * (uL + am) / (4.0 * am) - dvL) + dvL;
    } else {
        pLplus = L->p * dvL / uL;
        uLplus = dvL;
    }
    cx dvR = 0.5 * (uR - cabs_cs(uR));
    if (fabs(creal(MR)) <= 1.0) {
        pRminus = R->p * (MR - 1.0) * (MR - 1.0) * (2.0 + MR) * 0.25;
        uRminus = alphaR * (-(uR - am) * (uR - am) / (4.0 * am) - dvR) + dvR;
    } else {
        pRminus = R->p * dvR / uR;
        uRminus = dvR;
    }
    cx massL = uLplus * L->rho, massR = uRminus * R->rho;
    cx m = massL + massR, ph = pLplus + pRminus;
    cx dp = 10.0 * cabs_cs(L->p - R->p) / cmin_cs(L->p, R->p);
    cx s = 0.5 * cmin_cs(1.0, dp);
    cx rvV = massL * uL + massR * uR;
    cx rvD = 0.5 * (m * (uL + uR) - cabs_cs(m) * (uR - uL));
    cx rv2 = (0.5 + s) * rvV + (0.5 - s) * rvD;
    int up = creal(m) >= 0.0;
    F[0] = m;
    global_cs(f, rv2 + ph, m * (up ? vL : vR), m * (up ? wL : wR), &F[1], &F[2], &F[3]);
    F[4] = m * (up ? HL : HR);
    F[5] = m * (up ? L->YB : R->YB);
}

/* face.cu:268-342: the SBP-ASF flux over the four stencil states, each with its own p (here
 * from the equation of state); the alpha products are kept as written. */
static void sbp_asf_cs(const CS* g, const double* f, cx* F)
{
    cx v[10][4], w[10][4];
    for (int i = 0; i < 4; ++i) {
        cx velx, vely, velz;
        local_cs(f, g[i].vx, g[i].vy, g[i].vz, &velx, &vely, &velz);
        const cx rho = g[i].rho;
        v[0][i] = rho;                 w[0][i] = velx;
        v[1][i] = velx * rho;          w[1][i] = velx;
        v[2][i] = vely * rho;          w[2][i] = velx;
        v[3][i] = velz * rho;          w[3][i] = velx;
        v[4][i] = g[i].e * rho;        w[4][i] = velx;
        v[5][i] = velx * velx * rho;   w[5][i] = velx;
        v[6][i] = vely * vely * rho;   w[6][i] = velx;
        v[7][i] = velz * velz * rho;   w[7][i] = velx;
        v[8][i] = g[i].p;              w[8][i] = velx;
        v[9][i] = g[i].p;              w[9][i] = 1.0;
    }
    cx fc[10], fe[10];
    for (int j = 0; j < 10; ++j) {
        fc[j] = (1.0 / 12.0) * (-v[j][0] * w[j][0] + 7.0 * v[j][1] * w[j][1] + 7.0 * v[j][2] * w[j][2] -
                                v[j][3] * w[j][3]);
        fe[j] = (1.0 / 12.0) * (-v[j][0] * w[j][2] - v[j][2] * w[j][0] + 8.0 * v[j][1] * w[j][2] +
                                8.0 * v[j][2] * w[j][1] - v[j][1] * w[j][3] - v[j][3] * w[j][1]);
    }
    const double a_mass = 1.0, a_mom = 0.5, a_ie = 0.5, a_ke = 0.0, a_p = 0.0;
    const cx mass = a_mass * fc[0] + (1.0 - a_mass) * fe[0];
    F[0] = mass;
    const cx mx = a_mom * fc[1] + (1.0 - a_mom) * fe[1] + (a_p * fc[9] + (1.0 - a_p) * fe[9]);
    const cx my = a_mom * fc[2] + (1.0 - a_mom) * fe[2];
    const cx mz = a_mom * fc[3] + (1.0 - a_mom) * fe[3];
    global_cs(f, mx, my, mz, &F[1], &F[2], &F[3]);
    F[4] = a_ie * fc[4] + (1.0 - a_ie) * fe[4] +
           0.5 * (a_ke * fc[5] + (1.0 - a_ke) * fe[5] + a_ke * fc[6] + (1.0 - a_ke) * fe[6] +
                  a_ke * fc[7] + (1.0 - a_ke) * fe[7]) +
           a_p * fc[8] + (1.0 - a_p) * fe[8];
    F[5] = mass * (creal(mass) >= 0.0 ? g[1].YB : g[2].YB);
}

/* prims: 4 states x (rho, e, YB, vx, vy, vz); frame: n, t1, t2; dir: 24 values.
 * Fre = Re F(x + i h dir), jv = Im F(x + i h dir) / h.  x_order 1 or 2: AUSMDV on the
 * reconstructed (x_order 2) or adjacent states; x_order 0: the SBP-ASF flux. */
void orc_face_flux_cs(const double* prims, const double* frame, int x_order, const double* dir,
                      double h, double* Fre, double* jv)
{
    CS st[4];
    for (int c = 0; c < 4; ++c) {
        const double* p = prims + 6 * c;
        const double* v = dir + 6 * c;
        st[c].rho = p[0] + I * h * v[0];
        st[c].e = p[1] + I * h * v[1];
        st[c].YB = p[2] + I * h * v[2];
        st[c].vx = p[3] + I * h * v[3];
        st[c].vy = p[4] + I * h * v[4];
        st[c].vz = p[5] + I * h * v[5];
    }
    if (x_order == 0) {
        cx F[6];
        for (int c = 0; c < 4; ++c) rhoe_cs(&st[c]);
        sbp_asf_cs(st, frame, F);
        for (int q = 0; q < 6; ++q) {
            Fre[q] = creal(F[q]);
            jv[q] = cimag(F[q]) / h;
        }
        return;
    }
    CS L = st[1], R = st[2]; /* face.cu:347-371 */
    if (x_order > 1) {
        interp_cs(st[0].rho, st[1].rho, st[2].rho, st[3].rho, &L.rho, &R.rho);
        interp_cs(st[0].e, st[1].e, st[2].e, st[3].e, &L.e, &R.e);
        interp_cs(st[0].YB, st[1].YB, st[2].YB, st[3].YB, &L.YB, &R.YB);
        interp_cs(st[0].vx, st[1].vx, st[2].vx, st[3].vx, &L.vx, &R.vx);
        interp_cs(st[0].vy, st[1].vy, st[2].vy, st[3].vy, &L.vy, &R.vy);
        interp_cs(st[0].vz, st[1].vz, st[2].vz, st[3].vz, &L.vz, &R.vz);
    }
    rhoe_cs(&L);
    rhoe_cs(&R);
    cx F[6];
    ausmdv_cs(&L, &R, frame, F);
    for (int q = 0; q < 6; ++q) {
        Fre[q] = creal(F[q]);
        jv[q] = cimag(F[q]) / h;
    }
}
