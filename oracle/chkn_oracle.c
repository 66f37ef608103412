/* chkn_oracle.c -- CPU restatement of chicken's explicit FV time step.
 *
 * TEST INFRASTRUCTURE ONLY (see chkn_oracle.h).  The product never links this file.
 *
 * Plain C99 in the reference's number type (FP64, or FP32 for liborc_f32.so), built with
 * -ffp-contract=off so that every a*b+c is two roundings, as in the reference's gcc -O2
 * x86-64 build.  Where the reference's FP32 build evaluates an expression in double (a
 * double literal or argument), this code does the same with an explicit cast.  Every expression keeps the reference's operand
 * order and association; comments give the reference file:line each piece restates
 * (paths relative to the reference's src/chicken/).
 *
 * Pinning (what checks this code, see tests/test_oracle_*.py):
 *   - geometry, EOS, encode/decode, limiter, MUSCL, AUSMDV: closed forms and invariants
 *     (axis-aligned boxes, consistency F(W,W) = Euler flux, supersonic upwinding,
 *     mirror-state zero mass flux, s(a,a)=1, linear data reproduced exactly);
 *   - flux on general states, hex volumes on warped cells, whole multi-step trajectories:
 *     golden vectors written by the reference's own sources (tests/golden/), bit for bit.
 */
#include "chkn_oracle.h"

#include <float.h>
#include <math.h>
#include <tgmath.h> /* sqrt, fabs, fmin, fmax, exp of the real type */
#ifdef I
#undef I
#endif

/* The number type (number.cu:10-45): double, or float in the FP32 build (ORC_FP32, the
 * reference's FLOAT_NUMBERS).  Literals are the reference's number constants. */
#ifdef ORC_FP32
typedef float real;
#define REAL_MAX FLT_MAX
#else
typedef double real;
#define REAL_MAX DBL_MAX
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------ */
/* Constants: number.cu:29-45, gas.cu:21-49 (ideal air, the IDEAL_AIR build).            */

static const real zero = 0.0, half = 0.5, one_quarter = 0.25;
static const real one_eighth = 0.125, one = 1.0, two = 2.0, three = 3.0, four = 4.0;
static const real six = 6.0, seven = 7.0, eight = 8.0;
#define ONE_SIXTH ((real)1 / (real)6)
#define FOUR_SIXTH ((real)4 / (real)6)
#define ONE_TWELFTH ((real)1 / (real)12)

/* gas.cu:20-48: the gas model is chosen when the oracle is built, as the reference's is. */
#ifdef ORC_AB_REACTING_GAS
static const real GAS_G = (real)(6.0 / 5.0);
static const real GAS_R = (real)287.0;
static const real GAS_Q = (real)300000.0;
#else
static const real GAS_G = (real)1.4;
static const real GAS_R = (real)287.1;
static const real GAS_Q = (real)0.0;
#endif
/* gas.cu:40: Cv = R / (g - 1.0), the subtraction and division in double. */
#define GAS_CV ((real)((double)GAS_R / ((double)GAS_G - 1.0)))

static char g_err[512];
const char* orc_last_error(void) { return g_err; }
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }

/* ------------------------------------------------------------------------------------ */
/* Types.                                                                                 */

typedef struct { real x, y, z; } V3;
typedef struct { real rho, e, p, T, YB, a; V3 vel; } FS; /* FlowState, flow.cu:38-40 */

typedef struct {
    V3 pos; real vol, iL, jL, kL;
    FS fs;
    int face[6];
} Cell;

typedef struct {
    V3 pos, n, t1, t2; real area;
    real F[6];
    int L1, L0, R0, R1;
    FS fs; /* face-average state, face.cu:373 */
    /* viscous: boundary code (-1 interior), cloud of cells and faces, LSQ weights */
    int bcode, bfun, nc, nf, cc[2], fcl[8];
    real TWall, wx[10], wy[10], wz[10];
} Face;

typedef struct {
    int bi, bj, bk, nic, njc, nkc, active;
    int ncells, n0c[6], n1c[6], nghost[6], first_ghost[6], ntot;
    int nif, njf, nkf, nfaces;
    int bc[6], bcfs[6], bcfun[6];
    real twall[6];
    real* twf[6]; /* TWall per boundary face, or NULL */
    Cell* cells;
    Face* faces;
    V3* vtx;
    real* Q;  /* 2 levels x ncells x 6, cell-major (Q[level*ncells+i][j]) */
    real* dQ; /* 3 levels */
} Blk;

/* spline.cu:20-109, 12 knots. */
#define ORC_BL_N 11
typedef struct { real xs[ORC_BL_N + 1], ys[ORC_BL_N + 1], a[ORC_BL_N], b[ORC_BL_N], c[ORC_BL_N]; } Spline;
typedef struct { real p_e; Spline T, velx; int set; } BLProfile; /* config.cu:33-93 */

struct orc_sim {
    BLProfile bl;
    int nblocks, nib, njb, nkb;
    int* blk_ids;
    Blk* b;
    int nfs;
    FS* inflow;
    int x_order, flux_calc, cfl_count, max_step, viscous, reacting, time_scheme;
    real max_time, dt_init;
    int n_cfl;
    real* cfl_t;
    real* cfl_v;
    real t, dt;
    int step;
};

/* ------------------------------------------------------------------------------------ */
/* Vector3 (vector3.cu).                                                                  */

static real dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; } /* :72-75 */

/* vector3.cu:97-109 */
static V3 to_local(V3 v, V3 n, V3 t1, V3 t2)
{
    V3 r; r.x = dot3(v, n); r.y = dot3(v, t1); r.z = dot3(v, t2); return r;
}

/* vector3.cu:111-119 */
static V3 to_global(V3 v, V3 n, V3 t1, V3 t2)
{
    V3 r;
    r.x = v.x * n.x + v.y * t1.x + v.z * t2.x;
    r.y = v.x * n.y + v.y * t1.y + v.z * t2.y;
    r.z = v.x * n.z + v.y * t1.z + v.z * t2.z;
    return r;
}

/* vector3.cu:77-93 */
static V3 normalize(V3 v)
{
    real m = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (m > zero) { v.x /= m; v.y /= m; v.z /= m; }
    else { v.x = v.y = v.z = zero; }
    const real small = (real)1.0e-30;
    if (fabs(v.x) < small) v.x = zero;
    if (fabs(v.y) < small) v.y = zero;
    if (fabs(v.z) < small) v.z = zero;
    return v;
}

/* vector3.cu:150-157, cross(v3, v1, v2) */
static V3 cross_v(V3 v1, V3 v2)
{
    V3 r;
    r.x = v1.y * v2.z - v2.y * v1.z;
    r.y = v2.x * v1.z - v1.x * v2.z;
    r.z = v1.x * v2.y - v2.x * v1.y;
    return r;
}

/* vector3.cu:179-187, cross_product(a, b, c) */
static void cross_c(real ax, real ay, real az, real bx, real by, real bz,
                    real* cx, real* cy, real* cz)
{
    *cx = ay * bz - az * by;
    *cy = az * bx - ax * bz;
    *cz = ax * by - ay * bx;
}

/* vector3.cu:196-245 */
static void quad_properties(V3 p0, V3 p1, V3 p2, V3 p3, V3* c, V3* n, V3* t1, V3* t2,
                            real* area)
{
    c->x = one_quarter * (p0.x + p1.x + p2.x + p3.x);
    c->y = one_quarter * (p0.y + p1.y + p2.y + p3.y);
    c->z = one_quarter * (p0.z + p1.z + p2.z + p3.z);
    real p01x = p1.x - p0.x + p2.x - p3.x;
    real p01y = p1.y - p0.y + p2.y - p3.y;
    real p01z = p1.z - p0.z + p2.z - p3.z;
    real p03x = p3.x - p0.x + p2.x - p1.x;
    real p03y = p3.y - p0.y + p2.y - p1.y;
    real p03z = p3.z - p0.z + p2.z - p1.z;
    real vx, vy, vz;
    cross_c(p01x, p01y, p01z, p03x, p03y, p03z, &vx, &vy, &vz);
    vx *= one_quarter; vy *= one_quarter; vz *= one_quarter;
    *area = sqrt(vx * vx + vy * vy + vz * vz);
    if (*area > (real)1.0e-20) {
        n->x = vx / *area; n->y = vy / *area; n->z = vz / *area;
        V3 t = {p01x, p01y, p01z};
        *t1 = normalize(t);
        *t2 = normalize(cross_v(*n, *t1));
    } else {
        n->x = one; n->y = zero; n->z = zero;
        t1->x = zero; t1->y = one; t1->z = zero;
        t2->x = zero; t2->y = zero; t2->z = one;
    }
}

/* vector3.cu:259-270 */
static real tet_volume(V3 p0, V3 p1, V3 p2, V3 p3)
{
    real dx01 = p1.x - p0.x, dy01 = p1.y - p0.y, dz01 = p1.z - p0.z;
    real dx02 = p2.x - p0.x, dy02 = p2.y - p0.y, dz02 = p2.z - p0.z;
    real cx, cy, cz;
    cross_c(dx01, dy01, dz01, dx02, dy02, dz02, &cx, &cy, &cz);
    real dx03 = p3.x - p0.x, dy03 = p3.y - p0.y, dz03 = p3.z - p0.z;
    return (dx03 * cx + dy03 * cy + dz03 * cz) / six;
}

/* vector3.cu:309-348 with true_centroid = true, as called from hex_cell_properties */
static void pyramid_props(V3 p0, V3 p1, V3 p2, V3 p3, V3 p4, V3* cen, real* vol)
{
    V3 pmB;
    pmB.x = one_quarter * (p0.x + p1.x + p2.x + p3.x);
    pmB.y = one_quarter * (p0.y + p1.y + p2.y + p3.y);
    pmB.z = one_quarter * (p0.z + p1.z + p2.z + p3.z);
    V3 q[4][4] = {{p0, p1, pmB, p4}, {p1, p2, pmB, p4}, {p2, p3, pmB, p4}, {p3, p0, pmB, p4}};
    real v = zero; V3 m = {zero, zero, zero};
    for (int t = 0; t < 4; ++t) {
        /* tetrahedron_properties, vector3.cu:272-281 */
        V3 tc;
        tc.x = one_quarter * (q[t][0].x + q[t][1].x + q[t][2].x + q[t][3].x);
        tc.y = one_quarter * (q[t][0].y + q[t][1].y + q[t][2].y + q[t][3].y);
        tc.z = one_quarter * (q[t][0].z + q[t][1].z + q[t][2].z + q[t][3].z);
        real tv = tet_volume(q[t][0], q[t][1], q[t][2], q[t][3]);
        v += tv;
        tc.x *= tv; tc.y *= tv; tc.z *= tv;
        m.x += tc.x; m.y += tc.y; m.z += tc.z;
    }
    if (fabs(v) > zero) { m.x /= v; m.y /= v; m.z /= v; }
    *cen = m; *vol = v;
}

/* vector3.cu:350-433 with true_centroid = false (block.cu:642-646) */
static void hex_props(const V3 p[8], V3* cen, real* vol, real* iL, real* jL,
                      real* kL)
{
    cen->x = one_eighth * (p[0].x + p[1].x + p[2].x + p[3].x + p[4].x + p[5].x + p[6].x + p[7].x);
    cen->y = one_eighth * (p[0].y + p[1].y + p[2].y + p[3].y + p[4].y + p[5].y + p[6].y + p[7].y);
    cen->z = one_eighth * (p[0].z + p[1].z + p[2].z + p[3].z + p[4].z + p[5].z + p[6].z + p[7].z);
#define MID4(a, b, c, d, r) do { \
        (r).x = one_quarter * (p[a].x + p[b].x + p[c].x + p[d].x); \
        (r).y = one_quarter * (p[a].y + p[b].y + p[c].y + p[d].y); \
        (r).z = one_quarter * (p[a].z + p[b].z + p[c].z + p[d].z); } while (0)
    V3 pmN, pmE, pmS, pmW, pmT, pmB;
    MID4(3, 2, 6, 7, pmN);
    MID4(1, 2, 6, 5, pmE);
    MID4(0, 1, 5, 4, pmS);
    MID4(0, 3, 7, 4, pmW);
    MID4(4, 5, 6, 7, pmT);
    MID4(0, 1, 2, 3, pmB);
#undef MID4
    real dx, dy, dz;
    dx = pmE.x - pmW.x; dy = pmE.y - pmW.y; dz = pmE.z - pmW.z;
    *iL = sqrt(dx * dx + dy * dy + dz * dz);
    dx = pmN.x - pmS.x; dy = pmN.y - pmS.y; dz = pmN.z - pmS.z;
    *jL = sqrt(dx * dx + dy * dy + dz * dz);
    dx = pmT.x - pmB.x; dy = pmT.y - pmB.y; dz = pmT.z - pmB.z;
    *kL = sqrt(dx * dx + dy * dy + dz * dz);
    static const int pyr[6][4] = {{6, 7, 3, 2}, {5, 6, 2, 1}, {4, 5, 1, 0},
                                  {7, 4, 0, 3}, {7, 6, 5, 4}, {0, 1, 2, 3}};
    real v = zero; V3 m = {zero, zero, zero};
    for (int q = 0; q < 6; ++q) {
        V3 sc; real sv;
        pyramid_props(p[pyr[q][0]], p[pyr[q][1]], p[pyr[q][2]], p[pyr[q][3]], *cen, &sc, &sv);
        v += sv;
        sc.x *= sv; sc.y *= sv; sc.z *= sv;
        m.x += sc.x; m.y += sc.y; m.z += sc.z;
    }
    if ((v < zero && fabs(v) < (real)1.0e-12) || (v >= zero && v < (real)1.0e-20)) v = zero;
    *vol = v;
}

/* ------------------------------------------------------------------------------------ */
/* Gas and flow state (gas.cu, flow.cu).                                                  */

static void update_from_pT(FS* s) /* gas.cu:119-126 */
{
    s->e = s->T * GAS_CV - s->YB * GAS_Q;
    s->rho = s->p / (s->T * GAS_R);
    s->a = sqrt(GAS_G * GAS_R * s->T);
}

/* spline.cu:26-73: natural cubic spline; the interior second derivatives solve a
 * tridiagonal system by Gauss-Jordan elimination with partial pivoting (rsla.cu:201-227).
 * Returns -1 when a pivot is below 1e-16. */
static int spline_set(Spline* sp, const real* x, const real* y)
{
    enum { n = ORC_BL_N, m = ORC_BL_N - 1 };
    real h[n], A[m][m + 1], sigma[n + 1];
    for (int i = 0; i <= n; ++i) { sp->xs[i] = x[i]; sp->ys[i] = y[i]; }
    for (int i = 0; i < n; ++i) h[i] = sp->xs[i + 1] - sp->xs[i];
    for (int r = 0; r < m; ++r) for (int q = 0; q <= m; ++q) A[r][q] = 0.0;
    for (int k = 0; k < m; ++k) {
        int i = k + 1;
        if (k > 0) A[k][k - 1] = h[i - 1];
        A[k][k] = (real)(2.0 * (double)(h[i - 1] + h[i])); /* double literal */
        if (k < m - 1) A[k][k + 1] = h[i];
        A[k][m] = (real)(6.0 * (double)((sp->ys[i + 1] - sp->ys[i]) / h[i] - (sp->ys[i] - sp->ys[i - 1]) / h[i - 1]));
    }
    for (int j = 0; j < m; ++j) {
        int p = j;
        for (int i = j + 1; i < m; ++i) if (fabs(A[i][j]) > fabs(A[p][j])) p = i;
        if ((double)fabs(A[p][j]) < 1.0e-16) return -1; /* rsla.cu:202,211: a double threshold */
        if (p != j) for (int q = 0; q <= m; ++q) { real t = A[p][q]; A[p][q] = A[j][q]; A[j][q] = t; }
        real d = A[j][j];
        for (int q = 0; q <= m; ++q) A[j][q] /= d;
        for (int i = 0; i < m; ++i) {
            if (i == j) continue;
            real f = A[i][j];
            for (int q = 0; q <= m; ++q) A[i][q] -= f * A[j][q];
        }
    }
    sigma[0] = 0.0; sigma[n] = 0.0;
    for (int i = 0; i < m; ++i) sigma[i + 1] = A[i][m];
    for (int i = 0; i < n; ++i) {
        sp->a[i] = (sigma[i + 1] - sigma[i]) / (six * h[i]);
        sp->b[i] = sigma[i] / two;
        sp->c[i] = (sp->ys[i + 1] - sp->ys[i]) / h[i] - (two * sigma[i] + sigma[i + 1]) * h[i] / six;
    }
    return 0;
}

static real spline_eval(const Spline* sp, real x) /* spline.cu:86-99 */
{
    int i = 0;
    if (!(x < sp->xs[0])) {
        i = ORC_BL_N - 1;
        while (x < sp->xs[i] && i > 0) i -= 1;
    }
    double dx = (double)(x - sp->xs[i]); /* spline.cu:94-95: the piece in double */
    return (real)((((double)sp->a[i] * dx + (double)sp->b[i]) * dx + (double)sp->c[i]) * dx + (double)sp->ys[i]);
}

/* config.cu:37-60: p_e, then y, T, velx per sample. */
static int bl_set(BLProfile* bl, const real* v)
{
    real ys[ORC_BL_N + 1], Ts[ORC_BL_N + 1], us[ORC_BL_N + 1];
    bl->p_e = v[0];
    for (int i = 0; i <= ORC_BL_N; ++i) { ys[i] = v[1 + 3 * i]; Ts[i] = v[2 + 3 * i]; us[i] = v[3 + 3 * i]; }
    if (spline_set(&bl->T, ys, Ts) || spline_set(&bl->velx, ys, us)) return -1;
    bl->set = 1;
    return 0;
}

static void update_from_rhoe(FS* s);

/* gas.cu:102-111: the A->B reaction in an isolated blob of gas over dt. */
static void update_chemistry(FS* s, real dt)
{
#ifdef ORC_AB_REACTING_GAS
    const real alpha = (real)1000.0, Ti = (real)362.58;
#else
    const real alpha = (real)0.0, Ti = (real)0.0;
#endif
    real YA = one - s->YB;
    if (s->T > Ti) s->YB = one - YA * exp(-alpha * dt);
    update_from_rhoe(s);
}

static void update_from_rhoe(FS* s) /* gas.cu:128-135 */
{
    s->T = (s->e + s->YB * GAS_Q) / GAS_CV;
    s->p = s->rho * GAS_R * s->T;
    s->a = sqrt(GAS_G * GAS_R * s->T);
}

static void encode(const FS* s, real* U) /* flow.cu:227-243 */
{
    real rho = s->rho;
    U[0] = rho;
    U[1] = rho * s->vel.x;
    U[2] = rho * s->vel.y;
    U[3] = rho * s->vel.z;
    real ke = half * (s->vel.x * s->vel.x + s->vel.y * s->vel.y + s->vel.z * s->vel.z);
    U[4] = rho * (s->e + ke);
    U[5] = rho * s->YB;
}

static int decode(FS* s, const real* U) /* flow.cu:245-268 */
{
    for (int i = 0; i < 6; ++i) if (isnan(U[i])) return 1;
    real rho = U[0];
    if (rho <= zero) return 1;
    real dinv = one / rho;
    s->vel.x = U[1] * dinv; s->vel.y = U[2] * dinv; s->vel.z = U[3] * dinv;
    real e = U[4] * dinv;
    real ke = half * (s->vel.x * s->vel.x + s->vel.y * s->vel.y + s->vel.z * s->vel.z);
    e -= ke;
    s->rho = rho;
    s->e = e;
    s->YB = U[5] * dinv;
    update_from_rhoe(s);
    return 0;
}

static void set_as_average(FS* s, const FS* a, const FS* b) /* flow.cu:220-225 */
{
    s->rho = half * (a->rho + b->rho);
    s->e = half * (a->e + b->e);
    s->YB = half * (a->YB + b->YB);
    update_from_rhoe(s);
    s->vel.x = half * (a->vel.x + b->vel.x);
    s->vel.y = half * (a->vel.y + b->vel.y);
    s->vel.z = half * (a->vel.z + b->vel.z);
}

/* bcs.cu:321-352: state of inflow function fn at pos.  The boundary-layer profile clamps to
 * its end samples (config.cu:73-93); supersonic_vortex and manufactured_solution are fixed
 * placeholder states in the reference. */
static FS inflow_function(const orc_sim* s, int fn, V3 pos)
{
    FS f;
    memset(&f, 0, sizeof f);
    if (fn == 2) {
        const BLProfile* bl = &s->bl;
        real y = pos.y, T, u;
        if (y <= bl->T.xs[0]) { T = bl->T.ys[0]; u = bl->velx.ys[0]; }
        else if (y >= bl->T.xs[ORC_BL_N]) { T = bl->T.ys[ORC_BL_N]; u = bl->velx.ys[ORC_BL_N]; }
        else { T = spline_eval(&bl->T, y); u = spline_eval(&bl->velx, y); }
        f.p = bl->p_e; f.YB = 0.0; f.T = T;
        update_from_pT(&f);
        f.vel.x = u; f.vel.y = 0.0; f.vel.z = 0.0;
    } else {
        f.p = 100.0e3; f.T = 300.0; f.YB = 0.0;
        update_from_pT(&f);
    }
    return f;
}

/* ------------------------------------------------------------------------------------ */
/* Face flux (face.cu).                                                                   */

static real van_albada(real a, real b) /* face.cu:72-79 */
{
    const real eps = (real)1.0e-12;
    return (a * b + fabs(a * b) + eps) / (a * a + b * b + eps);
}

/* face.cu:81-96 */
static void interp_l2r2(real qL1, real qL0, real qR0, real qR1, real* qL, real* qR)
{
    real delLminus = (qL0 - qL1);
    real del = (qR0 - qL0);
    real delRplus = (qR1 - qR0);
    real sL = van_albada(delLminus, del);
    real sR = van_albada(del, delRplus);
    *qL = qL0 + sL * one_eighth * (three * del + delLminus);
    *qR = qR0 - sR * one_eighth * (delRplus + three * del);
}

/* face.cu:163-265 */
static void ausmdv(const FS* fsL, const FS* fsR, V3 n, V3 t1, V3 t2, real* F)
{
    V3 velL = to_local(fsL->vel, n, t1, t2);
    V3 velR = to_local(fsR->vel, n, t1, t2);
    real rhoL = fsL->rho, pL = fsL->p, pLrL = pL / rhoL;
    real velxL = velL.x, velyL = velL.y, velzL = velL.z;
    real uL = fsL->e, aL = fsL->a;
    real keL = half * (velxL * velxL + velyL * velyL + velzL * velzL);
    real HL = uL + pLrL + keL;
    real YBL = fsL->YB;
    real rhoR = fsR->rho, pR = fsR->p, pRrR = pR / rhoR;
    real velxR = velR.x, velyR = velR.y, velzR = velR.z;
    real uR = fsR->e, aR = fsR->a;
    real keR = half * (velxR * velxR + velyR * velyR + velzR * velzR);
    real HR = uR + pR / rhoR + keR;
    real YBR = fsR->YB;
    real alphaL = two * pLrL / (pLrL + pRrR);
    real alphaR = two * pRrR / (pLrL + pRrR);
    real am = fmax(aL, aR);
    real ML = velxL / am;
    real MR = velxR / am;
    real pLplus, velxLplus;
    real dvelxL = half * (velxL + fabs(velxL));
    if (fabs(ML) <= one) {
        pLplus = pL * (ML + one) * (ML + one) * (two - ML) * one_quarter;
        velxLplus = alphaL * ((velxL + am) * (velxL + am) / (four * am) - dvelxL) + dvelxL;
    } else {
        pLplus = pL * dvelxL / velxL;
        velxLplus = dvelxL;
    }
    real pRminus, velxRminus;
    real dvelxR = half * (velxR - fabs(velxR));
    if (fabs(MR) <= one) {
        pRminus = pR * (MR - one) * (MR - one) * (two + MR) * one_quarter;
        velxRminus = alphaR * (-(velxR - am) * (velxR - am) / (four * am) - dvelxR) + dvelxR;
    } else {
        pRminus = pR * dvelxR / velxR;
        velxRminus = dvelxR;
    }
    real massL = velxLplus * rhoL;
    real massR = velxRminus * rhoR;
    real mass_half = massL + massR;
    real p_half = pLplus + pRminus;
    real dp = pL - pR;
    const real K_SWITCH = 10.0;
    dp = K_SWITCH * fabs(dp) / fmin(pL, pR);
    real s = half * fmin(one, dp);
    real rvel2_AUSMV = massL * velxL + massR * velxR;
    real rvel2_AUSMD = half * (mass_half * (velxL + velxR) - fabs(mass_half) * (velxR - velxL));
    real rvel2_half = (half + s) * rvel2_AUSMV + (half - s) * rvel2_AUSMD;
    F[0] = mass_half;
    real vely = (mass_half >= zero) ? velyL : velyR;
    real velz = (mass_half >= zero) ? velzL : velzR;
    V3 mom = {rvel2_half + p_half, mass_half * vely, mass_half * velz};
    mom = to_global(mom, n, t1, t2);
    F[1] = mom.x; F[2] = mom.y; F[3] = mom.z;
    real H = (mass_half >= zero) ? HL : HR;
    F[4] = mass_half * H;
    real YB = (mass_half >= zero) ? YBL : YBR;
    F[5] = mass_half * YB;
}

/* face.cu:268-342 */
static void sbp_asf(const FS* s0, const FS* s1, const FS* s2, const FS* s3, V3 n, V3 t1, V3 t2,
                    real* F)
{
    const FS* g[4] = {s0, s1, s2, s3};
    V3 vel[4];
    for (int i = 0; i < 4; ++i) vel[i] = to_local(g[i]->vel, n, t1, t2);
    real v[10][4], w[10][4];
    for (int i = 0; i < 4; ++i) {
        real rho = g[i]->rho, p = g[i]->p, e = g[i]->e;
        real velx = vel[i].x, vely = vel[i].y, velz = vel[i].z;
        v[0][i] = rho;             w[0][i] = velx;
        v[1][i] = velx * rho;      w[1][i] = velx;
        v[2][i] = vely * rho;      w[2][i] = velx;
        v[3][i] = velz * rho;      w[3][i] = velx;
        v[4][i] = e * rho;         w[4][i] = velx;
        v[5][i] = velx * velx * rho; w[5][i] = velx;
        v[6][i] = vely * vely * rho; w[6][i] = velx;
        v[7][i] = velz * velz * rho; w[7][i] = velx;
        v[8][i] = p;               w[8][i] = velx;
        v[9][i] = p;               w[9][i] = 1;
    }
    real f_c[10], f_e[10];
    for (int j = 0; j < 10; ++j) {
        f_c[j] = ONE_TWELFTH * (-v[j][0] * w[j][0] + seven * v[j][1] * w[j][1]
                                + seven * v[j][2] * w[j][2] - v[j][3] * w[j][3]);
        f_e[j] = ONE_TWELFTH * (-v[j][0] * w[j][2] - v[j][2] * w[j][0] + eight * v[j][1] * w[j][2]
                                + eight * v[j][2] * w[j][1] - v[j][1] * w[j][3]
                                - v[j][3] * w[j][1]);
    }
    const real alpha_mass = 1.0, alpha_mom = 0.5, alpha_ie = 0.5, alpha_ke = 0.0, alpha_p = 0.0;
    real mass_flux = alpha_mass * f_c[0] + (one - alpha_mass) * f_e[0];
    F[0] = mass_flux;
    real mom_x = alpha_mom * f_c[1] + (one - alpha_mom) * f_e[1]
                   + (alpha_p * f_c[9] + (one - alpha_p) * f_e[9]);
    real mom_y = alpha_mom * f_c[2] + (one - alpha_mom) * f_e[2];
    real mom_z = alpha_mom * f_c[3] + (one - alpha_mom) * f_e[3];
    V3 mom = {mom_x, mom_y, mom_z};
    mom = to_global(mom, n, t1, t2);
    F[1] = mom.x; F[2] = mom.y; F[3] = mom.z;
    F[4] = alpha_ie * f_c[4] + (one - alpha_ie) * f_e[4]
           + half * (alpha_ke * f_c[5] + (one - alpha_ke) * f_e[5]
                     + alpha_ke * f_c[6] + (one - alpha_ke) * f_e[6]
                     + alpha_ke * f_c[7] + (one - alpha_ke) * f_e[7])
           + alpha_p * f_c[8] + (one - alpha_p) * f_e[8];
    F[5] = mass_flux * ((mass_flux >= zero) ? g[1]->YB : g[2]->YB);
}

/* face.cu:347-378 */
static void face_flux(Face* f, const FS* L1, const FS* L0, const FS* R0, const FS* R1,
                      int flux_calc, int x_order)
{
    if (flux_calc == 0) {
        FS L = *L0, R = *R0;
        if (x_order > 1) {
            interp_l2r2(L1->rho, L0->rho, R0->rho, R1->rho, &L.rho, &R.rho);
            interp_l2r2(L1->e, L0->e, R0->e, R1->e, &L.e, &R.e);
            interp_l2r2(L1->YB, L0->YB, R0->YB, R1->YB, &L.YB, &R.YB);
            update_from_rhoe(&L);
            update_from_rhoe(&R);
            interp_l2r2(L1->vel.x, L0->vel.x, R0->vel.x, R1->vel.x, &L.vel.x, &R.vel.x);
            interp_l2r2(L1->vel.y, L0->vel.y, R0->vel.y, R1->vel.y, &L.vel.y, &R.vel.y);
            interp_l2r2(L1->vel.z, L0->vel.z, R0->vel.z, R1->vel.z, &L.vel.z, &R.vel.z);
        }
        ausmdv(&L, &R, f->n, f->t1, f->t2, f->F);
        set_as_average(&f->fs, &L, &R);
    } else if (flux_calc == 1) {
        sbp_asf(L1, L0, R0, R1, f->n, f->t1, f->t2, f->F);
        set_as_average(&f->fs, L0, R0);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Block indexing and configuration (config.cu:152-225, block.cu:194-564).               */

enum { IM = 0, IP = 1, JM = 2, JP = 3, KM = 4, KP = 5 };

static int cidx(const Blk* b, int i, int j, int k) { return k * b->nic * b->njc + j * b->nic + i; }
static int ghost(const Blk* b, int f, int i0, int i1, int depth) /* config.cu:195-201 */
{
    return b->first_ghost[f] + b->n0c[f] * b->n1c[f] * depth + i1 * b->n0c[f] + i0;
}
static int ifi(const Blk* b, int i, int j, int k) { return i * b->njc * b->nkc + k * b->njc + j; }
static int jfi(const Blk* b, int i, int j, int k) { return b->nif + j * b->nic * b->nkc + k * b->nic + i; }
static int kfi(const Blk* b, int i, int j, int k) { return b->nif + b->njf + k * b->nic * b->njc + j * b->nic + i; }
static int vidx(const Blk* b, int i, int j, int k) { return k * (b->nic + 1) * (b->njc + 1) + j * (b->nic + 1) + i; }

/* Stencil cell for position m along a line of n cells: interior index, or ghost of the
 * minus/plus face at depth -m-1 / m-n.  This is the rule the explicit cases of
 * block.cu:272-316 (i), :373-417 (j), :474-518 (k) encode, including n == 2. */
static int line_cell(const Blk* b, int dir, int m, int a, int c)
{
    int n = (dir == 0) ? b->nic : (dir == 1) ? b->njc : b->nkc;
    int fm = 2 * dir, fp = 2 * dir + 1;
    /* (a, c) are the transverse indices in the order ghostCellIndex expects (i0, i1). */
    if (m < 0) return ghost(b, fm, a, c, -m - 1);
    if (m >= n) return ghost(b, fp, a, c, m - n);
    if (dir == 0) return cidx(b, m, a, c);
    if (dir == 1) return cidx(b, a, m, c);
    return cidx(b, a, c, m);
}

static int is_wall(int code)
{
    return code == ORC_WALL_SLIP || code == ORC_WALL_NOSLIP_ADIABATIC || code == ORC_WALL_NOSLIP_FIXED_T;
}

/* Clouds of cells and faces for the viscous gradients and the face's boundary code:
 * block.cu:317-358 (i-faces), :418-459 (j-faces), :519-560 (k-faces).  A wall face uses
 * only its interior cell; another boundary face also its first ghost cell. */
/* TWall of boundary face t of a fixed-T side: the side's value, overwritten by the file's
 * per-face values when given (block.cu:566-611). */
static real side_twall(const Blk* b, int side, int t) { return b->twf[side] ? b->twf[side][t] : b->twall[side]; }

static void set_cloud(Blk* b, Face* f, int side_bc, int c_minus, int c_plus, const int* fl, int nfl)
{
    f->bcode = side_bc;
    f->bfun = 0;
    f->TWall = 0.0;  /* the caller sets it on boundary faces */
    f->nc = 0;
    if (c_minus >= 0) f->cc[f->nc++] = c_minus;
    if (c_plus >= 0) f->cc[f->nc++] = c_plus;
    f->nf = nfl;
    for (int q = 0; q < nfl; ++q) f->fcl[q] = fl[q];
}

static void configure_clouds(Blk* b)
{
    int nic = b->nic, njc = b->njc, nkc = b->nkc;
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i <= nic; ++i) {
                Face* f = &b->faces[ifi(b, i, j, k)];
                if (i == 0) {
                    int fl[4] = {jfi(b, i, j, k), jfi(b, i, j + 1, k), kfi(b, i, j, k), kfi(b, i, j, k + 1)};
                    int g = is_wall(b->bc[IM]) ? -1 : ghost(b, IM, j, k, 0);
                    set_cloud(b, f, b->bc[IM], g, cidx(b, i, j, k), fl, 4);
                    f->TWall = side_twall(b, IM, j + njc * k); f->bfun = b->bcfun[IM];
                } else if (i == nic) {
                    int fl[4] = {jfi(b, i - 1, j, k), jfi(b, i - 1, j + 1, k), kfi(b, i - 1, j, k), kfi(b, i - 1, j, k + 1)};
                    int g = is_wall(b->bc[IP]) ? -1 : ghost(b, IP, j, k, 0);
                    set_cloud(b, f, b->bc[IP], cidx(b, i - 1, j, k), g, fl, 4);
                    f->TWall = side_twall(b, IP, j + njc * k); f->bfun = b->bcfun[IP];
                } else {
                    int fl[8] = {jfi(b, i - 1, j, k), jfi(b, i - 1, j + 1, k), kfi(b, i - 1, j, k), kfi(b, i - 1, j, k + 1),
                                 jfi(b, i, j, k), jfi(b, i, j + 1, k), kfi(b, i, j, k), kfi(b, i, j, k + 1)};
                    set_cloud(b, f, -1, cidx(b, i - 1, j, k), cidx(b, i, j, k), fl, 8);
                }
            }
    for (int k = 0; k < nkc; ++k)
        for (int i = 0; i < nic; ++i)
            for (int j = 0; j <= njc; ++j) {
                Face* f = &b->faces[jfi(b, i, j, k)];
                if (j == 0) {
                    int fl[4] = {ifi(b, i, j, k), ifi(b, i + 1, j, k), kfi(b, i, j, k), kfi(b, i, j, k + 1)};
                    int g = is_wall(b->bc[JM]) ? -1 : ghost(b, JM, i, k, 0);
                    set_cloud(b, f, b->bc[JM], g, cidx(b, i, j, k), fl, 4);
                    f->TWall = side_twall(b, JM, i + nic * k); f->bfun = b->bcfun[JM];
                } else if (j == njc) {
                    int fl[4] = {ifi(b, i, j - 1, k), ifi(b, i + 1, j - 1, k), kfi(b, i, j - 1, k), kfi(b, i, j - 1, k + 1)};
                    int g = is_wall(b->bc[JP]) ? -1 : ghost(b, JP, i, k, 0);
                    set_cloud(b, f, b->bc[JP], cidx(b, i, j - 1, k), g, fl, 4);
                    f->TWall = side_twall(b, JP, i + nic * k); f->bfun = b->bcfun[JP];
                } else {
                    int fl[8] = {ifi(b, i, j - 1, k), ifi(b, i + 1, j - 1, k), kfi(b, i, j - 1, k), kfi(b, i, j - 1, k + 1),
                                 ifi(b, i, j, k), ifi(b, i + 1, j, k), kfi(b, i, j, k), kfi(b, i, j, k + 1)};
                    set_cloud(b, f, -1, cidx(b, i, j - 1, k), cidx(b, i, j, k), fl, 8);
                }
            }
    for (int j = 0; j < njc; ++j)
        for (int i = 0; i < nic; ++i)
            for (int k = 0; k <= nkc; ++k) {
                Face* f = &b->faces[kfi(b, i, j, k)];
                if (k == 0) {
                    int fl[4] = {ifi(b, i, j, k), ifi(b, i + 1, j, k), jfi(b, i, j, k), jfi(b, i, j + 1, k)};
                    int g = is_wall(b->bc[KM]) ? -1 : ghost(b, KM, i, j, 0);
                    set_cloud(b, f, b->bc[KM], g, cidx(b, i, j, k), fl, 4);
                    f->TWall = side_twall(b, KM, i + nic * j); f->bfun = b->bcfun[KM];
                } else if (k == nkc) {
                    int fl[4] = {ifi(b, i, j, k - 1), ifi(b, i + 1, j, k - 1), jfi(b, i, j, k - 1), jfi(b, i, j + 1, k - 1)};
                    int g = is_wall(b->bc[KP]) ? -1 : ghost(b, KP, i, j, 0);
                    set_cloud(b, f, b->bc[KP], cidx(b, i, j, k - 1), g, fl, 4);
                    f->TWall = side_twall(b, KP, i + nic * j); f->bfun = b->bcfun[KP];
                } else {
                    int fl[8] = {ifi(b, i, j, k - 1), ifi(b, i + 1, j, k - 1), jfi(b, i, j, k - 1), jfi(b, i, j + 1, k - 1),
                                 ifi(b, i, j, k), ifi(b, i + 1, j, k), jfi(b, i, j, k), jfi(b, i, j + 1, k)};
                    set_cloud(b, f, -1, cidx(b, i, j, k - 1), cidx(b, i, j, k), fl, 8);
                }
            }
}

static int configure(Blk* b) /* config.cu:152-185, block.cu:194-564 */
{
    int nic = b->nic, njc = b->njc, nkc = b->nkc;
    if (nic < 2) { set_err("Too few cells in the i-index direction."); return -1; }
    if (njc < 2) { set_err("Too few cells in the j-index direction."); return -1; }
    if (nkc < 2) { set_err("Too few cells in the k-index direction."); return -1; }
    b->ncells = nic * njc * nkc;
    b->n0c[IM] = njc; b->n1c[IM] = nkc; b->n0c[IP] = njc; b->n1c[IP] = nkc;
    b->n0c[JM] = nic; b->n1c[JM] = nkc; b->n0c[JP] = nic; b->n1c[JP] = nkc;
    b->n0c[KM] = nic; b->n1c[KM] = njc; b->n0c[KP] = nic; b->n1c[KP] = njc;
    int ng = 0;
    for (int f = 0; f < 6; ++f) {
        b->nghost[f] = b->active ? 2 * b->n0c[f] * b->n1c[f] : 0;
        b->first_ghost[f] = (f > 0) ? b->first_ghost[f - 1] + b->nghost[f - 1] : b->ncells;
        ng += b->nghost[f];
    }
    b->ntot = b->ncells + ng;
    b->nif = (nic + 1) * njc * nkc;
    b->njf = nic * (njc + 1) * nkc;
    b->nkf = nic * njc * (nkc + 1);
    b->nfaces = b->nif + b->njf + b->nkf;
    b->cells = (Cell*)calloc((size_t)b->ntot, sizeof(Cell));
    b->faces = (Face*)calloc((size_t)b->nfaces, sizeof(Face));
    b->vtx = (V3*)calloc((size_t)(nic + 1) * (njc + 1) * (nkc + 1), sizeof(V3));
    if (b->active) {
        b->Q = (real*)calloc((size_t)b->ncells * 2 * 6, sizeof(real));
        b->dQ = (real*)calloc((size_t)b->ncells * 3 * 6, sizeof(real));
    }
    if (!b->cells || !b->faces || !b->vtx || (b->active && (!b->Q || !b->dQ))) {
        set_err("allocation failed"); return -1;
    }
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i < nic; ++i) {
                Cell* c = &b->cells[cidx(b, i, j, k)];
                c->face[IM] = ifi(b, i, j, k); c->face[IP] = ifi(b, i + 1, j, k);
                c->face[JM] = jfi(b, i, j, k); c->face[JP] = jfi(b, i, j + 1, k);
                c->face[KM] = kfi(b, i, j, k); c->face[KP] = kfi(b, i, j, k + 1);
            }
    /* Only active blocks have ghost cells; an inactive block never evaluates fluxes. */
    if (!b->active) return 0;
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i <= nic; ++i) {
                Face* f = &b->faces[ifi(b, i, j, k)];
                f->L1 = line_cell(b, 0, i - 2, j, k); f->L0 = line_cell(b, 0, i - 1, j, k);
                f->R0 = line_cell(b, 0, i, j, k);     f->R1 = line_cell(b, 0, i + 1, j, k);
            }
    for (int k = 0; k < nkc; ++k)
        for (int i = 0; i < nic; ++i)
            for (int j = 0; j <= njc; ++j) {
                Face* f = &b->faces[jfi(b, i, j, k)];
                f->L1 = line_cell(b, 1, j - 2, i, k); f->L0 = line_cell(b, 1, j - 1, i, k);
                f->R0 = line_cell(b, 1, j, i, k);     f->R1 = line_cell(b, 1, j + 1, i, k);
            }
    for (int j = 0; j < njc; ++j)
        for (int i = 0; i < nic; ++i)
            for (int k = 0; k <= nkc; ++k) {
                Face* f = &b->faces[kfi(b, i, j, k)];
                f->L1 = line_cell(b, 2, k - 2, i, j); f->L0 = line_cell(b, 2, k - 1, i, j);
                f->R0 = line_cell(b, 2, k, i, j);     f->R1 = line_cell(b, 2, k + 1, i, j);
            }
    configure_clouds(b);
    return 0;
}

/* block.cu:633-783 */
static void compute_geometry(Blk* b)
{
    int nic = b->nic, njc = b->njc, nkc = b->nkc;
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i < nic; ++i) {
                V3 p[8];
                p[0] = b->vtx[vidx(b, i, j, k)];         p[1] = b->vtx[vidx(b, i + 1, j, k)];
                p[2] = b->vtx[vidx(b, i + 1, j + 1, k)]; p[3] = b->vtx[vidx(b, i, j + 1, k)];
                p[4] = b->vtx[vidx(b, i, j, k + 1)];     p[5] = b->vtx[vidx(b, i + 1, j, k + 1)];
                p[6] = b->vtx[vidx(b, i + 1, j + 1, k + 1)]; p[7] = b->vtx[vidx(b, i, j + 1, k + 1)];
                Cell* c = &b->cells[cidx(b, i, j, k)];
                hex_props(p, &c->pos, &c->vol, &c->iL, &c->jL, &c->kL);
            }
    /* Face vertex orders, block.cu:267-270 (i), :368-371 (j), :469-472 (k). */
    for (int k = 0; k < nkc; ++k)
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i <= nic; ++i) {
                Face* f = &b->faces[ifi(b, i, j, k)];
                quad_properties(b->vtx[vidx(b, i, j, k)], b->vtx[vidx(b, i, j + 1, k)],
                                b->vtx[vidx(b, i, j + 1, k + 1)], b->vtx[vidx(b, i, j, k + 1)],
                                &f->pos, &f->n, &f->t1, &f->t2, &f->area);
            }
    for (int k = 0; k < nkc; ++k)
        for (int i = 0; i < nic; ++i)
            for (int j = 0; j <= njc; ++j) {
                Face* f = &b->faces[jfi(b, i, j, k)];
                quad_properties(b->vtx[vidx(b, i + 1, j, k)], b->vtx[vidx(b, i, j, k)],
                                b->vtx[vidx(b, i, j, k + 1)], b->vtx[vidx(b, i + 1, j, k + 1)],
                                &f->pos, &f->n, &f->t1, &f->t2, &f->area);
            }
    for (int j = 0; j < njc; ++j)
        for (int i = 0; i < nic; ++i)
            for (int k = 0; k <= nkc; ++k) {
                Face* f = &b->faces[kfi(b, i, j, k)];
                quad_properties(b->vtx[vidx(b, i, j, k)], b->vtx[vidx(b, i + 1, j, k)],
                                b->vtx[vidx(b, i + 1, j + 1, k)], b->vtx[vidx(b, i, j + 1, k)],
                                &f->pos, &f->n, &f->t1, &f->t2, &f->area);
            }
    if (!b->active) return;
    /* Ghost positions and lengths, block.cu:656-780: g0 = f + d, g1 = f + 3d, d = f - c0. */
    for (int dir = 0; dir < 3; ++dir)
        for (int side = 0; side < 2; ++side) {
            int na = (dir == 0) ? njc : nic, nc = (dir == 2) ? njc : nkc;
            for (int c2 = 0; c2 < nc; ++c2)
                for (int a = 0; a < na; ++a) {
                    int fi;
                    if (dir == 0) fi = ifi(b, side ? nic : 0, a, c2);
                    else if (dir == 1) fi = jfi(b, a, side ? njc : 0, c2);
                    else fi = kfi(b, a, c2, side ? nkc : 0);
                    Face* f = &b->faces[fi];
                    Cell* c0 = &b->cells[side ? f->L0 : f->R0];
                    Cell* g0 = &b->cells[side ? f->R0 : f->L0];
                    Cell* g1 = &b->cells[side ? f->R1 : f->L1];
                    g0->iL = c0->iL; g0->jL = c0->jL; g0->kL = c0->kL;
                    V3 d = {f->pos.x - c0->pos.x, f->pos.y - c0->pos.y, f->pos.z - c0->pos.z};
                    g0->pos.x = f->pos.x + d.x; g0->pos.y = f->pos.y + d.y; g0->pos.z = f->pos.z + d.z;
                    g1->iL = c0->iL; g1->jL = c0->jL; g1->kL = c0->kL;
                    d.x *= three; d.y *= three; d.z *= three; /* Vector3::mul(number) */
                    g1->pos.x = f->pos.x + d.x; g1->pos.y = f->pos.y + d.y; g1->pos.z = f->pos.z + d.z;
                }
        }
}

/* ------------------------------------------------------------------------------------ */
/* Boundary conditions (bcs.cu:16-544).                                                   */

static int boundary_face(const Blk* b, int bf, int a, int c)
{
    switch (bf) {
    case IM: return ifi(b, 0, a, c);
    case IP: return ifi(b, b->nic, a, c);
    case JM: return jfi(b, a, 0, c);
    case JP: return jfi(b, a, b->njc, c);
    case KM: return kfi(b, a, c, 0);
    default: return kfi(b, a, c, b->nkc);
    }
}

static void apply_bc(orc_sim* s, int ib, int bf)
{
    Blk* b = &s->b[ib];
    int plus = bf & 1;
    int na = (bf < 2) ? b->njc : b->nic;
    int nc = (bf < 4) ? b->nkc : b->njc;
    const Blk* o = NULL;
    if (b->bc[bf] == ORC_EXCHANGE) { /* bcs.cu:135-254, with periodic wrap :147 etc. */
        int oi = b->bi, oj = b->bj, ok = b->bk;
        if (bf == IM) { oi = b->bi - 1; if (oi < 0) oi = s->nib - 1; }
        if (bf == IP) { oi = b->bi + 1; if (oi >= s->nib) oi = 0; }
        if (bf == JM) { oj = b->bj - 1; if (oj < 0) oj = s->njb - 1; }
        if (bf == JP) { oj = b->bj + 1; if (oj >= s->njb) oj = 0; }
        if (bf == KM) { ok = b->bk - 1; if (ok < 0) ok = s->nkb - 1; }
        if (bf == KP) { ok = b->bk + 1; if (ok >= s->nkb) ok = 0; }
        o = &s->b[s->blk_ids[(oi * s->njb + oj) * s->nkb + ok]];
    }
    for (int c = 0; c < nc; ++c)
        for (int a = 0; a < na; ++a) {
            Face* f = &b->faces[boundary_face(b, bf, a, c)];
            int g0 = plus ? f->R0 : f->L0, g1 = plus ? f->R1 : f->L1;
            const FS* in = &b->cells[plus ? f->L0 : f->R0].fs;
            switch (b->bc[bf]) {
            case ORC_WALL_SLIP: case ORC_WALL_NOSLIP_ADIABATIC: case ORC_WALL_NOSLIP_FIXED_T: {
                FS r = *in; /* bcs.cu:26-37 */
                V3 v = to_local(r.vel, f->n, f->t1, f->t2);
                v.x = -(v.x);
                r.vel = to_global(v, f->n, f->t1, f->t2);
                b->cells[g0].fs = r;
                b->cells[g1].fs = r;
                break;
            }
            case ORC_EXCHANGE: {
                /* Ghost depth d = the other block's interior cell at depth d from the shared
                 * face (other face's left/right cells, bcs.cu:155-158). */
                int m0, m1;
                if (bf == IM) { m0 = cidx(o, o->nic - 1, a, c); m1 = cidx(o, o->nic - 2, a, c); }
                else if (bf == IP) { m0 = cidx(o, 0, a, c); m1 = cidx(o, 1, a, c); }
                else if (bf == JM) { m0 = cidx(o, a, o->njc - 1, c); m1 = cidx(o, a, o->njc - 2, c); }
                else if (bf == JP) { m0 = cidx(o, a, 0, c); m1 = cidx(o, a, 1, c); }
                else if (bf == KM) { m0 = cidx(o, a, c, o->nkc - 1); m1 = cidx(o, a, c, o->nkc - 2); }
                else { m0 = cidx(o, a, c, 0); m1 = cidx(o, a, c, 1); }
                b->cells[g0].fs = o->cells[m0].fs;
                b->cells[g1].fs = o->cells[m1].fs;
                break;
            }
            case ORC_INFLOW: /* bcs.cu:257-319 */
                b->cells[g0].fs = s->inflow[b->bcfs[bf]];
                b->cells[g1].fs = s->inflow[b->bcfs[bf]];
                break;
            case ORC_INFLOW_FUNCTION: /* bcs.cu:355-432: the function at each ghost's position */
                b->cells[g0].fs = inflow_function(s, b->bcfun[bf], b->cells[g0].pos);
                b->cells[g1].fs = inflow_function(s, b->bcfun[bf], b->cells[g1].pos);
                break;
            case ORC_OUTFLOW: /* bcs.cu:435-503 */
                b->cells[g0].fs = *in;
                b->cells[g1].fs = *in;
                break;
            default:
                break;
            }
        }
}

/* bcs.cu:506-544 */
static void apply_all_bcs(orc_sim* s)
{
#pragma omp parallel for schedule(dynamic)
    for (int ib = 0; ib < s->nblocks; ++ib) {
        if (!s->b[ib].active) continue;
        for (int bf = 0; bf < 6; ++bf) apply_bc(s, ib, bf);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Cell-level operations (cell.cu, block.cu:978-1172).                                    */

/* cell.cu:155-164 */
static real local_dt(const Blk* b, const Cell* c, real cfl)
{
    V3 in = b->faces[c->face[IM]].n, jn = b->faces[c->face[JM]].n, kn = b->faces[c->face[KM]].n;
    real is = c->iL / (fabs(dot3(c->fs.vel, in)) + c->fs.a);
    real js = c->jL / (fabs(dot3(c->fs.vel, jn)) + c->fs.a);
    real ks = c->kL / (fabs(dot3(c->fs.vel, kn)) + c->fs.a);
    return cfl * fmin(fmin(is, js), ks);
}

/* cell.cu:186-217 */
static void eval_dudt(const Blk* b, const Cell* c, real* d)
{
    real vol_inv = one / c->vol;
    const Face* fim = &b->faces[c->face[IM]]; const Face* fip = &b->faces[c->face[IP]];
    const Face* fjm = &b->faces[c->face[JM]]; const Face* fjp = &b->faces[c->face[JP]];
    const Face* fkm = &b->faces[c->face[KM]]; const Face* fkp = &b->faces[c->face[KP]];
    for (int i = 0; i < 6; ++i) {
        real si = fim->area * fim->F[i] - fip->area * fip->F[i]
                    + fjm->area * fjm->F[i] - fjp->area * fjp->F[i]
                    + fkm->area * fkm->F[i] - fkp->area * fkp->F[i];
        d[i] = vol_inv * si;
    }
}

/* block.cu:1018-1028 */
static void block_fluxes(orc_sim* s, Blk* b)
{
    for (int fi = 0; fi < b->nfaces; ++fi) {
        Face* f = &b->faces[fi];
        face_flux(f, &b->cells[f->L1].fs, &b->cells[f->L0].fs, &b->cells[f->R0].fs,
                  &b->cells[f->R1].fs, s->flux_calc, s->x_order);
    }
}

/* block.cu:1090-1160; stage 4 is the corrector of the two-stage predictor-corrector
 * (src/puffin/cell.d:294-303: U1 = U0 + 0.5*dt*(dUdt0 + dUdt1)); its predictor is stage 1. */
static int block_update(Blk* b, int stage, real dt)
{
    int bad = 0, n = b->ncells;
    for (int i = 0; i < n; ++i) {
        Cell* c = &b->cells[i];
        real* d0 = &b->dQ[(size_t)i * 6];
        real* d1 = &b->dQ[((size_t)n + i) * 6];
        real* d2 = &b->dQ[((size_t)2 * n + i) * 6];
        real* U0 = &b->Q[(size_t)i * 6];
        real* U1 = &b->Q[((size_t)n + i) * 6];
        if (stage == 1) {
            eval_dudt(b, c, d0);
            for (int j = 0; j < 6; ++j) U1[j] = U0[j] + dt * d0[j];
        } else if (stage == 2) {
            eval_dudt(b, c, d1);
            for (int j = 0; j < 6; ++j) U1[j] = U0[j] + one_quarter * dt * (d0[j] + d1[j]);
        } else if (stage == 4) {
            eval_dudt(b, c, d1);
            for (int j = 0; j < 6; ++j) U1[j] = U0[j] + half * dt * (d0[j] + d1[j]);
        } else {
            eval_dudt(b, c, d2);
            for (int j = 0; j < 6; ++j)
                U1[j] = U0[j] + dt * (ONE_SIXTH * d0[j] + ONE_SIXTH * d1[j] + FOUR_SIXTH * d2[j]);
        }
        bad += decode(&c->fs, U1);
    }
    return bad;
}

/* rsla.cu:53-72: 3x3 inverse by cofactors; returns -1 when |det| <= very_small_value. */
static int m_inverse3(real c[3][3], real ci[3][3], real very_small_value)
{
    real det = c[0][0] * (c[1][1] * c[2][2] - c[1][2] * c[2][1])
                 - c[0][1] * (c[1][0] * c[2][2] - c[1][2] * c[2][0])
                 + c[0][2] * (c[1][0] * c[2][1] - c[1][1] * c[2][0]);
    if (fabs(det) <= very_small_value) return -1;
    real one_over_det = one / det;
    ci[0][0] = (c[1][1] * c[2][2] - c[1][2] * c[2][1]) * one_over_det;
    ci[0][1] = (c[0][2] * c[2][1] - c[0][1] * c[2][2]) * one_over_det;
    ci[0][2] = (c[0][1] * c[1][2] - c[0][2] * c[1][1]) * one_over_det;
    ci[1][0] = (c[1][2] * c[2][0] - c[1][0] * c[2][2]) * one_over_det;
    ci[1][1] = (c[0][0] * c[2][2] - c[0][2] * c[2][0]) * one_over_det;
    ci[1][2] = (c[0][2] * c[1][0] - c[0][0] * c[1][2]) * one_over_det;
    ci[2][0] = (c[1][0] * c[2][1] - c[1][1] * c[2][0]) * one_over_det;
    ci[2][1] = (c[0][1] * c[2][0] - c[0][0] * c[2][1]) * one_over_det;
    ci[2][2] = (c[0][0] * c[1][1] - c[0][1] * c[1][0]) * one_over_det;
    return 0;
}

/* block.cu:28-86: least-squares weights of the face's cloud, face centre as reference.
 * A singular normal matrix leaves the weights at zero (the reference reports it on stderr
 * and carries on: its failure counter is never incremented). */
static void setup_lsq(Blk* b, Face* f)
{
    V3 pos[10];
    int n = 0;
    for (int q = 0; q < f->nc; ++q) pos[n++] = b->cells[f->cc[q]].pos;
    for (int q = 0; q < f->nf; ++q) pos[n++] = b->faces[f->fcl[q]].pos;
    real w2[10], dx[10], dy[10], dz[10];
    real x0 = f->pos.x, y0 = f->pos.y, z0 = f->pos.z;
    for (int i = 0; i < n; ++i) {
        dx[i] = pos[i].x - x0;
        dy[i] = pos[i].y - y0;
        dz[i] = pos[i].z - z0;
        w2[i] = (real)(1.0 / (double)(dx[i] * dx[i] + dy[i] * dy[i] + dz[i] * dz[i])); /* block.cu:54, in double */
    }
    real xx = zero, xy = zero, xz = zero, yy = zero, yz = zero, zz = zero;
    for (int i = 0; i < n; ++i) {
        xx += w2[i] * dx[i] * dx[i];
        xy += w2[i] * dx[i] * dy[i];
        xz += w2[i] * dx[i] * dz[i];
        yy += w2[i] * dy[i] * dy[i];
        yz += w2[i] * dy[i] * dz[i];
        zz += w2[i] * dz[i] * dz[i];
    }
    real m[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}}, mi[3][3];
    for (int i = 0; i < 10; ++i) f->wx[i] = f->wy[i] = f->wz[i] = zero;
    if (m_inverse3(m, mi, (real)1.0e-16)) return;
    for (int i = 0; i < n; ++i) {
        f->wx[i] = w2[i] * (mi[0][0] * dx[i] + mi[0][1] * dy[i] + mi[0][2] * dz[i]);
        f->wy[i] = w2[i] * (mi[1][0] * dx[i] + mi[1][1] * dy[i] + mi[1][2] * dz[i]);
        f->wz[i] = w2[i] * (mi[2][0] * dx[i] + mi[2][1] * dy[i] + mi[2][2] * dz[i]);
    }
}

/* bcs.cu:777-797 (the inflow-function case is not built). */
static void viscous_bc(const orc_sim* s, Face* f)
{
    if (f->bcode == ORC_INFLOW_FUNCTION) { /* bcs.cu:790-792 */
        f->fs = inflow_function(s, f->bfun, f->pos);
        return;
    }
    if (f->bcode == ORC_WALL_NOSLIP_ADIABATIC) {
        f->fs.vel.x = 0.0; f->fs.vel.y = 0.0; f->fs.vel.z = 0.0;
    } else if (f->bcode == ORC_WALL_NOSLIP_FIXED_T) {
        f->fs.vel.x = 0.0; f->fs.vel.y = 0.0; f->fs.vel.z = 0.0;
        f->fs.T = f->TWall;
    }
}

/* gas.cu:95-100, Sutherland's law for air. */
static void trans_coeffs(real T, real* mu, real* k)
{
    const real mu_ref = (real)1.716e-5, T_ref = (real)273.0, S = (real)111.0, Prandtl = (real)0.72;
    const real Cp = GAS_CV + GAS_R;
    *mu = mu_ref * sqrt(T / T_ref) * (T / T_ref) * (T_ref + S) / (T + S);
    *k = Cp * *mu / Prandtl;
}

/* block.cu:88-162 */
static void add_viscous(Blk* b, Face* f)
{
    FS cl[10];
    int n = 0;
    for (int q = 0; q < f->nc; ++q) cl[n++] = b->cells[f->cc[q]].fs;
    for (int q = 0; q < f->nf; ++q) cl[n++] = b->faces[f->fcl[q]].fs;
    FS fs0 = f->fs;
    V3 gT = {zero, zero, zero}, gx = {zero, zero, zero}, gy = {zero, zero, zero}, gz = {zero, zero, zero};
    real dq = zero;
    for (int i = 0; i < n; ++i) {
        real wx = f->wx[i], wy = f->wy[i], wz = f->wz[i];
        dq = cl[i].T - fs0.T;
        gT.x += wx * dq; gT.y += wy * dq; gT.z += wz * dq;
        dq = cl[i].vel.x - fs0.vel.x;
        gx.x += wx * dq; gx.y += wy * dq; gx.z += wz * dq;
        dq = cl[i].vel.y - fs0.vel.y;
        gy.x += wx * dq; gy.y += wy * dq; gy.z += wz * dq;
        dq = cl[i].vel.z - fs0.vel.z;
        gz.x += wx * dq; gz.y += wy * dq; gz.z += wz * dq;
    }
    real mu, k;
    trans_coeffs(fs0.T, &mu, &k);
    real lmbda = -two / three * mu;
    real tau_xx = two * mu * gx.x + lmbda * (gx.x + gy.y + gz.z);
    real tau_yy = two * mu * gy.y + lmbda * (gx.x + gy.y + gz.z);
    real tau_zz = two * mu * gz.z + lmbda * (gx.x + gy.y + gz.z);
    real tau_xy = mu * (gx.y + gy.x);
    real tau_xz = mu * (gx.z + gz.x);
    real tau_yz = mu * (gy.z + gz.y);
    real qx = k * gT.x, qy = k * gT.y, qz = k * gT.z;
    V3 nn = f->n;
    f->F[1] -= tau_xx * nn.x + tau_xy * nn.y + tau_xz * nn.z;
    f->F[2] -= tau_xy * nn.x + tau_yy * nn.y + tau_yz * nn.z;
    f->F[3] -= tau_xz * nn.x + tau_yz * nn.y + tau_zz * nn.z;
    f->F[4] -= (tau_xx * fs0.vel.x + tau_xy * fs0.vel.y + tau_xz * fs0.vel.z + qx) * nn.x
               + (tau_xy * fs0.vel.x + tau_yy * fs0.vel.y + tau_yz * fs0.vel.z + qy) * nn.y
               + (tau_xz * fs0.vel.x + tau_yz * fs0.vel.y + tau_zz * fs0.vel.z + qz) * nn.z;
}

/* config.cu:263-269 */
static real schedule_value(const orc_sim* s, real t)
{
    int i = s->n_cfl - 1;
    while (i > 0 && t < s->cfl_t[i]) i--;
    return s->cfl_v[i];
}

/* ------------------------------------------------------------------------------------ */
/* Public API.                                                                            */

orc_sim* orc_create(int nblocks, int nib, int njb, int nkb, const int* blk_ids,
                    const int* blk_ijk, const int* blk_dims, const int* active,
                    const int* bc_codes, const int* bc_fs, int n_flow_states,
                    const double* flow_states, const double* vertices, const double* iovars,
                    const orc_params* p)
{
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    if (!s) { set_err("allocation failed"); return NULL; }
    s->nblocks = nblocks; s->nib = nib; s->njb = njb; s->nkb = nkb;
    s->blk_ids = (int*)malloc(sizeof(int) * (size_t)nib * njb * nkb);
    memcpy(s->blk_ids, blk_ids, sizeof(int) * (size_t)nib * njb * nkb);
    s->nfs = n_flow_states;
    s->inflow = (FS*)calloc((size_t)(n_flow_states > 0 ? n_flow_states : 1), sizeof(FS));
    for (int i = 0; i < n_flow_states; ++i) { /* config.cu:398-406 */
        FS* f = &s->inflow[i];
        f->p = (real)flow_states[6 * i + 0]; f->T = (real)flow_states[6 * i + 1];
        f->YB = (real)flow_states[6 * i + 2];
        update_from_pT(f);
        f->vel.x = (real)flow_states[6 * i + 3]; f->vel.y = (real)flow_states[6 * i + 4];
        f->vel.z = (real)flow_states[6 * i + 5];
    }
    s->x_order = p->x_order; s->flux_calc = p->flux_calc; s->cfl_count = p->cfl_count;
    s->viscous = p->viscous;
    s->reacting = p->reacting;
    s->time_scheme = p->time_scheme;
    /* Configuration values are held in the number type (config.cu:494-504). */
    s->max_step = p->max_step; s->max_time = (real)p->max_time; s->dt_init = (real)p->dt_init;
    s->n_cfl = p->n_cfl;
    s->cfl_t = (real*)malloc(sizeof(real) * (size_t)p->n_cfl);
    s->cfl_v = (real*)malloc(sizeof(real) * (size_t)p->n_cfl);
    for (int i = 0; i < p->n_cfl; ++i) { s->cfl_t[i] = (real)p->cfl_times[i]; s->cfl_v[i] = (real)p->cfl_values[i]; }
    if (p->bl_profile) {
        real prof[1 + 3 * (ORC_BL_N + 1)];
        for (int i = 0; i < 1 + 3 * (ORC_BL_N + 1); ++i) prof[i] = (real)p->bl_profile[i];
        if (bl_set(&s->bl, prof)) {
            set_err("Gauss-Jordan elimination failed while solving for sigma."); orc_destroy(s); return NULL;
        }
    }
    s->b = (Blk*)calloc((size_t)nblocks, sizeof(Blk));
    size_t voff = 0, ooff = 0;
    int twall_used = 0;
    for (int ib = 0; ib < nblocks; ++ib) {
        Blk* b = &s->b[ib];
        b->bi = blk_ijk[3 * ib]; b->bj = blk_ijk[3 * ib + 1]; b->bk = blk_ijk[3 * ib + 2];
        b->nic = blk_dims[3 * ib]; b->njc = blk_dims[3 * ib + 1]; b->nkc = blk_dims[3 * ib + 2];
        b->active = active[ib];
        for (int f = 0; f < 6; ++f) {
            b->bc[f] = bc_codes[6 * ib + f]; b->bcfs[f] = bc_fs[6 * ib + f];
            b->twall[f] = p->bc_twall ? (real)p->bc_twall[6 * ib + f] : zero;
            b->bcfun[f] = p->bc_fun ? p->bc_fun[6 * ib + f] : 0;
            if (b->bc[f] == ORC_WALL_NOSLIP_FIXED_T && isnan(b->twall[f])) {
                int nt = f < 2 ? b->njc * b->nkc : (f < 4 ? b->nic * b->nkc : b->nic * b->njc);
                if (!p->twall_faces || twall_used + nt > p->n_twall_faces) {
                    set_err("TWall per face: twall_faces holds too few values"); orc_destroy(s); return NULL;
                }
                b->twf[f] = (real*)malloc(sizeof(real) * (size_t)nt);
                for (int q = 0; q < nt; ++q) b->twf[f][q] = (real)p->twall_faces[twall_used + q];
                twall_used += nt;
            }
            if (b->active && b->bc[f] == ORC_INFLOW_FUNCTION && b->bcfun[f] == 2 && !s->bl.set) {
                set_err("Boundary-layer file is not present in current directory."); orc_destroy(s); return NULL;
            }
        }
        if (configure(b)) { orc_destroy(s); return NULL; }
        size_t nv = (size_t)(b->nic + 1) * (b->njc + 1) * (b->nkc + 1);
        for (size_t v = 0; v < nv; ++v) {
            b->vtx[v].x = (real)vertices[voff + 3 * v]; b->vtx[v].y = (real)vertices[voff + 3 * v + 1];
            b->vtx[v].z = (real)vertices[voff + 3 * v + 2];
        }
        voff += 3 * nv;
        /* readFlow, block.cu:882-907: 13 IO variables, cell order i fastest. */
        int n = b->ncells;
        for (int c = 0; c < n; ++c) {
            Cell* cl = &b->cells[c];
            const double* io = iovars + ooff;
            cl->pos.x = (real)io[0 * n + c]; cl->pos.y = (real)io[1 * n + c]; cl->pos.z = (real)io[2 * n + c];
            cl->vol = (real)io[3 * n + c];
            cl->fs.p = (real)io[4 * n + c]; cl->fs.T = (real)io[5 * n + c]; cl->fs.rho = (real)io[6 * n + c];
            cl->fs.e = (real)io[7 * n + c]; cl->fs.YB = (real)io[8 * n + c]; cl->fs.a = (real)io[9 * n + c];
            cl->fs.vel.x = (real)io[10 * n + c]; cl->fs.vel.y = (real)io[11 * n + c];
            cl->fs.vel.z = (real)io[12 * n + c];
        }
        ooff += 13 * (size_t)n;
        compute_geometry(b);
        if (s->viscous && b->active) /* simulate.cu:201 */
            for (int fi = 0; fi < b->nfaces; ++fi) setup_lsq(b, &b->faces[fi]);
    }
    /* Exchange partners must exist (a missing block would be index -1 in the reference). */
    for (int ib = 0; ib < nblocks; ++ib)
        for (int f = 0; f < 6; ++f)
            if (s->b[ib].active && s->b[ib].bc[f] == ORC_EXCHANGE) {
                Blk* b = &s->b[ib];
                int oi = b->bi + (f == IP) - (f == IM), oj = b->bj + (f == JP) - (f == JM);
                int ok = b->bk + (f == KP) - (f == KM);
                oi = (oi + nib) % nib; oj = (oj + njb) % njb; ok = (ok + nkb) % nkb;
                int id = s->blk_ids[(oi * njb + oj) * nkb + ok];
                if (id < 0 || id >= nblocks) { set_err("exchange boundary without a neighbour block"); orc_destroy(s); return NULL; }
                Blk* o = &s->b[id];
                int ok_dims = (f < 2) ? (o->njc == b->njc && o->nkc == b->nkc)
                              : (f < 4) ? (o->nic == b->nic && o->nkc == b->nkc)
                                        : (o->nic == b->nic && o->njc == b->njc);
                if (!ok_dims) { set_err("exchange neighbour has mismatched face dimensions"); orc_destroy(s); return NULL; }
            }
    s->t = (real)p->t0; s->dt = s->dt_init; s->step = 0;
    /* simulate.cu:196-203: encode level 0 from the file's flow states. */
    for (int ib = 0; ib < nblocks; ++ib) {
        Blk* b = &s->b[ib];
        if (!b->active) continue;
        for (int c = 0; c < b->ncells; ++c) encode(&b->cells[c].fs, &b->Q[(size_t)c * 6]);
    }
    return s;
}

void orc_destroy(orc_sim* s)
{
    if (!s) return;
    if (s->b)
        for (int ib = 0; ib < s->nblocks; ++ib) {
            free(s->b[ib].cells); free(s->b[ib].faces); free(s->b[ib].vtx);
            free(s->b[ib].Q); free(s->b[ib].dQ);
            for (int f = 0; f < 6; ++f) free(s->b[ib].twf[f]);
        }
    free(s->b); free(s->blk_ids); free(s->inflow); free(s->cfl_t); free(s->cfl_v);
    free(s);
}

/* simulate.cu:205-335 (without plotting; the caller writes snapshots). */
int orc_march(orc_sim* s, int nsteps, double* dt_log, int* bad_count)
{
    int done = 0;
    while (done < nsteps && s->step < s->max_step && s->t < s->max_time) {
        if (s->step > 0 && (s->step % s->cfl_count) == 0) { /* :208-221 */
            real smallest = REAL_MAX;
            real cfl = schedule_value(s, s->t);
            for (int ib = 0; ib < s->nblocks; ++ib) {
                Blk* b = &s->b[ib];
                if (!b->active) continue;
                real bmin = REAL_MAX; /* block.cu:978-990 */
                for (int c = 0; c < b->ncells; ++c) bmin = fmin(bmin, local_dt(b, &b->cells[c], cfl));
                smallest = fmin(smallest, bmin);
            }
            s->dt = fmin((real)1.5 * s->dt, smallest);
        }
        const int nstages = s->time_scheme == 1 ? 2 : 3;
        for (int q = 1; q <= nstages; ++q) { /* :223-287 */
            const int stage = (s->time_scheme == 1 && q == 2) ? 4 : q;
            apply_all_bcs(s);
            int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : bad)
            for (int ib = 0; ib < s->nblocks; ++ib) {
                Blk* b = &s->b[ib];
                if (!b->active) continue;
                block_fluxes(s, b);
                if (s->viscous) { /* simulate.cu:235-238 */
                    for (int fi = 0; fi < b->nfaces; ++fi) viscous_bc(s, &b->faces[fi]);
                    for (int fi = 0; fi < b->nfaces; ++fi) add_viscous(b, &b->faces[fi]);
                }
                bad += block_update(b, stage, s->dt);
            }
            if (bad > 0) {
                if (bad_count) *bad_count = bad;
                return 100 * q + (bad > 99 ? 99 : bad);
            }
        }
        for (int ib = 0; ib < s->nblocks; ++ib) { /* :288-294, copy level 1 -> 0 */
            Blk* b = &s->b[ib];
            if (!b->active) continue;
            memcpy(b->Q, b->Q + (size_t)b->ncells * 6, sizeof(real) * (size_t)b->ncells * 6);
        }
        if (s->reacting) /* :296-303, block.cu:1174-1184 */
            for (int ib = 0; ib < s->nblocks; ++ib) {
                Blk* b = &s->b[ib];
                if (!b->active) continue;
                for (int c = 0; c < b->ncells; ++c) {
                    update_chemistry(&b->cells[c].fs, s->dt);
                    encode(&b->cells[c].fs, &b->Q[(size_t)c * 6]);
                }
            }
        if (dt_log) dt_log[done] = (double)s->dt;
        s->t += s->dt;
        s->step += 1;
        done += 1;
    }
    if (bad_count) *bad_count = 0;
    return 0;
}

int orc_step_count(const orc_sim* s) { return s->step; }
double orc_time(const orc_sim* s) { return (double)s->t; }
double orc_dt(const orc_sim* s) { return (double)s->dt; }

/* cell.cu:133-153 */
void orc_get_iovars(const orc_sim* s, int blk, double* out)
{
    const Blk* b = &s->b[blk];
    int n = b->ncells;
    for (int c = 0; c < n; ++c) {
        const Cell* cl = &b->cells[c];
        out[0 * n + c] = cl->pos.x; out[1 * n + c] = cl->pos.y; out[2 * n + c] = cl->pos.z;
        out[3 * n + c] = cl->vol;
        out[4 * n + c] = cl->fs.p; out[5 * n + c] = cl->fs.T; out[6 * n + c] = cl->fs.rho;
        out[7 * n + c] = cl->fs.e; out[8 * n + c] = cl->fs.YB; out[9 * n + c] = cl->fs.a;
        out[10 * n + c] = cl->fs.vel.x; out[11 * n + c] = cl->fs.vel.y; out[12 * n + c] = cl->fs.vel.z;
    }
}

void orc_get_conserved(const orc_sim* s, int blk, double* out)
{
    const Blk* b = &s->b[blk];
    int n = b->ncells;
    if (!b->active) { memset(out, 0, sizeof(double) * 6 * (size_t)n); return; }
    for (int c = 0; c < n; ++c)
        for (int j = 0; j < 6; ++j) out[(size_t)j * n + c] = b->Q[(size_t)c * 6 + j];
}

void orc_get_dudt(const orc_sim* s, int blk, int level, double* out)
{
    const Blk* b = &s->b[blk];
    int n = b->ncells;
    if (!b->active) { memset(out, 0, sizeof(double) * 6 * (size_t)n); return; }
    for (int c = 0; c < n; ++c)
        for (int j = 0; j < 6; ++j) out[(size_t)j * n + c] = b->dQ[((size_t)level * n + c) * 6 + j];
}

static FS fs_from9(const double* w)
{
    FS f;
    f.rho = (real)w[0]; f.e = (real)w[1]; f.p = (real)w[2]; f.T = (real)w[3]; f.YB = (real)w[4];
    f.a = (real)w[5];
    f.vel.x = (real)w[6]; f.vel.y = (real)w[7]; f.vel.z = (real)w[8];
    return f;
}

void orc_face_flux(const double* states, const double* frame, int x_order, int flux_calc, double* F)
{
    Face f;
    memset(&f, 0, sizeof f);
    f.n.x = (real)frame[0]; f.n.y = (real)frame[1]; f.n.z = (real)frame[2];
    f.t1.x = (real)frame[3]; f.t1.y = (real)frame[4]; f.t1.z = (real)frame[5];
    f.t2.x = (real)frame[6]; f.t2.y = (real)frame[7]; f.t2.z = (real)frame[8];
    FS s0 = fs_from9(states), s1 = fs_from9(states + 9), s2 = fs_from9(states + 18),
       s3 = fs_from9(states + 27);
    face_flux(&f, &s0, &s1, &s2, &s3, flux_calc, x_order);
    for (int i = 0; i < 6; ++i) F[i] = f.F[i];
}

void orc_hex_props(const double* p, double* out)
{
    V3 v[8];
    for (int i = 0; i < 8; ++i) { v[i].x = (real)p[3 * i]; v[i].y = (real)p[3 * i + 1]; v[i].z = (real)p[3 * i + 2]; }
    V3 c; real vol, iL, jL, kL;
    hex_props(v, &c, &vol, &iL, &jL, &kL);
    out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = vol; out[4] = iL; out[5] = jL; out[6] = kL;
}

void orc_quad_props(const double* p, double* out)
{
    V3 v[4];
    for (int i = 0; i < 4; ++i) { v[i].x = (real)p[3 * i]; v[i].y = (real)p[3 * i + 1]; v[i].z = (real)p[3 * i + 2]; }
    V3 c, n, t1, t2; real area;
    quad_properties(v[0], v[1], v[2], v[3], &c, &n, &t1, &t2, &area);
    out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = n.x; out[4] = n.y; out[5] = n.z;
    out[6] = t1.x; out[7] = t1.y; out[8] = t1.z; out[9] = t2.x; out[10] = t2.y; out[11] = t2.z;
    out[12] = area;
}

int orc_decode(const double* U, double* W)
{
    FS f;
    memset(&f, 0, sizeof f);
    real u[6];
    for (int i = 0; i < 6; ++i) u[i] = (real)U[i];
    int bad = decode(&f, u);
    W[0] = f.rho; W[1] = f.e; W[2] = f.p; W[3] = f.T; W[4] = f.YB; W[5] = f.a;
    W[6] = f.vel.x; W[7] = f.vel.y; W[8] = f.vel.z;
    return bad;
}

void orc_encode(const double* W, double* U)
{
    FS f = fs_from9(W);
    real u[6];
    encode(&f, u);
    for (int i = 0; i < 6; ++i) U[i] = (double)u[i];
}
