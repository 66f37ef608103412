/* chkn_oracle.h -- CPU restatement of chicken's explicit time step (test infrastructure).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it, and only as the
 * checker of the HIP path.  The product (libchkn) never links or calls it.
 *
 * The restatement follows the reference's CPU flavour (march_in_time_using_cpu_only,
 * src/chicken/simulate.cu:182-342) operation for operation, in the reference's number type
 * (FP64, or FP32 as its FLOAT_NUMBERS build), so that a build without FMA contraction
 * reproduces the reference's arithmetic bit for bit.  The interface below is double
 * throughout; the FP32 builds round inputs to float and widen outputs exactly.
 * Each function in chkn_oracle.c cites the reference file:line it restates.
 */
#ifndef CHKN_ORACLE_H
#define CHKN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Built four times: liborc.so for ideal air and liborc_ab.so (ORC_AB_REACTING_GAS) for the
 * reacting A->B gas, as the reference selects its gas model at build time (gas.cu:20-48),
 * and their FP32 twins liborc_f32.so and liborc_ab_f32.so (ORC_FP32, number.cu:10-27),
 * which have no complex-step probe. */

/* Boundary-condition codes, face.cu:22-36. */
enum { ORC_WALL_SLIP = 0, ORC_WALL_NOSLIP_ADIABATIC = 1, ORC_WALL_NOSLIP_FIXED_T = 2,
       ORC_EXCHANGE = 3, ORC_INFLOW = 4, ORC_OUTFLOW = 5, ORC_INFLOW_FUNCTION = 6 };

/* Simulation parameters (config.cu:312-344). */
typedef struct {
    int x_order;          /* 1 or 2 */
    int flux_calc;        /* 0 = ausmdv, 1 = sbp_asf */
    int cfl_count;
    int max_step;
    double max_time;
    double dt_init;
    int n_cfl;            /* cfl schedule, config.cu:250-269 */
    const double* cfl_times;
    const double* cfl_values;
    double t0;            /* starting time (times.data) */
    int viscous;          /* add viscous fluxes, block.cu:88-162 */
    const double* bc_twall; /* 6 per block, TWall of fixed-T walls (may be NULL); NaN: per face */
    const int* bc_fun;    /* 6 per block, function of inflow_function sides (may be NULL) */
    const double* bl_profile; /* p_e, then (y, T, velx) x 12 (config.cu:37-60), or NULL */
    int n_twall_faces;    /* TWall per face of the NaN sides, block and side order, each side */
    const double* twall_faces; /* in the TWall-file order (block.cu:566-611) */
    int reacting;         /* the A->B reaction after every step (simulate.cu:296-303) */
    int time_scheme;      /* 0 = TVD-RK3 (simulate.cu:223-287); 1 = two-stage predictor-
                             corrector (Lorikeet, src/puffin/cell.d:294-303): not part of
                             chicken, restated from the sibling code */
} orc_params;

typedef struct orc_sim orc_sim;

/* Build a simulation from flattened arrays.
 *  blk_ids   : nib*njb*nkb ints, index [(i*njb + j)*nkb + k] (config.json "blk_ids"[i][j][k])
 *  blk_ijk   : 3 ints per block (position in block array)
 *  blk_dims  : 3 ints per block (nic, njc, nkc)
 *  active    : 1 int per block
 *  bc_codes  : 6 ints per block, face order iminus..kplus
 *  bc_fs     : 6 ints per block, inflow flow-state index
 *  flow_states: 6 doubles per state (p, T, YB, velx, vely, velz); rho, e, a are derived
 *               with update_from_pT exactly as read_config_file does (config.cu:396-406)
 *  vertices  : per block (nic+1)(njc+1)(nkc+1)*3 doubles, xyz interleaved, i fastest
 *  iovars    : per block 13*ncells doubles in flow-file order (cell.cu:66-68)
 * Returns NULL on error (message via orc_last_error()). */
orc_sim* orc_create(int nblocks, int nib, int njb, int nkb, const int* blk_ids,
                    const int* blk_ijk, const int* blk_dims, const int* active,
                    const int* bc_codes, const int* bc_fs,
                    int n_flow_states, const double* flow_states,
                    const double* vertices, const double* iovars,
                    const orc_params* params);
void orc_destroy(orc_sim* s);
const char* orc_last_error(void);

/* Run the march loop (simulate.cu:205-335) for at most nsteps more steps, stopping at
 * max_step/max_time as the reference does.  Records dt per step into dt_log if not NULL
 * (length >= nsteps).  Returns 0 on success, or 100*stage + min(count,99) on the first
 * stage with bad cells (the reference throws "Stage k bad cell count: n").  The count
 * itself is returned through *bad_count if not NULL. */
int orc_march(orc_sim* s, int nsteps, double* dt_log, int* bad_count);

int orc_step_count(const orc_sim* s);
double orc_time(const orc_sim* s);
double orc_dt(const orc_sim* s);
/* 13*ncells doubles of the block's IO variables, flow-file order. */
void orc_get_iovars(const orc_sim* s, int blk, double* out);
/* Conserved quantities, level 0: 6*ncells doubles, quantity-major. */
void orc_get_conserved(const orc_sim* s, int blk, double* out);
/* Residual dU/dt of the last stage-1 evaluation: 6*ncells doubles, quantity-major. */
void orc_get_dudt(const orc_sim* s, int blk, int level, double* out);

/* Pointwise probes. */
/* states: 4 flow states (L1, L0, R0, R1), each 9 doubles (rho, e, p, T, YB, a, vx, vy, vz)
 * frame : 9 doubles (n, t1, t2).  F: 6 doubles. */
void orc_face_flux(const double* states, const double* frame, int x_order, int flux_calc,
                   double* F);
/* p: 8 vertices xyz.  out: pos(3), volume, iLen, jLen, kLen. */
void orc_hex_props(const double* p, double* out);
/* p: 4 vertices xyz.  out: centroid(3), n(3), t1(3), t2(3), area. */
void orc_quad_props(const double* p, double* out);
/* U: 6 doubles -> W: 9 doubles (rho, e, p, T, YB, a, vx, vy, vz); returns bad flag. */
int orc_decode(const double* U, double* W);
void orc_encode(const double* W, double* U);

/* Complex-step face flux (chkn_oracle_cs.c).  prims: 4 states (L1, L0, R0, R1) of
 * (rho, e, YB, vx, vy, vz); frame: n, t1, t2; dir: 24 doubles.  Fre = Re F(x + i h dir),
 * jv = Im F(x + i h dir) / h, the directional derivative of the flux. */
void orc_face_flux_cs(const double* prims, const double* frame, int x_order, const double* dir,
                      double h, double* Fre, double* jv);

#ifdef __cplusplus
}
#endif
#endif
