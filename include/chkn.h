/* chkn.h -- C-ABI of libchkn, the MI355X (gfx950) explicit finite-volume step of chicken.
 *
 * The reference has no plugin or FFI surface for its time step: the step is chosen at
 * compile time between march_in_time_using_cpu_only and march_in_time_using_gpu
 * (src/chicken/main.cu:71-80) and runs between initialize_simulation and
 * finalize_simulation on process-global state (src/chicken/simulate.cu:32-46).  The entry
 * points below are that seam as a library: each cites the reference function it replaces.
 * Plain pointers and sizes only; all device memory is owned by the handle.
 *
 * Return codes: 0 = success; > 0 = number of bad cells found in the first failing stage
 * (the reference throws "Stage k bad cell count: n", simulate.cu:243-245, :264-266,
 * :285-287; the same text is available from chkn_last_error); < 0 = error class below.
 */
#ifndef CHKN_H
#define CHKN_H

#ifdef __cplusplus
extern "C" {
#endif

#define CHKN_OK 0
#define CHKN_ERR_ARG (-1)     /* invalid argument or unsupported configuration */
#define CHKN_ERR_IO (-2)      /* job file missing or malformed */
#define CHKN_ERR_CONFIG (-3)  /* inconsistent configuration, e.g. gas-model mismatch */
#define CHKN_ERR_HIP (-4)     /* HIP runtime error */
#define CHKN_ERR_STATE (-5)   /* call not valid in the handle's current state */

/* Boundary-condition codes, identical to src/chicken/face.cu:26-32. */
#define CHKN_BC_WALL_WITH_SLIP 0
#define CHKN_BC_WALL_NO_SLIP_ADIABATIC 1
#define CHKN_BC_WALL_NO_SLIP_FIXED_T 2
#define CHKN_BC_EXCHANGE 3
#define CHKN_BC_INFLOW 4
#define CHKN_BC_OUTFLOW 5
#define CHKN_BC_INFLOW_FUNCTION 6

/* Number of IO variables per cell in flow files, src/chicken/cell.cu:66-84. */
#define CHKN_N_IOVARS 13

typedef struct chkn_sim chkn_sim;

/* Array form of a job: what config.json plus the grid and flow files hold
 * (src/chicken/config.cu:347-525, block.cu:785-930).  Block data are concatenated in block
 * order.  flow_states holds (p, T, YB, velx, vely, velz) per state; rho, e and a follow
 * from update_from_pT exactly as in config.cu:398-406. */
typedef struct {
    int nblocks, nib, njb, nkb;
    const int* blk_ids;      /* nib*njb*nkb, index (i*njb + j)*nkb + k; -1 = no block */
    const int* blk_ijk;      /* 3 per block */
    const int* blk_dims;     /* 3 per block: nic, njc, nkc */
    const int* active;       /* 1 per block */
    const int* bc_codes;     /* 6 per block, face order iminus, iplus, jminus, jplus, kminus, kplus */
    const int* bc_fs;        /* 6 per block: inflow flow-state index */
    int n_flow_states;
    const double* flow_states;
    const double* vertices;  /* per block (nic+1)(njc+1)(nkc+1) xyz triples, i fastest */
    const double* iovars;    /* per block 13*ncells, flow-file order */
    int x_order;             /* 1 or 2 */
    int flux_calc;           /* 0 = ausmdv, 1 = sbp_asf (face.cu:268-342, face-list stage) */
    int t_order;             /* parsed and ignored, as config.cu:491 */
    int cfl_count, max_step, print_count;
    double max_time, dt_init;
    int n_cfl;
    const double* cfl_times;
    const double* cfl_values;
    int n_dt_plot;
    const double* dt_plot_times;
    const double* dt_plot_values;
    double t0;               /* start time (times.data entry for the start index) */
    int start_tindx;
    int reacting;            /* 1: the A->B reaction after every step (simulate.cu:296-303) */
    int viscous;             /* 1 = add viscous fluxes (block.cu:88-162; simulate.cu:201,235-238) */
    const double* bc_twall;  /* 6 per block: TWall of wall_no_slip_fixed_T sides (NaN: TWall
                                given per face in twall_faces); may be NULL without such walls */
    const int* bc_fun;       /* 6 per block: function of inflow_function sides (face.cu:51-59:
                                1 supersonic_vortex, 2 laminar_boundary_layer,
                                3 manufactured_solution); may be NULL without such sides */
    const double* bl_profile;  /* laminar boundary-layer profile (config.cu:31-60): p_e, then
                                  y, T, velx at each of 12 samples (37 numbers); NULL if unused */
    int n_twall_faces;
    const double* twall_faces; /* TWall per boundary face of every fixed-T side with NaN in
                                  bc_twall, blocks in order, sides in order, faces in the
                                  TWall-file order (block.cu:566-611); NULL if none */
    const double* gas_model;   /* gamma, R, q, alpha, Ti, Cv of the job (config.json
                                  "gas_model"); must be the library's built model
                                  (config.cu:369-386); NULL: take the built model */
    int time_scheme;           /* CHKN_TIME_RK3 (the reference's step, simulate.cu:223-287)
                                  or CHKN_TIME_PC: two-stage predictor-corrector, U1 = U0 +
                                  dt L0 then U = U0 + 0.5 dt (L0 + L1) (the sibling solver
                                  Lorikeet, src/puffin/cell.d:294-303; chicken parses
                                  t_order and ignores it, config.cu:491).  config.json key
                                  "time_scheme": "rk3" (default) or "predictor_corrector" */
} chkn_desc;

#define CHKN_TIME_RK3 0
#define CHKN_TIME_PC 1

/* Progress of the march, filled by chkn_step and chkn_march. */
typedef struct {
    int step;          /* steps taken since creation (SimState::step) */
    double t;          /* simulation time */
    double dt;         /* time step used by the last step */
    int bad_stage;     /* 1..3 if a stage produced bad cells, else 0 */
    int bad_count;
    int steps_done;    /* steps taken by this call */
} chkn_report;

/* Launch options (performance only; results do not depend on them). */
typedef struct {
    int tile_i, tile_j;   /* cell tile per work-group in i and j; 0 = library default */
    int tile_k;           /* cells marched in k per work-group; 0 = default */
    int per_block_launch; /* 1 = one launch per block per stage (the reference's launch
                             structure, simulate.cu:605-625); 0 = one launch per stage */
    int use_graph;        /* 1 = replay steps from a captured hipGraph */
    int persistent;       /* 1 = persistent work-groups loop over the tile table */
    int tile_order;       /* 0 = k-chunk layers outermost (plane-major tile table);
                             1 = the k-chunks of one (i, j) column consecutive */
    int vertex_frames;    /* 1 = the stage kernel forms face frames from the vertices (3 words
                             per vertex read, more arithmetic) instead of reading the stored
                             10 words per face; built for 32x8 and 16x16 tiles, x_order 2,
                             FP64 */
} chkn_launch_opts;

/* Replaces initialize_simulation (simulate.cu:55-155), array form. */
int chkn_create(const chkn_desc* desc, chkn_sim** out);

/* Ensemble form of initialize_simulation (simulate.cu:55-155): ninst independent jobs in one
 * handle.  Each instance keeps its own block array, inflow states, time, time step, CFL and
 * plot schedules; all instances must share x_order, viscous and cfl_count.  Every stage of every
 * instance is computed by one launch over all blocks (the reference runs one job per
 * process and one launch per block, simulate.cu:605-625).  Blocks are numbered globally:
 * instance q owns blocks [first, first+count) of chkn_instance_blocks.  chkn_step advances
 * all instances in lockstep; an instance that reached its max_step or max_time idles.  A bad
 * cell in any instance stops the handle; the reported count sums the instances. */
int chkn_create_batch(const chkn_desc* descs, int ninst, chkn_sim** out);
int chkn_num_instances(chkn_sim* sim);
int chkn_instance_blocks(chkn_sim* sim, int inst, int* first, int* count);
/* Step, time and time step of one instance (chkn_get_report reports instance 0). */
int chkn_get_instance_report(chkn_sim* sim, int inst, chkn_report* rep);

/* Replaces initialize_simulation (simulate.cu:55-155) reading the job directory's
 * config.json, grid and flow files (binary = 1: .bin, 0: .gz) at time index tindx. */
int chkn_load_job(const char* job_dir, int tindx, int binary, chkn_sim** out);

/* Advance up to nsteps reference time steps (simulate.cu:205-308, without snapshot
 * writes), stopping at max_step or max_time.  rep may be NULL. */
int chkn_step(chkn_sim* sim, int nsteps, chkn_report* rep);

/* Replaces march_in_time_using_cpu_only/_gpu (simulate.cu:182-342): march to max_step or
 * max_time, writing snapshots on the dt_plot schedule and printing the progress line
 * "Step=... t=... dt=... cfl=... WC=...s WCtFT=...s WCtMS=...s" every print_count steps
 * (simulate.cu:311-326).  Needs a job directory. */
int chkn_march(chkn_sim* sim, int binary, chkn_report* rep);

/* Replaces finalize_simulation (simulate.cu:768-792): final snapshot if steps are
 * unwritten. */
int chkn_finalize(chkn_sim* sim, int binary);

/* Replaces write_flow_data (simulate.cu:158-179). */
int chkn_write_flow(chkn_sim* sim, int tindx, int binary);

/* The block's 13 IO variables (Block::writeFlow order, block.cu:943-954), 13*ncells. */
int chkn_get_iovars(chkn_sim* sim, int blk, double* out);

/* Conserved quantities of the block, 6*ncells, quantity-major (U0 of the next step). */
int chkn_get_conserved(chkn_sim* sim, int blk, double* out);

/* Replace the block's state by primitive IO variables (13*ncells, only p..velz are read);
 * the conserved state is re-encoded as at start-up (simulate.cu:196-203). */
int chkn_set_iovars(chkn_sim* sim, int blk, const double* iovars);

int chkn_set_launch_opts(chkn_sim* sim, const chkn_launch_opts* opts);
int chkn_get_report(chkn_sim* sim, chkn_report* rep);
int chkn_num_blocks(chkn_sim* sim);
int chkn_block_dims(chkn_sim* sim, int blk, int* dims3);
long long chkn_active_cells(chkn_sim* sim);

/* Issue all further work of the handle on `stream`, a hipStream_t the caller owns (for
 * example the current stream of a torch device; NULL is the null stream), so the handle's
 * work is ordered with the caller's without device-wide synchronisation.  Work already
 * queued completes first.  The caller keeps the stream alive while the handle uses it; on the
 * null stream, segments run without hipGraph capture.  Each entry point still returns only
 * after its work on the stream is done. */
int chkn_set_stream(chkn_sim* sim, void* stream);
/* Back to the handle's own stream (after the work on the caller's stream completes). */
int chkn_reset_stream(chkn_sim* sim);

/* Device time of the last chkn_step call's kernels, measured with HIP events on the
 * handle's stream (ms), and the number of stage launches it issued. */
int chkn_last_timing(chkn_sim* sim, double* ms, long long* launches);

/* Bytes of the build's number type: 8, or 4 in the FP32 flavours (libchkn_f32.so,
 * libchkn_ab_f32.so; the reference's FLOAT_NUMBERS build, number.cu:10-27).  The C-ABI takes
 * and returns doubles in every flavour; an FP32 build rounds inputs to float and widens
 * outputs exactly. */
int chkn_real_size(void);

/* ---- Gradients of an unrolled march (no counterpart in the reference, which has no
 * adjoint; these serve differentiable pipelines built on the step).  Flat state arrays
 * concatenate every block's 6*ncells conserved quantities (quantity-major per block, block
 * order; inactive blocks hold zeros) and live in device memory. */

/* Doubles in a flat state array. */
long long chkn_state_size(chkn_sim* sim);
/* Copy the conserved state of every block into / out of a flat device array.  Writing sets
 * each active block's primitive state to decode(U) (flow.cu:245-268). */
int chkn_read_conserved_dev(chkn_sim* sim, double* dst_dev);
int chkn_write_conserved_dev(chkn_sim* sim, const double* src_dev);
/* Number of inflow flow states over all instances (state indices are global: instance q's
 * states follow those of instances 0..q-1). */
int chkn_num_flow_states(chkn_sim* sim);
/* Inflow flow state `state` (global index over instances) as (rho, e, YB, vx, vy, vz); the
 * reference derives rho and e from p and T (config.cu:398-406, gas.cu:119-126). */
int chkn_set_inflow_state(chkn_sim* sim, int state, const double* w6);
int chkn_get_inflow_state(chkn_sim* sim, int state, double* w6);
/* Areas of a block's boundary faces on side 0..5 (iminus .. kplus), ordered with the
 * first transverse index fastest (j,k for i-sides; i,k for j-sides; i,j for k-sides). */
int chkn_boundary_face_areas(chkn_sim* sim, int blk, int side, double* out);
/* Advance nsteps steps keeping a checkpoint every checkpoint_every (1..64) steps (max_step
 * and max_time do not stop it).  The time steps follow the CFL schedule (dt_in NULL) or are
 * given, dt_in[step*ninst + inst]; those used are written to dt_out when it is not NULL.
 * Either time scheme (TVD-RK3 or predictor-corrector) and either flux (AUSMDV or SBP-ASF);
 * inviscid jobs, reacting or not (the reaction is reversed between the steps). */
int chkn_grad_forward(chkn_sim* sim, int nsteps, int checkpoint_every, const double* dt_in,
                      double* dt_out);
/* Device memory the forward pass may fill with the stage states of its last checkpoint
 * segments (resident tapes), which the backward pass then reverses without recomputing them:
 * bytes < 0 (the default) allows half of the free device memory at chkn_grad_forward, 0
 * recomputes every segment.  The gradient does not depend on it.  No counterpart in the
 * reference, which has no adjoint. */
int chkn_grad_set_tape_budget(chkn_sim* sim, long long bytes);
/* Form of each reversed stage: separate = 0 (the default) lets the k-direction face pass
 * finish the cells itself when no block exchanges through a k side; separate = 1 always
 * launches the face pass of every direction and then a per-cell gather.  Both give the same
 * gradient bit for bit.  No counterpart in the reference, which has no adjoint. */
int chkn_grad_set_gather_form(chkn_sim* sim, int separate);
/* The form the handle's reversed stages run in: 0 fused, 1 separate (asked for, or a block
 * exchanges through a k side); < 0 on error. */
int chkn_grad_gather_form(chkn_sim* sim);
/* Checkpoint segments of the last chkn_grad_forward held on resident tapes. */
int chkn_grad_resident_segments(chkn_sim* sim);
/* Number of successful chkn_grad_forward calls on the handle: chkn_grad_backward reverses the
 * latest one only, so a caller that holds several unrolls can tell whether its tape is still
 * the one on the handle. */
long long chkn_grad_generation(chkn_sim* sim);
/* Reverse the last chkn_grad_forward with its time steps held fixed: lambda_dev = dLoss/dU
 * of the final state (flat, device); grad_u0_dev receives dLoss/dU of the state the forward
 * started from (flat, device); grad_inflow (host, 6 per inflow state, may be NULL)
 * receives dLoss/d(rho, e, YB, vx, vy, vz) of every inflow state.  The handle's state is
 * left at the final state. */
int chkn_grad_backward(chkn_sim* sim, const double* lambda_dev, double* grad_u0_dev,
                       double* grad_inflow);
/* Test hook: vector-Jacobian products of the face flux at npts points on the device.
 * in: per point 4 x 6 stencil primitives (L1, L0, R0, R1; rho, e, YB, vx, vy, vz), the 9
 * frame components (n, t1, t2) and 6 flux adjoints; out: 24 stencil adjoints per point.
 * x_order 1 or 2: AUSMDV (face.cu:347-371) at that reconstruction order; 0: the SBP-ASF flux
 * (face.cu:268-342, p from the equation of state of each stencil state). */
int chkn_flux_vjp_points(int npts, int x_order, const double* in, double* out);
/* The same evaluation without any device allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  in and out are host arrays as above; they are staged
 * through a file-scope device array of 4096 points (2 MB, held by each GPU the library runs on
 * until the process ends), so npts above 4096 is refused (CHKN_ERR_ARG).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int chkn_flux_vjp_points_noalloc(int npts, int x_order, const double* in, double* out);

/* Whether one block of nic x njc x nkc cells fits the kernels' 32-bit index and offset
 * arithmetic (no counterpart in the reference, whose CPU arrays are std::vector): every
 * per-direction face count and the vertex count below 2^32 / sizeof(number) (2^29 in FP64),
 * and for viscous jobs 4 x (all faces) as well.  CHKN_OK, or CHKN_ERR_ARG with the reason in
 * chkn_last_error(NULL).  chkn_create applies the same check to every block. */
int chkn_check_block_dims(int nic, int njc, int nkc, int viscous);

/* Console verbosity of chkn_load_job and chkn_march (Config::verbosity, config.cu:312-344;
 * set from chkn-run --verbosity): 0 prints only the march's progress lines every print_count
 * steps (simulate.cu:311-326); > 0 also echoes the configuration and the march's start and
 * end. */
int chkn_set_verbosity(int level);

/* Digest of the sources the library was built from (no counterpart in the reference): the
 * first 16 hex digits of SHA-256 over the .h, .hip and .cpp files of gdtk-rmhd_amd/csrc in
 * name order followed by this header, so that a caller can tell a library built from other
 * sources. */
const char* chkn_source_digest(void);

/* Device memory allocated by the handle's set-up, in bytes (the counterpart of the
 * reference's "Bytes allocated on CPU", simulate.cu:89-90). */
long long chkn_device_bytes(chkn_sim* sim);

/* Error text of the last failing call on this handle (or of the last failed create/load
 * when sim is NULL). */
const char* chkn_last_error(chkn_sim* sim);

void chkn_destroy(chkn_sim* sim);

#ifdef __cplusplus
}
#endif
#endif
