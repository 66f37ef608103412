// chkn_runtime.h -- host-side state of a libchkn handle, shared by the stepping runtime
// (chkn_sim.cpp) and the adjoint runtime (chkn_grad.cpp).  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/chkn.h"
#include "chkn_bl.h"
#include "chkn_internal.h"

namespace chkn {
namespace rt {

constexpr int kMaxSeg = 64;  // steps per segment (device control slots)

// Error text of the last failed create/load on this thread.
extern thread_local std::string g_create_error;
// Console verbosity of job loading and marching (chkn_set_verbosity).
extern int g_verbosity;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw HipError(std::string("HIP error ") + hipGetErrorString(e_) + " at " #expr);   \
    } while (0)

struct HostBlock {
    int inst = 0;
    int bi, bj, bk, nic, njc, nkc;
    bool active;
    int bc[6], bc_fs[6];
    int n;
    std::vector<real> io0;  // IO variables as given at start (flow part used while fresh)
    std::vector<real> pos;  // 3*n computed centroids
    std::vector<real> vol;  // n computed volumes
    // device
    real* W = nullptr;  // 3 buffers (active) or 1 (inactive)
    real* U = nullptr;
    real* L = nullptr;
    real* gi = nullptr;
    real* gj = nullptr;
    real* gk = nullptr;
    real* volinv = nullptr;
    real* len = nullptr;
    real* vtx = nullptr;  // [3][nv] vertices (active blocks)
    int tile_first = 0, tile_count = 0;
    // face-list stage (viscous jobs; first-order jobs with verbatim gas states)
    real twall[6] = {0, 0, 0, 0, 0, 0};
    int bc_fun[6] = {0, 0, 0, 0, 0, 0};      // inflow_function sides: function index
    std::vector<real> twall_faces[6];      // fixed-T sides with TWall per face
    real* side[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // device side data
    bool gas_differs = false;  // some cell's file p, T, a differ from the equation of state
    real* F = nullptr;       // 6 x nft face fluxes
    real* FS = nullptr;      // nft x 4 face T, vx, vy, vz (viscous)
    real* fpos = nullptr;    // nft x 4 face centres (viscous)
    real* cpos = nullptr;    // n x 4 cell centres (viscous)
    real* gas = nullptr;     // 3 x n file p, T, a (only when gas_differs)
    // adjoint (allocated by the first gradient call)
    double* tape = nullptr;    // (3C+1) W slots of the segment being reversed
    double* tape2 = nullptr;   // second tape: the next segment is recomputed into it meanwhile
    double* rtape = nullptr;   // resident tapes: nres_alloc segments of (3C+1) W slots each
    double* ckpt = nullptr;    // checkpoints: U then W, 12n doubles each
    double* adj = nullptr;     // A0, P, D: 18n doubles
    double* Gc = nullptr;      // per-cell direction sums of face stencil adjoints: 18 x n
    double* GB = nullptr;      // ghost-line adjoints: 12 x (lines of the six sides)
};

// Time control of one instance (SimState of simulate.cu:32-39 plus its schedules).
struct Instance {
    int first_block = 0, nblocks = 0;
    int nib = 0, njb = 0, nkb = 0;
    std::vector<int> blk_ids;  // global block ids
    int cfl_count = 10, max_step = 100, print_count = 20;
    real max_time = real(1e-3), dt_init = real(1e-6);
    std::vector<real> cfl_t, cfl_v, plot_t, plot_v;
    std::string job;
    real t = 0.0, dt = 0.0, t_plot = 0.0, sigmin = 0.0;
    int step = 0, next_plot_indx = 1, steps_since_last_plot = 0;
    bool have_sig = false;
    bool reacting = false;  // the A->B reaction after every step (simulate.cu:296-303)
    bool running() const { return step < max_step && t < max_time; }
};

// Tape of the last chkn_grad_forward call: checkpoints every C steps, the dt of every step.
struct GradState {
    int C = 0;                    // checkpoint interval (steps)
    int nsteps = 0;               // steps recorded
    int nckpt_alloc = 0;          // checkpoint slots allocated per block
    int tape_C = 0;               // segment length the tapes are sized for
    std::vector<double> dts;      // [nsteps][ninst]
    std::vector<double> t0;       // time of each instance at the start
    int step0 = 0, gstep0 = 0;    // counters at the start
    bool valid = false;
    long long generation = 0;     // number of successful chkn_grad_forward calls on the handle
    // face/cell work tables and inflow patches (device)
    int nface_chunks = 0, ncell_chunks = 0, ninflow_cols = 0;  // nface_chunks: face patches
    // Patches are ordered i and j first (nface_ij of them), then k.  fused_ok: no active block
    // exchanges through a k side, so the k patches can finish their cells (the fused form);
    // separate_gather: the caller asked for the separate gather launch anyway.
    int nface_ij = 0;
    bool fused_ok = false;
    bool separate_gather = false;
    void* d_face_chunks = nullptr;
    void* d_cell_chunks = nullptr;
    void* d_inflow_cols = nullptr;
    double* d_inflow_acc = nullptr;  // 6 x ninflow_cols
    double* d_dts = nullptr;         // dts on the device, [nsteps][ninst]
    double* d_chem = nullptr;        // reacting handles: exp(-alpha dt) of every step, [nsteps][ninst]
                                     // (-1 for instances without the reaction)
    // Two tapes: segment j-1 is recomputed on stream rec into one while segment j is reversed
    // from the other on the handle's stream.
    bool two_tapes = false;
    hipStream_t rec = nullptr;
    hipEvent_t ev_rev[2] = {nullptr, nullptr};
    BlockDev* d_tab[2] = {nullptr, nullptr};  // block tables with W at tape / tape2
    // Resident tapes: the last nres segments of the unroll write their stage states during
    // the forward pass and are reversed without recomputation.  tape_budget: bytes they may
    // take (-1: half of the free device memory at chkn_grad_forward, 0: none).
    long long tape_budget = -1;
    int nres = 0, nres_alloc = 0, rtape_C = 0;
    std::vector<BlockDev*> d_rtab;  // block tables with W at each resident tape
    std::vector<int> inflow_col_state;  // flow-state index of each inflow column
};

struct GraphKey {
    int K;
    bool dtred;
    bool fresh;  // the first stage reads the file gas states
    bool operator<(const GraphKey& o) const
    {
        if (K != o.K) return K < o.K;
        if (dtred != o.dtred) return dtred < o.dtred;
        return fresh < o.fresh;
    }
};

}  // namespace rt
}  // namespace chkn

struct chkn_sim {
    using HostBlock = chkn::rt::HostBlock;
    using Instance = chkn::rt::Instance;
    using GraphKey = chkn::rt::GraphKey;
    using BlockDev = chkn::BlockDev;
    using Tile = chkn::Tile;
    // configuration
    std::vector<HostBlock> blocks;
    std::vector<Instance> inst;
    std::vector<chkn::real> inflow_w;  // 6 per state: rho, e, YB, vx, vy, vz (all instances)
    int x_order = 2, flux_calc = 0;
    int time_scheme = 0;  // CHKN_TIME_RK3 or CHKN_TIME_PC (shared by the instances)
    bool viscous = false;
    bool any_reacting = false;
    // Gas states (p, T, a) of the inflow states as configured (update_from_pT), and whether
    // the equation of state of their (rho, e) gives other values.
    std::vector<chkn::real> inflow_gas;
    std::vector<char> inflow_differs;
    int gstep = 0;       // lockstep step counter shared by the instances
    bool fresh = true;   // no step taken since the state was set from IO variables
    bool failed = false;
    // device
    hipStream_t stream = nullptr;      // the stream all work goes to: own_stream or the caller's
    hipStream_t own_stream = nullptr;  // created by the handle
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    BlockDev* d_blocks = nullptr;
    Tile* d_tiles = nullptr;
    chkn::real* d_inflow = nullptr;
    chkn::real* d_inflow_gas = nullptr;
    // face-list stage work tables
    int nsplit_faces = 0, nsplit_cells = 0, nsplit_patches = 0;
    chkn::FaceChunk* d_split_faces = nullptr;
    chkn::FacePatch* d_split_patches = nullptr;
    chkn::CellChunk* d_split_cells = nullptr;
    // Control area, kMaxSeg step slots: dt[S][N], sigmin[S][N], run[S][N], bad[S][3].
    char* d_ctl = nullptr;
    char* h_ctl = nullptr;
    size_t ctl_bytes = 0;
    int ntiles = 0;
    // launch options
    chkn_launch_opts opts{32, 8, 0, 0, 0, 0, 0, 0};
    int kc = 8;
    int grid_limit = 0;
    int ncu = 0;           // compute units of the device
    std::map<GraphKey, hipGraphExec_t> graphs;
    // timing of the last chkn_step call
    double last_ms = 0.0;
    std::chrono::steady_clock::time_point march_start;  // wall clock of the progress lines
    long long device_bytes = 0;  // device memory allocated at set-up
    long long last_launches = 0;
    std::string err;
    // host copy of the device block table (the adjoint re-points W at its tape)
    std::vector<chkn::BlockDev> bdev;
    std::unique_ptr<chkn::rt::GradState> grad;
    std::vector<double>* dt_log = nullptr;  // when set, do_steps appends [step][inst] dts
    // When set, the stages read and write their states in a tape (block table tape_tab, W
    // slots from tape_base on) instead of the ring; do_steps advances tape_base.
    const chkn::BlockDev* tape_tab = nullptr;
    int tape_base = 0;
};


namespace chkn {
namespace rt {

int set_error(chkn_sim* s, int code, const std::string& msg);
void dfree(void* p);
void dalloc(chkn_sim* s, void** p, size_t bytes);
// Views of the control area for step slot q (N instances).
real* ctl_dt(char* base, int q, int N);
unsigned long long* ctl_sig(char* base, int q, int N);
int* ctl_run(char* base, int q, int N);
int* ctl_bad(char* base, int q, int N);
real* ctl_chem(char* base, int q, int N);
// Fill the chemistry factors of step slot q from the instances' dt (-1: not reacting).
void fill_chem(chkn_sim* s, int q);
StageArgs stage_args(chkn_sim* s, int q);
// Stage launches per step: 3 (TVD-RK3) or 2 (predictor-corrector).
int stages_per_step(const chkn_sim* s);
// Runs K steps whose dt and run flags are already in the host control area.
void run_segment(chkn_sim* s, int K, bool dtred_last);
int do_steps(chkn_sim* s, int nsteps, bool plot_writes, bool binary, chkn_report* rep);
void upload_blocks(chkn_sim* s);
void free_grad(chkn_sim* s);
// One stage of every active block: fused tile kernel, or the face-list form where the job
// needs it (file_gas: the first stage after the state was read from IO variables).
void launch_stage_form(chkn_sim* s, const StageArgs& a, int stage, bool dtred, bool file_gas,
                       hipStream_t st);
long long launches_per_stage(chkn_sim* s, bool file_gas);
// Whether the first stage after the state was read must use the file gas states.
bool fresh_needs_file_gas(const chkn_sim* s);

// Face-list stage set-up (chkn_split.cpp).
// Per block after its geometry: face fluxes, and for viscous jobs face states and the
// cell and face centres (fpos, [3][nft]) the least-squares gradients are taken over.
void split_block_setup(chkn_sim* s, HostBlock& b, BlockDev& bd, const std::vector<real>& fpos);
// Per-side data of an active block: inflow-function ghost and face states evaluated at their
// positions (prof: the instance's boundary-layer profile, or null), TWall per face.
void side_data_setup(HostBlock& b, BlockDev& bd, const std::vector<real>& fpos, const bl::Profile* prof);
// The file gas states of block b from its IO variables (re-run when they are replaced).
void split_block_gas(chkn_sim* s, HostBlock& b, BlockDev& bd);
// Inflow gas states from the configured p, T, YB of every flow state.
void split_inflow_gas(chkn_sim* s, const std::vector<real>& pTYB);
// Work tables over every active block, after all blocks are set up.
void split_tables(chkn_sim* s);
void split_free(chkn_sim* s);
// An inflow state replaced by (rho, e, YB, v): its gas state follows from the equation of state.
void split_set_inflow(chkn_sim* s, int state);

template <typename F>
int guarded(chkn_sim* s, F&& fn)
{
    try {
        return fn();
    } catch (const HipError& e) {
        return set_error(s, CHKN_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        return set_error(s, CHKN_ERR_ARG, e.what());
    } catch (const std::exception& e) {
        return set_error(s, CHKN_ERR_IO, e.what());
    }
}

}  // namespace rt
}  // namespace chkn
