// chkn_internal.h -- device data layout and kernel launch interface shared by the host
// runtime (chkn_sim.cpp) and the gfx950 kernels (chkn_kernels.hip).
//
// HBM layout (per active block of n = nic*njc*nkc cells, structure of arrays, `real`:
// FP64, or FP32 in the FP32 flavour):
//   W[slots][6][n] primitive stage states (rho, e, YB, vx, vy, vz): a 3-slot ring while
//                stepping, or the 3C+1 stage states of a C-step segment while recording a
//                tape for the adjoint; slot r starts at W + r*wstride
//   U[6][n]      conserved quantities at the start of the step (updated in place, stage 3)
//   L[6][n]      stage-1 residual dU/dt, then (L0/6 + L1/6) after stage 2
//   face geometry per direction, 10 components x nfaces: area, n(3), t1(3), t2(3)
//     i-faces: (nic+1) x njc x nkc, index i + (nic+1)*(j + njc*k)
//     j-faces: nic x (njc+1) x nkc, index i + nic*(j + (njc+1)*k)
//     k-faces: nic x njc x (nkc+1), index i + nic*(j + njc*k)
//   volinv[n] = 1/volume, lengths[3][n] (iLength, jLength, kLength)
//   vtx[3][nv]   vertex x, y, z, nv = (nic+1)(njc+1)(nkc+1), index i + (nic+1)*(j + (njc+1)*k)
// Face-list stage (viscous jobs; first-order jobs whose file or inflow gas states are not
// reproduced by the equation of state), faces numbered i-faces, then j-faces (+nfi), then
// k-faces (+nfi+nfj), nft faces in all:
//   F[6][nft]    face fluxes (not area-weighted), viscous part subtracted in place
//   FS[nft][4]   face T, vx, vy, vz after the viscous boundary conditions (32-byte records)
//   fpos[nft][4], cpos[n][4] face and cell centres x, y, z (+1 unused) for the clouds
//   gas[3][n]    p, T, a as read from the flow file
// Inactive blocks hold one W slot (wstride 0: every slot is the file state) and the
// geometry needed by exchange neighbours.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chkn_real.h"

namespace chkn {

constexpr int kNQ = 6;      // conserved quantities, flow.cu:17-31
constexpr int kFrameN = 10; // area + n + t1 + t2

struct BlockDev {
    int nic, njc, nkc, active;
    int inst;        // index of the instance (independent job) the block belongs to
    int bc[6];       // face.cu:26-32 codes
    int bc_arg[6];   // inflow flow-state index, or neighbour block id for exchange
    long long n;     // cells
    long long nfi, nfj, nfk;
    real* W;               // slot 0 of the primitive stage states
    long long wstride;     // values between slots: 6n (active) or 0 (inactive)
    real* U;
    real* L;
    const real* gi;
    const real* gj;
    const real* gk;
    const real* volinv;
    const real* len;
    // Vertices [3][nv] (x, y, z), nv = (nic+1)(njc+1)(nkc+1), index i + (nic+1)*(j + (njc+1)*k):
    // the stage kernel forms face frames from them (active blocks).
    const real* vtx;
    long long nv;
    // adjoint
    int rp[6];             // block whose opposite side exchanges into this side, or -1
    int icol;              // first inflow line of this block in the inflow accumulator
    int icol_side[6];      // offset of each inflow side's lines from icol, or -1
    double* Gc;            // per direction d, the sum over the block's own faces of the adjoint
                           // reaching each cell through their stencils: [3][6][n]
    double* GB;            // adjoint reaching each side's ghost lines: per side [2 depths][6][nt]
    double* adj;           // A0 (adjoint of U at the step boundary), P, D, D': [4][6][n]
    // face-list stage
    real* F;
    real* FS;
    const real* fpos;
    const real* cpos;
    const real* gas;       // null when the block's file state is not needed
    real twall[6];         // TWall of each wall_no_slip_fixed_T side
    // Per side, or null: inflow_function sides hold the ghost states of depth 0 and 1
    // ([2][9][nt]: rho, e, YB, vx, vy, vz, p, T, a) then the face state ([4][nt]: T, vx, vy,
    // vz); fixed-T walls with TWall per face hold [nt] values.  nt = faces of the side,
    // index a + na*c as in the ghost lines.
    const real* side[6];
};

struct Tile {
    int blk;
    int i0, j0, k0;
    int nk;
};

// Per-step control data in device memory, written by the host before a segment of steps.
// One slot per instance for dt, run flag and signal-time minimum; one bad-cell counter per
// stage.
struct StepCtl {
    const real* dt;                   // [ninst] time step of each instance
    const int* run;                   // [ninst] 1 if the instance takes this step
    unsigned long long* sigmin_bits;  // [ninst] min over cells of the cfl-free signal time
    int* bad;                         // [3] bad cells per stage
    const real* chem;                 // [ninst] exp(-alpha dt) of reacting instances, else -1
};

struct StageArgs {
    const BlockDev* blocks;
    const Tile* tiles;
    int ntiles;
    int tile_base;         // offset into the tile table (per-block launches)
    const real* inflow;  // 6 values per flow state: rho, e, YB, vx, vy, vz
    StepCtl ctl;
    int win, wout;         // W slots read / written by this stage
};

// Work items of the adjoint kernels: a patch of faces of one direction, up to 256 cells of one
// block, one boundary column (line of ghost cells) of an inflow boundary.
// Patch: faces at line positions m0-1 .. m0+kAdjPatchM-2 (m = face between cells m-1 and m)
// of the kAdjPatchT lines with first transverse index a0 .. a0+kAdjPatchT-1 and second
// transverse index c, in direction dir; it owns the cells m0 .. m0+kAdjPatchM-4 of those lines
// (every face whose stencil holds an owned cell is in the patch).
constexpr int kAdjPatchM = 16, kAdjPatchT = 16, kAdjPatchOwn = kAdjPatchM - 3;
struct AdjPatch {
    int blk, dir, m0, a0, c;
};
struct FaceChunk {
    int blk, dir, f0, f1;  // faces [f0, f1) of direction dir, at most 128
};
// Faces of direction dir at k-level k whose (i, j) lie in the kPatchI x kPatchJ patch at
// (i0, j0).
constexpr int kPatchI = 16, kPatchJ = 16;
struct FacePatch {
    int blk, dir, i0, j0, k;
};
struct CellChunk {
    int blk, c0, c1;       // cells [c0, c1), at most 256
};
struct InflowCol {
    int blk, side, t;  // t: transverse index a + na*c of the line
};

struct AdjArgs {
    const BlockDev* blocks;   // W must point at the tape: slot = stage state being reversed
    const double* inflow;
    const AdjPatch* patches;
    const CellChunk* cells;
    const InflowCol* icols;
    double* inflow_acc;       // [6][nicols]
    int nicols;
    int slot;                 // W slot of the stage state W_s
    const double* dt;         // [ninst] time step of the step being reversed
    const double* dt_next;    // [ninst] time step of the step reversed next, or null (stage 1)
    double seed_p, seed_d;    // a reversed step starts from P = seed_p dt A0, D = seed_d dt A0:
                              // 1/6 and 4/6 (TVD-RK3), 1/2 and 1/2 (predictor-corrector)
    int seed_p_stored;        // 1: the stage-1 seeds store P (the predictor-corrector's first
                              // reversed stage reads it); 0: TVD-RK3's stage 3 forms
                              // P = seed_p dt A0 from the A0 it reads anyway
    const double* chem;       // [ninst] exp(-alpha dt) of the step being reversed (< 0: no
                              // reaction), for the reacting A->B gas
    int dr, dw;               // component offsets in adj of the D this stage reads (faces) and
                              // writes (cells): 12 or 18; equal in the separate-gather form
};

// Adjoint of one stage's residual: face pass (per patch, every face's stencil adjoints summed
// per owned cell into Gc, ghost-line adjoints into GB), then the per-cell gather with the
// stage's combination, in reverse order -- TVD-RK3 stages 3, 2, 1; predictor-corrector the
// corrector (gathered as stage 2: D = P + dt Ubar) then the predictor (stage 1) -- and the
// inflow lines; the stage-1 gather also seeds the next reversed step when dt_next is given.
// seed: P = seed_p dt A0, D = seed_d dt A0 (start of the first reversed step).
hipError_t launch_adj_seed(const AdjArgs& a, int ncell_chunks, hipStream_t s);
// Reacting jobs: A0 = adjoint of U after the step's reaction becomes the adjoint of U before
// it (chem_vjp, with the state after the reaction in W slot a.slot), then the seeds of the
// reversed step as launch_adj_seed.
hipError_t launch_adj_chem_seed(const AdjArgs& a, int ncell_chunks, hipStream_t s);
// epi 0: the face pass of the separate-gather form (sums to Gc and GB); epi = stage 1..3: k
// patches of the fused form, whose owned cells finish that stage (the i and j patches of the
// stage ran before with epi 0).
hipError_t launch_adj_faces(const AdjArgs& a, int npatches, int x_order, int flux_calc, int epi, hipStream_t s);
hipError_t launch_adj_gather(const AdjArgs& a, int ncell_chunks, int stage, hipStream_t s);
// U and W = decode(U) of every active block from U (after the conserved state was replaced).
hipError_t launch_decode_all(const BlockDev* blocks, const CellChunk* cells, int ncell_chunks,
                             hipStream_t s);
// Point evaluation of face_flux_vjp (test hook): in = per point 24 prims, 9 frame, 6 Fbar;
// out = 24 stencil adjoints.
hipError_t launch_flux_vjp_points(const double* in, double* out, int n, int x_order, hipStream_t s);
// The same with host in/out staged through a file-scope device array instead of hipMalloc'd
// buffers: no allocation, at most kNoAllocPoints points (hipErrorInvalidValue beyond).
constexpr int kNoAllocPoints = 4096;
hipError_t flux_vjp_points_noalloc(const double* in, double* out, int n, int x_order);

// Face-list form of one stage: every face of every active block in one pass (convective
// flux, and for viscous jobs the face state), then the viscous pass (gradients and viscous
// flux, viscous jobs only), then a per-cell pass (flux divergence, RK stage, decode).  Cell
// gas states (p, T, a) come from the flow file for inactive blocks, and for active blocks
// while file_gas is set (first stage after the state was read); inflow ghost cells carry the
// configured p, T, a (inflow_gas, 3 per flow state).
struct SplitArgs {
    const BlockDev* blocks;
    const real* inflow;
    const real* inflow_gas;
    const FaceChunk* faces;  // at most 256 faces each
    const FacePatch* patches;  // the viscous pass's work items
    const CellChunk* cells;  // at most 256 cells each
    StepCtl ctl;
    int win, wout;
    int file_gas;
};
hipError_t launch_split_faces(const SplitArgs& a, int nface_chunks, int x_order, int flux_calc,
                              bool viscous, hipStream_t s);
hipError_t launch_split_viscous(const SplitArgs& a, int npatches, hipStream_t s);
hipError_t launch_split_update(const SplitArgs& a, int ncell_chunks, int stage, bool dt_reduce,
                               hipStream_t s);

// The A->B reaction after a step on W slot `slot` and U of every active block of the
// reacting instances (ctl.chem >= 0); with dt_reduce also the signal-time minimum of every
// instance (then taken here rather than in the stage-3 launch).
hipError_t launch_chemistry(const BlockDev* blocks, const CellChunk* cells, int ncell_chunks,
                            const StepCtl& ctl, int slot, bool dt_reduce, hipStream_t s);

// Launch one stage over tiles [tile_base, tile_base + ntiles) of the table.
// stage 1..3 (TVD-RK3) or 4 (corrector of the predictor-corrector, after stage 1); x_order
// 1 or 2; tile shape (ti, tj) must be built; vertex_frames 1 forms
// the face frames from the block's vertices (built for 32x8 and 16x16 at x_order 2, FP64).
hipError_t launch_stage(int stage, int x_order, int ti, int tj, int vertex_frames, bool dt_reduce,
                        bool persistent, int grid_limit, const StageArgs& a, hipStream_t s);
bool tile_shape_supported(int ti, int tj);
bool vertex_frames_supported(int ti, int tj, int x_order);
// Work-groups of a tile shape's stage kernel resident per CU (0 if unknown).
int stage_slots_per_cu(int ti, int tj, int x_order);
// The stage kernel maps work-group b to XCD b mod kXcds and gives each XCD a contiguous
// range of the tile table.
constexpr int kXcds = 8;

// Minimum over the block's cells of the cfl-free signal time of its W[0] state, folded into
// *sigmin_bits (used when no stage-3 launch has reduced the current state).
hipError_t launch_signal_min(const BlockDev* block, int ncells, unsigned long long* sigmin_bits,
                             hipStream_t s);

}  // namespace chkn
