// chkn_internal.h -- device data layout and kernel launch interface shared by the host
// runtime (chkn_sim.cpp) and the gfx950 kernels (chkn_kernels.hip).
//
// HBM layout (per active block of n = nic*njc*nkc cells, structure of arrays, FP64):
//   W[3][6][n]   primitive stage states (rho, e, YB, vx, vy, vz), a 3-buffer ring
//   U[6][n]      conserved quantities at the start of the step (updated in place, stage 3)
//   L[6][n]      stage-1 residual dU/dt, then (L0/6 + L1/6) after stage 2
//   face geometry per direction, 10 components x nfaces: area, n(3), t1(3), t2(3)
//     i-faces: (nic+1) x njc x nkc, index i + (nic+1)*(j + njc*k)
//     j-faces: nic x (njc+1) x nkc, index i + nic*(j + (njc+1)*k)
//     k-faces: nic x njc x (nkc+1), index i + nic*(j + njc*k)
//   volinv[n] = 1/volume, lengths[3][n] (iLength, jLength, kLength)
// Inactive blocks hold only W (all three buffers equal to the file state) and geometry
// needed by exchange neighbours.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
    double* W[3];
    double* U;
    double* L;
    const double* gi;
    const double* gj;
    const double* gk;
    const double* volinv;
    const double* len;
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
    const double* dt;                 // [ninst] time step of each instance
    const int* run;                   // [ninst] 1 if the instance takes this step
    unsigned long long* sigmin_bits;  // [ninst] min over cells of the cfl-free signal time
    int* bad;                         // [3] bad cells per stage
};

struct StageArgs {
    const BlockDev* blocks;
    const Tile* tiles;
    int ntiles;
    int tile_base;         // offset into the tile table (per-block launches)
    const double* inflow;  // 6 doubles per flow state: rho, e, YB, vx, vy, vz
    StepCtl ctl;
    int win, wout;         // W ring indices read / written by this stage
};

// Launch one stage over tiles [tile_base, tile_base + ntiles) of the table.
// stage in 1..3; x_order 1 or 2; tile shape (ti, tj) must be a built variant.
hipError_t launch_stage(int stage, int x_order, int ti, int tj, bool dt_reduce, bool persistent,
                        int grid_limit, const StageArgs& a, hipStream_t s);
bool tile_shape_supported(int ti, int tj);

// Minimum over the block's cells of the cfl-free signal time of its W[0] state, folded into
// *sigmin_bits (used when no stage-3 launch has reduced the current state).
hipError_t launch_signal_min(const BlockDev* block, int ncells, unsigned long long* sigmin_bits,
                             hipStream_t s);

}  // namespace chkn
