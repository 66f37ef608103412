// chkn_kernels.hip -- gfx950 stage kernel of the chicken explicit step (FP64).
//
// One launch per RK stage covers every tile of every active block (or one block's tiles
// when the reference's per-block launch structure is requested).  A work-group owns a
// TI x TJ column of cells and marches up a chunk of k-planes.  With the plane's states (and
// a 2-cell halo) staged in LDS, it evaluates the plane's (TI+1)*TJ i-faces and TI*(TJ+1)
// j-faces as one flattened list, lanes of a wave taking whichever face type falls in their
// slot; the area-weighted fluxes go to LDS.  Each thread then evaluates its column's k-face
// one plane up (the face below was the previous plane's), forms its cell's flux divergence
// in the reference's summation order (cell.cu:206-213), applies the TVD-RK3 stage
// (block.cu:1090-1160), decodes (flow.cu:245-268) and writes the next stage's primitive
// state, while the work-group loads the next plane's halo into LDS (two barriers per
// plane).  Ghost cells are never stored: a stencil index outside the
// block is resolved on the fly by its boundary condition (bcs.cu:16-503), with exchange
// reading the neighbour block's current state directly.
//
// Build with -ffp-contract=off: every product and sum rounds exactly as in the reference.
#include "chkn_internal.h"
#include "chkn_physics.h"
#include "chkn_stencil.h"

namespace chkn {

namespace {

template <bool VF>
__device__ __forceinline__ Prim stencil_prim(const StageArgs& A, const BlockDev* B,
                                             const real* __restrict__ W, int n_line, int base,
                                             int stride, int dir, int m, int a, int c)
{
    if (m >= 0 && m < n_line) return ld_prim(W, B->n, base + m * stride);
    return ghost_prim<VF>(A.blocks, A.inflow, A.win, B, dir, m, a, c);
}

// LDS layout of a tile: the k-plane of primitive states with a 2-cell halo in i and j
// (rectangle of (TI+4) x (TJ+4) positions; only the cross-shaped part is filled), then the
// area-weighted fluxes of the plane's i-faces and j-faces.
template <int TI, int TJ>
struct TileLds {
    static constexpr int PW = TI + 4;
    static constexpr int NP = PW * (TJ + 4);
    static constexpr int NFI = (TI + 1) * TJ;
    static constexpr int NFJ = TI * (TJ + 1);
    static constexpr int NLOAD = (TI + 4) * TJ + 4 * TI;  // filled plane positions
    static constexpr int DOUBLES = 6 * NP + 6 * NFI + 6 * NFJ;
};

// Load one plane position (ii, jj relative to the tile origin) for plane k into LDS:
// an interior cell, a ghost cell across an i- or j-boundary, or nothing if no face of the
// tile can reference it.  skip_own: the tile's own interior cells are already in place (the
// threads that own them stored them from their k-stencil loads of the previous plane).
template <int TI, int TJ, bool VF>
__device__ __forceinline__ void fill_plane_position(const StageArgs& A, const BlockDev* B,
                                                    const real* __restrict__ Win, int i0,
                                                    int j0, int k, int p, real* Wp, bool skip_own)
{
    using L = TileLds<TI, TJ>;
    int ii, jj;
    if (p < (TI + 4) * TJ) {
        jj = p / (TI + 4);
        ii = p % (TI + 4) - 2;
    } else {
        const int q = p - (TI + 4) * TJ;
        const int r = q / TI;
        jj = (r < 2) ? r - 2 : TJ + r - 2;
        ii = q % TI;
    }
    const int gi = i0 + ii, gj = j0 + jj;
    const int nic = B->nic, njc = B->njc;
    const bool in_i = (gi >= 0) && (gi < nic), in_j = (gj >= 0) && (gj < njc);
    Prim w;
    if (in_i && in_j) {
        if (skip_own && ii >= 0 && ii < TI && jj >= 0 && jj < TJ) return;
        w = ld_prim(Win, B->n, gi + nic * (gj + njc * k));
    } else if (in_j && gi >= -2 && gi <= nic + 1) {
        w = ghost_prim<VF>(A.blocks, A.inflow, A.win, B, 0, gi, gj, k);
    } else if (in_i && gj >= -2 && gj <= njc + 1) {
        w = ghost_prim<VF>(A.blocks, A.inflow, A.win, B, 1, gj, gi, k);
    } else {
        return;
    }
    const int o = (jj + 2) * L::PW + (ii + 2);
    Wp[o] = w.rho;
    Wp[L::NP + o] = w.e;
    Wp[2 * L::NP + o] = w.YB;
    Wp[3 * L::NP + o] = w.vx;
    Wp[4 * L::NP + o] = w.vy;
    Wp[5 * L::NP + o] = w.vz;
}

__device__ __forceinline__ Prim lds_prim(const real* Wp, int np, int o)
{
    Prim p;
    p.rho = Wp[o];
    p.e = Wp[np + o];
    p.YB = Wp[2 * np + o];
    p.vx = Wp[3 * np + o];
    p.vy = Wp[4 * np + o];
    p.vz = Wp[5 * np + o];
    return p;
}

__device__ __forceinline__ void lds_put(real* Wp, int np, int o, const Prim& p)
{
    Wp[o] = p.rho;
    Wp[np + o] = p.e;
    Wp[2 * np + o] = p.YB;
    Wp[3 * np + o] = p.vx;
    Wp[4 * np + o] = p.vy;
    Wp[5 * np + o] = p.vz;
}

// Area-weighted flux of the k-face (i, j, kf) between cells kf-1 and kf from its stencil
// states (cells kf-2 .. kf+1 of the column).
template <int XORDER, bool VF>
__device__ __forceinline__ void k_face_flux(const BlockDev* B, int i, int j, int kf, const Prim& s0,
                                            const Prim& s1, const Prim& s2, const Prim& s3,
                                            real AF[6])
{
    Frame fr;
    real area;
    if (VF) {
        face_frame_v<2>(B, i, j, kf, fr, area);
    } else {
        const int fi = i + B->nic * (j + B->njc * kf);
        fr = ld_frame(B->gk, B->nfk, fi);
        area = gld(B->gk, boff(fi));
    }
    real F[6];
    face_flux<XORDER>(s0, s1, s2, s3, fr, F);
#pragma unroll
    for (int q = 0; q < 6; ++q) AF[q] = area * F[q];
}

// One tile: TI x TJ columns of cells marching planes k0 .. k0+nk-1.  Per plane, after the
// plane's states are in LDS: (1) the plane's i- and j-faces as one flattened list, fluxes to
// LDS; barrier; (2) each thread's cell: its column's k-face one plane up, the flux
// divergence, the RK stage and decode, while all threads fill the next plane's positions
// (nothing in this part reads them: a thread's own cell is read back by that thread only,
// and it stores its k-stencil cell k+1 there itself); barrier.  VF: face frames formed from
// the block's vertices (same bits as the stored geometry; 3 words per vertex read instead of
// 10 per face, for more arithmetic).
template <int STAGE, int TI, int TJ, int XORDER, bool DTRED, bool VF>
__device__ __forceinline__ void process_tile(const StageArgs& A, const Tile tile, real* lds)
{
    using L = TileLds<TI, TJ>;
    constexpr int NT = TI * TJ;
    constexpr int NFI = L::NFI, NFJ = L::NFJ, NP = L::NP, PW = L::PW;
    real* Wp = lds;               // [6][TJ+4][TI+4]
    real* AFi = Wp + 6 * NP;      // [6][TJ][TI+1]
    real* AFj = AFi + 6 * NFI;    // [6][TJ+1][TI]

    const BlockDev* B = A.blocks + tile.blk;
    const int nic = B->nic, njc = B->njc;
    const long long n = B->n;
    const real* __restrict__ Win = B->W + A.win * B->wstride;
    const int t = threadIdx.x;
    const int ti = t % TI, tj = t / TI;
    const int i = tile.i0 + ti, j = tile.j0 + tj;
    const bool own = (i < nic) && (j < njc);
    const int kbeg = tile.k0, kend = tile.k0 + tile.nk;
    const real dt = A.ctl.dt[B->inst];
    const int vsj = nic + 1, vsk = (nic + 1) * (njc + 1);
    const int o = (tj + 2) * PW + ti + 2;  // the thread's own cell in the LDS plane
    int bad = 0;
    unsigned long long sigmin = ~0ull;

    // The column's k-faces: the stencil of face k+1 is cells k-1 .. k+2, of which cell k is
    // read from the LDS plane.
    const int nkc = B->nkc, kbase = i + nic * j, kstride = nic * njc;
    real AFkm[6] = {0, 0, 0, 0, 0, 0};
    if (own) {
        const Prim s0 = stencil_prim<VF>(A, B, Win, nkc, kbase, kstride, 2, kbeg - 2, i, j);
        const Prim s1 = stencil_prim<VF>(A, B, Win, nkc, kbase, kstride, 2, kbeg - 1, i, j);
        const Prim s2 = stencil_prim<VF>(A, B, Win, nkc, kbase, kstride, 2, kbeg, i, j);
        const Prim s3 = stencil_prim<VF>(A, B, Win, nkc, kbase, kstride, 2, kbeg + 1, i, j);
        k_face_flux<XORDER, VF>(B, i, j, kbeg, s0, s1, s2, s3, AFkm);
    }
    // The first plane, with halo and ghost cells.
    for (int p = t; p < L::NLOAD; p += NT)
        fill_plane_position<TI, TJ, VF>(A, B, Win, tile.i0, tile.j0, kbeg, p, Wp, false);
    __syncthreads();

    for (int k = kbeg; k < kend; ++k) {
        // i- and j-faces of the plane, one flattened list, stencils read from LDS.
        for (int f = t; f < NFI + NFJ; f += NT) {
            int base, stride, fi, gi, gj;
            real* dst;
            int lstride;
            bool valid, idir;
            if (f < NFI) {
                const int fa = f % (TI + 1), fb = f / (TI + 1);
                gi = tile.i0 + fa; gj = tile.j0 + fb;
                valid = (gi <= nic) && (gj < njc);
                base = (fb + 2) * PW + fa;  // cells fa-2 .. fa+1 of row fb
                stride = 1;
                idir = true;
                fi = gi + (nic + 1) * (gj + njc * k);
                dst = AFi + fb * (TI + 1) + fa; lstride = NFI;
            } else {
                const int h = f - NFI;
                const int fa = h % TI, fb = h / TI;
                gi = tile.i0 + fa; gj = tile.j0 + fb;
                valid = (gi < nic) && (gj <= njc);
                base = fb * PW + fa + 2;    // rows fb-2 .. fb+1 of column fa
                stride = PW;
                idir = false;
                fi = gi + nic * (gj + (njc + 1) * k);
                dst = AFj + fb * TI + fa; lstride = NFJ;
            }
            if (valid) {
                Frame fr;
                real area;
                if (VF) {
                    // i-face vertices (i,j,k) (i,j+1,k) (i,j+1,k+1) (i,j,k+1); j-face
                    // (i+1,j,k) (i,j,k) (i,j,k+1) (i+1,j,k+1).
                    const int v = gi + vsj * gj + vsk * k;
                    if (idir) vertex_frame(B, v, vsj, vsj + vsk, vsk, fr, area);
                    else vertex_frame(B, v + 1, -1, vsk - 1, vsk, fr, area);
                } else {
                    const real* g = idir ? B->gi : B->gj;
                    const long long nf = idir ? B->nfi : B->nfj;
                    fr = ld_frame(g, nf, fi);
                    area = gld(g, boff(fi));
                }
                const Prim s0 = lds_prim(Wp, NP, base);
                const Prim s1 = lds_prim(Wp, NP, base + stride);
                const Prim s2 = lds_prim(Wp, NP, base + 2 * stride);
                const Prim s3 = lds_prim(Wp, NP, base + 3 * stride);
                real F[6];
                face_flux<XORDER>(s0, s1, s2, s3, fr, F);
#pragma unroll
                for (int q = 0; q < 6; ++q) dst[q * lstride] = area * F[q];
            }
        }
        __syncthreads();
        Prim sk{0, 0, 0, 0, 0, 0};
        if (own) sk = lds_prim(Wp, NP, o);
        // The next plane's positions other than the tile's own cells.
        if (k + 1 < kend) {
            for (int p = t; p < L::NLOAD; p += NT)
                fill_plane_position<TI, TJ, VF>(A, B, Win, tile.i0, tile.j0, k + 1, p, Wp, true);
        }
        if (own) {
            // The column's k-face one plane up, evaluated after the plane's i- and j-faces so
            // that its flux is not held in registers through the face loop.
            real AFkp[6];
            {
                const Prim s2 = stencil_prim<VF>(A, B, Win, nkc, kbase, kstride, 2, k + 1, i, j);
                const Prim s3 = stencil_prim<VF>(A, B, Win, nkc, kbase, kstride, 2, k + 2, i, j);
                const Prim s0 = stencil_prim<VF>(A, B, Win, nkc, kbase, kstride, 2, k - 1, i, j);
                k_face_flux<XORDER, VF>(B, i, j, k + 1, s0, sk, s2, s3, AFkp);
                if (k + 1 < kend) lds_put(Wp, NP, o, s2);  // the next plane's own cell
            }
            const int cidx = i + nic * (j + njc * k);
            const real vinv = gld(B->volinv, boff(cidx));
            real d[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                real s = AFi[q * NFI + tj * (TI + 1) + ti] - AFi[q * NFI + tj * (TI + 1) + ti + 1];
                s = s + AFj[q * NFJ + tj * TI + ti];
                s = s - AFj[q * NFJ + (tj + 1) * TI + ti];
                s = s + AFkm[q];
                s = s - AFkp[q];
                d[q] = vinv * s;
            }
            Prim w;
            const int flag = stage_update<STAGE>(B, n, cidx, dt, d, A.wout, w);
            bad += flag;
            if (DTRED && !flag) {
                const unsigned long long bits = signal_bits<VF>(B, n, i, j, k, cidx, w);
                sigmin = bits < sigmin ? bits : sigmin;
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) AFkm[q] = AFkp[q];
        }
        __syncthreads();
    }
    // Bad-cell count of the stage (decode flags, flow.cu:249-252).
    if (__ballot(bad != 0)) {
        int cnt = bad;
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(&A.ctl.bad[stage_position(STAGE) - 1], cnt);
    }
    if (DTRED) {
        const unsigned long long wm = wave_min_u64(sigmin);
        if ((threadIdx.x & 63) == 0 && wm != ~0ull) atomicMin(&A.ctl.sigmin_bits[B->inst], wm);
    }
}

template <int STAGE, int TI, int TJ, int XORDER, bool DTRED, bool VF>
__global__ __launch_bounds__(TI* TJ) void stage_kernel(StageArgs A0, const BlockDev* __restrict__ blocks,
                                                      const Tile* __restrict__ tiles, int persistent)
{
    __shared__ real lds[TileLds<TI, TJ>::DOUBLES];
    // The block and tile tables are read-only during the launch: passed as restrict-qualified
    // arguments, their loads can be scalar loads that no store of the kernel can clobber.
    StageArgs A = A0;
    A.blocks = blocks;
    A.tiles = tiles;
    // XCD-aware tile mapping: work-group b runs on XCD b % 8, and each XCD takes a
    // contiguous range of the (spatially ordered) tile table, so halo cells shared by
    // neighbouring tiles are re-read from the same L2.
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int nt = A.ntiles;
    const int xcd = b & 7, local = b >> 3, per_xcd = G >> 3;
    const int range = (nt + 7) >> 3;
    int t0 = xcd * range;
    int t1 = t0 + range;
    if (t1 > nt) t1 = nt;
    for (int tt = t0 + local; tt < t1; tt += per_xcd) {
        const Tile tile = A.tiles[A.tile_base + tt];
        // Instances that have reached their end time skip the step (uniform per tile).
        if (A.ctl.run[A.blocks[tile.blk].inst])
            process_tile<STAGE, TI, TJ, XORDER, DTRED, VF>(A, tile, lds);
        if (!persistent) break;
    }
}

template <int TI, int TJ, int XORDER, bool VF>
hipError_t launch_shape(int stage, bool dtred, bool persistent, int grid_limit,
                        const StageArgs& a, hipStream_t s)
{
    int range = (a.ntiles + 7) / 8;
    int grid = 8 * range;
    if (persistent && grid_limit > 0 && grid > grid_limit) grid = (grid_limit / 8) * 8;
    if (grid < 8) grid = 8;
    dim3 blk(TI * TJ);
    int p = persistent ? 1 : 0;
    if (stage == 1)
        hipLaunchKernelGGL((stage_kernel<1, TI, TJ, XORDER, false, VF>), dim3(grid), blk, 0, s, a, a.blocks, a.tiles, p);
    else if (stage == 2)
        hipLaunchKernelGGL((stage_kernel<2, TI, TJ, XORDER, false, VF>), dim3(grid), blk, 0, s, a, a.blocks, a.tiles, p);
    else if (stage == 3 && dtred)
        hipLaunchKernelGGL((stage_kernel<3, TI, TJ, XORDER, true, VF>), dim3(grid), blk, 0, s, a, a.blocks, a.tiles, p);
    else if (stage == 3)
        hipLaunchKernelGGL((stage_kernel<3, TI, TJ, XORDER, false, VF>), dim3(grid), blk, 0, s, a, a.blocks, a.tiles, p);
    else if (dtred)
        hipLaunchKernelGGL((stage_kernel<4, TI, TJ, XORDER, true, VF>), dim3(grid), blk, 0, s, a, a.blocks, a.tiles, p);
    else
        hipLaunchKernelGGL((stage_kernel<4, TI, TJ, XORDER, false, VF>), dim3(grid), blk, 0, s, a, a.blocks, a.tiles, p);
    return hipGetLastError();
}

// Shapes with the vertex-frame form built (FP64, second-order reconstruction).
template <int TI, int TJ, int XORDER>
constexpr bool vertex_frames_built()
{
#ifdef CHKN_FP32
    return false;
#else
    return XORDER == 2 && ((TI == 32 && TJ == 8) || (TI == 16 && TJ == 16));
#endif
}

template <int TI, int TJ, int XORDER>
hipError_t launch_frames(int vertex_frames, int stage, bool dtred, bool persistent, int grid_limit,
                          const StageArgs& a, hipStream_t s)
{
    if constexpr (vertex_frames_built<TI, TJ, XORDER>()) {
        if (vertex_frames) return launch_shape<TI, TJ, XORDER, true>(stage, dtred, persistent, grid_limit, a, s);
    } else {
        if (vertex_frames) return hipErrorInvalidValue;
    }
    return launch_shape<TI, TJ, XORDER, false>(stage, dtred, persistent, grid_limit, a, s);
}

template <int TI, int TJ>
hipError_t launch_order(int vertex_frames, int stage, int x_order, bool dtred, bool persistent, int grid_limit,
                        const StageArgs& a, hipStream_t s)
{
    if (x_order > 1) return launch_frames<TI, TJ, 2>(vertex_frames, stage, dtred, persistent, grid_limit, a, s);
    return launch_frames<TI, TJ, 1>(vertex_frames, stage, dtred, persistent, grid_limit, a, s);
}

__global__ __launch_bounds__(256) void signal_min_kernel(const BlockDev* B,
                                                        unsigned long long* sigmin_bits)
{
    const long long n = B->n;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = ~0ull;
    if (c < n) {
        const int nic = B->nic, njc = B->njc;
        const int i = c % nic, j = (c / nic) % njc, k = c / (nic * njc);
        v = signal_bits<false>(B, n, i, j, k, c, ld_prim(B->W, n, c));
    }
    v = wave_min_u64(v);
    if ((threadIdx.x & 63) == 0 && v != ~0ull) atomicMin(sigmin_bits, v);
}

// W = decode(U) for every active block (after its conserved state was replaced).
__global__ __launch_bounds__(256) void decode_all_kernel(const BlockDev* blocks, const CellChunk* cells)
{
    const CellChunk ch = cells[blockIdx.x];
    const BlockDev* B = blocks + ch.blk;
    const long long n = B->n;
    const int cidx = ch.c0 + (int)threadIdx.x;
    if (cidx >= ch.c1) return;
    real U[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) U[q] = B->U[(long long)q * n + cidx];
    Prim w{0, 0, 0, 0, 0, 0};
    decode(U, w);
    real* W = B->W;
    W[cidx] = w.rho;
    W[n + cidx] = w.e;
    W[2 * n + cidx] = w.YB;
    W[3 * n + cidx] = w.vx;
    W[4 * n + cidx] = w.vy;
    W[5 * n + cidx] = w.vz;
}

}  // namespace

hipError_t launch_decode_all(const BlockDev* blocks, const CellChunk* cells, int ncell_chunks,
                             hipStream_t s)
{
    if (ncell_chunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(decode_all_kernel, dim3(ncell_chunks), dim3(256), 0, s, blocks, cells);
    return hipGetLastError();
}

hipError_t launch_signal_min(const BlockDev* block, int ncells, unsigned long long* sigmin_bits,
                             hipStream_t s)
{
    if (ncells <= 0) return hipSuccess;
    hipLaunchKernelGGL(signal_min_kernel, dim3((ncells + 255) / 256), dim3(256), 0, s, block,
                       sigmin_bits);
    return hipGetLastError();
}

template <int TI, int TJ>
int slots_shape(int x_order)
{
    int n = 0;
    const hipError_t e =
        x_order > 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, stage_kernel<2, TI, TJ, 2, false, false>, TI * TJ, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, stage_kernel<2, TI, TJ, 1, false, false>, TI * TJ, 0);
    return e == hipSuccess ? n : 0;
}

int stage_slots_per_cu(int ti, int tj, int x_order)
{
    if (ti == 32 && tj == 8) return slots_shape<32, 8>(x_order);
    if (ti == 64 && tj == 4) return slots_shape<64, 4>(x_order);
    if (ti == 16 && tj == 16) return slots_shape<16, 16>(x_order);
    if (ti == 64 && tj == 8) return slots_shape<64, 8>(x_order);
    if (ti == 16 && tj == 4) return slots_shape<16, 4>(x_order);
    if (ti == 8 && tj == 8) return slots_shape<8, 8>(x_order);
    return 0;
}

bool tile_shape_supported(int ti, int tj)
{
    return (ti == 32 && tj == 8) || (ti == 64 && tj == 4) || (ti == 16 && tj == 16) ||
           (ti == 64 && tj == 8) || (ti == 16 && tj == 4) || (ti == 8 && tj == 8);
}

bool vertex_frames_supported(int ti, int tj, int x_order)
{
    if (ti == 32 && tj == 8) return x_order > 1 ? vertex_frames_built<32, 8, 2>() : vertex_frames_built<32, 8, 1>();
    if (ti == 16 && tj == 16) return x_order > 1 ? vertex_frames_built<16, 16, 2>() : vertex_frames_built<16, 16, 1>();
    return false;
}

hipError_t launch_stage(int stage, int x_order, int ti, int tj, int vertex_frames, bool dt_reduce, bool persistent,
                        int grid_limit, const StageArgs& a, hipStream_t s)
{
    if (a.ntiles <= 0) return hipSuccess;
    if (ti == 32 && tj == 8) return launch_order<32, 8>(vertex_frames, stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    if (ti == 64 && tj == 4) return launch_order<64, 4>(vertex_frames, stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    if (ti == 16 && tj == 16) return launch_order<16, 16>(vertex_frames, stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    if (ti == 64 && tj == 8) return launch_order<64, 8>(vertex_frames, stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    if (ti == 16 && tj == 4) return launch_order<16, 4>(vertex_frames, stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    if (ti == 8 && tj == 8) return launch_order<8, 8>(vertex_frames, stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    return hipErrorInvalidValue;
}

}  // namespace chkn
