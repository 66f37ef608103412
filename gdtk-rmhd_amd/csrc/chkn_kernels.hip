// chkn_kernels.hip -- gfx950 stage kernel of the chicken explicit step (FP64).
//
// One launch per RK stage covers every tile of every active block (or one block's tiles
// when the reference's per-block launch structure is requested).  A work-group owns a
// TI x TJ column of cells and marches up a chunk of k-planes.  Per plane it evaluates a
// single flattened list of faces -- the (TI+1)*TJ i-faces and TI*(TJ+1) j-faces of the
// plane and the TI*TJ k-faces one plane up -- so every face flux of the tile is computed
// exactly once by one call site, with lanes of a wave working on whichever face type
// falls in their slot.  Area-weighted fluxes go to LDS; the owning thread then forms the
// cell's flux divergence in the reference's summation order (cell.cu:206-213), applies
// the TVD-RK3 stage (block.cu:1090-1160), decodes (flow.cu:245-268) and writes the next
// stage's primitive state.  Ghost cells are never stored: a stencil index outside the
// block is resolved on the fly by its boundary condition (bcs.cu:16-503), with exchange
// reading the neighbour block's current state directly.
//
// Build with -ffp-contract=off: every product and sum rounds exactly as in the reference.
#include "chkn_internal.h"
#include "chkn_physics.h"

namespace chkn {

namespace {

__device__ __forceinline__ Prim ld_prim(const double* __restrict__ W, int n, int c)
{
    Prim p;
    p.rho = W[c];
    p.e = W[n + c];
    p.YB = W[2 * n + c];
    p.vx = W[3 * n + c];
    p.vy = W[4 * n + c];
    p.vz = W[5 * n + c];
    return p;
}

__device__ __forceinline__ Frame ld_frame(const double* __restrict__ g, int nf, int f)
{
    Frame r;
    r.nx = g[1 * nf + f];
    r.ny = g[2 * nf + f];
    r.nz = g[3 * nf + f];
    r.t1x = g[4 * nf + f];
    r.t1y = g[5 * nf + f];
    r.t1z = g[6 * nf + f];
    r.t2x = g[7 * nf + f];
    r.t2y = g[8 * nf + f];
    r.t2z = g[9 * nf + f];
    return r;
}

// Linear cell index of line position m in direction dir with transverse indices (a, c):
// dir 0 -> (m, a, c), dir 1 -> (a, m, c), dir 2 -> (a, c, m).
__device__ __forceinline__ int cell_index(int nic, int njc, int dir, int m, int a, int c)
{
    if (dir == 0) return m + nic * (a + njc * c);
    if (dir == 1) return a + nic * (m + njc * c);
    return a + nic * (c + njc * m);
}

// Boundary face (side 0 = minus, 1 = plus) of the line through (a, c) in direction dir.
__device__ __forceinline__ void boundary_face(const BlockDev* B, int dir, int side, int a, int c,
                                              const double*& g, int& nf, int& fi)
{
    if (dir == 0) {
        g = B->gi; nf = (int)B->nfi;
        fi = (side ? B->nic : 0) + (B->nic + 1) * (a + B->njc * c);
    } else if (dir == 1) {
        g = B->gj; nf = (int)B->nfj;
        fi = a + B->nic * ((side ? B->njc : 0) + (B->njc + 1) * c);
    } else {
        g = B->gk; nf = (int)B->nfk;
        fi = a + B->nic * (c + B->njc * (side ? B->nkc : 0));
    }
}

// Ghost-cell state for a stencil position outside the block (m < 0 or m >= n).
// Depth d = -m-1 on the minus side and m-n on the plus side (block.cu:272-316).
__device__ __noinline__ Prim ghost_prim(const StageArgs& A, const BlockDev* B, int dir, int m,
                                        int a, int c)
{
    int n = dir == 0 ? B->nic : (dir == 1 ? B->njc : B->nkc);
    int side = (m < 0) ? 0 : 1;
    int depth = side ? m - n : -m - 1;
    int bf = 2 * dir + side;
    int code = B->bc[bf];
    if (code == 3) {  // exchange: neighbour's interior cell at the same depth (bcs.cu:135-254)
        const BlockDev* O = A.blocks + B->bc_arg[bf];
        int no = dir == 0 ? O->nic : (dir == 1 ? O->njc : O->nkc);
        int mo = side ? depth : no - 1 - depth;
        return ld_prim(O->W[A.win], (int)O->n, cell_index(O->nic, O->njc, dir, mo, a, c));
    }
    if (code == 4) {  // inflow: the configured flow state (bcs.cu:257-319)
        const double* s = A.inflow + 6 * B->bc_arg[bf];
        Prim p;
        p.rho = s[0]; p.e = s[1]; p.YB = s[2]; p.vx = s[3]; p.vy = s[4]; p.vz = s[5];
        return p;
    }
    Prim in = ld_prim(B->W[A.win], (int)B->n,
                      cell_index(B->nic, B->njc, dir, side ? n - 1 : 0, a, c));
    if (code == 5) return in;  // outflow copies the adjacent interior cell (bcs.cu:435-503)
    // Walls: both ghost layers reflect the adjacent interior cell (bcs.cu:16-132).
    const double* g; int nf, fi;
    boundary_face(B, dir, side, a, c, g, nf, fi);
    return reflect(in, ld_frame(g, nf, fi));
}

__device__ __forceinline__ Prim stencil_prim(const StageArgs& A, const BlockDev* B,
                                             const double* __restrict__ W, int n_line, int base,
                                             int stride, int dir, int m, int a, int c)
{
    if (m >= 0 && m < n_line) return ld_prim(W, (int)B->n, base + m * stride);
    return ghost_prim(A, B, dir, m, a, c);
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}

template <int STAGE, int TI, int TJ, int XORDER, bool DTRED>
__device__ __forceinline__ void process_tile(const StageArgs& A, const Tile tile, double* lds)
{
    constexpr int NT = TI * TJ;
    constexpr int NFI = (TI + 1) * TJ;
    constexpr int NFJ = TI * (TJ + 1);
    constexpr int NFK = TI * TJ;
    double* AFi = lds;                      // [6][TJ][TI+1]
    double* AFj = AFi + 6 * NFI;            // [6][TJ+1][TI]
    double* AFk = AFj + 6 * NFJ;            // [2][6][TJ][TI]

    const BlockDev* B = A.blocks + tile.blk;
    const int nic = B->nic, njc = B->njc;
    const int n = (int)B->n;
    const double* __restrict__ Win = B->W[A.win];
    const int t = threadIdx.x;
    const int ti = t % TI, tj = t / TI;
    const int i = tile.i0 + ti, j = tile.j0 + tj;
    const bool own = (i < nic) && (j < njc);
    const int kbeg = tile.k0, kend = tile.k0 + tile.nk;
    const double dt = A.ctl->dt;
    int bad = 0;
    unsigned long long sigmin = ~0ull;

    for (int kk = kbeg - 1; kk < kend; ++kk) {
        const bool plane = (kk >= kbeg);
        const int nfaces = (plane ? NFI + NFJ : 0) + NFK;
        const int fofs = plane ? 0 : NFI + NFJ;
        double* AFkp = AFk + ((kk + 1) & 1) * 6 * NFK;
        for (int f = t; f < nfaces; f += NT) {
            const int fg = f + fofs;
            int dir, m0, a, c, base, stride, n_line, nf, fi;
            const double* g;
            double* dst;
            int lds_stride;
            bool valid;
            if (fg < NFI) {  // i-face (gi, gj, kk) between cells gi-1 and gi
                const int fa = fg % (TI + 1), fb = fg / (TI + 1);
                const int gi = tile.i0 + fa, gj = tile.j0 + fb;
                valid = (gi <= nic) && (gj < njc);
                dir = 0; m0 = gi; a = gj; c = kk;
                base = nic * (gj + njc * kk); stride = 1; n_line = nic;
                g = B->gi; nf = (int)B->nfi; fi = gi + (nic + 1) * (gj + njc * kk);
                dst = AFi + fb * (TI + 1) + fa; lds_stride = NFI;
            } else if (fg < NFI + NFJ) {  // j-face (gi, gj, kk)
                const int h = fg - NFI;
                const int fa = h % TI, fb = h / TI;
                const int gi = tile.i0 + fa, gj = tile.j0 + fb;
                valid = (gi < nic) && (gj <= njc);
                dir = 1; m0 = gj; a = gi; c = kk;
                base = gi + nic * njc * kk; stride = nic; n_line = njc;
                g = B->gj; nf = (int)B->nfj; fi = gi + nic * (gj + (njc + 1) * kk);
                dst = AFj + fb * TI + fa; lds_stride = NFJ;
            } else {  // k-face (gi, gj, kk+1) of the thread-column layout
                const int h = fg - NFI - NFJ;
                const int fa = h % TI, fb = h / TI;
                const int gi = tile.i0 + fa, gj = tile.j0 + fb;
                valid = (gi < nic) && (gj < njc);
                dir = 2; m0 = kk + 1; a = gi; c = gj;
                base = gi + nic * gj; stride = nic * njc; n_line = B->nkc;
                g = B->gk; nf = (int)B->nfk; fi = gi + nic * (gj + njc * (kk + 1));
                dst = AFkp + fb * TI + fa; lds_stride = NFK;
            }
            if (valid) {
                Prim s0 = stencil_prim(A, B, Win, n_line, base, stride, dir, m0 - 2, a, c);
                Prim s1 = stencil_prim(A, B, Win, n_line, base, stride, dir, m0 - 1, a, c);
                Prim s2 = stencil_prim(A, B, Win, n_line, base, stride, dir, m0, a, c);
                Prim s3 = stencil_prim(A, B, Win, n_line, base, stride, dir, m0 + 1, a, c);
                const Frame fr = ld_frame(g, nf, fi);
                const double area = g[fi];
                double F[6];
                face_flux<XORDER>(s0, s1, s2, s3, fr, F);
#pragma unroll
                for (int q = 0; q < 6; ++q) dst[q * lds_stride] = area * F[q];
            }
        }
        __syncthreads();
        if (plane && own) {
            const double* AFkm = AFk + (kk & 1) * 6 * NFK;
            const int cidx = i + nic * (j + njc * kk);
            const double vinv = B->volinv[cidx];
            double d[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                double s = AFi[q * NFI + tj * (TI + 1) + ti] - AFi[q * NFI + tj * (TI + 1) + ti + 1];
                s = s + AFj[q * NFJ + tj * TI + ti];
                s = s - AFj[q * NFJ + (tj + 1) * TI + ti];
                s = s + AFkm[q * NFK + tj * TI + ti];
                s = s - AFkp[q * NFK + tj * TI + ti];
                d[q] = vinv * s;
            }
            double* __restrict__ U = B->U;
            double* __restrict__ L = B->L;
            double Un[6];
            if (STAGE == 1) {
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    Un[q] = U[q * n + cidx] + dt * d[q];
                    L[q * n + cidx] = d[q];
                }
            } else if (STAGE == 2) {
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const double l0 = L[q * n + cidx];
                    Un[q] = U[q * n + cidx] + 0.25 * dt * (l0 + d[q]);
                    L[q * n + cidx] = kOneSixth * l0 + kOneSixth * d[q];
                }
            } else {
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    Un[q] = U[q * n + cidx] + dt * (L[q * n + cidx] + kFourSixth * d[q]);
                    U[q * n + cidx] = Un[q];
                }
            }
            Prim w;
            const int flag = decode(Un, w);
            bad += flag;
            double* __restrict__ Wout = B->W[A.wout];
            Wout[cidx] = w.rho;
            Wout[n + cidx] = w.e;
            Wout[2 * n + cidx] = w.YB;
            Wout[3 * n + cidx] = w.vx;
            Wout[4 * n + cidx] = w.vy;
            Wout[5 * n + cidx] = w.vz;
            if (DTRED && !flag) {
                const double T = eos_T(w.e, w.YB);
                const double sa = eos_a(T);
                const Frame fi_ = ld_frame(B->gi, (int)B->nfi, i + (nic + 1) * (j + njc * kk));
                const Frame fj_ = ld_frame(B->gj, (int)B->nfj, i + nic * (j + (njc + 1) * kk));
                const Frame fk_ = ld_frame(B->gk, (int)B->nfk, cidx);
                const double sig = local_signal(w, sa, B->len[cidx], B->len[n + cidx],
                                                B->len[2 * n + cidx], fi_, fj_, fk_);
                const unsigned long long bits = (unsigned long long)__double_as_longlong(sig);
                sigmin = bits < sigmin ? bits : sigmin;
            }
        }
        __syncthreads();
    }
    // Bad-cell count of the stage (decode flags, flow.cu:249-252).
    if (__ballot(bad != 0)) {
        int cnt = bad;
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(&A.ctl->bad[STAGE - 1], cnt);
    }
    if (DTRED) {
        const unsigned long long wm = wave_min_u64(sigmin);
        if ((threadIdx.x & 63) == 0 && wm != ~0ull) atomicMin(&A.ctl->sigmin_bits, wm);
    }
}

template <int STAGE, int TI, int TJ, int XORDER, bool DTRED>
__global__ __launch_bounds__(TI* TJ) void stage_kernel(StageArgs A, int persistent)
{
    constexpr int LDS_DOUBLES = 6 * ((TI + 1) * TJ + TI * (TJ + 1) + 2 * TI * TJ);
    __shared__ double lds[LDS_DOUBLES];
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
        process_tile<STAGE, TI, TJ, XORDER, DTRED>(A, tile, lds);
        if (!persistent) break;
    }
}

template <int TI, int TJ, int XORDER>
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
        hipLaunchKernelGGL((stage_kernel<1, TI, TJ, XORDER, false>), dim3(grid), blk, 0, s, a, p);
    else if (stage == 2)
        hipLaunchKernelGGL((stage_kernel<2, TI, TJ, XORDER, false>), dim3(grid), blk, 0, s, a, p);
    else if (dtred)
        hipLaunchKernelGGL((stage_kernel<3, TI, TJ, XORDER, true>), dim3(grid), blk, 0, s, a, p);
    else
        hipLaunchKernelGGL((stage_kernel<3, TI, TJ, XORDER, false>), dim3(grid), blk, 0, s, a, p);
    return hipGetLastError();
}

template <int TI, int TJ>
hipError_t launch_order(int stage, int x_order, bool dtred, bool persistent, int grid_limit,
                        const StageArgs& a, hipStream_t s)
{
    if (x_order > 1) return launch_shape<TI, TJ, 2>(stage, dtred, persistent, grid_limit, a, s);
    return launch_shape<TI, TJ, 1>(stage, dtred, persistent, grid_limit, a, s);
}

__global__ __launch_bounds__(256) void signal_min_kernel(const BlockDev* B, StepCtl* ctl)
{
    const int n = (int)B->n;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = ~0ull;
    if (c < n) {
        const int nic = B->nic, njc = B->njc;
        const int i = c % nic, j = (c / nic) % njc, k = c / (nic * njc);
        const Prim w = ld_prim(B->W[0], n, c);
        const double sa = eos_a(eos_T(w.e, w.YB));
        const Frame fi_ = ld_frame(B->gi, (int)B->nfi, i + (nic + 1) * (j + njc * k));
        const Frame fj_ = ld_frame(B->gj, (int)B->nfj, i + nic * (j + (njc + 1) * k));
        const Frame fk_ = ld_frame(B->gk, (int)B->nfk, c);
        const double sig = local_signal(w, sa, B->len[c], B->len[n + c], B->len[2 * n + c], fi_, fj_, fk_);
        v = (unsigned long long)__double_as_longlong(sig);
    }
    v = wave_min_u64(v);
    if ((threadIdx.x & 63) == 0 && v != ~0ull) atomicMin(&ctl->sigmin_bits, v);
}

}  // namespace

hipError_t launch_signal_min(const BlockDev* block, int ncells, StepCtl* ctl, hipStream_t s)
{
    if (ncells <= 0) return hipSuccess;
    hipLaunchKernelGGL(signal_min_kernel, dim3((ncells + 255) / 256), dim3(256), 0, s, block, ctl);
    return hipGetLastError();
}

bool tile_shape_supported(int ti, int tj)
{
    return (ti == 32 && tj == 8) || (ti == 64 && tj == 4) || (ti == 16 && tj == 16) ||
           (ti == 64 && tj == 8);
}

hipError_t launch_stage(int stage, int x_order, int ti, int tj, bool dt_reduce, bool persistent,
                        int grid_limit, const StageArgs& a, hipStream_t s)
{
    if (a.ntiles <= 0) return hipSuccess;
    if (ti == 32 && tj == 8) return launch_order<32, 8>(stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    if (ti == 64 && tj == 4) return launch_order<64, 4>(stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    if (ti == 16 && tj == 16) return launch_order<16, 16>(stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    if (ti == 64 && tj == 8) return launch_order<64, 8>(stage, x_order, dt_reduce, persistent, grid_limit, a, s);
    return hipErrorInvalidValue;
}

}  // namespace chkn
