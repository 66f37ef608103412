// chkn_adjoint.hip -- gfx950 kernels of the discrete adjoint of the chicken step (FP64).
//
// One reversed RK stage s with stage state W_s (read from the tape) and residual adjoint D
// (the adjoint of dU/dt of that stage, per cell):
//   1. face pass, per patch of 16 x 16 faces of one direction (16 consecutive faces along 16
//      lines): every face forms Fbar = A (D_R/V_R - D_L/V_L) from its two adjacent cells
//      (ghost sides contribute nothing) and the vector-Jacobian product of its flux
//      (face_flux_vjp) with respect to its four stencil states; the 24 values go to LDS, and
//      each of the patch's 13 owned cells per line sums the four that reach it (faces m-1,
//      m, m+1, m+2 as R1, R0, L0, L1) into Gc[dir].  The thread of a line's first or last
//      cell also sums the values that land on the block's ghost cells of that line, per
//      ghost depth, into GB;
//   2. gather: every cell adds its three direction sums, the ghost-line adjoints that come
//      back to it (wall and outflow ghosts are copies of it; exchange ghosts of other blocks
//      read it), maps the primitive adjoint to the conserved one (decode_vjp) and applies the
//      transpose of the TVD-RK3 stage combination (block.cu:1090-1160).  A cell next to an
//      inflow boundary also adds the adjoint of its line of inflow ghosts into that line's
//      accumulator (the gradient of the inflow state).
// Fused form (no block exchanges through its k sides): the i and j patches run first (step 1
// for those directions), then the k patches, whose owned cells end with step 2 for themselves
// -- the i and j sums from memory, the k sum and the k ghost lines from the patch's own LDS --
// with the same operations in the same order as the separate gather.  A k patch reads D of the
// cells its neighbouring patches finish, so D is double-buffered: a stage reads it from one
// buffer (A.dr) and its cells write the next stage's into the other (A.dw).
// No atomics: every sum has one writer and a fixed order, so the gradient is deterministic.
// Gradients need no bit-exact arithmetic, so this file is built with floating-point
// contraction, reciprocal arithmetic and approximate (Newton-refined reciprocal) division
// allowed (the stage kernels are not); the helpers the gather and the fused k pass share turn
// contraction off, so that the two forms round alike and give the same gradient bit for bit.
//
// Reversed TVD-RK3 step with A0 = adjoint of U^{n+1} on entry (seed: P = dt/6 A0, D = 4dt/6 A0;
// stage 3 forms that P from A0 itself, so the stage-1 seeds store D only):
//   stage 3 (W2): A0 += U2bar; D = P + dt/4 U2bar; P = D
//   stage 2 (W1): A0 += U1bar; D = P + dt U1bar
//   stage 1 (W0): A0 += U0bar        -> A0 = adjoint of U^n; seed of the next reversed step
// Reversed predictor-corrector step (U1 = U0 + dt L0, U = U0 + dt/2 (L0 + L1); seed: P = D =
// dt/2 A0): the corrector (W1) as stage 2 above, A0 += U1bar, D = P + dt U1bar; the
// predictor (W0) as stage 1.
#include <mutex>
#include <vector>

#include "../../include/chkn.h"
#include "chkn_adjoint.h"
#include "chkn_internal.h"
#include "chkn_physics.h"
#include "chkn_stencil.h"

namespace chkn {

namespace {

// Staging of the allocation-free point evaluation (flux_vjp_points_noalloc): the inputs
// (39 values per point) then the outputs (24 per point) of up to kNoAllocPoints points, in
// a file-scope device array (2 MB) that each GPU the library runs on holds until the process
// ends.  Every call on a GPU shares that GPU's copy: calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[(39 + 24) * kNoAllocPoints];

__device__ __forceinline__ int line_len(const BlockDev* B, int d)
{
    return d == 0 ? B->nic : (d == 1 ? B->njc : B->nkc);
}

__device__ __forceinline__ long long nfaces(const BlockDev* B, int d)
{
    return d == 0 ? B->nfi : (d == 1 ? B->nfj : B->nfk);
}

// Face index of line position m (face between cells m-1 and m) for transverse (a, c).
__device__ __forceinline__ int face_index(const BlockDev* B, int d, int m, int a, int c)
{
    if (d == 0) return m + (B->nic + 1) * (a + B->njc * c);
    if (d == 1) return a + B->nic * (m + (B->njc + 1) * c);
    return a + B->nic * (c + B->njc * m);
}

__device__ __forceinline__ const double* geom(const BlockDev* B, int d)
{
    return d == 0 ? B->gi : (d == 1 ? B->gj : B->gk);
}

// Ghost-line buffer of side `side` of a block: [2 depths][6][nt] per side, sides in order.
__device__ __forceinline__ int side_lines(const BlockDev* B, int side)
{
    const int d = side >> 1;
    return d == 0 ? B->njc * B->nkc : (d == 1 ? B->nic * B->nkc : B->nic * B->njc);
}
__device__ __forceinline__ double* ghost_buffer(const BlockDev* B, int side)
{
    long long off = 0;
    for (int s = 0; s < side; ++s) off += 12LL * side_lines(B, s);
    return B->GB + off;
}

// The stencil adjoints of the face at line position m (between cells m-1 and m) of the line
// (a, c) in direction d, left in column t of st: st[6p+q][t] = adjoint of stencil slot p
// (L1, L0, R0, R1), quantity q.  Zero for faces outside the block.
// FLUX 0: AUSMDV on MUSCL (XORDER 2) or adjacent states; FLUX 1: the SBP-ASF flux.
// KEEP: the stencil's R1 state (cell m + 1) is also returned in r1, and the function returns
// whether it loaded the stencil at all (faces with zero Fbar do not).
template <int XORDER, int FLUX, bool KEEP = false>
__device__ __forceinline__ bool face_vjp(const AdjArgs& A, const BlockDev* B, int d, int m, int a,
                                         int c, double (&st)[24][kAdjPatchM * kAdjPatchT], int t, Prim& r1)
{
    const int nline = line_len(B, d);
    const int na = d == 0 ? B->njc : B->nic;
    const int nc = d == 2 ? B->njc : B->nkc;
    const bool valid = m >= 0 && m <= nline && a < na && c < nc;
    double Fb[6] = {0, 0, 0, 0, 0, 0};
    bool any = false;
    const long long n = B->n;
    const int nic = B->nic, njc = B->njc;
    const long long nf = nfaces(B, d);
    const double* g = geom(B, d);
    const int fi = valid ? face_index(B, d, m, a, c) : 0;
    if (valid) {
        const double* D = B->adj + (long long)A.dr * n;
        const double area = g[fi];
        const int cl = (m - 1 >= 0) ? cell_index(nic, njc, d, m - 1, a, c) : -1;
        const int cr = (m < nline) ? cell_index(nic, njc, d, m, a, c) : -1;
        const double vl = cl >= 0 ? B->volinv[cl] : 0.0, vr = cr >= 0 ? B->volinv[cr] : 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double dl = cl >= 0 ? D[(long long)q * n + cl] : 0.0;
            const double dr = cr >= 0 ? D[(long long)q * n + cr] : 0.0;
            Fb[q] = area * (vr * dr - vl * dl);
            any = any || (Fb[q] != 0.0);
        }
    }
    if (!any) {
#pragma unroll
        for (int k = 0; k < 24; ++k) st[k][t] = 0.0;
        return false;
    }
    const double* W = B->W + A.slot * B->wstride;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int pos = m - 2 + p;
        const Prim sp = (pos >= 0 && pos < nline) ? ld_prim(W, n, cell_index(nic, njc, d, pos, a, c))
                                                  : ghost_prim(A.blocks, A.inflow, A.slot, B, d, pos, a, c);
        st[6 * p + 0][t] = sp.rho;
        st[6 * p + 1][t] = sp.e;
        st[6 * p + 2][t] = sp.YB;
        st[6 * p + 3][t] = sp.vx;
        st[6 * p + 4][t] = sp.vy;
        st[6 * p + 5][t] = sp.vz;
        if (KEEP && p == 3) r1 = sp;
    }
    // Compiler fence: the stencil values are read back from LDS where they are needed rather
    // than forwarded from the stores above and held in registers through the flux adjoint.
    __asm__ __volatile__("" ::: "memory");
    if (FLUX == 1) {
        Prim sp[4];
#pragma unroll
        for (int p = 0; p < 4; ++p)
            sp[p] = Prim{st[6 * p][t], st[6 * p + 1][t], st[6 * p + 2][t], st[6 * p + 3][t], st[6 * p + 4][t],
                         st[6 * p + 5][t]};
        double bs[4][6];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 6; ++q) bs[p][q] = 0.0;
        sbp_asf_vjp(sp, ld_frame(g, nf, fi), Fb, bs);
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 6; ++q) st[6 * p + q][t] = bs[p][q];
        return true;
    }
    // Reconstructed face states (face.cu:347-371).
    double qL[6], qR[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        if (XORDER > 1) {
            muscl(st[q][t], st[6 + q][t], st[12 + q][t], st[18 + q][t], qL[q], qR[q]);
        } else {
            qL[q] = st[6 + q][t];
            qR[q] = st[12 + q][t];
        }
    }
    const Frame fr = ld_frame(g, nf, fi);
    const double TL = eos_T(qL[1], qL[2]), TR = eos_T(qR[1], qR[2]);
    const double pL = eos_p(qL[0], TL), pR = eos_p(qR[0], TR);
    const double aL = eos_a(TL), aR = eos_a(TR);
    double bL[8] = {0, 0, 0, 0, 0, 0, 0, 0}, bR[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    ausmdv_vjp(qL[0], qL[1], pL, aL, qL[2], qL[3], qL[4], qL[5], qR[0], qR[1], pR, aR, qR[2], qR[3],
               qR[4], qR[5], fr, Fb, bL, bR);
    double side[2][6];
    {
        const double bT = bL[2] * qL[0] * kR + bL[3] * kGamma * kR / (2.0 * aL);
        side[0][0] = bL[0] + bL[2] * kR * TL;
        side[0][1] = bL[1] + bT / kCv;
        side[0][2] = bL[4] + bT * kQ / kCv;
        side[0][3] = bL[5]; side[0][4] = bL[6]; side[0][5] = bL[7];
    }
    {
        const double bT = bR[2] * qR[0] * kR + bR[3] * kGamma * kR / (2.0 * aR);
        side[1][0] = bR[0] + bR[2] * kR * TR;
        side[1][1] = bR[1] + bT / kCv;
        side[1][2] = bR[4] + bT * kQ / kCv;
        side[1][3] = bR[5]; side[1][4] = bR[6]; side[1][5] = bR[7];
    }
    __asm__ __volatile__("" ::: "memory");
    // Each quantity's four stencil values are read before its four adjoints replace them.
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        double gq[4] = {0.0, 0.0, 0.0, 0.0};
        if (XORDER > 1) {
            muscl_vjp(st[q][t], st[6 + q][t], st[12 + q][t], st[18 + q][t], side[0][q], side[1][q], gq);
        } else {
            gq[1] = side[0][q];
            gq[2] = side[1][q];
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) st[6 * p + q][t] = gq[p];
    }
    return true;
}

// What the ghosts of one side of a line return to the cell next to that side, gb = the sum of
// both ghost depths: wall and outflow ghosts are copies of the cell (walls through the
// reflection), inflow ghosts hold the inflow state, whose line accumulator this cell owns.
__device__ __forceinline__ void own_side_return(const AdjArgs& A, const BlockDev* B, int d, int side, int a,
                                                int c, int line, const double gb[6], double Wb[6])
{
#pragma clang fp contract(off)
    const int code = B->bc[2 * d + side];
    if (code == CHKN_BC_INFLOW) {
        const long long col = B->icol + B->icol_side[2 * d + side] + line;
#pragma unroll
        for (int q = 0; q < 6; ++q) A.inflow_acc[(long long)q * A.nicols + col] += gb[q];
    } else if (code == CHKN_BC_OUTFLOW) {
#pragma unroll
        for (int q = 0; q < 6; ++q) Wb[q] += gb[q];
    } else {
        const double* g; int nf, fi;
        boundary_face(B, d, side, a, c, g, nf, fi);
        reflect_vjp(gb, ld_frame(g, nf, fi), Wb);
    }
}

__device__ __forceinline__ bool returns_to_cell(int code)
{
    return code != CHKN_BC_EXCHANGE && code != CHKN_BC_INFLOW_FUNCTION;
}

// The ghost-line adjoints of direction d that come back to cell ijk, from the ghost buffers:
// the block's own sides, then the exchange ghosts of other blocks (or this one, periodically)
// that read the cell -- the plus-side ghosts of P at depth m read cells m = 0, 1; the
// minus-side ghosts of Q at depth nline-1-m read cells m = nline-1, nline-2.
__device__ __forceinline__ void line_returns(const AdjArgs& A, const BlockDev* B, int d, const int ijk[3],
                                             double Wb[6])
{
#pragma clang fp contract(off)
    const int m = ijk[d];
    const int a = d == 0 ? ijk[1] : ijk[0];
    const int c = d == 2 ? ijk[1] : ijk[2];
    const int nline = line_len(B, d);
    const int line = a + (d == 0 ? B->njc : B->nic) * c;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        if (m != (side ? nline - 1 : 0) || !returns_to_cell(B->bc[2 * d + side])) continue;
        const double* gbuf = ghost_buffer(B, 2 * d + side);
        const int nt = side_lines(B, 2 * d + side);
        double gb[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) gb[q] = gbuf[(long long)q * nt + line] + gbuf[(long long)(6 + q) * nt + line];
        own_side_return(A, B, d, side, a, c, line, gb, Wb);
    }
    const int P = B->rp[2 * d];
    if (P >= 0 && m <= 1) {
        const BlockDev* O = A.blocks + P;
        const double* gbuf = ghost_buffer(O, 2 * d + 1) + (long long)(6 * m) * side_lines(O, 2 * d + 1);
        const int nt = side_lines(O, 2 * d + 1);
#pragma unroll
        for (int q = 0; q < 6; ++q) Wb[q] += gbuf[(long long)q * nt + line];
    }
    const int Q = B->rp[2 * d + 1];
    if (Q >= 0 && m >= nline - 2) {
        const BlockDev* O = A.blocks + Q;
        const int depth = nline - 1 - m;
        const double* gbuf = ghost_buffer(O, 2 * d) + (long long)(6 * depth) * side_lines(O, 2 * d);
        const int nt = side_lines(O, 2 * d);
#pragma unroll
        for (int q = 0; q < 6; ++q) Wb[q] += gbuf[(long long)q * nt + line];
    }
}

// What a cell's stage combination reads besides its adjoint: the stage state and A0, P.
struct CellIn {
    Prim w;
    double a0[6], p[6];
};

// P is read at stage 2 only: stage 3 (TVD-RK3's first reversed stage) forms it from A0.
template <int STAGE>
__device__ __forceinline__ void load_cell(const AdjArgs& A, const BlockDev* B, int cidx, CellIn& c,
                                          bool load_w = true)
{
    const long long n = B->n;
    const unsigned o = boff(cidx);
    if (load_w) c.w = ld_prim(B->W + A.slot * B->wstride, n, cidx);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        c.a0[q] = gld(B->adj + (long long)q * n, o);
        c.p[q] = STAGE == 2 ? gld(B->adj + (long long)(6 + q) * n, o) : 0.0;
    }
}

// The cell's adjoint of W_s to that of U_s (decode_vjp), then the transposed stage
// combination (block.cu:1090-1160); D is written to the buffer A.dw.
template <int STAGE>
__device__ __forceinline__ void finish_cell(const AdjArgs& A, const BlockDev* B, int cidx, const double Wb[6],
                                            const CellIn& c)
{
#pragma clang fp contract(off)
    const long long n = B->n;
    const unsigned o = boff(cidx);
    double Ub[6];
    decode_vjp(c.w, Wb, Ub);
    const double dt = A.dt[B->inst];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        double* A0 = B->adj + (long long)q * n;
        double* Pp = B->adj + (long long)(6 + q) * n;
        double* D = B->adj + (long long)(A.dw + q) * n;
        const double a0 = c.a0[q] + Ub[q];
        gst(A0, o, a0);
        if (STAGE == 3) {
            // The step's seed P = seed_p dt A0 (as the seeds form it), A0 before this stage.
            const double v = A.seed_p * dt * c.a0[q] + 0.25 * dt * Ub[q];
            gst(D, o, v);
            gst(Pp, o, v);
        } else if (STAGE == 2) {
            gst(D, o, c.p[q] + dt * Ub[q]);
        } else if (A.dt_next) {
            const double dtn = A.dt_next[B->inst];
            if (A.seed_p_stored) gst(Pp, o, A.seed_p * dtn * a0);
            gst(D, o, A.seed_d * dtn * a0);
        }
    }
}

// Face pass over one patch (see AdjPatch).  Lanes run along i in every direction so that the
// stencil, D and geometry loads of a wave are contiguous: in direction 0 a lane is a face
// position along the line, in directions 1 and 2 a line.
// Three work-groups per CU fit the LDS (48 KB each), three waves per SIMD.
// EPI 0: the owned cells' sums go to Gc[dir] and the ghost lines' to GB.  EPI = s (1..3), k
// patches of the fused form only: each owned cell finishes stage s itself (see the header).
template <int XORDER, int FLUX, int EPI>
__global__ __launch_bounds__(kAdjPatchM * kAdjPatchT) void adj_patch_kernel(AdjArgs Ain, const BlockDev* __restrict__ blocks,
                                                                          const AdjPatch* __restrict__ patches)
{
    // Read-only tables as restrict-qualified arguments: scalar loads no store can clobber.
    AdjArgs A = Ain;
    A.blocks = blocks;
    A.patches = patches;
    __shared__ double st[24][kAdjPatchM * kAdjPatchT];
    const int t = (int)threadIdx.x;
    const AdjPatch P = A.patches[blockIdx.x];
    const BlockDev* B = A.blocks + P.blk;
    const int d = EPI > 0 ? 2 : P.dir;
    const int tm = d == 0 ? t % kAdjPatchM : t / kAdjPatchT;  // face position in the patch
    const int ta = d == 0 ? t / kAdjPatchM : t % kAdjPatchT;  // line in the patch
    const int a = P.a0 + ta, c = P.c;
    // In the fused form the face at position tm (face m - 1) has the owned cell m as its R1
    // stencil state: kept from the stencil load instead of read again.
    Prim own_w;
    const bool have_w = face_vjp<XORDER, FLUX, (EPI > 0)>(A, B, d, P.m0 - 1 + tm, a, c, st, t, own_w);
    // Owned cell m = m0 + tm of line (a, c): faces m-1 .. m+2 sit at positions tm .. tm+3.
    const int nline = line_len(B, d);
    const int na = d == 0 ? B->njc : B->nic;
    const int m = P.m0 + tm;
    const bool own = tm < kAdjPatchOwn && m < nline && a < na;
    const long long n = B->n;
    const int cidx = own ? cell_index(B->nic, B->njc, d, m, a, c) : 0;
    // Fused form: the owned cell's i and j sums and its stage inputs are loaded before the
    // barrier, while the registers of the flux adjoint are free and other waves finish theirs.
    double gij[6];
    CellIn cin;
    if constexpr (EPI > 0) {
        if (own) {
            const unsigned o = boff(cidx);
#pragma unroll
            for (int q = 0; q < 6; ++q)
                gij[q] = gld_nt(B->Gc + (long long)q * n, o) + gld_nt(B->Gc + (long long)(6 + q) * n, o);
            load_cell<EPI>(A, B, cidx, cin, !have_w);
            if (have_w) cin.w = own_w;
        }
    }
    __syncthreads();
    if (!own) return;
    auto lane = [&](int pos) { return d == 0 ? pos + kAdjPatchM * ta : ta + kAdjPatchT * pos; };
    const int l3 = lane(tm), l2 = lane(tm + 1), l1 = lane(tm + 2), l0 = lane(tm + 3);
    const int line = a + na * c;
    if constexpr (EPI > 0) {
        // The cell (i, j, k) = (a, c, m): its three direction sums, the ghost lines of i and j
        // from the buffers, those of k from this patch, then the stage combination.
        double Wb[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) Wb[q] = gij[q] + (((st[18 + q][l3] + st[12 + q][l2]) + st[6 + q][l1]) + st[q][l0]);
        const int ijk[3] = {a, c, m};
        line_returns(A, B, 0, ijk, Wb);
        line_returns(A, B, 1, ijk, Wb);
        if (m == 0 && returns_to_cell(B->bc[4])) {
            const int f0 = lane(tm + 1), f1 = lane(tm + 2);
            double gb[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) gb[q] = (st[6 + q][f0] + st[q][f1]) + st[q][f0];
            own_side_return(A, B, 2, 0, a, c, line, gb, Wb);
        }
        if (m == nline - 1 && returns_to_cell(B->bc[5])) {
            const int fa = lane(tm + 1), fb = lane(tm + 2);
            double gb[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) gb[q] = (st[18 + q][fa] + st[12 + q][fb]) + st[18 + q][fb];
            own_side_return(A, B, 2, 1, a, c, line, gb, Wb);
        }
        finish_cell<EPI>(A, B, cidx, Wb, cin);
        return;
    }
    double* Gd = B->Gc + 6 * (long long)d * n;
#pragma unroll
    for (int q = 0; q < 6; ++q)
        gst_nt(Gd + (long long)q * n, boff(cidx), ((st[18 + q][l3] + st[12 + q][l2]) + st[6 + q][l1]) + st[q][l0]);
    // Ghost lines: what lands on the ghosts at depth 0 and 1 of the line's minus side (faces 0
    // and 1 as L0, L1 and face 0 as L1) and plus side (faces n-1 and n as R1, R0 and face n as
    // R1).
    if (m == 0) {
        const int f0 = lane(tm + 1), f1 = lane(tm + 2);
        double* gb = ghost_buffer(B, 2 * d);
        const int nt = side_lines(B, 2 * d);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            gb[(long long)q * nt + line] = st[6 + q][f0] + st[q][f1];
            gb[(long long)(6 + q) * nt + line] = st[q][f0];
        }
    }
    if (m == nline - 1) {
        const int fa = lane(tm + 1), fb = lane(tm + 2);
        double* gb = ghost_buffer(B, 2 * d + 1);
        const int nt = side_lines(B, 2 * d + 1);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            gb[(long long)q * nt + line] = st[18 + q][fa] + st[12 + q][fb];
            gb[(long long)(6 + q) * nt + line] = st[18 + q][fb];
        }
    }
}

template <int STAGE>
__global__ __launch_bounds__(256) void adj_gather_kernel(AdjArgs Ain, const BlockDev* __restrict__ blocks,
                                                         const CellChunk* __restrict__ cells)
{
    AdjArgs A = Ain;
    A.blocks = blocks;
    A.cells = cells;
    const CellChunk ch = A.cells[blockIdx.x];
    const BlockDev* B = A.blocks + ch.blk;
    const long long n = B->n;
    const int cidx = ch.c0 + (int)threadIdx.x;
    if (cidx >= ch.c1) return;
    const int nic = B->nic, njc = B->njc;
    const int ijk[3] = {cidx % nic, (cidx / nic) % njc, cidx / (nic * njc)};
    const unsigned o = boff(cidx);
    double Wb[6];
#pragma unroll
    for (int q = 0; q < 6; ++q)
        Wb[q] = (gld_nt(B->Gc + (long long)q * n, o) + gld_nt(B->Gc + (long long)(6 + q) * n, o)) +
                gld_nt(B->Gc + (long long)(12 + q) * n, o);
#pragma unroll
    for (int d = 0; d < 3; ++d) line_returns(A, B, d, ijk, Wb);
    CellIn cin;
    load_cell<STAGE>(A, B, cidx, cin);
    finish_cell<STAGE>(A, B, cidx, Wb, cin);
}

__global__ __launch_bounds__(256) void adj_seed_kernel(AdjArgs Ain, const BlockDev* __restrict__ blocks,
                                                       const CellChunk* __restrict__ cells)
{
    AdjArgs A = Ain;
    A.blocks = blocks;
    A.cells = cells;
    const CellChunk ch = A.cells[blockIdx.x];
    const BlockDev* B = A.blocks + ch.blk;
    const long long n = B->n;
    const int cidx = ch.c0 + (int)threadIdx.x;
    if (cidx >= ch.c1) return;
    const double dt = A.dt[B->inst];
    const double* A0 = B->adj;
    double* Pp = B->adj + 6 * (long long)n;
    double* D = B->adj + (long long)A.dr * n;  // read by the next face pass
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const long long o = (long long)q * n + cidx;
        Pp[o] = A.seed_p * dt * A0[o];
        D[o] = A.seed_d * dt * A0[o];
    }
}

__global__ __launch_bounds__(256) void adj_chem_seed_kernel(AdjArgs Ain, const BlockDev* __restrict__ blocks,
                                                            const CellChunk* __restrict__ cells)
{
    AdjArgs A = Ain;
    A.blocks = blocks;
    A.cells = cells;
    const CellChunk ch = A.cells[blockIdx.x];
    const BlockDev* B = A.blocks + ch.blk;
    const long long n = B->n;
    const int cidx = ch.c0 + (int)threadIdx.x;
    if (cidx >= ch.c1) return;
    const double dt = A.dt[B->inst];
    double* A0 = B->adj;
    double* Pp = B->adj + 6 * (long long)n;
    double* D = B->adj + (long long)A.dr * n;  // read by the next face pass
    double Ub[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) Ub[q] = A0[(long long)q * n + cidx];
    chem_vjp(ld_prim(B->W + A.slot * B->wstride, n, cidx), A.chem[B->inst], Ub);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const long long o = (long long)q * n + cidx;
        A0[o] = Ub[q];
        Pp[o] = A.seed_p * dt * Ub[q];
        D[o] = A.seed_d * dt * Ub[q];
    }
}

__global__ __launch_bounds__(64) void flux_vjp_points_kernel(const double* in, double* out, int npts,
                                                            int x_order)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    const double* p = in + 39 * (long long)t;
    Prim s[4];
    for (int c = 0; c < 4; ++c)
        s[c] = Prim{p[6 * c], p[6 * c + 1], p[6 * c + 2], p[6 * c + 3], p[6 * c + 4], p[6 * c + 5]};
    const Frame f{p[24], p[25], p[26], p[27], p[28], p[29], p[30], p[31], p[32]};
    double b[4][6] = {};
    if (x_order == 0) sbp_asf_vjp(s, f, p + 33, b);
    else if (x_order > 1) face_flux_vjp<2>(s[0], s[1], s[2], s[3], f, p + 33, b);
    else face_flux_vjp<1>(s[0], s[1], s[2], s[3], f, p + 33, b);
    for (int c = 0; c < 4; ++c)
        for (int q = 0; q < 6; ++q) out[24 * (long long)t + 6 * c + q] = b[c][q];
}

}  // namespace

hipError_t launch_adj_seed(const AdjArgs& a, int ncell_chunks, hipStream_t s)
{
    if (ncell_chunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_seed_kernel, dim3(ncell_chunks), dim3(256), 0, s, a, a.blocks, a.cells);
    return hipGetLastError();
}

hipError_t launch_adj_chem_seed(const AdjArgs& a, int ncell_chunks, hipStream_t s)
{
    if (ncell_chunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_chem_seed_kernel, dim3(ncell_chunks), dim3(256), 0, s, a, a.blocks, a.cells);
    return hipGetLastError();
}

namespace {
template <int XORDER, int FLUX>
void launch_patches(const AdjArgs& a, int npatches, int epi, hipStream_t s)
{
    const dim3 grid(npatches), blk(kAdjPatchM * kAdjPatchT);
    if (epi == 3) hipLaunchKernelGGL((adj_patch_kernel<XORDER, FLUX, 3>), grid, blk, 0, s, a, a.blocks, a.patches);
    else if (epi == 2) hipLaunchKernelGGL((adj_patch_kernel<XORDER, FLUX, 2>), grid, blk, 0, s, a, a.blocks, a.patches);
    else if (epi == 1) hipLaunchKernelGGL((adj_patch_kernel<XORDER, FLUX, 1>), grid, blk, 0, s, a, a.blocks, a.patches);
    else hipLaunchKernelGGL((adj_patch_kernel<XORDER, FLUX, 0>), grid, blk, 0, s, a, a.blocks, a.patches);
}
}  // namespace

hipError_t launch_adj_faces(const AdjArgs& a, int npatches, int x_order, int flux_calc, int epi, hipStream_t s)
{
    if (npatches <= 0) return hipSuccess;
    if (flux_calc == 1) launch_patches<2, 1>(a, npatches, epi, s);
    else if (x_order > 1) launch_patches<2, 0>(a, npatches, epi, s);
    else launch_patches<1, 0>(a, npatches, epi, s);
    return hipGetLastError();
}

hipError_t launch_adj_gather(const AdjArgs& a, int ncell_chunks, int stage, hipStream_t s)
{
    if (ncell_chunks <= 0) return hipSuccess;
    if (stage == 3) hipLaunchKernelGGL(adj_gather_kernel<3>, dim3(ncell_chunks), dim3(256), 0, s, a, a.blocks, a.cells);
    else if (stage == 2) hipLaunchKernelGGL(adj_gather_kernel<2>, dim3(ncell_chunks), dim3(256), 0, s, a, a.blocks, a.cells);
    else hipLaunchKernelGGL(adj_gather_kernel<1>, dim3(ncell_chunks), dim3(256), 0, s, a, a.blocks, a.cells);
    return hipGetLastError();
}

hipError_t launch_flux_vjp_points(const double* in, double* out, int n, int x_order, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(flux_vjp_points_kernel, dim3((n + 63) / 64), dim3(64), 0, s, in, out, n,
                       x_order);
    return hipGetLastError();
}

// Device address of the current device's copy of `scratch`, looked up once per device with
// that device current.  In host code the array's name is only a stand-in for the symbol:
// kernels and copies take the address returned here.
static hipError_t scratch_address(double** out)
{
    static std::mutex mu;
    static std::vector<double*> addr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = static_cast<double*>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

hipError_t flux_vjp_points_noalloc(const double* in, double* out, int n, int x_order)
{
    if (n <= 0 || n > kNoAllocPoints) return hipErrorInvalidValue;
    double* base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e != hipSuccess) return e;
    double* din = base;
    double* dout = base + 39 * (size_t)kNoAllocPoints;
    e = hipMemcpy(din, in, sizeof(double) * 39 * (size_t)n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    e = launch_flux_vjp_points(din, dout, n, x_order, nullptr);
    if (e != hipSuccess) return e;
    // The null stream orders the copy after the kernel; every output read back was written by
    // this call (the kernel writes all 24 values of each of its n points).
    return hipMemcpy(out, dout, sizeof(double) * 24 * (size_t)n, hipMemcpyDeviceToHost);
}

}  // namespace chkn
