// chkn_adjoint.hip -- gfx950 kernels of the discrete adjoint of the chicken step (FP64).
//
// One reversed RK stage s with stage state W_s (read from the tape) and residual adjoint D
// (the adjoint of dU/dt of that stage, per cell):
//   1. face pass: every face of every active block forms Fbar = A (D_R/V_R - D_L/V_L) from
//      its two adjacent cells (ghost sides contribute nothing) and the vector-Jacobian
//      product of its flux (face_flux_vjp) with respect to its four stencil states, stored
//      per face in G (24 values: L1, L0, R0, R1 x rho, e, YB, vx, vy, vz);
//   2. gather: every cell sums, in a fixed order, the stencil adjoints of the faces whose
//      stencils contain it -- its own block's faces, the wall and outflow ghosts copied from
//      it, and the exchange ghosts other blocks read from it -- then maps the primitive
//      adjoint to the conserved one (decode_vjp) and applies the transpose of the TVD-RK3
//      stage combination (block.cu:1090-1160).  A cell next to an inflow boundary also adds
//      the adjoints that land on its line of inflow ghosts into that line's accumulator (the
//      gradient of the inflow state).
// No atomics: every sum has one writer and a fixed order, so the gradient is deterministic.
// Gradients need no bit-exact arithmetic, so this file is built with floating-point
// contraction and reciprocal arithmetic allowed (the stage kernels are not).
//
// Reversed step with A0 = adjoint of U^{n+1} on entry (seed: P = dt/6 A0, D = 4dt/6 A0):
//   stage 3 (W2): A0 += U2bar; D = P + dt/4 U2bar; P = D
//   stage 2 (W1): A0 += U1bar; D = P + dt U1bar
//   stage 1 (W0): A0 += U0bar        -> A0 = adjoint of U^n; seed of the next reversed step
#include <mutex>
#include <vector>

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

__device__ __forceinline__ double* gface(const BlockDev* B, int d)
{
    long long off = d == 0 ? 0 : (d == 1 ? B->nfi : B->nfi + B->nfj);
    return B->G + 24 * off;
}

// Adds stencil slot p (0 L1, 1 L0, 2 R0, 3 R1) of face fi of direction d to acc.  G is
// written once by the face pass and read once here: both sides use non-temporal accesses so
// that it does not evict the tape states and residual adjoints the passes re-read.
__device__ __forceinline__ void add_slot(const BlockDev* B, int d, int fi, int p, double acc[6])
{
    const double* g = gface(B, d);
    const long long nf = nfaces(B, d);
    const unsigned o = boff(fi);
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] += gld_nt(g + (6 * p + q) * nf, o);
}

constexpr int kFaceThreads = 128;

// One thread per face.  The four stencil states are staged in LDS (one column per thread,
// read back only by that thread) so that they need not stay in registers across the
// flux adjoint: reconstruction reads them, AUSMDV's adjoint runs on the reconstructed
// states alone, and the limiter adjoints read them again and write G directly.
template <int XORDER>
__global__ __launch_bounds__(kFaceThreads) void adj_face_kernel(AdjArgs A)
{
    __shared__ double st[24][kFaceThreads];
    const int t = (int)threadIdx.x;
    const FaceChunk ch = A.faces[blockIdx.x];
    const BlockDev* B = A.blocks + ch.blk;
    const int d = ch.dir;
    const long long nf = nfaces(B, d);
    const int fi = ch.f0 + t;
    if (fi >= ch.f1) return;
    const int nic = B->nic, njc = B->njc;
    int m, a, c;
    if (d == 0) {
        m = fi % (nic + 1);
        const int r = fi / (nic + 1);
        a = r % njc;
        c = r / njc;
    } else if (d == 1) {
        a = fi % nic;
        const int r = fi / nic;
        m = r % (njc + 1);
        c = r / (njc + 1);
    } else {
        a = fi % nic;
        const int r = fi / nic;
        c = r % njc;
        m = r / njc;
    }
    const int nline = line_len(B, d);
    const long long n = B->n;
    const double* D = B->adj + 12 * (long long)n;
    const double* g = geom(B, d);
    const double area = g[fi];
    const int cl = (m - 1 >= 0) ? cell_index(nic, njc, d, m - 1, a, c) : -1;
    const int cr = (m < nline) ? cell_index(nic, njc, d, m, a, c) : -1;
    const double vl = cl >= 0 ? B->volinv[cl] : 0.0, vr = cr >= 0 ? B->volinv[cr] : 0.0;
    double Fb[6];
    bool any = false;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double dl = cl >= 0 ? D[(long long)q * n + cl] : 0.0;
        const double dr = cr >= 0 ? D[(long long)q * n + cr] : 0.0;
        Fb[q] = area * (vr * dr - vl * dl);
        any = any || (Fb[q] != 0.0);
    }
    double* G = gface(B, d);
    if (!any) {
#pragma unroll
        for (int k = 0; k < 24; ++k) gst_nt(G + (long long)k * nf, boff(fi), 0.0);
        return;
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
        for (int p = 0; p < 4; ++p) gst_nt(G + (long long)(6 * p + q) * nf, boff(fi), gq[p]);
    }
}

// Adjoint reaching the ghost line of boundary side `side` at transverse (a, c): the sum over
// both ghost layers, in the same order the gather uses for walls and outflow.
__device__ __forceinline__ void ghost_line_adjoint(const BlockDev* B, int side, int a, int c,
                                                   double acc[6])
{
    const int d = side >> 1;
    const int nline = line_len(B, d);
    if ((side & 1) == 0) {
        const int f0 = face_index(B, d, 0, a, c), f1 = face_index(B, d, 1, a, c);
        add_slot(B, d, f0, 1, acc);  // depth 0 as L0 of face 0
        add_slot(B, d, f1, 0, acc);  // depth 0 as L1 of face 1
        add_slot(B, d, f0, 0, acc);  // depth 1 as L1 of face 0
    } else {
        const int fa = face_index(B, d, nline - 1, a, c), fb = face_index(B, d, nline, a, c);
        add_slot(B, d, fa, 3, acc);  // depth 0 as R1 of face n-1
        add_slot(B, d, fb, 2, acc);  // depth 0 as R0 of face n
        add_slot(B, d, fb, 3, acc);  // depth 1 as R1 of face n
    }
}

template <int STAGE>
__global__ __launch_bounds__(256) void adj_gather_kernel(AdjArgs A)
{
    const CellChunk ch = A.cells[blockIdx.x];
    const BlockDev* B = A.blocks + ch.blk;
    const long long n = B->n;
    const int cidx = ch.c0 + (int)threadIdx.x;
    if (cidx >= ch.c1) return;
    const int nic = B->nic, njc = B->njc;
    const int ijk[3] = {cidx % nic, (cidx / nic) % njc, cidx / (nic * njc)};
    double Wb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int m = ijk[d];
        const int a = d == 0 ? ijk[1] : ijk[0];
        const int c = d == 2 ? ijk[1] : ijk[2];
        const int nline = line_len(B, d);
        // This block's faces whose stencils hold the cell: m-1 (as R1), m (R0), m+1 (L0),
        // m+2 (L1).
        if (m - 1 >= 0) add_slot(B, d, face_index(B, d, m - 1, a, c), 3, Wb);
        add_slot(B, d, face_index(B, d, m, a, c), 2, Wb);
        add_slot(B, d, face_index(B, d, m + 1, a, c), 1, Wb);
        if (m + 2 <= nline) add_slot(B, d, face_index(B, d, m + 2, a, c), 0, Wb);
        // Wall and outflow ghosts of this block copy the cell next to the boundary.
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            if (m != (side ? nline - 1 : 0)) continue;
            const int code = B->bc[2 * d + side];
            if (code == 3 || code == 6) continue;
            double gb[6] = {0, 0, 0, 0, 0, 0};
            ghost_line_adjoint(B, 2 * d + side, a, c, gb);
            if (code == 4) {
                // Inflow ghosts hold the inflow state: this cell owns its line's accumulator.
                const int na = d == 0 ? njc : nic;
                const long long col = B->icol + B->icol_side[2 * d + side] + a + na * c;
#pragma unroll
                for (int q = 0; q < 6; ++q) A.inflow_acc[(long long)q * A.nicols + col] += gb[q];
            } else if (code == 5) {
#pragma unroll
                for (int q = 0; q < 6; ++q) Wb[q] += gb[q];
            } else {
                const double* g; int nf, fi;
                boundary_face(B, d, side, a, c, g, nf, fi);
                reflect_vjp(gb, ld_frame(g, nf, fi), Wb);
            }
        }
        // Exchange ghosts of other blocks (or this one, periodically) that read the cell.
        const int P = B->rp[2 * d];      // its plus-side ghosts at depth m read cells m = 0, 1
        if (P >= 0 && m <= 1) {
            const BlockDev* O = A.blocks + P;
            const int no = line_len(O, d);
            if (m == 0) {
                add_slot(O, d, face_index(O, d, no - 1, a, c), 3, Wb);
                add_slot(O, d, face_index(O, d, no, a, c), 2, Wb);
            } else {
                add_slot(O, d, face_index(O, d, no, a, c), 3, Wb);
            }
        }
        const int Q = B->rp[2 * d + 1];  // its minus-side ghosts at depth nline-1-m
        if (Q >= 0 && m >= nline - 2) {
            const BlockDev* O = A.blocks + Q;
            if (m == nline - 1) {
                add_slot(O, d, face_index(O, d, 0, a, c), 1, Wb);
                add_slot(O, d, face_index(O, d, 1, a, c), 0, Wb);
            } else {
                add_slot(O, d, face_index(O, d, 0, a, c), 0, Wb);
            }
        }
    }
    const Prim w = ld_prim(B->W + A.slot * B->wstride, n, cidx);
    double Ub[6];
    decode_vjp(w, Wb, Ub);
    const double dt = A.dt[B->inst];
    const unsigned o = boff(cidx);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        double* A0 = B->adj + (long long)q * n;
        double* Pp = B->adj + (long long)(6 + q) * n;
        double* D = B->adj + (long long)(12 + q) * n;
        const double a0 = gld(A0, o) + Ub[q];
        gst(A0, o, a0);
        if (STAGE == 3) {
            const double v = gld(Pp, o) + 0.25 * dt * Ub[q];
            gst(D, o, v);
            gst(Pp, o, v);
        } else if (STAGE == 2) {
            gst(D, o, gld(Pp, o) + dt * Ub[q]);
        } else if (A.dt_next) {
            const double dtn = A.dt_next[B->inst];
            gst(Pp, o, kOneSixth * dtn * a0);
            gst(D, o, kFourSixth * dtn * a0);
        }
    }
}

__global__ __launch_bounds__(256) void adj_seed_kernel(AdjArgs A)
{
    const CellChunk ch = A.cells[blockIdx.x];
    const BlockDev* B = A.blocks + ch.blk;
    const long long n = B->n;
    const int cidx = ch.c0 + (int)threadIdx.x;
    if (cidx >= ch.c1) return;
    const double dt = A.dt[B->inst];
    const double* A0 = B->adj;
    double* Pp = B->adj + 6 * (long long)n;
    double* D = B->adj + 12 * (long long)n;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const long long o = (long long)q * n + cidx;
        Pp[o] = kOneSixth * dt * A0[o];
        D[o] = kFourSixth * dt * A0[o];
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
    if (x_order > 1) face_flux_vjp<2>(s[0], s[1], s[2], s[3], f, p + 33, b);
    else face_flux_vjp<1>(s[0], s[1], s[2], s[3], f, p + 33, b);
    for (int c = 0; c < 4; ++c)
        for (int q = 0; q < 6; ++q) out[24 * (long long)t + 6 * c + q] = b[c][q];
}

}  // namespace

hipError_t launch_adj_seed(const AdjArgs& a, int ncell_chunks, hipStream_t s)
{
    if (ncell_chunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_seed_kernel, dim3(ncell_chunks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adj_faces(const AdjArgs& a, int nface_chunks, int x_order, hipStream_t s)
{
    if (nface_chunks <= 0) return hipSuccess;
    if (x_order > 1) hipLaunchKernelGGL(adj_face_kernel<2>, dim3(nface_chunks), dim3(kFaceThreads), 0, s, a);
    else hipLaunchKernelGGL(adj_face_kernel<1>, dim3(nface_chunks), dim3(kFaceThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adj_gather(const AdjArgs& a, int ncell_chunks, int stage, hipStream_t s)
{
    if (ncell_chunks <= 0) return hipSuccess;
    if (stage == 3) hipLaunchKernelGGL(adj_gather_kernel<3>, dim3(ncell_chunks), dim3(256), 0, s, a);
    else if (stage == 2) hipLaunchKernelGGL(adj_gather_kernel<2>, dim3(ncell_chunks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(adj_gather_kernel<1>, dim3(ncell_chunks), dim3(256), 0, s, a);
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
