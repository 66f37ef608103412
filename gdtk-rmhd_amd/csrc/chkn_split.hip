// chkn_split.hip -- gfx950 kernels of the face-list form of a stage (FP64).
//
// The fused stage kernel forms each cell's flux divergence from face fluxes it evaluates on
// the fly.  Two cases need the face fluxes as a stored list instead:
//  * viscous jobs: the viscous flux of a face (block.cu:88-162) takes least-squares
//    gradients over a cloud made of the cells beside the face and the transverse faces of
//    those cells, so it needs the face states of its neighbours;
//  * first-order jobs whose cells or inflow ghost cells hold gas states (p, T, a) that the
//    equation of state does not reproduce from (rho, e): the reference's first-order flux
//    reads p and a of the two cells beside the face verbatim (face.cu:356-371), as read from
//    the flow file (cell.cu:112-131) or configured for the inflow (config.cu:398-406);
//  * the SBP-ASF flux (face.cu:268-342), which reads p of all four stencil cells verbatim.
// Three launches per stage, every active block of every instance in each:
//   faces:   one thread per face: convective flux and, for viscous jobs, the face state
//            set_as_average(L, R) (flow.cu:220-225) with the viscous boundary conditions
//            (bcs.cu:777-797);
//   viscous: one thread per face: gradients of T and velocity over the face's cloud, then
//            the viscous stress and heat flux subtracted from the face flux;
//   update:  one thread per cell: flux divergence in the reference's order (cell.cu:206-213),
//            TVD-RK3 stage, decode, next stage state, bad-cell count, signal-time minimum.
//
// Build with -ffp-contract=off: every product and sum rounds exactly as in the reference.
#include "chkn_internal.h"
#include "chkn_physics.h"
#include "chkn_stencil.h"

namespace chkn {

namespace {

constexpr int kChunk = 256;

// Two reals: the half of a 4-value record one load moves.
#ifdef CHKN_FP32
using real2 = float2;
__device__ __forceinline__ real2 make_real2(real a, real b) { return make_float2(a, b); }
#else
using real2 = double2;
__device__ __forceinline__ real2 make_real2(real a, real b) { return make_double2(a, b); }
#endif

// gas.cu:36-48: Cp and Sutherland's law for air.
constexpr real kCp = kCv + kR;
constexpr real kMuRef = real(1.716e-5);
constexpr real kTRef = real(273.0);
constexpr real kSuth = real(111.0);
constexpr real kPrandtl = real(0.72);

struct Gas {
    real p, T, a;
};

__device__ __forceinline__ Gas eos_gas(const Prim& w)  // gas.cu:77-83
{
    Gas g;
    g.T = eos_T(w.e, w.YB);
    g.p = eos_p(w.rho, g.T);
    g.a = eos_a(g.T);
    return g;
}

// Gas state of cell c of block B in state w: the flow file's p, T, a where the reference
// still holds them (cells of inactive blocks; active cells before their first update),
// otherwise the equation of state of (rho, e).  A block whose file values agree with the
// equation of state carries no gas array.
__device__ __forceinline__ Gas cell_gas(const SplitArgs& A, const BlockDev* B, const Prim& w,
                                        int c)
{
    if (B->gas && (A.file_gas || !B->active)) {
        const long long n = B->n;
        const unsigned o = boff(c);
        return Gas{gld(B->gas, o), gld(B->gas + n, o), gld(B->gas + 2 * n, o)};
    }
    return eos_gas(w);
}

struct Point {
    Prim w;
    Gas g;
};

// Stencil position m of the line (dir, a, c): an interior cell, or the ghost cell of the
// side's boundary condition (bcs.cu:16-503) with the gas state the reference copies into it.
template <bool GAS>
__device__ __forceinline__ Point point(const SplitArgs& A, const BlockDev* B, int dir, int m,
                                       int a, int c)
{
    const int nl = dir == 0 ? B->nic : (dir == 1 ? B->njc : B->nkc);
    Point r;
    if (m >= 0 && m < nl) {
        const int idx = cell_index(B->nic, B->njc, dir, m, a, c);
        r.w = ld_prim(B->W + A.win * B->wstride, B->n, idx);
        if (GAS) r.g = cell_gas(A, B, r.w, idx);
        return r;
    }
    const int side = m < 0 ? 0 : 1;
    const int depth = side ? m - nl : -m - 1;
    const int bf = 2 * dir + side;
    const int code = B->bc[bf];
    if (code == 3) {  // exchange: the neighbour's cell at the same depth
        const BlockDev* O = A.blocks + B->bc_arg[bf];
        const int no = dir == 0 ? O->nic : (dir == 1 ? O->njc : O->nkc);
        const int idx = cell_index(O->nic, O->njc, dir, side ? depth : no - 1 - depth, a, c);
        r.w = ld_prim(O->W + A.win * O->wstride, O->n, idx);
        if (GAS) r.g = cell_gas(A, O, r.w, idx);
        return r;
    }
    if (code == 4) {  // inflow: the configured flow state
        const real* s = A.inflow + 6 * B->bc_arg[bf];
        r.w = Prim{s[0], s[1], s[2], s[3], s[4], s[5]};
        if (GAS) {
            const real* g = A.inflow_gas + 3 * B->bc_arg[bf];
            r.g = Gas{g[0], g[1], g[2]};
        }
        return r;
    }
    if (code == 6) {  // inflow function: the state at the ghost cell's position
        const int na = dir == 0 ? B->njc : B->nic;
        const int nt = na * (dir == 2 ? B->njc : B->nkc);
        const CHKN_GMEM real* s = gmem(B->side[bf]) + (long long)depth * 9 * nt + (a + na * c);
        r.w = Prim{s[0], s[nt], s[2 * nt], s[3 * nt], s[4 * nt], s[5 * nt]};
        if (GAS) r.g = Gas{s[6 * nt], s[7 * nt], s[8 * nt]};
        return r;
    }
    const int idx = cell_index(B->nic, B->njc, dir, side ? nl - 1 : 0, a, c);
    r.w = ld_prim(B->W + A.win * B->wstride, B->n, idx);
    if (GAS) r.g = cell_gas(A, B, r.w, idx);
    if (code == 5) return r;  // outflow copies the interior cell
    const real* g;          // walls reflect it; p, T, a are copied unchanged
    int nf, fi;
    boundary_face(B, dir, side, a, c, g, nf, fi);
    r.w = reflect(r.w, ld_frame(g, nf, fi));
    return r;
}

// Face f of direction dir: line position m (between cells m-1 and m), transverse indices
// (a, c) of its line, and its index in the block's face list.
struct FacePos {
    int m, a, c, ff;
};

__device__ __forceinline__ FacePos face_pos(const BlockDev* B, int dir, int f)
{
    const int nic = B->nic, njc = B->njc;
    FacePos p;
    if (dir == 0) {  // i + (nic+1)*(j + njc*k)
        p.m = f % (nic + 1);
        const int r = f / (nic + 1);
        p.a = r % njc;
        p.c = r / njc;
        p.ff = f;
    } else if (dir == 1) {  // i + nic*(j + (njc+1)*k)
        p.a = f % nic;
        const int r = f / nic;
        p.m = r % (njc + 1);
        p.c = r / (njc + 1);
        p.ff = (int)B->nfi + f;
    } else {  // i + nic*(j + njc*k)
        p.a = f % nic;
        const int r = f / nic;
        p.c = r % njc;
        p.m = r / njc;
        p.ff = (int)(B->nfi + B->nfj) + f;
    }
    return p;
}

__device__ __forceinline__ int line_length(const BlockDev* B, int dir)
{
    return dir == 0 ? B->nic : (dir == 1 ? B->njc : B->nkc);
}

__device__ __forceinline__ bool is_wall(int code) { return code >= 0 && code <= 2; }

// FLUX 0: AUSMDV on the MUSCL (XORDER 2) or first-order states, face state the average of
// the two reconstructed states; FLUX 1: SBP-ASF, face state the average of the two cells
// beside the face (face.cu:373-376).
template <int XORDER, bool VISC, int FLUX>
__global__ __launch_bounds__(kChunk) void split_face_kernel(SplitArgs A0, const BlockDev* __restrict__ blocks,
                                                            const FaceChunk* __restrict__ faces)
{
    // Read-only tables as restrict-qualified arguments: scalar loads no store can clobber.
    SplitArgs A = A0;
    A.blocks = blocks;
    A.faces = faces;
    const FaceChunk ch = A.faces[blockIdx.x];
    const BlockDev* B = A.blocks + ch.blk;
    const int f = ch.f0 + (int)threadIdx.x;
    if (f >= ch.f1 || !A.ctl.run[B->inst]) return;
    const int dir = ch.dir;
    const FacePos p = face_pos(B, dir, f);
    const real* g = dir == 0 ? B->gi : (dir == 1 ? B->gj : B->gk);
    const long long nf = dir == 0 ? B->nfi : (dir == 1 ? B->nfj : B->nfk);
    const Frame fr = ld_frame(g, nf, f);
    real rhoL, rhoR, eL, eR, YBL, YBR, vxL, vxR, vyL, vyR, vzL, vzR, pL, pR, aL, aR;
    real F[6];
    if (FLUX == 1) {
        Prim s[4];
        real ps[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const Point pt = point<true>(A, B, dir, p.m - 2 + q, p.a, p.c);
            s[q] = pt.w;
            ps[q] = pt.g.p;
        }
        sbp_asf(s, ps, fr, F);
        rhoL = s[1].rho; eL = s[1].e; YBL = s[1].YB; vxL = s[1].vx; vyL = s[1].vy; vzL = s[1].vz;
        rhoR = s[2].rho; eR = s[2].e; YBR = s[2].YB; vxR = s[2].vx; vyR = s[2].vy; vzR = s[2].vz;
    } else if (XORDER > 1) {
        // face.cu:359-368: rho, e, YB, then the equation of state, then the velocity.
        const Prim s0 = point<false>(A, B, dir, p.m - 2, p.a, p.c).w;
        const Prim s1 = point<false>(A, B, dir, p.m - 1, p.a, p.c).w;
        const Prim s2 = point<false>(A, B, dir, p.m, p.a, p.c).w;
        const Prim s3 = point<false>(A, B, dir, p.m + 1, p.a, p.c).w;
        muscl(s0.rho, s1.rho, s2.rho, s3.rho, rhoL, rhoR);
        muscl(s0.e, s1.e, s2.e, s3.e, eL, eR);
        muscl(s0.YB, s1.YB, s2.YB, s3.YB, YBL, YBR);
        muscl(s0.vx, s1.vx, s2.vx, s3.vx, vxL, vxR);
        muscl(s0.vy, s1.vy, s2.vy, s3.vy, vyL, vyR);
        muscl(s0.vz, s1.vz, s2.vz, s3.vz, vzL, vzR);
        const real TL = eos_T(eL, YBL), TR = eos_T(eR, YBR);
        pL = eos_p(rhoL, TL);
        pR = eos_p(rhoR, TR);
        aL = eos_a(TL);
        aR = eos_a(TR);
    } else {
        const Point L = point<true>(A, B, dir, p.m - 1, p.a, p.c);
        const Point R = point<true>(A, B, dir, p.m, p.a, p.c);
        rhoL = L.w.rho; eL = L.w.e; YBL = L.w.YB; vxL = L.w.vx; vyL = L.w.vy; vzL = L.w.vz;
        rhoR = R.w.rho; eR = R.w.e; YBR = R.w.YB; vxR = R.w.vx; vyR = R.w.vy; vzR = R.w.vz;
        pL = L.g.p; aL = L.g.a;
        pR = R.g.p; aR = R.g.a;
    }
    if (FLUX == 0)
        ausmdv(rhoL, eL, pL, aL, YBL, vxL, vyL, vzL, rhoR, eR, pR, aR, YBR, vxR, vyR, vzR, fr, F);
    const long long nft = B->nfi + B->nfj + B->nfk;
#pragma unroll
    for (int q = 0; q < 6; ++q) gmem(B->F)[q * nft + p.ff] = F[q];
    if (VISC) {
        // Face state: average of the two reconstructed states, T from the mean e and YB.
        const real half = real(0.5);
        real T = eos_T(half * (eL + eR), half * (YBL + YBR));
        real vx = half * (vxL + vxR), vy = half * (vyL + vyR), vz = half * (vzL + vzR);
        const int nl = line_length(B, dir);
        const int side = (p.m == 0) ? 0 : ((p.m == nl) ? 1 : -1);
        if (side >= 0) {
            const int bf = 2 * dir + side;
            const int code = B->bc[bf];
            const CHKN_GMEM real* sd = gmem(B->side[bf]);
            const int t = p.a + (dir == 0 ? B->njc : B->nic) * p.c;
            if (code == 1 || code == 2) {  // no-slip walls; fixed-T walls also set T
                vx = real(0);
                vy = real(0);
                vz = real(0);
                if (code == 2) T = sd ? sd[t] : B->twall[bf];
            } else if (code == 6) {  // inflow function: its state at the face centre
                const int nt = (dir == 0 ? B->njc : B->nic) * (dir == 2 ? B->njc : B->nkc);
                T = sd[18 * nt + t];
                vx = sd[19 * nt + t];
                vy = sd[20 * nt + t];
                vz = sd[21 * nt + t];
            }
        }
        const unsigned o = boff(4 * p.ff);
        gst(B->FS, o, T);
        gst(B->FS + 1, o, vx);
        gst(B->FS + 2, o, vy);
        gst(B->FS + 3, o, vz);
    }
}

// Index in the face list of the face on side hi (0 or 1) in direction d of cell (i, j, k).
__device__ __forceinline__ int cell_face(const BlockDev* B, int d, int hi, int i, int j, int k)
{
    const int nic = B->nic, njc = B->njc;
    if (d == 0) return (i + hi) + (nic + 1) * (j + njc * k);
    if (d == 1) return (int)B->nfi + i + nic * ((j + hi) + (njc + 1) * k);
    return (int)(B->nfi + B->nfj) + i + nic * (j + njc * (k + hi));
}

// 4-value records: (x, y, z, unused) centres and (T, vx, vy, vz) face states.
__device__ __forceinline__ void ld_rec4(const real* base, long long r, real& a, real& b,
                                        real& c, real& d)
{
    const unsigned o = boff((int)(4 * r));
    a = gld(base, o); b = gld(base + 1, o); c = gld(base + 2, o); d = gld(base + 3, o);
}
__device__ __forceinline__ void ld_rec(const real* base, long long r, real& x, real& y,
                                       real& z)
{
    const unsigned o = boff((int)(4 * r));
    x = gld(base, o); y = gld(base + 1, o); z = gld(base + 2, o);
}

// The cloud of face p (block.cu:317-358, :418-459, :519-560), in the order the reference
// lists it: the cells beside the face, minus side first -- on a boundary the interior cell
// and, unless the side is a wall, the first ghost cell -- then for each of those interior
// cells its faces in the two other directions, lower direction first, low side first.
// visit(kind, index): kind 0 = cell at line position index (a ghost when outside the
// block), kind 1 = face index in the face list.
template <class Visit>
__device__ __forceinline__ void visit_cloud(const BlockDev* B, int dir, const FacePos& p,
                                            Visit&& visit)
{
    const int nl = line_length(B, dir);
    if (p.m == 0) {
        if (!is_wall(B->bc[2 * dir])) visit(0, -1);
        visit(0, 0);
    } else if (p.m == nl) {
        visit(0, nl - 1);
        if (!is_wall(B->bc[2 * dir + 1])) visit(0, nl);
    } else {
        visit(0, p.m - 1);
        visit(0, p.m);
    }
    const int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
    for (int mc = p.m - 1; mc <= p.m; ++mc) {
        if (mc < 0 || mc >= nl) continue;
        int i, j, k;
        if (dir == 0) { i = mc; j = p.a; k = p.c; }
        else if (dir == 1) { i = p.a; j = mc; k = p.c; }
        else { i = p.a; j = p.c; k = mc; }
        visit(1, cell_face(B, d1, 0, i, j, k));
        visit(1, cell_face(B, d1, 1, i, j, k));
        visit(1, cell_face(B, d2, 0, i, j, k));
        visit(1, cell_face(B, d2, 1, i, j, k));
    }
}

// Offset of a cloud point from the face centre (fx, fy, fz) and its inverse squared
// distance (block.cu:40-47).  Ghost cells sit at f + (f - c) for the interior cell c
// beside the boundary face (block.cu:656-780).
__device__ __forceinline__ void cloud_offset(const BlockDev* B, int dir, const FacePos& p,
                                             real fx, real fy, real fz,
                                             int kind, int idx, real& dx, real& dy,
                                             real& dz, real& w2)
{
    real x, y, z;
    if (kind == 1) {
        ld_rec(B->fpos, idx, x, y, z);
    } else {
        const int nl = line_length(B, dir);
        const bool ghost = idx < 0 || idx >= nl;
        const int c = cell_index(B->nic, B->njc, dir, ghost ? (idx < 0 ? 0 : nl - 1) : idx, p.a, p.c);
        ld_rec(B->cpos, c, x, y, z);
        if (ghost) {
            const real ex = fx - x, ey = fy - y, ez = fz - z;
            x = fx + ex;
            y = fy + ey;
            z = fz + ez;
        }
    }
    dx = x - fx;
    dy = y - fy;
    dz = z - fz;
    w2 = (real)(1.0 / (double)(dx * dx + dy * dy + dz * dz));  // block.cu:54 divides in double
}

// Viscous flux of every face (block.cu:88-162): least-squares gradients of T and velocity
// over the face's cloud, with the weights of block.cu:28-86 (normal matrix inverted by
// cofactors, rsla.cu:53-72) formed on the fly from the cell and face centres -- two passes
// over the cloud, the second recomputing each offset -- rather than read from memory.
// A work-group takes a kPatchI x kPatchJ patch of faces of one direction at one k-level; the patch
// table runs block, k, patch, direction, and work-group b (on XCD b % 8) takes the b/8-th
// entry of XCD b%8's contiguous range, so the clouds of neighbouring patches, which share
// most of their faces and cells, are read through one L2.
__global__ __launch_bounds__(kChunk) void split_viscous_kernel(SplitArgs A0, const BlockDev* __restrict__ blocks,
                                                               const FacePatch* __restrict__ patches, int npatches)
{
    SplitArgs A = A0;
    A.blocks = blocks;
    A.patches = patches;
    const int per_xcd = (npatches + 7) >> 3;
    const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (item >= npatches) return;
    const FacePatch ch = A.patches[item];
    const BlockDev* B = A.blocks + ch.blk;
    const int dir = ch.dir;
    const int i = ch.i0 + (int)threadIdx.x % kPatchI, j = ch.j0 + (int)threadIdx.x / kPatchI;
    const int nic = B->nic, njc = B->njc;
    if (i > nic - (dir != 0) || j > njc - (dir != 1) || !A.ctl.run[B->inst]) return;
    const int f = dir == 0 ? i + (nic + 1) * (j + njc * ch.k)
                           : (dir == 1 ? i + nic * (j + (njc + 1) * ch.k) : i + nic * (j + njc * ch.k));
    const FacePos p = face_pos(B, dir, f);
    const long long nft = B->nfi + B->nfj + B->nfk;
    const real* __restrict__ FS = B->FS;
    real fx, fy, fz;
    ld_rec(B->fpos, p.ff, fx, fy, fz);
    // Normal matrix of the cloud.
    real xx = 0.0, xy = 0.0, xz = 0.0, yy = 0.0, yz = 0.0, zz = 0.0;
    visit_cloud(B, dir, p, [&](int kind, int idx) {
        real dx, dy, dz, w2;
        cloud_offset(B, dir, p, fx, fy, fz, kind, idx, dx, dy, dz, w2);
        xx += w2 * dx * dx;
        xy += w2 * dx * dy;
        xz += w2 * dx * dz;
        yy += w2 * dy * dy;
        yz += w2 * dy * dz;
        zz += w2 * dz * dz;
    });
    // Inverse by cofactors of the symmetric matrix [[xx,xy,xz],[xy,yy,yz],[xz,yz,zz]]; a
    // singular matrix (|det| <= 1e-16) gives zero weights, as in the reference.
    real m00 = 0.0, m01 = 0.0, m02 = 0.0, m10 = 0.0, m11 = 0.0, m12 = 0.0, m20 = 0.0,
           m21 = 0.0, m22 = 0.0;
    const real det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    if (!(fabs(det) <= real(1.0e-16))) {  // rsla.cu:54-59
        const real r = real(1) / det;
        m00 = (yy * zz - yz * yz) * r;
        m01 = (xz * yz - xy * zz) * r;
        m02 = (xy * yz - xz * yy) * r;
        m10 = (yz * xz - xy * zz) * r;
        m11 = (xx * zz - xz * xz) * r;
        m12 = (xz * xy - xx * yz) * r;
        m20 = (xy * yz - yy * xz) * r;
        m21 = (xy * xz - xx * yz) * r;
        m22 = (xx * yy - xy * xy) * r;
    }
    real T0, vx0, vy0, vz0;
    ld_rec4(FS, p.ff, T0, vx0, vy0, vz0);
    real gTx = 0.0, gTy = 0.0, gTz = 0.0, gxx = 0.0, gxy = 0.0, gxz = 0.0;
    real gyx = 0.0, gyy = 0.0, gyz = 0.0, gzx = 0.0, gzy = 0.0, gzz = 0.0;
    visit_cloud(B, dir, p, [&](int kind, int idx) {
        real dx, dy, dz, w2;
        cloud_offset(B, dir, p, fx, fy, fz, kind, idx, dx, dy, dz, w2);
        const real wx = w2 * (m00 * dx + m01 * dy + m02 * dz);
        const real wy = w2 * (m10 * dx + m11 * dy + m12 * dz);
        const real wz = w2 * (m20 * dx + m21 * dy + m22 * dz);
        real T, vx, vy, vz;
        if (kind == 1) {
            ld_rec4(FS, idx, T, vx, vy, vz);
        } else {
            const Point pt = point<true>(A, B, dir, idx, p.a, p.c);
            T = pt.g.T;
            vx = pt.w.vx;
            vy = pt.w.vy;
            vz = pt.w.vz;
        }
        real dq = T - T0;
        gTx += wx * dq; gTy += wy * dq; gTz += wz * dq;
        dq = vx - vx0;
        gxx += wx * dq; gxy += wy * dq; gxz += wz * dq;
        dq = vy - vy0;
        gyx += wx * dq; gyy += wy * dq; gyz += wz * dq;
        dq = vz - vz0;
        gzx += wx * dq; gzy += wy * dq; gzz += wz * dq;
    });
    // Transport coefficients at the face temperature (gas.cu:95-100), stresses and heat
    // flux (block.cu:134-161).
    const real mu = kMuRef * sqrt(T0 / kTRef) * (T0 / kTRef) * (kTRef + kSuth) / (T0 + kSuth);
    const real kcond = kCp * mu / kPrandtl;
    const real two = real(2);
    const real lmbda = -two / real(3) * mu;
    const real tau_xx = two * mu * gxx + lmbda * (gxx + gyy + gzz);
    const real tau_yy = two * mu * gyy + lmbda * (gxx + gyy + gzz);
    const real tau_zz = two * mu * gzz + lmbda * (gxx + gyy + gzz);
    const real tau_xy = mu * (gxy + gyx);
    const real tau_xz = mu * (gxz + gzx);
    const real tau_yz = mu * (gyz + gzy);
    const real qx = kcond * gTx, qy = kcond * gTy, qz = kcond * gTz;
    const real* gf = dir == 0 ? B->gi : (dir == 1 ? B->gj : B->gk);
    const long long nf = dir == 0 ? B->nfi : (dir == 1 ? B->nfj : B->nfk);
    const unsigned of = boff(f);
    const real nx = gld(gf + nf, of), ny = gld(gf + 2 * nf, of), nz = gld(gf + 3 * nf, of);
    CHKN_GMEM real* __restrict__ F = gmem(B->F);
    F[nft + p.ff] -= tau_xx * nx + tau_xy * ny + tau_xz * nz;
    F[2 * nft + p.ff] -= tau_xy * nx + tau_yy * ny + tau_yz * nz;
    F[3 * nft + p.ff] -= tau_xz * nx + tau_yz * ny + tau_zz * nz;
    F[4 * nft + p.ff] -= (tau_xx * vx0 + tau_xy * vy0 + tau_xz * vz0 + qx) * nx +
                         (tau_xy * vx0 + tau_yy * vy0 + tau_yz * vz0 + qy) * ny +
                         (tau_xz * vx0 + tau_yz * vy0 + tau_zz * vz0 + qz) * nz;
}

template <int STAGE, bool DTRED>
__global__ __launch_bounds__(kChunk) void split_update_kernel(SplitArgs A0, const BlockDev* __restrict__ blocks,
                                                              const CellChunk* __restrict__ cells)
{
    SplitArgs A = A0;
    A.blocks = blocks;
    A.cells = cells;
    const CellChunk ch = A.cells[blockIdx.x];
    const BlockDev* B = A.blocks + ch.blk;
    if (!A.ctl.run[B->inst]) return;  // uniform over the work-group
    const int cidx = ch.c0 + (int)threadIdx.x;
    int bad = 0;
    unsigned long long sigmin = ~0ull;
    if (cidx < ch.c1) {
        const int nic = B->nic, njc = B->njc;
        const long long n = B->n;
        const int i = cidx % nic, j = (cidx / nic) % njc, k = cidx / (nic * njc);
        const long long nfi = B->nfi, nfj = B->nfj, nft = nfi + B->nfj + B->nfk;
        const int fim = i + (nic + 1) * (j + njc * k);
        const int fjm = i + nic * (j + (njc + 1) * k);
        const int fkm = cidx;
        const CHKN_GMEM real* gi = gmem(B->gi);
        const CHKN_GMEM real* gj = gmem(B->gj);
        const CHKN_GMEM real* gk = gmem(B->gk);
        const real aim = gi[fim], aip = gi[fim + 1];
        const real ajm = gj[fjm], ajp = gj[fjm + nic];
        const real akm = gk[fkm], akp = gk[fkm + nic * njc];
        const real vinv = B->volinv[cidx];
        real d[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const CHKN_GMEM real* Fq = gmem(B->F) + q * nft;
            real s = aim * Fq[fim] - aip * Fq[fim + 1];
            s = s + ajm * Fq[nfi + fjm];
            s = s - ajp * Fq[nfi + fjm + nic];
            s = s + akm * Fq[nfi + nfj + fkm];
            s = s - akp * Fq[nfi + nfj + fkm + nic * njc];
            d[q] = vinv * s;
        }
        Prim w;
        const int flag = stage_update<STAGE>(B, n, cidx, A.ctl.dt[B->inst], d, A.wout, w);
        bad = flag;
        if (DTRED && !flag) sigmin = signal_bits(B, n, i, j, k, cidx, w);
    }
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

// The A->B reaction of every cell after a step (simulate.cu:296-303, block.cu:1174-1184,
// gas.cu:102-111): YA = 1 - YB; where T > Ti, YB = 1 - YA exp(-alpha dt), with the factor
// computed on the host for the instance's dt; then U = encode(state) (level 0, every cell of
// a reacting instance).
template <bool DTRED>
__global__ __launch_bounds__(kChunk) void chemistry_kernel(const BlockDev* __restrict__ blocks,
                                                           const CellChunk* __restrict__ cells, StepCtl ctl,
                                                           int slot)
{
    const CellChunk ch = cells[blockIdx.x];
    const BlockDev* B = blocks + ch.blk;
    if (!ctl.run[B->inst]) return;  // uniform over the work-group
    const real factor = ctl.chem[B->inst];
    const int cidx = ch.c0 + (int)threadIdx.x;
    unsigned long long sigmin = ~0ull;
    if (cidx < ch.c1) {
        const long long n = B->n;
        real* W = B->W + slot * B->wstride;
        Prim w = ld_prim(W, n, cidx);
        if (factor >= real(0)) {
            const real T = eos_T(w.e, w.YB);
            const real YA = real(1) - w.YB;
            if (T > kTi) {
                w.YB = real(1) - YA * factor;
                gst(W + 2 * n, boff(cidx), w.YB);
            }
            real U[6];
            encode(w, U);
#pragma unroll
            for (int q = 0; q < 6; ++q) gst(B->U + q * n, boff(cidx), U[q]);
        }
        if (DTRED) {
            const int nic = B->nic, njc = B->njc;
            sigmin = signal_bits(B, n, cidx % nic, (cidx / nic) % njc, cidx / (nic * njc), cidx, w);
        }
    }
    if (DTRED) {
        const unsigned long long wm = wave_min_u64(sigmin);
        if ((threadIdx.x & 63) == 0 && wm != ~0ull) atomicMin(&ctl.sigmin_bits[B->inst], wm);
    }
}

}  // namespace

hipError_t launch_chemistry(const BlockDev* blocks, const CellChunk* cells, int ncell_chunks,
                            const StepCtl& ctl, int slot, bool dt_reduce, hipStream_t s)
{
    if (ncell_chunks <= 0) return hipSuccess;
    const dim3 grid(ncell_chunks), blk(kChunk);
    if (dt_reduce) hipLaunchKernelGGL((chemistry_kernel<true>), grid, blk, 0, s, blocks, cells, ctl, slot);
    else hipLaunchKernelGGL((chemistry_kernel<false>), grid, blk, 0, s, blocks, cells, ctl, slot);
    return hipGetLastError();
}

hipError_t launch_split_faces(const SplitArgs& a, int nface_chunks, int x_order, int flux_calc,
                              bool viscous, hipStream_t s)
{
    if (nface_chunks <= 0) return hipSuccess;
    const dim3 grid(nface_chunks), blk(kChunk);
    if (flux_calc == 1) {
        if (viscous) hipLaunchKernelGGL((split_face_kernel<2, true, 1>), grid, blk, 0, s, a, a.blocks, a.faces);
        else hipLaunchKernelGGL((split_face_kernel<2, false, 1>), grid, blk, 0, s, a, a.blocks, a.faces);
    } else if (x_order > 1) {
        if (viscous) hipLaunchKernelGGL((split_face_kernel<2, true, 0>), grid, blk, 0, s, a, a.blocks, a.faces);
        else hipLaunchKernelGGL((split_face_kernel<2, false, 0>), grid, blk, 0, s, a, a.blocks, a.faces);
    } else {
        if (viscous) hipLaunchKernelGGL((split_face_kernel<1, true, 0>), grid, blk, 0, s, a, a.blocks, a.faces);
        else hipLaunchKernelGGL((split_face_kernel<1, false, 0>), grid, blk, 0, s, a, a.blocks, a.faces);
    }
    return hipGetLastError();
}

hipError_t launch_split_viscous(const SplitArgs& a, int npatches, hipStream_t s)
{
    if (npatches <= 0) return hipSuccess;
    const int grid = 8 * ((npatches + 7) / 8);
    hipLaunchKernelGGL(split_viscous_kernel, dim3(grid), dim3(kChunk), 0, s, a, a.blocks, a.patches, npatches);
    return hipGetLastError();
}

hipError_t launch_split_update(const SplitArgs& a, int ncell_chunks, int stage, bool dt_reduce,
                               hipStream_t s)
{
    if (ncell_chunks <= 0) return hipSuccess;
    const dim3 grid(ncell_chunks), blk(kChunk);
    if (stage == 1) hipLaunchKernelGGL((split_update_kernel<1, false>), grid, blk, 0, s, a, a.blocks, a.cells);
    else if (stage == 2) hipLaunchKernelGGL((split_update_kernel<2, false>), grid, blk, 0, s, a, a.blocks, a.cells);
    else if (stage == 3 && dt_reduce) hipLaunchKernelGGL((split_update_kernel<3, true>), grid, blk, 0, s, a, a.blocks, a.cells);
    else if (stage == 3) hipLaunchKernelGGL((split_update_kernel<3, false>), grid, blk, 0, s, a, a.blocks, a.cells);
    else if (dt_reduce) hipLaunchKernelGGL((split_update_kernel<4, true>), grid, blk, 0, s, a, a.blocks, a.cells);
    else hipLaunchKernelGGL((split_update_kernel<4, false>), grid, blk, 0, s, a, a.blocks, a.cells);
    return hipGetLastError();
}

}  // namespace chkn
