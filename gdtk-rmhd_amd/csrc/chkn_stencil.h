// chkn_stencil.h -- device-side access to block data shared by the stage kernels and the
// adjoint kernels: SoA loads of primitive states and face frames, line indexing, and the
// ghost-cell states of the boundary conditions.
#pragma once

#include "chkn_geometry.h"
#include "chkn_internal.h"
#include "chkn_physics.h"

namespace chkn {

// Global-memory view of a device pointer held in a block table or argument struct.  The
// compiler cannot tell such pointers from LDS ones and would access them with flat_*
// instructions, which also count against the LDS wait counter: every wait for an LDS read
// would then also wait for the outstanding HBM loads.  Through this view they are global_*.
#define CHKN_GMEM __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ CHKN_GMEM T* gmem(T* p)
{
    return (CHKN_GMEM T*)p;
}

// Element at byte offset `off` of a global array.  With a uniform base and a 32-bit offset
// the compiler issues one global load with the base in scalar registers and the offset in a
// single VGPR shared by all arrays indexed alike, instead of a 64-bit address per load.
// Every per-block array component (one of the 6 of a W slot, U or L; one of the 10 of a
// direction's face geometry; one of the 3 of the vertices) is below 4 GiB; component bases
// are formed in 64-bit arithmetic.
__device__ __forceinline__ unsigned boff(int c)
{
    return (unsigned)c * (unsigned)sizeof(real);
}
__device__ __forceinline__ real gld(const real* base, unsigned off)
{
    return *(const CHKN_GMEM real*)((const CHKN_GMEM char*)gmem(base) + off);
}
__device__ __forceinline__ void gst(real* base, unsigned off, real v)
{
    *(CHKN_GMEM real*)((CHKN_GMEM char*)gmem(base) + off) = v;
}
// Non-temporal forms, for data a stage streams through once (U and L of the RK stage and
// the stores of the next stage's state): they are kept from evicting the W stencil lines
// that neighbouring planes and tiles read again shortly.
__device__ __forceinline__ real gld_nt(const real* base, unsigned off)
{
    return __builtin_nontemporal_load((const CHKN_GMEM real*)((const CHKN_GMEM char*)gmem(base) + off));
}
__device__ __forceinline__ void gst_nt(real* base, unsigned off, real v)
{
    __builtin_nontemporal_store(v, (CHKN_GMEM real*)((CHKN_GMEM char*)gmem(base) + off));
}

__device__ __forceinline__ Prim ld_prim(const real* W, long long n, int c)
{
    const unsigned o = boff(c);
    Prim p;
    p.rho = gld(W, o);
    p.e = gld(W + n, o);
    p.YB = gld(W + 2 * n, o);
    p.vx = gld(W + 3 * n, o);
    p.vy = gld(W + 4 * n, o);
    p.vz = gld(W + 5 * n, o);
    return p;
}

__device__ __forceinline__ Frame ld_frame(const real* g, long long nf, int f)
{
    const unsigned o = boff(f);
    Frame r;
    r.nx = gld(g + 1 * nf, o);
    r.ny = gld(g + 2 * nf, o);
    r.nz = gld(g + 3 * nf, o);
    r.t1x = gld(g + 4 * nf, o);
    r.t1y = gld(g + 5 * nf, o);
    r.t1z = gld(g + 6 * nf, o);
    r.t2x = gld(g + 7 * nf, o);
    r.t2y = gld(g + 8 * nf, o);
    r.t2z = gld(g + 9 * nf, o);
    return r;
}

// Vertex v = i + (nic+1)*(j + (njc+1)*k) of a block.
__device__ __forceinline__ geom::V3 ld_vertex(const BlockDev* B, int v)
{
    const unsigned o = boff(v);
    const long long nv = B->nv;
    return geom::V3{gld(B->vtx, o), gld(B->vtx + nv, o), gld(B->vtx + 2 * nv, o)};
}

// Area and frame of the face whose vertices are v0 + d[0..3] (in the reference's vertex
// order), formed as the host geometry pass forms them (quad_properties, vector3.cu:196-245),
// so the bits are those of the stored face geometry.
__device__ __forceinline__ void vertex_frame(const BlockDev* B, int v0, int d1, int d2, int d3,
                                             Frame& f, real& area)
{
    geom::Quad q;
    geom::quad_frame(ld_vertex(B, v0), ld_vertex(B, v0 + d1), ld_vertex(B, v0 + d2),
                     ld_vertex(B, v0 + d3), q);
    area = q.area;
    f.nx = q.n.x; f.ny = q.n.y; f.nz = q.n.z;
    f.t1x = q.t1.x; f.t1y = q.t1.y; f.t1z = q.t1.z;
    f.t2x = q.t2.x; f.t2y = q.t2.y; f.t2z = q.t2.z;
}

// Frame of the dir-face (i, j, k): i-face vertex order block.cu:267-270, j-face :368-371,
// k-face :469-472 (face index conventions as the stored geometry).
template <int DIR>
__device__ __forceinline__ void face_frame_v(const BlockDev* B, int i, int j, int k, Frame& f, real& area)
{
    const int sj = B->nic + 1, sk = (B->nic + 1) * (B->njc + 1);
    const int v = i + sj * j + sk * k;
    if (DIR == 0) vertex_frame(B, v, sj, sj + sk, sk, f, area);
    else if (DIR == 1) vertex_frame(B, v + 1, -1, sk - 1, sk, f, area);
    else vertex_frame(B, v, 1, 1 + sj, sj, f, area);
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
                                              const real*& g, int& nf, int& fi)
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

// Ghost-cell state for a stencil position outside the block (m < 0 or m >= n).  Inlined:
// an out-of-line call would make every caller preserve its live registers across the
// call ABI, which costs about 30 VGPRs and one wave per SIMD of occupancy.
// Depth d = -m-1 on the minus side and m-n on the plus side (block.cu:272-316).
// VFRAME: wall frames formed from the block's vertices rather than read from the stored
// face geometry (same bits).
template <bool VFRAME = false>
__device__ __forceinline__ Prim ghost_prim(const BlockDev* blocks, const real* inflow, int win,
                                           const BlockDev* B, int dir, int m, int a, int c)
{
    int n = dir == 0 ? B->nic : (dir == 1 ? B->njc : B->nkc);
    int side = (m < 0) ? 0 : 1;
    int depth = side ? m - n : -m - 1;
    int bf = 2 * dir + side;
    int code = B->bc[bf];
    if (code == 3) {  // exchange: neighbour's interior cell at the same depth (bcs.cu:135-254)
        const BlockDev* O = blocks + B->bc_arg[bf];
        int no = dir == 0 ? O->nic : (dir == 1 ? O->njc : O->nkc);
        int mo = side ? depth : no - 1 - depth;
        return ld_prim(O->W + win * O->wstride, O->n, cell_index(O->nic, O->njc, dir, mo, a, c));
    }
    if (code == 4) {  // inflow: the configured flow state (bcs.cu:257-319)
        const CHKN_GMEM real* s = gmem(inflow) + 6 * B->bc_arg[bf];
        Prim p;
        p.rho = s[0]; p.e = s[1]; p.YB = s[2]; p.vx = s[3]; p.vy = s[4]; p.vz = s[5];
        return p;
    }
    if (code == 6) {  // inflow function: states at the ghost positions (bcs.cu:355-432)
        const int na = dir == 0 ? B->njc : B->nic;
        const int nt = na * (dir == 2 ? B->njc : B->nkc);
        const CHKN_GMEM real* s = gmem(B->side[bf]) + (long long)depth * 9 * nt + (a + na * c);
        Prim p;
        p.rho = s[0]; p.e = s[nt]; p.YB = s[2 * nt]; p.vx = s[3 * nt]; p.vy = s[4 * nt]; p.vz = s[5 * nt];
        return p;
    }
    Prim in = ld_prim(B->W + win * B->wstride, B->n,
                      cell_index(B->nic, B->njc, dir, side ? n - 1 : 0, a, c));
    if (code == 5) return in;  // outflow copies the adjacent interior cell (bcs.cu:435-503)
    // Walls: both ghost layers reflect the adjacent interior cell (bcs.cu:16-132).
    if (VFRAME) {
        Frame f;
        real area;
        if (dir == 0) face_frame_v<0>(B, side ? B->nic : 0, a, c, f, area);
        else if (dir == 1) face_frame_v<1>(B, a, side ? B->njc : 0, c, f, area);
        else face_frame_v<2>(B, a, c, side ? B->nkc : 0, f, area);
        return reflect(in, f);
    }
    const real* g; int nf, fi;
    boundary_face(B, dir, side, a, c, g, nf, fi);
    return reflect(in, ld_frame(g, nf, fi));
}

// Bit pattern of a non-negative real as an unsigned 64-bit integer: non-negative values
// order as their patterns, so integer minima give the minimum value.
__device__ __forceinline__ unsigned long long order_bits(double x)
{
    return (unsigned long long)__double_as_longlong(x);
}
__device__ __forceinline__ unsigned long long order_bits(float x)
{
    return (unsigned long long)__float_as_uint(x);
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}

// Stages of the update are numbered 1, 2, 3 for TVD-RK3 and 1, 4 for the two-stage
// predictor-corrector; the bad-cell counter of a stage is its position in the step.
__host__ __device__ constexpr int stage_position(int stage) { return stage == 4 ? 2 : stage; }

// RK stage of cell cidx from its flux divergence d.  TVD-RK3 (block.cu:1090-1160): stage 1
// keeps d in L, stage 2 replaces L by L/6 + d/6, stage 3 writes the new U.  Predictor-
// corrector (src/puffin/cell.d:294-303): stage 1 as RK3's, stage 4 writes
// U + 0.5 dt (L + d).  Decodes the stage result (flow.cu:245-268) into W slot wout and
// returns the decode's bad-cell flag.
template <int STAGE>
__device__ __forceinline__ int stage_update(const BlockDev* B, long long n, int cidx, real dt,
                                            const real d[6], int wout, Prim& w)
{
    real* U = B->U;
    real* Lr = B->L;
    const unsigned o = boff(cidx);
    real Un[6];
    if (STAGE == 1) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            Un[q] = gld_nt(U + q * n, o) + dt * d[q];
            gst_nt(Lr + q * n, o, d[q]);
        }
    } else if (STAGE == 2) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const real l0 = gld_nt(Lr + q * n, o);
            Un[q] = gld_nt(U + q * n, o) + real(0.25) * dt * (l0 + d[q]);
            gst_nt(Lr + q * n, o, kOneSixth * l0 + kOneSixth * d[q]);
        }
    } else if (STAGE == 3) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            Un[q] = gld_nt(U + q * n, o) + dt * (gld_nt(Lr + q * n, o) + kFourSixth * d[q]);
            gst_nt(U + q * n, o, Un[q]);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            Un[q] = gld_nt(U + q * n, o) + real(0.5) * dt * (gld_nt(Lr + q * n, o) + d[q]);
            gst_nt(U + q * n, o, Un[q]);
        }
    }
    const int flag = decode(Un, w);
    real* Wout = B->W + wout * B->wstride;
    gst_nt(Wout, o, w.rho);
    gst_nt(Wout + n, o, w.e);
    gst_nt(Wout + 2 * n, o, w.YB);
    gst_nt(Wout + 3 * n, o, w.vx);
    gst_nt(Wout + 4 * n, o, w.vy);
    gst_nt(Wout + 5 * n, o, w.vz);
    return flag;
}

// Bits of the cfl-free signal time of cell (i, j, k) = cidx in state w (cell.cu:155-164);
// non-negative reals order as their bit patterns, so integer minima give the minimum time.
template <bool VFRAME = false>
__device__ __forceinline__ unsigned long long signal_bits(const BlockDev* B, long long n, int i, int j,
                                                         int k, int cidx, const Prim& w)
{
    const int nic = B->nic, njc = B->njc;
    const real T = eos_T(w.e, w.YB);
    const real sa = eos_a(T);
    Frame fi_, fj_, fk_;
    if (VFRAME) {
        real area;
        face_frame_v<0>(B, i, j, k, fi_, area);
        face_frame_v<1>(B, i, j, k, fj_, area);
        face_frame_v<2>(B, i, j, k, fk_, area);
    } else {
        fi_ = ld_frame(B->gi, B->nfi, i + (nic + 1) * (j + njc * k));
        fj_ = ld_frame(B->gj, B->nfj, i + nic * (j + (njc + 1) * k));
        fk_ = ld_frame(B->gk, B->nfk, cidx);
    }
    const unsigned o = boff(cidx);
    const real sig = local_signal(w, sa, gld(B->len, o), gld(B->len + n, o), gld(B->len + 2 * n, o),
                                  fi_, fj_, fk_);
    return order_bits(sig);
}

}  // namespace chkn
