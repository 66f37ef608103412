// chkn_geometry.h -- cell and face geometry of a structured hexahedral block, in the
// build's real type (the reference computes it in its number type).
//
// Face centroid, unit normal/tangents and area follow quad_properties
// (vector3.cu:196-245); cell centroid, volume and i/j/k lengths follow
// hex_cell_properties with true_centroid = false (vector3.cu:350-433, block.cu:640-647),
// whose volume is the sum of six tetragonal dipyramids (J. Grandy 1997), each the sum of
// four tetrahedra about the base mid-point.  Operand order matches the reference so that
// the values are bit-identical to its computeGeometry (block.cu:633-783).  The face frame
// part is host and device code: the stage kernels can form a face's frame from its four
// vertices with the same operations, so the result has the same bits as the host pre-pass.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "chkn_real.h"

#ifndef CHKN_HD
#define CHKN_HD __host__ __device__ __forceinline__
#endif

namespace chkn {
namespace geom {

struct V3 {
    real x, y, z;
};

CHKN_HD V3 normalized(V3 v)  // vector3.cu:77-93
{
    real m = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (m > real(0)) { v.x /= m; v.y /= m; v.z /= m; }
    else { v.x = v.y = v.z = real(0); }
    const real small = real(1.0e-30);
    if (fabs(v.x) < small) v.x = real(0);
    if (fabs(v.y) < small) v.y = real(0);
    if (fabs(v.z) < small) v.z = real(0);
    return v;
}

struct Quad {
    V3 c, n, t1, t2;
    real area;
};

// Area, unit normal and tangents of the quadrilateral p0 p1 p2 p3 (vector3.cu:205-244).
CHKN_HD void quad_frame(const V3& p0, const V3& p1, const V3& p2, const V3& p3, Quad& q)
{
    const real quarter = real(0.25);
    // Edge-pair sums along the two quad directions.
    real ax = p1.x - p0.x + p2.x - p3.x, ay = p1.y - p0.y + p2.y - p3.y, az = p1.z - p0.z + p2.z - p3.z;
    real bx = p3.x - p0.x + p2.x - p1.x, by = p3.y - p0.y + p2.y - p1.y, bz = p3.z - p0.z + p2.z - p1.z;
    real vx = ay * bz - az * by, vy = az * bx - ax * bz, vz = ax * by - ay * bx;
    vx *= quarter; vy *= quarter; vz *= quarter;
    q.area = sqrt(vx * vx + vy * vy + vz * vz);
    if (q.area > real(1.0e-20)) {
        q.n = V3{vx / q.area, vy / q.area, vz / q.area};
        q.t1 = normalized(V3{ax, ay, az});
        // t2 = n x t1, written as cross(v3, v1, v2) of vector3.cu:150-157 with v1 = n, v2 = t1.
        V3 c{q.n.y * q.t1.z - q.t1.y * q.n.z, q.t1.x * q.n.z - q.n.x * q.t1.z,
             q.n.x * q.t1.y - q.t1.x * q.n.y};
        q.t2 = normalized(c);
    } else {
        q.n = V3{real(1), real(0), real(0)};
        q.t1 = V3{real(0), real(1), real(0)};
        q.t2 = V3{real(0), real(0), real(1)};
    }
}

inline Quad quad(const V3& p0, const V3& p1, const V3& p2, const V3& p3)
{
    Quad q;
    const real quarter = real(0.25);
    q.c.x = quarter * (p0.x + p1.x + p2.x + p3.x);
    q.c.y = quarter * (p0.y + p1.y + p2.y + p3.y);
    q.c.z = quarter * (p0.z + p1.z + p2.z + p3.z);
    quad_frame(p0, p1, p2, p3, q);
    return q;
}

inline real tet_vol(const V3& p0, const V3& p1, const V3& p2, const V3& p3)  // vector3.cu:259-270
{
    real ax = p1.x - p0.x, ay = p1.y - p0.y, az = p1.z - p0.z;
    real bx = p2.x - p0.x, by = p2.y - p0.y, bz = p2.z - p0.z;
    real cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    real dx = p3.x - p0.x, dy = p3.y - p0.y, dz = p3.z - p0.z;
    return (dx * cx + dy * cy + dz * cz) / real(6);
}

// Volume of the dipyramid over base (b0, b1, b2, b3) with apex `apex` (vector3.cu:309-348).
inline real pyramid_vol(const V3& b0, const V3& b1, const V3& b2, const V3& b3, const V3& apex)
{
    const real quarter = real(0.25);
    V3 m{quarter * (b0.x + b1.x + b2.x + b3.x), quarter * (b0.y + b1.y + b2.y + b3.y),
         quarter * (b0.z + b1.z + b2.z + b3.z)};
    real v = real(0);
    v += tet_vol(b0, b1, m, apex);
    v += tet_vol(b1, b2, m, apex);
    v += tet_vol(b2, b3, m, apex);
    v += tet_vol(b3, b0, m, apex);
    return v;
}

struct Hex {
    V3 c;
    real vol, iL, jL, kL;
};

inline V3 mid4(const V3& a, const V3& b, const V3& c, const V3& d)
{
    const real quarter = real(0.25);
    return V3{quarter * (a.x + b.x + c.x + d.x), quarter * (a.y + b.y + c.y + d.y),
              quarter * (a.z + b.z + c.z + d.z)};
}

inline real dist(const V3& a, const V3& b)
{
    real dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// p: the 8 vertices in the cell order of block.cu:243-250.
inline Hex hex(const V3 p[8])
{
    Hex h;
    const real eighth = real(0.125);
    h.c.x = eighth * (p[0].x + p[1].x + p[2].x + p[3].x + p[4].x + p[5].x + p[6].x + p[7].x);
    h.c.y = eighth * (p[0].y + p[1].y + p[2].y + p[3].y + p[4].y + p[5].y + p[6].y + p[7].y);
    h.c.z = eighth * (p[0].z + p[1].z + p[2].z + p[3].z + p[4].z + p[5].z + p[6].z + p[7].z);
    V3 north = mid4(p[3], p[2], p[6], p[7]);
    V3 east = mid4(p[1], p[2], p[6], p[5]);
    V3 south = mid4(p[0], p[1], p[5], p[4]);
    V3 west = mid4(p[0], p[3], p[7], p[4]);
    V3 top = mid4(p[4], p[5], p[6], p[7]);
    V3 bottom = mid4(p[0], p[1], p[2], p[3]);
    h.iL = dist(east, west);
    h.jL = dist(north, south);
    h.kL = dist(top, bottom);
    real v = real(0);
    v += pyramid_vol(p[6], p[7], p[3], p[2], h.c);
    v += pyramid_vol(p[5], p[6], p[2], p[1], h.c);
    v += pyramid_vol(p[4], p[5], p[1], p[0], h.c);
    v += pyramid_vol(p[7], p[4], p[0], p[3], h.c);
    v += pyramid_vol(p[7], p[6], p[5], p[4], h.c);
    v += pyramid_vol(p[0], p[1], p[2], p[3], h.c);
    if ((v < real(0) && fabs(v) < real(1.0e-12)) || (v >= real(0) && v < real(1.0e-20))) v = real(0);
    h.vol = v;
    return h;
}

}  // namespace geom
}  // namespace chkn
