"""Geometry helpers for job scripts: Vector3, TFIVolume and StructuredGrid.

These provide the names chicken job scripts use from the gdtk geometry package
(gdtk.geom.vector3.Vector3, gdtk.geom.volume.TFIVolume, gdtk.geom.sgrid.StructuredGrid),
restricted to what a chicken job needs: hexahedral volumes with straight edges, uniform
vertex distribution, sub-grids and the grid file formats of sgrid.py:250-289.
The solver recomputes all cell and face geometry from the vertices (block.cu:633-783),
so only the vertex coordinates written here matter to a run.
"""
import gzip

import numpy as np


class Vector3:
    """3-vector whose components may be scalars or numpy arrays."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x=0.0, y=0.0, z=0.0):
        if isinstance(x, Vector3):
            x, y, z = x.x, x.y, x.z
        elif isinstance(x, (list, tuple)) and not np.isscalar(x):
            x, y, z = (list(x) + [0.0, 0.0, 0.0])[:3]
        elif isinstance(x, dict):
            x, y, z = x.get("x", 0.0), x.get("y", 0.0), x.get("z", 0.0)
        self.x, self.y, self.z = x, y, z

    def __add__(self, o):
        return Vector3(self.x + o.x, self.y + o.y, self.z + o.z)

    def __sub__(self, o):
        return Vector3(self.x - o.x, self.y - o.y, self.z - o.z)

    def __mul__(self, s):
        return Vector3(self.x * s, self.y * s, self.z * s)

    __rmul__ = __mul__

    def __truediv__(self, s):
        return Vector3(self.x / s, self.y / s, self.z / s)

    def __neg__(self):
        return Vector3(-self.x, -self.y, -self.z)

    def __abs__(self):
        return np.sqrt(self.x * self.x + self.y * self.y + self.z * self.z)

    def dot(self, o):
        return self.x * o.x + self.y * o.y + self.z * o.z

    def unit(self):
        return self / abs(self)

    def __repr__(self):
        return "Vector3(x={}, y={}, z={})".format(self.x, self.y, self.z)


def cross(a, b):
    return Vector3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x)


class TFIVolume:
    """Hexahedral volume with straight edges, parameterised by (r, s, t) in [0,1]^3.

    For straight edges the transfinite interpolation of the bounding surfaces is the
    trilinear blend of the eight corners, which is what is evaluated here.
    """

    def __init__(self, p000, p100, p110, p010, p001, p101, p111, p011):
        self.p = [Vector3(p) for p in (p000, p100, p110, p010, p001, p101, p111, p011)]

    def __call__(self, r, s, t):
        p000, p100, p110, p010, p001, p101, p111, p011 = self.p
        w = [(1 - r) * (1 - s) * (1 - t), r * (1 - s) * (1 - t), r * s * (1 - t), (1 - r) * s * (1 - t),
             (1 - r) * (1 - s) * t, r * (1 - s) * t, r * s * t, (1 - r) * s * t]
        x = sum(wi * pi.x for wi, pi in zip(w, self.p))
        y = sum(wi * pi.y for wi, pi in zip(w, self.p))
        z = sum(wi * pi.z for wi, pi in zip(w, self.p))
        return Vector3(x, y, z)


class StructuredGrid:
    """Vertices of a 3D structured block, arrays of shape (niv, njv, nkv)."""

    def __init__(self, pvolume=None, niv=2, njv=2, nkv=2, empty=False, x=None, y=None, z=None,
                 label=""):
        self.label = label
        self.dimensions = 3
        if x is not None:
            self.vertices = Vector3(np.asarray(x, dtype=float), np.asarray(y, dtype=float),
                                    np.asarray(z, dtype=float))
            self.niv, self.njv, self.nkv = self.vertices.x.shape
            return
        if empty:
            self.vertices = None
            self.niv = self.njv = self.nkv = 0
            return
        self.niv, self.njv, self.nkv = niv, njv, nkv
        r = np.linspace(0.0, 1.0, niv)
        s = np.linspace(0.0, 1.0, njv)
        t = np.linspace(0.0, 1.0, nkv)
        R, S, T = np.meshgrid(r, s, t, indexing="ij")
        p = pvolume(R, S, T)
        self.vertices = Vector3(np.asarray(p.x, dtype=float) * np.ones_like(R),
                                np.asarray(p.y, dtype=float) * np.ones_like(R),
                                np.asarray(p.z, dtype=float) * np.ones_like(R))

    def subgrid(self, i0=0, j0=0, k0=0, niv=1, njv=1, nkv=1):
        """Copy of vertices [i0:i0+niv, j0:j0+njv, k0:k0+nkv] (sgrid.py:136-171)."""
        v = self.vertices
        return StructuredGrid(x=v.x[i0:i0 + niv, j0:j0 + njv, k0:k0 + nkv].copy(),
                              y=v.y[i0:i0 + niv, j0:j0 + njv, k0:k0 + nkv].copy(),
                              z=v.z[i0:i0 + niv, j0:j0 + njv, k0:k0 + nkv].copy())

    def xyz_flat(self):
        """(niv*njv*nkv, 3) array with i fastest, then j, then k (the file order)."""
        v = self.vertices
        return np.stack([v.x.transpose().ravel(), v.y.transpose().ravel(),
                         v.z.transpose().ravel()], axis=1)

    def write_to_binary_file(self, file_name):
        """float64 stream [dims, 0, 0], [niv, njv, nkv], then x y z per vertex (sgrid.py:275-289)."""
        data = np.zeros((self.niv * self.njv * self.nkv + 2, 3), dtype=float)
        data[0, :] = [float(self.dimensions), 0.0, 0.0]
        data[1, :] = [float(self.niv), float(self.njv), float(self.nkv)]
        data[2:, :] = self.xyz_flat()
        data.tofile(file_name)

    def write_to_gzip_file(self, file_name):
        """Text header and one 'x y z' line per vertex (sgrid.py:250-273)."""
        with gzip.open(file_name, "wt") as f:
            f.write("structured_grid 1.0\n")
            f.write("label: %s\n" % self.label)
            f.write("dimensions: %d\n" % self.dimensions)
            f.write("niv: %d\n" % self.niv)
            f.write("njv: %d\n" % self.njv)
            f.write("nkv: %d\n" % self.nkv)
            np.savetxt(f, self.xyz_flat())

    @staticmethod
    def read_from_binary_file(file_name):
        data = np.fromfile(file_name, dtype=float).reshape(-1, 3)
        niv, njv, nkv = (int(v) for v in data[1])
        xyz = data[2:]
        shp = (nkv, njv, niv)
        return StructuredGrid(x=xyz[:, 0].reshape(shp).transpose(), y=xyz[:, 1].reshape(shp).transpose(),
                              z=xyz[:, 2].reshape(shp).transpose())

    @staticmethod
    def read_from_gzip_file(file_name):
        with gzip.open(file_name, "rt") as f:
            f.readline()
            f.readline()
            f.readline()
            niv = int(f.readline().split()[1])
            njv = int(f.readline().split()[1])
            nkv = int(f.readline().split()[1])
            xyz = np.loadtxt(f, max_rows=niv * njv * nkv).reshape(-1, 3)
        shp = (nkv, njv, niv)
        return StructuredGrid(x=xyz[:, 0].reshape(shp).transpose(), y=xyz[:, 1].reshape(shp).transpose(),
                              z=xyz[:, 2].reshape(shp).transpose())


def cell_centres(grid):
    """Vertex-average cell centres, arrays of shape (nic, njc, nkc)."""
    v = grid.vertices
    out = []
    for a in (v.x, v.y, v.z):
        s = (a[:-1, :-1, :-1] + a[1:, :-1, :-1] + a[1:, 1:, :-1] + a[:-1, 1:, :-1]
             + a[:-1, :-1, 1:] + a[1:, :-1, 1:] + a[1:, 1:, 1:] + a[:-1, 1:, 1:])
        out.append(0.125 * s)
    return Vector3(*out)


def hex_volumes(grid):
    """Cell volumes from two-tetrahedra-free divergence formula (for the flow-file 'vol')."""
    v = grid.vertices
    P = [np.stack([a[di:a.shape[0] - 1 + di, dj:a.shape[1] - 1 + dj, dk:a.shape[2] - 1 + dk]
                   for a in (v.x, v.y, v.z)], axis=-1)
         for (di, dj, dk) in ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                              (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))]
    # Volume as the sum of six tetrahedra sharing the 0-6 diagonal.
    def tet(a, b, c, d):
        return np.einsum("...i,...i->...", d - a, np.cross(b - a, c - a)) / 6.0
    p0, p1, p2, p3, p4, p5, p6, p7 = P
    vol = (tet(p0, p1, p2, p6) + tet(p0, p2, p3, p6) + tet(p0, p3, p7, p6)
           + tet(p0, p7, p4, p6) + tet(p0, p4, p5, p6) + tet(p0, p5, p1, p6))
    return vol
