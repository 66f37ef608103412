"""Post-processing of chicken job directories: readers, probes, slices and VTK output.

The public functions mirror chkn_post.py (read_config :73-90, read_grids :92-113,
read_flow_blocks :115-128, read_block_of_flow_data :130-152, reshape_flow_data_arrays
:154-169, find_nearest_cell :171-195, write_slice_list_to_gnuplot_file :199-229,
write_pvd_file :233-250, write_vtk_files :252-301).  State lives in a ``PostData``
object instead of module globals; ``main`` provides the chkn-post command line.
"""
import gzip
import json
import math
import os
import sys

import numpy as np

from .geom import StructuredGrid


class PostData:
    def __init__(self):
        self.config = {}
        self.times = {}
        self.grids = {}
        self.flows = {}

    # ------------------------------------------------------------------ readers
    def read_config(self, jobDir):
        with open(os.path.join(jobDir, "config.json")) as fp:
            self.config = json.load(fp)
        self.times = {}
        with open(os.path.join(jobDir, "times.data")) as fp:
            for line in fp:
                if not line.strip() or line[0] == "#":
                    continue
                items = line.split()
                self.times[int(items[0])] = float(items[1])

    def read_grids(self, jobDir, binaryData):
        c = self.config
        for k in range(c["nkb"]):
            for j in range(c["njb"]):
                for i in range(c["nib"]):
                    name = os.path.join(jobDir, "grid", "grid-%04d-%04d-%04d" % (i, j, k))
                    name += ".bin" if binaryData else ".gz"
                    if os.path.exists(name):
                        self.grids["%d,%d,%d" % (i, j, k)] = (
                            StructuredGrid.read_from_binary_file(name) if binaryData
                            else StructuredGrid.read_from_gzip_file(name))

    def read_flow_blocks(self, jobDir, tindx, binaryData):
        c = self.config
        d = os.path.join(jobDir, "flow", "t%04d" % tindx)
        if not os.path.exists(d):
            raise RuntimeError("Cannot find flow directory: " + d)
        for k in range(c["nkb"]):
            for j in range(c["njb"]):
                for i in range(c["nib"]):
                    name = os.path.join(d, "flow-%04d-%04d-%04d" % (i, j, k))
                    name += ".bin" if binaryData else ".gz"
                    if os.path.exists(name):
                        self.flows["%d,%d,%d" % (i, j, k)] = self.read_block_of_flow_data(name, binaryData)

    def read_block_of_flow_data(self, fileName, binaryData):
        """One flat array per IO variable (cell order i fastest)."""
        if binaryData:
            data = np.fromfile(fileName, dtype=float)
        else:
            with gzip.open(fileName, "rt") as f:
                data = np.loadtxt(f)
        names = self.config["iovar_names"]
        data = data.reshape(len(names), -1)
        return {v: data[q, :] for q, v in enumerate(names)}

    def reshape_flow_data_arrays(self):
        """Flat arrays -> [i][j][k] arrays per block."""
        c = self.config
        for ib in range(c["nib"]):
            for jb in range(c["njb"]):
                for kb in range(c["nkb"]):
                    key = "%d,%d,%d" % (ib, jb, kb)
                    if c["blk_ids"][ib][jb][kb] >= 0 and key in self.flows:
                        fd = self.flows[key]
                        for var in fd:
                            fd[var] = fd[var].reshape((c["nkcs"][kb], c["njcs"][jb], c["nics"][ib])).transpose()

    def find_nearest_cell(self, x, y, z):
        best = {"ib": -1, "jb": -1, "kb": -1, "ic": -1, "jc": -1, "kc": -1,
                "distance": sys.float_info.max}
        for key, fd in self.flows.items():
            ib, jb, kb = (int(s) for s in key.split(","))
            d2 = (fd["posx"] - x) ** 2 + (fd["posy"] - y) ** 2 + (fd["posz"] - z) ** 2
            idx = np.unravel_index(int(np.argmin(d2)), d2.shape)
            dist = math.sqrt(float(d2[idx]))
            if dist < best["distance"]:
                best.update(ib=ib, jb=jb, kb=kb, ic=int(idx[0]), jc=int(idx[1]), kc=int(idx[2]),
                            distance=dist)
        return best

    # ------------------------------------------------------------------ writers
    def write_slice_list_to_gnuplot_file(self, sliceList, fileName):
        c = self.config
        with open(fileName, "wt") as f:
            f.write("# " + " ".join(c["iovar_names"]) + " \n")
            for s in sliceList:
                ib, jb, kb = s["ib"], s["jb"], s["kb"]
                if c["blk_ids"][ib][jb][kb] < 0:
                    continue
                fd = self.flows["%d,%d,%d" % (ib, jb, kb)]
                if s["dir"] == "i":
                    cells = [(ic, s["jc"], s["kc"]) for ic in range(c["nics"][ib])]
                elif s["dir"] == "j":
                    cells = [(s["ic"], jc, s["kc"]) for jc in range(c["njcs"][jb])]
                else:
                    cells = [(s["ic"], s["jc"], kc) for kc in range(c["nkcs"][kb])]
                for (ic, jc, kc) in cells:
                    f.write("".join("%e " % fd[v][ic][jc][kc] for v in c["iovar_names"]) + "\n")

    def write_pvd_file(self, jobDir, tindxList, timesList):
        plotDir = os.path.join(jobDir, "plot")
        os.makedirs(plotDir, exist_ok=True)
        with open(os.path.join(plotDir, "flow.pvd"), "w") as fp:
            fp.write('<?xml version="1.0"?>\n')
            fp.write('<VTKFile type="Collection" version="0.1" byte_order="BigEndian">\n<Collection>\n')
            for tindx, ts in zip(tindxList, timesList):
                fp.write('<DataSet timestep="%.18e" group="" part="0" file="flow-t%04d.pvts"/>\n' % (ts, tindx))
            fp.write("</Collection>\n</VTKFile>\n")

    def write_vtk_files(self, jobDir, tindx):
        """flow-tNNNN.pvts plus one .vts piece per block; flows must be flat (not reshaped)."""
        c = self.config
        plotDir = os.path.join(jobDir, "plot")
        os.makedirs(plotDir, exist_ok=True)
        wn = (sum(c["nics"]) + 1, sum(c["njcs"]) + 1, sum(c["nkcs"]) + 1)
        names = c["iovar_names"]
        with open(os.path.join(plotDir, "flow-t%04d.pvts" % tindx), "w") as fp:
            fp.write('<VTKFile type="PStructuredGrid" version="0.1" byte_order="BigEndian">\n')
            fp.write('<PStructuredGrid WholeExtent="0 %d 0 %d 0 %d" GhostLevel="0">\n<PCellData>\n'
                     % (wn[0] - 1, wn[1] - 1, wn[2] - 1))
            for v in names:
                fp.write('<PDataArray Name="%s" type="Float64" NumberOfComponents="1" format="ascii" />\n' % v)
            fp.write('<PDataArray Name="Mach" type="Float64" NumberOfComponents="1" format="ascii" />\n')
            fp.write('<PDataArray Name="vel.vector" type="Float64" NumberOfComponents="3" format="ascii" />\n')
            fp.write('</PCellData>\n<PPoints>\n<PDataArray type="Float64" NumberOfComponents="3" format="ascii" />\n</PPoints>\n')
            sk = 0
            for k in range(c["nkb"]):
                sj = 0
                for j in range(c["njb"]):
                    si = 0
                    for i in range(c["nib"]):
                        key = "%d,%d,%d" % (i, j, k)
                        piece = "flow-t%04d-%04d-%04d-%04d.vts" % (tindx, i, j, k)
                        if key in self.grids and key in self.flows:
                            g = self.grids[key]
                            ext = (si, si + g.niv - 1, sj, sj + g.njv - 1, sk, sk + g.nkv - 1)
                            fp.write('<Piece Extent="%d %d %d %d %d %d" Source="%s" />\n' % (ext + (piece,)))
                            self._write_vts(os.path.join(plotDir, piece), g, self.flows[key], wn, ext)
                        si += c["nics"][i]
                    sj += c["njcs"][j]
                sk += c["nkcs"][k]
            fp.write("</PStructuredGrid>\n</VTKFile>\n")

    def _write_vts(self, fileName, g, fd, wn, ext):
        with open(fileName, "w") as fp:
            fp.write('<VTKFile type="StructuredGrid" version="0.1" byte_order="BigEndian">\n')
            fp.write('<StructuredGrid WholeExtent="0 %d 0 %d 0 %d">\n' % (wn[0] - 1, wn[1] - 1, wn[2] - 1))
            fp.write('<Piece Extent="%d %d %d %d %d %d">\n<CellData>\n' % ext)
            for v in self.config["iovar_names"]:
                fp.write('<DataArray Name="%s" type="Float64" NumberOfComponents="1" format="ascii">\n' % v)
                fp.write("".join("%g\n" % x for x in np.ravel(fd[v])))
                fp.write("</DataArray>\n")
            vx, vy, vz = (np.ravel(fd[k]) for k in ("velx", "vely", "velz"))
            mach = np.sqrt(vx * vx + vy * vy + vz * vz) / np.ravel(fd["a"])
            fp.write('<DataArray Name="Mach" type="Float64" NumberOfComponents="1" format="ascii">\n')
            fp.write("".join("%g\n" % x for x in mach))
            fp.write('</DataArray>\n<DataArray Name="vel.vector" type="Float64" NumberOfComponents="3" format="ascii">\n')
            fp.write("".join("%g %g %g\n" % t for t in zip(vx, vy, vz)))
            fp.write("</DataArray>\n</CellData>\n<Points>\n")
            fp.write('<DataArray type="Float64" NumberOfComponents="3" format="ascii">\n')
            fp.write("".join("%g %g %g\n" % tuple(p) for p in g.xyz_flat()))
            fp.write("</DataArray>\n</Points>\n</Piece>\n</StructuredGrid>\n</VTKFile>\n")


def main(argv=None):
    """chkn-post --job=<jobName> [--binary] [--tindx=<int|all|last>] [--probe=x,y,z]"""
    import getopt
    argv = sys.argv[1:] if argv is None else argv
    opts, _ = getopt.getopt(argv, "hf:bt:p:", ["help", "job=", "binary", "tindx=", "probe="])
    uo = dict(opts)
    if not opts or "--help" in uo or "-h" in uo:
        print("Usage: chkn-post --job=<jobName> [--binary] [--tindx=<int|all|last>] [--probe=x,y,z]")
        return 0
    job = uo.get("--job", uo.get("-f"))
    binary = "--binary" in uo or "-b" in uo
    pd = PostData()
    pd.read_config(job)
    sel = uo.get("--tindx", uo.get("-t", "last"))
    tlist = sorted(pd.times)
    tindxs = tlist if sel == "all" else ([tlist[-1]] if sel == "last" else [int(sel)])
    pd.read_grids(job, binary)
    for t in tindxs:
        pd.read_flow_blocks(job, t, binary)
        if "--probe" in uo or "-p" in uo:
            x, y, z = (float(v) for v in uo.get("--probe", uo.get("-p")).split(","))
            pd.reshape_flow_data_arrays()
            c = pd.find_nearest_cell(x, y, z)
            fd = pd.flows["%d,%d,%d" % (c["ib"], c["jb"], c["kb"])]
            print(" ".join("%s=%g" % (v, fd[v][c["ic"]][c["jc"]][c["kc"]]) for v in pd.config["iovar_names"]))
        else:
            pd.write_vtk_files(job, t)
    if "--probe" not in uo and "-p" not in uo:
        pd.write_pvd_file(job, tindxs, [pd.times[t] for t in tindxs])
    return 0
