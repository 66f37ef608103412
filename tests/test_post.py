"""chkn.post (the chkn-post equivalent) on a job directory written by chkn.prep, on the CPU.

VTK output: the .pvts index names a piece per block with the block's extent in the whole
grid, and each .vts piece holds the block's vertices and the flow file's cell values (to the
'%g' text precision), Mach = |v|/a; the .pvd collection lists the time index.  The probe
reports the cell nearest the point (checked by brute force).  Binary and gzipped-text job
directories read the same.
"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from chkn import cases, post, prep


def _floats(elem):
    return np.array(elem.text.split(), dtype=float)


@pytest.mark.parametrize("binary", [True, False])
def test_vtk_pieces_hold_grid_and_flow(tmp_path, binary):
    job = cases.mixed_bc_array(max_step=5)
    d = str(tmp_path)
    prep.write_initial_files(binary, d)
    args = ["--job=" + d, "--tindx=0"] + (["--binary"] if binary else [])
    assert post.main(args) == 0
    plot = os.path.join(d, "plot")
    idx = ET.parse(os.path.join(plot, "flow-t0000.pvts")).getroot()
    pieces = idx.find("PStructuredGrid").findall("Piece")
    assert len(pieces) == len(job.blocks)
    offsets = {"i": np.concatenate([[0], np.cumsum(prep.config.nics)]),
               "j": np.concatenate([[0], np.cumsum(prep.config.njcs)]),
               "k": np.concatenate([[0], np.cumsum(prep.config.nkcs)])}
    for b, fb in zip(job.blocks, prep.fluidBlocksList):
        src = "flow-t0000-%04d-%04d-%04d.vts" % (b.i, b.j, b.k)
        piece = [p for p in pieces if p.get("Source") == src]
        assert len(piece) == 1
        ext = [int(v) for v in piece[0].get("Extent").split()]
        assert ext == [offsets["i"][b.i], offsets["i"][b.i] + b.nic, offsets["j"][b.j],
                       offsets["j"][b.j] + b.njc, offsets["k"][b.k], offsets["k"][b.k] + b.nkc]
        root = ET.parse(os.path.join(plot, src)).getroot()
        cd = {a.get("Name"): a for a in root.iter("DataArray") if a.get("Name")}
        io = np.asarray(b.iovars)
        for q, name in enumerate(("posx", "posy", "posz", "vol", "p", "T", "rho", "e", "YB", "a",
                                  "velx", "vely", "velz")):
            np.testing.assert_allclose(_floats(cd[name]), io[q], rtol=1e-5, atol=1e-12)
        mach = np.sqrt(io[10] ** 2 + io[11] ** 2 + io[12] ** 2) / io[9]
        np.testing.assert_allclose(_floats(cd["Mach"]), mach, rtol=1e-5, atol=1e-12)
        pts = _floats(root.find("StructuredGrid/Piece/Points/DataArray")).reshape(-1, 3)
        np.testing.assert_allclose(pts, np.asarray(b.vertices).reshape(-1, 3), rtol=1e-5, atol=1e-9)
    pvd = ET.parse(os.path.join(plot, "flow.pvd")).getroot()
    ds = pvd.find("Collection").findall("DataSet")
    assert [x.get("file") for x in ds] == ["flow-t0000.pvts"]


def test_probe_reports_the_nearest_cell(tmp_path, capsys):
    job = cases.mixed_bc_array(max_step=5)
    d = str(tmp_path)
    prep.write_initial_files(True, d)
    point = (0.37, 0.11, 0.05)
    assert post.main(["--job=" + d, "--binary", "--tindx=0", "--probe=%g,%g,%g" % point]) == 0
    out = capsys.readouterr().out.strip().splitlines()[-1]
    got = dict(item.split("=") for item in out.split())
    best, bio, bc = None, None, None
    for b in job.blocks:
        io = np.asarray(b.iovars)
        dist = (io[0] - point[0]) ** 2 + (io[1] - point[1]) ** 2 + (io[2] - point[2]) ** 2
        c = int(np.argmin(dist))
        if best is None or dist[c] < best:
            best, bio, bc = dist[c], io, c
    assert float(got["rho"]) == pytest.approx(bio[6, bc], rel=1e-5)
    assert float(got["posx"]) == pytest.approx(bio[0, bc], rel=1e-5)
