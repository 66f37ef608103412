"""The job-preparation front-end (chkn.prep): job scripts through the chkn-prep command line,
the block-array checks, makeFBArray's split and the config.json text the solver parses."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from chkn import cases, prep
from chkn.job import JobData

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gdtk-rmhd_amd")

SCRIPT = '''
config.title = "script test"
config.max_time = 1.0e-3
config.max_step = 7
add_cfl_value(0.0, 0.4)
add_dt_plot(0.0, 5.0e-4, 1.0e-5)
fs = FlowState(p=5.0e3, T=300.0, velx=700.0)
vol = TFIVolume(p000=Vector3(0, 0, 0), p100=Vector3(1, 0, 0), p110=Vector3(1, 0.5, 0),
                p010=Vector3(0, 0.5, 0), p001=Vector3(0, 0, 0.5), p101=Vector3(1, 0, 0.5),
                p111=Vector3(1, 0.5, 0.5), p011=Vector3(0, 0.5, 0.5))
grid = StructuredGrid(pvolume=vol, niv=9, njv=5, nkv=4)
makeFBArray(ni=2, nj=2, grid=grid, initialState=fs,
            bcs={"iminus": InflowBC(fs), "iplus": OutflowBC()}, label="b")
'''


def _files(d):
    out = {}
    for root, _, names in os.walk(d):
        for n in names:
            p = os.path.join(root, n)
            out[os.path.relpath(p, d)] = open(p, "rb").read()
    return out


def test_chkn_prep_command_line_runs_a_job_script(tmp_path):
    job = tmp_path / "scr"
    (tmp_path / "scr.py").write_text(SCRIPT)
    r = subprocess.run([sys.executable, "-m", "chkn.prep", "--job=scr", "--binary"], cwd=tmp_path,
                       env=dict(os.environ, PYTHONPATH=PKG), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    cfg = json.load(open(job / "config.json"))
    assert cfg["title"] == "script test" and cfg["max_step"] == 7 and cfg["cfl_values"] == [0.4]
    assert (cfg["nib"], cfg["njb"], cfg["nkb"]) == (2, 2, 1)
    assert cfg["nics"] == [4, 4] and cfg["njcs"] == [2, 2] and cfg["nkcs"] == [3]
    b0 = cfg["fluid_blocks"][0]
    assert b0["bcs"]["iplus"]["tag"] == "exchange" and b0["bcs"]["jminus"]["tag"] == "wall_with_slip"
    # The two i=0 blocks share the script's InflowBC object: each registers its flow state and
    # the shared object keeps the last index, as with the reference's prep.
    assert cfg["n_flow_states"] == 2
    assert b0["bcs"]["iminus"] == {"tag": "inflow", "flow_state_index": 1}
    assert cfg["fluid_blocks"][2]["bcs"]["iminus"] == b0["bcs"]["iminus"]
    # The same description written in-process gives the same files.
    prep.reset()
    ns = prep.script_namespace()
    exec(SCRIPT, prep.__dict__)
    assert ns  # the script's names are the module's
    d2 = tmp_path / "inproc"
    prep.write_initial_files(True, str(d2))
    assert _files(job) == _files(d2)


def test_help_and_missing_job():
    assert prep.main(["--help"]) == 0
    with pytest.raises(SystemExit):
        prep.main(["--binary"])


def test_block_array_checks():
    cases.sod_box(4, 4, 4, max_step=1)
    g = cases.box_grid(4, 4, 4, origin=(1.0, 0.0, 0.0))
    prep.FluidBlock(i=0, j=0, k=0, grid=g, initialState=prep.FlowState())
    with pytest.raises(RuntimeError, match="both sit at array position"):
        prep.config_json_text()
    cases.sod_box(4, 4, 4, max_step=1)
    prep.FluidBlock(i=1, j=0, k=0, grid=cases.box_grid(4, 5, 4), initialState=prep.FlowState())
    with pytest.raises(RuntimeError, match="cells along j"):
        prep.config_json_text()


def test_make_fb_array_splits_cells_and_places_exchanges():
    prep.reset()
    g = cases.box_grid(10, 6, 5)
    blocks = prep.makeFBArray(i0=1, ni=3, nj=2, nk=1, grid=g, initialState=prep.FlowState(),
                              bcs={prep.Face.kminus: prep.WallNoSlipAdiabaticBC()})
    assert [(b.i, b.j, b.k) for b in blocks] == [(1, 0, 0), (2, 0, 0), (3, 0, 0),
                                                 (1, 1, 0), (2, 1, 0), (3, 1, 0)]
    assert [b.nic for b in blocks[:3]] == [3, 3, 4]  # the last block takes the remainder
    assert all(b.njc == 3 and b.nkc == 5 for b in blocks)
    # Vertices of the sub-grids are the parent's.
    np.testing.assert_array_equal(blocks[4].grid.vertices.x, g.vertices.x[3:7, 3:7, :])
    assert blocks[0].bcs["iminus"].tag == "wall_with_slip" and blocks[0].bcs["iplus"].tag == "exchange"
    assert blocks[2].bcs["iplus"].tag == "wall_with_slip" and blocks[3].bcs["jminus"].tag == "exchange"
    assert all(b.bcs["kminus"].tag == "wall_no_slip_adiabatic" for b in blocks)
    with pytest.raises(ValueError):
        prep.makeFBArray(ni=20, grid=g, initialState=prep.FlowState())


def test_integer_face_key_iplus_also_sets_jplus():
    # The reference's lookup matches Face.iplus for both the iplus and the jplus side.
    prep.reset()
    bcs = prep._resolve_bcs({prep.Face.iplus: prep.OutflowBC(), prep.Face.kplus: prep.OutflowBC()})
    assert bcs["iplus"].tag == bcs["jplus"].tag == bcs["kplus"].tag == "outflow"
    assert bcs["jminus"].tag == "wall_with_slip"
    bcs = prep._resolve_bcs({"iplus": prep.OutflowBC()})
    assert bcs["jplus"].tag == "wall_with_slip"


def test_initial_state_forms_agree():
    # A FlowState, per-cell arrays and a function of position giving the same state.
    prep.reset()
    fs = prep.FlowState(p=2.0e4, T=310.0, velx=12.0, velz=-3.0)
    g = cases.box_grid(3, 2, 2)
    a = prep.FluidBlock(grid=g, initialState=fs)
    arr = {k: np.full((3, 2, 2), v) for k, v in zip(prep.FLOW_KEYS, fs.cell_values())}
    b = prep.FluidBlock(i=1, grid=g, initialState=arr)
    c = prep.FluidBlock(i=2, grid=g, initialState=lambda x, y, z: fs)
    for blk in (b, c):
        np.testing.assert_array_equal(blk.iovars(), a.iovars())
    jd = JobData.from_prep()
    assert len(jd.blocks) == 3


def test_schedules_keep_increasing_times(capsys):
    prep.reset()
    assert prep.add_cfl_value(0.0, 0.5) == 1
    assert prep.add_cfl_value(1.0e-3, 0.7) == 2
    assert prep.add_cfl_value(1.0e-3, 0.9) == 2
    assert "ignored" in capsys.readouterr().out
    prep.config.cfl_list[0] = (0.0, 0.0)
    prep.FluidBlock(grid=cases.box_grid(2, 2, 2), initialState=prep.FlowState())
    with pytest.raises(ValueError, match="CFL values must be positive"):
        prep.config_json_text()
