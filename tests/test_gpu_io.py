"""Job-directory I/O of the drop-in on the GPU: gzipped-text grids and flow files, snapshots
by the dt_plot schedule, times.data, restart from a time index, and the chkn-run command line.

tests/golden/march_text.npz holds what the reference's own driver wrote for the same job
(written by tests/golden/make_golden.py): the times.data text and the SHA-256 of every flow
file it wrote, for the march from time index 0 and for the restart from time index 2.  The
files written here must be the same text, byte for byte.
"""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

from chkn import prep
from chkn.lib import Sim

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CHKN_RUN = os.path.join(ROOT, "gdtk-rmhd_amd", "chkn", "chkn-run")


def _golden_module():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _check(job_dir, gold, times_key, names_key, digests_key, first_tindx, mg):
    with open(os.path.join(job_dir, "times.data")) as f:
        assert f.read() == str(gold[times_key])
    got = mg.written_flow_digests(job_dir, first_tindx)
    want = dict(zip(gold[names_key].tolist(), gold[digests_key].tolist()))
    assert sorted(got) == sorted(want)
    bad = [k for k in want if got[k] != want[k]]
    assert not bad, bad


def _run_cli(*args):
    r = subprocess.run([CHKN_RUN] + list(args), capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "bad cell" not in r.stderr and "Error" not in r.stderr
    return r


def test_text_march_and_restart_through_chkn_run(tmp_path):
    mg = _golden_module()
    gold = np.load(os.path.join(HERE, "golden", "march_text.npz"))
    mg.text_march_job()
    d = str(tmp_path)
    prep.write_initial_files(False, d)
    _run_cli("--job=" + d)
    _check(d, gold, "times", "names", "digests", 1, mg)
    _run_cli("--job=" + d, "--tindx=2")
    _check(d, gold, "times_restart", "names_restart", "digests_restart", 3, mg)


def test_text_march_through_the_library(tmp_path):
    mg = _golden_module()
    gold = np.load(os.path.join(HERE, "golden", "march_text.npz"))
    mg.text_march_job()
    d = str(tmp_path)
    prep.write_initial_files(False, d)
    sim = Sim(job_dir=d, binary=False)
    sim.march(binary=False)
    sim.finalize(binary=False)
    sim.close()
    _check(d, gold, "times", "names", "digests", 1, mg)


def test_reacting_gas_cli_matches_reference_build(tmp_path):
    # chkn-run-ab (the A->B gas build) marches the reacting fixture job to its max_step and
    # writes the final snapshot: bit for bit the reference AB build's fields.
    from chkn.job import read_flow_file
    from test_oracle_golden import build_step_job
    jd, steps, _ = build_step_job("ab_reacting", str(tmp_path))
    assert jd.max_step == steps
    gold = np.load(os.path.join(HERE, "golden", "step_ab_reacting.npz"))
    r = subprocess.run([CHKN_RUN + "-ab", "--job=" + str(tmp_path), "--binary"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0 and "Error" not in r.stderr
    last = max(int(line.split()[0]) for line in open(tmp_path / "times.data") if line[0] != "#")
    for b, blk in enumerate(jd.blocks):
        path = tmp_path / "flow" / ("t%04d" % last) / ("flow-%04d-%04d-%04d.bin" % (blk.i, blk.j, blk.k))
        io = read_flow_file(str(path), blk.ncells)
        assert np.array_equal(io.view(np.uint64), np.ascontiguousarray(gold["io%d" % b]).view(np.uint64))
