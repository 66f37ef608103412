"""The launch-shape autotuner: it times real candidates, returns the fastest, caches it per
problem shape, and the option set it picks gives the same fields as the default launch."""
import json

import numpy as np
import pytest

from chkn import cases, tune
from chkn.lib import Sim

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_timed_candidate_and_caches(tmp_path):
    job = cases.ramp_array(seed=5, lo=6, hi=14, max_step=12)
    cands = tune.candidates(tiles=[(32, 8), (16, 4)], tile_k=[0, 4], orders=[0, 1],
                            persistent=[0, 1])
    cache = str(tmp_path / "tune.json")
    best, table = tune.autotune(job, steps=4, warmup=1, cands=cands, cache_path=cache)
    assert len(table) == len(cands)
    assert all(ms > 0 for ms, _ in table)
    assert best in [o for _, o in table[:8]]  # re-timed among the sweep's eight fastest
    assert [ms for ms, _ in table] == sorted(ms for ms, _ in table)
    with open(cache) as f:
        assert tune.shape_key([job]) in json.load(f)
    best2, _ = tune.autotune(job, cands=cands[:1], cache_path=cache)  # served from the cache
    assert best2 == best

    a, b = Sim(job), Sim(job)
    b.set_launch(**best)
    a.step(12)
    b.step(12)
    for blk in range(len(job.blocks)):
        assert np.array_equal(a.iovars(blk), b.iovars(blk))
    a.close()
    b.close()
