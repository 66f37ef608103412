"""The allocation-free face-flux VJP entry point (chkn_flux_vjp_points_noalloc) against the
allocating one (chkn_flux_vjp_points): identical outputs on two calls in a row, the second
smaller than the first (the shared device array keeps the first call's contents), and the
size limit refused with an error."""
import numpy as np
import pytest

from chkn import lib

pytestmark = pytest.mark.gpu


def _points(n, seed):
    rng = np.random.default_rng(seed)
    prims = np.empty((n, 4, 6))
    prims[:, :, 0] = rng.uniform(0.2, 2.0, (n, 4))          # rho
    prims[:, :, 1] = rng.uniform(1.5e5, 4.0e5, (n, 4))      # e
    prims[:, :, 2] = 0.0                                    # YB
    prims[:, :, 3:] = rng.uniform(-600.0, 900.0, (n, 4, 3))  # velocity
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    t1 = np.cross(nrm, rng.normal(size=(n, 3)))
    t1 /= np.linalg.norm(t1, axis=1)[:, None]
    t2 = np.cross(nrm, t1)
    frames = np.concatenate([nrm, t1, t2], axis=1)
    fbar = rng.normal(size=(n, 6))
    return prims, frames, fbar


@pytest.mark.parametrize("x_order", [2, 1])
def test_noalloc_matches_allocating_entry_point(x_order):
    for n, seed in ((4096, 1), (97, 2)):  # the second call smaller than the first
        prims, frames, fbar = _points(n, seed)
        ref = lib.flux_vjp_points(prims, frames, fbar, x_order=x_order)
        got = lib.flux_vjp_points(prims, frames, fbar, x_order=x_order, noalloc=True)
        assert np.array_equal(got.view(np.int64), ref.view(np.int64)), n
        assert np.all(np.isfinite(got)) and np.any(got != 0.0)


def test_noalloc_refuses_more_points_than_its_array():
    prims, frames, fbar = _points(4097, 3)
    with pytest.raises(lib.ChknError, match="allocation-free limit of 4096"):
        lib.flux_vjp_points(prims, frames, fbar, noalloc=True)
