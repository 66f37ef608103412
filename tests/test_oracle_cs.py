"""Pins the complex-step restatement of the face flux (oracle/chkn_oracle_cs.c), which the
adjoint kernels are checked against: its real part is the real oracle flux, and its
directional derivatives agree with central differences of the real oracle flux."""
import numpy as np

import oracle_bind as ob


def _case(rng, speed):
    prims = np.empty((4, 6))
    for c in range(4):
        prims[c] = [0.5 * (1 + 0.2 * rng.uniform(-1, 1)), 2.0e5 * (1 + 0.2 * rng.uniform(-1, 1)),
                    0.0, speed * rng.uniform(-1, 1), speed * rng.uniform(-1, 1),
                    speed * rng.uniform(-1, 1)]
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t1 = rng.normal(size=3)
    t1 -= t1.dot(n) * n
    t1 /= np.linalg.norm(t1)
    frame = np.concatenate([n, t1, np.cross(n, t1)])
    return prims, frame


def _states9(prims):
    out = []
    for rho, e, YB, vx, vy, vz in prims:
        T = (e + YB * 0.0) / (287.1 / (1.4 - 1.0))
        out.append([rho, e, rho * 287.1 * T, T, YB, np.sqrt(1.4 * 287.1 * T), vx, vy, vz])
    return np.array(out)


def test_real_part_is_the_oracle_flux():
    rng = np.random.default_rng(0)
    for trial in range(50):
        prims, frame = _case(rng, [100.0, 400.0, 1500.0][trial % 3])
        d = rng.normal(size=24)
        for xo in (1, 2):
            Fre, _ = ob.face_flux_cs(prims, frame, d, x_order=xo)
            F = ob.face_flux(_states9(prims), frame, x_order=xo)
            # Equal bit for bit except O(h^2) cross terms where the real part is exactly 0.
            assert np.all((Fre == F) | (np.abs(Fre - F) < 1e-50)), (trial, xo, Fre - F)


def test_complex_step_matches_central_differences():
    rng = np.random.default_rng(1)
    worst = 0.0
    for trial in range(60):
        prims, frame = _case(rng, [100.0, 400.0, 1500.0][trial % 3])
        scale = np.abs(prims).ravel() + 1.0
        d = rng.normal(size=24) * scale
        d[2::6] = 0.0  # YB = 0 sits on the limiter's kink at zero differences
        _, jv = ob.face_flux_cs(prims, frame, d)
        h = 1.0e-7
        Fp = ob.face_flux(_states9(prims + h * d.reshape(4, 6)), frame)
        Fm = ob.face_flux(_states9(prims - h * d.reshape(4, 6)), frame)
        fd = (Fp - Fm) / (2 * h)
        ref = np.max(np.abs(jv)) + 1e-300
        worst = max(worst, float(np.max(np.abs(fd - jv)) / ref))
    assert worst < 1e-5, worst


def test_sbp_asf_complex_step():
    # The SBP-ASF flux (x_order 0 in the complex-step entry point): its real part is the real
    # oracle's flux_calc 1 flux, and its directional derivatives agree with central
    # differences of it.
    rng = np.random.default_rng(2)
    worst = 0.0
    for trial in range(40):
        prims, frame = _case(rng, [60.0, 250.0, 900.0][trial % 3])
        scale = np.abs(prims).ravel() + 1.0
        d = rng.normal(size=24) * scale
        Fre, jv = ob.face_flux_cs(prims, frame, d, x_order=0)
        F = ob.face_flux(_states9(prims), frame, flux_calc=1)
        assert np.all((Fre == F) | (np.abs(Fre - F) < 1e-50)), (trial, Fre - F)
        h = 1.0e-7
        Fp = ob.face_flux(_states9(prims + h * d.reshape(4, 6)), frame, flux_calc=1)
        Fm = ob.face_flux(_states9(prims - h * d.reshape(4, 6)), frame, flux_calc=1)
        fd = (Fp - Fm) / (2 * h)
        ref = np.max(np.abs(jv)) + 1e-300
        worst = max(worst, float(np.max(np.abs(fd - jv)) / ref))
    assert worst < 1e-6, worst
