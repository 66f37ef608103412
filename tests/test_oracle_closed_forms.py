"""Closed-form and invariant checks that pin the oracle independently of the golden vectors.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign, a swapped
operand) fails it: the limiter identities, exact reproduction of linear data, flux
consistency with the Euler flux, pure upwinding for supersonic states, zero mass flux
for mirrored states, exact box geometry, encode/decode round trips and bad-cell flags.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob

G, R = 1.4, 287.1
CV = R / (G - 1.0)


def state(rho, T, v, YB=0.0):
    p = rho * R * T
    return np.array([rho, T * CV, p, T, YB, math.sqrt(G * R * T), v[0], v[1], v[2]])


def euler_flux(s, n):
    rho, e, p, T, YB, a = s[:6]
    v = s[6:9]
    un = float(np.dot(v, n))
    H = e + p / rho + 0.5 * float(np.dot(v, v))
    return np.array([rho * un, *(rho * un * v + p * n), rho * un * H, rho * un * YB])


def frame(theta=0.3, phi=1.1):
    n = np.array([math.cos(theta) * math.cos(phi), math.sin(theta) * math.cos(phi), math.sin(phi)])
    t1 = np.cross(n, [0.2, 0.7, -0.3])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    return n, np.concatenate([n, t1, t2])


@pytest.mark.parametrize("mach", [0.0, 0.3, 0.9, 1.5, 4.0, -0.6, -2.5])
def test_consistency_with_euler_flux(mach):
    n, fr = frame()
    a = math.sqrt(G * R * 300.0)
    s = state(1.2, 300.0, mach * a * n + np.array([3.0, -2.0, 1.0]))
    F = ob.face_flux(np.tile(s, 4), fr, 2)
    Fe = euler_flux(s, n)
    assert np.allclose(F, Fe, rtol=1e-12, atol=1e-9 * np.abs(Fe).max())


def test_supersonic_is_pure_upwind():
    n, fr = frame(0.7, -0.2)
    a = math.sqrt(G * R * 300.0)
    sL = state(0.8, 300.0, 3.0 * a * n + np.array([5.0, -7.0, 2.0]))
    sR = state(2.0, 300.0, 2.5 * a * n)
    F = ob.face_flux(np.concatenate([sL, sL, sR, sR]), fr, 1)
    assert np.allclose(F, euler_flux(sL, n), rtol=1e-12)
    sL2 = state(0.8, 300.0, -2.5 * a * n)
    sR2 = state(2.0, 300.0, -3.0 * a * n + np.array([1.0, 4.0, -3.0]))
    F = ob.face_flux(np.concatenate([sL2, sL2, sR2, sR2]), fr, 1)
    assert np.allclose(F, euler_flux(sR2, n), rtol=1e-12)


def test_mirror_states_have_zero_mass_flux():
    n, fr = frame(1.3, 0.4)
    v = np.array([120.0, -40.0, 75.0])
    s = state(1.1, 310.0, v)
    m = s.copy()
    m[6:9] = v - 2.0 * np.dot(v, n) * n
    F = ob.face_flux(np.concatenate([s, s, m, m]), fr, 2)
    assert abs(F[0]) <= 1e-12 * 1.1 * abs(np.dot(v, n))
    assert abs(F[5]) == 0.0


def test_linear_data_reconstructs_exactly():
    # rho = 1, 2, 3, 4 at constant T: every limiter is 1 and qL = qR = 2.5 exactly, so the
    # flux is the Euler flux of the face state (pressure only, since the gas is at rest).
    n, fr = frame(0.4, 0.9)
    sts = [state(r, 300.0, np.zeros(3)) for r in (1.0, 2.0, 3.0, 4.0)]
    F = ob.face_flux(np.concatenate(sts), fr, 2)
    face = state(2.5, 300.0, np.zeros(3))
    assert F[0] == 0.0 and F[4] == 0.0
    assert np.allclose(F, euler_flux(face, n), rtol=1e-13, atol=1e-9)


def test_geometry_unit_box_and_affine():
    p = np.array([[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0], [0, 0, 5], [2, 0, 5], [2, 3, 5],
                  [0, 3, 5]], dtype=float)
    out = ob.hex_props(p)
    assert np.allclose(out[:3], [1.0, 1.5, 2.5])
    assert out[3] == pytest.approx(30.0, rel=1e-15)
    assert np.allclose(out[4:], [2.0, 3.0, 5.0])
    q = ob.quad_props(p[[0, 1, 2, 3]])  # bottom face, counter-clockwise seen from +z
    assert q[12] == pytest.approx(6.0, rel=1e-15)
    assert np.allclose(q[3:6], [0, 0, 1]) and np.allclose(q[6:9], [1, 0, 0]) and np.allclose(q[9:12], [0, 1, 0])
    A = np.array([[1.0, 0.3, 0.0], [0.0, 1.0, 0.2], [0.1, 0.0, 1.0]])
    pa = p @ A.T + np.array([3.0, -1.0, 2.0])
    assert ob.hex_props(pa)[3] == pytest.approx(30.0 * np.linalg.det(A), rel=1e-13)


def test_encode_decode_roundtrip_and_bad_cells():
    s = state(0.7, 410.0, np.array([300.0, -20.0, 5.0]), YB=0.25)
    U = ob.encode(s)
    bad, w = ob.decode(U)
    assert bad == 0
    assert np.allclose(w, s, rtol=1e-14)
    for broken in ([0.0, 1, 1, 1, 1, 0], [-1e-3, 1, 1, 1, 1, 0], [1.0, np.nan, 1, 1, 1, 0]):
        assert ob.decode(np.array(broken, dtype=float))[0] == 1
