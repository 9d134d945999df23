"""Pins of the Yoshida push (JuMC/src/electron.jl:38-62) and of the staggered
field gather (MCStreamer/src/gridfields.jl:94-190, here in 3D Cartesian form)."""
import ctypes

import numpy as np
import pytest

from mcstreamer_amd import constants as co
from oracle import refcpu as R

QM = -co.elementary_charge / co.electron_mass


def _push(field, x, v, dt):
    x = np.array(x, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    fl = ctypes.c_int32(0)
    R.lib().ref_yoshida(ctypes.byref(field), x.ctypes.data_as(R._PD), v.ctypes.data_as(R._PD),
                        float(dt), ctypes.byref(fl))
    return x, v, fl.value


def test_yoshida_uniform_field_is_exact_parabola():
    rng = np.random.default_rng(3)
    E = np.array([1e5, -3e5, 2.45e6])
    f = R.uniform_field(E)
    a = QM * E
    for _ in range(200):
        x0 = rng.normal(size=3) * 1e-4
        v0 = rng.normal(size=3) * 1e6
        dt = 10 ** rng.uniform(-14, -11)
        x, v, _ = _push(f, x0, v0, dt)
        np.testing.assert_allclose(v, v0 + a * dt, rtol=1e-13, atol=1e-9 * np.abs(v0).max())
        np.testing.assert_allclose(x, x0 + v0 * dt + 0.5 * a * dt * dt, rtol=1e-13,
                                   atol=1e-13 * (np.abs(x0).max() + np.abs(v0).max() * dt))


def _mesh_linear_faces(n, h, g):
    """Face arrays holding a linear function of the face coordinates."""
    m = R.Mesh(n, h)
    nx, ny, nz = n

    def fill(arr, shape, off):
        sx, sy, sz = shape
        k, j, i = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        X = (i + off[0]) * h
        Y = (j + off[1]) * h
        Z = (k + off[2]) * h
        arr[:] = (g[0] * X + g[1] * Y + g[2] * Z + g[3]).reshape(-1)

    # EX at x faces (i h), cell-centred in y, z with one ghost: centre of ghost index j is (j-1/2) h
    fill(m.ex, (nx + 1, ny + 2, nz + 2), (0.0, -0.5, -0.5))
    fill(m.ey, (nx + 2, ny + 1, nz + 2), (-0.5, 0.0, -0.5))
    fill(m.ez, (nx + 2, ny + 2, nz + 1), (-0.5, -0.5, 0.0))
    return m


def test_staggered_gather_reproduces_linear_fields():
    n, h = (6, 5, 7), 1e-5
    g = (1e9, -2e9, 3e9, 5e4)
    m = _mesh_linear_faces(n, h, g)
    f = R.mesh_field(m)
    rng = np.random.default_rng(1)
    for _ in range(500):
        x = rng.uniform(0.01, 0.99, 3) * np.array(n) * h
        e = np.zeros(3)
        fl = ctypes.c_int32(0)
        R.lib().ref_efield(ctypes.byref(f), x.ctypes.data_as(R._PD), e.ctypes.data_as(R._PD),
                           ctypes.byref(fl))
        expect = g[0] * x[0] + g[1] * x[1] + g[2] * x[2] + g[3]
        np.testing.assert_allclose(e, [expect] * 3, rtol=1e-11)


def test_gather_zero_outside_and_field_flag():
    n, h = (4, 4, 4), 1e-5
    m = R.Mesh(n, h)
    m.ez[:] = 2e8
    f = R.mesh_field(m)
    for x, inside in (([-1e-7, 2e-5, 2e-5], False), ([2e-5, 2e-5, 4e-5], False), ([2e-5, 2e-5, 2e-5], True)):
        x = np.array(x)
        e = np.zeros(3)
        fl = ctypes.c_int32(0)
        R.lib().ref_efield(ctypes.byref(f), x.ctypes.data_as(R._PD), e.ctypes.data_as(R._PD),
                           ctypes.byref(fl))
        if inside:
            assert e[2] == pytest.approx(2e8) and fl.value == 2
        else:
            assert np.all(e == 0.0) and fl.value == 0


def test_faces_from_linear_potential_are_uniform():
    n, h = (6, 4, 8), 2e-6
    m = R.Mesh(n, h)
    nx, ny, nz = n
    k, j, i = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    a, b, c = 3e3, -1e3, 2e4
    m.u[:] = (a * (i - 0.5) * h + b * (j - 0.5) * h + c * (k - 0.5) * h).reshape(-1)
    R.efield_faces(m, -5e6)
    np.testing.assert_allclose(m.ex, -a, rtol=1e-9)
    np.testing.assert_allclose(m.ey, -b, rtol=1e-9)
    np.testing.assert_allclose(m.ez, -5e6 - c, rtol=1e-9)
