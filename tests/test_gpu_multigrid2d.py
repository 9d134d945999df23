"""The two-dimensional solver (Multigrid.jl with N = 2) on the GPU against the
CPU restatement: Cartesian and cylindrical connectors, Dirichlet/Neumann mixes,
one- and two-sweep smoothing, absolute and relative tolerances.  The cell
arithmetic is the restatement's, so with equal iteration counts the solutions
agree bitwise (only the residual norm's summation order differs) when the
coarse levels run as sweeps; with the tail tabulated (the default) the cycle
below its start level is one matrix-vector product, which rounds differently:
the solutions agree to 1e-12."""
import ctypes

import numpy as np
import pytest

from oracle import refcpu as R

pytestmark = pytest.mark.gpu

D, N = 0, 1


@pytest.mark.parametrize("tail", ["matrix", "1", "0"])
@pytest.mark.parametrize("n,bc,levels,smooth,cyl,tol", [
    ((32, 64), (N, D, D, D), 3, (2, 2), 1, -1e-10),
    ((64, 32), (N, N, D, N), 4, (2, 2), 0, -1e-10),
    ((48, 80), (D, N, N, D), 3, (1, 2), 2, -1e-9),
    ((256, 512), (N, D, D, D), 6, (2, 2), 1, -1e-10),
    ((128, 128), (N, D, D, D), 5, (2, 1), 1, 1e-6),
    ((96, 64), (N, D, D, D), 3, (2, 0), 1, -1e-8),
])
def test_solve_2d_matches_restatement(monkeypatch, n, bc, levels, smooth, cyl, tol, tail):
    """(tail "1": the coarse levels of every V-cycle in one workgroup; "0": one
    launch per phase on every level; "matrix": the tail tabulated.)"""
    from mcstreamer_amd import _lib
    from mcstreamer_amd import multigrid as mg
    monkeypatch.setenv("MCS_MG2_TAIL", "0" if tail == "0" else "1")
    monkeypatch.setenv("MCS_MG_TAIL_MATRIX", "1" if tail == "matrix" else "0")
    rng = np.random.default_rng(17 + sum(n))
    g = (n[0] + 2) * (n[1] + 2)
    b = rng.normal(size=g)
    u0 = rng.normal(size=g) * 0.1
    faces = [mg.BottomBnd, mg.TopBnd, mg.LeftBnd, mg.RightBnd]
    conn = mg.CylindricalConnector(cyl) if cyl else mg.CartesianConnector()
    conf = mg.MGConfig(bc=list(zip(faces, bc)), s=0.9, levels=levels, tolerance=tol, maxiter=50,
                       smooth1=smooth[0], smooth2=smooth[1], ndim=2, conn=conn)
    ws = mg.Workspace(conf, n)
    u_g = u0.copy()
    it = ctypes.c_int32()
    res = ctypes.c_double()
    assert _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u_g), _lib.dptr(b), ctypes.byref(it),
                                        ctypes.byref(res)) == 0
    u_r = u0.copy()
    rc, it_r, res_r = R.mg_solve(n, bc, 0.9, levels, smooth[0], smooth[1], tol, 50, u_r, b, cyl=cyl)
    assert rc == 0
    print(f"n={n} cyl={cyl} iters gpu {it.value} ref {it_r} resid {res.value:.3e} {res_r:.3e}", flush=True)
    assert it.value == it_r
    if tail == "matrix":
        np.testing.assert_allclose(u_g, u_r, rtol=0, atol=1e-12 * np.abs(u_r).max())
        # the residual's rms moves by at most |L| (8 with the connector's factors
        # below 2) times the largest change of u, plus its own rounding
        assert abs(res.value - res_r) <= 8.0 * np.abs(u_g - u_r).max() + 1e-10 * res_r
    else:
        np.testing.assert_array_equal(u_g, u_r)
        assert res.value == pytest.approx(res_r, rel=1e-10)


def test_wrapper_2d_axisymmetric_solution():
    """The host wrapper on (nr+2, nz+2) arrays: u = (1 - r^2) sin(pi z) to second
    order, as the restatement's pin of the same problem."""
    from mcstreamer_amd import multigrid as mg
    m = 64
    h = 1.0 / m
    c = (np.arange(m) + 0.5) * h
    Rr, Z = np.meshgrid(c, c, indexing="ij")   # array axis 0 = r, axis 1 = z
    ue = (1 - Rr ** 2) * np.sin(np.pi * Z)
    f = (4 + np.pi ** 2 * (1 - Rr ** 2)) * np.sin(np.pi * Z)
    b = np.zeros((m + 2, m + 2))
    b[1:-1, 1:-1] = f
    x = np.zeros_like(b)
    conf = mg.MGConfig(bc=[(mg.BottomBnd, mg.Neumann), (mg.TopBnd, mg.Dirichlet), (mg.LeftBnd, mg.Dirichlet),
                           (mg.RightBnd, mg.Dirichlet)], s=h * h, levels=4, tolerance=1e-12, ndim=2,
                       conn=mg.CylindricalConnector(1))
    ws = mg.allocate(conf, x)
    it, res = mg.solve(conf, x, b, ws)
    assert it >= 2
    assert np.abs(x[1:-1, 1:-1] - ue).max() < 2e-3


def test_graph_replays_match_one_launch_per_phase(monkeypatch):
    """The fmg correction replayed from a HIP graph (captured once per (u, b)),
    three solves in a row with host work and synchronisation between them, each
    bitwise equal to the same solve launched phase by phase."""
    from mcstreamer_amd import _lib
    from mcstreamer_amd import multigrid as mg
    n, levels = (96, 128), 4
    g = (n[0] + 2) * (n[1] + 2)
    rng = np.random.default_rng(5)
    bs = [rng.normal(size=g) * (1 + k) for k in range(3)]
    out = {}
    for graph in ("0", "1"):
        monkeypatch.setenv("MCS_MG_GRAPH", graph)
        conf = mg.MGConfig(bc=[(mg.BottomBnd, mg.Neumann), (mg.TopBnd, mg.Dirichlet), (mg.LeftBnd, mg.Dirichlet),
                               (mg.RightBnd, mg.Dirichlet)], s=0.7, levels=levels, tolerance=-1e-10, maxiter=50,
                           ndim=2, conn=mg.CylindricalConnector(1))
        ws = mg.Workspace(conf, n)
        res = []
        for b in bs:
            u = np.zeros(g)
            it = ctypes.c_int32()
            rr = ctypes.c_double()
            assert _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u), _lib.dptr(b), ctypes.byref(it),
                                                ctypes.byref(rr)) == 0
            res.append((it.value, u.copy()))
            ws.ctx.synchronize()
        out[graph] = res
    for (i0, u0), (i1, u1) in zip(out["0"], out["1"]):
        assert i0 == i1 and i0 >= 2
        np.testing.assert_array_equal(u0, u1)
