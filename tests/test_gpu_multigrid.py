"""GPU parity of Multigrid.solve (Multigrid/src/Multigrid.jl:449-467) against the
oracle restatement: same iteration count; the solution equal to 1e-12 relative.
With the tails run as sweeps (MCS_MG_TAIL_MATRIX=0) the only reordered sum is the
residual norm of the convergence test; with the tabulated tails (the default)
the cycle below the tail's start level is one matrix-vector product, which
rounds differently from the sweeps in the last bits."""
import numpy as np
import pytest

from mcstreamer_amd import multigrid as mg
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

D, N = mg.Dirichlet, mg.Neumann


def _cfg(bc, s, levels, tol, maxiter=50):
    faces = [mg.BottomBnd, mg.TopBnd, mg.LeftBnd, mg.RightBnd, mg.FrontBnd, mg.BackBnd]
    return mg.MGConfig(bc=list(zip(faces, bc)), s=s, levels=levels, tolerance=tol, maxiter=maxiter)


@pytest.mark.parametrize("tailmat", ["0", "1"])
@pytest.mark.parametrize("n,bc,levels,tol", [
    ((8, 8, 8), (D,) * 6, 2, 1e-12),
    ((16, 8, 32), (N, N, N, N, D, D), 2, 1e-11),
    ((32, 32, 64), (N, N, N, N, D, D), 4, -1e-8),
    ((64, 64, 128), (N, N, N, N, D, D), 5, -1e-8),
    ((16, 16, 16), (N, D, D, N, D, N), 3, 1e-10),
])
def test_solve_parity(monkeypatch, n, bc, levels, tol, tailmat):
    monkeypatch.setenv("MCS_MG_TAIL_MATRIX", tailmat)
    rng = np.random.default_rng(sum(n))
    g = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    b = rng.normal(size=g)
    u0 = rng.normal(size=g) * 0.1
    s = 0.9
    conf = _cfg(bc, s, levels, tol)
    ws = mg.Workspace(conf, n)
    from mcstreamer_amd import _lib
    import ctypes
    u_g = u0.copy()
    it = ctypes.c_int32()
    res = ctypes.c_double()
    rc = _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u_g), _lib.dptr(b), ctypes.byref(it), ctypes.byref(res))
    assert rc == 0
    u_r = u0.copy()
    rc_r, it_r, res_r = R.mg_solve(n, bc, s, levels, 2, 2, tol, 50, u_r, b)
    assert rc_r == 0
    print(f"n={n} iters gpu {it.value} ref {it_r} resid {res.value:.3e} {res_r:.3e}", flush=True)
    assert it.value == it_r
    np.testing.assert_allclose(u_g, u_r, rtol=1e-12, atol=1e-12 * np.abs(u_r).max())
    if tailmat == "0":
        assert res.value == pytest.approx(res_r, rel=1e-10)
    else:
        # the converged residual is a difference of nearly equal terms: its rms
        # moves by at most |L| = 12 times the largest change of u (plus the
        # rounding of its own sum)
        assert abs(res.value - res_r) <= 12.0 * np.abs(u_g - u_r).max() + 1e-10 * res_r


@pytest.mark.parametrize("n,bc,levels,smooth,prec", [
    ((8, 8, 8), (D,) * 6, 2, (2, 2), "f64"),
    ((32, 32, 64), (N, N, N, N, D, D), 4, (2, 2), "f64"),
    ((128, 128, 256), (N, N, N, N, D, D), 6, (2, 2), "f64"),
    ((48, 40, 56), (D, N, D, N, N, D), 3, (1, 2), "f64"),
    ((64, 64, 128), (N, N, N, N, D, D), 5, (2, 1), "f64"),
    ((64, 64, 128), (N, D, N, N, D, D), 5, (2, 2), "f32"),
])
def test_tabulated_tail_matches_sweeps(monkeypatch, n, bc, levels, smooth, prec):
    """The tabulated tails (mgv! below the tail's start level from x = 0, and
    fmg!'s coarse part, each a matrix over the start level's right-hand side,
    tabulated by running the tail on unit right-hand sides) give the solution of
    the tails run as sweeps: same iteration count, equal to 1e-13 relative in
    FP64 (a few units in the last place of the coarse corrections) and to the
    FP32 correction's accuracy in precision mode; and not bitwise, which shows
    that the matrix path ran."""
    from mcstreamer_amd import _lib
    import ctypes
    rng = np.random.default_rng(17 + sum(n))
    g = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    b = rng.normal(size=g)
    u0 = rng.normal(size=g) * 0.1
    out = {}
    for tm in ("0", "1"):
        monkeypatch.setenv("MCS_MG_TAIL_MATRIX", tm)
        faces = [mg.BottomBnd, mg.TopBnd, mg.LeftBnd, mg.RightBnd, mg.FrontBnd, mg.BackBnd]
        conf = mg.MGConfig(bc=list(zip(faces, bc)), s=0.9, levels=levels, tolerance=-1e-10, maxiter=50,
                           smooth1=smooth[0], smooth2=smooth[1], precision=prec)
        ws = mg.Workspace(conf, n)
        u = u0.copy()
        it = ctypes.c_int32()
        res = ctypes.c_double()
        assert _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u), _lib.dptr(b), ctypes.byref(it),
                                            ctypes.byref(res)) == 0
        out[tm] = (u, it.value, res.value)
        print(f"tailmat={tm} n={n} iters {it.value} resid {res.value:.3e}", flush=True)
    assert out["0"][1] == out["1"][1]
    scale = np.abs(out["0"][0]).max()
    rtol = 1e-13 if prec == "f64" else 1e-7
    np.testing.assert_allclose(out["1"][0], out["0"][0], rtol=0, atol=rtol * scale)
    assert not np.array_equal(out["0"][0], out["1"][0])


def test_solve_nonconverged_reports():
    n = (8, 8, 8)
    conf = _cfg((D,) * 6, 1.0, 2, 1e-300, maxiter=1)
    ws = mg.Workspace(conf, n)
    g = 10 * 10 * 10
    b = np.random.default_rng(1).normal(size=g)
    u = np.zeros((n[0] + 2, n[1] + 2, n[2] + 2))
    bb = b.reshape(n[2] + 2, n[1] + 2, n[0] + 2).transpose(2, 1, 0)
    it, res = mg.solve(conf, u, bb, ws)
    assert it == 2  # maxiter + 1 attempts, as the reference's while loop (Multigrid.jl:452-455)


def test_graph_replay_matches_eager(monkeypatch):
    """The cycle body is replayed from a HIP graph; every replay (several solves
    with host work in between) must equal the eagerly launched path bitwise."""
    from mcstreamer_amd import _lib
    import ctypes
    n = (32, 32, 64)
    g = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    conf = _cfg((N, N, N, N, D, D), 0.9, 4, -1e-9)
    outs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("MCS_MG_GRAPH", mode)
        ws = mg.Workspace(conf, n)
        res = []
        for rep in range(3):
            rng = np.random.default_rng(rep)
            b = rng.normal(size=g)
            u = np.zeros(g)
            it = ctypes.c_int32()
            r = ctypes.c_double()
            assert _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u), _lib.dptr(b), ctypes.byref(it), ctypes.byref(r)) == 0
            res.append((u.copy(), it.value))
        outs[mode] = res
    for (ue, ie), (ug, ig) in zip(outs["0"], outs["1"]):
        assert ie == ig
        np.testing.assert_array_equal(ue, ug)


@pytest.mark.parametrize("n,bc,levels", [
    ((32, 32, 64), (N, N, N, N, D, D), 4),
    ((64, 64, 128), (N, N, N, N, D, D), 5),
    ((16, 16, 16), (N, D, D, N, D, N), 3),
])
def test_fp32_smoother_in_fp64_refinement(n, bc, levels):
    """Precision mode: FP32 correction cycles inside FP64 refinement reach the same
    1e-10 relative residual as the FP64 solver, and the same solution to the
    accuracy that residual allows (the error of L u = -s b is bounded by the
    residual over the smallest eigenvalue; compared here at 1e-6 relative)."""
    from mcstreamer_amd import _lib
    import ctypes
    rng = np.random.default_rng(7 + sum(n))
    g = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    b = rng.normal(size=g)
    s = 0.9
    out = {}
    for prec in ("f64", "f32"):
        faces = [mg.BottomBnd, mg.TopBnd, mg.LeftBnd, mg.RightBnd, mg.FrontBnd, mg.BackBnd]
        conf = mg.MGConfig(bc=list(zip(faces, bc)), s=s, levels=levels, tolerance=-1e-10, maxiter=50,
                           precision=prec)
        ws = mg.Workspace(conf, n)
        u = np.zeros(g)
        it = ctypes.c_int32()
        res = ctypes.c_double()
        rc = _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u), _lib.dptr(b), ctypes.byref(it),
                                          ctypes.byref(res))
        assert rc == 0
        out[prec] = (u.copy(), it.value, res.value)
        print(f"{prec} n={n} iters {it.value} resid {res.value:.3e}", flush=True)
    brms = np.sqrt((b.reshape(n[2] + 2, n[1] + 2, n[0] + 2)[1:-1, 1:-1, 1:-1] ** 2).mean())
    for prec in ("f64", "f32"):
        assert out[prec][2] < 1e-10 * brms * s
    # the residual of the FP32-mode solution, recomputed independently with numpy
    u32 = out["f32"][0].reshape(n[2] + 2, n[1] + 2, n[0] + 2)
    bb = b.reshape(n[2] + 2, n[1] + 2, n[0] + 2)
    lap = (u32[1:-1, 1:-1, 2:] + u32[1:-1, 1:-1, :-2] + u32[1:-1, 2:, 1:-1] + u32[1:-1, :-2, 1:-1]
           + u32[2:, 1:-1, 1:-1] + u32[:-2, 1:-1, 1:-1] - 6 * u32[1:-1, 1:-1, 1:-1])
    r = s * bb[1:-1, 1:-1, 1:-1] + lap
    assert np.sqrt((r ** 2).mean()) < 1.01e-10 * brms * s
    u64 = out["f64"][0]
    assert not np.array_equal(out["f32"][0], u64)   # the FP32 correction path ran
    scale = np.abs(u64).max()
    assert np.abs(out["f32"][0] - u64).max() < 1e-6 * scale


@pytest.mark.parametrize("n,bc,levels,smooth,prec,knobs", [
    ((32, 32, 64), (N, N, N, N, D, D), 4, (2, 2), "f64", {}),
    ((48, 40, 56), (D, N, D, N, N, D), 3, (1, 2), "f64", {}),
    ((64, 64, 128), (N, N, N, N, D, D), 5, (2, 1), "f64", {}),
    ((32, 32, 64), (N, N, N, N, D, D), 4, (2, 2), "f32", {}),
    ((32, 32, 64), (N, D, N, N, D, D), 4, (2, 2), "f64", {"MCS_MG_SM_BIG": "0", "MCS_MG_SM_TILE": "8"}),
    ((48, 40, 56), (D, N, D, N, N, D), 3, (2, 2), "f64", {"MCS_MG_SM_BIG": "0"}),
    # more tiles than run at once: later tiles must not see what earlier ones wrote
    ((64, 64, 128), (N, N, N, N, D, D), 5, (2, 2), "f64", {"MCS_MG_SM_TILE": "8"}),
    ((128, 128, 256), (N, N, N, N, D, D), 6, (2, 2), "f64", {"MCS_MG_TAIL_LEVELS": "13"}),
    ((128, 128, 256), (N, N, N, N, D, D), 6, (2, 2), "f64",
     {"MCS_MG_TAIL_LEVELS": "3", "MCS_MG_SM_MAXCELLS": "100000000", "MCS_MG_SM_TILE": "16"}),
])
def test_tiled_smoother_matches_per_pass(monkeypatch, n, bc, levels, smooth, prec, knobs):
    """The tiled smoother (a level's pre-smoothing with the restricted residual,
    and the prolongation with post-smoothing, one or two sweeps per launch, halos
    recomputed in LDS) equals one launch per red-black pass bitwise, on every
    level but the coarsest (tail capped to one level unless stated), for tile sizes
    that are not powers of two, for one- and two-sweep smoothing, and on grids of
    more tiles than the device runs at once."""
    from mcstreamer_amd import _lib
    import ctypes
    rng = np.random.default_rng(3 + sum(n))
    g = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    b = rng.normal(size=g)
    u0 = rng.normal(size=g) * 0.1
    faces = [mg.BottomBnd, mg.TopBnd, mg.LeftBnd, mg.RightBnd, mg.FrontBnd, mg.BackBnd]
    monkeypatch.setenv("MCS_MG_TAIL_LEVELS", "1")
    monkeypatch.delenv("MCS_MG_SM_TILE", raising=False)
    monkeypatch.delenv("MCS_MG_SM_BIG", raising=False)
    monkeypatch.delenv("MCS_MG_SM_MAXCELLS", raising=False)
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    out = {}
    for fused in ("0", "1"):
        monkeypatch.setenv("MCS_MG_FUSED", fused)
        conf = mg.MGConfig(bc=list(zip(faces, bc)), s=0.9, levels=levels, tolerance=-1e-9, maxiter=50,
                           smooth1=smooth[0], smooth2=smooth[1], precision=prec)
        ws = mg.Workspace(conf, n)
        u = u0.copy()
        it = ctypes.c_int32()
        res = ctypes.c_double()
        rc = _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u), _lib.dptr(b), ctypes.byref(it),
                                          ctypes.byref(res))
        assert rc == 0
        out[fused] = (u, it.value)
        print(f"fused={fused} n={n} iters {it.value} resid {res.value:.3e}", flush=True)
    assert out["0"][1] == out["1"][1]
    np.testing.assert_array_equal(out["0"][0], out["1"][0])


@pytest.mark.parametrize("n,levels,prec", [((64, 64, 128), 5, "f64"), ((128, 128, 256), 6, "f64"),
                                          ((64, 64, 128), 5, "f32")])
def test_split_level0_matches_ghost_layout(monkeypatch, n, levels, prec):
    """The level-0 correction in the colour-split layout (residual, restriction,
    prolongation, passes and x += z indexing it) gives bitwise the solution of the
    ghost layout."""
    from mcstreamer_amd import _lib
    import ctypes
    rng = np.random.default_rng(5 + sum(n))
    g = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    b = rng.normal(size=g)
    u0 = rng.normal(size=g) * 0.1
    faces = [mg.BottomBnd, mg.TopBnd, mg.LeftBnd, mg.RightBnd, mg.FrontBnd, mg.BackBnd]
    out = {}
    for split in ("0", "1"):
        monkeypatch.setenv("MCS_MG_SPLIT0", split)
        conf = mg.MGConfig(bc=list(zip(faces, (N, D, N, N, D, D))), s=0.9, levels=levels, tolerance=-1e-9,
                           maxiter=50, precision=prec)
        ws = mg.Workspace(conf, n)
        u = u0.copy()
        it = ctypes.c_int32()
        res = ctypes.c_double()
        assert _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u), _lib.dptr(b), ctypes.byref(it),
                                            ctypes.byref(res)) == 0
        out[split] = (u, it.value, res.value)
        print(f"split={split} n={n} iters {it.value} resid {res.value:.3e}", flush=True)
    assert out["0"][1] == out["1"][1] and out["0"][2] == out["1"][2]
    np.testing.assert_array_equal(out["0"][0], out["1"][0])


@pytest.mark.parametrize("n,bc,levels,smooth,prec,knobs", [
    ((128, 128, 256), (N, N, N, N, D, D), 6, (2, 2), "f64", {}),
    ((64, 64, 128), (N, D, N, N, D, D), 5, (2, 2), "f32", {}),
    ((48, 40, 56), (D, N, D, N, N, D), 3, (1, 2), "f64", {}),
    ((64, 64, 128), (N, N, N, N, D, D), 5, (2, 1), "f64", {"MCS_MG_TAIL_MAXCELLS": "8192"}),
    ((32, 32, 64), (N, N, N, N, D, D), 4, (2, 2), "f64", {"MCS_MG_TAIL_MAXCELLS": "100000000"}),
])
def test_fmg_tail_matches_grid_path(monkeypatch, n, bc, levels, smooth, prec, knobs):
    """fmg!'s restrictions, coarse solve and coarse-to-fine V-cycles of the tail
    levels in one launch give bitwise the solution of one launch per step."""
    from mcstreamer_amd import _lib
    import ctypes
    rng = np.random.default_rng(11 + sum(n))
    g = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    b = rng.normal(size=g)
    u0 = rng.normal(size=g) * 0.1
    faces = [mg.BottomBnd, mg.TopBnd, mg.LeftBnd, mg.RightBnd, mg.FrontBnd, mg.BackBnd]
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    # the sweep tails on both sides (the tabulated one is compared separately)
    monkeypatch.setenv("MCS_MG_TAIL_MATRIX", "0")
    out = {}
    for ft in ("0", "1"):
        monkeypatch.setenv("MCS_MG_FMG_TAIL", ft)
        conf = mg.MGConfig(bc=list(zip(faces, bc)), s=0.9, levels=levels, tolerance=-1e-9, maxiter=50,
                           smooth1=smooth[0], smooth2=smooth[1], precision=prec)
        ws = mg.Workspace(conf, n)
        u = u0.copy()
        it = ctypes.c_int32()
        res = ctypes.c_double()
        assert _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u), _lib.dptr(b), ctypes.byref(it),
                                            ctypes.byref(res)) == 0
        out[ft] = (u, it.value, res.value)
        print(f"fmg_tail={ft} n={n} iters {it.value} resid {res.value:.3e}", flush=True)
    assert out["0"][1] == out["1"][1] and out["0"][2] == out["1"][2]
    np.testing.assert_array_equal(out["0"][0], out["1"][0])


@pytest.mark.parametrize("n,levels,prec", [((128, 128, 256), 6, "f64"), ((64, 64, 128), 5, "f32")])
def test_split_passes_two_cells_per_thread(monkeypatch, n, levels, prec):
    """The colour-split level-0 passes with two cells of a row per thread (pairs
    loaded and stored together) give bitwise the solution of one cell per
    thread."""
    from mcstreamer_amd import _lib
    import ctypes
    rng = np.random.default_rng(23 + sum(n))
    g = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
    b = rng.normal(size=g)
    u0 = rng.normal(size=g) * 0.1
    faces = [mg.BottomBnd, mg.TopBnd, mg.LeftBnd, mg.RightBnd, mg.FrontBnd, mg.BackBnd]
    out = {}
    for single in ("1", "0"):
        monkeypatch.setenv("MCS_MG_RBGS_SINGLE", single)
        conf = mg.MGConfig(bc=list(zip(faces, (D, N, N, D, D, D))), s=0.9, levels=levels, tolerance=-1e-9,
                           maxiter=50, precision=prec)
        ws = mg.Workspace(conf, n)
        u = u0.copy()
        it = ctypes.c_int32()
        res = ctypes.c_double()
        assert _lib.lib().mcs_mg_solve_host(ws.handle, _lib.dptr(u), _lib.dptr(b), ctypes.byref(it),
                                            ctypes.byref(res)) == 0
        out[single] = (u, it.value, res.value)
        print(f"single={single} n={n} iters {it.value} resid {res.value:.3e}", flush=True)
    assert out["0"][1] == out["1"][1] and out["0"][2] == out["1"][2]
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
