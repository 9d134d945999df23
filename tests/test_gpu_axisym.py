"""The axisymmetric (r, z) grid on the GPU against the oracle: cylindrical
density! (bitwise: per cell, terms in particle order), poisson!'s charge, solve
and face fields (bitwise), the particle step in the axisymmetric FieldInterp
with the ion tracker (particles and the tracker's array bitwise), and a few
Poisson + advance + repack steps."""
import ctypes

import numpy as np
import pytest

from mcstreamer_amd import axisym, jumc
from mcstreamer_amd import constants as co
from oracle import refcpu as R
from tests.common import gas_table, ref_table

pytestmark = pytest.mark.gpu

RR, LL, MM, NN = 0.64e-3, 1.28e-3, 64, 128


def _states(n, seed, grow=1.05):
    """Electrons over a cylinder a little larger than the grid (so some fall in
    the z ghost rows or outside), some inactive, 1 eV Maxwellian."""
    rng = np.random.default_rng(seed)
    r = RR * grow * np.sqrt(rng.uniform(size=n))
    th = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-0.03 * LL, LL * 1.03, n)
    x = np.stack([r * np.cos(th), r * np.sin(th), z], axis=1)
    v = rng.normal(size=(n, 3)) * np.sqrt(co.eV / co.electron_mass)
    w = rng.integers(1, 4, n).astype(np.float64)
    s = -np.log1p(-rng.uniform(size=n))
    act = (rng.uniform(size=n) > 0.05).astype(np.uint8)
    return x, v, w, s, act


def _setup(n=20000, seed=1, o2=0.2):
    proc, energy, rate, maxrate = gas_table(o2=o2)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    T = ref_table(proc, energy, rate, maxrate)
    x, v, w, s, act = _states(n, seed)
    popl = jumc.Population(4 * n, jumc.ParticleStates(x, v, w, s, act), table, seed=11)
    P = R.Pop(x, v, w, s, act, cap=4 * n)
    grid = axisym.Grid(RR, LL, MM, NN)
    fields = axisym.GridFields(grid)
    G = R.CGrid(RR, LL, MM, NN)
    return table, T, popl, P, grid, fields, G


def _shape(G, name):
    return (G.N + 3, G.M + 3) if name in ("er", "ez") else (G.N + 2, G.M + 2)


def _smooth_potential(G, amp=40.0, seed=3):
    """A potential with fields of a few MV/m: a Gaussian bump plus random noise."""
    rng = np.random.default_rng(seed)
    dr, dz = G.steps[0], G.steps[1]
    I = (np.arange(G.M + 2)[None, :] - 0.5) * dr
    J = (np.arange(G.N + 2)[:, None] - 0.5) * dz
    u = amp * np.exp(-(I / (0.3 * RR)) ** 2 - ((J - 0.6 * LL) / (0.2 * LL)) ** 2)
    return u + 0.01 * amp * rng.normal(size=u.shape)


def test_geometry_matches_oracle():
    f = axisym.GridFields(axisym.Grid(RR, LL, MM, NN))
    np.testing.assert_array_equal(f.steps(), R.CGrid(RR, LL, MM, NN).steps)
    g = axisym.Grid(3e-3, 1e-2, 300, 1000)
    np.testing.assert_array_equal(axisym.GridFields(g).steps(), R.CGrid(3e-3, 1e-2, 300, 1000).steps)


def test_density_matches_oracle_bitwise():
    table, T, popl, P, grid, fields, G = _setup(n=50000)
    rng = np.random.default_rng(5)
    q0 = rng.normal(size=_shape(G, "qfixed")) * 1e15
    fields.set("qfixed", q0)
    G.qfixed[:] = q0.ravel()
    axisym.density(fields, "qfixed", popl, 1.0)
    R.cyl_density(P, G, "qfixed", 1.0)
    axisym.density(fields, "qpart", popl, -1.0)
    R.cyl_density(P, G, "qpart", -1.0)
    axisym.density(fields, "qpart", popl, 0.5)      # accumulates into what is there
    R.cyl_density(P, G, "qpart", 0.5)
    for k in ("qfixed", "qpart"):
        np.testing.assert_array_equal(fields.get(k).ravel(), getattr(G, k), err_msg=k)
    # the z ghost rows received the particles just outside in z, as the reference's arrays do
    qp = fields.get("qpart")
    assert np.abs(qp[0]).sum() > 0 and np.abs(qp[-1]).sum() > 0


def test_density_long_runs_bitwise():
    """Cells holding thousands of particles (summed by a wavefront each, in chunks
    of 64 with a partial last chunk) next to short runs."""
    proc, energy, rate, maxrate = gas_table()
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    rng = np.random.default_rng(9)
    n = 60000
    x = rng.normal(size=(n, 3)) * 2e-5
    x[:, 2] += 0.5 * LL
    x[:7000] = rng.uniform(0.0, RR * 0.7, size=(7000, 3))       # spread: short runs
    v = np.zeros((n, 3))
    w = rng.uniform(0.5, 2.0, n)
    s = np.ones(n)
    act = np.ones(n, dtype=np.uint8)
    popl = jumc.Population(n, jumc.ParticleStates(x, v, w, s, act), table)
    P = R.Pop(x, v, w, s, act, cap=n)
    fields = axisym.GridFields(axisym.Grid(RR, LL, MM, NN))
    G = R.CGrid(RR, LL, MM, NN)
    axisym.density(fields, "qpart", popl, -1.0)
    R.cyl_density(P, G, "qpart", -1.0)
    got = fields.get("qpart").ravel()
    np.testing.assert_array_equal(got, G.qpart)
    assert (G.qpart != 0).sum() > 100
    # the densest cell holds far more than one wavefront's chunk
    drf, dzf = G.steps[2], G.steps[3]
    i = np.floor(np.hypot(x[:, 0], x[:, 1]) / drf).astype(int)
    j = np.floor(x[:, 2] / dzf).astype(int)
    _, c = np.unique(i * 100000 + j, return_counts=True)
    assert c.max() > 2000


def test_cyl_resample_matches_oracle_bitwise():
    """Canonical (cell, id) order and resample! per (r, z) cell, twice (merged cells,
    splits, the outside bucket, inactive particles dropped), with the p and
    wtotal each resample leaves on the grid."""
    table, T, popl, P, grid, fields, G = _setup(n=60000, seed=8)
    x, v, w, s, act = _states(60000, 8)
    rng = np.random.default_rng(12)
    # a dense knot of light particles (merged) among heavy ones (split where thin)
    x[:20000] = rng.normal(size=(20000, 3)) * 1.5e-5 + np.array([1e-4, 0.0, 0.5 * LL])
    w[20000:] = rng.integers(1, 9, 40000)
    popl.upload(jumc.ParticleStates(x, v, w, s, act))
    P = R.Pop(x, v, w, s, act, cap=4 * 60000)
    merged = splits = 0
    pz, wz = fields.resample_fields()
    assert not pz.any() and not wz.any()     # zeros before the first resample
    for step in range(2):
        pc, wc = R.cyl_resample_counts(P, G, 10, 15)
        st = axisym.resample(popl, fields, 10, 15, step=step)
        rc, rs = R.cyl_resample(P, G, 10, 15, seed=popl.seed, step=step)
        # GridFields.p and .wtotal as this resample! left them
        pg, wg = fields.resample_fields()
        np.testing.assert_array_equal(pg, pc)
        np.testing.assert_array_equal(wg, wc)
        assert wg.max() > 0.0 and (step > 0 or pg.max() > 15)   # the first resample merges crowded cells
        print(f"resample {step}: n {st.n_in} -> {st.n_out}, merged {st.cells_merged}, splits {st.splits}", flush=True)
        assert rc == 0
        assert (st.n_out, st.cells_merged, st.clones, st.splits, st.removed) == \
            (rs.n_out, rs.cells_merged, rs.clones, rs.splits, rs.removed)
        merged += rs.cells_merged
        splits += rs.splits
    assert merged > 0 and splits > 0
    g = popl.download()
    r = P.view()
    np.testing.assert_array_equal(g.id, r["id"])
    for k in ("x", "v", "w", "s", "active"):
        np.testing.assert_array_equal(getattr(g, k), r[k], err_msg=k)


def test_faces_match_oracle_bitwise():
    G = R.CGrid(RR, LL, MM, NN)
    fields = axisym.GridFields(axisym.Grid(RR, LL, MM, NN))
    u = _smooth_potential(G)
    fields.set("u", u)
    G.u[:] = u.ravel()
    eb = -5e6
    from mcstreamer_amd import _lib
    _lib.check(fields.ctx.handle, _lib.lib().mcs_cyl_efield_faces(fields.handle, eb))
    R.cyl_faces(G, eb)
    for k in ("er", "ez"):
        np.testing.assert_array_equal(fields.get(k).ravel(), getattr(G, k), err_msg=k)


def _ref_poisson(G, P, conf, eb, fields=None, freebc=False):
    """The oracle's poisson!; with the free boundary it takes the library's f_n
    (pinned against scipy separately), so that the chains compare bitwise."""
    f = axisym.freebc_coeffs(fields) if freebc else None
    return R.cyl_poisson(G, P, conf, eb, f=f, freebc=freebc)


def test_freebc_coefficients_match_scipy():
    for (Rr, L, M, N) in [(RR, LL, MM, NN), (0.015, 0.1, 1536, 10240)]:
        fields = axisym.GridFields(axisym.Grid(Rr, L, M, N))
        np.testing.assert_allclose(axisym.freebc_coeffs(fields), R.freebc_coeffs(R.CGrid(Rr, L, M, N)),
                                   rtol=2e-14)


@pytest.mark.parametrize("tailmat", ["0", "1"])
@pytest.mark.parametrize("freebc", [False, True])
def test_poisson_matches_oracle_bitwise(monkeypatch, freebc, tailmat):
    """poisson! bitwise with the solver's coarse levels as sweeps; with its
    tabulated tail (the default) the potential, the fields and the charge (whose
    column at r = R takes the free boundary's value, from the first solve) to
    1e-12 of their largest magnitude."""
    monkeypatch.setenv("MCS_MG_TAIL_MATRIX", tailmat)
    table, T, popl, P, grid, fields, G = _setup(n=40000, seed=2)
    # ions where the electrons are, less a few: a net charge of the electrons' order
    axisym.density(fields, "qfixed", popl, 0.9)
    R.cyl_density(P, G, "qfixed", 0.9)
    conf = axisym.poisson_config(grid, tolerance=-1e-10)
    ws = axisym.mgmod.Workspace(conf, (MM, NN))
    eb = -5e6
    it, res = axisym.poisson(fields, popl, eb, ws, freebc=freebc)
    it_ref = _ref_poisson(G, P, conf, eb, fields, freebc)
    print(f"freebc {freebc}: mg iterations {it} / oracle {it_ref}", flush=True)
    assert it == it_ref
    for k in ("q", "u", "er", "ez"):
        if tailmat == "0":
            np.testing.assert_array_equal(fields.get(k).ravel(), getattr(G, k), err_msg=k)
        else:
            ref = getattr(G, k)
            np.testing.assert_allclose(fields.get(k).ravel(), ref, rtol=0, atol=1e-12 * np.abs(ref).max(),
                                       err_msg=k)
    assert np.abs(G.er).max() > 1.0


def test_advance_in_axisymmetric_field_matches_oracle():
    table, T, popl, P, grid, fields, G = _setup(n=20000, seed=4)
    u = _smooth_potential(G, amp=60.0)
    fields.set("u", u)
    G.u[:] = u.ravel()
    from mcstreamer_amd import _lib
    _lib.check(fields.ctx.handle, _lib.lib().mcs_cyl_efield_faces(fields.handle, -5e6))
    R.cyl_faces(G, -5e6)
    field, tracker = axisym.FieldInterp(fields), axisym.CollisionTracker(fields)
    dt = 1e-12
    for step in range(3):
        st = jumc.advance(popl, field, dt, tracker, step=step)
        rs = R.advance_cyl(P, T, G, dt, track=True, seed=popl.seed, step=step)
        print(f"step {step}: n {rs.n_out} births {rs.births} deaths {rs.deaths}", flush=True)
        assert (st.n_out, st.births, st.deaths, st.flights) == (rs.n_out, rs.births, rs.deaths, rs.flights)
        assert list(st.events)[:len(T.kind)] == list(rs.events)[:len(T.kind)]
    assert rs.births > 0 and rs.deaths > 0
    g = popl.download()
    r = P.view()
    np.testing.assert_array_equal(g.id, r["id"])
    np.testing.assert_array_equal(g.active, r["active"])
    for k in ("x", "v", "w", "s"):
        np.testing.assert_array_equal(getattr(g, k), r[k], err_msg=k)
    qf = fields.get("qfixed").ravel()
    np.testing.assert_array_equal(qf, G.qfixed)   # per-cell sums in parent order, as the oracle's loop
    assert np.abs(G.qfixed).max() > 0


def test_axisymmetric_streamer_steps():
    """poisson! (with the free boundary) + advance! (field and tracker of the same
    grid) + repack! + resample!, four steps, against the same chain in the oracle."""
    table, T, popl, P, grid, fields, G = _setup(n=20000, seed=6)
    axisym.density(fields, "qfixed", popl, 1.0)
    R.cyl_density(P, G, "qfixed", 1.0)
    conf = axisym.poisson_config(grid)
    ws = axisym.mgmod.Workspace(conf, (MM, NN))
    field, tracker = axisym.FieldInterp(fields), axisym.CollisionTracker(fields)
    eb, dt = -5e6, 1e-12
    for step in range(4):
        it, _ = axisym.poisson(fields, popl, eb, ws)
        it_ref = _ref_poisson(G, P, conf, eb, fields, freebc=True)
        st = jumc.advance(popl, field, dt, tracker, step=step)
        rs = R.advance_cyl(P, T, G, dt, track=True, seed=popl.seed, step=step)
        jumc.repack(popl)
        R.repack(P)
        ss = axisym.resample(popl, fields, 10, 15, step=step)
        rcr, rr = R.cyl_resample(P, G, 10, 15, seed=popl.seed, step=step)
        assert ss.n_out == rr.n_out
        print(f"step {step}: mg {it}/{it_ref} n {P.n} births {rs.births} merged {rr.cells_merged}", flush=True)
        assert it == it_ref
        assert (st.births, st.deaths) == (rs.births, rs.deaths)
    g = popl.download()
    r = P.view()
    np.testing.assert_array_equal(g.id, r["id"])
    # vectors to 1e-12 of each particle's |x| or |v|, scalars to 1e-12 relative
    for k in ("x", "v"):
        err = np.abs(getattr(g, k) - r[k]).max(axis=1)
        assert np.all(err <= 1e-12 * np.linalg.norm(r[k], axis=1)), k
    for k in ("w", "s"):
        np.testing.assert_allclose(getattr(g, k), r[k], rtol=1e-12, atol=0, err_msg=k)
    np.testing.assert_allclose(fields.get("u").ravel(), G.u, rtol=1e-10, atol=1e-12 * np.abs(G.u).max())


def test_argument_errors():
    """Grid sizes, a solver that does not match the grid, the free boundary
    without Dirichlet at r = R, and resample! before the canonical sort."""
    from mcstreamer_amd import _lib
    from mcstreamer_amd import multigrid as mg
    with pytest.raises(ValueError):
        axisym.Grid(RR, LL, 1, NN)
    fields = axisym.GridFields(axisym.Grid(RR, LL, MM, NN))
    it = ctypes.c_int32()
    res = ctypes.c_double()
    wrong_n = axisym.poisson_config(axisym.Grid(RR, LL, 2 * MM, NN))
    ws = mg.Workspace(wrong_n, (2 * MM, NN))
    assert _lib.lib().mcs_cyl_poisson(fields.handle, ws.handle, 0.0, 1, ctypes.byref(it),
                                      ctypes.byref(res)) == _lib.MCS_E_ARG
    neumann_wall = mg.MGConfig(bc=[(mg.LeftBnd, mg.Dirichlet), (mg.RightBnd, mg.Dirichlet),
                                   (mg.TopBnd, mg.Neumann), (mg.BottomBnd, mg.Neumann)],
                               s=1.0, levels=3, tolerance=1e-3, ndim=2, conn=mg.CylindricalConnector(1))
    ws2 = mg.Workspace(neumann_wall, (MM, NN))
    assert _lib.lib().mcs_cyl_poisson(fields.handle, ws2.handle, 0.0, 1, ctypes.byref(it),
                                      ctypes.byref(res)) == _lib.MCS_E_ARG
    assert b"Dirichlet" in _lib.lib().mcs_last_error(fields.ctx.handle)
    table, T, popl, P, grid, f2, G = _setup(n=1000, seed=2)
    st = _lib.ResampleStats()
    assert _lib.lib().mcs_cyl_resample(popl.handle, f2.handle, 10, 15, 0, 0, 0, ctypes.byref(st)) == _lib.MCS_E_ARG
