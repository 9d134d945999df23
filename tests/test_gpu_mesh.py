"""GPU parity of the mesh side of the step against the oracle: canonical cell
sort, NGP/CIC deposition (density!, poisson.jl:65-84), charge assembly and
face fields (poisson.jl:19-57), the fused step with the mesh gather and the ion
tracker (gridfields.jl:115-224), per-cell resampling (resample.jl:7-84),
repack!, the population reductions and Swarm.resample!.

Bar: integer/index results (order, ids, counts) bit-exact; floating results
bit-exact where the operation order is the same on both sides (sort, deposit,
faces, push, the ion tracker's per-cell sums in parent order)."""
import numpy as np
import pytest

from mcstreamer_amd import constants as co
from mcstreamer_amd import jumc, streamer
from oracle import refcpu as R
from tests.common import gas_table, random_particles, ref_table, smooth_potential

pytestmark = pytest.mark.gpu


def _pair(n, h, npart, seed, p_inactive=0.0, p_outside=0.0, o2=0.2, cap=None):
    L = np.array(n, dtype=float) * h
    x, v, w, s, act = random_particles(npart, L, seed, p_inactive, p_outside)
    cap = cap or 4 * npart + 64
    proc, energy, rate, maxrate = gas_table(o2=o2)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    pop = jumc.Population(cap, jumc.ParticleStates(x, v, w, s, act, np.arange(npart, dtype=np.uint64), npart),
                          table, seed=seed)
    fields = streamer.GridFields(streamer.Grid3(n, h))
    P = R.Pop(x, v, w, s, act, cap=cap)
    M = R.Mesh(n, h)
    T = ref_table(proc, energy, rate, maxrate)
    return pop, fields, P, M, T


def _same_pop(pop, P, exact=True):
    g = pop.download()
    r = P.view()
    assert len(g) == P.n
    np.testing.assert_array_equal(g.id, r["id"])
    np.testing.assert_array_equal(g.active, r["active"])
    for k in ("x", "v", "w", "s"):
        if exact:
            np.testing.assert_array_equal(getattr(g, k), r[k], err_msg=k)
        else:
            np.testing.assert_allclose(getattr(g, k), r[k], rtol=1e-12, atol=0, err_msg=k)


@pytest.mark.parametrize("n,npart", [((8, 6, 10), 3000), ((16, 16, 32), 40000)])
def test_sort_by_cell_parity(n, npart):
    pop, fields, P, M, _ = _pair(n, 1e-5, npart, 1, p_inactive=0.2, p_outside=0.05)
    streamer.sort_by_cell(pop, fields)
    R.sort_by_cell(P, M)
    _same_pop(pop, P)


def test_sort_handles_long_segments():
    # many particles in very few cells: exercises the block-wide segment sort
    n, h = (4, 4, 4), 1e-5
    pop, fields, P, M, _ = _pair(n, h, 5000, 2)
    x = np.full((5000, 3), 0.5e-5)
    x[::3] = 2.5e-5
    ps = pop.download()
    ps.x[:] = x
    rng = np.random.default_rng(0)
    perm = rng.permutation(5000).astype(np.uint64)
    ps.id = perm
    pop.upload(ps)
    P = R.Pop(x, ps.v, ps.w, ps.s, ps.active, perm, cap=P.cap)
    streamer.sort_by_cell(pop, fields)
    R.sort_by_cell(P, M)
    _same_pop(pop, P)


@pytest.mark.parametrize("scheme,blob,dense,lane_listed", [
    ("ngp", False, None, None), ("cic", False, None, None), ("cic", True, None, None),
    # a wavefront per listed cell over most cells; one lane per cell everywhere
    ("cic", True, 8, None), ("cic", False, 4, None), ("cic", True, 8, -1), ("cic", True, 64, -1),
    # the first call lists the dense cells, the second (their count above 5) does not
    ("cic", True, 8, 5)])
def test_deposit_parity(scheme, blob, dense, lane_listed, monkeypatch):
    if dense is not None:
        monkeypatch.setenv("MCS_DEP_DENSE", str(dense))
    if lane_listed is not None:
        monkeypatch.setenv("MCS_DEP_LANE_LISTED", str(lane_listed))
    n, h = (12, 10, 14), 2e-6
    pop, fields, P, M, _ = _pair(n, h, 20000, 3, p_outside=0.02)
    if blob:
        # a dense core and sparse edges: cells on both sides of the wavefront
        # threshold of the CIC gather
        ps = pop.download()
        rng = np.random.default_rng(11)
        ps.x[:] = np.array(n) * h / 2 + rng.normal(size=ps.x.shape) * 1.5 * h
        pop.upload(ps)
        P = R.Pop(ps.x, ps.v, ps.w, ps.s, ps.active, ps.id, cap=P.cap)
    streamer.sort_by_cell(pop, fields)
    R.sort_by_cell(P, M)
    streamer.density(fields, "qpart", pop, -1.0, scheme)
    R.deposit(P, M, 1, 0 if scheme == "ngp" else 1, -1.0)
    np.testing.assert_array_equal(fields.get("qpart"), M.qpart)
    if lane_listed is not None and lane_listed > 0:
        # again, with the other gather (the deposit adds to the array)
        streamer.density(fields, "qpart", pop, -1.0, scheme)
        R.deposit(P, M, 1, 1, -1.0)
        np.testing.assert_array_equal(fields.get("qpart"), M.qpart)
    # charge conservation (exact up to rounding) for the particles inside
    inside = np.all((P.view()["x"] >= 0) & (P.view()["x"] < np.array(n) * h), axis=1)
    if scheme == "ngp":
        total = -P.view()["w"][inside].sum()
        assert M.qpart.sum() * h ** 3 == pytest.approx(total, rel=1e-12)


def test_assemble_and_faces_parity():
    n, h = (10, 12, 16), 5e-6
    pop, fields, P, M, _ = _pair(n, h, 100, 4)
    rng = np.random.default_rng(5)
    for name, arr in (("qfixed", M.qfixed), ("qpart", M.qpart), ("u", M.u)):
        arr[:] = rng.normal(size=arr.size)
        fields.set(name, arr)
    from mcstreamer_amd import _lib
    _lib.check(fields.ctx.handle, _lib.lib().mcs_assemble_charge(fields.handle))
    _lib.check(fields.ctx.handle, _lib.lib().mcs_efield_faces(fields.handle, -5e6))
    R.assemble_charge(M)
    R.efield_faces(M, -5e6)
    np.testing.assert_array_equal(fields.get("q"), M.q)
    for name in ("ex", "ey", "ez"):
        np.testing.assert_array_equal(fields.get(name), getattr(M, name), err_msg=name)


def test_advance_mesh_field_and_tracker_parity():
    n, h = (16, 16, 24), 5e-6
    pop, fields, P, M, T = _pair(n, h, 20000, 6, p_outside=0.01, o2=0.25)
    M.u[:] = smooth_potential(n, h, amp=3.0)
    fields.set("u", M.u)
    from mcstreamer_amd import _lib
    _lib.check(fields.ctx.handle, _lib.lib().mcs_efield_faces(fields.handle, -8e6))
    R.efield_faces(M, -8e6)
    fld = streamer.FieldInterp(fields)
    trk = streamer.CollisionTracker(fields)
    f = R.mesh_field(M)
    for step in range(3):
        gs = jumc.advance(pop, fld, 1e-12, trk, step=step)
        rs = R.advance(P, T, f, 1e-12, M, seed=pop.seed, step=step)
        print(f"step {step}: births {rs.births} deaths {rs.deaths}", flush=True)
        assert list(gs.events)[:len(T.kind)] == list(rs.events)[:len(T.kind)]
        assert (gs.births, gs.deaths, gs.flights) == (rs.births, rs.deaths, rs.flights)
    _same_pop(pop, P)
    # the tracker's deposits are added per cell in parent order: bitwise
    np.testing.assert_array_equal(fields.get("qfixed").ravel(), M.qfixed)
    assert rs.births > 0 and rs.deaths > 0


@pytest.mark.parametrize("seed", [7, 8])
def test_resample_cells_parity(seed):
    n, h = (6, 6, 6), 1e-5
    rng = np.random.default_rng(seed)
    # clustered particles: some cells far above nmax, some below ntarget with w > 2
    npart = 6000
    centers = rng.uniform(0.1, 0.9, size=(5, 3)) * np.array(n) * h
    x = centers[rng.integers(0, 5, npart)] + rng.normal(size=(npart, 3)) * 0.4 * h
    x = np.clip(x, 1e-9, np.array(n) * h - 1e-9)
    sparse = rng.uniform(size=(300, 3)) * np.array(n) * h
    x = np.concatenate([x, sparse])
    npart = x.shape[0]
    v = rng.normal(size=(npart, 3)) * 1e5
    w = rng.integers(1, 9, npart).astype(float)
    s = rng.uniform(size=npart)
    act = np.ones(npart, dtype=np.uint8)
    proc, energy, rate, maxrate = gas_table()
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    pop = jumc.Population(3 * npart, jumc.ParticleStates(x, v, w, s, act, np.arange(npart, dtype=np.uint64), npart),
                          table, seed=seed)
    fields = streamer.GridFields(streamer.Grid3(n, h))
    P = R.Pop(x, v, w, s, act, cap=3 * npart)
    M = R.Mesh(n, h)
    streamer.sort_by_cell(pop, fields)
    R.sort_by_cell(P, M)
    w_before = P.view()["w"].sum()
    gst = streamer.resample(pop, fields, 8, 64, step=3)
    rc, rst = R.resample_cells(P, M, 8, 64, seed=seed, step=3)
    assert rc == 0
    assert (gst.n_out, gst.cells_merged, gst.clones, gst.splits, gst.removed) == \
        (rst.n_out, rst.cells_merged, rst.clones, rst.splits, rst.removed)
    assert rst.cells_merged > 0 and rst.splits > 0
    _same_pop(pop, P)
    assert P.view()["w"].sum() == pytest.approx(w_before, rel=1e-12)
    # the cell order and the starts carry over: a second resample agrees too
    gst = streamer.resample(pop, fields, 8, 64, step=4)
    rc, rst = R.resample_cells(P, M, 8, 64, seed=seed, step=4)
    _same_pop(pop, P)


def test_repack_moments_swarm_resample_parity():
    proc, energy, rate, maxrate = gas_table()
    T = ref_table(proc, energy, rate, maxrate)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    rng = np.random.default_rng(9)
    for npart in (1, 255, 256, 257, 5000, 70001):
        x = rng.normal(size=(npart, 3))
        v = rng.normal(size=(npart, 3)) * 1e6
        w = rng.uniform(1, 2, npart)
        s = rng.uniform(size=npart)
        act = (rng.uniform(size=npart) < 0.6).astype(np.uint8)
        ids = np.arange(npart, dtype=np.uint64)
        pop = jumc.Population(3 * npart + 10, jumc.ParticleStates(x, v, w, s, act, ids, npart), table, seed=2)
        P = R.Pop(x, v, w, s, act, ids, cap=3 * npart + 10)
        m_g = pop.moments(table)
        m_r = R.moments(P, T)
        assert (m_g.nparticles, m_g.nactive) == (m_r.nparticles, m_r.nactive)
        for k in ("weight", "wenergy", "maxenergy", "sum_energy"):
            assert getattr(m_g, k) == pytest.approx(getattr(m_r, k), rel=1e-12, abs=1e-300)
        for k in ("sum_x", "sum_x2", "sum_v"):
            ref = np.array(getattr(m_r, k)[:])
            np.testing.assert_allclose(getattr(m_g, k)[:], ref, rtol=1e-11,
                                       atol=1e-12 * np.abs(getattr(P, k[4:] if k != "sum_x2" else "x")).sum())
        np.testing.assert_allclose(m_g.collrate[:len(proc)], m_r.collrate[:len(proc)], rtol=1e-12)
        jumc.repack(pop)
        R.repack(P)
        _same_pop(pop, P)
        # global resample towards a target below and above the current count
        for step, n0 in ((0, max(1, P.n // 2)), (1, 2 * P.n + 1)):
            from mcstreamer_amd import swarm
            did = swarm.resample(pop, n0, 0.1, step=step)
            rc, did_r = R.swarm_resample(P, n0, 0.1, seed=2, step=step)
            assert did == bool(did_r)
            _same_pop(pop, P)


def test_sort_handles_huge_segments():
    # one cell with 12000 particles and 9000 outside the mesh: segments above the
    # block-sort size take the grid-wide rank path; order must equal the oracle's
    n, h = (4, 4, 4), 1e-5
    N = 30000
    pop, fields, P, M, _ = _pair(n, h, N, 4)
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 4e-5, size=(N, 3))
    x[:12000] = 1.1e-5 + rng.uniform(0, 0.8e-5, size=(12000, 3))
    x[12000:21000, 2] = 5e-5 + rng.uniform(0, 1e-4, size=9000)
    ps = pop.download()
    ps.x[:] = x
    perm = rng.permutation(N).astype(np.uint64)
    ps.id = perm
    pop.upload(ps)
    P = R.Pop(x, ps.v, ps.w, ps.s, ps.active, perm, cap=P.cap)
    streamer.sort_by_cell(pop, fields)
    R.sort_by_cell(P, M)
    _same_pop(pop, P)
