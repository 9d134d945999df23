"""Photoionization on the device against the oracle (JuMC/src/photon.jl:11-56,
electron.jl:88-105,141-144, collisions.jl:117-141, mixed_population.jl:24-54,
gridfields.jl:219-222): advance! over MultiPopulation(electrons, photons) with
PhotoEmission rows, the photon loop of ZhelezniakCollisions and the ion tracker.

Bar: bitwise.  Electrons and photons (x, v, nu, w, s, active, id), the
tracker's array and every count of both loops equal the oracle's after each
step, on the 3D mesh field and on the axisymmetric grid; photon repack! and the
photon moments equal the oracle's and numpy's; the entry points refuse the
misuses the reference cannot express (a photoemission table without photons,
photons where electrons belong)."""
import numpy as np
import pytest

from mcstreamer_amd import _lib, axisym, jumc
from mcstreamer_amd import constants as co
from mcstreamer_amd import streamer as S
from oracle import refcpu as R
from tests.common import gas_table, ref_table, smooth_potential

pytestmark = pytest.mark.gpu


def _tables(photo_rescale=1.0, mult=2.0, pw=1.0):
    proc, energy, rate, maxrate = gas_table(ne=800, o2=0.25, photo=True, photo_rescale=photo_rescale,
                                            photon_multiplier=mult, photon_weight=pw)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    T = ref_table(proc, energy, rate, maxrate)
    zc = jumc.ZhelezniakCollisions.for_table(table)
    return proc, table, T, zc


def _electrons(n, lo, hi, seed, kT_ev=20.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, size=(n, 3))
    v = rng.normal(size=(n, 3)) * np.sqrt(kT_ev * co.eV / co.electron_mass)
    w = rng.integers(1, 5, size=n).astype(np.float64)
    s = -np.log1p(-rng.uniform(size=n))
    act = (rng.uniform(size=n) > 0.03).astype(np.uint8)
    return x, v, w, s, act


def _check_stats(g, r, nproc, tag):
    got = (g.n_in, g.n_out, g.births, g.deaths, g.flights, list(g.events)[:nproc], g.flags)
    want = (r.n_in, r.n_out, r.births, r.deaths, r.flights, list(r.events)[:nproc], r.flags)
    assert got == want, tag


def _compare(mp, E, P, rtol=0.0):
    """Ids and flags bitwise; states bitwise (rtol 0) or to rtol relative to
    each array's largest magnitude."""
    def same(a, b, msg):
        if rtol == 0.0:
            np.testing.assert_array_equal(a, b, err_msg=msg)
        else:
            np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * np.abs(b).max(initial=0.0), err_msg=msg)

    g, r = mp.electron.download(), E.view()
    np.testing.assert_array_equal(g.id, r["id"])
    np.testing.assert_array_equal(g.active, r["active"])
    for k in ("x", "v", "w", "s"):
        same(getattr(g, k), r[k], "electron " + k)
    g, r = mp.photon.download(), P.view()
    assert g.next_id == P.next_id
    np.testing.assert_array_equal(g.id, r["id"])
    np.testing.assert_array_equal(g.active, r["active"])
    for k in ("x", "v", "nu", "w", "s"):
        same(getattr(g, k), r[k], "photon " + k)


@pytest.mark.parametrize("tracked", [True, False])
def test_mixed_step_on_mesh_matches_oracle(tracked):
    proc, table, T, zc = _tables()
    n, h = (24, 24, 32), 2e-5
    L = np.array(n) * h
    x, v, w, s, act = _electrons(6000, 0.15 * L, 0.85 * L, seed=21)
    cap = 40 * len(w)
    el = jumc.Population(cap, jumc.ParticleStates(x, v, w, s, act), table, seed=17)
    ph = jumc.PhotonPopulation(200 * len(w), collisions=zc)
    mp = jumc.MultiPopulation(el, ph)
    E = R.Pop(x, v, w, s, act, cap=cap)
    P = R.Pop.photons(cap=200 * len(w))
    fields = S.GridFields(S.Grid3(n, h))
    M = R.Mesh(n, h)
    u = smooth_potential(n, h, amp=20.0)
    fields.set("u", u)
    M.u[:] = u
    eb = -4e6
    _lib.check(fields.ctx.handle, _lib.lib().mcs_efield_faces(fields.handle, eb))
    R.efield_faces(M, eb)
    field = S.FieldInterp(fields)
    tracker = S.CollisionTracker(fields) if tracked else None
    dt = 4e-13
    for step in range(4):
        est, pst = jumc.advance(mp, field, dt, tracker)
        rest, rpst = R.advance_mixed(E, P, T, zc.numax, R.mesh_field(M), dt, M if tracked else None,
                                     seed=el.seed, step=step)
        print(f"step {step}: electrons {rest.n_out}+{rpst.births} photons {rpst.n_in}->{rpst.n_out} "
              f"absorbed {rpst.events[0]}", flush=True)
        _check_stats(est, rest, len(proc), f"electron loop, step {step}")
        _check_stats(pst, rpst, 1, f"photon loop, step {step}")
    assert rpst.n_out > 0 and rpst.events[0] > 0 and rpst.births > 0
    _compare(mp, E, P)
    if tracked:
        np.testing.assert_array_equal(fields.get("qfixed"), M.qfixed)
        assert np.abs(M.qfixed).max() > 0


def test_mixed_step_on_axisymmetric_grid_matches_oracle():
    proc, table, T, zc = _tables(photo_rescale=2.0, mult=1.0)
    RR, LL, MM, NN = 0.32e-3, 0.64e-3, 32, 64
    rng = np.random.default_rng(8)
    nn = 6000
    r = RR * 0.9 * np.sqrt(rng.uniform(size=nn))
    th = rng.uniform(0, 2 * np.pi, nn)
    z = rng.uniform(0.1 * LL, 0.9 * LL, nn)
    x = np.stack([r * np.cos(th), r * np.sin(th), z], axis=1)
    _, v, w, s, act = _electrons(nn, 0.0, 1.0, seed=9)
    cap = 40 * nn
    el = jumc.Population(cap, jumc.ParticleStates(x, v, w, s, act), table, seed=23)
    ph = jumc.PhotonPopulation(200 * nn, collisions=zc)
    mp = jumc.MultiPopulation(el, ph)
    E = R.Pop(x, v, w, s, act, cap=cap)
    P = R.Pop.photons(cap=200 * nn)
    grid = axisym.Grid(RR, LL, MM, NN)
    fields = axisym.GridFields(grid)
    G = R.CGrid(RR, LL, MM, NN)
    I = (np.arange(MM + 2)[None, :] - 0.5) * G.steps[0]
    J = (np.arange(NN + 2)[:, None] - 0.5) * G.steps[1]
    u = 30.0 * np.exp(-(I / (0.3 * RR)) ** 2 - ((J - 0.6 * LL) / (0.2 * LL)) ** 2)
    fields.set("u", u)
    G.u[:] = u.ravel()
    _lib.check(fields.ctx.handle, _lib.lib().mcs_cyl_efield_faces(fields.handle, -4e6))
    R.cyl_faces(G, -4e6)
    field, tracker = axisym.FieldInterp(fields), axisym.CollisionTracker(fields)
    dt = 4e-13
    for step in range(4):
        est, pst = jumc.advance(mp, field, dt, tracker)
        rest, rpst = R.advance_cyl_mixed(E, P, T, zc.numax, G, dt, track=True, seed=el.seed, step=step)
        print(f"step {step}: electrons {rest.n_out}+{rpst.births} photons {rpst.n_in}->{rpst.n_out}", flush=True)
        _check_stats(est, rest, len(proc), f"electron loop, step {step}")
        _check_stats(pst, rpst, 1, f"photon loop, step {step}")
    assert rpst.n_out > 0 and rpst.births > 0
    _compare(mp, E, P)
    np.testing.assert_array_equal(fields.get("qfixed").ravel(), G.qfixed)


def test_photon_repack_and_moments():
    proc, table, T, zc = _tables()
    rng = np.random.default_rng(4)
    n = 50_000
    x = rng.uniform(-1e-3, 1e-3, size=(n, 3))
    d = rng.normal(size=(n, 3))
    v = d / np.linalg.norm(d, axis=1)[:, None] * co.c
    nu = zc.numax * rng.uniform(size=n)
    w = rng.integers(1, 4, size=n).astype(np.float64)
    s = -np.log1p(-rng.uniform(size=n))
    act = (rng.uniform(size=n) > 0.4).astype(np.uint8)
    ph = jumc.PhotonPopulation(n, jumc.PhotonStates(x, v, nu, w, s, act), collisions=zc)
    m = ph.moments()
    a = act == 1
    assert (m.nparticles, m.nactive) == (n, int(a.sum()))
    assert m.weight == pytest.approx(w[a].sum(), rel=1e-14)
    np.testing.assert_allclose(list(m.sum_x), x[a].sum(axis=0), rtol=1e-12, atol=1e-18)
    assert m.wenergy == 0.0 and m.maxenergy == 0.0 and m.sum_energy == 0.0
    jumc.repack(ph)
    P = R.Pop.photons(cap=n, x=x, v=v, nu=nu, w=w, s=s, active=act)
    R.repack(P)
    g, r = ph.download(), P.view()
    assert len(g) == P.n == a.sum()
    np.testing.assert_array_equal(g.id, r["id"])
    for k in ("x", "v", "nu", "w", "s", "active"):
        np.testing.assert_array_equal(getattr(g, k), r[k], err_msg=k)


def test_photon_misuse_is_refused():
    proc, table, T, zc = _tables()
    x, v, w, s, act = _electrons(100, 0.0, 1e-4, seed=2)
    el = jumc.Population(1000, jumc.ParticleStates(x, v, w, s, act), table, seed=1)
    with pytest.raises(_lib.MCSError, match="photoemission"):
        jumc.advance(el, jumc.FixedField(-1e6), 1e-13)
    with pytest.raises(_lib.MCSError, match="photoemission"):
        jumc.advance(jumc.MultiPopulation(el, None), jumc.FixedField(-1e6), 1e-13)
    ph = jumc.PhotonPopulation(10, collisions=zc)
    h = ph.handle
    with pytest.raises(_lib.MCSError, match="photon"):
        _lib.check(ph.ctx.handle, _lib.lib().mcs_pop_upload(h, 0, None, None, None, None, None, None, 0))
    with pytest.raises(_lib.MCSError, match="photon"):
        _lib.check(ph.ctx.handle, _lib.lib().mcs_swarm_resample(h, 10, 0.1, 0, 0, 0, None))
    # a photon population too small for the emission: the capacity error names photons
    x, v, w, s, act = _electrons(4000, 0.0, 1e-4, seed=3, kT_ev=60.0)
    el = jumc.Population(100_000, jumc.ParticleStates(x, v, w, s, act), table, seed=1)
    ph = jumc.PhotonPopulation(4, collisions=zc)
    with pytest.raises(_lib.MCSError, match="Maximum number of photons"):
        for _ in range(5):
            jumc.advance(jumc.MultiPopulation(el, ph), jumc.FixedField(-5e6), 1e-12)


def test_streamer_steps_with_photons_match_oracle():
    """Whole 3D streamer steps with photoionization: poisson! (CIC deposit,
    multigrid, faces), advance! over electrons and photons with the tracker,
    photon repack!, the canonical sort and resample!, against the same chain in
    the oracle; counts bit-exact every step, ids bitwise, particle and photon
    states to 1e-12 relative (the solver's tabulated tail rounds the potential
    differently from the oracle's sweeps in the last bits)."""
    n, npairs, cap = (32, 32, 64), 20_000, 400_000
    cfg = S.make_config1(n=n, npairs=npairs, cap=cap, seed=5,
                         photo=dict(rescale=1.0, photon_multiplier=8.0, photon_weight=1.0))
    # the seed at 20 eV instead of at rest, so that the first steps photoemit
    ps = cfg["popl"].download()
    rng = np.random.default_rng(6)
    v = rng.normal(size=ps.v.shape) * np.sqrt(20.0 * co.eV / co.electron_mass)
    cfg["popl"].upload(jumc.ParticleStates(ps.x, v, ps.w, ps.s, ps.active, ps.id, ps.next_id))
    sim = S.Streamer(cfg["popl"], cfg["fields"], cfg["ws"], cfg["eb"], cfg["dt"], cfg["ntarget"], cfg["nmax"],
                     photons=cfg["photons"])
    ps = sim.popl.download()
    E = R.Pop(ps.x, ps.v, ps.w, ps.s, ps.active, ps.id, cap=cap, next_id=ps.next_id)
    P = R.Pop.photons(cap=cap)
    M = R.Mesh(n, 5e-6)
    M.qfixed[:] = sim.fields.get("qfixed")
    proc = cfg["table"].proc
    T = ref_table(proc, cfg["table"].energy, cfg["table"].rate, cfg["table"].maxrate)
    numax = cfg["photons"].collisions.numax
    conf = cfg["conf"]
    R.sort_by_cell(E, M)
    for i in range(3):
        out = sim.step()
        M.qpart[:] = 0.0
        R.deposit(E, M, 1, 1, -1.0)
        R.assemble_charge(M)
        rc, it, res = R.mg_solve(n, conf.bc, conf.s, conf.levels, conf.smooth1, conf.smooth2, conf.tolerance,
                                 conf.maxiter, M.u, M.q)
        R.efield_faces(M, sim.eb)
        est, pst = R.advance_mixed(E, P, T, numax, R.mesh_field(M), sim.dt, M, seed=sim.popl.seed, step=i)
        R.repack(P)
        R.sort_by_cell(E, M)
        rc2, rst = R.resample_cells(E, M, sim.ntarget, sim.nmax, seed=sim.popl.seed, step=i)
        print(f"step {i}: electrons {E.n} photons {P.n} emitted {pst.n_out - pst.n_in} absorbed {pst.events[0]}",
              flush=True)
        assert out["mg_iters"] == it
        _check_stats(out["advance"], est, len(proc), f"electron loop, step {i}")
        _check_stats(out["photons"], pst, 1, f"photon loop, step {i}")
        assert out["resample"].n_out == rst.n_out
    assert pst.n_out > 0
    _compare(jumc.MultiPopulation(sim.popl, sim.photons), E, P, rtol=1e-12)
    np.testing.assert_allclose(sim.fields.get("qfixed").ravel(), M.qfixed, rtol=1e-12,
                               atol=1e-12 * np.abs(M.qfixed).max())
