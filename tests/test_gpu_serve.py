"""Served swarm requests (continuous batching) on the device.

Each request's result must be the same alone and in the mix (bitwise: every slot
draws from its own Philox key and is checked at its own step counts), the plain
one-request-at-a-time path must agree with the batched one, and admission must
wait for a free slab instead of failing."""
import numpy as np
import pytest

from mcstreamer_amd import constants as co
from mcstreamer_amd.serve import GasTables, PlainServer, Policy, Request, SwarmServer, _field, request_dt
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

SEED = 99


def _policy(**kw):
    p = Policy(relax_steps=60, block_steps=20, min_blocks=3, max_steps=400, chunk=20,
               capacity_factor=2.0, capacity_pad=2048, compact_every=4, max_events_per_step=1e12)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _requests():
    return [
        Request(0, 0.0, 120.0, 0.1, 1500, 0.03),
        Request(1, 0.0, 800.0, 0.25, 800, 0.05),
        Request(2, 0.0, 40.0, 0.0, 2500, 0.04),
        Request(3, 0.0, 300.0, 0.2, 600, 0.03),
        Request(4, 0.0, 15.0, 0.3, 1200, 0.05),
    ]


@pytest.fixture(scope="module")
def tables():
    return GasTables()


def _key(res):
    return {r.rid: r for r in res}


def test_retired_requests_equal_lone_oracle_trajectories(tables):
    """Served with the FP64 store: at its retirement step, each request's
    particles (active ones, in slab order) equal, bitwise, a lone oracle
    population keyed (seed, request id) stepped as many times -- thermal start
    (phase INIT), then advance! and Swarm.resample! per step (Swarm.jl:124-142)
    -- and its drift-velocity estimate equals the oracle's per-step means within
    1e-12 (the device sums in another order, compensated)."""
    reqs = _requests()[:4]
    srv = SwarmServer(400_000, max_slots=8, seed=SEED, policy=_policy(), tables=tables, precision="f64")
    snaps = {}

    def grab(slot, res):
        snaps[res.rid] = (srv.batch.download(slot), srv.batch.tally(slot))

    res, left = srv.run(reqs, on_retire=grab)
    assert not left and len(res) == len(reqs)
    pol = _policy()
    vth = float(np.sqrt(pol.init_kT_ev * co.eV / co.electron_mass))
    for r in reqs:
        got, t = snaps[r.rid]
        table = tables.get(r.o2)
        T = R.Table(table.energy, table.rate, table.kind, table.param, table.maxrate, egrid=table.egrid)
        P = R.swarm_init(r.n0, SEED, r.rid, vth, cap=pol.capacity(r.n0) * 4)
        f = R.uniform_field(_field(r.en_td))
        vz = []
        for st in range(int(t.steps)):
            R.advance(P, T, f, request_dt(r.en_td), None, seed=SEED, instance=r.rid, step=st)
            m = R.moments(P, T)
            vz.append(m.sum_v[2] / m.nactive)
            R.swarm_resample(P, r.n0, pol.tol, seed=SEED, instance=r.rid, step=st)
        ref = P.view()
        a, g = ref["active"] == 1, got.active == 1
        print(f"request {r.rid}: {t.steps} steps, actives {a.sum()} / {g.sum()}", flush=True)
        np.testing.assert_array_equal(got.id[g], ref["id"][a])
        for key in ("x", "v", "w", "s"):
            np.testing.assert_array_equal(getattr(got, key)[g], ref[key][a], err_msg=key)
        rec = np.array(vz[pol.relax_steps:])
        np.testing.assert_allclose(t.sum[0], rec.sum(), rtol=1e-12)


def test_results_same_alone_and_in_mix(tables):
    reqs = _requests()
    srv = SwarmServer(400_000, max_slots=8, seed=SEED, policy=_policy(), tables=tables)
    mixed, left = srv.run(reqs)
    assert not left and len(mixed) == len(reqs)
    mixed = _key(mixed)
    for r in (reqs[0], reqs[2], reqs[4]):
        alone, _ = SwarmServer(100_000, max_slots=2, seed=SEED, policy=_policy(),
                               tables=tables).run([r])
        a, m = alone[0], mixed[r.rid]
        print(f"request {r.rid}: steps {a.steps} / {m.steps}, W {a.estimates['drift_velocity']}", flush=True)
        assert a.steps == m.steps and a.flags == 0 and m.flags == 0
        for name in a.estimates:
            assert a.estimates[name] == m.estimates[name], name


def test_plain_path_agrees_with_batched(tables):
    # converged by the minimum step rule (large target): both stop at the same
    # step and see the same trajectory (the plain path without repack!)
    pol = _policy(compact_every=0)
    r = Request(7, 0.0, 200.0, 0.15, 1000, 0.9)
    batched, _ = SwarmServer(100_000, max_slots=2, seed=SEED, policy=_policy(),
                             tables=tables).run([r])
    steps, est = PlainServer(seed=SEED, policy=pol, tables=tables).solve(r)
    b = batched[0]
    assert steps == b.steps == pol.relax_steps + pol.min_blocks * pol.block_steps
    for name in est:
        np.testing.assert_allclose(est[name][0], b.estimates[name][0], rtol=1e-10)
        np.testing.assert_allclose(est[name][1], b.estimates[name][1], rtol=1e-6)


def test_admission_waits_for_free_slab(tables):
    # room for two requests of this footprint at a time: the others wait, then
    # take the freed slabs; every request completes with the mix-independent result
    pol = _policy()
    reqs = [Request(k, 0.0, 100.0 + 50 * k, 0.1, 1000, 0.9) for k in range(4)]
    cap = pol.capacity(1000)
    slab = (cap + 511) // 512 * 512
    srv = SwarmServer(2 * slab + 100, max_slots=8, seed=SEED, policy=pol, tables=tables)
    res, left = srv.run(reqs)
    assert not left and len(res) == 4
    res = sorted(res, key=lambda x: x.rid)
    assert res[2].admitted >= min(res[0].finished, res[1].finished)
    alone, _ = SwarmServer(slab + 100, max_slots=1, seed=SEED, policy=pol, tables=tables).run([reqs[3]])
    assert alone[0].estimates == res[3].estimates

