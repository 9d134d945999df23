"""Served swarm requests (continuous batching) on the device.

Each request's result must be the same alone and in the mix (bitwise: every slot
draws from its own Philox key and is checked at its own step counts), the plain
one-request-at-a-time path must agree with the batched one, and admission must
wait for a free slab instead of failing."""
import numpy as np
import pytest

from mcstreamer_amd.serve import GasTables, PlainServer, Policy, Request, SwarmServer

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

