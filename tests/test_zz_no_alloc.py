"""The allocation-free moments entry point (mcs_pop_moments_noalloc) against
mcs_pop_moments through the C ABI: every field bitwise equal, on two calls in a
row with the second on a smaller population (the shared scratch keeps the first
call's partial records, which must not leak into the second), with and
without a collision table, FP64 and split-precision stores."""
import ctypes

import numpy as np
import pytest

from mcstreamer_amd import _lib, jumc
from mcstreamer_amd.streamer import Grid3, GridFields
from tests.common import gas_table, maxwell_states

pytestmark = pytest.mark.gpu


def _moments(fn, pop, table):
    m = _lib.Moments()
    rc = fn(pop.handle, table.handle if table is not None else None, ctypes.byref(m))
    jumc.check(pop.ctx.handle, rc)
    return m


def _assert_same(a, b):
    for name, _ in _lib.Moments._fields_:
        x, y = getattr(a, name), getattr(b, name)
        if hasattr(x, "__len__"):
            assert list(x) == list(y), name
        else:
            assert x == y, name


def _pop(n, seed, table, precision="f64", mesh=None, cap=None):
    x, v, w, s, act = maxwell_states(n, kT_ev=3.0, seed=seed, spread=2e-5)
    if mesh is not None:
        x = np.abs(x) % 1.5e-4 + 1e-6
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 4.0, n)
    act = (rng.uniform(size=n) < 0.9).astype(np.uint8)
    ps = jumc.ParticleStates(x, v, w, s, act, np.arange(n, dtype=np.uint64), n)
    return jumc.Population(cap or n, ps, table, seed=seed, precision=precision, mesh=mesh)


@pytest.mark.parametrize("with_table", [True, False])
def test_noalloc_moments_equal_moments_two_calls(with_table):
    proc, energy, rate, maxrate = gas_table(o2=0.2)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    t = table if with_table else None
    L = _lib.lib()
    # largest first (more blocks than the reduction's cap), then smaller ones
    for n, seed in ((400_000, 1), (70_001, 2), (300, 3)):
        print(f"noalloc moments: n {n} table {with_table}", flush=True)
        pop = _pop(n, seed, table)
        ref = _moments(L.mcs_pop_moments, pop, t)
        got = _moments(L.mcs_pop_moments_noalloc, pop, t)
        _assert_same(got, ref)
        assert got.nparticles == n and got.nactive > 0


def test_noalloc_moments_same_population_shrinks_between_calls():
    proc, energy, rate, maxrate = gas_table(o2=0.0)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    L = _lib.lib()
    pop = _pop(200_000, 5, table, cap=200_000)
    first = _moments(L.mcs_pop_moments_noalloc, pop, table)
    _assert_same(first, _moments(L.mcs_pop_moments, pop, table))
    x, v, w, s, act = maxwell_states(1000, kT_ev=1.0, seed=6)
    pop.upload(jumc.ParticleStates(x, v, w, s, act, np.arange(1000, dtype=np.uint64), 1000))
    second = _moments(L.mcs_pop_moments_noalloc, pop, table)
    _assert_same(second, _moments(L.mcs_pop_moments, pop, table))
    assert second.nparticles == 1000 and second.weight == 1000.0
    # an empty population reads zeros without touching the device
    pop.upload(jumc.ParticleStates(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0),
                                   np.zeros(0, dtype=np.uint8), np.zeros(0, dtype=np.uint64), 0))
    empty = _moments(L.mcs_pop_moments_noalloc, pop, table)
    assert empty.nparticles == 0 and empty.weight == 0.0


def test_noalloc_moments_split_store():
    proc, energy, rate, maxrate = gas_table(o2=0.2)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    fields = GridFields(Grid3((32, 32, 32), 5e-6))
    L = _lib.lib()
    for n, seed in ((150_000, 7), (5_000, 8)):
        pop = _pop(n, seed, table, precision="f32", mesh=fields)
        _assert_same(_moments(L.mcs_pop_moments_noalloc, pop, table), _moments(L.mcs_pop_moments, pop, table))
