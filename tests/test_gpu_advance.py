"""GPU parity of the fused particle step (JuMC.advance!, mixed_population.jl:24-54)
against the oracle on the same seeded inputs.  Bar: event counts, birth/death
counts and ids bit-exact; states equal to 1e-12 relative (they are in fact
expected bitwise, since host and device share the random stream and the
elementary functions and both round every operation)."""
import numpy as np
import pytest

from mcstreamer_amd import constants as co
from mcstreamer_amd import jumc, gases
from oracle import refcpu as R
from tests.common import gas_table, maxwell_states, ref_table

pytestmark = pytest.mark.gpu


def _run_pair(n, steps, o2, en_td, dt, seed, egrid="log", cap=None, kT=1.0):
    proc, energy, rate, maxrate = gas_table(egrid=egrid, o2=o2)
    cap = cap or 8 * n + 1000
    x, v, w, s, act = maxwell_states(n, kT_ev=kT, seed=seed)
    # oracle
    T = ref_table(proc, energy, rate, maxrate, egrid)
    P = R.Pop(x, v, w, s, act, cap=cap)
    E = en_td * co.Td * gases.N_GAS
    f = R.uniform_field((0.0, 0.0, -E))
    # device
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid=egrid)
    pop = jumc.Population(cap, jumc.ParticleStates(x, v, w, s, act, np.arange(n, dtype=np.uint64), n),
                          table, seed=seed)
    field = jumc.FixedField(-E)
    for step in range(steps):
        rs = R.advance(P, T, f, dt, None, seed=seed, step=step)
        gs = jumc.advance(pop, field, dt)
        print(f"step {step}: n {P.n} births {rs.births} deaths {rs.deaths}", flush=True)
        assert gs.births == rs.births
        assert gs.deaths == rs.deaths
        assert gs.flights == rs.flights
        assert list(gs.events)[: len(proc)] == list(rs.events)[: len(proc)]
    return P, pop


def _compare(P, pop, exact=True):
    ref = P.view()
    got = pop.download()
    assert len(got) == P.n
    np.testing.assert_array_equal(got.id, ref["id"])
    np.testing.assert_array_equal(got.active, ref["active"])
    for k in ("x", "v", "w", "s"):
        a, b = getattr(got, k), ref[k]
        if exact:
            np.testing.assert_array_equal(a, b, err_msg=k)
        else:
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=0, err_msg=k)


def test_advance_swarm_n2_parity():
    P, pop = _run_pair(n=3000, steps=5, o2=0.0, en_td=100.0, dt=1e-12, seed=11)
    _compare(P, pop)


def test_advance_attachment_and_births_parity():
    # strong field so that ionization births and attachment deaths both occur;
    # dead particles keep flying for the rest of the step (reference quirk)
    P, pop = _run_pair(n=2000, steps=4, o2=0.3, en_td=600.0, dt=2e-12, seed=5, kT=8.0)
    _compare(P, pop)
    assert (P.view()["active"] == 0).any()


def test_advance_uniform_grid_parity():
    P, pop = _run_pair(n=1500, steps=3, o2=0.2, en_td=300.0, dt=1e-12, seed=3, egrid="uniform")
    _compare(P, pop)


def test_advance_capacity_error():
    proc, energy, rate, maxrate = gas_table(o2=0.0)
    n = 256
    x, v, w, s, act = maxwell_states(n, kT_ev=60.0, seed=1)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    pop = jumc.Population(n, jumc.ParticleStates(x, v, w, s, act), table, seed=1)
    with pytest.raises(Exception, match="Maximum number of particles"):
        for _ in range(50):
            jumc.advance(pop, jumc.FixedField(-2e7), 1e-11)
