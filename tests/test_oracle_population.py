"""Pins of the population bookkeeping: repack! (JuMC/src/population.jl:191-215)
against the rank-matching formulation the GPU uses and against brute force,
the reductions (population.jl:114-158, Swarm.jl:126-156) against numpy, and
Swarm.resample! (Swarm/src/Swarm.jl:193-223) against its expected counts."""
import numpy as np
import pytest

from mcstreamer_amd import constants as co
from oracle import refcpu as R
from tests.common import gas_table, ref_table


def _pop(n, p_active, seed, cap=None):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3)) * 1e6
    act = (rng.uniform(size=n) < p_active).astype(np.uint8)
    return R.Pop(x, v, rng.uniform(1, 3, n), rng.uniform(size=n), act, np.arange(n, dtype=np.uint64) * 7,
                 cap=cap)


def _two_pointer(active):
    """Literal restatement of population.jl:191-215 on an index list."""
    idx = list(range(len(active)))
    act = list(active)
    l = len(act) - 1
    while l >= 0 and not act[l]:
        l -= 1
    i = 0
    while i <= l:
        if not act[i]:
            idx[i], act[i] = idx[l], act[l]
            l -= 1
            while l >= 0 and not act[l]:
                l -= 1
        i += 1
    return idx[: l + 1]


def _rank_match(active):
    active = np.asarray(active, dtype=bool)
    A = int(active.sum())
    holes = np.flatnonzero(~active[:A])
    srcs = np.flatnonzero(active[A:])[::-1] + A
    out = np.arange(A)
    out[holes] = srcs
    return out.tolist()


@pytest.mark.parametrize("seed", range(20))
def test_repack_matches_rank_matching_and_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 300))
    act = (rng.uniform(size=n) < rng.uniform(0.05, 0.95)).astype(np.uint8)
    assert _two_pointer(act) == _rank_match(act)
    P = R.Pop(np.zeros((n, 3)), np.zeros((n, 3)), np.arange(n, dtype=float), np.zeros(n), act)
    R.repack(P)
    v = P.view()
    assert P.n == int(act.sum())
    np.testing.assert_array_equal(v["w"], np.array(_rank_match(act), dtype=float))
    assert np.all(v["active"] == 1)


def test_moments_against_numpy():
    proc, energy, rate, maxrate = gas_table(ne=500)
    T = ref_table(proc, energy, rate, maxrate)
    P = _pop(5000, 0.7, 1)
    m = R.moments(P, T)
    v = P.view()
    a = v["active"] == 1
    eng = 0.5 * co.electron_mass * (v["v"] ** 2).sum(axis=1)
    assert m.nactive == a.sum() and m.nparticles == 5000
    assert m.weight == pytest.approx(v["w"][a].sum(), rel=1e-13)
    assert m.wenergy == pytest.approx((v["w"] * eng)[a].sum(), rel=1e-13)
    assert m.maxenergy == pytest.approx(eng.max(), rel=1e-15)  # over all slots (population.jl:157)
    np.testing.assert_allclose(m.sum_x[:], v["x"][a].sum(axis=0), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(m.sum_x2[:], (v["x"][a] ** 2).sum(axis=0), rtol=1e-12)
    np.testing.assert_allclose(m.sum_v[:], v["v"][a].sum(axis=0), rtol=1e-9, atol=1e-3)
    e_grid = energy
    cr = np.zeros(len(proc))
    for e in eng[a]:
        k = min(max(np.searchsorted(e_grid, e, side="right") - 1, 0), e_grid.size - 2)
        w = (e_grid[k + 1] - e) / (e_grid[k + 1] - e_grid[k])
        cr += w * rate[:, k] + (1 - w) * rate[:, k + 1]
    np.testing.assert_allclose(m.collrate[: len(proc)], cr, rtol=1e-11)


def test_swarm_resample_drop_and_clone():
    P = _pop(20000, 1.0, 2)
    rc, did = R.swarm_resample(P, 10000, 0.1, seed=3, step=0)
    assert did == 1 and rc == 0
    kept = int(P.view()["active"].sum())
    assert abs(kept - 10000) < 5 * np.sqrt(20000 * 0.25)
    P = _pop(8000, 1.0, 3, cap=20000)
    rc, did = R.swarm_resample(P, 10000, 0.1, seed=3, step=1)
    assert did == 1
    added = P.n - 8000
    assert abs(added - 2000) < 5 * np.sqrt(2000)
    # copies keep the parent's state and get new ids
    v = P.view()
    assert len(set(v["id"].tolist())) == P.n
    P = _pop(10500, 1.0, 4)
    rc, did = R.swarm_resample(P, 10000, 0.1)
    assert did == 0 and P.n == 10500


def test_swarm_init_draw_order_and_distribution():
    # the first particles against Box-Muller written with numpy's own log/sincos on
    # the same uniforms (phase INIT = 3, step 0): pins the draw order and formula
    from scipy import stats
    vth = np.sqrt(2.0 * co.eV / co.electron_mass)
    seed, inst = 77, 5
    P = R.swarm_init(50_000, seed, inst, vth)
    v = P.view()["v"]
    s = P.view()["s"]
    for pid in (0, 1, 4999):
        u = R.stream_u01(seed, inst, 3, pid, 0, 5)
        r1 = np.sqrt(-2.0 * np.log1p(-u[0]))
        r2 = np.sqrt(-2.0 * np.log1p(-u[2]))
        want = vth * np.array([r1 * np.cos(2 * np.pi * u[1]), r1 * np.sin(2 * np.pi * u[1]),
                               r2 * np.cos(2 * np.pi * u[3])])
        np.testing.assert_allclose(v[pid], want, rtol=1e-12, atol=1e-12 * vth)
        np.testing.assert_allclose(s[pid], -np.log1p(-u[4]), rtol=1e-13)
    assert P.n == 50_000 and P.next_id == 50_000
    np.testing.assert_array_equal(P.view()["id"], np.arange(50_000, dtype=np.uint64))
    assert (P.view()["x"] == 0).all() and (P.view()["w"] == 1).all()
    # Maxwellian: each component N(0, vth^2), s ~ Exp(1); mean energy 3/2 kT
    for d in range(3):
        assert stats.kstest(v[:, d] / vth, "norm").pvalue > 1e-3
    assert stats.kstest(s, "expon").pvalue > 1e-3
    e = 0.5 * co.electron_mass * (v ** 2).sum(axis=1) / co.eV
    assert abs(e.mean() - 3.0) < 5 * np.sqrt(6.0 / 50_000)   # var of (kT/2) chi2_3 = 1.5 (kT)^2
