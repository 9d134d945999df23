"""Swarm.swarm for one E/N on the device (Swarm/src/Swarm.jl:124-169) against the
oracle's chain (advance!, the position moments, Swarm.resample! per step):

- sigma^2(t) of every step within 1e-10 (the device's compensated sums differ
  from the oracle's sequential ones by rounding), the diffusion coefficients as
  numpy's polyfit of the oracle's sigma^2 over the last l steps, mobility, nu,
  alpha, eta and the rates k from the oracle's collision-rate sums;
- the Kolmogorov-Smirnov stop (Swarm.jl:274-306): device energies bitwise those
  of the downloaded velocities, the stop taken at the first comparison when
  pval is below every p-value, never when it is above, and the p-value of the
  first comparison equal to ks_pvalue of the oracle's energies;
- swarm_scan / main / stats end to end on two E/N values and two replicas."""
import numpy as np
import pytest

from mcstreamer_amd import constants as co
from mcstreamer_amd import gases, jumc
from mcstreamer_amd import swarm as S
from oracle import refcpu as R
from tests.common import gas_table, maxwell_states, ref_table

pytestmark = pytest.mark.gpu

EN, O2, DT, N0 = 300.0, 0.2, 1e-12, 2000


def _setup(seed=21, n=N0, cap=600_000):
    proc, energy, rate, maxrate = gas_table(o2=O2)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    x, v, w, s, act = maxwell_states(n, kT_ev=1.0, seed=seed)
    popl = jumc.Population(cap, jumc.ParticleStates(x, v, w, s, act), table, seed=seed)
    P = R.Pop(x, v, w, s, act, cap=cap)
    T = ref_table(proc, energy, rate, maxrate)
    return proc, table, popl, P, T


def test_swarm_coefficients_match_oracle():
    nt, l = 300, 100
    proc, table, popl, P, T = _setup()
    r = S.swarm(popl, EN, DT, nt, N0, gases.N_GAS, l=l, ntest=0)
    assert r["steps"] == nt and r["stopped"] is None
    eb = EN * gases.N_GAS * co.Td
    f = R.uniform_field((0.0, 0.0, -eb))
    sig = np.zeros((nt, 3))
    for st in range(nt):
        R.advance(P, T, f, DT, None, seed=popl.seed, step=st)
        m = R.moments(P, T)
        X = np.array(m.sum_x[:]) / m.nactive
        X2 = np.array(m.sum_x2[:]) / m.nactive
        sig[st] = X2 - X ** 2
        R.swarm_resample(P, N0, 0.1, seed=popl.seed, step=st)
    np.testing.assert_allclose(r["sigma2"], sig, rtol=1e-10)
    t = np.arange(nt) * DT
    slope = np.array([np.polyfit(t[nt - l:], sig[nt - l:, d], 1)[0] for d in range(3)])
    D = slope / 2
    np.testing.assert_allclose(r["diffusion_l"], D[2], rtol=1e-8)
    np.testing.assert_allclose(r["diffusion_t"], (D[0] + D[1]) / 2, rtol=1e-8)
    m = R.moments(P, T)
    W = m.sum_v[2] / m.nactive
    k = np.array(m.collrate[: len(proc)]) / m.nactive
    g = S.gains(proc)
    np.testing.assert_allclose(r["drift_velocity"], W, rtol=1e-12)
    np.testing.assert_allclose(r["mobility"], W / eb, rtol=1e-12)
    np.testing.assert_allclose(r["k"], k, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(r["nu"], g @ k, rtol=1e-10)
    np.testing.assert_allclose(r["alpha"], np.maximum(g, 0) @ k / abs(W), rtol=1e-12)
    np.testing.assert_allclose(r["eta"], -np.minimum(g, 0) @ k / abs(W), rtol=1e-12)
    print(f"W {W:.4e} D_l {D[2]:.4e} D_t {(D[0] + D[1]) / 2:.4e} alpha {r['alpha']:.4e} eta {r['eta']:.4e}",
          flush=True)
    assert r["alpha"] > 0 and r["eta"] > 0 and D[2] > 0


def test_energies_match_downloaded_velocities():
    proc, table, popl, P, T = _setup(seed=5, n=3000)
    jumc.advance(popl, jumc.FixedField(-3e6), 1e-12)
    ps = popl.download()
    a = ps.active == 1
    v = ps.v[a]
    want = (0.5 * co.electron_mass) * ((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    got = popl.energies()
    np.testing.assert_allclose(got, want, rtol=1e-15)
    np.testing.assert_allclose(got, 0.5 * co.electron_mass * (v ** 2).sum(axis=1), rtol=1e-14)


def test_ks_stop_control_flow_and_pvalue():
    nt = 400
    # pval below every p-value: the first comparison (step 2 ntest) stops the run
    proc, table, popl, P, T = _setup(seed=31)
    r = S.swarm(popl, EN, DT, nt, N0, gases.N_GAS, l=50, ntest=50, pval=-1.0)
    assert r["stopped"] == "ks" and r["steps"] == 100
    # above every p-value: it never stops
    proc, table, popl, P, T = _setup(seed=31)
    r = S.swarm(popl, EN, DT, nt, N0, gases.N_GAS, l=50, ntest=50, pval=2.0)
    assert r["stopped"] is None and r["steps"] == nt
    # the first comparison's p-value from the device equals the oracle's
    proc, table, popl, P, T = _setup(seed=33)
    eb = EN * gases.N_GAS * co.Td
    f = R.uniform_field((0.0, 0.0, -eb))
    kst = S.KSTester()
    samples = []
    for st in range(100):
        jumc.advance(popl, jumc.FixedField(-eb), DT, step=popl.step)
        R.advance(P, T, f, DT, None, seed=popl.seed, step=st)
        if (st + 1) % 50 == 0:
            kst.test(popl, 2.0)
            ref = P.view()
            v = ref["v"][ref["active"] == 1]
            samples.append((0.5 * co.electron_mass) * ((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]))
        S.resample(popl, N0, 0.1, step=popl.step)
        R.swarm_resample(P, N0, 0.1, seed=popl.seed, step=st)
        popl.step += 1
    np.testing.assert_array_equal(kst.prev, samples[1])
    p = S.ks_pvalue(samples[1], samples[0])
    print(f"KS p-value between steps 50 and 100: {p:.4f}", flush=True)
    assert 0.0 <= p <= 1.0


def test_scan_main_and_stats_end_to_end():
    proc, energy, rate, maxrate = gas_table(o2=O2)
    table = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    st = S.main(table, gases.N_GAS, nsamples=2, en=[100.0, 400.0], n0=500, trelax=1e-11, maxp=20_000,
                ntest=0, l=20)
    for name in S.COEFFICIENTS:
        assert st[name].shape == (2,) and np.all(np.isfinite(st[name])), name
        assert np.all(st[name + "_err"] >= 0), name
    assert st["k"].shape == (2, len(proc))
    assert st["mobility"][0] > 0 and st["alpha"][1] > st["alpha"][0]
