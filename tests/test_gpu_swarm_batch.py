"""GPU parity of the batched swarm path (many Swarm.swarm instances,
Swarm/src/Swarm.jl:102-142, in one set of launches) against the oracle.

Each slot must follow exactly the trajectory of a lone oracle population keyed
(seed, instance): advance! (mixed_population.jl:24-54) then resample!
(Swarm.jl:193-223) per step.  Slabs are compacted every few steps, which only
drops inactive particles and keeps order, so the comparison is over the active
particles in slab order.  Bar: ids and states bit-exact; per-step tallies
(mean v_z, energy, ionization and attachment rate over the actives) to 1e-12
relative (the device sums in another order, with compensation)."""
import numpy as np
import pytest

from mcstreamer_amd import constants as co
from mcstreamer_amd import gases, jumc
from mcstreamer_amd.swarm import SwarmBatch, tally_estimates
from oracle import refcpu as R
from tests.common import gas_table, maxwell_states, ref_table

pytestmark = pytest.mark.gpu

SEED = 2024


def _instances():
    # uneven counts, E/N, gas mixes and time steps, as in the batched workload
    return [
        dict(n=700, n0=700, en=100.0, o2=0.0, dt=1e-12, inst=7, kT=1.0),
        dict(n=1500, n0=1200, en=600.0, o2=0.3, dt=1e-12 * 100 / 600, inst=123456, kT=6.0),
        dict(n=300, n0=500, en=30.0, o2=0.2, dt=1e-12 * 100 / 30, inst=99, kT=0.5),
        dict(n=64, n0=64, en=1000.0, o2=0.0, dt=1e-13, inst=3, kT=20.0),
    ]


def _tables():
    out = {}
    for o2 in (0.0, 0.2, 0.3):
        proc, energy, rate, maxrate = gas_table(o2=o2)
        out[o2] = (proc, energy, rate, maxrate)
    return out


def _oracle_means(m, proc, nact):
    kinds = np.array([p.code for p in proc])
    cr = np.array(m.collrate[: len(proc)])
    ion = cr[kinds == 3].sum()
    att = cr[kinds == 4].sum()
    return np.array([m.sum_v[2], m.sum_energy, ion, att]) / nact


def _run(block, compact_every, steps=24, relax=4, bsteps=5, only=None):
    tabs = _tables()
    insts = _instances() if only is None else [_instances()[k] for k in only]
    batch = SwarmBatch(total_capacity=200_000, max_slots=16, seed=SEED)
    batch.configure(block=block, compact_every=compact_every)
    gtabs = {o2: jumc.CollisionTable(*tabs[o2], egrid="log") for o2 in tabs}
    slots, oracles = [], []
    for k, d in enumerate(insts):
        x, v, w, s, act = maxwell_states(d["n"], kT_ev=d["kT"], seed=100 + d["inst"])
        E = d["en"] * co.Td * gases.N_GAS
        slot = batch.add(gtabs[d["o2"]], jumc.ParticleStates(x, v, w, s, act), (0.0, 0.0, -E), d["dt"],
                         d["n0"], capacity=30_000, instance=d["inst"], relax_steps=relax,
                         block_steps=bsteps)
        slots.append(slot)
        P = R.Pop(x, v, w, s, act, cap=60_000)
        T = ref_table(*tabs[d["o2"]], "log")
        oracles.append((P, T, R.uniform_field((0.0, 0.0, -E)), tabs[d["o2"]][0]))
    batch.step(steps)
    for k, d in enumerate(insts):
        P, T, f, proc = oracles[k]
        means, sig = [], []
        for st in range(steps):
            R.advance(P, T, f, d["dt"], None, seed=SEED, instance=d["inst"], step=st)
            m = R.moments(P, T)
            means.append(_oracle_means(m, proc, m.nactive))
            X = np.array(m.sum_x[:]) / m.nactive
            sig.append(np.array(m.sum_x2[:]) / m.nactive - X ** 2)
            R.swarm_resample(P, d["n0"], 0.1, seed=SEED, instance=d["inst"], step=st)
        d = dict(d, sigma2=np.array(sig))
        yield batch, slots[k], P, np.array(means), d


def _check(batch, slot, P, means, relax, bsteps, steps):
    t = batch.tally(slot)
    assert t.flags == 0
    assert t.steps == steps
    ref = P.view()
    a = ref["active"] == 1
    got = batch.download(slot)
    g = got.active == 1
    print(f"slot {slot}: oracle actives {a.sum()} device actives {g.sum()} n {t.n}", flush=True)
    np.testing.assert_array_equal(got.id[g], ref["id"][a])
    for key in ("x", "v", "w", "s"):
        np.testing.assert_array_equal(getattr(got, key)[g], ref[key][a], err_msg=key)
    np.testing.assert_allclose(np.array(t.last), means[-1], rtol=1e-12, atol=0)
    rec = means[relax:]
    assert t.nrecorded == len(rec)
    np.testing.assert_allclose(np.array(t.sum), rec.sum(axis=0), rtol=1e-12)
    nb = len(rec) // bsteps
    assert t.nblocks == nb
    bm = rec[: nb * bsteps].reshape(nb, bsteps, 4).mean(axis=1)
    np.testing.assert_allclose(np.array(t.bm_s1), bm.sum(axis=0), rtol=1e-12)
    np.testing.assert_allclose(np.array(t.bm_s2), (bm ** 2).sum(axis=0), rtol=1e-12)


def _check_diffusion(t, d, steps):
    """D = slope / 2 of numpy's least-squares line through the oracle's sigma^2 of
    the last min(steps, 100) steps (Swarm.jl:144-150)."""
    m = min(steps, 100)
    tt = np.arange(steps - m, steps) * d["dt"]
    want = np.array([np.polyfit(tt, d["sigma2"][steps - m:, q], 1)[0] / 2 for q in range(3)])
    assert t.window == m
    np.testing.assert_allclose(np.array(t.diffusion), want, rtol=1e-7, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("block,compact_every", [(256, 4), (128, 1), (256, 0)])
def test_batch_matches_lone_oracle(block, compact_every):
    steps, relax, bsteps = 24, 4, 5
    for batch, slot, P, means, d in _run(block, compact_every, steps, relax, bsteps):
        _check(batch, slot, P, means, relax, bsteps, steps)
        _check_diffusion(batch.tally(slot), d, steps)


def test_batch_diffusion_window_wraps():
    # 150 steps: the fit covers the last 100 of them, the ring has wrapped
    steps = 150
    for batch, slot, P, means, d in _run(256, 4, steps, 4, 5, only=[0, 2]):
        t = batch.tally(slot)
        _check_diffusion(t, d, steps)
        print(f"slot {slot}: D {np.array(t.diffusion)}", flush=True)
        assert np.all(np.array(t.diffusion) > 0)


def test_batch_slot_independent_of_mix():
    # the same instance alone and in the mix gives identical particles and tallies
    steps = 16
    alone = list(_run(256, 4, steps=steps, only=[1]))[0]
    mixed = list(_run(256, 4, steps=steps))[1]
    ta, tm = alone[0].tally(alone[1]), mixed[0].tally(mixed[1])
    for f in ("steps", "n", "nactive", "nrecorded", "nblocks"):
        assert getattr(ta, f) == getattr(tm, f)
    assert list(ta.sum) == list(tm.sum)
    ga, gm = alone[0].download(alone[1]), mixed[0].download(mixed[1])
    np.testing.assert_array_equal(ga.id, gm.id)
    np.testing.assert_array_equal(ga.x, gm.x)
    np.testing.assert_array_equal(ga.v, gm.v)


def test_batch_add_remove_between_steps():
    # slots admitted and retired between steps: a later instance lands in a freed
    # slab and still follows its own oracle trajectory from its own step 0
    tabs = _tables()
    batch = SwarmBatch(total_capacity=70_000, max_slots=4, seed=SEED)
    gt = jumc.CollisionTable(*tabs[0.2], egrid="log")
    T = ref_table(*tabs[0.2], "log")
    E = 200.0 * co.Td * gases.N_GAS
    f = R.uniform_field((0.0, 0.0, -E))
    dt = 5e-13

    def mk(inst, n):
        x, v, w, s, act = maxwell_states(n, kT_ev=2.0, seed=inst)
        return jumc.ParticleStates(x, v, w, s, act), R.Pop(x, v, w, s, act, cap=60_000)

    pa, Pa = mk(11, 800)
    pb, Pb = mk(12, 900)
    sa = batch.add(gt, pa, (0, 0, -E), dt, 800, capacity=30_000, instance=11)
    sb = batch.add(gt, pb, (0, 0, -E), dt, 900, capacity=30_000, instance=12)
    batch.step(5)
    batch.remove(sa)
    pc, Pc = mk(13, 700)
    sc = batch.add(gt, pc, (0, 0, -E), dt, 700, capacity=30_000, instance=13)
    batch.step(7)
    for P, inst, n0, nst, slot in ((Pb, 12, 900, 12, sb), (Pc, 13, 700, 7, sc)):
        for st in range(nst):
            R.advance(P, T, f, dt, None, seed=SEED, instance=inst, step=st)
            R.swarm_resample(P, n0, 0.1, seed=SEED, instance=inst, step=st)
        ref = P.view()
        a = ref["active"] == 1
        got = batch.download(slot)
        g = got.active == 1
        np.testing.assert_array_equal(got.id[g], ref["id"][a])
        np.testing.assert_array_equal(got.x[g], ref["x"][a])
        np.testing.assert_array_equal(got.v[g], ref["v"][a])
    # a fourth instance that does not fit is refused with a capacity error
    pd, _ = mk(14, 100)
    with pytest.raises(Exception, match="no free slab"):
        batch.add(gt, pd, (0, 0, -E), dt, 100, capacity=40_000, instance=14)
    est = tally_estimates(batch.tally(sb))
    assert set(est) == {"drift_velocity", "mean_energy", "ionization_rate", "attachment_rate", "diffusion_l",
                        "diffusion_t", "diffusion"}


def test_device_thermal_init_matches_oracle():
    # the x == NULL form of add and mcs_pop_init_thermal draw exactly the oracle's
    # initial electrons (phase INIT of the instance key)
    tabs = _tables()
    gt = jumc.CollisionTable(*tabs[0.0], egrid="log")
    kT = 1.0
    vth = np.sqrt(kT * co.eV / co.electron_mass)
    batch = SwarmBatch(total_capacity=50_000, max_slots=2, seed=SEED)
    slot = batch.add(gt, 3000, (0, 0, -1e6), 1e-12, 3000, capacity=10_000, instance=41, init_kT_ev=kT)
    got = batch.download(slot)
    ref = R.swarm_init(3000, SEED, 41, vth).view()
    for key in ("x", "v", "w", "s", "active", "id"):
        np.testing.assert_array_equal(getattr(got, key), ref[key], err_msg=key)
    pop = jumc.Population(5000, None, gt, seed=SEED, instance=41)
    pop.init_thermal(3000, kT)
    p = pop.download()
    np.testing.assert_array_equal(p.v, ref["v"])
    np.testing.assert_array_equal(p.s, ref["s"])
    assert p.next_id == 3000


def _constant_nu_table(nu, ne=64, emax_ev=1e5):
    from mcstreamer_amd.lxcat import Process
    e = np.linspace(0.0, emax_ev, ne) * co.eV
    return jumc.CollisionTable([Process("ELASTIC", 0.0)], e, np.full((1, ne), nu), nu, egrid="uniform")


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_constant_frequency_drift_closed_form(precision):
    # elastic scattering at constant nu, zero mass ratio: <v_z> -> qE/(m nu),
    # here time-averaged over steps after relaxation, in both store precisions
    nu, E, dt = 2e12, 2e5, 1e-12
    t = _constant_nu_table(nu)
    batch = SwarmBatch(total_capacity=200_000, max_slots=4, seed=SEED, precision=precision)
    n = 20000
    slots = [batch.add(t, n, (0.0, 0.0, E), dt, n, capacity=40_000, instance=k, relax_steps=20,
                       block_steps=20, init_kT_ev=0.1) for k in range(2)]
    batch.step(220)
    W = -co.elementary_charge * E / (co.electron_mass * nu)
    for slot in slots:
        est = tally_estimates(batch.tally(slot))
        mean, err = est["drift_velocity"]
        print(f"{precision} slot {slot}: W {mean:.6e} analytic {W:.6e} rel err {err:.2e}", flush=True)
        assert batch.tally(slot).flags == 0
        assert abs(mean - W) < 5 * err * abs(W) + 2e-3 * abs(W)


def test_f32_transport_matches_f64_within_error_bars():
    # the split-precision store gives the same transport coefficients as the
    # FP64 one within the statistical errors (independent random streams)
    tabs = _tables()
    res = {}
    for prec in ("f64", "f32"):
        batch = SwarmBatch(total_capacity=400_000, max_slots=4, seed=SEED, precision=prec)
        gt = {o2: jumc.CollisionTable(*tabs[o2], egrid="log") for o2 in (0.0, 0.3)}
        slots = []
        for k, (en, o2) in enumerate(((100.0, 0.0), (500.0, 0.3))):
            E = en * co.Td * gases.N_GAS
            slots.append(batch.add(gt[o2], 8000, (0.0, 0.0, -E), 1e-12 * 100 / en, 8000, capacity=60_000,
                                   instance=50 + k, relax_steps=300, block_steps=50))
        batch.step(1300)
        res[prec] = [tally_estimates(batch.tally(s)) for s in slots]
        assert all(batch.tally(s).flags == 0 for s in slots)
    for k in range(2):
        for name in ("drift_velocity", "mean_energy", "ionization_rate"):
            (a, ea), (b, eb) = res["f64"][k][name], res["f32"][k][name]
            print(f"instance {k} {name}: f64 {a:.5e} (+-{ea:.1e}) f32 {b:.5e} (+-{eb:.1e})", flush=True)
            if not np.isfinite(ea) or a == 0:
                continue
            assert abs(a - b) <= 5 * np.hypot(ea * abs(a), eb * abs(b)) + 2e-3 * abs(a), name
