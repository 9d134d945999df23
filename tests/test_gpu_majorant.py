"""The banded majorant of the FP32 batched swarm step (mcs_swarm_batch_majorant).

The reference samples the next collision at the table's one maxrate
(collisions.jl:159-192, mixed_population.jl:35-44).  Any majorant at or above the
real collision rate gives the same process of real collisions (the null-collision
method is exact for every such majorant), so the banded form must:
- bound the real rate: each band's majorant is at least the summed rate of the
  non-null rows at every energy of the band, the rates being linear between
  nodes (checked here on the table itself and on a dense energy sample);
- give the same transport as the global form within the statistical errors
  (independent draws), at low, middle and high E/N with attachment;
- keep a slot's trajectory independent of the other slots.
"""
import numpy as np
import pytest

from mcstreamer_amd import constants as co
from mcstreamer_amd import gases, jumc
from mcstreamer_amd.swarm import SwarmBatch, tally_estimates
from tests.common import gas_table

pytestmark = pytest.mark.gpu

SEED = 31


def test_band_majorants_bound_the_real_rates():
    proc, energy, rate, maxrate = gas_table(o2=0.2)
    gt = jumc.CollisionTable(proc, energy, rate, maxrate, egrid="log")
    cap, maj = gt.majorants()
    kinds = np.array([p.code for p in proc])
    real = rate[kinds != 0].sum(axis=0)          # summed non-null rows per node
    assert np.isinf(cap[-1]) and maj[-1] == maxrate
    assert np.all(np.diff(cap[:-1]) > 0) and np.allclose(cap[1:-1] / cap[:-2], 2.0)
    assert cap[-2] == pytest.approx(energy[-1])
    assert np.all(np.diff(maj) >= 0) and np.all(maj <= maxrate)
    # dense sample of every band, rates linear between nodes
    e = np.geomspace(energy[0], energy[-1], 200_000)
    r = np.interp(e, energy, real)
    for k in range(len(cap) - 1):
        inside = e <= cap[k]
        if inside.any():
            assert r[inside].max() <= maj[k], k
        # and the majorant is tight: within the node's step of the band's real maximum
        nodes = energy <= cap[k]
        upto = min(int(nodes.sum()) + 1, len(energy))
        assert maj[k] == pytest.approx(min(real[:upto].max() * (1 + 1e-5), maxrate), rel=1e-12)
    # far below the global majorant at low energy: the point of the bands
    assert maj[0] < 0.2 * maxrate


def _transport(majorant, cases, steps, relax, block):
    tabs = {o2: gas_table(o2=o2) for o2 in sorted({c[1] for c in cases})}
    batch = SwarmBatch(total_capacity=100_000 * len(cases), max_slots=len(cases), seed=SEED, precision="f32",
                       majorant=majorant)
    gt = {o2: jumc.CollisionTable(*tabs[o2], egrid="log") for o2 in tabs}
    slots = []
    for k, (en, o2) in enumerate(cases):
        E = en * co.Td * gases.N_GAS
        slots.append(batch.add(gt[o2], 10_000, (0.0, 0.0, -E), 1e-12 * 100 / en, 10_000, capacity=80_000,
                               instance=70 + k, relax_steps=relax, block_steps=block))
    batch.step(steps)
    out = [tally_estimates(batch.tally(s)) for s in slots]
    assert all(batch.tally(s).flags == 0 for s in slots)
    return out


def test_banded_transport_matches_global_within_error_bars():
    cases = ((10.0, 0.2), (60.0, 0.3), (400.0, 0.1))
    res = {m: _transport(m, cases, steps=1400, relax=400, block=50) for m in ("global", "banded")}
    for k, (en, o2) in enumerate(cases):
        for name in ("drift_velocity", "mean_energy", "ionization_rate", "attachment_rate"):
            (a, ea), (b, eb) = res["global"][k][name], res["banded"][k][name]
            print(f"E/N {en:g} Td O2 {o2}: {name} global {a:.5e} (+-{ea:.1e}) banded {b:.5e} (+-{eb:.1e})",
                  flush=True)
            if not (np.isfinite(ea) and np.isfinite(eb)) or a == 0:
                continue
            assert abs(a - b) <= 5 * np.hypot(ea * abs(a), eb * abs(b)) + 2e-3 * abs(a), (en, name)


def test_banded_slot_independent_of_mix():
    # a slot's draws depend on its own key and particles only, with bands too
    tabs = {o2: gas_table(o2=o2) for o2 in (0.0, 0.2)}
    got = []
    for mix in (False, True):
        batch = SwarmBatch(total_capacity=60_000, max_slots=4, seed=SEED, precision="f32", majorant="banded")
        gt = {o2: jumc.CollisionTable(*tabs[o2], egrid="log") for o2 in tabs}
        if mix:
            batch.add(gt[0.0], 3000, (0.0, 0.0, -500 * co.Td * gases.N_GAS), 2e-13, 3000, capacity=20_000,
                      instance=5)
        s = batch.add(gt[0.2], 2000, (0.0, 0.0, -20 * co.Td * gases.N_GAS), 5e-12, 2000, capacity=20_000,
                      instance=6)
        batch.step(40)
        got.append(batch.download(s))
    a, b = got
    np.testing.assert_array_equal(a.id, b.id)
    np.testing.assert_array_equal(a.x, b.x)
    np.testing.assert_array_equal(a.v, b.v)


def test_banded_refused_for_fp64_batches():
    batch = SwarmBatch(total_capacity=1000, max_slots=1, seed=SEED, precision="f64")
    with pytest.raises(Exception, match="banded majorant"):
        batch.set_majorant("banded")


def test_banded_energy_distribution_matches_global():
    # the electron energy distribution after relaxation, body and tail: a
    # two-sample KS test on the energies, and the fractions above 2, 3 and 4 times
    # the mean within their binomial errors (the tail carries the rare processes)
    from mcstreamer_amd.swarm import ks_pvalue
    tabs = {0.2: gas_table(o2=0.2)}
    eng = {}
    for mode in ("global", "banded"):
        batch = SwarmBatch(total_capacity=120_000, max_slots=2, seed=SEED + (mode == "banded"), precision="f32",
                           majorant=mode)
        gt = jumc.CollisionTable(*tabs[0.2], egrid="log")
        s = batch.add(gt, 40_000, (0.0, 0.0, -25 * co.Td * gases.N_GAS), 4e-12, 40_000, capacity=100_000,
                      instance=81)
        batch.step(600)
        ps = batch.download(s)
        v = ps.v[ps.active.astype(bool)]
        eng[mode] = 0.5 * co.electron_mass * (v ** 2).sum(axis=1)
    a, b = eng["global"], eng["banded"]
    p = ks_pvalue(a, b)
    mean = 0.5 * (a.mean() + b.mean())
    print(f"KS p-value {p:.3f}; mean energy {a.mean() / co.eV:.4f} / {b.mean() / co.eV:.4f} eV", flush=True)
    assert p > 1e-3
    for f in (2.0, 3.0, 4.0):
        fa, fb = (a > f * mean).mean(), (b > f * mean).mean()
        err = np.sqrt(fa * (1 - fa) / len(a) + fb * (1 - fb) / len(b))
        print(f"above {f:g} x mean: {fa:.5f} / {fb:.5f} (+-{err:.5f})", flush=True)
        assert abs(fa - fb) <= 5 * err + 1e-4
