"""Host mirror of Swarm.jl (Swarm/src/Swarm.jl) on top of libmcs.

  swarm(en; ...) (Swarm.jl:48-186) -> swarm
  resample!(popl, n, tol) (:193-223) -> resample
  gains(proc) (:229-239), collrate (:242-248) -> gains, collrates
Per step: advance! with a FixedField, the position tallies (:126-130) from one
device reduction, and the global resample towards n0.  The KS-test early stop
(:132-134) is not reproduced: runs are fixed-length.
"""
import ctypes

import numpy as np

from . import _lib
from . import constants as co
from ._lib import check, dptr, u8ptr, u64ptr
from .jumc import FixedField, ParticleStates, advance, context


def resample(popl, n, tol=0.1, step=None):
    """Swarm.resample!: drop with probability (ncur-n)/ncur, or add Poisson copies."""
    did = ctypes.c_int32()
    s = popl.step if step is None else int(step)
    check(popl.ctx.handle, _lib.lib().mcs_swarm_resample(
        popl.handle, int(n), float(tol), popl.seed, popl.instance, s, ctypes.byref(did)))
    return bool(did.value)


def gains(proc):
    return np.array([-1 if p.kind == "ATTACHMENT" else (1 if p.kind == "IONIZATION" else 0)
                     for p in proc])


def swarm_step(popl, field, dt, n0, tol=0.1):
    """One iteration of the loop at Swarm.jl:124-142 (without the KS test)."""
    st = advance(popl, field, dt, step=popl.step, stats=True)
    resample(popl, n0, tol, step=popl.step)
    popl.step += 1
    return st


def swarm(popl, en_td, dt, nt, n0, n_gas, l=100, tol=0.1):
    """Run nt steps and return the transport coefficients of Swarm.jl:144-169:
    mobility, diffusion (OLS of sigma^2 over the last l steps), nu, alpha, eta
    and the per-process rates k (table order)."""
    eb = en_td * n_gas * co.Td
    field = FixedField(-eb)
    table = popl.collisions
    t = np.arange(nt) * dt
    sigma2 = np.zeros((nt, 3))
    last = nt
    for i in range(nt):
        advance(popl, field, dt, step=popl.step)
        m = popl.moments()
        n = m.nactive
        X = np.array(m.sum_x[:]) / n
        X2 = np.array(m.sum_x2[:]) / n
        sigma2[i] = X2 - X ** 2
        resample(popl, n0, tol, step=popl.step)
        popl.step += 1
        if popl.nparticles() > 0.9 * popl.capacity:
            last = i + 1
            break
    i = last
    A = np.stack([np.ones(l), t[i - l:i]], axis=1)
    coef, *_ = np.linalg.lstsq(A, sigma2[i - l:i], rcond=None)
    D = coef[1] / 2
    m = popl.moments(table)
    W = m.sum_v[2] / m.nactive
    k1 = np.array(m.collrate[: len(table.proc)]) / m.nactive
    g = gains(table.proc)
    return {
        "en": en_td,
        "mobility": -W / (-eb),
        "drift_velocity": W,
        "diffusion_t": (D[0] + D[1]) / 2,
        "diffusion_l": D[2],
        "diffusion": (D[0] + D[1] + D[2]) / 3,
        "nu": float(g @ k1),
        "alpha": float(np.maximum(g, 0) @ k1 / abs(W)),
        "eta": float(-np.minimum(g, 0) @ k1 / abs(W)),
        "k": k1,
        "mean_energy": m.sum_energy / m.nactive,
        "steps": i,
    }


TALLY_NAMES = ("drift_velocity", "mean_energy", "ionization_rate", "attachment_rate")


class SwarmBatch:
    """Many independent swarm instances (Swarm.swarm, Swarm.jl:102-142) advanced
    together on one device.

    Each instance ("slot") owns a slab of one particle store, its own collision
    table, uniform field, time step and resample target n0, and draws from the
    Philox key (seed, instance), so its trajectory is the one a lone
    Population(seed=seed, instance=instance) would follow under swarm_step.  A
    step of every slot is advance! + the per-step tallies (mean v_z, mean energy,
    mean ionization and attachment rate over the actives) + resample!; every
    `compact_every` steps each slab is compacted (stable, actives only).
    """

    def __init__(self, total_capacity, max_slots, seed=0, precision="f64", ctx=None):
        self.ctx = ctx or context()
        self.seed = int(seed)
        self.precision = precision
        prec = {"f64": _lib.PREC_F64, "f32": _lib.PREC_F32}[precision]
        h = ctypes.c_void_p()
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_create(
            self.ctx.handle, int(total_capacity), int(max_slots), self.seed, prec, ctypes.byref(h)))
        self.handle = h
        self.max_slots = int(max_slots)
        self._tables = {}

    def configure(self, block=256, compact_every=8):
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_config(self.handle, int(block), int(compact_every)))

    def autotune(self, candidates=((256, 4), (128, 4), (256, 8), (128, 8), (256, 16), (256, 2)),
                 min_steps=2):
        """Project-written autotuner of the launch shape: runs real steps under each
        (workgroup size, compaction cadence) candidate, one compaction per trial,
        and keeps the fastest per step.  The choice never changes results (a slot's
        trajectory is the same under every shape and cadence); the trial steps count
        as ordinary steps.  Returns {candidate: seconds per step}."""
        import time as _t
        timings = {}
        for blk, ce in candidates:
            self.configure(blk, ce)
            self.step(1)                      # capture the graphs for this shape
            self.ctx.synchronize()
            k = max(min_steps, ce)
            t0 = _t.perf_counter()
            self.step(k)
            self.ctx.synchronize()
            timings[(blk, ce)] = (_t.perf_counter() - t0) / k
        best = min(timings, key=timings.get)
        self.configure(*best)
        self.tuned = best
        return timings

    def add(self, table, particles, e, dt, n0, capacity, instance, tol=0.1, relax_steps=0,
            block_steps=100, init_kT_ev=1.0):
        """Admit one instance; returns its slot.  `particles` is a ParticleStates, or
        an int n for n electrons at the origin with a Maxwellian of temperature
        init_kT_ev drawn on the device.  Raises MCSError (capacity) when no slab of
        `capacity` particles is free."""
        prm = _lib.SwarmParams()
        for d in range(3):
            prm.e[d] = float(e[d])
        prm.dt = float(dt)
        prm.n0 = int(n0)
        prm.tol = float(tol)
        prm.capacity = int(capacity)
        prm.instance = int(instance)
        prm.relax_steps = int(relax_steps)
        prm.block_steps = int(block_steps)
        prm.init_vth = float(np.sqrt(init_kT_ev * co.eV / co.electron_mass))
        slot = ctypes.c_int32()
        if isinstance(particles, (int, np.integer)):
            check(self.ctx.handle, _lib.lib().mcs_swarm_batch_add(
                self.handle, table.handle, ctypes.byref(prm), int(particles), None, None, None, None,
                None, ctypes.byref(slot)))
        else:
            ps = particles
            check(self.ctx.handle, _lib.lib().mcs_swarm_batch_add(
                self.handle, table.handle, ctypes.byref(prm), len(ps), dptr(ps.x), dptr(ps.v),
                dptr(ps.w), dptr(ps.s), u8ptr(ps.active), ctypes.byref(slot)))
        self._tables[slot.value] = table
        return slot.value

    def remove(self, slot):
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_remove(self.handle, int(slot)))
        self._tables.pop(int(slot), None)

    def step(self, nsteps=1):
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_step(self.handle, int(nsteps)))

    def tally(self, slot):
        t = _lib.SwarmTally()
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_tally(self.handle, int(slot), ctypes.byref(t)))
        return t

    def tallies(self, nslots=None):
        n = self.max_slots if nslots is None else int(nslots)
        arr = (_lib.SwarmTally * max(n, 1))()
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_tally_all(self.handle, n, arr))
        return list(arr)[:n]

    def download(self, slot):
        t = self.tally(slot)
        n = int(t.n)
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        w = np.zeros(n)
        s = np.zeros(n)
        act = np.zeros(n, dtype=np.uint8)
        ids = np.zeros(n, dtype=np.uint64)
        nn = ctypes.c_int64()
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_download(
            self.handle, int(slot), n, ctypes.byref(nn), dptr(x), dptr(v), dptr(w), dptr(s),
            u8ptr(act), u64ptr(ids)))
        return ParticleStates(x, v, w, s, act, ids)

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().mcs_swarm_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def tally_estimates(t):
    """Time-averaged tallies of one slot and their relative statistical errors
    (standard error of the block means over |mean|; inf with fewer than 2 blocks)."""
    out = {}
    nrec = max(int(t.nrecorded), 1)
    nb = int(t.nblocks)
    for q, name in enumerate(TALLY_NAMES):
        mean = t.sum[q] / nrec
        if nb >= 2:
            m1 = t.bm_s1[q] / nb
            var = max(t.bm_s2[q] / nb - m1 * m1, 0.0) * nb / (nb - 1)
            err = np.sqrt(var / nb) / abs(m1) if m1 != 0 else np.inf
        else:
            err = np.inf
        out[name] = (mean, err)
    return out
