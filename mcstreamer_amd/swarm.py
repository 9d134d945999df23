"""Host mirror of Swarm.jl (Swarm/src/Swarm.jl) on top of libmcs.

  main(; nsamples, en) (Swarm.jl:34-45)   -> main
  swarm(en; ...) (:48-186)               -> swarm_scan (over E/N), swarm (one E/N)
  resample!(popl, n, tol) (:193-223)     -> resample
  gains(proc) (:229-239), collrate (:242-248) -> gains, the moments' collrate sums
  stats(dfs) (:253-269)                  -> stats
  KSTester, kstest (:274-306)            -> KSTester, ks_statistic, ks_pvalue
Per step: advance! with a FixedField, the position tallies (:126-130) from one
device reduction, the KS stop every ntest steps on the device-computed energies,
and the global resample towards n0.
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


def ks_statistic(x, y):
    """Two-sample Kolmogorov-Smirnov statistics (delta, delta+, delta-) as the
    reference's test computes them (HypothesisTests' ksstats, called from
    Swarm.jl:297): the pooled sample in stable sorted order, the running sum of
    +1/n_x for points of x and -1/n_y for points of y, its maximum and minus
    its minimum.  (With tied values this evaluates the difference inside a tie
    as well, like the reference.)"""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    nx, ny = len(x), len(y)
    order = np.argsort(np.concatenate([x, y]), kind="stable")
    steps = np.concatenate([np.full(nx, 1.0 / nx), np.full(ny, -1.0 / ny)])[order]
    cdf = np.cumsum(steps)
    dp = float(cdf.max())
    dn = float(-cdf.min())
    return max(dp, dn), dp, dn


def kolmogorov_sf(t):
    """P(K > t) of the Kolmogorov distribution (the limit law of sqrt(n) D_n):
    for t > 1 the alternating series 2 sum_{k>=1} (-1)^(k-1) exp(-2 k^2 t^2), for
    t <= 1 one minus the theta-transformed series
    sqrt(2 pi)/t sum_{k>=1} exp(-(2k-1)^2 pi^2 / (8 t^2)) (Kolmogorov 1933)."""
    t = float(t)
    if t <= 0.0:
        return 1.0
    if t <= 1.0:
        acc = 0.0
        for k in range(1, 60):
            term = np.exp(-((2 * k - 1) ** 2) * np.pi ** 2 / (8.0 * t * t))
            acc += term
            if term < 1e-300:
                break
        return 1.0 - np.sqrt(2.0 * np.pi) / t * acc
    acc = 0.0
    for k in range(1, 200):
        term = np.exp(-2.0 * k * k * t * t)
        acc += term if k % 2 else -term
        if term < 1e-300:
            break
    return 2.0 * acc


def ks_pvalue(x, y):
    """Two-sided p-value of the approximate two-sample KS test
    (ApproximateTwoSampleKSTest): P(K > sqrt(n_x n_y / (n_x + n_y)) delta)."""
    d, _, _ = ks_statistic(x, y)
    n = len(x) * len(y) / (len(x) + len(y))
    return kolmogorov_sf(np.sqrt(n) * d)


class KSTester:
    """The early stop of Swarm.jl:274-306: each test compares the energies of the
    active particles with those of the previous test; the run stops when the two
    samples cannot be told apart (p-value above pval)."""

    def __init__(self):
        self.prev = None

    def test(self, popl, pval):
        eng = popl.energies()
        stop = False
        if self.prev is not None and len(self.prev) > 0 and len(eng) > 0:
            if ks_pvalue(eng, self.prev) > pval:
                stop = True
        self.prev = eng
        return stop


def swarm(popl, en_td, dt, nt, n0, n_gas, l=100, tol=0.1, ntest=1000, pval=0.75, resampling=True):
    """One E/N of Swarm.swarm (Swarm.jl:124-169): up to nt steps of advance!, the
    position tallies, the KS stop every ntest steps, resample! towards n0 and the
    capacity stop; then the transport coefficients: mobility, diffusion (OLS of
    sigma^2 over the last l steps), nu, alpha, eta and the per-process rates k
    (table order).  The field points along -z (electrons drift along +z)."""
    eb = en_td * n_gas * co.Td
    field = FixedField(-eb)
    table = popl.collisions
    t = np.arange(nt) * dt
    sigma2 = np.zeros((nt, 3))
    kst = KSTester()
    stopped = None
    i = 0
    for i in range(1, nt + 1):
        advance(popl, field, dt, step=popl.step)
        m = popl.moments()
        n = m.nactive
        X = np.array(m.sum_x[:]) / n
        X2 = np.array(m.sum_x2[:]) / n
        sigma2[i - 1] = X2 - X ** 2
        if ntest and i % ntest == 0 and kst.test(popl, pval):
            stopped = "ks"
            break
        if resampling:
            resample(popl, n0, tol, step=popl.step)
        popl.step += 1
        if popl.nparticles() > 0.9 * popl.capacity:
            stopped = "capacity"
            break
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
        "stopped": stopped,
        "sigma2": sigma2[:i].copy(),
    }


COEFFICIENTS = ("mobility", "diffusion_l", "diffusion_t", "diffusion", "nu", "alpha", "eta")


def swarm_scan(en, table, n_gas, n0=10_000, trelax=2e-9, maxp=10 ** 7, ntest=1000, pval=0.75,
               resampling=True, seed=0, instance0=0, l=100, ctx=None):
    """Swarm.swarm over a list of E/N values (Swarm.jl:48-186): dt = 1/maxrate/4,
    nt = ceil(trelax/dt); for each E/N a fresh population of n0 electrons at rest
    at the origin (its own Philox instance: instance0 + index), run by `swarm`.
    Returns the table's columns: en, the coefficients and the rates k (one array
    per column, k as an (len(en), nproc) array)."""
    from .jumc import Population

    dt = 1.0 / table.maxrate / 4.0
    nt = int(np.ceil(trelax / dt))
    cols = {name: [] for name in ("en",) + COEFFICIENTS}
    cols["k"] = []
    for j, e in enumerate(en):
        popl = Population(maxp, None, table, seed=seed, instance=instance0 + j, ctx=ctx)
        popl.init_thermal(n0, 0.0)
        r = swarm(popl, float(e), dt, nt, n0, n_gas, l=l, ntest=ntest, pval=pval, resampling=resampling)
        for name in ("en",) + COEFFICIENTS:
            cols[name].append(r[name])
        cols["k"].append(r["k"])
    return {k: np.array(v) for k, v in cols.items()}


def stats(dfs):
    """Swarm.stats (Swarm.jl:253-269): per column the mean over the replicas and,
    except for en, the standard error std/sqrt(nsamples) (std with the n - 1
    divisor) as name + "_err"."""
    names = list(dfs[0])
    if any(list(d) != names for d in dfs):
        raise ValueError("replicas with different columns")
    out = {}
    n = len(dfs)
    for name in names:
        stack = np.array([np.asarray(d[name], dtype=np.float64) for d in dfs])
        out[name] = stack.mean(axis=0)
        if name != "en":
            out[name + "_err"] = stack.std(axis=0, ddof=1) / np.sqrt(n)
    return out


def main(table, n_gas, nsamples=10, en=None, **kw):
    """Swarm.main (Swarm.jl:34-45): nsamples independent replicas of swarm_scan
    (replica r uses Philox instances r * len(en) + j) and their statistics."""
    en = 10 ** np.linspace(-1, 3, 100) if en is None else np.asarray(en, dtype=np.float64)
    dfs = [swarm_scan(en, table, n_gas, instance0=r * len(en), **kw) for r in range(nsamples)]
    return stats(dfs)


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

    `majorant` is the null-collision majorant (mcs_swarm_batch_majorant):
    "global", the table's maxrate as in the reference, or "banded" (FP32 only),
    the maxrate of the particle's energy band; an FP32 batch defaults to "banded",
    an FP64 batch is always "global".
    """

    def __init__(self, total_capacity, max_slots, seed=0, precision="f64", ctx=None, majorant=None):
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
        self.set_majorant(majorant or ("banded" if precision == "f32" else "global"))

    def set_majorant(self, mode):
        """"global" or "banded" (see the class docstring); the step graph is
        captured again at the next step."""
        m = {"global": _lib.MAJORANT_GLOBAL, "banded": _lib.MAJORANT_BANDED}[mode]
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_majorant(self.handle, m))
        self.majorant = mode

    def configure(self, block=256, compact_every=8):
        check(self.ctx.handle, _lib.lib().mcs_swarm_batch_config(self.handle, int(block), int(compact_every)))

    def autotune(self, candidates=((256, 4), (128, 4), (256, 8), (128, 8), (256, 16), (256, 2)),
                 min_steps=2, log=None):
        """Project-written autotuner of the launch shape: runs real steps under each
        (workgroup size, compaction cadence) candidate, one compaction per trial,
        and keeps the fastest per step.  The populations change as they relax, so
        the candidates run twice, forward and then in reverse order, and each is
        scored by the mean of its two trials.  The choice never changes results (a
        slot's trajectory is the same under every shape and cadence); the trial
        steps count as ordinary steps.  log(message), when given, hears each trial
        as it ends.  Returns {candidate: seconds per step}."""
        import time as _t
        trials = {c: [] for c in candidates}
        for order in (list(candidates), list(reversed(candidates))):
            for blk, ce in order:
                self.configure(blk, ce)
                self.step(1)                      # capture the graphs for this shape
                self.ctx.synchronize()
                k = max(min_steps, ce)
                t0 = _t.perf_counter()
                self.step(k)
                self.ctx.synchronize()
                trials[(blk, ce)].append((_t.perf_counter() - t0) / k)
                if log:
                    log(f"autotune {blk} threads, compaction every {ce}: {trials[(blk, ce)][-1] * 1e3:.2f} ms "
                        "per step")
        timings = {c: sum(v) / len(v) for c, v in trials.items()}
        best = min(timings, key=timings.get)
        self.configure(*best)
        self.tuned = best
        self.tune_steps = sum(2 * (max(min_steps, ce) + 1) for _, ce in candidates)
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
    (standard error of the block means over |mean|; inf with fewer than 2 blocks),
    and the diffusion coefficients of the last steps' least-squares fit
    (Swarm.jl:144-150; no error estimate, as in the reference, which takes errors
    from replicas)."""
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
    if hasattr(t, "diffusion"):
        D = np.array(t.diffusion[:])
        out["diffusion_l"] = (float(D[2]), np.inf)
        out["diffusion_t"] = (float((D[0] + D[1]) / 2), np.inf)
        out["diffusion"] = (float(D.sum() / 3), np.inf)
    return out
