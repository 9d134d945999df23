"""Served swarm-coefficient requests on one device (continuous batching).

A request asks for the transport coefficients of Swarm.swarm (Swarm/src/Swarm.jl:
102-169) at one E/N and gas mix: drift velocity W = <v_z>, mean energy, and the
ionization and attachment rates, time-averaged after a relaxation period, until
the relative statistical error of W reaches the request's target.

SwarmServer keeps every admitted population resident in one SwarmBatch: requests
that arrive while others run are admitted as soon as a slab of their footprint
is free, share the same launches, and retire independently.  A request's result
does not depend on what else is in the batch (each slot draws from its own
Philox key and is checked at its own step counts).

PlainServer is the comparison point: one plain FP64 launch sequence per request
(advance!, the per-step tallies from one device reduction, resample!), requests
served one at a time in arrival order.
"""
import math
import os
import tempfile
import time
from dataclasses import dataclass, field

import numpy as np

from . import constants as co
from . import gases, lxcat
from ._lib import MCSError
from .jumc import CollisionTable, FixedField, Population, advance, context, repack
from .swarm import SwarmBatch, resample, tally_estimates

MIN_TD, MAX_TD = 10.0, 1000.0


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    en_td: float            # E/N in Td
    o2: float               # attacher fraction
    n0: int                 # initial count = resample target
    target: float           # relative statistical error wanted on W


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    converged: bool
    flags: int
    estimates: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def draw_requests(rate, duration, seed=0, n0_range=(1e3, 1e6), target_range=(0.005, 0.02)):
    """Open-loop Poisson arrivals; E/N log-uniform on 10-1000 Td, O2 fraction
    uniform on 0-0.3, n0 log-uniform on n0_range, target uniform on target_range."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        en = 10 ** rng.uniform(np.log10(MIN_TD), np.log10(MAX_TD))
        o2 = rng.uniform(0.0, 0.3)
        n0 = int(round(10 ** rng.uniform(np.log10(n0_range[0]), np.log10(n0_range[1]))))
        tg = rng.uniform(*target_range)
        out.append(Request(len(out), t, en, o2, n0, tg))


def request_dt(en_td):
    """dt = 1e-12 s * (100 Td / (E/N))."""
    return 1e-12 * 100.0 / en_td


class GasTables:
    """Collision tables of the N2-like + O2-like mix on the 4000-point log grid,
    one per O2 fraction rounded to `quantum` (table build is setup, not step)."""

    def __init__(self, quantum=0.005, ctx=None, ne=4000):
        self.ctx = ctx or context()
        self.quantum = quantum
        self.e_ev = np.geomspace(1e-3, 1e3, ne)
        self.path = os.path.join(tempfile.gettempdir(), f"mcs_gas_serve_{os.getpid()}_{ne}.json")
        gases.write_gas(self.path, self.e_ev, with_attacher=True)
        self._cache = {}

    def key(self, o2):
        return round(round(o2 / self.quantum) * self.quantum, 6)

    def get(self, o2):
        k = self.key(o2)
        if k not in self._cache:
            proc, rate, maxrate, _ = lxcat.load_lxcat([self.path], gases.densities(k), self.e_ev * co.eV)
            self._cache[k] = CollisionTable(proc, self.e_ev * co.eV, rate, maxrate, egrid="log", ctx=self.ctx)
        return self._cache[k]

    def warm(self):
        for j in range(int(round(0.3 / self.quantum)) + 1):
            self.get(j * self.quantum)


@dataclass
class Policy:
    """Convergence and sizing rules shared by the batched and the plain server."""
    relax_steps: int = 2000
    block_steps: int = 100
    min_blocks: int = 10
    max_steps: int = 20000
    chunk: int = 100               # steps between convergence checks
    tol: float = 0.1               # resample tolerance (Swarm.jl:193)
    capacity_factor: float = 2.0   # slab = factor * n0 + pad particles
    capacity_pad: int = 4096
    compact_every: int = 4
    init_kT_ev: float = 1.0
    # admission also keeps the expected (null-)collision events per batch step
    # below this budget, so that admitted requests progress at a bounded step time
    max_events_per_step: float = 2e8

    def capacity(self, n0):
        return int(self.capacity_factor * n0) + self.capacity_pad

    def done(self, steps, est_w_err, target):
        if steps >= self.max_steps:
            return True
        return steps >= self.relax_steps + self.min_blocks * self.block_steps and est_w_err <= target


def _field(en_td):
    return (0.0, 0.0, -en_td * co.Td * gases.N_GAS)


class SwarmServer:
    """Continuous batching of swarm requests in one SwarmBatch."""

    def __init__(self, total_capacity, max_slots=4096, seed=0, policy=None, tables=None,
                 block=256, precision="f64", ctx=None):
        self.ctx = ctx or context()
        self.policy = policy or Policy()
        self.tables = tables or GasTables(ctx=self.ctx)
        self.batch = SwarmBatch(total_capacity, max_slots, seed=seed, precision=precision, ctx=self.ctx)
        self.batch.configure(block=block, compact_every=self.policy.compact_every)
        self.total_capacity = int(total_capacity)
        self.max_slots = int(max_slots)
        self.steps_run = 0
        self.particle_steps = 0.0

    def work(self, r):
        """Expected events per step of a request: n0 * maxrate * dt."""
        return float(r.n0) * self.tables.get(r.o2).maxrate * request_dt(r.en_td)

    def _admit(self, r, now):
        P = self.policy
        table = self.tables.get(r.o2)
        return self.batch.add(table, int(r.n0), _field(r.en_td), request_dt(r.en_td), r.n0,
                              P.capacity(r.n0), instance=r.rid, tol=P.tol,
                              relax_steps=P.relax_steps, block_steps=P.block_steps,
                              init_kT_ev=P.init_kT_ev)

    def run(self, requests, wall_limit=None, log=None, on_retire=None):
        """Serve `requests` (sorted by arrival) in real time; returns (results,
        unfinished requests).  With wall_limit, stops admitting and stepping after
        that many seconds.  on_retire(slot, result) is called for each retired
        request while its slab still holds its particles."""
        P = self.policy
        pending = list(requests)
        waiting, active, results = [], {}, []
        t0 = time.perf_counter()
        head = 0
        load = 0.0
        while head < len(pending) or waiting or active:
            now = time.perf_counter() - t0
            if wall_limit is not None and now > wall_limit:
                break
            while head < len(pending) and pending[head].arrival <= now:
                waiting.append(pending[head])
                head += 1
            # admission by footprint, in arrival order (no overtaking)
            while waiting and len(active) < self.max_slots:
                r = waiting[0]
                if P.capacity(r.n0) > self.total_capacity:
                    raise ValueError(f"request {r.rid} can never fit: n0={r.n0}")
                if active and load + self.work(r) > P.max_events_per_step:
                    break
                try:
                    slot = self._admit(r, now)
                except MCSError as e:
                    if "no free slab" in str(e) or "no free slot" in str(e):
                        break
                    raise
                active[slot] = (r, now)
                load += self.work(r)
                waiting.pop(0)
            if not active:
                if head < len(pending):
                    time.sleep(max(0.0, min(pending[head].arrival - now, 0.05)))
                continue
            self.batch.step(P.chunk)
            self.steps_run += P.chunk
            tallies = self.batch.tallies(max(active) + 1)
            now = time.perf_counter() - t0
            for slot in list(active):
                t = tallies[slot]
                r, adm = active[slot]
                self.particle_steps += float(r.n0) * P.chunk
                est = tally_estimates(t)
                if t.flags or P.done(t.steps, est["drift_velocity"][1], r.target):
                    results.append(Result(r.rid, r.arrival, adm, now, int(t.steps),
                                          est["drift_velocity"][1] <= r.target, int(t.flags), est))
                    if on_retire:
                        on_retire(slot, results[-1])
                    self.batch.remove(slot)
                    del active[slot]
                    load -= self.work(r)
                    if log:
                        log(results[-1])
        unfinished = [r for r, _ in active.values()] + waiting + pending[head:]
        return results, unfinished


def _host_tally(means, relax, bsteps):
    """Per-step means -> the SwarmTally fields used by tally_estimates."""

    class T:
        pass

    t = T()
    m = np.asarray(means)
    rec = m[relax:] if len(m) > relax else np.zeros((0, 4))
    nb = len(rec) // bsteps
    t.nrecorded = len(rec)
    t.nblocks = nb
    t.sum = [math.fsum(rec[:, q]) for q in range(4)] if len(rec) else [0.0] * 4
    bm = rec[: nb * bsteps].reshape(nb, bsteps, 4).mean(axis=1) if nb else np.zeros((0, 4))
    t.bm_s1 = list(bm.sum(axis=0)) if nb else [0.0] * 4
    t.bm_s2 = list((bm ** 2).sum(axis=0)) if nb else [0.0] * 4
    return t


class PlainServer:
    """One plain FP64 launch sequence per request, one request at a time: per step
    advance!, the tallies from one device reduction (mcs_pop_moments), resample!;
    repack! every compact_every steps."""

    def __init__(self, seed=0, policy=None, tables=None, ctx=None):
        self.ctx = ctx or context()
        self.policy = policy or Policy()
        self.tables = tables or GasTables(ctx=self.ctx)
        self.seed = int(seed)

    def solve(self, r):
        P = self.policy
        table = self.tables.get(r.o2)
        kinds = np.array([p.code for p in table.proc])
        pop = Population(P.capacity(r.n0) * 4, None, table, seed=self.seed, instance=r.rid, ctx=self.ctx)
        pop.init_thermal(r.n0, P.init_kT_ev)
        f = FixedField(_field(r.en_td)[2])
        dt = request_dt(r.en_td)
        means = []
        steps = 0
        while True:
            advance(pop, f, dt, step=pop.step)
            m = pop.moments(table)
            na = max(m.nactive, 1)
            cr = np.array(m.collrate[: len(kinds)])
            means.append((m.sum_v[2] / na, m.sum_energy / na, cr[kinds == 3].sum() / na,
                          cr[kinds == 4].sum() / na))
            resample(pop, r.n0, P.tol, step=pop.step)
            pop.step += 1
            steps += 1
            if P.compact_every and steps % P.compact_every == 0:
                repack(pop)
            if steps % P.chunk == 0:
                est = tally_estimates(_host_tally(means, P.relax_steps, P.block_steps))
                if P.done(steps, est["drift_velocity"][1], r.target):
                    return steps, est

    def run(self, requests, wall_limit=None):
        """Serve `requests` (sorted by arrival) in real time, one at a time in
        arrival order; returns (results, unfinished).  With wall_limit, stops
        starting requests after that many seconds (the one running completes)."""
        results, t0 = [], time.perf_counter()
        for k, r in enumerate(requests):
            now = time.perf_counter() - t0
            if wall_limit is not None and now > wall_limit:
                return results, list(requests[k:])
            if r.arrival > now:
                if wall_limit is not None and r.arrival > wall_limit:
                    return results, list(requests[k:])
                time.sleep(r.arrival - now)
            start = time.perf_counter() - t0
            steps, est = self.solve(r)
            self.ctx.synchronize()
            fin = time.perf_counter() - t0
            results.append(Result(r.rid, r.arrival, start, fin, steps,
                                  est["drift_velocity"][1] <= r.target, 0, est))
        return results, []

    def service_times(self, requests):
        """Measured service time of each request run alone on the device."""
        out = []
        for r in requests:
            t = time.perf_counter()
            steps, est = self.solve(r)
            self.ctx.synchronize()
            out.append((time.perf_counter() - t, steps, est))
        return out


def fifo_latencies(requests, service):
    """Latencies of a single FIFO server with the given service times."""
    free, lat = 0.0, []
    for r, s in zip(requests, service):
        start = max(r.arrival, free)
        free = start + s
        lat.append(free - r.arrival)
    return np.array(lat)


def percentiles(lat):
    lat = np.asarray(lat, dtype=np.float64)
    if lat.size == 0:
        return {"p50": None, "p99": None}
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))}
