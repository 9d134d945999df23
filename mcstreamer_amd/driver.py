"""The reference's input-file driver for axisymmetric runs (MCStreamer/src/main.jl:15-197
`main`, :217-315 `nsteps`, timesteps.jl:5-28 `TimeStepper`/`atstep`), on libmcs.

    run = driver.setup("input.toml")    # main(finput; run=false)
    run.nsteps(k)                        # nsteps(...) for k steps
    driver.main("input.toml")            # the whole run to tmax

The TOML schema is the reference's (samples/input.toml): eb and field_units,
nair, n, maxp, ntarget, nmax, dt, tmax, cross_sections (LXCat JSON paths,
relative to the input file), rng_seed, initial_weight, poisson_levels,
[domain] R L M N, [seed] z0 z1 w, [interval] poisson output resample,
[photoionization] photon_weight photon_multiplier (main.jl:59-68): with a
PHOTOEMISSION process in the cross sections the run carries a photon population
of maxp slots (main.jl:113-122) and steps MultiPopulation(electrons, photons).
Sections this build does not carry (init_files, denoise, needle) raise
NotImplementedError rather than being ignored.  The seed positions come from
numpy's generator seeded with rng_seed (the reference draws them from Julia's
global RNG; the distribution, sampseg, is the same); everything after the
seed draws from the Philox streams keyed by rng_seed.  Output (every
interval.output, main.jl:245-249) is {j:04d}.h5 with the fields of
SavedGridFields (io.jl:9-44) in the layout of the reference's HDF5 export
(plot.jl:210-240, axisym.h5export), which HDF5 readers (HDF5.jl, h5py) open
directly, in place of the reference's JLD2 file {j:04d}.jld (Julia type
records, which only JLD2 reads); beside it {j:04d}.npz holds the ghosted
arrays u, er, ez, qfixed, qpart.
"""
import json
import os

import numpy as np

from . import axisym, lxcat
from . import constants as co
from .jumc import (CollisionTable, MultiPopulation, ParticleStates, PhotonPopulation, PhotonStates, Population,
                   ZhelezniakCollisions, advance, context, repack)

UNSUPPORTED = ("init_files", "denoise", "needle")


class TimeStepper:
    """TimeStepper and atstep (timesteps.jl:5-28): when t >= step * i, call f(i)
    and advance i.  `start` is stored but, as in the reference's atstep, unused."""

    def __init__(self, step, start=0.0):
        self.i = 0
        self.start = start
        self.step = step

    def atstep(self, f, t):
        if t >= self.step * self.i:
            f(self.i)
            self.i += 1


def geteb(inp):
    """geteb (main.jl:346-357): -eb times the unit's scale."""
    units = inp.get("field_units", "Td")
    nair = float(inp.get("nair", co.nair))
    scale = {"Td": co.Td * nair, "kV/cm": 1e5, "MV/m": 1e6, "V/m": 1.0, "": 1.0}[units]
    return -float(inp["eb"]) * scale


def sampseg(z0, z1, w, n, rng):
    """sampseg (main.jl:321-340): n points around the segment z0..z1 on the axis
    with a Gaussian blur of width w."""
    assert z1 >= z0
    l = z1 - z0
    p = l / (l + np.sqrt(2 * np.pi) * w)
    seg = rng.uniform(size=n) < p
    x = rng.normal(size=(n, 3)) * w
    u = rng.uniform(size=n)
    x[seg, 2] = z0 + u[seg] * l
    cap = ~seg
    x[cap, 2] += np.where(x[cap, 2] > 0, z1, z0)
    return x


def load_input(path):
    import tomli

    with open(path, "rb") as fd:
        inp = tomli.load(fd)
    bad = [k for k in UNSUPPORTED if k in inp]
    if bad:
        raise NotImplementedError(f"{path}: sections {bad} are not supported by this build")
    return inp


class Run:
    """The state main() builds for an axisymmetric run."""

    def __init__(self, finput, ctx=None, energy_points=100_000):
        inp = load_input(finput)
        self.input = inp
        self.finput = os.path.abspath(finput)
        self.energy_points = int(energy_points)
        self.ctx = ctx or context()
        self.outpath = os.path.splitext(os.path.abspath(finput))[0]
        seed = int(inp.get("rng_seed", 0))
        nair = float(inp.get("nair", co.nair))
        dom = inp["domain"]
        self.grid = axisym.Grid(float(dom["R"]), float(dom["L"]), int(dom["M"]), int(dom["N"]))
        self.fields = axisym.GridFields(self.grid, ctx=self.ctx)
        self.ntarget, self.nmax = int(inp["ntarget"]), int(inp["nmax"])
        densities = {"N2": nair * 0.79, "O2": nair * 0.21}
        energy = np.linspace(0.0, 1000 * co.eV, energy_points)   # LinRange(0, 1000 eV, 100_000)
        files = [os.path.join(os.path.dirname(os.path.abspath(finput)), f) for f in inp["cross_sections"]]
        photo = inp.get("photoionization", {})
        proc, rate, maxrate, _ = lxcat.load_lxcat(files, densities, energy,
                                                  photon_weight=float(photo.get("photon_weight", 1.0)),
                                                  photon_multiplier=float(photo.get("photon_multiplier", 1.0)))
        self.table = CollisionTable(proc, energy, rate, maxrate, egrid="uniform", ctx=self.ctx)
        self.dt = float(inp["dt"])
        self.tmax = float(inp["tmax"])
        self.eb = geteb(inp)
        s = inp["seed"]
        w, z0 = float(s["w"]), float(s["z0"])
        z1 = float(s.get("z1", z0))
        n = int(inp["n"])
        rng = np.random.default_rng(seed)
        x = sampseg(z0, z1, w, n, rng)
        wt = float(inp.get("initial_weight", 1.0))
        sclock = -np.log1p(-rng.uniform(size=n))   # ElectronState's nextcoll()
        self.popl = Population(int(inp["maxp"]), ParticleStates(x, np.zeros((n, 3)), np.full(n, wt), sclock,
                                                                  np.ones(n, dtype=np.uint8)),
                               self.table, seed=seed, ctx=self.ctx)
        # photo-ionization only with photo-emission (main.jl:113-122)
        pcolls = ZhelezniakCollisions.for_table(self.table)
        self.photons = None
        if pcolls is not None:
            self.photons = PhotonPopulation(int(inp["maxp"]), collisions=pcolls, ctx=self.ctx)
        self.mpopl = MultiPopulation(self.popl, self.photons)
        axisym.density(self.fields, "qfixed", self.popl, 1.0)     # main.jl:130-132
        levels = inp.get("poisson_levels")
        self.conf = axisym.poisson_config(self.grid, levels=None if levels is None else int(levels))
        self.ws = axisym.mgmod.Workspace(self.conf, (self.grid.M, self.grid.N), ctx=self.ctx)
        self.efield = axisym.FieldInterp(self.fields)
        self.tracker = axisym.CollisionTracker(self.fields)
        iv = inp["interval"]
        self.poisson_ts = TimeStepper(float(iv["poisson"]))
        self.output_ts = TimeStepper(float(iv["output"]))
        self.resample_ts = TimeStepper(float(iv["resample"]), self.dt)
        self.i = 0
        self.log = []

    def save_fields(self, j):
        os.makedirs(self.outpath, exist_ok=True)
        sv = axisym.saved_fields(self.fields)
        arrs = {k: sv[k] for k in ("u", "er", "ez", "qfixed", "qpart")}
        meta = json.dumps({"t": self.i * self.dt, "step": self.i, "grid": [self.grid.R, self.grid.L,
                                                                           self.grid.M, self.grid.N]})
        np.savez(os.path.join(self.outpath, f"{j:04d}.npz"), meta=np.array(meta), **arrs)
        axisym.h5export(self.fields, os.path.join(self.outpath, f"{j:04d}.h5"), saved=sv)

    def nsteps(self, k, output=True):
        """k steps of nsteps (main.jl:238-315): Poisson, output, advance!,
        diagnostics, resample (repack! first), repack!."""
        for _ in range(k):
            self.i += 1
            t = (self.i - 1) * self.dt
            self.poisson_ts.atstep(lambda j: axisym.poisson(self.fields, self.popl, self.eb, self.ws), t)
            if output:
                self.output_ts.atstep(self.save_fields, t)
            if self.photons is None:
                st = advance(self.popl, self.efield, self.dt, self.tracker, step=self.i - 1)
                pst = None
            else:
                st, pst = advance(self.mpopl, self.efield, self.dt, self.tracker, step=self.i - 1)
            m = self.popl.moments()
            rec = {"t": t, "n": st.n_out, "births": st.births, "deaths": st.deaths, "nactive": m.nactive,
                   "weight": m.weight, "mean_energy_ev": (m.wenergy / m.weight / co.eV) if m.weight else 0.0}
            if pst is not None:
                pm = self.photons.moments()
                rec.update(photons=pm.nactive, photon_weight=pm.weight if pm.nactive else 0.0,
                           photons_emitted=pst.n_out - pst.n_in, photoelectrons=pst.births, n=st.n_out + pst.births)

            def do_resample(j):
                if self.photons is not None:
                    repack(self.photons)       # main.jl:285
                repack(self.popl)
                rs = axisym.resample(self.popl, self.fields, self.ntarget, self.nmax, step=self.i - 1)
                rec["resample"] = (rs.n_in, rs.n_out)
                rec["weight_resampled"] = self.popl.moments().weight

            self.resample_ts.atstep(do_resample, t)
            repack(self.popl)
            self.log.append(rec)
        return self.log[-1] if self.log else None


def setup(finput, ctx=None, **kw):
    return Run(finput, ctx=ctx, **kw)


def main(finput, ctx=None):
    """main(finput) (main.jl:15-197) to tmax."""
    run = Run(finput, ctx=ctx)
    run.nsteps(int(run.tmax // run.dt))
    return run


CHECKPOINT_FORMAT = "mcstreamer_amd axisymmetric checkpoint 1"


def save(run, path):
    """The state of `run` after its last step to one .npz (no pickles): the
    particles in their order with ids and the next id, qfixed, the potential the
    next solve starts from, the face fields of the last solve, the step and the
    TimeSteppers' counters.  The input
    file is re-read on load (its path is stored)."""
    ps = run.popl.download()
    meta = {"format": CHECKPOINT_FORMAT, "finput": run.finput, "energy_points": run.energy_points, "i": run.i,
            "ts": [run.poisson_ts.i, run.output_ts.i, run.resample_ts.i], "next_id": int(ps.next_id)}
    extra = {}
    if run.photons is not None:
        ph = run.photons.download()
        meta["photon_next_id"] = int(ph.next_id)
        extra = {f"photon_{k}": getattr(ph, k) for k in ("x", "v", "nu", "w", "s", "active", "id")}
    np.savez(path, meta=np.array(json.dumps(meta)), x=ps.x, v=ps.v, w=ps.w, s=ps.s, active=ps.active, id=ps.id,
             qfixed=run.fields.get("qfixed"), u=run.fields.get("u"), er=run.fields.get("er"),
             ez=run.fields.get("ez"), **extra)


def load(path, ctx=None):
    """A Run continuing the one saved in `path` (the Philox streams are keyed by
    particle id and step, so the continuation is the uninterrupted run's)."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        if meta.get("format") != CHECKPOINT_FORMAT:
            raise ValueError(f"{path}: not a {CHECKPOINT_FORMAT} file")
        arrs = {k: z[k] for k in ("x", "v", "w", "s", "active", "id", "qfixed", "u", "er", "ez")}
        phot = {k: z[f"photon_{k}"] for k in ("x", "v", "nu", "w", "s", "active", "id")} \
            if "photon_next_id" in meta else None
    run = Run(meta["finput"], ctx=ctx, energy_points=int(meta["energy_points"]))
    if (phot is None) != (run.photons is None):
        raise ValueError(f"{path}: photon population does not match the input's cross sections")
    if phot is not None:
        run.photons.upload(PhotonStates(phot["x"], phot["v"], phot["nu"], phot["w"], phot["s"], phot["active"],
                                        phot["id"], int(meta["photon_next_id"])))
    run.popl.upload(ParticleStates(arrs["x"], arrs["v"], arrs["w"], arrs["s"], arrs["active"], arrs["id"],
                                   int(meta["next_id"])))
    run.fields.set("qfixed", arrs["qfixed"])
    for k in ("u", "er", "ez"):
        run.fields.set(k, arrs[k])
    run.i = int(meta["i"])
    run.poisson_ts.i, run.output_ts.i, run.resample_ts.i = (int(v) for v in meta["ts"])
    return run

