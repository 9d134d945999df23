"""Checkpoint and restart of a streamer run: the particle state in its current
(canonical) order with ids and the next id, the ion charge (qfixed), the
potential the next Poisson solve starts from (u), the step counter, the
collision table, the multigrid configuration and the run's parameters, in one
.npz file (arrays, plus the metadata as a JSON string; no pickles).

The random streams are counter-based -- Philox keyed by (seed, instance, phase,
particle id, step) -- so a run restored from a checkpoint continues bitwise as
the uninterrupted run would (the driver's output and restart state of
MCStreamer/src/main.jl:217-315 and io.jl, here for the FP64 store).
Everything else a step uses (qpart, q, the face fields, the cell order) is
rebuilt from this state at the start of the next step.
"""
import json

import numpy as np

from . import multigrid as mgmod
from .jumc import CollisionTable, ParticleStates, Population, context
from .lxcat import Process
from .streamer import Grid3, GridFields, Streamer

FORMAT = "mcstreamer_amd checkpoint 1"


def save(sim, path):
    """Write the state of Streamer `sim` (FP64 store) to `path` (.npz)."""
    popl, fields, ws = sim.popl, sim.fields, sim.ws
    if popl.precision != "f64":
        raise ValueError("checkpoints cover the FP64 particle store")
    conf = getattr(ws, "conf", None)
    if conf is None:
        raise ValueError("the workspace does not carry its multigrid configuration")
    tab = popl.collisions
    ps = popl.download()
    meta = {
        "format": FORMAT,
        "grid": {"n": list(fields.grid.n), "h": fields.grid.h},
        "run": {"eb": sim.eb, "dt": sim.dt, "ntarget": sim.ntarget, "nmax": sim.nmax,
                "scheme": sim.scheme, "step": sim.i},
        "population": {"capacity": popl.capacity, "seed": popl.seed, "instance": popl.instance,
                       "step": popl.step, "next_id": int(ps.next_id)},
        "table": {"maxrate": tab.maxrate, "egrid": "log" if tab.egrid == 1 else "uniform",
                  "proc": [{"kind": p.kind, "param": p.param, "signature": p.signature} for p in tab.proc]},
        "multigrid": {"bc": [int(b) for b in conf.bc], "s": conf.s, "levels": conf.levels,
                      "tolerance": conf.tolerance, "smooth1": conf.smooth1, "smooth2": conf.smooth2,
                      "maxiter": conf.maxiter, "precision": "f32" if conf.precision == 1 else "f64"},
    }
    np.savez(path, meta=np.array(json.dumps(meta)), x=ps.x, v=ps.v, w=ps.w, s=ps.s, active=ps.active,
             id=ps.id, qfixed=fields.get("qfixed"), u=fields.get("u"), energy=tab.energy, rate=tab.rate)


def load(path, ctx=None):
    """A Streamer continuing the run saved in `path`."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        if meta.get("format") != FORMAT:
            raise ValueError(f"{path}: not a {FORMAT} file")
        arrs = {k: z[k] for k in ("x", "v", "w", "s", "active", "id", "qfixed", "u", "energy", "rate")}
    ctx = ctx or context()
    t = meta["table"]
    procs = [Process(p["kind"], p["param"], p["signature"]) for p in t["proc"]]
    table = CollisionTable(procs, arrs["energy"], arrs["rate"], t["maxrate"], egrid=t["egrid"], ctx=ctx)
    g = meta["grid"]
    grid = Grid3(g["n"], g["h"])
    fields = GridFields(grid, ctx=ctx)
    fields.set("qfixed", arrs["qfixed"])
    fields.set("u", arrs["u"])
    pm = meta["population"]
    popl = Population(pm["capacity"], None, table, seed=pm["seed"], instance=pm["instance"], ctx=ctx)
    popl.upload(ParticleStates(arrs["x"], arrs["v"], arrs["w"], arrs["s"], arrs["active"], arrs["id"],
                               pm["next_id"]))
    popl.step = pm["step"]
    m = meta["multigrid"]
    faces = [mgmod.BottomBnd, mgmod.TopBnd, mgmod.LeftBnd, mgmod.RightBnd, mgmod.FrontBnd, mgmod.BackBnd]
    conf = mgmod.MGConfig(bc=list(zip(faces, m["bc"])), s=m["s"], levels=m["levels"], tolerance=m["tolerance"],
                          smooth1=m["smooth1"], smooth2=m["smooth2"], maxiter=m["maxiter"],
                          precision=m["precision"])
    ws = mgmod.Workspace(conf, grid.n, ctx=ctx)
    r = meta["run"]
    sim = Streamer(popl, fields, ws, r["eb"], r["dt"], r["ntarget"], r["nmax"], scheme=r["scheme"])
    sim.i = r["step"]
    sim.table = table
    return sim
