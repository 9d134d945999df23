"""Benchmark of the per-step particle-mesh loop on one MI355X.

Default workload (N = 1): the FP64 streamer step on a 128x128x256 mesh with
h = 5 um, a 1e6-pair Gaussian seed, CIC deposition, V(2,2) multigrid to 1e-8
relative residual, per-cell merge/split (8..64) and dt = 2e-13 s.  One step =
deposit + Poisson solve + face fields + fused push/collisions (mesh gather, ion
tracker) + diagnostics + canonical cell sort + per-cell resample.

Prints one JSON line.  `value` is particle-steps/s of whole steps with the
state resident in HBM.  `phase_share` splits the step by HIP events.
`roofline` is for the step's largest share, the multigrid solve: algorithmic
bytes per solve over its duration (HIP events around each solve);
`advance_roofline` is the fused step's (128 B per particle-step), with its VALU
issue rate beside it.  `late_run` times steps of the same run near its end.
`cpu_baseline` times whole oracle steps of the same state on every thread the
OpenMP build of the oracle gets, with the single-threaded figure beside it.
`swarm` holds the swarm half of the metric: batched swarm solves/s (FP32, banded
null-collision majorant, and the same batch with the table's one maxrate)
against one plain FP64 launch sequence per instance, and served requests
(p50/p99 latency, solves/s) at the configuration's arrival rate and at a rate
the device sustains, against a plain FP64 server on the same stream.
"""
import argparse
import json
import os
import tempfile
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "particle-steps/s & %HBM peak; swarm solves/s, p99 vs plain at <=1% stat err"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PSTEP_F64 = 128.0


# Algorithmic HBM bytes of one fmg! correction pass per finest cell (FP64): the
# finest residual (24 B) and update (24 B), plus the V(2,2) cycles of the FMG
# sweep — 165 B per cell for the finest cycle (4 RBGS sweeps x 24 B, residual
# 24, restriction 8, prolongation 16, times 8/7 over the coarser levels), and
# 8/7 again for the cycles that start at coarser levels.
BYTES_PER_CELL_FMG_PASS = 24.0 + 24.0 + 165.0 * 8.0 / 7.0


def vcycle_roofline(passes_per_s, n):
    cells = float(n[0]) * n[1] * n[2]
    ach = passes_per_s * BYTES_PER_CELL_FMG_PASS * cells / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "bytes_per_pass": BYTES_PER_CELL_FMG_PASS * cells,
            "note": "one pass = one fmg! correction (the counted V-cycles); whole-solve time over algorithmic bytes"}


def measured_traffic(kernel, units):
    """HBM bytes per launch of `kernel` from the FETCH_SIZE/WRITE_SIZE counter runs
    summarised under profiles/ (bytes per particle-step there, times the
    particle-steps of one launch here); None when no measurement is recorded."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fd:
            rec = json.load(fd)[kernel]
        return rec["hbm_bytes_per_particle_step"] * units
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline_streamer(sim, table, omp, min_s=10.0, max_steps=30):
    """Whole oracle steps from the device state, one after another, until at least
    min_s seconds of CPU work (bounded sample): the OpenMP build on every thread
    OpenMP offers (omp), or the single-threaded checker."""
    from oracle import refcpu as R

    prev = R.use_openmp(omp)
    try:
        popl, fields = sim.popl, sim.fields
        ps = popl.download()
        n0 = len(ps)
        P = R.Pop(ps.x, ps.v, ps.w, ps.s, ps.active, ps.id, cap=int(n0 * 3) + 1000, next_id=ps.next_id)
        grid = fields.grid
        M = R.Mesh(grid.n, grid.h)
        M.qfixed[:] = fields.get("qfixed")
        M.u[:] = fields.get("u")
        T = R.Table(table.energy, table.rate, table.kind, table.param, table.maxrate, egrid=table.egrid)
        conf = sim.ws_conf
        psteps, steps, iters = 0, 0, []
        t0 = time.perf_counter()
        el = 0.0
        while steps < max_steps and (steps == 0 or el < min_s):
            R.sort_by_cell(P, M)
            M.qpart[:] = 0.0
            R.deposit(P, M, 1, 1, -1.0)
            R.assemble_charge(M)
            rc, it, _ = R.mg_solve(grid.n, conf.bc, conf.s, conf.levels, conf.smooth1, conf.smooth2,
                                   conf.tolerance, conf.maxiter, M.u, M.q)
            R.efield_faces(M, sim.eb)
            f = R.mesh_field(M)
            psteps += P.n
            R.advance(P, T, f, sim.dt, M, seed=popl.seed, step=10_000 + steps)
            R.sort_by_cell(P, M)
            R.resample_cells(P, M, sim.ntarget, sim.nmax, seed=popl.seed, step=10_000 + steps)
            iters.append(it)
            steps += 1
            el = time.perf_counter() - t0
        cores = R.threads()
    finally:
        R.use_openmp(prev)
    return {"value": psteps / el, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} full streamer steps of the oracle from the device state after the timed steps "
                      f"({n0} particles at the start, {grid.n[0]}x{grid.n[1]}x{grid.n[2]} mesh, "
                      f"{sum(iters)} fmg iterations), {el:.1f} s on {cores} thread(s)"
                      + (" (OpenMP build: particle loop in ordered chunks, threaded grid loops, the deposit "
                         "as a per-cell gather; the canonical sort and the resample stay sequential)"
                         if omp else "")}


def read_profile_json(name, key):
    """A record of profiles/<name> (measured by rocprofv3 counter runs), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as fd:
            return json.load(fd)[key]
    except (OSError, KeyError, ValueError):
        return None


# MI355X engine clock and SIMD count; a wave64 FP64 VALU instruction issues over 4
# cycles (16 FP64 lanes per SIMD and cycle: 78.6 TFLOP/s FMA peak)
CLOCK_HZ = 2.4e9
SIMDS = 1024
VALU_PEAK_WAVE_INSTR_PER_S = SIMDS * CLOCK_HZ / 4.0

PHASES = ("mg_solve", "advance", "track", "deposit", "sort_count", "sort_scatter", "sort_gather", "resample",
          "faces")


def run_streamer(args):
    from mcstreamer_amd import streamer
    from mcstreamer_amd.jumc import context

    ctx = context()
    n = tuple(args.mesh)
    cfg = streamer.make_config1(n=n, h=args.h, npairs=args.npairs, seed=args.seed)
    sim = streamer.Streamer(cfg["popl"], cfg["fields"], cfg["ws"], cfg["eb"], cfg["dt"],
                            cfg["ntarget"], cfg["nmax"], scheme="cic")
    sim.ws_conf = cfg["conf"]
    for i in range(args.warmup):
        out = sim.step()
        print(f"warmup {i}: n={out['advance'].n_in} mg_iters={out['mg_iters']}", file=sys.stderr, flush=True)
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
    psteps = 0
    vcycles = 0
    iters = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = sim.step()
        psteps += out["advance"].n_in
        vcycles += max(out["mg_iters"] - 1, 0)
        iters.append(out["mg_iters"])
    ctx.synchronize()
    el = time.perf_counter() - t0
    ctx.profile(False)
    kt = {}
    for k in PHASES + ("mg_rbgs", "mg_residual", "mg_restrict", "mg_prolong"):
        ms, cnt = ctx.profile_read(k)
        kt[k] = {"ms_total": ms, "launches": cnt}
    cells = float(n[0]) * n[1] * n[2]
    # the dominant phase: the multigrid solve (HIP events around each mcs_mg_solve_mesh)
    solve_s = kt["mg_solve"]["ms_total"] / 1e3
    nsolve = max(kt["mg_solve"]["launches"], 1)
    solve_bytes = (vcycles * BYTES_PER_CELL_FMG_PASS + len(iters) * 16.0) * cells / nsolve
    solve_ach = solve_bytes / (solve_s / nsolve) / 1e9 if solve_s > 0 else 0.0
    mg_traffic = read_profile_json("hbm_traffic.json", "mg_solve")
    # the fused step
    adv_ms, adv_n = kt["advance"]["ms_total"], kt["advance"]["launches"]
    adv_avg_s = adv_ms / 1e3 / max(adv_n, 1)
    pps = psteps / max(args.steps, 1)
    bytes_per_launch = BYTES_PER_PSTEP_F64 * pps
    achieved = bytes_per_launch / adv_avg_s / 1e9
    adv_traffic = read_profile_json("hbm_traffic.json", "k_advance<1,1>")
    valu = read_profile_json("valu_counters.json", "k_advance<1,1>")
    valu_obj = None
    if valu:
        ips = valu["valu_instr_per_particle_step"] * pps / adv_avg_s
        valu_obj = {"bound": "valu", "achieved": ips, "peak": VALU_PEAK_WAVE_INSTR_PER_S,
                    "unit": "wave64 VALU instructions/s", "frac": ips / VALU_PEAK_WAVE_INSTR_PER_S,
                    "note": "SQ_INSTS_VALU per particle-step (rocprofv3 counters, profiles/valu_counters.json) "
                            "x particle-steps per launch / HIP-event launch time; peak = 1024 SIMDs x 2.4 GHz / "
                            "4 cycles, the FP64 issue rate (FP32 and integer instructions issue faster)"}
    step_ms = el / args.steps * 1e3
    result = {
        "metric": METRIC,
        "value": psteps / el,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"FP64 streamer step, {n[0]}x{n[1]}x{n[2]} mesh h={args.h:g} m, "
                               f"{args.npairs:g}-pair Gaussian seed, CIC, V(2,2) to 1e-8 rel, "
                               "merge/split 8-64, dt=2e-13 s",
                   "mesh": list(n), "particles_per_step": pps, "fmg_iterations": iters,
                   "parallelism": "single GPU"},
        "phase_share": {k: kt[k]["ms_total"] / (el * 1e3) for k in PHASES},
        "vcycles_per_s": vcycles / solve_s if solve_s > 0 else None,
        "vcycle_note": "fmg! corrections (V-cycle passes) per second of solve time (HIP events around the solves)",
        "roofline": {"bound": "hbm", "achieved": solve_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": solve_ach / HBM_PEAK_GBS,
                     "traffic": mg_traffic["hbm_bytes_per_correction"] * vcycles / nsolve if mg_traffic else None,
                     "kernel": "multigrid solve (mcs_mg_solve_mesh: the k_mg_* kernels, the step's largest share)",
                     "avg_launch_ms": solve_s / nsolve * 1e3, "bytes_per_launch": solve_bytes,
                     "bytes_note": "per fmg! correction 24+24 B/cell (finest residual, update) + 165 B/cell x 8/7 "
                                   "(V(2,2) cycles of the FMG sweep), + 16 B/cell for the final residual check"},
        "advance_roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS,
                             "traffic": adv_traffic["hbm_bytes_per_particle_step"] * pps if adv_traffic else None,
                             "kernel": "k_advance<1,1> (fused push + collisions, mesh gather, ion tracker log)",
                             "avg_launch_ms": adv_avg_s * 1e3, "bytes_per_launch": bytes_per_launch,
                             "valu": valu_obj},
        "kernels_ms": kt,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline_streamer(sim, cfg["table"], omp=True)
        result["cpu_baseline"]["single_core"] = cpu_baseline_streamer(sim, cfg["table"], omp=False)
        print(f"cpu baseline: {result['cpu_baseline']['value']:.3e} on {result['cpu_baseline']['cores']} threads, "
              f"{result['cpu_baseline']['single_core']['value']:.3e} on 1", file=sys.stderr, flush=True)
    else:
        result["cpu_baseline"] = None
    if args.late_steps > 0:
        # the same run further on: untimed steps up to late_at, then late_steps timed
        done = args.warmup + args.steps
        while done < args.late_at:
            sim.step(diagnostics=False)
            done += 1
        ctx.synchronize()
        lp = 0
        t0 = time.perf_counter()
        for _ in range(args.late_steps):
            out = sim.step()
            lp += out["advance"].n_in
        ctx.synchronize()
        lel = time.perf_counter() - t0
        result["late_run"] = {"from_step": done, "steps": args.late_steps, "ms_per_step": lel / args.late_steps * 1e3,
                              "particle_steps_per_s": lp / lel, "particles_per_step": lp / args.late_steps,
                              "fmg_iterations_last": out["mg_iters"]}
        print(f"late run: steps {done}-{done + args.late_steps}, {lel / args.late_steps * 1e3:.2f} ms/step",
              file=sys.stderr, flush=True)
    if not args.no_swarm:
        result["swarm"] = swarm_block(args)
    return result


def run_swarm(args):
    """Config-0 style swarm: 1e4 electrons at 100 Td in the N2-like gas."""
    from mcstreamer_amd import constants as co
    from mcstreamer_amd import gases, jumc, lxcat

    ctx = jumc.context()
    e_ev = np.geomspace(1e-3, 1e3, 4000)
    path = os.path.join(tempfile.gettempdir(), f"mcs_bench_gas_c0_{os.getpid()}.json")
    gases.write_gas(path, e_ev, with_attacher=False)
    proc, rate, maxrate, _ = lxcat.load_lxcat([path], gases.densities(0.0), e_ev * co.eV)
    table = jumc.CollisionTable(proc, e_ev * co.eV, rate, maxrate, egrid="log")
    n0 = args.swarm_n
    rng = np.random.default_rng(args.seed)
    v = rng.normal(size=(n0, 3)) * np.sqrt(co.eV / co.electron_mass)
    ps = jumc.ParticleStates(np.zeros((n0, 3)), v, np.ones(n0), -np.log1p(-rng.uniform(size=n0)),
                             np.ones(n0, dtype=np.uint8))
    pop = jumc.Population(10 * n0, ps, table, seed=args.seed)
    field = jumc.FixedField(-100 * co.Td * gases.N_GAS)
    from mcstreamer_amd import swarm

    for _ in range(args.warmup):
        swarm.swarm_step(pop, field, 1e-12, n0)
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
    psteps = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = swarm.swarm_step(pop, field, 1e-12, n0)
        psteps += st.n_in
    ctx.synchronize()
    el = time.perf_counter() - t0
    ms, cnt = ctx.profile_read("advance")
    avg = ms / 1e3 / max(cnt, 1)
    ach = BYTES_PER_PSTEP_F64 * psteps / max(args.steps, 1) / avg / 1e9
    return {"metric": METRIC, "value": psteps / el, "unit": "particle-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"FP64 swarm, {n0} electrons, 100 Td, N2-like gas, dt=1e-12 s"},
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": avg * 1e3},
            "cpu_baseline": None}


def c3_instances(n_inst, seed):
    """Batched-swarm instances: E/N log-uniform on 10-1000 Td, O2 fraction uniform
    on 0-0.3, n0 log-uniform on 1e3-1e6, dt = 1e-12 s * 100 Td / (E/N)."""
    from mcstreamer_amd.serve import Request

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        en = 10 ** rng.uniform(1.0, 3.0)
        o2 = rng.uniform(0.0, 0.3)
        n0 = int(round(10 ** rng.uniform(3.0, 6.0)))
        out.append(Request(i, 0.0, en, o2, n0, 0.01))
    return out


def run_batch(args):
    """1024 independent swarm instances in one batch (graph-replayed steps,
    autotuned launch shape), against one plain FP64 launch sequence per instance."""
    from mcstreamer_amd import jumc, swarm
    from mcstreamer_amd.serve import GasTables, Policy, _field, request_dt

    ctx = jumc.context()
    insts = c3_instances(args.instances, args.seed)
    pol = Policy()
    tables = GasTables(ctx=ctx)
    total = sum(cdiv512(pol.capacity(r.n0)) for r in insts)
    b = swarm.SwarmBatch(total, len(insts), seed=args.seed, precision=args.precision)
    t0 = time.perf_counter()
    for r in insts:
        b.add(tables.get(r.o2), r.n0, _field(r.en_td), request_dt(r.en_td), r.n0, pol.capacity(r.n0),
              instance=r.rid, relax_steps=pol.relax_steps, block_steps=pol.block_steps)
    ctx.synchronize()
    t_admit = time.perf_counter() - t0
    print(f"admitted {len(insts)} instances, capacity {total}, {t_admit:.1f} s", file=sys.stderr, flush=True)
    tune = b.autotune(log=lambda m: print(m, file=sys.stderr, flush=True)) if not args.no_autotune else {}
    print(f"autotune: {tune}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        b.step(1)
    ctx.synchronize()
    n_before = sum(t.nactive for t in b.tallies(len(insts)))
    t0 = time.perf_counter()
    b.step(args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    print(f"batched: {args.steps} steps in {el:.2f} s", file=sys.stderr, flush=True)
    tl = b.tallies(len(insts))
    n_after = sum(t.nactive for t in tl)
    flags = sum(1 for t in tl if t.flags)
    psteps = 0.5 * (n_before + n_after) * args.steps
    events = sum(r.n0 * tables.get(r.o2).maxrate * request_dt(r.en_td) for r in insts)
    t_step = el / args.steps
    solves = len(insts) / (args.solve_steps * t_step)
    # one plain FP64 launch sequence per instance (advance!, tallies, resample!)
    plain = None
    if not args.no_plain:
        tp = 0.0
        for k, r in enumerate(insts):
            if k % 128 == 0:
                print(f"plain path: instance {k}", file=sys.stderr, flush=True)
            tab = tables.get(r.o2)
            pop = jumc.Population(2 * pol.capacity(r.n0), None, tab, seed=args.seed, instance=r.rid)
            pop.init_thermal(r.n0, pol.init_kT_ev)
            f = jumc.FixedField(_field(r.en_td)[2])
            swarm.swarm_step(pop, f, request_dt(r.en_td), r.n0)
            pop.moments(tab)
            ctx.synchronize()
            t1 = time.perf_counter()
            for _ in range(2):
                swarm.swarm_step(pop, f, request_dt(r.en_td), r.n0)
                pop.moments(tab)
            ctx.synchronize()
            tp += (time.perf_counter() - t1) / 2
            del pop
        plain = {"s_per_step_all_instances": tp, "solves_per_s": len(insts) / (args.solve_steps * tp)}
    ach = BYTES_PER_PSTEP_F64 * psteps / el / 1e9
    return {"metric": METRIC, "value": psteps / el, "unit": "particle-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 split (f64 tallies)" if args.precision == "f32" else "f64",
            "data": "synthetic",
            "config": {"workload": f"batched swarms: {len(insts)} instances, E/N log-uniform 10-1000 Td, "
                                   "O2 0-0.3, n0 log-uniform 1e3-1e6, dt=1e-12 s*100/(E/N)",
                       "sum_n0": int(sum(r.n0 for r in insts)), "events_per_step": events,
                       "solve_steps": args.solve_steps, "parallelism": "single GPU"},
            "swarm_solves_per_s": solves,
            "plain_fp64_per_instance": plain,
            "speedup_vs_plain": (solves / plain["solves_per_s"]) if plain else None,
            "autotune_s_per_step": {f"{k[0]}x{k[1]}": v for k, v in tune.items()},
            "tuned": list(getattr(b, "tuned", ())), "admit_s": t_admit, "slots_flagged": flags,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "whole batched step (128 B per particle-step over the step time)"},
            "cpu_baseline": None}


def swarm_block(args):
    """The swarm half of the metric, bounded so that the default run stays a few
    minutes: config 3 (batched FP32 swarms against one plain FP64 launch sequence
    per instance) and config 4 (served requests against a plain FP64 server, both
    in real time on the same Poisson streams)."""
    from mcstreamer_amd import jumc, swarm
    from mcstreamer_amd.serve import (GasTables, PlainServer, Policy, SwarmServer, _field, draw_requests,
                                      percentiles, request_dt)

    ctx = jumc.context()
    pol = Policy()
    t00 = time.perf_counter()

    def note(msg):
        print(f"[swarm {time.perf_counter() - t00:6.1f} s] {msg}", file=sys.stderr, flush=True)

    tables = GasTables(ctx=ctx)
    tables.warm()
    note("gas tables ready")
    out = {}
    # ---- config 3
    insts = c3_instances(args.instances, args.seed)
    total = sum(cdiv512(pol.capacity(r.n0)) for r in insts)
    b = swarm.SwarmBatch(total, len(insts), seed=args.seed, precision="f32")
    for r in insts:
        b.add(tables.get(r.o2), r.n0, _field(r.en_td), request_dt(r.en_td), r.n0, pol.capacity(r.n0),
              instance=r.rid, relax_steps=pol.relax_steps, block_steps=pol.block_steps)
    ctx.synchronize()
    note(f"c3: {len(insts)} instances admitted, capacity {total}")
    tune = b.autotune(log=note)
    note(f"c3: autotuned {getattr(b, 'tuned', None)}")
    # timed at the stage the plain sample is timed at (after c3_plain_relax steps)
    done = b.tune_steps
    b.step(max(3, args.c3_plain_relax - done))
    ctx.synchronize()
    note("c3: relaxed")
    t0 = time.perf_counter()
    b.step(args.c3_steps)
    ctx.synchronize()
    t_step = (time.perf_counter() - t0) / args.c3_steps
    flagged = sum(1 for t in b.tallies(len(insts)) if t.flags)
    print(f"c3 batched ({b.majorant} majorant): {t_step * 1e3:.1f} ms per step", file=sys.stderr, flush=True)
    # the same batch with the reference's one maxrate as the majorant
    b.set_majorant("global")
    b.step(2)
    ctx.synchronize()
    t0 = time.perf_counter()
    b.step(args.c3_steps)
    ctx.synchronize()
    t_step_global = (time.perf_counter() - t0) / args.c3_steps
    del b
    print(f"c3 batched (global majorant): {t_step_global * 1e3:.1f} ms per step", file=sys.stderr, flush=True)
    # plain FP64 per instance: a stratified sample (every k-th instance by n0), each
    # timed over c3_plain_steps steps after c3_plain_relax steps
    order = sorted(insts, key=lambda r: r.n0)
    k = max(len(order) // args.c3_plain_sample, 1)
    sample = order[k // 2::k][: args.c3_plain_sample]
    per_step = []
    for r in sample:
        tab = tables.get(r.o2)
        # the plain server's launch sequence: capacity and repack! cadence as there
        pop = jumc.Population(4 * pol.capacity(r.n0), None, tab, seed=args.seed, instance=r.rid, ctx=ctx)
        pop.init_thermal(r.n0, pol.init_kT_ev)
        f = jumc.FixedField(_field(r.en_td)[2])

        def plain_step(k):
            swarm.swarm_step(pop, f, request_dt(r.en_td), r.n0)
            pop.moments(tab)
            if pol.compact_every and (k + 1) % pol.compact_every == 0:
                jumc.repack(pop)

        for k in range(args.c3_plain_relax):
            plain_step(k)
        ctx.synchronize()
        t1 = time.perf_counter()
        for k in range(args.c3_plain_steps):
            plain_step(args.c3_plain_relax + k)
        ctx.synchronize()
        per_step.append((time.perf_counter() - t1) / args.c3_plain_steps)
        note(f"c3 plain: instance {r.rid} n0 {r.n0} {per_step[-1] * 1e3:.2f} ms/step")
        del pop
    plain_solve_s = float(np.mean(per_step)) * args.solve_steps
    batched = len(insts) / (args.solve_steps * t_step)
    batched_global = len(insts) / (args.solve_steps * t_step_global)
    out["c3"] = {
        "workload": f"{len(insts)} instances, E/N log-uniform 10-1000 Td, O2 0-0.3, n0 log-uniform 1e3-1e6, "
                    f"{args.solve_steps} steps each, dt=1e-12 s*100/(E/N)",
        "timed_from_step": max(done + 3, args.c3_plain_relax),
        "batched_f32_solves_per_s": batched, "batched_ms_per_step": t_step * 1e3, "slots_flagged": flagged,
        "majorant": "banded: the null-collision clock of each flight runs at the maxrate of the particle's "
                    "energy band (the same process of real collisions as the table's one maxrate)",
        "batched_f32_global_majorant_solves_per_s": batched_global,
        "batched_global_majorant_ms_per_step": t_step_global * 1e3,
        "plain_f64_solves_per_s": 1.0 / plain_solve_s,
        "plain_method": f"stratified sample of {len(sample)} instances by n0, each timed over "
                        f"{args.c3_plain_steps} steps after {args.c3_plain_relax} (advance!, tallies, resample!, repack! every "
                        f"{pol.compact_every} steps); "
                        f"mean step time x {args.solve_steps} steps per solve",
        "speedup_vs_plain": batched * plain_solve_s,
        "speedup_vs_plain_global_majorant": batched_global * plain_solve_s}
    print(f"c3: batched {batched:.3f} solves/s, plain {1.0 / plain_solve_s:.3f} solves/s", file=sys.stderr,
          flush=True)
    # ---- config 4: the configuration's rate, and a rate the device sustains
    c4 = {}
    for label, rate, dur in (("config_rate", args.rate, args.c4_seconds), ("sustained", args.c4_low_rate,
                                                                           args.c4_low_seconds)):
        reqs = draw_requests(rate, dur, seed=args.seed + 1)
        srv = SwarmServer(args.serve_capacity, max_slots=4096, seed=args.seed, policy=pol, tables=tables,
                          precision="f32")
        note(f"c4 {label}: {len(reqs)} requests at {rate:g}/s")
        t0 = time.perf_counter()
        done, left = srv.run(reqs, wall_limit=dur * 1.5,
                             log=lambda r: note(f"c4: request {r.rid} retired after {r.steps} steps"))
        el = time.perf_counter() - t0
        rec = {"rate_req_per_s": rate, "seconds": dur, "arrived": len(reqs), "completed": len(done),
               "unfinished": len(left), "converged": sum(1 for r in done if r.converged),
               "solves_per_s": len(done) / el, "latency_s": percentiles([r.latency for r in done])}
        if label == "sustained":
            ps = PlainServer(seed=args.seed, policy=pol, tables=tables, ctx=ctx)
            note("c4 plain server")
            t0 = time.perf_counter()
            pdone, pleft = ps.run(reqs, wall_limit=dur * 1.5)
            pel = time.perf_counter() - t0
            rec["plain_f64"] = {"completed": len(pdone), "unfinished": len(pleft), "solves_per_s": len(pdone) / pel,
                                "latency_s": percentiles([r.latency for r in pdone]),
                                "method": "one plain FP64 launch sequence per request, served one at a time in "
                                          "arrival order in real time on the same stream"}
        c4[label] = rec
        print(f"c4 {label}: {len(done)}/{len(reqs)} done, p50/p99 {rec['latency_s']}", file=sys.stderr, flush=True)
    out["c4"] = c4
    return out


def cdiv512(n):
    return (n + 511) // 512 * 512


def run_serve(args):
    """Served swarm requests: open-loop Poisson arrivals, continuous batching with
    footprint and work-budget admission, retirement at the target error on W."""
    from mcstreamer_amd import jumc
    from mcstreamer_amd.serve import (GasTables, PlainServer, Policy, SwarmServer, draw_requests,
                                      fifo_latencies, percentiles)

    ctx = jumc.context()
    reqs = draw_requests(args.rate, args.serve_seconds, seed=args.seed)
    pol = Policy()
    tables = GasTables(ctx=ctx)
    tables.warm()
    srv = SwarmServer(args.serve_capacity, max_slots=4096, seed=args.seed, policy=pol, tables=tables,
                      precision=args.precision)
    t0 = time.perf_counter()
    done, unfinished = srv.run(reqs, wall_limit=args.serve_seconds * args.serve_drain,
                               log=lambda r: print(f"retired {r.rid} after {r.steps} steps, "
                                                   f"latency {r.latency:.2f} s", file=sys.stderr,
                                                   flush=True))
    el = time.perf_counter() - t0
    lat = [r.latency for r in done]
    out = {"metric": METRIC, "value": len(done) / el, "unit": "swarm solves/s", "n_gpus": 1,
           "steps": srv.steps_run, "warmup": 0, "ms_per_step": el / max(srv.steps_run, 1) * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 split (f64 tallies)" if args.precision == "f32" else "f64",
           "data": "synthetic",
           "config": {"workload": f"served swarm requests, Poisson {args.rate:g} req/s for "
                                  f"{args.serve_seconds:g} s, target error on W 0.5-2%",
                      "arrived": len(reqs), "completed": len(done), "unfinished": len(unfinished),
                      "converged": sum(1 for r in done if r.converged),
                      "policy": pol.__dict__, "parallelism": "single GPU"},
           "latency_s": percentiles(lat), "particle_steps_per_s": srv.particle_steps / el,
           "cpu_baseline": None}
    if args.serve_plain > 0:
        ps = PlainServer(seed=args.seed, policy=pol, tables=tables)
        sample = [r for r in reqs[: args.serve_plain]]
        svc = ps.service_times(sample)
        st = [s for s, _, _ in svc]
        out["plain_fp64_per_request"] = {
            "requests": len(sample), "solves_per_s": len(sample) / sum(st),
            "latency_s_fifo": percentiles(fifo_latencies(sample, st)),
            "batched_latency_s_same_requests": percentiles([r.latency for r in done if r.rid < len(sample)])}
    return out


BYTES_PER_PSTEP_F32 = 72.0


def run_precision(args):
    """The precision form of the streamer step (256x256x512 cells of 2.5 um, 4e7
    super-particles): split store, FP32 step, FP32 correction cycles inside FP64
    refinement to 1e-10, against an FP64 run from the same seed state."""
    from mcstreamer_amd import streamer
    from mcstreamer_amd.jumc import context

    ctx = context()
    runs = {}
    for prec in ("f32", "f64"):
        t0 = time.perf_counter()
        cfg = streamer.make_config2(precision=prec, npairs=args.c2_pairs, seed=args.seed)
        sim = streamer.Streamer(cfg["popl"], cfg["fields"], cfg["ws"], cfg["eb"], cfg["dt"], cfg["ntarget"],
                                cfg["nmax"], scheme="cic")
        ctx.synchronize()
        print(f"{prec}: built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        for i in range(args.warmup):
            out = sim.step()
            print(f"{prec} warmup {i}: n={out['advance'].n_in} mg_iters={out['mg_iters']}", file=sys.stderr,
                  flush=True)
        ctx.synchronize()
        ctx.profile_reset()
        ctx.profile(True)
        psteps, vcyc, events, iters = 0, 0, None, []
        t0 = time.perf_counter()
        for i in range(args.steps):
            out = sim.step()
            psteps += out["advance"].n_in
            vcyc += max(out["mg_iters"] - 1, 0)
            iters.append(out["mg_iters"])
            ev = np.array(list(out["advance"].events)[:len(cfg["table"].proc)], dtype=np.float64)
            events = ev if events is None else events + ev
        ctx.synchronize()
        el = time.perf_counter() - t0
        ctx.profile(False)
        adv_ms, adv_n = ctx.profile_read("advance")
        f = cfg["fields"]
        runs[prec] = dict(el=el, psteps=psteps, vcycles=vcyc, events=events, iters=iters,
                          resid=out["mg_resid"], nactive=out["moments"].nactive,
                          ex=f.get("ex"), ey=f.get("ey"), ez=f.get("ez"), u=f.get("u"), q=f.get("q"),
                          adv_avg_s=adv_ms / 1e3 / max(adv_n, 1))
        print(f"{prec}: {args.steps} steps in {el:.2f} s, n={out['advance'].n_out}", file=sys.stderr, flush=True)
        del sim, cfg, f, out
        import gc
        gc.collect()
        ctx.synchronize()
    a, b = runs["f32"], runs["f64"]
    num = sum(float(np.sum((a[k] - b[k]) ** 2)) for k in ("ex", "ey", "ez"))
    den = sum(float(np.sum(b[k] ** 2)) for k in ("ex", "ey", "ez"))
    bpl = BYTES_PER_PSTEP_F32 * a["psteps"] / max(args.steps, 1)
    ach = bpl / a["adv_avg_s"] / 1e9
    return {"metric": METRIC, "value": a["psteps"] / a["el"], "unit": "particle-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": a["el"] / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 split (f64 refinement residual/update)", "data": "synthetic",
            "config": {"workload": "precision streamer step, 256x256x512 mesh h=2.5e-6 m, "
                                   f"{args.c2_pairs:g}-pair Gaussian seed (sigma 100 um), CIC, "
                                   "FP32 V(2,2) inside FP64 refinement to 1e-10 rel, merge/split 8-64, "
                                   "dt=2e-13 s", "mesh": [256, 256, 512],
                       "particles_per_step": a["psteps"] / max(args.steps, 1), "parallelism": "single GPU"},
            "vcycles_per_s": a["vcycles"] / a["el"],
            "vcycle_roofline": vcycle_roofline(a["vcycles"] / a["el"], (256, 256, 512)),
            "fp64_same_workload": {"particle_steps_per_s": b["psteps"] / b["el"],
                                   "ms_per_step": b["el"] / args.steps * 1e3},
            "errors_vs_fp64": {
                "field_l2_rel": float(np.sqrt(num / den)),
                "potential_l2_rel": float(np.linalg.norm(a["u"] - b["u"]) / np.linalg.norm(b["u"])),
                "total_charge_rel": float(abs(a["q"].sum() - b["q"].sum()) / np.abs(b["q"]).sum()),
                "electron_count_rel": float((a["nactive"] - b["nactive"]) / b["nactive"]),
                "events_rel": [float((x - y) / y) if y else None for x, y in zip(a["events"], b["events"])],
                "mg_iters_f32": a["iters"], "mg_iters_f64": b["iters"],
                "final_resid_f32": a["resid"], "final_resid_f64": b["resid"]},
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "k_advance32<1,1> (split store, FP32 push + collisions)",
                         "avg_launch_ms": a["adv_avg_s"] * 1e3, "bytes_per_launch": bpl},
            "cpu_baseline": None}


BYTES_PER_UPDATE_GS = 32   # per updated cell of a colour pass: read u, b and the neighbours' share, write u


def run_mg2d(args):
    """The reference's axisymmetric Poisson solve (MCStreamer/src/main.jl:130-145:
    CylindricalConnector{1}, Neumann on the axis, Dirichlet elsewhere, V(2,2),
    s = dz^2/eps0) on its production grid (R = 1.5 cm, L = 10 cm, 1536 x 10240
    cells) with a synthetic streamer-head charge; inputs resident on the device.
    `value` is cell-solves/s (cells x solves per second)."""
    import ctypes
    import torch
    from mcstreamer_amd import _lib
    from mcstreamer_amd import constants as co
    from mcstreamer_amd import multigrid as mg
    from mcstreamer_amd.jumc import context

    ctx = context()
    M, N = args.mg2d_grid
    R, L = 0.015, 0.1
    dr, dz = R / M, L / N
    levels = args.mg2d_levels
    rc = (np.arange(M) + 0.5) * dr
    zc = (np.arange(N) + 0.5) * dz
    # a space-charge layer at the head of an axial channel (C/m^3)
    q = 1e3 * np.exp(-(rc[:, None] / 2e-4) ** 2) * np.exp(-((zc[None, :] - 0.06) / 5e-4) ** 2)
    b = np.zeros((M + 2, N + 2))
    b[1:-1, 1:-1] = q
    conf = mg.MGConfig(bc=[(mg.BottomBnd, mg.Neumann), (mg.TopBnd, mg.Dirichlet), (mg.LeftBnd, mg.Dirichlet),
                           (mg.RightBnd, mg.Dirichlet)], s=dz * dz / co.epsilon_0, levels=levels,
                       tolerance=-1e-8, smooth1=2, smooth2=2, ndim=2, conn=mg.CylindricalConnector(1))
    ws = mg.Workspace(conf, (M, N), ctx=ctx)
    dev = torch.device("cuda", 0)
    bf = torch.from_numpy(np.ascontiguousarray(b.T).reshape(-1)).to(dev)
    u0 = torch.zeros_like(bf)
    u = torch.zeros_like(bf)
    torch.cuda.synchronize()
    it = ctypes.c_int32()
    res = ctypes.c_double()

    def solve():
        u.copy_(u0)
        torch.cuda.synchronize()
        rc_ = _lib.lib().mcs_mg_solve_device(ws.handle, ctypes.c_void_p(u.data_ptr()),
                                              ctypes.c_void_p(bf.data_ptr()), ctypes.byref(it), ctypes.byref(res))
        if rc_ != 0:
            _lib.check(ctx.handle, rc_)

    for _ in range(args.warmup):
        solve()
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
    t_solve = 0.0
    iters = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        solve()
        ctx.synchronize()
        t_solve += time.perf_counter() - t0
        iters.append(it.value)
    ctx.profile(False)
    # the finest-level sweep timed on its own: the solves above replay the cycle
    # from a HIP graph (no per-launch records), so two more solves launch phase
    # by phase on a second workspace
    os.environ["MCS_MG_GRAPH"] = "0"
    try:
        ws_eager = mg.Workspace(conf, (M, N), ctx=ctx)
    finally:
        del os.environ["MCS_MG_GRAPH"]
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(2):
        u.copy_(u0)
        torch.cuda.synchronize()
        _lib.check(ctx.handle, _lib.lib().mcs_mg_solve_device(ws_eager.handle, ctypes.c_void_p(u.data_ptr()),
                                                              ctypes.c_void_p(bf.data_ptr()), ctypes.byref(it),
                                                              ctypes.byref(res)))
    ctx.synchronize()
    ctx.profile(False)
    sw_ms, sw_n = ctx.profile_read("mg2_sweep0")
    sweep_s = sw_ms / 1e3 / max(sw_n, 1)
    cells = float(M) * N
    bytes_sweep = BYTES_PER_UPDATE_GS * cells   # two colour passes update every cell once
    ach = bytes_sweep / sweep_s / 1e9 if sweep_s > 0 else 0.0
    result = {
        "metric": "cell-solves/s (axisymmetric Poisson, production grid)",
        "value": cells * args.steps / t_solve,
        "unit": "cell-solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_solve / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"axisymmetric multigrid solve, {M}x{N} cells (R=1.5 cm, L=10 cm), "
                               f"CylindricalConnector{{1}}, V(2,2), {levels} levels, 1e-8 relative residual, "
                               "Neumann axis, Dirichlet elsewhere", "grid": [M, N], "levels": levels,
                   "fmg_iterations": iters, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k2_gs on the finest level (one sweep = two colour passes)",
                     "avg_launch_ms": sweep_s * 1e3, "bytes_per_launch": bytes_sweep},
    }
    if not args.no_cpu_baseline:
        from oracle import refcpu as Rf
        # bounded sample: whole solves of the same problem, one after another, until
        # at least 10 s of single-threaded work
        bq = np.ascontiguousarray(b.T).reshape(-1)
        nsol, itq = 0, []
        t0 = time.perf_counter()
        el = 0.0
        while nsol == 0 or (el < 10.0 and nsol < 20):
            uq = np.zeros_like(bq)
            rcq, iq, _ = Rf.mg_solve((M, N), (mg.Neumann, mg.Dirichlet, mg.Dirichlet, mg.Dirichlet),
                                     dz * dz / co.epsilon_0, levels, 2, 2, -1e-8, 50, uq, bq, cyl=1)
            itq.append(iq)
            nsol += 1
            el = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": cells * nsol / el, "unit": "cell-solves/s", "cores": 1,
                                  "kind": "port",
                                  "sample": f"{nsol} solves of the same problem ({itq[0]} fmg iterations each), "
                                            f"single-threaded restatement, {el:.1f} s"}
    else:
        result["cpu_baseline"] = None
    return result


def run_axisym(args):
    """The reference's production geometry (MCStreamer/samples/input.toml: R = 1.5 cm,
    L = 10 cm, 1536 x 10240 cells, eb = 15 kV/cm, dt = 1e-12 s, Poisson every step):
    one step is poisson! (cylindrical density!, charge, solve_freebc!'s two solves
    of the axisymmetric multigrid with 8 levels and the reference's absolute
    tolerance 1e8 e around the free radial boundary, faces), advance! in the
    axisymmetric FieldInterp with the ion tracker, repack!, and resample! per
    (r, z) cell with the sample input's ntarget = 10, nmax = 15.  `--axisym-n`
    electrons over as many ions fill a column (see below); C1 gas (N2-like +
    20 % O2-like)."""
    from mcstreamer_amd import axisym, gases, jumc, lxcat
    from mcstreamer_amd import constants as co

    ctx = jumc.context()
    e_ev = np.geomspace(1e-3, 1e3, 4000)
    path = os.path.join(tempfile.gettempdir(), f"mcs_bench_gas_c1_{os.getpid()}.json")
    gases.write_gas(path, e_ev, with_attacher=True)
    proc, rate, maxrate, _ = lxcat.load_lxcat([path], gases.densities(0.2), e_ev * co.eV)
    table = jumc.CollisionTable(proc, e_ev * co.eV, rate, maxrate, egrid="log")
    M, N = args.mg2d_grid
    grid = axisym.Grid(0.015, 0.1, M, N)
    fields = axisym.GridFields(grid)
    n = int(args.axisym_n)
    rng = np.random.default_rng(args.seed)
    # a plasma column of radius 4 mm over z = 1..5 cm, r uniform so that every cell
    # of the column holds about n / (410 x 4096) = 12 electrons: inside the
    # resampling window (ntarget 10, nmax 15 of the sample input), so the
    # population keeps its size from step to step
    r = rng.uniform(0.0, 4e-3, n)
    th = rng.uniform(0.0, 2 * np.pi, n)
    x = np.stack([r * np.cos(th), r * np.sin(th), rng.uniform(1e-2, 5e-2, n)], axis=1)
    del r, th
    v = np.zeros((n, 3))
    s = -np.log1p(-rng.uniform(size=n))
    popl = jumc.Population(3 * n, jumc.ParticleStates(x, v, np.ones(n), s, np.ones(n, dtype=np.uint8)), table,
                           seed=args.seed)
    del x, v, s
    axisym.density(fields, "qfixed", popl, 1.0)
    # 8 levels (6 x 40 coarsest): the reference's default 9 leaves 3 cells in r, an
    # odd size for which redblack2 (redblack.jl:38-46) visits the ghost column
    conf = axisym.poisson_config(grid, levels=args.mg2d_levels)
    ws = axisym.mgmod.Workspace(conf, (M, N))
    field, tracker = axisym.FieldInterp(fields), axisym.CollisionTracker(fields)
    eb, dt = -1.5e6, 1e-12   # geteb (main.jl:346-357) of eb = 15 kV/cm

    def step(i):
        it, _ = axisym.poisson(fields, popl, eb, ws)
        st = jumc.advance(popl, field, dt, tracker, step=i)
        jumc.repack(popl)
        axisym.resample(popl, fields, 10, 15, step=i)
        return it, st

    for i in range(args.warmup):
        it, st = step(i)
        print(f"warmup {i}: n={st.n_in} mg_iters={it} births={st.births}", file=sys.stderr, flush=True)
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(True)
    psteps, iters = 0, []
    t0 = time.perf_counter()
    for i in range(args.steps):
        it, st = step(args.warmup + i)
        psteps += st.n_in
        iters.append(it)
    ctx.synchronize()
    el = time.perf_counter() - t0
    ctx.profile(False)
    kt = {}
    for k in ("advance", "cyl_density", "cyl_faces", "cyl_freebc", "mg2_sweep0", "cyl_sort", "resample"):
        ms, cnt = ctx.profile_read(k)
        kt[k] = {"ms_total": ms, "launches": cnt}
    adv_s = kt["advance"]["ms_total"] / 1e3 / max(kt["advance"]["launches"], 1)
    bytes_per_launch = BYTES_PER_PSTEP_F64 * psteps / max(args.steps, 1)
    ach = bytes_per_launch / adv_s / 1e9 if adv_s > 0 else 0.0
    result = {
        "metric": "particle-steps/s (axisymmetric streamer step, production grid)",
        "value": psteps / el, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"FP64 axisymmetric streamer step, {M}x{N} cells (R=1.5 cm, L=10 cm), "
                               f"{n:g} electrons in a 4 mm x 4 cm column (~12 per cell), eb=15 kV/cm, dt=1e-12 s, "
                               f"poisson! with solve_freebc! ({conf.levels} levels, tol 1e8 e) + advance! "
                               "(cylindrical field, tracker) + repack! + resample! (10..15 per cell)",
                   "grid": [M, N], "particles_per_step": psteps / max(args.steps, 1), "fmg_iterations": iters,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_advance<2,1> (fused push + collisions, axisymmetric gather, ion tracker)",
                     "avg_launch_ms": adv_s * 1e3, "bytes_per_launch": bytes_per_launch},
        "kernels_ms": kt,
    }
    if not args.no_cpu_baseline:
        from oracle import refcpu as Rf
        # bounded sample: one oracle poisson! on the full grid from the device state,
        # and oracle advance! steps of the first particles until 8 s; the rate is
        # particles / (poisson time + per-particle advance time x particles)
        ps = popl.download()
        nn = len(ps)
        G = Rf.CGrid(0.015, 0.1, M, N)
        G.qfixed[:] = fields.get("qfixed").ravel()
        G.u[:] = fields.get("u").ravel()
        P = Rf.Pop(ps.x, ps.v, ps.w, ps.s, ps.active, ps.id, cap=3 * nn + 1000, next_id=ps.next_id)
        t0 = time.perf_counter()
        itq = Rf.cyl_poisson(G, P, conf, eb, freebc=True)
        t_poisson = time.perf_counter() - t0
        k = min(nn, 200_000)
        S = Rf.Pop(ps.x[:k], ps.v[:k], ps.w[:k], ps.s[:k], ps.active[:k], ps.id[:k], cap=3 * k + 1000,
                   next_id=ps.next_id)
        done, t0 = 0, time.perf_counter()
        el_a = 0.0
        while done == 0 or el_a < 8.0:
            done += S.n
            Rf.advance_cyl(S, Rf.Table(table.energy, table.rate, table.kind, table.param, table.maxrate,
                                       egrid=table.egrid), G, dt, track=True, seed=args.seed, step=10_000 + done)
            Rf.repack(S)
            el_a = time.perf_counter() - t0
        per_p = el_a / done
        result["cpu_baseline"] = {
            "value": nn / (t_poisson + per_p * nn), "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"one oracle poisson! on the full grid ({itq} fmg iterations, {t_poisson:.1f} s) plus "
                      f"{done} oracle particle-steps of the first {k} particles ({el_a:.1f} s), scaled to the "
                      f"{nn} particles of one step; single-threaded"}
    else:
        result["cpu_baseline"] = None
    return result


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device copy of known size: its FETCH_SIZE/WRITE_SIZE counters give
    the factor that turns the counters of the step's kernels into bytes."""
    import torch

    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy calibration: {reps} x {nbytes} bytes read and written", file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["streamer", "swarm", "batch", "serve", "precision", "mg2d", "axisym",
                                           "swarmblock"], default="streamer")
    ap.add_argument("--c2-pairs", type=int, default=40_000_000)
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--solve-steps", type=int, default=5000)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-drain", type=float, default=1.5)
    ap.add_argument("--serve-capacity", type=int, default=200_000_000)
    ap.add_argument("--serve-plain", type=int, default=0)
    ap.add_argument("--precision", choices=["f64", "f32"], default="f32",
                    help="store and step precision of the batched/served swarm workloads")
    ap.add_argument("--mesh", type=int, nargs=3, default=[128, 128, 256])
    ap.add_argument("--mg2d-grid", type=int, nargs=2, default=[1536, 10240])
    ap.add_argument("--mg2d-levels", type=int, default=8)
    ap.add_argument("--axisym-n", type=float, default=2e7)
    ap.add_argument("--h", type=float, default=5e-6)
    ap.add_argument("--npairs", type=int, default=1_000_000)
    ap.add_argument("--swarm-n", type=int, default=10_000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-swarm", action="store_true",
                    help="leave the swarm block (batched and served swarms) out of the streamer line")
    ap.add_argument("--late-at", type=int, default=480, help="step of the run at which the late sample starts")
    ap.add_argument("--late-steps", type=int, default=20, help="timed steps of the late sample (0: none)")
    ap.add_argument("--c3-steps", type=int, default=10)
    ap.add_argument("--c3-plain-sample", type=int, default=32)
    ap.add_argument("--c3-plain-relax", type=int, default=100)
    ap.add_argument("--c3-plain-steps", type=int, default=20)
    ap.add_argument("--c4-seconds", type=float, default=15.0)
    ap.add_argument("--c4-low-rate", type=float, default=1.0)
    ap.add_argument("--c4-low-seconds", type=float, default=30.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the particle-mesh step is measured at N=1 "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)
    res = {"streamer": run_streamer, "swarm": run_swarm, "batch": run_batch,
           "serve": run_serve, "precision": run_precision, "mg2d": run_mg2d,
           "axisym": run_axisym, "swarmblock": swarm_block}[args.workload](args)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
