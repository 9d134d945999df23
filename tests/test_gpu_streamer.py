"""GPU parity of whole streamer steps (MCStreamer.nsteps, main.jl:217-315, with
Poisson and resampling every step) against the oracle composed in the same
order: deposit, charge, multigrid solve, faces, advance with the mesh field and
ion tracker, canonical sort, per-cell resample.  Ids, counts and events are
bit-exact; floating state is compared at 1e-12 relative (the ion tracker's
atomics may reorder its sums)."""
import numpy as np
import pytest

from mcstreamer_amd import multigrid as mgmod
from mcstreamer_amd import streamer
from oracle import refcpu as R
from tests.common import ref_table

pytestmark = pytest.mark.gpu


def test_streamer_steps_parity():
    n = (16, 16, 32)
    cfg = streamer.make_config1(n=n, h=5e-6, npairs=20000, seed=3, cap=200000)
    sim = streamer.Streamer(cfg["popl"], cfg["fields"], cfg["ws"], cfg["eb"], cfg["dt"] * 5,
                            cfg["ntarget"], cfg["nmax"], scheme="cic")
    # oracle copy of the initial state (after the seed's sort and ion deposit)
    ps = sim.popl.download()
    P = R.Pop(ps.x, ps.v, ps.w, ps.s, ps.active, ps.id, cap=200000, next_id=ps.next_id)
    M = R.Mesh(n, 5e-6)
    M.qfixed[:] = sim.fields.get("qfixed")
    t = cfg["table"]
    T = R.Table(t.energy, t.rate, t.kind, t.param, t.maxrate, egrid=t.egrid)
    conf = cfg["conf"]
    R.sort_by_cell(P, M)
    for i in range(4):
        out = sim.step()
        M.qpart[:] = 0.0
        R.deposit(P, M, 1, 1, -1.0)
        R.assemble_charge(M)
        rc, it, res = R.mg_solve(n, conf.bc, conf.s, conf.levels, conf.smooth1, conf.smooth2,
                                 conf.tolerance, conf.maxiter, M.u, M.q)
        R.efield_faces(M, sim.eb)
        st = R.advance(P, T, R.mesh_field(M), sim.dt, M, seed=sim.popl.seed, step=i)
        R.sort_by_cell(P, M)
        rc2, rst = R.resample_cells(P, M, sim.ntarget, sim.nmax, seed=sim.popl.seed, step=i)
        a = out["advance"]
        print(f"step {i}: n {P.n} mg {it}/{out['mg_iters']} births {st.births} deaths {st.deaths}", flush=True)
        assert out["mg_iters"] == it
        assert (a.births, a.deaths) == (st.births, st.deaths)
        assert list(a.events)[:len(T.kind)] == list(st.events)[:len(T.kind)]
        assert out["resample"].n_out == rst.n_out
    g = sim.popl.download()
    r = P.view()
    np.testing.assert_array_equal(g.id, r["id"])
    for k in ("x", "v", "w", "s"):
        np.testing.assert_allclose(getattr(g, k), r[k], rtol=1e-12, atol=0, err_msg=k)
    np.testing.assert_allclose(sim.fields.get("u"), M.u, rtol=1e-10, atol=1e-12 * np.abs(M.u).max())
