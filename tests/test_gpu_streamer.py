"""GPU parity of whole streamer steps (MCStreamer.nsteps, main.jl:217-315, with
Poisson and resampling every step) against the oracle composed in the same
order: deposit, charge, multigrid solve, faces, advance with the mesh field and
ion tracker, canonical sort, per-cell resample.  Ids, counts and events are
bit-exact; floating state is compared at 1e-12 relative (the 3D solver's
tabulated tail and its residual norms round differently from the oracle's
sweeps and sums, which moves the potential in the last bits)."""
import numpy as np
import pytest

from mcstreamer_amd import multigrid as mgmod
from mcstreamer_amd import streamer
from oracle import refcpu as R
from tests.common import ref_table

pytestmark = pytest.mark.gpu


def _steps_vs_oracle(n, npairs, steps, dt_factor, cap, seed=3):
    """`steps` whole streamer steps on the GPU and the same chain in the oracle;
    ids, counts, events and resample outcomes bit-exact on the way; returns the
    simulation and the oracle's population and mesh."""
    cfg = streamer.make_config1(n=n, h=5e-6, npairs=npairs, seed=seed, cap=cap)
    sim = streamer.Streamer(cfg["popl"], cfg["fields"], cfg["ws"], cfg["eb"], cfg["dt"] * dt_factor,
                            cfg["ntarget"], cfg["nmax"], scheme="cic")
    # oracle copy of the initial state (after the seed's sort and ion deposit)
    ps = sim.popl.download()
    P = R.Pop(ps.x, ps.v, ps.w, ps.s, ps.active, ps.id, cap=cap, next_id=ps.next_id)
    M = R.Mesh(n, 5e-6)
    M.qfixed[:] = sim.fields.get("qfixed")
    t = cfg["table"]
    T = R.Table(t.energy, t.rate, t.kind, t.param, t.maxrate, egrid=t.egrid)
    conf = cfg["conf"]
    R.sort_by_cell(P, M)
    for i in range(steps):
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
    return sim, P, M


def _compare_state(sim, P, M, u_rtol):
    g = sim.popl.download()
    r = P.view()
    np.testing.assert_array_equal(g.id, r["id"])
    np.testing.assert_array_equal(g.active, r["active"])
    # vectors to 1e-12 of each particle's |x| or |v| (a component near zero is a
    # difference of large terms), scalars to 1e-12 relative
    for k in ("x", "v"):
        a, b = getattr(g, k), r[k]
        err = np.abs(a - b).max(axis=1)
        bound = 1e-12 * np.linalg.norm(b, axis=1)
        assert np.all(err <= bound), f"{k}: {np.count_nonzero(err > bound)} particles, worst {np.max(err / np.maximum(bound, 1e-300)):.3g} x bound"
    for k in ("w", "s"):
        np.testing.assert_allclose(getattr(g, k), r[k], rtol=1e-12, atol=0, err_msg=k)
    np.testing.assert_allclose(sim.fields.get("u").ravel(), M.u, rtol=u_rtol, atol=u_rtol * np.abs(M.u).max())
    np.testing.assert_allclose(sim.fields.get("qfixed").ravel(), M.qfixed, rtol=1e-12,
                               atol=1e-12 * np.abs(M.qfixed).max())


def test_streamer_steps_parity():
    sim, P, M = _steps_vs_oracle((16, 16, 32), 20000, 4, 5, 200000)
    _compare_state(sim, P, M, 1e-10)


def test_streamer_steps_parity_full_config1():
    """Two whole steps of the 128 x 128 x 256 workload with its 1e6-pair seed:
    the production paths that switch on size (colour-split finest level, tiled
    smoother, dense CIC gather) against the oracle at their real shape."""
    sim, P, M = _steps_vs_oracle((128, 128, 256), 1_000_000, 2, 1, 2_000_000, seed=7)
    assert P.n > 900_000   # the seed's pairs after two per-cell resamples
    _compare_state(sim, P, M, 1e-10)
