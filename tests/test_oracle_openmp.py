"""The oracle's OpenMP build (the all-core CPU baseline) against its
single-threaded build on the same inputs: the chunked particle loop (births in
parent order, the tracker's deposits in particle order), the per-cell gather
deposit and the threaded grid loops give bitwise the same particles, charge and
fields; the multigrid solve the same iteration count (its residual norm is the
one sum that is not threaded)."""
import numpy as np
import pytest

from oracle import refcpu as R
from tests.common import gas_table, random_particles, ref_table, smooth_potential


@pytest.fixture
def both():
    yield
    R.use_openmp(False)


def _chain(omp, n, h, npart, seed):
    prev = R.use_openmp(omp)
    try:
        proc, energy, rate, maxrate = gas_table(o2=0.25)
        T = ref_table(proc, energy, rate, maxrate)
        L = np.array(n) * h
        x, v, w, s, act = random_particles(npart, L, seed, p_inactive=0.05, p_outside=0.02, kT_ev=8.0)
        P = R.Pop(x, v, w, s, act, cap=4 * npart)
        M = R.Mesh(n, h)
        M.u[:] = smooth_potential(n, h, amp=3.0)
        R.efield_faces(M, -8e6)
        out = {}
        for step in range(2):
            st = R.advance(P, T, R.mesh_field(M), 2e-12, M, seed=seed, step=step)
            out[f"births{step}"] = (st.births, st.deaths, st.flights, list(st.events)[: len(proc)])
            R.sort_by_cell(P, M)
            M.qpart[:] = 0.0
            R.deposit(P, M, 1, 1, -1.0)
            R.deposit(P, M, 0, 0, 0.5)
            R.assemble_charge(M)
            rc, it, res = R.mg_solve(n, [1, 1, 1, 1, 0, 0], h * h / 8.854e-12, 3, 2, 2, -1e-9, 50, M.u, M.q)
            out[f"mg{step}"] = it
            R.efield_faces(M, -8e6)
            R.resample_cells(P, M, 8, 64, seed=seed, step=step)
        out["pop"] = P.view()
        for k in ("qfixed", "qpart", "q", "u", "ex", "ey", "ez"):
            out[k] = getattr(M, k).copy()
        return out
    finally:
        R.use_openmp(prev)


def test_openmp_build_matches_the_checker(both):
    n, h = (16, 16, 24), 5e-6
    a = _chain(False, n, h, 20000, 4)
    b = _chain(True, n, h, 20000, 4)
    for step in range(2):
        assert a[f"births{step}"] == b[f"births{step}"]
        assert a[f"mg{step}"] == b[f"mg{step}"]
    assert a["births0"][0] > 0 and a["births1"][0] > 0
    for k in ("id", "active", "x", "v", "w", "s"):
        np.testing.assert_array_equal(a["pop"][k], b["pop"][k], err_msg=k)
    for k in ("qfixed", "qpart", "q"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    for k in ("u", "ex", "ey", "ez"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-12, atol=1e-12 * np.abs(a[k]).max(), err_msg=k)
