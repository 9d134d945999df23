"""Pins of the oracle's mesh operators by mathematics, on the CPU.

- The canonical (cell, id) order of ref_sort_by_cell / ref_cyl_sort_by_cell
  against numpy's lexsort over independently computed cell keys, and the cell
  starts against a cumulative bincount.  The cell keys follow cellindex
  (MCStreamer/src/grid.jl:84-90) in 3D Cartesian form and resample!'s ghosted
  per-cell arrays (MCStreamer/src/resample.jl:21-23) in (r, z).
- The cloud-in-cell deposit (the trilinear form of density!,
  MCStreamer/src/poisson.jl:65-84, with the cell-centred staggering of
  gridfields.jl:94-109) against the tent assignment function
  W(x) = prod_d max(0, 1 - |x_d - c_d| / h), a different formula for the same
  weights; charge and first-moment conservation; the special points (cell
  centre, cell corner); the shares that fall outside the mesh.
- The reciprocal in normdivrem (gridfields.jl:94-95, @fastmath x / y read as
  x * (1/y)): how often it moves a cell index against the true quotient, and
  that the interpolated field does not change beyond rounding when it does.

Each pin is checked for sensitivity: a plausible mistake (a transposed cell key,
a dropped half-cell stagger) made in a numpy copy of the expectation must fail
the same comparison.
"""
import ctypes

import numpy as np
import pytest

from oracle import refcpu as R


def _pop(x, ids=None, active=None, w=None):
    n = len(x)
    return R.Pop(x, np.zeros((n, 3)), np.ones(n) if w is None else w, np.ones(n),
                 np.ones(n, dtype=np.uint8) if active is None else active, ids)


def _keys3(x, n, h):
    """cellindex in 3D: floor(x_d / h) per axis, flat i + nx (j + ny k); outside -> ncells."""
    nx, ny, nz = n
    f = np.floor(x / h)
    ok = np.all((f >= 0) & (f < np.array(n)), axis=1)
    fi = np.where(ok[:, None], f, 0).astype(np.int64)
    key = fi[:, 0] + nx * (fi[:, 1] + ny * fi[:, 2])
    return np.where(ok, key, nx * ny * nz)


def _sorted_expectation(keys, ids, active):
    a = np.flatnonzero(active)
    order = a[np.lexsort((ids[a], keys[a]))]
    return order


# ----------------------------------------------------------- canonical sort, 3D


@pytest.mark.parametrize("seed", range(6))
def test_sort_by_cell_is_lexsort_of_cell_then_id(seed):
    rng = np.random.default_rng(seed)
    n = (5, 7, 3)   # unequal sides: a transposed key orders differently
    h = 1e-5
    L = np.array(n) * h
    npart = int(rng.integers(50, 400))
    x = rng.uniform(-0.15, 1.15, size=(npart, 3)) * L    # some outside on every face
    ids = rng.permutation(10 * npart)[:npart].astype(np.uint64)  # ids not in slot order
    act = (rng.uniform(size=npart) < 0.8).astype(np.uint8)
    P = _pop(x, ids, act)
    M = R.Mesh(n, h)
    R.sort_by_cell(P, M)
    v = P.view()

    keys = _keys3(x, n, h)
    order = _sorted_expectation(keys, ids, act)
    assert P.n == int(act.sum())
    np.testing.assert_array_equal(v["id"], ids[order])
    np.testing.assert_array_equal(v["x"], x[order])
    ncells = int(np.prod(n))
    cs = np.zeros(ncells + 2, dtype=np.int64)
    cs[1:] = np.cumsum(np.bincount(keys[act == 1], minlength=ncells + 1))
    np.testing.assert_array_equal(M.cell_start, cs)
    # outside particles sit last, in id order
    out = keys[order] == ncells
    assert not np.any(np.diff(out.astype(int)) < 0)

    # sensitivity: a key with x and z exchanged gives another order
    f = np.floor(x / h).astype(np.int64)
    ok = keys < ncells
    tkeys = np.where(ok, f[:, 2] + n[2] * (f[:, 1] + n[1] * f[:, 0]), ncells)
    assert not np.array_equal(ids[_sorted_expectation(tkeys, ids, act)], v["id"])


def test_sort_by_cell_empty_and_all_inactive():
    M = R.Mesh((4, 4, 4), 1.0)
    P = _pop(np.zeros((0, 3)))
    R.sort_by_cell(P, M)
    assert P.n == 0 and not M.cell_start.any()
    P = _pop(np.full((9, 3), 0.5), active=np.zeros(9, dtype=np.uint8))
    R.sort_by_cell(P, M)
    assert P.n == 0 and not M.cell_start.any()


# ----------------------------------------------------- canonical sort, (r, z)


@pytest.mark.parametrize("seed", range(6))
def test_cyl_sort_by_cell_is_lexsort_of_cell_then_id(seed):
    rng = np.random.default_rng(100 + seed)
    Rr, Lz, M, N = 1e-3, 3e-3, 6, 9
    G = R.CGrid(Rr, Lz, M, N)
    drf, dzf = G.steps[2], G.steps[3]
    npart = int(rng.integers(50, 400))
    rho = rng.uniform(0, 1.4 * Rr, npart)
    phi = rng.uniform(0, 2 * np.pi, npart)
    z = rng.uniform(-0.3 * Lz, 1.3 * Lz, npart)
    x = np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)
    ids = rng.permutation(10 * npart)[:npart].astype(np.uint64)
    act = (rng.uniform(size=npart) < 0.8).astype(np.uint8)
    P = _pop(x, ids, act)
    R.cyl_sort_by_cell(P, G)
    v = P.view()

    ncg = (M + 2) * (N + 2)
    i = np.floor(np.hypot(x[:, 0], x[:, 1]) / drf).astype(np.int64)
    j = np.floor(x[:, 2] / dzf).astype(np.int64)
    ok = (i >= -1) & (i <= M) & (j >= -1) & (j <= N)
    keys = np.where(ok, (i + 1) + (M + 2) * (j + 1), ncg)
    order = _sorted_expectation(keys, ids, act)
    np.testing.assert_array_equal(v["id"], ids[order])
    cs = np.zeros(ncg + 2, dtype=np.int64)
    cs[1:] = np.cumsum(np.bincount(keys[act == 1], minlength=ncg + 1))
    np.testing.assert_array_equal(G.cell_start, cs)

    tkeys = np.where(ok, (j + 1) + (N + 2) * (i + 1), ncg)
    assert not np.array_equal(ids[_sorted_expectation(tkeys, ids, act)], v["id"])


# ------------------------------------------------------------ CIC deposit


def _tent_deposit(x, w, val, n, h):
    """sum_p w val / dV * prod_d max(0, 1 - |x_d - (i_d + 1/2) h| / h), interior
    cells only, for particles inside the mesh.  Ghosted layout (x fastest)."""
    nx, ny, nz = n
    out = np.zeros((nz + 2, ny + 2, nx + 2))
    dV = h * h * h
    inside = _keys3(x, n, h) < nx * ny * nz
    for p in np.flatnonzero(inside):
        c = np.floor(x[p] / h).astype(int)
        for k in range(c[2] - 1, c[2] + 2):
            for j in range(c[1] - 1, c[1] + 2):
                for i in range(c[0] - 1, c[0] + 2):
                    if not (0 <= i < nx and 0 <= j < ny and 0 <= k < nz):
                        continue
                    cc = (np.array([i, j, k]) + 0.5) * h
                    W = np.prod(np.maximum(0.0, 1.0 - np.abs(x[p] - cc) / h))
                    out[k + 1, j + 1, i + 1] += w[p] * val / dV * W
    return out.reshape(-1)


def _oracle_cic(x, w, val, n, h):
    P = _pop(x, w=w)
    M = R.Mesh(n, h)
    R.deposit(P, M, 1, 1, val)
    return M.qpart


@pytest.mark.parametrize("seed", range(4))
def test_cic_equals_tent_assignment(seed):
    rng = np.random.default_rng(200 + seed)
    n, h = (6, 5, 4), 2.5e-6
    L = np.array(n) * h
    npart = 150
    x = rng.uniform(-0.1, 1.1, size=(npart, 3)) * L
    w = rng.uniform(0.5, 3.0, npart)
    got = _oracle_cic(x, w, 1.0, n, h)
    ref = _tent_deposit(x, w, 1.0, n, h)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * scale)
    # sensitivity: the same tent without the half-cell stagger is far off
    shifted = _tent_deposit(x + 0.5 * h, w, 1.0, n, h)
    assert np.abs(shifted - got).max() > 1e-2 * scale


def test_cic_conserves_charge_and_first_moment_inside():
    rng = np.random.default_rng(7)
    n, h = (8, 8, 8), 1e-5
    L = np.array(n) * h
    # interior: every one of the eight cells of a particle's cloud lies in the mesh,
    # also after the half-cell shift of the sensitivity check below
    x = rng.uniform(0.5 * h, L - 1.0 * h, size=(300, 3))
    w = rng.uniform(0.5, 3.0, 300)
    val = -2.0
    q = _oracle_cic(x, w, val, n, h).reshape(n[2] + 2, n[1] + 2, n[0] + 2)
    dV = h ** 3
    total = (w * val).sum()
    np.testing.assert_allclose(q.sum() * dV, total, rtol=1e-13)
    k, j, i = np.meshgrid(np.arange(n[2] + 2), np.arange(n[1] + 2), np.arange(n[0] + 2), indexing="ij")
    centres = [(i - 0.5) * h, (j - 0.5) * h, (k - 0.5) * h]   # ghost-inclusive index -> centre
    for d in range(3):
        np.testing.assert_allclose((q * centres[d]).sum() * dV, (w * val * x[:, d]).sum(), rtol=1e-12)
    # sensitivity: a half-cell shift moves the first moment by h/2 times the charge
    qs = _oracle_cic(x + 0.5 * h, w, val, n, h).reshape(q.shape)
    assert abs((qs * centres[0]).sum() * dV - (w * val * x[:, 0]).sum()) > 0.4 * h * abs(total)


def test_cic_special_points():
    n, h = (4, 4, 4), 1.0
    # a particle at a cell centre puts everything into that cell
    q = _oracle_cic(np.array([[1.5, 2.5, 0.5]]), np.array([2.0]), 1.0, n, h).reshape(6, 6, 6)
    assert q[1, 3, 2] == 2.0 and np.count_nonzero(q) == 1
    # a particle at an interior cell corner shares equally among the eight cells around it
    q = _oracle_cic(np.array([[2.0, 1.0, 3.0]]), np.array([1.0]), 1.0, n, h).reshape(6, 6, 6)
    nz = np.argwhere(q)
    assert len(nz) == 8 and np.all(q[q != 0] == 0.125)
    assert sorted(set(nz[:, 2])) == [2, 3] and sorted(set(nz[:, 1])) == [1, 2] and sorted(set(nz[:, 0])) == [3, 4]
    # offsets (0.25, 0.5, 0.75) of a cell from a centre: weights (0.75, 0.25) x (0.5, 0.5) x (0.25, 0.75)
    q = _oracle_cic(np.array([[1.75, 2.0, 2.25]]), np.array([1.0]), 1.0, n, h).reshape(6, 6, 6)
    assert q[2, 2, 2] == 0.75 * 0.5 * 0.25 and q[3, 3, 3] == 0.25 * 0.5 * 0.75


def test_cic_drops_shares_outside_the_mesh():
    n, h = (4, 4, 4), 1.0
    # x = 0.25 h: the cloud spans centres -0.5 (outside, share 0.25) and 0.5 (share 0.75)
    q = _oracle_cic(np.array([[0.25, 2.0, 2.0]]), np.array([1.0]), 1.0, n, h)
    assert abs(q.sum() - 0.75) < 1e-15
    # a corner particle keeps only the octant inside
    q = _oracle_cic(np.array([[0.25, 0.25, 0.25]]), np.array([1.0]), 1.0, n, h)
    assert abs(q.sum() - 0.75 ** 3) < 1e-15
    # a particle outside the mesh deposits nothing, even where its cloud would overlap
    q = _oracle_cic(np.array([[-0.25, 2.0, 2.0], [4.1, 1.0, 1.0]]), np.array([1.0, 1.0]), 1.0, n, h)
    assert not q.any()
    # inactive particles deposit nothing
    P = _pop(np.array([[1.5, 1.5, 1.5]]), active=np.zeros(1, dtype=np.uint8))
    M = R.Mesh(n, h)
    R.deposit(P, M, 1, 1, 1.0)
    assert not M.qpart.any()


def test_ngp_equals_cell_binning():
    rng = np.random.default_rng(11)
    n, h = (5, 6, 7), 3e-6
    L = np.array(n) * h
    x = rng.uniform(-0.1, 1.1, size=(400, 3)) * L
    w = rng.uniform(0.5, 3.0, 400)
    P = _pop(x, w=w)
    M = R.Mesh(n, h)
    R.deposit(P, M, 0, 0, 3.0)
    keys = _keys3(x, n, h)
    ncells = int(np.prod(n))
    ok = keys < ncells
    ref = np.zeros((n[2] + 2, n[1] + 2, n[0] + 2))
    f = np.floor(x[ok] / h).astype(int)
    np.add.at(ref, (f[:, 2] + 1, f[:, 1] + 1, f[:, 0] + 1), w[ok] * 3.0 / h ** 3)
    np.testing.assert_allclose(M.qfixed, ref.reshape(-1), rtol=1e-14, atol=0)


# ------------------------------------------------- normdivrem's reciprocal


def _field_np(ex, ey, ez, n, h, x, divide):
    """FieldInterp in 3D (gridfields.jl:94-190) in numpy, normdivrem by x / h
    (divide) or x * (1/h)."""
    nx, ny, nz = n
    q = x / h if divide else x * (1.0 / h)
    m = np.minimum(np.trunc(q).astype(np.int64), np.array(n) - 1)
    d = q - m
    m1 = np.where(d < 0.5, m, m + 1)
    d1 = np.where(d < 0.5, d + 0.5, d - 0.5)

    def tri(a, sx, sxy, ix, iy, iz, dx, dy, dz):
        b = ix + sx * iy + sxy * iz
        c00 = a[b] * (1 - dx) + a[b + 1] * dx
        c10 = a[b + sx] * (1 - dx) + a[b + sx + 1] * dx
        c01 = a[b + sxy] * (1 - dx) + a[b + sxy + 1] * dx
        c11 = a[b + sxy + sx] * (1 - dx) + a[b + sxy + sx + 1] * dx
        return (c00 * (1 - dy) + c10 * dy) * (1 - dz) + (c01 * (1 - dy) + c11 * dy) * dz

    e0 = tri(ex, nx + 1, (nx + 1) * (ny + 2), m[:, 0], m1[:, 1], m1[:, 2], d[:, 0], d1[:, 1], d1[:, 2])
    e1 = tri(ey, nx + 2, (nx + 2) * (ny + 1), m1[:, 0], m[:, 1], m1[:, 2], d1[:, 0], d[:, 1], d1[:, 2])
    e2 = tri(ez, nx + 2, (nx + 2) * (ny + 2), m1[:, 0], m1[:, 1], m[:, 2], d1[:, 0], d1[:, 1], d[:, 2])
    return np.stack([e0, e1, e2], axis=1), m


def test_normdivrem_reciprocal_is_continuous_across_the_cells_it_moves():
    rng = np.random.default_rng(5)
    n, h = (16, 16, 32), 5e-6
    nx, ny, nz = n
    M = R.Mesh(n, h)
    M.ex[:] = rng.normal(size=M.ex.size) * 1e6
    M.ey[:] = rng.normal(size=M.ey.size) * 1e6
    M.ez[:] = rng.normal(size=M.ez.size) * 1e6
    # a dense sample at cell edges (where the two forms can disagree) and inside
    k = rng.integers(1, np.array(n), size=(200000, 3))
    x = k * h
    x[::2] += rng.uniform(-2.0, 2.0, size=x[::2].shape) * np.spacing(x[::2])
    x[1::4] = rng.uniform(0.0, 1.0, size=x[1::4].shape) * np.array(n) * h
    x = np.clip(x, 1e-12, np.array(n) * h * (1 - 1e-12))
    e_div, m_div = _field_np(M.ex, M.ey, M.ez, n, h, x, True)
    e_rec, m_rec = _field_np(M.ex, M.ey, M.ez, n, h, x, False)
    moved = np.any(m_div != m_rec, axis=1)
    assert moved.sum() > 0   # the sample does reach positions where the forms differ
    assert moved.mean() < 0.5
    scale = np.abs(M.ez).max()
    assert np.abs(e_div - e_rec).max() < 1e-9 * scale

    # the oracle's gather is the reciprocal form, value for value
    f = R.mesh_field(M)
    e = np.zeros(3)
    fl = ctypes.c_int32(0)
    for p in rng.choice(len(x), 300, replace=False):
        R.lib().ref_efield(ctypes.byref(f), R._d(np.ascontiguousarray(x[p])), R._d(e), ctypes.byref(fl))
        np.testing.assert_array_equal(e, e_rec[p])
