"""Pins of the two-dimensional multigrid restatement (Multigrid.jl with N = 2)
and of its cylindrical connector (Multigrid.jl:31-39): the connector's operator
on r^2 and z^2 (closed forms 4 and 2), the solve against a dense numpy solve of
the same 5-point operator with the connector and the ghost rules of
boundaries.jl, the transfer weights, the coarse inverse of the Cartesian
operator that buildmat builds, and second order on a manufactured axisymmetric
solution."""
import numpy as np
import pytest

from oracle import refcpu as R

D, N = 0, 1


def conn(cyl, d, sgn, i):
    """Connector factor of the neighbour at offset sgn along d of the cell with
    interior index i along d (CylindricalConnector{cyl}: x (1 + sgn / (2 i + 1)))."""
    return 1.0 + sgn / (2 * i + 1) if cyl == d + 1 else 1.0


def dense_operator2(n, bc, cyl):
    nx, ny = n
    A = np.zeros((nx * ny, nx * ny))
    for j in range(ny):
        for i in range(nx):
            r = i + nx * j
            A[r, r] -= 4
            for d, s in ((0, 1), (1, 1), (0, -1), (1, -1)):
                c = [i, j]
                c[d] += s
                w = conn(cyl, d, s, (i, j)[d])
                if c[d] < 0 or c[d] >= n[d]:
                    face = 2 * d + (1 if s > 0 else 0)
                    A[r, r] += w * (-1.0 if bc[face] == D else 1.0)
                else:
                    A[r, c[0] + nx * c[1]] += w
    return A


def interior2(a, n):
    return a.reshape(n[1] + 2, n[0] + 2)[1:-1, 1:-1].reshape(-1)


def with_ghosts2(v, n):
    a = np.zeros((n[1] + 2, n[0] + 2))
    a[1:-1, 1:-1] = v.reshape(n[1], n[0])
    return a.reshape(-1)


@pytest.mark.parametrize("cyl", [1, 2])
def test_connector_on_closed_forms(cyl):
    """With r the cylindrical coordinate in cell units (centres at i + 1/2), the
    discrete operator gives 4 on r^2 and 2 on z^2 wherever the stencil stays in
    the interior (at the axis the outer neighbour has weight 0)."""
    n = (16, 12)
    rd = cyl - 1
    jj, ii = np.meshgrid(np.arange(n[1] + 2), np.arange(n[0] + 2), indexing="ij")
    coord = [ii - 0.5, jj - 0.5]   # interior cell i (0-based) sits at array index i + 1
    for u_of, expect in ((coord[rd] ** 2, 4.0), (coord[1 - rd] ** 2, 2.0)):
        u = u_of.reshape(-1).astype(np.float64)
        r = np.zeros_like(u)
        R.mg_op2("residual", R.n32(n), cyl, r, u, np.zeros_like(u), 0.0)
        L = interior2(r, n).reshape(n[1], n[0])
        # interior rows and columns whose neighbours are interior (the axis side
        # along r reads its ghost with weight 0)
        sl = [slice(0, n[0] - 1), slice(1, n[1] - 1)] if rd == 0 else [slice(1, n[0] - 1), slice(0, n[1] - 1)]
        np.testing.assert_allclose(L[sl[1], sl[0]], expect, rtol=1e-12)


@pytest.mark.parametrize("n,bc,levels,cyl", [
    ((8, 16), (N, D, D, D), 2, 1),
    ((16, 8), (N, N, D, N), 2, 0),
    ((16, 32), (D, N, N, D), 3, 2),
    ((32, 32), (N, D, D, D), 4, 1),
])
def test_solve_matches_dense(n, bc, levels, cyl):
    rng = np.random.default_rng(3)
    s = 0.8
    b = with_ghosts2(rng.normal(size=np.prod(n)), n)
    u = np.zeros_like(b)
    rc, it, res = R.mg_solve(n, bc, s, levels, 2, 2, 1e-13, 50, u, b, cyl=cyl)
    assert rc == 0 and it >= 2
    A = dense_operator2(n, bc, cyl)
    exact = np.linalg.solve(A, -s * interior2(b, n))
    np.testing.assert_allclose(interior2(u, n), exact, rtol=0, atol=1e-10 * np.abs(exact).max())


def test_transfer_weights():
    nf = R.n32([4, 4])
    r = np.full(6 * 6, 2.5)
    rh = np.zeros(4 * 4)
    R.mg_op2("restrict", nf, rh, r)
    assert np.all(interior2(rh, (2, 2)) == 4 * 2.5)   # 4 x mean (the h = 1 scaling of 2D)
    rh = np.full(4 * 4, 1.75)
    f = np.zeros(6 * 6)
    R.mg_op2("interpolate", nf, f, rh, 0)
    assert np.all(interior2(f, (4, 4)) == 1.75)
    # a linear coarse field is reproduced exactly at the fine centres
    jc, ic = np.meshgrid(np.arange(6), np.arange(6), indexing="ij")
    lin = lambda X, Y: 2.0 * X - 1.0 * Y + 3.0
    rh = lin((ic - 0.5) * 2, (jc - 0.5) * 2).reshape(-1)
    nf = R.n32([8, 8])
    f = np.zeros(10 * 10)
    R.mg_op2("interpolate", nf, f, rh, 0)
    jf, i_f = np.meshgrid(np.arange(10), np.arange(10), indexing="ij")
    expect = lin(i_f - 0.5, jf - 0.5).reshape(-1)
    np.testing.assert_allclose(interior2(f, (8, 8)), interior2(expect, (8, 8)), rtol=1e-14)


def test_coarse_inverse_is_cartesian():
    n = R.n32([2, 4])
    bc = R.n32([N, D, D, N])
    ainv = np.zeros(8 * 8)
    assert R.mg_op2("coarse_inverse", n, bc, ainv) == 0
    A = dense_operator2((2, 4), tuple(bc), 0)
    np.testing.assert_allclose(ainv.reshape(8, 8) @ A, np.eye(8), atol=1e-13)


def test_second_order_axisymmetric():
    """u = (1 - r^2) sin(pi z) on r in [0, 1], z in [0, 1]: -lap u =
    (4 + pi^2 (1 - r^2)) sin(pi z); Neumann on the axis, Dirichlet elsewhere."""
    errs = []
    for m in (16, 32, 64):
        n = (m, m)
        h = 1.0 / m
        c = (np.arange(m) + 0.5) * h
        Z, Rr = np.meshgrid(c, c, indexing="ij")
        ue = (1 - Rr ** 2) * np.sin(np.pi * Z)
        f = (4 + np.pi ** 2 * (1 - Rr ** 2)) * np.sin(np.pi * Z)
        b = with_ghosts2(f.reshape(-1), n)
        u = np.zeros_like(b)
        rc, it, _ = R.mg_solve(n, (N, D, D, D), h * h, 3, 2, 2, 1e-12, 50, u, b, cyl=1)
        assert rc == 0
        errs.append(np.abs(interior2(u, n) - ue.reshape(-1)).max())
    rate = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rate > 1.8), errs
