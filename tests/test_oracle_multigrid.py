"""Pins of the multigrid restatement (Multigrid/src/Multigrid.jl): the solve
against a dense numpy solve of the same 7-point operator with the ghost rules of
boundaries.jl, the restriction/prolongation weights, the coarse inverse, the
cycle's contraction and second-order convergence on a manufactured solution."""
import numpy as np
import pytest

from oracle import refcpu as R

D, N = 0, 1


def dense_operator(n, bc):
    nx, ny, nz = n
    M = nx * ny * nz
    A = np.zeros((M, M))

    def lin(i, j, k):
        return i + nx * (j + ny * k)

    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = lin(i, j, k)
                A[r, r] -= 6
                for d, s in ((0, 1), (1, 1), (2, 1), (0, -1), (1, -1), (2, -1)):
                    c = [i, j, k]
                    c[d] += s
                    if c[d] < 0 or c[d] >= n[d]:
                        face = 2 * d + (1 if s > 0 else 0)
                        A[r, r] += -1.0 if bc[face] == D else 1.0
                    else:
                        A[r, lin(*c)] += 1
    return A


def interior(a, n):
    return a.reshape(n[2] + 2, n[1] + 2, n[0] + 2)[1:-1, 1:-1, 1:-1].reshape(-1)


def with_ghosts(v, n):
    a = np.zeros((n[2] + 2, n[1] + 2, n[0] + 2))
    a[1:-1, 1:-1, 1:-1] = v.reshape(n[2], n[1], n[0])
    return a.reshape(-1)


@pytest.mark.parametrize("n,bc,levels", [
    ((4, 4, 8), (N, N, N, N, D, D), 1),
    ((8, 8, 8), (D, D, D, D, D, D), 2),
    ((8, 4, 16), (N, D, D, N, D, N), 1),
    ((16, 16, 16), (N, N, N, N, D, D), 3),
])
def test_solve_matches_dense(n, bc, levels):
    rng = np.random.default_rng(0)
    s = 0.7
    b = with_ghosts(rng.normal(size=np.prod(n)), n)
    u = np.zeros_like(b)
    rc, it, res = R.mg_solve(n, bc, s, levels, 2, 2, 1e-13, 50, u, b)
    assert rc == 0 and it >= 2
    A = dense_operator(n, bc)
    exact = np.linalg.solve(A, -s * interior(b, n))
    np.testing.assert_allclose(interior(u, n), exact, rtol=0, atol=1e-10 * np.abs(exact).max())


def test_restrict_and_prolong_weights():
    nf = R.n32([4, 4, 4])
    r = np.full((6 * 6 * 6), 2.5)
    rh = np.zeros(4 * 4 * 4)
    R.mg_op("restrict", nf, rh, r)
    assert np.all(interior(rh, (2, 2, 2)) == 4 * 2.5)  # 4 x mean keeps the h = 1 scaling
    # prolongation of a constant with Neumann ghosts is that constant
    rh = np.full(4 * 4 * 4, 1.75)
    f = np.zeros(6 * 6 * 6)
    R.mg_op("interpolate", nf, f, rh, 0)
    assert np.all(interior(f, (4, 4, 4)) == 1.75)
    # prolongation of a linear coarse field is exact at the fine centres
    nc = (4, 4, 4)
    kc, jc, ic = np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij")
    lin = lambda X, Y, Z: 2.0 * X - 1.0 * Y + 0.5 * Z + 3.0
    # coarse centre of ghost-inclusive index i is at (i - 1/2) * 2 in fine units
    rh = lin((ic - 0.5) * 2, (jc - 0.5) * 2, (kc - 0.5) * 2).reshape(-1)
    nf = R.n32([8, 8, 8])
    f = np.zeros(10 * 10 * 10)
    R.mg_op("interpolate", nf, f, rh, 0)
    kf, jf, i_f = np.meshgrid(np.arange(10), np.arange(10), np.arange(10), indexing="ij")
    expect = lin(i_f - 0.5, jf - 0.5, kf - 0.5).reshape(-1)
    np.testing.assert_allclose(interior(f, (8, 8, 8)), interior(expect, (8, 8, 8)), rtol=1e-14)


def test_coarse_inverse():
    n = R.n32([2, 2, 4])
    bc = R.n32([N, N, D, N, D, D])
    ainv = np.zeros(16 * 16)
    assert R.mg_op("coarse_inverse", n, bc, ainv) == 0
    A = dense_operator((2, 2, 4), tuple(bc))
    np.testing.assert_allclose(ainv.reshape(16, 16) @ A, np.eye(16), atol=1e-13)


def test_cycle_contraction():
    n = (32, 32, 32)
    rng = np.random.default_rng(1)
    b = with_ghosts(rng.normal(size=np.prod(n)), n)
    res = []
    for it in range(1, 5):
        u = np.zeros_like(b)
        R.mg_solve(n, (N, N, N, N, D, D), 1.0, 4, 2, 2, 0.0, it - 1, u, b)
        r = np.zeros_like(b)
        R.mg_op("residual", R.n32(n), r, u, b, 1.0)
        res.append(np.linalg.norm(r))
    ratios = np.array(res[1:]) / np.array(res[:-1])
    assert np.all(ratios < 0.2), ratios


def test_second_order_on_manufactured_solution():
    errs = []
    for m in (8, 16, 32):
        n = (m, m, m)
        h = 1.0 / m
        c = (np.arange(m) + 0.5) * h
        Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
        ue = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
        f = 3 * np.pi ** 2 * ue  # -lap u = f ; L u + s b = 0 with s = h^2, b = f
        b = with_ghosts(f.reshape(-1), n)
        u = np.zeros_like(b)
        R.mg_solve(n, (D,) * 6, h * h, 2, 2, 2, 1e-12, 50, u, b)
        errs.append(np.abs(interior(u, n) - ue.reshape(-1)).max())
    rate = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rate > 1.8), errs
