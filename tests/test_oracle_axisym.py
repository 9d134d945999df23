"""The oracle's axisymmetric grid (MCStreamer/src/grid.jl, gridfields.jl:115-224,
poisson.jl:6-84) pinned without the reference: exact floor against rational
arithmetic, the LinRange geometry against its closed forms, the staggered
bilinear field against linear fields it must reproduce, density! against a
brute-force placement with exact cell indices, the face fields against a
linear potential, the push in a uniform radial or axial field against the
exact parabola, and solve_freebc! (freebc.jl) against scipy's DST-II/III and
against a point charge between grounded planes with no outer wall."""
from fractions import Fraction
import math

import numpy as np

from mcstreamer_amd import constants as co
from mcstreamer_amd.axisym import Grid
from oracle import refcpu as R


def test_fld_is_the_exact_floor():
    rng = np.random.default_rng(0)
    ys = rng.uniform(1e-7, 1e-5, 400)
    ks = rng.integers(-3, 300, 400)
    xs = [float(k) * y for k, y in zip(ks, ys)]               # quotients at or next to integers
    xs += [np.nextafter(x, -np.inf) for x in xs[:200]]
    xs += list(rng.uniform(-1e-5, 3e-3, 400))
    ys = list(ys) * 3
    for x, y in zip(xs, ys[:len(xs)]):
        exact = math.floor(Fraction(x) / Fraction(float(y)))
        assert R.fld(x, y) == exact, (x, y)
    assert R.fld(0.3, 0.1) == 2.0            # 0.3 / 0.1 rounds to 3.0
    assert R.fld(-0.0, 1.0) == 0.0
    assert R.fld(-1e-20, 1.0) == -1.0


def test_geometry_matches_linrange_closed_forms():
    for (Rr, L, M, N) in [(1e-3, 2e-3, 8, 16), (0.64e-3, 1.28e-3, 128, 256), (3e-3, 1e-2, 300, 1000)]:
        g = R.CGrid(Rr, L, M, N)
        dr, dz, drf, dzf = g.steps
        py = Grid(Rr, L, M, N)
        assert (dr, dz, drf, dzf) == (py.dr, py.dz, py.drf, py.dzf)
        assert drf == Rr / M and dzf == L / N
        assert abs(dr - Rr / M) <= 8 * np.finfo(float).eps * Rr / M
        assert abs(dz - L / N) <= 8 * np.finfo(float).eps * L / N
        # cell volumes: dV(i) = 2 pi rc[i] dr dz with rc[i] = (i - 1/2) dr; they tile the cylinder
        rc = (np.arange(M) + 0.5) * (Rr / M)
        np.testing.assert_allclose(g.dv, 2 * np.pi * rc * dr * dz, rtol=1e-13)
        np.testing.assert_allclose(g.dv.sum(), np.pi * Rr * Rr * dz, rtol=1e-12)
        assert g.dv[0] == py.dV(1) and g.dv[-1] == py.dV(M)


def _fill_linear(g, a, b, c):
    """er = a r on its faces (storage I at r = (I - 1) dr), ez = b z + c r on its
    faces (storage J at z = (J - 1) dz, storage I at the centre r = (I - 1/2) dr)."""
    dr, dz = g.steps[0], g.steps[1]
    I = np.arange(g.M + 3)[None, :]
    J = np.arange(g.N + 3)[:, None]
    g.er[:] = (a * (I - 1) * dr + 0.0 * J).ravel()
    g.ez[:] = (b * (J - 1) * dz + c * (I - 0.5) * dr).ravel()


def test_field_reproduces_linear_fields():
    g = R.CGrid(1e-3, 2e-3, 8, 16)
    a, b, c = 3e9, -2e9, 5e8
    _fill_linear(g, a, b, c)
    rng = np.random.default_rng(1)
    for _ in range(500):
        r = rng.uniform(1e-9, 1e-3 * (1 - 1e-12))
        th = rng.uniform(0, 2 * np.pi)
        z = rng.uniform(1e-9, 2e-3 * (1 - 1e-12))
        x = np.array([r * np.cos(th), r * np.sin(th), z])
        e, fl = R.cyl_efield(g, x)
        rr = math.hypot(x[0], x[1])
        scale = abs(a) * 1e-3 + abs(b) * 2e-3 + abs(c) * 1e-3
        np.testing.assert_allclose(e, [a * x[0], a * x[1], b * z + c * rr], rtol=0, atol=1e-12 * scale)
        assert fl == 0


def test_field_zero_outside_and_limit_flag():
    g = R.CGrid(1e-3, 2e-3, 8, 16)
    g.er[:] = 1e7
    g.ez[:] = 1e7
    for x in ([0.0, 0.0, 1e-3], [1e-3, 0.0, 1e-3], [0.8e-3, 0.8e-3, 1e-3], [1e-4, 0.0, 0.0],
              [1e-4, 0.0, 2e-3], [1e-4, 0.0, -1e-6]):
        e, fl = R.cyl_efield(g, np.array(x))
        assert list(e) == [0.0, 0.0, 0.0] and fl == 0
    e, fl = R.cyl_efield(g, np.array([3e-4, 4e-4, 1e-3]))
    np.testing.assert_allclose(e, [0.6e7, 0.8e7, 1e7], rtol=1e-14)
    assert fl == 0
    g.ez[:] = 1.5e8
    e, fl = R.cyl_efield(g, np.array([3e-4, 4e-4, 1e-3]))
    assert fl == 2


def _pop(x, w=None, active=None):
    n = len(x)
    w = np.ones(n) if w is None else w
    act = np.ones(n, dtype=np.uint8) if active is None else active
    return R.Pop(np.asarray(x, dtype=np.float64), np.zeros((n, 3)), w, np.ones(n), act, cap=max(n, 1))


def test_density_places_by_exact_cell_index_in_order():
    Rr, L, M, N = 1e-3, 2e-3, 8, 16
    g = R.CGrid(Rr, L, M, N)
    drf, dzf = g.steps[2], g.steps[3]
    rng = np.random.default_rng(2)
    n = 3000
    r = rng.uniform(0, 1.1 * Rr, n)
    th = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-2 * dzf, L + 2 * dzf, n)
    x = np.stack([r * np.cos(th), r * np.sin(th), z], axis=1)
    x[:50, 2] = np.round(x[:50, 2] / dzf) * dzf      # on z faces
    w = rng.uniform(0.5, 2.0, n)
    act = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    P = _pop(x, w, act)
    g.qfixed[:] = rng.normal(size=g.qfixed.size)
    start = g.qfixed.copy()
    R.cyl_density(P, g, "qfixed", -1.0)
    exp = start.copy()
    used = 0
    for k in range(n):
        if not act[k]:
            continue
        rho = math.sqrt(x[k, 0] * x[k, 0] + x[k, 1] * x[k, 1])
        i = math.floor(Fraction(rho) / Fraction(drf)) + 1
        j = math.floor(Fraction(float(x[k, 2])) / Fraction(dzf)) + 1
        if not (1 <= i <= M and 0 <= j <= N + 1):
            continue
        used += 1
        c = i + (M + 2) * j
        exp[c] = exp[c] + (w[k] * -1.0) / g.dv[i - 1]
    assert used > n // 2
    np.testing.assert_array_equal(g.qfixed, exp)
    # charge: sum over cells of (added density) dV is the weight placed
    added = (g.qfixed - start).reshape(N + 2, M + 2)[:, 1:M + 1]
    np.testing.assert_allclose((added * g.dv[None, :]).sum(), -sum(
        w[k] for k in range(n) if act[k] and math.floor(math.sqrt(x[k, 0] ** 2 + x[k, 1] ** 2) / drf) < M
        and -1 <= math.floor(x[k, 2] / dzf) <= N), rtol=1e-12)


def test_faces_of_a_linear_potential_and_ghost_copies():
    Rr, L, M, N = 1e-3, 2e-3, 8, 16
    g = R.CGrid(Rr, L, M, N)
    dr, dz = g.steps[0], g.steps[1]
    al, be, eb = 2.0e3, -5.0e3, -1e6
    I = np.arange(M + 2)[None, :]
    J = np.arange(N + 2)[:, None]
    g.u[:] = (al * (I - 0.5) * dr + be * (J - 0.5) * dz).ravel()
    g.er[:] = 7.0
    g.ez[:] = 7.0
    R.cyl_faces(g, eb)
    er = g.er.reshape(N + 3, M + 3)
    ez = g.ez.reshape(N + 3, M + 3)
    np.testing.assert_allclose(er[1:N + 1, 1:M + 2], -al, rtol=1e-9)
    np.testing.assert_allclose(ez[1:N + 2, 1:M + 1], eb - be, rtol=1e-12)
    np.testing.assert_array_equal(ez[:, 0], ez[:, 1])
    np.testing.assert_array_equal(ez[:, M + 1], ez[:, M])
    np.testing.assert_array_equal(er[0, :], er[1, :])
    np.testing.assert_array_equal(er[N + 1, :], er[N, :])
    assert (er[1:N + 1, M + 2] == 7.0).all() and (ez[N + 2, 1:M + 1] == 7.0).all()   # never written


def test_charge_is_rounded_through_float32():
    g = R.CGrid(1e-3, 2e-3, 8, 16)
    rng = np.random.default_rng(3)
    g.qfixed[:] = rng.normal(size=g.qfixed.size) * 1e18
    g.qpart[:] = rng.normal(size=g.qpart.size) * 1e18
    g.q[:] = 5.0
    R.cyl_charge(g)
    q = g.q.reshape(18, 10)
    inner = q[1:17, 1:9]
    assert (inner.astype(np.float32).astype(np.float64) == inner).all()
    exact = (co.elementary_charge * (g.qfixed + g.qpart)).reshape(18, 10)[1:17, 1:9]
    assert (np.abs(inner - exact) <= 2.0 ** -24 * np.abs(exact) * (1 + 1e-12)).all()
    assert (q[0, :] == 5.0).all() and (q[:, 0] == 5.0).all()


def _quiet_table():
    # one null process with a tiny maximum rate: no collision within the test's dt
    energy = np.linspace(0.0, 1e3, 16) * co.eV
    rate = np.full((1, 16), 1e-30)
    return R.Table(energy, rate, [0], [0.0], 1e-30, egrid=0)


def test_push_in_uniform_axial_and_radial_fields_is_the_parabola():
    g = R.CGrid(1e-3, 2e-3, 8, 16)
    T = _quiet_table()
    dt = 2e-13
    qm = -co.elementary_charge / co.electron_mass
    # axial: er = 0, ez = E0 everywhere
    E0 = 3e6
    g.er[:] = 0.0
    g.ez[:] = E0
    x0 = np.array([[2e-4, -1e-4, 1e-3]])
    v0 = np.array([[1e5, 2e5, -3e5]])
    P = R.Pop(x0.copy(), v0.copy(), np.ones(1), np.ones(1), np.ones(1, dtype=np.uint8), cap=4)
    R.advance_cyl(P, T, g, dt, track=False)
    acc = np.array([0.0, 0.0, qm * E0])
    np.testing.assert_allclose(P.x[0], x0[0] + v0[0] * dt + 0.5 * acc * dt * dt, rtol=1e-12)
    np.testing.assert_allclose(P.v[0], v0[0] + acc * dt, rtol=1e-12)
    # radial: er = a on every face, ez = 0; a particle at rest on the x axis moves along x only
    a = -4e6
    g.er[:] = a
    g.ez[:] = 0.0
    P = R.Pop(np.array([[5e-4, 0.0, 1e-3]]), np.zeros((1, 3)), np.ones(1), np.ones(1),
              np.ones(1, dtype=np.uint8), cap=4)
    R.advance_cyl(P, T, g, dt, track=False)
    assert P.x[0, 1] == 0.0 and P.v[0, 1] == 0.0 and P.x[0, 2] == 1e-3
    np.testing.assert_allclose(P.x[0, 0], 5e-4 + 0.5 * qm * a * dt * dt, rtol=1e-12)
    np.testing.assert_allclose(P.v[0, 0], qm * a * dt, rtol=1e-12)


def _cyl_conf(G, tol=-1e-12, levels=4):
    from mcstreamer_amd import multigrid as mgmod
    return mgmod.MGConfig(
        bc=[(mgmod.LeftBnd, mgmod.Dirichlet), (mgmod.RightBnd, mgmod.Dirichlet),
            (mgmod.TopBnd, mgmod.Dirichlet), (mgmod.BottomBnd, mgmod.Neumann)],
        s=G.steps[1] ** 2 / co.epsilon_0, levels=levels, tolerance=tol, smooth1=2, smooth2=2, maxiter=50,
        ndim=2, conn=mgmod.CylindricalConnector(1))


def test_freebc_transforms_are_fftw_r2r():
    """The boundary value of solve_freebc! against scipy's DST-II/III, which are
    FFTW's RODFT10/RODFT01 (same definitions, same normalisation)."""
    from scipy.fft import dst
    G = R.CGrid(1e-3, 2e-3, 16, 64)
    rng = np.random.default_rng(4)
    G.u[:] = rng.normal(size=G.u.size)
    f = R.freebc_coeffs(G)
    got = R.freebc_bcond(G, f)
    M, N, dr = G.M, G.N, G.steps[0]
    u = G.u.reshape(N + 2, M + 2)
    grad = (u[1:N + 1, M + 1] - u[1:N + 1, M]) / dr
    am = (-(dr / 2) * dst(grad, type=2)) / f
    am1 = np.zeros(N)
    am1[-1] = am[-1]
    exp = dst(am, type=3) / 2 + dst(am1, type=3)
    np.testing.assert_allclose(got, exp, rtol=1e-11, atol=1e-12 * np.abs(exp).max())


def test_freebc_coefficients_large_argument_limit():
    """f_n / (L k_n / 2) = I1/I0 + K1/K0 tends to 2 - 1/(4 x^2) + ... for large x."""
    G = R.CGrid(1.5e-2, 0.1, 8, 10240)
    f = R.freebc_coeffs(G)
    k = np.pi / G.L * np.arange(1, G.N + 1)
    x = k * G.R
    ratio = f / (G.L * k / 2)
    big = x > 200
    np.testing.assert_allclose(ratio[big], 2 - 1 / (4 * x[big] ** 2), rtol=1e-9)
    assert (np.diff(ratio[big]) > 0).all()


def test_freebc_reproduces_point_charge_between_grounded_planes():
    """A compact charge Q on the axis between the grounded planes z = 0 and z = L,
    no outer wall: phi = Q / (pi eps0 L) sum_n sin(k z0) sin(k z) K0(k r).  With the
    free boundary the solution matches it across the grid (to discretisation); the
    homogeneous solve alone (a grounded wall at r = R) does not."""
    from scipy.special import k0
    Rr, L, M, N = 1e-3, 2e-3, 64, 128
    Q, z0, sig = 1e-13, 1e-3, 3e-5
    out = {}
    for fb in (True, False):
        G = R.CGrid(Rr, L, M, N)
        dr, dz = G.steps[0], G.steps[1]
        rc = (np.arange(M) + 0.5) * dr
        zc = (np.arange(N) + 0.5) * dz
        rho = Q / ((2 * np.pi) ** 1.5 * sig ** 3) * np.exp(-(rc[None, :] ** 2 + (zc[:, None] - z0) ** 2) / (2 * sig ** 2))
        qf = np.zeros((N + 2, M + 2))
        qf[1:N + 1, 1:M + 1] = rho / co.elementary_charge
        G.qfixed[:] = qf.ravel()
        P = R.Pop(np.zeros((0, 3)), np.zeros((0, 3)), [], [], [], cap=1)
        R.cyl_poisson(G, P, _cyl_conf(G), 0.0, freebc=fb)
        out[fb] = G.u.reshape(N + 2, M + 2)[1:N + 1, 1:M + 1]
    n = np.arange(1, 400)[:, None]
    kk = n * np.pi / L
    cols = [M // 2, M - 1]
    for c in cols:
        r = rc[c]
        phi = Q / (np.pi * co.epsilon_0 * L) * (np.sin(kk * z0) * np.sin(kk * zc[None, :]) * k0(kk * r)).sum(axis=0)
        sel = np.abs(phi) > 0.2 * np.abs(phi).max()
        err_free = np.abs(out[True][sel, c] - phi[sel]) / np.abs(phi[sel])
        err_wall = np.abs(out[False][sel, c] - phi[sel]) / np.abs(phi[sel])
        assert err_free.max() < 0.02, (c, err_free.max())
        assert err_wall.max() > 0.2, (c, err_wall.max())


def test_cyl_resample_per_cell_rules():
    """resample! on the (r, z) grid: cells over nmax end with ntarget equal weights
    and their total weight; thin cells split integer weights above 2; particles
    outside the per-cell arrays are kept as they are."""
    Rr, L, M, N = 1e-3, 2e-3, 8, 16
    G = R.CGrid(Rr, L, M, N)
    drf, dzf = G.steps[2], G.steps[3]
    rng = np.random.default_rng(7)
    n = 4000
    r = np.concatenate([rng.uniform(0, 0.3 * drf, 300), rng.uniform(0, 1.3 * Rr, n - 300)])
    th = rng.uniform(0, 2 * np.pi, n)
    z = np.concatenate([np.full(300, 0.5 * dzf + 5 * dzf), rng.uniform(-2 * dzf, L + 2 * dzf, n - 300)])
    x = np.stack([r * np.cos(th), r * np.sin(th), z], axis=1)
    w = rng.integers(1, 6, n).astype(np.float64)
    P = R.Pop(x, np.zeros((n, 3)), w, np.ones(n), np.ones(n, dtype=np.uint8), cap=4 * n)

    def cells(xx):
        i = np.floor(np.hypot(xx[:, 0], xx[:, 1]) / drf).astype(int) + 1
        j = np.floor(xx[:, 2] / dzf).astype(int) + 1
        ok = (i <= M + 1) & (j >= 0) & (j <= N + 1)
        return np.where(ok, i + (M + 2) * j, -1)

    c0 = cells(x)
    rc, st = R.cyl_resample(P, G, 10, 15, seed=3, step=1)
    assert rc == 0
    v = P.view()
    c1 = cells(v["x"])
    assert st.cells_merged >= 1
    for c in np.unique(c0):
        w_in = w[c0 == c].sum()
        w_out = v["w"][c1 == c].sum()
        np.testing.assert_allclose(w_out, w_in, rtol=1e-12)
        k = (c1 == c).sum()
        if c >= 0 and (c0 == c).sum() > 15:
            assert k == 10 and np.allclose(v["w"][c1 == c], w_in / 10, rtol=1e-15)
        if c < 0:
            assert k == (c0 == c).sum()
    # canonical order: cells ascending, outside last
    key = np.where(c1 < 0, 10 ** 9, c1)
    assert (np.diff(key) >= 0).all()
