"""Host mirror of MCStreamer's axisymmetric (r, z) geometry on top of libmcs.

  Grid(R, L, M, N) (grid.jl:4-90)           -> Grid
  GridFields (gridfields.jl:1-84)           -> GridFields: qfixed, qpart, q, u, er, ez
  FieldInterp (gridfields.jl:86-190)        -> FieldInterp
  CollisionTracker (gridfields.jl:193-224)  -> CollisionTracker
  density! (poisson.jl:65-84)               -> density
  poisson! (poisson.jl:6-58)                -> poisson
  solve_freebc! (freebc.jl:4-36)            -> poisson(freebc=True)
  resample! (resample.jl:7-84)              -> sort_by_cell + resample
  the solver of main.jl:134-145             -> poisson_config
  SavedGridFields (io.jl:9-44), h5export    -> saved_fields, h5export
    (plot.jl:210-240)
Particles stay in 3D Cartesian coordinates; the grid is indexed by
r = sqrt(x^2 + y^2) and z.  Arrays come back as (rows j, columns i) = (z, r),
the reference's column-major (i, j) layout read row-major.
"""
import ctypes

import numpy as np

from . import _lib
from . import constants as co
from . import multigrid as mgmod
from ._lib import check, dptr
from .jumc import context

CARR = {"qfixed": 0, "qpart": 1, "q": 2, "u": 3, "er": 4, "ez": 5}


def _lerpi(j, d, a, b):
    t = j / d
    return (1.0 - t) * a + t * b


class Grid:
    """Grid(R, L, M, N): M cells in r over [0, R], N in z over [0, L].  dr and dz
    are the steps of the cell-centre LinRanges (grid.jl:74-75), drf and dzf those
    of the face LinRanges (cellindex divides by these)."""

    def __init__(self, R, L, M, N):
        self.R, self.L, self.M, self.N = float(R), float(L), int(M), int(N)
        if self.M < 2 or self.N < 2:
            raise ValueError("M and N must be >= 2")
        r0, r1 = self._centres(self.R, self.M)
        z0, z1 = self._centres(self.L, self.N)
        self._rc = (r0, r1)
        self.dr = (r1 - r0) / (self.M - 1)
        self.dz = (z1 - z0) / (self.N - 1)
        self.drf = (self.R - 0.0) / self.M
        self.dzf = (self.L - 0.0) / self.N

    @staticmethod
    def _centres(X, n):
        f1, f2 = _lerpi(0, n, 0.0, X), _lerpi(1, n, 0.0, X)
        fn, fn1 = _lerpi(n - 1, n, 0.0, X), _lerpi(n, n, 0.0, X)
        return 0.5 * (f2 + f1), 0.5 * (fn1 + fn)

    @property
    def rc(self):
        return np.array([_lerpi(i, self.M - 1, *self._rc) for i in range(self.M)])

    @property
    def zc(self):
        z0, z1 = self._centres(self.L, self.N)
        return np.array([_lerpi(j, self.N - 1, z0, z1) for j in range(self.N)])

    @property
    def rf(self):
        return np.array([_lerpi(i, self.M, 0.0, self.R) for i in range(self.M + 1)])

    @property
    def zf(self):
        return np.array([_lerpi(j, self.N, 0.0, self.L) for j in range(self.N + 1)])

    def dV(self, i):
        """dV(grid, i) = 2 pi rc[i] dr dz (grid.jl:76), i 1-based."""
        return 2.0 * 3.141592653589793 * float(self.rc[i - 1]) * self.dr * self.dz


class GridFields:
    """Device arrays of one axisymmetric grid (mcs_cgrid)."""

    def __init__(self, grid, ctx=None):
        self.ctx = ctx or context()
        self.grid = grid
        h = ctypes.c_void_p()
        check(self.ctx.handle, _lib.lib().mcs_cgrid_create(self.ctx.handle, grid.R, grid.L, grid.M, grid.N,
                                                           ctypes.byref(h)))
        self.handle = h

    def shape(self, name):
        M, N = self.grid.M, self.grid.N
        return (N + 3, M + 3) if name in ("er", "ez") else (N + 2, M + 2)

    def steps(self):
        s = np.zeros(4)
        check(self.ctx.handle, _lib.lib().mcs_cgrid_steps(self.handle, dptr(s)))
        return s

    def get(self, name):
        a = np.zeros(self.shape(name))
        check(self.ctx.handle, _lib.lib().mcs_cgrid_get(self.handle, CARR[name], dptr(a)))
        return a

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape(name):
            raise ValueError(f"{name} has shape {self.shape(name)}")
        check(self.ctx.handle, _lib.lib().mcs_cgrid_set(self.handle, CARR[name], dptr(a)))

    def zero(self, name):
        check(self.ctx.handle, _lib.lib().mcs_cgrid_zero(self.handle, CARR[name]))

    def resample_fields(self):
        """(p, wtotal) as the last resample left them (gridfields.jl:41-48,
        resample.jl:16-30, 63-72), shape (N+2, M+2); zeros before any resample."""
        M, N = self.grid.M, self.grid.N
        p = np.zeros((N + 2, M + 2), dtype=np.int64)
        w = np.zeros((N + 2, M + 2))
        check(self.ctx.handle, _lib.lib().mcs_cgrid_resample_fields(
            self.handle, p.ctypes.data_as(_lib._PI64), dptr(w)))
        return p, w

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().mcs_cgrid_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class FieldInterp:
    """The axisymmetric FieldInterp: bilinear er and ez on the staggered faces,
    er rotated into x and y (gridfields.jl:115-190).  jumc.advance dispatches on
    `axisymmetric`."""

    axisymmetric = True

    def __init__(self, fields):
        self.fields = fields


class CollisionTracker:
    """+w/dV at ionizations, -w/dV at attachments into qfixed (gridfields.jl:203-218)."""

    axisymmetric = True

    def __init__(self, fields):
        self.fields = fields


def density(fields, which, popl, val=1.0):
    """density! (poisson.jl:65-84): (w val)/dV(i) of each active particle into the
    NGP cell of `which` ('qfixed' or 'qpart'), in particle order per cell."""
    check(popl.ctx.handle, _lib.lib().mcs_cyl_density(popl.handle, fields.handle, CARR[which], float(val)))


def poisson_config(grid, levels=None, tolerance=1.0e8 * co.elementary_charge, maxiter=50):
    """MGConfig of main.jl:134-145: Dirichlet at z = 0, z = L and r = R, Neumann
    on the axis, s = dz^2 / eps0, CylindricalConnector{1}, V(2,2).  `levels`
    defaults to the deepest hierarchy of even sizes (the reference's default is 9)."""
    if levels is None:
        levels = 0
        while all((n >> (levels + 1)) >= 2 and n % (1 << (levels + 1)) == 0 for n in (grid.M, grid.N)):
            levels += 1
    return mgmod.MGConfig(
        bc=[(mgmod.LeftBnd, mgmod.Dirichlet), (mgmod.RightBnd, mgmod.Dirichlet),
            (mgmod.TopBnd, mgmod.Dirichlet), (mgmod.BottomBnd, mgmod.Neumann)],
        s=grid.dz * grid.dz / co.epsilon_0, levels=levels, tolerance=tolerance, smooth1=2, smooth2=2,
        maxiter=maxiter, ndim=2, conn=mgmod.CylindricalConnector(1))


def freebc_coeffs(fields):
    """f_n of solve_freebc! (freebc.jl:16-18) for n = 1..N."""
    f = np.zeros(fields.grid.N)
    check(fields.ctx.handle, _lib.lib().mcs_cgrid_freebc_coeffs(fields.handle, dptr(f)))
    return f


def poisson(fields, popl, eb, ws, freebc=True):
    """poisson! (poisson.jl:6-58): qpart from the electrons (val -1), q = e (qfixed
    + qpart) rounded through FP32, solve_freebc! (freebc.jl:4-36) for u from the
    previous u (freebc=False: its first, homogeneous solve only), face fields with
    eb on z.  Returns (iterations of the solves together, residual)."""
    fields.zero("qpart")
    density(fields, "qpart", popl, -1.0)
    it = ctypes.c_int32()
    res = ctypes.c_double()
    rc = _lib.lib().mcs_cyl_poisson(fields.handle, ws.handle, float(eb), int(bool(freebc)), ctypes.byref(it),
                                    ctypes.byref(res))
    if rc not in (_lib.MCS_OK, _lib.MCS_E_NONCONVERGED):
        check(fields.ctx.handle, rc)
    ws.last_iters, ws.last_resid = it.value, res.value
    return it.value, res.value


def sort_by_cell(popl, fields):
    """Canonical (cell, id) order over resample!'s cells; inactive particles dropped."""
    check(popl.ctx.handle, _lib.lib().mcs_cyl_sort_by_cell(popl.handle, fields.handle))


def resample(popl, fields, ntarget, nmax, step=0):
    """resample! (resample.jl:7-84) per (r, z) cell; sorts first."""
    sort_by_cell(popl, fields)
    st = _lib.ResampleStats()
    check(popl.ctx.handle, _lib.lib().mcs_cyl_resample(popl.handle, fields.handle, int(ntarget), int(nmax),
                                                       popl.seed, popl.instance, int(step), ctypes.byref(st)))
    return st


def saved_fields(fields):
    """SavedGridFields(fields) (io.jl:9-44): the grid's arrays as the reference saves
    them, each in this module's (rows j, columns i) layout with its ghosts:
    qfixed, qpart, q0, q, u, er, ez, p, wtotal.  q0 is poisson!'s charge before
    the denoiser, ((qfixed + qpart) - ne) e with no fluid electrons
    (poisson.jl:19-27); q is the same rounded through Float32 (the NullDenoiser's
    input), as the last poisson left it."""
    qfixed, qpart = fields.get("qfixed"), fields.get("qpart")
    p, wtotal = fields.resample_fields()
    return {"qfixed": qfixed, "qpart": qpart, "q0": ((qfixed + qpart) - 0.0) * co.elementary_charge,
            "q": fields.get("q"), "u": fields.get("u"), "er": fields.get("er"), "ez": fields.get("ez"),
            "p": p, "wtotal": wtotal}


def h5export(fields, filename, saved=None):
    """h5export (MCStreamer/src/plot.jl:210-240): an HDF5 file with group `grid`
    (R, L as Float64 scalars, M, N as Int64 scalars, rc, rf, zc, zf as vectors) and
    the datasets qfixed, qpart, q0, q, u, p, wtotal (cells 1:M x 1:N), er
    (1:M+1 x 1:N) and ez (1:M x 1:N+1).  HDF5.jl stores a Julia (M, N) matrix with
    the dimensions (N, M) of its column-major memory, which is the numpy array of
    shape (N, M) here; p is Int64, the rest Float64.  `saved` (saved_fields'
    dictionary) is read from the device when not given."""
    from . import h5

    g = fields.grid
    M, N = g.M, g.N
    sv = saved if saved is not None else saved_fields(fields)
    tree = {"grid": {"R": np.float64(g.R), "L": np.float64(g.L), "M": np.int64(M), "N": np.int64(N),
                     "rc": g.rc, "rf": g.rf, "zc": g.zc, "zf": g.zf}}
    for k in ("qfixed", "qpart", "q0", "q", "u", "wtotal"):
        tree[k] = sv[k][1:N + 1, 1:M + 1]
    tree["p"] = np.asarray(sv["p"], dtype=np.int64)[1:N + 1, 1:M + 1]
    tree["er"] = sv["er"][1:N + 1, 1:M + 2]
    tree["ez"] = sv["ez"][1:N + 2, 1:M + 1]
    h5.write(filename, tree)
