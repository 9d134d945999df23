"""Host mirror of Multigrid.jl (Multigrid/src/Multigrid.jl) on top of libmcs.

  MGConfig (Multigrid.jl:41-74)    -> MGConfig
  allocate(conf, x) (:302-324)     -> allocate -> Workspace (device hierarchy)
  solve(conf, x, b, ws) (:449-467) -> solve (x updated in place, host arrays)
Boundaries are given per face like the reference's tuple of (boundary,
condition) pairs; BottomBnd/TopBnd are the first array dimension (x),
LeftBnd/RightBnd the second (y), FrontBnd/BackBnd the third (z)
(boundaries.jl:110-115).  Ghost planes are applied in the order x-, x+, y-,
y+, z-, z+.  Two-dimensional grids (the reference's (r, z) solver,
MCStreamer/src/main.jl:139) take four faces and a connector:
CartesianConnector() or CylindricalConnector(d), d the 1-based cylindrical
dimension as in CylindricalConnector{d} (Multigrid.jl:31-39).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr
from .jumc import context

BottomBnd, TopBnd, LeftBnd, RightBnd, FrontBnd, BackBnd = range(6)
Dirichlet, Neumann = _lib.BC_DIRICHLET, _lib.BC_NEUMANN


class CartesianConnector:
    """Multigrid.jl:31-32: neighbour terms enter unchanged."""
    cyl = 0


class CylindricalConnector:
    """Multigrid.jl:35-39: along dimension d (1-based) the neighbour at offset
    s of interior cell i enters as x (1 + s / (2 i + 1)) (0-based i)."""

    def __init__(self, d=1):
        if d not in (1, 2):
            raise ValueError("the cylindrical dimension of a 2D grid is 1 or 2")
        self.cyl = int(d)


class MGConfig:
    def __init__(self, bc, s, levels, tolerance, smooth1=2, smooth2=2, maxiter=50,
                 precision="f64", ndim=3, conn=None):
        self.ndim = int(ndim)
        if self.ndim not in (2, 3):
            raise ValueError("ndim is 2 or 3")
        nf = 2 * self.ndim
        faces = [None] * nf
        for bnd, cond in bc:
            faces[bnd] = cond
        if any(f is None for f in faces):
            raise ValueError(f"{self.ndim}D solve needs a condition on all {nf} faces")
        self.conn = conn or CartesianConnector()
        if self.conn.cyl and self.ndim != 2:
            raise ValueError("the cylindrical connector is for 2D grids")
        self.bc = faces
        self.s = float(s)
        self.levels = int(levels)
        self.tolerance = float(tolerance)
        self.smooth1 = int(smooth1)
        self.smooth2 = int(smooth2)
        self.maxiter = int(maxiter)
        self.precision = {"f64": _lib.PREC_F64, "f32": _lib.PREC_F32}[precision]

    def cstruct(self, n):
        c = _lib.MGConfigC()
        for d in range(3):
            c.n[d] = int(n[d]) if d < self.ndim else 1
        for f in range(2 * self.ndim):
            c.bc[f] = int(self.bc[f])
        c.ndim = self.ndim
        c.cyl = self.conn.cyl
        c.s = self.s
        c.levels = self.levels
        c.smooth1 = self.smooth1
        c.smooth2 = self.smooth2
        c.tolerance = self.tolerance
        c.maxiter = self.maxiter
        c.precision = self.precision
        return c


class Workspace:
    """Device level hierarchy and coarse inverse for one mesh size."""

    def __init__(self, conf, n, ctx=None):
        self.ctx = ctx or context()
        self.conf = conf
        self.n = tuple(int(v) for v in n)
        cfg = conf.cstruct(self.n)
        h = ctypes.c_void_p()
        check(self.ctx.handle, _lib.lib().mcs_mg_create(self.ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.last_iters = 0
        self.last_resid = 0.0

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().mcs_mg_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def allocate(conf, x):
    """x is the (nx+2, ny+2, nz+2) (or (nx+2, ny+2) in 2D) array with ghosts (any
    memory order accepted through its shape); the workspace fits its interior size."""
    shape = np.shape(x)
    return Workspace(conf, tuple(s - 2 for s in shape))


def _as_flat(a, n):
    """Flatten an (nx+2, ny+2[, nz+2]) array so that x runs fastest."""
    a = np.asarray(a, dtype=np.float64)
    want = tuple(v + 2 for v in n)
    if a.shape != want:
        raise ValueError(f"array shape must be {want}")
    return np.ascontiguousarray(a.transpose(*range(len(n) - 1, -1, -1))).reshape(-1)


def solve(conf, x, b, ws):
    """Solve L x + s b = 0; x (initial guess) is overwritten with the solution.
    Returns (iterations, final residual rms) like the reference's fmg! loop."""
    n = ws.n
    xf = _as_flat(x, n)
    bf = _as_flat(b, n)
    it = ctypes.c_int32()
    res = ctypes.c_double()
    rc = _lib.lib().mcs_mg_solve_host(ws.handle, dptr(xf), dptr(bf), ctypes.byref(it), ctypes.byref(res))
    if rc not in (_lib.MCS_OK, _lib.MCS_E_NONCONVERGED):
        check(ws.ctx.handle, rc)
    x[...] = xf.reshape(tuple(v + 2 for v in reversed(n))).transpose(*range(len(n) - 1, -1, -1))
    ws.last_iters, ws.last_resid = it.value, res.value
    return it.value, res.value
