"""Host mirror of the MCStreamer step (MCStreamer/src/main.jl:217-315) on a 3D
Cartesian mesh, on top of libmcs.

  Grid / GridFields (grid.jl:4-100, gridfields.jl:1-84) -> Grid3, GridFields
  FieldInterp (gridfields.jl:86-190)      -> FieldInterp
  CollisionTracker (gridfields.jl:193-224) -> CollisionTracker
  density! (poisson.jl:65-84)              -> density
  poisson! (poisson.jl:6-58)               -> poisson
  resample! (resample.jl:7-84)             -> resample
  nsteps (main.jl:217-315)                 -> Streamer.step / Streamer.nsteps
The population is kept in canonical cell order (sorted by cell, then id), which
replaces repack! where only the set of particles matters.
"""
import ctypes

import numpy as np

from . import _lib
from . import constants as co
from ._lib import check, dptr
from .jumc import MultiPopulation, advance, context, repack
from . import multigrid as mgmod


class Grid3:
    """Uniform 3D box [0, n h)^3 with cell centres at (i + 1/2) h."""

    def __init__(self, n, h):
        self.n = tuple(int(v) for v in n)
        self.h = float(h)

    @property
    def ncells(self):
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def dV(self):
        return (self.h * self.h) * self.h


class GridFields:
    """Device arrays of one mesh: qfixed, qpart, q, u, ex, ey, ez (mcs_mesh)."""

    NAMES = {"qfixed": _lib.ARR_QFIXED, "qpart": _lib.ARR_QPART, "q": _lib.ARR_Q,
             "u": _lib.ARR_U, "ex": _lib.ARR_EX, "ey": _lib.ARR_EY, "ez": _lib.ARR_EZ}

    def __init__(self, grid, ctx=None):
        self.ctx = ctx or context()
        self.grid = grid
        n = (ctypes.c_int32 * 3)(*grid.n)
        h = ctypes.c_void_p()
        check(self.ctx.handle, _lib.lib().mcs_mesh_create(self.ctx.handle, n, grid.h, ctypes.byref(h)))
        self.handle = h

    def shape(self, name):
        nx, ny, nz = self.grid.n
        return {"ex": (nx + 1, ny + 2, nz + 2), "ey": (nx + 2, ny + 1, nz + 2),
                "ez": (nx + 2, ny + 2, nz + 1)}.get(name, (nx + 2, ny + 2, nz + 2))

    def get(self, name):
        """Array with x fastest, returned as shape (nz', ny', nx') C-order."""
        sx, sy, sz = self.shape(name)
        a = np.zeros(sx * sy * sz)
        check(self.ctx.handle, _lib.lib().mcs_mesh_get(self.handle, self.NAMES[name], dptr(a)))
        return a

    def set(self, name, flat):
        a = np.ascontiguousarray(flat, dtype=np.float64).reshape(-1)
        sx, sy, sz = self.shape(name)
        if a.size != sx * sy * sz:
            raise ValueError(f"{name} needs {sx * sy * sz} values")
        check(self.ctx.handle, _lib.lib().mcs_mesh_set(self.handle, self.NAMES[name], dptr(a)))

    def zero(self, name):
        check(self.ctx.handle, _lib.lib().mcs_mesh_zero(self.handle, self.NAMES[name]))

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().mcs_mesh_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class FieldInterp:
    """Staggered trilinear interpolation of the face fields (gridfields.jl:115-190)."""

    def __init__(self, fields):
        self.fields = fields

    def cstruct(self):
        f = _lib.Field()
        f.kind = _lib.FIELD_MESH
        f.mesh = self.fields.handle
        return f


class CollisionTracker:
    """Ion charge left at ionization (+w/dV) and attachment (-w/dV) events."""

    def __init__(self, fields):
        self.fields = fields

    def mesh_handle(self):
        return self.fields.handle


def sort_by_cell(popl, fields):
    check(popl.ctx.handle, _lib.lib().mcs_sort_by_cell(popl.handle, fields.handle))


def density(fields, which, popl, val=1.0, scheme="cic"):
    """density! (poisson.jl:65-84): add w*val/dV of each particle to `which`
    ('qfixed' or 'qpart').  The population must be in cell order."""
    sch = {"ngp": _lib.DEPOSIT_NGP, "cic": _lib.DEPOSIT_CIC}[scheme]
    check(popl.ctx.handle, _lib.lib().mcs_deposit(popl.handle, fields.handle,
                                                  GridFields.NAMES[which], sch, float(val)))


def poisson(fields, popl, eb, ws, scheme="cic"):
    """poisson! (poisson.jl:6-58) without denoiser and free boundary: deposit
    the electrons, q = e (qfixed + qpart), solve, face fields with eb on z."""
    fields.zero("qpart")
    density(fields, "qpart", popl, -1.0, scheme)
    check(fields.ctx.handle, _lib.lib().mcs_assemble_charge(fields.handle))
    it = ctypes.c_int32()
    res = ctypes.c_double()
    rc = _lib.lib().mcs_mg_solve_mesh(ws.handle, fields.handle, ctypes.byref(it), ctypes.byref(res))
    if rc not in (_lib.MCS_OK, _lib.MCS_E_NONCONVERGED):
        check(fields.ctx.handle, rc)
    ws.last_iters, ws.last_resid = it.value, res.value
    check(fields.ctx.handle, _lib.lib().mcs_efield_faces(fields.handle, float(eb)))
    return it.value, res.value


def resample(popl, fields, ntarget, nmax, step=0):
    """resample! (resample.jl:7-84) per cell on the canonical order."""
    st = _lib.ResampleStats()
    check(popl.ctx.handle, _lib.lib().mcs_resample_cells(
        popl.handle, fields.handle, int(ntarget), int(nmax), popl.seed, popl.instance, int(step),
        ctypes.byref(st)))
    return st


class Streamer:
    """The per-step loop of nsteps (main.jl:217-315) with Poisson and resampling
    every step: deposit + solve + faces, advance (mesh field, ion tracker),
    diagnostics, canonical sort (the repack), per-cell resample.  With a photon
    population (a table with PhotoEmission rows) the step advances
    MultiPopulation(electrons, photons) and repacks the photons before the
    resample (main.jl:285)."""

    def __init__(self, popl, fields, ws, eb, dt, ntarget, nmax, scheme="cic", photons=None):
        self.popl = popl
        self.photons = photons
        self.mpopl = MultiPopulation(popl, photons) if photons is not None else None
        self.fields = fields
        self.ws = ws
        self.eb = float(eb)
        self.dt = float(dt)
        self.ntarget = int(ntarget)
        self.nmax = int(nmax)
        self.scheme = scheme
        self.i = 0
        self.field = FieldInterp(fields)
        self.tracker = CollisionTracker(fields)
        sort_by_cell(popl, fields)

    def step(self, diagnostics=True):
        it, res = poisson(self.fields, self.popl, self.eb, self.ws, self.scheme)
        pst = None
        if self.mpopl is None:
            st = advance(self.popl, self.field, self.dt, self.tracker, step=self.i)
        else:
            st, pst = advance(self.mpopl, self.field, self.dt, self.tracker, step=self.i)
        mom = self.popl.moments() if diagnostics else None
        if self.photons is not None:
            repack(self.photons)
        sort_by_cell(self.popl, self.fields)
        rs = resample(self.popl, self.fields, self.ntarget, self.nmax, step=self.i)
        self.i += 1
        out = dict(mg_iters=it, mg_resid=res, advance=st, moments=mom, resample=rs)
        if pst is not None:
            out["photons"] = pst
        return out

    def nsteps(self, n):
        out = None
        for _ in range(n):
            out = self.step()
        return out


def seed_pairs(grid, n, center, sigma, rng):
    """Neutral Gaussian seed: n electrons at rest (main.jl:83-92 with v0 = 0)."""
    x = np.asarray(center, dtype=np.float64)[None, :] + rng.normal(size=(n, 3)) * sigma
    v = np.zeros((n, 3))
    w = np.ones(n)
    s = -np.log1p(-rng.uniform(size=n))
    act = np.ones(n, dtype=np.uint8)
    return x, v, w, s, act


def make_streamer(n=(128, 128, 256), h=5e-6, npairs=1_000_000, seed=7, levels=None, cap=None,
                  precision="f64", tolerance=-1e-8, sigma=50e-6, ctx=None, photo=None):
    """A streamer workload on an n-cell box of cell size h: C0 gas + 20% attacher,
    E = 5 MV/m along z, neutral Gaussian seed of npairs (width sigma) at
    (Lx/2, Ly/2, 0.8 Lz), CIC, V(2,2) cycles to `tolerance` (negative: relative),
    merge/split 8..64 per cell, dt = 2e-13 s.  precision "f32" gives the split
    particle store, the FP32 step and FP32 correction cycles inside FP64
    refinement; "f64" the FP64 path.  photo = dict(rescale=, photon_multiplier=,
    photon_weight=) adds the gas's photoemission and a photon population of the
    same capacity (returned as "photons")."""
    from . import gases, lxcat
    from .jumc import CollisionTable, PhotonPopulation, Population, ParticleStates, ZhelezniakCollisions
    import os
    import tempfile

    ctx = ctx or context()
    grid = Grid3(n, h)
    e_ev = np.geomspace(1e-3, 1e3, 4000)
    ph = dict(photo or {})
    path = os.path.join(tempfile.gettempdir(), f"mcs_gas_c1_{int(photo is not None)}_{os.getpid()}.json")
    gases.write_gas(path, e_ev, with_attacher=True, with_photo=photo is not None,
                    photo_rescale=float(ph.get("rescale", 0.05)))
    proc, rate, maxrate, _ = lxcat.load_lxcat([path], gases.densities(0.2), e_ev * co.eV,
                                              photon_weight=float(ph.get("photon_weight", 1.0)),
                                              photon_multiplier=float(ph.get("photon_multiplier", 1.0)))
    table = CollisionTable(proc, e_ev * co.eV, rate, maxrate, egrid="log", ctx=ctx)
    rng = np.random.default_rng(seed)
    L = np.array(n, dtype=np.float64) * h
    x, v, w, s, act = seed_pairs(grid, npairs, (L[0] / 2, L[1] / 2, 0.8 * L[2]), sigma, rng)
    cap = cap or max(4 * npairs, 1 << 20)
    fields = GridFields(grid, ctx=ctx)
    popl = Population(cap, None, table, seed=seed, ctx=ctx, precision=precision, mesh=fields)
    popl.upload(ParticleStates(x, v, w, s, act))
    del x, v, w, s, act
    # ions of the neutral seed: NGP deposit of +1 per electron into qfixed (main.jl:130-132)
    sort_by_cell(popl, fields)
    density(fields, "qfixed", popl, 1.0, "ngp")
    if levels is None:
        levels = 0
        while all((nn >> (levels + 1)) >= 2 and nn % (1 << (levels + 1)) == 0 for nn in n) and \
                (n[0] >> (levels + 1)) % 2 == 0:
            levels += 1
    eb = -5e6
    s_ = h * h / co.epsilon_0
    conf = mgmod.MGConfig(
        bc=[(mgmod.BottomBnd, mgmod.Neumann), (mgmod.TopBnd, mgmod.Neumann),
            (mgmod.LeftBnd, mgmod.Neumann), (mgmod.RightBnd, mgmod.Neumann),
            (mgmod.FrontBnd, mgmod.Dirichlet), (mgmod.BackBnd, mgmod.Dirichlet)],
        s=s_, levels=levels, tolerance=tolerance, smooth1=2, smooth2=2, maxiter=50, precision=precision)
    ws = mgmod.Workspace(conf, n, ctx=ctx)
    dt = 2e-13
    zc = ZhelezniakCollisions.for_table(table)
    photons = PhotonPopulation(cap, collisions=zc, ctx=ctx) if zc is not None else None
    return dict(grid=grid, table=table, popl=popl, fields=fields, ws=ws, conf=conf, eb=eb, dt=dt,
                ntarget=8, nmax=64, photons=photons)


def make_config1(n=(128, 128, 256), h=5e-6, npairs=1_000_000, seed=7, levels=None, cap=None, ctx=None, photo=None):
    """The FP64 streamer workload: 128x128x256 cells of 5 um, 1e6-pair seed of
    width 50 um, V(2,2) to 1e-8 relative residual.  The default capacity (16
    particles per seed pair) holds the avalanche of the full 500-step run."""
    return make_streamer(n, h, npairs, seed, levels, cap or 16 * npairs, "f64", -1e-8, 50e-6 * (h / 5e-6), ctx,
                         photo)


def make_config2(precision="f32", n=(256, 256, 512), h=2.5e-6, npairs=40_000_000, seed=7, levels=None,
                 cap=None, sigma=100e-6, ctx=None):
    """The precision form on the same 0.64 x 0.64 x 1.28 mm box at 2.5 um cells:
    4e7 seed pairs (width 100 um, so that no cell starts above the 64-particle
    merge threshold and the super-particle count stays near 4e7), Poisson to
    1e-10 relative residual.  precision "f64" builds the FP64 reference run."""
    return make_streamer(n, h, npairs, seed, levels, cap or int(npairs * 1.5), precision, -1e-10, sigma,
                         ctx)
