"""Host mirror of the JuMC API on top of libmcs.

Reference names map as follows (Julia's `!` dropped):
  CollisionTable (JuMC/src/collisions.jl:86-99)   -> CollisionTable
  Population (population.jl:4-36)                 -> Population
  advance! (mixed_population.jl:24-54)            -> advance
  repack! (population.jl:191-215)                 -> repack
  nparticles/actives/weight/meanenergy/maxenergy (population.jl:63-158)
  FixedField (Swarm/src/Swarm.jl:28-32)           -> FixedField
  PhotonState / ZhelezniakCollisions (photon.jl)  -> PhotonPopulation / ZhelezniakCollisions
  MultiPopulation (mixed_population.jl:4-18)      -> MultiPopulation
All compute runs in libmcs on the GPU; this module only marshals arguments.
"""
import ctypes
import math

import numpy as np

from . import _lib
from . import constants as co
from ._lib import check, dptr, i32ptr, u8ptr, u64ptr


class Context:
    """One HIP device and stream (mcs_ctx)."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        rc = L.mcs_ctx_create(int(device), ctypes.byref(h))
        if rc != _lib.MCS_OK:
            raise _lib.MCSError(rc, f"cannot create a context on device {device}")
        self.handle = h
        self.device = device

    def synchronize(self):
        check(self.handle, _lib.lib().mcs_ctx_synchronize(self.handle))

    def stream_ptr(self):
        s = ctypes.c_void_p()
        check(self.handle, _lib.lib().mcs_ctx_stream(self.handle, ctypes.byref(s)))
        return s.value

    def profile(self, enable=True):
        check(self.handle, _lib.lib().mcs_ctx_profile(self.handle, 1 if enable else 0))

    def profile_read(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(self.handle, _lib.lib().mcs_ctx_profile_read(self.handle, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def profile_reset(self):
        check(self.handle, _lib.lib().mcs_ctx_profile_reset(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().mcs_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def context():
    """The process-wide default context on device 0."""
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class CollisionTable:
    """Rate table on an energy grid: proc, energy (J), rate (nproc x ne), maxrate."""

    def __init__(self, proc, energy, rate, maxrate, egrid="uniform", ctx=None):
        self.ctx = ctx or context()
        self.proc = list(proc)
        self.energy = np.ascontiguousarray(energy, dtype=np.float64)
        self.rate = np.ascontiguousarray(rate, dtype=np.float64)
        self.maxrate = float(maxrate)
        self.egrid = {"uniform": _lib.EGRID_UNIFORM, "log": _lib.EGRID_LOG}[egrid]
        nproc, ne = self.rate.shape
        if nproc != len(self.proc) or ne != self.energy.size:
            raise ValueError("rate must have shape (len(proc), len(energy))")
        self.kind = np.array([p.code for p in self.proc], dtype=np.int32)
        self.param = np.array([p.param for p in self.proc], dtype=np.float64)
        emajor = np.ascontiguousarray(self.rate.T)
        h = ctypes.c_void_p()
        check(self.ctx.handle, _lib.lib().mcs_table_create(
            self.ctx.handle, nproc, ne, self.egrid, dptr(self.energy), dptr(emajor),
            i32ptr(self.kind), dptr(self.param), self.maxrate, ctypes.byref(h)))
        self.handle = h
        for j, p in enumerate(self.proc):
            if p.code == _lib.PROC_PHOTOEMISSION:
                numin, numax, mult, pw = p.photo
                check(self.ctx.handle, _lib.lib().mcs_table_set_photoemission(
                    h, j, math.log(numin), math.log(numax), float(mult), float(pw)))

    def photoemission(self):
        """The first PhotoEmission process, or None (main.jl:115)."""
        for p in self.proc:
            if p.code == _lib.PROC_PHOTOEMISSION:
                return p
        return None

    def majorants(self):
        """The energy bands of the banded majorant (mcs_table_majorants):
        (caps in J, majorant rates in 1/s), the last band unbounded."""
        nb = ctypes.c_int32()
        check(self.ctx.handle, _lib.lib().mcs_table_majorants(self.handle, 0, ctypes.byref(nb), None, None))
        cap = np.zeros(nb.value)
        rate = np.zeros(nb.value)
        check(self.ctx.handle, _lib.lib().mcs_table_majorants(self.handle, nb.value, ctypes.byref(nb), dptr(cap),
                                                              dptr(rate)))
        return cap, rate

    def __len__(self):
        return len(self.proc)

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().mcs_table_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class FixedField:
    """Uniform field (0, 0, ez) (Swarm/src/Swarm.jl:28-32)."""

    def __init__(self, ez, ex=0.0, ey=0.0):
        self.e = (float(ex), float(ey), float(ez))

    def cstruct(self):
        f = _lib.Field()
        f.kind = _lib.FIELD_UNIFORM
        for d in range(3):
            f.e[d] = self.e[d]
        f.mesh = None
        return f


class ParticleStates:
    """Host copy of a population's columns (x, v: n x 3; w, s; active; id)."""

    def __init__(self, x, v, w, s, active, id=None, next_id=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = self.x.shape[0]
        self.v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        self.w = np.ascontiguousarray(w, dtype=np.float64).reshape(n)
        self.s = np.ascontiguousarray(s, dtype=np.float64).reshape(n)
        self.active = np.ascontiguousarray(active, dtype=np.uint8).reshape(n)
        self.id = None if id is None else np.ascontiguousarray(id, dtype=np.uint64).reshape(n)
        self.next_id = next_id

    def __len__(self):
        return self.x.shape[0]


class Population:
    """A device-resident particle population with capacity `max_particles`
    (JuMC.Population, population.jl:26-36).  `seed` and `instance` key the
    per-particle Philox streams; `step` counts the advance! calls."""

    def __init__(self, max_particles, particles=None, collisions=None, seed=0, instance=0, ctx=None,
                 precision="f64", mesh=None):
        self.ctx = ctx or context()
        self.collisions = collisions
        self.seed = int(seed)
        self.instance = int(instance)
        self.step = 0
        self.precision = precision
        h = ctypes.c_void_p()
        if precision == "f32":
            # split store on the grid of `mesh` (a GridFields)
            if mesh is None:
                raise ValueError("a split-precision population needs the mesh of its grid")
            check(self.ctx.handle, _lib.lib().mcs_pop_create_split(self.ctx.handle, int(max_particles),
                                                                   mesh.handle, ctypes.byref(h)))
        elif precision == "f64":
            check(self.ctx.handle, _lib.lib().mcs_pop_create(self.ctx.handle, int(max_particles),
                                                             ctypes.byref(h)))
        else:
            raise ValueError(f"unknown precision {precision!r}")
        self.handle = h
        self.capacity = int(max_particles)
        if particles is not None:
            self.upload(particles)

    def upload(self, ps, next_id=0):
        n = len(ps)
        nid = ps.next_id if ps.next_id is not None else next_id
        check(self.ctx.handle, _lib.lib().mcs_pop_upload(
            self.handle, n, dptr(ps.x), dptr(ps.v), dptr(ps.w), dptr(ps.s), u8ptr(ps.active),
            u64ptr(ps.id) if ps.id is not None else None, int(nid)))

    def init_thermal(self, n, kT_ev=1.0):
        """n electrons at the origin with a Maxwellian of temperature kT_ev, drawn on
        the device from this population's (seed, instance) key; ids 0..n-1."""
        vth = float(np.sqrt(kT_ev * co.eV / co.electron_mass))
        check(self.ctx.handle, _lib.lib().mcs_pop_init_thermal(self.handle, int(n), self.seed,
                                                                self.instance, vth))

    def download(self):
        n = self.nparticles()
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        w = np.zeros(n)
        s = np.zeros(n)
        act = np.zeros(n, dtype=np.uint8)
        ids = np.zeros(n, dtype=np.uint64)
        nn = ctypes.c_int64()
        nid = ctypes.c_uint64()
        check(self.ctx.handle, _lib.lib().mcs_pop_download(
            self.handle, n, ctypes.byref(nn), dptr(x), dptr(v), dptr(w), dptr(s), u8ptr(act),
            u64ptr(ids), ctypes.byref(nid)))
        return ParticleStates(x, v, w, s, act, ids, nid.value)

    def nparticles(self):
        n = ctypes.c_int64()
        check(self.ctx.handle, _lib.lib().mcs_pop_size(self.handle, ctypes.byref(n)))
        return n.value

    def energies(self):
        """Energies (J) of the active particles in slot order."""
        n = ctypes.c_int64()
        out = np.zeros(max(self.nparticles(), 1))
        check(self.ctx.handle, _lib.lib().mcs_pop_energies(self.handle, out.size, ctypes.byref(n), dptr(out)))
        return out[: n.value].copy()

    def moments(self, table=None):
        m = _lib.Moments()
        check(self.ctx.handle, _lib.lib().mcs_pop_moments(
            self.handle, table.handle if table is not None else None, ctypes.byref(m)))
        return m

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().mcs_pop_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class ZhelezniakCollisions:
    """The photons' collision table (photon.jl:27-43): one PhotoIonization
    process at the photon's own rate nu, maxrate numax."""

    def __init__(self, numax):
        self.numax = float(numax)

    @classmethod
    def for_table(cls, table):
        """ZhelezniakCollisions(exp(log_numax)) of the table's first PhotoEmission
        row (main.jl:115-118), or None without such a row."""
        p = table.photoemission()
        return None if p is None else cls(math.exp(math.log(p.photo[1])))


class PhotonStates:
    """Host copy of a photon population's columns (PhotonState, photon.jl:11-18)."""

    def __init__(self, x, v, nu, w, s, active, id=None, next_id=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = self.x.shape[0]
        self.v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        self.nu = np.ascontiguousarray(nu, dtype=np.float64).reshape(n)
        self.w = np.ascontiguousarray(w, dtype=np.float64).reshape(n)
        self.s = np.ascontiguousarray(s, dtype=np.float64).reshape(n)
        self.active = np.ascontiguousarray(active, dtype=np.uint8).reshape(n)
        self.id = None if id is None else np.ascontiguousarray(id, dtype=np.uint64).reshape(n)
        self.next_id = next_id

    def __len__(self):
        return self.x.shape[0]


class PhotonPopulation:
    """A device-resident photon population (Population of PhotonState,
    main.jl:116-121), empty unless `photons` are given."""

    def __init__(self, max_particles, photons=None, collisions=None, ctx=None):
        self.ctx = ctx or context()
        self.collisions = collisions
        h = ctypes.c_void_p()
        check(self.ctx.handle, _lib.lib().mcs_pop_create_photons(self.ctx.handle, int(max_particles),
                                                                 ctypes.byref(h)))
        self.handle = h
        self.capacity = int(max_particles)
        if photons is not None:
            self.upload(photons)

    def upload(self, ps, next_id=0):
        nid = ps.next_id if ps.next_id is not None else next_id
        check(self.ctx.handle, _lib.lib().mcs_photons_upload(
            self.handle, len(ps), dptr(ps.x), dptr(ps.v), dptr(ps.nu), dptr(ps.w), dptr(ps.s), u8ptr(ps.active),
            u64ptr(ps.id) if ps.id is not None else None, int(nid)))

    def download(self):
        n = self.nparticles()
        x, v = np.zeros((n, 3)), np.zeros((n, 3))
        nu, w, s = np.zeros(n), np.zeros(n), np.zeros(n)
        act = np.zeros(n, dtype=np.uint8)
        ids = np.zeros(n, dtype=np.uint64)
        nn, nid = ctypes.c_int64(), ctypes.c_uint64()
        check(self.ctx.handle, _lib.lib().mcs_photons_download(
            self.handle, n, ctypes.byref(nn), dptr(x), dptr(v), dptr(nu), dptr(w), dptr(s), u8ptr(act),
            u64ptr(ids), ctypes.byref(nid)))
        return PhotonStates(x, v, nu, w, s, act, ids, nid.value)

    def nparticles(self):
        n = ctypes.c_int64()
        check(self.ctx.handle, _lib.lib().mcs_pop_size(self.handle, ctypes.byref(n)))
        return n.value

    def moments(self, table=None):
        m = _lib.Moments()
        check(self.ctx.handle, _lib.lib().mcs_pop_moments(self.handle, None, ctypes.byref(m)))
        return m

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().mcs_pop_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class MultiPopulation:
    """MultiPopulation(:electron => electron, :photon => photon)
    (mixed_population.jl:4-18); the electrons' (seed, instance, step) key every
    stream of the step."""

    def __init__(self, electron, photon=None):
        self.electron = electron
        self.photon = photon

    def __getitem__(self, name):
        return {"electron": self.electron, "photon": self.photon}[name]

    @property
    def step(self):
        return self.electron.step

    @step.setter
    def step(self, value):
        self.electron.step = value


def advance(popl, efield, dt, tracker=None, *, step=None, stats=True):
    """advance! (mixed_population.jl:24-54) on the GPU.  Returns StepStats, or
    (electron StepStats, photon StepStats) for a MultiPopulation.

    `efield` is a FixedField, a streamer FieldInterp or an axisymmetric
    FieldInterp (axisym); `tracker` is None (VoidCollisionTracker) or a
    CollisionTracker of the field's grid."""
    if isinstance(popl, MultiPopulation):
        return _advance_mixed(popl, efield, dt, tracker, step=step)
    st = _lib.StepStats()
    s = popl.step if step is None else int(step)
    if getattr(efield, "axisymmetric", False):
        if tracker is not None and tracker.fields is not efield.fields:
            raise ValueError("the tracker must belong to the field's grid")
        check(popl.ctx.handle, _lib.lib().mcs_advance_cyl(
            popl.handle, popl.collisions.handle, efield.fields.handle, float(dt), int(tracker is not None),
            popl.seed, popl.instance, s, ctypes.byref(st) if stats else None))
        if step is None:
            popl.step += 1
        return st
    fld = efield.cstruct()
    trk = tracker.mesh_handle() if tracker is not None else None
    check(popl.ctx.handle, _lib.lib().mcs_advance(
        popl.handle, popl.collisions.handle, ctypes.byref(fld), float(dt), trk,
        popl.seed, popl.instance, s, ctypes.byref(st) if stats else None))
    if step is None:
        popl.step += 1
    return st


def _advance_mixed(mpopl, efield, dt, tracker, *, step=None):
    el, ph = mpopl.electron, mpopl.photon
    numax = ph.collisions.numax if ph is not None and ph.collisions is not None else 0.0
    est, pst = _lib.StepStats(), _lib.StepStats()
    s = el.step if step is None else int(step)
    phh = ph.handle if ph is not None else None
    if getattr(efield, "axisymmetric", False):
        if tracker is not None and tracker.fields is not efield.fields:
            raise ValueError("the tracker must belong to the field's grid")
        check(el.ctx.handle, _lib.lib().mcs_advance_cyl_mixed(
            el.handle, phh, el.collisions.handle, numax, efield.fields.handle, float(dt), int(tracker is not None),
            el.seed, el.instance, s, ctypes.byref(est), ctypes.byref(pst)))
    else:
        fld = efield.cstruct()
        trk = tracker.mesh_handle() if tracker is not None else None
        check(el.ctx.handle, _lib.lib().mcs_advance_mixed(
            el.handle, phh, el.collisions.handle, numax, ctypes.byref(fld), float(dt), trk, el.seed, el.instance,
            s, ctypes.byref(est), ctypes.byref(pst)))
    if step is None:
        el.step += 1
    return est, pst


def repack(popl):
    check(popl.ctx.handle, _lib.lib().mcs_repack(popl.handle))


def nparticles(popl):
    return popl.nparticles()


def actives(popl):
    return popl.moments().nactive


def weight(popl):
    return popl.moments().weight


def meanenergy(popl):
    m = popl.moments()
    return m.wenergy / m.weight


def maxenergy(popl):
    return popl.moments().maxenergy
