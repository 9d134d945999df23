"""Shared builders for the tests: synthetic tables, populations, meshes."""
import os
import tempfile

import numpy as np

from mcstreamer_amd import constants as co
from mcstreamer_amd import gases, lxcat
from oracle import refcpu as R

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def gas_table(ne=4000, egrid="log", o2=0.0, tmpdir=None, photo=False, photon_weight=1.0, photon_multiplier=1.0,
              photo_rescale=0.05):
    """(proc, energy_J, rate, maxrate) of the synthetic N2-like (+O2-like) gas,
    with photoemission (absorbed by the O2-like gas) when `photo`."""
    if egrid == "log":
        e_ev = np.geomspace(1e-3, 1e3, ne)
    else:
        e_ev = np.linspace(0.0, 1e3, ne)
    d = tmpdir or tempfile.gettempdir()
    path = os.path.join(d, f"mcs_gas_{egrid}_{ne}_{int(o2 > 0)}_{int(photo)}_{photo_rescale}_{os.getpid()}.json")
    gases.write_gas(path, e_ev, with_attacher=o2 > 0, with_photo=photo, photo_rescale=photo_rescale)
    proc, rate, maxrate, _ = lxcat.load_lxcat([path], gases.densities(o2), e_ev * co.eV,
                                              photon_weight=photon_weight, photon_multiplier=photon_multiplier)
    return proc, e_ev * co.eV, rate, maxrate


def ref_table(proc, energy, rate, maxrate, egrid="log"):
    photo = {j: p.photo for j, p in enumerate(proc) if p.kind == "PHOTOEMISSION"}
    return R.Table(energy, rate, [p.code for p in proc], [p.param for p in proc], maxrate,
                   egrid=1 if egrid == "log" else 0, photo=photo or None)


def maxwell_states(n, kT_ev=1.0, seed=0, x0=(0.0, 0.0, 0.0), spread=0.0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3)) * np.sqrt(kT_ev * co.eV / co.electron_mass)
    x = np.tile(np.asarray(x0, dtype=np.float64), (n, 1)) + rng.normal(size=(n, 3)) * spread
    s = -np.log1p(-rng.uniform(size=n))
    return x, v, np.ones(n), s, np.ones(n, dtype=np.uint8)


def mesh_geometry(n, h):
    nx, ny, nz = n
    return np.array(n, dtype=np.float64) * h


def random_particles(n, L, seed, p_inactive=0.0, p_outside=0.0, kT_ev=2.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, size=(n, 3)) * L
    out = rng.uniform(size=n) < p_outside
    x[out] = x[out] + rng.choice([-1.0, 1.0], size=(out.sum(), 1)) * L * 0.6
    v = rng.normal(size=(n, 3)) * np.sqrt(kT_ev * co.eV / co.electron_mass)
    w = rng.integers(1, 6, size=n).astype(np.float64)
    s = -np.log1p(-rng.uniform(size=n))
    act = (rng.uniform(size=n) >= p_inactive).astype(np.uint8)
    return x, v, w, s, act


def smooth_potential(n, h, amp=0.5, seed=0):
    """A smooth cell-centred potential with ghosts (x fastest), volts."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    k, j, i = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 3)
    u = amp * (np.sin(2 * np.pi * i / nx + ph[0]) * np.cos(2 * np.pi * j / ny + ph[1])
               + np.sin(np.pi * k / nz + ph[2]))
    return u.reshape(-1)
