"""Shared builders for the tests: synthetic tables, populations, meshes."""
import os

import numpy as np

from mcstreamer_amd import constants as co
from mcstreamer_amd import gases, lxcat
from oracle import refcpu as R

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def gas_table(ne=4000, egrid="log", o2=0.0, tmpdir=None):
    """(proc, energy_J, rate, maxrate) of the synthetic N2-like (+O2-like) gas."""
    if egrid == "log":
        e_ev = np.geomspace(1e-3, 1e3, ne)
    else:
        e_ev = np.linspace(0.0, 1e3, ne)
    d = tmpdir or os.path.join(HERE, "..", "_work")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"gas_{egrid}_{ne}_{int(o2 > 0)}.json")
    gases.write_gas(path, e_ev, with_attacher=o2 > 0)
    proc, rate, maxrate, _ = lxcat.load_lxcat([path], gases.densities(o2), e_ev * co.eV)
    return proc, e_ev * co.eV, rate, maxrate


def ref_table(proc, energy, rate, maxrate, egrid="log"):
    return R.Table(energy, rate, [p.code for p in proc], [p.param for p in proc], maxrate,
                   egrid=1 if egrid == "log" else 0)


def maxwell_states(n, kT_ev=1.0, seed=0, x0=(0.0, 0.0, 0.0), spread=0.0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3)) * np.sqrt(kT_ev * co.eV / co.electron_mass)
    x = np.tile(np.asarray(x0, dtype=np.float64), (n, 1)) + rng.normal(size=(n, 3)) * spread
    s = -np.log1p(-rng.uniform(size=n))
    return x, v, np.ones(n), s, np.ones(n, dtype=np.uint8)
