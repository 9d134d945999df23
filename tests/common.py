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


# ---- reading HDF5 files back with the HDF5 library ---------------------------
# The image carries the HDF5 library (h5py 3.3 for a Python 3.9 interpreter, and
# h5dump / h5diff).  The files this package writes are read back with it as a
# separate process; tests that need it skip where it is absent.
H5PY = "/opt/conda/bin/python3.9"
H5DUMP = "/opt/conda/bin/h5dump"
H5DIFF = "/opt/conda/bin/h5diff"

_H5READ = r"""
import sys, h5py, numpy as np
out = {}
meta = []
def visit(name, obj):
    if isinstance(obj, h5py.Dataset):
        out[name] = obj[()]
        meta.append(f"{name}\t{obj.dtype.str}\t{obj.shape}")
with h5py.File(sys.argv[1], "r") as f:
    f.visititems(visit)
np.savez(sys.argv[2], **{k.replace("/", "|"): np.asarray(v) for k, v in out.items()})
print("\n".join(meta))
"""


def have_h5py():
    import subprocess
    if not os.path.exists(H5PY):
        return False
    r = subprocess.run([H5PY, "-c", "import h5py"], capture_output=True)
    return r.returncode == 0


def h5read(path):
    """{dataset path: array} and {dataset path: (dtype str, shape)} of an HDF5
    file, read by the HDF5 library (h5py)."""
    import subprocess
    out = path + ".read.npz"
    r = subprocess.run([H5PY, "-c", _H5READ, path, out], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"h5py could not read {path}:\n{r.stderr}")
    meta = {}
    for line in r.stdout.strip().splitlines():
        name, dt, shape = line.split("\t")
        meta[name] = (dt, shape)
    with np.load(out) as z:
        data = {k.replace("|", "/"): z[k] for k in z.files}
    os.remove(out)
    return data, meta


def h5write_lib(path, tree):
    """The same tree written by the HDF5 library itself (h5py), for h5diff."""
    import subprocess
    flat = {}

    def walk(prefix, t):
        for k, v in t.items():
            if isinstance(v, dict):
                walk(prefix + k + "/", v)
            else:
                flat[prefix + k] = np.asarray(v)

    walk("", tree)
    src = path + ".src.npz"
    np.savez(src, **{k.replace("/", "|"): v for k, v in flat.items()})
    script = ("import sys, h5py, numpy as np\n"
              "z = np.load(sys.argv[1])\n"
              "with h5py.File(sys.argv[2], 'w') as f:\n"
              "    for k in z.files:\n"
              "        f.create_dataset(k.replace('|', '/'), data=z[k])\n")
    r = subprocess.run([H5PY, "-c", script, src, path], capture_output=True, text=True)
    os.remove(src)
    if r.returncode != 0:
        raise RuntimeError(r.stderr)
