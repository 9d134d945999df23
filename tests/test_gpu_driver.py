"""The input-file driver on the GPU: a small axisymmetric run from a TOML file in
the reference's schema (MCStreamer/src/main.jl), steps, output files, and the
weight bookkeeping of resample!."""
import json
import os

import numpy as np
import pytest

from mcstreamer_amd import driver, gases
from tests.common import h5read, have_h5py

pytestmark = pytest.mark.gpu

TOML = """
eb = 50.0
field_units = "kV/cm"
n = 4000
maxp = 200000
ntarget = 10
nmax = 15
dt = 1e-12
tmax = 4e-12
cross_sections = ["gas.json"]
rng_seed = 3
[domain]
R = 1e-3
L = 2e-3
M = 32
N = 64
[seed]
z0 = 1e-3
w = 1e-4
[interval]
poisson = 1e-12
output = 2e-12
resample = 1e-12
"""


def _gas(path, photo=False):
    e = np.geomspace(1e-3, 1e3, 400)
    items = gases.n2_like_items(e) + gases.o2_like_items(e)
    if photo:
        items += gases.photo_items(e, rescale=0.5)
    for it in items:
        it["target"] = {"N2like": "N2", "O2like": "O2"}[it["target"]]
        if "absorber" in it:
            it["absorber"] = "O2"
    with open(path, "w") as f:
        json.dump(items, f)


def test_driver_runs_an_axisymmetric_input(tmp_path):
    _gas(tmp_path / "gas.json")
    p = tmp_path / "input.toml"
    p.write_text(TOML)
    run = driver.setup(str(p), energy_points=4000)
    assert run.eb == -5e6
    w0 = run.popl.moments().weight
    log = run.nsteps(4)
    print(run.log, flush=True)
    assert len(run.log) == 4
    # output at t = 0 and 2 ps
    out = sorted(os.listdir(run.outpath))
    assert out == ["0000.h5", "0000.npz", "0001.h5", "0001.npz"]
    with np.load(os.path.join(run.outpath, "0001.npz"), allow_pickle=False) as z:
        assert z["u"].shape == (66, 34) and z["er"].shape == (67, 35)
        assert np.abs(z["er"]).max() > 0
        u1, er1, qfixed1 = z["u"], z["er"], z["qfixed"]
    # the HDF5 export of the same output (plot.jl:210-240), read by the HDF5 library
    print(f"HDF5 library present to read the export: {have_h5py()}", flush=True)
    if have_h5py():
        data, meta = h5read(os.path.join(run.outpath, "0001.h5"))
        assert meta["u"] == ("<f8", "(64, 32)") and meta["er"] == ("<f8", "(64, 33)")
        assert meta["ez"] == ("<f8", "(65, 32)") and meta["p"] == ("<i8", "(64, 32)")
        assert int(data["grid/M"]) == 32 and int(data["grid/N"]) == 64
        np.testing.assert_array_equal(data["u"], u1[1:65, 1:33])
        np.testing.assert_array_equal(data["er"], er1[1:65, 1:34])
        np.testing.assert_array_equal(data["qfixed"], qfixed1[1:65, 1:33])
        # written after the first resample: p counts the cells' particles (+ splits)
        assert data["p"].sum() > 0 and data["wtotal"].sum() > 0
    # resample! keeps the weight (main.jl:300-306 asserts it to 1e-5)
    assert log["nactive"] > 0 and run.log[0]["weight"] == w0
    for r in run.log:
        np.testing.assert_allclose(r["weight_resampled"], r["weight"], rtol=1e-12)
    assert run.log[0]["resample"][1] < run.log[0]["resample"][0]   # the seed's dense cells merged


def test_driver_checkpoint_continues_bitwise(tmp_path):
    _gas(tmp_path / "gas.json")
    p = tmp_path / "input.toml"
    p.write_text(TOML)
    a = driver.setup(str(p), energy_points=4000)
    a.nsteps(2)
    ck = str(tmp_path / "ck.npz")
    driver.save(a, ck)
    a.nsteps(3)
    b = driver.load(ck)
    assert b.i == 2
    b.nsteps(3)
    ga, gb = a.popl.download(), b.popl.download()
    assert len(ga) == len(gb) and ga.next_id == gb.next_id
    np.testing.assert_array_equal(ga.id, gb.id)
    for k in ("x", "v", "w", "s", "active"):
        np.testing.assert_array_equal(getattr(ga, k), getattr(gb, k), err_msg=k)
    for k in ("u", "qfixed", "er", "ez"):
        np.testing.assert_array_equal(a.fields.get(k), b.fields.get(k), err_msg=k)



def test_driver_photoionization_run_and_checkpoint(tmp_path):
    """[photoionization] with a PHOTOEMISSION process: photons are emitted, fly,
    are absorbed into photoelectrons, and a checkpoint (electrons and photons)
    continues bitwise."""
    _gas(tmp_path / "gas.json", photo=True)
    p = tmp_path / "input.toml"
    p.write_text(TOML + "\n[photoionization]\nphoton_weight = 1.0\nphoton_multiplier = 4.0\n")
    a = driver.setup(str(p), energy_points=4000)
    assert a.photons is not None and a.table.photoemission().photo[2] == 4.0
    a.nsteps(3, output=False)
    ck = str(tmp_path / "ck.npz")
    driver.save(a, ck)
    a.nsteps(4, output=False)
    print([(r["photons"], r["photons_emitted"], r["photoelectrons"]) for r in a.log], flush=True)
    assert sum(r["photons_emitted"] for r in a.log) > 0
    assert sum(r["photoelectrons"] for r in a.log) > 0
    b = driver.load(ck)
    b.nsteps(4, output=False)
    for pa, pb, keys in ((a.popl, b.popl, ("x", "v", "w", "s", "active", "id")),
                         (a.photons, b.photons, ("x", "v", "nu", "w", "s", "active", "id"))):
        ga, gb = pa.download(), pb.download()
        assert len(ga) == len(gb) and ga.next_id == gb.next_id
        for k in keys:
            np.testing.assert_array_equal(getattr(ga, k), getattr(gb, k), err_msg=k)
    np.testing.assert_array_equal(a.fields.get("qfixed"), b.fields.get("qfixed"))
