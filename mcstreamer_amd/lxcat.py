"""Collision-table builder: JuMC.load_lxcat (JuMC/src/lxcat.jl:17-156).

Reads LXCat-style JSON items ({"target", "kind", "data": [[eV, m^2], ...],
"threshold" | "mass_ratio", optional "rescale", "weight_scale", "comment"}),
interpolates each cross section linearly with flat extrapolation onto the
energy grid (Interpolations' Gridded(Linear()) + Flat(), lxcat.jl:63-64),
forms nu_j = N_target v(eps) sigma_j(eps), converts EFFECTIVE into ELASTIC by
subtracting the target's excitations and ionizations (ensure_elastic,
lxcat.jl:146-156), appends the null process maxrate - sum(nu), and orders the
processes by their integral over the grid, descending (lxcat.jl:114-133).

Targets are visited in order of first appearance in the files; the reference
iterates a Julia Dict, whose order depends on hashing.  The order only matters
for ties of the integral and for the `origperm` labels.
"""
import json
from dataclasses import dataclass

import numpy as np

from . import constants as co
from . import _lib

KINDS = {
    "NULL": _lib.PROC_NULL,
    "ELASTIC": _lib.PROC_ELASTIC,
    "EXCITATION": _lib.PROC_EXCITATION,
    "IONIZATION": _lib.PROC_IONIZATION,
    "ATTACHMENT": _lib.PROC_ATTACHMENT,
    "PHOTOEMISSION": _lib.PROC_PHOTOEMISSION,
}


@dataclass(frozen=True)
class Process:
    """One collision process (JuMC/src/electron.jl:70-105, collisions.jl:80)."""

    kind: str           # ELASTIC, EXCITATION, IONIZATION, ATTACHMENT, PHOTOEMISSION or NULL
    param: float        # threshold in J, or the mass ratio for ELASTIC
    signature: str = ""
    # PhotoEmission(numin, numax, photon_multiplier, photon_weight) (electron.jl:88-105)
    photo: tuple = None

    @property
    def code(self):
        return KINDS[self.kind]


def _signature(item):
    product = item.get("product", "")
    return f"e + {item['target']} -> {product}... ({item['kind']})"


def load_lxcat(fnames, densities, energy, photon_weight=1.0, photon_multiplier=1.0):
    """Return (proc, rate, maxrate, origperm) like JuMC.load_lxcat.

    `energy` is the node grid in J; `rate` has shape (nproc, len(energy)).  A
    PHOTOEMISSION item's cross section is multiplied by photon_multiplier
    (lxcat.jl:46-49) and its process is PhotoEmission(c N_absorber chi_min,
    c N_absorber chi_max, photon_multiplier, photon_weight) (lxcat.jl:85-92).
    """
    energy = np.asarray(energy, dtype=np.float64)
    db = []
    for fname in fnames:
        with open(fname, "r") as fd:
            db.extend(json.load(fd))
    v = np.sqrt(2.0 * energy / co.electron_mass)
    targets = {}
    for item in db:
        dens = float(densities.get(item["target"], 0.0))
        if dens == 0.0:
            continue
        item = dict(item)
        e0 = np.array([d[0] * co.eV for d in item["data"]], dtype=np.float64)
        cs0 = np.array([d[1] for d in item["data"]], dtype=np.float64)
        if "3-body" in item.get("comment", ""):
            cs0 = cs0 * (densities[item["target"]] / co.centi ** -3)
        if "rescale" in item:
            cs0 = cs0 * item["rescale"]
        if "weight_scale" in item:
            cs0 = cs0 * item["weight_scale"]
        if item["kind"] == "PHOTOEMISSION":
            cs0 = cs0 * photon_multiplier
        sigma = np.interp(energy, e0, cs0)
        item["nu"] = dens * v * sigma
        targets.setdefault(item["target"], []).append(item)

    proc = []
    rows = []
    for _, ps in targets.items():
        for p in ps:  # ensure_elastic
            if p["kind"] != "EFFECTIVE":
                continue
            for p2 in ps:
                if p2 is not p and p2["kind"] in ("EXCITATION", "IONIZATION"):
                    p["nu"] = p["nu"] - p2["nu"]
            p["kind"] = "ELASTIC"
        for p in ps:
            if p["kind"] == "ELASTIC":
                proc.append(Process("ELASTIC", float(p["mass_ratio"]), _signature(p)))
            elif p["kind"] == "PHOTOEMISSION":
                nabs = float(densities[p["absorber"]])
                numin = co.c * nabs * float(p["chi_min"])
                numax = co.c * nabs * float(p["chi_max"])
                proc.append(Process("PHOTOEMISSION", 0.0, _signature(p),
                                    (numin, numax, float(photon_multiplier), float(photon_weight))))
            else:
                proc.append(Process(p["kind"], float(p["threshold"]) * co.eV, _signature(p)))
            rows.append(p["nu"])
    nprocs = len(rows)
    rate = np.zeros((nprocs + 1, energy.size), dtype=np.float64)
    for i, r in enumerate(rows):
        rate[i, :] = r
    nutotal = np.zeros(energy.size, dtype=np.float64)
    for i in range(nprocs + 1):
        nutotal = nutotal + rate[i, :]
    maxrate = float(np.max(nutotal))
    rate[nprocs, :] = maxrate - nutotal
    proc.append(Process("NULL", 0.0, "null"))
    inteng = np.array([float(np.sum(rate[i, :])) for i in range(nprocs + 1)])
    perm = np.argsort(-inteng, kind="stable")
    origperm = np.argsort(perm, kind="stable")
    proc = [proc[i] for i in perm]
    rate = np.ascontiguousarray(rate[perm, :])
    return proc, rate, maxrate, origperm
