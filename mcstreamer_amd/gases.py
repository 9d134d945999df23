"""Synthetic gases of the benchmark workloads, written in the LXCat JSON layout
that load_lxcat reads (JuMC/data/LXCat-June2013.json has the same fields).

N2-like: elastic 6e-20 m^2 (mass ratio 3.9e-5); excitations at 6.2, 8.4 and
11.0 eV with sigma = 2e-21 (1 - eps_th/eps); ionization I = 15.6 eV with
sigma = 2.5e-20 ln(eps/I)/(eps/I).  O2-like attacher: sigma = 1.5e-22
exp(-(eps - 6.5)^2 / 2) m^2 (centre 6.5 eV, width 1 eV).  Photoemission (the
layout of the reference's PHOTOEMISSION items, lxcat.jl:85-92): the N2-like
ionization cross section times `rescale`, absorbed by the O2-like gas with
chi between 1.09e-23 and 6.21e-21 m^2 (Zhelezniak's range at 300 K).
"""
import json
import os

import numpy as np

from . import constants as co

N_GAS = 2.45e25  # m^-3


def log_grid(n=4000, emin_ev=1e-3, emax_ev=1e3):
    """Log-spaced energy nodes in J (the 4000-point grid of the swarm workload)."""
    return np.geomspace(emin_ev, emax_ev, n) * co.eV


def uniform_grid(n=4000, emax_ev=1e3):
    return np.linspace(0.0, emax_ev, n) * co.eV


def n2_like_items(nodes_ev):
    e = np.asarray(nodes_ev, dtype=np.float64)
    items = [
        {
            "target": "N2like",
            "kind": "ELASTIC",
            "mass_ratio": 3.9e-5,
            "comment": "synthetic constant elastic cross section",
            "data": [[float(x), 6e-20] for x in e],
        }
    ]
    for th in (6.2, 8.4, 11.0):
        sig = np.where(e > th, 2e-21 * (1.0 - th / np.where(e > 0, e, 1.0)), 0.0)
        items.append(
            {
                "target": "N2like",
                "kind": "EXCITATION",
                "threshold": th,
                "product": f"N2like*({th})",
                "comment": "synthetic excitation",
                "data": [[float(x), float(s)] for x, s in zip(e, sig)],
            }
        )
    ion = 15.6
    r = np.where(e > ion, e / ion, 1.0)
    sig = np.where(e > ion, 2.5e-20 * np.log(r) / r, 0.0)
    items.append(
        {
            "target": "N2like",
            "kind": "IONIZATION",
            "threshold": ion,
            "product": "N2like^+",
            "comment": "synthetic ionization",
            "data": [[float(x), float(s)] for x, s in zip(e, sig)],
        }
    )
    return items


def o2_like_items(nodes_ev):
    e = np.asarray(nodes_ev, dtype=np.float64)
    sig = 1.5e-22 * np.exp(-((e - 6.5) ** 2) / 2.0)
    return [
        {
            "target": "O2like",
            "kind": "ATTACHMENT",
            "threshold": 0.0,
            "product": "O2like^-",
            "comment": "synthetic Gaussian attachment",
            "data": [[float(x), float(s)] for x, s in zip(e, sig)],
        }
    ]


CHI_MIN, CHI_MAX = 1.09e-23, 6.21e-21


def photo_items(nodes_ev, rescale=0.05):
    e = np.asarray(nodes_ev, dtype=np.float64)
    ion = 15.6
    r = np.where(e > ion, e / ion, 1.0)
    sig = np.where(e > ion, 2.5e-20 * np.log(r) / r, 0.0)
    return [
        {
            "target": "N2like",
            "kind": "PHOTOEMISSION",
            "absorber": "O2like",
            "chi_min": CHI_MIN,
            "chi_max": CHI_MAX,
            "rescale": rescale,
            "comment": "synthetic photoemission: the ionization cross section, rescaled",
            "data": [[float(x), float(s)] for x, s in zip(e, sig)],
        }
    ]


def write_gas(path, nodes_ev, with_attacher=True, with_photo=False, photo_rescale=0.05):
    items = n2_like_items(nodes_ev)
    if with_attacher:
        items += o2_like_items(nodes_ev)
    if with_photo:
        items += photo_items(nodes_ev, photo_rescale)
    with open(path, "w") as f:
        json.dump(items, f)
    return path


def densities(o2_fraction=0.0, n_gas=N_GAS):
    d = {"N2like": n_gas * (1.0 - o2_fraction)}
    if o2_fraction > 0.0:
        d["O2like"] = n_gas * o2_fraction
    return d


def default_gas_file(directory, nodes_ev, with_attacher=True, tag="gas"):
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, f"{tag}_{len(nodes_ev)}.json")
    if not os.path.exists(path):
        write_gas(path, nodes_ev, with_attacher)
    return path
