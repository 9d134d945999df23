"""Host mirror of Swarm.jl (Swarm/src/Swarm.jl) on top of libmcs.

  swarm(en; ...) (Swarm.jl:48-186) -> swarm
  resample!(popl, n, tol) (:193-223) -> resample
  gains(proc) (:229-239), collrate (:242-248) -> gains, collrates
Per step: advance! with a FixedField, the position tallies (:126-130) from one
device reduction, and the global resample towards n0.  The KS-test early stop
(:132-134) is not reproduced: runs are fixed-length.
"""
import ctypes

import numpy as np

from . import _lib
from . import constants as co
from ._lib import check
from .jumc import FixedField, advance


def resample(popl, n, tol=0.1, step=None):
    """Swarm.resample!: drop with probability (ncur-n)/ncur, or add Poisson copies."""
    did = ctypes.c_int32()
    s = popl.step if step is None else int(step)
    check(popl.ctx.handle, _lib.lib().mcs_swarm_resample(
        popl.handle, int(n), float(tol), popl.seed, popl.instance, s, ctypes.byref(did)))
    return bool(did.value)


def gains(proc):
    return np.array([-1 if p.kind == "ATTACHMENT" else (1 if p.kind == "IONIZATION" else 0)
                     for p in proc])


def swarm_step(popl, field, dt, n0, tol=0.1):
    """One iteration of the loop at Swarm.jl:124-142 (without the KS test)."""
    st = advance(popl, field, dt, step=popl.step, stats=True)
    resample(popl, n0, tol, step=popl.step)
    popl.step += 1
    return st


def swarm(popl, en_td, dt, nt, n0, n_gas, l=100, tol=0.1):
    """Run nt steps and return the transport coefficients of Swarm.jl:144-169:
    mobility, diffusion (OLS of sigma^2 over the last l steps), nu, alpha, eta
    and the per-process rates k (table order)."""
    eb = en_td * n_gas * co.Td
    field = FixedField(-eb)
    table = popl.collisions
    t = np.arange(nt) * dt
    sigma2 = np.zeros((nt, 3))
    last = nt
    for i in range(nt):
        advance(popl, field, dt, step=popl.step)
        m = popl.moments()
        n = m.nactive
        X = np.array(m.sum_x[:]) / n
        X2 = np.array(m.sum_x2[:]) / n
        sigma2[i] = X2 - X ** 2
        resample(popl, n0, tol, step=popl.step)
        popl.step += 1
        if popl.nparticles() > 0.9 * popl.capacity:
            last = i + 1
            break
    i = last
    A = np.stack([np.ones(l), t[i - l:i]], axis=1)
    coef, *_ = np.linalg.lstsq(A, sigma2[i - l:i], rcond=None)
    D = coef[1] / 2
    m = popl.moments(table)
    W = m.sum_v[2] / m.nactive
    k1 = np.array(m.collrate[: len(table.proc)]) / m.nactive
    g = gains(table.proc)
    return {
        "en": en_td,
        "mobility": -W / (-eb),
        "drift_velocity": W,
        "diffusion_t": (D[0] + D[1]) / 2,
        "diffusion_l": D[2],
        "diffusion": (D[0] + D[1] + D[2]) / 3,
        "nu": float(g @ k1),
        "alpha": float(np.maximum(g, 0) @ k1 / abs(W)),
        "eta": float(-np.minimum(g, 0) @ k1 / abs(W)),
        "k": k1,
        "mean_energy": m.sum_energy / m.nactive,
        "steps": i,
    }
