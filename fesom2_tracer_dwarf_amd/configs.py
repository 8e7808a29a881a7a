"""Named synthetic workloads and their initial conditions.

The analytic fields follow the reference's analytic dwarf
(src/fesom_analytic.F90:306-331): T = 15 + 5 sin(2 pi x/Lx) cos(2 pi y/Ly)
- 10 z_n with z_n = (k-1)/(nl-2) from surface to the deepest layer; S = 34 +
a Gaussian blob times (1 - z_n), here of amplitude 1 with a seeded centre and
width. Passive tracers are seeded Gaussian blobs, 0/1 step functions and
white noise in [0, 1].
"""
from __future__ import annotations

import numpy as np

WORKLOADS = {
    # 20x20 periodic nodes, 10 interfaces of 10 m, flat bottom, 2 tracers, 10 steps
    "analytic_20x20x10": dict(nx=20, ny=20, nl=10, dz=10.0, Lx=1.0e5, Ly=1.0e5, lev=(10, 10), ntr=2, steps=10),
    # 1000x1000 periodic nodes, 64 interfaces, bottom level U{3..64}, 2 tracers, 100 steps
    "periodic_1000x1000x64": dict(nx=1000, ny=1000, nl=64, dz=10.0, Lx=2.0e6, Ly=2.0e6, lev=(3, 64), ntr=2,
                                  steps=100),
}


def _zn(L):
    return np.arange(L, dtype=np.float64) / max(L - 1, 1)


def tracer_T(coord, L, Lx, Ly):
    x, y = coord[0::2], coord[1::2]
    col = 15.0 + 5.0 * np.sin(2.0 * np.pi * x / Lx) * np.cos(2.0 * np.pi * y / Ly)
    return np.asfortranarray(col[None, :] - 10.0 * _zn(L)[:, None])


def _periodic_d(a, c, Lp):
    d = a - c
    if Lp:
        d = d - Lp * np.round(d / Lp)
    return d


def tracer_blob(coord, L, Lx, Ly, rng, base=0.0, amp=1.0, periodic=True, depth_decay=True):
    x, y = coord[0::2], coord[1::2]
    xc, yc = rng.uniform(0.2, 0.8) * Lx, rng.uniform(0.2, 0.8) * Ly
    w = (0.08 + 0.08 * rng.uniform()) * Lx
    dx = _periodic_d(x, xc, Lx if periodic else 0.0)
    dy = _periodic_d(y, yc, Ly if periodic else 0.0)
    g = np.exp(-(dx * dx + dy * dy) / (w * w))
    prof = (1.0 - _zn(L)) if depth_decay else np.ones(L)
    return np.asfortranarray(base + amp * g[None, :] * prof[:, None])


def tracer_step(coord, L, Lx, Ly, rng):
    x, y = coord[0::2], coord[1::2]
    x0, y0 = rng.uniform(0.25, 0.5) * Lx, rng.uniform(0.25, 0.5) * Ly
    inside = ((x >= x0) & (x < x0 + 0.4 * Lx) & (y >= y0) & (y < y0 + 0.4 * Ly)).astype(np.float64)
    return np.asfortranarray(np.repeat(inside[None, :], L, axis=0))


def tracer_noise(coord, L, rng):
    n = coord.size // 2
    return np.asfortranarray(rng.uniform(0.0, 1.0, size=(L, n)))


def initial_tracers(coord, nl, Lx, Ly, ntr, seed=0, periodic=True):
    """T, S, then passive tracers cycling blob / step / noise: list of (nl-1, nod2D)."""
    L = nl - 1
    rng = np.random.default_rng(seed)
    out = [tracer_T(coord, L, Lx, Ly)]
    if ntr > 1:
        out.append(tracer_blob(coord, L, Lx, Ly, rng, base=34.0, amp=1.0, periodic=periodic))
    kinds = ("step", "blob", "noise")
    i = 0
    while len(out) < ntr:
        k = kinds[i % 3]
        if k == "step":
            out.append(tracer_step(coord, L, Lx, Ly, rng))
        elif k == "blob":
            out.append(tracer_blob(coord, L, Lx, Ly, rng, periodic=periodic, depth_decay=False))
        else:
            out.append(tracer_noise(coord, L, rng))
        i += 1
    return out[:ntr]
