"""Seeded random inputs in the reference storage for kernel-level parity."""
from __future__ import annotations

import numpy as np

from codens_amd import layout as L

GAM = 1.4


def random_mesh(dims, seed, jitter=0.01):
    lay, sn, vol, _ = random_mesh_xyz(dims, seed, jitter)
    return lay, sn, vol


def random_mesh_xyz(dims, seed, jitter=0.01):
    """random_mesh and its node coordinates (x, y, z), stored like the cells."""
    lay = L.Layout(dims)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.zeros(lay.ip11); y = np.zeros(lay.ip11); z = np.zeros(lay.ip11)
    for b in lay.blocks:
        K, J, I = np.meshgrid(np.arange(b.nkd), np.arange(b.njd), np.arange(b.nid), indexing="ij")
        b.view(x)[:] = I * 0.1 + jitter * rng.standard_normal(I.shape)
        b.view(y)[:] = J * 0.13 + jitter * rng.standard_normal(I.shape)
        b.view(z)[:] = K * 0.07 + jitter * rng.standard_normal(I.shape)
    sn, vol = L.metrics(lay, x, y, z)
    return lay, sn, vol, (x, y, z)


def random_state(n, rng, pinfl=0.0, amp=0.1):
    v = np.zeros((5, n))
    rho = 1 + amp * rng.standard_normal(n)
    u = 0.5 + amp * rng.standard_normal(n)
    vv = amp * rng.standard_normal(n)
    w = amp * rng.standard_normal(n)
    p = 1 / GAM + 0.5 * amp * rng.standard_normal(n)
    v[0] = rho; v[1] = rho * u; v[2] = rho * vv; v[3] = rho * w
    v[4] = (p - pinfl) / (GAM - 1) + pinfl + 0.5 * rho * (u * u + vv * vv + w * w)
    return v


def all_thermo(v, pinfl=0.0):
    """p, T, c on every stored entry (so ghost entries hold valid values)."""
    p = (GAM - 1) * (v[4] - 0.5 * (v[1] ** 2 + v[2] ** 2 + v[3] ** 2) / v[0] - pinfl)
    t = GAM * p / v[0]
    return p, t, np.sqrt(t)


def relerr(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))
