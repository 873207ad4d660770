"""Seeded synthetic models and acquisition geometries.

These build the benchmark workloads (2D acoustic constant density on a
Gaussian anomaly and on a layered model, batched variable density on a
layered model with elliptical inclusions) from numpy's default_rng(seed),
so the same inputs can be fed to the library and to a CPU reference.
All arrays are (nx, nz) with x the first (fast) axis, as in the reference.
"""
from __future__ import annotations

import numpy as np


def gaussian_anomaly(nx, nz, h, v0=2000.0, dv=500.0, xc=1500.0, zc=1000.0, sigma=300.0):
    x = np.arange(nx)[:, None] * h[0]
    z = np.arange(nz)[None, :] * h[1]
    return v0 + dv * np.exp(-((x - xc) ** 2 + (z - zc) ** 2) / (2 * sigma ** 2))


def layered(nx, nz, h, rng, nlayers=8, vmin=1500.0, vmax=4000.0):
    """nlayers layers with speeds U[vmin, vmax] sorted with depth; each
    interface is z_k + sum_{m=1..3} A_m sin(2 pi x / L_m + phi_m) with
    A in [10, 60] m, L in [500, 3000] m, phi in [0, 2 pi)."""
    speeds = np.sort(rng.uniform(vmin, vmax, nlayers))
    x = np.arange(nx) * h[0]
    z = np.arange(nz) * h[1]
    depth = (nz - 1) * h[1]
    vp = np.full((nx, nz), speeds[0])
    for k in range(1, nlayers):
        zk = k * depth / nlayers
        A = rng.uniform(10, 60, 3)
        L = rng.uniform(500, 3000, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        iface = zk + sum(A[m] * np.sin(2 * np.pi * x / L[m] + ph[m]) for m in range(3))
        vp[z[None, :] >= iface[:, None]] = speeds[k]
    return vp


def add_ellipses(vp, h, rng, n=200, amp=0.15, amin=20.0, amax=200.0):
    """n ellipses (semi-axes U[amin, amax] m, rotation U[0, pi)) scaling vp
    by (1 +- amp)."""
    nx, nz = vp.shape
    x = np.arange(nx)[:, None] * h[0]
    z = np.arange(nz)[None, :] * h[1]
    out = vp.copy()
    for _ in range(n):
        cx, cz = rng.uniform(0, (nx - 1) * h[0]), rng.uniform(0, (nz - 1) * h[1])
        a, b = rng.uniform(amin, amax, 2)
        th = rng.uniform(0, np.pi)
        sgn = rng.choice([-1.0, 1.0])
        i0, i1 = max(int((cx - amax) / h[0]), 0), min(int((cx + amax) / h[0]) + 2, nx)
        j0, j1 = max(int((cz - amax) / h[1]), 0), min(int((cz + amax) / h[1]) + 2, nz)
        xx, zz = x[i0:i1] - cx, z[:, j0:j1] - cz
        u = xx * np.cos(th) + zz * np.sin(th)
        w = -xx * np.sin(th) + zz * np.cos(th)
        inside = (u / a) ** 2 + (w / b) ** 2 <= 1
        out[i0:i1, j0:j1][inside] *= 1 + sgn * amp
    return out


def gardner(vp, rng, noise=0.05):
    """rho = 310 vp^0.25 (1 + noise U[-1, 1])."""
    return 310.0 * vp ** 0.25 * (1 + noise * rng.uniform(-1, 1, vp.shape))


def gaussian_smooth(m, sigma_cells):
    """Separable Gaussian smoothing with reflecting edges."""
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(m, sigma_cells, mode="reflect")


def ricker_tf(nt, dt, f0, t0):
    t = np.arange(nt) * dt
    arg = (np.pi * f0 * (t - t0)) ** 2
    return (1 - 2 * arg) * np.exp(-arg)


def surface_receivers(n, nx, h, halo, z):
    """n receivers evenly spaced over the x-range outside the C-PML, depth z."""
    x = np.linspace(halo * h[0], (nx - 1 - halo) * h[0], n)
    return np.stack([x, np.full(n, z)], 1)
