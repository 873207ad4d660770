"""C-PML coefficient profiles computed on the host, as the reference does.

Reference: src/models/cpmlcoeffs.jl:18-100 and init_bdc!,
src/models/acoustic/acou_init_bc.jl:7-40.
"""
from __future__ import annotations

import numpy as np


def _profile(halo: int, dt: float, npower: float, d0: float, alpha_max: float, half: bool):
    """calc_Kab_CPML_axis (cpmlcoeffs.jl:43-100) with K = 1.

    Returns (a, b) of length 2*halo (half grid) or 2*(halo+1) (on grid),
    the left side first.  Distances are evaluated in float64 like Julia's
    LinRange and converted by the caller."""
    shift = 0.5 if half else 0.0
    size = halo if half else halo + 1
    dist = shift + np.arange(size, dtype=np.float64)
    if halo != 0:
        left, right = dist[::-1] / halo, dist / halo
    else:
        left, right = dist[::-1], dist
    out_a, out_b = [], []
    for nd in (left, right):
        d = d0 * nd ** npower
        alpha = alpha_max * (1.0 - nd)
        b = np.exp(-(d + alpha) * dt)
        with np.errstate(invalid="ignore", divide="ignore"):
            a = d * (b - 1.0) / (d + alpha)
        out_a.append(a)
        out_b.append(b)
    return np.concatenate(out_a), np.concatenate(out_b)


def cpml_axis(vel_max, dt, halo, rcoef, thickness, f0, dtype):
    """compute_CPML_coefficientsAxis! (cpmlcoeffs.jl:18-41) for one axis."""
    T = np.dtype(dtype).type
    alpha_max = T(np.pi * T(f0))
    npower = 2.0
    if halo == 0:
        d0 = T(0.0)
    else:
        d0 = T(-(npower + 1) * T(vel_max) * np.log(T(rcoef)) / (2.0 * T(thickness)))
    a, b = _profile(halo, float(T(dt)), npower, float(d0), float(alpha_max), half=False)
    ah, bh = _profile(halo, float(T(dt)), npower, float(d0), float(alpha_max), half=True)
    return {"a": a.astype(T), "a_h": ah.astype(T), "b": b.astype(T), "b_h": bh.astype(T)}


def init_bdc(vel_max, dt, spacing, halo, rcoef, freeboundtop, f0, dtype):
    """init_bdc! (acou_init_bc.jl:7-40): one coefficient set per axis; with a
    free top the first half of the last axis gets a = 0, b = 1."""
    axes = [cpml_axis(vel_max, dt, halo, rcoef, h * halo, f0, dtype) for h in spacing]
    if freeboundtop and axes:
        last = axes[-1]
        for key, val in (("a", 0.0), ("a_h", 0.0), ("b", 1.0), ("b_h", 1.0)):
            last[key][: len(last[key]) // 2] = val
    return axes
