"""Band-limited spreading of off-grid sources and receivers over grid taps.

Reference: src/utils/utils.jl:62-214 (kaiser, modd, coeffsinc1D,
spread_positions) as used by the elastic possrcrec_scaletf
(src/models/elastic/ela_models.jl:6-90).  A point becomes, per dimension, a
Kaiser-windowed sinc over the taps i0-r-1 .. i0+r+1 around its nearest grid
point (weights |c| <= 1e-15 dropped), taps beyond the grid are folded back
across the boundary (with a sign flip when `mirror`), coincident taps are
summed, and the 2D weights are the tensor product of the 1D ones.
"""
from __future__ import annotations

import numpy as np

KAISER_R = 4
KAISER_BETA = 6.31


def coeffsinc1d(x0, dx, nx, r=KAISER_R, beta=KAISER_BETA, xstart=0.0, mirror=False, xbl=0.0, xbr=None):
    """1D taps of a band-limited delta at x0 on the grid xstart + dx*(0..nx-1).
    Returns (1-based indices, weights), indices ascending."""
    if xbr is None:
        xbr = (nx - 1) * dx
    xs = xstart + dx * np.arange(nx)
    i0 = int(np.argmin(np.abs(x0 - xs))) + 1
    idx = np.arange(i0 - r - 1, i0 + r + 2)
    xc = (idx - 1) * dx + xstart
    d = xc - x0
    rr = r * dx
    win = np.where(np.abs(d) <= rr, np.i0(beta * np.sqrt(np.clip(1.0 - (d / rr) ** 2, 0.0, None))) / np.i0(beta), 0.0)
    coe = win * np.sinc(d / dx)
    keep = ~(np.abs(coe) <= 1e-15)
    idx, xc, coe = idx[keep], xc[keep], coe[keep]
    left, right = xc < xbl, xc > xbr
    folded = left | right
    if np.any(folded):
        xm = np.where(left, 2 * xbl - xc, 2 * xbr - xc)
        near = np.argmin(np.abs(xm[:, None] - xs[None, :]), axis=1) + 1
        idx = np.where(folded, near, idx)
        if mirror:
            coe = np.where(folded, -coe, coe)
    uniq, inv = np.unique(idx, return_inverse=True)
    w = np.zeros(len(uniq))
    np.add.at(w, inv, coe)
    return uniq.astype(np.int64), w


def spread_positions(positions, spacing, gridsize, shift=(0.0, 0.0), mirror=False, r=KAISER_R, beta=KAISER_BETA):
    """spread_positions with freesurfposition=:ongridbound: per point, the
    taps (ntaps, 2) 1-based and their weights, first dimension fastest."""
    positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
    ndim = positions.shape[1]
    out = []
    for p in positions:
        per = [coeffsinc1d(p[d], spacing[d], gridsize[d], r, beta, shift[d], mirror, 0.0,
                           (gridsize[d] - 1) * spacing[d]) for d in range(ndim)]
        ii, jj = np.meshgrid(per[0][0], per[1][0], indexing="ij")
        wi, wj = np.meshgrid(per[0][1], per[1][1], indexing="ij")
        ij = np.stack([ii.ravel(order="F"), jj.ravel(order="F")], axis=1)
        out.append((ij, (wi * wj).ravel(order="F")))
    return out


def nearest_taps(idx):
    """One unit tap per point at its nearest grid point (sincinterp = false)."""
    return [(np.asarray(idx[k:k + 1], dtype=np.int64), np.ones(1)) for k in range(idx.shape[0])]
