"""L2 misfit and its adjoint source, src/inversion/misfits/L2Misfit.jl."""
from __future__ import annotations

import numpy as np


class L2Misfit:
    """L2Misfit, L2Misfit.jl:2-20.

    observed: (nt, nrec); invcov: (nt, nt) matrix, a 1-D diagonal, or None
    for the identity; windows: list of 1-based inclusive (first, last) pairs.
    """

    def __init__(self, observed, invcov=None, windows=()):
        observed = np.array(observed, copy=True)
        if observed.ndim not in (2, 3):
            raise ValueError("Observed data must be a 2D or 3D array!")
        nt = observed.shape[0]
        if invcov is not None:
            invcov = np.asarray(invcov)
            if invcov.ndim == 2:
                assert invcov.shape[0] == invcov.shape[1], "Inverse covariance matrix must be square!"
                assert invcov.shape[0] == nt, "Size of inverse covariance matrix must match the number of timesteps!"
            else:
                assert invcov.shape == (nt,), "Size of inverse covariance matrix must match the number of timesteps!"
        for w in windows:
            assert 1 <= w[0] <= w[1] <= nt, "Windows indices must be between 1 and maximum number of timesteps!"
        self.observed = observed
        self.invcov = invcov
        self.windows = list(windows)

    def _residuals(self, syn):
        res = np.asarray(syn, dtype=np.float64) - self.observed.astype(np.float64)
        if self.windows:
            mask = np.zeros(res.shape[0])
            for a, b in self.windows:
                mask[a - 1:b] = 1.0
            res = res * mask.reshape((-1,) + (1,) * (res.ndim - 1))
        return res

    def _apply_invcov(self, res):
        if self.invcov is None:
            return res
        if self.invcov.ndim == 1:
            return res * self.invcov.reshape((-1,) + (1,) * (res.ndim - 1))
        return np.tensordot(self.invcov, res, axes=(1, 0))

    def calcmisfit(self, recs) -> float:
        """½ rᵀ C⁻¹ r, L2Misfit.jl:24-62 (summed over components for 3D data)."""
        res = self._residuals(recs.seismograms)
        return 0.5 * float(np.sum(res * self._apply_invcov(res)))

    def dchi_du(self, recs) -> np.ndarray:
        """∂χ/∂u = C⁻¹ r, L2Misfit.jl:64-95."""
        return self._apply_invcov(self._residuals(recs.seismograms))
