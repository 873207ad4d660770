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


def _xcorr(u, v):
    """DSP.xcorr(u, v) of two equal-length series (padmode :none): the
    2n-1 lags of sum_j u[j] v[j - k], k = -(n-1)..n-1, i.e. conv(u, reverse(v))."""
    return np.convolve(u, v[::-1])


def _trapezoid(x, y):
    """NumericalIntegration.integrate(x, y) with its default trapezoidal rule."""
    return float(np.sum(np.diff(x) * (y[1:] + y[:-1]) / 2))


class CCTSMisfit:
    """CCTSMisfit, src/inversion/misfits/CCTSMisfit.jl:1-15: cross-correlation
    traveltime misfit.  For each trace the delay dtau = (argmax |xcorr(obs,
    syn)| - (nt - 1)) dt, with the 1-based argmax of the 2nt-1 lags exactly
    as the reference forms it (:17-28); misfit = sum dtau^2 / 2.  `std` is
    stored as the reference stores it (it enters neither formula)."""

    def __init__(self, dt, observed, std=1.0):
        observed = np.array(observed, copy=True)
        if observed.ndim not in (2, 3):
            raise ValueError("Observed data must be a 2D or 3D array!")
        assert dt > 0, "Timestep size must be positive!"
        assert std > 0, "Standard deviation must be positive!"
        self.dt = dt
        self.std = std
        self.observed = observed

    def _traces(self, syn):
        """(observed, synthetic) column pairs in the reference's loop order
        (receivers outer, components inner)."""
        syn = np.asarray(syn)
        if syn.ndim == 2:
            for r in range(syn.shape[1]):
                yield (r,), self.observed[:, r], syn[:, r]
        else:
            for r in range(syn.shape[2]):
                for d in range(syn.shape[1]):
                    yield (d, r), self.observed[:, d, r], syn[:, d, r]

    def _delay(self, o, s):
        nt = s.shape[0]
        k = int(np.argmax(np.abs(_xcorr(np.asarray(o, np.float64), np.asarray(s, np.float64))))) + 1
        return (k - (nt - 1)) * self.dt

    def calcmisfit(self, recs) -> float:
        """CCTSMisfit.jl:17-43."""
        return float(sum(self._delay(o, s) ** 2 / 2 for _, o, s in self._traces(recs.seismograms)))

    def dchi_du(self, recs) -> np.ndarray:
        """∂χ_∂u, CCTSMisfit.jl:46-80: dtau / N * ds/dt with the centred
        time derivative (first sample s[2] / 2dt, last sample 0) and N its
        trapezoidal energy over t = dt, 2dt, ..., nt dt."""
        syn = np.asarray(recs.seismograms, dtype=np.float64)
        dt = self.dt
        nt = syn.shape[0]
        ds = np.zeros_like(syn)
        ds[1:-1] = (syn[2:] - syn[:-2]) / (2 * dt)
        ds[0] = syn[1] / (2 * dt)
        times = dt * np.arange(1, nt + 1)
        for idx, o, s in self._traces(syn):
            col = (slice(None),) + idx
            tau = self._delay(o, s)
            ds[col] = (tau / _trapezoid(times, ds[col] ** 2)) * ds[col]
        return ds
