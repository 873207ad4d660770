"""Sources, receivers, shots and source time functions.

Reference: src/shots/sources.jl:8-26, src/shots/receivers.jl:100-145,
src/shots/shot.jl:211-216, src/utils/utils.jl:6-20.
"""
from __future__ import annotations

import numpy as np


def rickerstf(t, t0, f0):
    """Ricker wavelet (second derivative of a Gaussian), utils.jl:6."""
    arg = (np.pi * f0 * (np.asarray(t) - t0)) ** 2
    return (1 - 2 * arg) * np.exp(-arg)


def gaussderivstf(t, t0, f0):
    """First derivative of a Gaussian, utils.jl:13."""
    t = np.asarray(t)
    return (t - t0) * np.exp(-(np.pi * f0 * (t - t0)) ** 2)


def gaussstf(t, t0, f0):
    """Gaussian, utils.jl:20."""
    t = np.asarray(t)
    return -np.exp(-(np.pi * f0 * (t - t0)) ** 2) / (2 * (np.pi * f0) ** 2)


class ScalarSources:
    """ScalarSources, sources.jl:8-26: positions (nsrc, ndim) in metres,
    tf (nt, nsrc) source time functions, domfreq the dominant frequency."""

    def __init__(self, positions, tf, domfreq):
        positions = np.atleast_2d(np.asarray(positions))
        tf = np.asarray(tf)
        if tf.ndim == 1:
            tf = tf[:, None]
        assert positions.shape[0] > 0, "There must be at least one source!"
        assert positions.shape[0] == tf.shape[1], "Number of sources do not match between positions and time-functions!"
        self.positions = positions
        self.tf = tf
        self.domfreq = domfreq


class ScalarReceivers:
    """ScalarReceivers, receivers.jl:100-145: seismograms (nt, nrec) are
    allocated here and filled in place by swforward/swgradient."""

    def __init__(self, positions, nt: int, dtype=np.float64):
        positions = np.atleast_2d(np.asarray(positions))
        assert positions.shape[0] > 0, "There must be at least one receiver!"
        self.positions = positions
        self.seismograms = np.zeros((int(nt), positions.shape[0]), dtype=dtype)


class ScalarShot:
    """ScalarShot, shot.jl:211-216."""

    def __init__(self, srcs: ScalarSources, recs: ScalarReceivers):
        self.srcs = srcs
        self.recs = recs
