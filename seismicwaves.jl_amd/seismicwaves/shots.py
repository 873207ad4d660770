"""Sources, receivers, shots and source time functions.

Reference: src/shots/sources.jl:8-91, src/shots/receivers.jl:8-110,
src/shots/shot.jl:9-42, src/utils/utils.jl:6-20.
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
    """ScalarReceivers, receivers.jl:8-53: seismograms (nt, nrec) are
    allocated here and filled in place by swforward/swgradient."""

    def __init__(self, positions, nt: int, dtype=np.float64):
        positions = np.atleast_2d(np.asarray(positions))
        assert positions.shape[0] > 0, "There must be at least one receiver!"
        self.positions = positions
        self.seismograms = np.zeros((int(nt), positions.shape[0]), dtype=dtype)


class ScalarShot:
    """ScalarShot, shot.jl:9-14."""

    def __init__(self, srcs: ScalarSources, recs: ScalarReceivers):
        self.srcs = srcs
        self.recs = recs


class MomentTensor2D:
    """MomentTensor2D, sources.jl:33-37."""

    def __init__(self, Mxx, Mzz, Mxz):
        self.Mxx, self.Mzz, self.Mxz = float(Mxx), float(Mzz), float(Mxz)


class MomentTensorSources:
    """MomentTensorSources, sources.jl:54-76: positions (nsrc, 2), tf
    (nt, nsrc), one MomentTensor2D per source, dominant frequency."""

    def __init__(self, positions, tf, momtens, domfreq):
        positions = np.atleast_2d(np.asarray(positions))
        tf = np.asarray(tf)
        if tf.ndim == 1:
            tf = tf[:, None]
        momtens = list(momtens)
        assert positions.shape[0] > 0, "There must be at least one source!"
        assert positions.shape[0] == tf.shape[1], "Number of sources do not match between positions and time-functions!"
        assert len(momtens) == positions.shape[0], "Number of moment tensors must match number of sources!"
        self.positions = positions
        self.tf = tf
        self.momtens = momtens
        self.domfreq = domfreq


class ExternalForceSources:
    """ExternalForceSources, sources.jl:78-91: tf (nt, ncomp, nsrc)."""

    def __init__(self, positions, tf, domfreq):
        positions = np.atleast_2d(np.asarray(positions))
        tf = np.asarray(tf)
        assert positions.shape[0] > 0, "There must be at least one source!"
        assert tf.ndim == 3 and positions.shape[0] == tf.shape[2], \
            "Number of sources do not match between positions and time-functions!"
        assert tf.shape[1] == positions.shape[1], \
            "Number of components do not match between time-functions and positions!"
        self.positions = positions
        self.tf = tf
        self.domfreq = domfreq


class VectorReceivers:
    """VectorReceivers, receivers.jl:64-100: seismograms (nt, ndim, nrec)."""

    def __init__(self, positions, nt: int, ndim: int = 2, dtype=np.float64):
        positions = np.atleast_2d(np.asarray(positions))
        assert positions.shape[0] > 0, "There must be at least one receiver!"
        self.positions = positions
        self.seismograms = np.zeros((int(nt), int(ndim), positions.shape[0]), dtype=dtype)


class MomentTensorShot:
    """MomentTensorShot, shot.jl:23-28."""

    def __init__(self, srcs: MomentTensorSources, recs: VectorReceivers):
        self.srcs = srcs
        self.recs = recs


class ExternalForceShot:
    """ExternalForceShot, shot.jl:37-42."""

    def __init__(self, srcs: ExternalForceSources, recs: VectorReceivers):
        self.srcs = srcs
        self.recs = recs
