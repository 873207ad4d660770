"""CCTSMisfit (CCTSMisfit.jl): cross-correlation traveltime misfit and its
adjoint source, pinned by what the definitions fix.

* A trace delayed by d samples against itself: the |xcorr| peak sits at lag
  d, the 1-based index nt + d of the 2nt - 1 lags, so the reference's
  dtau = (argmax - (nt - 1)) dt is (d + 1) dt (the reference's index
  arithmetic: identical traces give dt, not 0) and the misfit is
  nrec ((d + 1) dt)^2 / 2.  A delay of the synthetic instead flips the sign.
* The adjoint source dtau / N * ds/dt against the closed-form derivative of
  a finely sampled Ricker wavelet and its energy integral (to the accuracy
  of the centred difference and the trapezoidal rule)."""
import numpy as np

import seismicwaves as sw


def _ricker(t, f0, t0):
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def _ricker_dt(t, f0, t0):
    b = (np.pi * f0) ** 2
    x = t - t0
    return np.exp(-b * x * x) * (-4 * b * x - 2 * b * x * (1 - 2 * b * x * x))


class _Recs:
    def __init__(self, s):
        self.seismograms = s


def test_delay_and_misfit_of_shifted_traces():
    nt, dt, f0 = 600, 1e-3, 15.0
    t = dt * np.arange(nt)
    syn = np.stack([_ricker(t, f0, 0.1 + 0.01 * r) for r in range(4)], 1)
    for d in (0, 7, 23):
        obs = np.zeros_like(syn)
        obs[d:] = syn[:nt - d]                      # observed = synthetic delayed by d samples
        m = sw.CCTSMisfit(dt, obs)
        assert np.isclose(m.calcmisfit(_Recs(syn)), 4 * ((d + 1) * dt) ** 2 / 2, rtol=1e-12)
        # the synthetic delayed instead: lag -d
        m2 = sw.CCTSMisfit(dt, syn)
        assert np.isclose(m2.calcmisfit(_Recs(obs)), 4 * ((1 - d) * dt) ** 2 / 2, rtol=1e-12, atol=1e-18)


def test_vector_data_sums_components():
    nt, dt = 400, 1e-3
    t = dt * np.arange(nt)
    s = _ricker(t, 12.0, 0.12)
    syn = np.zeros((nt, 2, 3))
    obs = np.zeros((nt, 2, 3))
    for r in range(3):
        for c in range(2):
            d = 3 * r + c
            syn[:, c, r] = s
            obs[d:, c, r] = s[:nt - d]
    m = sw.CCTSMisfit(dt, obs)
    want = sum(((3 * r + c + 1) * dt) ** 2 / 2 for r in range(3) for c in range(2))
    assert np.isclose(m.calcmisfit(_Recs(syn)), want, rtol=1e-12)


def test_adjoint_source_against_closed_form():
    nt, dt, f0, t0, d = 4000, 2.5e-5, 100.0, 0.03, 40
    t = dt * np.arange(nt)
    syn = _ricker(t, f0, t0)[:, None]
    obs = np.zeros_like(syn)
    obs[d:] = syn[:nt - d]
    m = sw.CCTSMisfit(dt, obs)
    adj = m.dchi_du(_Recs(syn))[:, 0]
    tau = (d + 1) * dt
    ds = _ricker_dt(t, f0, t0)
    # energy of ds/dt over t = dt .. nt dt (the derivative at sample k is
    # attached to time (k + 1) dt by the reference's time axis)
    tt = dt * np.arange(1, nt + 1)
    energy = np.trapezoid(ds ** 2, tt)
    want = tau / energy * ds
    inner = slice(1, nt - 1)
    err = np.max(np.abs(adj[inner] - want[inner])) / np.max(np.abs(want))
    assert err < 1e-4, err
    assert adj[-1] == 0.0
