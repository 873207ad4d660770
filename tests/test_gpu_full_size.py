"""Parity at the full grid sizes of the benchmark configurations, through
properties that do not need the oracle (which would take hours there):

* linearity: doubling the source time function doubles every seismogram
  (every operation of a step is linear, and scaling by 2 is exact in binary
  floating point except where a value passes through the subnormal range,
  as at the far edge of the numerical wavefront: to 1e-6 of the peak for the
  acoustic schemes, 2e-5 for elastic, where the free surface carries those
  last-bit differences on to later times), and doubling the data residual
  doubles the gradient;
* checkpointing: the replay interval does not change the gradient (bit for
  bit), including storing every level;
* batching: a shot run inside a batch gives the seismograms it gives alone.

Grids: VD 2000 x 800 (64-shot batched gradient), elastic 1500 x 600,
CD 2D 1000 x 1000, CD 3D 600 x 600 x 400; a few shots each."""
import numpy as np
import pytest

import seismicwaves as sw
from seismicwaves import synthetic as syn

pytestmark = pytest.mark.gpu

F32 = np.float32


def _smooth_model(shape, lo, hi, seed, sigma=20.0):
    rng = np.random.default_rng(seed)
    if len(shape) == 3:
        # a smooth depth profile (a 3D random field would take long to smooth)
        prof = syn.gaussian_smooth(rng.uniform(lo, hi, shape[2]), sigma)
        m = np.broadcast_to(prof, shape)
    else:
        m = syn.gaussian_smooth(rng.uniform(lo, hi, shape), sigma)
    m = lo + (m - m.min()) / (m.max() - m.min()) * (hi - lo)
    return np.ascontiguousarray(m, dtype=F32)


def _double(shots, make):
    """The same shots with the source time function doubled."""
    return [make(sh, 2 * sh.srcs.tf) for sh in shots]


def _assert_doubled(a, b, tol=1e-6):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert np.max(np.abs(a)) > 0
    np.testing.assert_allclose(b, 2 * a, rtol=0, atol=tol * np.max(np.abs(b)))


def _assert_gradient_linear(g1, g2, keys):
    for k in keys:
        a, b = np.asarray(g1[k], np.float64), np.asarray(g2[k], np.float64)
        assert np.max(np.abs(a)) > 0
        err = np.linalg.norm(b - 2 * a) / np.linalg.norm(2 * a)
        print(f"gradient {k}: rel-L2 of grad(2r) - 2 grad(r) = {err:.2e}", flush=True)
        assert err <= 1e-5


# --------------------------------------------------------------------------
# VD 2D, 2000 x 800
# --------------------------------------------------------------------------
def _vd_problem():
    nx, nz, h, dt, f0, halo = 2000, 800, (5.0, 5.0), 5e-4, 12.0, 20
    vp = _smooth_model((nx, nz), 1800.0, 3500.0, 2)
    rho = _smooth_model((nx, nz), 1900.0, 2600.0, 3)
    nts = (3000, 2600)
    tf = syn.ricker_tf(max(nts), dt, f0, 1.2 / f0)[:, None].astype(F32)
    rec = syn.surface_receivers(500, nx, h, halo, 10.0)
    shots = [sw.ScalarShot(sw.ScalarSources(np.array([[x, 10.0]]), tf[:nt], f0),
                           sw.ScalarReceivers(rec, nt, dtype=F32)) for x, nt in zip((2500.0, 7300.0), nts)]
    params = sw.InputParametersAcoustic(max(nts), dt, (nx, nz), h, sw.CPMLBoundaryConditionParameters(halo=halo))
    return params, sw.VpRhoAcousticVDMaterialProperties(vp, rho), shots


def _scalar_shot(sh, tf):
    return sw.ScalarShot(sw.ScalarSources(sh.srcs.positions, tf, sh.srcs.domfreq),
                         sw.ScalarReceivers(sh.recs.positions, sh.recs.seismograms.shape[0], dtype=F32))


def test_vd_full_size_properties():
    print("[full size VD 2000x800]", flush=True)
    params, mp, shots = _vd_problem()
    rp = sw.RunParameters(erroronPPW=False)
    sw.swforward(params, mp, shots, runparams=rp)
    seis = [sh.recs.seismograms.copy() for sh in shots]
    shots2 = _double(shots, _scalar_shot)
    sw.swforward(params, mp, shots2, runparams=rp)
    for a, sh in zip(seis, shots2):
        _assert_doubled(a, sh.recs.seismograms)
    # one shot alone = the same shot inside the batch, bit for bit
    alone = [_scalar_shot(shots[1], shots[1].srcs.tf)]
    sw.swforward(params, mp, alone, runparams=rp)
    assert np.array_equal(alone[0].recs.seismograms, seis[1])
    # gradient: linear in the residual, independent of the replay interval
    obs = [0.9 * s for s in seis]
    obs2 = [2 * o - s for o, s in zip(obs, seis)]
    g = {}
    for name, ob, cf in (("r", obs, 50), ("2r", obs2, 50), ("r_all", obs, 1)):
        g[name] = sw.swgradient(params, mp, shots, [sw.L2Misfit(o.astype(F32)) for o in ob], runparams=rp,
                                gradparams=sw.GradParameters(check_freq=cf))
    _assert_gradient_linear(g["r"], g["2r"], ("vp", "rho"))
    for k in ("vp", "rho"):
        assert np.array_equal(g["r"][k], g["r_all"][k])


# --------------------------------------------------------------------------
# Elastic 2D, 1500 x 600
# --------------------------------------------------------------------------
def _el_problem(kind):
    nx, nz, h, dt, f0, halo, nt = 1500, 600, (5.0, 5.0), 4e-4, 15.0, 20, 3000
    vp = _smooth_model((nx, nz), 2200.0, 4000.0, 4)
    vs = (vp / 1.8).astype(F32)
    rho = _smooth_model((nx, nz), 2000.0, 2500.0, 5)
    mu = (rho * vs ** 2).astype(F32)
    lam = (rho * vp ** 2 - 2 * mu).astype(F32)
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0).astype(F32)
    rec = syn.surface_receivers(300, nx, h, halo, 0.0)
    shots = []
    for x in (2000.0, 5200.0):
        src = np.array([[x, 10.0]])
        if kind == "mt":
            shots.append(sw.MomentTensorShot(
                sw.MomentTensorSources(src, tf[:, None], [sw.MomentTensor2D(1e10, 1e10, 0.0)], f0),
                sw.VectorReceivers(rec, nt, dtype=F32)))
        else:
            stf = np.zeros((nt, 2, 1), dtype=F32)
            stf[:, 1, 0] = 1e10 * tf
            shots.append(sw.ExternalForceShot(sw.ExternalForceSources(src, stf, f0),
                                              sw.VectorReceivers(rec, nt, dtype=F32)))
    params = sw.InputParametersElastic(nt, dt, (nx, nz), h, sw.CPMLBoundaryConditionParameters(halo=halo))
    return params, sw.ElasticIsoMaterialProperties(lam, mu, rho), shots


def _vector_shot(sh, tf):
    nt = sh.recs.seismograms.shape[0]
    recs = sw.VectorReceivers(sh.recs.positions, nt, dtype=F32)
    if isinstance(sh, sw.MomentTensorShot):
        return sw.MomentTensorShot(sw.MomentTensorSources(sh.srcs.positions, tf, sh.srcs.momtens, sh.srcs.domfreq),
                                   recs)
    return sw.ExternalForceShot(sw.ExternalForceSources(sh.srcs.positions, tf, sh.srcs.domfreq), recs)


@pytest.mark.parametrize("kind", ["mt", "ef"])
def test_elastic_full_size_properties(kind):
    print(f"[full size elastic 1500x600 {kind}]", flush=True)
    params, mp, shots = _el_problem(kind)
    rp = sw.RunParameters(erroronPPW=False)
    sw.swforward(params, mp, shots, runparams=rp)
    seis = [sh.recs.seismograms.copy() for sh in shots]
    shots2 = _double(shots, _vector_shot)
    sw.swforward(params, mp, shots2, runparams=rp)
    for a, sh in zip(seis, shots2):
        _assert_doubled(a, sh.recs.seismograms, tol=2e-5)
    obs = [0.9 * s for s in seis]
    obs2 = [2 * o - s for o, s in zip(obs, seis)]
    g = {}
    for name, ob, cf in (("r", obs, 50), ("2r", obs2, 50), ("r_cf", obs, 37)):
        g[name] = sw.swgradient(params, mp, shots, [sw.L2Misfit(o.astype(F32)) for o in ob], runparams=rp,
                                gradparams=sw.GradParameters(check_freq=cf))
    _assert_gradient_linear(g["r"], g["2r"], ("rho", "lambda", "mu"))
    for k in ("rho", "lambda", "mu"):
        assert np.array_equal(g["r"][k], g["r_cf"][k])


# --------------------------------------------------------------------------
# CD 2D, 1000 x 1000, and CD 3D, 600 x 600 x 400
# --------------------------------------------------------------------------
def _cd_problem(n, h, dt, f0, nt, src, rec, halo=20, freetop=True):
    vp = _smooth_model(n, 1500.0, 4000.0, 6, sigma=10.0)
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0)[:, None].astype(F32)
    shots = [sw.ScalarShot(sw.ScalarSources(src, tf, f0), sw.ScalarReceivers(rec, nt, dtype=F32))]
    params = sw.InputParametersAcoustic(nt, dt, n, h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, freeboundtop=freetop))
    return params, sw.VpAcousticCDMaterialProperties(vp), shots


def _cd_properties(params, mp, shots, cfs):
    rp = sw.RunParameters(erroronPPW=False)
    sw.swforward(params, mp, shots, runparams=rp)
    seis = [sh.recs.seismograms.copy() for sh in shots]
    shots2 = _double(shots, _scalar_shot)
    sw.swforward(params, mp, shots2, runparams=rp)
    for a, sh in zip(seis, shots2):
        _assert_doubled(a, sh.recs.seismograms)
    obs = [0.9 * s for s in seis]
    obs2 = [2 * o - s for o, s in zip(obs, seis)]
    g = {}
    for name, ob, cf in (("r", obs, cfs[0]), ("2r", obs2, cfs[0]), ("r_cf", obs, cfs[1])):
        g[name] = sw.swgradient(params, mp, shots, [sw.L2Misfit(o.astype(F32)) for o in ob], runparams=rp,
                                gradparams=sw.GradParameters(check_freq=cf))
    _assert_gradient_linear(g["r"], g["2r"], ("vp",))
    assert np.array_equal(g["r"]["vp"], g["r_cf"]["vp"])


def test_cd2d_full_size_properties():
    print("[full size CD 2D 1000x1000]", flush=True)
    n, h = (1000, 1000), (5.0, 5.0)
    rec = syn.surface_receivers(400, n[0], h, 20, 10.0)
    _cd_properties(*_cd_problem(n, h, 5e-4, 12.0, 4000, np.array([[2500.0, 10.0]]), rec), cfs=(50, 1))


def test_cd3d_full_size_properties():
    print("[full size CD 3D 600x600x400]", flush=True)
    n, h = (600, 600, 400), (10.0, 10.0, 10.0)
    off = (np.arange(20) - 10) * 50.0
    rec = np.array([[1500.0 + a, 2250.0 + b, 240.0] for a in off for b in off])
    _cd_properties(*_cd_problem(n, h, 1e-3, 7.0, 400, np.array([[1500.0, 2250.0, 250.0]]), rec, freetop=False),
                   cfs=(50, 37))
