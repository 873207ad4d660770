"""The L2 adjoint source and misfit formed on the device (swx_set_observed +
swx_adjoint_l2) against the host path (L2Misfit.dchi_du / calcmisfit
passed to swx_adjoint): bitwise equal gradients, misfits equal to the
rounding of a different summation order (rel 1e-13), for every physics,
with and without windows and a diagonal inverse covariance."""
import numpy as np
import pytest

import seismicwaves as sw
from seismicwaves import synthetic as syn

pytestmark = pytest.mark.gpu

H, DT, F0 = 10.0, 1e-3, 8.0


def _pair(params, mp_true, mp0, make_shots, weighted, keys, dtype):
    obs_shots = make_shots()
    sw.swforward(params, mp_true, obs_shots, runparams=sw.RunParameters(erroronPPW=False))
    misfits = []
    for k, sh in enumerate(obs_shots):
        o = sh.recs.seismograms.astype(dtype)
        nt = o.shape[0]
        if weighted:
            invcov = np.linspace(0.5, 2.0, nt) * (1.0 + 0.3 * k)
            misfits.append(sw.L2Misfit(o, invcov=invcov, windows=[(5, nt // 3), (nt // 2, nt - 7)]))
        else:
            misfits.append(sw.L2Misfit(o))
    out = []
    for device in (True, False):
        ws = sw.build_wavesim(params, mp0, runparams=sw.RunParameters(erroronPPW=False),
                              gradparams=sw.GradParameters(check_freq=17, compute_misfit=True), gradient=True)
        ws.device_misfit = device
        out.append(sw.swgradient(ws, mp0, make_shots(), misfits))
    (gd, md), (gh, mh) = out
    for k in keys:
        assert np.max(np.abs(gh[k])) > 0
        assert np.array_equal(gd[k], gh[k]), k
    print(f"misfit device {md:.15e} host {mh:.15e}", flush=True)
    assert abs(md - mh) <= 1e-13 * abs(mh)


def _scalar_shots(n, nts, dtype):
    tf = syn.ricker_tf(max(nts), DT, F0, 1.2 / F0)

    def make():
        out = []
        for k, nt in enumerate(nts):
            src = np.array([[(0.3 + 0.2 * k) * (n[0] - 1) * H] + [0.4 * (m - 1) * H for m in n[1:]]])
            rec = np.array([[x] + [0.2 * (m - 1) * H for m in n[1:]] for x in np.linspace(100.0, (n[0] - 11) * H, 9)])
            out.append(sw.ScalarShot(sw.ScalarSources(src, tf[:nt, None].astype(dtype), F0),
                                     sw.ScalarReceivers(rec, nt, dtype=dtype)))
        return out
    return make


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("weighted", [False, True])
def test_cd2d_device_l2(dtype, weighted):
    n = (130, 100)
    rng = np.random.default_rng(1)
    vt = syn.layered(*n, (H, H), rng, nlayers=4, vmin=1800.0, vmax=2600.0).astype(dtype)
    v0 = syn.gaussian_smooth(vt, 4.0).astype(dtype)
    params = sw.InputParametersAcoustic(220, DT, n, (H, H), sw.CPMLBoundaryConditionParameters(halo=10))
    _pair(params, sw.VpAcousticCDMaterialProperties(vt), sw.VpAcousticCDMaterialProperties(v0),
          _scalar_shots(n, [220, 180], dtype), weighted, ("vp",), dtype)


@pytest.mark.parametrize("weighted", [False, True])
def test_vd2d_device_l2(weighted):
    dtype = np.float32
    n = (140, 90)
    rng = np.random.default_rng(2)
    vt = syn.layered(*n, (H, H), rng, nlayers=4, vmin=1800.0, vmax=2600.0).astype(dtype)
    rt = syn.gardner(vt, rng).astype(dtype)
    mp0 = sw.VpRhoAcousticVDMaterialProperties(syn.gaussian_smooth(vt, 4.0).astype(dtype),
                                               syn.gaussian_smooth(rt, 4.0).astype(dtype))
    params = sw.InputParametersAcoustic(200, DT, n, (H, H), sw.CPMLBoundaryConditionParameters(halo=8))
    _pair(params, sw.VpRhoAcousticVDMaterialProperties(vt, rt), mp0, _scalar_shots(n, [200, 150, 170], dtype),
          weighted, ("vp", "rho"), dtype)


def test_cd3d_device_l2():
    dtype = np.float32
    n = (60, 50, 40)
    vt = np.full(n, 2000.0, dtype=dtype)
    vt[:, :, 20:] = 2400.0
    v0 = np.full(n, 2100.0, dtype=dtype)
    params = sw.InputParametersAcoustic(120, DT, n, (H, H, H),
                                        sw.CPMLBoundaryConditionParameters(halo=5, freeboundtop=False))
    _pair(params, sw.VpAcousticCDMaterialProperties(vt), sw.VpAcousticCDMaterialProperties(v0),
          _scalar_shots(n, [120, 100], dtype), True, ("vp",), dtype)


@pytest.mark.parametrize("weighted", [False, True])
def test_elastic_device_l2(weighted):
    dtype = np.float32
    n = (120, 80)
    rho = np.full(n, 2200.0, dtype=dtype)
    vp = np.full(n, 2800.0, dtype=dtype)
    vp[:, 40:] = 3200.0
    vs = (vp / 1.8).astype(dtype)
    mu = (rho * vs ** 2).astype(dtype)
    lam = (rho * vp ** 2 - 2 * mu).astype(dtype)
    mp_true = sw.ElasticIsoMaterialProperties(lam, mu, rho)
    mp0 = sw.ElasticIsoMaterialProperties(*(syn.gaussian_smooth(a, 6.0).astype(dtype) for a in (lam, mu, rho)))
    nt = 200
    tf = syn.ricker_tf(nt, DT, F0, 1.2 / F0).astype(dtype)
    rec = np.array([[x, 0.0] for x in np.linspace(120.0, 1050.0, 7)])

    def make():
        return [sw.MomentTensorShot(sw.MomentTensorSources(np.array([[x, 200.0]]), tf[:, None],
                                                           [sw.MomentTensor2D(1e3, 1e3, 0.0)], F0),
                                    sw.VectorReceivers(rec, nt, dtype=dtype)) for x in (400.0, 700.0)]
    params = sw.InputParametersElastic(nt, DT, n, (H, H), sw.CPMLBoundaryConditionParameters(halo=8))
    _pair(params, mp_true, mp0, make, weighted, ("rho", "lambda", "mu"), dtype)
