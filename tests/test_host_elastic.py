"""Host-side logic of the elastic path (CPU only): point spreading, staggered
material averages and their adjoint, checks of the wave simulation."""
import numpy as np
import pytest

import seismicwaves as sw


def test_spread_positions_matches_oracle(oracle_mod):
    rng = np.random.default_rng(0)
    for _ in range(100):
        p = rng.random((2, 2)) * [395.0, 295.0]
        for shift in [(0.0, 0.0), (2.5, 0.0), (0.0, 2.5), (2.5, 2.5)]:
            for mirror in (False, True):
                a = sw.spread_positions(p, (5.0, 5.0), (80, 60), shift, mirror)
                b = oracle_mod.spread_positions(p, (5.0, 5.0), (80, 60), shift, mirror)
                for (ia, ca), (ib, cb) in zip(a, b):
                    np.testing.assert_array_equal(ia, ib)
                    np.testing.assert_allclose(ca, cb, rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("method", [sw.ARITHMETIC, sw.HARMONIC])
def test_staggered_averages_and_adjoint(oracle_mod, method):
    rng = np.random.default_rng(1)
    m = 1.0 + rng.random((9, 7))
    for dims in ([1], [2], [1, 2]):
        np.testing.assert_allclose(sw.interp_stag(m, dims, method), oracle_mod.interp_dims(m, dims, method),
                                   rtol=1e-14)
        g = rng.standard_normal(sw.interp_stag(m, dims, method).shape)
        np.testing.assert_allclose(sw.back_interp(method, m, g, dims), oracle_mod.back_interp_dims(m, g, dims, method),
                                   rtol=1e-13)
        # back_interp is the transpose of the Jacobian of the average
        dm = rng.standard_normal(m.shape) * 1e-7
        lin = np.sum(g * (sw.interp_stag(m + dm, dims, method) - sw.interp_stag(m - dm, dims, method)) / 2)
        assert lin == pytest.approx(np.sum(sw.back_interp(method, m, g, dims) * dm), rel=1e-6)


def _sim(n=(40, 30), vp=3000.0, vs=1700.0, rho=2000.0, dt=1e-3, h=10.0, **run):
    params = sw.InputParametersElastic(100, dt, n, (h, h), sw.CPMLBoundaryConditionParameters(halo=5))
    mu = np.full(n, rho * vs ** 2)
    lam = np.full(n, rho * vp ** 2) - 2 * mu
    mp = sw.ElasticIsoMaterialProperties(lam, mu, np.full(n, rho))
    ws = sw.build_wavesim(params, mp, runparams=sw.RunParameters(**run))
    return ws, mp


def test_model_fields_match_oracle(oracle_mod):
    rng = np.random.default_rng(2)
    ws, mp = _sim()
    mp = sw.ElasticIsoMaterialProperties(*(a * (1 + 0.1 * rng.random(a.shape)) for a in (mp.lam, mp.mu, mp.rho)))
    ws.set_matprop(mp)
    for a, b in zip(ws.model_fields(), oracle_mod.el_model_fields(mp.lam, mp.mu, mp.rho)):
        np.testing.assert_allclose(a, b, rtol=1e-15)


def test_elastic_checks():
    ws, mp = _sim()
    ws.set_matprop(mp)
    with pytest.raises(AssertionError, match="Courant"):
        _sim(dt=3e-3)[0].set_matprop(_sim(dt=3e-3)[1])
    bad = sw.ElasticIsoMaterialProperties(-mp.lam, mp.mu, mp.rho)
    with pytest.raises(AssertionError, match="λ must be positive"):
        ws.set_matprop(bad)
    with pytest.raises(AssertionError, match="Density must be positive"):
        ws.set_matprop(sw.ElasticIsoMaterialProperties(mp.lam, mp.mu, 0 * mp.rho))
    # points per wavelength from the S velocity
    shot = sw.ExternalForceShot(sw.ExternalForceSources(np.array([[200.0, 150.0]]), np.zeros((100, 2, 1)), 20.0),
                                sw.VectorReceivers(np.array([[100.0, 0.0]]), 100))
    with pytest.raises(AssertionError, match="points per wavelength"):
        ws.check_numerics(shot)
    with pytest.raises(ValueError, match="inconsistent"):
        sw.api.check_sim_consistency(ws, mp, [sw.ScalarShot(sw.ScalarSources(np.array([[1.0, 1.0]]),
                                                                             np.zeros((10, 1)), 5.0),
                                                            sw.ScalarReceivers(np.array([[1.0, 1.0]]), 10))])
    with pytest.raises(AssertionError, match="moment tensors"):
        sw.MomentTensorSources(np.array([[1.0, 1.0]]), np.zeros((10, 1)), [], 5.0)
    with pytest.raises(AssertionError, match="components"):
        sw.ExternalForceSources(np.array([[1.0, 1.0]]), np.zeros((10, 3, 1)), 5.0)
