"""Host-side logic of the product package (no GPU needed): C-PML profiles,
grid-point lookup, staggered interpolation, muting, misfits and checks,
compared with the oracle's independent restatement and the reference's own
test values."""
import numpy as np
import pytest

import seismicwaves as sw
from seismicwaves import api
from seismicwaves.cpml import init_bdc


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("halo", [0, 1, 5, 20])
@pytest.mark.parametrize("freetop", [False, True])
def test_cpml_matches_oracle(oracle_mod, dtype, halo, freetop):
    vp = np.full((60, 50), 2500.0).astype(dtype)
    a = init_bdc(float(vp.max()), 1e-3, (5.0, 7.5), halo, 1e-4, freetop, 11.0, dtype)
    b = oracle_mod.init_bdc(vp, 1e-3, (5.0, 7.5), halo, 1e-4, freetop, 11.0, dtype)
    for ax in range(2):
        for k in ("a", "a_h", "b", "b_h"):
            assert a[ax][k].dtype == dtype
            np.testing.assert_allclose(a[ax][k], b[ax][k], rtol=1e-6 if dtype == np.float32 else 1e-13, atol=0)


def test_nearest_grid_points_round_ties_up(oracle_mod):
    pos = np.array([[0.0, 0.0], [2.5, 7.4999], [5.0, 12.5], [97.5, 2.49]])
    got = sw.find_nearest_grid_points(pos, (5.0, 5.0))
    np.testing.assert_array_equal(got, [[1, 1], [2, 2], [2, 4], [21, 1]])
    np.testing.assert_array_equal(got, oracle_mod.find_nearest_grid_points(pos, (5.0, 5.0)))


def test_interpolations_reference_values():
    """test_interpolations.jl:5-60."""
    m = np.array([1.0, 2.0, 3.0, 4.0])
    assert np.array_equal(api.interp_stag(m, 1, sw.ARITHMETIC), [1.5, 2.5, 3.5])
    np.testing.assert_allclose(api.interp_stag(m, 1, sw.HARMONIC), [1.333, 2.4, 3.428], atol=1e-3)
    g = np.array([0.1, 0.2, 0.3])
    exp = np.zeros(4)
    exp[:-1] += g * 0.5
    exp[1:] += g * 0.5
    np.testing.assert_allclose(api.back_interp(sw.ARITHMETIC, m, g, 1), exp)
    gg = 1 / m[:-1] + 1 / m[1:]
    exp = np.zeros(4)
    exp[:-1] += g * (-2 / gg ** 2) * (-1 / m[:-1] ** 2)
    exp[1:] += g * (-2 / gg ** 2) * (-1 / m[1:] ** 2)
    np.testing.assert_allclose(api.back_interp(sw.HARMONIC, m, g, 1), exp)


def test_interpolations_2d_and_adjointness():
    rng = np.random.default_rng(3)
    m = 1 + rng.random((7, 5))
    for dim in (1, 2):
        for meth in (sw.ARITHMETIC, sw.HARMONIC):
            st = api.interp_stag(m, dim, meth)
            assert st.shape == tuple(s - (1 if k == dim - 1 else 0) for k, s in enumerate(m.shape))
            # back_interp is the transpose of the Jacobian of interp_stag
            g = rng.random(st.shape)
            dm = 1e-6 * rng.random(m.shape)
            lin = np.sum(g * (api.interp_stag(m + dm, dim, meth) - api.interp_stag(m - dm, dim, meth)) / 2)
            adj = np.sum(api.back_interp(meth, m, g, dim) * dm)
            assert np.isclose(lin, adj, rtol=1e-6)


def test_mute_matches_oracle(oracle_mod):
    rng = np.random.default_rng(1)
    g = rng.random((40, 30))
    pts = np.array([[50.0, 20.0], [0.0, 145.0]])
    a = sw.mute_around_points(g.copy(), pts, (5.0, 5.0), 4)
    b = oracle_mod.mute_around_points(g.copy(), pts, (5.0, 5.0), 4)
    np.testing.assert_allclose(a, b, rtol=1e-15)
    assert np.array_equal(sw.mute_around_points(g.copy(), pts, (5.0, 5.0), 0), g)


def test_l2_misfit_and_adjoint_source(oracle_mod):
    rng = np.random.default_rng(2)
    syn, obs = rng.random((50, 3)), rng.random((50, 3))
    rec = sw.ScalarReceivers(np.zeros((3, 2)), 50)
    rec.seismograms[:] = syn
    m = sw.L2Misfit(obs)
    r = syn - obs
    assert np.isclose(m.calcmisfit(rec), 0.5 * np.sum(r * r))
    np.testing.assert_allclose(m.dchi_du(rec), r)
    C = np.diag(rng.random(50))
    mc = sw.L2Misfit(obs, invcov=C, windows=[(5, 20)])
    assert np.isclose(mc.calcmisfit(rec), oracle_mod.l2_misfit(syn, obs, C, [(5, 20)]))
    np.testing.assert_allclose(mc.dchi_du(rec), oracle_mod.l2_dchi_du(syn, obs, C, [(5, 20)]))
    md = sw.L2Misfit(obs, invcov=np.diag(C), windows=[(5, 20)])
    np.testing.assert_allclose(md.dchi_du(rec), mc.dchi_du(rec))
    with pytest.raises(AssertionError):
        sw.L2Misfit(obs, invcov=np.eye(49))


def _params(n=(60, 50), halo=10, freetop=True, nt=100, dt=1e-3, h=(10.0, 10.0)):
    return sw.InputParametersAcoustic(nt, dt, n, h, sw.CPMLBoundaryConditionParameters(halo=halo, freeboundtop=freetop))


def test_wavesim_checks_mirror_reference():
    rp = sw.RunParameters()
    mp = sw.VpAcousticCDMaterialProperties(np.full((60, 50), 2000.0))
    with pytest.raises(AssertionError):            # n >= 2halo+3 on C-PML axes
        sw.build_wavesim(_params(n=(20, 50)), mp, runparams=rp)
    with pytest.raises(ValueError):                # unknown backend symbol
        sw.build_wavesim(_params(), mp, runparams=sw.RunParameters(parall="threads"))
    ws = sw.build_wavesim(_params(), mp, runparams=rp)
    with pytest.raises(AssertionError):            # CFL
        ws.check_matprop(sw.VpAcousticCDMaterialProperties(np.full((60, 50), 9000.0)))
    with pytest.raises(AssertionError):            # positive vp
        ws.check_matprop(sw.VpAcousticCDMaterialProperties(np.full((60, 50), -1.0)))
    with pytest.raises(AssertionError):            # shape
        ws.check_matprop(sw.VpAcousticCDMaterialProperties(np.full((60, 49), 2000.0)))
    ws.check_matprop(mp)
    with pytest.raises(AssertionError):            # outside grid
        ws.check_positions(np.array([[-1.0, 10.0]]))
    with pytest.raises(AssertionError):
        sw.build_wavesim(_params(), mp, runparams=rp, gradparams=sw.GradParameters(), gradient=False)
    with pytest.raises(AssertionError):
        sw.GradParameters(check_freq=0)


def test_source_scaling_matches_oracle(oracle_mod):
    rng = np.random.default_rng(5)
    vp = 2000 + 100 * rng.random((60, 50))
    ws = sw.build_wavesim(_params(), sw.VpAcousticCDMaterialProperties(vp), runparams=sw.RunParameters())
    ws.matprop = sw.VpAcousticCDMaterialProperties(vp)
    tf = rng.random((100, 2))
    si = np.array([[10, 12], [30, 40]])
    np.testing.assert_allclose(ws.scale_srctf(tf.copy(), si),
                               oracle_mod.scale_srctf_cd(tf, (10.0, 10.0), 1e-3, vp, si), rtol=1e-14)
