"""1D acoustic, constant and variable density, on the GPU through the C ABI
against the FP64 oracle: seismograms and gradients as relative L2 error,
<= 1e-10 for the FP64 GPU mode and <= 1e-4 for FP32.  Batches of shots of
uneven length with their own dominant frequency (C-PML per shot), sources
and receivers on the outermost cells too, every replay interval bitwise
equal, and snapshots (snapevery)."""
import numpy as np
import pytest

import seismicwaves as sw
from seismicwaves import synthetic as syn

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-10, np.float32: 1e-4}


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


@pytest.mark.parametrize("vd", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forward_analytic_configs(oracle_mod, vd, dtype):
    """The reference's 1D analytic setups (test_analytical_vs_numerical_
    acoustic_*_density.jl, halo 20): GPU seismograms against the oracle."""
    c0, rho0, nx, dx, f0, halo = 1000.0, 1500.0, 501, 2.5, 5.0, 20
    nt = 5000 if vd else 1000
    dt = 0.99 * dx / c0 * (6 / 7 if vd else 1.0)
    print(f"[1d {'vd' if vd else 'cd'} analytic config {np.dtype(dtype).name}]", flush=True)
    lx = (nx - 1) * dx
    t = dt * np.arange(nt)
    tf = (oracle_mod.gaussderivstf(t, 2 / f0, f0) if vd else oracle_mod.rickerstf(t, 2 / f0, f0))[:, None]
    src, rec = np.array([[lx / 2]]), np.array([[lx / 3], [lx / 5]])
    params = sw.InputParametersAcoustic(nt, dtype(dt), (nx,), (dx,),
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=False))
    shots = [sw.ScalarShot(sw.ScalarSources(src, tf.astype(dtype), f0), sw.ScalarReceivers(rec, nt, dtype=dtype))]
    osh = [dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)]
    if vd:
        mp = sw.VpRhoAcousticVDMaterialProperties(np.full(nx, c0, dtype), np.full(nx, rho0, dtype))
        ref = oracle_mod.acoustic_vd1d_forward(np.full(nx, c0), np.full(nx, rho0), (dx,), dt, nt, halo, 1e-4, False,
                                               osh)[0]
    else:
        mp = sw.VpAcousticCDMaterialProperties(np.full(nx, c0, dtype))
        ref = oracle_mod.acoustic_cd1d_forward(np.full(nx, c0), (dx,), dt, nt, halo, 1e-4, False, osh)[0]
    sw.swforward(params, mp, shots, runparams=sw.RunParameters())
    err = rel_l2(shots[0].recs.seismograms, ref)
    print(f"rel-L2 seismograms = {err:.3e}")
    assert np.max(np.abs(ref)) > 0
    assert err <= TOL[dtype]


def _batch(rng, nx, h, nts, edge):
    """Shots of uneven length, each with its own dominant frequency; with
    edge, the first shot's source and some receivers sit on the outermost
    cells."""
    tf = syn.ricker_tf(max(nts), 1e-3, 10.0, 0.12)
    shots = []
    for k, nt in enumerate(nts):
        xs = 0.0 if (edge and k == 0) else (0.3 + 0.2 * k) * (nx - 1) * h
        rec = [[x * h] for x in rng.integers(1, nx - 1, 4)] + ([[0.0], [(nx - 1) * h]] if edge else [])
        shots.append(dict(src_pos=np.array([[xs], [xs + 7 * h]]), srctf=np.stack([tf[:nt], -0.5 * tf[:nt]], 1),
                          f0=10.0 * (1 + 0.4 * k), rec_pos=np.array(rec, dtype=float), nt=nt))
    return shots


def _api_shots(osh, dtype):
    return [sw.ScalarShot(sw.ScalarSources(o["src_pos"], o["srctf"].astype(dtype), o["f0"]),
                          sw.ScalarReceivers(o["rec_pos"], o["nt"], dtype=dtype)) for o in osh]


@pytest.mark.parametrize("vd", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("freetop", [False, True])
def test_gradient_batched_uneven_shots(oracle_mod, vd, dtype, freetop):
    nx, h, dt, halo = 300, 10.0, 1e-3, 15
    nts = [700, 513, 640]
    print(f"[1d {'vd' if vd else 'cd'} gradient {np.dtype(dtype).name} freetop {freetop}]", flush=True)
    rng = np.random.default_rng(31 + vd)
    vt = syn.gaussian_smooth(rng.uniform(1900.0, 2600.0, nx), 3.0)
    rt = syn.gaussian_smooth(rng.uniform(1800.0, 2500.0, nx), 3.0)
    v0, r0 = syn.gaussian_smooth(vt, 8.0), syn.gaussian_smooth(rt, 8.0)
    osh = _batch(rng, nx, h, nts, edge=True)
    args = ((h,), dt, max(nts), halo, 1e-4, freetop)
    if vd:
        obs = oracle_mod.acoustic_vd1d_forward(vt, rt, *args, osh)
        gref, mref, sref = oracle_mod.acoustic_vd1d_gradient(v0, r0, *args, [dict(o, observed=b) for o, b in
                                                                            zip(osh, obs)], check_freq=37)
        mp = sw.VpRhoAcousticVDMaterialProperties(v0.astype(dtype), r0.astype(dtype))
    else:
        obs = oracle_mod.acoustic_cd1d_forward(vt, *args, osh)
        gref, mref, sref = oracle_mod.acoustic_cd1d_gradient(v0, *args, [dict(o, observed=b) for o, b in
                                                                        zip(osh, obs)], check_freq=37)
        mp = sw.VpAcousticCDMaterialProperties(v0.astype(dtype))
    params = sw.InputParametersAcoustic(max(nts), dtype(dt), (nx,), (h,),
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=freetop))
    grads = []
    for cf in (1, 37, 300):
        shots = _api_shots(osh, dtype)
        g, mis = sw.swgradient(params, mp, shots, [sw.L2Misfit(b.astype(dtype)) for b in obs],
                               runparams=sw.RunParameters(erroronPPW=False),
                               gradparams=sw.GradParameters(check_freq=cf, compute_misfit=True))
        for sh, ref in zip(shots, sref):
            assert rel_l2(sh.recs.seismograms, ref) <= TOL[dtype]
        assert abs(mis - mref) <= TOL[dtype] * abs(mref)
        for key in g:
            assert np.max(np.abs(gref[key])) > 0
            if dtype == np.float64:
                assert rel_l2(g[key], gref[key]) <= TOL[dtype], key
            else:
                # FP32: the edge source cell (x = 0) holds a running sum of
                # the injections, and its correlation term is a difference
                # of those sums: rounding noise in FP32.  The FP64 run checks
                # that cell.
                keep = np.ones(nx, bool)
                keep[0] = False
                assert rel_l2(g[key][keep], gref[key][keep]) <= TOL[dtype], key
        grads.append(g)
    for g in grads[1:]:
        for key in g:
            assert np.array_equal(g[key], grads[0][key]), key


@pytest.mark.parametrize("vd", [False, True])
def test_snapshots(oracle_mod, vd):
    nx, h, dt, halo, every = 250, 10.0, 1e-3, 10, 40
    nts = [210, 165]
    rng = np.random.default_rng(41)
    vp = syn.gaussian_smooth(rng.uniform(1900.0, 2600.0, nx), 3.0)
    rho = syn.gaussian_smooth(rng.uniform(1800.0, 2500.0, nx), 3.0)
    osh = _batch(rng, nx, h, nts, edge=True)
    params = sw.InputParametersAcoustic(max(nts), dt, (nx,), (h,),
                                        sw.CPMLBoundaryConditionParameters(halo=halo, freeboundtop=False))
    mp = sw.VpRhoAcousticVDMaterialProperties(vp, rho) if vd else sw.VpAcousticCDMaterialProperties(vp)
    shots = _api_shots(osh, np.float64)
    snaps = sw.swforward(params, mp, shots, runparams=sw.RunParameters(snapevery=every, erroronPPW=False))
    names = ["pcur", "vcur"] if vd else ["pcur"]
    shapes = [(nx,), (nx - 1,)] if vd else [(nx,)]
    for got, o in zip(snaps, osh):
        if vd:
            run = lambda o=o: oracle_mod.acoustic_vd1d_forward(vp, rho, (h,), dt, o["nt"], halo, 1e-4, False, [o])
        else:
            run = lambda o=o: oracle_mod.acoustic_cd1d_forward(vp, (h,), dt, o["nt"], halo, 1e-4, False, [o])
        ref = oracle_mod.with_snapshots(every, o["nt"], shapes, np.float64, run)
        assert sorted(got) == sorted(ref)
        for it, fields in ref.items():
            k = 0
            for name in names:
                g = got[it][name]
                for arr in (g if isinstance(g, list) else [g]):
                    assert arr.shape == fields[k].shape
                    assert rel_l2(arr, fields[k]) <= 1e-10, (name, it)
                    k += 1
