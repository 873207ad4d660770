"""2D isotropic elastic on the GPU, through the C ABI, against the FP64
oracle: seismograms and gradients ("rho", "lambda", "mu") as relative L2
error, <= 1e-10 for FP64 and <= 1e-4 for FP32."""
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


def _media(n, seed):
    rng = np.random.default_rng(seed)
    rho = syn.gaussian_smooth(2000 + 300 * rng.random(n), 2.0)
    vp = syn.gaussian_smooth(2600 + 500 * rng.random(n), 2.0)
    vs = vp / np.sqrt(3) * syn.gaussian_smooth(0.9 + 0.1 * rng.random(n), 2.0)
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    return lam, mu, rho


def _problem(dtype, kind="mt", nts=(300, 260), n=(90, 70), freetop=True, seed=4, halo=10, h=10.0):
    dt, f0 = 1e-3, 6.0
    lam, mu, rho = _media(n, seed)
    lam0, mu0, rho0 = [np.full(n, float(a.mean())) for a in (lam, mu, rho)]
    ntmax = max(nts)
    tf = syn.ricker_tf(ntmax, dt, f0, 1.2 / f0)
    shots, osh = [], []
    for k, nt in enumerate(nts):
        src = np.array([[(30 + 17 * k) * h + 3.0, (8 + 3 * k) * h + 1.5]])
        rec = np.array([[x * h + 2.0, 2 * h] for x in range(12, 78, 6)] + [[45 * h, 0.0], [20 * h + 4.0, 50 * h]])
        if kind == "mt":
            mom = [sw.MomentTensor2D(1.0e3, 0.6e3, 0.4e3)]
            srcs = sw.MomentTensorSources(src, tf[:nt, None].astype(dtype), mom, f0)
            sh = sw.MomentTensorShot(srcs, sw.VectorReceivers(rec, nt, dtype=dtype))
            o = dict(src_pos=src, srctf=tf[:nt, None] * 1.0, moments=np.array([[1.0e3, 0.6e3, 0.4e3]]), f0=f0,
                     rec_pos=rec, nt=nt)
        else:
            stf = np.zeros((nt, 2, 1))
            stf[:, 0, 0] = 1e3 * tf[:nt]
            stf[:, 1, 0] = -7e2 * tf[:nt]
            srcs = sw.ExternalForceSources(src, stf.astype(dtype), f0)
            sh = sw.ExternalForceShot(srcs, sw.VectorReceivers(rec, nt, dtype=dtype))
            o = dict(src_pos=src, srctf=stf, f0=f0, rec_pos=rec, nt=nt)
        shots.append(sh)
        osh.append(o)
    params = sw.InputParametersElastic(ntmax, dt, n, (h, h),
                                       sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=freetop))
    return dict(params=params, true=(lam, mu, rho), init=(lam0, mu0, rho0), shots=shots, osh=osh, h=h, dt=dt,
                halo=halo, freetop=freetop, nts=nts)


def _matprop(m, dtype):
    return sw.ElasticIsoMaterialProperties(*[a.astype(dtype) for a in m])


def _oracle_args(p):
    return ((p["h"], p["h"]), p["dt"], max(p["nts"]), p["halo"], 1e-4, p["freetop"])


@pytest.mark.parametrize("dtype,kind,freetop", [(np.float64, "mt", True), (np.float32, "mt", True),
                                                (np.float64, "ef", False), (np.float32, "ef", True)])
def test_forward_batched(oracle_mod, dtype, kind, freetop):
    print(f"[el2d forward {np.dtype(dtype).name} {kind} freetop={freetop}]", flush=True)
    p = _problem(dtype, kind=kind, freetop=freetop)
    sw.swforward(p["params"], _matprop(p["true"], dtype), p["shots"], runparams=sw.RunParameters())
    ref = oracle_mod.elastic_forward(*p["true"], *_oracle_args(p), p["osh"])
    for sh, r in zip(p["shots"], ref):
        err = rel_l2(sh.recs.seismograms, r)
        print(f"rel-L2 seismograms = {err:.3e}")
        assert np.max(np.abs(r)) > 0
        assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype,kind,freetop", [(np.float64, "mt", True), (np.float32, "mt", True),
                                                (np.float64, "ef", False), (np.float32, "ef", True)])
def test_gradient_batched(oracle_mod, dtype, kind, freetop):
    print(f"[el2d gradient {np.dtype(dtype).name} {kind} freetop={freetop}]", flush=True)
    p = _problem(dtype, kind=kind, freetop=freetop)
    obs = oracle_mod.elastic_forward(*p["true"], *_oracle_args(p), p["osh"])
    mis = [sw.L2Misfit(o.astype(dtype)) for o in obs]
    grad, misfit = sw.swgradient(p["params"], _matprop(p["init"], dtype), p["shots"], mis,
                                 runparams=sw.RunParameters(),
                                 gradparams=sw.GradParameters(check_freq=37, compute_misfit=True))
    osh = [dict(o, observed=ob) for o, ob in zip(p["osh"], obs)]
    gref, mref, _ = oracle_mod.elastic_gradient(*p["init"], *_oracle_args(p), osh, check_freq=37)
    for key in ("rho", "lambda", "mu"):
        err = rel_l2(grad[key], gref[key])
        print(f"rel-L2 gradient {key} = {err:.3e}")
        assert np.max(np.abs(gref[key])) > 0
        assert err <= TOL[dtype]
    assert abs(misfit - mref) <= TOL[dtype] * abs(mref)


def test_checkpoint_interval_bitwise_and_mute(oracle_mod):
    """The checkpoint interval does not change results; per-shot muting goes
    through the per-shot gradient path and matches the oracle."""
    p = _problem(np.float64, kind="mt", nts=(240,))
    obs = oracle_mod.elastic_forward(*p["true"], *_oracle_args(p), p["osh"])
    mis = [sw.L2Misfit(o) for o in obs]
    mp = _matprop(p["init"], np.float64)
    outs = [sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=cf)) for cf in (1, 13, 100)]
    for g in outs[1:]:
        for k in ("rho", "lambda", "mu"):
            assert np.array_equal(g[k], outs[0][k])
    gm = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                       gradparams=sw.GradParameters(check_freq=50, mute_radius_src=4, mute_radius_rec=2))
    osh = [dict(o, observed=ob) for o, ob in zip(p["osh"], obs)]
    gref, _, _ = oracle_mod.elastic_gradient(*p["init"], *_oracle_args(p), osh, check_freq=50, mute_radius_src=4,
                                             mute_radius_rec=2)
    for k in ("rho", "lambda", "mu"):
        assert rel_l2(gm[k], gref[k]) <= 1e-10


@pytest.mark.parametrize("fwd_variant", ["1", "2"])
def test_point_edge_cases(oracle_mod, monkeypatch, fwd_variant):
    """Sources next to the grid edges (folded and mirrored taps, taps on
    never-updated stress cells), colocated receivers, receivers on the free
    surface and on the last rows, nearest-point (non-sinc) mode; with the
    two-launch steps and with the fused tile steps (stresses in LDS, edge
    stress cells accumulating in HBM)."""
    monkeypatch.setenv("SWX_EL2D_VARIANT_FWD", fwd_variant)
    monkeypatch.setenv("SWX_EL2D_VARIANT_ADJ", fwd_variant)
    n, h, dt, nt, f0, halo = (64, 50), (10.0, 10.0), 1e-3, 220, 10.0, 6
    lam, mu, rho = _media(n, 11)
    lam0, mu0, rho0 = [np.full(n, float(a.mean())) for a in (lam, mu, rho)]
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0)
    mom = np.array([[1e3, 2e3, 0.5e3], [-1e3, 1e3, 0.0], [0.3e3, 0.3e3, 1e3]])
    for sinc in (True, False):
        # the nearest-point mode puts every family on the same index, so the
        # far corner is outside the staggered grids there
        far = [630.0, 490.0] if sinc else [600.0, 460.0]
        src = np.array([[12.0, 0.0], [317.0, 251.0], far])
        rec = np.array([[100.0, 0.0], [100.0, 0.0], [0.0, 200.0], far, [255.0, 123.0], [5.0, 3.0]])
        osh = dict(src_pos=src, srctf=np.repeat(tf[:, None], 3, axis=1), moments=mom, f0=f0, rec_pos=rec)
        args = (h, dt, nt, halo, 1e-4, True)
        obs = oracle_mod.elastic_forward(lam, mu, rho, *args, [osh], sincinterp=sinc)[0]
        osh["observed"] = obs
        gref, _, sref = oracle_mod.elastic_gradient(lam0, mu0, rho0, *args, [osh], check_freq=30, sincinterp=sinc)
        params = sw.InputParametersElastic(nt, dt, n, h, sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4))
        shot = sw.MomentTensorShot(
            sw.MomentTensorSources(src, np.repeat(tf[:, None], 3, axis=1), [sw.MomentTensor2D(*m) for m in mom], f0),
            sw.VectorReceivers(rec, nt))
        ws = sw.build_wavesim(params, sw.ElasticIsoMaterialProperties(lam0, mu0, rho0),
                              runparams=sw.RunParameters(erroronPPW=False), gradparams=sw.GradParameters(check_freq=30),
                              gradient=True, sincinterp=sinc)
        g = sw.run_swgradient(ws, sw.ElasticIsoMaterialProperties(lam0, mu0, rho0), [shot], [sw.L2Misfit(obs)])
        assert rel_l2(shot.recs.seismograms, sref[0]) <= 1e-10
        for k in ("rho", "lambda", "mu"):
            assert rel_l2(g[k], gref[k]) <= 1e-10, (sinc, k)


@pytest.mark.parametrize("kind,freetop", [("mt", True), ("ef", False)])
def test_kernel_variants_agree_bitwise(oracle_mod, monkeypatch, kind, freetop):
    """Every forward and adjoint kernel variant the autotuner may pick (the
    stress/displacement step as interior + other-block launches or as one
    launch each; with the XCD-aware tile map, with the C-PML/edge tiles
    first or with the round-robin block order; z-marching strips of 128/256
    columns; the fused tile step with the stresses in LDS) gives the
    same seismograms and gradients bit for bit, on a grid
    that needs several strips and a partial last one, and matches the
    oracle."""
    p = _problem(np.float64, kind=kind, nts=(200, 170), n=(301, 96), freetop=freetop)
    obs = oracle_mod.elastic_forward(*p["true"], *_oracle_args(p), p["osh"])
    mis = [sw.L2Misfit(o) for o in obs]
    mp = _matprop(p["init"], np.float64)
    outs = []
    for v, w, xcd, order in (("0", "0", "1", "1"), ("1", "1", "1", "1"), ("1", "1", "1", "0"), ("0", "0", "1", "0"),
                             ("1", "1", "0", "0"), ("0", "0", "0", "0"), ("102", "104", "1", "1"),
                             ("104", "102", "1", "1"), ("0", "104", "1", "1"), ("1", "0", "1", "1"),
                             ("2", "1", "1", "1"), ("2", "2", "0", "0"), ("1", "2", "1", "1")):
        print(f"[el2d kernel variants {kind} forward {v} adjoint {w} xcd map {xcd} frame first {order}]", flush=True)
        monkeypatch.setenv("SWX_EL2D_VARIANT_FWD", v)
        monkeypatch.setenv("SWX_EL2D_VARIANT_ADJ", w)
        monkeypatch.setenv("SWX_EL2D_XCD", xcd)
        monkeypatch.setenv("SWX_EL2D_ORDER", order)
        g = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=23))
        outs.append((g, [sh.recs.seismograms.copy() for sh in p["shots"]]))
    for g, seis in outs[1:]:
        for k in ("rho", "lambda", "mu"):
            assert np.array_equal(g[k], outs[0][0][k]), k
        assert all(np.array_equal(a, b) for a, b in zip(seis, outs[0][1]))
    osh = [dict(o, observed=ob) for o, ob in zip(p["osh"], obs)]
    gref, _, _ = oracle_mod.elastic_gradient(*p["init"], *_oracle_args(p), osh, check_freq=23)
    for k in ("rho", "lambda", "mu"):
        assert rel_l2(outs[-1][0][k], gref[k]) <= 1e-10, k
