"""3D acoustic constant density on the GPU, through the C ABI, against the
FP64 oracle: seismograms and gradients as relative L2 error, <= 1e-10 for
FP64 and <= 1e-4 for FP32."""
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


def _problem(dtype, nts=(230, 260), n=(70, 52, 45), freetop=True, seed=9, halo=8):
    h, dt, f0 = (10.0, 10.0, 10.0), 1e-3, 8.0
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    vtrue = 2000 + 400 * rng.random(n)
    vtrue = syn.gaussian_smooth(vtrue, 2.0)
    v0 = np.full(n, float(vtrue.mean()))
    v0[1, 1, 1] = float(vtrue.max())
    ntmax = max(nts)
    tf = syn.ricker_tf(ntmax, dt, f0, 1.2 / f0)[:, None]
    shots, osh = [], []
    for k, nt in enumerate(nts):
        src = np.array([[(20 + 15 * k) * h[0], 25 * h[1], 12 * h[2]]])
        rec = np.array([[x * h[0], y * h[1], 30.0] for x in range(12, 60, 6) for y in (15, 26, 37)])
        shots.append(sw.ScalarShot(sw.ScalarSources(src, tf[:nt].astype(dtype), f0),
                                   sw.ScalarReceivers(rec, nt, dtype=dtype)))
        osh.append(dict(src_pos=src, srctf=tf[:nt], f0=f0, rec_pos=rec, nt=nt))
    params = sw.InputParametersAcoustic(ntmax, dt, n, h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=freetop))
    return dict(params=params, vtrue=vtrue, v0=v0, shots=shots, osh=osh, h=h, dt=dt, halo=halo, freetop=freetop,
                nts=nts)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forward_batched(oracle_mod, dtype):
    print(f"[cd3d forward {np.dtype(dtype).name}]", flush=True)
    p = _problem(dtype)
    sw.swforward(p["params"], sw.VpAcousticCDMaterialProperties(p["vtrue"].astype(dtype)), p["shots"],
                 runparams=sw.RunParameters())
    ref = oracle_mod.acoustic_cd3d_forward(p["vtrue"], p["h"], p["dt"], max(p["nts"]), p["halo"], 1e-4,
                                           p["freetop"], p["osh"])
    for sh, r in zip(p["shots"], ref):
        err = rel_l2(sh.recs.seismograms, r)
        print(f"rel-L2 seismograms = {err:.3e}")
        assert np.max(np.abs(r)) > 0
        assert err <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gradient_batched(oracle_mod, dtype):
    print(f"[cd3d gradient {np.dtype(dtype).name}]", flush=True)
    p = _problem(dtype)
    obs = oracle_mod.acoustic_cd3d_forward(p["vtrue"], p["h"], p["dt"], max(p["nts"]), p["halo"], 1e-4,
                                           p["freetop"], p["osh"])
    mis = [sw.L2Misfit(o.astype(dtype)) for o in obs]
    grad, misfit = sw.swgradient(p["params"], sw.VpAcousticCDMaterialProperties(p["v0"].astype(dtype)), p["shots"],
                                 mis, runparams=sw.RunParameters(),
                                 gradparams=sw.GradParameters(check_freq=40, compute_misfit=True))
    osh = [dict(o, observed=ob) for o, ob in zip(p["osh"], obs)]
    gref, mref, _ = oracle_mod.acoustic_cd3d_gradient(p["v0"], p["h"], p["dt"], max(p["nts"]), p["halo"], 1e-4,
                                                      p["freetop"], osh, check_freq=40)
    err = rel_l2(grad["vp"], gref["vp"])
    print(f"rel-L2 gradient = {err:.3e}")
    assert np.max(np.abs(gref["vp"])) > 0
    assert err <= TOL[dtype]
    assert abs(misfit - mref) <= TOL[dtype] * abs(mref)


def test_zchunk_and_checkpoints_agree_bitwise(oracle_mod, monkeypatch):
    """The z-chunking of the marching kernel and the checkpoint interval do
    not change results."""
    p = _problem(np.float32, nts=(200,))
    rng = np.random.default_rng(0)
    mis = [sw.L2Misfit(rng.standard_normal(sh.recs.seismograms.shape).astype(np.float32) * 1e-3) for sh in p["shots"]]
    mp = sw.VpAcousticCDMaterialProperties(p["v0"].astype(np.float32))
    outs = []
    for chunk, cf in (("32", 1), ("7", 50), ("45", 13)):
        monkeypatch.setenv("SWX_CD3D_ZCHUNK", chunk)
        outs.append(sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                                  gradparams=sw.GradParameters(check_freq=cf))["vp"])
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_point_edge_cases(oracle_mod):
    """Colocated sources and receivers on the outermost planes."""
    n, h, dt, nt, f0, halo = (40, 36, 30), (10.0, 10.0, 10.0), 1e-3, 180, 8.0, 6
    rng = np.random.default_rng(3)
    vtrue = 2000 + 300 * rng.random(n)
    v0 = np.full(n, 2100.0)
    tf = np.stack([syn.ricker_tf(nt, dt, f0, 1.2 / f0), 0.5 * syn.ricker_tf(nt, dt, f0, 1.4 / f0)], 1)
    src = np.array([[200.0, 180.0, 150.0], [200.0, 180.0, 150.0]])
    rec = np.array([[100.0, 100.0, 0.0], [100.0, 100.0, 0.0], [0.0, 200.0, 100.0], [390.0, 0.0, 290.0],
                    [250.0, 150.0, 120.0]])
    osh = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    obs = oracle_mod.acoustic_cd3d_forward(vtrue, h, dt, nt, halo, 1e-4, True, [osh])[0]
    osh["observed"] = obs
    gref, _, sref = oracle_mod.acoustic_cd3d_gradient(v0, h, dt, nt, halo, 1e-4, True, [osh], check_freq=1)
    params = sw.InputParametersAcoustic(nt, dt, n, h, sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4))
    shots = [sw.ScalarShot(sw.ScalarSources(src, tf, f0), sw.ScalarReceivers(rec, nt))]
    g = sw.swgradient(params, sw.VpAcousticCDMaterialProperties(v0), shots, [sw.L2Misfit(obs)],
                      runparams=sw.RunParameters(), gradparams=sw.GradParameters(check_freq=1))
    assert rel_l2(shots[0].recs.seismograms, sref[0]) <= 1e-10
    assert rel_l2(g["vp"], gref["vp"]) <= 1e-10


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_edge_sources_with_replay(oracle_mod, dtype):
    """Sources and receivers on the outer faces, edges and a corner (cells
    that keep level it-2 plus the injection, acoustic3D_xPU.jl:155-157) with
    checkpoint replay, against the oracle; the gradient is bitwise the same
    for every replay interval."""
    n, h, dt, nt, f0, halo = (40, 36, 30), (10.0, 10.0, 10.0), 1e-3, 160, 8.0, 6
    print(f"[cd3d edge sources {np.dtype(dtype).name}]", flush=True)
    rng = np.random.default_rng(23)
    vtrue = 2000 + 300 * rng.random(n)
    v0 = np.full(n, 2100.0)
    tf = np.stack([syn.ricker_tf(nt, dt, f0, 1.2 / f0) * a for a in (1.0, -0.6, 0.8)], 1)
    src = np.array([[0.0, 0.0, 0.0], [200.0, 0.0, 150.0], [390.0, 180.0, 290.0]])
    rec = np.array([[0.0, 0.0, 0.0], [200.0, 0.0, 150.0], [100.0, 350.0, 120.0], [250.0, 150.0, 120.0],
                    [390.0, 180.0, 290.0]])
    osh = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    obs = oracle_mod.acoustic_cd3d_forward(vtrue, h, dt, nt, halo, 1e-4, False, [osh])[0]
    osh["observed"] = obs
    gref, _, sref = oracle_mod.acoustic_cd3d_gradient(v0, h, dt, nt, halo, 1e-4, False, [osh], check_freq=13)
    params = sw.InputParametersAcoustic(nt, dtype(dt), n, h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=False))
    grads = []
    for cf in (1, 13, 40):
        shots = [sw.ScalarShot(sw.ScalarSources(src, tf.astype(dtype), f0), sw.ScalarReceivers(rec, nt, dtype=dtype))]
        g = sw.swgradient(params, sw.VpAcousticCDMaterialProperties(v0.astype(dtype)), shots,
                          [sw.L2Misfit(obs.astype(dtype))], runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=cf))
        assert rel_l2(shots[0].recs.seismograms, sref[0]) <= TOL[dtype]
        if dtype == np.float64:
            assert rel_l2(g["vp"], gref["vp"]) <= TOL[dtype]
        else:
            # An edge source cell holds the running sum of every other
            # injection, and its correlation term is a second difference of
            # those sums: in FP32 that difference is rounding noise (it would
            # be in the reference's FP32 run too).  FP32 is compared on the
            # other cells; the FP64 run above covers these cells at 1e-10.
            keep = np.ones(gref["vp"].shape, bool)
            keep[tuple((oracle_mod.find_nearest_grid_points(src, h) - 1).T)] = False
            assert rel_l2(g["vp"][keep], gref["vp"][keep]) <= TOL[dtype]
        grads.append(g["vp"])
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def _interior_problem(dtype, nt=150):
    """A grid wide enough for 64 x 8 column tiles clear of the x/y C-PML strips
    (the test-free interior path), with points on some of their rows, so
    that in the z chunks that hold them the same tiles take the general path."""
    n, h, dt, f0, halo = (150, 44, 36), (10.0, 10.0, 10.0), 1e-3, 8.0, 8
    rng = np.random.default_rng(21)
    vtrue = syn.gaussian_smooth(2000 + 400 * rng.random(n), 2.0)
    v0 = np.full(n, float(vtrue.mean()))
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0)[:, None]
    src = np.array([[900.0, 200.0, 120.0]])
    rec = np.array([[x * h[0], 280.0, 30.0] for x in range(70, 131, 5)])
    shots = [sw.ScalarShot(sw.ScalarSources(src, tf.astype(dtype), f0), sw.ScalarReceivers(rec, nt, dtype=dtype))]
    osh = [dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec, nt=nt)]
    params = sw.InputParametersAcoustic(nt, dt, n, h, sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4))
    return dict(params=params, vtrue=vtrue, v0=v0, shots=shots, osh=osh, h=h, dt=dt, halo=halo, nt=nt)


@pytest.mark.parametrize("chunk,variant", [("24", None), ("7", None), ("7", "204"), ("7", "202")])
def test_interior_tiles_against_oracle(oracle_mod, monkeypatch, chunk, variant):
    """FP64 seismograms and gradient on a grid whose interior tiles take the
    test-free path; with 7 planes per block some chunks lie wholly inside the
    z-interior and take the straight-line plane update (launch shapes 204 and
    202: staged through LDS)."""
    print(f"[cd3d interior tiles float64, z chunk {chunk}, launch shape {variant}]", flush=True)
    monkeypatch.setenv("SWX_CD3D_ZCHUNK", chunk)
    if variant:
        monkeypatch.setenv("SWX_CD3D_VARIANT", variant)
    p = _interior_problem(np.float64)
    obs = oracle_mod.acoustic_cd3d_forward(p["vtrue"], p["h"], p["dt"], p["nt"], p["halo"], 1e-4, True, p["osh"])
    sw.swforward(p["params"], sw.VpAcousticCDMaterialProperties(p["vtrue"]), p["shots"], runparams=sw.RunParameters())
    assert np.max(np.abs(obs[0])) > 0
    assert rel_l2(p["shots"][0].recs.seismograms, obs[0]) <= 1e-10
    osh = [dict(o, observed=ob) for o, ob in zip(p["osh"], obs)]
    gref, _, _ = oracle_mod.acoustic_cd3d_gradient(p["v0"], p["h"], p["dt"], p["nt"], p["halo"], 1e-4, True, osh,
                                                   check_freq=30)
    g = sw.swgradient(p["params"], sw.VpAcousticCDMaterialProperties(p["v0"]), p["shots"],
                      [sw.L2Misfit(o) for o in obs], runparams=sw.RunParameters(),
                      gradparams=sw.GradParameters(check_freq=30))
    err = rel_l2(g["vp"], gref["vp"])
    print(f"rel-L2 gradient = {err:.3e}")
    assert np.max(np.abs(gref["vp"])) > 0
    assert err <= 1e-10


@pytest.mark.parametrize("chunk", ["32", "7"])
def test_interior_path_agrees_bitwise(monkeypatch, chunk):
    """The test-free interior path gives the general path's results bit for
    bit (seismograms and gradient, FP32), with 64 x 8, 64 x 4 and 64 x 2 tiles, in one
    launch or two, and with the LDS-staged march."""
    p = _interior_problem(np.float32)
    rng = np.random.default_rng(1)
    mis = [sw.L2Misfit(rng.standard_normal(sh.recs.seismograms.shape).astype(np.float32) * 1e-3) for sh in p["shots"]]
    mp = sw.VpAcousticCDMaterialProperties(p["vtrue"].astype(np.float32))
    monkeypatch.setenv("SWX_CD3D_ZCHUNK", chunk)
    outs = []
    for variant, lean in (("8", "1"), ("8", "0"), ("4", "1"), ("4", "0"), ("2", "1"), ("108", "1"), ("104", "1"),
                          ("102", "1"), ("208", "1"), ("204", "1"), ("202", "1")):
        print(f"[cd3d z chunk {chunk}, launch shape {variant}, interior path {lean}]", flush=True)
        monkeypatch.setenv("SWX_CD3D_VARIANT", variant)
        monkeypatch.setenv("SWX_CD3D_LEAN", lean)
        g = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=20))["vp"]
        sw.swforward(p["params"], mp, p["shots"], runparams=sw.RunParameters())
        outs.append((g, p["shots"][0].recs.seismograms.copy()))
    assert np.max(np.abs(outs[0][0])) > 0
    for g, seis in outs[1:]:
        assert np.array_equal(outs[0][0], g)
        assert np.array_equal(outs[0][1], seis)
