"""2D acoustic variable density (staggered) on the GPU, through the C ABI,
against the FP64 oracle: seismograms and vp/rho gradients as relative L2
error, <= 1e-10 for the FP64 GPU mode and <= 1e-4 for FP32."""
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


def _problem(dtype, nts=(420, 500, 333), freetop=True, seed=3, n=(150, 110), method=sw.ARITHMETIC):
    nx, nz = n
    h, dt, f0, halo = (10.0, 10.0), 1e-3, 8.0, 12
    rng = np.random.default_rng(seed)
    vtrue = syn.add_ellipses(syn.layered(nx, nz, h, rng, vmin=1800, vmax=2800), h, rng, n=10, amin=30, amax=120)
    rtrue = syn.gardner(vtrue, rng)
    v0 = syn.gaussian_smooth(vtrue, 4.0)
    r0 = syn.gaussian_smooth(rtrue, 4.0)
    ntmax = max(nts)
    tf = syn.ricker_tf(ntmax, dt, f0, 1.2 / f0)[:, None]
    xs = rng.uniform(halo * h[0], (nx - 1 - halo) * h[0], len(nts))
    shots, oshots = [], []
    for k, (x, nt) in enumerate(zip(xs, nts)):
        rec = syn.surface_receivers(31 + 7 * k, nx, h, halo, 30.0)
        src = np.array([[x, 20.0]])
        shots.append(sw.ScalarShot(sw.ScalarSources(src, tf[:nt].astype(dtype), f0),
                                   sw.ScalarReceivers(rec, nt, dtype=dtype)))
        oshots.append(dict(src_pos=src, srctf=tf[:nt], f0=f0, rec_pos=rec, nt=nt))
    params = sw.InputParametersAcoustic(ntmax, dt, (nx, nz), h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=freetop))
    return dict(params=params, vtrue=vtrue, rtrue=rtrue, v0=v0, r0=r0, shots=shots, oshots=oshots, h=h, dt=dt,
                halo=halo, freetop=freetop, nts=nts, method=method)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forward_batched(oracle_mod, dtype):
    print(f"[vd2d forward batched {np.dtype(dtype).name}]", flush=True)
    p = _problem(dtype)
    mp = sw.VpRhoAcousticVDMaterialProperties(p["vtrue"].astype(dtype), p["rtrue"].astype(dtype))
    sw.swforward(p["params"], mp, p["shots"], runparams=sw.RunParameters())
    ref = oracle_mod.acoustic_vd_forward(p["vtrue"], p["rtrue"], p["h"], p["dt"], max(p["nts"]), p["halo"], 1e-4,
                                         p["freetop"], p["oshots"])
    for sh, r in zip(p["shots"], ref):
        err = rel_l2(sh.recs.seismograms, r)
        print(f"rel-L2 seismograms = {err:.3e}")
        assert np.max(np.abs(r)) > 0
        assert err <= TOL[dtype]


def _oracle_grad(oracle_mod, p, obs, cf, **kw):
    osh = []
    for o, ob in zip(p["oshots"], obs):
        d = dict(o)
        d["observed"] = ob
        osh.append(d)
    return oracle_mod.acoustic_vd_gradient(p["v0"], p["r0"], p["h"], p["dt"], max(p["nts"]), p["halo"], 1e-4,
                                           p["freetop"], osh, check_freq=cf, method=p["method"], **kw)


def _observed(oracle_mod, p):
    return oracle_mod.acoustic_vd_forward(p["vtrue"], p["rtrue"], p["h"], p["dt"], max(p["nts"]), p["halo"], 1e-4,
                                          p["freetop"], p["oshots"], method=p["method"])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gradient_batched_uneven_shots(oracle_mod, dtype):
    print(f"[vd2d gradient batched {np.dtype(dtype).name}]", flush=True)
    p = _problem(dtype)
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o.astype(dtype)) for o in obs]
    mp = sw.VpRhoAcousticVDMaterialProperties(p["v0"].astype(dtype), p["r0"].astype(dtype))
    grad, misfit = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                                 gradparams=sw.GradParameters(check_freq=50, compute_misfit=True))
    gref, mref, sref = _oracle_grad(oracle_mod, p, obs, 50)
    for sh, s in zip(p["shots"], sref):
        assert rel_l2(sh.recs.seismograms, s) <= TOL[dtype]
    for key in ("vp", "rho"):
        err = rel_l2(grad[key], gref[key])
        print(f"rel-L2 gradient {key} = {err:.3e}")
        assert np.max(np.abs(gref[key])) > 0
        assert err <= TOL[dtype]
    assert abs(misfit - mref) <= TOL[dtype] * abs(mref)


def test_gradient_harmonic_and_mute(oracle_mod):
    p = _problem(np.float64, nts=(300, 260), method=sw.HARMONIC, freetop=False)
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o) for o in obs]
    mp = sw.VpRhoAcousticVDMaterialProperties(p["v0"], p["r0"], interp_method=sw.HARMONIC)
    gp = sw.GradParameters(check_freq=37, mute_radius_src=4, mute_radius_rec=2)
    grad = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(), gradparams=gp)
    gref, _, _ = _oracle_grad(oracle_mod, p, obs, 37, mute_radius_src=4, mute_radius_rec=2)
    for key in ("vp", "rho"):
        assert rel_l2(grad[key], gref[key]) <= 1e-10


def test_checkpoint_intervals_agree_bitwise(oracle_mod):
    p = _problem(np.float32, nts=(300, 257))
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o.astype(np.float32)) for o in obs]
    mp = sw.VpRhoAcousticVDMaterialProperties(p["v0"].astype(np.float32), p["r0"].astype(np.float32))
    outs = [sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=cf)) for cf in (1, 50, 9)]
    for o in outs[1:]:
        assert np.array_equal(o["vp"], outs[0]["vp"]) and np.array_equal(o["rho"], outs[0]["rho"])


def test_kernel_variants_agree_bitwise(oracle_mod, monkeypatch):
    """Every forward and adjoint kernel variant the autotuner may pick (2-D
    tiles of height 8/16/32 as interior + other-tile launches or as one
    launch, two steps per launch on tiles of height 8/16/32, z-marching
    strips of 128/256 columns, two shots per block on tiles of height 8/16)
    gives the same seismograms
    and gradients bit for bit, and matches the oracle.  The uneven shots
    (one ends on an odd step) and the odd replay length (39) make the
    two-step launches end with single steps; check_freq = 1 stores every
    level."""
    p = _problem(np.float64, nts=(300, 257), n=(301, 150))
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o) for o in obs]
    mp = sw.VpRhoAcousticVDMaterialProperties(p["v0"], p["r0"])
    outs = []
    for v, w, cf in (("8", "8", 40), ("16", "16", 40), ("32", "102", 40), ("102", "104", 40), ("104", "8", 40),
                     ("208", "16", 40), ("216", "8", 40), ("232", "16", 40), ("216", "16", 1), ("208", "8", 7),
                     ("316", "316", 40), ("308", "308", 9), ("416", "16", 40), ("408", "316", 7),
                     ("332", "16", 40)):
        print(f"[vd2d kernel variants forward {v} adjoint {w} check_freq {cf}]", flush=True)
        monkeypatch.setenv("SWX_VD2D_TILE_FWD", v)
        monkeypatch.setenv("SWX_VD2D_TILE_ADJ", w)
        g = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=cf))
        outs.append((g, [sh.recs.seismograms.copy() for sh in p["shots"]]))
    for g, seis in outs[1:]:
        assert np.array_equal(g["vp"], outs[0][0]["vp"]) and np.array_equal(g["rho"], outs[0][0]["rho"])
        assert all(np.array_equal(a, b) for a, b in zip(seis, outs[0][1]))
    gref = _oracle_grad(oracle_mod, p, obs, 40)[0]
    assert rel_l2(outs[-1][0]["vp"], gref["vp"]) <= 1e-10


@pytest.mark.parametrize("case", ["colocated", "edge_points", "halo0"])
def test_point_edge_cases(oracle_mod, case):
    dtype = np.float64
    nx, nz, h, dt, nt, f0 = 90, 70, (10.0, 10.0), 1e-3, 260, 9.0
    halo = 0 if case == "halo0" else 10
    rng = np.random.default_rng(12)
    vtrue = 2000 + 300 * rng.random((nx, nz))
    rtrue = 1800 + 400 * rng.random((nx, nz))
    v0, r0 = np.full((nx, nz), 2100.0), np.full((nx, nz), 2000.0)
    tf = np.stack([syn.ricker_tf(nt, dt, f0, 1.2 / f0), 0.5 * syn.ricker_tf(nt, dt, f0, 1.5 / f0)], 1)
    if case == "colocated":
        src = np.array([[450.0, 300.0], [450.0, 300.0]])
        rec = np.array([[200.0, 400.0], [200.0, 400.0], [600.0, 420.0]])
    elif case == "edge_points":
        src = np.array([[450.0, 0.0], [0.0, 200.0]])
        rec = np.stack([np.arange(15, 75, 6) * 10.0, np.zeros(10)], 1)
        rec = np.vstack([rec, [[0.0, 350.0], [890.0, 350.0], [400.0, 690.0]]])
    else:
        src = np.array([[450.0, 300.0], [300.0, 200.0]])
        rec = np.stack([np.arange(15, 75, 6) * 10.0, np.full(10, 500.0)], 1)
    osh = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    obs = oracle_mod.acoustic_vd_forward(vtrue, rtrue, h, dt, nt, halo, 1e-4, True, [osh])[0]
    osh["observed"] = obs
    gref, _, sref = oracle_mod.acoustic_vd_gradient(v0, r0, h, dt, nt, halo, 1e-4, True, [osh], check_freq=30)
    params = sw.InputParametersAcoustic(nt, dt, (nx, nz), h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=True))
    shots = [sw.ScalarShot(sw.ScalarSources(src, tf, f0), sw.ScalarReceivers(rec, nt))]
    g = sw.swgradient(params, sw.VpRhoAcousticVDMaterialProperties(v0, r0), shots, [sw.L2Misfit(obs)],
                      runparams=sw.RunParameters(), gradparams=sw.GradParameters(check_freq=30))
    assert rel_l2(shots[0].recs.seismograms, sref[0]) <= 1e-10
    for key in ("vp", "rho"):
        assert rel_l2(g[key], gref[key]) <= 1e-10


def test_pair_kernel_odd_shot_count(monkeypatch):
    """Two shots per block with an odd number of shots (the last block has
    one) and uneven lengths: bitwise the seismograms and gradients of the
    one-shot-per-block kernel."""
    p = _problem(np.float64, nts=(300, 257, 281), n=(301, 150))
    sw.swforward(p["params"], sw.VpRhoAcousticVDMaterialProperties(p["vtrue"], p["rtrue"]), p["shots"],
                 runparams=sw.RunParameters())
    mis = [sw.L2Misfit(sh.recs.seismograms.copy()) for sh in p["shots"]]
    mp = sw.VpRhoAcousticVDMaterialProperties(p["v0"], p["r0"])
    outs = []
    for v in ("316", "416"):
        monkeypatch.setenv("SWX_VD2D_TILE_FWD", v)
        g = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=40))
        outs.append((g, [sh.recs.seismograms.copy() for sh in p["shots"]]))
    assert np.array_equal(outs[0][0]["vp"], outs[1][0]["vp"])
    assert np.array_equal(outs[0][0]["rho"], outs[1][0]["rho"])
    assert all(np.array_equal(a, b) for a, b in zip(outs[0][1], outs[1][1]))


def test_shot_group_orders_agree_bitwise(monkeypatch):
    """The block orders that take the shots in groups of 2, 4 and 8 (tile
    codes + 1000, 2000, 3000) give the plain order's seismograms and
    gradients bit for bit, with 8 shots of uneven length (some end inside a
    group while the others go on), in one launch or two per step; the
    autotuner's pick (no variant pinned) agrees too."""
    nts = (300, 257, 281, 240, 299, 263, 270, 251)
    p = _problem(np.float32, nts=nts, n=(301, 150))
    sw.swforward(p["params"], sw.VpRhoAcousticVDMaterialProperties(p["vtrue"].astype(np.float32),
                                                                    p["rtrue"].astype(np.float32)),
                 p["shots"], runparams=sw.RunParameters())
    mis = [sw.L2Misfit(sh.recs.seismograms.copy()) for sh in p["shots"]]
    mp = sw.VpRhoAcousticVDMaterialProperties(p["v0"].astype(np.float32), p["r0"].astype(np.float32))
    outs = []
    for v, w in (("316", "316"), ("1316", "2316"), ("2316", "3316"), ("3316", "1316"), ("1016", "2016"),
                 ("2008", "1316"), ("3332", "16"), (None, None)):
        print(f"[vd2d shot groups forward {v} adjoint {w}]", flush=True)
        for env, val in (("SWX_VD2D_TILE_FWD", v), ("SWX_VD2D_TILE_ADJ", w)):
            if val is None:
                monkeypatch.delenv(env, raising=False)
            else:
                monkeypatch.setenv(env, val)
        g = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=40))
        outs.append((g, [sh.recs.seismograms.copy() for sh in p["shots"]]))
    assert np.max(np.abs(outs[0][0]["vp"])) > 0
    for g, seis in outs[1:]:
        assert np.array_equal(g["vp"], outs[0][0]["vp"]) and np.array_equal(g["rho"], outs[0][0]["rho"])
        assert all(np.array_equal(a, b) for a, b in zip(seis, outs[0][1]))
