"""2D acoustic constant density on the GPU, through the C ABI, against the
FP64 oracle: seismograms and gradients as relative L2 error, <= 1e-10 for
the FP64 GPU mode and <= 1e-4 for FP32."""
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
    # an all-zero reference (receivers on a pressure-release edge) must be matched exactly
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def make_shots(pos_list, tf, rec_list, f0, nts, dtype):
    shots = []
    for pos, recp, nt in zip(pos_list, rec_list, nts):
        srcs = sw.ScalarSources(np.atleast_2d(pos), tf[:nt].astype(dtype), f0)
        shots.append(sw.ScalarShot(srcs, sw.ScalarReceivers(np.atleast_2d(recp), nt, dtype=dtype)))
    return shots


def oracle_shots(shots):
    return [dict(src_pos=s.srcs.positions.astype(np.float64), srctf=s.srcs.tf.astype(np.float64),
                 f0=float(s.srcs.domfreq), rec_pos=s.recs.positions.astype(np.float64)) for s in shots]


# ---------------------------------------------------------------------------
# Forward: the plain forward configuration (301 x 201, nt = 2000)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forward_gaussian_anomaly(oracle_mod, dtype):
    nx, nz, h, dt, nt, f0 = 301, 201, (10.0, 10.0), 1e-3, 2000, 10.0
    print(f"[cd2d forward 301x201 nt=2000 {np.dtype(dtype).name}]", flush=True)
    vp = syn.gaussian_anomaly(nx, nz, h)
    tf = syn.ricker_tf(nt, dt, f0, 2 / f0)[:, None]
    rec = np.stack([30.0 * np.arange(101), np.full(101, 20.0)], 1)
    params = sw.InputParametersAcoustic(nt, dtype(dt), (nx, nz), h,
                                        sw.CPMLBoundaryConditionParameters(halo=20, rcoef=1e-4, freeboundtop=False))
    shots = make_shots([[1500.0, 20.0]], tf, [rec], f0, [nt], dtype)
    sw.swforward(params, sw.VpAcousticCDMaterialProperties(vp.astype(dtype)), shots, runparams=sw.RunParameters())
    ref = oracle_mod.acoustic_cd_forward(vp, h, dt, nt, 20, 1e-4, False, oracle_shots(shots))[0]
    err = rel_l2(shots[0].recs.seismograms, ref)
    print(f"rel-L2 seismograms = {err:.3e}")
    assert np.max(np.abs(ref)) > 0
    assert err <= TOL[dtype]


# ---------------------------------------------------------------------------
# Gradient: batched shots of uneven length with checkpoint replay
# ---------------------------------------------------------------------------
def _grad_problem(dtype, nts=(420, 500, 333), freetop=True, seed=7, n=(150, 110)):
    nx, nz = n
    h, dt, f0, halo = (10.0, 10.0), 1e-3, 8.0, 12
    rng = np.random.default_rng(seed)
    vtrue = syn.layered(nx, nz, h, rng, vmin=1800, vmax=3000)
    v0 = syn.gaussian_smooth(vtrue, 4.0)
    ntmax = max(nts)
    tf = syn.ricker_tf(ntmax, dt, f0, 1.2 / f0)[:, None]
    xs = rng.uniform(halo * h[0], (nx - 1 - halo) * h[0], len(nts))
    recs = [syn.surface_receivers(37 + 5 * k, nx, h, halo, 30.0) for k in range(len(nts))]
    shots = make_shots([[x, 20.0] for x in xs], tf, recs, f0, nts, dtype)
    params = sw.InputParametersAcoustic(ntmax, dtype(dt), (nx, nz), h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=freetop))
    return dict(params=params, vtrue=vtrue, v0=v0, shots=shots, h=h, dt=dt, halo=halo, freetop=freetop, nts=nts)


def _observed(oracle_mod, p):
    out = []
    for sh, nt in zip(p["shots"], p["nts"]):
        o = oracle_shots([sh])
        out.append(oracle_mod.acoustic_cd_forward(p["vtrue"], p["h"], p["dt"], nt, p["halo"], 1e-4, p["freetop"], o)[0])
    return out


def _oracle_gradient(oracle_mod, p, obs, check_freq):
    tot, mis, seis = None, 0.0, []
    for sh, nt, ob in zip(p["shots"], p["nts"], obs):
        o = oracle_shots([sh])
        o[0]["observed"] = ob
        g, m, s = oracle_mod.acoustic_cd_gradient(p["v0"], p["h"], p["dt"], nt, p["halo"], 1e-4, p["freetop"], o,
                                                  check_freq=min(check_freq, nt - 1))
        tot = g["vp"] if tot is None else tot + g["vp"]
        mis += m
        seis.append(s[0])
    return tot, mis, seis


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gradient_batched_uneven_shots(oracle_mod, dtype):
    print(f"[cd2d gradient batched 3 shots {np.dtype(dtype).name}]", flush=True)
    p = _grad_problem(dtype)
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o.astype(dtype)) for o in obs]
    grad, misfit = sw.swgradient(p["params"], sw.VpAcousticCDMaterialProperties(p["v0"].astype(dtype)), p["shots"], mis,
                                 runparams=sw.RunParameters(),
                                 gradparams=sw.GradParameters(check_freq=50, compute_misfit=True))
    gref, mref, sref = _oracle_gradient(oracle_mod, p, obs, 50)
    for sh, s in zip(p["shots"], sref):
        assert rel_l2(sh.recs.seismograms, s) <= TOL[dtype]
    err = rel_l2(grad["vp"], gref)
    print(f"rel-L2 gradient = {err:.3e}; misfit {misfit} vs {mref}")
    assert np.max(np.abs(gref)) > 0
    assert err <= TOL[dtype]
    assert abs(misfit - mref) <= TOL[dtype] * abs(mref)


def test_gradient_checkpoint_intervals_agree_bitwise(oracle_mod):
    """Store-all (check_freq=1), check_freq=50 and a ragged check_freq=7
    give bit-identical gradients (the replay is deterministic)."""
    p = _grad_problem(np.float32, nts=(300, 257))
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o.astype(np.float32)) for o in obs]
    mp = sw.VpAcousticCDMaterialProperties(p["v0"].astype(np.float32))
    outs = []
    for cf in (1, 50, 7):
        outs.append(sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                                  gradparams=sw.GradParameters(check_freq=cf))["vp"])
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def _deep_points_problem(dtype):
    """Sources and receivers deep inside the grid, so that tiles clear of the
    C-PML strips step through the test-free interior path except on the rows
    that hold points."""
    p = _grad_problem(dtype, nts=(300, 260))
    h = p["h"]
    tf = syn.ricker_tf(300, p["dt"], 8.0, 1.2 / 8.0)[:, None]
    recs = [np.array([[x * h[0] + 3.0, 70 * h[1]] for x in range(20, 131, 7)]),
            np.array([[95 * h[0], z * h[1]] for z in range(20, 91, 5)])]
    p["shots"] = make_shots([[900.0, 500.0], [412.0, 630.0]], tf, recs, 8.0, p["nts"], dtype)
    return p


def test_interior_tiles_and_rows_agree_bitwise(oracle_mod, monkeypatch):
    """1, 2 or 4 rows per thread, with or without the interior-tile path, give
    the same seismograms and gradient bit for bit (FP32), and the FP64
    default matches the oracle."""
    p = _deep_points_problem(np.float32)
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o.astype(np.float32)) for o in obs]
    mp = sw.VpAcousticCDMaterialProperties(p["v0"].astype(np.float32))
    outs = []
    for rows, lean in (("2", "1"), ("2", "0"), ("1", "1"), ("1", "0"), ("4", "1"), ("4", "0")):
        print(f"[cd2d rows per thread {rows}, interior path {lean}]", flush=True)
        monkeypatch.setenv("SWX_CD2D_ROWS", rows)
        monkeypatch.setenv("SWX_CD2D_LEAN", lean)
        g = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=23))["vp"]
        sw.swforward(p["params"], mp, p["shots"], runparams=sw.RunParameters())
        outs.append((g, [sh.recs.seismograms.copy() for sh in p["shots"]]))
    assert np.max(np.abs(outs[0][0])) > 0
    for g, seis in outs[1:]:
        assert np.array_equal(g, outs[0][0])
        assert all(np.array_equal(a, b) for a, b in zip(seis, outs[0][1]))
    monkeypatch.delenv("SWX_CD2D_ROWS")
    monkeypatch.delenv("SWX_CD2D_LEAN")
    p = _deep_points_problem(np.float64)
    gref, mref, _ = _oracle_gradient(oracle_mod, p, obs, 23)
    g = sw.swgradient(p["params"], sw.VpAcousticCDMaterialProperties(p["v0"]), p["shots"],
                      [sw.L2Misfit(o) for o in obs], runparams=sw.RunParameters(),
                      gradparams=sw.GradParameters(check_freq=23))["vp"]
    assert rel_l2(g, gref) <= 1e-10


def test_batched_equals_sequential():
    """One batched launch over three shots = three single-shot runs."""
    p = _grad_problem(np.float32)
    rng = np.random.default_rng(1)
    mis = [sw.L2Misfit(rng.standard_normal(sh.recs.seismograms.shape).astype(np.float32) * 1e-3) for sh in p["shots"]]
    mp = sw.VpAcousticCDMaterialProperties(p["v0"].astype(np.float32))
    gp = sw.GradParameters(check_freq=40)
    gb = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(), gradparams=gp)["vp"]
    seis_b = [sh.recs.seismograms.copy() for sh in p["shots"]]
    gs = np.zeros_like(gb)
    for k, sh in enumerate(p["shots"]):
        gs += sw.swgradient(p["params"], mp, [sh], [mis[k]], runparams=sw.RunParameters(), gradparams=gp)["vp"]
        assert np.array_equal(sh.recs.seismograms, seis_b[k])
    np.testing.assert_allclose(gb, gs, rtol=1e-5, atol=1e-6 * np.max(np.abs(gs)))


def test_gradient_mute_per_shot(oracle_mod):
    p = _grad_problem(np.float64, nts=(300, 280))
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o) for o in obs]
    gp = sw.GradParameters(check_freq=30, mute_radius_src=5, mute_radius_rec=2)
    grad = sw.swgradient(p["params"], sw.VpAcousticCDMaterialProperties(p["v0"]), p["shots"], mis,
                         runparams=sw.RunParameters(), gradparams=gp)["vp"]
    tot = None
    for sh, nt, ob in zip(p["shots"], p["nts"], obs):
        o = oracle_shots([sh])
        o[0]["observed"] = ob
        g, _, _ = oracle_mod.acoustic_cd_gradient(p["v0"], p["h"], p["dt"], nt, p["halo"], 1e-4, p["freetop"], o,
                                                  check_freq=30, mute_radius_src=5, mute_radius_rec=2)
        tot = g["vp"] if tot is None else tot + g["vp"]
    assert rel_l2(grad, tot) <= 1e-10


@pytest.mark.parametrize("case", ["colocated", "edge_receivers", "halo0"])
def test_point_edge_cases(oracle_mod, case):
    """Several sources on one cell (serial injection order), receivers on
    the outermost row (edge values are never updated by the stencil and
    accumulate the adjoint injection) and a grid without C-PML."""
    dtype = np.float64
    nx, nz, h, dt, nt, f0 = 90, 70, (10.0, 10.0), 1e-3, 260, 9.0
    halo = 0 if case == "halo0" else 10
    rng = np.random.default_rng(11)
    vtrue = 2000 + 300 * rng.random((nx, nz))
    v0 = np.full((nx, nz), 2100.0)
    tf = np.stack([syn.ricker_tf(nt, dt, f0, 1.2 / f0), 0.5 * syn.ricker_tf(nt, dt, f0, 1.5 / f0)], 1)
    if case == "colocated":
        src = np.array([[450.0, 300.0], [450.0, 300.0]])
        rec = np.array([[200.0, 400.0], [200.0, 400.0], [600.0, 420.0]])
    elif case == "edge_receivers":
        src = np.array([[450.0, 300.0], [300.0, 200.0]])
        rec = np.stack([np.arange(15, 75, 6) * 10.0, np.zeros(10)], 1)
        rec = np.vstack([rec, [[0.0, 350.0], [890.0, 350.0]]])
    else:
        src = np.array([[450.0, 300.0], [300.0, 200.0]])
        rec = np.stack([np.arange(15, 75, 6) * 10.0, np.full(10, 500.0)], 1)
    shot = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    obs = oracle_mod.acoustic_cd_forward(vtrue, h, dt, nt, halo, 1e-4, True, [shot])[0]
    shot["observed"] = obs
    gref, mref, sref = oracle_mod.acoustic_cd_gradient(v0, h, dt, nt, halo, 1e-4, True, [shot], check_freq=1)
    params = sw.InputParametersAcoustic(nt, dt, (nx, nz), h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=True))
    shots = [sw.ScalarShot(sw.ScalarSources(src, tf, f0), sw.ScalarReceivers(rec, nt))]
    g = sw.swgradient(params, sw.VpAcousticCDMaterialProperties(v0), shots, [sw.L2Misfit(obs)],
                      runparams=sw.RunParameters(), gradparams=sw.GradParameters(check_freq=1))
    assert rel_l2(shots[0].recs.seismograms, sref[0]) <= 1e-10
    assert rel_l2(g["vp"], gref["vp"]) <= 1e-10


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_edge_sources_with_replay(oracle_mod, dtype):
    """Sources and receivers on the outermost rows and columns (cells the
    stencil never updates, which keep level it-2 plus the injection,
    acoustic2D_xPU.jl:122-124) with checkpoint replay: seismograms and
    gradient against the oracle, and the gradient bitwise the same for
    every replay interval (a replay reproduces the forward exactly)."""
    nx, nz, h, dt, nt, f0, halo = 90, 70, (10.0, 10.0), 1e-3, 240, 9.0, 8
    print(f"[cd2d edge sources {np.dtype(dtype).name}]", flush=True)
    rng = np.random.default_rng(21)
    vtrue = 2000 + 300 * rng.random((nx, nz))
    v0 = np.full((nx, nz), 2100.0)
    tf = np.stack([syn.ricker_tf(nt, dt, f0, 1.2 / f0) * a for a in (1.0, 0.7, -0.5, 0.9)], 1)
    src = np.array([[0.0, 0.0], [300.0, 0.0], [890.0, 350.0], [400.0, 690.0]])
    rec = np.array([[0.0, 0.0], [300.0, 0.0], [0.0, 420.0], [620.0, 690.0], [450.0, 300.0], [890.0, 350.0]])
    shot = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    obs = oracle_mod.acoustic_cd_forward(vtrue, h, dt, nt, halo, 1e-4, True, [shot])[0]
    shot["observed"] = obs
    gref, _, sref = oracle_mod.acoustic_cd_gradient(v0, h, dt, nt, halo, 1e-4, True, [shot], check_freq=17)
    params = sw.InputParametersAcoustic(nt, dtype(dt), (nx, nz), h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=True))
    grads = []
    for cf in (1, 17, 50):
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


def test_misfit_api_matches_gradient_misfit(oracle_mod):
    p = _grad_problem(np.float64, nts=(300, 300))
    obs = _observed(oracle_mod, p)
    mis = [sw.L2Misfit(o) for o in obs]
    mp = sw.VpAcousticCDMaterialProperties(p["v0"])
    m1 = sw.swmisfit(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters())
    _, m2 = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(compute_misfit=True, check_freq=20))
    assert np.isclose(m1, m2, rtol=1e-12)


def _edge_problem(dtype):
    nx, nz, h, dt, nt, f0, halo = 90, 70, (10.0, 10.0), 1e-3, 241, 9.0, 8
    rng = np.random.default_rng(23)
    vtrue = 2000 + 300 * rng.random((nx, nz))
    v0 = np.full((nx, nz), 2100.0)
    tf = np.stack([syn.ricker_tf(nt, dt, f0, 1.2 / f0) * a for a in (1.0, 0.7, -0.5)], 1)
    src = np.array([[0.0, 0.0], [890.0, 350.0], [400.0, 690.0]])
    rec = np.array([[0.0, 0.0], [300.0, 0.0], [620.0, 690.0], [450.0, 300.0], [890.0, 350.0]])
    params = sw.InputParametersAcoustic(nt, dtype(dt), (nx, nz), h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=True))
    shots = [sw.ScalarShot(sw.ScalarSources(src, tf.astype(dtype), f0), sw.ScalarReceivers(rec, nt, dtype=dtype))]
    return dict(params=params, vtrue=vtrue, v0=v0, shots=shots)


@pytest.mark.parametrize("case", ["deep_points", "uneven_shots", "edge_points"])
def test_two_levels_per_launch_agree_bitwise(oracle_mod, monkeypatch, case):
    """Two forward / replay levels per launch (cd2d_step2_kernel, interior
    and C-PML/edge/point tiles) give the one-level launches' seismograms,
    snapshots and gradient bit for bit (FP32): with points on interior rows,
    with a batch whose shots end on an odd and an even level, with sources
    and receivers on the outermost cells and a free top; odd replay
    intervals and snapshot spacing.  The FP64 two-level run matches the
    oracle."""
    print(f"[cd2d two levels per launch: {case}]", flush=True)
    if case == "deep_points":
        p = _deep_points_problem(np.float32)
        obs = _observed(oracle_mod, p)
    elif case == "uneven_shots":
        p = _grad_problem(np.float32, nts=(300, 257, 281))
        obs = _observed(oracle_mod, p)
    else:
        p = _edge_problem(np.float32)
        sw.swforward(p["params"], sw.VpAcousticCDMaterialProperties(p["vtrue"].astype(np.float32)), p["shots"],
                     runparams=sw.RunParameters())
        obs = [sh.recs.seismograms.astype(np.float64) for sh in p["shots"]]
    mis = [sw.L2Misfit(o.astype(np.float32)) for o in obs]
    mp = sw.VpAcousticCDMaterialProperties(p["v0"].astype(np.float32))
    outs = []
    for steps, lean, cf in (("1", "1", 23), ("2", "1", 23), ("2", "0", 23), ("2", "1", 1), ("2", "1", 50)):
        print(f"[cd2d levels per launch {steps}, interior path {lean}, check_freq {cf}]", flush=True)
        monkeypatch.setenv("SWX_CD2D_STEPS", steps)
        monkeypatch.setenv("SWX_CD2D_LEAN", lean)
        g = sw.swgradient(p["params"], mp, p["shots"], mis, runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=cf))["vp"]
        snaps = sw.swforward(p["params"], mp, p["shots"], runparams=sw.RunParameters(snapevery=7))
        outs.append((g, [sh.recs.seismograms.copy() for sh in p["shots"]], snaps))
    assert np.max(np.abs(outs[0][0])) > 0
    for g, seis, snaps in outs[1:]:
        assert np.array_equal(g, outs[0][0])
        assert all(np.array_equal(a, b) for a, b in zip(seis, outs[0][1]))
        for s1, s0 in zip(snaps, outs[0][2]):
            assert sorted(s1) == sorted(s0)
            for it in s0:
                assert np.array_equal(s1[it]["pcur"], s0[it]["pcur"]), it
    if case == "deep_points":
        monkeypatch.setenv("SWX_CD2D_STEPS", "2")
        monkeypatch.setenv("SWX_CD2D_LEAN", "1")
        p = _deep_points_problem(np.float64)
        gref, _, sref = _oracle_gradient(oracle_mod, p, obs, 23)
        g = sw.swgradient(p["params"], sw.VpAcousticCDMaterialProperties(p["v0"]), p["shots"],
                          [sw.L2Misfit(o) for o in obs], runparams=sw.RunParameters(),
                          gradparams=sw.GradParameters(check_freq=23))["vp"]
        assert rel_l2(g, gref) <= 1e-10
        for sh, s in zip(p["shots"], sref):
            assert rel_l2(sh.recs.seismograms, s) <= 1e-10


@pytest.mark.parametrize("n,halo,freetop", [((37, 7), 2, True), ((65, 9), 0, False), ((5, 4), 0, True),
                                            ((130, 17), 4, False)])
def test_two_levels_per_launch_small_grids(oracle_mod, monkeypatch, n, halo, freetop):
    """Grids narrower or lower than one tile (partial tiles, rings outside
    the grid, every cell an edge or C-PML cell), FP64: one and two levels
    per launch agree bit for bit and match the oracle."""
    print(f"[cd2d two levels, small grid {n} halo {halo} freetop {freetop}]", flush=True)
    h, dt, nt, f0 = (10.0, 10.0), 1e-3, 61, 9.0
    rng = np.random.default_rng(sum(n) + halo)
    vtrue = 2000 + 300 * rng.random(n)
    v0 = np.full(n, 2100.0)
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0)[:, None]
    src = np.array([[(n[0] // 2) * h[0], (n[1] // 2) * h[1]]])
    rec = np.array([[0.0, 0.0], [(n[0] - 1) * h[0], (n[1] - 1) * h[1]], [(n[0] // 3) * h[0], (n[1] // 2) * h[1]]])
    shot = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    obs = oracle_mod.acoustic_cd_forward(vtrue, h, dt, nt, halo, 1e-4, freetop, [shot])[0]
    gref, _, sref = oracle_mod.acoustic_cd_gradient(v0, h, dt, nt, halo, 1e-4, freetop, [dict(shot, observed=obs)],
                                                    check_freq=3)
    params = sw.InputParametersAcoustic(nt, dt, n, h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=freetop))
    outs = []
    for steps in ("1", "2"):
        monkeypatch.setenv("SWX_CD2D_STEPS", steps)
        shots = [sw.ScalarShot(sw.ScalarSources(src, tf, f0), sw.ScalarReceivers(rec, nt))]
        g = sw.swgradient(params, sw.VpAcousticCDMaterialProperties(v0), shots, [sw.L2Misfit(obs)],
                          runparams=sw.RunParameters(erroronPPW=False), gradparams=sw.GradParameters(check_freq=3))
        outs.append((g["vp"], shots[0].recs.seismograms.copy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert np.max(np.abs(gref["vp"])) > 0
    assert rel_l2(outs[1][0], gref["vp"]) <= 1e-10
    assert rel_l2(outs[1][1], sref[0]) <= 1e-10
