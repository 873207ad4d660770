"""Randomised parity sweep on the GPU: seeded random configurations of every
physics (grid sizes around the width of a few interior tiles, C-PML halo,
free surface on or off, several shots of uneven length, sources and
receivers anywhere in the grid, replay interval), FP64 seismograms and
gradients against the FP64 oracle at relative L2 <= 1e-10.  The shots of a
batch carry different dominant frequencies, so each has its own C-PML
coefficients (alpha = pi f0 (1 - x), acou_init_bc.jl:19-30).

The sizes are chosen so that each grid has tiles that take the test-free
interior paths as well as C-PML, edge and point tiles; the positions put
some points inside interior tiles.  Acoustic sources and receivers are put
on the outermost rows and columns half of the time (cells the stencil does
not update: the CD time step keeps level it-2 there, acoustic2D_xPU.jl:
122-124)."""
import numpy as np
import pytest

import seismicwaves as sw
from seismicwaves import synthetic as syn

pytestmark = pytest.mark.gpu

H, DT, F0 = 10.0, 1e-3, 8.0
TOL = 1e-10


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def _field(rng, shape, lo, hi, sigma=2.0):
    m = syn.gaussian_smooth(rng.uniform(lo, hi, shape), sigma)
    return lo + (m - m.min()) / max(m.max() - m.min(), 1e-12) * (hi - lo)


def _points(rng, k, n, margin, edge=0.0):
    """k positions (m) with every coordinate at least `margin` cells inside;
    with probability `edge` a point has one coordinate on the grid's first or
    last row/column instead."""
    pts = np.stack([rng.uniform(margin * H, (ni - 1 - margin) * H, k) for ni in n], 1)
    for q in range(k):
        if rng.random() < edge:
            ax = int(rng.integers(len(n)))
            pts[q, ax] = 0.0 if rng.random() < 0.5 else (n[ax] - 1) * H
    return pts


def _config(rng, nlo, nhi, halos):
    n = tuple(int(rng.integers(lo, hi)) for lo, hi in zip(nlo, nhi))
    halo = int(rng.choice(halos))
    freetop = bool(rng.random() < 0.5)
    nsh = int(rng.integers(1, 4))
    nts = [int(x) for x in rng.integers(100, 180, nsh)]
    cf = int(rng.choice([1, 7, 23, 50]))
    return n, halo, freetop, nts, cf


def _domfreq(k):
    """Dominant frequency of shot k of a batch (C-PML alpha per shot)."""
    return F0 * (1.0 + 0.45 * k)


def _scalar_shots(rng, n, nts, dtype):
    ntmax = max(nts)
    tf = syn.ricker_tf(ntmax, DT, F0, 1.2 / F0)
    shots, osh = [], []
    for k, nt in enumerate(nts):
        nsrc = int(rng.integers(1, 3))
        src = _points(rng, nsrc, n, 0.0, edge=0.5)
        stf = np.stack([tf[:nt] * (1.0 + 0.3 * k) for k in range(nsrc)], 1)
        # one receiver on the first source's cell: every shot records a wave
        # (a short shot's wave may not reach any other receiver)
        rec = np.vstack([_points(rng, int(rng.integers(3, 12)), n, 0.0, edge=0.5), src[:1]])
        shots.append(sw.ScalarShot(sw.ScalarSources(src, stf.astype(dtype), _domfreq(k)),
                                   sw.ScalarReceivers(rec, nt, dtype=dtype)))
        osh.append(dict(src_pos=src, srctf=stf, f0=_domfreq(k), rec_pos=rec, nt=nt))
    return shots, osh


def _acoustic_params(n, halo, freetop, nts):
    return sw.InputParametersAcoustic(max(nts), DT, n, (H,) * len(n),
                                      sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=freetop))


def _forward(params, mp, shots):
    """GPU seismograms of the shots on the model mp."""
    sw.swforward(params, mp, shots, runparams=sw.RunParameters(erroronPPW=False))
    return [sh.recs.seismograms.copy() for sh in shots]


@pytest.mark.parametrize("steps", ["1", "2"])
@pytest.mark.parametrize("seed", range(8))
def test_cd2d_random(oracle_mod, monkeypatch, seed, steps):
    """With one and with two levels per launch (SWX_CD2D_STEPS)."""
    monkeypatch.setenv("SWX_CD2D_STEPS", steps)
    rng = np.random.default_rng(100 + seed)
    n, halo, freetop, nts, cf = _config(rng, (130, 40), (220, 140), (0, 4, 9, 13))
    print(f"[random cd2d seed {seed}, {steps} level(s) per launch: n {n} halo {halo} freetop {freetop} nts {nts} "
          f"check_freq {cf}]", flush=True)
    vt = _field(rng, n, 1800.0, 2600.0)
    v0 = syn.gaussian_smooth(vt, 4.0)
    shots, osh = _scalar_shots(rng, n, nts, np.float64)
    params = _acoustic_params(n, halo, freetop, nts)
    seis = _forward(params, sw.VpAcousticCDMaterialProperties(vt), shots)
    obs, gref = [], None
    for o, nt in zip(osh, nts):
        ob = oracle_mod.acoustic_cd_forward(vt, (H, H), DT, nt, halo, 1e-4, freetop, [o])[0]
        obs.append(ob)
        g, _, _ = oracle_mod.acoustic_cd_gradient(v0, (H, H), DT, nt, halo, 1e-4, freetop, [dict(o, observed=ob)],
                                                  check_freq=min(cf, nt - 1))
        gref = g["vp"] if gref is None else gref + g["vp"]
    for s, ob in zip(seis, obs):
        assert np.max(np.abs(ob)) > 0
        assert rel_l2(s, ob) <= TOL
    g = sw.swgradient(params, sw.VpAcousticCDMaterialProperties(v0), shots, [sw.L2Misfit(o) for o in obs],
                      runparams=sw.RunParameters(erroronPPW=False), gradparams=sw.GradParameters(check_freq=cf))
    assert np.max(np.abs(gref)) > 0
    assert rel_l2(g["vp"], gref) <= TOL


@pytest.mark.parametrize("seed", range(5))
def test_cd3d_random(oracle_mod, seed):
    rng = np.random.default_rng(200 + seed)
    n, halo, freetop, nts, cf = _config(rng, (140, 20, 20), (190, 40, 34), (3, 6, 8))
    nts = [nt - 40 for nt in nts]
    print(f"[random cd3d seed {seed}: n {n} halo {halo} freetop {freetop} nts {nts} check_freq {cf}]", flush=True)
    vt = _field(rng, n, 1800.0, 2600.0, sigma=1.5)
    v0 = syn.gaussian_smooth(vt, 3.0)
    shots, osh = _scalar_shots(rng, n, nts, np.float64)
    params = _acoustic_params(n, halo, freetop, nts)
    seis = _forward(params, sw.VpAcousticCDMaterialProperties(vt), shots)
    obs = oracle_mod.acoustic_cd3d_forward(vt, (H, H, H), DT, max(nts), halo, 1e-4, freetop, osh)
    for s, ob in zip(seis, obs):
        assert np.max(np.abs(ob)) > 0
        assert rel_l2(s, ob) <= TOL
    gref, _, _ = oracle_mod.acoustic_cd3d_gradient(v0, (H, H, H), DT, max(nts), halo, 1e-4, freetop,
                                                   [dict(o, observed=ob) for o, ob in zip(osh, obs)], check_freq=cf)
    g = sw.swgradient(params, sw.VpAcousticCDMaterialProperties(v0), shots, [sw.L2Misfit(o) for o in obs],
                      runparams=sw.RunParameters(erroronPPW=False), gradparams=sw.GradParameters(check_freq=cf))
    assert np.max(np.abs(gref["vp"])) > 0
    assert rel_l2(g["vp"], gref["vp"]) <= TOL


@pytest.mark.parametrize("seed", range(8))
def test_vd2d_random(oracle_mod, seed):
    rng = np.random.default_rng(300 + seed)
    n, halo, freetop, nts, cf = _config(rng, (140, 60), (230, 140), (0, 5, 12))
    print(f"[random vd2d seed {seed}: n {n} halo {halo} freetop {freetop} nts {nts} check_freq {cf}]", flush=True)
    vt = _field(rng, n, 1800.0, 2800.0)
    rt = _field(rng, n, 1900.0, 2500.0)
    v0, r0 = syn.gaussian_smooth(vt, 4.0), syn.gaussian_smooth(rt, 4.0)
    shots, osh = _scalar_shots(rng, n, nts, np.float64)
    params = _acoustic_params(n, halo, freetop, nts)
    seis = _forward(params, sw.VpRhoAcousticVDMaterialProperties(vt, rt), shots)
    obs = oracle_mod.acoustic_vd_forward(vt, rt, (H, H), DT, max(nts), halo, 1e-4, freetop, osh)
    for s, ob in zip(seis, obs):
        assert np.max(np.abs(ob)) > 0
        assert rel_l2(s, ob) <= TOL
    gref, _, _ = oracle_mod.acoustic_vd_gradient(v0, r0, (H, H), DT, max(nts), halo, 1e-4, freetop,
                                                 [dict(o, observed=ob) for o, ob in zip(osh, obs)], check_freq=cf)
    g = sw.swgradient(params, sw.VpRhoAcousticVDMaterialProperties(v0, r0), shots, [sw.L2Misfit(o) for o in obs],
                      runparams=sw.RunParameters(erroronPPW=False), gradparams=sw.GradParameters(check_freq=cf))
    for k in ("vp", "rho"):
        assert np.max(np.abs(gref[k])) > 0
        assert rel_l2(g[k], gref[k]) <= TOL, k


@pytest.mark.parametrize("seed", range(8))
def test_elastic_random(oracle_mod, seed):
    rng = np.random.default_rng(400 + seed)
    n, halo, freetop, nts, cf = _config(rng, (140, 50), (220, 110), (4, 10, 14))
    kind = "mt" if seed % 2 == 0 else "ef"
    print(f"[random elastic seed {seed}: {kind} n {n} halo {halo} freetop {freetop} nts {nts} check_freq {cf}]",
          flush=True)
    rho = _field(rng, n, 2000.0, 2300.0)
    vp = _field(rng, n, 2600.0, 3100.0)
    vs = vp / np.sqrt(3) * _field(rng, n, 0.9, 1.0)
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    m0 = [syn.gaussian_smooth(a, 4.0) for a in (lam, mu, rho)]
    ntmax = max(nts)
    tf = syn.ricker_tf(ntmax, DT, 6.0, 1.2 / 6.0)
    shots, osh = [], []
    for k, nt in enumerate(nts):
        f0 = 6.0 * (1.0 + 0.45 * k)
        src = _points(rng, 1, n, 2.0)
        rec = _points(rng, int(rng.integers(3, 10)), n, 0.0)
        if kind == "mt":
            mom = rng.uniform(-1.0, 1.0, 3) * 1e3
            shots.append(sw.MomentTensorShot(
                sw.MomentTensorSources(src, tf[:nt, None], [sw.MomentTensor2D(*mom)], f0),
                sw.VectorReceivers(rec, nt)))
            osh.append(dict(src_pos=src, srctf=tf[:nt, None] * 1.0, moments=mom[None, :], f0=f0, rec_pos=rec,
                            nt=nt))
        else:
            stf = np.zeros((nt, 2, 1))
            stf[:, 0, 0] = rng.uniform(-1, 1) * 1e3 * tf[:nt]
            stf[:, 1, 0] = rng.uniform(-1, 1) * 1e3 * tf[:nt]
            shots.append(sw.ExternalForceShot(sw.ExternalForceSources(src, stf, f0), sw.VectorReceivers(rec, nt)))
            osh.append(dict(src_pos=src, srctf=stf, f0=f0, rec_pos=rec, nt=nt))
    params = sw.InputParametersElastic(ntmax, DT, n, (H, H),
                                       sw.CPMLBoundaryConditionParameters(halo=halo, rcoef=1e-4, freeboundtop=freetop))
    args = ((H, H), DT, ntmax, halo, 1e-4, freetop)
    seis = _forward(params, sw.ElasticIsoMaterialProperties(lam, mu, rho), shots)
    obs = oracle_mod.elastic_forward(lam, mu, rho, *args, osh)
    for s, ob in zip(seis, obs):
        assert np.max(np.abs(ob)) > 0
        assert rel_l2(s, ob) <= TOL
    gref, _, _ = oracle_mod.elastic_gradient(*m0, *args, [dict(o, observed=ob) for o, ob in zip(osh, obs)],
                                             check_freq=cf)
    g = sw.swgradient(params, sw.ElasticIsoMaterialProperties(*m0), shots, [sw.L2Misfit(o) for o in obs],
                      runparams=sw.RunParameters(erroronPPW=False), gradparams=sw.GradParameters(check_freq=cf))
    for k in ("rho", "lambda", "mu"):
        assert np.max(np.abs(gref[k])) > 0
        assert rel_l2(g[k], gref[k]) <= TOL, k
