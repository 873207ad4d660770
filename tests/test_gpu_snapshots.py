"""Snapshots of the forward run (RunParameters.snapevery; LinearSnapshotter,
snapshotter.jl:1-41) taken on the device against the oracle's, FP64 to
rel-L2 1e-10 per field, for a batch of shots of uneven length: "pcur" (CD
2D/3D), "pcur" and "vcur" (VD), "ucur" and "σ" (elastic, both source
kinds).  The seismograms of a run with snapshots are bitwise those of a
run without."""
import numpy as np
import pytest

import seismicwaves as sw
from seismicwaves import synthetic as syn

pytestmark = pytest.mark.gpu

H, DT, F0, EVERY = 10.0, 1e-3, 8.0, 25
TOL = 1e-10


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def _check(snaps, refs, names):
    assert len(snaps) == len(refs)
    for got, ref in zip(snaps, refs):
        assert sorted(got) == sorted(ref)
        for it, fields in ref.items():
            k = 0
            for name in names:
                g = got[it][name]
                g = g if isinstance(g, list) else [g]
                for arr in g:
                    assert arr.shape == fields[k].shape, (name, it)
                    if np.max(np.abs(fields[k])) > 0:
                        assert _rel(arr, fields[k]) <= TOL, (name, it)
                    else:
                        assert np.max(np.abs(arr)) == 0
                    k += 1


def _acoustic(n, nts, oracle_mod, vd):
    rng = np.random.default_rng(7 + len(n))
    vp = syn.gaussian_smooth(rng.uniform(1900.0, 2600.0, n), 2.0)
    rho = syn.gaussian_smooth(rng.uniform(1900.0, 2400.0, n), 2.0)
    tf = syn.ricker_tf(max(nts), DT, F0, 1.2 / F0)
    shots, osh = [], []
    for k, nt in enumerate(nts):
        src = np.array([[(0.35 + 0.2 * k) * (n[0] - 1) * H] + [0.45 * (m - 1) * H for m in n[1:]]])
        rec = np.array([[f * (n[0] - 1) * H] + [0.3 * (m - 1) * H for m in n[1:]] for f in (0.25, 0.55)])
        shots.append(sw.ScalarShot(sw.ScalarSources(src, tf[:nt, None], F0 * (1 + 0.3 * k)),
                                   sw.ScalarReceivers(rec, nt)))
        osh.append(dict(src_pos=src, srctf=tf[:nt, None], f0=F0 * (1 + 0.3 * k), rec_pos=rec, nt=nt))
    bc = sw.CPMLBoundaryConditionParameters(halo=6, freeboundtop=len(n) == 2)
    params = sw.InputParametersAcoustic(max(nts), DT, n, (H,) * len(n), bc)
    mp = sw.VpRhoAcousticVDMaterialProperties(vp, rho) if vd else sw.VpAcousticCDMaterialProperties(vp)
    snaps = sw.swforward(params, mp, shots, runparams=sw.RunParameters(snapevery=EVERY, erroronPPW=False))
    seis = [sh.recs.seismograms.copy() for sh in shots]
    plain = [sw.ScalarShot(sh.srcs, sw.ScalarReceivers(sh.recs.positions, sh.recs.seismograms.shape[0]))
             for sh in shots]
    sw.swforward(params, mp, plain, runparams=sw.RunParameters(erroronPPW=False))
    for a, b in zip(seis, plain):
        assert np.array_equal(a, b.recs.seismograms)
    refs = []
    for o in osh:
        if vd:
            nx, nz = n
            shapes = [n, (nx - 1, nz), (nx, nz - 1)]
            run = lambda o=o: oracle_mod.acoustic_vd_forward(vp, rho, (H, H), DT, o["nt"], 6, 1e-4, True, [o])
        elif len(n) == 3:
            shapes = [n]
            run = lambda o=o: oracle_mod.acoustic_cd3d_forward(vp, (H,) * 3, DT, o["nt"], 6, 1e-4, False, [o])
        else:
            shapes = [n]
            run = lambda o=o: oracle_mod.acoustic_cd_forward(vp, (H, H), DT, o["nt"], 6, 1e-4, True, [o])
        refs.append(oracle_mod.with_snapshots(EVERY, o["nt"], shapes, np.float64, run))
    _check(snaps, refs, ["pcur", "vcur"] if vd else ["pcur"])


def test_cd2d_snapshots(oracle_mod):
    _acoustic((90, 70), [130, 101], oracle_mod, vd=False)


def test_cd3d_snapshots(oracle_mod):
    _acoustic((40, 36, 30), [80, 60], oracle_mod, vd=False)


def test_vd2d_snapshots(oracle_mod):
    _acoustic((90, 70), [130, 101], oracle_mod, vd=True)


@pytest.mark.parametrize("kind", ["mt", "ef"])
def test_elastic_snapshots(oracle_mod, kind):
    n = (100, 70)
    rng = np.random.default_rng(3)
    rho = syn.gaussian_smooth(rng.uniform(2000.0, 2300.0, n), 2.0)
    vp = syn.gaussian_smooth(rng.uniform(2700.0, 3100.0, n), 2.0)
    mu = rho * (vp / 1.8) ** 2
    lam = rho * vp ** 2 - 2 * mu
    nts = [120, 95]
    tf = syn.ricker_tf(max(nts), DT, 6.0, 1.2 / 6.0)
    rec = np.array([[x, 0.0] for x in (150.0, 420.0, 700.0)])
    shots, osh = [], []
    for k, nt in enumerate(nts):
        # the first source sits on the free surface (mirrored taps, edge stress cells)
        src = np.array([[(0.3 + 0.3 * k) * (n[0] - 1) * H, 0.0 if k == 0 else 180.0]])
        if kind == "mt":
            mom = np.array([1e3, 0.6e3, 0.4e3])
            shots.append(sw.MomentTensorShot(sw.MomentTensorSources(src, tf[:nt, None], [sw.MomentTensor2D(*mom)], 6.0),
                                             sw.VectorReceivers(rec, nt)))
            osh.append(dict(src_pos=src, srctf=tf[:nt, None], moments=mom[None, :], f0=6.0, rec_pos=rec, nt=nt))
        else:
            stf = np.zeros((nt, 2, 1))
            stf[:, 0, 0] = 0.7e3 * tf[:nt]
            stf[:, 1, 0] = 1e3 * tf[:nt]
            shots.append(sw.ExternalForceShot(sw.ExternalForceSources(src, stf, 6.0), sw.VectorReceivers(rec, nt)))
            osh.append(dict(src_pos=src, srctf=stf, f0=6.0, rec_pos=rec, nt=nt))
    params = sw.InputParametersElastic(max(nts), DT, n, (H, H), sw.CPMLBoundaryConditionParameters(halo=6))
    snaps = sw.swforward(params, sw.ElasticIsoMaterialProperties(lam, mu, rho), shots,
                         runparams=sw.RunParameters(snapevery=EVERY, erroronPPW=False))
    nx, nz = n
    shapes = [(nx - 1, nz), (nx, nz - 1), (nx, nz), (nx, nz), (nx - 1, nz - 1)]
    refs = [oracle_mod.with_snapshots(EVERY, o["nt"], shapes, np.float64,
                                      lambda o=o: oracle_mod.elastic_forward(lam, mu, rho, (H, H), DT, o["nt"], 6,
                                                                             1e-4, True, [o]))
            for o in osh]
    _check(snaps, refs, ["ucur", "σ"])
