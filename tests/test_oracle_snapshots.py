"""The oracle's snapshot sink (LinearSnapshotter, snapshotter.jl:1-41)
pinned by the seismograms of the same run: record_receivers! reads the
field the snapshot holds right after the step, so the snapshot of step it
sampled at the receivers is the seismogram row it (CD 2D, VD pressure,
elastic ux/uz with nearest-point receivers), bit for bit."""
import numpy as np

H, DT = 10.0, 1e-3


def _src(o, nt, x=250.0, z=220.0):
    return dict(src_pos=np.array([[x, z]]), srctf=o.rickerstf(np.arange(nt) * DT, 0.06, 15.0)[:, None], f0=15.0)


def test_cd2d_snapshot_at_receivers_is_trace(oracle_mod):
    o = oracle_mod
    n, nt = (60, 50), 90
    vp = np.linspace(1800.0, 2400.0, n[0] * n[1]).reshape(n, order="F")
    rec = np.array([[x, z] for x in (150.0, 300.0, 420.0) for z in (120.0, 300.0)])
    sh = dict(_src(o, nt), rec_pos=rec)
    tr = {}
    snaps = o.with_snapshots(20, nt, [n], np.float64,
                             lambda: tr.setdefault(0, o.acoustic_cd_forward(vp, (H, H), DT, nt, 5, 1e-4, True, [sh])[0]))
    ri = o.find_nearest_grid_points(rec, (H, H)) - 1
    assert sorted(snaps) == [20, 40, 60, 80]
    for it, (p,) in snaps.items():
        assert np.array_equal(p[ri[:, 0], ri[:, 1]], tr[0][it - 1])


def test_vd2d_snapshot_at_receivers_is_trace(oracle_mod):
    o = oracle_mod
    n, nt = (60, 50), 90
    vp = np.full(n, 2000.0)
    rho = np.linspace(1900.0, 2300.0, n[0] * n[1]).reshape(n, order="F")
    rec = np.array([[x, 200.0] for x in (100.0, 300.0, 450.0)])
    sh = dict(_src(o, nt), rec_pos=rec, nt=nt)
    tr = {}
    shapes = [n, (n[0] - 1, n[1]), (n[0], n[1] - 1)]
    snaps = o.with_snapshots(30, nt, shapes, np.float64,
                             lambda: tr.setdefault(0, o.acoustic_vd_forward(vp, rho, (H, H), DT, nt, 5, 1e-4, True,
                                                                            [sh])[0]))
    ri = o.find_nearest_grid_points(rec, (H, H)) - 1
    for it, (p, vx, vz) in snaps.items():
        assert np.array_equal(p[ri[:, 0], ri[:, 1]], tr[0][it - 1])
        assert np.max(np.abs(vx)) > 0 and np.max(np.abs(vz)) > 0


def test_elastic_snapshot_at_receivers_is_trace(oracle_mod):
    o = oracle_mod
    n, nt = (60, 50), 90
    rho = np.full(n, 2200.0)
    mu = rho * 1600.0 ** 2
    lam = rho * 2800.0 ** 2 - 2 * mu
    rec = np.array([[x, 100.0] for x in (100.0, 310.0, 450.0)])
    sh = dict(_src(o, nt), moments=np.array([[1e3, 1e3, 0.0]]), rec_pos=rec, nt=nt)
    tr = {}
    shapes = [(n[0] - 1, n[1]), (n[0], n[1] - 1), n, n, (n[0] - 1, n[1] - 1)]
    snaps = o.with_snapshots(30, nt, shapes, np.float64,
                             lambda: tr.setdefault(0, o.elastic_forward(lam, mu, rho, (H, H), DT, nt, 5, 1e-4, True,
                                                                        [sh], sincinterp=False)[0]))
    ri = o.find_nearest_grid_points(rec, (H, H)) - 1
    for it, (ux, uz, sxx, szz, sxz) in snaps.items():
        assert np.array_equal(ux[ri[:, 0], ri[:, 1]], tr[0][it - 1, 0])
        assert np.array_equal(uz[ri[:, 0], ri[:, 1]], tr[0][it - 1, 1])
        assert np.all(szz[:, 0] == 0)          # free surface: sigma_zz = 0 on the top row
        assert np.max(np.abs(sxx)) > 0 and np.max(np.abs(sxz)) > 0
