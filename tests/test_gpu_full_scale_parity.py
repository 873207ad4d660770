"""Accuracy of the GPU path against the FP64 oracle on the benchmark
workloads at their full grid sizes and full lengths.

* CD 2D 1000 x 1000, nt 4000, one shot, check_freq 50: FP64 GPU to 1e-10
  and FP32 GPU to 1e-4 (rel-L2) on seismograms, gradient and misfit.
* The same problem in FP64: the returned "vp" gradient against central
  differences of the misfit (-dJ/dvp, the reference's sign) on 20 seeded
  cells at least 10 cells from the source, the receivers and the C-PML.
* VD 2D 2000 x 800, the 64-shot batch of nt 3000..5000 in one plan: the
  per-shot gradients and seismograms of its shortest and its longest shot
  against the oracle run on those shots alone, FP32 to 1e-4.
* Elastic 2D 1500 x 600, nt 5000: one explosive moment-tensor shot and one
  vertical-force shot, FP32 to 1e-4 on seismograms and the rho, lambda, mu
  gradients.

The inputs are the FP32 values of each workload; the FP64 runs (GPU and
oracle) see the same values widened, so the FP32 figures measure FP32
arithmetic over thousands of steps.  Observed data come from a forward run
on the true model.  Every test prints its errors."""
import numpy as np
import pytest

import bench
import seismicwaves as sw

pytestmark = pytest.mark.gpu

F32, F64 = np.float32, np.float64


def _rel(a, b):
    a = np.asarray(a, F64)
    b = np.asarray(b, F64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _scalar_shot(sh, dtype, nt=None):
    nt = sh.recs.seismograms.shape[0] if nt is None else nt
    return sw.ScalarShot(sw.ScalarSources(sh.srcs.positions, np.asarray(sh.srcs.tf[:nt], dtype), sh.srcs.domfreq),
                         sw.ScalarReceivers(sh.recs.positions, nt, dtype=dtype))


def _oracle_scalar_shot(sh, observed):
    nt = observed.shape[0]
    return dict(src_pos=sh.srcs.positions, srctf=np.asarray(sh.srcs.tf[:nt], F64), f0=sh.srcs.domfreq,
                rec_pos=sh.recs.positions, nt=nt, observed=np.asarray(observed, F64))


# --------------------------------------------------------------------------
# CD 2D, 1000 x 1000, nt 4000
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cd2d():
    wl = bench.workload_cd2d_grad_1000()
    p = wl["params"]
    sw.swforward(p, wl["mp_true"], wl["shots"], runparams=sw.RunParameters())
    return wl, wl["shots"][0].recs.seismograms.copy()          # observed, FP32


def test_cd2d_1000_nt4000_fp64_and_fp32_vs_oracle(cd2d, oracle_mod):
    wl, obs = cd2d
    p = wl["params"]
    sh = wl["shots"][0]
    print("[full scale CD 2D 1000x1000 nt 4000: oracle FP64]", flush=True)
    v0 = np.asarray(wl["mp0"].vp, F32)
    bc = p.boundcond
    gref, mref, sref = oracle_mod.acoustic_cd_gradient(v0.astype(F64), p.gridspacing, p.dt, p.ntimesteps, bc.halo,
                                                       bc.rcoef, bc.freeboundtop, [_oracle_scalar_shot(sh, obs)],
                                                       check_freq=50)
    for dtype, tol in ((F64, 1e-10), (F32, 1e-4)):
        shot = _scalar_shot(sh, dtype)
        g, mis = sw.swgradient(p, sw.VpAcousticCDMaterialProperties(v0.astype(dtype)), [shot],
                               [sw.L2Misfit(obs.astype(dtype))], runparams=sw.RunParameters(),
                               gradparams=sw.GradParameters(check_freq=50, compute_misfit=True))
        es, eg, em = _rel(shot.recs.seismograms, sref[0]), _rel(g["vp"], gref["vp"]), abs(mis - mref) / abs(mref)
        print(f"CD 2D 1000x1000 nt 4000 {np.dtype(dtype).name}: rel-L2 seismograms {es:.2e}, gradient {eg:.2e}, "
              f"misfit {em:.2e} (tolerance {tol:g})", flush=True)
        assert es <= tol and eg <= tol and em <= tol


def test_cd2d_1000_nt4000_gradient_vs_central_differences(cd2d):
    """-dJ/dvp by central differences (delta = 1e-3 vp) in FP64 on 20 seeded
    cells against the returned gradient, rtol 1e-4."""
    wl, obs = cd2d
    p = wl["params"]
    sh = wl["shots"][0]
    v0 = np.asarray(wl["mp0"].vp, F32).astype(F64)
    misfit = [sw.L2Misfit(obs.astype(F64))]
    rp = sw.RunParameters()
    ws_g = sw.build_wavesim(p, sw.VpAcousticCDMaterialProperties(v0), runparams=rp,
                            gradparams=sw.GradParameters(check_freq=50), gradient=True)
    g = sw.swgradient(ws_g, sw.VpAcousticCDMaterialProperties(v0), [_scalar_shot(sh, F64)], misfit)["vp"]
    ws_f = sw.build_wavesim(p, sw.VpAcousticCDMaterialProperties(v0), runparams=rp, gradient=False)

    def J(v):
        return sw.swmisfit(ws_f, sw.VpAcousticCDMaterialProperties(v), [_scalar_shot(sh, F64)], misfit)

    # 1-based cells: C-PML is i <= 20, i >= 981, j >= 981 (free top); source
    # (501, 3), receivers on row 3 from i = 21 to 980.  Within 4000 steps
    # (2 s) the wave path from the surface and back only reaches about 2 km
    # down: deeper cells have no sensitivity (the gradient underflows to 0
    # there and so does the misfit change), so cells are drawn from the
    # upper 300 rows and a cell is kept if its gradient is at least 1e-2 of
    # the largest one there (below that, the FP64 rounding of the misfit
    # difference, ~1e-16 J / (g delta), reaches 1e-5 of the difference).
    rng = np.random.default_rng(20)
    cells = []
    vmax = float(v0.max())
    gmax = float(np.max(np.abs(g[30:970, 13:300])))
    while len(cells) < 20:
        i, j = int(rng.integers(31, 971)), int(rng.integers(14, 301))
        if abs(i - 501) + abs(j - 3) >= 10 and v0[i - 1, j - 1] < vmax * (1 - 2e-3) and (i, j) not in cells \
                and abs(g[i - 1, j - 1]) >= 1e-2 * gmax:
            cells.append((i, j))
    errs = []
    for (i, j) in cells:
        d = 1e-3 * v0[i - 1, j - 1]
        vp, vm = v0.copy(), v0.copy()
        vp[i - 1, j - 1] += d
        vm[i - 1, j - 1] -= d
        fd = -(J(vp) - J(vm)) / (2 * d)
        gc = float(g[i - 1, j - 1])
        errs.append(abs(gc - fd) / abs(fd))
        print(f"cell ({i}, {j}): gradient {gc:.6e}, -dJ/dvp by central differences {fd:.6e}, "
              f"relative difference {errs[-1]:.2e}", flush=True)
    print(f"central differences on 20 cells: largest relative difference {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-4


# --------------------------------------------------------------------------
# VD 2D, 2000 x 800, 64 shots of nt 3000..5000 in one plan
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def vd2d():
    wl = bench.workload_vd2d_batched()
    p = wl["params"]
    rp = sw.RunParameters(**wl["run"])
    sw.swforward(p, wl["mp_true"], wl["shots"], runparams=rp)
    obs = [sh.recs.seismograms.copy() for sh in wl["shots"]]
    shots = [_scalar_shot(sh, F32) for sh in wl["shots"]]
    grads = sw.swgradient(p, wl["mp0"], shots, [sw.L2Misfit(o) for o in obs], runparams=rp,
                          gradparams=sw.GradParameters(check_freq=50), per_shot_gradients=True)
    return wl, obs, shots, grads


@pytest.mark.parametrize("which", ["shortest", "longest"])
def test_vd2d_64shot_batch_per_shot_fp32_vs_oracle(vd2d, oracle_mod, which):
    wl, obs, shots, grads = vd2d
    p = wl["params"]
    nts = [o.shape[0] for o in obs]
    k = int(np.argmin(nts) if which == "shortest" else np.argmax(nts))
    print(f"[full scale VD 2000x800: shot {k} of the 64-shot batch, nt {nts[k]}: oracle FP64]", flush=True)
    bc = p.boundcond
    mp = wl["mp0"]
    gref, _, sref = oracle_mod.acoustic_vd_gradient(np.asarray(mp.vp, F64), np.asarray(mp.rho, F64), p.gridspacing,
                                                    p.dt, p.ntimesteps, bc.halo, bc.rcoef, bc.freeboundtop,
                                                    [_oracle_scalar_shot(wl["shots"][k], obs[k])], check_freq=50)
    es = _rel(shots[k].recs.seismograms, sref[0])
    ev, er = _rel(grads[k]["vp"], gref["vp"]), _rel(grads[k]["rho"], gref["rho"])
    print(f"VD 2000x800 batched shot {k} (nt {nts[k]}) FP32: rel-L2 seismograms {es:.2e}, gradient vp {ev:.2e}, "
          f"rho {er:.2e} (tolerance 1e-4)", flush=True)
    assert es <= 1e-4 and ev <= 1e-4 and er <= 1e-4


# --------------------------------------------------------------------------
# Elastic 2D, 1500 x 600, nt 5000
# --------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["moment_tensor", "vertical_force"])
def test_elastic_1500x600_nt5000_fp32_vs_oracle(oracle_mod, kind):
    wl = bench.workload_el2d_batched(nt=5000)
    p = wl["params"]
    sh = wl["groups"][0 if kind == "moment_tensor" else 1][0]
    rp = sw.RunParameters(**wl["run"])
    sw.swforward(p, wl["mp_true"], [sh], runparams=rp)
    obs = sh.recs.seismograms.copy()
    print(f"[full scale elastic 1500x600 nt 5000, {kind} shot: oracle FP64]", flush=True)
    mp = wl["mp0"]
    bc = p.boundcond
    o = dict(src_pos=sh.srcs.positions, srctf=np.asarray(sh.srcs.tf, F64), f0=sh.srcs.domfreq,
             rec_pos=sh.recs.positions, observed=obs.astype(F64))
    if kind == "moment_tensor":
        o["moments"] = np.array([[m.Mxx, m.Mzz, m.Mxz] for m in sh.srcs.momtens], F64)
    gref, mref, sref = oracle_mod.elastic_gradient(np.asarray(mp.lam, F64), np.asarray(mp.mu, F64),
                                                   np.asarray(mp.rho, F64), p.gridspacing, p.dt, p.ntimesteps,
                                                   bc.halo, bc.rcoef, bc.freeboundtop, [o], check_freq=50)
    g, mis = sw.swgradient(p, mp, [sh], [sw.L2Misfit(obs)], runparams=rp,
                           gradparams=sw.GradParameters(check_freq=50, compute_misfit=True))
    es = _rel(sh.recs.seismograms, sref[0])
    e = {k: _rel(g[k], gref[k]) for k in ("rho", "lambda", "mu")}
    em = abs(mis - mref) / abs(mref)
    print(f"elastic 1500x600 nt 5000 {kind} FP32: rel-L2 seismograms {es:.2e}, gradient rho {e['rho']:.2e}, "
          f"lambda {e['lambda']:.2e}, mu {e['mu']:.2e}, misfit {em:.2e} (tolerance 1e-4)", flush=True)
    assert es <= 1e-4 and em <= 1e-4 and max(e.values()) <= 1e-4
