"""Batch splitting under a device memory budget, and the plan state after a
rejected shot upload.

* A call whose shots do not fit one plan in the HBM budget (state and
  checkpoints of every shot) runs as several smaller batches: same
  seismograms bit for bit, same gradient to rounding (the per-batch sums
  are added on the host), FP64.
* swx_set_shots / swx_set_shots_elastic that reject a batch (a receiver
  outside the grid) leave the plan without a batch: swx_forward then fails
  with SWX_ESTATE instead of running on half-rebuilt tables."""
import numpy as np
import pytest

import seismicwaves as sw
from seismicwaves import _native
from seismicwaves import synthetic as syn

pytestmark = pytest.mark.gpu

H, DT, F0, NT, HALO = 10.0, 1e-3, 8.0, 220, 10


def _cd2d_problem(nshots):
    n = (150, 110)
    rng = np.random.default_rng(11)
    vt = syn.layered(n[0], n[1], (H, H), rng, nlayers=4, vmin=1800.0, vmax=2600.0)
    v0 = syn.gaussian_smooth(vt, 4.0)
    tf = syn.ricker_tf(NT, DT, F0, 1.2 / F0)[:, None]
    rec = syn.surface_receivers(25, n[0], (H, H), HALO, 30.0)
    xs = np.linspace(300.0, 1200.0, nshots)

    def shots():
        return [sw.ScalarShot(sw.ScalarSources(np.array([[x, 40.0]]), tf, F0), sw.ScalarReceivers(rec, NT))
                for x in xs]
    params = sw.InputParametersAcoustic(NT, DT, n, (H, H), sw.CPMLBoundaryConditionParameters(halo=HALO))
    return params, vt, v0, shots


def test_gradient_split_into_batches_under_memory_budget():
    params, vt, v0, shots = _cd2d_problem(4)
    obs_shots = shots()
    sw.swforward(params, sw.VpAcousticCDMaterialProperties(vt), obs_shots, runparams=sw.RunParameters())
    misfit = [sw.L2Misfit(s.recs.seismograms) for s in obs_shots]
    gp = sw.GradParameters(check_freq=20)
    # one batch of 4 shots, unconstrained
    ref_shots = shots()
    gref = sw.swgradient(params, sw.VpAcousticCDMaterialProperties(v0), ref_shots, misfit,
                         runparams=sw.RunParameters(), gradparams=gp)
    # the footprint of a 1-shot plan at the most economical replay interval
    # (check_freq ~ sqrt(2 nt)); twice that holds two shots but not four
    p1 = _native.Plan(physics=_native.SWX_ACOUSTIC_CD, dtype=np.float64, n=params.gridsize, h=(H, H), dt=DT,
                      halo=HALO, freetop=True, max_shots=1, nt_max=NT, check_freq=20, gradient=True)
    inf = p1.info()
    one = inf.bytes_state + inf.bytes_checkpoints + inf.bytes_replay
    p1.close()
    budget = 2.2 * one
    ws = sw.build_wavesim(params, sw.VpAcousticCDMaterialProperties(v0),
                          runparams=sw.RunParameters(mem_budget=budget), gradparams=gp, gradient=True)
    split_shots = shots()
    g = sw.swgradient(ws, sw.VpAcousticCDMaterialProperties(v0), split_shots, misfit)
    print(f"1-shot plan {one} bytes, budget {budget:.0f}: batches of {ws._fit_batch} shots", flush=True)
    assert 1 <= ws._fit_batch < 4
    for a, b in zip(ref_shots, split_shots):
        assert np.array_equal(a.recs.seismograms, b.recs.seismograms)
    err = np.linalg.norm(g["vp"] - gref["vp"]) / np.linalg.norm(gref["vp"])
    print(f"gradient of split batches vs one batch: rel-L2 {err:.2e}", flush=True)
    assert err <= 1e-13


def test_rejected_shot_batch_leaves_plan_without_batch():
    params, vt, v0, shots = _cd2d_problem(2)
    ws = sw.build_wavesim(params, sw.VpAcousticCDMaterialProperties(vt), runparams=sw.RunParameters(),
                          gradient=False)
    ws.set_matprop(sw.VpAcousticCDMaterialProperties(vt))
    good = [ws.shot_desc(sh) for sh in shots()]
    plan = ws._ensure_plan(2)
    plan.set_shots(good)
    tr = np.zeros(sum(int(np.prod(d["tshape"])) for d in good))
    plan.forward(tr, save_state=False)
    assert np.max(np.abs(tr)) > 0
    bad = [dict(d) for d in good]
    bad[1]["rec_idx"] = bad[1]["rec_idx"].copy()
    bad[1]["rec_idx"][3, 0] = params.gridsize[0] + 5        # receiver outside the grid
    with pytest.raises(AssertionError, match="not inside the grid"):
        plan.set_shots(bad)
    with pytest.raises(_native.SwxError, match="swx_set_shots must be called"):
        plan.forward(tr, save_state=False)
    # a valid batch makes the plan usable again, with the same results
    plan.set_shots(good)
    tr2 = np.zeros_like(tr)
    plan.forward(tr2, save_state=False)
    assert np.array_equal(tr, tr2)


def test_rejected_elastic_batch_leaves_plan_without_batch():
    n = (120, 90)
    lam = np.full(n, 2500.0 * 2800.0 ** 2 / 3)
    mu = np.full(n, 2500.0 * 1600.0 ** 2)
    rho = np.full(n, 2500.0)
    nt = 120
    tf = syn.ricker_tf(nt, DT, F0, 1.2 / F0)
    rec = syn.surface_receivers(9, n[0], (H, H), HALO, 0.0)
    shot = sw.MomentTensorShot(sw.MomentTensorSources(np.array([[600.0, 300.0]]), tf[:, None],
                                                      [sw.MomentTensor2D(1e3, 1e3, 0.0)], F0),
                               sw.VectorReceivers(rec, nt))
    params = sw.InputParametersElastic(nt, DT, n, (H, H), sw.CPMLBoundaryConditionParameters(halo=HALO))
    mp = sw.ElasticIsoMaterialProperties(lam, mu, rho)
    ws = sw.build_wavesim(params, mp, runparams=sw.RunParameters(erroronPPW=False), gradient=False)
    ws.set_matprop(mp)
    good = [ws.shot_desc(shot)]
    plan = ws._ensure_plan(1)
    plan.set_shots_elastic(good)
    tr = np.zeros(int(np.prod(good[0]["tshape"])))
    plan.forward(tr, save_state=False)
    bad = dict(good[0])
    ij, c = bad["rec"][1][2]
    ij = np.array(ij, copy=True).reshape(-1, 2)
    ij[0, 0] = n[0] + 3                                     # a uz receiver tap outside the grid
    bad["rec"] = [bad["rec"][0], list(bad["rec"][1])]
    bad["rec"][1][2] = (ij, c)
    with pytest.raises(AssertionError):
        plan.set_shots_elastic([bad])
    with pytest.raises(_native.SwxError, match="swx_set_shots must be called"):
        plan.forward(tr, save_state=False)
    plan.set_shots_elastic(good)
    tr2 = np.zeros_like(tr)
    plan.forward(tr2, save_state=False)
    assert np.array_equal(tr, tr2)
