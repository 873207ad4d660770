"""swx_forward_noalloc, the forward of one 2D acoustic constant-density shot
run in the library's static scratch array (no device allocation), against
swx_forward on a 1-shot plan: bitwise equal seismograms on two calls in a
row, the second on a smaller grid with fewer steps and receivers, in FP64
and FP32.  A shot larger than the scratch array is refused with
SWX_ENOMEM before any device work."""
import numpy as np
import pytest

import seismicwaves as sw
from seismicwaves import _native
from seismicwaves import synthetic as syn

pytestmark = pytest.mark.gpu

H, DT, F0 = 10.0, 1e-3, 8.0


def _problem(n, nt, nrec, halo, freetop, seed, dtype):
    rng = np.random.default_rng(seed)
    vp = syn.layered(n[0], n[1], (H, H), rng, nlayers=5, vmin=1800.0, vmax=2800.0).astype(dtype)
    tf = syn.ricker_tf(nt, DT, F0, 1.2 / F0)
    # two sources on one cell (added in index order) and one elsewhere;
    # receivers spread over the grid, one on the outermost row
    src = np.array([[0.4 * (n[0] - 1) * H, 0.3 * (n[1] - 1) * H]] * 2 + [[0.7 * (n[0] - 1) * H, 0.5 * (n[1] - 1) * H]])
    stf = np.stack([tf, -0.5 * tf, 0.8 * tf], 1).astype(dtype)
    rec = np.stack([rng.uniform(0, (n[0] - 1) * H, nrec), rng.uniform(0, (n[1] - 1) * H, nrec)], 1)
    rec[0, 1] = 0.0
    params = sw.InputParametersAcoustic(nt, DT, n, (H, H),
                                        sw.CPMLBoundaryConditionParameters(halo=halo, freeboundtop=freetop))
    mp = sw.VpAcousticCDMaterialProperties(vp)
    shot = sw.ScalarShot(sw.ScalarSources(src, stf, F0), sw.ScalarReceivers(rec, nt, dtype=dtype))
    ws = sw.build_wavesim(params, mp, runparams=sw.RunParameters(erroronPPW=False), gradient=False)
    ws.set_matprop(mp)
    return ws, ws.shot_desc(shot), ws.model_fields()[0]


def _plan_forward(ws, desc):
    plan = ws._ensure_plan(1)
    plan.set_shots([desc])
    nt, nrec = desc["tshape"]
    tr = np.zeros(nt * nrec, dtype=ws.dtype)
    plan.forward(tr, save_state=False)
    return tr.reshape((nt, nrec), order="F")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_forward_equals_plan_forward(dtype):
    for n, nt, nrec, halo, freetop, seed in (((310, 230), 420, 40, 20, True, 1), ((170, 120), 260, 17, 8, False, 2)):
        print(f"[noalloc forward {np.dtype(dtype).name} n {n} nt {nt} nrec {nrec} halo {halo} freetop {freetop}]",
              flush=True)
        ws, desc, fact = _problem(n, nt, nrec, halo, freetop, seed, dtype)
        ref = _plan_forward(ws, desc)
        out = _native.forward_noalloc(dtype=dtype, n=n, h=ws.spacing, dt=float(ws.dt), halo=halo, fact=fact,
                                      shot=desc)
        assert np.max(np.abs(ref)) > 0
        assert out.shape == ref.shape
        assert np.array_equal(out, ref)


def test_noalloc_refuses_shot_larger_than_scratch():
    n, nt = (3000, 3000), 10
    ws, desc, fact = _problem(n, nt, 5, 20, True, 3, np.float32)
    assert 4 * n[0] * n[1] * 4 > _native.NOALLOC_SCRATCH_BYTES
    with pytest.raises(MemoryError, match="scratch"):
        _native.forward_noalloc(dtype=np.float32, n=n, h=ws.spacing, dt=float(ws.dt), halo=20, fact=fact, shot=desc)
