"""Benchmark: FP32 forward + adjoint gradient throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full gradient of the workload's shots: forward with
device-resident checkpoints (traces copied to the host, as swgradient fills
the seismograms), L2 adjoint source and misfit on the device, adjoint with
replay and gradient correlation; shots are batched as far as the HBM budget
allows.  All inputs are uploaded to HBM before the timed region.  The JSON line carries

  value        GCell-updates/s = 2 * cells * sum(nt of every shot) / time
               (one forward and one adjoint update per cell and time step;
               replay is not counted as work),
  roofline     the dominant kernel's algorithmic bytes per launch over its
               average launch time (HIP events around every launch on the
               plan's stream) against MI355X's 8 TB/s,
  one_shot     the same kernels run one shot at a time (the plain 1-shot
               path), on a bounded sample of the shots,
  cpu_baseline the oracle's OpenMP restatement of the reference's threads
               backend on a bounded sample, same precision.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.join(REPO, "seismicwaves.jl_amd")
for _p in (REPO, PKG_ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "GCell-updates/s fwd+adjoint (batched shots) & % HBM roofline vs plain 1-shot path"
HBM_PEAK_GBS = 8000.0


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
def workload_vd2d_batched():
    """Batched 2D acoustic variable-density gradient w.r.t. vp and rho:
    2000 x 800, dh = 5 m, dt = 0.5 ms, 12 Hz Ricker, 64 shots at seeded x
    along z = 10 m, nt per shot uniform in 3000..5000, 500 receivers per
    shot at z = 10 m, L2 misfit, checkpoint every 50 steps.  vp: 8 seeded
    layers (1500-4000 m/s) plus 200 elliptical inclusions of +-15%; rho:
    Gardner with +-5% noise.  The starting model is the true one smoothed
    with sigma = 100 m."""
    import seismicwaves as sw
    from seismicwaves import synthetic as syn
    nx, nz = 2000, 800
    h = (5.0, 5.0)
    dt, f0, halo, nshots = 5e-4, 12.0, 20, 64
    rng = np.random.default_rng(2)
    vtrue = syn.add_ellipses(syn.layered(nx, nz, h, rng), h, rng, n=200)
    rtrue = syn.gardner(vtrue, rng)
    v0 = syn.gaussian_smooth(vtrue, 100.0 / h[0])
    r0 = syn.gaussian_smooth(rtrue, 100.0 / h[0])
    nts = rng.integers(3000, 5001, nshots)
    xs = rng.uniform(halo * h[0], (nx - 1 - halo) * h[0], nshots)
    ntmax = int(nts.max())
    tf = syn.ricker_tf(ntmax, dt, f0, 1.2 / f0)[:, None].astype(np.float32)
    rec = syn.surface_receivers(500, nx, h, halo, 10.0)
    shots = [sw.ScalarShot(sw.ScalarSources(np.array([[x, 10.0]]), tf[:nt], f0),
                           sw.ScalarReceivers(rec, int(nt), dtype=np.float32)) for x, nt in zip(xs, nts)]
    params = sw.InputParametersAcoustic(ntmax, dt, (nx, nz), h, sw.CPMLBoundaryConditionParameters(halo=halo))
    f32 = np.float32
    return dict(name="acoustic_vd_2d_gradient_2000x800_64shots_nt3000-5000", params=params, physics="vd2d",
                mp_true=sw.VpRhoAcousticVDMaterialProperties(vtrue.astype(f32), rtrue.astype(f32)),
                mp0=sw.VpRhoAcousticVDMaterialProperties(v0.astype(f32), r0.astype(f32)),
                shots=shots, check_freq=50,
                # overlapping -15% inclusions push the slowest cells below
                # 10 points per wavelength: warn instead of failing, as
                # RunParameters(erroronPPW=false) does in the reference
                run={"erroronPPW": False},
                config={"nx": nx, "nz": nz, "dh_m": 5.0, "dt_s": dt, "f0_hz": f0, "shots": nshots,
                        "nt_min": int(nts.min()), "nt_max": ntmax, "sum_nt": int(nts.sum()), "receivers": 500,
                        "check_freq": 50, "halo": halo})


def workload_cd2d_grad_1000(nt=4000):
    """2D acoustic constant-density gradient: 1000 x 1000, dh = 5 m,
    dt = 0.5 ms, nt = 4000, seeded 8-layer model, start model smoothed with
    sigma = 100 m, one 12 Hz Ricker shot at (2500 m, 10 m), 400 receivers at
    10 m depth, checkpoint every 50 steps with recompute."""
    import seismicwaves as sw
    from seismicwaves import synthetic as syn
    nx = nz = 1000
    h = (5.0, 5.0)
    dt, f0, halo = 5e-4, 12.0, 20
    rng = np.random.default_rng(1)
    vtrue = syn.layered(nx, nz, h, rng)
    v0 = syn.gaussian_smooth(vtrue, 100.0 / h[0])
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0)[:, None].astype(np.float32)
    rec = syn.surface_receivers(400, nx, h, halo, 10.0)
    params = sw.InputParametersAcoustic(nt, dt, (nx, nz), h, sw.CPMLBoundaryConditionParameters(halo=halo))
    shots = [sw.ScalarShot(sw.ScalarSources(np.array([[2500.0, 10.0]]), tf, f0),
                           sw.ScalarReceivers(rec, nt, dtype=np.float32))]
    return dict(name=f"acoustic_cd_2d_gradient_1000x1000_nt{nt}", params=params, physics="cd2d",
                mp_true=sw.VpAcousticCDMaterialProperties(vtrue.astype(np.float32)),
                mp0=sw.VpAcousticCDMaterialProperties(v0.astype(np.float32)), shots=shots, check_freq=50, run={},
                config={"nx": nx, "nz": nz, "dh_m": 5.0, "dt_s": dt, "nt": nt, "shots": 1, "receivers": 400,
                        "check_freq": 50, "halo": halo})


def workload_cd3d_600x600x400(nshots=8, nt=3000):
    """3D acoustic constant-density gradient: 600 x 600 x 400 cells of 10 m,
    dt = 1 ms, nt = 3000, C-PML (20 cells) on all faces, 7 Hz Ricker.  vp:
    8 seeded layers (1500-5000 m/s with depth) plus 50 ellipsoidal
    inclusions of +-10%; the starting model is the layered profile smoothed
    over 100 m in depth.  Shots on a 2 x 4 grid at 250 m depth (just below
    the top C-PML), each with a 100 x 100 patch of receivers at 10 m spacing
    centred on it at 240 m depth;
    checkpoint every 50 steps.  The shots run in batches of the size whose
    state and checkpoints fit the HBM budget (two of these shots per plan)."""
    import seismicwaves as sw
    from seismicwaves import synthetic as syn
    nx, ny, nz = 600, 600, 400
    h = (10.0, 10.0, 10.0)
    dt, f0, halo = 1e-3, 7.0, 20
    rng = np.random.default_rng(5)
    z = np.arange(nz) * h[2]
    ifaces = np.sort(rng.uniform(0.05, 0.95, 7)) * z[-1]
    speeds = np.sort(rng.uniform(1500.0, 5000.0, 8))
    prof = speeds[np.searchsorted(ifaces, z)].astype(np.float32)
    vtrue = np.empty((nx, ny, nz), dtype=np.float32)
    vtrue[:] = prof
    X, Y, Z = (np.arange(n) * dh for n, dh in zip((nx, ny, nz), h))
    for _ in range(50):
        c = rng.uniform([0, 0, 0], [X[-1], Y[-1], Z[-1]])
        ax = rng.uniform(60.0, 400.0, 3)
        sl = [slice(max(0, int((c[d] - ax[d]) / h[d])), min(n, int((c[d] + ax[d]) / h[d]) + 1))
              for d, n in enumerate((nx, ny, nz))]
        gx, gy, gz = np.meshgrid(X[sl[0]], Y[sl[1]], Z[sl[2]], indexing="ij")
        inside = ((gx - c[0]) / ax[0]) ** 2 + ((gy - c[1]) / ax[1]) ** 2 + ((gz - c[2]) / ax[2]) ** 2 <= 1.0
        vtrue[sl[0], sl[1], sl[2]] *= np.where(inside, np.float32(1.0 + rng.uniform(-0.1, 0.1)), np.float32(1.0))
    w = np.exp(-0.5 * (np.arange(-30, 31) * h[2] / 100.0) ** 2)
    sprof = np.convolve(np.pad(prof.astype(np.float64), 30, mode="edge"), w / w.sum(), mode="valid")
    v0 = np.empty_like(vtrue)
    v0[:] = sprof.astype(np.float32)
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0)[:, None].astype(np.float32)
    xs = np.array([1500.0, 4500.0])
    ys = np.array([750.0, 2250.0, 3750.0, 5250.0])
    pos = [(x, y) for y in ys for x in xs][:nshots]
    off = (np.arange(100) - 50) * 10.0
    shots = []
    for x, y in pos:
        rec = np.array([[x + a, y + b, 240.0] for a in off for b in off])
        shots.append(sw.ScalarShot(sw.ScalarSources(np.array([[x, y, 250.0]]), tf, f0),
                                   sw.ScalarReceivers(rec, nt, dtype=np.float32)))
    params = sw.InputParametersAcoustic(nt, dt, (nx, ny, nz), h,
                                        sw.CPMLBoundaryConditionParameters(halo=halo, freeboundtop=False))
    return dict(name=f"acoustic_cd_3d_gradient_600x600x400_{nshots}shots_nt{nt}", params=params, physics="cd3d",
                mp_true=sw.VpAcousticCDMaterialProperties(vtrue), mp0=sw.VpAcousticCDMaterialProperties(v0),
                shots=shots, check_freq=50, run={},
                config={"nx": nx, "ny": ny, "nz": nz, "dh_m": 10.0, "dt_s": dt, "f0_hz": f0, "nt": nt,
                        "shots": nshots, "of_shots": 8, "receivers": 10000, "check_freq": 50, "halo": halo})


def workload_vd2d_batched_small():
    """The batched VD workload cut to 8 shots of nt 300..500 (for profiling
    runs, where every dispatch is instrumented)."""
    wl = workload_vd2d_batched()
    import seismicwaves as sw
    rng = np.random.default_rng(5)
    shots = []
    for sh in wl["shots"][:8]:
        nt = int(rng.integers(300, 501))
        shots.append(sw.ScalarShot(sw.ScalarSources(sh.srcs.positions, sh.srcs.tf[:nt], sh.srcs.domfreq),
                                   sw.ScalarReceivers(sh.recs.positions, nt, dtype=np.float32)))
    wl["shots"] = shots
    wl["name"] = "acoustic_vd_2d_gradient_2000x800_8shots_nt300-500"
    nts = [s.recs.seismograms.shape[0] for s in shots]
    wl["config"].update({"shots": 8, "nt_min": min(nts), "nt_max": max(nts), "sum_nt": sum(nts)})
    return wl


def workload_vd2d_batched_nt10():
    """The batched VD workload with every shot's nt cut by ten (314..499):
    the same grid, shots, receivers and checkpoint interval, so a launch
    steps the same cells of the same shots and the share of active shots
    over the run is the same; for counter-collection runs, where every
    dispatch is instrumented."""
    wl = workload_vd2d_batched()
    import seismicwaves as sw
    shots = []
    for sh in wl["shots"]:
        nt = sh.recs.seismograms.shape[0] // 10
        shots.append(sw.ScalarShot(sw.ScalarSources(sh.srcs.positions, sh.srcs.tf[:nt], sh.srcs.domfreq),
                                   sw.ScalarReceivers(sh.recs.positions, nt, dtype=np.float32)))
    wl["shots"] = shots
    nts = [s.recs.seismograms.shape[0] for s in shots]
    wl["name"] = f"acoustic_vd_2d_gradient_2000x800_64shots_nt{min(nts)}-{max(nts)}"
    wl["config"].update({"nt_min": min(nts), "nt_max": max(nts), "sum_nt": sum(nts)})
    return wl


def workload_el2d_batched(nt=5000):
    """Batched 2D isotropic elastic gradient w.r.t. rho, lambda, mu:
    1500 x 600, dh = 5 m, dt = 0.4 ms, nt = 5000, free surface on top, C-PML
    (20 cells) elsewhere.  vp: 8 seeded layers 1800-4500 m/s, vs = vp/1.8
    with +-10% noise, Gardner density.  32 shots of 15 Hz Ricker sources at
    seeded x along z = 10 m: 16 explosive moment tensors (Mxx = Mzz = 1e10)
    and 16 vertical forces, run as two calls (one source kind per call, as
    the reference requires); 200-600 receivers per shot at z = 0;
    checkpoint every 50 steps.  The starting model is the true one smoothed
    with sigma = 100 m."""
    import seismicwaves as sw
    from seismicwaves import synthetic as syn
    nx, nz = 1500, 600
    h = (5.0, 5.0)
    dt, f0, halo, nshots = 4e-4, 15.0, 20, 32
    rng = np.random.default_rng(4)
    vp = syn.layered(nx, nz, h, rng, vmin=1800.0, vmax=4500.0)
    vs = vp / 1.8 * (1 + 0.1 * rng.uniform(-1, 1, vp.shape))
    rho = syn.gardner(vp, rng)
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    sm = lambda m: syn.gaussian_smooth(m, 100.0 / h[0])  # noqa: E731
    f32 = np.float32
    tf = syn.ricker_tf(nt, dt, f0, 1.2 / f0).astype(f32)
    xs = rng.uniform(halo * h[0], (nx - 1 - halo) * h[0], nshots)
    nrecs = rng.integers(200, 601, nshots)
    groups = [[], []]
    for k, (x, nr) in enumerate(zip(xs, nrecs)):
        rec = syn.surface_receivers(int(nr), nx, h, halo, 0.0)
        src = np.array([[x, 10.0]])
        if k < nshots // 2:
            srcs = sw.MomentTensorSources(src, tf[:, None], [sw.MomentTensor2D(1e10, 1e10, 0.0)], f0)
            groups[0].append(sw.MomentTensorShot(srcs, sw.VectorReceivers(rec, nt, dtype=f32)))
        else:
            stf = np.zeros((nt, 2, 1), dtype=f32)
            stf[:, 1, 0] = 1e10 * tf
            groups[1].append(sw.ExternalForceShot(sw.ExternalForceSources(src, stf, f0),
                                                  sw.VectorReceivers(rec, nt, dtype=f32)))
    params = sw.InputParametersElastic(nt, dt, (nx, nz), h, sw.CPMLBoundaryConditionParameters(halo=halo))
    mpt = sw.ElasticIsoMaterialProperties(lam.astype(f32), mu.astype(f32), rho.astype(f32))
    mp0 = sw.ElasticIsoMaterialProperties(sm(lam).astype(f32), sm(mu).astype(f32), sm(rho).astype(f32))
    return dict(name=f"elastic_2d_gradient_1500x600_32shots_nt{nt}", params=params, physics="el2d",
                mp_true=mpt, mp0=mp0, shots=groups[0] + groups[1], groups=groups, check_freq=50,
                run={"erroronPPW": False},
                config={"nx": nx, "nz": nz, "dh_m": 5.0, "dt_s": dt, "f0_hz": f0, "shots": nshots, "nt": nt,
                        "receivers": f"{int(nrecs.min())}-{int(nrecs.max())}", "check_freq": 50, "halo": halo,
                        "sources": "16 explosive moment tensors + 16 vertical forces (two calls)"})


WORKLOADS = {"vd2d_batched": workload_vd2d_batched, "cd2d_grad_1000": workload_cd2d_grad_1000,
             "vd2d_batched_small": workload_vd2d_batched_small, "el2d_batched": workload_el2d_batched,
             "vd2d_batched_nt10": workload_vd2d_batched_nt10, "cd3d_600x600x400": workload_cd3d_600x600x400,
             "el2d_batched_small": lambda: workload_el2d_batched(nt=500),
             # counter-collection runs (every dispatch instrumented): the same
             # grids, shots and receivers with short time loops
             "el2d_batched_nt100": lambda: workload_el2d_batched(nt=100),
             "cd3d_600x600x400_nt60": lambda: workload_cd3d_600x600x400(nt=60),
             "cd2d_grad_1000_nt400": lambda: workload_cd2d_grad_1000(nt=400)}


# ---------------------------------------------------------------------------
# Algorithmic bytes per launch of the step kernels (each array counted once
# per launch, perfect halo reuse).  e = bytes per value.
# ---------------------------------------------------------------------------
def _strip_cells(n, halo):
    # C-PML memory variables touched per launch: psi (2h per axis) and xi
    # (2h+2 per axis) over the face of each axis, each read and written once.
    N = int(np.prod(n))
    return sum((2 * halo + 2 * halo + 2) * (N // nd) for nd in n)


def _el_strip_cells(n, halo):
    # the eight elastic C-PML memory variables: per axis two arrays of 2h
    # and two of 2h+2 rows (ela_models.jl:276-299)
    nx, nz = n
    return (8 * halo + 4) * nz + (8 * halo + 4) * nx


def _launches_two_step(first, nt, ntmax, S, checkpoints):
    """Launches of the two-step forward schedule that touch a shot of nt
    steps: pairs (it, it+1) from `first` on, except a single step where it
    is a checkpoint level or it+1 would pass ntmax."""
    n, it = 0, first
    while it <= min(nt, ntmax):
        two = (not checkpoints or it % S != 0) and it + 1 <= ntmax
        n += 1
        it += 2 if two else 1
    return n


def alg_bytes(physics, n, halo, e, nts, S, variant=(0, 0)):
    """Returns {kernel family: (total bytes of one gradient, launches)} for
    family 0 (forward/replay step kernel) and 1 (adjoint step kernel);
    `variant` holds the kernel variants the plan picked."""
    N = int(np.prod(n))
    ntmax = max(nts)
    strip = _strip_cells(n, halo) * 2 * e
    fwd_b = fwd_n = adj_b = adj_n = 0
    # VD shots whose model factors one block (4xx: two shots per block) or
    # one group of the block order (code + 1000 k: 2^k shots) read together
    vd_groups = tuple(2 if 400 <= v % 1000 < 500 else 1 << (v // 1000) for v in variant)
    for nt in nts:
        if physics == "vd2d":
            # the final nt - floor(nt/S)*S steps are stored during the
            # forward; every earlier segment replays S-1 steps
            last = nt - (nt // S) * S
            rep = (nt // S) * (S - 1)
        else:
            kl = (nt - 1) // S
            last = nt - kl * S              # steps of the last segment: stored during the forward
            rep = nt - last                 # steps replayed from checkpoints
        if physics == "el2d":
            # The ideal elastic step keeps the stresses on chip: forward r ux, uz
            # at two levels (4e), lambda, mu, mu_ihjh, dt2/rho_x, dt2/rho_z (5e),
            # w ux, uz (2e) = 11e; + w the u history (13e) on stored steps.
            # Adjoint 11e + r 3 history levels x 2 components (6e) + 5 gradient
            # fields read and written (10e) = 27e.  C-PML: the 8 psi arrays read
            # and written once per step.  A variant with two launches per step
            # (stress, displacement) is priced at the same bytes per step.
            # (one launch per step: the z-marching variants > 100 and the
            # fused tile step 2)
            march_f, march_a = variant[0] > 100 or variant[0] == 2, variant[1] > 100 or variant[1] == 2
            last = nt - (nt // S) * S
            rep = (nt // S) * (S - 1)
            est = _el_strip_cells(n, halo) * 2 * e
            fwd_b += (nt - last) * 11 * e * N + (last + rep) * 13 * e * N + (nt + rep) * est
            fwd_n += (1 if march_f else 2) * (nt + rep)
            adj_b += nt * 27 * e * N + nt * est
            adj_n += (1 if march_a else 2) * nt
            continue
        if physics == "vd2d" and (vd_groups[0] > 1 or vd_groups[1] > 1):
            # two shots per block (4xx) or shot groups in the block order
            # (code + 1000 k): the model factors m0, m1x, m1z (3e) are read
            # once per group of shots and step (counted below per group); per
            # shot forward r/w p, vx, vz (6e) + w p-history, adjoint 14e
            pf, pa = vd_groups[0] > 1, vd_groups[1] > 1
            fwd_b += ((nt + rep) * (6 if pf else 9) + (last + rep)) * e * N
            fwd_b += (nt + rep) * strip
            fwd_n += nt + rep
            adj_b += nt * (14 if pa else 17) * e * N + nt * strip
            adj_n += nt
            continue
        if physics == "vd2d" and 200 <= variant[0] < 300:
            # two steps per launch: a launch reads p, vx, vz, m0, m1x, m1z and
            # writes p, vx, vz once (9e) whether it advances the shot one or two
            # steps; + w p-history (1e) per stored step
            nl_f = _launches_two_step(1, nt, ntmax, S, checkpoints=True)
            nl_r = (nt // S) * ((S - 1 + 1) // 2)
            fwd_b += (nl_f + nl_r) * (9 * e * N + strip) + (last + rep) * e * N
            fwd_n += nl_f + nl_r
        elif physics == "vd2d":
            # forward: r p, vx, vz, m0, m1x, m1z; w p, vx, vz (9e); + w p-history (10e)
            fwd_b += (nt - last) * 9 * e * N + last * 10 * e * N + nt * strip
            fwd_b += rep * 10 * e * N + rep * strip
            fwd_n += nt + rep
            # adjoint: 9e + r p^it, p^it-1 (2e) + grad_m0/m1x/m1z read-modify-write (6e)
            adj_b += nt * 17 * e * N + nt * strip
            adj_n += nt
        else:
            # CD: r p^n, p^n-1, fact; w p^n+1 (4e); + w D (5e).  Adjoint with
            # the fused correlation: + r D, grad read-modify-write (7e).
            if physics == "cd2d" and len(variant) > 1 and variant[1] % 10 == 2:
                # two levels per forward / replay launch: r p^n, p^n-1, fact
                # once per launch (3e), w each level (1e) and its D (1e);
                # launches start at levels 1, 3, 5, ... of the forward and
                # of every replayed segment
                nl = (nt + 1) // 2 + kl * ((S + 1) // 2)
                fwd_b += nl * 3 * e * N + (nt + rep) * e * N + (last + rep) * e * N + (nt + rep) * strip
                fwd_n += nl
            else:
                fwd_b += (nt - last) * 4 * e * N + last * 5 * e * N + nt * strip
                fwd_b += rep * 5 * e * N + rep * strip
                fwd_n += nt + rep
            nseg = kl + 1
            if physics == "cd2d" and len(variant) > 1 and variant[1] // 10 == 2:
                # two adjoint levels per launch: r a^{n+1}, a^{n+2}, fact once
                # (3e), w both levels (2e), the gradient read and written once
                # (2e) when either level correlates and one history level per
                # correlating level (1e); the top level of every segment does
                # not correlate (its correlation is in the next segment's
                # separate kernel), a lone level takes a one-level launch (7e,
                # or 4e at a segment top)
                for k in range(nseg):
                    L = min((k + 1) * S, nt) - k * S
                    for top in range(L, 0, -2):
                        if top >= 2:
                            ncorr = 2 if top < L else 1
                            adj_b += 5 * e * N + 2 * e * N + ncorr * e * N + 2 * strip
                            adj_n += 1
                        else:
                            adj_b += (7 if top < L else 4) * e * N + strip
                            adj_n += 1
            else:
                adj_b += (nt - nseg) * 7 * e * N + nseg * 4 * e * N + nt * strip
                adj_n += nt
    if physics == "vd2d":
        # model factors once per group and step the group steps (its
        # longest shot's forward and replay steps; adjoint steps)
        gf, ga = vd_groups
        if gf > 1:
            for k in range(0, len(nts), gf):
                fwd_b += 3 * e * N * max(nt + (nt // S) * (S - 1) for nt in nts[k:k + gf])
        if ga > 1:
            for k in range(0, len(nts), ga):
                adj_b += 3 * e * N * max(nts[k:k + ga])
    return {0: (fwd_b, fwd_n), 1: (adj_b, adj_n)}


def kernel_names(physics, variant):
    """Names of the forward and adjoint step kernels the plan's autotuner
    picked (variant codes of swx_plan_info.kernel_variant)."""
    if physics == "vd2d":
        def name(kind, v):
            v %= 1000      # the thousands give the shot groups of the block order
            if v >= 400:   # two shots per block, one launch
                return f"vd2d_{kind}_pair_kernel<float, {v - 400}>"
            if v >= 300:   # interior and other tiles in one launch
                return f"vd2d_{kind}_kernel<float, {v - 300}, true>"
            if v >= 200:   # two steps per launch
                return f"vd2d_{kind}2_kernel<float, {v - 200}, {512 if v - 200 > 8 else 256}>"
            if v > 100:
                return f"vd2d_{kind}_march<float, {v - 100}>"
            # interior tiles and the other tiles: two kernels per step
            return f"vd2d_{kind}_lean<float, {v}> + vd2d_{kind}_kernel<float, {v}, false>"
        return name("fwd", variant[0]), name("adj", variant[1])
    if physics == "el2d":
        def name(adj, v):
            if v > 100:
                return f"el_march_kernel<float, {v - 100}, {'true' if adj else 'false'}>"
            if v == 2:   # stresses in LDS, one launch per step
                return f"el_tile_kernel<float, {'true' if adj else 'false'}>"
            tail = " (adjoint)" if adj else " (forward)"
            if v == 1:   # interior and other blocks in one launch per kernel
                return "el_stress_kernel<float, true> + el_disp_kernel<float, true>" + tail
            return ("el_stress_lean<float> + el_stress_kernel<float, false> + el_disp_lean<float> + "
                    "el_disp_kernel<float, false>" + tail)
        return name(False, variant[0]), name(True, variant[1])
    if physics == "cd3d":
        # [launch shape, planes per z chunk]
        v = variant[0] or 4
        if v > 200:   # one launch, interior march staged through LDS
            k = f"cd3d_step_kernel<float, {v - 200}, 3>"
        elif v > 100:   # interior tiles and the others in two launches
            k = f"cd3d_step_kernel<float, {v - 100}, 1> + cd3d_step_kernel<float, {v - 100}, 2>"
        else:
            k = f"cd3d_step_kernel<float, {v}, 0>"
        return f"{k} (forward)", f"{k} (adjoint)"
    # [rows per thread, levels per launch: 10 x adjoint + forward / replay]
    k = f"cd2d_step_kernel<float, {variant[0] or 2}>"
    two = variant[1] if len(variant) > 1 else 0
    fwd = "cd2d_step2_kernel<float, false>" if two % 10 == 2 else k
    adj = "cd2d_step2_kernel<float, true>" if two // 10 == 2 else k
    return f"{fwd} (forward)", f"{adj} (adjoint)"


def traffic_from_profiles(workload, kernel):
    """HBM bytes per dispatch of `kernel` on `workload` as recorded in the
    rocprof summaries under profiles/ (a line `TRAFFIC <workload> <kernel>
    <bytes>`), or None when no summary covers it."""
    import glob
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*.md"))):
        for line in open(path):
            parts = line.strip().strip("`").split()
            if len(parts) >= 4 and parts[0] == "TRAFFIC" and parts[1] == workload and \
                    " ".join(parts[2:-1]) == kernel:
                return int(float(parts[-1]))
    return None


# ---------------------------------------------------------------------------
# Gradient driver on an already-populated plan
# ---------------------------------------------------------------------------
class Runner:
    """One gradient of a group of shots per step, through the same calls as
    seismicwaves.swgradient: batches as large as the HBM budget allows
    (one plan, reused), forward with device checkpoints and the traces
    copied to the host, the L2 adjoint source and misfit formed on the
    device, adjoint with replay and correlation.  With one batch the shots
    and observed data stay on the device between steps, as in an inversion
    loop over a fixed acquisition."""

    def __init__(self, wl, shots, max_batch=None):
        import seismicwaves as sw
        p = wl["params"]
        self.wl = wl
        self.ws = sw.build_wavesim(p, wl["mp0"], runparams=sw.RunParameters(max_batch=max_batch, **wl["run"]),
                                   gradparams=sw.GradParameters(check_freq=wl["check_freq"]), gradient=True)
        self.ws.set_matprop(wl["mp0"])
        self.descs = [self.ws.shot_desc(sh) for sh in shots]
        self.observed = [np.asfortranarray(sh.recs.seismograms, dtype=np.float32).ravel(order="F") for sh in shots]
        self.batches = [b for b, _ in self.ws.batches(len(self.descs))]
        self.plan = self.ws._plan
        size = sum(int(np.prod(s)) for s in self.ws.grad_fields_shapes())
        self.grad = np.zeros(size, dtype=np.float32)
        self.gsum = np.zeros(size, dtype=np.float32)
        self._bufs = []
        for b in self.batches:
            ntr = sum(int(np.prod(self.descs[k]["tshape"])) for k in b)
            self._bufs.append((np.zeros(ntr, dtype=np.float32), np.concatenate([self.observed[k] for k in b])))
        self._loaded = -1
        if len(self.batches) == 1:
            self._load(0)

    def _load(self, i):
        if self._loaded == i:
            return
        self.ws.upload_shots(self.plan, [self.descs[k] for k in self.batches[i]])
        self.plan.set_observed(self._bufs[i][1])
        self._loaded = i

    def step(self):
        """One gradient; `device_ms` is the time between stream events
        around each forward and adjoint loop of it (no per-launch events)."""
        self.gsum[:] = 0
        self.device_ms = 0.0
        for i, b in enumerate(self.batches):
            self._load(i)
            self.plan.forward(self._bufs[i][0], True)
            self.device_ms += self.plan.info().ms_forward
            self.plan.adjoint_l2(self.grad, False, len(b))
            self.device_ms += self.plan.info().ms_adjoint
            self.gsum += self.grad

    def timed_kernels(self):
        """One gradient with HIP events around every step-kernel launch;
        returns the kernel times and launch counts and the device time."""
        self.plan.set_kernel_timing(True)
        ms, n = [0.0, 0.0], [0, 0]
        dev = 0.0
        for i, b in enumerate(self.batches):
            self._load(i)
            self.plan.forward(self._bufs[i][0], True)
            inf = self.plan.info()
            dev += inf.ms_forward
            for k in range(2):
                ms[k] += inf.ms_kernel[k]
                n[k] += inf.n_kernel[k]
            self.plan.adjoint_l2(self.grad, False, len(b))
            inf = self.plan.info()
            dev += inf.ms_adjoint
            for k in range(2):
                ms[k] += inf.ms_kernel[k]
                n[k] += inf.n_kernel[k]
        self.plan.set_kernel_timing(False)
        return ms, n, dev

    def close(self):
        self.plan.close()


# ---------------------------------------------------------------------------
# CPU baseline: oracle OpenMP restatement of the reference's threads backend
# ---------------------------------------------------------------------------
def cpu_baseline(wl, nshots_sample=4, nt_sample=1000):
    import oracle
    oracle.build()
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    p = wl["params"]
    if wl["physics"] == "el2d":
        return cpu_baseline_elastic(wl, cores, nt_sample=300)
    if wl["physics"] == "cd3d":
        return cpu_baseline_cd3d(wl, cores)
    nx, nz = p.gridsize
    h = p.gridspacing
    dt = float(p.dt)
    halo, rcoef, ft = p.boundcond.halo, p.boundcond.rcoef, p.boundcond.freeboundtop
    T = np.float32
    cf = wl["check_freq"]
    vp = np.asarray(wl["mp0"].vp, dtype=np.float64)
    shots = wl["shots"][:nshots_sample]
    rng = np.random.default_rng(0)
    el = 0.0
    for sh in shots:
        f0 = float(sh.srcs.domfreq)
        cp = oracle.init_bdc(vp, dt, h, halo, rcoef, ft, f0, T)
        si = oracle.find_nearest_grid_points(sh.srcs.positions, h)
        ri = oracle.find_nearest_grid_points(sh.recs.positions, h)
        adj = (rng.standard_normal((nt_sample, ri.shape[0])) * 1e-3).astype(T)
        if wl["physics"] == "vd2d":
            rho = np.asarray(wl["mp0"].rho, dtype=np.float64)
            m0, m1x, m1z = oracle.vd_model_fields(vp, rho, dt, "arithmetic", T)
            stf = oracle.scale_srctf_vd(sh.srcs.tf[:nt_sample], h, dt, vp, rho, si, T)
            t0 = time.perf_counter()
            oracle.vd2d_gradient(m0, m1x, m1z, h, dt, halo, cp, nt_sample, si, stf, ri, adj, cf, T)
        else:
            stf = oracle.scale_srctf_cd(sh.srcs.tf[:nt_sample], h, dt, vp, si, T)
            fact = ((T(dt) ** 2) * vp.astype(T) ** 2).astype(T)
            t0 = time.perf_counter()
            oracle.cd2d_gradient(fact, h, dt, halo, cp, nt_sample, si, stf, ri, adj, cf, T)
        el += time.perf_counter() - t0
    work = 2.0 * nx * nz * nt_sample * len(shots)
    return {"value": round(work / el / 1e9, 4), "unit": "GCell-updates/s", "cores": cores, "kind": "port",
            "sample": f"FP32 gradient of the same model: {len(shots)} shot(s) x nt={nt_sample} "
                      f"(check_freq={cf}), sequential shots, OpenMP threads={cores}; C restatement of the "
                      f"reference's threads backend (not the Julia binary); {el:.2f} s"}


def cpu_baseline_cd3d(wl, cores, nt_sample=100, cf=20):
    """The oracle's FP32 OpenMP 3D CD gradient on the workload's first shot,
    nt_sample steps with a checkpoint every cf steps."""
    import oracle
    p = wl["params"]
    h, dt = p.gridspacing, float(p.dt)
    bc = p.boundcond
    T = np.float32
    vp = np.asarray(wl["mp0"].vp, dtype=np.float64)
    sh = wl["shots"][0]
    cp = oracle.init_bdc(vp, dt, h, bc.halo, bc.rcoef, bc.freeboundtop, float(sh.srcs.domfreq), T)
    si = oracle.find_nearest_grid_points(sh.srcs.positions, h)
    ri = oracle.find_nearest_grid_points(sh.recs.positions, h)
    stf = oracle.scale_srctf_cd(sh.srcs.tf[:nt_sample], h, dt, vp, si, T)
    fact = ((T(dt) ** 2) * vp.astype(T) ** 2).astype(T)
    adj = (np.random.default_rng(0).standard_normal((nt_sample, ri.shape[0])) * 1e-3).astype(T)
    t0 = time.perf_counter()
    oracle.cd3d_gradient(fact, h, dt, bc.halo, cp, nt_sample, si, stf, ri, adj, cf, T)
    el = time.perf_counter() - t0
    cells = int(np.prod(p.gridsize))
    return {"value": round(2.0 * cells * nt_sample / el / 1e9, 4), "unit": "GCell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"FP32 gradient of the same model: 1 shot x nt={nt_sample} (check_freq={cf}), OpenMP "
                      f"threads={cores}; C restatement of the reference's threads backend (not the Julia "
                      f"binary); {el:.2f} s"}


def cpu_baseline_elastic(wl, cores, nt_sample=300):
    """The oracle's FP32 OpenMP elastic gradient on one moment-tensor shot of
    the workload, nt_sample steps."""
    import oracle
    p = wl["params"]
    h, dt = p.gridspacing, float(p.dt)
    bc = p.boundcond
    T = np.float32
    mp = wl["mp0"]
    sh = wl["groups"][0][0]
    m = sh.srcs.momtens[0]
    osh = dict(src_pos=sh.srcs.positions, srctf=sh.srcs.tf[:nt_sample], moments=np.array([[m.Mxx, m.Mzz, m.Mxz]]),
               f0=float(sh.srcs.domfreq), rec_pos=sh.recs.positions, nt=nt_sample)
    fields, cpml, kind, fams, stf, mom = oracle._el_prepare(mp.lam, mp.mu, mp.rho, h, dt, bc.halo, bc.rcoef,
                                                           bc.freeboundtop, osh, True, T)
    nrec = sh.recs.positions.shape[0]
    adj = (np.random.default_rng(0).standard_normal((nt_sample, 2, nrec)) * 1e-3).astype(T)
    t0 = time.perf_counter()
    oracle.el2d_gradient(fields, h, dt, bc.halo, bc.freeboundtop, cpml, nt_sample, kind, fams, mom, stf, adj,
                         wl["check_freq"], T)
    el = time.perf_counter() - t0
    nx, nz = p.gridsize
    return {"value": round(2.0 * nx * nz * nt_sample / el / 1e9, 4), "unit": "GCell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"FP32 elastic gradient of the same model: 1 moment-tensor shot x nt={nt_sample} "
                      f"(check_freq={wl['check_freq']}), OpenMP threads={cores}; C restatement of the "
                      f"reference's threads backend (not the Julia binary); {el:.2f} s"}


# ---------------------------------------------------------------------------
def warm_process(workload):
    """Runs one gradient step of the same workload in a child process that
    then exits, before the measurement.  On a freshly started box the first
    process that runs a large workload runs the step kernels slower: the
    64-shot VD gradient measured 68.1 GCell-updates/s in the first process
    and 74.8 in the next one on the same box (adjoint launch 1176 -> 998 us,
    forward 654 -> 631 us); touching the memory from a child process without
    running the workload did not always reach the second state.  A device
    that has run anything before, as a serving or inversion host has, is in
    the second state.  Nothing of the measured work is done here: the child
    builds its own plan and data."""
    import subprocess
    t0 = time.perf_counter()
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", workload, "--steps", "1", "--warmup", "0",
           "--no-cpu", "--no-one-shot", "--timed-steps-only", "--no-warm-process"]
    # the child's progress lines go to this process's stderr
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    if r.returncode != 0:
        return f"failed ({r.returncode})"
    return f"one {workload} gradient step in a child process, {time.perf_counter() - t0:.0f} s"


def hbm_triad(bytes_each=4 << 30, reps=20):
    """The device's practical HBM ceiling: a STREAM triad a = b + s c on
    float4 vectors (seismicwaves.jl_amd/probe/hbm_triad.hip), 3 arrays of
    `bytes_each`, best and mean GB/s over `reps` launches of 8 blocks per CU
    and of 1 block per 4 KiB."""
    import ctypes
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "seismicwaves.jl_amd", "probe",
                        "libswx_probe.so")
    lib = ctypes.CDLL(path)
    lib.swx_probe_hbm_triad_mode.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_double)]
    out = {}
    for mode, tag in ((0, ""), (1, "_nt_stores"), (3, "_nt_loads_stores")):
        for name, blocks in (("8_blocks_per_cu", 0), ("1_block_per_4KiB", bytes_each // 4096)):
            r = (ctypes.c_double * 2)()
            rc = lib.swx_probe_hbm_triad_mode(bytes_each, reps, int(blocks), mode, r)
            if rc != 0:
                raise RuntimeError(f"swx_probe_hbm_triad_mode failed ({rc})")
            out[name + tag] = {"best_GBs": round(r[0], 1), "mean_GBs": round(r[1], 1)}
    return {"probe": "hbm_triad", "bytes_per_array": bytes_each, "reps": reps, "peak_GBs": 8000.0, **out}


def launch_gap(launches=2000, reps=5):
    """What a HIP graph of a time loop saves: a chain of `launches` dependent
    three-point passes over n floats (seismicwaves.jl_amd/probe/hbm_triad.hip,
    swx_probe_launch_chain), stream-launched and replayed from a captured
    graph; us per pass (best of `reps` chains) and of one pass alone."""
    import ctypes
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "seismicwaves.jl_amd", "probe",
                        "libswx_probe.so")
    lib = ctypes.CDLL(path)
    lib.swx_probe_launch_chain.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double)]
    rows = []
    for n in (1 << 12, 1 << 18, 1 << 20, 1 << 22, 1 << 24):
        r = (ctypes.c_double * 4)()
        rc = lib.swx_probe_launch_chain(n, launches, reps, r)
        if rc != 0:
            raise RuntimeError(f"swx_probe_launch_chain failed ({rc})")
        rows.append({"floats": n, "stream_us_per_pass": round(r[0], 2), "graph_us_per_pass": round(r[1], 2),
                     "one_pass_us": round(r[2], 2), "capture_instantiate_ms": round(r[3], 1)})
        print(json.dumps(rows[-1]), flush=True)
    return {"probe": "launch_chain", "launches": launches, "reps": reps, "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="vd2d_batched", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-one-shot", action="store_true", help="skip the plain 1-shot path")
    ap.add_argument("--forward-only", action="store_true",
                    help="time only the forward pass with checkpointing (for counter-collection runs): no "
                         "observed-data pass, no adjoint, no roofline")
    ap.add_argument("--timed-steps-only", action="store_true",
                    help="stop after the timed steps (for counter-collection runs, where every dispatch is "
                         "instrumented): no per-launch timing pass, no roofline")
    ap.add_argument("--one-shot-sample", type=int, default=2, help="shots timed on the 1-shot path")
    ap.add_argument("--synthetic-observed", action="store_true",
                    help="seeded noise as observed data instead of a forward run on the true model (for "
                         "counter-collection runs: then the only step-kernel dispatches are the gradients')")
    ap.add_argument("--no-warm-process", action="store_true",
                    help="measure without first running one step of the workload in a child process")
    ap.add_argument("--hbm-triad", action="store_true",
                    help="only measure the HBM ceiling: a STREAM triad on 3 x 4 GiB (probe/libswx_probe.so)")
    ap.add_argument("--launch-gap", action="store_true",
                    help="only measure launch gaps: dependent passes stream-launched against a replayed HIP graph "
                         "(probe/libswx_probe.so)")
    args = ap.parse_args()
    if args.launch_gap:
        print(json.dumps(launch_gap()))
        return
    if args.hbm_triad:
        print(json.dumps(hbm_triad()))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build targets one GPU: shots would be sharded across GPUs with one "
                                   "gradient all-reduce, which is described but not run here",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)  # noqa: E731
    warm = None
    if not args.no_warm_process:
        warm = warm_process(args.workload)
        log("warm-up process:", warm)
    import seismicwaves as sw
    wl = WORKLOADS[args.workload]()
    p = wl["params"]
    log("workload", wl["name"])
    # observed data from the true model (outside the timed region)
    if args.forward_only:
        runs = [Runner(wl, g) for g in wl.get("groups", [wl["shots"]])]
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for r in runs:
                for i in range(len(r.batches)):
                    r._load(i)
                    r.plan.forward(r._bufs[i][0], True)
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        for r in runs:
            r.close()
        print(json.dumps({"workload": wl["name"], "forward_only": True, "steps": args.steps,
                          "ms_per_forward": round(ms, 3)}))
        return
    if args.synthetic_observed:
        # seeded noise in place of the true model's traces: the bytes a
        # gradient moves do not depend on the data
        rng = np.random.default_rng(7)
        for g in wl.get("groups", [wl["shots"]]):
            for sh in g:
                sh.recs.seismograms[...] = 1e-3 * rng.standard_normal(sh.recs.seismograms.shape)
    else:
        for g in wl.get("groups", [wl["shots"]]):
            sw.swforward(p, wl["mp_true"], g, runparams=sw.RunParameters(**wl["run"]))
    log("observed data ready")
    # one plan per call: the elastic workload runs its two source kinds as
    # two calls, as the reference requires
    runs = [Runner(wl, g) for g in wl.get("groups", [wl["shots"]])]
    nts = [d["nt"] for r in runs for d in r.descs]
    for _ in range(args.warmup):
        for r in runs:
            r.step()
    log("warmup done")
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(args.steps):
        for r in runs:
            r.step()
            dev_ms += r.device_ms
    wall = time.perf_counter() - t0
    dev_ms /= max(args.steps, 1)
    ms_step = wall / args.steps * 1e3
    log(f"timed {args.steps} steps: {ms_step:.1f} ms/step")
    if args.timed_steps_only:
        for r in runs:
            r.close()
        print(json.dumps({"workload": wl["name"], "steps": args.steps, "ms_per_step": round(ms_step, 3)}))
        return
    kms, kn = [0.0, 0.0], [0, 0]
    for r in runs:
        m, n, _ = r.timed_kernels()
        for k in range(2):
            kms[k] += m[k]
            kn[k] += n[k]
    info = runs[0].plan.info()
    batches = [len(b) for r in runs for b in r.batches]
    S = info.segment
    for r in runs:
        r.close()
    cells = int(np.prod(p.gridsize))
    value = 2.0 * cells * sum(nts) / (ms_step / 1e3) / 1e9
    ab = alg_bytes(wl["physics"], p.gridsize, p.boundcond.halo, 4, nts, S, list(info.kernel_variant))
    fam = {}
    for k, name in ((0, "forward/replay step"), (1, "adjoint step + correlation")):
        b, _ = ab[k]
        avg_ms = kms[k] / max(kn[k], 1)
        per_launch = b / max(kn[k], 1)     # one launch steps every active shot of the batch
        fam[k] = {"name": name, "launches": kn[k], "avg_launch_us": round(avg_ms * 1e3, 2),
                  "alg_bytes_per_launch": int(per_launch),
                  "achieved_GBs": round(per_launch / (avg_ms / 1e3) / 1e9, 1), "total_ms": round(kms[k], 1)}
    dom = max(fam, key=lambda k: fam[k]["total_ms"])
    kern = kernel_names(wl["physics"], list(info.kernel_variant))
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered model, Ricker sources, observed data from a forward run)",
        "config": {"workload": wl["name"], **wl["config"], "segment_used": S, "batches": batches,
                   "kernel_variant": [int(v) for v in info.kernel_variant]},
        "roofline": {"bound": "hbm", "achieved": fam[dom]["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(fam[dom]["achieved_GBs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(wl["name"], kern[dom]),
                     "kernel": kern[dom], "avg_launch_us": fam[dom]["avg_launch_us"],
                     "alg_bytes_per_launch": fam[dom]["alg_bytes_per_launch"]},
        "kernels": {kern[k]: fam[k] for k in fam},
        "device_ms_per_step": round(dev_ms, 1),
        "warm_process": warm or "off",
    }
    if not args.no_one_shot and len(wl["shots"]) > 1:
        # the same kernels on a plan holding one shot, shots run in sequence
        sample = wl.get("groups", [wl["shots"]])[0][:args.one_shot_sample]
        one = Runner(wl, sample, max_batch=1)
        one.step()    # warm-up
        t0 = time.perf_counter()
        one.step()
        el = time.perf_counter() - t0
        one.close()
        v1 = 2.0 * cells * sum(d["nt"] for d in one.descs) / el / 1e9
        out["one_shot"] = {"value": round(v1, 3), "unit": "GCell-updates/s",
                           "sample": f"{len(sample)} shots run one at a time on a 1-shot plan",
                           "batched_over_one_shot": round(value / v1, 3)}
        log("one-shot path done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl)
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 2)
        log("cpu baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
